"""GPU parity: the HIP path through the C ABI against the FP64 CPU oracle.

Tolerances are relative max-norm errors, max|gpu - ref| / scale, with the
field's own maximum as scale for rho and T and the velocity scale
max(|u|, |v|, |w|) for each velocity component:
  * FP64 GPU path vs oracle: <= 1e-13 (the two compute the same operations
    in the same order; the test also reports whether they are bit-exact);
  * FP32 shifted GPU path vs FP64: <= 1e-5 on u, v, w, T and 1e-4 relative
    on the mean Nusselt numbers.
"""
import numpy as np
import pytest

from helpers import field_errors, rel_max_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-13
FP32_TOL = 1e-5
NU_TOL = 1e-4


def _pair(oracle_mod, hip, dim, n, Ra, precision, **kw):
    o = oracle_mod.Oracle(dim, n, Ra, **{k: v for k, v in kw.items()
                                         if k in ("Pr", "Ma", "Thot", "Tcold", "A")})
    skw = {k: v for k, v in kw.items() if k in ("Pr", "Ma", "Thot", "Tcold")}
    if "A" in kw:
        skw["stretch"] = kw["A"]
    s = hip.Solver(dim, n, Ra, precision=precision, **skw)
    return o, s


def _smooth_fields(dim, n, seed):
    rng = np.random.default_rng(seed)
    t = (np.arange(n) + 0.5) / n
    if dim == 2:
        Y, X = np.meshgrid(t, t, indexing="ij")
        base = [np.sin(np.pi * X) * np.sin(np.pi * Y)]
        amp = rng.uniform(0.5, 1.0, size=4)
        u = 0.02 * amp[0] * np.sin(2 * np.pi * Y) * base[0]
        v = -0.02 * amp[1] * np.sin(2 * np.pi * X) * base[0]
        T = 1.0 - X + 0.05 * amp[2] * np.sin(3 * np.pi * X) * np.cos(np.pi * Y)
        return u, v, None, T
    Z, Y, X = np.meshgrid(t, t, t, indexing="ij")
    amp = rng.uniform(0.5, 1.0, size=4)
    b = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
    u = 0.02 * amp[0] * np.sin(2 * np.pi * Z) * b
    v = 0.015 * amp[1] * np.cos(np.pi * X) * b
    w = -0.02 * amp[2] * np.sin(2 * np.pi * X) * b
    T = 1.0 - Y + 0.05 * amp[3] * np.sin(3 * np.pi * Y) * np.cos(np.pi * Z)
    return u, v, w, T


@pytest.mark.parametrize("dim,n,Ra", [(2, 33, 1e4), (2, 65, 1e5), (2, 24, 1e6),
                                      (3, 17, 1e5), (3, 12, 1e6)])
def test_fp64_rest_start_matches_oracle(oracle_mod, hip, dim, n, Ra):
    print(f"\n[fp64 rest] dim={dim} n={n} Ra={Ra}", flush=True)
    o, s = _pair(oracle_mod, hip, dim, n, Ra, hip.PREC_FP64)
    o.init_rest()
    s.init_rest()
    done = 0
    for target in (1, 2, 10, 100):
        o.step(target - done)
        s.step(target - done)
        done = target
        errs = field_errors(s.fields(), o.fields())
        exact = all(np.array_equal(a, b) for a, b in zip(s.fields(), o.fields()))
        print(f"  itc={target} errs={errs} bit-exact={exact}", flush=True)
        for k, e in errs.items():
            assert e <= FP64_TOL, (target, k, e)
    assert s.itc == o.itc == 100


@pytest.mark.parametrize("dim,n,Ra,precision", [(2, 33, 1e5, 1), (3, 13, 1e5, 1),
                                                (2, 40, 1e5, 0), (3, 16, 1e5, 0)])
def test_tanh_grid_matches_oracle(oracle_mod, hip, dim, n, Ra, precision):
    """The tanh grid map (LBM_GRID_TANH, an extension): the grid, weights
    and trajectory against the oracle's tanh grid, FP64 to 1e-13 and FP32
    (the 3D case on the z-march) to 1e-5, with the Nusselt numbers."""
    print(f"\n[tanh grid] dim={dim} n={n} precision={precision}", flush=True)
    o = oracle_mod.Oracle(dim, n, Ra, A=2.0, grid_kind=1)
    s = hip.Solver(dim, n, Ra, stretch=2.0, grid_kind=hip.GRID_TANH, precision=precision)
    assert np.array_equal(s.grid(), o.coords())
    o.init_rest()
    s.init_rest()
    o.step(300)
    s.step(300)
    errs = field_errors(s.fields(), o.fields())
    print(f"  errs={errs}", flush=True)
    tol = FP64_TOL if precision == hip.PREC_FP64 else FP32_TOL
    for k, e in errs.items():
        assert e <= tol, (k, e)
    for a, b in zip(s.nusselt(), o.nusselt()):
        assert abs(a - b) <= (1e-12 if precision == hip.PREC_FP64 else NU_TOL) * abs(b)
    s.close()
    o.close()


@pytest.mark.parametrize("dim,n,Ra,seed", [(2, 29, 1e5, 1), (3, 14, 1e5, 2)])
def test_fp64_warm_start_matches_oracle(oracle_mod, hip, dim, n, Ra, seed):
    print(f"\n[fp64 warm] dim={dim} n={n} Ra={Ra}", flush=True)
    o, s = _pair(oracle_mod, hip, dim, n, Ra, hip.PREC_FP64)
    u, v, w, T = _smooth_fields(dim, n, seed)
    o.init_fields(u, v, w, T)
    s.init_fields(u, v, w, T)
    # the initial fields are returned unchanged before the first step
    g0 = s.fields()
    assert np.array_equal(g0[1], u) and np.array_equal(g0[4], T)
    for nsteps in (1, 19):
        o.step(nsteps)
        s.step(nsteps)
        errs = field_errors(s.fields(), o.fields())
        print(f"  itc={o.itc} errs={errs}", flush=True)
        for k, e in errs.items():
            assert e <= FP64_TOL, (k, e)


@pytest.mark.parametrize("dim,n,Ra,A", [(2, 33, 1e4, 3.2), (2, 41, 1e5, 1.5),
                                        (3, 17, 1e5, 3.2)])
def test_fp64_check_and_nusselt_match_oracle(oracle_mod, hip, dim, n, Ra, A):
    print(f"\n[fp64 check/Nu] dim={dim} n={n} Ra={Ra} A={A}", flush=True)
    o, s = _pair(oracle_mod, hip, dim, n, Ra, hip.PREC_FP64, A=A)
    o.init_rest()
    s.init_rest()
    for k in range(2):
        o.step(500)
        s.step(500)
        eo = o.check()
        es = s.check()
        print(f"  check {k}: oracle={eo} gpu={es}", flush=True)
        if k == 0:
            assert es == (1.0, 1.0)  # the snapshot starts at zero
        assert abs(es[0] - eo[0]) <= 1e-12 * abs(eo[0])
        assert abs(es[1] - eo[1]) <= 1e-12 * abs(eo[1])
    nh, nc = s.nusselt()
    oh, oc = o.nusselt()
    print(f"  Nu gpu=({nh}, {nc}) oracle=({oh}, {oc})", flush=True)
    assert abs(nh - oh) <= 1e-12 * abs(oh)
    assert abs(nc - oc) <= 1e-12 * abs(oc)


@pytest.mark.parametrize("dim,n,Ra,steps", [(2, 65, 1e5, 2000), (3, 21, 1e5, 500)])
def test_fp32_shifted_matches_oracle(oracle_mod, hip, dim, n, Ra, steps):
    print(f"\n[fp32 vs oracle] dim={dim} n={n} Ra={Ra} steps={steps}", flush=True)
    o, s = _pair(oracle_mod, hip, dim, n, Ra, hip.PREC_FP32_SHIFTED)
    o.init_rest()
    s.init_rest()
    o.step(steps)
    s.step(steps)
    errs = field_errors(s.fields(), o.fields())
    print(f"  errs={errs}", flush=True)
    names = ("u", "v", "w", "T") if dim == 3 else ("u", "v", "T")
    for k in names:
        assert errs[k] <= FP32_TOL, (k, errs[k])
    nh, nc = s.nusselt()
    oh, oc = o.nusselt()
    print(f"  Nu gpu=({nh}, {nc}) oracle=({oh}, {oc})", flush=True)
    assert abs(nh - oh) <= NU_TOL * abs(oh)
    assert abs(nc - oc) <= NU_TOL * abs(oc)


def test_run_stops_at_itc_max_and_checks(hip):
    print("\n[run loop] 2D 33 fp32", flush=True)
    s = hip.Solver(2, 33, 1e4, itc_max=5000, check_every=2000, epsU=1e-12, epsT=1e-12)
    s.init_rest()
    code, itc, eu, et = s.run()
    assert code == hip.LBM_ERR_NOT_CONVERGED
    assert itc == 5000
    # the last check ran at itc = 4000; itc_max is not a check point
    assert 0 < eu < 1 and 0 < et < 1


def test_save_load_state_is_exact(hip):
    print("\n[checkpoint] 2D 40 fp32", flush=True)
    s = hip.Solver(2, 40, 1e5)
    s.init_rest()
    s.step(50)
    buf = s.save_state()
    s.step(50)
    a = s.fields()
    s.load_state(buf)
    assert s.itc == 50
    s.step(50)
    b = s.fields()
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_call_order_and_arguments(hip):
    print("\n[errors]", flush=True)
    s = hip.Solver(2, 17, 1e4)
    with pytest.raises(hip.LbmError) as e:
        s.step(1)
    assert e.value.code == hip.LBM_ERR_STATE
    with pytest.raises(hip.LbmError) as e:
        hip.Solver(2, 17, 1e4, Ma=1e-9)  # tau_f -> 1/2
    assert e.value.code == hip.LBM_ERR_ARG


def _rel_l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def test_large_2d_fp32_vs_fp64_gpu(hip):
    """1025^2 at Ra = 1e6 (configuration C2) over the reference's 20 000 steps
    (2dim_mono.f90:31): the FP32 path against the FP64 GPU path (itself
    checked against the oracle at small sizes above), at 2000 steps and at
    20 000.  Bars: relative L2 error <= 1e-5 for the velocity and for T, and
    both walls' Nu within 1e-4.  The max-norm error is bounded at 1e-5 up to
    2000 steps and at 2e-5 over the whole run: through this start-up
    transient the rounding of the FP32-stored populations alone accumulates
    to 1.4-2.0e-5 in the largest pointwise velocity error (a build that
    rounds the FP64 path's stored populations to the FP32 shifted format,
    with FP64 arithmetic otherwise, gives 2.0e-5), so a 1e-5 max-norm bar at
    20 000 steps needs FP64 storage (LBM_PREC_FP64)."""
    print("\n[2D 1025 fp32 vs fp64] 2000 and 20000 steps", flush=True)
    a = hip.Solver(2, 1025, 1e6, precision=hip.PREC_FP32_SHIFTED)
    b = hip.Solver(2, 1025, 1e6, precision=hip.PREC_FP64)
    a.init_rest()
    b.init_rest()
    for steps, max_bar in ((2000, FP32_TOL), (20000, 2 * FP32_TOL)):
        a.step(steps - a.itc)
        b.step(steps - b.itc)
        fa, fb = a.fields(), b.fields()
        errs = field_errors(fa, fb)
        l2u = float(np.sqrt((np.sum((fa[1] - fb[1]) ** 2) + np.sum((fa[2] - fb[2]) ** 2)) /
                            (np.sum(fb[1] ** 2) + np.sum(fb[2] ** 2))))
        l2t = _rel_l2(fa[4], fb[4])
        na, nb = a.nusselt(), b.nusselt()
        nu_err = max(abs(x - y) / abs(y) for x, y in zip(na, nb))
        print(f"  itc={steps}: max-norm {errs}; relative L2 velocity {l2u:.3g} T {l2t:.3g}; "
              f"Nu fp32={na} fp64={nb} (rel {nu_err:.3g})", flush=True)
        assert l2u <= FP32_TOL and l2t <= FP32_TOL, (steps, l2u, l2t)
        for k in ("u", "v", "T"):
            assert errs[k] <= max_bar, (steps, k, errs[k])
        assert nu_err <= NU_TOL, (steps, na, nb)


def test_large_3d_fp32_vs_fp64_gpu(hip):
    """256^3 at Ra = 1e5 after 2000 steps (the reference's 3D itc_max)."""
    print("\n[3D 256 fp32 vs fp64] 2000 steps", flush=True)
    a = hip.Solver(3, 256, 1e5, precision=hip.PREC_FP32_SHIFTED)
    a.init_rest()
    a.step(2000)
    fa = a.fields()
    na = a.nusselt()
    a.close()
    b = hip.Solver(3, 256, 1e5, precision=hip.PREC_FP64)
    b.init_rest()
    b.step(2000)
    fb = b.fields()
    nb = b.nusselt()
    b.close()
    errs = field_errors(fa, fb)
    print(f"  errs={errs} Nu fp32={na} fp64={nb}", flush=True)
    for k in ("u", "v", "w", "T"):
        assert errs[k] <= FP32_TOL, (k, errs[k])
    assert abs(na[0] - nb[0]) <= NU_TOL * abs(nb[0])


def test_autotune_restores_state(hip):
    print("\n[autotune] 3D 40", flush=True)
    s = hip.Solver(3, 40, 1e5)
    s.init_rest()
    s.step(30)
    before = s.fields()
    s.autotune(3)
    assert s.itc == 30
    after = s.fields()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    # stepping on from the restored state equals stepping without the tuner:
    # the same 30 steps with the default launch, then the tuned one
    ref = hip.Solver(3, 40, 1e5)
    ref.init_rest()
    ref.step(30)
    ref.set_launch(**s.get_launch())
    s.step(10)
    ref.step(10)
    for x, y in zip(s.fields(), ref.fields()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("dim,n,bx,by", [(3, 32, 64, 4), (3, 64, 32, 2), (2, 256, 64, 4)])
def test_tile_orders_give_identical_results(hip, dim, n, bx, by):
    """Tile order only changes which workgroup computes a node, never the
    arithmetic: every order must give bit-identical fields."""
    print(f"\n[tile order] dim={dim} n={n} block=({bx},{by})", flush=True)
    outs = []
    for remap in (0, 1, 2):
        s = hip.Solver(dim, n, 1e5)
        s.set_launch(bx, by, remap=remap)
        s.init_rest()
        s.step(25)
        outs.append(s.fields())
        s.close()
    for other in outs[1:]:
        for x, y in zip(outs[0], other):
            assert np.array_equal(x, y)



@pytest.mark.parametrize("n,by,remap", [
    (12, 2, 0),     # one plane chunk per 4 planes, walls in every tile
    (20, 4, 0),     # far wall at tile row 3
    (68, 4, 1),     # a partial second column of lanes, far wall at lane 3
    (72, 2, 2),
    (96, 4, 0),
    (132, 4, 2),    # y bands per XCD
    (24, 8, 0),
    (136, 8, 2),    # 17 tile rows: XCD bands of 2 and 3 rows
    (24, 6, 0),     # far wall at tile row 5
    (68, 6, 2),     # partial last tile row (2 of 6 rows), 12 tile rows in bands
    (132, 6, 2),
    (72, 4, 2),     # 18 tile rows in bands of 2 and 3
    # any n: rows not 16-byte aligned and the nodes the march cannot hold
    # (the far wall in a tile's first column or row, tiles of 2-3 columns)
    # updated by the rim launch
    (17, 4, 0),     # far wall in row 0 of the last tile row: rim row
    (13, 6, 0),     # the same with 6-row tiles
    (21, 4, 1),     # partial quad at the x edge, far wall at tile row 0
    (37, 2, 0),
    (33, 8, 0),
    (65, 4, 2),     # far wall in column 0 of a tile: rim column and row
    (66, 4, 0),     # a 2-column last tile: rim columns
    (67, 2, 2),     # a 3-column last tile, rim row
    (70, 6, 0),     # 6-column last tile with a partial quad
    (131, 8, 2),
    (257, 4, 2),    # the reference's default 3D grid (3dim_mono.f90:6)
])
def test_lds_zmarch_variant_gives_identical_results(hip, n, by, remap):
    """Variant 1 stages the populations in LDS and marches in z; it uses the
    same stencil values and weights as variant 0, but interpolates the c_z
    != 0 populations in-plane first (as each plane arrives) and then along
    z, and the compiler fuses a different handful of multiply-adds.  So the
    two agree to FP32 rounding, not to the bit: ~1e-7 of the velocity scale
    after 24 steps, while a wrong stencil point or plane shows as an error
    near 1e-2.  Bound: 1e-6 relative to each field's largest value, walls and
    partial tiles included; accuracy against the oracle is checked
    separately."""
    print(f"\n[z-march] n={n} by={by} remap={remap}", flush=True)
    outs = []
    for variant, bx in ((0, 64), (1, 64)):
        s = hip.Solver(3, n, 1e5)
        s.set_launch(bx, by, remap=remap, variant=variant)
        assert s.get_launch()["variant"] == variant
        s.init_rest()
        s.step(24)
        outs.append(s.fields())
        s.close()
    vscale = max(np.abs(outs[0][k]).max() for k in (1, 2, 3))
    for name, x, y, scale in zip(("rho", "u", "v", "w", "T"), outs[0], outs[1],
                                 (1.0, vscale, vscale, vscale, np.abs(outs[0][4]).max())):
        err = np.abs(x - y).max() / scale
        print(f"  {name}: {err:.3g}", flush=True)
        assert err <= 1e-6, (name, err)


@pytest.mark.parametrize("n,steps,by", [(16, 300, 2), (36, 200, 4), (40, 200, 6),
                                        (17, 300, 4), (21, 200, 2), (13, 300, 6),
                                        (33, 150, 8)])
def test_lds_zmarch_variant_matches_oracle(oracle_mod, hip, n, steps, by):
    """Variant 1 against the FP64 oracle, to the FP32 bar of the plain path."""
    print(f"\n[z-march vs oracle] n={n} steps={steps} by={by}", flush=True)
    s = hip.Solver(3, n, 1e5)
    s.set_launch(64, by, remap=0, variant=1)
    o = oracle_mod.Oracle(3, n, 1e5)
    s.init_rest()
    o.init_rest()
    s.step(steps)
    o.step(steps)
    errs = field_errors(s.fields(), o.fields())
    print(f"  {errs}", flush=True)
    for k in ("rho", "u", "v", "w", "T"):
        assert errs[k] <= 1e-5, (k, errs[k])
    nh, nc = s.nusselt()
    oh, oc = o.nusselt()
    assert abs(nh - oh) <= 1e-4 * abs(oh) and abs(nc - oc) <= 1e-4 * abs(oc)
    s.close()
    o.close()


@pytest.mark.parametrize("n,remap", [(17, 0), (64, 1), (65, 1), (66, 0), (129, 1), (130, 0),
                                     (257, 1)])
def test_ymarch_variant_gives_identical_results(hip, n, remap):
    """The 2D y-march (variant 1: one wave per 64 columns, marching in y)
    against variant 0 after 24 steps: the same stencil values and weights,
    with the x stage of the diagonal populations first, so they agree to
    FP32 rounding (bound 1e-6 of each field's largest value).  n = 65, 129,
    257 put the far x wall on lane 0 of the last strip; n = 66, 130 on lane
    1."""
    print(f"\n[y-march] n={n} remap={remap}", flush=True)
    outs = []
    for variant, bx, by in ((0, 64, 4), (1, 64, 1)):
        s = hip.Solver(2, n, 1e5)
        s.set_launch(bx, by, remap=remap, variant=variant)
        assert s.get_launch()["variant"] == variant
        s.init_rest()
        s.step(24)
        outs.append(s.fields())
        s.close()
    vscale = max(np.abs(outs[0][k]).max() for k in (1, 2))
    for name, x, y, scale in zip(("rho", "u", "v", "w", "T"), outs[0], outs[1],
                                 (1.0, vscale, vscale, 1.0, np.abs(outs[0][4]).max())):
        err = np.abs(x - y).max() / scale
        print(f"  {name}: {err:.3g}", flush=True)
        assert err <= 1e-6, (name, err)


@pytest.mark.parametrize("n,steps", [(33, 600), (65, 400)])
def test_ymarch_variant_matches_oracle(oracle_mod, hip, n, steps):
    """The 2D y-march against the FP64 oracle, to the FP32 bar of the plain path."""
    print(f"\n[y-march vs oracle] n={n} steps={steps}", flush=True)
    s = hip.Solver(2, n, 1e5)
    s.set_launch(64, 1, remap=1, variant=1)
    o = oracle_mod.Oracle(2, n, 1e5)
    s.init_rest()
    o.init_rest()
    s.step(steps)
    o.step(steps)
    errs = field_errors(s.fields(), o.fields())
    print(f"  {errs}", flush=True)
    for k in ("rho", "u", "v", "T"):
        assert errs[k] <= 1e-5, (k, errs[k])
    nh, nc = s.nusselt()
    oh, oc = o.nusselt()
    assert abs(nh - oh) <= 1e-4 * abs(oh) and abs(nc - oc) <= 1e-4 * abs(oc)
    s.close()
    o.close()


def test_lds_zmarch_variant_is_rejected_where_it_cannot_run(hip):
    cases = [(dict(dim=2, n=64), (64, 4)),            # 2D: by = 1
             (dict(dim=2, n=64, precision=hip.PREC_FP64), (64, 1)),
             (dict(dim=3, n=64, precision=hip.PREC_FP64), (64, 4)),
             (dict(dim=3, n=64), (128, 2)),          # bx must be 64
             (dict(dim=3, n=64), (64, 16)),          # by 2, 4, 6 or 8
             (dict(dim=3, n=64), (64, 3)),
             (dict(dim=3, n=5), (64, 2))]            # n >= 6
    for kw, (bx, by) in cases:
        s = hip.Solver(kw["dim"], kw["n"], 1e5, precision=kw.get("precision", hip.PREC_FP32_SHIFTED))
        with pytest.raises(hip.LbmError):
            s.set_launch(bx, by, variant=1)
        s.close()


@pytest.mark.parametrize("n", [5, 9, 33, 61, 62, 63, 121, 122, 129, 181, 257])
def test_two_step_variant_equals_the_plain_kernel(hip, n):
    """Variant 2 runs two steps per launch on 60 x TY tiles held in
    registers and LDS (an odd step count ends with one plain launch).  Each
    node update is k_step's arithmetic on the values k_step reads, so the
    whole state (populations, fields, itc) must equal variant 0's to the
    bit: every tile height, both tile orders, even and odd step counts.
    n = 61, 121, 181 leave a last tile of one column (clamped to start at
    n - 2, overlapping its neighbour); n = 62, 122 one of two columns; n = 5,
    9 a single tile smaller than the region."""
    print(f"\n[two-step] n={n}", flush=True)
    ref = {}
    for steps in (1, 2, 7, 20):
        s = hip.Solver(2, n, 1e5, stretch=2.0)
        s.set_launch(64, 4, remap=1, variant=0)
        s.init_rest()
        s.step(steps)
        ref[steps] = (s.save_state(), s.fields())
        s.close()
    for ty in (8, 12, 16):
        for remap in (0, 1):
            for steps in (1, 2, 7, 20):
                s = hip.Solver(2, n, 1e5, stretch=2.0)
                s.set_launch(64, ty, remap=remap, variant=2)
                assert s.get_launch()["variant"] == 2
                s.init_rest()
                s.step(steps)
                state, fields = s.save_state(), s.fields()
                s.close()
                assert np.array_equal(state, ref[steps][0]), (ty, remap, steps)
                for x, y in zip(fields, ref[steps][1]):
                    assert np.array_equal(x, y), (ty, remap, steps)


@pytest.mark.parametrize("n,steps,ty", [(65, 401, 8), (129, 400, 16)])
def test_two_step_variant_matches_oracle(oracle_mod, hip, n, steps, ty):
    """Variant 2 against the FP64 oracle directly, to the FP32 bar of the
    plain path (an odd count ends with a one-step pass)."""
    print(f"\n[two-step vs oracle] n={n} steps={steps} ty={ty}", flush=True)
    s = hip.Solver(2, n, 1e5)
    s.set_launch(64, ty, remap=1, variant=2)
    o = oracle_mod.Oracle(2, n, 1e5)
    s.init_rest()
    o.init_rest()
    s.step(steps)
    o.step(steps)
    errs = field_errors(s.fields(), o.fields())
    print(f"  {errs}", flush=True)
    for k in ("rho", "u", "v", "T"):
        assert errs[k] <= 1e-5, (k, errs[k])
    nh, nc = s.nusselt()
    oh, oc = o.nusselt()
    assert abs(nh - oh) <= 1e-4 * abs(oh) and abs(nc - oc) <= 1e-4 * abs(oc)
    s.close()
    o.close()


def test_two_step_variant_is_rejected_where_it_cannot_run(hip):
    cases = [(dict(dim=3, n=32), (64, 8)),
             (dict(dim=2, n=64, precision=hip.PREC_FP64), (64, 8)),
             (dict(dim=2, n=64), (64, 4)),           # by 8, 12 or 16
             (dict(dim=2, n=64), (128, 8))]          # bx = 64
    for kw, (bx, by) in cases:
        s = hip.Solver(kw["dim"], kw["n"], 1e5, precision=kw.get("precision", hip.PREC_FP32_SHIFTED))
        with pytest.raises(hip.LbmError):
            s.set_launch(bx, by, variant=2)
        s.close()


@pytest.mark.parametrize("dim,n,prec,launch", [
    (2, 33, "fp32", None), (2, 64, "fp64", None), (3, 16, "fp32", (64, 4, 0, 1)),
    (2, 65, "fp32", (64, 1, 1, 1)), (2, 129, "fp32", (64, 16, 1, 2))])
def test_graph_replays_match_direct_launches(hip, dim, n, prec, launch):
    """Grids up to 2^20 nodes replay a check interval (check_every = 2000
    steps) as a HIP graph of the same launches; any other step count is
    launched directly.  Three replays in a row, with direct launches, checks
    and field reads between them, must each match the direct run bit for bit."""
    print(f"\n[graph] dim={dim} n={n} {prec} launch={launch}", flush=True)
    p = hip.PREC_FP64 if prec == "fp64" else hip.PREC_FP32_SHIFTED
    a = hip.Solver(dim, n, 1e5, precision=p)
    b = hip.Solver(dim, n, 1e5, precision=p)
    for s in (a, b):
        if launch:
            s.set_launch(launch[0], launch[1], remap=launch[2], variant=launch[3])
        s.init_rest()
    for r in range(3):
        a.step(2000)                  # graph replay
        b.step(1200)                  # direct launches
        b.step(800)
        assert a.itc == b.itc == 4000 * r + 2000
        fa, fb = a.fields(), b.fields()
        for name, x, y in zip(("rho", "u", "v", "w", "T"), fa, fb):
            assert np.array_equal(x, y), (r, name)
        assert a.check() == b.check()
        a.step(1)                     # direct work between replays
        b.step(1)
        a.step(1999)
        b.step(1999)
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)
    a.close()
    b.close()


@pytest.mark.parametrize("dim,n,prec,launch,steps,tol", [
    (2, 65, "fp64", None, 400, 1e-11),
    (3, 16, "fp64", None, 150, 1e-11),
    (2, 65, "fp32", (64, 4, 1, 0), 400, FP32_TOL),
    (2, 66, "fp32", (64, 1, 1, 1), 400, FP32_TOL),
    (2, 125, "fp32", (64, 12, 0, 2), 400, FP32_TOL),
    (3, 32, "fp32", (64, 4, 2, 1), 300, FP32_TOL),
    (3, 32, "fp32", (64, 4, 1, 0), 300, FP32_TOL),
])
def test_trajectory_keeps_the_cavity_symmetries(hip, dim, n, prec, launch, steps, tol):
    """The cavity's mirror and half-turn symmetries (cavity_symmetry.py;
    walls at +-1/2 so that the half turns are exact) hold for the GPU
    trajectory to rounding: FP64 to 1e-11, the FP32 paths (the plain
    kernel, the 2D y-march and the 3D z-march) to the FP32 bar.  A lane,
    tile or halo index that is wrong on one side of a tile or of the grid
    breaks them."""
    from cavity_symmetry import cavity_asymmetry
    p = hip.PREC_FP64 if prec == "fp64" else hip.PREC_FP32_SHIFTED
    s = hip.Solver(dim, n, 1e5, Thot=0.5, Tcold=-0.5, stretch=2.0, precision=p)
    if launch:
        s.set_launch(launch[0], launch[1], remap=launch[2], variant=launch[3])
    print(f"\n[symmetry] {dim}D n={n} {prec} launch={launch}", flush=True)
    s.init_rest()
    s.step(steps)
    f = s.fields()
    s.close()
    assert np.abs(f[1] if dim == 2 else f[3]).max() > 1e-4
    a = cavity_asymmetry(f, dim, t_sum=0.0)
    print(f"  {a}", flush=True)
    for name, val in a.items():
        assert val < tol, (name, val)


@pytest.mark.parametrize("dim,n,prec,want", [
    (3, 32, "fp32", dict(bx=64, by=4, remap=2, variant=1)),
    (3, 30, "fp32", dict(bx=64, by=4, remap=2, variant=1)),    # any n
    (3, 257, "fp32", dict(bx=64, by=4, remap=2, variant=1)),
    (3, 32, "fp64", dict(bx=64, by=4, remap=1, variant=0)),
    (2, 129, "fp32", dict(bx=32, by=2, remap=1, variant=0)),
    (2, 129, "fp64", dict(bx=32, by=2, remap=1, variant=0)),
    (2, 513, "fp32", dict(bx=64, by=8, remap=1, variant=2)),
    (2, 513, "fp64", dict(bx=64, by=4, remap=1, variant=0)),
])
def test_default_launch_shapes(hip, dim, n, prec, want):
    """lbm_create's default launch (the z-march for 3D FP32 where it
    applies, narrow blocks for small 2D grids), as include/lbm_capi.h
    states it."""
    p = hip.PREC_FP64 if prec == "fp64" else hip.PREC_FP32_SHIFTED
    s = hip.Solver(dim, n, 1e5, precision=p)
    got = s.get_launch()
    s.close()
    assert {k: got[k] for k in want} == want


@pytest.mark.parametrize("dim,n,prec", [(2, 33, "fp64"), (2, 65, "fp32"), (3, 32, "fp32"),
                                        (3, 17, "fp64"), (3, 21, "fp32")])
def test_copy_state_same_precision_continues_exactly(hip, dim, n, prec):
    """lbm_copy_state between two handles of one precision (a plain device
    copy): the copy steps on bit for bit like the original."""
    p = hip.PREC_FP64 if prec == "fp64" else hip.PREC_FP32_SHIFTED
    a = hip.Solver(dim, n, 1e5, precision=p)
    b = hip.Solver(dim, n, 1e5, precision=p)
    a.init_rest()
    a.step(150)
    b.copy_state_from(a)
    assert b.itc == 150
    a.step(100)
    b.step(100)
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)
    assert a.check() == b.check()
    a.close()
    b.close()


@pytest.mark.parametrize("dim,n,steps", [(2, 65, 500), (3, 24, 200), (3, 21, 200)])
def test_copy_state_promotes_fp32_to_fp64(hip, dim, n, steps):
    """An FP32 run promoted to FP64 half way tracks a pure FP64 run from the
    same start to within the FP32 bar (3D: the FP32 population rows are
    padded to 32 elements, the FP64 ones are not)."""
    f32 = hip.Solver(dim, n, 1e5)
    f64 = hip.Solver(dim, n, 1e5, precision=hip.PREC_FP64)
    ref = hip.Solver(dim, n, 1e5, precision=hip.PREC_FP64)
    f32.init_rest()
    ref.init_rest()
    f32.step(steps // 2)
    f64.copy_state_from(f32)
    for x, y in zip(f32.fields(), f64.fields()):
        assert np.array_equal(x, y)   # the FP64 fields are copied as they are
    f64.step(steps - steps // 2)
    ref.step(steps)
    errs = field_errors(f64.fields(), ref.fields())
    print(f"\n[promote] {dim}D n={n}: {errs}", flush=True)
    assert max(errs.values()) <= FP32_TOL
    for s in (f32, f64, ref):
        s.close()


def test_copy_state_rejects_mismatched_handles(hip):
    a = hip.Solver(2, 33, 1e5)
    b = hip.Solver(2, 35, 1e5, precision=hip.PREC_FP64)
    c = hip.Solver(2, 33, 1e4, precision=hip.PREC_FP64)
    with pytest.raises(hip.LbmError) as e:
        b.copy_state_from(a)     # a not initialised
    assert e.value.code == hip.LBM_ERR_STATE
    a.init_rest()
    for dst in (b, c, a):
        with pytest.raises(hip.LbmError) as e:
            dst.copy_state_from(a)
        assert e.value.code == hip.LBM_ERR_ARG
    for s in (a, b, c):
        s.close()


def test_fp32_then_fp64_reaches_the_fp64_steady_state(hip):
    """The 129^2, Ra = 1e4 cavity: FP32 to a residual of 2e-6 (above its
    floor near 1.2e-6), then FP64 from the promoted state to the reference's
    1e-8.  It ends on the pure FP64 steady state (the oracle's, bit for bit,
    in test_gpu_steady.py)."""
    print("\n[fp32 -> fp64] 129^2 Ra 1e4 to 1e-8", flush=True)
    ref = hip.Solver(2, 129, 1e4, precision=hip.PREC_FP64, epsU=1e-8, epsT=1e-8,
                     itc_max=2000000)
    ref.init_rest()
    st, itc_ref, _, _ = ref.run()
    assert st == hip.LBM_OK and itc_ref == 536000
    a = hip.Solver(2, 129, 1e4, epsU=2e-6, epsT=2e-6, itc_max=2000000)
    a.init_rest()
    st, itc32, _, _ = a.run()
    assert st == hip.LBM_OK
    b = hip.Solver(2, 129, 1e4, precision=hip.PREC_FP64, epsU=1e-8, epsT=1e-8,
                   itc_max=2000000)
    b.copy_state_from(a)
    st, itc, eu, et = b.run()
    print(f"  FP32 to 2e-6 at {itc32}, FP64 to 1e-8 at {itc} (pure FP64: {itc_ref})", flush=True)
    assert st == hip.LBM_OK and eu <= 1e-8 and et <= 1e-8
    errs = field_errors(b.fields(), ref.fields())
    nb, nr = b.nusselt(), ref.nusselt()
    print(f"  fields vs the FP64 steady state {errs}, Nu {abs(nb[0] / nr[0] - 1):.2e}",
          flush=True)
    assert max(errs.values()) <= FP32_TOL
    assert abs(nb[0] / nr[0] - 1) <= 1e-6
    for s in (ref, a, b):
        s.close()


@pytest.mark.parametrize("dim,n", [(2, 65), (3, 24)])
def test_copy_state_demotes_fp64_to_fp32(hip, dim, n):
    """An FP64 state demoted to FP32 storage (each population rounded once)
    steps on within the FP32 bar of the FP64 run it came from."""
    f64 = hip.Solver(dim, n, 1e5, precision=hip.PREC_FP64)
    f32 = hip.Solver(dim, n, 1e5)
    f64.init_rest()
    f64.step(200)
    f32.copy_state_from(f64)
    assert f32.itc == 200
    f64.step(100)
    f32.step(100)
    errs = field_errors(f32.fields(), f64.fields())
    print(f"\n[demote] {dim}D n={n}: {errs}", flush=True)
    assert max(errs.values()) <= FP32_TOL
    nu32, nu64 = f32.nusselt(), f64.nusselt()
    assert abs(nu32[0] / nu64[0] - 1) <= NU_TOL
    f64.close()
    f32.close()
