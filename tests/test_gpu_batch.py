"""Batches of independent 2D cavities (lbm_batch_*): one launch per time step
advances every active cavity.  Each cavity must follow the single-cavity
semantics (natural-convection_2dim_mono.f90 per cavity): the batch is
compared with independent single-cavity solvers on the same parameters and
with the FP64 oracle.
"""
import numpy as np
import pytest

from helpers import field_errors

pytestmark = pytest.mark.gpu

# uneven sizes, stretch factors, Rayleigh and Prandtl numbers
CAVITIES = [dict(n=17, Ra=1e3, Pr=0.71, stretch=1.0),
            dict(n=33, Ra=1e5, Pr=7.0, stretch=2.5),
            dict(n=24, Ra=1e6, Pr=0.7, stretch=3.2),
            dict(n=65, Ra=3e4, Pr=2.0, stretch=1.7),
            dict(n=40, Ra=1e4, Pr=0.71, stretch=3.2)]


def _single(hip, c, **kw):
    c = dict(c)
    n = c.pop("n")
    return hip.Solver(2, n, c.pop("Ra"), **c, **kw)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_batch_matches_single_cavities(hip):
    print("\n[batch] 5 cavities vs single solvers, 300 steps", flush=True)
    b = hip.Batch(CAVITIES)
    b.set_launch(64, 4)
    b.init_rest()
    b.step(300)
    for i, c in enumerate(CAVITIES):
        s = _single(hip, c)
        s.set_launch(64, 4, remap=0)
        s.init_rest()
        s.step(300)
        got, ref = b[i].fields(), s.fields()
        assert b[i].itc == s.itc == 300
        errs = [_rel(x, y) for x, y in zip(got, ref) if np.abs(y).max() > 0]
        print(f"  cavity {i} n={c['n']}: max rel diff {max(errs):.3g}", flush=True)
        # the same node code in another kernel: FP32 rounding at most
        assert max(errs) <= 1e-6, (i, errs)
        nb, ns = b[i].nusselt(), s.nusselt()
        assert abs(nb[0] - ns[0]) <= 1e-6 * abs(ns[0])
        s.close()
    b.close()


def test_batch_ymarch_matches_single_cavities_to_the_bit(hip):
    """The batch's variant 1 runs the same y-march code as a single cavity's
    variant 1, over a table dealt to the XCDs: every cavity must match its
    single-cavity run bit for bit (the chunk split does not change any
    node's arithmetic)."""
    print("\n[batch y-march] 5 cavities vs single y-march solvers, 300 steps", flush=True)
    cav = CAVITIES + [dict(n=129, Ra=1e6, Pr=0.71, stretch=2.0)]
    b = hip.Batch(cav)
    b.set_launch(64, 1, variant=1)
    assert b.get_launch()["variant"] == 1
    b.init_rest()
    b.step(300)
    for i, c in enumerate(cav):
        s = _single(hip, c)
        s.set_launch(64, 1, remap=1, variant=1)
        s.init_rest()
        s.step(300)
        got, ref = b[i].fields(), s.fields()
        for x, y in zip(got, ref):
            assert np.array_equal(x, y), (i, c)
        s.close()
    b.close()


def test_batch_matches_oracle(hip, oracle_mod):
    print("\n[batch] vs oracle, 2 cavities, 1000 steps", flush=True)
    cav = [CAVITIES[0], CAVITIES[1]]
    b = hip.Batch(cav)
    b.init_rest()
    b.step(1000)
    for i, c in enumerate(cav):
        o = oracle_mod.Oracle(2, c["n"], c["Ra"], Pr=c["Pr"], A=c["stretch"])
        o.init_rest()
        o.step(1000)
        errs = field_errors(b[i].fields(), o.fields())
        print(f"  cavity {i}: {errs}", flush=True)
        for k in ("rho", "u", "v", "T"):
            assert errs[k] <= 1e-5, (i, k, errs[k])
        nh, nc = b[i].nusselt()
        oh, oc = o.nusselt()
        assert abs(nh - oh) <= 1e-4 * abs(oh) and abs(nc - oc) <= 1e-4 * abs(oc)
        o.close()
    b.close()


def test_batch_run_stops_each_cavity_on_its_own_rule(hip):
    print("\n[batch] run with per-cavity eps and itc_max", flush=True)
    cav = [dict(n=17, Ra=1e3, epsU=1e-2, epsT=1e-2, itc_max=40000),
           dict(n=24, Ra=1e4, epsU=1e-9, epsT=1e-9, itc_max=6000),     # stops at itc_max
           dict(n=20, Ra=1e4, stretch=1.5, epsU=1e-3, epsT=1e-3, itc_max=40000)]
    b = hip.Batch(cav, check_every=1000)
    b.init_rest()
    itc, eu, et, st = b.run()
    print(f"  itc={itc.tolist()} status={st.tolist()} errU={eu.tolist()}", flush=True)
    assert st.tolist() == [hip.LBM_OK, hip.LBM_ERR_NOT_CONVERGED, hip.LBM_OK]
    assert itc[1] == 6000
    assert b.active()[0] == 0
    for i, c in enumerate(cav):
        s = _single(hip, c, check_every=1000)
        s.set_launch(*[b.get_launch()[k] for k in ("bx", "by")], remap=0)
        s.init_rest()
        code, sitc, seu, set_ = s.run()
        assert code == st[i] and sitc == itc[i], (i, code, sitc)
        assert abs(seu - eu[i]) <= 1e-5 * seu and abs(set_ - et[i]) <= 1e-5 * set_
        s.close()
    b.close()


def test_batch_steps_cavities_initialised_one_by_one(hip):
    """Cavities set up through their own handles (lbm_init_fields on
    lbm_batch_cavity, no lbm_batch_init_rest) are active: a batch step
    advances every one of them like a single-cavity solver.  After
    lbm_batch_run every cavity is retired, lbm_batch_step refuses with
    LBM_ERR_STATE, and re-initialising one cavity makes it, and only it,
    active again."""
    print("\n[batch] init_fields on every cavity, then batch steps", flush=True)
    cav = CAVITIES[:4]
    b = hip.Batch(cav)
    b.set_launch(64, 4)
    with pytest.raises(hip.LbmError) as e:
        b.step(1)
    assert e.value.code == hip.LBM_ERR_STATE
    rng = np.random.default_rng(11)
    init = []
    for i, c in enumerate(cav):
        n = c["n"]
        u = 1e-3 * rng.standard_normal((n, n))
        v = 1e-3 * rng.standard_normal((n, n))
        T = 0.5 + 0.1 * rng.standard_normal((n, n))
        init.append((u, v, T))
        b[i].init_fields(u, v, None, T)
    assert b.active() == (len(cav), sum(c["n"] ** 2 for c in cav))
    b.step(40)
    for i, c in enumerate(cav):
        assert b[i].itc == 40
        s = _single(hip, c)
        s.set_launch(64, 4, remap=0)
        u, v, T = init[i]
        s.init_fields(u, v, None, T)
        s.step(40)
        errs = [_rel(x, y) for x, y in zip(b[i].fields(), s.fields()) if np.abs(y).max() > 0]
        assert max(errs) <= 1e-6, (i, errs)
        s.close()
    b.run()
    assert b.active()[0] == 0
    with pytest.raises(hip.LbmError) as e:
        b.step(1)
    assert e.value.code == hip.LBM_ERR_STATE
    b[2].init_rest()
    assert b.active() == (1, cav[2]["n"] ** 2)
    itc0 = b[1].itc
    b.step(10)
    assert b[2].itc == 10 and b[1].itc == itc0
    b.close()


def test_batch_warm_start_of_one_cavity(hip):
    print("\n[batch] init_fields on one cavity", flush=True)
    b = hip.Batch(CAVITIES[:3])
    b.init_rest()
    n = CAVITIES[1]["n"]
    rng = np.random.default_rng(5)
    u = 1e-3 * rng.standard_normal((n, n))
    T = 0.5 + 0.1 * rng.standard_normal((n, n))
    b[1].init_fields(u, None, None, T)
    b.step(50)
    s = _single(hip, CAVITIES[1])
    s.set_launch(64, 4, remap=0)
    s.init_fields(u, None, None, T)
    s.step(50)
    errs = [_rel(x, y) for x, y in zip(b[1].fields(), s.fields()) if np.abs(y).max() > 0]
    assert max(errs) <= 1e-6, errs
    assert b[0].itc == 50 and b[1].itc == 50
    s.close()
    b.close()


def test_batch_autotune_restores_state(hip):
    print("\n[batch] autotune", flush=True)
    b = hip.Batch(CAVITIES)
    b.init_rest()
    b.step(20)
    before = [b[i].fields() for i in range(len(CAVITIES))]
    ms = b.autotune(3)
    assert ms > 0
    for i in range(len(CAVITIES)):
        assert b[i].itc == 20
        for x, y in zip(before[i], b[i].fields()):
            assert np.array_equal(x, y)
    cav, nodes = b.active()
    assert cav == len(CAVITIES) and nodes == sum(c["n"] ** 2 for c in CAVITIES)
    b.close()


def test_batch_mall_groups_keep_every_cavity_exact(hip):
    """A batch whose step traffic exceeds half the 256 MB MALL runs in groups
    of whole cavities, each advancing up to 64 steps before the next group
    starts.  Every cavity's arithmetic is unchanged: the same cavities split
    into batches small enough to run ungrouped (the same kernel, launch
    shape and ordinary stores) must give the same fields to the bit, over
    full and partial group chunks and an odd count; and each cavity agrees
    with its single-cavity run to FP32 rounding."""
    cav = [dict(n=513 if i % 3 else 257, Ra=10 ** (3 + 3 * i / 13), Pr=0.7 + 0.4 * i,
                stretch=1.0 + 0.1 * i) for i in range(14)]
    print(f"\n[batch MALL groups] {len(cav)} cavities, "
          f"{sum(c['n'] ** 2 for c in cav) / 1e6:.2f} M nodes", flush=True)
    split = [cav[k:k + 3] for k in range(0, len(cav), 3)]
    # each part is below the 128 MB group limit of step traffic (112 B per
    # node), the whole batch well above it
    assert all(sum(c["n"] ** 2 for c in h) * 112 < 128 * 2 ** 20 for h in split)
    assert sum(c["n"] ** 2 for c in cav) * 112 > 256 * 2 ** 20
    a = hip.Batch(cav)
    parts = [hip.Batch(h) for h in split]
    for x in [a] + parts:
        x.set_launch(128, 4)
        x.init_rest()
        x.step(70)
        x.step(7)
    for i in range(len(cav)):
        part, j = parts[i // 3], i % 3
        assert a[i].itc == part[j].itc == 77
        for x, y in zip(a[i].fields(), part[j].fields()):
            assert np.array_equal(x, y), i
    for i in (0, 1, 13):
        s = _single(hip, cav[i])
        s.set_launch(128, 4, remap=0)
        s.init_rest()
        s.step(77)
        errs = [_rel(x, y) for x, y in zip(a[i].fields(), s.fields()) if np.abs(y).max() > 0]
        print(f"  cavity {i} n={cav[i]['n']}: max rel diff vs single {max(errs):.3g}", flush=True)
        assert max(errs) <= 1e-6, (i, errs)
        s.close()
    for x in [a] + parts:
        x.close()


@pytest.mark.skip(reason="variant 2 (the persistent launch) is switched off: its first two GPU runs, in lbm_batch_autotune, stopped making progress and the cause is not established")
@pytest.mark.parametrize("bx,by", [(64, 8), (64, 1), (32, 2)])
def test_batch_persistent_launch_matches_per_step_launches(hip, bx, by):
    """Variant 2 advances the whole batch through every lbm_batch_step call
    in one persistent launch (a work queue of (step, tile) items with
    per-tile step counters).  Each node's update is the same code as the
    per-step launches of variant 0 with the same block, so every cavity must
    agree with variant 0 to the bit, over an odd and an even step count and
    with by = 1 (tiles of one row, whose wall stencils reach two tile rows)."""
    print(f"\n[batch persistent] {bx}x{by} vs per-step launches", flush=True)
    cav = CAVITIES + [dict(n=129, Ra=1e6, Pr=0.71, stretch=2.0)]
    a, p = hip.Batch(cav), hip.Batch(cav)
    a.set_launch(bx, by, variant=0)
    p.set_launch(bx, by, variant=2)
    assert p.get_launch()["variant"] == 2
    for x in (a, p):
        x.init_rest()
        x.step(301)
        x.step(6)
    for i in range(len(cav)):
        assert a[i].itc == p[i].itc == 307
        for x, y in zip(a[i].fields(), p[i].fields()):
            assert np.array_equal(x, y), (i, cav[i])
    a.close()
    p.close()


@pytest.mark.skip(reason="variant 2 (the persistent launch) is switched off: its first two GPU runs, in lbm_batch_autotune, stopped making progress and the cause is not established")
def test_batch_persistent_launch_with_mall_groups(hip):
    """A batch above half the MALL runs its persistent launch group by group
    in chunks of 64 steps (ticket order: chunk, group, step, tile).  Every
    cavity must match the grouped per-step launches to the bit, over full
    and partial chunks."""
    cav = [dict(n=513 if i % 3 else 257, Ra=10 ** (3 + 3 * i / 13), Pr=0.7 + 0.4 * i,
                stretch=1.0 + 0.1 * i) for i in range(14)]
    print(f"\n[batch persistent, MALL groups] {len(cav)} cavities", flush=True)
    a, p = hip.Batch(cav), hip.Batch(cav)
    a.set_launch(128, 4, variant=0)
    p.set_launch(128, 4, variant=2)
    for x in (a, p):
        x.init_rest()
        x.step(70)
        x.step(7)
    for i in range(len(cav)):
        assert a[i].itc == p[i].itc == 77
        for x, y in zip(a[i].fields(), p[i].fields()):
            assert np.array_equal(x, y), i
    a.close()
    p.close()


@pytest.mark.skip(reason="variant 2 (the persistent launch) is switched off: its first two GPU runs, in lbm_batch_autotune, stopped making progress and the cause is not established")
def test_batch_persistent_run_stops_like_single_cavities(hip):
    """lbm_batch_run on variant 2: one persistent launch per check window,
    each cavity stopped by its own rule, the same iteration counts,
    statuses and residuals as variant 0."""
    print("\n[batch persistent] run with per-cavity eps and itc_max", flush=True)
    cav = [dict(n=17, Ra=1e3, epsU=1e-2, epsT=1e-2, itc_max=40000),
           dict(n=24, Ra=1e4, epsU=1e-9, epsT=1e-9, itc_max=6000),
           dict(n=20, Ra=1e4, stretch=1.5, epsU=1e-3, epsT=1e-3, itc_max=40000)]
    res = []
    for variant in (0, 2):
        b = hip.Batch(cav, check_every=1000)
        b.set_launch(64, 4, variant=variant)
        b.init_rest()
        res.append(b.run())
        b.close()
    for x, y in zip(res[0], res[1]):
        assert np.array_equal(x, y)
    assert res[1][3].tolist() == [hip.LBM_OK, hip.LBM_ERR_NOT_CONVERGED, hip.LBM_OK]


# cavities the resident variant takes (edges up to 70): odd and even edges,
# a last node group of every size, the far wall on any lane
RESIDENT = [dict(n=9, Ra=1e4, Pr=0.71, stretch=1.0),
            dict(n=17, Ra=1e3, Pr=0.71, stretch=1.0),
            dict(n=33, Ra=1e5, Pr=7.0, stretch=2.5),
            dict(n=24, Ra=1e6, Pr=0.7, stretch=3.2),
            dict(n=64, Ra=3e4, Pr=2.0, stretch=1.7),
            dict(n=65, Ra=1e6, Pr=0.71, stretch=2.2),
            dict(n=70, Ra=1e4, Pr=4.0, stretch=3.2)]


def _resident_vs_single(hip, cav, chunks, launch_single=(64, 4)):
    """Runs cav as a resident batch (variant 3) over the step counts in
    chunks, and each cavity alone (variant 0) over the same chunks; returns
    the largest relative difference of the fields per cavity and whether
    all of them are equal to the bit."""
    b = hip.Batch(cav)
    b.set_launch(512, 1, variant=3)
    assert b.get_launch()["variant"] == 3
    b.init_rest()
    for n in chunks:
        b.step(n)
    worst, exact = [], True
    for i, c in enumerate(cav):
        s = _single(hip, c)
        s.set_launch(*launch_single, remap=0)
        s.init_rest()
        for n in chunks:
            s.step(n)
        assert b[i].itc == s.itc == sum(chunks)
        got, ref = b[i].fields(), s.fields()
        exact = exact and all(np.array_equal(x, y) for x, y in zip(got, ref))
        worst.append(max(_rel(x, y) for x, y in zip(got, ref) if np.abs(y).max() > 0))
        s.close()
    b.close()
    return worst, exact


def test_resident_batch_matches_single_cavities(hip):
    """Variant 3 keeps each cavity in one workgroup for all the steps of an
    lbm_batch_step call.  Every step runs the single-cavity node code on the
    same values in the same order, so each cavity equals its own variant-0
    solver to the bit: checked over one step, an odd and an even chunk (the
    final buffer parity), and a chunk longer than one launch's 4096 steps."""
    for chunks in ((1,), (137, 164), (4101,)):
        print(f"\n[batch resident] {len(RESIDENT)} cavities, chunks {chunks}", flush=True)
        worst, exact = _resident_vs_single(hip, RESIDENT, chunks)
        print(f"  max rel diff per cavity {['%.2g' % w for w in worst]}, bit-exact {exact}",
              flush=True)
        assert exact, worst


def test_resident_batch_matches_oracle(hip, oracle_mod):
    print("\n[batch resident] vs oracle, 2 cavities, 1000 steps", flush=True)
    cav = [RESIDENT[2], RESIDENT[5]]
    b = hip.Batch(cav)
    b.set_launch(512, 1, variant=3)
    b.init_rest()
    b.step(1000)
    for i, c in enumerate(cav):
        o = oracle_mod.Oracle(2, c["n"], c["Ra"], Pr=c["Pr"], A=c["stretch"])
        o.init_rest()
        o.step(1000)
        errs = field_errors(b[i].fields(), o.fields())
        print(f"  cavity {i} n={c['n']}: {errs}", flush=True)
        for k in ("rho", "u", "v", "T"):
            assert errs[k] <= 1e-5, (i, k, errs[k])
        nh, nc = b[i].nusselt()
        oh, oc = o.nusselt()
        assert abs(nh - oh) <= 1e-4 * abs(oh) and abs(nc - oc) <= 1e-4 * abs(oc)
        o.close()
    b.close()


def test_resident_batch_run_stops_like_single_cavities(hip):
    """lbm_batch_run on variant 3: one resident launch per check window, each
    cavity stopped by its own rule, with the statuses and iteration counts of
    its own lbm_run and its residuals to FP32 rounding."""
    print("\n[batch resident] run with per-cavity eps and itc_max", flush=True)
    cav = [dict(n=17, Ra=1e3, epsU=1e-2, epsT=1e-2, itc_max=40000),
           dict(n=24, Ra=1e4, epsU=1e-9, epsT=1e-9, itc_max=6000),     # stops at itc_max
           dict(n=64, Ra=1e4, stretch=1.5, epsU=1e-3, epsT=1e-3, itc_max=40000)]
    b = hip.Batch(cav, check_every=1000)
    b.set_launch(512, 1, variant=3)
    b.init_rest()
    itc, eu, et, st = b.run()
    print(f"  itc={itc.tolist()} status={st.tolist()} errU={eu.tolist()}", flush=True)
    assert st.tolist() == [hip.LBM_OK, hip.LBM_ERR_NOT_CONVERGED, hip.LBM_OK]
    assert b.active()[0] == 0
    for i, c in enumerate(cav):
        s = _single(hip, c, check_every=1000)
        s.set_launch(64, 4, remap=0)
        s.init_rest()
        code, sitc, seu, set_ = s.run()
        assert code == st[i] and sitc == itc[i], (i, code, sitc)
        assert abs(seu - eu[i]) <= 1e-5 * seu and abs(set_ - et[i]) <= 1e-5 * set_
        s.close()
    b.close()


def test_resident_variant_needs_small_cavities(hip):
    """Variant 3 is refused when a cavity's edge is above 70, and the
    autotuner then leaves it out; with small cavities it is a candidate and
    the autotuner restores every cavity's state."""
    b = hip.Batch([dict(n=33, Ra=1e4), dict(n=71, Ra=1e4)])
    with pytest.raises(hip.LbmError) as e:
        b.set_launch(512, 1, variant=3)
    assert e.value.code == hip.LBM_ERR_ARG
    b.init_rest()
    b.autotune(3)
    assert b.get_launch()["variant"] != 3
    b.close()
    b = hip.Batch(RESIDENT)
    b.init_rest()
    b.step(21)
    before = [b[i].fields() for i in range(len(RESIDENT))]
    b.autotune(3)
    print(f"\n[batch resident] autotuned launch {b.get_launch()}", flush=True)
    for i in range(len(RESIDENT)):
        assert b[i].itc == 21
        for x, y in zip(before[i], b[i].fields()):
            assert np.array_equal(x, y)
    b.close()


@pytest.mark.parametrize("ty", [8, 16])
def test_two_step_batch_matches_single_cavities_to_the_bit(hip, ty):
    """Variant 4 runs two steps per launch on 60 x ty tiles of every cavity
    (the single-cavity variant 2 over the batch's table): each cavity equals
    its own variant-0 solver to the bit, over an even and an odd chunk."""
    cav = CAVITIES + [dict(n=129, Ra=1e6, Pr=0.71, stretch=2.0), dict(n=121, Ra=1e5)]
    print(f"\n[batch two-step] {len(cav)} cavities, tiles 60 x {ty}", flush=True)
    b = hip.Batch(cav)
    b.set_launch(64, ty, variant=4)
    assert b.get_launch()["variant"] == 4
    b.init_rest()
    b.step(120)
    b.step(37)
    for i, c in enumerate(cav):
        s = _single(hip, c)
        s.set_launch(64, 4, remap=0)
        s.init_rest()
        s.step(157)
        assert b[i].itc == s.itc == 157
        for x, y in zip(b[i].fields(), s.fields()):
            assert np.array_equal(x, y), (i, c)
        assert np.array_equal(b[i].save_state(), s.save_state()), (i, c)
        s.close()
    b.close()


def test_two_step_batch_with_mall_groups(hip):
    """A batch above half the MALL runs variant 4 group by group, 64 steps
    (32 launches) per group at a time: every cavity equals its variant-0
    single-cavity run to the bit over full and partial chunks."""
    cav = [dict(n=513 if i % 3 else 257, Ra=10 ** (3 + 3 * i / 13), Pr=0.7 + 0.4 * i,
                stretch=1.0 + 0.1 * i) for i in range(14)]
    print(f"\n[batch two-step, MALL groups] {len(cav)} cavities", flush=True)
    b = hip.Batch(cav)
    b.set_launch(64, 12, variant=4)
    b.init_rest()
    b.step(70)
    b.step(7)
    for i in (0, 1, 5, 13):
        s = _single(hip, cav[i])
        s.set_launch(64, 4, remap=0)
        s.init_rest()
        s.step(77)
        for x, y in zip(b[i].fields(), s.fields()):
            assert np.array_equal(x, y), i
        s.close()
    b.close()


def test_two_step_batch_run_matches_one_step_launches(hip):
    """lbm_batch_run on variant 4 (two steps per launch, check windows of
    1000 steps and an itc_max that is not a multiple of two launches'
    worth): the same statuses, iteration counts and residuals as variant
    0, to the bit."""
    print("\n[batch two-step] run with per-cavity eps and itc_max", flush=True)
    cav = [dict(n=17, Ra=1e3, epsU=1e-2, epsT=1e-2, itc_max=40000),
           dict(n=24, Ra=1e4, epsU=1e-9, epsT=1e-9, itc_max=6001),
           dict(n=129, Ra=1e4, stretch=1.5, epsU=1e-3, epsT=1e-3, itc_max=40000)]
    res = []
    for bx, by, variant in ((64, 4, 0), (64, 8, 4)):
        b = hip.Batch(cav, check_every=1000)
        b.set_launch(bx, by, variant=variant)
        b.init_rest()
        res.append(b.run())
        fields = [b[i].fields() for i in range(len(cav))]
        b.close()
        res[-1] = res[-1] + (fields,)
    for x, y in zip(res[0][:4], res[1][:4]):
        assert np.array_equal(x, y)
    for fa, fb in zip(res[0][4], res[1][4]):
        for x, y in zip(fa, fb):
            assert np.array_equal(x, y)
    assert res[1][0].tolist()[1] == 6001
    assert res[1][3].tolist() == [hip.LBM_OK, hip.LBM_ERR_NOT_CONVERGED, hip.LBM_OK]
