"""Host logic of the learned warm start (surrogate.py, lbmnn.py) -- CPU only:
seeded cavity sampling, the network's input encoding, the Nusselt-band
rule, the Nusselt stopping rule and the aggregation of warm-versus-cold
results."""
import numpy as np

import lbmhip
import lbmnn
import surrogate as S


def test_sample_cavities_is_seeded_and_in_range():
    a = S.sample_cavities(500, 11)
    assert a == S.sample_cavities(500, 11)
    assert a != S.sample_cavities(500, 12)
    lr = np.log10([c["Ra"] for c in a])
    pr = np.array([c["Pr"] for c in a])
    st = np.array([c["stretch"] for c in a])
    assert all(c["n"] == S.GRID for c in a)
    assert S.LOG_RA[0] <= lr.min() and lr.max() <= S.LOG_RA[1]
    assert S.PR[0] <= pr.min() and pr.max() <= S.PR[1]
    assert S.STRETCH[0] <= st.min() and st.max() <= S.STRETCH[1]
    # uniform draws cover the ranges
    assert lr.max() - lr.min() > 2.8 and pr.max() - pr.min() > 6.0


def test_inputs_map_the_trained_ranges_onto_unit_interval():
    x = np.linspace(0.0, 2.0, 8)
    lo = lbmnn.inputs(10 ** S.LOG_RA[0], S.PR[0], x, 2.0)
    hi = lbmnn.inputs(10 ** S.LOG_RA[1], S.PR[1], x, 2.0)
    assert lo.shape == (4, 8, 8) and lo.dtype == np.float32
    assert np.allclose(lo[0], -1) and np.allclose(hi[0], 1)
    assert np.allclose(lo[1], -1) and np.allclose(hi[1], 1)
    # channel 2 varies with i (fastest index), channel 3 with j
    assert np.allclose(lo[2][0], np.linspace(-1, 1, 8)) and np.allclose(lo[2][:, 3], lo[2][0, 3])
    assert np.allclose(lo[3][:, 0], np.linspace(-1, 1, 8)) and np.allclose(lo[3][3], lo[3][3, 0])


def test_targets_scale_velocity_by_a_quarter_of_free_fall():
    u, v, T = np.full((4, 4), 0.02), np.full((4, 4), -0.01), np.full((4, 4), 0.5)
    t = lbmnn.targets(u, v, T, 0.04)
    assert np.allclose(t[0], 0.5) and np.allclose(t[1], -0.25) and np.allclose(t[2], 0.5)
    info = {"gbeta": 1e-5, "H": 160.0}
    assert np.isclose(lbmnn.free_fall_velocity(info), 0.04)
    assert np.isclose(lbmnn.velocity_scale(info), 0.01)
    back = lbmnn.fields(lbmnn.targets(u, v, T, 0.01), 0.01)
    for a, b in zip(back, (u, v, T)):
        assert np.allclose(a, b, rtol=1e-6)


def test_erf_nodes_match_the_oracle_grid(oracle_mod):
    for n, A in ((64, 1.0), (64, 2.5), (129, 3.2), (257, 1.7), (128, 3.2)):
        o = oracle_mod.Oracle(2, n, 1e4, A=A)
        x, H = lbmnn.erf_nodes(n, A)
        assert np.allclose(x, o.coords(), rtol=1e-13, atol=0)
        assert np.isclose(H, o.params()["H"], rtol=1e-13)
        o.close()


def test_resample_is_exact_for_fields_linear_in_each_axis():
    """Piecewise-linear interpolation reproduces a linear function of the
    coordinate exactly, over the nodes and the walls; so T = 1 - x/H (which
    meets the hot, cold and adiabatic wall values resample() adds) comes
    through unchanged on any target grid."""
    G = 16
    xs, Hs = lbmnn.erf_nodes(G, 2.0)
    xt, Ht = lbmnn.erf_nodes(41, 1.3)
    ss, st = xs / Hs, xt / Ht
    M = lbmnn._lerp_matrix(np.concatenate(([0.0], ss, [1.0])), st)
    assert np.allclose(M.sum(axis=1), 1.0) and (M >= 0).all()
    assert (np.count_nonzero(M, axis=1) <= 2).all()
    f = lambda x: 0.3 - 1.7 * x   # noqa: E731
    src = np.concatenate(([0.0], ss, [1.0]))
    assert np.allclose(M @ f(src), f(st), rtol=0, atol=1e-14)
    # T linear in x from 1 (hot wall) to 0 (cold wall), constant in y: exact
    y = np.zeros((3, G, G))
    y[2] = 1.0 - ss[None, :]
    r = lbmnn.resample(y, ss, st)
    assert r.shape == (3, 41, 41)
    assert np.allclose(r[2], 1.0 - st[None, :], atol=1e-14)
    assert np.all(r[:2] == 0.0)


def test_resample_keeps_the_source_nodes_and_the_wall_values():
    G = 16
    xs, Hs = lbmnn.erf_nodes(G, 2.0)
    ss = xs / Hs
    rng = np.random.default_rng(3)
    y = rng.standard_normal((3, G, G))
    assert np.allclose(lbmnn.resample(y, ss, ss), y, rtol=0, atol=1e-15)
    # a target node half-way between the wall and the first source node gets
    # the mean of the wall value and that node's value
    st = np.array([0.5 * ss[0], ss[3]])
    r = lbmnn.resample(y, ss, st)
    assert np.isclose(r[0, 1, 0], 0.5 * y[0, 3, 0])          # u: 0 on the wall x = 0
    assert np.isclose(r[2, 1, 0], 0.5 * (1.0 + y[2, 3, 0]))  # T: 1 on the hot wall
    assert np.isclose(r[2, 0, 1], y[2, 0, 3])                # T: adiabatic wall y = 0
    assert np.isclose(r[1, 0, 1], 0.5 * y[1, 0, 3])          # v: 0 on the wall y = 0


def test_resample3_keeps_nodes_linear_fields_and_wall_values():
    """The cube's resampling (the coarse-cube warm start): identity on the
    source nodes, exact for T linear in y (the heated axis, the middle
    index of [k][j][i]), and the wall values of the cube's boundary
    conditions half a step outside the first node: no slip everywhere, the
    hot wall at y = 0, zero normal gradient of T on the x and z walls."""
    n = 9
    xs, Hs = lbmnn.erf_nodes(n, 2.0)
    ss = xs / Hs
    rng = np.random.default_rng(4)
    f = [rng.standard_normal((n, n, n)) for _ in range(4)]
    for a, b in zip(lbmnn.resample3(*f, ss, ss), f):
        assert np.allclose(a, b, rtol=0, atol=1e-14)
    xt, Ht = lbmnn.erf_nodes(17, 1.4)
    st = xt / Ht
    K, J, I = np.meshgrid(ss, ss, ss, indexing="ij")
    zero = np.zeros((n, n, n))
    r = lbmnn.resample3(zero, zero, zero, 1.0 - J, ss, st)
    Kt, Jt, It = np.meshgrid(st, st, st, indexing="ij")
    assert np.allclose(r[3], 1.0 - Jt, rtol=0, atol=1e-14)
    assert all(np.all(c == 0.0) for c in r[:3])
    p = np.array([0.5 * ss[0], ss[4]])   # half-way to the wall, and a source node
    u, v, w, T = lbmnn.resample3(*f, ss, p)
    assert np.isclose(u[1, 1, 0], 0.5 * f[0][4, 4, 0])            # u = 0 on x = 0
    assert np.isclose(w[0, 1, 1], 0.5 * f[2][0, 4, 4])            # w = 0 on z = 0
    assert np.isclose(T[1, 0, 1], 0.5 * (1.0 + f[3][4, 0, 4]))    # T = 1 on y = 0
    assert np.isclose(T[1, 1, 0], f[3][4, 4, 0])                  # dT/dx = 0 on x = 0
    assert np.isclose(T[0, 1, 1], f[3][0, 4, 4])                  # dT/dz = 0 on z = 0


def test_extrude_puts_the_plane_on_the_cubes_y_z_axes():
    n = 5
    u = np.arange(25.0).reshape(5, 5)
    v, T = -u, u / 25.0
    u3, v3, w3, T3 = lbmnn.extrude(u, v, T, n)
    assert u3.shape == v3.shape == w3.shape == T3.shape == (n, n, n)
    assert np.all(u3 == 0)
    k, j, i = 3, 1, 4   # [k][j][i], i fastest: z index k, heated-axis index j
    assert v3[k, j, i] == u[k, j] and w3[k, j, i] == v[k, j] and T3[k, j, i] == T[k, j]
    assert np.all(T3 == T3[:, :, :1])


def test_band_entry_is_the_last_entry_into_the_band():
    nu = np.array([1.0, 1.5, 2.0, 2.0025, 1.99, 2.001, 2.0004, 2.0])
    trace = np.stack([np.arange(1, 9) * 2000, np.arange(1, 9) * 0.5, nu], axis=1)
    # leaves the band at 1.99 (0.5 %), re-enters at 2.001 and stays
    assert S.band_entry(trace, 2.0, 1e-3) == (12000, 3.0)
    assert S.band_entry(trace, 2.0, 2e-2) == (6000, 1.5)
    assert S.band_entry(trace[:5], 2.0, 1e-3) is None          # ends outside
    nan = trace.copy()
    nan[-1, 2] = np.nan
    assert S.band_entry(nan, 2.0, 1e-3) is None


def test_settle_estimate_is_the_geometric_tail():
    """For Nu_k = Nu_inf + A r^k with 0 < r < 1 the bound is the exact
    remaining drift: the samples SETTLE_SPACING checks apart shrink by
    rho = r^SETTLE_SPACING, and |d| rho / (1 - rho) = A r^k."""
    span = (S.SETTLE_CHECKS - 1) * S.SETTLE_SPACING + 1
    for r, A in ((0.95, 0.1), (0.8, -0.3), (0.99, 0.02)):
        k = np.arange(span + 10)
        nu = 5.0 + A * r ** k
        for j in range(span, span + 11):
            est = S.settle_estimate(nu[:j])
            exact = abs(nu[j - 1] - 5.0) / nu[j - 1]
            assert np.isclose(est, exact, rtol=1e-8), (r, j, est, exact)
        # one check short of a full window
        assert S.settle_estimate(nu[:span - 1]) == np.inf
    assert S.settle_estimate([2.0] * span) == 0.0
    # an alternating approach (r < 0), growing differences (a transient),
    # ratios that disagree, non-finite values: no bound
    alt = np.repeat(5.0 + 0.1 * (-0.6) ** np.arange(S.SETTLE_CHECKS), S.SETTLE_SPACING)
    assert S.settle_estimate(alt) == np.inf
    assert S.settle_estimate(1.0 + 0.001 * 1.1 ** np.arange(span)) == np.inf
    w = 5.0 + 0.1 * 0.9 ** np.arange(span)
    w[-1] = w[-1 - S.SETTLE_SPACING] - 0.1 * (w[-1 - 2 * S.SETTLE_SPACING] - w[-1 - S.SETTLE_SPACING])
    assert S.settle_estimate(w) == np.inf
    nan = 5.0 + 0.1 * 0.9 ** np.arange(span)
    nan[-1 - S.SETTLE_SPACING] = np.nan
    assert S.settle_estimate(nan) == np.inf


def test_settle_entry_is_the_first_check_where_the_rule_holds_long_enough():
    r, A = 0.93, 0.2
    k = np.arange(1, 241)
    nu = 4.0 - A * r ** k
    trace = np.stack([k * 2000, k * 0.25, nu], axis=1)
    est = np.array([S.settle_estimate(nu[:j]) for j in range(1, 241)])
    ok = est <= S.SETTLE_TOL
    H = S.SETTLE_HOLD
    first = next(j for j in range(H - 1, 240) if ok[j - H + 1:j + 1].all())
    itc, wall, got = S.settle_entry(trace)
    assert (itc, wall, got) == (trace[first, 0], trace[first, 1], trace[first, 2])
    # the drift left at the stop is within the rule's bound
    assert abs(got / 4.0 - 1.0) <= S.SETTLE_TOL * 1.0001
    # one check before, it had not yet held SETTLE_HOLD checks in a row
    assert S.settle_entry(trace[:first]) is None
    # the span traced_run keeps is enough to decide at the latest check
    assert S.settle_entry(trace[first + 1 - S.SETTLE_SPAN:first + 1]) is not None
    assert S.settle_entry(trace[first - S.SETTLE_SPAN:first]) is None
    # a run that never settles
    assert S.settle_entry(np.stack([k, k, 1.0 + 0.01 * k], axis=1)) is None


def test_settle_entry_does_not_stop_on_a_flat_stretch_between_two_modes(monkeypatch):
    """A recorded cold run of a 257^2 cavity (tests/golden/nu_trace_257_flat.npz,
    written by make_nu_trace_257_flat.py) whose Nu flattens for about 20
    checks between two modes: the rule's setting does not stop it early
    (it runs to the residual rule), while six values four checks apart with
    a bound of 1e-4 held for two checks stop it on the flat stretch, 5.9e-3
    away."""
    import json
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "nu_trace_257_flat.npz"))
    trace, nu = d["trace"], float(d["nu_converged"])
    assert json.loads(str(d["cavity"]))["n"] == 257
    # the rule's setting: no stop before the residual rule ends the run, or
    # one within 0.1% of the converged Nu
    got = S.settle_entry(trace)
    assert got is None or abs(got[2] / nu - 1.0) <= 1e-3
    # six values four checks apart, bound 1e-4, held two checks
    monkeypatch.setattr(S, "SETTLE_CHECKS", 6)
    monkeypatch.setattr(S, "SETTLE_SPACING", 4)
    monkeypatch.setattr(S, "SETTLE_HOLD", 2)
    early = S.settle_entry(trace, tol=1e-4)
    assert early is not None and abs(early[2] / nu - 1.0) > 5e-3
    assert got is None or early[0] < got[0]


def _row(cs, ws, ci, wi, cb, wb, status=lbmhip.LBM_OK, fallback=False, dev=1e-5):
    return dict(cold_status=status, warm_status=lbmhip.LBM_OK, cold_s=cs, warm_s=ws,
                cold_itc=ci, warm_itc=wi, cold_band_s=cb[1], cold_band_itc=cb[0],
                warm_band_s=wb[1], warm_band_itc=wb[0], fallback=fallback, nu_rel_dev=dev)


def test_summary_counts_only_cavities_the_cold_run_judged():
    rows = [_row(4.0, 2.0, 40000, 20000, (30000, 3.0), (4000, 0.5)),
            _row(2.0, 2.0, 20000, 20000, (10000, 1.0), (4000, 0.5), dev=3e-4),
            dict(cold_status=lbmhip.LBM_ERR_NOT_CONVERGED)]
    s = S.summarise(rows)
    assert s["cavities"] == 3 and s["judged"] == 2 and s["fallbacks"] == 0
    assert np.isclose(s["stop_wall_ratio"], 6.0 / 4.0)
    assert np.isclose(s["stop_itc_ratio"], 60000 / 40000)
    assert np.isclose(s["band_wall_ratio"], 4.0 / 1.0)
    assert np.isclose(s["band_itc_ratio"], 40000 / 8000)
    assert np.isclose(s["nu_rel_dev_max"], 3e-4)
    assert s["settle_cavities"] == 0 and s["settle_wall_ratio"] is None
    assert S.summarise([dict(cold_status=lbmhip.LBM_ERR_DIVERGED)])["judged"] == 0
    rows[0].update(cold_settle_s=2.0, warm_settle_s=0.4, cold_settle_itc=20000,
                   warm_settle_itc=4000, cold_settle_nu_dev=2e-4, warm_settle_nu_dev=1.2e-3)
    rows[1].update(cold_settle_s=None, warm_settle_s=0.5)
    s = S.summarise(rows)
    assert s["settle_cavities"] == 1
    assert np.isclose(s["settle_wall_ratio"], 5.0) and np.isclose(s["settle_itc_ratio"], 5.0)
    assert np.isclose(s["settle_nu_dev_max"], 1.2e-3) and s["settle_nu_dev_over_1e3"] == 1


def test_data_files_round_trip(tmp_path):
    """save_data keeps targets to half an int16 step (2^-14) and the kept
    cavities' parameters; load_data rebuilds the inputs net_inputs() gives
    a cavity on a grid other than the network's."""
    rng = np.random.default_rng(4)
    G = 16
    cav = S.sample_cavities(3, 11, n=257)
    y = rng.uniform(-3.9, 3.9, (3, 3, G, G)).astype(np.float32)
    rec = dict(count=3, seed=11, cavities=[[c["Ra"], c["Pr"], c["stretch"]] for c in cav])
    S.save_data(str(tmp_path / "d.npz"), y, rec)
    x2, y2, recs = S.load_data([str(tmp_path / "d.npz")])
    assert np.abs(y2 - y).max() <= 2.0 ** -14 * (1 + 1e-6)
    assert recs == [dict(count=3, seed=11)]
    for i, c in enumerate(cav):
        xc, Hc = lbmnn.erf_nodes(G, c["stretch"])
        np.testing.assert_array_equal(x2[i], lbmnn.inputs(c["Ra"], c["Pr"], xc, Hc))
    import pytest
    with pytest.raises(ValueError):
        S.save_data(str(tmp_path / "e.npz"), np.full((1, 3, G, G), 4.0), dict(cavities=[[1e4, 1, 2]]))
