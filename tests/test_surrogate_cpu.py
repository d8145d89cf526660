"""Host logic of the learned warm start (surrogate.py, lbmnn.py) -- CPU only:
seeded cavity sampling, the network's input encoding, the Nusselt-band
rule and the aggregation of warm-versus-cold results."""
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
    assert S.summarise([dict(cold_status=lbmhip.LBM_ERR_DIVERGED)])["judged"] == 0
