"""The batch's variant 5: one cooperative launch runs every step of an
lbm_batch_step call over the tiles of every cavity, with a grid-wide barrier
between steps.  Each cavity keeps the single-cavity loop semantics
(natural-convection_2dim_mono.f90:307-330 per cavity): the batch must equal
variant 0 (one launch per step, the same node code over the same tile
table) to the bit, across MALL groups, the per-launch step cap and
lbm_batch_run's per-cavity stopping, and single-cavity runs (the same
arithmetic compiled into another kernel) to FP32 rounding, within COOP_TOL.
"""
import numpy as np
import pytest

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="variant 5 (the cooperative launch) is switched off: after three "
                            "passing GPU runs, one run stopped making progress inside "
                            "test_coop_batch_with_mall_groups and the cause is not established"),
]

# relative max-norm difference of the fields against single-cavity runs
# (the FP32 field bar against the FP64 oracle is 1e-5)
COOP_TOL = 1e-5

CAVITIES = [dict(n=17, Ra=1e3, Pr=0.71, stretch=1.0),
            dict(n=33, Ra=1e5, Pr=7.0, stretch=2.5),
            dict(n=24, Ra=1e6, Pr=0.7, stretch=3.2),
            dict(n=65, Ra=3e4, Pr=2.0, stretch=1.7),
            dict(n=40, Ra=1e4, Pr=0.71, stretch=3.2),
            dict(n=129, Ra=1e6, Pr=0.71, stretch=2.0),
            dict(n=121, Ra=1e5)]


def _single(hip, c, **kw):
    c = dict(c)
    n = c.pop("n")
    s = hip.Solver(2, n, c.pop("Ra"), **c, **kw)
    s.set_launch(64, 4, remap=0)
    return s


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _batches(hip, cav, bx, by, calls):
    """The same cavities through variant 0 (one launch per step, the same
    node code over the same tile table) and variant 5."""
    out = []
    for variant in (0, 5):
        b = hip.Batch(cav)
        b.set_launch(bx, by, variant=variant)
        assert b.get_launch()["variant"] == variant
        b.init_rest()
        for k in calls:
            b.step(k)
        out.append(b)
    return out


def _assert_coop_equals_per_step(hip, cav, b0, b5, idx, steps, single=()):
    diffs = {}
    for i in idx:
        f0, f5 = b0[i].fields(), b5[i].fields()
        diffs[i] = max(_rel(x, y) for x, y in zip(f5, f0) if np.abs(y).max() > 0)
        print(f"  cavity {i} n={cav[i]['n']}: max rel diff vs variant 0 {diffs[i]:.3g}",
              flush=True)
    for i in idx:
        assert b0[i].itc == b5[i].itc == steps
        for x, y in zip(b5[i].fields(), b0[i].fields()):
            assert np.array_equal(x, y), (i, cav[i], diffs[i])
        assert np.array_equal(b5[i].save_state(), b0[i].save_state()), (i, cav[i])
    for i in single:
        s = _single(hip, cav[i])
        s.init_rest()
        s.step(steps)
        errs = [_rel(x, y) for x, y in zip(b5[i].fields(), s.fields()) if np.abs(y).max() > 0]
        print(f"  cavity {i}: max rel diff vs single cavity {max(errs):.3g}", flush=True)
        assert max(errs) <= COOP_TOL, (i, errs)
        s.close()


@pytest.mark.parametrize("bx,by", [(64, 8), (128, 4)])
def test_coop_batch_matches_per_step_launches_to_the_bit(hip, bx, by):
    print(f"\n[batch coop] {len(CAVITIES)} cavities, block {bx} x {by}, 120 + 37 steps",
          flush=True)
    b0, b5 = _batches(hip, CAVITIES, bx, by, (120, 37))
    _assert_coop_equals_per_step(hip, CAVITIES, b0, b5, range(len(CAVITIES)), 157,
                                 single=(0, 5))
    b0.close()
    b5.close()


def test_coop_batch_past_the_per_launch_cap(hip):
    """4100 steps take two cooperative launches (4096 + 4): the buffer
    parity and the last step's macroscopic fields carry over."""
    cav = CAVITIES[:3]
    print("\n[batch coop] 3 cavities, 4100 steps in one call", flush=True)
    b0, b5 = _batches(hip, cav, 64, 8, (4100,))
    _assert_coop_equals_per_step(hip, cav, b0, b5, range(len(cav)), 4100)
    b0.close()
    b5.close()


def test_coop_batch_with_mall_groups(hip):
    """A batch above half the MALL runs its groups in 64-step chunks inside
    the cooperative launch: every cavity equals variant 0 (which runs the
    same chunks, one launch per step) to the bit over a full and a partial
    chunk, and over a one-step call."""
    cav = [dict(n=513 if i % 3 else 257, Ra=10 ** (3 + 3 * i / 13), Pr=0.7 + 0.4 * i,
                stretch=1.0 + 0.1 * i) for i in range(14)]
    print(f"\n[batch coop, MALL groups] {len(cav)} cavities, 70 + 1 + 6 steps", flush=True)
    b0, b5 = _batches(hip, cav, 64, 8, (70, 1, 6))
    _assert_coop_equals_per_step(hip, cav, b0, b5, range(len(cav)), 77, single=(1,))
    b0.close()
    b5.close()


def test_coop_batch_run_matches_per_step_launches(hip):
    """lbm_batch_run on variant 5 (check windows of 1000 steps, an itc_max
    that is not a multiple of the window): the same statuses, iteration
    counts, residuals and fields as variant 0, to the bit."""
    print("\n[batch coop] run with per-cavity eps and itc_max", flush=True)
    cav = [dict(n=17, Ra=1e3, epsU=1e-2, epsT=1e-2, itc_max=40000),
           dict(n=24, Ra=1e4, epsU=1e-9, epsT=1e-9, itc_max=6001),
           dict(n=129, Ra=1e4, stretch=1.5, epsU=1e-3, epsT=1e-3, itc_max=40000)]
    res = []
    for bx, by, variant in ((64, 8, 0), (64, 8, 5)):
        b = hip.Batch(cav, check_every=1000)
        b.set_launch(bx, by, variant=variant)
        b.init_rest()
        out = b.run()
        res.append(out + ([b[i].fields() for i in range(len(cav))],))
        b.close()
    for x, y in zip(res[0][:4], res[1][:4]):   # itc, errU, errT, status
        assert np.array_equal(x, y)
    for fa, fb in zip(res[0][4], res[1][4]):
        for x, y in zip(fb, fa):
            assert np.array_equal(x, y)
    assert res[1][0].tolist()[1] == 6001
    assert res[1][3].tolist() == [hip.LBM_OK, hip.LBM_ERR_NOT_CONVERGED, hip.LBM_OK]
