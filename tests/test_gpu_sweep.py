"""Configuration C4 on its own workload: the batched sweep of 64 seeded 2D
cavities (edge from {65, 129, 257, 513}, A ~ U[1, 2.5], log10 Ra ~ U[3, 6],
Pr ~ U[0.7, 7]; lbmhip.sweep_cavities, seed 20261203) in one batch with the
autotuned block shape.  Each cavity follows the single-cavity semantics of
natural-convection_2dim_mono.f90, so every cavity of the FP32 batch is
compared with an FP64 single-cavity run of the same parameters (fields to
1e-5 relative, both walls' Nu to 1e-4), and the batch's per-cavity stopping
rule with single-cavity lbm_run calls."""
import numpy as np
import pytest

from helpers import field_errors

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5
NU_TOL = 1e-4


def _solver(hip, c, **kw):
    c = dict(c)
    return hip.Solver(2, c.pop("n"), c.pop("Ra"), **c, **kw)


def test_sweep_batch_matches_fp64_single_cavities(hip):
    cav = hip.sweep_cavities()
    assert len(cav) == 64
    assert sorted({c["n"] for c in cav}) == [65, 129, 257, 513]
    steps = 2000
    print(f"\n[C4 sweep] 64 cavities, {sum(c['n'] ** 2 for c in cav) / 1e6:.2f} M nodes, "
          f"{steps} steps", flush=True)
    b = hip.Batch(cav)
    b.init_rest()
    b.autotune(3)
    print(f"  launch {b.get_launch()}", flush=True)
    b.step(steps)
    worst = {"u": 0.0, "v": 0.0, "T": 0.0, "nu": 0.0}
    for i, c in enumerate(cav):
        assert b[i].itc == steps
        s = _solver(hip, c, precision=hip.PREC_FP64)
        s.init_rest()
        s.step(steps)
        errs = field_errors(b[i].fields(), s.fields())
        nb, ns = b[i].nusselt(), s.nusselt()
        nu = max(abs(x - y) / abs(y) for x, y in zip(nb, ns))
        s.close()
        for k in ("u", "v", "T"):
            worst[k] = max(worst[k], errs[k])
            assert errs[k] <= FP32_TOL, (i, c, k, errs[k])
        worst["nu"] = max(worst["nu"], nu)
        assert nu <= NU_TOL, (i, c, nb, ns)
    print(f"  worst over the 64 cavities: {worst}", flush=True)
    b.close()


def test_sweep_batch_run_stops_like_single_cavities(hip):
    """lbm_batch_run over the 64 cavities with eps = 0.15 and itc_max = 8000
    (checks every 2000 steps; the residuals of these cavities at 8000 steps
    range over 0.005-0.45, so some stop by the rule and others at itc_max):
    a sample of two cavities per edge length gets the same status and
    iteration count as its own lbm_run, and the same residuals to FP32
    rounding."""
    cav = [dict(c, epsU=0.15, epsT=0.15, itc_max=8000) for c in hip.sweep_cavities()]
    print("\n[C4 sweep] lbm_batch_run, eps 0.15, itc_max 8000", flush=True)
    b = hip.Batch(cav)
    b.init_rest()
    itc, eu, et, st = b.run()
    print(f"  statuses {np.bincount(st, minlength=5).tolist()} itc {np.bincount(itc // 2000).tolist()}",
          flush=True)
    assert {hip.LBM_OK, hip.LBM_ERR_NOT_CONVERGED} <= set(st.tolist())
    launch = b.get_launch()
    sample = []
    for n in (65, 129, 257, 513):
        sample += [i for i, c in enumerate(cav) if c["n"] == n][:2]
    for i in sample:
        s = _solver(hip, cav[i])
        s.set_launch(launch["bx"], launch["by"], remap=0)
        s.init_rest()
        code, sitc, seu, set_ = s.run()
        s.close()
        print(f"  cavity {i} n={cav[i]['n']}: batch ({st[i]}, {itc[i]}, {eu[i]:.4g}, {et[i]:.4g})"
              f" single ({code}, {sitc}, {seu:.4g}, {set_:.4g})", flush=True)
        assert code == st[i] and sitc == itc[i], i
        assert abs(seu - eu[i]) <= 1e-4 * seu and abs(set_ - et[i]) <= 1e-4 * set_, i
    b.close()
