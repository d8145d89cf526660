"""Configuration C5 on its own workloads: the learned warm start (weights
data/warm257w28.npz: base width 28, 516 687 parameters, on a 128 x 128 erf
grid, trained on 2028 steady 257 x 257 fields of the batched solver) on
held-out seeded 257 x 257 cavities and on the 128^3
cube (Ra 1e5, Pr 0.71, A 3.2) from the extruded 2D prediction of its
cross-section, each run warm and cold to the same stopping rule
(natural-convection_2dim_mono.f90:307-330: both residuals <= eps on
snapshots 2000 steps apart).  The cold start is the judge: the warm run
must reach the cold run's converged Nusselt number within 0.1%, without
falling back to a cold start.  Both gains are printed: to the check from
which the hot-wall Nu stays within 0.1% of the converged value (the Nu
band), and to the stopping rule itself."""
import os

import pytest

import lbmhip
import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net257():
    path = os.path.join(os.path.dirname(S.WEIGHTS), "warm257w28.npz")
    net = S.load_net(path)
    yield net
    net.close()


def _report(label, s):
    print(f"  {label}: Nu band {s['band_itc_ratio']:.2f}x in steps, "
          f"{s['band_wall_ratio']:.2f}x in wall time; stopping rule "
          f"{s['stop_itc_ratio']:.2f}x in steps, {s['stop_wall_ratio']:.2f}x in wall time; "
          f"largest converged-Nu deviation {s['nu_rel_dev_max']:.2e}", flush=True)
    print(f"  {label}: Nusselt stopping rule {s['settle_itc_ratio']:.2f}x in steps, "
          f"{s['settle_wall_ratio']:.2f}x in wall time (both runs stopped by it in "
          f"{s['settle_by_nu']} of {s['settle_cavities']}); largest distance from the "
          f"converged Nu at its stops {s['settle_nu_dev_max']:.2e}", flush=True)


def test_c5_held_out_257_cavities(net257):
    """Eight held-out 257^2 cavities (sample_cavities(8, 7): log10 Ra ~
    U[3, 6], Pr ~ U[0.7, 7], A ~ U[1, 2.5]; the draws of each parameter come
    in one block, so these share Ra with the first eight of the bench's 100
    and have other Pr and A)."""
    cavs = S.sample_cavities(8, 7, n=257)
    print("\n[C5] 8 held-out 257^2 cavities, warm vs cold to eps 1e-6", flush=True)
    rows, s = S.compare(net257, cavs, itc_max=2000000)
    for r in rows:
        print(f"  Ra {r['Ra']:.3g} Pr {r['Pr']:.2f} A {r['stretch']:.2f}: stop "
              f"{r['cold_itc']} -> {r['warm_itc']}, band {r['cold_band_itc']} -> "
              f"{r['warm_band_itc']}, dNu {r['nu_rel_dev']:.1e}", flush=True)
    _report("257^2", s)
    assert s["judged"] == len(cavs) and s["fallbacks"] == 0
    assert s["nu_rel_dev_max"] <= 1e-3
    assert s["band_cavities"] == len(cavs)
    assert s["band_itc_ratio"] >= 3.0
    assert s["stop_itc_ratio"] >= 1.0
    # the Nusselt stopping rule stops every run within 0.1% of the
    # converged Nu
    assert s["settle_cavities"] == len(cavs)
    assert s["settle_nu_dev_max"] <= 1e-3


def test_c5_nusselt_rule_run_live(net257):
    """traced_run(stop="nu") on three held-out 257^2 cavities, cold and
    warm: the run ends at the Nusselt rule's entry in the full run's trace
    (the rule reads only earlier checks, and the trajectory is the same to
    the bit), no later than the residual rule, within 0.1% of the converged
    Nu."""
    print("\n[C5] Nusselt stopping rule, run live", flush=True)
    for cav in S.sample_cavities(3, 7, n=257):
        for net in (None, net257):
            full = S.traced_run(cav, net, itc_max=2000000)
            live = S.traced_run(cav, net, itc_max=2000000, stop="nu")
            ent = S.settle_entry(full["trace"])
            want = int(ent[0]) if ent is not None else int(full["itc"])
            dev = abs(live["nu"] / full["nu"] - 1.0)
            print(f"  Ra {cav['Ra']:.3g} {'warm' if net else 'cold'}: residual rule at "
                  f"{full['itc']}, Nusselt rule at {live['itc']}, Nu {dev:.1e} from the "
                  f"converged value", flush=True)
            assert full["status"] == lbmhip.LBM_OK and live["status"] == lbmhip.LBM_OK
            assert live["itc"] == want and live["itc"] <= full["itc"]
            assert dev <= 1e-3


def test_c5_cube_128(net257):
    """The 128^3 cube at its rule eps = 2e-6 (its FP32 velocity residual
    wanders between 1e-6 and 2e-6 for hundreds of thousands of steps),
    started from a 64^3 cube that the extruded prediction starts
    (surrogate.warm_fields); the coarse run's time counts as warm time."""
    cav = dict(dim=3, n=128, Ra=1e5, Pr=0.71, stretch=3.2)
    print("\n[C5] 128^3 cube, warm vs cold to eps 2e-6", flush=True)
    rows, s = S.compare(net257, [cav], eps=2e-6, itc_max=1200000,
                        log=lambda m: print("   ", m, flush=True))
    r = rows[0]
    print(f"  stop {r['cold_itc']} -> {r['warm_itc']} steps ({r['cold_s']:.1f} -> "
          f"{r['warm_s']:.1f} s), Nu {r['cold_nu']:.5f} / {r['warm_nu']:.5f}", flush=True)
    _report("128^3", s)
    assert s["judged"] == 1 and s["fallbacks"] == 0
    assert s["nu_rel_dev_max"] <= 1e-3
    assert r["warm_band_itc"] * 3 <= r["cold_band_itc"]
    assert s["band_wall_ratio"] >= 3.0
    assert r["warm_itc"] <= r["cold_itc"]
