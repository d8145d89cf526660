"""The learned warm start on the GPU (surrogate.py): the pipeline from the
batched solver's steady fields to a trained network, and what the trained
network (data/warm64.npz, trained on 4096 seeded cavities, seed 20261203)
gains over the cold start on held-out cavities (seed 7).

The cold start is the judge: the warm start runs the same solver to the
same stopping rule (eps 1e-6 on snapshots 2000 steps apart), and must end
on the same hot-wall Nusselt number within 0.1 %.  The gain is measured on
the time to reach that Nusselt number for good (the first check from which
on it stays within 0.1 % of the cold run's converged value), and the time to
the stopping rule is reported beside it.
"""
import os

import numpy as np
import pytest

import lbmhip
import surrogate as S

pytestmark = pytest.mark.gpu


def test_pipeline_trains_on_solver_fields():
    """Steady fields from the batched solver -> network targets -> Adam."""
    print("\n[surrogate pipeline] 12 cavities, itc_max 20000", flush=True)
    x, y, rec = S.generate(12, 5, itc_max=20000, per_batch=12, keep_tol=10.0, log=lambda s: None)
    assert x.shape == (12, 4, S.GRID, S.GRID) and y.shape == (12, 3, S.GRID, S.GRID)
    assert np.isfinite(x).all() and np.isfinite(y).all()
    # T spans the wall values; the velocity channels are O(1) in units of V
    assert y[:, 2].min() >= -0.05 and y[:, 2].max() <= 1.05
    assert 0.01 < np.abs(y[:, :2]).max() < 4.0
    import lbmnn
    net = lbmnn.Net(grid=S.GRID, batch_max=4, seed=3)
    hist = S.train(net, x, y, epochs=20, batch=4, lr=2e-3, log=lambda s: None)
    print(f"  loss {hist[0]['loss']:.4g} -> {hist[-1]['loss']:.4g}", flush=True)
    assert hist[-1]["loss"] < hist[0]["loss"] / 5
    net.close()


@pytest.fixture(scope="module")
def trained():
    assert os.path.exists(S.WEIGHTS), S.WEIGHTS
    net = S.load_net(S.WEIGHTS)
    yield net
    net.close()


def test_warm_start_reaches_the_same_nusselt_faster(trained):
    cavs = S.sample_cavities(8, 7)
    print("\n[surrogate warm vs cold] 8 held-out cavities, seed 7", flush=True)
    rows, s = S.compare(trained, cavs)
    for r in rows:
        print(f"  Ra {r['Ra']:.3g} Pr {r['Pr']:.2f} A {r['stretch']:.2f}: itc {r['cold_itc']} -> "
              f"{r['warm_itc']}, Nu band at {r['cold_band_itc']} -> {r['warm_band_itc']}, "
              f"dNu {r['nu_rel_dev']:.1e}", flush=True)
    print(f"  wall to the Nu band {s['band_wall_ratio']:.2f}x, to the stopping rule "
          f"{s['stop_wall_ratio']:.2f}x", flush=True)
    assert s["judged"] == len(cavs) and s["fallbacks"] == 0
    assert s["nu_rel_dev_max"] <= 1e-3
    assert s["band_cavities"] == len(cavs)
    assert s["band_wall_ratio"] >= 3.0 and s["band_itc_ratio"] >= 3.0
    # the stopping rule is reached no later than from rest
    assert s["stop_itc_ratio"] >= 1.0


def test_untrained_network_gains_less(trained):
    """The same warm start with the network's initial weights: the gain above
    comes from training, not from the restart itself."""
    import lbmnn
    cavs = S.sample_cavities(4, 7)
    print("\n[surrogate untrained] 4 held-out cavities", flush=True)
    raw = lbmnn.Net(grid=S.GRID, batch_max=4, seed=20261203)
    _, s0 = S.compare(raw, cavs)
    _, s1 = S.compare(trained, cavs)
    raw.close()
    print(f"  band wall ratio untrained {s0['band_wall_ratio']}, trained {s1['band_wall_ratio']:.2f}",
          flush=True)
    assert s1["fallbacks"] == 0
    assert s0["band_wall_ratio"] is None or s1["band_wall_ratio"] > 2 * s0["band_wall_ratio"]


@pytest.fixture(scope="module")
def trained257():
    path = os.path.join(os.path.dirname(S.WEIGHTS), "warm257.npz")
    assert os.path.exists(path), path
    net = S.load_net(path)
    yield net
    net.close()


def test_warm_start_resampled_onto_257_grids(trained257):
    """The network runs on its 64 x 64 erf grid and its output is resampled
    onto the 257 x 257 nodes (weights fine-tuned on 257 x 257 steady fields
    sampled at the network's nodes)."""
    cavs = S.sample_cavities(3, 7, n=257)
    print("\n[surrogate 257] 3 held-out cavities, seed 7", flush=True)
    rows, s = S.compare(trained257, cavs, itc_max=2000000)
    for r in rows:
        print(f"  Ra {r['Ra']:.3g} Pr {r['Pr']:.2f}: Nu band at {r['cold_band_itc']} -> "
              f"{r['warm_band_itc']}, dNu {r['nu_rel_dev']:.1e}", flush=True)
    print(f"  steps to the Nu band {s['band_itc_ratio']:.2f}x, wall {s['band_wall_ratio']:.2f}x",
          flush=True)
    assert s["judged"] == len(cavs) and s["fallbacks"] == 0
    assert s["nu_rel_dev_max"] <= 1e-3
    assert s["band_cavities"] == len(cavs)
    assert s["band_itc_ratio"] >= 1.5


def test_warm_start_extruded_onto_a_cube(trained257):
    """The 3D cube from the 2D prediction of its (y, z) cross-section: the
    solver reaches the cold start's steady Nusselt number from it."""
    cav = dict(dim=3, n=32, Ra=1e5, Pr=0.71, stretch=3.2)
    print("\n[surrogate cube] 32^3, Ra 1e5", flush=True)
    rows, s = S.compare(trained257, [cav], itc_max=1000000)
    r = rows[0]
    print(f"  itc {r['cold_itc']} -> {r['warm_itc']}, Nu band at {r['cold_band_itc']} -> "
          f"{r['warm_band_itc']}, dNu {r['nu_rel_dev']:.1e}", flush=True)
    assert s["judged"] == 1 and s["fallbacks"] == 0
    assert s["nu_rel_dev_max"] <= 1e-3
    assert r["warm_band_itc"] <= r["cold_band_itc"]
