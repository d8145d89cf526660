"""The HIP warm-start network (include/lbm_nn.h) against its numpy oracle.

The HIP network multiplies BF16 operands on the matrix cores with FP32
accumulation; the oracle applies the same BF16 roundings with FP64
accumulation.  The two then differ only where an FP32-vs-FP64 difference
tips a later BF16 rounding, a few BF16 units in isolated elements: the
bounds are 2^-7 of the output scale for the largest forward difference and
1e-2 for the relative L2 distance of the gradients (a wrong index or a
missing term is far larger).
"""
import numpy as np
import pytest

import nn_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nnmod():
    import lbmnn
    return lbmnn


def _data(G, B, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros((B, 4, G, G), dtype=np.float32)
    xi = np.linspace(0.02, 0.98, G)
    X, Y = np.meshgrid(xi, xi)
    for b in range(B):
        x[b] = [np.full((G, G), rng.uniform(3, 6)), np.full((G, G), rng.uniform(0.7, 7)), X, Y]
    t = rng.standard_normal((B, 3, G, G)).astype(np.float32) * 0.3
    return x, t


@pytest.mark.parametrize("G,B,W", [(16, 3, 16), (32, 2, 16), (16, 2, 8), (32, 2, 28)])
def test_forward_matches_oracle(nnmod, G, B, W):
    print(f"\n[nn forward] G={G} B={B} width={W}", flush=True)
    net = nnmod.Net(grid=G, batch_max=4, seed=7, width=W)
    p = net.get_params()
    assert p.size == O.param_count(W) == 657 * W * W + 57 * W + 3
    x, _ = _data(G, B, 1)
    y = net.forward(x)
    ref = O.forward(p, x, quant=True)
    scale = np.abs(ref).max()
    err = np.abs(y - ref)
    print(f"  max {err.max() / scale:.3g} rms {np.sqrt(np.mean(err ** 2)) / scale:.3g}", flush=True)
    assert err.max() <= 2 ** -7 * scale
    assert np.array_equal(y, net.forward(x))      # deterministic
    net.close()


@pytest.mark.parametrize("W", [16, 8, 28])
def test_gradient_matches_oracle(nnmod, W):
    print(f"\n[nn gradient] G=16 B=2 width={W}", flush=True)
    G, B = 16, 2
    net = nnmod.Net(grid=G, batch_max=2, seed=3, width=W)
    p = net.get_params()
    x, t = _data(G, B, 2)
    loss, g = net.loss_grad(x, t)
    rloss, rg = O.loss_grad(p, x, t, quant=True)
    rel = np.linalg.norm(g - rg) / np.linalg.norm(rg)
    print(f"  loss {loss:.6g} vs {rloss:.6g}; gradient relative L2 {rel:.3g}", flush=True)
    assert abs(loss - rloss) <= 1e-3 * rloss
    assert rel <= 1e-2
    # per layer, so that no layer hides behind a large one
    o = 0
    for ci, co, k, _, _ in O.layers(W):
        n = co * ci * k * k + co
        lr = np.linalg.norm(g[o:o + n] - rg[o:o + n]) / np.linalg.norm(rg[o:o + n])
        assert lr <= 3e-2, (ci, co, lr)
        o += n
    assert np.array_equal(net.get_params(), p)    # loss_grad does not step
    net.close()


def test_training_reduces_the_loss(nnmod):
    """Adam on a smooth target that depends on the inputs: the loss falls by
    more than 10x in 300 steps, and set_params restores a starting point."""
    print("\n[nn train] G=16", flush=True)
    G, B = 16, 8
    net = nnmod.Net(grid=G, batch_max=B, seed=9)
    p0 = net.get_params()
    x, _ = _data(G, B, 4)
    X, Y = x[:, 2], x[:, 3]
    ra = (x[:, 0] - 4.5) / 1.5
    t = np.stack([np.sin(np.pi * X) * np.sin(2 * np.pi * Y) * (1 + 0.3 * ra),
                  -np.sin(2 * np.pi * X) * np.sin(np.pi * Y) * (1 + 0.3 * ra),
                  1 - X + 0.1 * ra * np.sin(np.pi * X)], axis=1).astype(np.float32)
    first = net.train_step(x, t, 2e-3)
    for _ in range(299):
        last = net.train_step(x, t, 2e-3)
    print(f"  loss {first:.4g} -> {last:.4g}", flush=True)
    assert last < first / 10
    net.set_params(p0)
    assert abs(net.loss_grad(x, t)[0] - first) <= 1e-6 * first
    net.close()


def test_adam_step_matches_oracle(nnmod):
    """lbm_nn_train_step's Adam update (k_adam) against the oracle's textbook
    step, fed the HIP network's own gradient (lbm_nn_loss_grad at the same
    parameters: the forward and backward passes are deterministic), over
    three steps with the oracle carrying m, v and t.  Then set_params clears
    the state: the next step from the starting point repeats the first one
    to the bit.  FP32 update arithmetic: 1e-6 of lr per parameter (a wrong
    bias correction or eps placement is off by at least 10% of lr), plus
    one FP32 unit of the stored parameter."""
    print("\n[nn adam] G=16 B=2, 3 steps", flush=True)
    G, B, lr = 16, 2, 1e-3
    net = nnmod.Net(grid=G, batch_max=B, seed=5)
    x, t = _data(G, B, 6)
    p0 = net.get_params()
    p = p0.astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    firsts = None
    for step in range(1, 4):
        pin = net.get_params()
        _, g = net.loss_grad(x, t)
        assert np.array_equal(net.get_params(), pin)
        net.train_step(x, t, lr)
        got = net.get_params()
        want, m, v = O.adam_step(pin, g, m, v, step, lr)
        # one FP32 unit of the stored parameter plus the update's rounding
        tol = 2.0 ** -23 * np.abs(want) + 1e-6 * lr
        err = np.abs(got - want)
        moved = np.abs(got - pin).max()
        print(f"  step {step}: max |p - oracle| / tol = {(err / tol).max():.3g} "
              f"(largest move {moved:.3g})", flush=True)
        assert np.all(err <= tol)
        assert moved > 0.5 * lr
        if step == 1:
            firsts = got
    net.set_params(p0)
    net.train_step(x, t, lr)
    assert np.array_equal(net.get_params(), firsts)
    net.close()
