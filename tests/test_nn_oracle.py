"""Pins of the network oracle (oracle/nn_oracle.py) -- CPU only.

The oracle is the checker of the HIP network, so it is pinned by what the
mathematics fixes: brute-force loops for the convolution (padding, stride,
kernel order), finite differences of the FP64 loss for every part of the
backward pass, the adjoint identity of the upsampling, known BF16 roundings,
and the C-ABI exports of include/lbm_nn.h.
"""
import os
import re

import numpy as np
import pytest

import nn_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bf16_rounding_known_values():
    cases = {1.0: 1.0, 1.0 + 2 ** -8: 1.0,              # tie -> even (1.0)
             1.0 + 3 * 2 ** -8: 1.0 + 2 ** -6,          # tie -> even (odd mantissa rounds up)
             1.0 + 2 ** -8 + 2 ** -20: 1.0 + 2 ** -7,   # above the tie
             -2.5: -2.5, 3.0e-3: None}
    for x, want in cases.items():
        got = float(O.bf16(np.array([x], dtype=np.float32))[0])
        if want is None:   # half a unit in the last of 8 significant bits
            assert abs(got - x) <= abs(x) * 2 ** -8
        else:
            assert got == want, (x, got, want)


@pytest.mark.parametrize("k,stride", [(3, 1), (3, 2), (1, 1)])
def test_conv_matches_brute_force_loops(k, stride):
    rng = np.random.default_rng(k * 10 + stride)
    B, C, H, Co = 2, 3, 6, 4
    X = rng.standard_normal((B, C, H, H))
    W = rng.standard_normal((Co, C, k, k))
    b = rng.standard_normal(Co)
    got = O.conv(X, W, b, stride, quant=False)
    pad = k // 2
    Ho = H // stride
    ref = np.zeros((B, Co, Ho, Ho))
    for bb in range(B):
        for o in range(Co):
            for yo in range(Ho):
                for xo in range(Ho):
                    s = b[o]
                    for c in range(C):
                        for ky in range(k):
                            for kx in range(k):
                                y, x = yo * stride + ky - pad, xo * stride + kx - pad
                                if 0 <= y < H and 0 <= x < H:
                                    s += W[o, c, ky, kx] * X[bb, c, y, x]
                    ref[bb, o, yo, xo] = s
    assert np.allclose(got, ref, rtol=0, atol=1e-12)


def test_upsample_adjoint():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 3, 4, 4))
    y = rng.standard_normal((2, 3, 8, 8))
    assert abs(np.sum(O.up2(x) * y) - np.sum(x * O.up2_bwd(y))) < 1e-12


@pytest.mark.parametrize("width", [16, 8])
def test_loss_gradient_matches_finite_differences(width):
    """Directional derivatives of the FP64 loss along random directions, and
    single-parameter derivatives of every layer's first weight and bias, at
    the default base width and at another one (the concatenations split at
    2W and 4W channels)."""
    rng = np.random.default_rng(11)
    G, B = 16, 2
    p = np.zeros(O.param_count(width))
    o = 0
    for ci, co, k, _, _ in O.layers(width):
        nw = co * ci * k * k
        p[o:o + nw] = rng.uniform(-1, 1, nw) * np.sqrt(6.0 / (ci * k * k))
        p[o + nw:o + nw + co] = 0.05 * rng.standard_normal(co)
        o += nw + co
    x = rng.standard_normal((B, 4, G, G))
    t = rng.standard_normal((B, 3, G, G))
    loss, g = O.loss_grad(p, x, t, quant=False)
    # A step along a random direction moves every pre-activation, and at
    # h = 1e-6 some ReLU input of the width-8 net changes sign inside the
    # difference interval; h = 1e-7 stays clear of every kink and its FP64
    # rounding (about 1e-16 |loss| / h) is far below the tolerance.
    h = 1e-7
    for _ in range(3):
        d = rng.standard_normal(p.size)
        num = (O.loss_grad(p + h * d, x, t, False)[0] - O.loss_grad(p - h * d, x, t, False)[0]) / (2 * h)
        assert abs(num - g @ d) <= 1e-6 * max(1.0, abs(num)), (num, g @ d)
    h = 1e-6
    o = 0
    for ci, co, k, _, _ in O.layers(width):
        for idx in (o, o + co * ci * k * k):   # first weight, first bias
            e = np.zeros(p.size)
            e[idx] = 1.0
            num = (O.loss_grad(p + h * e, x, t, False)[0] - O.loss_grad(p - h * e, x, t, False)[0]) / (2 * h)
            assert abs(num - g[idx]) <= 1e-6 * max(1e-3, abs(num)) + 1e-9, (idx, num, g[idx])
        o += co * ci * k * k + co


def test_quantised_forward_is_a_bf16_perturbation():
    rng = np.random.default_rng(5)
    p = 0.2 * rng.standard_normal(O.param_count())
    x = rng.standard_normal((1, 4, 16, 16))
    y0 = O.forward(p, x, quant=False)
    y1 = O.forward(p, x, quant=True)
    rel = np.abs(y1 - y0).max() / np.abs(y0).max()
    assert 1e-5 < rel < 5e-2


def test_header_declares_the_binding_list(hip):
    import lbmnn
    text = open(os.path.join(ROOT, "include", "lbm_nn.h")).read()
    names = sorted(set(re.findall(r"^\s*int32_t\s*(lbm_nn_\w+)\s*\(", text, flags=re.M)))
    assert names == sorted(lbmnn.EXPORTS)
    lib = hip.load()
    for n in names:
        assert hasattr(lib, n), n
    assert O.param_count() == 169107


def test_parameter_count_closed_form():
    """657 W^2 + 57 W + 3 parameters at base width W (the layer list of
    include/lbm_nn.h); the width is recovered from a vector's size."""
    for w in (4, 8, 16, 24, 28, 32):
        assert O.param_count(w) == 657 * w * w + 57 * w + 3
        assert O.width_of(np.zeros(O.param_count(w))) == w
    assert O.param_count(28) == 516687


def test_adam_closed_forms():
    """The oracle's Adam step against what the update rule fixes.
    A gradient held constant for t steps gives bias-corrected moments equal
    to g and g^2 exactly, so every step moves p by lr g / (|g| + eps): the
    bias corrections are pinned (without them the first step is 10x (m) or
    sqrt(1000)x (v) off).  A gradient of 1e-9 moves p by lr / 11 (eps added
    outside the square root, not inside: that would give lr 1e-9 /
    sqrt(1e-18 + 1e-8)).  A zero gradient does not move p."""
    lr = 1e-2
    g = np.array([3.0, -0.5, 1e-9, 0.0, -2e-3])
    p = np.array([1.0, -1.0, 0.5, 0.25, 0.0])
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    want_step = lr * g / (np.abs(g) + 1e-8)
    for t in range(1, 6):
        p1, m, v = O.adam_step(p, g, m, v, t, lr)
        np.testing.assert_allclose(p - p1, want_step, rtol=1e-12, atol=1e-18)
        p = p1
    step = 0.5 - O.adam_step(np.array([0.5]), np.array([1e-9]), np.zeros(1), np.zeros(1),
                             1, lr)[0][0]
    assert abs(step - lr / 11) <= 1e-12 * lr
    # the moments themselves: after one step from zero, (1-b1) g and (1-b2) g^2
    _, m1, v1 = O.adam_step(p, g, np.zeros_like(g), np.zeros_like(g), 1, lr)
    np.testing.assert_allclose(m1, 0.1 * g, rtol=1e-15)
    np.testing.assert_allclose(v1, 0.001 * g * g, rtol=1e-12)
