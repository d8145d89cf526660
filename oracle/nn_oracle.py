"""numpy restatement of the warm-start network (include/lbm_nn.h), at any
base width W (the width follows from the parameter vector's size).

TEST INFRASTRUCTURE ONLY: the checker of the HIP network.  Only tests/ may
import this module.

The layer list, parameter layout (per layer: weights [out][in][ky][kx], then
biases), zero padding, nearest x2 upsampling, channel concatenation (the
upsampled map first, then the encoder map) and mean squared error follow
the header.  With quant=True the BF16 roundings of the HIP products are
applied at the same points (weights and im2col inputs in the forward
products; the output gradient dZ, the weights and the inputs in the
backward products), with FP64 accumulation elsewhere; with quant=False
everything is FP64, which the finite-difference tests use.
"""
import numpy as np

def layers(width=16):
    """(cin, cout, kernel, stride, relu) of the 8 convolutions at base width W."""
    W = width
    return [(4, W, 3, 1, True), (W, 2 * W, 3, 2, True), (2 * W, 4 * W, 3, 2, True),
            (4 * W, 4 * W, 3, 2, True), (8 * W, 4 * W, 3, 1, True), (6 * W, 2 * W, 3, 1, True),
            (3 * W, W, 3, 1, True), (W, 3, 1, 1, False)]


LAYERS = layers(16)


def param_count(width=16):
    return sum(co * ci * k * k + co for ci, co, k, _, _ in layers(width))


def width_of(p):
    """The base width of a flat parameter vector (its size fixes it)."""
    n = np.asarray(p).size
    for w in range(4, 257):
        if param_count(w) == n:
            return w
    raise ValueError(f"no width has {n} parameters")


def split_params(p):
    out, o = [], 0
    for ci, co, k, s, relu in layers(width_of(p)):
        nw = co * ci * k * k
        out.append((p[o:o + nw].reshape(co, ci, k, k), p[o + nw:o + nw + co]))
        o += nw + co
    return out


def bf16(a):
    """Round to BF16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _q(a, quant):
    return bf16(a) if quant else np.asarray(a, dtype=np.float64)


def conv(X, W, b, stride, quant):
    """Zero-padded convolution (cross-correlation) of X[B][C][H][W]."""
    k = W.shape[2]
    pad = k // 2
    Xq, Wq = _q(X, quant), _q(W, quant)
    B, C, H, _ = X.shape
    Ho = H // stride
    Xp = np.pad(Xq, ((0, 0), (0, 0), (pad, pad), (pad, pad)))
    out = np.zeros((B, W.shape[0], Ho, Ho))
    for ky in range(k):
        for kx in range(k):
            patch = Xp[:, :, ky:ky + stride * Ho:stride, kx:kx + stride * Ho:stride]
            out += np.einsum("oi,bihw->bohw", Wq[:, :, ky, kx], patch)
    return out + np.asarray(b, dtype=np.float64)[None, :, None, None]


def conv_bwd(dZ, X, W, stride, quant, need_dx=True):
    """Gradients of conv() with respect to W, b and X, given dZ."""
    k = W.shape[2]
    pad = k // 2
    dZq, Xq, Wq = _q(dZ, quant), _q(X, quant), _q(W, quant)
    B, C, H, _ = X.shape
    Ho = dZ.shape[2]
    Xp = np.pad(Xq, ((0, 0), (0, 0), (pad, pad), (pad, pad)))
    dW = np.zeros(W.shape)
    dXp = np.zeros(Xp.shape)
    for ky in range(k):
        for kx in range(k):
            sl = (slice(None), slice(None), slice(ky, ky + stride * Ho, stride),
                  slice(kx, kx + stride * Ho, stride))
            dW[:, :, ky, kx] = np.einsum("bohw,bihw->oi", dZq, Xp[sl])
            if need_dx:
                dXp[sl] += np.einsum("oi,bohw->bihw", Wq[:, :, ky, kx], dZq)
    db = np.asarray(dZ, dtype=np.float64).sum(axis=(0, 2, 3))
    dX = dXp[:, :, pad:pad + H, pad:pad + H] if need_dx else None
    return dW, db, dX


def up2(X):
    return X.repeat(2, axis=2).repeat(2, axis=3)


def up2_bwd(dY):
    B, C, H, W = dY.shape
    return dY.reshape(B, C, H // 2, 2, W // 2, 2).sum(axis=(3, 5))


def forward(p, x, quant=True, keep=False):
    """y = net(x) for x[B][4][G][G]; keep=True also returns the activations."""
    P = split_params(np.asarray(p, dtype=np.float64))
    relu = lambda z: np.maximum(z, 0.0)
    a1 = relu(conv(x, *P[0], 1, quant))
    a2 = relu(conv(a1, *P[1], 2, quant))
    a3 = relu(conv(a2, *P[2], 2, quant))
    a4 = relu(conv(a3, *P[3], 2, quant))
    c5 = np.concatenate([up2(a4), a3], axis=1)
    a5 = relu(conv(c5, *P[4], 1, quant))
    c6 = np.concatenate([up2(a5), a2], axis=1)
    a6 = relu(conv(c6, *P[5], 1, quant))
    c7 = np.concatenate([up2(a6), a1], axis=1)
    a7 = relu(conv(c7, *P[6], 1, quant))
    y = conv(a7, *P[7], 1, quant)
    if keep:
        return y, dict(x=np.asarray(x, dtype=np.float64), a1=a1, a2=a2, a3=a3, a4=a4,
                       c5=c5, a5=a5, c6=c6, a6=a6, c7=c7, a7=a7)
    return y


def loss_grad(p, x, t, quant=True):
    """Mean squared error over all outputs and its parameter gradient."""
    P = split_params(np.asarray(p, dtype=np.float64))
    W = width_of(p)
    L = layers(W)
    y, A = forward(p, x, quant, keep=True)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    grads = [None] * len(L)
    mask = lambda dA, a: dA * (a > 0)

    def bwd(i, dZ, X, need_dx=True):
        Wt, _ = P[i]
        dW, db, dX = conv_bwd(dZ, X, Wt, L[i][3], quant, need_dx)
        grads[i] = (dW, db)
        return dX

    da7 = bwd(7, dy, A["a7"])
    dc7 = bwd(6, mask(da7, A["a7"]), A["c7"])
    da6, da1 = up2_bwd(dc7[:, :2 * W]), dc7[:, 2 * W:]
    dc6 = bwd(5, mask(da6, A["a6"]), A["c6"])
    da5, da2 = up2_bwd(dc6[:, :4 * W]), dc6[:, 4 * W:]
    dc5 = bwd(4, mask(da5, A["a5"]), A["c5"])
    da4, da3 = up2_bwd(dc5[:, :4 * W]), dc5[:, 4 * W:]
    da3 = da3 + bwd(3, mask(da4, A["a4"]), A["a3"])
    da2 = da2 + bwd(2, mask(da3, A["a3"]), A["a2"])
    da1 = da1 + bwd(1, mask(da2, A["a2"]), A["a1"])
    bwd(0, mask(da1, A["a1"]), A["x"], need_dx=False)
    g = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma and Ba 2015, Algorithm 1) at step count t >= 1:
    m <- b1 m + (1-b1) g, v <- b2 v + (1-b2) g^2, then
    p <- p - lr (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps).  FP64; returns
    the new (p, m, v).  This is the update lbm_nn_train_step documents
    (beta1 0.9, beta2 0.999, eps 1e-8, the state cleared by set_params)."""
    g = np.asarray(g, dtype=np.float64)
    m = b1 * np.asarray(m, dtype=np.float64) + (1.0 - b1) * g
    v = b2 * np.asarray(v, dtype=np.float64) + (1.0 - b2) * g * g
    mhat = m / (1.0 - b1 ** t)
    vhat = v / (1.0 - b2 ** t)
    return np.asarray(p, dtype=np.float64) - lr * mhat / (np.sqrt(vhat) + eps), m, v
