"""ctypes binding of the learned warm start (include/lbm_nn.h) in
liblbm_hip.so, and the helpers that turn cavity parameters into network
inputs and network outputs into the fields lbm_init_fields() takes.

Input channels: log10(Ra), Pr and the node coordinates x/H, y/H of the
stretched grid; output channels: u/V, v/V and T with V = sqrt(g beta H) / 4,
a quarter of the cavity's free-fall velocity in LB units.  There is no
fallback: a missing library or device raises.
"""
import ctypes

import numpy as np

import lbmhip

EXPORTS = ["lbm_nn_create", "lbm_nn_create_width", "lbm_nn_destroy", "lbm_nn_param_count",
           "lbm_nn_get_params",
           "lbm_nn_set_params", "lbm_nn_forward", "lbm_nn_loss_grad", "lbm_nn_train_step"]

_sig_done = False


def _lib():
    global _sig_done
    L = lbmhip.load()
    if not _sig_done:
        H = ctypes.c_void_p
        FP = ctypes.POINTER(ctypes.c_float)
        I32, I64 = ctypes.c_int32, ctypes.c_int64
        sig = {
            "lbm_nn_create": [I32, I32, ctypes.c_uint64, I32, ctypes.POINTER(H)],
            "lbm_nn_create_width": [I32, I32, I32, ctypes.c_uint64, I32, ctypes.POINTER(H)],
            "lbm_nn_destroy": [H],
            "lbm_nn_param_count": [H, ctypes.POINTER(I64)],
            "lbm_nn_get_params": [H, FP],
            "lbm_nn_set_params": [H, FP],
            "lbm_nn_forward": [H, I32, FP, FP],
            "lbm_nn_loss_grad": [H, I32, FP, FP, FP, FP],
            "lbm_nn_train_step": [H, I32, FP, FP, ctypes.c_float, FP],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = I32
        _sig_done = True
    return L


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Net:
    """The encoder-decoder on a G x G grid with base width `width`."""

    def __init__(self, grid=64, batch_max=32, seed=1, device=0, width=16):
        self._L = _lib()
        self.h = ctypes.c_void_p()
        lbmhip._check(self._L.lbm_nn_create_width(grid, width, batch_max, seed, device,
                                                  ctypes.byref(self.h)))
        self.grid, self.batch_max, self.width = grid, batch_max, width

    def close(self):
        if self.h:
            self._L.lbm_nn_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def param_count(self):
        c = ctypes.c_int64()
        lbmhip._check(self._L.lbm_nn_param_count(self.h, ctypes.byref(c)))
        return c.value

    def get_params(self):
        p = np.zeros(self.param_count, dtype=np.float32)
        lbmhip._check(self._L.lbm_nn_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = _f32(p)
        assert p.size == self.param_count
        lbmhip._check(self._L.lbm_nn_set_params(self.h, _fp(p)))

    def forward(self, x):
        x = _f32(x)
        b = x.shape[0]
        y = np.zeros((b, 3, self.grid, self.grid), dtype=np.float32)
        lbmhip._check(self._L.lbm_nn_forward(self.h, b, _fp(x), _fp(y)))
        return y

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        loss = ctypes.c_float()
        g = np.zeros(self.param_count, dtype=np.float32)
        lbmhip._check(self._L.lbm_nn_loss_grad(self.h, x.shape[0], _fp(x), _fp(t),
                                               ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def train_step(self, x, t, lr):
        x, t = _f32(x), _f32(t)
        loss = ctypes.c_float()
        lbmhip._check(self._L.lbm_nn_train_step(self.h, x.shape[0], _fp(x), _fp(t),
                                                float(lr), ctypes.byref(loss)))
        return loss.value


def free_fall_velocity(info):
    """U = sqrt(g beta H) of a cavity (lbm_get_info values)."""
    return float(np.sqrt(info["gbeta"] * info["H"]))


def velocity_scale(info):
    """V = U / 4, the velocity unit of the network's outputs: steady
    velocities reach about U / 4, so u/V, v/V and T all span about [-1, 1]
    and weigh alike in the squared error."""
    return 0.25 * free_fall_velocity(info)


def inputs(Ra, Pr, x, H):
    """Network input [4][G][G] of one cavity, each channel mapped onto
    [-1, 1] over the trained ranges: (log10 Ra - 4.5) / 1.5,
    (Pr - 3.85) / 3.15, and 2 x/H - 1, 2 y/H - 1 on the node grid
    (x: node coordinates, H: cavity size)."""
    G = len(x)
    xi = 2.0 * np.asarray(x) / H - 1.0
    X, Y = np.meshgrid(xi, xi)   # [j][i]
    return np.stack([np.full((G, G), (np.log10(Ra) - 4.5) / 1.5),
                     np.full((G, G), (Pr - 3.85) / 3.15), X, Y]).astype(np.float32)


def targets(u, v, T, V):
    """Network target [3][G][G] of one cavity's steady fields (V: the
    velocity scale, velocity_scale())."""
    return np.stack([u / V, v / V, T]).astype(np.float32)


def fields(y, V):
    """(u, v, T) in LB units from one network output y[3][G][G]."""
    y = np.asarray(y, dtype=np.float64)
    return y[0] * V, y[1] * V, y[2]


def erf_nodes(n, A):
    """Node coordinates x[n] and cavity size H (LB units) of the erf grid of
    n nodes with stretch A, the map of lbm_create (after
    natural-convection_2dim_mono.f90:85-116): X_i = (erf(A (i/(n+1) - 1/2))
    / erf(A/2) + 1) / 2 for i = 0..n+1, dx0 = X_1 - X_0, L = 1/dx0,
    x_i = (X_i - dx0/2) L for i = 1..n, H = L - 1."""
    from math import erf
    X = np.array([0.5 * (erf(A * (i / (n + 1.0) - 0.5)) / erf(0.5 * A) + 1.0)
                  for i in range(n + 2)])
    dx0 = X[1] - X[0]
    L = 1.0 / dx0
    return (X[1:n + 1] - 0.5 * dx0) * L, L - 1.0


def _lerp_matrix(src, dst):
    """Dense [len(dst)][len(src)] matrix of piecewise-linear interpolation
    from increasing points src to points dst inside [src[0], src[-1]]."""
    src, dst = np.asarray(src, np.float64), np.asarray(dst, np.float64)
    idx = np.clip(np.searchsorted(src, dst, side="right") - 1, 0, len(src) - 2)
    w = (dst - src[idx]) / (src[idx + 1] - src[idx])
    M = np.zeros((len(dst), len(src)))
    rows = np.arange(len(dst))
    M[rows, idx] = 1.0 - w
    M[rows, idx + 1] = w
    return M


def resample(y, s_src, s_dst, t_hot=1.0, t_cold=0.0):
    """A network output y[3][G][G] (u/V, v/V, T on the nodes s_src = x/H of
    its own grid, both axes alike) on another grid of the same cavity, the
    nodes s_dst = x/H: piecewise-linear in each axis over the nodes plus the
    four walls, where the wall values are the boundary conditions of the
    cavity (u = v = 0 on every wall, T = t_hot at x = 0 and t_cold at
    x = H, zero normal gradient of T at y = 0 and y = H).  Returns
    [3][len(s_dst)][len(s_dst)]."""
    y = np.asarray(y, np.float64)
    G = y.shape[-1]
    s = np.concatenate(([0.0], np.asarray(s_src, np.float64), [1.0]))
    aug = np.zeros((3, G + 2, G + 2))
    aug[:, 1:-1, 1:-1] = y
    aug[2, 1:-1, 0] = t_hot
    aug[2, 1:-1, -1] = t_cold
    aug[2, 0, :] = aug[2, 1, :]
    aug[2, -1, :] = aug[2, -2, :]
    M = _lerp_matrix(s, s_dst)
    return np.stack([M @ aug[c] @ M.T for c in range(3)])


def resample3(u, v, w, T, s_src, s_dst, t_hot=1.0, t_cold=0.0):
    """3D cube fields [k][j][i] on the nodes s_src = x/H (every axis alike)
    onto the nodes s_dst: piecewise-linear in each axis over the nodes plus
    the six walls, whose values are the cube's boundary conditions (no-slip
    everywhere, T = t_hot at y = 0 and t_cold at y = H, zero normal gradient
    of T on the x and z walls).  Returns (u, v, w, T), each
    [len(s_dst)]^3."""
    n = len(s_src)
    s = np.concatenate(([0.0], np.asarray(s_src, np.float64), [1.0]))
    M = _lerp_matrix(s, s_dst)
    out = []
    for c, f in enumerate((u, v, w, T)):
        aug = np.zeros((n + 2, n + 2, n + 2))
        aug[1:-1, 1:-1, 1:-1] = np.asarray(f, np.float64).reshape(n, n, n)
        if c == 3:
            aug[1:-1, 0, 1:-1] = t_hot       # y = 0 (the j axis is the middle index)
            aug[1:-1, -1, 1:-1] = t_cold
            aug[:, :, 0] = aug[:, :, 1]      # x walls
            aug[:, :, -1] = aug[:, :, -2]
            aug[0] = aug[1]                  # z walls
            aug[-1] = aug[-2]
            aug[0, 0, :] = aug[-1, 0, :] = t_hot   # the edges of the heated walls
            aug[0, -1, :] = aug[-1, -1, :] = t_cold
        out.append(np.einsum("ka,jb,ic,abc->kji", M, M, M, aug, optimize=True))
    return tuple(out)


def extrude(u, v, T, n):
    """The 3D cube's fields from a 2D prediction on its (y, z) plane: the
    3D cavity has its heated walls at y = 0 and y = H and gravity along +z,
    so the 2D (x, y) become (y, z).  Every x slice gets the same plane:
    u3 = 0, v3 = u, w3 = v, T3 = T, each [n][n][n] with i fastest."""
    shape = (n, n, n)
    return (np.zeros(shape), np.broadcast_to(u[:, :, None], shape).copy(),
            np.broadcast_to(v[:, :, None], shape).copy(),
            np.broadcast_to(T[:, :, None], shape).copy())
