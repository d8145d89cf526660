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

EXPORTS = ["lbm_nn_create", "lbm_nn_destroy", "lbm_nn_param_count", "lbm_nn_get_params",
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
    """The encoder-decoder on a G x G grid."""

    def __init__(self, grid=64, batch_max=32, seed=1, device=0):
        self._L = _lib()
        self.h = ctypes.c_void_p()
        lbmhip._check(self._L.lbm_nn_create(grid, batch_max, seed, device, ctypes.byref(self.h)))
        self.grid, self.batch_max = grid, batch_max

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
