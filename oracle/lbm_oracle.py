"""ctypes binding of the CPU FP64 oracle (liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker that the HIP product is compared
against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py may import this module.  The oracle restates the reference solver
(natural-convection_2dim_mono.f90 / natural-convection_3dim_mono.f90); see
lbm_oracle.h for the function-by-function citations.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblbm_oracle.so")
_lib = None

PARAM_NAMES = ["L_LB", "dt", "dx0", "nu", "kappa", "gbeta", "tauf", "Snu",
               "Sq", "sigk", "H"]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.c_double
        DP = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, D]
        L.orc_create_grid.restype = P
        L.orc_create_grid.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, D,
                                      ctypes.c_int]
        L.orc_destroy.argtypes = [P]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        for name in ["orc_dim", "orc_n", "orc_qf", "orc_qg"]:
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = ctypes.c_int
        L.orc_itc.argtypes = [P]
        L.orc_itc.restype = ctypes.c_long
        L.orc_params.argtypes = [P, DP]
        L.orc_coords.argtypes = [P, DP]
        L.orc_stencil.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.orc_array.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)]
        L.orc_array.restype = DP
        L.orc_init_fields.argtypes = [P, DP, DP, DP, DP]
        for name in ["orc_init_rest", "orc_collision_u", "orc_collision_t",
                     "orc_interpolate", "orc_bounceback_u", "orc_bounceback_t",
                     "orc_macro_u", "orc_macro_t"]:
            getattr(L, name).argtypes = [P]
        L.orc_step.argtypes = [P, ctypes.c_long]
        L.orc_check.argtypes = [P, DP, DP]
        L.orc_run.argtypes = [P, D, D, ctypes.c_long, ctypes.c_int, DP, DP]
        L.orc_run.restype = ctypes.c_long
        L.orc_nusselt.argtypes = [P, DP, DP]
        L.orc_interpolateF.argtypes = [D] * 7
        L.orc_interpolateF.restype = D
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().orc_set_threads(int(n))


class Oracle:
    """One cavity of the FP64 CPU oracle.  Arrays are numpy views of the
    oracle's own memory, shaped (n, n[, n], Q) for populations and
    (n, n[, n]) for fields with index order [k, j, i] (i fastest)."""

    def __init__(self, dim, n, Ra, Pr=0.71, Ma=0.1, Thot=1.0, Tcold=0.0,
                 A=3.2, grid_kind=0):
        self._l = lib()
        self.h = self._l.orc_create_grid(dim, n, Ra, Pr, Ma, Thot, Tcold, A, grid_kind)
        if not self.h:
            raise ValueError("orc_create rejected the arguments")
        self.dim, self.n = dim, n
        self.Thot, self.Tcold = Thot, Tcold

    def close(self):
        if self.h:
            self._l.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.n,) * self.dim

    @property
    def qf(self):
        return self._l.orc_qf(self.h)

    @property
    def qg(self):
        return self._l.orc_qg(self.h)

    @property
    def itc(self):
        return self._l.orc_itc(self.h)

    def params(self):
        out = np.zeros(len(PARAM_NAMES))
        self._l.orc_params(self.h, _dp(out))
        return dict(zip(PARAM_NAMES, out.tolist()))

    def coords(self):
        out = np.zeros(self.n)
        self._l.orc_coords(self.h, _dp(out))
        return out

    def stencil(self):
        out = np.zeros((self.n, 3), dtype=np.int32)
        self._l.orc_stencil(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def array(self, name):
        cnt = ctypes.c_long(0)
        p = self._l.orc_array(self.h, name.encode(), ctypes.byref(cnt))
        if not p:
            raise KeyError(name)
        flat = np.ctypeslib.as_array(p, shape=(cnt.value,))
        nn = self.n ** self.dim
        if cnt.value == nn:
            return flat.reshape(self.shape)
        return flat.reshape(self.shape + (cnt.value // nn,))

    def init_rest(self):
        self._l.orc_init_rest(self.h)

    def init_fields(self, u, v, w, T):
        arrs = []
        ptrs = []
        for a in (u, v, w, T):
            if a is None:
                ptrs.append(None)
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                arrs.append(a)
                ptrs.append(_dp(a))
        self._l.orc_init_fields(self.h, *ptrs)

    def collision_u(self):
        self._l.orc_collision_u(self.h)

    def collision_t(self):
        self._l.orc_collision_t(self.h)

    def interpolate(self):
        self._l.orc_interpolate(self.h)

    def bounceback_u(self):
        self._l.orc_bounceback_u(self.h)

    def bounceback_t(self):
        self._l.orc_bounceback_t(self.h)

    def macro_u(self):
        self._l.orc_macro_u(self.h)

    def macro_t(self):
        self._l.orc_macro_t(self.h)

    def step(self, n=1):
        self._l.orc_step(self.h, int(n))

    def check(self):
        eu, et = ctypes.c_double(), ctypes.c_double()
        self._l.orc_check(self.h, ctypes.byref(eu), ctypes.byref(et))
        return eu.value, et.value

    def run(self, epsU, epsT, itc_max, check_every=2000):
        eu, et = ctypes.c_double(), ctypes.c_double()
        itc = self._l.orc_run(self.h, epsU, epsT, int(itc_max), int(check_every),
                              ctypes.byref(eu), ctypes.byref(et))
        return itc, eu.value, et.value

    def nusselt(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._l.orc_nusselt(self.h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def fields(self):
        """Copies of (rho, u, v, w, T)."""
        return tuple(self.array(k).copy() for k in ("rho", "u", "v", "w", "T"))


def interpolateF(x0, x1, x2, x, f0, f1, f2):
    return lib().orc_interpolateF(x0, x1, x2, x, f0, f1, f2)
