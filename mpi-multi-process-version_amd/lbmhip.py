"""ctypes binding of liblbm_hip.so, the C ABI of include/lbm_capi.h.

The Python mirror of the reference's driver-level interface: a Solver owns
one lbm_handle and exposes the reference's steps (initial, the time-loop
body, check, gather) as methods.  Fields use numpy arrays shaped (n, n) or
(n, n, n) with index order [k, j, i] (i fastest, the reference's Fortran
order).  There is no fallback: if the HIP library is missing or no GPU is
visible, construction raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_hip.so")

LBM_OK = 0
LBM_ERR_ARG = 1
LBM_ERR_HIP = 2
LBM_ERR_DIVERGED = 3
LBM_ERR_NOT_CONVERGED = 4
LBM_ERR_STATE = 5
LBM_ERR_NOMEM = 6

GRID_ERF = 0
GRID_TANH = 1
PREC_FP32_SHIFTED = 0
PREC_FP64 = 1

INFO_NAMES = ["L_LB", "dt", "nu", "kappa", "gbeta", "tauf", "Snu", "Sq",
              "sigk", "H"]

EXPORTS = ["lbm_create", "lbm_destroy", "lbm_get_grid", "lbm_get_info",
           "lbm_init_rest", "lbm_init_fields", "lbm_step", "lbm_check", "lbm_check_noalloc",
           "lbm_nusselt", "lbm_get_fields", "lbm_run", "lbm_get_itc",
           "lbm_time_steps", "lbm_synchronize", "lbm_state_size",
           "lbm_save_state", "lbm_load_state", "lbm_copy_state", "lbm_set_launch",
           "lbm_get_launch", "lbm_autotune", "lbm_bytes_per_update",
           "lbm_batch_create", "lbm_batch_destroy", "lbm_batch_cavity",
           "lbm_batch_init_rest", "lbm_batch_step", "lbm_batch_check",
           "lbm_batch_nusselt", "lbm_batch_run", "lbm_batch_active",
           "lbm_batch_time_steps", "lbm_batch_set_launch",
           "lbm_batch_get_launch", "lbm_batch_autotune",
           "lbm_status_string", "lbm_last_error"]


class LbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (status {code})")
        self.code = code


class Launch(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("bx", "by", "remap", "occupancy", "variant")]


class Params(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("nx", ctypes.c_int32),
                ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("Ra", ctypes.c_double), ("Pr", ctypes.c_double),
                ("Ma", ctypes.c_double), ("Thot", ctypes.c_double),
                ("Tcold", ctypes.c_double), ("stretch", ctypes.c_double),
                ("grid_kind", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("epsU", ctypes.c_double), ("epsT", ctypes.c_double),
                ("check_every", ctypes.c_int32), ("device", ctypes.c_int32),
                ("itc_max", ctypes.c_int64)]


_lib = None


def load():
    """Loads liblbm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(LIB_PATH)
        H = ctypes.c_void_p
        DP = ctypes.POINTER(ctypes.c_double)
        I32 = ctypes.c_int32
        I64 = ctypes.c_int64
        sig = {
            "lbm_create": [ctypes.POINTER(Params), ctypes.POINTER(H)],
            "lbm_destroy": [H],
            "lbm_get_grid": [H, DP, DP, DP],
            "lbm_get_info": [H, DP],
            "lbm_init_rest": [H],
            "lbm_init_fields": [H, DP, DP, DP, DP],
            "lbm_step": [H, I64],
            "lbm_check": [H, DP, DP],
            "lbm_nusselt": [H, DP, DP],
            "lbm_check_noalloc": [H, DP, DP],
            "lbm_get_fields": [H, DP, DP, DP, DP, DP],
            "lbm_run": [H, ctypes.POINTER(I64), DP, DP],
            "lbm_get_itc": [H, ctypes.POINTER(I64)],
            "lbm_time_steps": [H, I64, DP],
            "lbm_synchronize": [H],
            "lbm_state_size": [H, ctypes.POINTER(I64)],
            "lbm_save_state": [H, ctypes.c_void_p, I64],
            "lbm_load_state": [H, ctypes.c_void_p, I64],
            "lbm_copy_state": [H, H],
            "lbm_bytes_per_update": [H, DP],
            "lbm_set_launch": [H, ctypes.POINTER(Launch)],
            "lbm_get_launch": [H, ctypes.POINTER(Launch)],
            "lbm_autotune": [H, I32, DP],
            "lbm_batch_create": [ctypes.POINTER(Params), I32, ctypes.POINTER(H)],
            "lbm_batch_destroy": [H],
            "lbm_batch_init_rest": [H],
            "lbm_batch_step": [H, I64],
            "lbm_batch_check": [H, DP, DP],
            "lbm_batch_nusselt": [H, DP, DP],
            "lbm_batch_run": [H, ctypes.POINTER(I64), DP, DP, ctypes.POINTER(I32)],
            "lbm_batch_active": [H, ctypes.POINTER(I32), ctypes.POINTER(I64)],
            "lbm_batch_time_steps": [H, I64, DP],
            "lbm_batch_set_launch": [H, ctypes.POINTER(Launch)],
            "lbm_batch_get_launch": [H, ctypes.POINTER(Launch)],
            "lbm_batch_autotune": [H, I32, DP],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = I32
        L.lbm_batch_cavity.argtypes = [H, I32]
        L.lbm_batch_cavity.restype = H
        L.lbm_status_string.argtypes = [I32]
        L.lbm_status_string.restype = ctypes.c_char_p
        L.lbm_last_error.argtypes = []
        L.lbm_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(code, allow=()):
    if code != LBM_OK and code not in allow:
        L = load()
        raise LbmError(code, L.lbm_last_error().decode() or
                       L.lbm_status_string(code).decode())
    return code


class Solver:
    """One cavity on the GPU (the reference's `commondata` + main loop)."""

    def __init__(self, dim, n, Ra, Pr=0.71, Ma=0.1, Thot=1.0, Tcold=0.0,
                 stretch=3.2, grid_kind=GRID_ERF, precision=PREC_FP32_SHIFTED,
                 epsU=1e-6, epsT=1e-6, check_every=2000, itc_max=20000,
                 device=0):
        self._L = load()
        p = Params(dim=dim, nx=n, ny=n, nz=n if dim == 3 else 1, Ra=Ra, Pr=Pr,
                   Ma=Ma, Thot=Thot, Tcold=Tcold, stretch=stretch,
                   grid_kind=grid_kind, precision=precision, epsU=epsU,
                   epsT=epsT, check_every=check_every, device=device,
                   itc_max=itc_max)
        self.h = ctypes.c_void_p()
        _check(self._L.lbm_create(ctypes.byref(p), ctypes.byref(self.h)))
        self.params = p
        self.dim, self.n = dim, n
        self._owned = True

    @classmethod
    def _borrow(cls, handle, params):
        """A view of a cavity owned by a Batch (never destroyed by the view)."""
        s = cls.__new__(cls)
        s._L = load()
        s.h = ctypes.c_void_p(handle)
        s.params = params
        s.dim, s.n = params.dim, params.nx
        s._owned = False
        return s

    def close(self):
        if self.h and self._owned:
            self._L.lbm_destroy(self.h)
        self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.n,) * self.dim

    @property
    def itc(self):
        v = ctypes.c_int64()
        _check(self._L.lbm_get_itc(self.h, ctypes.byref(v)))
        return v.value

    def info(self):
        out = np.zeros(len(INFO_NAMES))
        _check(self._L.lbm_get_info(self.h, _dp(out)))
        return dict(zip(INFO_NAMES, out.tolist()))

    def grid(self):
        x = np.zeros(self.n)
        _check(self._L.lbm_get_grid(self.h, _dp(x), None, None))
        return x

    def init_rest(self):
        _check(self._L.lbm_init_rest(self.h))

    def init_fields(self, u, v, w, T):
        arrs = [None if a is None else
                np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
                for a in (u, v, w, T)]
        _check(self._L.lbm_init_fields(self.h, *[_dp(a) for a in arrs]))

    def step(self, nsteps=1):
        _check(self._L.lbm_step(self.h, int(nsteps)))

    def check(self, allow_diverged=False):
        eu, et = ctypes.c_double(), ctypes.c_double()
        _check(self._L.lbm_check(self.h, ctypes.byref(eu), ctypes.byref(et)),
               allow=(LBM_ERR_DIVERGED,) if allow_diverged else ())
        return eu.value, et.value

    def nusselt(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(self._L.lbm_nusselt(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def fields(self):
        """(rho, u, v, w, T) as new FP64 arrays."""
        outs = [np.zeros(self.shape) for _ in range(5)]
        _check(self._L.lbm_get_fields(self.h, *[_dp(o) for o in outs]))
        return tuple(outs)

    def run(self):
        """The reference's time loop; returns (status, itc, errU, errT)."""
        itc = ctypes.c_int64()
        eu, et = ctypes.c_double(), ctypes.c_double()
        code = self._L.lbm_run(self.h, ctypes.byref(itc), ctypes.byref(eu),
                               ctypes.byref(et))
        _check(code, allow=(LBM_ERR_NOT_CONVERGED, LBM_ERR_DIVERGED))
        return code, itc.value, eu.value, et.value

    def time_steps(self, nsteps):
        ms = ctypes.c_double()
        _check(self._L.lbm_time_steps(self.h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def synchronize(self):
        _check(self._L.lbm_synchronize(self.h))

    def bytes_per_update(self):
        b = ctypes.c_double()
        _check(self._L.lbm_bytes_per_update(self.h, ctypes.byref(b)))
        return b.value

    def set_launch(self, bx, by, remap=1, occupancy=0, variant=0):
        cfg = Launch(bx=bx, by=by, remap=remap, occupancy=occupancy, variant=variant)
        _check(self._L.lbm_set_launch(self.h, ctypes.byref(cfg)))

    def get_launch(self):
        """dict of the launch configuration that runs."""
        cfg = Launch()
        _check(self._L.lbm_get_launch(self.h, ctypes.byref(cfg)))
        return {k: getattr(cfg, k) for k, _ in Launch._fields_}

    def autotune(self, steps_per_trial=10):
        ms = ctypes.c_double()
        _check(self._L.lbm_autotune(self.h, int(steps_per_trial), ctypes.byref(ms)))
        return ms.value

    def save_state(self):
        size = ctypes.c_int64()
        _check(self._L.lbm_state_size(self.h, ctypes.byref(size)))
        buf = np.zeros(size.value, dtype=np.uint8)
        _check(self._L.lbm_save_state(self.h, buf.ctypes.data_as(ctypes.c_void_p),
                                      size.value))
        return buf

    def load_state(self, buf):
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        _check(self._L.lbm_load_state(self.h, buf.ctypes.data_as(ctypes.c_void_p),
                                      buf.size))

    def copy_state_from(self, src):
        """This solver takes src's state (lbm_copy_state), converted to this
        solver's precision: e.g. an FP64 solver continuing an FP32 run."""
        _check(self._L.lbm_copy_state(self.h, src.h))


class Batch:
    """Independent 2D cavities of any sizes and parameters, advanced
    together: one kernel launch per time step covers every active cavity.
    `cavities` is a list of dicts of Solver keyword arguments (n, Ra, Pr,
    stretch, epsU, epsT, itc_max, ...).  batch[i] is a Solver view of
    cavity i for the read-only calls and per-cavity initialisation."""

    def __init__(self, cavities, device=0, check_every=2000):
        self._L = load()
        self.count = len(cavities)
        arr = (Params * self.count)()
        for i, c in enumerate(cavities):
            c = dict(c)
            n = c.pop("n")
            arr[i] = Params(dim=2, nx=n, ny=n, nz=1, Ra=c.pop("Ra"), Pr=c.pop("Pr", 0.71),
                            Ma=c.pop("Ma", 0.1), Thot=c.pop("Thot", 1.0),
                            Tcold=c.pop("Tcold", 0.0), stretch=c.pop("stretch", 3.2),
                            grid_kind=c.pop("grid_kind", GRID_ERF),
                            precision=PREC_FP32_SHIFTED, epsU=c.pop("epsU", 1e-6),
                            epsT=c.pop("epsT", 1e-6), check_every=check_every,
                            device=device, itc_max=c.pop("itc_max", 20000))
            if c:
                raise TypeError(f"unknown cavity parameters {sorted(c)}")
        self.params = arr
        self.h = ctypes.c_void_p()
        _check(self._L.lbm_batch_create(arr, self.count, ctypes.byref(self.h)))
        self.cavities = [Solver._borrow(self._L.lbm_batch_cavity(self.h, i), arr[i])
                         for i in range(self.count)]

    def close(self):
        if self.h:
            for c in self.cavities:
                c.h = ctypes.c_void_p()
            self._L.lbm_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return self.count

    def __getitem__(self, i):
        return self.cavities[i]

    def init_rest(self):
        _check(self._L.lbm_batch_init_rest(self.h))

    def step(self, nsteps=1):
        _check(self._L.lbm_batch_step(self.h, int(nsteps)))

    def check(self):
        eu, et = np.zeros(self.count), np.zeros(self.count)
        _check(self._L.lbm_batch_check(self.h, _dp(eu), _dp(et)))
        return eu, et

    def nusselt(self):
        a, b = np.zeros(self.count), np.zeros(self.count)
        _check(self._L.lbm_batch_nusselt(self.h, _dp(a), _dp(b)))
        return a, b

    def run(self):
        """Each cavity to its own eps / itc_max; returns (itc, errU, errT,
        status) arrays."""
        itc = np.zeros(self.count, dtype=np.int64)
        eu, et = np.zeros(self.count), np.zeros(self.count)
        st = np.zeros(self.count, dtype=np.int32)
        _check(self._L.lbm_batch_run(
            self.h, itc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dp(eu), _dp(et),
            st.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return itc, eu, et, st

    def active(self):
        k, m = ctypes.c_int32(), ctypes.c_int64()
        _check(self._L.lbm_batch_active(self.h, ctypes.byref(k), ctypes.byref(m)))
        return k.value, m.value

    def time_steps(self, nsteps):
        ms = ctypes.c_double()
        _check(self._L.lbm_batch_time_steps(self.h, int(nsteps), ctypes.byref(ms)))
        return ms.value

    def set_launch(self, bx, by, variant=0):
        cfg = Launch(bx=bx, by=by, remap=0, occupancy=0, variant=variant)
        _check(self._L.lbm_batch_set_launch(self.h, ctypes.byref(cfg)))

    def get_launch(self):
        cfg = Launch()
        _check(self._L.lbm_batch_get_launch(self.h, ctypes.byref(cfg)))
        return {k: getattr(cfg, k) for k, _ in Launch._fields_}

    def autotune(self, steps_per_trial=10):
        ms = ctypes.c_double()
        _check(self._L.lbm_batch_autotune(self.h, int(steps_per_trial), ctypes.byref(ms)))
        return ms.value


# The batched sweep's seed (configuration C4: 64 cavities).
SWEEP_SEED = 20261203


def sweep_cavities(count=64, seed=SWEEP_SEED):
    """The batched sweep's cavities: edge from {65, 129, 257, 513}, erf
    stretch factor A ~ U[1, 2.5], log10 Ra ~ U[3, 6], Pr ~ U[0.7, 7], drawn
    in that order per cavity from numpy's PCG64 with the given seed."""
    rng = np.random.default_rng(seed)
    return [dict(n=int(rng.choice([65, 129, 257, 513])), stretch=float(rng.uniform(1.0, 2.5)),
                 Ra=float(10.0 ** rng.uniform(3.0, 6.0)), Pr=float(rng.uniform(0.7, 7.0)))
            for _ in range(count)]

