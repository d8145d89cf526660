"""lbm_check_noalloc: lbm_check with its reduction's partial sums in the
library's static device scratch array instead of the handle's buffers (for
callers that cannot allocate).  Called through the C ABI, it must give
lbm_check's results bit for bit, on two calls in a row (the second on a
smaller grid, so fewer partial sums than the first left in the array) and
over repeated checks of one handle (the snapshot update is the same)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _twins(hip, dim, n, Ra, steps, precision):
    a = hip.Solver(dim, n, Ra, precision=precision)
    b = hip.Solver(dim, n, Ra, precision=precision)
    for s in (a, b):
        s.init_rest()
        s.step(steps)
    return a, b


def _noalloc(hip, s):
    L = hip.load()
    eu, et = ctypes.c_double(), ctypes.c_double()
    rc = L.lbm_check_noalloc(s.h, ctypes.byref(eu), ctypes.byref(et))
    assert rc == hip.LBM_OK, L.lbm_last_error().decode()
    return eu.value, et.value


@pytest.mark.parametrize("precision", [0, 1])
def test_check_noalloc_matches_check(hip, precision):
    # a large 3D grid (433 reduction blocks), then a small 2D one (5 blocks)
    for dim, n, Ra, steps in ((3, 48, 1e5, 30), (2, 33, 1e4, 50)):
        print(f"\n[check_noalloc] dim={dim} n={n} precision={precision}", flush=True)
        a, b = _twins(hip, dim, n, Ra, steps, precision)
        for k in range(3):
            ra = a.check()
            rb = _noalloc(hip, b)
            print(f"  check {k}: {ra} vs {rb}", flush=True)
            assert ra == rb, (k, ra, rb)
            if k == 0:
                assert ra == (1.0, 1.0)     # the first check after init is 1
            a.step(7)
            b.step(7)
        for x, y in zip(a.fields(), b.fields()):
            assert np.array_equal(x, y)
        a.close()
        b.close()


def test_check_noalloc_rejects_an_uninitialised_handle(hip):
    s = hip.Solver(2, 17, 1e4)
    L = hip.load()
    eu, et = ctypes.c_double(), ctypes.c_double()
    assert L.lbm_check_noalloc(s.h, ctypes.byref(eu), ctypes.byref(et)) == hip.LBM_ERR_STATE
    s.close()
