"""CPU model of the batch's resident launch (variant 3, k_step_batch_res in
csrc/lbm_step_res.h, launched by launch_step_batch_res_f32 in
csrc/lbm_step_f32.hip).  The constants are read from the sources, and the
kernel's indexing is replayed on the host for every edge it accepts.
Checks that
  * the launcher's choice of nodes per thread K covers every node: thread t
    owns q = t + T k, k < K, and every q < n^2 is owned exactly once;
  * every LDS index stays inside the static array: the eight flow planes
    (and the four thermal ones in the same place), the stencil and heat
    tables after them, and the array fits the 160 KB of one CU;
  * every stencil point a node's interpolation reads (the one-sided wall
    stencils of natural-convection_2dim_mono.f90:242-252 included, along x
    and y) is a node of the same cavity, i.e. inside its LDS plane;
  * the packed (i, j) of a node (i | j << 16) round-trips;
  * the output buffer after s steps is the one the cavity's parity expects.
"""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "mpi-multi-process-version_amd", "csrc")


def _const(name):
    src = open(os.path.join(CSRC, "lbm_step_res.h")).read()
    m = re.search(rf"constexpr int {name} = (\d+);", src)
    return int(m.group(1))


T = _const("kResThreads")
NMAX = _const("kResMaxN")
KMAX = (NMAX * NMAX + T - 1) // T
PLANES = 8 * NMAX * NMAX
TABLES = 11 * NMAX


def _instantiated_k():
    """The K values launch_step_batch_res_f32 dispatches to, in order."""
    src = open(os.path.join(CSRC, "lbm_step_f32.hip")).read()
    body = src[src.index("hipError_t launch_step_batch_res_f32"):]
    body = body[:body.index("return hipGetLastError();")]
    ks = [int(x) for x in re.findall(r"launch_res<(\d+)>", body)]
    if "launch_res<kResMaxK>" in body:
        ks.append(KMAX)
    return ks


def _launch_k(n):
    """Mirror of the launcher: the first instantiated K >= the nodes per
    thread of an n x n cavity."""
    need = (n * n + T - 1) // T
    for k in _instantiated_k():
        if need <= k:
            return k
    return None


def _stencil(n):
    """Offsets (o0, o2) of stencil points 0 and 2 per grid index: centred
    inside, one-sided (2, 1, 3) and (n-1, n, n-2) at the walls."""
    off = np.zeros((n, 2), dtype=np.int64)
    for i in range(n):
        if i == 0:
            off[i] = (1, 2)
        elif i == n - 1:
            off[i] = (-1, -2)
        else:
            off[i] = (-1, 1)
    return off


def test_constants_and_capacity():
    assert T == 512 and NMAX == 70 and KMAX == 10
    # the static LDS array: planes + tables, 4-byte floats, within one CU
    assert (PLANES + TABLES) * 4 <= 160 * 1024
    # the table region: off (2 per index), wt (6), tref (1), tcorr (2)
    assert 2 * NMAX + 6 * NMAX + NMAX + 2 * NMAX == TABLES
    # int2 offsets start 8-byte aligned
    assert (PLANES * 4) % 8 == 0


@pytest.mark.parametrize("n", list(range(5, NMAX + 1)))
def test_ownership_and_lds_bounds(n):
    N = n * n
    K = _launch_k(n)
    assert K is not None and K <= KMAX
    owners = np.zeros(N, dtype=np.int64)
    for t in range(T):
        for k in range(K):
            q = t + T * k
            if q < N:
                owners[q] += 1
                i, j = q % n, q // n
                packed = i | (j << 16)
                assert (packed & 0xFFFF, packed >> 16) == (i, j)
    assert np.all(owners == 1)
    # flow planes 0..7 and thermal planes 0..3 at (a - 1) * N + q
    assert 8 * N <= PLANES
    off = _stencil(n)
    q = np.arange(N)
    i, j = q % n, q // n
    for dxs in (off[i, 0], off[i, 1], np.zeros(N, dtype=np.int64)):
        for dys in (off[j, 0], off[j, 1], np.zeros(N, dtype=np.int64)):
            ii, jj = i + dxs, j + dys
            assert ii.min() >= 0 and ii.max() < n and jj.min() >= 0 and jj.max() < n
            r = q + dxs + n * dys
            assert r.min() >= 0 and r.max() < N


@pytest.mark.parametrize("cur,steps", [(0, 1), (1, 1), (0, 2), (1, 4101), (0, 4096)])
def test_output_buffer_parity(cur, steps):
    """args[2c + p] reads buffer p and writes 1 - p; after `steps` steps from
    buffer cur the state must be in buffer cur ^ (steps & 1)."""
    pin = lambda ai: ai & 1
    pout = lambda ai: 1 - (ai & 1)
    ai = 2 * 3 + cur
    final = pout(ai) if steps & 1 else pout(ai ^ 1)
    assert pin(ai) == cur
    assert final == cur ^ (steps & 1)
