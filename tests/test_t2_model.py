"""CPU model of the two-step pass (launch variant 2 of a 2D FP32 cavity,
k_step_t2 in csrc/lbm_step_t2.h, launched by launch_step2_f32 in
csrc/lbm_step_f32.hip).  For every tile height and a range of grid sizes,
replays the tiling and the two update regions on the host and checks that
  * the output tiles cover every node of the cavity (clamped tiles may
    overlap their neighbour, which writes the same values);
  * every stencil point of a node updated in the first step (centred, or
    one-sided at the walls, natural-convection_2dim_mono.f90:242-252) holds
    P(k) in the tile's region, and every stencil point of an output node
    was updated in the first step, so no node reads a value the pass has
    not produced;
  * every LDS index (region plane q = lane + 64 r) stays inside the planes,
    and the region rows split evenly over the four waves.
"""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = open(os.path.join(ROOT, "mpi-multi-process-version_amd", "csrc", "lbm_step_t2.h")).read()
W = int(re.search(r"constexpr int kT2Width = (\d+);", SRC).group(1))
OUT = W - 4
THREADS = int(re.search(r"constexpr int kT2Threads = (\d+);", SRC).group(1))


def stencil_points(i, n):
    """Grid indices of the three interpolation points of index i."""
    if i == 0:
        return (1, 0, 2)
    if i == n - 1:
        return (n - 2, n - 1, n - 3)
    return (i - 1, i, i + 1)


def tiles(n, ty):
    gx, gy = -(-n // OUT), -(-n // ty)
    for t in range(gx * gy):
        yield min((t % gx) * OUT, n - 2), min((t // gx) * ty, n - 2)


@pytest.mark.parametrize("ty", [8, 12, 16])
@pytest.mark.parametrize("n", [5, 6, 9, 17, 33, 59, 60, 61, 62, 63, 64, 65, 119, 120, 121,
                               122, 125, 129, 181, 257])
def test_tiles_cover_and_stencils_stay_inside(n, ty):
    H = ty + 4
    assert H % 4 == 0 and (THREADS // 64) == 4
    covered = np.zeros((n, n), dtype=np.int64)
    for x0, y0 in tiles(n, ty):
        # region: lanes l -> i = x0 - 2 + l, rows r -> j = y0 - 2 + r
        def region(m):
            return {(x0 - 2 + l, y0 - 2 + r) for l in range(m, W - m) for r in range(m, H - m)
                    if 0 <= x0 - 2 + l < n and 0 <= y0 - 2 + r < n}
        r0, r1, r2 = region(0), region(1), region(2)
        for (i, j) in r2:
            covered[j, i] += 1
        for upd, src in ((r1, r0), (r2, r1)):
            for (i, j) in upd:
                for pi in stencil_points(i, n):
                    for pj in stencil_points(j, n):
                        assert (pi, pj) in src, (n, ty, x0, y0, i, j, pi, pj)
        for (i, j) in r0:
            q = (i - (x0 - 2)) + W * (j - (y0 - 2))
            assert 0 <= q < W * H
    assert covered.min() >= 1
