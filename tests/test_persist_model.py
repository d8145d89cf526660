"""CPU model of the batch's persistent launch (variant 2,
k_step_batch_persist in csrc/lbm_step_f32.hip): the ticket decoding and
the dependency rule, replayed on the host for the tile tables that the
batch builds.  Checks that
  * the tickets 0 .. steps*ntiles - 1 decode to every (step, tile) item
    exactly once, and the next ticket to none;
  * every tile an item waits for (the tiles within the stencil's reach,
    one step earlier) holds a smaller ticket, so a wait never targets a
    ticket that no running workgroup holds and the launch cannot deadlock;
  * the tiles waited for cover every node the item's stencils read (the
    one-sided wall stencils reach two nodes inward) and every node whose
    previous-step readers the item's writes would overtake;
  * every counter index stays inside the cavity's block of the table.
"""
import itertools

import numpy as np
import pytest


def decode(k, steps, depth, groups, ntiles):
    """Mirror of persist_decode: ticket k -> (table entry, step 1..steps)."""
    T, DT = ntiles, depth * ntiles
    nfull = steps // depth
    if k < nfull * DT:
        chunk, r, d = k // DT, k % DT, depth
    else:
        chunk, r, d = nfull, k - nfull * DT, steps - nfull * depth
        if r >= d * T:
            return None
    g = 0
    while g + 1 < len(groups) - 1 and r >= d * groups[g + 1]:
        g += 1
    g0, ng = groups[g], groups[g + 1] - groups[g]
    rr = r - d * g0
    return g0 + rr % ng, chunk * depth + rr // ng + 1


def build_table(sizes, bx, by, group_nodes):
    """The variant-0/2 tile table of rebuild(): cavities largest first, each
    cavity's tiles y-major; MALL groups of whole cavities."""
    order = sorted(range(len(sizes)), key=lambda c: -sizes[c])
    tiles, groups, cav, g = [], [0], {}, 0.0
    for c in order:
        n = sizes[c]
        if g > 0 and g + n * n > group_nodes:
            groups.append(len(tiles))
            g = 0.0
        g += n * n
        gx, gy = -(-n // bx), -(-n // by)
        cav[c] = (len(tiles), gx, gy)
        for ty in range(gy):
            for tx in range(gx):
                tiles.append((c, tx, ty))
    groups.append(len(tiles))
    return tiles, groups, cav


def stencil_rows(j, n):
    """Rows node row j reads: its 3-point stencil (one-sided at the walls)."""
    if j == 0:
        return [1, 0, 2]
    if j == n - 1:
        return [n - 2, n - 1, n - 3]
    return [j - 1, j, j + 1]


@pytest.mark.parametrize("sizes,bx,by,steps,depth,group_nodes", [
    ([17, 33, 24, 65, 40], 64, 8, 7, 7, 1e9),
    ([17, 33, 24, 65, 40], 64, 1, 5, 5, 1e9),      # 1-row tiles: two rows of reach
    ([17, 33, 24, 65, 40], 32, 2, 9, 9, 1e9),
    ([65, 129, 65, 129, 33, 17], 64, 4, 11, 4, 130 ** 2),   # MALL groups, partial chunk
    ([257, 129, 65], 128, 4, 3, 64, 200 ** 2),
])
def test_persistent_schedule(sizes, bx, by, steps, depth, group_nodes):
    tiles, groups, cav = build_table(sizes, bx, by, group_nodes)
    T = len(tiles)
    ry = 1 if by >= 2 else 2
    ticket = {}
    for k in range(steps * T):
        it = decode(k, steps, depth, groups, T)
        assert it is not None and it not in ticket, (k, it)
        ticket[it] = k
    assert decode(steps * T, steps, depth, groups, T) is None
    assert len(ticket) == steps * T
    for (e, s), k in ticket.items():
        c, tx, ty = tiles[e]
        first, gx, gy = cav[c]
        n = sizes[c]
        waits = set()
        if s > 1:
            for dy, dx in itertools.product(range(-ry, ry + 1), (-1, 0, 1)):
                ty2, tx2 = ty + dy, tx + dx
                if 0 <= ty2 < gy and 0 <= tx2 < gx:
                    e2 = first + ty2 * gx + tx2
                    assert first <= e2 < first + gx * gy and tiles[e2] == (c, tx2, ty2)
                    assert ticket[(e2, s - 1)] < k
                    waits.add((tx2, ty2))
            # reads: every stencil point of every node of the tile
            for j in range(ty * by, min(n, (ty + 1) * by)):
                for jj in stencil_rows(j, n):
                    for i in (tx * bx, min(n, (tx + 1) * bx) - 1):
                        for ii in stencil_rows(i, n):
                            assert (ii // bx, jj // by) in waits
            # writes: every node that reads one of this tile's nodes
            for j in range(n):
                for jj in stencil_rows(j, n):
                    if ty * by <= jj < (ty + 1) * by:
                        assert (tx, j // by) in waits or j // by == ty
