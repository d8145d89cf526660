"""A CPU model of the 2D y-march's indexing (csrc/lbm_step_ym.h) -- CPU only.

Restates the LDS layout, the per-lane fetch and commit of a row (own values
and the three halo columns), the register ring with its wall-row copies and
the tile tables (one cavity, and the batch's table dealt to the 8 XCDs), and
checks for odd, even and 64k+1 grids:
  * every LDS index written or read lies in its population's row;
  * every value a lane reads from LDS or from the ring is the stencil point
    its interpolation needs (population, row, column);
  * every strip chunk is run by exactly one workgroup and every row of a
    chunk once, with the wall copies at k = 0 and k = n-1.
"""
import itertools
import math

import pytest

C2 = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]
NP = 14
PITCH, C0 = 72, 4


def vel(p):
    return C2[p if p < 9 else p - 9]


def rest(p):
    return p in (0, 9)


def xonly(p):
    return vel(p)[0] != 0 and vel(p)[1] == 0


def diag(p):
    return vel(p)[0] != 0 and vel(p)[1] != 0


def yonly(p):
    return vel(p)[0] == 0 and vel(p)[1] != 0


def slot(p):
    same = xonly if xonly(p) else diag
    return sum(1 for q in range(p) if same(q))


def region(p):
    return 0 if xonly(p) else 4 * PITCH


def stencil(i, n):
    if i == 0:
        return 1, 2
    if i == n - 1:
        return -1, -2
    return -1, 1


def window(k, n):
    return min(max(k - 1, 0), n - 3)


def halo_col(i0, h):
    return (i0 - 2, i0 - 1, i0 + 64)[h % 3]


def rows_per_chunk(strip_rows):
    return max(8, math.ceil(strip_rows / 4096.0))


ORDER = {0: (0, 1, 2), -1: (1, 0, 2), 1: (1, 2, 0)}


def march(n, i0, k0, k1, order=ORDER):
    lds = {}
    checked = 0
    lanes = range(64)

    def commit(xo, kk):
        kk = min(kk, n - 1)
        pops = [p for p in range(NP) if (xonly(p) if xo else diag(p))]
        assert [slot(p) for p in pops] == [0, 1, 2, 3]
        for l in lanes:
            i = min(i0 + l, n - 1)
            for p in pops:
                idx = region(p) + slot(p) * PITCH + C0 + l
                assert 0 <= idx < 8 * PITCH
                lds[idx] = (p, kk, i)
            if l < 12:
                hc = halo_col(i0, l)
                if 0 <= hc < n:
                    p = pops[l // 3]
                    col = (2, 3, 68)[l % 3]
                    lds[region(p) + slot(p) * PITCH + col] = (p, kk, hc)

    def xread(p, l, i, dx, m):
        nonlocal checked
        col = C0 + l + dx
        assert 2 <= col <= 68, (p, l, dx)
        got = lds.get(region(p) + slot(p) * PITCH + col)
        if i0 + l < n:
            assert got == (p, m, i + dx), (p, m, i, dx, got)
            checked += 1

    ring = {l: {p: [None] * 3 for p in range(NP) if vel(p)[1] != 0} for l in lanes}

    def ring_in(e, m):
        for l in lanes:
            i = min(i0 + l, n - 1)
            for p in ring[l]:
                if diag(p):
                    for d in (0,) + stencil(i, n):
                        xread(p, l, i, d, m)
                ring[l][p][e] = (p, min(m, n - 1))

    ws = window(k0, n)
    for s in range(3):
        commit(False, ws + s)
        ring_in(s, ws + s)
    commit(True, k0)
    visited = []

    def row(W, SL, k):
        nonlocal checked
        visited.append(k)
        assert W == (-1 if k == 0 else (1 if k == n - 1 else 0))
        more = k + 1 < k1
        assert SL == (more and window(k + 1, n) != window(k, n))
        oy = stencil(k, n)
        want = (k + oy[0], k, k + oy[1])
        for l in lanes:
            i = min(i0 + l, n - 1)
            for p in range(NP):
                if rest(p):
                    continue
                if vel(p)[1] != 0:
                    assert tuple(ring[l][p][e][1] for e in order[W]) == want, (p, k)
                    checked += 1
                else:
                    for d in (0,) + stencil(i, n):
                        xread(p, l, i, d, k)
        if more:
            commit(True, k + 1)
        if SL:
            commit(False, k + 2)
            for l in lanes:
                for p in ring[l]:
                    e = ring[l][p]
                    ring[l][p] = [e[1], e[2], None]
            ring_in(2, k + 2)

    k = k0
    if k == 0:
        row(-1, False, 0)
        k = 1
    kslide = min(k1 - 1, n - 2)
    while k < kslide:
        row(0, True, k)
        k += 1
    if k < min(k1, n - 1):
        row(0, False, k)
        k += 1
    if k == n - 1 and k < k1:
        row(1, False, k)
    assert visited == list(range(k0, k1))
    return checked


def single_tiles(n, remap):
    gx = (n + 63) // 64
    kc = min(n, rows_per_chunk(gx * n))
    gy = (n + kc - 1) // kc
    G = gx * gy
    seen = []
    for P in range(G):
        L = P
        if remap == 1:
            g8 = G >> 3
            if P < (g8 << 3):
                L = (P & 7) * g8 + (P >> 3)
        tx, ch = L % gx, L // gx
        if ch * kc < n:
            seen.append((tx, ch))
    assert sorted(seen) == sorted(itertools.product(range(gx), range(gy)))
    return kc, gx, gy


@pytest.mark.parametrize("n", [5, 17, 64, 65, 66, 127, 129, 130, 257])
def test_ymarch_indexing_model(n):
    kc, gx, gy = single_tiles(n, 1)
    single_tiles(n, 0)
    total = 0
    for tx in sorted({0, gx - 1}):
        for ch in sorted({0, 1, gy - 2, gy - 1} & set(range(gy))):
            k0 = ch * kc
            total += march(n, tx * 64, k0, min(n, k0 + kc))
    assert total > 0


def test_ymarch_model_catches_a_wrong_ring_order():
    flat = {0: (0, 1, 2), -1: (0, 1, 2), 1: (0, 1, 2)}
    with pytest.raises(AssertionError):
        march(17, 0, 0, 8, order=flat)
    with pytest.raises(AssertionError):
        march(17, 0, 8, 17, order=flat)
    assert march(17, 0, 0, 8) > 0


def test_batch_table_deals_every_chunk_once():
    """The batch's variant-1 table: (cavity, strip, chunk) entries, largest
    cavity first, dealt to the XCDs in contiguous runs and padded to a
    multiple of 8 with empty entries."""
    sizes = [65, 513, 129, 257, 65, 513, 257]
    order = sorted(range(len(sizes)), key=lambda c: -sizes[c])
    kc = rows_per_chunk(sum(((n + 63) // 64) * n for n in sizes))
    tiles = [(c, tx, ch) for c in order for ch in range((sizes[c] + kc - 1) // kc)
             for tx in range((sizes[c] + 63) // 64)]
    per = (len(tiles) + 7) // 8
    dealt = [None] * (8 * per)
    for L, t in enumerate(tiles):
        dealt[(L % per) * 8 + L // per] = t
    assert len(dealt) % 8 == 0
    assert sorted(t for t in dealt if t) == sorted(tiles)
    # XCD x = slot mod 8 runs a contiguous run of the list
    for x in range(8):
        run = [tiles.index(t) for t in dealt[x::8] if t]
        assert run == list(range(x * per, x * per + len(run)))
