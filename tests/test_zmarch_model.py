"""A CPU model of k_step_zm's indexing (csrc/lbm_step_zm.h) -- CPU only.

The kernel stages rows in LDS from per-thread plans, keeps the c_z != 0
populations in a three-entry register ring and peels the two wall planes
into their own copies of the plane body.  None of that can fault or go
stale without showing up somewhere in the GPU parity tests, but those only
see the end result.  This model restates the layout, the staging plans, the
workgroup-to-tile map and the march in Python and checks, for the launch
shapes the GPU tests use:
  * every LDS index written or read lies inside its population's region,
    and every global element loaded lies inside the population planes;
  * every value a lane reads (from LDS or from the ring) is the one the
    interpolation needs: the right population, plane, row and column;
  * the barriers order every LDS access: no LDS element is written and
    read between the same two barriers (where the waves of a workgroup run
    in any order);
  * the tile map hands every (x tile, y tile, z group) to exactly one
    workgroup, and the march visits every plane of a group once, with the
    wall copies at k = 0 and k = n-1 and the ring shift exactly when the z
    window moves.
"""
import itertools

import pytest

# D3Q19 velocities in the kernel's order (slots 0..18), then D3Q7 (19..25)
C = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
     (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1),
     (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)]
NP = 26


def vel(p):
    return C[p if p < 19 else p - 19]


def zpop(p):
    return vel(p)[2] != 0


def purez(p):
    return zpop(p) and vel(p)[0] == 0 and vel(p)[1] == 0


def staged(p):
    return p not in (0, 19) and not purez(p)


class Layout:
    def __init__(self, by):
        self.by = by
        self.hx = {p: int(vel(p)[0] != 0) for p in range(NP)}
        self.hy = {p: int(vel(p)[1] != 0) for p in range(NP)}
        self.P = {p: 72 if self.hx[p] else 64 for p in range(NP)}
        self.C0 = {p: 4 if self.hx[p] else 0 for p in range(NP)}
        self.H = {p: by + 2 * self.hy[p] for p in range(NP)}
        self.off, o = {}, 0
        for p in range(NP):
            self.off[p] = o
            o += self.P[p] * self.H[p] if staged(p) else 0
        self.total = o
        self.rows = {z: [(p, rr) for p in range(NP) if staged(p) and zpop(p) == z
                         for rr in range(self.H[p])] for z in (False, True)}

    def ninstr(self, z):
        return (len(self.rows[z]) + 4 * self.by - 1) // (4 * self.by)

    def region(self, p):
        return self.off[p], self.off[p] + self.P[p] * self.H[p]

    def at(self, p, l, r, dx, dy):
        row, col = r + self.hy[p] + dy, self.C0[p] + l + dx
        assert 0 <= row < self.H[p], (p, r, dy)
        assert (3 if self.hx[p] else 0) <= col <= (68 if self.hx[p] else 63), (p, l, dx)
        return self.off[p] + row * self.P[p] + col


def stencil(i, n):
    """Offsets of stencil points 0 and 2 (point 1 is the node)."""
    if i == 0:
        return 1, 2
    if i == n - 1:
        return -1, -2
    return -1, 1


def window(k, n):
    return min(max(k - 1, 0), n - 3)


def planes_per_group(n, by):
    cols = ((n + 63) // 64) * ((n + by - 1) // by)
    nzc = max(1, min((3072 + cols - 1) // cols, max(1, n // 4)))
    return (n + nzc - 1) // nzc


def region(n, by):
    """zm_region: the z-march covers i < xz, j < yz; k_step_rim the rest."""
    xz = n - 1 if (n % 64 == 1 and n > 64) else n
    yz = n - 1 if (n - 1) % by == 0 else n
    return xz, yz


def rim_nodes(n, xz, yz):
    """k_step_rim's thread -> node map, replayed."""
    cols, per = (n - xz) * n, (n - xz) * n + (n - yz) * xz
    out = []
    for t in range(per * n):
        k, r = divmod(t, per)
        if r < cols:
            i, j = xz + r % (n - xz), r // (n - xz)
        else:
            r -= cols
            i, j = r % xz, yz + r // xz
        out.append((i, j, k))
    return out


def tile_map(n, by, remap):
    """Workgroup -> (tx, ty, tz) of zm_tile, with the launcher's grid size."""
    kc = planes_per_group(n, by)
    nzc = (n + kc - 1) // kc
    xz, yz = region(n, by)
    gx, gy = (xz + 63) // 64, (yz + by - 1) // by
    remap = 0 if (remap == 2 and gy < 8) else remap
    ntiles = 8 * gx * ((gy + 7) // 8) * nzc if remap == 2 else gx * gy * nzc
    out = {}
    for P in range(ntiles):
        if remap == 2:
            x, q = P & 7, P >> 3
            b0 = x * gy // 8
            band = (x + 1) * gy // 8 - b0
            tx, rr = q % gx, q // gx
            ty, tz = b0 + rr % band, rr // band
        else:
            L = P
            if remap == 1:
                g8 = ntiles >> 3
                if P < (g8 << 3):
                    L = (P & 7) * g8 + (P >> 3)
            tx, rr = L % gx, L // gx
            ty, tz = rr % gy, rr // gy
        if tz < nzc:
            assert (tx, ty, tz) not in out.values()
            out[P] = (tx, ty, tz)
    assert sorted(out.values()) == sorted(itertools.product(range(gx), range(gy), range(nzc)))
    return kc, gx, gy, nzc


KERNEL_ORDER = {0: (0, 1, 2), -1: (1, 0, 2), 1: (1, 2, 0)}


def march(n, by, tx, ty, tz, kc, order=KERNEL_ORDER, buffers=2, barriers=1):
    """Replays one workgroup's march; returns the number of checked reads.
    order: ring entries read as stencil points (o0, node, o2) by the body
    copy for k = 0 (-1), inside (0) and k = n-1 (+1).  buffers: staging
    buffers (plane k reads buffer (k - k0) % buffers); barriers: per plane
    (2: after the plane's reads and after its writes; 1: after its writes
    only).  The kernel has two buffers and one barrier.
    Reads are checked for the lanes at both tile edges, around the middle
    and around the last live lane of a partial tile; the staging covers
    every lane.  Every LDS access is tagged with the number of barriers
    before it, and an element written and read under the same count is a
    race between waves."""
    L = Layout(by)
    xz, yz = region(n, by)
    pitch = (n + 31) // 32 * 32                  # padded population row
    pstride = (pitch * n * n + 63) // 64 * 64    # population plane stride
    i0, j0 = tx * 64, ty * by
    k0, k1 = tz * kc, min(n, (tz + 1) * kc)
    assert xz - i0 >= 1 and yz - j0 >= 1
    lds = {}
    epoch = 0
    wrote, read, races = {}, {}, []   # (buffer, LDS index) -> last barrier count
    checked = 0
    # halo ownership: every halo value of a phase has a thread
    for z in (False, True):
        nh = sum(2 * L.H[p] for p in range(NP) if staged(p) and zpop(p) == z and L.hx[p])
        assert nh <= 64 * by

    def barrier():
        nonlocal epoch
        epoch += 1

    def write(b, idx, tag):
        if read.get((b, idx)) == epoch:
            races.append((b, idx, epoch))
        lds[(b, idx)] = tag
        wrote[(b, idx)] = epoch

    def load(b, idx):
        if wrote.get((b, idx)) == epoch:
            races.append((b, idx, epoch))
        read[(b, idx)] = epoch
        return lds.get((b, idx))

    def stage(z, kk, b):
        kk = min(kk, n - 1)
        rl, covered = L.rows[z], set()
        for w, l in itertools.product(range(by), range(64)):
            g, c4 = l >> 4, 4 * (l & 15)
            ii = min(i0 + c4, pitch - 4)
            for q in range(L.ninstr(z)):
                R = min(4 * (q * by + w) + g, len(rl) - 1)
                covered.add(R)
                p, rr = rl[R]
                jj = min(max(j0 - L.hy[p] + rr, 0), n - 1)
                base = L.off[p] + rr * L.P[p] + L.C0[p] + (ii - i0)
                lo, hi = L.region(p)
                g_idx = p * pstride + kk * pitch * n + jj * pitch + ii   # padded rows
                assert lo <= base and base + 3 < hi and 0 <= g_idx and g_idx + 3 < NP * pstride
                assert g_idx % 4 == 0 and ii + 3 < pitch     # aligned, inside the padded row
                for e in range(4):
                    # values past the grid's edge are padding
                    write(b, base + e, (p, kk, jj, ii + e) if ii + e < n else None)
            h = l + 64 * w
            for p in range(NP):
                if not (staged(p) and zpop(p) == z and L.hx[p]):
                    continue
                cnt = 2 * L.H[p]
                if 0 <= h < cnt:
                    rr, side = h >> 1, h & 1
                    jj, ih = j0 - L.hy[p] + rr, (i0 + 64 if side else i0 - 1)
                    if 0 <= jj < n and 0 <= ih < n:
                        idx = L.off[p] + rr * L.P[p] + (68 if side else 3)
                        write(b, idx, (p, kk, jj, ih))
                h -= cnt
        assert covered == set(range(len(rl)))

    def inplane(p, l, r, i, j, m, b):
        """The in-plane stage of z-population p at plane m: its LDS reads."""
        nonlocal checked
        cx, cy, _ = vel(p)
        ax_off = stencil(i, n) if cx else stencil(j, n)
        for d in (0,) + ax_off:
            dx, dy = (d, 0) if cx else (0, d)
            got = load(b, L.at(p, l, r, dx, dy))
            assert got == (p, m, j + dy, i + dx), (p, m, i, j, dx, dy, got)
            checked += 1

    last = min(63, xz - 1 - i0)
    cols = {0, 1, 2, 3, 30, 31, 32, 33, 60, 61, 62, 63, last - 1, last, min(63, last + 1)}
    lanes = [(l, r) for r in range(by) for l in sorted(c for c in cols if 0 <= c < 64)]
    ring = {}   # (lane, p) -> [tag of entry 0, 1, 2]

    def ring_in(entry, m, b):
        for l, r in lanes:
            live = i0 + l < xz and j0 + r < yz
            i, j = min(i0 + l, xz - 1), min(j0 + r, yz - 1)
            for p in range(NP):
                if not zpop(p):
                    continue
                if purez(p):
                    tag = (p, min(m, n - 1))
                else:
                    if live:
                        inplane(p, l, r, i, j, m, b)
                    tag = (p, m)
                ring.setdefault((l, r), {}).setdefault(p, [None] * 3)[entry] = tag

    ws = window(k0, n)
    for s in range(3):
        stage(True, ws + s, 0)
        barrier()
        ring_in(s, ws + s, 0)
        barrier()
    stage(False, k0, 0)
    barrier()

    visited = []

    def plane(W, SL, k):
        nonlocal checked
        visited.append((W, SL, k))
        assert W == (-1 if k == 0 else (1 if k == n - 1 else 0))
        more = k + 1 < k1
        assert SL == (more and window(k + 1, n) != window(k, n))
        E = order[W]
        b = (k - k0) % buffers
        zo = stencil(k, n)
        want_planes = (k + zo[0], k, k + zo[1])
        for l, r in lanes:
            live = i0 + l < xz and j0 + r < yz
            i, j = min(i0 + l, xz - 1), min(j0 + r, yz - 1)
            xo, yo = stencil(i, n), stencil(j, n)
            for p in range(NP):
                if p in (0, 19):
                    continue
                if zpop(p):
                    ent = ring[(l, r)][p]
                    assert tuple(ent[e][1] for e in E) == want_planes, (p, k, ent)
                    checked += 1
                    continue
                cx, cy, _ = vel(p)
                for dx, dy in itertools.product((0,) + (xo if cx else ()), (0,) + (yo if cy else ())):
                    idx = L.at(p, l, r, dx, dy)
                    if live:
                        assert load(b, idx) == (p, k, j + dy, i + dx), (p, k, i, j, dx, dy)
                        checked += 1
        nb = (b + 1) % buffers
        if barriers == 2:
            barrier()   # every read of this plane's staging is done
        if more:
            stage(False, k + 1, nb)
        if SL:
            stage(True, k + 2, nb)
        barrier()
        if SL:
            for lr in ring:
                for p in ring[lr]:
                    e = ring[lr][p]
                    ring[lr][p] = [e[1], e[2], None]
            ring_in(2, k + 2, nb)

    # the kernel's loop structure
    k = k0
    if k == 0:
        plane(-1, False, 0)
        k = 1
    kslide = min(k1 - 1, n - 2)
    while k < kslide:
        plane(0, True, k)
        k += 1
    if k < min(k1, n - 1):
        plane(0, False, k)
        k += 1
    if k == n - 1 and k < k1:
        plane(1, False, k)
    assert [v[2] for v in visited] == list(range(k0, k1))
    assert not races, f"LDS race: {races[:3]}"
    return checked


SHAPES = [(12, 2, 0), (20, 4, 0), (68, 4, 1), (72, 2, 2), (96, 4, 0), (132, 4, 2), (24, 8, 0),
          (136, 8, 2), (24, 6, 0), (68, 6, 2), (132, 6, 2), (72, 4, 2), (16, 2, 0), (36, 4, 0),
          (40, 6, 0),
          # any n: rows not 16-byte aligned, partial quads at the edge, the far
          # wall in a tile's first column (n % 64 == 1) or row ((n-1) % by == 0),
          # tiles of 2-3 columns (n % 64 == 2, 3)
          (6, 2, 0), (7, 4, 0), (9, 4, 0), (13, 6, 0), (17, 4, 0), (21, 4, 1), (33, 8, 0),
          (37, 2, 0), (65, 4, 2), (66, 4, 0), (67, 2, 2), (70, 6, 0), (129, 4, 2), (131, 8, 2),
          (257, 4, 2), (257, 6, 2)]


@pytest.mark.parametrize("n,by,remap", SHAPES)
def test_zmarch_indexing_model(n, by, remap):
    kc, gx, gy, nzc = tile_map(n, by, remap)
    # the first and last tiles on each axis, and their neighbours on the
    # smaller grids (the largest ones take minutes otherwise)
    edge = lambda g: sorted(({0, 1, g - 2, g - 1} if n < 200 else {0, g - 1}) & set(range(g)))
    total = 0
    for tx, ty, tz in itertools.product(edge(gx), edge(gy), edge(nzc)):
        total += march(n, by, tx, ty, tz, kc)
    assert total > 0


def test_zmarch_model_catches_a_wrong_ring_order():
    """The model is not vacuous: a body that read the wall planes' z points
    in window order (no peeled copies) is caught at k = 0 and at k = n-1."""
    n, by = 12, 2
    kc, gx, gy, nzc = tile_map(n, by, 0)
    flat = {0: (0, 1, 2), -1: (0, 1, 2), 1: (0, 1, 2)}
    with pytest.raises(AssertionError):
        march(n, by, 0, 0, 0, kc, order=flat)
    with pytest.raises(AssertionError):
        march(n, by, 0, 0, nzc - 1, kc, order=flat)
    assert march(n, by, 0, 0, 0, kc) > 0


@pytest.mark.parametrize("n,by", [(12, 2), (72, 4)])
def test_zmarch_model_catches_a_missing_barrier(n, by):
    """With one staging buffer, one barrier per plane lets the next plane's
    rows land while slower waves still read the current ones: caught.  One
    buffer needs a second barrier, after the plane's reads."""
    kc, gx, gy, nzc = tile_map(n, by, 0)
    with pytest.raises(AssertionError, match="LDS race"):
        march(n, by, 0, 0, 0, kc, buffers=1, barriers=1)
    assert march(n, by, 0, 0, 0, kc, buffers=1, barriers=2) > 0


def test_tile_map_covers_every_tile_once_on_the_bench_grid():
    for by, remap in ((4, 2), (6, 2), (8, 2), (4, 0), (4, 1)):
        tile_map(256, by, remap)


@pytest.mark.parametrize("n,by", [(n, by) for n, by, _ in SHAPES])
def test_zmarch_and_rim_cover_every_node_once(n, by):
    """The live lanes of the z-march tiles and the rim kernel's threads
    together update every node of every plane exactly once; the rim is
    empty when n % 4 == 0 (even by)."""
    xz, yz = region(n, by)
    gx, gy = (xz + 63) // 64, (yz + by - 1) // by
    inner = {(i, j) for i in range(min(gx * 64, xz)) for j in range(min(gy * by, yz))}
    assert inner == {(i, j) for i in range(xz) for j in range(yz)}
    rim = rim_nodes(n, xz, yz)
    assert len(rim) == len(set(rim))
    rim_plane = {(i, j) for i, j, k in rim if k == 0}
    assert {(i, j, k) for i, j in rim_plane for k in range(n)} == set(rim)
    assert not (rim_plane & inner)
    assert rim_plane | inner == {(i, j) for i in range(n) for j in range(n)}
    if n % 4 == 0:
        assert not rim

