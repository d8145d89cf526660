// lbm_step_zm.h -- LDS-staged, z-marching variant of the 3D FP32 step
// (launch variant 1).
//
// The plain kernel (k_step) reads 156 values per node through the L1; the
// stencil rows it shares with other workgroups (the y halo and the planes
// k +- 1) are L1 misses, and the L1's miss handling, not HBM, bounds it.
// Here a workgroup owns a 64 x BY column of nodes over a range of planes and
// marches it in +z.  Each plane's post-collision populations are staged in
// LDS once, one slot per population:
//   * populations with c_z = 0 (except the rest ones): plane k, with a one-
//     node halo in x if c_x != 0 and in y if c_y != 0, interpolated from
//     LDS when node k is computed;
//   * populations with c_z != 0 and an in-plane component: the newest plane
//     of the z window, with their x or y halo.  As a plane arrives, each
//     lane interpolates it in-plane at its own (i, j) and keeps the result
//     (and its own value) in a three-entry register ring; node k then
//     interpolates along z from the ring.  The tensor-product weights are
//     those of k_step, applied in-plane first;
//   * populations with c = (0, 0, +-1) have no in-plane neighbours: each
//     lane loads its own value of the arriving plane into the ring;
//   * the two rest populations are loaded by each lane, one plane ahead.
// The ring holds the planes of the window [ws, ws + 2], ws = clamp(k - 1,
// 0, n - 3), in order: a step to the next plane shifts it and interpolates
// the arriving plane, except at the walls, where the window does not move.
// The wall planes k = 0 and k = n-1 (one-sided z stencils) run their own
// copies of the plane body, which take the stencil points from the ring in
// their order; the hot copy takes them in window order and always shifts.
// Rows are staged four at a time: a lane loads four consecutive x values
// (one 16-byte load), so a wave moves four 64-wide rows per instruction, and
// the rows of the next planes are fetched into registers while the current
// plane is computed.  The results agree with k_step to FP32 rounding (the
// order of the tensor-product stages and the fused multiply-adds differ).
//
// All waves run one copy of the loop: each thread's staging offsets are
// computed once (a plan) and the loads and LDS writes of a plane are
// straight-line code.
//
// The staging is double-buffered: plane k reads one buffer and writes the
// rows of plane k+1 (and of the plane entering the z window) into the
// other, once its node is collided and before its stores are issued; a wait
// for every load of the plane follows, then the stores and one barrier.
// The stores are issued inside a branch (live lanes only), and after such a
// branch the compiler's count of outstanding memory operations no longer
// tells the stores from the loads: a load consumed after the stores would
// wait for the stores to complete too.  Consuming every load before the
// stores leaves the stores in flight while the next plane computes.
//
// Any n from 6 up, as long as the 26 population planes stay within 32-bit
// element offsets (checked by the host), FP32 storage, 3D, bx = 64,
// by = BY.  The population rows are padded to a pitch of a multiple of 32
// elements, so every row starts on a 128-byte line and a lane loads four
// of its x values with one aligned 16-byte load; a quad clamped at the
// padded row's end lands in LDS at its true columns.  The march
// covers i < xz, j < yz: the far-wall node's one-sided stencil reaches two
// nodes inward, beyond the one-node halo, when it would sit in the first
// column of a tile (n % 64 == 1), so then the last column is left out
// (xz = n - 1); likewise the last row when it would be a tile's first row
// ((n - 1) % BY == 0, yz = n - 1).  The first rim_wgs
// workgroups of the launch update the nodes left out with the single-node
// code of k_step (zm_rim_nodes), several per thread.
#pragma once

#include "lbm_step.h"

namespace lbm {

// A native 4-float vector (HIP's float4 is a union type that the compiler
// does not keep in registers when it sits in an array).
typedef float zm_f4 __attribute__((ext_vector_type(4)));
typedef float zm_f2 __attribute__((ext_vector_type(2)));

// Slots p = 0..25: flow populations 0..18, then thermal 0..6 (whose
// velocities are the first seven flow velocities).  The population plane
// of slot p is plane p of the buffer.  A staged region is H rows of pitch
// P: x columns i0 .. i0+63 at C0 .. C0+63, and for populations with an x
// halo, column i0-1 at 3 and column i0+64 at 68 (so rows stay 16-byte
// aligned).
template <int BY> struct ZmLayout {
    static constexpr int NP = 26;
    static constexpr int vel(int p) { return p < 19 ? p : p - 19; }
    static constexpr int cx(int p) { return Flow<3>::c[vel(p)][0]; }
    static constexpr int cy(int p) { return Flow<3>::c[vel(p)][1]; }
    static constexpr int cz(int p) { return Flow<3>::c[vel(p)][2]; }
    static constexpr bool zpop(int p) { return cz(p) != 0; }
    static constexpr bool purez(int p) { return zpop(p) && cx(p) == 0 && cy(p) == 0; }
    static constexpr bool staged(int p) { return p != 0 && p != 19 && !purez(p); }
    static constexpr int hx(int p) { return cx(p) != 0 ? 1 : 0; }
    static constexpr int hy(int p) { return cy(p) != 0 ? 1 : 0; }
    static constexpr int P(int p) { return hx(p) ? 72 : 64; }
    static constexpr int C0(int p) { return hx(p) ? 4 : 0; }
    static constexpr int H(int p) { return BY + 2 * hy(p); }
    static constexpr int psize(int p) { return staged(p) ? P(p) * H(p) : 0; }
    static constexpr int off(int p) {
        int o = 0;
        for (int q = 0; q < p; q++) o += psize(q);
        return o;
    }
    static constexpr int total() { return off(NP); }
    // register ring indices: over the c_z != 0 populations, over the staged
    // ones among them, and over the (0, 0, +-1) ones
    static constexpr int zi(int p) {
        int c = 0;
        for (int q = 0; q < p; q++) c += zpop(q) ? 1 : 0;
        return c;
    }
    static constexpr int oi(int p) {
        int c = 0;
        for (int q = 0; q < p; q++) c += (zpop(q) && staged(q)) ? 1 : 0;
        return c;
    }
    static constexpr int pzi(int p) {
        int c = 0;
        for (int q = 0; q < p; q++) c += purez(q) ? 1 : 0;
        return c;
    }
    static constexpr int nz() { return zi(NP); }
    static constexpr int nzo() { return oi(NP); }
    static constexpr int npz() { return pzi(NP); }
    // Rows of a phase (Z: the staged c_z != 0 populations of the arriving
    // plane; !Z: the in-plane populations of plane k), in slot order.
    // Staging instruction q of wave w gives lane group g (16 lanes of 4 x
    // values each) row R = 4 (q BY + w) + g, clamped to the last row (a
    // repeated row writes the same values).
    static constexpr int nrows(bool z) {
        int c = 0;
        for (int p = 0; p < NP; p++)
            if (staged(p) && zpop(p) == z) c += H(p);
        return c;
    }
    static constexpr int ninstr(bool z) { return (nrows(z) + 4 * BY - 1) / (4 * BY); }
};

// A thread's staging plan for one phase, fixed for the whole march: for
// each staging instruction, the element offset of its four values at plane
// 0 and their LDS index; the same for its halo value, if any (thread t owns
// halo value t: population, region row, side).
template <int BY, bool Z> struct ZmPlan {
    using L = ZmLayout<BY>;
    uint32_t goff[L::ninstr(Z)];
    uint32_t lds[L::ninstr(Z)];
    uint32_t hoff;
    uint32_t hlds;   // ~0u: no halo value
};

template <int BY, bool Z>
__device__ __forceinline__ ZmPlan<BY, Z> zm_plan(const StepArgs<float> &a, int i0, int j0,
                                                 int w, int l) {
    using L = ZmLayout<BY>;
    const int n = a.n;
    const uint32_t plane = (uint32_t)a.plane, pitch = (uint32_t)a.pitch;
    const int g = l >> 4, c4 = 4 * (l & 15);
    // a group past the padded row loads its last four columns (padding, or
    // nodes i >= n that no live lane reads), into their own LDS columns;
    // i0 and the pitch are multiples of 32, so the load stays aligned
    const int ii = min(i0 + c4, (int)pitch - 4);
    ZmPlan<BY, Z> pl;
    static_for<0, L::ninstr(Z)>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        const int R = min(4 * (q * BY + w) + g, L::nrows(Z) - 1);
        uint32_t go = 0, ld = 0;
        int cum = 0;
        static_for<0, L::NP>([&](auto Pp) {
            constexpr int p = decltype(Pp)::value;
            if constexpr (L::staged(p) && L::zpop(p) == Z) {
                if (R >= cum && R < cum + L::H(p)) {
                    const int rr = R - cum;
                    const int jj = min(max(j0 - L::hy(p) + rr, 0), n - 1);
                    go = (uint32_t)p * plane + (uint32_t)jj * pitch + ii;
                    ld = (uint32_t)(L::off(p) + rr * L::P(p) + L::C0(p) + (ii - i0));
                }
                cum += L::H(p);
            }
        });
        pl.goff[q] = go;
        pl.lds[q] = ld;
    });
    pl.hoff = 0;
    pl.hlds = ~0u;
    int h = l + 64 * w;
    static_for<0, L::NP>([&](auto Pp) {
        constexpr int p = decltype(Pp)::value;
        if constexpr (L::staged(p) && L::zpop(p) == Z && L::hx(p) == 1) {
            constexpr int cnt = 2 * L::H(p);
            const bool mine = h >= 0 && h < cnt;
            const int rr = h >> 1, side = h & 1;
            const int jj = j0 - L::hy(p) + rr, ih = side ? i0 + 64 : i0 - 1;
            const bool inside = jj >= 0 && jj < n && ih >= 0 && ih < n;
            if (mine && inside) {
                pl.hoff = (uint32_t)p * plane + (uint32_t)jj * pitch + ih;
                pl.hlds = (uint32_t)(L::off(p) + rr * L::P(p) + (side ? 68 : 3));
            }
            h -= cnt;
        }
    });
    return pl;
}

// Registers holding a thread's staged values of one phase while in flight.
template <int BY, bool Z> struct ZmStage {
    using L = ZmLayout<BY>;
    zm_f4 quad[L::ninstr(Z)];
    float halo;
};

// Loads the thread's values of phase Z at plane kk (clamped to the grid)
// into registers.  Straight-line code: every load is issued, so the waits
// before the compute count them as in flight.
template <int BY, bool Z>
__device__ __forceinline__ void zm_fetch(const StepArgs<float> &a, int kk,
                                         const ZmPlan<BY, Z> &pl, ZmStage<BY, Z> &s) {
    using L = ZmLayout<BY>;
    kk = min(kk, a.n - 1);
    const size_t po = (size_t)kk * (size_t)a.pitch * (size_t)a.n;
    static_for<0, L::ninstr(Z)>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        s.quad[q] = *reinterpret_cast<const zm_f4 *>(a.pin + (po + pl.goff[q]));
    });
    // threads without a halo value load the plane origin (unused)
    s.halo = a.pin[po + pl.hoff];
}

// The (0, 0, +-1) populations of the lane's own column at plane kk (clamped).
template <int BY>
__device__ __forceinline__ void zm_fetch_purez(const StepArgs<float> &a, uint32_t col, int kk,
                                               float *v) {
    using L = ZmLayout<BY>;
    kk = min(kk, a.n - 1);
    const uint32_t node = col + (uint32_t)kk * (uint32_t)a.pitch * (uint32_t)a.n;
    static_for<0, L::NP>([&](auto Pp) {
        constexpr int p = decltype(Pp)::value;
        if constexpr (L::purez(p)) v[L::pzi(p)] = ldp(a.pin + p * a.plane, node);
    });
}

// Writes fetched values into their LDS region as two 8-byte writes per
// group of four (fewer bank conflicts than one 16-byte write).
template <int BY, bool Z>
__device__ __forceinline__ void zm_commit(float *lds, const ZmPlan<BY, Z> &pl,
                                          const ZmStage<BY, Z> &s) {
    using L = ZmLayout<BY>;
    static_for<0, L::ninstr(Z)>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        *reinterpret_cast<zm_f2 *>(&lds[pl.lds[q]]) = s.quad[q].xy;
        *reinterpret_cast<zm_f2 *>(&lds[pl.lds[q] + 2]) = s.quad[q].zw;
    });
    if (pl.hlds != ~0u) lds[pl.hlds] = s.halo;
}

// Value of slot p at x, y offsets (dx, dy) from the lane's node.
template <int BY, int p>
__device__ __forceinline__ float zm_at(const float *lds, int l, int r, int dx, int dy) {
    using L = ZmLayout<BY>;
    return lds[L::off(p) + (r + L::hy(p) + dy) * L::P(p) + L::C0(p) + l + dx];
}

// Value of slot p at stencil point S1 along axis A1 and S2 along axis A2
// (in-plane axes; point 1 is the node itself; A < 0 means no axis).
template <int BY, int p, int A1, int S1, int A2, int S2>
__device__ __forceinline__ float zm_pt(const float *lds, int l, int r,
                                       const AxisStencil<float> *st) {
    int dx = 0, dy = 0;
    if constexpr (A1 == 0 && S1 != 1) dx = S1 == 0 ? st[0].o0 : st[0].o2;
    if constexpr (A2 == 0 && S2 != 1) dx = S2 == 0 ? st[0].o0 : st[0].o2;
    if constexpr (A1 == 1 && S1 != 1) dy = S1 == 0 ? st[1].o0 : st[1].o2;
    if constexpr (A2 == 1 && S2 != 1) dy = S2 == 0 ? st[1].o0 : st[1].o2;
    return zm_at<BY, p>(lds, l, r, dx, dy);
}

// interp_pop of a c_z = 0 population with its operands from LDS: the same
// expressions in the same order.
template <int BY, int p>
__device__ __forceinline__ float zm_interp_plane(const float *lds, int l, int r, float fc,
                                                 const AxisStencil<float> *st) {
    using L = ZmLayout<BY>;
    constexpr int c[2] = {L::cx(p), L::cy(p)};
    if constexpr (c[0] == 0 || c[1] == 0) {
        constexpr int ax = c[0] ? 0 : 1;
        const float *w = c[ax] > 0 ? st[ax].wp : st[ax].wm;
        return lerp3<false>(w, zm_pt<BY, p, ax, 0, -1, 1>(lds, l, r, st), fc,
                            zm_pt<BY, p, ax, 2, -1, 1>(lds, l, r, st));
    } else {
        // outer stage along x, inner along y
        const float *wy = c[1] > 0 ? st[1].wp : st[1].wm;
        const float *wx = c[0] > 0 ? st[0].wp : st[0].wm;
        const float c0 = lerp3<false>(wy, zm_pt<BY, p, 0, 0, 1, 0>(lds, l, r, st),
                                      zm_pt<BY, p, 0, 0, 1, 1>(lds, l, r, st),
                                      zm_pt<BY, p, 0, 0, 1, 2>(lds, l, r, st));
        const float c1 = lerp3<false>(wy, zm_pt<BY, p, 0, 1, 1, 0>(lds, l, r, st), fc,
                                      zm_pt<BY, p, 0, 1, 1, 2>(lds, l, r, st));
        const float c2 = lerp3<false>(wy, zm_pt<BY, p, 0, 2, 1, 0>(lds, l, r, st),
                                      zm_pt<BY, p, 0, 2, 1, 1>(lds, l, r, st),
                                      zm_pt<BY, p, 0, 2, 1, 2>(lds, l, r, st));
        return lerp3<false>(wx, c0, c1, c2);
    }
}

// Puts the arriving plane into ring entry S: for each c_z != 0 population
// its in-plane interpolation at the lane's (i, j) and, for the staged ones,
// its own value (the (0, 0, +-1) values come from pz).
template <int BY, int S, int NZ, int NZO>
__device__ __forceinline__ void zm_ring_in(const float *lds, int l, int r,
                                           const AxisStencil<float> *st, const float *pz,
                                           float (&zr)[NZ][3], float (&zo)[NZO][3]) {
    using L = ZmLayout<BY>;
    static_for<0, L::NP>([&](auto Pp) {
        constexpr int p = decltype(Pp)::value;
        if constexpr (L::purez(p)) {
            zr[L::zi(p)][S] = pz[L::pzi(p)];
        } else if constexpr (L::zpop(p)) {
            constexpr int ax = L::cx(p) ? 0 : 1;
            constexpr int cc = ax == 0 ? L::cx(p) : L::cy(p);
            const float *w = cc > 0 ? st[ax].wp : st[ax].wm;
            const float own = zm_at<BY, p>(lds, l, r, 0, 0);
            zo[L::oi(p)][S] = own;
            zr[L::zi(p)][S] = lerp3<false>(w, zm_pt<BY, p, ax, 0, -1, 1>(lds, l, r, st), own,
                                           zm_pt<BY, p, ax, 2, -1, 1>(lds, l, r, st));
        }
    });
}

// Ring move to the next plane: entry 0 <- 1, 1 <- 2 (entry 2 is then
// refilled from the arriving plane).
template <int NZ, int NZO>
__device__ __forceinline__ void zm_ring_shift(float (&zr)[NZ][3], float (&zo)[NZO][3]) {
    static_for<0, NZ>([&](auto Z) {
        constexpr int z = decltype(Z)::value;
        zr[z][0] = zr[z][1];
        zr[z][1] = zr[z][2];
    });
    static_for<0, NZO>([&](auto Z) {
        constexpr int z = decltype(Z)::value;
        zo[z][0] = zo[z][1];
        zo[z][1] = zo[z][2];
    });
}

__device__ __forceinline__ int zm_window(int k, int n) { return min(max(k - 1, 0), n - 3); }

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt not waited on): every vector memory
// operation issued so far has completed.
__device__ __forceinline__ void zm_wait_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// The z stencil of plane k (the same for every lane), read through the
// constant address space so that it is fetched with scalar loads: a vector
// load here would be counted behind the previous plane's stores and wait for
// them.  The tables are never written by a step.
__device__ __forceinline__ AxisStencil<float> zm_stencil_z(const StepArgs<float> &a, int k) {
    typedef const __attribute__((address_space(4))) int *coff_t;
    typedef const __attribute__((address_space(4))) float *cwt_t;
    const coff_t off = (coff_t)a.off + 2 * k;   // (o0, o2) pairs
    const cwt_t wt = (cwt_t)a.wt + 6 * k;
    AxisStencil<float> s;
    s.o0 = off[0];
    s.o2 = off[1];
#pragma unroll
    for (int b = 0; b < 3; b++) {
        s.wp[b] = wt[b];
        s.wm[b] = wt[3 + b];
    }
    return s;
}

// Tile (tx, ty) and z group tz of hardware workgroup P; false for the
// workgroups past the last z group.  remap 2 (gy >= 8): workgroups are
// dispatched round-robin over the 8 XCDs, and XCD x = P % 8 owns the tile
// rows [x gy / 8, (x + 1) gy / 8), swept x fastest, then y, then z, so the
// x and y neighbour tiles that re-read a tile's rows run on the same XCD at
// the same time and find them in its L2.  The grid gives every XCD the
// largest share; the surplus workgroups of the smaller bands return at once.
// Other remaps: the x-fastest order of logical_tile.
__device__ __forceinline__ bool zm_tile(int P, const StepArgs<float> &a, int nzc, int &tx,
                                        int &ty, int &tz) {
    if (a.remap == 2) {
        const int x = P & 7, q = P >> 3;
        const int b0 = x * a.gy / 8, band = (x + 1) * a.gy / 8 - b0;
        tx = q % a.gx;
        const int rr = q / a.gx;
        ty = b0 + rr % band;
        tz = rr / band;
    } else {
        const int L = logical_tile(P, a.ntiles, a.remap);
        tx = L % a.gx;
        const int rr = L / a.gx;
        ty = rr % a.gy;
        tz = rr / a.gy;
    }
    return tz < nzc;
}

// The nodes the march leaves out (i >= xz or j >= yz, every k), thread t
// of `threads` taking nodes t, t + threads, ...: per plane, the columns
// i >= xz of every row, then the rows j >= yz of the columns i < xz.
template <bool MACRO, bool NT>
__device__ __forceinline__ void zm_rim_nodes(const StepArgs<float> &a, int xz, int yz, long t,
                                             long threads) {
    const int n = a.n;
    const long cols = (long)(n - xz) * n, per = cols + (long)(n - yz) * xz;
    for (; t < per * n; t += threads) {
        const int k = (int)(t / per);
        long r = t - (long)k * per;
        int i, j;
        if (r < cols) {
            i = xz + (int)(r % (n - xz));
            j = (int)(r / (n - xz));
        } else {
            r -= cols;
            i = (int)(r % xz);
            j = yz + (int)(r / xz);
        }
        step_node<3, float, float, true, MACRO, NT>(a, i, j, k);
    }
}

template <int BY, bool MACRO, bool NT>
__global__ __launch_bounds__(64 * BY) void k_step_zm(StepArgs<float> a, int kc, int xz, int yz,
                                                     int rim_wgs) {
    using L = ZmLayout<BY>;
    using FL = Flow<3>;
    using HT = Heat<3>;
    constexpr int QF = FL::Q, QG = HT::Q;
    constexpr int NZ = L::nz(), NZO = L::nzo();
    // two staging buffers: plane k reads buffer (k - k0) & 1 and stages
    // plane k + 1 (and the plane entering the z window) into the other
    __shared__ zm_f4 lds4[2 * L::total() / 4];
    float *lds = reinterpret_cast<float *>(lds4);
    const int n = a.n;
    if ((int)blockIdx.x < rim_wgs) {
        // the nodes outside the marched region, with the single-node code;
        // these few workgroups come first and run beside the march
        zm_rim_nodes<MACRO, NT>(a, xz, yz,
                                (long)blockIdx.x * (64 * BY) + threadIdx.x + 64 * threadIdx.y,
                                (long)rim_wgs * 64 * BY);
        return;
    }
    int tx, ty, tc;
    if (!zm_tile((int)blockIdx.x - rim_wgs, a, (n + kc - 1) / kc, tx, ty, tc)) return;
    const int i0 = tx * 64, j0 = ty * BY;
    const int k0 = tc * kc, k1 = min(n, k0 + kc);
    const int l = threadIdx.x;
    const int r = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one row per wave
    const bool live = i0 + l < xz && j0 + r < yz;
    // lanes past the marched region stage and compute like its last node
    // and store nothing (every lane must reach the barriers)
    const int i = min(i0 + l, xz - 1), j = min(j0 + r, yz - 1);
    // population indices (rows padded to the pitch) and field indices
    const uint32_t sy = a.pitch, sz2 = (uint32_t)a.pitch * n;
    const uint32_t col = i + sy * j;
    const uint32_t fcol = i + (uint32_t)n * j, fsz = (uint32_t)n * n;
    AxisStencil<float> st[3];
    st[0] = load_stencil(a, i);
    st[1] = load_stencil(a, j);
    const ZmPlan<BY, false> pin_ = zm_plan<BY, false>(a, i0, j0, r, l);
    const ZmPlan<BY, true> pz = zm_plan<BY, true>(a, i0, j0, r, l);

    // z ring of plane k0: the planes of its window, in order
    float zr[NZ][3], zo[NZO][3];
    {
        const int ws = zm_window(k0, n);
        static_for<0, 3>([&](auto S) {
            constexpr int s = decltype(S)::value;
            ZmStage<BY, true> sz;
            float pzv[L::npz()];
            zm_fetch<BY, true>(a, ws + s, pz, sz);
            zm_fetch_purez<BY>(a, col, ws + s, pzv);
            zm_commit<BY, true>(lds, pz, sz);
            lds_barrier();
            zm_ring_in<BY, s>(lds, l, r, st, pzv, zr, zo);
            lds_barrier();
        });
        ZmStage<BY, false> sp;
        zm_fetch<BY, false>(a, k0, pin_, sp);
        zm_commit<BY, false>(lds, pin_, sp);
    }
    lds_barrier();

    // constant along z (the heated axis is y)
    const HeatShift<float> hs = load_heat_shift<3, true>(a, {i, j, 0});
    // the rest populations of the next plane, loaded one plane ahead
    float rest_f = ldp(a.pin, col + sz2 * k0);
    float rest_g = ldp(a.pin + QF * a.plane, col + sz2 * k0);

    // One plane.  WALL: -1 for k = 0 (stencil planes 1, 0, 2 = ring entries
    // 1, 0, 2), +1 for k = n-1 (n-2, n-1, n-3 = entries 1, 2, 0), 0 inside
    // (entries 0, 1, 2).  SLIDE: the window moves to the next plane, whose
    // arriving plane is fetched and interpolated (always inside a z group
    // below n-2; never at the walls or on a group's last plane).
    auto plane = [&](auto WALL, auto SLIDE, int k) {
        constexpr int W = decltype(WALL)::value;
        constexpr bool SL = decltype(SLIDE)::value;
        constexpr int E0 = W == 0 ? 0 : (W < 0 ? 1 : 1);
        constexpr int E1 = W == 0 ? 1 : (W < 0 ? 0 : 2);
        constexpr int E2 = W == 0 ? 2 : (W < 0 ? 2 : 0);
        const int kn = k + 1;
        const bool more = kn < k1;
        const int b = (k - k0) & 1;
        const float *cur = lds + b * L::total();
        float *nxt = lds + (b ^ 1) * L::total();
        st[2] = zm_stencil_z(a, k);
        const uint32_t node = col + sz2 * k;
        const int idx[3] = {i, j, k};
        float pcf[QF], pcg[QG], f[QF], g[QG];
        // Every value this plane reads was loaded during the previous one, so
        // the loads issued here (the next planes' rows and rest populations)
        // stay in flight through the whole compute; the in-plane prefetch is
        // issued unconditionally (clamped plane) so that the waits before the
        // commit count a fixed number of loads.
        pcf[0] = rest_f;
        pcg[0] = rest_g;
        {
            const uint32_t nn = col + sz2 * (uint32_t)min(kn, n - 1);
            rest_f = ldp(a.pin, nn);
            rest_g = ldp(a.pin + QF * a.plane, nn);
        }
        ZmStage<BY, false> np;
        ZmStage<BY, true> nzs;
        float pzn[L::npz()];
        zm_fetch<BY, false>(a, kn, pin_, np);
        if constexpr (SL) {
            const int znew = k + 2;   // the plane that enters the window
            zm_fetch<BY, true>(a, znew, pz, nzs);
            zm_fetch_purez<BY>(a, col, znew, pzn);
        }
        __builtin_amdgcn_sched_barrier(0);

        static_for<1, L::NP>([&](auto Pp) {
            constexpr int p = decltype(Pp)::value;
            if constexpr (p != QF) {
                float v;
                if constexpr (L::purez(p)) v = zr[L::zi(p)][E1];
                else if constexpr (L::zpop(p)) v = zo[L::oi(p)][E1];
                else v = zm_at<BY, p>(cur, l, r, 0, 0);
                if constexpr (p < QF) pcf[p] = v;
                else pcg[p - QF] = v;
            }
        });
        float sp_f = 0.0f, sp_g = 0.0f;
        static_for<0, QF>([&](auto A) { sp_f += pcf[decltype(A)::value]; });
        static_for<0, QG>([&](auto A) { sp_g += pcg[decltype(A)::value]; });
        f[0] = pcf[0];
        g[0] = pcg[0];
        static_for<1, L::NP>([&](auto Pp) {
            constexpr int p = decltype(Pp)::value;
            if constexpr (p != QF) {
                float v;
                if constexpr (L::zpop(p)) {
                    constexpr int z = L::zi(p);
                    const float *w = L::cz(p) > 0 ? st[2].wp : st[2].wm;
                    v = lerp3<false>(w, zr[z][E0], zr[z][E1], zr[z][E2]);
                } else if constexpr (p < QF) {
                    v = zm_interp_plane<BY, p>(cur, l, r, pcf[p], st);
                } else {
                    v = zm_interp_plane<BY, p>(cur, l, r, pcg[p - QF], st);
                }
                if constexpr (p < QF) f[p] = v;
                else g[p - QF] = v;
            }
        });
        float fp[QF], gp[QG];
        if (live)
            node_tail_collide<3, float, float, true, MACRO, NT>(a, fcol + fsz * k, idx, pcf, pcg,
                                                                f, g, sp_f, sp_g, hs, fp, gp);
        // The other buffer was last read before the previous plane's barrier
        // (its in-plane rows by the previous plane's compute, its z rows by
        // the ring_in of the plane before that), and is read next after this
        // plane's barrier: one barrier per plane orders every LDS write.
        if (more) zm_commit<BY, false>(nxt, pin_, np);
        if constexpr (SL) zm_commit<BY, true>(nxt, pz, nzs);
        zm_wait_loads();
        if (live) node_store<3, float, NT>(a.pout, a.plane, node, fp, gp);
        lds_barrier();
        if constexpr (SL) {
            zm_ring_shift(zr, zo);
            zm_ring_in<BY, 2>(nxt, l, r, st, pzn, zr, zo);
        }
    };

    using std::integral_constant;
    int k = k0;
    if (k == 0) {
        plane(integral_constant<int, -1>{}, std::false_type{}, 0);
        k = 1;
    }
    // planes whose next plane moves the window: k + 1 < k1 and k + 1 <= n - 2
    const int kslide = min(k1 - 1, n - 2);
    for (; k < kslide; k++) plane(integral_constant<int, 0>{}, std::true_type{}, k);
    // an inner plane without a move: the group's last plane, or n - 2
    if (k < min(k1, n - 1)) {
        plane(integral_constant<int, 0>{}, std::false_type{}, k);
        k++;
    }
    if (k == n - 1 && k < k1) plane(integral_constant<int, 1>{}, std::false_type{}, k);
}

}  // namespace lbm
