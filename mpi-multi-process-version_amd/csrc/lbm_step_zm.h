// lbm_step_zm.h -- LDS-staged, z-marching variant of the 3D FP32 step
// (launch variant 1).
//
// The plain kernel (k_step) reads 156 values per node through the L1; the
// stencil rows it shares with other workgroups (the y halo and the planes
// k +- 1) are L1 misses, and the L1's miss handling, not HBM, bounds it.
// Here a workgroup owns a 64 x BY column of nodes over a range of planes and
// marches it in +z.  The post-collision populations it needs are staged in
// LDS once per plane:
//   * populations with c_z = 0 (except the rest ones): plane k, with a one-
//     node halo in x if c_x != 0 and in y if c_y != 0;
//   * populations with c_z != 0: a ring of three planes, the window
//     [ws, ws + 2] with ws = clamp(k - 1, 0, n - 3), which holds every z
//     stencil point of plane k, one-sided wall stencils included;
//   * the two rest populations are read by each lane directly.
// Rows are staged four at a time: a lane loads four consecutive x values
// (one 16-byte load), so a wave moves four 64-wide rows per instruction, and
// the rows of the next plane are fetched into registers while the current
// plane is computed.  Every stencil read is an LDS read.  The interpolation
// expressions and their operands are those of k_step; the results agree
// with it to FP32 rounding (the compiler fuses a different handful of
// multiply-adds in the two kernels).
//
// All waves run one copy of the loop: each thread's staging offsets are
// computed once (a plan) and the loads and LDS writes of a plane are
// straight-line code.
//
// Preconditions (checked by the host): FP32 storage, 3D, bx = 64, by = BY,
// n % 4 == 0 (16-byte aligned rows; the far-wall node then never sits in
// the first column or row of a tile, so its one-sided stencil stays inside
// the staged halo) and n <= 544 (32-bit element offsets over 26 planes).
#pragma once

#include "lbm_step.h"

namespace lbm {

// A native 4-float vector (HIP's float4 is a union type that the compiler
// does not keep in registers when it sits in an array).
typedef float zm_f4 __attribute__((ext_vector_type(4)));
typedef float zm_f2 __attribute__((ext_vector_type(2)));

// Staging slots p = 0..25: flow populations 0..18, then thermal 0..6
// (whose velocities are the first seven flow velocities).  The population
// plane of slot p is plane p of the buffer.  A staged region is H rows of
// pitch P: x columns i0 .. i0+63 at C0 .. C0+63, and for populations with
// an x halo, column i0-1 at 3 and column i0+64 at 68 (so rows stay 16-byte
// aligned).
template <int BY> struct ZmLayout {
    static constexpr int NP = 26;
    static constexpr int vel(int p) { return p < 19 ? p : p - 19; }
    static constexpr int cx(int p) { return Flow<3>::c[vel(p)][0]; }
    static constexpr int cy(int p) { return Flow<3>::c[vel(p)][1]; }
    static constexpr int cz(int p) { return Flow<3>::c[vel(p)][2]; }
    static constexpr bool staged(int p) { return p != 0 && p != 19; }
    static constexpr bool zpop(int p) { return cz(p) != 0; }
    static constexpr int hx(int p) { return cx(p) != 0 ? 1 : 0; }
    static constexpr int hy(int p) { return cy(p) != 0 ? 1 : 0; }
    static constexpr int P(int p) { return hx(p) ? 72 : 64; }
    static constexpr int C0(int p) { return hx(p) ? 4 : 0; }
    static constexpr int H(int p) { return BY + 2 * hy(p); }
    static constexpr int slots(int p) { return zpop(p) ? 3 : 1; }
    static constexpr int psize(int p) { return staged(p) ? P(p) * H(p) : 0; }
    static constexpr int off(int p) {
        int o = 0;
        for (int q = 0; q < p; q++) o += psize(q) * slots(q);
        return o;
    }
    static constexpr int total() { return off(NP); }
    // Rows of a phase (Z: the c_z != 0 populations at a new plane; !Z: the
    // in-plane populations at plane k), in slot order.  Staging instruction
    // q of wave w gives lane group g (16 lanes of 4 x values each) row
    // R = 4 (q BY + w) + g, clamped to the last row (a repeated row writes
    // the same values).
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
// 0 and their LDS index in ring slot 0 (low 16 bits) with the region size
// (high 16 bits, the ring stride); the same for its halo value, if any
// (thread t owns halo value t: population, region row, side).
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
    const uint32_t plane = (uint32_t)a.plane;
    const int g = l >> 4, c4 = 4 * (l & 15);
    const int ii = min(i0 + c4, n - 4);   // n % 4 == 0: groups never straddle the edge
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
                    go = (uint32_t)p * plane + (uint32_t)jj * n + ii;
                    ld = (uint32_t)(L::off(p) + rr * L::P(p) + L::C0(p) + c4) |
                         ((uint32_t)L::psize(p) << 16);
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
                pl.hoff = (uint32_t)p * plane + (uint32_t)jj * n + ih;
                pl.hlds = (uint32_t)(L::off(p) + rr * L::P(p) + (side ? 68 : 3)) |
                          ((uint32_t)L::psize(p) << 16);
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
    const size_t po = (size_t)kk * (size_t)a.n * (size_t)a.n;
    static_for<0, L::ninstr(Z)>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        s.quad[q] = *reinterpret_cast<const zm_f4 *>(a.pin + (po + pl.goff[q]));
    });
    // threads without a halo value load the plane origin (unused)
    s.halo = a.pin[po + pl.hoff];
}

// Writes fetched values into their LDS slot (Z: ring slot of the plane),
// as two 8-byte writes per group of four (fewer bank conflicts than one
// 16-byte write).
template <int BY, bool Z>
__device__ __forceinline__ void zm_commit(float *lds, const ZmPlan<BY, Z> &pl,
                                          const ZmStage<BY, Z> &s, int slot) {
    using L = ZmLayout<BY>;
    static_for<0, L::ninstr(Z)>([&](auto Q) {
        constexpr int q = decltype(Q)::value;
        const uint32_t idx = (pl.lds[q] & 0xffffu) + (Z ? (uint32_t)slot * (pl.lds[q] >> 16) : 0u);
        *reinterpret_cast<zm_f2 *>(&lds[idx]) = s.quad[q].xy;
        *reinterpret_cast<zm_f2 *>(&lds[idx + 2]) = s.quad[q].zw;
    });
    if (pl.hlds != ~0u)
        lds[(pl.hlds & 0xffffu) + (Z ? (uint32_t)slot * (pl.hlds >> 16) : 0u)] = s.halo;
}

// Value of slot p at x, y offsets (dx, dy) from the lane's node, in ring
// slot zs (0 for populations staged on one plane).
template <int BY, int p>
__device__ __forceinline__ float zm_at(const float *lds, int l, int r, int dx, int dy,
                                       int zs) {
    using L = ZmLayout<BY>;
    return lds[L::off(p) + L::psize(p) * zs + (r + L::hy(p) + dy) * L::P(p) + L::C0(p) + l +
               dx];
}

// Value of slot p at stencil point S1 along axis A1 and S2 along axis A2
// (point 1 is the node itself; A < 0 means no axis).  zsl[s] is the ring
// slot of the plane of z stencil point s.
template <int BY, int p, int A1, int S1, int A2, int S2>
__device__ __forceinline__ float zm_pt(const float *lds, int l, int r,
                                       const AxisStencil<float> *st, const int *zsl) {
    using L = ZmLayout<BY>;
    int dx = 0, dy = 0, zs = zsl[1];
    if constexpr (A1 == 0 && S1 != 1) dx = S1 == 0 ? st[0].o0 : st[0].o2;
    if constexpr (A2 == 0 && S2 != 1) dx = S2 == 0 ? st[0].o0 : st[0].o2;
    if constexpr (A1 == 1 && S1 != 1) dy = S1 == 0 ? st[1].o0 : st[1].o2;
    if constexpr (A2 == 1 && S2 != 1) dy = S2 == 0 ? st[1].o0 : st[1].o2;
    if constexpr (A1 == 2 && S1 != 1) zs = zsl[S1];
    if constexpr (A2 == 2 && S2 != 1) zs = zsl[S2];
    return zm_at<BY, p>(lds, l, r, dx, dy, L::zpop(p) ? zs : 0);
}

// interp_pop with its operands from LDS: the same expressions in the same
// order.
template <int BY, int p>
__device__ __forceinline__ float zm_interp(const float *lds, int l, int r, float fc,
                                           const AxisStencil<float> *st, const int *zsl) {
    using L = ZmLayout<BY>;
    constexpr int c[3] = {L::cx(p), L::cy(p), L::cz(p)};
    constexpr int nz = (c[0] != 0) + (c[1] != 0) + (c[2] != 0);
    if constexpr (nz == 1) {
        constexpr int ax = c[0] ? 0 : (c[1] ? 1 : 2);
        const float *w = c[ax] > 0 ? st[ax].wp : st[ax].wm;
        return lerp3<false>(w, zm_pt<BY, p, ax, 0, -1, 1>(lds, l, r, st, zsl), fc,
                            zm_pt<BY, p, ax, 2, -1, 1>(lds, l, r, st, zsl));
    } else {
        constexpr int q = c[0] ? 0 : 1;   // outer stage axis
        constexpr int pp = c[2] ? 2 : 1;  // inner stage axis
        const float *wpp = c[pp] > 0 ? st[pp].wp : st[pp].wm;
        const float *wqq = c[q] > 0 ? st[q].wp : st[q].wm;
        const float c0 = lerp3<false>(wpp, zm_pt<BY, p, q, 0, pp, 0>(lds, l, r, st, zsl),
                                      zm_pt<BY, p, q, 0, pp, 1>(lds, l, r, st, zsl),
                                      zm_pt<BY, p, q, 0, pp, 2>(lds, l, r, st, zsl));
        const float c1 = lerp3<false>(wpp, zm_pt<BY, p, q, 1, pp, 0>(lds, l, r, st, zsl), fc,
                                      zm_pt<BY, p, q, 1, pp, 2>(lds, l, r, st, zsl));
        const float c2 = lerp3<false>(wpp, zm_pt<BY, p, q, 2, pp, 0>(lds, l, r, st, zsl),
                                      zm_pt<BY, p, q, 2, pp, 1>(lds, l, r, st, zsl),
                                      zm_pt<BY, p, q, 2, pp, 2>(lds, l, r, st, zsl));
        return lerp3<false>(wqq, c0, c1, c2);
    }
}

__device__ __forceinline__ int zm_window(int k, int n) { return min(max(k - 1, 0), n - 3); }

template <int BY, bool MACRO, bool NT>
__global__ __launch_bounds__(64 * BY) void k_step_zm(StepArgs<float> a, int kc) {
    using L = ZmLayout<BY>;
    using FL = Flow<3>;
    using HT = Heat<3>;
    constexpr int QF = FL::Q, QG = HT::Q;
    __shared__ zm_f4 lds4[L::total() / 4];
    float *lds = reinterpret_cast<float *>(lds4);
    const int n = a.n;
    int tx, ty, tc;
    tile_coords((int)blockIdx.x, a.ntiles, a.gx, a.gy, a.remap, tx, ty, tc);
    const int i0 = tx * 64, j0 = ty * BY;
    const int k0 = tc * kc, k1 = min(n, k0 + kc);
    const int l = threadIdx.x;
    const int r = __builtin_amdgcn_readfirstlane((int)threadIdx.y);  // one row per wave
    const bool live = i0 + l < n && j0 + r < n;
    // lanes past the grid edge stage and compute like the last node and
    // store nothing (every lane must reach the barriers)
    const int i = min(i0 + l, n - 1), j = min(j0 + r, n - 1);
    AxisStencil<float> st[3];
    st[0] = load_stencil(a, i);
    st[1] = load_stencil(a, j);
    const ZmPlan<BY, false> pin_ = zm_plan<BY, false>(a, i0, j0, r, l);
    const ZmPlan<BY, true> pz = zm_plan<BY, true>(a, i0, j0, r, l);

    {   // prologue: the z window of plane k0 and the in-plane rows of k0
        const int ws = zm_window(k0, n);
        for (int zz = ws; zz < ws + 3; zz++) {
            ZmStage<BY, true> sz;
            zm_fetch<BY, true>(a, zz, pz, sz);
            zm_commit<BY, true>(lds, pz, sz, zz % 3);
        }
        ZmStage<BY, false> sp;
        zm_fetch<BY, false>(a, k0, pin_, sp);
        zm_commit<BY, false>(lds, pin_, sp, 0);
    }
    __syncthreads();

    const uint32_t sy = n, sz2 = (uint32_t)n * n;
    // constant along z (the heated axis is y)
    const HeatShift<float> hs = load_heat_shift<3, true>(a, {i, j, 0});
    for (int k = k0; k < k1; k++) {
        const int kn = k + 1;
        const bool more = kn < k1;
        const int wsn = zm_window(kn, n);
        const bool newz = more && wsn != zm_window(k, n);
        st[2] = load_stencil(a, k);
        const int zsl[3] = {(k + st[2].o0) % 3, k % 3, (k + st[2].o2) % 3};
        const uint32_t node = i + sy * j + sz2 * k;
        const int idx[3] = {i, j, k};
        float pcf[QF], pcg[QG], f[QF], g[QG];
        // This plane's own loads go out before the next plane's rows: vector
        // loads complete in order, so the compute below waits only for them
        // and the prefetch stays in flight behind it.  The prefetch is issued
        // unconditionally (clamped planes) for the same reason.
        pcf[0] = ldp(a.pin, node);
        pcg[0] = ldp(a.pin + QF * a.plane, node);
        __builtin_amdgcn_sched_barrier(0);
        ZmStage<BY, false> np;
        ZmStage<BY, true> nzs;
        zm_fetch<BY, false>(a, kn, pin_, np);
        zm_fetch<BY, true>(a, wsn + 2, pz, nzs);
        __builtin_amdgcn_sched_barrier(0);

        static_for<1, L::NP>([&](auto Pp) {
            constexpr int p = decltype(Pp)::value;
            if constexpr (L::staged(p)) {
                const float v = zm_at<BY, p>(lds, l, r, 0, 0, L::zpop(p) ? k % 3 : 0);
                if constexpr (p < QF) pcf[p] = v;
                else pcg[p - QF] = v;
            }
        });
        float sp_f = 0.0f, sp_g = 0.0f;
        static_for<0, QF>([&](auto A) { sp_f += pcf[decltype(A)::value]; });
        static_for<0, QG>([&](auto A) { sp_g += pcg[decltype(A)::value]; });
        f[0] = pcf[0];
        g[0] = pcg[0];
        static_for<1, QF>([&](auto A) {
            constexpr int al = decltype(A)::value;
            f[al] = zm_interp<BY, al>(lds, l, r, pcf[al], st, zsl);
        });
        static_for<1, QG>([&](auto A) {
            constexpr int al = decltype(A)::value;
            g[al] = zm_interp<BY, QF + al>(lds, l, r, pcg[al], st, zsl);
        });
        if (live)
            node_tail<3, float, float, true, MACRO, NT>(a, node, idx, pcf, pcg, f, g, sp_f, sp_g,
                                                        hs);

        __syncthreads();   // every read of plane k's staging is done
        if (more) zm_commit<BY, false>(lds, pin_, np, 0);
        if (newz) zm_commit<BY, true>(lds, pz, nzs, (wsn + 2) % 3);
        __syncthreads();
    }
}

}  // namespace lbm
