// lbm_step_ym.h -- y-marching variant of the 2D FP32 step (launch variant 1
// in 2D, and the batch's variant 1).
//
// The plain kernel reads about 60 stencil values per node through the L1,
// and at grids beyond the MALL its L1 and texture path (TA, TD), not HBM,
// bound it.  Here one wave owns a strip of 64 columns of a cavity over a
// chunk of rows and marches it in +y, one node per lane.  Each population
// value is loaded from HBM once per strip:
//   * c = (+-1, 0): the row k+1 is staged in LDS during row k (with two
//     halo columns on the left and one on the right, which cover the
//     one-sided wall stencils) and interpolated in x from LDS at row k+1;
//   * c = (+-1, +-1): each row is staged in LDS as it enters the y window
//     [ws, ws + 2], ws = clamp(k - 1, 0, n - 3); each lane interpolates it
//     in x at once and keeps the result and its own value in a three-entry
//     register ring, from which row k interpolates in y;
//   * c = (0, +-1): each lane keeps its own values of the window's rows in
//     the ring (no neighbours in x);
//   * the rest populations are loaded by each lane, one row ahead.
// The tensor product is applied x first (the plain kernel: y first), so
// the two agree to FP32 rounding, not to the bit.
//
// The ring holds the window's rows in order; inside the cavity a step to
// the next row shifts it and takes in the arriving row.  The wall rows
// k = 0 and k = n-1 (one-sided y stencils) run their own copies of the row
// body, which read the ring in their stencil order; the hot copy always
// shifts.  A workgroup is one wave, so the LDS handoffs need no cross-wave
// barrier; lds_barrier orders them within the wave.
//
// No alignment or size preconditions beyond n >= 5: loads are per lane, and
// lanes past the cavity's last column compute the last column's node and
// store nothing.
#pragma once

#include "lbm_step.h"

namespace lbm {

// Slots p = 0..13: flow populations 0..8, then thermal 0..4 (whose
// velocities are the first five flow velocities).
struct YmLayout {
    static constexpr int NP = 14;
    static constexpr int vel(int p) { return p < 9 ? p : p - 9; }
    static constexpr int cx(int p) { return Flow<2>::c[vel(p)][0]; }
    static constexpr int cy(int p) { return Flow<2>::c[vel(p)][1]; }
    static constexpr bool rest(int p) { return p == 0 || p == 9; }
    static constexpr bool xonly(int p) { return cx(p) != 0 && cy(p) == 0; }
    static constexpr bool yonly(int p) { return cx(p) == 0 && cy(p) != 0; }
    static constexpr bool diag(int p) { return cx(p) != 0 && cy(p) != 0; }
    static constexpr bool ypop(int p) { return cy(p) != 0; }
    // LDS row of a staged population: columns i0-2, i0-1 at 2, 3; i0..i0+63
    // at 4..67; i0+64 at 68
    static constexpr int PITCH = 72, C0 = 4;
    // slot of a population in its LDS region (the c = (+-1, 0) ones and the
    // diagonal ones are staged in two regions of four rows each)
    static constexpr int slot(int p) {
        int c = 0;
        for (int q = 0; q < p; q++) c += (xonly(p) ? xonly(q) : diag(q)) ? 1 : 0;
        return c;
    }
    static constexpr int region(int p) { return xonly(p) ? 0 : 4 * PITCH; }
    static constexpr int LDS = 8 * PITCH;
    // register ring indices: over the c_y != 0 populations, and over the
    // diagonal ones (whose own values are kept beside their x stage)
    static constexpr int yi(int p) {
        int c = 0;
        for (int q = 0; q < p; q++) c += ypop(q) ? 1 : 0;
        return c;
    }
    static constexpr int di(int p) {
        int c = 0;
        for (int q = 0; q < p; q++) c += diag(q) ? 1 : 0;
        return c;
    }
    static constexpr int ny() { return yi(NP); }
    static constexpr int nd() { return di(NP); }
};

// Halo value h (0..11) of a region: population slot h / 3, column
// i0-2, i0-1, i0+64 for h % 3 = 0, 1, 2.  Lane h < 12 owns it.
__device__ __forceinline__ int ym_halo_col(int i0, int h) {
    const int s = h % 3;
    return s == 0 ? i0 - 2 : (s == 1 ? i0 - 1 : i0 + 64);
}

// The populations of one region (XO: the c = (+-1, 0) ones; else the
// diagonal ones) at row kk (clamped): the lane's own value and its halo value.
template <bool XO>
__device__ __forceinline__ void ym_fetch(const StepArgs<float> &a, int kk, int i, int i0, int l,
                                         float *own, float &halo) {
    using L = YmLayout;
    kk = min(kk, a.n - 1);
    const uint32_t row = (uint32_t)kk * (uint32_t)a.n;
    static_for<0, L::NP>([&](auto Pp) {
        constexpr int p = decltype(Pp)::value;
        if constexpr (XO ? L::xonly(p) : L::diag(p))
            own[L::slot(p)] = ldp(a.pin + p * a.plane, row + (uint32_t)i);
    });
    // lanes 0..11 load a halo value; the others (and halos outside the
    // cavity) load the lane's own node of the first staged population
    const int h = min(l, 11);
    const int hc = ym_halo_col(i0, h);
    const bool inside = l < 12 && hc >= 0 && hc < a.n;
    int hp = 0, s = 0;
    static_for<0, L::NP>([&](auto Pp) {
        constexpr int p = decltype(Pp)::value;
        if constexpr (XO ? L::xonly(p) : L::diag(p)) {
            if (h / 3 == s) hp = p;
            s++;
        }
    });
    halo = ldp(a.pin + hp * a.plane, row + (uint32_t)(inside ? hc : i));
}

template <bool XO>
__device__ __forceinline__ void ym_commit(float *lds, int i0, int l, int n, const float *own,
                                          float halo) {
    using L = YmLayout;
    float *base = lds + (XO ? 0 : 4 * L::PITCH);
    static_for<0, 4>([&](auto S) {
        constexpr int s = decltype(S)::value;
        base[s * L::PITCH + L::C0 + l] = own[s];
    });
    if (l < 12) {
        const int hc = ym_halo_col(i0, l);
        if (hc >= 0 && hc < n) {
            const int m = l % 3;
            base[(l / 3) * L::PITCH + (m == 0 ? 2 : (m == 1 ? 3 : 68))] = halo;
        }
    }
}

// x interpolation of staged population p at the lane's column (own: the
// lane's value; dx offsets from the x stencil).
template <int p>
__device__ __forceinline__ float ym_xinterp(const float *lds, int l, float own,
                                            const AxisStencil<float> &sx) {
    using L = YmLayout;
    const float *row = lds + L::region(p) + L::slot(p) * L::PITCH + L::C0 + l;
    const float *w = L::cx(p) > 0 ? sx.wp : sx.wm;
    return lerp3<false>(w, row[sx.o0], own, row[sx.o2]);
}

// The arriving row (diagonal populations from LDS, c = (0, +-1) own values
// from yown) into ring entry E.
template <int E, int NY, int ND>
__device__ __forceinline__ void ym_ring_in(const float *lds, int l, const AxisStencil<float> &sx,
                                           const float *yown, float (&yr)[NY][3],
                                           float (&dov)[ND][3]) {
    using L = YmLayout;
    static_for<0, L::NP>([&](auto Pp) {
        constexpr int p = decltype(Pp)::value;
        if constexpr (L::yonly(p)) {
            yr[L::yi(p)][E] = yown[L::yi(p) - L::di(p)];
        } else if constexpr (L::diag(p)) {
            const float own = lds[L::region(p) + L::slot(p) * L::PITCH + L::C0 + l];
            dov[L::di(p)][E] = own;
            yr[L::yi(p)][E] = ym_xinterp<p>(lds, l, own, sx);
        }
    });
}

// The c = (0, +-1) populations of the lane's column at row kk (clamped), in
// ring order.
__device__ __forceinline__ void ym_fetch_y(const StepArgs<float> &a, int kk, int i, float *v) {
    using L = YmLayout;
    kk = min(kk, a.n - 1);
    const uint32_t node = (uint32_t)kk * (uint32_t)a.n + (uint32_t)i;
    static_for<0, L::NP>([&](auto Pp) {
        constexpr int p = decltype(Pp)::value;
        if constexpr (L::yonly(p)) v[L::yi(p) - L::di(p)] = ldp(a.pin + p * a.plane, node);
    });
}

__device__ __forceinline__ int ym_window(int k, int n) { return min(max(k - 1, 0), n - 3); }

// Rows [k0, k1) of the strip of columns [i0, i0 + 64) of one cavity.
template <bool MACRO, bool NT>
__device__ __forceinline__ void ym_march(const StepArgs<float> &a, float *lds, int i0, int k0,
                                         int k1) {
    using L = YmLayout;
    using FL = Flow<2>;
    using HT = Heat<2>;
    constexpr int QF = FL::Q, QG = HT::Q;
    constexpr int NY = L::ny(), ND = L::nd(), NYO = NY - ND;
    const int n = a.n;
    const int l = threadIdx.x;
    const bool live = i0 + l < n;
    const int i = min(i0 + l, n - 1);
    const AxisStencil<float> sx = load_stencil(a, i);
    // constant along y (the heated axis is x)
    const HeatShift<float> hs = load_heat_shift<2, true>(a, {i, 0, 0});

    float yr[NY][3], dov[ND][3];
    {
        const int ws = ym_window(k0, n);
        static_for<0, 3>([&](auto S) {
            constexpr int s = decltype(S)::value;
            float dv[4], yv[NYO], dh;
            ym_fetch<false>(a, ws + s, i, i0, l, dv, dh);
            ym_fetch_y(a, ws + s, i, yv);
            ym_commit<false>(lds, i0, l, n, dv, dh);
            lds_barrier();
            ym_ring_in<s>(lds, l, sx, yv, yr, dov);
            lds_barrier();
        });
        float xv[4], xh;
        ym_fetch<true>(a, k0, i, i0, l, xv, xh);
        ym_commit<true>(lds, i0, l, n, xv, xh);
    }
    lds_barrier();

    // Loads run two rows ahead.  Buffer b of a row k holds what row k
    // commits at its end (c = (+-1, 0) row k+1; diagonal and c = (0, +-1)
    // row k+2) and the rest populations of row k+1; row k issues the loads
    // of row k+1's buffer before it computes, so each load has a whole row
    // of compute and one wait in flight.  Rows alternate between two
    // buffers (B0 on even steps of the march, B1 on odd ones).
    struct Buf {
        float xv[4], xh, dv[4], dh, yv[NYO], rf, rg;
    };
    auto load = [&](Buf &b, int k) {
        ym_fetch<true>(a, k + 1, i, i0, l, b.xv, b.xh);
        ym_fetch<false>(a, k + 2, i, i0, l, b.dv, b.dh);
        ym_fetch_y(a, k + 2, i, b.yv);
        const uint32_t nn = (uint32_t)min(k + 1, n - 1) * (uint32_t)n + (uint32_t)i;
        b.rf = ldp(a.pin, nn);
        b.rg = ldp(a.pin + QF * a.plane, nn);
    };
    Buf B0, B1;
    float rest_f = ldp(a.pin, (uint32_t)k0 * (uint32_t)n + (uint32_t)i);
    float rest_g = ldp(a.pin + QF * a.plane, (uint32_t)k0 * (uint32_t)n + (uint32_t)i);
    load(B0, k0);

    // One row.  WALL: -1 for k = 0 (stencil rows 1, 0, 2 = ring entries
    // 1, 0, 2), +1 for k = n-1 (n-2, n-1, n-3 = entries 1, 2, 0), 0 inside.
    // SLIDE: the window moves to the next row, whose arriving row is taken
    // into the ring.  CUR: the row's buffer (the other one is loaded).
    auto row = [&](auto WALL, auto SLIDE, auto CUR, int k) {
        constexpr int W = decltype(WALL)::value;
        constexpr bool SL = decltype(SLIDE)::value;
        constexpr int E0 = W == 0 ? 0 : 1;
        constexpr int E1 = W == 0 ? 1 : (W < 0 ? 0 : 2);
        constexpr int E2 = W == 0 ? 2 : (W < 0 ? 2 : 0);
        Buf &cur = decltype(CUR)::value ? B1 : B0;
        Buf &nxt = decltype(CUR)::value ? B0 : B1;
        const bool more = k + 1 < k1;
        const AxisStencil<float> sy = load_stencil(a, k);
        const uint32_t node = (uint32_t)k * (uint32_t)n + (uint32_t)i;
        const int idx[3] = {i, k, 0};
        float pcf[QF], pcg[QG], f[QF], g[QG];
        pcf[0] = rest_f;
        pcg[0] = rest_g;
        load(nxt, k + 1);
        __builtin_amdgcn_sched_barrier(0);

        static_for<1, L::NP>([&](auto Pp) {
            constexpr int p = decltype(Pp)::value;
            if constexpr (!L::rest(p)) {
                float v;
                if constexpr (L::yonly(p)) v = yr[L::yi(p)][E1];
                else if constexpr (L::diag(p)) v = dov[L::di(p)][E1];
                else v = lds[L::region(p) + L::slot(p) * L::PITCH + L::C0 + l];
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
            if constexpr (!L::rest(p)) {
                float v;
                if constexpr (L::ypop(p)) {
                    constexpr int y = L::yi(p);
                    const float *w = L::cy(p) > 0 ? sy.wp : sy.wm;
                    v = lerp3<false>(w, yr[y][E0], yr[y][E1], yr[y][E2]);
                } else if constexpr (p < QF) {
                    v = ym_xinterp<p>(lds, l, pcf[p], sx);
                } else {
                    v = ym_xinterp<p>(lds, l, pcg[p - QF], sx);
                }
                if constexpr (p < QF) f[p] = v;
                else g[p - QF] = v;
            }
        });
        if (live)
            // 2D: unpadded rows, so the population and field indices agree
            node_tail<2, float, float, true, MACRO, NT>(a, node, node, idx, pcf, pcg, f, g, sp_f,
                                                        sp_g, hs);

        lds_barrier();   // every read of this row's staging is done
        if (more) ym_commit<true>(lds, i0, l, n, cur.xv, cur.xh);
        if constexpr (SL) ym_commit<false>(lds, i0, l, n, cur.dv, cur.dh);
        lds_barrier();
        if constexpr (SL) {
            static_for<0, NY>([&](auto Y) {
                constexpr int y = decltype(Y)::value;
                yr[y][0] = yr[y][1];
                yr[y][1] = yr[y][2];
            });
            static_for<0, ND>([&](auto D) {
                constexpr int d = decltype(D)::value;
                dov[d][0] = dov[d][1];
                dov[d][1] = dov[d][2];
            });
            ym_ring_in<2>(lds, l, sx, cur.yv, yr, dov);
        }
        rest_f = cur.rf;
        rest_g = cur.rg;
    };

    using std::integral_constant;
    using std::false_type;
    using std::true_type;
    const integral_constant<int, 0> in;
    int k = k0;
    if (k == 0) {
        row(integral_constant<int, -1>{}, false_type{}, false_type{}, 0);
        k = 1;
        B0 = B1;   // the next row is an even step again (waits for B1's loads once)
    }
    // rows whose next row moves the window: k + 1 < k1 and k + 1 <= n - 2
    const int kslide = min(k1 - 1, n - 2);
    for (; k + 1 < kslide; k += 2) {
        row(in, true_type{}, false_type{}, k);
        row(in, true_type{}, true_type{}, k + 1);
    }
    if (k < kslide) {
        row(in, true_type{}, false_type{}, k);
        k++;
        B0 = B1;
    }
    // an inner row without a move: the chunk's last row, or n - 2
    if (k < min(k1, n - 1)) {
        row(in, false_type{}, false_type{}, k);
        k++;
        B0 = B1;
    }
    if (k == n - 1 && k < k1) row(integral_constant<int, 1>{}, false_type{}, false_type{}, k);
}

// One cavity: workgroup b runs strip tx = L % gx, chunk L / gx of logical
// tile L; with remap 1 each XCD (workgroups b = x mod 8) sweeps one
// contiguous range of logical tiles, so the x neighbours that read a
// strip's edge columns run on the same XCD.
template <bool MACRO, bool NT>
__global__ __launch_bounds__(64) void k_step_ym(StepArgs<float> a, int kc) {
    __shared__ float lds[YmLayout::LDS];
    const int L = logical_tile((int)blockIdx.x, a.ntiles, a.remap);
    const int tx = L % a.gx, ch = L / a.gx;
    const int k0 = ch * kc;
    if (k0 >= a.n) return;
    ym_march<MACRO, NT>(a, lds, tx * 64, k0, min(a.n, k0 + kc));
}

// A batch: workgroup b takes table entry b = (cavity, strip, chunk,
// argument index at parity 0); cavity < 0 marks padding.
template <bool MACRO, bool NT>
__global__ __launch_bounds__(64) void k_step_ym_batch(const StepArgs<float> *__restrict__ args,
                                                      const int4 *__restrict__ tiles, int parity,
                                                      int kc) {
    __shared__ float lds[YmLayout::LDS];
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;
    const StepArgs<float> a = args[t.w ^ parity];
    const int k0 = t.z * kc;
    ym_march<MACRO, NT>(a, lds, t.y * 64, k0, min(a.n, k0 + kc));
}

}  // namespace lbm
