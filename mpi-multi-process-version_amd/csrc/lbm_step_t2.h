// lbm_step_t2.h -- two time steps per pass for the 2D FP32 step (launch
// variant 2 of a single 2D cavity).
//
// A 2D FP32 step moves 112 bytes per node through HBM for a few hundred
// register operations; a workgroup that keeps a tile in registers can run
// two steps on it before writing it back, so a pass moves about 65 bytes per
// node and step.  A workgroup of four waves owns an output tile of 60 x TY
// nodes (columns x0 .. x0+59, rows y0 .. y0+TY-1) and holds the region two
// nodes wider on every side, 64 x (TY + 4) nodes: lane l of wave w holds
// column x0 - 2 + l of the region rows w, w + 4, ... (K = (TY + 4) / 4 rows
// per lane).  A pass
//   1. loads P(k) of the region's nodes (those inside the cavity);
//   2. runs step k+1 on the region shrunk by one node on every side, the
//      stencils reading P(k) from LDS (the exchange of the resident batch:
//      eight flow planes, then four thermal ones, interp_pop in k_step's
//      operand order, the wall rules, node_collide);
//   3. runs step k+2 on the output tile from the step-(k+1) values;
//   4. stores P(k+2) of the output tile.
// Every node update is k_step's arithmetic on the values k_step would read,
// so a pass gives the bits of two k_step launches.  Reach: a centred
// stencil reaches one node, and the one-sided wall stencils reach two nodes
// inward; tiles start at x0 = min(60 tx, n - 2) (y0 likewise), so the far
// wall is never a tile's first column or row and its two inward points lie
// in the region a step has updated.  Tiles clamped that way overlap their
// neighbour, which writes the same values to the shared nodes.
//
// The host flips the buffer parity once per pass; an odd step count ends
// with a one-step pass.
//
// The LDS exchange of a step is written out here as in k_step_batch_res
// rather than shared through one device function: with the step body in a
// separate inlined function, the compiler gave the resident kernel 18 to 37
// more VGPRs at 64^2 (and spills), so each kernel keeps its own copy.
#pragma once

#include "lbm_step.h"

namespace lbm {

constexpr int kT2Width = 64;          // region columns (one wave per region row)
constexpr int kT2Out = kT2Width - 4;  // output columns per tile
constexpr int kT2Threads = 256;

// Per-index tables of the region's columns and rows in LDS: stencil offsets,
// six weights and (columns only) Tref and the two streaming corrections.
template <int H> struct T2Tables {
    int2 xoff[kT2Width];
    float xwt[6 * kT2Width];
    float xtref[kT2Width];
    float xtcorr[2 * kT2Width];
    int2 yoff[H];
    float ywt[6 * H];
};

__device__ __forceinline__ AxisStencil<float> t2_stencil(const int2 *off, const float *wt, int r) {
    AxisStencil<float> s;
    const int2 o = off[r];
    s.o0 = o.x;
    s.o2 = o.y;
#pragma unroll
    for (int b = 0; b < 3; b++) {
        s.wp[b] = wt[6 * r + b];
        s.wm[b] = wt[6 * r + 3 + b];
    }
    return s;
}

// One pass of NSTEP (2, or 1 for the last step of an odd count) steps over
// tile (tx, ty) of the cavity of a (the whole body of the single-cavity and
// the batched kernels).  A one-step pass updates the output tile from the
// loaded region directly.
template <int TY, int NSTEP, bool MACRO, bool NT>
__device__ __forceinline__ void t2_pass(const StepArgs<float> &a, int tx, int ty) {
    using FL = Flow<2>;
    using HT = Heat<2>;
    constexpr int QF = FL::Q, QG = HT::Q;
    constexpr int W = kT2Width, H = TY + 4, K = H / 4, PS = W * H;
    static_assert(H % 4 == 0, "region rows must split over the four waves");
    __shared__ __attribute__((aligned(16))) float lds[8 * PS];
    __shared__ __attribute__((aligned(16))) T2Tables<H> tab;
    const int n = a.n;
    const int x0 = min(tx * kT2Out, n - 2), y0 = min(ty * TY, n - 2);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int i = x0 - 2 + lane;
    const bool col_in = i >= 0 && i < n;

    for (int c = threadIdx.x; c < W; c += kT2Threads) {
        const int gi = min(max(x0 - 2 + c, 0), n - 1);
        tab.xoff[c] = a.off[gi];
        tab.xtref[c] = a.tref[gi];
        tab.xtcorr[2 * c] = a.tcorr[2 * gi];
        tab.xtcorr[2 * c + 1] = a.tcorr[2 * gi + 1];
    }
    for (int c = threadIdx.x; c < 6 * W; c += kT2Threads)
        tab.xwt[c] = a.wt[6 * min(max(x0 - 2 + c / 6, 0), n - 1) + c % 6];
    for (int r = threadIdx.x; r < H; r += kT2Threads) tab.yoff[r] = a.off[min(max(y0 - 2 + r, 0), n - 1)];
    for (int c = threadIdx.x; c < 6 * H; c += kT2Threads)
        tab.ywt[c] = a.wt[6 * min(max(y0 - 2 + c / 6, 0), n - 1) + c % 6];

    // slot k: region row r = w + 4 k, cavity row j = y0 - 2 + r
    float sf[K][QF], sg[K][QG], fprev[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int j = y0 - 2 + w + 4 * k;
        if (col_in && j >= 0 && j < n) {
            const uint32_t node = (uint32_t)i + (uint32_t)n * (uint32_t)j;
            static_for<0, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                sf[k][al] = ldp(a.pin + al * a.plane, node);
            });
            static_for<0, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                sg[k][al] = ldp(a.pin + (QF + al) * a.plane, node);
            });
        }
    }
    const uint32_t stride[3] = {1u, (uint32_t)W, 0u};
    // nodes of slot k inside the region shrunk by m on every side (and in
    // the cavity)
    auto inside = [&](int k, int m) {
        const int r = w + 4 * k, j = y0 - 2 + r;
        return col_in && j >= 0 && j < n && lane >= m && lane < W - m && r >= m && r < H - m;
    };

    // One step: the nodes of the region shrunk by m - 1 feed it, those of the
    // region shrunk by m are updated (m = 1, then 2).
    auto step = [&](auto M, auto MC) {
        constexpr int m = decltype(M)::value;
        constexpr bool MCR = decltype(MC)::value;
        lds_barrier();   // the previous step's thermal reads (and the tables)
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!inside(k, m - 1)) continue;
            const int q = lane + W * (w + 4 * k);
            static_for<1, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                lds[(al - 1) * PS + q] = sf[k][al];
            });
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!inside(k, m)) continue;
            const int r = w + 4 * k;
            const int q = lane + W * r;
            const int idx[3] = {i, y0 - 2 + r, 0};
            AxisStencil<float> st[3];
            st[0] = t2_stencil(tab.xoff, tab.xwt, lane);
            st[1] = t2_stencil(tab.yoff, tab.ywt, r);
            float sp = 0.0f;
            static_for<0, QF>([&](auto A) { sp += sf[k][decltype(A)::value]; });
            float f[QF];
            f[0] = sf[k][0];
            static_for<1, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                f[al] = interp_pop<FL::c[al][0], FL::c[al][1], 0, false>(
                    lds + (al - 1) * PS, (uint32_t)q, sf[k][al], st, stride);
            });
            node_walls_flow<2>(n, idx, sf[k], f);
            static_for<0, QF>([&](auto A) { sf[k][decltype(A)::value] = f[decltype(A)::value]; });
            fprev[k] = sp;   // the sum of P(k)'s flow populations, for F(k)
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!inside(k, m - 1)) continue;
            const int q = lane + W * (w + 4 * k);
            static_for<1, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                lds[(al - 1) * PS + q] = sg[k][al];
            });
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!inside(k, m)) continue;
            const int r = w + 4 * k;
            const int q = lane + W * r;
            const int j = y0 - 2 + r;
            const int idx[3] = {i, j, 0};
            AxisStencil<float> st[3];
            st[0] = t2_stencil(tab.xoff, tab.xwt, lane);
            st[1] = t2_stencil(tab.yoff, tab.ywt, r);
            const HeatShift<float> hs{tab.xtref[lane], tab.xtcorr[2 * lane],
                                      tab.xtcorr[2 * lane + 1]};
            float sp = 0.0f;
            static_for<0, QG>([&](auto A) { sp += sg[k][decltype(A)::value]; });
            float g[QG];
            g[0] = sg[k][0];
            static_for<1, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                g[al] = interp_pop<FL::c[al][0], FL::c[al][1], 0, false>(
                    lds + (al - 1) * PS, (uint32_t)q, sg[k][al], st, stride);
            });
            node_walls_heat<2, true>(a, idx, sg[k], g, hs);
            const float Fp = lagged_force<true>(a.gbeta, fprev[k], sp, hs.tloc);
            float fp[QF], gp[QG];
            node_collide<2, float, float, true, MCR>(a, (uint32_t)i + (uint32_t)n * (uint32_t)j,
                                                     sf[k], g, Fp, hs.tloc, fp, gp);
            static_for<0, QF>([&](auto A) { sf[k][decltype(A)::value] = fp[decltype(A)::value]; });
            static_for<0, QG>([&](auto A) { sg[k][decltype(A)::value] = gp[decltype(A)::value]; });
        }
    };
    if constexpr (NSTEP == 2) step(std::integral_constant<int, 1>{}, std::false_type{});
    step(std::integral_constant<int, 2>{}, std::integral_constant<bool, MACRO>{});

#pragma unroll
    for (int k = 0; k < K; k++) {
        if (!inside(k, 2)) continue;
        const uint32_t node = (uint32_t)i + (uint32_t)n * (uint32_t)(y0 - 2 + w + 4 * k);
        static_for<0, QF>([&](auto A) {
            constexpr int al = decltype(A)::value;
            stp<NT>(a.pout + al * a.plane, node, sf[k][al]);
        });
        static_for<0, QG>([&](auto A) {
            constexpr int al = decltype(A)::value;
            stp<NT>(a.pout + (QF + al) * a.plane, node, sg[k][al]);
        });
    }
}

// Tile (tx, ty) of workgroup P: the x-fastest order, or, with remap 1, one
// contiguous range of tiles per XCD (logical_tile).
template <int TY, int NSTEP, bool MACRO, bool NT>
__global__ __launch_bounds__(kT2Threads) void k_step_t2(StepArgs<float> a) {
    const int L = logical_tile((int)blockIdx.x, a.ntiles, a.remap);
    t2_pass<TY, NSTEP, MACRO, NT>(a, L % a.gx, L / a.gx);
}

// The batch's variant 4: workgroup b takes entry b of a table of
// (cavity, tile x, tile y, argument index at parity 0) and runs one pass
// on it with args[tiles[b].w ^ parity].
template <int TY, int NSTEP, bool MACRO, bool NT>
__global__ __launch_bounds__(kT2Threads) void k_step_t2_batch(
    const StepArgs<float> *__restrict__ args, const int4 *__restrict__ tiles, int parity) {
    const int4 t = tiles[blockIdx.x];
    t2_pass<TY, NSTEP, MACRO, NT>(args[t.w ^ parity], t.y, t.z);
}

}  // namespace lbm
