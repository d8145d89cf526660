// lbm_step_res.h -- the batch's resident launch (variant 3): one workgroup
// per small 2D cavity keeps the cavity's whole state on the chip and steps
// it many times in one launch.
//
// A 2D FP32 cavity of n <= 70 holds at most 70^2 x 14 populations (274 KB),
// which fit the registers of one workgroup of 512 threads: thread t owns the
// nodes q = t, t + 512, ... (q = i + n j, at most ten) and keeps their 14
// populations, so no population goes through HBM between the launch's first
// load and its last store.  Streaming goes through LDS: per step
//   1. every thread writes its nodes' non-rest flow populations P(k) into
//      eight LDS planes of n^2 values, and after a barrier interpolates its
//      nodes' f(k) from them with the stencil expressions of k_step
//      (interp_pop, the same operand order) and applies the flow wall rules;
//   2. the same for the four non-rest thermal populations, in the first four
//      planes (the flow reads are finished at the barrier before);
//   3. collides its nodes in registers (node_collide, k_step's arithmetic).
// A barrier before each step's first LDS write ends the previous step's
// thermal reads.  The per-index stencil tables (offsets, weights, the heat
// shift) are copied into LDS once.  Each step of each cavity therefore runs
// the same operations on the same values as the single-cavity kernel and
// gives the same bits.  The last step of the launch writes the FP64
// macroscopic fields (the MACRO step of k_step); the populations are then
// stored to the buffer the cavity's parity expects after `steps` steps.
//
// Limits (checked on the host): 2D, FP32 shifted storage, unpadded rows
// (pitch n), n <= kResMaxN.  LDS: 8 n^2 floats + 11 n table values, 160 KB
// at n = 70, one workgroup per CU.
#pragma once

#include "lbm_step.h"

namespace lbm {

constexpr int kResThreads = 512;
constexpr int kResMaxN = 70;
constexpr int kResMaxK = (kResMaxN * kResMaxN + kResThreads - 1) / kResThreads;   // 10
constexpr int kResPlaneFloats = 8 * kResMaxN * kResMaxN;
// per grid index: stencil offsets (int2), six weights, Tref, two corrections
constexpr int kResTableFloats = 11 * kResMaxN;

// Nodes per thread for a largest cavity of n^2 nodes.
inline int res_nodes_per_thread(int n) { return (n * n + kResThreads - 1) / kResThreads; }

// args[table[b].w ^ parity] is the cavity of workgroup b reading its
// current buffer; args[table[b].w ^ parity ^ 1].pout is that same buffer.  K >= the
// nodes per thread of every cavity in the table.
template <int K>
__global__ __launch_bounds__(kResThreads) void k_step_batch_res(
    const StepArgs<float> *__restrict__ args, const int4 *__restrict__ table, int parity,
    int steps) {
    using FL = Flow<2>;
    using HT = Heat<2>;
    constexpr int QF = FL::Q, QG = HT::Q;
    __shared__ __attribute__((aligned(16))) float lds[kResPlaneFloats + kResTableFloats];
    const int ai = table[blockIdx.x].w ^ parity;
    StepArgs<float> a = args[ai];
    float *const pfinal = (steps & 1) ? a.pout : args[ai ^ 1].pout;
    const int n = a.n, N = n * n;
    const int t = threadIdx.x;

    // the stencil and heat-shift tables in LDS; the node code reads them
    // through the argument block's pointers
    int2 *s_off = reinterpret_cast<int2 *>(lds + kResPlaneFloats);
    float *s_wt = lds + kResPlaneFloats + 2 * kResMaxN;
    float *s_tref = s_wt + 6 * kResMaxN;
    float *s_tcorr = s_tref + kResMaxN;
    for (int x = t; x < n; x += kResThreads) {
        s_off[x] = a.off[x];
        s_tref[x] = a.tref[x];
        s_tcorr[2 * x] = a.tcorr[2 * x];
        s_tcorr[2 * x + 1] = a.tcorr[2 * x + 1];
    }
    for (int x = t; x < 6 * n; x += kResThreads) s_wt[x] = a.wt[x];
    a.off = s_off;
    a.wt = s_wt;
    a.tref = s_tref;
    a.tcorr = s_tcorr;

    // the nodes' populations P(k), then f(k), g(k) between the phases
    float sf[K][QF], sg[K][QG], fprev[K];
    int ij[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int q = t + kResThreads * k;
        ij[k] = q < N ? ((q % n) | ((q / n) << 16)) : -1;
        if (q < N) {
            static_for<0, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                sf[k][al] = ldp(a.pin + al * a.plane, (uint32_t)q);
            });
            static_for<0, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                sg[k][al] = ldp(a.pin + (QF + al) * a.plane, (uint32_t)q);
            });
        }
    }
    const uint32_t stride[3] = {1u, (uint32_t)n, 0u};

    auto step = [&](auto MACRO) {
        constexpr bool MC = decltype(MACRO)::value;
        lds_barrier();   // the previous step's thermal reads (and the tables)
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (ij[k] < 0) continue;
            const int q = t + kResThreads * k;
            static_for<1, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                lds[(al - 1) * N + q] = sf[k][al];
            });
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (ij[k] < 0) continue;
            const int q = t + kResThreads * k;
            const int idx[3] = {ij[k] & 0xffff, ij[k] >> 16, 0};
            AxisStencil<float> st[3];
            st[0] = load_stencil(a, idx[0]);
            st[1] = load_stencil(a, idx[1]);
            float sp = 0.0f;
            static_for<0, QF>([&](auto A) { sp += sf[k][decltype(A)::value]; });
            float f[QF];
            f[0] = sf[k][0];
            static_for<1, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                f[al] = interp_pop<FL::c[al][0], FL::c[al][1], 0, false>(
                    lds + (al - 1) * N, (uint32_t)q, sf[k][al], st, stride);
            });
            node_walls_flow<2>(n, idx, sf[k], f);
            static_for<0, QF>([&](auto A) { sf[k][decltype(A)::value] = f[decltype(A)::value]; });
            fprev[k] = sp;   // the sum of P(k)'s flow populations, for F(k)
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (ij[k] < 0) continue;
            const int q = t + kResThreads * k;
            static_for<1, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                lds[(al - 1) * N + q] = sg[k][al];
            });
        }
        lds_barrier();
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (ij[k] < 0) continue;
            const int q = t + kResThreads * k;
            const int idx[3] = {ij[k] & 0xffff, ij[k] >> 16, 0};
            AxisStencil<float> st[3];
            st[0] = load_stencil(a, idx[0]);
            st[1] = load_stencil(a, idx[1]);
            const HeatShift<float> hs = load_heat_shift<2, true>(a, idx);
            float sp = 0.0f;
            static_for<0, QG>([&](auto A) { sp += sg[k][decltype(A)::value]; });
            float g[QG];
            g[0] = sg[k][0];
            static_for<1, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                g[al] = interp_pop<FL::c[al][0], FL::c[al][1], 0, false>(
                    lds + (al - 1) * N, (uint32_t)q, sg[k][al], st, stride);
            });
            node_walls_heat<2, true>(a, idx, sg[k], g, hs);
            const float Fp = lagged_force<true>(a.gbeta, fprev[k], sp, hs.tloc);
            float fp[QF], gp[QG];
            node_collide<2, float, float, true, MC>(a, (uint32_t)q, sf[k], g, Fp, hs.tloc, fp, gp);
            static_for<0, QF>([&](auto A) { sf[k][decltype(A)::value] = fp[decltype(A)::value]; });
            static_for<0, QG>([&](auto A) { sg[k][decltype(A)::value] = gp[decltype(A)::value]; });
        }
    };
    for (int s = 1; s < steps; s++) step(std::false_type{});
    step(std::true_type{});

#pragma unroll
    for (int k = 0; k < K; k++) {
        if (ij[k] < 0) continue;
        const int q = t + kResThreads * k;
        static_for<0, QF>([&](auto A) {
            constexpr int al = decltype(A)::value;
            stp(pfinal + al * a.plane, (uint32_t)q, sf[k][al]);
        });
        static_for<0, QG>([&](auto A) {
            constexpr int al = decltype(A)::value;
            stp(pfinal + (QF + al) * a.plane, (uint32_t)q, sg[k][al]);
        });
    }
}

}  // namespace lbm
