// lbm_step_f32.hip -- FP32 shifted-population instantiation of the fused
// step (the production path).  Built with default FMA contraction.
#include "lbm_internal.h"
#include "lbm_step_res.h"
#include "lbm_step_t2.h"
#include "lbm_step_ym.h"
#include "lbm_step_zm.h"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cmath>

namespace lbm {

template <int DIM, int WPE, bool NT>
static void launch_k(bool macro, dim3 grid, dim3 block, const StepArgs<float> &a,
                     hipStream_t s) {
    if (macro)
        hipLaunchKernelGGL((k_step<DIM, float, float, true, true, WPE, NT>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((k_step<DIM, float, float, true, false, WPE, NT>), grid, block, 0, s,
                           a);
}

template <int DIM, int WPE>
static void launch_k(bool macro, bool nt, dim3 grid, dim3 block, const StepArgs<float> &a,
                     hipStream_t s) {
    if (nt) launch_k<DIM, WPE, true>(macro, grid, block, a, s);
    else launch_k<DIM, WPE, false>(macro, grid, block, a, s);
}

// population bytes read and written by one FP32 step
static double step_bytes_f32(int dim, int n) {
    const double nodes = dim == 2 ? (double)n * n : (double)n * n * n;
    return nodes * (dim == 2 ? 14 : 26) * 4.0 * 2.0;
}

hipError_t launch_step_zm(bool macro, const StepArgs<float> &a0, const LaunchShape &ls,
                          hipStream_t s);

hipError_t launch_step_ym(bool macro, const StepArgs<float> &a0, const LaunchShape &ls,
                          hipStream_t s);

hipError_t launch_step_f32(int dim, bool macro, const StepArgs<float> &a0, const LaunchShape &ls,
                           hipStream_t s) {
    if (ls.variant == 1)
        return dim == 3 ? launch_step_zm(macro, a0, ls, s) : launch_step_ym(macro, a0, ls, s);
    StepArgs<float> a = a0;
    dim3 block(ls.bx, ls.by, 1);
    a.gx = (a.n + ls.bx - 1) / ls.bx;
    a.gy = (a.n + ls.by - 1) / ls.by;
    a.ntiles = a.gx * a.gy * (dim == 3 ? a.n : 1);
    a.remap = (ls.remap == 2 && a.gy % 8) ? 0 : ls.remap;
    dim3 grid(a.ntiles);
    const bool nt = stream_stores(step_bytes_f32(dim, a.n));
    if (dim == 2) launch_k<2, 0>(macro, nt, grid, block, a, s);
    else if (ls.occ == 5) launch_k<3, 5>(macro, nt, grid, block, a, s);
    else if (ls.occ == 6) launch_k<3, 6>(macro, nt, grid, block, a, s);
    else launch_k<3, 0>(macro, nt, grid, block, a, s);
    return hipGetLastError();
}

// Launch variant 1 (3D): LDS-staged z-march, one 64 x by column of nodes
// per workgroup over kc planes.  kc is chosen so that the launch has about
// 3072 workgroups (12 per CU; at 2 per CU resident, six rounds).  Measured
// at 256^3, 64 x 4: 0.724 ms per step at 2048 workgroups, 0.715 at 3072,
// 0.713 at 3584, 0.719 at 4096, 0.735 at 8192, 0.780 at 512 (longer
// groups re-read fewer window planes but balance worse).
int zm_planes_per_group(int n, int by) {
    const long cols = (long)((n + 63) / 64) * ((n + by - 1) / by);
    long nzc = (3072 + cols - 1) / cols;
    nzc = std::max(1L, std::min(nzc, (long)std::max(1, n / 4)));
    return (int)((n + nzc - 1) / nzc);
}

template <int BY>
static void launch_zm(bool macro, bool nt, dim3 grid, dim3 block, const StepArgs<float> &a, int kc,
                      int xz, int yz, int rim_wgs, hipStream_t s) {
    if (macro && nt)
        hipLaunchKernelGGL((k_step_zm<BY, true, true>), grid, block, 0, s, a, kc, xz, yz, rim_wgs);
    else if (macro)
        hipLaunchKernelGGL((k_step_zm<BY, true, false>), grid, block, 0, s, a, kc, xz, yz, rim_wgs);
    else if (nt)
        hipLaunchKernelGGL((k_step_zm<BY, false, true>), grid, block, 0, s, a, kc, xz, yz, rim_wgs);
    else
        hipLaunchKernelGGL((k_step_zm<BY, false, false>), grid, block, 0, s, a, kc, xz, yz,
                           rim_wgs);
}

// The region the z-march covers, i < xz and j < yz (k_step_zm's header).
void zm_region(int n, int by, int &xz, int &yz) {
    xz = (n % 64 == 1 && n > 64) ? n - 1 : n;
    yz = ((n - 1) % by == 0) ? n - 1 : n;
}

hipError_t launch_step_zm(bool macro, const StepArgs<float> &a0, const LaunchShape &ls,
                          hipStream_t s) {
    if (ls.bx != 64 || (ls.by != 2 && ls.by != 4 && ls.by != 6 && ls.by != 8))
        return hipErrorInvalidValue;
    StepArgs<float> a = a0;
    const int kc = zm_planes_per_group(a.n, ls.by);
    const int nzc = (a.n + kc - 1) / kc;
    int xz, yz;
    zm_region(a.n, ls.by, xz, yz);
    a.gx = (xz + 63) / 64;
    a.gy = (yz + ls.by - 1) / ls.by;
    // remap 2 splits the tile rows into 8 XCD bands (k_step_zm's zm_tile);
    // every XCD gets the largest band's share of workgroups
    a.remap = (ls.remap == 2 && a.gy < 8) ? 0 : ls.remap;
    a.ntiles = a.remap == 2 ? 8 * a.gx * ((a.gy + 7) / 8) * nzc : a.gx * a.gy * nzc;
    // the nodes outside the marched region: one per thread in the launch's
    // first workgroups, a multiple of 8 of them so that the march keeps its
    // workgroup-to-XCD map
    // the nodes outside the marched region (none when n % 4 == 0): a few
    // workgroups at the head of the launch, each thread taking several
    // nodes, and a multiple of 8 of them so that the march keeps its
    // workgroup-to-XCD map; up to 64, about 8 nodes per thread
    const long rim = ((long)(a.n - xz) * a.n + (long)(a.n - yz) * xz) * a.n;
    const long per_wg = 8L * 64 * ls.by;
    const int rim_wgs = (int)std::min<long>(64, ((rim + per_wg - 1) / per_wg + 7) / 8 * 8);
    dim3 block(64, ls.by, 1), grid(a.ntiles + rim_wgs);
    const bool nt = stream_stores(step_bytes_f32(3, a.n));
    if (ls.by == 2) launch_zm<2>(macro, nt, grid, block, a, kc, xz, yz, rim_wgs, s);
    else if (ls.by == 6) launch_zm<6>(macro, nt, grid, block, a, kc, xz, yz, rim_wgs, s);
    else if (ls.by == 8) launch_zm<8>(macro, nt, grid, block, a, kc, xz, yz, rim_wgs, s);
    else launch_zm<4>(macro, nt, grid, block, a, kc, xz, yz, rim_wgs, s);
    return hipGetLastError();
}

// Launch variant 2 (2D): two steps per pass on tiles of 60 x TY output
// nodes (TY = 8, 12 or 16), k_step_t2 / k_step_t2_batch; a one-step pass
// (the last step of an odd count) always writes the macroscopic fields.
bool t2_shape_ok(int bx, int by) { return bx == kT2Width && (by == 8 || by == 12 || by == 16); }

int t2_out_columns() { return kT2Out; }

template <int TY, int NSTEP, bool MACRO, bool NT>
static void launch_t2_kernel(dim3 grid, const StepArgs<float> *args, const int4 *tiles,
                             int parity, const StepArgs<float> &a, hipStream_t s) {
    const dim3 block(kT2Threads);
    if (tiles)
        hipLaunchKernelGGL((k_step_t2_batch<TY, NSTEP, MACRO, NT>), grid, block, 0, s, args,
                           tiles, parity);
    else
        hipLaunchKernelGGL((k_step_t2<TY, NSTEP, MACRO, NT>), grid, block, 0, s, a);
}

template <int TY>
static void launch_t2(int nstep, bool macro, bool nt, dim3 grid, const StepArgs<float> *args,
                      const int4 *tiles, int parity, const StepArgs<float> &a, hipStream_t s) {
    if (nstep == 1) {
        if (nt) launch_t2_kernel<TY, 1, true, true>(grid, args, tiles, parity, a, s);
        else launch_t2_kernel<TY, 1, true, false>(grid, args, tiles, parity, a, s);
    } else if (macro) {
        if (nt) launch_t2_kernel<TY, 2, true, true>(grid, args, tiles, parity, a, s);
        else launch_t2_kernel<TY, 2, true, false>(grid, args, tiles, parity, a, s);
    } else {
        if (nt) launch_t2_kernel<TY, 2, false, true>(grid, args, tiles, parity, a, s);
        else launch_t2_kernel<TY, 2, false, false>(grid, args, tiles, parity, a, s);
    }
}

static void launch_t2_ty(int ty, int nstep, bool macro, bool nt, dim3 grid,
                         const StepArgs<float> *args, const int4 *tiles, int parity,
                         const StepArgs<float> &a, hipStream_t s) {
    if (ty == 8) launch_t2<8>(nstep, macro, nt, grid, args, tiles, parity, a, s);
    else if (ty == 12) launch_t2<12>(nstep, macro, nt, grid, args, tiles, parity, a, s);
    else launch_t2<16>(nstep, macro, nt, grid, args, tiles, parity, a, s);
}

hipError_t launch_step2_f32(int nstep, bool macro, const StepArgs<float> &a0,
                            const LaunchShape &ls, hipStream_t s) {
    if (!t2_shape_ok(ls.bx, ls.by) || nstep < 1 || nstep > 2) return hipErrorInvalidValue;
    StepArgs<float> a = a0;
    a.gx = (a.n + kT2Out - 1) / kT2Out;
    a.gy = (a.n + ls.by - 1) / ls.by;
    a.ntiles = a.gx * a.gy;
    a.remap = ls.remap == 0 ? 0 : 1;
    const bool nt = stream_stores(step_bytes_f32(2, a.n));
    launch_t2_ty(ls.by, nstep, macro, nt, dim3(a.ntiles), nullptr, nullptr, 0, a, s);
    return hipGetLastError();
}

hipError_t launch_step2_batch_f32(int nstep, bool macro, const StepArgs<float> *args,
                                  const int4 *tiles, int ntiles, int parity, int ty, bool nt,
                                  hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    if (!t2_shape_ok(kT2Width, ty) || nstep < 1 || nstep > 2) return hipErrorInvalidValue;
    launch_t2_ty(ty, nstep, macro, nt, dim3(ntiles), args, tiles, parity, StepArgs<float>{}, s);
    return hipGetLastError();
}

// Launch variant 1 (2D): y-march, one wave per strip of 64 columns over
// kc rows; kc is chosen for about 4096 one-wave workgroups (16 per CU), but
// at least 8 rows, so that the three rows a chunk loads beyond its own
// stay a small share of its traffic.
int ym_rows_per_chunk(double strip_rows) {
    const long kc = (long)std::ceil(strip_rows / 4096.0);
    return (int)std::max(8L, kc);
}

hipError_t launch_step_ym(bool macro, const StepArgs<float> &a0, const LaunchShape &ls,
                          hipStream_t s) {
    if (ls.bx != 64 || ls.by != 1) return hipErrorInvalidValue;
    StepArgs<float> a = a0;
    a.gx = (a.n + 63) / 64;
    const int kc = std::min(a.n, ym_rows_per_chunk((double)a.gx * a.n));
    a.gy = (a.n + kc - 1) / kc;
    a.ntiles = a.gx * a.gy;
    a.remap = ls.remap == 0 ? 0 : 1;
    dim3 block(64, 1, 1), grid(a.ntiles);
    const bool nt = stream_stores(step_bytes_f32(2, a.n));
    if (macro && nt) hipLaunchKernelGGL((k_step_ym<true, true>), grid, block, 0, s, a, kc);
    else if (macro) hipLaunchKernelGGL((k_step_ym<true, false>), grid, block, 0, s, a, kc);
    else if (nt) hipLaunchKernelGGL((k_step_ym<false, true>), grid, block, 0, s, a, kc);
    else hipLaunchKernelGGL((k_step_ym<false, false>), grid, block, 0, s, a, kc);
    return hipGetLastError();
}

hipError_t launch_step_batch_ym_f32(bool macro, const StepArgs<float> *args, const int4 *tiles,
                                    int ntiles, int parity, int kc, bool nt, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    dim3 block(64, 1, 1), grid(ntiles);
    if (macro && nt)
        hipLaunchKernelGGL((k_step_ym_batch<true, true>), grid, block, 0, s, args, tiles, parity,
                           kc);
    else if (macro)
        hipLaunchKernelGGL((k_step_ym_batch<true, false>), grid, block, 0, s, args, tiles, parity,
                           kc);
    else if (nt)
        hipLaunchKernelGGL((k_step_ym_batch<false, true>), grid, block, 0, s, args, tiles, parity,
                           kc);
    else
        hipLaunchKernelGGL((k_step_ym_batch<false, false>), grid, block, 0, s, args, tiles,
                           parity, kc);
    return hipGetLastError();
}

template <bool MACRO, bool NT>
__global__ __launch_bounds__(1024) void k_step_batch(const StepArgs<float> *__restrict__ args,
                                                     const int4 *__restrict__ tiles, int parity) {
    const int4 t = tiles[blockIdx.x];
    const StepArgs<float> a = args[t.w ^ parity];
    const int i = t.y * (int)blockDim.x + (int)threadIdx.x;
    const int j = t.z * (int)blockDim.y + (int)threadIdx.y;
    if (i >= a.n || j >= a.n) return;
    step_node<2, float, float, true, MACRO, NT>(a, i, j, 0);
}

hipError_t launch_step_batch_f32(bool macro, const StepArgs<float> *args, const int4 *tiles,
                                 int ntiles, int parity, int bx, int by, bool nt, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    dim3 block(bx, by, 1), grid(ntiles);
    if (macro && nt)
        hipLaunchKernelGGL((k_step_batch<true, true>), grid, block, 0, s, args, tiles, parity);
    else if (macro)
        hipLaunchKernelGGL((k_step_batch<true, false>), grid, block, 0, s, args, tiles, parity);
    else if (nt)
        hipLaunchKernelGGL((k_step_batch<false, true>), grid, block, 0, s, args, tiles, parity);
    else
        hipLaunchKernelGGL((k_step_batch<false, false>), grid, block, 0, s, args, tiles, parity);
    return hipGetLastError();
}

// The batch's variant 3 (lbm_step_res.h): one workgroup per cavity, steps
// time steps per launch.  Nodes per thread are instantiated for K = 1, 2,
// 4, 6, 8, 9 and 10 (64^2 cavities take 8, 65^2 take 9, 70^2 take 10); the
// launch uses the smallest K that covers every listed cavity.
template <int K>
static void launch_res(const StepArgs<float> *args, const int4 *table, int ncav, int parity,
                       int steps, hipStream_t s) {
    hipLaunchKernelGGL((k_step_batch_res<K>), dim3(ncav), dim3(kResThreads), 0, s, args, table,
                       parity, steps);
}

int res_max_edge() { return kResMaxN; }

hipError_t launch_step_batch_res_f32(const StepArgs<float> *args, const int4 *table, int ncav,
                                     int parity, int steps, int max_n, hipStream_t s) {
    if (ncav <= 0 || steps <= 0) return hipSuccess;
    if (max_n < 3 || max_n > kResMaxN) return hipErrorInvalidValue;
    const int k = res_nodes_per_thread(max_n);
    if (k <= 1) launch_res<1>(args, table, ncav, parity, steps, s);
    else if (k <= 2) launch_res<2>(args, table, ncav, parity, steps, s);
    else if (k <= 4) launch_res<4>(args, table, ncav, parity, steps, s);
    else if (k <= 6) launch_res<6>(args, table, ncav, parity, steps, s);
    else if (k <= 8) launch_res<8>(args, table, ncav, parity, steps, s);
    else if (k <= 9) launch_res<9>(args, table, ncav, parity, steps, s);
    else launch_res<kResMaxK>(args, table, ncav, parity, steps, s);
    return hipGetLastError();
}

// ---------------------------------------------------------------------
// The batch's variant 2: one persistent launch advances every cavity of
// the batch through `steps` time steps.  Work items are (step, tile) pairs
// handed out by one atomic ticket counter, in the order
//   chunk of `depth` steps, MALL group, step, tile of the group,
// so every item a tile's update depends on (the tiles around it, one step
// earlier) holds a smaller ticket.  A workgroup takes a ticket, waits until
// the tiles within the stencil's reach have finished the previous step
// (per-tile step counters), updates its tile with the single-cavity node
// code and publishes its own counter.  Waits only ever target smaller
// tickets, which running workgroups hold, so the launch completes whatever
// the number of resident workgroups.  Every wait is bounded in time
// (pl.spin_ticks of the constant-rate wall clock): past it the workgroup
// raises the abort flag, every workgroup then stops taking tickets, and the
// host reports the error.
//
// Memory ordering across the XCDs (each with its own L2): the counters are
// polled with relaxed device-scope loads, which read the coherent value
// without cache maintenance, and one device-scope acquire fence follows the
// wait (on gfx950 it invalidates the L1 and L2 lines that may be stale).
// Polling with acquire loads would invalidate the XCD's L2 on every poll
// and starve the workgroups that do the work.  Before the release, every
// wave waits for its own stores (vmcnt(0)), so that the lead wave's
// device-scope release (an L2 write-back) covers the whole tile.
//
// Reach: a node's stencil spans its neighbours, and the one-sided wall
// stencils reach two nodes inward, so with tiles of at least 2 rows (and
// bx >= 32) the tiles within one tile in x and y cover it; 1-row tiles need
// two rows of tiles (ry = 2).  The same set covers the write-after-read
// hazard: a tile's step-s output overwrites the buffer that its neighbours
// read at step s - 1.

// s_waitcnt immediate for vmcnt(0) with expcnt and lgkmcnt left at their
// maxima (gfx9 encoding: vmcnt [3:0] and [15:14], expcnt [6:4], lgkmcnt
// [11:8]).
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ bool persist_decode(const PersistPlan &pl, long k, int &e, int &s) {
    const long T = pl.ntiles, DT = (long)pl.depth * T;
    const int nfull = pl.steps / pl.depth;
    long r;
    int chunk, d;
    if (k < (long)nfull * DT) {
        chunk = (int)(k / DT);
        r = k - (long)chunk * DT;
        d = pl.depth;
    } else {
        chunk = nfull;
        r = k - (long)nfull * DT;
        d = pl.steps - nfull * pl.depth;
        if (r >= (long)d * T) return false;
    }
    int g = 0;
    while (g + 1 < pl.ngroups && r >= (long)d * pl.groups[g + 1]) g++;
    const long g0 = pl.groups[g], ng = pl.groups[g + 1] - g0;
    const long rr = r - (long)d * g0;
    s = chunk * pl.depth + (int)(rr / ng) + 1;
    e = (int)(g0 + rr % ng);
    return true;
}

template <bool NT>
__global__ __launch_bounds__(1024) void k_step_batch_persist(
    const StepArgs<float> *__restrict__ args, PersistPlan pl) {
    __shared__ int item[2];
    const bool lead = threadIdx.x == 0 && threadIdx.y == 0;
    for (;;) {
        if (lead) {
            int e = -1, s = 0;
            if (!__hip_atomic_load(&pl.ctl[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                const long k = __hip_atomic_fetch_add(&pl.ctl[0], 1, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT);
                if (!persist_decode(pl, k, e, s)) e = -1;
            }
            if (e >= 0 && s > 1) {
                const int4 t = pl.tiles[e];
                const int4 ci = pl.cav[t.x];   // first entry, gx, gy
                const uint64_t t0 = wall_clock64();
                bool ok = true;
                for (int dy = -pl.ry; dy <= pl.ry && ok; dy++) {
                    const int ty = t.z + dy;
                    if (ty < 0 || ty >= ci.z) continue;
                    for (int dx = -1; dx <= 1 && ok; dx++) {
                        const int tx = t.y + dx;
                        if (tx < 0 || tx >= ci.y) continue;
                        const int *flag = pl.done + ci.x + ty * ci.y + tx;
                        while (__hip_atomic_load(flag, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) < s - 1) {
                            if (wall_clock64() - t0 > pl.spin_ticks ||
                                __hip_atomic_load(&pl.ctl[1], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)) {
                                __hip_atomic_store(&pl.ctl[1], 1, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                                ok = false;
                                break;
                            }
                            __builtin_amdgcn_s_sleep(2);
                        }
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (!ok) e = -1;
            }
            item[0] = e;
            item[1] = s;
        }
        __syncthreads();
        const int e = item[0], s = item[1];
        if (e < 0) break;
        const int4 t = pl.tiles[e];
        const StepArgs<float> a = args[t.w ^ ((pl.parity + s - 1) & 1)];
        const int i = t.y * (int)blockDim.x + (int)threadIdx.x;
        const int j = t.z * (int)blockDim.y + (int)threadIdx.y;
        if (i < a.n && j < a.n) {
            if (s == pl.steps) step_node<2, float, float, true, true, NT>(a, i, j, 0);
            else step_node<2, float, float, true, false, NT>(a, i, j, 0);
        }
        // every wave's stores of the tile have reached the L2 (vmcnt(0))
        // before the lead wave's release writes the L2 back
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __syncthreads();
        if (lead) __hip_atomic_store(pl.done + e, s, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void k_persist_reset(int *__restrict__ ctl, int *__restrict__ done, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) ctl[0] = 0;   // the ticket counter; the abort flag ctl[1] is sticky
    for (int k = t; k < n; k += gridDim.x * blockDim.x) done[k] = 0;
}

const void *step_batch_persist_kernel(bool nt) {
    return nt ? (const void *)k_step_batch_persist<true> : (const void *)k_step_batch_persist<false>;
}

hipError_t launch_step_batch_persist_f32(const StepArgs<float> *args, const PersistPlan &pl,
                                         int bx, int by, int nwg, bool nt, hipStream_t s) {
    if (pl.ntiles <= 0 || pl.steps <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_persist_reset, dim3(std::min(64, (pl.ntiles + 255) / 256)), dim3(256),
                       0, s, pl.ctl, pl.done, pl.ntiles);
    dim3 block(bx, by, 1), grid(nwg);
    if (nt) hipLaunchKernelGGL(k_step_batch_persist<true>, grid, block, 0, s, args, pl);
    else hipLaunchKernelGGL(k_step_batch_persist<false>, grid, block, 0, s, args, pl);
    return hipGetLastError();
}

// ---------------------------------------------------------------------
// The batch's variant 5: one cooperative launch runs every step of an
// lbm_batch_step call (up to the host's per-launch cap) over the tile table
// of every cavity.  The grid holds only workgroups that are resident at
// once (hipLaunchCooperativeKernel refuses a larger one).  Each workgroup
// strides over the tiles of the current MALL group and updates them with
// the single-cavity node code; the grid-wide barrier (cooperative groups'
// grid sync: an agent-scope release and acquire around a device-memory
// counter) orders every tile's step-s stores, across the XCDs' L2s, before
// any step-(s + 1) load, which also covers the write-after-read hazard
// (step s + 1 overwrites the buffer step s read).  Nothing waits on data,
// so the launch ends once every workgroup has passed every barrier.
//
// Barriers: one after each step that the same group's next step follows.
// When the launch moves on to the next MALL group, a group's next step
// comes after the other groups' chunks, and with depth >= 2 each of those
// holds a barrier of its own, so the move needs none; with depth 1 every
// step ends with one.
//
// A launch that writes the macroscopic fields runs its last step in a tail
// of its own, after one more barrier, rather than as a branch in the loop:
// with both node updates in one loop body the compiler shares code between
// them and fuses fewer products into FMAs, so the results would differ
// from variant 0's in the last bits.  Separate, each update compiles as in
// k_step_batch and the results equal variant 0's to the bit.
template <bool MACRO, bool NT>
__device__ __forceinline__ void coop_tiles(const StepArgs<float> *__restrict__ args,
                                           const int4 *__restrict__ tiles, int t0, int t1, int p) {
    for (int e = t0 + (int)blockIdx.x; e < t1; e += (int)gridDim.x) {
        const int4 t = tiles[e];
        const StepArgs<float> a = args[t.w ^ p];
        const int i = t.y * (int)blockDim.x + (int)threadIdx.x;
        const int j = t.z * (int)blockDim.y + (int)threadIdx.y;
        if (i < a.n && j < a.n) step_node<2, float, float, true, MACRO, NT>(a, i, j, 0);
    }
}

template <bool NT>
__global__ __launch_bounds__(1024) void k_step_batch_coop(const StepArgs<float> *__restrict__ args,
                                                          CoopPlan pl) {
    namespace cg = cooperative_groups;
    const cg::grid_group grid = cg::this_grid();
    if (!grid.is_valid()) return;   // not a cooperative launch: no barrier to meet
    // steps of the loop; the tail runs the last one when it writes macros
    const int nloop = pl.macro_last ? pl.steps - 1 : pl.steps;
    for (int c0 = 0; c0 < nloop; c0 += pl.depth) {
        const int d = min(pl.depth, nloop - c0);
        for (int g = 0; g < pl.ngroups; g++) {
            for (int s = 0; s < d; s++) {
                const int step = c0 + s;
                coop_tiles<false, NT>(args, pl.tiles, pl.groups[g], pl.groups[g + 1],
                                      pl.parity ^ (step & 1));
                const bool last = step == nloop - 1 && g == pl.ngroups - 1;
                const bool group_move = s == d - 1 && pl.ngroups > 1 && d >= 2;
                if (!last && !group_move) grid.sync();
            }
        }
    }
    if (pl.macro_last) {
        if (nloop > 0) grid.sync();
        const int p = pl.parity ^ ((pl.steps - 1) & 1);
        for (int g = 0; g < pl.ngroups; g++)
            coop_tiles<true, NT>(args, pl.tiles, pl.groups[g], pl.groups[g + 1], p);
    }
}

const void *step_batch_coop_kernel(bool nt) {
    return nt ? (const void *)k_step_batch_coop<true> : (const void *)k_step_batch_coop<false>;
}

hipError_t launch_step_batch_coop_f32(const StepArgs<float> *args, const CoopPlan &pl, int bx,
                                      int by, int nwg, bool nt, hipStream_t s) {
    if (pl.steps <= 0 || pl.ngroups <= 0 || nwg <= 0) return hipSuccess;
    if (pl.depth < 1) return hipErrorInvalidValue;
    CoopPlan plan = pl;
    const StepArgs<float> *a = args;
    void *kargs[] = {(void *)&a, (void *)&plan};
    return hipLaunchCooperativeKernel(step_batch_coop_kernel(nt), dim3(nwg), dim3(bx, by, 1),
                                      kargs, 0, s);
}

hipError_t launch_prologue_f32(int dim, const StepArgs<float> &a,
                               const StepArgs<double> &ad, hipStream_t s) {
    long nn = dim == 2 ? (long)a.n * a.n : (long)a.n * a.n * a.n;
    dim3 block(256);
    dim3 grid((unsigned)((nn + 255) / 256));
    if (dim == 2) hipLaunchKernelGGL((k_prologue<2, float, float, true>), grid, block, 0, s, a, ad);
    else hipLaunchKernelGGL((k_prologue<3, float, float, true>), grid, block, 0, s, a, ad);
    return hipGetLastError();
}

}  // namespace lbm
