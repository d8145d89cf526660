// lbm_internal.h -- state behind an lbm_handle and the launchers shared by
// the translation units of liblbm_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/lbm_capi.h"
#include "lbm_step.h"

namespace lbm {

// Block shape and tile order of the step kernel.
struct LaunchShape {
    int bx = 64, by = 4;
    int remap = 1;
    int occ = 0;           // 3D FP32: target waves per SIMD (0, 5 or 6)
    int variant = 0;       // FP32: 1 = LDS-staged z-march (3D, lbm_step_zm.h) or
                           // y-march (2D, lbm_step_ym.h); 2 = two steps per
                           // pass (2D, lbm_step_t2.h)
};

// Device-side state of one cavity.  Populations live in two ping-pong
// buffers of (Qf + Qg) SoA planes each; the FP64 macroscopic fields are
// written on the last step of every lbm_step call.
struct Ctx {
    lbm_params p{};
    int dim = 2, n = 0, qf = 9, qg = 5;
    int pitch = 0;              // population row pitch (elements): 3D FP32 storage
                                // pads rows to a multiple of 32 (128-byte lines);
                                // otherwise n
    long nn = 0, plane = 0;     // plane: population plane stride, pitch*n(*n)
                                // rounded up to 64 elements
    bool fp64 = false;          // FP64 storage and arithmetic
    double info[LBM_INFO_COUNT] = {};
    double dx0 = 0;
    std::vector<double> x;      // node coordinates (LB units), same on every axis
    std::vector<double> tref;   // conduction profile at each heated-axis index
    std::vector<int> st;        // 3-point stencils, 3 per grid index
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void *pbuf[2] = {nullptr, nullptr};
    int cur = 0;
    double *rho = nullptr, *u = nullptr, *v = nullptr, *w = nullptr, *T = nullptr;
    double *up = nullptr, *vp = nullptr, *wp = nullptr, *Tp = nullptr;
    double *force = nullptr;    // FP64 path only
    int2 *d_off = nullptr;
    void *d_wt = nullptr;       // weights in the storage type
    double *d_x = nullptr;      // node coordinates
    double *d_cvw = nullptr;    // control-volume widths / H
    float *d_tref = nullptr;    // FP32 conduction profile (shifted storage)
    float *d_tcorr = nullptr;   // its streaming correction, 2 per index
    double *d_part = nullptr;   // reduction partials
    double *d_red = nullptr;    // reduction results
    int red_blocks = 0;
    int64_t itc = 0;
    bool initialized = false;
    int64_t inits = 0;          // initialisations so far (init_rest/init_fields/
                                // load_state/copy_state), so a batch can tell a
                                // re-initialised cavity from one it retired
    LaunchShape shape{};
    StepArgs<float> a32{};
    StepArgs<double> a64{};
    // Small grids, where a launch costs more than the step: one check
    // interval of steps (gsteps) replayed as a HIP graph, graph[p] starting
    // from buffer p, captured for launch shape gshape.
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int64_t gsteps = 0;
    int gflip = 0;              // buffer parity change of one replay
    LaunchShape gshape{};
};

// Fills the per-launch constants of both argument blocks (pointers to the
// current buffers are set by the caller).
void fill_args(Ctx &c);

// Capacity of the memory-side last-level cache (MALL, Infinity Cache) of
// MI355X.  The next step reads what a step writes: while the written
// populations (half of a step's read + write traffic) fit in the MALL,
// ordinary stores leave them there for it; beyond that they only displace
// rows that are still to be read, and the step stores with streaming stores.
constexpr double kMallBytes = 256.0 * 1024 * 1024;
inline bool stream_stores(double bytes_per_step) { return 0.5 * bytes_per_step > kMallBytes; }

// Launchers (lbm_step_f32.hip / lbm_step_f64.hip).
hipError_t launch_step_f32(int dim, bool macro, const StepArgs<float> &a,
                           const LaunchShape &ls, hipStream_t s);
hipError_t launch_prologue_f32(int dim, const StepArgs<float> &a,
                               const StepArgs<double> &ad, hipStream_t s);
// Launch variant 2 of a 2D FP32 cavity: one pass of nstep (2, or 1 for the
// last step of an odd count) time steps (lbm_step_t2.h) from a.pin to
// a.pout, tiles of 60 x ls.by output nodes; macro: the pass's last step
// writes the FP64 fields (always, for a one-step pass).
hipError_t launch_step2_f32(int nstep, bool macro, const StepArgs<float> &a, const LaunchShape &ls,
                            hipStream_t s);
bool t2_shape_ok(int bx, int by);
// The batch's variant 4: one such pass over a table of (cavity, tile x,
// tile y, argument index at parity 0) entries, tiles of t2_out_columns() x
// ty output nodes; workgroup b uses args[tiles[b].w ^ parity].
hipError_t launch_step2_batch_f32(int nstep, bool macro, const StepArgs<float> *args,
                                  const int4 *tiles, int ntiles, int parity, int ty, bool nt,
                                  hipStream_t s);
int t2_out_columns();
hipError_t launch_step_f64(int dim, bool macro, const StepArgs<double> &a,
                           const LaunchShape &ls, hipStream_t s);
hipError_t launch_prologue_f64(int dim, const StepArgs<double> &a,
                               const StepArgs<double> &ad, hipStream_t s);

// Auxiliary kernels (lbm_aux.hip), all FP64.
hipError_t launch_init_rest(const Ctx &c);
hipError_t launch_residual(const Ctx &c);   // results: d_red[0..3]
// The same kernels with their partial sums and results in the library's
// static device scratch array of the current device instead of d_part and
// d_red; *sums receives the device address of the four sums.  Allocates
// nothing.
hipError_t launch_residual_scratch(const Ctx &c, double **sums);
// errU and errT from the four sums of launch_residual (lbm_check's
// arithmetic); LBM_ERR_DIVERGED if either is not finite.
int32_t residual_ratios(const double r[4], double *errU, double *errT);
hipError_t launch_nusselt(const Ctx &c);    // results: d_red[4..5]
// dst's current population buffer (and, for an FP64 dst, its force) from
// src's state, converted between the storages; same grid and parameters.
hipError_t launch_convert_state(const Ctx &dst, const Ctx &src);

// Batched step (lbm_step_f32.hip): one launch advances every cavity listed
// in the tile table.  tiles[b] = (cavity, tile x, tile y, argument index of
// the cavity at parity 0); workgroup b uses args[tiles[b].w ^ parity].  2D,
// FP32 shifted storage.
hipError_t launch_step_batch_f32(bool macro, const StepArgs<float> *args, const int4 *tiles,
                                 int ntiles, int parity, int bx, int by, bool nt, hipStream_t s);
// The batch's variant 1: the y-march over a table of (cavity, strip, chunk,
// argument index) entries, kc rows per chunk; cavity < 0 marks padding.
hipError_t launch_step_batch_ym_f32(bool macro, const StepArgs<float> *args, const int4 *tiles,
                                    int ntiles, int parity, int kc, bool nt, hipStream_t s);
// Rows per chunk of the y-march for a total of strip_rows (strips x rows).
int ym_rows_per_chunk(double strip_rows);

// The batch's variant 3: one launch of `steps` time steps of every cavity
// in table[0..ncav) (table[b].w ^ parity = argument index of cavity b reading
// its current buffer), one workgroup per cavity holding the cavity in
// registers and LDS (lbm_step_res.h); max_n = the largest edge, at most
// res_max_edge().
hipError_t launch_step_batch_res_f32(const StepArgs<float> *args, const int4 *table, int ncav,
                                     int parity, int steps, int max_n, hipStream_t s);
int res_max_edge();

// The batch's variant 2: one persistent launch of `steps` time steps over
// the variant-0 tile table (tiles[e] = (cavity, tile x, tile y, argument
// index at parity 0), each cavity's tiles contiguous, y-major), cut into
// MALL groups groups[0..ngroups] and chunks of `depth` steps.  cav[c] =
// (first table entry, gx, gy) of cavity c.  ctl[0] is the ticket counter,
// ctl[1] the sticky abort flag (a dependency wait timed out); done[e] the
// last step tile e completed in this launch.  parity is the buffer parity
// of the launch's first step.  A wait longer than spin_ticks aborts.
struct PersistPlan {
    const int4 *tiles;
    const int4 *cav;
    const int *groups;
    int ngroups, ntiles;
    int steps, depth;
    int parity;
    int ry;
    int *ctl;
    int *done;
    uint64_t spin_ticks;   // wait bound, in ticks of the device's wall clock
};
// Resets the ticket counter and the step counters, then launches nwg
// persistent workgroups of bx x by threads.
hipError_t launch_step_batch_persist_f32(const StepArgs<float> *args, const PersistPlan &pl,
                                         int bx, int by, int nwg, bool nt, hipStream_t s);
// The persistent kernel, for occupancy queries.
const void *step_batch_persist_kernel(bool nt);

// The batch's variant 5: one cooperative launch of `steps` time steps over
// the variant-0 tile table, cut into MALL groups groups[0..ngroups] and
// chunks of `depth` steps (each group runs a chunk's steps before the next
// group starts).  parity is the buffer parity of the launch's first step;
// the launch's last step writes the macroscopic fields when macro_last.
struct CoopPlan {
    const int4 *tiles;
    const int *groups;
    int ngroups;
    int steps, depth;
    int parity;
    int macro_last;
};
// Launches nwg workgroups of bx x by threads with hipLaunchCooperativeKernel
// (which refuses a grid that is not resident at once).
hipError_t launch_step_batch_coop_f32(const StepArgs<float> *args, const CoopPlan &pl, int bx,
                                      int by, int nwg, bool nt, hipStream_t s);
// The cooperative kernel, for occupancy queries.
const void *step_batch_coop_kernel(bool nt);

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 1024;

// Records the detail message returned by lbm_last_error() and returns code.
int32_t set_error(int32_t code, const std::string &msg);

}  // namespace lbm

struct lbm_ctx : public lbm::Ctx {};
