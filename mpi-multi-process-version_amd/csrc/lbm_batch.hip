// lbm_batch.hip -- batches of independent 2D cavities behind the C ABI
// (lbm_batch_* in include/lbm_capi.h).
//
// Each cavity is a full single-cavity context (grid, tables, buffers); the
// batch owns them and one HIP stream that all of them use.  A time step of
// the batch is one launch of k_step_batch over a table of tiles of every
// active cavity: workgroup b takes tile b, reads its cavity's step
// arguments from a device array (two per cavity, one per ping-pong parity)
// and runs the single-cavity node update on it.  Tiles are listed largest
// cavity first, so the big cavities start early and the small ones fill
// the tail of the launch.  Batches of small cavities (edge <= 70) can
// instead keep every cavity on the chip (variant 3, lbm_step_res.h): one
// workgroup per cavity holds its state in registers and LDS and runs all
// the steps of an lbm_batch_step call in one launch.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "lbm_internal.h"

using lbm::set_error;

struct lbm_batch_ctx {
    std::vector<lbm_handle> cav;
    // Cavities a step advances: every initialised cavity, except those
    // lbm_batch_run retired (converged, at itc_max or diverged) and that
    // have not been initialised again since (retired_at holds the cavity's
    // initialisation count at retirement, -1 if not retired).
    std::vector<char> active;
    std::vector<int64_t> retired_at;
    std::vector<int> base_cur;   // cur of each cavity when the tile table was built
    int device = 0;
    int check_every = 2000;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    lbm::StepArgs<float> *d_args = nullptr;   // [2c + p]: cavity c reading buffer p
    int4 *d_tiles = nullptr;
    int ntiles = 0, tiles_cap = 0;
    int parity = 0;   // launches since the table was built, mod 2
    bool dirty = true;
    // default block (lbm_batch_set_launch and lbm_batch_autotune change
    // it): 64 x 8 beat 64 x 4, 64 x 2, 32 x 4 and 32 x 2 on batches of 256
    // cavities of 64^2 (24.4 against 26.6 us per step), 128 of 129^2
    // (73.8 against 76.9) and 64 of 257^2 (137.8 against 145.7; 128 x 4:
    // 132.4)
    int bx = 64, by = 8;
    int variant = 0;   // 0: one node per thread; 1: y-march strips; 2: persistent;
                       // 3: resident (one workgroup per cavity); 4: two steps
                       // per launch on 60 x by tiles; 5: cooperative (one
                       // launch per call, grid barriers between steps)
    int kc = 0;        // variant 1: rows per chunk
    // variant 2 (one persistent launch per lbm_batch_step call): per-cavity
    // (first table entry, gx, gy), the MALL group offsets, the ticket
    // counter and abort flag, the per-tile step counters, and the resident
    // workgroup count for the current block shape
    int4 *d_cav = nullptr;
    int *d_groups = nullptr, *d_ctl = nullptr, *d_done = nullptr;
    int groups_cap = 0, done_cap = 0;
    int *h_ctl = nullptr;   // pinned: the abort flag, read at synchronising calls
    int nwg = 0, nwg_bx = 0, nwg_by = 0;
    uint64_t spin_ticks = 0;   // 2 s of the device's wall clock
    // variant 5: resident workgroups of the cooperative kernel for the
    // block shape nwg_coop_bx x nwg_coop_by (and store mode)
    int nwg_coop = 0, nwg_coop_bx = 0, nwg_coop_by = 0, nwg_coop_nt = -1;
    double *h_red = nullptr;   // pinned host memory: 4 residual sums per cavity
    // variant 0: tile-table offsets of the MALL groups (whole cavities in
    // table order), groups[g] .. groups[g + 1]; one group when all fit
    std::vector<int> groups;
    std::vector<double> group_nodes;   // nodes of each group
};

namespace {

// Variant 2, the persistent launch (k_step_batch_persist), is switched off.
// Its first two GPU runs, inside lbm_batch_autotune, stopped making
// progress: first with acquire loads in the dependency polls (each poll
// then invalidates the XCD's L2), then again with relaxed polls, one
// acquire fence per wait and a 2 s wall-clock bound on every wait, which
// should have turned any wait that never ends into an error.  The cause is
// not established, so lbm_batch_set_launch refuses the variant and the
// autotuner skips it.
constexpr bool kPersistentEnabled = false;

// Variant 5, the cooperative launch (k_step_batch_coop), is switched off.
// It passed its tests on three GPU runs, then one run stopped making
// progress inside its MALL-group test, where a variant-0 batch on another
// stream still had several hundred launches queued when the cooperative
// launch was issued.  Every workgroup meets the same sequence of grid
// barriers and none waits on data, so the cause is not established;
// lbm_batch_set_launch refuses the variant and the autotuner skips it.
constexpr bool kCoopEnabled = false;

// Variant 5 runs at most this many steps per cooperative launch (a
// multiple of kGroupDepth, so that MALL-group chunks never straddle two
// launches except at the end of a call).
constexpr int64_t kCoopStepsPerLaunch = 4096;

// Variant 3 runs at most this many steps per launch (a launch then lasts
// tens of milliseconds even for the largest resident cavities).
constexpr int64_t kResStepsPerLaunch = 4096;

// bytes a 2D FP32 step moves per node: 14 populations read and written
constexpr double kStepBytesPerNode = 14 * 4.0 * 2.0;
// MALL groups: whole cavities whose step traffic stays within half the
// MALL, each run kGroupDepth steps before the next group starts.  Measured
// on 64 mixed cavities (6.44 M nodes, 128 x 4 blocks), per step: 0.212 ms
// ungrouped; 0.197 ms at half the MALL and depth 64 (0.198 at depth 32,
// 0.205 at depth 8); 0.205 ms with groups of the whole MALL (depth 8),
// 0.209 at three eighths (depth 64) and 0.239 at a quarter (depth 16).
constexpr double kGroupBytes = 0.5 * lbm::kMallBytes;
constexpr int64_t kGroupDepth = 64;

int32_t hip_err(hipError_t e, const char *where) {
    return set_error(LBM_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define BATCH_TRY(expr)                                      \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_err(e_, #expr);     \
    } while (0)

void release(lbm_batch b) {
    for (lbm_handle h : b->cav) {
        if (!h) continue;
        h->stream = nullptr;   // the batch's stream, destroyed below
        (void)lbm_destroy(h);
    }
    if (b->d_args) (void)hipFree(b->d_args);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    void *persist[] = {b->d_cav, b->d_groups, b->d_ctl, b->d_done};
    for (void *q : persist)
        if (q) (void)hipFree(q);
    if (b->h_ctl) (void)hipHostFree(b->h_ctl);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->h_red) (void)hipHostFree(b->h_red);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int32_t check_batch(lbm_batch b) {
    if (!b) return set_error(LBM_ERR_ARG, "null batch");
    BATCH_TRY(hipSetDevice(b->device));
    return LBM_OK;
}

bool steppable(const lbm_batch b, size_t c) {
    const lbm_ctx &h = *b->cav[c];
    return h.initialized && b->retired_at[c] != h.inits;
}

// The active set outside lbm_batch_run: every steppable cavity.
void refresh_active(lbm_batch b) {
    for (size_t c = 0; c < b->cav.size(); c++) {
        const char a = steppable(b, c) ? 1 : 0;
        if (a != b->active[c]) {
            b->active[c] = a;
            b->dirty = true;
        }
    }
}

int32_t need_active(lbm_batch b) {
    refresh_active(b);
    for (char a : b->active)
        if (a) return LBM_OK;
    return set_error(LBM_ERR_STATE, "no cavity to step: initialise one (lbm_batch_init_rest, or "
                                    "lbm_init_rest/lbm_init_fields/lbm_load_state on "
                                    "lbm_batch_cavity) or re-initialise a retired one");
}

// Variant 3 holds each cavity in one workgroup: every cavity's edge must be
// at most lbm::res_max_edge().
bool resident_fits(const lbm_batch b) {
    for (lbm_handle h : b->cav)
        if (h->n > lbm::res_max_edge()) return false;
    return true;
}

bool table_stale(const lbm_batch b) {
    if (b->dirty) return true;
    for (size_t c = 0; c < b->cav.size(); c++)
        if (b->active[c] && b->cav[c]->cur != (b->base_cur[c] ^ b->parity)) return true;
    return false;
}

// Variant 2's device tables for the tile table just built: each cavity's
// first entry and tile grid, the MALL group offsets, one step counter per
// tile, and (once) the ticket counter and the abort flag.
int32_t persist_tables(lbm_batch b, const std::vector<int4> &tiles) {
    const size_t nc = b->cav.size();
    std::vector<int4> cav(nc, make_int4(0, 0, 0, 0));
    for (size_t e = 0; e < tiles.size(); e++) {
        const int4 t = tiles[e];
        if (t.y == 0 && t.z == 0) {
            const int n = b->cav[t.x]->n;
            cav[t.x] = make_int4((int)e, (n + b->bx - 1) / b->bx, (n + b->by - 1) / b->by, 0);
        }
    }
    if (!b->d_cav) BATCH_TRY(hipMalloc(&b->d_cav, sizeof(int4) * nc));
    BATCH_TRY(hipMemcpy(b->d_cav, cav.data(), sizeof(int4) * nc, hipMemcpyHostToDevice));
    if ((int)b->groups.size() > b->groups_cap) {
        if (b->d_groups) BATCH_TRY(hipFree(b->d_groups));
        b->d_groups = nullptr;
        b->groups_cap = 0;
        BATCH_TRY(hipMalloc(&b->d_groups, sizeof(int) * b->groups.size()));
        b->groups_cap = (int)b->groups.size();
    }
    BATCH_TRY(hipMemcpy(b->d_groups, b->groups.data(), sizeof(int) * b->groups.size(),
                        hipMemcpyHostToDevice));
    if (b->ntiles > b->done_cap) {
        if (b->d_done) BATCH_TRY(hipFree(b->d_done));
        b->d_done = nullptr;
        b->done_cap = 0;
        BATCH_TRY(hipMalloc(&b->d_done, sizeof(int) * (size_t)b->ntiles));
        b->done_cap = b->ntiles;
    }
    if (!b->d_ctl) {
        BATCH_TRY(hipMalloc(&b->d_ctl, 2 * sizeof(int)));
        BATCH_TRY(hipMemset(b->d_ctl, 0, 2 * sizeof(int)));
    }
    return LBM_OK;
}

// Variant 5's device copy of the MALL group offsets of the tile table just
// built.
int32_t coop_tables(lbm_batch b) {
    if ((int)b->groups.size() > b->groups_cap) {
        if (b->d_groups) BATCH_TRY(hipFree(b->d_groups));
        b->d_groups = nullptr;
        b->groups_cap = 0;
        BATCH_TRY(hipMalloc(&b->d_groups, sizeof(int) * b->groups.size()));
        b->groups_cap = (int)b->groups.size();
    }
    BATCH_TRY(hipMemcpy(b->d_groups, b->groups.data(), sizeof(int) * b->groups.size(),
                        hipMemcpyHostToDevice));
    return LBM_OK;
}

// Workgroups of variant 5's cooperative launch: every workgroup of the
// current block shape that the device holds at once (the launch is refused
// beyond that), and no more than the largest MALL group has tiles.
int32_t coop_workgroups(lbm_batch b, bool nt) {
    if (b->nwg_coop <= 0 || b->nwg_coop_bx != b->bx || b->nwg_coop_by != b->by ||
        b->nwg_coop_nt != (int)nt) {
        int coop = 0;
        BATCH_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, b->device));
        if (!coop)
            return set_error(LBM_ERR_HIP, "variant 5 needs cooperative launches, which this "
                                          "device does not support");
        int per_cu = 0, cus = 0;
        BATCH_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, lbm::step_batch_coop_kernel(nt), b->bx * b->by, 0));
        BATCH_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
        if (per_cu < 1)
            return set_error(LBM_ERR_ARG, "variant 5: no workgroup of this block shape fits a "
                                          "compute unit");
        b->nwg_coop = per_cu * cus;
        b->nwg_coop_bx = b->bx;
        b->nwg_coop_by = b->by;
        b->nwg_coop_nt = (int)nt;
    }
    return LBM_OK;
}

// Resident workgroups of the persistent kernel for the current block
// shape, over all compute units.
int32_t persist_workgroups(lbm_batch b, bool nt) {
    if (b->nwg > 0 && b->nwg_bx == b->bx && b->nwg_by == b->by) return LBM_OK;
    int per_cu = 0, cus = 0;
    BATCH_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, lbm::step_batch_persist_kernel(nt), b->bx * b->by, 0));
    BATCH_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    int khz = 0;
    BATCH_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    // a wait lasts a few tile updates (microseconds); 2 s only ends a
    // launch that cannot make progress
    b->spin_ticks = (uint64_t)std::max(khz, 1000) * 2000ULL;
    b->nwg = std::max(1, per_cu) * std::max(1, cus);
    b->nwg_bx = b->bx;
    b->nwg_by = b->by;
    return LBM_OK;
}

// LBM_ERR_HIP if a persistent launch of this batch gave up on a dependency
// wait (the fields are then invalid).  Synchronises the batch's stream.
int32_t persist_status(lbm_batch b) {
    if (!b->d_ctl) return LBM_OK;
    BATCH_TRY(hipMemcpyAsync(b->h_ctl, b->d_ctl + 1, sizeof(int), hipMemcpyDeviceToHost,
                             b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
    if (*b->h_ctl)
        return set_error(LBM_ERR_HIP, "persistent batch launch: a tile's dependency wait timed "
                                      "out and the launch stopped; the fields are invalid");
    return LBM_OK;
}

// Lists the tiles of every active cavity, largest cavity first.  Variant 1
// lists (cavity, strip, chunk) entries and then deals them to the 8 XCDs in
// contiguous runs (workgroup b runs on XCD b mod 8), so that the strips
// whose edge columns a strip reads run on its XCD; the table is padded to
// a multiple of 8 with empty entries.
int32_t rebuild(lbm_batch b) {
    BATCH_TRY(hipStreamSynchronize(b->stream));   // the old table may still be read
    std::vector<int> order(b->cav.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return b->cav[x]->n > b->cav[y]->n; });
    std::vector<int4> tiles;
    b->groups.assign(1, 0);
    b->group_nodes.clear();
    double gnodes = 0.0;
    if (b->variant == 1) {
        double strip_rows = 0.0;
        for (int c : order)
            if (b->active[c]) strip_rows += (double)((b->cav[c]->n + 63) / 64) * b->cav[c]->n;
        b->kc = lbm::ym_rows_per_chunk(strip_rows);
    }
    for (int c : order) {
        if (!b->active[c]) continue;
        const lbm_ctx &h = *b->cav[c];
        b->base_cur[c] = h.cur;
        if (b->variant == 3) {
            tiles.push_back(make_int4(c, 0, 0, 2 * c + h.cur));
        } else if (b->variant == 1) {
            const int gx = (h.n + 63) / 64, nch = (h.n + b->kc - 1) / b->kc;
            for (int ch = 0; ch < nch; ch++)
                for (int tx = 0; tx < gx; tx++)
                    tiles.push_back(make_int4(c, tx, ch, 2 * c + h.cur));
        } else {
            // a new MALL group when this cavity's step would take the
            // group's step traffic (14 FP32 populations read and written)
            // past kGroupBytes
            const double nodes = (double)h.n * h.n;
            if (gnodes > 0.0 && (gnodes + nodes) * kStepBytesPerNode > kGroupBytes) {
                b->groups.push_back((int)tiles.size());
                b->group_nodes.push_back(gnodes);
                gnodes = 0.0;
            }
            gnodes += nodes;
            const int tw = b->variant == 4 ? lbm::t2_out_columns() : b->bx;
            const int gx = (h.n + tw - 1) / tw, gy = (h.n + b->by - 1) / b->by;
            for (int ty = 0; ty < gy; ty++)
                for (int tx = 0; tx < gx; tx++)
                    tiles.push_back(make_int4(c, tx, ty, 2 * c + h.cur));
        }
    }
    if (b->variant == 1 && !tiles.empty()) {
        const size_t per = (tiles.size() + 7) / 8;
        std::vector<int4> dealt(8 * per, make_int4(-1, 0, 0, 0));
        for (size_t L = 0; L < tiles.size(); L++) dealt[(L % per) * 8 + L / per] = tiles[L];
        tiles.swap(dealt);
    }
    if ((int)tiles.size() > b->tiles_cap) {
        if (b->d_tiles) BATCH_TRY(hipFree(b->d_tiles));
        b->d_tiles = nullptr;
        b->tiles_cap = 0;
        BATCH_TRY(hipMalloc(&b->d_tiles, sizeof(int4) * tiles.size()));
        b->tiles_cap = (int)tiles.size();
    }
    if (!tiles.empty())
        BATCH_TRY(hipMemcpy(b->d_tiles, tiles.data(), sizeof(int4) * tiles.size(),
                            hipMemcpyHostToDevice));
    b->ntiles = (int)tiles.size();
    b->groups.push_back(b->ntiles);
    b->group_nodes.push_back(gnodes);
    if (b->variant == 2) {
        int32_t rc = persist_tables(b, tiles);
        if (rc) return rc;
    }
    if (b->variant == 5) {
        int32_t rc = coop_tables(b);
        if (rc) return rc;
    }
    b->parity = 0;
    b->dirty = false;
    return LBM_OK;
}

int32_t batch_steps(lbm_batch b, int64_t nsteps) {
    if (nsteps <= 0) return LBM_OK;
    for (size_t c = 0; c < b->cav.size(); c++)
        if (b->active[c] && !b->cav[c]->initialized)
            return set_error(LBM_ERR_STATE, "cavity " + std::to_string(c) + " not initialised");
    if (table_stale(b)) {
        int32_t rc = rebuild(b);
        if (rc) return rc;
    }
    double nodes = 0.0;   // active nodes
    for (size_t c = 0; c < b->cav.size(); c++)
        if (b->active[c]) nodes += (double)b->cav[c]->nn;
    const int ng = (int)b->groups.size() - 1;
    if (b->variant == 3) {
        // every cavity in one workgroup for up to kResStepsPerLaunch steps
        // per launch; the launch's last step writes the macroscopic fields
        int max_n = 0;
        for (size_t c = 0; c < b->cav.size(); c++)
            if (b->active[c]) max_n = std::max(max_n, b->cav[c]->n);
        for (int64_t s0 = 0; s0 < nsteps; s0 += kResStepsPerLaunch) {
            const int64_t d = std::min(kResStepsPerLaunch, nsteps - s0);
            hipError_t e = lbm::launch_step_batch_res_f32(b->d_args, b->d_tiles, b->ntiles,
                                                          b->parity, (int)d, max_n, b->stream);
            if (e != hipSuccess) return hip_err(e, "resident batch launch");
            b->parity ^= (int)(d & 1);
        }
        for (size_t c = 0; c < b->cav.size(); c++) {
            if (!b->active[c]) continue;
            b->cav[c]->cur ^= (int)(nsteps & 1);
            b->cav[c]->itc += nsteps;
        }
        return LBM_OK;
    }
    if (b->variant == 4) {
        // passes of two steps (one launch and one buffer flip each), a
        // one-step pass for an odd count; MALL groups run up to
        // kGroupDepth steps each before the next group starts
        int64_t flips = 0;
        const int64_t depth = ng <= 1 ? nsteps : kGroupDepth;
        for (int64_t s0 = 0; s0 < nsteps; s0 += depth) {
            const int64_t d = std::min(depth, nsteps - s0);
            int64_t launches = 0;
            for (int g = 0; g < ng; g++) {
                const int t0 = b->groups[g], cnt = b->groups[g + 1] - t0;
                const bool nt = lbm::stream_stores(
                    (ng <= 1 ? nodes : b->group_nodes[g]) * kStepBytesPerNode);
                launches = 0;
                for (int64_t s = 0; s < d;) {
                    const int nstep = d - s >= 2 ? 2 : 1;
                    hipError_t e = lbm::launch_step2_batch_f32(
                        nstep, s0 + s + nstep == nsteps, b->d_args, b->d_tiles + t0, cnt,
                        b->parity ^ (int)(launches & 1), b->by, nt, b->stream);
                    if (e != hipSuccess) return hip_err(e, "two-step batch launch");
                    s += nstep;
                    launches++;
                }
            }
            b->parity ^= (int)(launches & 1);
            flips += launches;
        }
        for (size_t c = 0; c < b->cav.size(); c++) {
            if (!b->active[c]) continue;
            b->cav[c]->cur ^= (int)(flips & 1);
            b->cav[c]->itc += nsteps;
        }
        return LBM_OK;
    }
    if (b->variant == 2) {
        // one persistent launch for all nsteps (several only if the ticket
        // count would pass 2^31)
        const bool nt = ng <= 1 && lbm::stream_stores(nodes * kStepBytesPerNode);
        int32_t rc = persist_workgroups(b, nt);
        if (rc) return rc;
        const int64_t per = std::max<int64_t>(1, (INT32_MAX - 2LL * b->nwg) / std::max(1, b->ntiles));
        for (int64_t s0 = 0; s0 < nsteps; s0 += per) {
            const int64_t d = std::min(per, nsteps - s0);
            lbm::PersistPlan pl{b->d_tiles, b->d_cav, b->d_groups, ng, b->ntiles, (int)d,
                                ng > 1 ? (int)kGroupDepth : (int)d, b->parity,
                                b->by >= 2 ? 1 : 2, b->d_ctl, b->d_done, b->spin_ticks};
            const int64_t items = d * (int64_t)b->ntiles;
            const int nwg = (int)std::min<int64_t>(b->nwg, items);
            hipError_t e = lbm::launch_step_batch_persist_f32(b->d_args, pl, b->bx, b->by, nwg,
                                                              nt, b->stream);
            if (e != hipSuccess) return hip_err(e, "persistent batch launch");
            b->parity ^= (int)(d & 1);
        }
        for (size_t c = 0; c < b->cav.size(); c++) {
            if (!b->active[c]) continue;
            b->cav[c]->cur ^= (int)(nsteps & 1);
            b->cav[c]->itc += nsteps;
        }
        return LBM_OK;
    }
    if (b->variant == 5) {
        // one cooperative launch per kCoopStepsPerLaunch steps; MALL groups
        // run kGroupDepth-step chunks inside it, as the per-step launches do
        const bool nt = ng <= 1 && lbm::stream_stores(nodes * kStepBytesPerNode);
        int32_t rc = coop_workgroups(b, nt);
        if (rc) return rc;
        int max_group = 0;
        for (int g = 0; g < ng; g++) max_group = std::max(max_group, b->groups[g + 1] - b->groups[g]);
        const int nwg = std::min(b->nwg_coop, max_group);
        for (int64_t s0 = 0; s0 < nsteps; s0 += kCoopStepsPerLaunch) {
            const int64_t d = std::min(kCoopStepsPerLaunch, nsteps - s0);
            const lbm::CoopPlan pl{b->d_tiles, b->d_groups, ng, (int)d,
                                   ng > 1 ? (int)std::min(kGroupDepth, d) : (int)d, b->parity,
                                   s0 + d == nsteps ? 1 : 0};
            hipError_t e = lbm::launch_step_batch_coop_f32(b->d_args, pl, b->bx, b->by, nwg, nt,
                                                           b->stream);
            if (e != hipSuccess) return hip_err(e, "cooperative batch launch");
            b->parity ^= (int)(d & 1);
        }
        for (size_t c = 0; c < b->cav.size(); c++) {
            if (!b->active[c]) continue;
            b->cav[c]->cur ^= (int)(nsteps & 1);
            b->cav[c]->itc += nsteps;
        }
        return LBM_OK;
    }
    auto launch = [&](int g0, int g1, int parity, bool macro, bool nt) {
        const int t0 = b->groups[g0], cnt = b->groups[g1] - t0;
        return b->variant == 1
                   ? lbm::launch_step_batch_ym_f32(macro, b->d_args, b->d_tiles + t0, cnt,
                                                   parity, b->kc, nt, b->stream)
                   : lbm::launch_step_batch_f32(macro, b->d_args, b->d_tiles + t0, cnt, parity,
                                                b->bx, b->by, nt, b->stream);
    };
    if (ng <= 1) {
        const bool nt = lbm::stream_stores(nodes * kStepBytesPerNode);
        for (int64_t s = 0; s < nsteps; s++) {
            hipError_t e = launch(0, ng, b->parity, s == nsteps - 1, nt);
            if (e != hipSuccess) return hip_err(e, "batched step launch");
            b->parity ^= 1;
        }
    } else {
        // Cavities are independent, so each MALL group runs up to
        // kGroupDepth steps before the next group starts: its steps after
        // the first read what the previous one wrote from the MALL (stores
        // stay ordinary, so that they are kept there) instead of from HBM.
        // A group of one cavity too large for the MALL streams its stores.
        for (int64_t s0 = 0; s0 < nsteps; s0 += kGroupDepth) {
            const int64_t d = std::min(kGroupDepth, nsteps - s0);
            for (int g = 0; g < ng; g++) {
                const bool nt = lbm::stream_stores(b->group_nodes[g] * kStepBytesPerNode);
                for (int64_t s = 0; s < d; s++) {
                    hipError_t e = launch(g, g + 1, b->parity ^ (int)(s & 1),
                                          s0 + s == nsteps - 1, nt);
                    if (e != hipSuccess) return hip_err(e, "batched step launch");
                }
            }
            b->parity ^= (int)(d & 1);
        }
    }
    for (size_t c = 0; c < b->cav.size(); c++) {
        if (!b->active[c]) continue;
        b->cav[c]->cur ^= (int)(nsteps & 1);
        b->cav[c]->itc += nsteps;
    }
    return LBM_OK;
}

}  // namespace

extern "C" {

int32_t lbm_batch_create(const lbm_params *params, int32_t count, lbm_batch *out) {
    if (!out) return set_error(LBM_ERR_ARG, "null output");
    *out = nullptr;
    if (!params || count < 1) return set_error(LBM_ERR_ARG, "need at least one cavity");
    for (int32_t c = 0; c < count; c++) {
        const lbm_params &p = params[c];
        if (p.dim != 2 || p.precision != LBM_PREC_FP32_SHIFTED)
            return set_error(LBM_ERR_ARG, "batched cavities are 2D with FP32 shifted storage");
        if (p.device != params[0].device || p.check_every != params[0].check_every)
            return set_error(LBM_ERR_ARG, "batched cavities share the device and check_every");
    }
    lbm_batch b = new (std::nothrow) lbm_batch_ctx();
    if (!b) return set_error(LBM_ERR_NOMEM, "batch");
    b->device = params[0].device;
    b->check_every = params[0].check_every;
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e != hipSuccess) {
        release(b);
        return hip_err(e, "batch stream");
    }
    b->cav.assign(count, nullptr);
    for (int32_t c = 0; c < count; c++) {
        int32_t rc = lbm_create(&params[c], &b->cav[c]);
        if (rc) {
            const std::string msg = "cavity " + std::to_string(c) + ": " + lbm_last_error();
            release(b);
            return set_error(rc, msg);
        }
        // every cavity works on the batch's stream
        (void)hipStreamDestroy(b->cav[c]->stream);
        b->cav[c]->stream = b->stream;
    }
    b->active.assign(count, 0);
    b->retired_at.assign(count, -1);
    b->base_cur.assign(count, 0);
    e = hipHostMalloc((void **)&b->h_red, sizeof(double) * 4 * (size_t)count);
    if (e != hipSuccess) b->h_red = nullptr;
    else if ((e = hipHostMalloc((void **)&b->h_ctl, sizeof(int))) != hipSuccess) b->h_ctl = nullptr;
    if (e != hipSuccess) {
        release(b);
        return hip_err(e, "batch residual buffer");
    }
    std::vector<lbm::StepArgs<float>> args(2 * (size_t)count);
    for (int32_t c = 0; c < count; c++)
        for (int p = 0; p < 2; p++) {
            lbm::StepArgs<float> a = b->cav[c]->a32;
            a.pin = (const float *)b->cav[c]->pbuf[p];
            a.pout = (float *)b->cav[c]->pbuf[1 - p];
            args[2 * c + p] = a;
        }
    e = hipMalloc(&b->d_args, sizeof(lbm::StepArgs<float>) * args.size());
    if (e == hipSuccess)
        e = hipMemcpy(b->d_args, args.data(), sizeof(lbm::StepArgs<float>) * args.size(),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release(b);
        return hip_err(e, "batch arguments");
    }
    *out = b;
    return LBM_OK;
}

int32_t lbm_batch_destroy(lbm_batch b) {
    if (!b) return LBM_OK;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    release(b);
    return LBM_OK;
}

lbm_handle lbm_batch_cavity(lbm_batch b, int32_t index) {
    if (!b || index < 0 || index >= (int32_t)b->cav.size()) {
        set_error(LBM_ERR_ARG, "cavity index out of range");
        return nullptr;
    }
    return b->cav[index];
}

int32_t lbm_batch_init_rest(lbm_batch b) {
    int32_t rc = check_batch(b);
    if (rc) return rc;
    for (size_t c = 0; c < b->cav.size(); c++)
        if ((rc = lbm_init_rest(b->cav[c]))) return rc;
    refresh_active(b);
    return LBM_OK;
}

int32_t lbm_batch_step(lbm_batch b, int64_t nsteps) {
    int32_t rc = check_batch(b);
    if (rc) return rc;
    if (nsteps < 0) return set_error(LBM_ERR_ARG, "nsteps < 0");
    if (nsteps == 0) return LBM_OK;
    if ((rc = need_active(b))) return rc;
    return batch_steps(b, nsteps);
}

int32_t lbm_batch_check(lbm_batch b, double *errU, double *errT) {
    int32_t rc = check_batch(b);
    if (rc) return rc;
    for (size_t c = 0; c < b->cav.size(); c++) {
        double eu = 0.0, et = 0.0;
        rc = lbm_check(b->cav[c], &eu, &et);
        if (errU) errU[c] = eu;
        if (errT) errT[c] = et;
        if (rc && rc != LBM_ERR_DIVERGED) return rc;
    }
    return persist_status(b);
}

int32_t lbm_batch_nusselt(lbm_batch b, double *nu_hot, double *nu_cold) {
    int32_t rc = check_batch(b);
    if (rc) return rc;
    for (size_t c = 0; c < b->cav.size(); c++) {
        double h = NAN, k = NAN;
        rc = lbm_nusselt(b->cav[c], &h, &k);
        if (nu_hot) nu_hot[c] = h;
        if (nu_cold) nu_cold[c] = k;
        if (rc && rc != LBM_ERR_DIVERGED) return rc;
    }
    return persist_status(b);
}

int32_t lbm_batch_run(lbm_batch b, int64_t *itc, double *errU, double *errT, int32_t *status) {
    int32_t rc = check_batch(b);
    if (rc) return rc;
    const size_t nc = b->cav.size();
    const int64_t every = b->check_every;
    std::vector<double> eu(nc, 100.0), et(nc, 100.0);   // 2dim_mono.f90:148-149
    std::vector<int32_t> st(nc, LBM_ERR_NOT_CONVERGED);
    for (size_t c = 0; c < nc; c++) {
        if (!b->cav[c]->initialized) return set_error(LBM_ERR_STATE, "cavity not initialised");
        b->active[c] = b->cav[c]->itc < b->cav[c]->p.itc_max;
        b->retired_at[c] = -1;
    }
    b->dirty = true;
    for (;;) {
        int64_t n = INT64_MAX;
        for (size_t c = 0; c < nc; c++) {
            if (!b->active[c]) continue;
            const lbm_ctx &h = *b->cav[c];
            n = std::min<int64_t>(n, std::min<int64_t>(every - h.itc % every, h.p.itc_max - h.itc));
        }
        if (n == INT64_MAX) break;   // no active cavity left
        if ((rc = batch_steps(b, n))) return rc;
        // lbm_check of every cavity due for one, with one wait: the
        // residual kernels and the copies of their sums are queued on the
        // batch's stream, then read together
        bool due = false;
        for (size_t c = 0; c < nc; c++) {
            if (!b->active[c] || b->cav[c]->itc % every != 0) continue;
            hipError_t e = lbm::launch_residual(*b->cav[c]);
            if (e == hipSuccess)
                e = hipMemcpyAsync(b->h_red + 4 * c, b->cav[c]->d_red, 4 * sizeof(double),
                                   hipMemcpyDeviceToHost, b->stream);
            if (e != hipSuccess) return hip_err(e, "batch residual");
            due = true;
        }
        if (due) {
            BATCH_TRY(hipStreamSynchronize(b->stream));
            if ((rc = persist_status(b))) return rc;
        }
        bool changed = false;
        for (size_t c = 0; c < nc; c++) {
            if (!b->active[c]) continue;
            lbm_ctx &h = *b->cav[c];
            bool done = false;
            if (h.itc % every == 0) {
                rc = lbm::residual_ratios(b->h_red + 4 * c, &eu[c], &et[c]);
                if (rc == LBM_ERR_DIVERGED) {
                    st[c] = LBM_ERR_DIVERGED;
                    done = true;
                } else if (rc) {
                    return rc;
                } else if (eu[c] <= h.p.epsU && et[c] <= h.p.epsT) {
                    st[c] = LBM_OK;
                    done = true;
                }
            }
            if (!done && h.itc >= h.p.itc_max) done = true;   // st stays NOT_CONVERGED
            if (done) {
                b->active[c] = 0;
                changed = true;
            }
        }
        if (changed) b->dirty = true;
    }
    BATCH_TRY(hipStreamSynchronize(b->stream));
    if ((rc = persist_status(b))) return rc;
    // every cavity has left the run: later steps skip it until it is
    // initialised again
    for (size_t c = 0; c < nc; c++) b->retired_at[c] = b->cav[c]->inits;
    for (size_t c = 0; c < nc; c++) {
        if (itc) itc[c] = b->cav[c]->itc;
        if (errU) errU[c] = eu[c];
        if (errT) errT[c] = et[c];
        if (status) status[c] = st[c];
    }
    return LBM_OK;
}

int32_t lbm_batch_active(lbm_batch b, int32_t *cavities, int64_t *nodes) {
    if (!b) return set_error(LBM_ERR_ARG, "null batch");
    refresh_active(b);
    int32_t k = 0;
    int64_t m = 0;
    for (size_t c = 0; c < b->cav.size(); c++)
        if (b->active[c]) {
            k++;
            m += b->cav[c]->nn;
        }
    if (cavities) *cavities = k;
    if (nodes) *nodes = m;
    return LBM_OK;
}

int32_t lbm_batch_time_steps(lbm_batch b, int64_t nsteps, double *ms) {
    int32_t rc = check_batch(b);
    if (rc) return rc;
    if (nsteps < 1 || !ms) return set_error(LBM_ERR_ARG, "nsteps >= 1 and ms required");
    if ((rc = need_active(b))) return rc;
    if (table_stale(b) && (rc = rebuild(b))) return rc;   // the timed region only launches
    BATCH_TRY(hipEventRecord(b->ev0, b->stream));
    if ((rc = batch_steps(b, nsteps))) return rc;
    BATCH_TRY(hipEventRecord(b->ev1, b->stream));
    BATCH_TRY(hipEventSynchronize(b->ev1));
    if ((rc = persist_status(b))) return rc;
    float t = 0.0f;
    BATCH_TRY(hipEventElapsedTime(&t, b->ev0, b->ev1));
    if (ms) *ms = t;
    return LBM_OK;
}

int32_t lbm_batch_set_launch(lbm_batch b, const lbm_launch *l) {
    if (!b || !l) return set_error(LBM_ERR_ARG, "null argument");
    if (l->bx < 32 || l->bx % 32 || l->by < 1 || l->bx * l->by > 1024 || (l->bx * l->by) % 64 ||
        l->remap != 0 || l->occupancy != 0 || l->variant < 0 || l->variant > 5 ||
        (l->variant == 1 && (l->bx != 64 || l->by != 1)) ||
        (l->variant == 3 && (l->bx != 512 || l->by != 1)) ||
        (l->variant == 4 && !lbm::t2_shape_ok(l->bx, l->by)))
        return set_error(LBM_ERR_ARG, "invalid batched launch configuration (variant 1: bx = 64, "
                                      "by = 1; variant 3: bx = 512, by = 1; variant 4: bx = 64, "
                                      "by = 8, 12 or 16)");
    if (l->variant == 3 && !resident_fits(b))
        return set_error(LBM_ERR_ARG, "variant 3 (resident cavities) needs every cavity's edge "
                                      "to be at most " + std::to_string(lbm::res_max_edge()));
    if (l->variant == 5 && !kCoopEnabled)
        return set_error(LBM_ERR_ARG, "variant 5 (the cooperative launch) is switched off: one "
                                      "GPU run stopped making progress inside it and the cause "
                                      "is not established");
    if (l->variant == 2 && !kPersistentEnabled)
        return set_error(LBM_ERR_ARG, "variant 2 (the persistent launch) is switched off: its "
                                      "first GPU runs stopped making progress and the cause is "
                                      "not established");
    b->bx = l->bx;
    b->by = l->by;
    b->variant = l->variant;
    b->dirty = true;
    return LBM_OK;
}

int32_t lbm_batch_get_launch(lbm_batch b, lbm_launch *l) {
    if (!b || !l) return set_error(LBM_ERR_ARG, "null argument");
    *l = lbm_launch{b->bx, b->by, 0, 0, b->variant};
    return LBM_OK;
}

int32_t lbm_batch_autotune(lbm_batch b, int32_t steps_per_trial, double *best_ms) {
    int32_t rc = check_batch(b);
    if (rc) return rc;
    if (steps_per_trial < 1) return set_error(LBM_ERR_ARG, "steps_per_trial < 1");
    if ((rc = need_active(b))) return rc;
    const size_t nc = b->cav.size();
    // host snapshot of every initialised cavity
    std::vector<std::vector<char>> saved(nc);
    for (size_t c = 0; c < nc; c++) {
        if (!b->cav[c]->initialized) continue;
        int64_t bytes = 0;
        if ((rc = lbm_state_size(b->cav[c], &bytes))) return rc;
        saved[c].resize((size_t)bytes);
        if ((rc = lbm_save_state(b->cav[c], saved[c].data(), bytes))) return rc;
    }
    // (bx, by, variant); variant 1 is the y-march, variant 2 the
    // persistent launch, variant 3 the resident cavities, variant 4 two
    // steps per launch, variant 5 the cooperative launch
    static const int cand[][3] = {{64, 1, 0},  {64, 2, 0},  {64, 4, 0},  {64, 8, 0},
                                  {128, 1, 0}, {128, 2, 0}, {128, 4, 0}, {256, 1, 0},
                                  {256, 2, 0}, {32, 8, 0},  {64, 1, 1},  {64, 4, 2},
                                  {64, 8, 2},  {128, 4, 2}, {32, 8, 2},  {512, 1, 3},
                                  {64, 8, 4},  {64, 12, 4}, {64, 16, 4}, {64, 8, 5},
                                  {128, 4, 5}, {64, 4, 5}};
    const int bx0 = b->bx, by0 = b->by, v0 = b->variant;
    int best_bx = bx0, best_by = by0, best_v = v0;
    double best = 1e300;
    for (const auto &cd : cand) {
        if (cd[2] == 2 && !kPersistentEnabled) continue;
        if (cd[2] == 5 && !kCoopEnabled) continue;
        if (cd[2] == 3 && !resident_fits(b)) continue;
        b->bx = cd[0];
        b->by = cd[1];
        b->variant = cd[2];
        b->dirty = true;
        double ms = 0.0;
        if ((rc = batch_steps(b, 1)) || (rc = lbm_batch_time_steps(b, steps_per_trial, &ms)))
            break;
        if (ms / steps_per_trial < best) {
            best = ms / steps_per_trial;
            best_bx = cd[0];
            best_by = cd[1];
            best_v = cd[2];
        }
    }
    b->bx = rc ? bx0 : best_bx;
    b->by = rc ? by0 : best_by;
    b->variant = rc ? v0 : best_v;
    b->dirty = true;
    BATCH_TRY(hipStreamSynchronize(b->stream));
    for (size_t c = 0; c < nc; c++) {
        if (saved[c].empty()) continue;
        // restoring a state is not a new initialisation: a retired cavity
        // stays retired
        const bool retired = b->retired_at[c] == b->cav[c]->inits;
        int32_t r2 = lbm_load_state(b->cav[c], saved[c].data(), (int64_t)saved[c].size());
        if (r2 && !rc) rc = r2;
        if (retired) b->retired_at[c] = b->cav[c]->inits;
    }
    if (rc) return rc;
    if (best_ms) *best_ms = best;
    return LBM_OK;
}

}  // extern "C"
