// lbm_capi.hip -- host side of the C ABI declared in include/lbm_capi.h.
//
// lbm_create replaces the reference's `mesh` + parameter part of `initial`
// (natural-convection_2dim_mono.f90:74-174, 3dim_mono.f90:122-229): the grid,
// the relaxation rates and the interpolation weights are formed once on the
// host in FP64 and copied to the device.  lbm_step replaces the loop body
// (2dim_mono.f90:309-325) by one fused kernel launch per step; the MPI halo
// exchange and domain cut of the reference have no counterpart because the
// whole cavity is resident on one GPU.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_internal.h"

using lbm::Ctx;

namespace {

thread_local std::string g_err;

int32_t fail(int32_t code, const std::string &msg) {
    g_err = msg;
    return code;
}

int32_t hip_fail(hipError_t e, const char *where) {
    return fail(LBM_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

// mesh (2dim_mono.f90:83-116): X_i = (map(A(i/(N+1) - 1/2))/map(A/2) + 1)/2,
// dx0 = X_1 - X_0, length_LB = 1/dx0, dt = dx0*length_LB,
// x_i = (X_i - dx0/2)*length_LB for i = 1..N.
void build_grid(Ctx &c) {
    const int n = c.n;
    const double A = c.p.stretch;
    std::vector<double> X(n + 2);
    for (int i = 0; i <= n + 1; i++) {
        double t = A * ((double)i / (double)(n + 1) - 0.5);
        X[i] = c.p.grid_kind == LBM_GRID_TANH
                   ? 0.5 * (std::tanh(t) / std::tanh(0.5 * A) + 1.0)
                   : 0.5 * (std::erf(t) / std::erf(0.5 * A) + 1.0);
    }
    const double dx0 = X[1] - X[0];
    const double L = 1.0 / dx0;
    c.dx0 = dx0;
    c.info[LBM_INFO_LLB] = L;
    c.info[LBM_INFO_DT] = dx0 * L;
    c.x.resize(n);
    for (int i = 1; i <= n; i++) c.x[i - 1] = (X[i] - dx0 / 2.0) * L;
    // one-sided stencils at the walls (2dim_mono.f90:242-278)
    c.st.resize(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        int *t = &c.st[3 * (size_t)i];
        if (i == 0) { t[0] = 1; t[1] = 0; t[2] = 2; }
        else if (i == n - 1) { t[0] = n - 2; t[1] = n - 1; t[2] = n - 3; }
        else { t[0] = i - 1; t[1] = i; t[2] = i + 1; }
    }
}

// initial, parameter part (2dim_mono.f90:151-162, 3dim_mono.f90:206-229).
void build_params(Ctx &c) {
    const double L = c.info[LBM_INFO_LLB], Pr = c.p.Pr, Ra = c.p.Ra;
    const double nu = (c.p.Ma * (L - 1.0) * std::sqrt(Pr)) / std::sqrt(3.0 * Ra);
    const double kappa = nu / Pr;
    const double H = L - 1.0;
    const double gbeta = Ra * nu * kappa / (H * H * H);
    const double tauf = nu * 3.0 + 0.5;
    const double Snu = 1.0 / tauf;
    const double Sq = 8.0 * (2.0 - Snu) / (8.0 - Snu);
    const double sigk = c.dim == 2 ? 1.0 / (0.5 + 4.0 * (1.0 / Snu - 0.5) / (3.0 * Pr))
                                   : 1.0 / (0.5 + (1.0 / Snu - 0.5) / (0.5 * Pr));
    c.info[LBM_INFO_NU] = nu;
    c.info[LBM_INFO_KAPPA] = kappa;
    c.info[LBM_INFO_GBETA] = gbeta;
    c.info[LBM_INFO_TAUF] = tauf;
    c.info[LBM_INFO_SNU] = Snu;
    c.info[LBM_INFO_SQ] = Sq;
    c.info[LBM_INFO_SIGK] = sigk;
    c.info[LBM_INFO_H] = H;
}

// Quadratic Lagrange weights of 2dim_mono.f90:649-660 for a lattice
// velocity component cv along an axis, at grid index i.
void lagrange_weights(const Ctx &c, int i, int cv, double *wt) {
    const int *t = &c.st[3 * (size_t)i];
    const double d = (double)cv * c.info[LBM_INFO_DT];
    const double x0 = c.x[t[0]] + d, x1 = c.x[t[1]] + d, x2 = c.x[t[2]] + d;
    const double X = c.x[t[1]];
    wt[0] = ((X - x1) * (X - x2)) / ((x0 - x1) * (x0 - x2));
    wt[1] = ((X - x0) * (X - x2)) / ((x1 - x0) * (x1 - x2));
    wt[2] = ((X - x0) * (X - x1)) / ((x2 - x0) * (x2 - x1));
}

double cv_width(const Ctx &c, int j) {
    const int n = c.n;
    const double H = c.info[LBM_INFO_H];
    if (j == 0) return 0.5 * (c.x[0] + c.x[1]);
    if (j == n - 1) return H - 0.5 * (c.x[n - 2] + c.x[n - 1]);
    return 0.5 * (c.x[j + 1] - c.x[j - 1]);
}

template <typename A> void set_consts(const Ctx &c, A &a, bool shifted) {
    using R = typename A::real;
    const double Snu = c.info[LBM_INFO_SNU], Sq = c.info[LBM_INFO_SQ];
    const double dt = c.info[LBM_INFO_DT];
    a.n = c.n;
    a.pitch = c.pitch;
    a.plane = c.plane;
    a.dt = R(dt);
    a.half_dt = R(0.5 * dt);
    a.gbeta = R(c.info[LBM_INFO_GBETA]);
    if (shifted) {
        // conduction profile Tref(x) = Thot + (Tcold - Thot) x/H, stored in
        // FP32; the wall values are taken relative to the stored numbers
        const bool f32 = std::is_same<R, float>::value;
        const double t0 = f32 ? (double)(float)c.tref[0] : c.tref[0];
        const double t1 = f32 ? (double)(float)c.tref[c.n - 1] : c.tref[c.n - 1];
        a.tw_hot = R(c.p.Thot - t0);
        a.tw_cold = R(c.p.Tcold - t1);
    } else {
        a.tw_hot = R(c.p.Thot);
        a.tw_cold = R(c.p.Tcold);
    }
    a.s_nu = R(Snu);
    a.s_q = R(Sq);
    a.sig_k = R(c.info[LBM_INFO_SIGK]);
    a.q_e = R(c.dim == 2 ? 1.9 : 1.2);  // 2dim_mono.f90:540-541, 3dim_mono.f90:693-695
    const double hn = 1.0 - 0.5 * Snu, hq = 1.0 - 0.5 * Sq;
    a.c_e = R(6.0 * hn);
    a.c_eps = R(-6.0 * hn);
    a.c_q = R(-hq);
    a.c_pxx = R(2.0 * hn);
    a.c_pxy = R(hn);
    a.c3_e = R(38.0 * hn);
    a.c3_eps = R(-11.0 * hn);
    a.c3_q = R(-(2.0 / 3.0) * hq);
    a.c3_9 = R(2.0 * hn);
    a.c3_10 = R(-hn);
    a.c3_11 = R(2.0 * hn);
    a.c3_12 = R(-hn);
    a.c3_14 = R(hn);
    a.c3_15 = R(hn);
    a.rho = c.rho;
    a.u = c.u;
    a.v = c.v;
    a.w = c.w;
    a.T = c.T;
}

void free_ctx(Ctx *c) {
    void *ptrs[] = {c->pbuf[0], c->pbuf[1], c->rho, c->u, c->v, c->w, c->T, c->up, c->vp,
                    c->wp, c->Tp, c->force, c->d_off, c->d_wt, c->d_x, c->d_cvw, c->d_tref, c->d_tcorr,
                    c->d_part, c->d_red};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipGraphExec_t &g : c->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

int32_t set_device(const Ctx &c) {
    HIP_TRY(hipSetDevice(c.p.device));
    return LBM_OK;
}

// Points the argument blocks at the current ping-pong buffers.
void bind_buffers(Ctx &c) {
    if (c.fp64) {
        c.a64.pin = (const double *)c.pbuf[c.cur];
        c.a64.pout = (double *)c.pbuf[1 - c.cur];
    } else {
        c.a32.pin = (const float *)c.pbuf[c.cur];
        c.a32.pout = (float *)c.pbuf[1 - c.cur];
    }
}

int32_t run_prologue(Ctx &c) {
    // The prologue writes the first post-collision populations into the
    // current buffer.
    hipError_t e;
    if (c.fp64) {
        lbm::StepArgs<double> a = c.a64;
        a.pout = (double *)c.pbuf[c.cur];
        e = lbm::launch_prologue_f64(c.dim, a, c.a64, c.stream);
    } else {
        lbm::StepArgs<float> a = c.a32;
        a.pout = (float *)c.pbuf[c.cur];
        e = lbm::launch_prologue_f32(c.dim, a, c.a64, c.stream);
    }
    if (e != hipSuccess) return hip_fail(e, "prologue launch");
    const size_t bytes = sizeof(double) * (size_t)c.nn;
    HIP_TRY(hipMemsetAsync(c.up, 0, bytes, c.stream));
    HIP_TRY(hipMemsetAsync(c.vp, 0, bytes, c.stream));
    HIP_TRY(hipMemsetAsync(c.Tp, 0, bytes, c.stream));
    if (c.wp) HIP_TRY(hipMemsetAsync(c.wp, 0, bytes, c.stream));
    c.itc = 0;
    c.initialized = true;
    c.inits++;
    return LBM_OK;
}

// Launch variant 1 (LDS-staged z-march) in 3D: its staging offsets are
// 32-bit element offsets over the 26 padded planes (n <= 544); any
// n from 6 up (the nodes its tiles cannot hold are updated by a few
// workgroups of the same launch with the single-node code, see k_step_zm).
// Launch variant 1 in 2D (the y-march) loads per lane and has no alignment
// condition: FP32 storage, bx = 64 and by = 1 (one wave per workgroup).
bool zm_launch_ok(const lbm_ctx &c, int bx, int by, int occ) {
    if (c.fp64 || bx != 64 || occ != 0) return false;
    if (c.dim == 2) return by == 1;
    return (by == 2 || by == 4 || by == 6 || by == 8) && c.n >= 6 &&
           26.0 * (double)c.plane < 4294967296.0;
}

int32_t check_handle(lbm_handle h, bool need_init) {
    if (!h) return fail(LBM_ERR_ARG, "null handle");
    if (need_init && !h->initialized)
        return fail(LBM_ERR_STATE, "populations not initialised (call lbm_init_rest or lbm_init_fields)");
    return set_device(*h);
}

// Launch variant 2 of a 2D FP32 cavity: passes of two steps (each flips the
// buffer parity once), then a one-step pass for an odd count; the last step
// writes the macroscopic fields.
int32_t launch_steps_t2(Ctx &c, int64_t nsteps) {
    for (int64_t s = 0; s < nsteps;) {
        const int nstep = nsteps - s >= 2 ? 2 : 1;
        bind_buffers(c);
        hipError_t e = lbm::launch_step2_f32(nstep, s + nstep == nsteps, c.a32, c.shape, c.stream);
        if (e != hipSuccess) return hip_fail(e, "two-step launch");
        c.cur = 1 - c.cur;
        c.itc += nstep;
        s += nstep;
    }
    return LBM_OK;
}

int32_t launch_steps(Ctx &c, int64_t nsteps) {
    if (c.shape.variant == 2) return launch_steps_t2(c, nsteps);
    for (int64_t s = 0; s < nsteps; s++) {
        const bool macro = (s == nsteps - 1);
        bind_buffers(c);
        hipError_t e = c.fp64 ? lbm::launch_step_f64(c.dim, macro, c.a64, c.shape, c.stream)
                              : lbm::launch_step_f32(c.dim, macro, c.a32, c.shape, c.stream);
        if (e != hipSuccess) return hip_fail(e, "step launch");
        c.cur = 1 - c.cur;
        c.itc++;
    }
    return LBM_OK;
}

// Grids up to this many nodes replay a check interval as a HIP graph: their
// step kernel runs for a few microseconds, about what a launch costs.
constexpr long kGraphMaxNodes = 1L << 20;

bool same_shape(const lbm::LaunchShape &a, const lbm::LaunchShape &b) {
    return a.bx == b.bx && a.by == b.by && a.remap == b.remap && a.occ == b.occ &&
           a.variant == b.variant;
}

// nsteps steps on the context's stream (asynchronous).  A check interval of
// a small grid goes through a graph captured from the same launches (the
// graph holds only kernel launches; each buffer parity has its own graph);
// any other count is launched directly.
int32_t do_steps(Ctx &c, int64_t nsteps) {
    if (nsteps < 2 || nsteps != c.p.check_every || c.nn > kGraphMaxNodes)
        return launch_steps(c, nsteps);
    if (c.gsteps != nsteps || !same_shape(c.gshape, c.shape)) {
        for (hipGraphExec_t &g : c.graph)
            if (g) {
                HIP_TRY(hipGraphExecDestroy(g));
                g = nullptr;
            }
        c.gsteps = nsteps;
        c.gshape = c.shape;
    }
    hipGraphExec_t &g = c.graph[c.cur];
    if (!g) {
        const int cur0 = c.cur;
        const int64_t itc0 = c.itc;
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
        const int32_t rc = launch_steps(c, nsteps);
        const hipError_t e = hipStreamEndCapture(c.stream, &graph);
        c.gflip = c.cur ^ cur0;
        c.cur = cur0;
        c.itc = itc0;
        if (rc || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc ? rc : hip_fail(e, "step graph capture");
        }
        const hipError_t ei = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            g = nullptr;
            return hip_fail(ei, "step graph instantiate");
        }
    }
    HIP_TRY(hipGraphLaunch(g, c.stream));
    c.cur ^= c.gflip;
    c.itc += nsteps;
    return LBM_OK;
}

}  // namespace

namespace lbm {
int32_t set_error(int32_t code, const std::string &msg) { return fail(code, msg); }
}  // namespace lbm

namespace lbm {

int32_t residual_ratios(const double r[4], double *errU, double *errT) {
    auto ratio = [](double num, double den) {
        if (num == 0.0 && den == 0.0) return 0.0;
        return std::sqrt(num) / std::sqrt(den);
    };
    const double eu = ratio(r[0], r[1]), et = ratio(r[2], r[3]);
    if (errU) *errU = eu;
    if (errT) *errT = et;
    if (!std::isfinite(eu) || !std::isfinite(et))
        return set_error(LBM_ERR_DIVERGED, "non-finite residual");
    return LBM_OK;
}

}  // namespace lbm

extern "C" {

int32_t lbm_create(const lbm_params *params, lbm_handle *out) {
    if (!params || !out) return fail(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    const lbm_params &p = *params;
    if (p.dim != 2 && p.dim != 3) return fail(LBM_ERR_ARG, "dim must be 2 or 3");
    if (p.nx < 5 || p.ny != p.nx || (p.dim == 3 && p.nz != p.nx))
        return fail(LBM_ERR_ARG, "grid must be square/cubic with nx >= 5");
    if (p.nx > 4096 || (p.dim == 3 && p.nx > 1024))
        return fail(LBM_ERR_ARG, "grid too large");
    {
        // kernels address a population plane with 32-bit byte offsets
        const double esz = p.precision == LBM_PREC_FP64 ? 8.0 : 4.0;
        const double pitch = (p.dim == 3 && p.precision != LBM_PREC_FP64)
                                 ? (double)((p.nx + 31) / 32 * 32) : (double)p.nx;
        const double nodes = p.dim == 2 ? pitch * p.nx : pitch * p.nx * p.nx;
        if (nodes * esz >= 4294967296.0)
            return fail(LBM_ERR_ARG, "grid too large for one population plane");
    }
    if (!(p.Ra > 0) || !(p.Pr > 0) || !(p.Ma > 0) || !(p.stretch > 0) ||
        !std::isfinite(p.Ra) || !std::isfinite(p.Pr) || !std::isfinite(p.Ma) ||
        !std::isfinite(p.stretch) || !std::isfinite(p.Thot) || !std::isfinite(p.Tcold) ||
        p.Thot == p.Tcold)
        return fail(LBM_ERR_ARG, "Ra, Pr, Ma, stretch must be positive and finite; Thot != Tcold");
    if (p.grid_kind != LBM_GRID_ERF && p.grid_kind != LBM_GRID_TANH)
        return fail(LBM_ERR_ARG, "unknown grid_kind");
    if (p.precision != LBM_PREC_FP32_SHIFTED && p.precision != LBM_PREC_FP64)
        return fail(LBM_ERR_ARG, "unknown precision");
    if (p.check_every < 1 || p.itc_max < 0) return fail(LBM_ERR_ARG, "check_every/itc_max");
    if (p.device < 0)
        return fail(LBM_ERR_ARG, "device must be a HIP device ordinal: this library has no CPU "
                                 "backend");

    lbm_ctx *c = new (std::nothrow) lbm_ctx();
    if (!c) return fail(LBM_ERR_NOMEM, "host allocation");
    c->p = p;
    c->dim = p.dim;
    c->n = p.nx;
    c->qf = p.dim == 2 ? 9 : 19;
    c->qg = p.dim == 2 ? 5 : 7;
    c->nn = p.dim == 2 ? (long)c->n * c->n : (long)c->n * c->n * c->n;
    // Default launch shape (lbm_set_launch and lbm_autotune change it).
    // Small 2D grids run a step in a few microseconds, and narrower blocks
    // shorten it: 32 x 2 is at or near the best of 64x4, 32x4, 32x2, 64x2
    // and 32x8 at 64^2 to 257^2 in both precisions (e.g. 129^2: FP64 5.15
    // against 5.69 us, FP32 3.13 against 3.37 us per step); at 513^2 the
    // shapes are equal.
    if (p.dim == 2 && c->n <= 384) {
        c->shape.bx = 32;
        c->shape.by = 2;
    }
    c->fp64 = p.precision == LBM_PREC_FP64;
    // Larger 2D FP32 grids: two steps per launch on 60 x 8 tiles (variant
    // 2).  Per step, with check intervals replayed as graphs: 513^2 7.21
    // against 8.14 us, 1025^2 17.0 against 21.6 us; 2049^2 72 against
    // 127 us (y-march) and 147 us (plain).  Up to 257^2 a pass has too few
    // tiles, and one-node-per-thread steps are faster (129^2: 3.1 against
    // 5.4 us).
    if (p.dim == 2 && c->n > 384 && !c->fp64) {
        c->shape.bx = 64;
        c->shape.by = 8;
        c->shape.remap = 1;
        c->shape.variant = 2;
    }
    // 3D FP32 population rows start on 128-byte lines (pitch a multiple of
    // 32 elements): the step kernels' row loads and streaming stores then
    // cover whole lines whatever n is
    c->pitch = (c->dim == 3 && !c->fp64) ? (c->n + 31) / 32 * 32 : c->n;
    c->plane = ((long)c->pitch * c->n * (c->dim == 3 ? c->n : 1) + 63) / 64 * 64;
    // 3D FP32: the LDS z-march (variant 1, 64 x 4, XCD tile bands) wherever
    // its preconditions hold (those of lbm_set_launch); at 256^3 it takes
    // 0.72 ms per step against 1.10 ms for the plain kernel
    if (zm_launch_ok(*c, 64, 4, 0) && c->dim == 3) {
        c->shape.bx = 64;
        c->shape.by = 4;
        c->shape.remap = 2;
        c->shape.variant = 1;
    }
    build_grid(*c);
    build_params(*c);
    const double tauf = c->info[LBM_INFO_TAUF];
    if (!(tauf > 0.5 + 1e-6) || !std::isfinite(tauf)) {
        delete c;
        return fail(LBM_ERR_ARG, "tau_f must exceed 1/2 (viscosity too small for this grid)");
    }

    int32_t rc = set_device(*c);
    if (rc) { delete c; return rc; }
    auto cleanup = [&](int32_t code) { free_ctx(c); delete c; return code; };
    hipError_t e;
#define ALLOC(ptr, bytes)                                                           \
    do {                                                                            \
        e = hipMalloc((void **)&(ptr), (bytes));                                    \
        if (e != hipSuccess) return cleanup(fail(LBM_ERR_NOMEM, std::string("hipMalloc ") + #ptr + ": " + hipGetErrorString(e))); \
    } while (0)
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return cleanup(hip_fail(e, "hipStreamCreate"));
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess || (e = hipEventCreate(&c->ev1)) != hipSuccess)
        return cleanup(hip_fail(e, "hipEventCreate"));
    const size_t esz = c->fp64 ? sizeof(double) : sizeof(float);
    const size_t pbytes = esz * (size_t)c->plane * (size_t)(c->qf + c->qg);
    ALLOC(c->pbuf[0], pbytes);
    ALLOC(c->pbuf[1], pbytes);
    // the padding between population planes is never written by a step:
    // zero it in both buffers so that a saved state does not depend on
    // which buffer holds the current populations
    for (void *b : c->pbuf)
        if ((e = hipMemset(b, 0, pbytes)) != hipSuccess)
            return cleanup(hip_fail(e, "population buffer clear"));
    const size_t fbytes = sizeof(double) * (size_t)c->nn;
    ALLOC(c->d_tref, sizeof(float) * (size_t)c->n);
    ALLOC(c->d_tcorr, sizeof(float) * 2 * (size_t)c->n);
    ALLOC(c->rho, fbytes);
    ALLOC(c->u, fbytes);
    ALLOC(c->v, fbytes);
    ALLOC(c->T, fbytes);
    ALLOC(c->up, fbytes);
    ALLOC(c->vp, fbytes);
    ALLOC(c->Tp, fbytes);
    if (c->dim == 3) {
        ALLOC(c->w, fbytes);
        ALLOC(c->wp, fbytes);
    }
    if (c->fp64) ALLOC(c->force, fbytes);
    ALLOC(c->d_off, sizeof(int2) * (size_t)c->n);
    ALLOC(c->d_wt, (c->fp64 ? sizeof(double) : sizeof(float)) * 6 * (size_t)c->n);
    ALLOC(c->d_x, sizeof(double) * (size_t)c->n);
    ALLOC(c->d_cvw, sizeof(double) * (size_t)c->n);
    long want = (c->nn + lbm::kRedThreads - 1) / lbm::kRedThreads;
    c->red_blocks = (int)std::min<long>(std::max<long>(want, 1), lbm::kRedMaxBlocks);
    ALLOC(c->d_part, sizeof(double) * 4 * (size_t)c->red_blocks);
    ALLOC(c->d_red, sizeof(double) * 8);
#undef ALLOC

    // stencil offsets and weights
    std::vector<int2> off(c->n);
    std::vector<double> wt(6 * (size_t)c->n);
    for (int i = 0; i < c->n; i++) {
        const int *t = &c->st[3 * (size_t)i];
        off[i] = make_int2(t[0] - i, t[2] - i);
        lagrange_weights(*c, i, +1, &wt[6 * (size_t)i]);
        lagrange_weights(*c, i, -1, &wt[6 * (size_t)i + 3]);
    }
    c->tref.resize(c->n);
    for (int i = 0; i < c->n; i++)
        c->tref[i] = p.Thot + (p.Tcold - p.Thot) * c->x[i] / c->info[LBM_INFO_H];
    std::vector<float> treff(c->tref.begin(), c->tref.end());
    // Streaming of the stored profile: with the FP32 weights W0, W2 used by
    // the kernel's difference-form interpolation, the shift w*Tref streams
    // to w*(Tref_i + W0 (Tref_0 - Tref_i) + W2 (Tref_2 - Tref_i)); the
    // bracket minus Tref_i is formed here in FP64 from the stored values.
    std::vector<float> tcorr(2 * (size_t)c->n);
    for (int i = 0; i < c->n; i++) {
        const int *t = &c->st[3 * (size_t)i];
        const double ti = treff[i], t0 = treff[t[0]], t2 = treff[t[2]];
        for (int s = 0; s < 2; s++) {
            const double w0 = (double)(float)wt[6 * (size_t)i + 3 * s];
            const double w2 = (double)(float)wt[6 * (size_t)i + 3 * s + 2];
            tcorr[2 * (size_t)i + s] = (float)(w0 * (t0 - ti) + w2 * (t2 - ti));
        }
    }
    if ((e = hipMemcpy(c->d_tref, treff.data(), sizeof(float) * c->n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_tcorr, tcorr.data(), sizeof(float) * tcorr.size(),
                       hipMemcpyHostToDevice)))
        return cleanup(hip_fail(e, "tref upload"));
    std::vector<double> cvw(c->n);
    for (int i = 0; i < c->n; i++) cvw[i] = cv_width(*c, i) / c->info[LBM_INFO_H];
    if ((e = hipMemcpy(c->d_off, off.data(), sizeof(int2) * c->n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_x, c->x.data(), sizeof(double) * c->n, hipMemcpyHostToDevice)) ||
        (e = hipMemcpy(c->d_cvw, cvw.data(), sizeof(double) * c->n, hipMemcpyHostToDevice)))
        return cleanup(hip_fail(e, "table upload"));
    if (c->fp64) {
        e = hipMemcpy(c->d_wt, wt.data(), sizeof(double) * wt.size(), hipMemcpyHostToDevice);
    } else {
        std::vector<float> wf(wt.begin(), wt.end());
        e = hipMemcpy(c->d_wt, wf.data(), sizeof(float) * wf.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return cleanup(hip_fail(e, "weight upload"));

    set_consts(*c, c->a64, false);
    c->a64.off = c->d_off;
    c->a64.wt = (const double *)(c->fp64 ? c->d_wt : nullptr);
    c->a64.force = c->force;
    set_consts(*c, c->a32, true);
    c->a32.off = c->d_off;
    c->a32.wt = (const float *)(c->fp64 ? nullptr : c->d_wt);
    c->a32.force = nullptr;
    c->a32.tref = c->d_tref;
    c->a32.tcorr = c->d_tcorr;
    c->a64.tcorr = nullptr;
    c->a64.tref = nullptr;

    *out = c;
    g_err.clear();
    return LBM_OK;
}

int32_t lbm_destroy(lbm_handle h) {
    if (!h) return LBM_OK;
    (void)hipSetDevice(h->p.device);
    (void)hipStreamSynchronize(h->stream);
    free_ctx(h);
    delete h;
    return LBM_OK;
}

int32_t lbm_get_grid(lbm_handle h, double *x, double *y, double *z) {
    if (!h) return fail(LBM_ERR_ARG, "null handle");
    double *outs[3] = {x, y, z};
    for (int a = 0; a < 3; a++)
        if (outs[a] && (a < h->dim)) std::memcpy(outs[a], h->x.data(), sizeof(double) * h->n);
    return LBM_OK;
}

int32_t lbm_get_info(lbm_handle h, double *info) {
    if (!h || !info) return fail(LBM_ERR_ARG, "null argument");
    std::memcpy(info, h->info, sizeof(h->info));
    return LBM_OK;
}

int32_t lbm_init_rest(lbm_handle h) {
    int32_t rc = check_handle(h, false);
    if (rc) return rc;
    hipError_t e = lbm::launch_init_rest(*h);
    if (e != hipSuccess) return hip_fail(e, "init launch");
    return run_prologue(*h);
}

int32_t lbm_init_fields(lbm_handle h, const double *u, const double *v, const double *w,
                        const double *T) {
    int32_t rc = check_handle(h, false);
    if (rc) return rc;
    if (!T) return fail(LBM_ERR_ARG, "T must not be NULL");
    const size_t bytes = sizeof(double) * (size_t)h->nn;
    std::vector<double> ones(h->nn, 1.0);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(h->rho, ones.data(), bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->T, T, bytes, hipMemcpyHostToDevice));
    if (u) HIP_TRY(hipMemcpy(h->u, u, bytes, hipMemcpyHostToDevice));
    else HIP_TRY(hipMemset(h->u, 0, bytes));
    if (v) HIP_TRY(hipMemcpy(h->v, v, bytes, hipMemcpyHostToDevice));
    else HIP_TRY(hipMemset(h->v, 0, bytes));
    if (h->w) {
        if (w) HIP_TRY(hipMemcpy(h->w, w, bytes, hipMemcpyHostToDevice));
        else HIP_TRY(hipMemset(h->w, 0, bytes));
    }
    return run_prologue(*h);
}

int32_t lbm_step(lbm_handle h, int64_t nsteps) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    if (nsteps < 0) return fail(LBM_ERR_ARG, "nsteps < 0");
    return do_steps(*h, nsteps);
}

int32_t lbm_check(lbm_handle h, double *errU, double *errT) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    hipError_t e = lbm::launch_residual(*h);
    if (e != hipSuccess) return hip_fail(e, "residual launch");
    double r[4];
    HIP_TRY(hipMemcpyAsync(r, h->d_red, sizeof r, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return lbm::residual_ratios(r, errU, errT);
}

int32_t lbm_check_noalloc(lbm_handle h, double *errU, double *errT) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    if (4 * (long)h->red_blocks + 4 > 4 * (long)lbm::kRedMaxBlocks + 4)
        return fail(LBM_ERR_ARG, "lbm_check_noalloc: more reduction blocks than its scratch "
                                 "array holds");
    double *sums = nullptr;
    hipError_t e = lbm::launch_residual_scratch(*h, &sums);
    if (e != hipSuccess) return hip_fail(e, "residual launch (static scratch)");
    double r[4];
    HIP_TRY(hipMemcpyAsync(r, sums, sizeof r, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return lbm::residual_ratios(r, errU, errT);
}

int32_t lbm_nusselt(lbm_handle h, double *nu_hot, double *nu_cold) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    hipError_t e = lbm::launch_nusselt(*h);
    if (e != hipSuccess) return hip_fail(e, "nusselt launch");
    double r[2];
    HIP_TRY(hipMemcpyAsync(r, h->d_red + 4, sizeof r, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (nu_hot) *nu_hot = r[0];
    if (nu_cold) *nu_cold = r[1];
    if (!std::isfinite(r[0]) || !std::isfinite(r[1]))
        return fail(LBM_ERR_DIVERGED, "non-finite Nusselt number");
    return LBM_OK;
}

int32_t lbm_get_fields(lbm_handle h, double *rho, double *u, double *v, double *w,
                       double *T) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    const size_t bytes = sizeof(double) * (size_t)h->nn;
    if (rho) HIP_TRY(hipMemcpy(rho, h->rho, bytes, hipMemcpyDeviceToHost));
    if (u) HIP_TRY(hipMemcpy(u, h->u, bytes, hipMemcpyDeviceToHost));
    if (v) HIP_TRY(hipMemcpy(v, h->v, bytes, hipMemcpyDeviceToHost));
    if (w) {
        if (h->w) HIP_TRY(hipMemcpy(w, h->w, bytes, hipMemcpyDeviceToHost));
        else std::memset(w, 0, bytes);
    }
    if (T) HIP_TRY(hipMemcpy(T, h->T, bytes, hipMemcpyDeviceToHost));
    return LBM_OK;
}

int32_t lbm_run(lbm_handle h, int64_t *itc, double *errU, double *errT) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    const int64_t every = h->p.check_every;
    double eu = 100.0, et = 100.0;   // 2dim_mono.f90:148-149
    int32_t status = LBM_OK;
    while ((eu > h->p.epsU || et > h->p.epsT) && h->itc < h->p.itc_max) {
        int64_t to_check = every - (h->itc % every);
        int64_t n = std::min<int64_t>(to_check, h->p.itc_max - h->itc);
        if ((rc = do_steps(*h, n))) return rc;
        if (h->itc % every == 0) {
            rc = lbm_check(h, &eu, &et);
            if (rc == LBM_ERR_DIVERGED) { status = rc; break; }
            if (rc) return rc;
        }
    }
    if (status == LBM_OK && (eu > h->p.epsU || et > h->p.epsT)) {
        status = LBM_ERR_NOT_CONVERGED;
        g_err = "itc_max reached before convergence";
    }
    if (itc) *itc = h->itc;
    if (errU) *errU = eu;
    if (errT) *errT = et;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return status;
}

int32_t lbm_get_itc(lbm_handle h, int64_t *itc) {
    if (!h || !itc) return fail(LBM_ERR_ARG, "null argument");
    *itc = h->itc;
    return LBM_OK;
}

int32_t lbm_time_steps(lbm_handle h, int64_t nsteps, double *ms) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    if (nsteps < 1 || !ms) return fail(LBM_ERR_ARG, "nsteps >= 1 and ms required");
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    if ((rc = do_steps(*h, nsteps))) return rc;
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipEventSynchronize(h->ev1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = t;
    return LBM_OK;
}

int32_t lbm_synchronize(lbm_handle h) {
    if (!h) return fail(LBM_ERR_ARG, "null handle");
    HIP_TRY(hipSetDevice(h->p.device));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return LBM_OK;
}

}  // extern "C"

// Checkpoint layout: header (magic, dim, n, precision, itc), the current
// population buffer, then rho, u, v, w, T and the snapshots up, vp, wp, Tp
// and (FP64) the force, all raw device bytes.
namespace {
struct StateHeader {
    char magic[8];
    int32_t dim, n, precision, pad;
    int64_t itc;
};
struct Segment {
    void *dev;
    size_t bytes;
};
std::vector<Segment> state_segments(lbm_handle h) {
    const size_t esz = h->fp64 ? sizeof(double) : sizeof(float);
    const size_t fb = sizeof(double) * (size_t)h->nn;
    std::vector<Segment> s;
    s.push_back({h->pbuf[h->cur], esz * (size_t)h->plane * (size_t)(h->qf + h->qg)});
    double *fields[] = {h->rho, h->u, h->v, h->w, h->T, h->up, h->vp, h->wp, h->Tp, h->force};
    for (double *p : fields)
        if (p) s.push_back({p, fb});
    return s;
}
}  // namespace

extern "C" {

int32_t lbm_state_size(lbm_handle h, int64_t *bytes) {
    if (!h || !bytes) return fail(LBM_ERR_ARG, "null argument");
    int64_t b = sizeof(StateHeader);
    for (auto &s : state_segments(h)) b += (int64_t)s.bytes;
    *bytes = b;
    return LBM_OK;
}

int32_t lbm_save_state(lbm_handle h, void *buf, int64_t bytes) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    int64_t need = 0;
    lbm_state_size(h, &need);
    if (!buf || bytes < need) return fail(LBM_ERR_ARG, "buffer too small");
    HIP_TRY(hipStreamSynchronize(h->stream));
    StateHeader hd{};
    std::memcpy(hd.magic, "LBMSTAT1", 8);
    hd.dim = h->dim;
    hd.n = h->n;
    hd.precision = h->p.precision;
    hd.itc = h->itc;
    char *o = (char *)buf;
    std::memcpy(o, &hd, sizeof hd);
    o += sizeof hd;
    for (auto &s : state_segments(h)) {
        HIP_TRY(hipMemcpy(o, s.dev, s.bytes, hipMemcpyDeviceToHost));
        o += s.bytes;
    }
    return LBM_OK;
}

int32_t lbm_load_state(lbm_handle h, const void *buf, int64_t bytes) {
    int32_t rc = check_handle(h, false);
    if (rc) return rc;
    int64_t need = 0;
    lbm_state_size(h, &need);
    if (!buf || bytes < need) return fail(LBM_ERR_ARG, "buffer too small");
    StateHeader hd;
    std::memcpy(&hd, buf, sizeof hd);
    if (std::memcmp(hd.magic, "LBMSTAT1", 8) != 0 || hd.dim != h->dim || hd.n != h->n ||
        hd.precision != h->p.precision)
        return fail(LBM_ERR_ARG, "checkpoint does not match this handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    const char *o = (const char *)buf + sizeof hd;
    for (auto &s : state_segments(h)) {
        HIP_TRY(hipMemcpy(s.dev, o, s.bytes, hipMemcpyHostToDevice));
        o += s.bytes;
    }
    h->itc = hd.itc;
    h->initialized = true;
    h->inits++;
    return LBM_OK;
}

int32_t lbm_copy_state(lbm_handle dst, lbm_handle src) {
    int32_t rc = check_handle(src, true);
    if (rc) return rc;
    if ((rc = check_handle(dst, false))) return rc;
    if (dst == src) return fail(LBM_ERR_ARG, "source and destination are the same handle");
    const lbm_params &a = src->p, &b = dst->p;
    if (a.dim != b.dim || a.nx != b.nx || a.Ra != b.Ra || a.Pr != b.Pr || a.Ma != b.Ma ||
        a.Thot != b.Thot || a.Tcold != b.Tcold || a.stretch != b.stretch ||
        a.grid_kind != b.grid_kind || a.device != b.device)
        return fail(LBM_ERR_ARG, "lbm_copy_state needs the same grid, physics and device");
    HIP_TRY(hipStreamSynchronize(src->stream));
    HIP_TRY(hipStreamSynchronize(dst->stream));
    hipError_t e = lbm::launch_convert_state(*dst, *src);
    if (e != hipSuccess) return hip_fail(e, "state conversion");
    const size_t fb = sizeof(double) * (size_t)src->nn;
    double *from[] = {src->rho, src->u, src->v, src->w, src->T, src->up, src->vp, src->wp, src->Tp};
    double *to[] = {dst->rho, dst->u, dst->v, dst->w, dst->T, dst->up, dst->vp, dst->wp, dst->Tp};
    for (int k = 0; k < 9; k++)
        if (from[k] && to[k])
            HIP_TRY(hipMemcpyAsync(to[k], from[k], fb, hipMemcpyDeviceToDevice, dst->stream));
    HIP_TRY(hipStreamSynchronize(dst->stream));
    dst->itc = src->itc;
    dst->initialized = true;
    dst->inits++;
    return LBM_OK;
}


int32_t lbm_set_launch(lbm_handle h, const lbm_launch *l) {
    if (!h || !l) return fail(LBM_ERR_ARG, "null argument");
    if (l->bx < 32 || l->bx % 32 || l->by < 1 || l->bx * l->by > 1024 || (l->bx * l->by) % 64 ||
        l->remap < 0 || l->remap > 2 ||
        (l->occupancy != 0 && l->occupancy != 5 && l->occupancy != 6) ||
        l->variant < 0 || l->variant > 2)
        return fail(LBM_ERR_ARG, "invalid launch configuration");
    if (l->variant == 2 && (h->dim != 2 || h->fp64 || l->occupancy != 0 ||
                            !lbm::t2_shape_ok(l->bx, l->by)))
        return fail(LBM_ERR_ARG, "variant 2 (two steps per pass) needs a 2D cavity with FP32 "
                                 "storage, bx = 64, by = 8, 12 or 16 and occupancy 0");
    if (l->variant == 1 && !zm_launch_ok(*h, l->bx, l->by, l->occupancy))
        return fail(LBM_ERR_ARG,
                    "variant 1 needs FP32 storage, bx = 64, occupancy 0 and, in 2D, by = 1; "
                    "in 3D, by = 2, 4, 6 or 8, n >= 6 and 26 padded population planes "
                    "within 32-bit element offsets (n <= 544)");
    h->shape.bx = l->bx;
    h->shape.by = l->by;
    h->shape.remap = l->remap;
    h->shape.occ = l->occupancy;
    h->shape.variant = l->variant;
    return LBM_OK;
}

int32_t lbm_get_launch(lbm_handle h, lbm_launch *l) {
    if (!h || !l) return fail(LBM_ERR_ARG, "null argument");
    l->bx = h->shape.bx;
    l->by = h->shape.by;
    l->remap = h->shape.remap;
    l->occupancy = (h->dim == 3 && h->p.precision == LBM_PREC_FP32_SHIFTED) ? h->shape.occ : 0;
    l->variant = h->shape.variant;
    return LBM_OK;
}

int32_t lbm_autotune(lbm_handle h, int32_t steps_per_trial, double *best_ms) {
    int32_t rc = check_handle(h, true);
    if (rc) return rc;
    if (steps_per_trial < 1) return fail(LBM_ERR_ARG, "steps_per_trial < 1");
    // snapshot of everything the trial steps modify
    std::vector<Segment> segs = state_segments(h);
    std::vector<void *> saved(segs.size(), nullptr);
    auto release = [&]() {
        for (void *p : saved)
            if (p) (void)hipFree(p);
    };
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (size_t k = 0; k < segs.size(); k++) {
        hipError_t e = hipMalloc(&saved[k], segs[k].bytes);
        if (e == hipSuccess)
            e = hipMemcpyAsync(saved[k], segs[k].dev, segs[k].bytes, hipMemcpyDeviceToDevice,
                               h->stream);
        if (e != hipSuccess) {
            release();
            return hip_fail(e, "autotune snapshot");
        }
    }
    const int cur0 = h->cur;
    const int64_t itc0 = h->itc;
    const lbm::LaunchShape start = h->shape;
    // candidates (bx, by), each with every tile order and, for the 3D FP32
    // kernel, each occupancy target
    static const int cand[][2] = {{64, 1}, {64, 2}, {64, 4}, {64, 6}, {64, 8}, {128, 1},
                                  {128, 2}, {128, 4}, {256, 1}, {256, 2}, {32, 8}};
    lbm::LaunchShape best = start;
    double best_t = 1e300;
    const int nocc = (h->dim == 3 && h->p.precision == LBM_PREC_FP32_SHIFTED) ? 3 : 1;
    for (const auto &cd : cand) {
        if (cd[0] > 2 * h->n && cd[0] > 64) continue;
        // variant 1 (occupancy fixed by its LDS) where it can run
        const int nvar = zm_launch_ok(*h, cd[0], cd[1], 0) ? 2 : 1;
        for (int combo = 0; combo < 3 * (nocc + nvar - 1); combo++) {
            static const int occs[3] = {0, 5, 6};
            const int sel = combo / 3;   // 0..nocc-1: variant 0 occupancies; nocc: variant 1
            h->shape.bx = cd[0];
            h->shape.by = cd[1];
            h->shape.remap = combo % 3;
            h->shape.occ = sel < nocc ? occs[sel] : 0;
            h->shape.variant = sel < nocc ? 0 : 1;
            double ms = 0;
            // one untimed warm-up launch per candidate, then the timed trial
            if ((rc = do_steps(*h, 1)) || (rc = lbm_time_steps(h, steps_per_trial, &ms))) {
                release();
                return rc;
            }
            ms /= steps_per_trial;
            if (ms < best_t) {
                best_t = ms;
                best = h->shape;
            }
        }
    }
    // variant 2 (2D FP32): two steps per pass, tiles of 60 x 8, 12 or 16
    if (h->dim == 2 && !h->fp64) {
        for (int by : {8, 12, 16})
            for (int remap = 0; remap < 2; remap++) {
                h->shape = lbm::LaunchShape{64, by, remap, 0, 2};
                double ms = 0;
                if ((rc = do_steps(*h, 2)) || (rc = lbm_time_steps(h, steps_per_trial, &ms))) {
                    release();
                    return rc;
                }
                ms /= steps_per_trial;
                if (ms < best_t) {
                    best_t = ms;
                    best = h->shape;
                }
            }
    }
    h->shape = best;
    // restore: the snapshot was taken from buffer cur0
    h->cur = cur0;
    segs = state_segments(h);
    for (size_t k = 0; k < segs.size(); k++) {
        hipError_t e = hipMemcpyAsync(segs[k].dev, saved[k], segs[k].bytes,
                                      hipMemcpyDeviceToDevice, h->stream);
        if (e != hipSuccess) {
            release();
            return hip_fail(e, "autotune restore");
        }
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    release();
    h->itc = itc0;
    if (best_ms) *best_ms = best_t;
    return LBM_OK;
}

int32_t lbm_bytes_per_update(lbm_handle h, double *bytes) {
    if (!h || !bytes) return fail(LBM_ERR_ARG, "null argument");
    const double esz = h->fp64 ? 8.0 : 4.0;
    // every population read once and written once; the FP64 path also
    // reads and writes its per-node force (the FP32 path recovers it from
    // the node's own populations)
    *bytes = 2.0 * esz * (h->qf + h->qg) + (h->fp64 ? 2.0 * sizeof(double) : 0.0);
    return LBM_OK;
}

const char *lbm_status_string(int32_t status) {
    switch (status) {
    case LBM_OK: return "ok";
    case LBM_ERR_ARG: return "invalid argument";
    case LBM_ERR_HIP: return "HIP runtime error";
    case LBM_ERR_DIVERGED: return "diverged (non-finite residual)";
    case LBM_ERR_NOT_CONVERGED: return "not converged within itc_max";
    case LBM_ERR_STATE: return "call out of order";
    case LBM_ERR_NOMEM: return "out of memory";
    default: return "unknown status";
    }
}

const char *lbm_last_error(void) { return g_err.c_str(); }

}  // extern "C"
