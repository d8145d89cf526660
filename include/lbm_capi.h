/*
 * lbm_capi.h -- C ABI of the MI355X (gfx950) coupled MRT lattice-Boltzmann
 * natural-convection solver.
 *
 * The reference (zzhao0930/MPI-Multi-Process-Version) has no plugin or FFI
 * layer: its boundary is the sequence of argument-less subroutine calls in
 * `program main` over the module-global state `commondata`
 * (natural-convection_2dim_mono.f90:1-72 and :283-380;
 *  natural-convection_3dim_mono.f90:1-77 and :347-443).  Each entry point
 * below replaces one group of those calls; the citation is on each
 * declaration.  A Fortran caller binds them with iso_c_binding (see
 * mpi-multi-process-version_amd/fortran/lbm_capi_mod.f90).
 *
 * Conventions
 *  - Every function returns an int32 status, LBM_OK (0) on success.  No C++
 *    exception crosses the ABI.  lbm_last_error() gives a detail message for
 *    the calling thread.
 *  - The handle owns all device memory.  Host arrays belong to the caller:
 *    contiguous FP64 arrays of nx*ny(*nz) values in Fortran order (i fastest,
 *    then j, then k), copied synchronously.  w may be NULL in 2D.
 *  - A handle is not thread-safe; each handle owns one HIP stream and
 *    separate handles may run concurrently.  lbm_step() is asynchronous;
 *    every getter synchronises the handle's stream.
 *  - Axes follow the reference: 2D heated x-walls (hot at i=1) with gravity
 *    along +y; 3D heated y-walls (hot at j=1) with gravity along +z.
 */
#ifndef LBM_CAPI_H
#define LBM_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define LBM_OK                 0
#define LBM_ERR_ARG            1  /* invalid argument or parameter set        */
#define LBM_ERR_HIP            2  /* HIP runtime error (no device, launch...)  */
#define LBM_ERR_DIVERGED       3  /* non-finite residual                      */
#define LBM_ERR_NOT_CONVERGED  4  /* itc_max reached before epsU/epsT         */
#define LBM_ERR_STATE          5  /* call out of order (e.g. step before init) */
#define LBM_ERR_NOMEM          6  /* device or host allocation failed         */

/* Grid maps.  The reference uses the erf map with constA = 3.2
 * (natural-convection_2dim_mono.f90:85-91). */
#define LBM_GRID_ERF   0  /* x_i = (erf(A(i/(N+1)-1/2))/erf(A/2)+1)/2        */
#define LBM_GRID_TANH  1  /* same with tanh in place of erf (extension)      */

#define LBM_PREC_FP32_SHIFTED  0  /* FP32 populations stored as deviations from
                                     the rest equilibrium (flow) and from the
                                     conduction profile (heat)                */
#define LBM_PREC_FP64          1  /* FP64 populations, the reference's form   */

/* Inputs of the reference's `commondata` module, made run-time
 * (natural-convection_2dim_mono.f90:6-35, constA at :85). */
typedef struct lbm_params {
    int32_t dim;          /* 2 or 3                                          */
    int32_t nx, ny, nz;   /* nodes per axis; the cavity is square/cubic:
                             nx == ny (== nz in 3D), nx >= 5               */
    double  Ra, Pr, Ma;   /* Rayleigh, Prandtl, Mach numbers                 */
    double  Thot, Tcold;  /* wall temperatures                               */
    double  stretch;      /* grid stretching factor A (reference: 3.2)       */
    int32_t grid_kind;    /* LBM_GRID_*                                      */
    int32_t precision;    /* LBM_PREC_*                                      */
    double  epsU, epsT;   /* convergence thresholds of lbm_run               */
    int32_t check_every;  /* residual window in steps (reference: 2000)      */
    int32_t device;       /* HIP device ordinal (>= 0; there is no CPU
                             backend)                                        */
    int64_t itc_max;      /* step limit of lbm_run                           */
} lbm_params;

typedef struct lbm_ctx *lbm_handle;

/* Derived parameters returned by lbm_get_info(), in this order. */
enum {
    LBM_INFO_LLB = 0,  /* length_LB = 1/dx0           (2dim_mono.f90:110) */
    LBM_INFO_DT,       /* dt = dx0*length_LB           (2dim_mono.f90:111) */
    LBM_INFO_NU,       /* viscosity_LB                 (2dim_mono.f90:152) */
    LBM_INFO_KAPPA,    /* thermal diffusivity          (2dim_mono.f90:153) */
    LBM_INFO_GBETA,    /* g*beta                       (2dim_mono.f90:154) */
    LBM_INFO_TAUF,     /* tau_f                        (2dim_mono.f90:157) */
    LBM_INFO_SNU,      /* S_nu                         (2dim_mono.f90:160) */
    LBM_INFO_SQ,       /* S_q                          (2dim_mono.f90:161) */
    LBM_INFO_SIGK,     /* sigma_k (2dim_mono.f90:162, 3dim_mono.f90:229)   */
    LBM_INFO_H,        /* cavity size in LB units = length_LB - 1          */
    LBM_INFO_COUNT
};

/* mesh + the parameter part of initial (2dim_mono.f90:74-134, :151-174;
 * 3dim_mono.f90:122-189, :206-229): validates the parameters (tau_f > 1/2,
 * square/cubic grid), builds the grid and the interpolation tables and
 * allocates the device state.  The populations are not set until
 * lbm_init_rest() or lbm_init_fields(). */
int32_t lbm_create(const lbm_params *params, lbm_handle *out);

/* Frees the device state (the reference's deallocate block,
 * 2dim_mono.f90:347-375). */
int32_t lbm_destroy(lbm_handle h);

/* Node coordinates in LB units, nx (ny, nz) values each (xGrid, y/yGrid, z
 * after 2dim_mono.f90:113-116).  z may be NULL in 2D; any may be NULL. */
int32_t lbm_get_grid(lbm_handle h, double *x, double *y, double *z);

/* Derived parameters, LBM_INFO_COUNT doubles. */
int32_t lbm_get_info(lbm_handle h, double *info);

/* The initial condition of `initial` (2dim_mono.f90:199-240;
 * 3dim_mono.f90:256-301): rho = 1, u = 0, T linear between the heated walls,
 * equilibrium populations; residual snapshots cleared; itc = 0. */
int32_t lbm_init_rest(lbm_handle h);

/* Warm start / restart from macroscopic fields: rho = 1 and the given
 * u, v, (w), T, equilibrium populations (the same formulas as
 * 2dim_mono.f90:227-240); snapshots cleared; itc = 0.  v and w may be NULL
 * (zero); T must not be NULL. */
int32_t lbm_init_fields(lbm_handle h, const double *u, const double *v,
                        const double *w, const double *T);

/* nsteps iterations of the time-loop body: collision_U, collision_T,
 * [update], interpolate, bounceback_u, bounceback_T, macro_u, macro_t
 * (2dim_mono.f90:309-325; 3dim_mono.f90:375-391).  Asynchronous. */
int32_t lbm_step(lbm_handle h, int64_t nsteps);

/* check (2dim_mono.f90:755-805; 3dim_mono.f90:1027-1082): relative L2
 * change of (u,v[,w]) and of T since the previous call, FP64; then the
 * snapshot is updated.  The first call after initialisation returns 1.
 * If a field and its change are both identically zero the ratio is
 * reported as 0 (the reference would compute 0/0). */
int32_t lbm_check(lbm_handle h, double *errU, double *errT);

/* lbm_check with nothing allocated by the library for the call: no
 * reference counterpart; for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  The same kernels and results as lbm_check, bit for
 * bit, with the reduction's partial sums (at most 1024 blocks x 4 FP64
 * values, the most lbm_create ever sets up) in one static device array of
 * 32.8 KB per GPU, held until the process ends, instead of the handle's
 * buffers.  Calls that overlap on one GPU (handles on different streams
 * or threads) share that array and race: serialise them. */
int32_t lbm_check_noalloc(lbm_handle h, double *errU, double *errT);

/* Mean Nusselt numbers of the hot and cold walls, FP64 (no reference
 * counterpart; the reference only writes fields, 2dim_mono.f90:818-843):
 * Nu = H dT/dn / (Thot - Tcold) from the one-sided quadratic through the
 * wall value and the first two nodes, averaged with control-volume
 * weights over the wall. */
int32_t lbm_nusselt(lbm_handle h, double *nu_hot, double *nu_cold);

/* gather_data (2dim_mono.f90:807-816; 3dim_mono.f90:1084-1095): the
 * macroscopic fields of the current step, FP64, Fortran order.  Any pointer
 * may be NULL to skip that field. */
int32_t lbm_get_fields(lbm_handle h, double *rho, double *u, double *v,
                       double *w, double *T);

/* The time loop with its stopping rule (2dim_mono.f90:307-330): steps in
 * windows of check_every, calling lbm_check at every multiple of
 * check_every, until both residuals are <= epsU/epsT (LBM_OK) or
 * itc >= itc_max (LBM_ERR_NOT_CONVERGED).  A non-finite residual returns
 * LBM_ERR_DIVERGED (the reference would exit as if converged).  itc, errU
 * and errT receive the final values. */
int32_t lbm_run(lbm_handle h, int64_t *itc, double *errU, double *errT);

/* Iteration counter (the reference's itc, 2dim_mono.f90:30). */
int32_t lbm_get_itc(lbm_handle h, int64_t *itc);

/* Runs nsteps steps like lbm_step and returns the elapsed time between two
 * HIP events recorded on the handle's stream around them, in ms
 * (the reference times with MPI_WTIME, 2dim_mono.f90:300-337). */
int32_t lbm_time_steps(lbm_handle h, int64_t nsteps, double *ms);

/* Blocks until the handle's stream is idle. */
int32_t lbm_synchronize(lbm_handle h);

/* Exact checkpoint of the whole solver state (populations, macroscopic
 * fields, residual snapshots, itc).  No reference counterpart (the
 * reference has no restart).  lbm_state_size gives the byte count. */
int32_t lbm_state_size(lbm_handle h, int64_t *bytes);
int32_t lbm_save_state(lbm_handle h, void *buf, int64_t bytes);
int32_t lbm_load_state(lbm_handle h, const void *buf, int64_t bytes);

/* dst takes src's whole solver state (populations, macroscopic fields,
 * residual snapshots, itc), converted between the storages when their
 * precisions differ: FP32 shifted f - w_a and g - w_a Tref to FP64 f and g
 * (each rounded once to FP64), or back (rounded once to FP32); an FP64 dst
 * of an FP32 src gets the lagged force g beta T rho from the fields.  One
 * precision: an exact copy.  Both handles need the same grid, physics
 * and device.  For running FP32 through a transient and FP64 to a residual
 * below the FP32 floor.  No reference counterpart (the reference has one
 * precision and no restart). */
int32_t lbm_copy_state(lbm_handle dst, lbm_handle src);

/* Launch configuration of the step kernel (no reference counterpart: the
 * reference's loops have no launch shape).  One lattice node per thread,
 * bx x by threads per workgroup (bx a multiple of 32, bx*by a multiple of 64
 * and at most 1024).
 *  remap: tile order.  Workgroups are dealt round-robin to the 8 XCDs, each
 *    with its own L2.  0 = x fastest, then y, then z; 1 = each XCD sweeps one
 *    contiguous range of that order; 2 = each XCD owns a band of tile rows in
 *    y and sweeps it plane by plane (variant 0: needs the number of y tiles
 *    divisible by 8, else 0 is used; variant 1 in 3D: at least 8 y tiles; in
 *    2D, variant 1 treats 2 as 1).
 *  occupancy (3D, FP32 storage): 0, or 5 or 6 resident waves per SIMD, which
 *    caps the registers per lane.
 *  variant: step-kernel variant.  0 = one node per thread, every stencil
 *    point read through the caches.  1 (FP32 storage, occupancy 0) = each
 *    workgroup marches a strip of 64 columns through the grid with the
 *    populations staged in LDS and a register ring:
 *      3D: a 64 x by column of nodes in z; bx = 64, by = 2, 4, 6 or 8,
 *          6 <= n <= 544 (when the far wall would sit in a tile's first
 *          column or row, that column or row is updated by a second,
 *          one-node-per-thread launch);
 *      2D: one wave per 64 columns, in y; bx = 64, by = 1.
 *    2 (2D, FP32 storage, occupancy 0) = two time steps per launch: each
 *    workgroup of 256 threads holds a 64 x (by + 4) region in registers and
 *    LDS and writes its 60 x by output tile after two steps; bx = 64,
 *    by = 8, 12 or 16 (an odd step count ends with a one-step launch).
 *    Its results equal variant 0's to the bit.
 *    The variants read the same stencil values and agree to FP32 rounding.
 * Defaults set by lbm_create: 3D FP32 with 6 <= n <= 544: variant
 * 1, 64 x 4, remap 2; 2D with n <= 384: 32 x 2, remap 1; 2D FP32 with
 * n > 384: variant 2, 64 x 8, remap 1; otherwise 64 x 4, remap 1,
 * variant 0. */
typedef struct lbm_launch {
    int32_t bx, by, remap, occupancy, variant;
} lbm_launch;
int32_t lbm_set_launch(lbm_handle h, const lbm_launch *launch);
int32_t lbm_get_launch(lbm_handle h, lbm_launch *launch);

/* Block-shape autotuner: times steps_per_trial steps for each candidate
 * shape and tile order, keeps the fastest, and restores the solver state
 * (populations, fields, itc) to what it was before the call.  best_ms gets
 * the fastest time per step.  No reference counterpart. */
int32_t lbm_autotune(lbm_handle h, int32_t steps_per_trial, double *best_ms);

/* Bytes moved per lattice update by one step kernel, as counted for the
 * roofline: every population read once and written once, plus, in FP64
 * storage, the per-node force read and written (16 B). */
int32_t lbm_bytes_per_update(lbm_handle h, double *bytes);

/* ------------------------------------------------------------------------
 * Batches of independent 2D cavities (no reference counterpart: the
 * reference runs one cavity per program; each cavity of a batch follows the
 * single-cavity semantics above, per cavity, natural-convection_2dim_mono.f90
 * throughout).  One kernel launch per time step advances every active cavity
 * of the batch, whatever their sizes, grids and parameters: workgroups are
 * dealt tiles of all cavities from one table.  2D, FP32 shifted storage, one
 * device.  Arrays of per-cavity results hold `count` values in batch order.
 */
typedef struct lbm_batch_ctx *lbm_batch;

/* Creates `count` cavities from params[0..count-1] (each as lbm_create;
 * all 2D, precision LBM_PREC_FP32_SHIFTED, the same device and the same
 * check_every). */
int32_t lbm_batch_create(const lbm_params *params, int32_t count, lbm_batch *out);
int32_t lbm_batch_destroy(lbm_batch b);

/* The handle of cavity `index`, owned by the batch.  Use it for the
 * read-only calls (lbm_get_grid, lbm_get_info, lbm_get_fields, lbm_nusselt,
 * lbm_get_itc, lbm_state_size, lbm_save_state) and to initialise one cavity
 * (lbm_init_rest, lbm_init_fields, lbm_load_state); step, check and run
 * through the batch. */
lbm_handle lbm_batch_cavity(lbm_batch b, int32_t index);

/* lbm_init_rest on every cavity. */
int32_t lbm_batch_init_rest(lbm_batch b);

/* nsteps time steps of every active cavity (asynchronous).  Active means
 * initialised (lbm_batch_init_rest, or lbm_init_rest / lbm_init_fields /
 * lbm_load_state on lbm_batch_cavity) and not retired by lbm_batch_run
 * since its last initialisation.  LBM_ERR_STATE if no cavity is active. */
int32_t lbm_batch_step(lbm_batch b, int64_t nsteps);

/* lbm_check on every cavity (errU, errT: count values each). */
int32_t lbm_batch_check(lbm_batch b, double *errU, double *errT);

/* lbm_nusselt on every cavity. */
int32_t lbm_batch_nusselt(lbm_batch b, double *nu_hot, double *nu_cold);

/* Runs every cavity with the stopping rule of lbm_run, each with its own
 * epsU, epsT and itc_max: every initialised cavity below its itc_max
 * steps with the others in windows of check_every; a cavity whose
 * residuals both reach its eps, whose itc reaches its itc_max, or whose
 * residual is non-finite leaves the batch.  At the end every cavity is
 * retired: lbm_batch_step skips it until it is initialised again.
 * itc, errU, errT, status (LBM_OK, LBM_ERR_NOT_CONVERGED or
 * LBM_ERR_DIVERGED): count values each, any may be NULL.  Returns LBM_OK
 * when the run completed, whatever the per-cavity statuses. */
int32_t lbm_batch_run(lbm_batch b, int64_t *itc, double *errU, double *errT, int32_t *status);

/* Number of active cavities (those lbm_batch_step advances) and their
 * total node count (the MLUPS numerator of one step). */
int32_t lbm_batch_active(lbm_batch b, int32_t *cavities, int64_t *nodes);

/* lbm_time_steps for the batch: nsteps batched steps between two HIP events
 * on the batch's stream, in ms. */
int32_t lbm_batch_time_steps(lbm_batch b, int64_t nsteps, double *ms);

/* Block shape of the batched launch (bx a multiple of 32, bx*by a multiple
 * of 64 and at most 1024; remap and occupancy must be 0), and its
 * autotuner, which tries every variant the batch can take (restores every
 * cavity's state).  Variants:
 *  0: one launch per step, one node per thread, tiles of every cavity;
 *  1: one launch per step, the 2D y-march (bx = 64, by = 1);
 *  2: reserved (a persistent work-queue launch, switched off: rejected);
 *  3: resident cavities (bx = 512, by = 1; every cavity's edge at most 70):
 *     one workgroup per cavity holds the cavity's populations in registers
 *     and LDS and runs all the steps of an lbm_batch_step call (up to 4096
 *     per launch) in one launch; results equal variant 0's to the bit;
 *  4: two steps per launch (bx = 64, by = 8, 12 or 16): tiles of 60 x by
 *     nodes of every cavity, each held with a two-node halo in registers
 *     and LDS for two steps (the single-cavity variant 2); results equal
 *     variant 0's to the bit;
 *  5: reserved (one cooperative launch per call with grid-wide barriers
 *     between steps; switched off: rejected). */
int32_t lbm_batch_set_launch(lbm_batch b, const lbm_launch *launch);
int32_t lbm_batch_get_launch(lbm_batch b, lbm_launch *launch);
int32_t lbm_batch_autotune(lbm_batch b, int32_t steps_per_trial, double *best_ms);

const char *lbm_status_string(int32_t status);
const char *lbm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
