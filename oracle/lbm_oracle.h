/*
 * lbm_oracle.h -- CPU FP64 restatement of the reference's coupled MRT
 * lattice-Boltzmann natural-convection solver (D2Q9 + D2Q5 and D3Q19 + D3Q7,
 * interpolation-supplemented streaming on an erf-stretched grid).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links,
 * loads or falls back to it.
 *
 * The reference is a set of Fortran 90 + MPI programs
 * (natural-convection_2dim_mono.f90 = "2M", natural-convection_3dim_mono.f90 =
 * "3M").  Each function below cites the subroutine it restates.  The domain
 * is one undivided block, which is the reference's np >= 2 semantics: every
 * wall is applied and the one-sided wall stencils are used on all walls.
 *
 * Layout: populations are AoS, direction fastest, node index
 * i + n*j (+ n*n*k) with i fastest (the reference's Fortran order).
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* Parameters returned by orc_params(), in this order. */
enum {
    ORC_P_LLB = 0,   /* length_LB = 1/dx0                       (2M:110) */
    ORC_P_DT,        /* dt = dx0 * length_LB                    (2M:111) */
    ORC_P_DX0,       /* smallest spacing of the unit-interval map (2M:96) */
    ORC_P_NU,        /* viscosity_LB                            (2M:152) */
    ORC_P_KAPPA,     /* kappa = nu/Pr                           (2M:153) */
    ORC_P_GBETA,     /* g*beta                                  (2M:154) */
    ORC_P_TAUF,      /* tau_f = 3 nu + 1/2                      (2M:157) */
    ORC_P_SNU,       /* S_nu = 1/tau_f                          (2M:160) */
    ORC_P_SQ,        /* S_q = 8(2-S_nu)/(8-S_nu)                (2M:161) */
    ORC_P_SIGK,      /* thermal relaxation sigma_k (2M:162 / 3M:229)    */
    ORC_P_H,         /* cavity height in LB units = length_LB - 1       */
    ORC_P_COUNT
};

orc *orc_create(int dim, int n, double Ra, double Pr, double Ma,
                double Thot, double Tcold, double A);
/* The same with a choice of grid map: 0 = erf (2M:85-91, what orc_create
 * uses), 1 = tanh in place of erf (an extension with no reference
 * counterpart). */
orc *orc_create_grid(int dim, int n, double Ra, double Pr, double Ma,
                     double Thot, double Tcold, double A, int grid_kind);
void orc_destroy(orc *s);
void orc_set_threads(int nthreads);   /* OpenMP threads for the sweeps (>=1) */
int  orc_get_threads(void);

int  orc_dim(const orc *s);
int  orc_n(const orc *s);
int  orc_qf(const orc *s);
int  orc_qg(const orc *s);
long orc_itc(const orc *s);
void orc_params(const orc *s, double *out /* ORC_P_COUNT */);
void orc_coords(const orc *s, double *out /* n node coordinates, LB units */);
void orc_stencil(const orc *s, int *idx /* 3*n, 0-based, reference order */);

/* Raw state access for tests: "f", "fpost", "g", "gpost", "rho", "u", "v",
 * "w", "T", "Fx", "Fy", "Fz", "up", "vp", "wp", "Tp".  Returns NULL for an
 * unknown name; *count receives the number of doubles. */
double *orc_array(orc *s, const char *name, long *count);

/* Initial state (2M:199-240, 3M:256-301): rho = 1, the given u, v, w, T,
 * f = f_eq, g = g_eq, snapshots zeroed, itc = 0.  orc_init_rest uses u = 0
 * and T linear between the heated walls (2M:208-212, 3M:265-271). */
void orc_init_rest(orc *s);
void orc_init_fields(orc *s, const double *u, const double *v,
                     const double *w, const double *T);

/* One reference subroutine each. */
void orc_collision_u(orc *s);   /* 2M:431-520, 3M:445-669 */
void orc_collision_t(orc *s);   /* 2M:522-569, 3M:671-730 */
void orc_interpolate(orc *s);   /* 2M:591-646, 3M:751-862 */
void orc_bounceback_u(orc *s);  /* 2M:662-696, 3M:878-943 */
void orc_bounceback_t(orc *s);  /* 2M:698-724, 3M:945-986 */
void orc_macro_u(orc *s);       /* 2M:726-739, 3M:988-1010 */
void orc_macro_t(orc *s);       /* 2M:741-753, 3M:1012-1025 */

/* Loop body of the time loop, nsteps times (2M:309-325, 3M:375-391). */
void orc_step(orc *s, long nsteps);
/* Windowed relative-L2 residuals + snapshot update (2M:755-805). */
void orc_check(orc *s, double *errU, double *errT);
/* Time loop with the reference's stopping rule (2M:307-330); returns itc. */
long orc_run(orc *s, double epsU, double epsT, long itc_max, int check_every,
             double *errU, double *errT);
/* Mean Nusselt number on the hot and cold walls (no reference counterpart;
 * one-sided quadratic wall gradient, control-volume weights). */
void orc_nusselt(const orc *s, double *nu_hot, double *nu_cold);

/* The reference's 3-point Lagrange formula (2M:649-660). */
double orc_interpolateF(double x0, double x1, double x2, double x,
                        double f0, double f1, double f2);

#ifdef __cplusplus
}
#endif
#endif
