/*
 * lbm_oracle.c -- CPU FP64 restatement of the reference's coupled MRT-LBM
 * natural-convection solver.  TEST INFRASTRUCTURE ONLY (see lbm_oracle.h):
 * the HIP product never links, loads or falls back to this file.
 *
 * Reference files: natural-convection_2dim_mono.f90 ("2M") and
 * natural-convection_3dim_mono.f90 ("3M").  The arithmetic follows the
 * reference's algorithm step by step; the moment matrices are built from
 * their polynomial definitions (orthogonal bases, so M^-1 = M^T diag(1/|row|^2))
 * and the boundary rules from the lattice vectors, instead of the reference's
 * hand-expanded sums.  Build with -ffp-contract=off (see Makefile).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define QMAX 19

/* ---------------------------------------------------------------- lattices */
/* Velocity sets in the reference's numbering: D2Q9 2M:58-60 (D2Q5 = its first
 * five), D3Q19 3M:60-63 (D3Q7 = its first seven). */
static const int C2[9][3] = {
    {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0},
    {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
static const int C3[19][3] = {
    {0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
    {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, 1, 0},  {1, -1, 0},
    {-1, -1, 0}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1},  {-1, 0, -1},
    {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};

/* Thermal inverse transforms N^-1 as integer numerators over a common
 * denominator (2M:553-557 with denominator 4, 3M:709-715 with 6). */
static const int NINV2[5][5] = {
    {4, 0, 0, -4, 0}, {0, 2, 0, 1, 1}, {0, 0, 2, 1, -1},
    {0, -2, 0, 1, 1}, {0, 0, -2, 1, -1}};
static const int NINV3[7][7] = {
    {6, 0, 0, 0, -6, 0, 0}, {0, 3, 0, 0, 1, 1, 1}, {0, -3, 0, 0, 1, 1, 1},
    {0, 0, 3, 0, 1, -2, 1}, {0, 0, -3, 0, 1, -2, 1}, {0, 0, 0, 3, 1, 1, -2},
    {0, 0, 0, -3, 1, 1, -2}};

struct orc {
    int dim, n, qf, qg;
    long nn;
    long itc;
    double Ra, Pr, Ma, Thot, Tcold, A;
    int grid_kind;   /* 0: erf map (the reference's), 1: tanh map */
    double p[ORC_P_COUNT];
    int c[QMAX][3];
    double wf[QMAX], wg[7];
    double M[QMAX][QMAX], norm[QMAX]; /* flow moment basis               */
    double N[7][7];                   /* thermal moment basis            */
    int ninv[7][7], ninv_den;
    double s[QMAX];                   /* flow relaxation rates           */
    double qrel[7];                   /* thermal relaxation rates        */
    double *x;                        /* node coordinates (same per axis) */
    int *st;                          /* 3-point stencil, 3 per index     */
    double *f, *fpost, *g, *gpost;
    double *rho, *u, *v, *w, *T, *Fx, *Fy, *Fz, *up, *vp, *wp, *Tp;
};

static int g_threads = 1;
void orc_set_threads(int nthreads) { g_threads = nthreads < 1 ? 1 : nthreads; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------ moment bases */
/* D2Q9 rows: rho, e, eps, jx, qx, jy, qy, pxx, pxy (Lallemand & Luo 2000);
 * these are the rows coded at 2M:445-453. */
static void moment_row_2d(const int *c, double *row) {
    double cx = c[0], cy = c[1], q = cx * cx + cy * cy;
    row[0] = 1;
    row[1] = -4 + 3 * q;
    row[2] = 4 - 10.5 * q + 4.5 * q * q;
    row[3] = cx;
    row[4] = (-5 + 3 * q) * cx;
    row[5] = cy;
    row[6] = (-5 + 3 * q) * cy;
    row[7] = cx * cx - cy * cy;
    row[8] = cx * cy;
}
/* D3Q19 rows (d'Humieres et al. 2002): rho, e, eps, jx, qx, jy, qy, jz, qz,
 * 3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz (3M:460-520). */
static void moment_row_3d(const int *c, double *row) {
    double cx = c[0], cy = c[1], cz = c[2], q = cx * cx + cy * cy + cz * cz;
    row[0] = 1;
    row[1] = 19 * q - 30;
    row[2] = (21 * q * q - 53 * q + 24) / 2;
    row[3] = cx;
    row[4] = (5 * q - 9) * cx;
    row[5] = cy;
    row[6] = (5 * q - 9) * cy;
    row[7] = cz;
    row[8] = (5 * q - 9) * cz;
    row[9] = 3 * cx * cx - q;
    row[10] = (3 * q - 5) * (3 * cx * cx - q);
    row[11] = cy * cy - cz * cz;
    row[12] = (3 * q - 5) * (cy * cy - cz * cz);
    row[13] = cx * cy;
    row[14] = cy * cz;
    row[15] = cx * cz;
    row[16] = (cy * cy - cz * cz) * cx;
    row[17] = (cz * cz - cx * cx) * cy;
    row[18] = (cx * cx - cy * cy) * cz;
}

/* ------------------------------------------------------------------- grid */
/* mesh (2M:74-134, 3M:122-189): erf-stretched unit interval, smallest
 * spacing dx0 = X1 - X0, length_LB = 1/dx0, dt = dx0*length_LB, nodes
 * x_i = (X_i - dx0/2)*length_LB for i = 1..n.  grid_kind 1 puts tanh in
 * place of erf in the map of 2M:85-91 (an extension: the reference has the
 * erf map only); everything after the map is the same. */
static void make_grid(orc *s) {
    int n = s->n;
    double A = s->A;
    double *X = (double *)malloc(sizeof(double) * (n + 2));
    for (int i = 0; i <= n + 1; i++) {
        double t = A * ((double)i / (double)(n + 1) - 0.5);
        X[i] = s->grid_kind == 1 ? 0.5 * (tanh(t) / tanh(0.5 * A) + 1.0)
                                 : 0.5 * (erf(t) / erf(0.5 * A) + 1.0);
    }
    double dx0 = X[1] - X[0];
    double L = 1.0 / dx0;
    s->p[ORC_P_DX0] = dx0;
    s->p[ORC_P_LLB] = L;
    s->p[ORC_P_DT] = dx0 * L;
    for (int i = 1; i <= n; i++) s->x[i - 1] = (X[i] - dx0 / 2.0) * L;
    free(X);
    /* 3-point stencils (2M:242-278): one-sided at the two walls. */
    for (int i = 0; i < n; i++) {
        int *t = s->st + 3 * i;
        if (i == 0) { t[0] = 1; t[1] = 0; t[2] = 2; }
        else if (i == n - 1) { t[0] = n - 2; t[1] = n - 1; t[2] = n - 3; }
        else { t[0] = i - 1; t[1] = i; t[2] = i + 1; }
    }
}

/* initial, parameter part (2M:151-162, 3M:206-229). */
static void make_params(orc *s) {
    double L = s->p[ORC_P_LLB], Pr = s->Pr, Ra = s->Ra;
    double nu = (s->Ma * (L - 1.0) * sqrt(Pr)) / sqrt(3.0 * Ra);
    double kappa = nu / Pr;
    double H = L - 1.0;
    double gbeta = Ra * nu * kappa / (H * H * H);
    double tauf = nu * 3.0 + 0.5;
    double Snu = 1.0 / tauf;
    double Sq = 8.0 * (2.0 - Snu) / (8.0 - Snu);
    double sigk = s->dim == 2 ? 1.0 / (0.5 + 4.0 * (1.0 / Snu - 0.5) / (3.0 * Pr))
                              : 1.0 / (0.5 + (1.0 / Snu - 0.5) / (0.5 * Pr));
    s->p[ORC_P_NU] = nu;
    s->p[ORC_P_KAPPA] = kappa;
    s->p[ORC_P_GBETA] = gbeta;
    s->p[ORC_P_TAUF] = tauf;
    s->p[ORC_P_SNU] = Snu;
    s->p[ORC_P_SQ] = Sq;
    s->p[ORC_P_SIGK] = sigk;
    s->p[ORC_P_H] = H;
    if (s->dim == 2) {
        /* S = diag(0, Snu, Snu, 0, Sq, 0, Sq, Snu, Snu)  (2M:465-473) */
        double sv[9] = {0, Snu, Snu, 0, Sq, 0, Sq, Snu, Snu};
        memcpy(s->s, sv, sizeof sv);
        double qv[5] = {1.0, sigk, sigk, 1.9, 1.9};          /* 2M:537-541 */
        memcpy(s->qrel, qv, sizeof qv);
    } else {
        /* 3M:542-560 */
        double sv[19] = {0, Snu, Snu, 0, Sq, 0, Sq, 0, Sq, Snu,
                         Snu, Snu, Snu, Snu, Snu, Snu, Sq, Sq, Sq};
        memcpy(s->s, sv, sizeof sv);
        double qv[7] = {1.0, sigk, sigk, sigk, 1.2, 1.2, 1.2};  /* 3M:689-695 */
        memcpy(s->qrel, qv, sizeof qv);
    }
}

orc *orc_create(int dim, int n, double Ra, double Pr, double Ma,
                double Thot, double Tcold, double A) {
    return orc_create_grid(dim, n, Ra, Pr, Ma, Thot, Tcold, A, 0);
}

orc *orc_create_grid(int dim, int n, double Ra, double Pr, double Ma,
                     double Thot, double Tcold, double A, int grid_kind) {
    if ((dim != 2 && dim != 3) || n < 3 || grid_kind < 0 || grid_kind > 1) return NULL;
    orc *s = (orc *)calloc(1, sizeof(orc));
    s->grid_kind = grid_kind;
    s->dim = dim;
    s->n = n;
    s->qf = dim == 2 ? 9 : 19;
    s->qg = dim == 2 ? 5 : 7;
    s->nn = dim == 2 ? (long)n * n : (long)n * n * n;
    s->Ra = Ra; s->Pr = Pr; s->Ma = Ma; s->Thot = Thot; s->Tcold = Tcold; s->A = A;
    for (int a = 0; a < s->qf; a++)
        for (int d = 0; d < 3; d++) s->c[a][d] = dim == 2 ? C2[a][d] : C3[a][d];
    if (dim == 2) {
        double wf[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
        memcpy(s->wf, wf, sizeof wf);
        double wg[5] = {0.5, 0.125, 0.125, 0.125, 0.125};
        memcpy(s->wg, wg, sizeof wg);
    } else {
        s->wf[0] = 1.0 / 3.0;
        for (int a = 1; a <= 6; a++) s->wf[a] = 1.0 / 18.0;
        for (int a = 7; a < 19; a++) s->wf[a] = 1.0 / 36.0;
        s->wg[0] = 0.5;
        for (int a = 1; a <= 6; a++) s->wg[a] = 1.0 / 12.0;
    }
    /* flow moment basis and its row norms */
    for (int a = 0; a < s->qf; a++) {
        double row[QMAX];
        if (dim == 2) moment_row_2d(s->c[a], row); else moment_row_3d(s->c[a], row);
        for (int r = 0; r < s->qf; r++) s->M[r][a] = row[r];
    }
    for (int r = 0; r < s->qf; r++) {
        double acc = 0;
        for (int a = 0; a < s->qf; a++) acc += s->M[r][a] * s->M[r][a];
        s->norm[r] = acc;
    }
    /* thermal basis: 1, c_x, c_y, (c_z), |c|^2, then c_x^2-c_y^2 (and
     * c_x^2-c_z^2 in 3D)  (2M:531-535, 3M:681-687) */
    for (int a = 0; a < s->qg; a++) {
        double cx = s->c[a][0], cy = s->c[a][1], cz = s->c[a][2];
        double q = cx * cx + cy * cy + cz * cz;
        if (dim == 2) {
            double r5[5] = {1, cx, cy, q, cx * cx - cy * cy};
            for (int r = 0; r < 5; r++) s->N[r][a] = r5[r];
        } else {
            double r7[7] = {1, cx, cy, cz, q, cx * cx - cy * cy, cx * cx - cz * cz};
            for (int r = 0; r < 7; r++) s->N[r][a] = r7[r];
        }
    }
    for (int a = 0; a < s->qg; a++)
        for (int r = 0; r < s->qg; r++)
            s->ninv[a][r] = dim == 2 ? NINV2[a][r] : NINV3[a][r];
    s->ninv_den = dim == 2 ? 4 : 6;

    s->x = (double *)calloc(n, sizeof(double));
    s->st = (int *)calloc(3 * (size_t)n, sizeof(int));
    make_grid(s);
    make_params(s);

    size_t nn = (size_t)s->nn;
    s->f = (double *)calloc(nn * s->qf, sizeof(double));
    s->fpost = (double *)calloc(nn * s->qf, sizeof(double));
    s->g = (double *)calloc(nn * s->qg, sizeof(double));
    s->gpost = (double *)calloc(nn * s->qg, sizeof(double));
    double **fields[] = {&s->rho, &s->u, &s->v, &s->w, &s->T, &s->Fx,
                         &s->Fy, &s->Fz, &s->up, &s->vp, &s->wp, &s->Tp};
    for (size_t k = 0; k < sizeof fields / sizeof fields[0]; k++)
        *fields[k] = (double *)calloc(nn, sizeof(double));
    orc_init_rest(s);
    return s;
}

void orc_destroy(orc *s) {
    if (!s) return;
    free(s->x); free(s->st);
    free(s->f); free(s->fpost); free(s->g); free(s->gpost);
    free(s->rho); free(s->u); free(s->v); free(s->w); free(s->T);
    free(s->Fx); free(s->Fy); free(s->Fz);
    free(s->up); free(s->vp); free(s->wp); free(s->Tp);
    free(s);
}

int orc_dim(const orc *s) { return s->dim; }
int orc_n(const orc *s) { return s->n; }
int orc_qf(const orc *s) { return s->qf; }
int orc_qg(const orc *s) { return s->qg; }
long orc_itc(const orc *s) { return s->itc; }
void orc_params(const orc *s, double *out) { memcpy(out, s->p, sizeof s->p); }
void orc_coords(const orc *s, double *out) { memcpy(out, s->x, sizeof(double) * s->n); }
void orc_stencil(const orc *s, int *idx) { memcpy(idx, s->st, sizeof(int) * 3 * s->n); }

double *orc_array(orc *s, const char *name, long *count) {
    long nn = s->nn;
    struct { const char *k; double *p; long c; } tab[] = {
        {"f", s->f, nn * s->qf},     {"fpost", s->fpost, nn * s->qf},
        {"g", s->g, nn * s->qg},     {"gpost", s->gpost, nn * s->qg},
        {"rho", s->rho, nn}, {"u", s->u, nn}, {"v", s->v, nn}, {"w", s->w, nn},
        {"T", s->T, nn},     {"Fx", s->Fx, nn}, {"Fy", s->Fy, nn}, {"Fz", s->Fz, nn},
        {"up", s->up, nn},   {"vp", s->vp, nn}, {"wp", s->wp, nn}, {"Tp", s->Tp, nn}};
    for (size_t k = 0; k < sizeof tab / sizeof tab[0]; k++)
        if (strcmp(tab[k].k, name) == 0) {
            if (count) *count = tab[k].c;
            return tab[k].p;
        }
    return NULL;
}

/* ---------------------------------------------------------------- initial */
/* Equilibria of 2M:227-240 / 3M:286-301 at every node from rho,u,v,w,T. */
static void set_equilibrium(orc *s) {
    double cs2inv_t = s->dim == 2 ? 4.0 : 6.0;
#pragma omp parallel for num_threads(g_threads)
    for (long nd = 0; nd < s->nn; nd++) {
        double uu = s->u[nd], vv = s->v[nd], ww = s->w[nd];
        double u2 = uu * uu + vv * vv + ww * ww;
        for (int a = 0; a < s->qf; a++) {
            double un = uu * s->c[a][0] + vv * s->c[a][1] + ww * s->c[a][2];
            s->f[nd * s->qf + a] =
                s->rho[nd] * s->wf[a] * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * u2);
        }
        for (int a = 0; a < s->qg; a++) {
            double un = uu * s->c[a][0] + vv * s->c[a][1] + ww * s->c[a][2];
            s->g[nd * s->qg + a] = s->wg[a] * s->T[nd] * (1.0 + cs2inv_t * un);
        }
    }
}

static void reset_common(orc *s) {
    s->itc = 0;
    size_t bytes = sizeof(double) * (size_t)s->nn;
    memset(s->up, 0, bytes); memset(s->vp, 0, bytes);
    memset(s->wp, 0, bytes); memset(s->Tp, 0, bytes);
    memset(s->Fx, 0, bytes); memset(s->Fy, 0, bytes); memset(s->Fz, 0, bytes);
    memset(s->fpost, 0, bytes * s->qf);
    memset(s->gpost, 0, bytes * s->qg);
    for (long nd = 0; nd < s->nn; nd++) s->rho[nd] = 1.0;
}

/* 2M:199-212 / 3M:256-271: rest, T linear across the heated direction
 * (x in 2D, y in 3D). */
void orc_init_rest(orc *s) {
    reset_common(s);
    int n = s->n;
    for (long nd = 0; nd < s->nn; nd++) {
        int ih = s->dim == 2 ? (int)(nd % n) : (int)((nd / n) % n);
        s->u[nd] = 0.0; s->v[nd] = 0.0; s->w[nd] = 0.0;
        s->T[nd] = (double)ih / (double)(n - 1) * (s->Tcold - s->Thot) + s->Thot;
    }
    set_equilibrium(s);
}

void orc_init_fields(orc *s, const double *u, const double *v, const double *w,
                     const double *T) {
    reset_common(s);
    for (long nd = 0; nd < s->nn; nd++) {
        s->u[nd] = u ? u[nd] : 0.0;
        s->v[nd] = v ? v[nd] : 0.0;
        s->w[nd] = (w && s->dim == 3) ? w[nd] : 0.0;
        s->T[nd] = T[nd];
    }
    set_equilibrium(s);
}

/* -------------------------------------------------------------- collision */
/* Equilibrium moments (2M:455-463, 3M:522-540). */
static void meq_flow(int dim, double rho, double u, double v, double w, double *me) {
    if (dim == 2) {
        double u2 = u * u + v * v;
        me[0] = rho;
        me[1] = -2.0 * rho + 3.0 * rho * u2;
        me[2] = rho - 3.0 * rho * u2;
        me[3] = rho * u;
        me[4] = -rho * u;
        me[5] = rho * v;
        me[6] = -rho * v;
        me[7] = rho * (u * u - v * v);
        me[8] = rho * (u * v);
    } else {
        double u2 = u * u + v * v + w * w;
        double pxx = 3.0 * u * u - u2, pww = v * v - w * w;
        me[0] = rho;
        me[1] = -11.0 * rho + 19.0 * rho * u2;
        me[2] = 3.0 * rho - 5.5 * rho * u2;
        me[3] = rho * u;
        me[4] = -(2.0 / 3.0) * rho * u;
        me[5] = rho * v;
        me[6] = -(2.0 / 3.0) * rho * v;
        me[7] = rho * w;
        me[8] = -(2.0 / 3.0) * rho * w;
        me[9] = rho * pxx;
        me[10] = -0.5 * rho * pxx;
        me[11] = rho * pww;
        me[12] = -0.5 * rho * pww;
        me[13] = rho * u * v;
        me[14] = rho * v * w;
        me[15] = rho * u * w;
        me[16] = 0.0;
        me[17] = 0.0;
        me[18] = 0.0;
    }
}

/* Guo forcing in moment space, (I - S/2) M F_Guo, for a general force
 * (2M:478-486 and 3M:565-583 are its Fx = 0 / Fx = Fy = 0 cases). */
static void src_flow(int dim, const double *sr, double u, double v, double w,
                     double Fx, double Fy, double Fz, double *src) {
    if (dim == 2) {
        double uF = u * Fx + v * Fy;
        src[0] = 0.0;
        src[1] = 6.0 * (1.0 - 0.5 * sr[1]) * uF;
        src[2] = -6.0 * (1.0 - 0.5 * sr[2]) * uF;
        src[3] = Fx;
        src[4] = -(1.0 - 0.5 * sr[4]) * Fx;
        src[5] = Fy;
        src[6] = -(1.0 - 0.5 * sr[6]) * Fy;
        src[7] = 2.0 * (1.0 - 0.5 * sr[7]) * (u * Fx - v * Fy);
        src[8] = (1.0 - 0.5 * sr[8]) * (u * Fy + v * Fx);
    } else {
        double uF = u * Fx + v * Fy + w * Fz;
        double a = 2.0 * u * Fx - v * Fy - w * Fz, b = v * Fy - w * Fz;
        src[0] = 0.0;
        src[1] = 38.0 * (1.0 - 0.5 * sr[1]) * uF;
        src[2] = -11.0 * (1.0 - 0.5 * sr[2]) * uF;
        src[3] = Fx;
        src[4] = -(2.0 / 3.0) * (1.0 - 0.5 * sr[4]) * Fx;
        src[5] = Fy;
        src[6] = -(2.0 / 3.0) * (1.0 - 0.5 * sr[6]) * Fy;
        src[7] = Fz;
        src[8] = -(2.0 / 3.0) * (1.0 - 0.5 * sr[8]) * Fz;
        src[9] = 2.0 * (1.0 - 0.5 * sr[9]) * a;
        src[10] = -(1.0 - 0.5 * sr[10]) * a;
        src[11] = 2.0 * (1.0 - 0.5 * sr[11]) * b;
        src[12] = -(1.0 - 0.5 * sr[12]) * b;
        src[13] = (1.0 - 0.5 * sr[13]) * (u * Fy + v * Fx);
        src[14] = (1.0 - 0.5 * sr[14]) * (v * Fz + w * Fy);
        src[15] = (1.0 - 0.5 * sr[15]) * (u * Fz + w * Fx);
        src[16] = 0.0;
        src[17] = 0.0;
        src[18] = 0.0;
    }
}

/* collision_U (2M:431-520, 3M:445-669).  Body force F = g*beta*T*rho along
 * +y in 2D (2M:475-476) and +z in 3D (3M:563), with T_ref = 0. */
void orc_collision_u(orc *s) {
    const int q = s->qf, dim = s->dim;
    const double dt = s->p[ORC_P_DT], gb = s->p[ORC_P_GBETA];
#pragma omp parallel for num_threads(g_threads)
    for (long nd = 0; nd < s->nn; nd++) {
        double *f = s->f + nd * q, *fp = s->fpost + nd * q;
        double m[QMAX], me[QMAX], src[QMAX], mp[QMAX];
        for (int r = 0; r < q; r++) {
            double acc = 0.0;
            for (int a = 0; a < q; a++) acc += s->M[r][a] * f[a];
            m[r] = acc;
        }
        double rho = s->rho[nd], uu = s->u[nd], vv = s->v[nd], ww = s->w[nd];
        meq_flow(dim, rho, uu, vv, ww, me);
        double Fb = gb * s->T[nd] * rho;
        s->Fx[nd] = 0.0;
        s->Fy[nd] = dim == 2 ? Fb : 0.0;
        s->Fz[nd] = dim == 3 ? Fb : 0.0;
        src_flow(dim, s->s, uu, vv, ww, s->Fx[nd], s->Fy[nd], s->Fz[nd], src);
        for (int r = 0; r < q; r++)
            mp[r] = m[r] - s->s[r] * (m[r] - me[r]) + src[r] * dt;
        for (int r = 0; r < q; r++) mp[r] = mp[r] / s->norm[r];
        for (int a = 0; a < q; a++) {
            double acc = 0.0;
            for (int r = 0; r < q; r++) acc += s->M[r][a] * mp[r];
            fp[a] = acc;
        }
        f[0] = fp[0]; /* rest population does not stream (2M:513-517) */
    }
}

/* collision_T (2M:522-569, 3M:671-730). */
void orc_collision_t(orc *s) {
    const int q = s->qg, dim = s->dim;
#pragma omp parallel for num_threads(g_threads)
    for (long nd = 0; nd < s->nn; nd++) {
        double *g = s->g + nd * q, *gp = s->gpost + nd * q;
        double nm[7], ne[7], np[7];
        for (int r = 0; r < q; r++) {
            double acc = 0.0;
            for (int a = 0; a < q; a++) acc += s->N[r][a] * g[a];
            nm[r] = acc;
        }
        double T = s->T[nd];
        ne[0] = T;
        ne[1] = T * s->u[nd];
        ne[2] = T * s->v[nd];
        if (dim == 2) {
            ne[3] = T * 0.5;
            ne[4] = 0.0;
        } else {
            ne[3] = T * s->w[nd];
            ne[4] = T * 0.5;
            ne[5] = 0.0;
            ne[6] = 0.0;
        }
        for (int r = 0; r < q; r++) np[r] = nm[r] - s->qrel[r] * (nm[r] - ne[r]);
        for (int a = 0; a < q; a++) {
            double acc = 0.0;
            for (int r = 0; r < q; r++) acc += (double)s->ninv[a][r] * np[r];
            gp[a] = acc / (double)s->ninv_den;
        }
        g[0] = gp[0]; /* 2M:562-566 */
    }
}

/* --------------------------------------------------------- interpolation */
double orc_interpolateF(double x0, double x1, double x2, double x,
                        double f0, double f1, double f2) {
    /* quadratic Lagrange interpolant through (x0,f0),(x1,f1),(x2,f2) at x
     * (2M:649-660) */
    return ((x - x1) * (x - x2)) / ((x0 - x1) * (x0 - x2)) * f0 +
           ((x - x0) * (x - x2)) / ((x1 - x0) * (x1 - x2)) * f1 +
           ((x - x0) * (x - x1)) / ((x2 - x0) * (x2 - x1)) * f2;
}

static inline long node_of(const orc *s, const int *ijk) {
    long n = s->n;
    return s->dim == 2 ? ijk[0] + n * ijk[1] : ijk[0] + n * (ijk[1] + n * (long)ijk[2]);
}

/* 1D interpolation of population a of array P (q per node) along axis ax,
 * at node ijk, for lattice velocity component ca along that axis. */
static double interp1(const orc *s, const double *P, int q, int a, const int *ijk,
                      int ax, int ca) {
    const int *t = s->st + 3 * ijk[ax];
    double d = (double)ca * s->p[ORC_P_DT];
    int id[3] = {ijk[0], ijk[1], ijk[2]};
    double fv[3];
    for (int b = 0; b < 3; b++) {
        id[ax] = t[b];
        fv[b] = P[node_of(s, id) * q + a];
    }
    return orc_interpolateF(s->x[t[0]] + d, s->x[t[1]] + d, s->x[t[2]] + d,
                            s->x[t[1]], fv[0], fv[1], fv[2]);
}

/* Two-stage tensor interpolation: first along axis p at the three stencil
 * positions of axis qa, then along qa (2M:604-617; 3M:767-824). */
static double interp2(const orc *s, const double *P, int q, int a, const int *ijk,
                      int p, int qa) {
    const int *t = s->st + 3 * ijk[qa];
    double vals[3];
    int id[3] = {ijk[0], ijk[1], ijk[2]};
    for (int b = 0; b < 3; b++) {
        id[qa] = t[b];
        vals[b] = interp1(s, P, q, a, id, p, s->c[a][p]);
    }
    double d = (double)s->c[a][qa] * s->p[ORC_P_DT];
    return orc_interpolateF(s->x[t[0]] + d, s->x[t[1]] + d, s->x[t[2]] + d,
                            s->x[t[1]], vals[0], vals[1], vals[2]);
}

/* interpolate / interploate (2M:591-646, 3M:751-862).  2D: every moving
 * population is interpolated along y then x.  3D: a flow direction is
 * interpolated inside its own coordinate plane -- yoz {5,6,15-18}: z then y;
 * xoz {1,2,11-14}: z then x; xoy {3,4,7-10}: y then x -- and a thermal
 * direction along its own axis only (3M:831-859). */
void orc_interpolate(orc *s) {
    const int n = s->n, dim = s->dim;
    const long nn = s->nn;
#pragma omp parallel for num_threads(g_threads)
    for (long nd = 0; nd < nn; nd++) {
        int ijk[3];
        ijk[0] = (int)(nd % n);
        ijk[1] = (int)((nd / n) % n);
        ijk[2] = dim == 3 ? (int)(nd / ((long)n * n)) : 0;
        for (int a = 1; a < s->qf; a++) {
            int p, qa;
            if (dim == 2) { p = 1; qa = 0; }
            else if (s->c[a][0] == 0) { p = 2; qa = 1; }      /* yoz */
            else if (s->c[a][1] == 0) { p = 2; qa = 0; }      /* xoz */
            else { p = 1; qa = 0; }                           /* xoy */
            s->f[nd * s->qf + a] = interp2(s, s->fpost, s->qf, a, ijk, p, qa);
        }
        for (int a = 1; a < s->qg; a++) {
            double val;
            if (dim == 2) {
                val = interp2(s, s->gpost, s->qg, a, ijk, 1, 0);
            } else {
                int ax = s->c[a][0] ? 0 : (s->c[a][1] ? 1 : 2);
                val = interp1(s, s->gpost, s->qg, a, ijk, ax, s->c[a][ax]);
            }
            s->g[nd * s->qg + a] = val;
        }
    }
}

/* ------------------------------------------------------------- boundaries */
static int opposite(const orc *s, int a) {
    for (int b = 0; b < s->qf; b++)
        if (s->c[b][0] == -s->c[a][0] && s->c[b][1] == -s->c[a][1] &&
            s->c[b][2] == -s->c[a][2])
            return b;
    return -1;
}

/* bounceback_u (2M:662-696, 3M:878-943): half-way bounce-back on every wall.
 * At a wall node, each population entering the fluid (c . n_inward > 0)
 * takes the post-collision value of its opposite at the same node. */
void orc_bounceback_u(orc *s) {
    const int n = s->n, dim = s->dim, q = s->qf;
    for (long nd = 0; nd < s->nn; nd++) {
        int ijk[3] = {(int)(nd % n), (int)((nd / n) % n),
                      dim == 3 ? (int)(nd / ((long)n * n)) : 0};
        for (int ax = 0; ax < dim; ax++) {
            int side = ijk[ax] == 0 ? 1 : (ijk[ax] == n - 1 ? -1 : 0);
            if (!side) continue;
            for (int a = 1; a < q; a++)
                if (s->c[a][ax] * side > 0)
                    s->f[nd * q + a] = s->fpost[nd * q + opposite(s, a)];
        }
    }
}

/* bounceback_T (2M:698-724, 3M:945-986): anti-bounce-back Dirichlet walls on
 * the heated axis (x in 2D, y in 3D; hot at index 0, cold at n-1):
 * g(a) = -g_post(opp a) + 2 w_a T_wall; adiabatic bounce-back elsewhere. */
void orc_bounceback_t(orc *s) {
    const int n = s->n, dim = s->dim, q = s->qg;
    const int hax = dim == 2 ? 0 : 1;
    for (long nd = 0; nd < s->nn; nd++) {
        int ijk[3] = {(int)(nd % n), (int)((nd / n) % n),
                      dim == 3 ? (int)(nd / ((long)n * n)) : 0};
        for (int ax = 0; ax < dim; ax++) {
            int side = ijk[ax] == 0 ? 1 : (ijk[ax] == n - 1 ? -1 : 0);
            if (!side) continue;
            for (int a = 1; a < q; a++) {
                if (s->c[a][ax] * side <= 0) continue;
                double gp = s->gpost[nd * q + opposite(s, a)];
                if (ax == hax) {
                    double Tw = side > 0 ? s->Thot : s->Tcold;
                    s->g[nd * q + a] = -gp + 2.0 * s->wg[a] * Tw;
                } else {
                    s->g[nd * q + a] = gp;
                }
            }
        }
    }
}

/* ---------------------------------------------------------------- macros */
/* macro_u (2M:726-739, 3M:988-1010): rho = sum f; u = (sum c f + F dt/2)/rho
 * with the force stored by this step's collision.  In 3D only w carries the
 * force term (3M:1000-1005). */
void orc_macro_u(orc *s) {
    const int q = s->qf;
    const double dt = s->p[ORC_P_DT];
#pragma omp parallel for num_threads(g_threads)
    for (long nd = 0; nd < s->nn; nd++) {
        const double *f = s->f + nd * q;
        double rho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0;
        for (int a = 0; a < q; a++) rho += f[a];
        for (int a = 0; a < q; a++) {
            if (s->c[a][0]) jx += s->c[a][0] > 0 ? f[a] : -f[a];
            if (s->c[a][1]) jy += s->c[a][1] > 0 ? f[a] : -f[a];
            if (s->c[a][2]) jz += s->c[a][2] > 0 ? f[a] : -f[a];
        }
        s->rho[nd] = rho;
        if (s->dim == 2) {
            s->u[nd] = (jx + 0.5 * dt * s->Fx[nd]) / rho;
            s->v[nd] = (jy + 0.5 * dt * s->Fy[nd]) / rho;
            s->w[nd] = 0.0;
        } else {
            s->u[nd] = jx / rho;
            s->v[nd] = jy / rho;
            s->w[nd] = (jz + 0.5 * dt * s->Fz[nd]) / rho;
        }
    }
}

/* macro_t (2M:741-753, 3M:1012-1025). */
void orc_macro_t(orc *s) {
    const int q = s->qg;
#pragma omp parallel for num_threads(g_threads)
    for (long nd = 0; nd < s->nn; nd++) {
        double T = 0.0;
        for (int a = 0; a < q; a++) T += s->g[nd * q + a];
        s->T[nd] = T;
    }
}

void orc_step(orc *s, long nsteps) {
    for (long it = 0; it < nsteps; it++) {
        s->itc++;
        orc_collision_u(s);
        orc_collision_t(s);
        orc_interpolate(s);
        orc_bounceback_u(s);
        orc_bounceback_t(s);
        orc_macro_u(s);
        orc_macro_t(s);
    }
}

/* check (2M:755-805, 3M:1027-1082): relative L2 change since the previous
 * check; the snapshot starts at zero, so the first check gives 1. */
void orc_check(orc *s, double *errU, double *errT) {
    double e1 = 0.0, e2 = 0.0, e3 = 0.0, e4 = 0.0;
    for (long nd = 0; nd < s->nn; nd++) {
        double du = s->u[nd] - s->up[nd], dv = s->v[nd] - s->vp[nd];
        e1 = e1 + du * du + dv * dv;
        e2 = e2 + s->u[nd] * s->u[nd] + s->v[nd] * s->v[nd];
        if (s->dim == 3) {
            double dw = s->w[nd] - s->wp[nd];
            e1 += dw * dw;
            e2 += s->w[nd] * s->w[nd];
        }
        s->up[nd] = s->u[nd];
        s->vp[nd] = s->v[nd];
        s->wp[nd] = s->w[nd];
    }
    *errU = sqrt(e1) / sqrt(e2);
    for (long nd = 0; nd < s->nn; nd++) {
        double dT = s->T[nd] - s->Tp[nd];
        e3 = e3 + dT * dT;
        e4 = e4 + s->T[nd] * s->T[nd];
        s->Tp[nd] = s->T[nd];
    }
    *errT = sqrt(e3) / sqrt(e4);
}

/* Main loop with the reference's stopping rule (2M:307-330): iterate while
 * either residual exceeds its threshold and itc < itc_max, checking every
 * check_every steps.  A NaN residual ends the loop, as in the reference. */
long orc_run(orc *s, double epsU, double epsT, long itc_max, int check_every,
             double *errU, double *errT) {
    double eu = 100.0, et = 100.0;
    while ((eu > epsU || et > epsT) && s->itc < itc_max) {
        orc_step(s, 1);
        if (s->itc % check_every == 0) orc_check(s, &eu, &et);
    }
    if (errU) *errU = eu;
    if (errT) *errT = et;
    return s->itc;
}

/* ---------------------------------------------------------------- Nusselt */
/* Wall-normal derivative at distance 0 of the quadratic through (0, Tw),
 * (d1, T1), (d2, T2). */
static double wall_gradient(double Tw, double d1, double T1, double d2, double T2) {
    return Tw * (-(d1 + d2) / (d1 * d2)) + T1 * d2 / (d1 * (d2 - d1)) -
           T2 * d1 / (d2 * (d2 - d1));
}

/* Control-volume width of node j on [0, H]. */
static double cv_width(const orc *s, int j) {
    int n = s->n;
    double H = s->p[ORC_P_H];
    if (j == 0) return 0.5 * (s->x[0] + s->x[1]);
    if (j == n - 1) return H - 0.5 * (s->x[n - 2] + s->x[n - 1]);
    return 0.5 * (s->x[j + 1] - s->x[j - 1]);
}

/* Mean Nusselt numbers on the two heated walls: Nu = H dT/dn / (Thot-Tcold)
 * with the one-sided quadratic wall gradient, averaged with control-volume
 * weights over the wall (2D: along y; 3D: over the x-z plane). */
void orc_nusselt(const orc *s, double *nu_hot, double *nu_cold) {
    const int n = s->n, dim = s->dim;
    const double H = s->p[ORC_P_H], dT = s->Thot - s->Tcold;
    const double d1h = s->x[0], d2h = s->x[1];
    const double d1c = H - s->x[n - 1], d2c = H - s->x[n - 2];
    double sh = 0.0, sc = 0.0;
    int outer = dim == 2 ? 1 : n;
    for (int k = 0; k < outer; k++) {
        for (int t = 0; t < n; t++) {
            long h1, h2, c1, c2;
            double wgt;
            if (dim == 2) { /* heated x-walls, wall coordinate j = t */
                h1 = 0 + (long)n * t; h2 = 1 + (long)n * t;
                c1 = (n - 1) + (long)n * t; c2 = (n - 2) + (long)n * t;
                wgt = cv_width(s, t) / H;
            } else {        /* heated y-walls, wall coordinates i = t, k */
                h1 = t + (long)n * (0 + (long)n * k);
                h2 = t + (long)n * (1 + (long)n * k);
                c1 = t + (long)n * ((n - 1) + (long)n * k);
                c2 = t + (long)n * ((n - 2) + (long)n * k);
                wgt = cv_width(s, t) / H * (cv_width(s, k) / H);
            }
            double gh = wall_gradient(s->Thot, d1h, s->T[h1], d2h, s->T[h2]);
            double gc = wall_gradient(s->Tcold, d1c, s->T[c1], d2c, s->T[c2]);
            sh += -H * gh / dT * wgt;
            sc += H * gc / dT * wgt;
        }
    }
    *nu_hot = sh;
    *nu_cold = sc;
}
