// lbm_step.h -- fused pull kernel of one time step of the coupled solver,
// templated on dimension, storage type and population shift.  Included by
// lbm_step_f32.hip (FP32 shifted populations, the production path) and
// lbm_step_f64.hip (FP64 populations in the reference's form, compiled
// without FMA contraction).
//
// One launch of k_step maps the reference's loop body (collision_U,
// collision_T, interpolate, bounceback_u, bounceback_T, macro_u, macro_t;
// natural-convection_2dim_mono.f90:311-325, 3dim_mono.f90:377-391) onto one
// pull sweep, rotated so that a launch starts where the populations are
// written: it reads the post-collision populations P(k) of every node's
// 3-point neighbourhood, and per node
//   1. interpolates them onto the node (2dim_mono.f90:591-646,
//      3dim_mono.f90:751-862), giving f(k), g(k);
//   2. applies the wall rules at wall nodes with the node's own P(k)
//      (2dim_mono.f90:662-724, 3dim_mono.f90:878-986);
//   3. forms rho(k), T(k) and u(k) = (sum c f(k) + F(k) dt/2)/rho(k) with
//      F(k) = g beta T(k-1) rho(k-1) (2dim_mono.f90:726-753).  The FP32 path
//      recovers rho(k-1) and T(k-1) as the zeroth moments of the node's own
//      P(k), which the collision conserves, so the one-step force lag costs
//      no storage; the FP64 path keeps F per node like the reference;
//   4. collides (MRT + Guo forcing in moment space, thermal MRT) with
//      F(k+1) = g beta T(k) rho(k) (2dim_mono.f90:431-569) and writes P(k+1).
// On the last launch of a batch the kernel also writes the FP64 macroscopic
// fields that the residual check, the Nusselt number and the export read.
//
// Layout: populations are SoA planes, plane a at p + a*plane, flow planes
// 0..Qf-1 then thermal planes Qf..Qf+Qg-1; node index i + P*j (+ P*n*k),
// with the row pitch P = pitch >= n (3D FP32 storage pads rows to 32
// elements, so that every row starts on a 128-byte line); the FP64
// macroscopic fields are unpadded, i + n*j (+ n*n*k).
// Shifted storage (SHIFT = true) keeps f - w_a and g - w_a*Tref(x), where
// Tref is the linear conduction profile between the heated walls, so FP32
// resolves the small deviations that carry the flow and the near-wall
// temperature that sets the Nusselt number.  Quadratic interpolation streams
// a linear profile exactly, so the shift costs one constant per thermal
// direction along the heated axis.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lbm_lattice.h"

namespace lbm {

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Per-launch constants.  Coefficients that depend only on the relaxation
// rates are formed on the host in FP64, in the same operation order as the
// oracle, and passed in.
// R is the arithmetic type and S the storage type of the populations.
template <typename R, typename S = R> struct StepArgs {
    using real = R;
    using store = S;
    const S *__restrict__ pin;
    S *__restrict__ pout;
    const int2 *__restrict__ off;  // per grid index: offsets of stencil points 0 and 2
    const R *__restrict__ wt;      // per grid index: w(+1)[3], w(-1)[3]
    double *__restrict__ rho;
    double *__restrict__ u;
    double *__restrict__ v;
    double *__restrict__ w;
    double *__restrict__ T;
    R *__restrict__ force;  // FP64 path: per-node body force F(k), in place
    long plane;
    int n;
    int pitch;             // population row pitch in elements (>= n; rows of the
                           // population planes are padded to it, the FP64
                           // fields are not)
    int gx, gy, ntiles;    // tile grid of the launch (x, y tiles; all tiles)
    int remap;             // 1: XCD-aware tile order
    const R *__restrict__ tref;  // shifted path: Tref at each heated-axis index
    R dt, half_dt, gbeta;
    // shifted path, per heated-axis index: interpolation of the stored Tref
    // minus Tref at the node, for c = +1 and c = -1 (exact for the stored
    // FP32 weights and Tref values, formed in FP64 on the host)
    const R *__restrict__ tcorr;
    R tw_hot, tw_cold;     // T_wall - Tref(wall node) (plain path: T_wall)
    R s_nu, s_q, sig_k, q_e;
    // Guo source coefficients (1 - s/2) times the lattice factor
    R c_e, c_eps, c_q, c_pxx, c_pxy;     // 2D: 6h, -6h, -(h_q), 2h, h
    R c3_e, c3_eps, c3_q, c3_9, c3_10, c3_11, c3_12, c3_14, c3_15;  // 3D
};

// ----------------------------------------------------------- collisions
// Equilibrium moments with a selectable "base" density for the parts that
// are linear in rho: base = rho (plain) or rho - 1 (shifted storage).
template <int DIM, typename R>
__device__ __forceinline__ void flow_meq(R base, R rho, R u, R v, R w, R *me) {
    if constexpr (DIM == 2) {
        R u2 = u * u + v * v;
        me[0] = base;
        me[1] = R(-2.0) * base + R(3.0) * rho * u2;
        me[2] = base - R(3.0) * rho * u2;
        me[3] = rho * u;
        me[4] = -rho * u;
        me[5] = rho * v;
        me[6] = -rho * v;
        me[7] = rho * (u * u - v * v);
        me[8] = rho * (u * v);
    } else {
        R u2 = u * u + v * v + w * w;
        R pxx = R(3.0) * u * u - u2, pww = v * v - w * w;
        me[0] = base;
        me[1] = R(-11.0) * base + R(19.0) * rho * u2;
        me[2] = R(3.0) * base - R(5.5) * rho * u2;
        me[3] = rho * u;
        me[4] = R(-(2.0 / 3.0)) * rho * u;
        me[5] = rho * v;
        me[6] = R(-(2.0 / 3.0)) * rho * v;
        me[7] = rho * w;
        me[8] = R(-(2.0 / 3.0)) * rho * w;
        me[9] = rho * pxx;
        me[10] = R(-0.5) * rho * pxx;
        me[11] = rho * pww;
        me[12] = R(-0.5) * rho * pww;
        me[13] = rho * u * v;
        me[14] = rho * v * w;
        me[15] = rho * u * w;
        me[16] = R(0);
        me[17] = R(0);
        me[18] = R(0);
    }
}

// Moment rows with a possibly non-zero collision increment: every relaxed
// row, plus the momentum row along the body force (j_y in 2D, j_z in 3D).
template <int DIM> constexpr bool flow_row_active(int r) {
    if constexpr (DIM == 2) return r != 0 && r != 3;
    else return r != 0 && r != 3 && r != 5;
}

// Opposite pairs (a, opp a) with a < opp a, in increasing a.  Every MRT row
// is either even (same coefficient for a and opp a) or odd (opposite
// coefficients) in the lattice velocity, so the FP32 path transforms pair
// sums s = f_a + f_opp and differences d = f_a - f_opp.
template <int DIM> constexpr int npairs() { return (Flow<DIM>::Q - 1) / 2; }
template <int DIM> constexpr int pair_first(int p) {
    int cnt = 0;
    for (int a = 1; a < Flow<DIM>::Q; a++)
        if (a < opposite<DIM>(a)) {
            if (cnt == p) return a;
            cnt++;
        }
    return -1;
}
template <int DIM> constexpr bool row_even(int r) {
    for (int a = 1; a < Flow<DIM>::Q; a++)
        if (Flow<DIM>::row(r, a) != Flow<DIM>::row(r, opposite<DIM>(a))) return false;
    return true;
}
// Rows relaxed towards equilibrium (s_r != 0).
template <int DIM> constexpr bool row_relaxed(int r) {
    if constexpr (DIM == 2) return r != 0 && r != 3 && r != 5;
    else return r != 0 && r != 3 && r != 5 && r != 7;
}

// m = M f, relax towards m_eq, add (I - S/2) M F_Guo dt, f_post = M^-1 m.
// The body force is along +y (2D) or +z (3D) with magnitude Fb.
// EXACT selects the operation order of the FP64 oracle (FP64 storage);
// otherwise the increment form on opposite pairs is used.
template <int DIM, typename R, bool EXACT, typename A>
__device__ __forceinline__ void collide_flow(const R *f, R base, R rho, R u, R v,
                                             R w, R Fb, const A &a, R *fp) {
    using FL = Flow<DIM>;
    constexpr int Q = FL::Q;
    constexpr int NP = npairs<DIM>();
    R m[Q], me[Q], src[Q];
    R ps[NP], pd[NP];
    if constexpr (EXACT) {
        static_for<0, Q>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            R acc = R(0);
            static_for<0, Q>([&](auto Aa) {
                constexpr int al = decltype(Aa)::value;
                constexpr double co = FL::row(r, al);
                if constexpr (co == 1.0) acc += f[al];
                else if constexpr (co == -1.0) acc -= f[al];
                else if constexpr (co != 0.0) acc += R(co) * f[al];
            });
            m[r] = acc;
        });
    } else {
        static_for<0, NP>([&](auto P) {
            constexpr int p = decltype(P)::value;
            constexpr int al = pair_first<DIM>(p), op = opposite<DIM>(al);
            ps[p] = f[al] + f[op];
            pd[p] = f[al] - f[op];
        });
        static_for<0, Q>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            if constexpr (row_relaxed<DIM>(r)) {
                constexpr bool ev = row_even<DIM>(r);
                constexpr double c0 = FL::row(r, 0);
                R acc = R(0);
                if constexpr (ev && c0 != 0.0) acc = R(c0) * f[0];
                static_for<0, NP>([&](auto P) {
                    constexpr int p = decltype(P)::value;
                    constexpr double co = FL::row(r, pair_first<DIM>(p));
                    const R x = ev ? ps[p] : pd[p];
                    if constexpr (co == 1.0) acc += x;
                    else if constexpr (co == -1.0) acc -= x;
                    else if constexpr (co != 0.0) acc += R(co) * x;
                });
                m[r] = acc;
            }
        });
    }
    flow_meq<DIM, R>(base, rho, u, v, w, me);
    R s[Q];
    if constexpr (DIM == 2) {
        // S = diag(0, Snu, Snu, 0, Sq, 0, Sq, Snu, Snu)  (2dim_mono.f90:465-473)
        const R sn = a.s_nu, sq = a.s_q;
        s[0] = 0; s[1] = sn; s[2] = sn; s[3] = 0; s[4] = sq;
        s[5] = 0; s[6] = sq; s[7] = sn; s[8] = sn;
        // Guo source with F = (0, Fb)  (2dim_mono.f90:478-486)
        R uF = v * Fb;
        src[0] = 0;
        src[1] = a.c_e * uF;
        src[2] = a.c_eps * uF;
        src[3] = 0;
        src[4] = 0;
        src[5] = Fb;
        src[6] = a.c_q * Fb;
        src[7] = a.c_pxx * (-(v * Fb));
        src[8] = a.c_pxy * (u * Fb);
    } else {
        // 3dim_mono.f90:542-560
        const R sn = a.s_nu, sq = a.s_q;
        s[0] = 0; s[1] = sn; s[2] = sn; s[3] = 0; s[4] = sq; s[5] = 0;
        s[6] = sq; s[7] = 0; s[8] = sq;
        s[9] = sn; s[10] = sn; s[11] = sn; s[12] = sn; s[13] = sn; s[14] = sn;
        s[15] = sn; s[16] = sq; s[17] = sq; s[18] = sq;
        // Guo source with F = (0, 0, Fb)  (3dim_mono.f90:563-583)
        R uF = w * Fb, am = -(w * Fb);
        src[0] = 0; src[1] = a.c3_e * uF; src[2] = a.c3_eps * uF;
        src[3] = 0; src[4] = 0; src[5] = 0; src[6] = 0;
        src[7] = Fb; src[8] = a.c3_q * Fb;
        src[9] = a.c3_9 * am; src[10] = a.c3_10 * am;
        src[11] = a.c3_11 * am; src[12] = a.c3_12 * am;
        src[13] = 0; src[14] = a.c3_14 * (v * Fb); src[15] = a.c3_15 * (u * Fb);
        src[16] = 0; src[17] = 0; src[18] = 0;
    }
    if constexpr (EXACT) {
        // FP64: m_post = m - S(m - m_eq) + src dt, f_post = M^-1 m_post, in
        // the operation order of the FP64 oracle.
        static_for<0, Q>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            R mp = m[r] - s[r] * (m[r] - me[r]) + src[r] * a.dt;
            m[r] = mp / flow_norm<DIM>(r);
        });
        static_for<0, Q>([&](auto Aa) {
            constexpr int al = decltype(Aa)::value;
            R acc = R(0);
            static_for<0, Q>([&](auto Rr) {
                constexpr int r = decltype(Rr)::value;
                constexpr double co = FL::row(r, al);
                if constexpr (co == 1.0) acc += m[r];
                else if constexpr (co == -1.0) acc -= m[r];
                else if constexpr (co != 0.0) acc += R(co) * m[r];
            });
            fp[al] = acc;
        });
    } else {
        // the same update written as an increment,
        // f_post = f + M^-1 (-S(m - m_eq) + src dt).  The conserved moments
        // have a zero increment, so mass and momentum pass through exactly
        // instead of being scaled every step by the rounded 1/|row|^2.
        static_for<0, Q>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            if constexpr (flow_row_active<DIM>(r)) {
                if constexpr (row_relaxed<DIM>(r))
                    m[r] = (src[r] * a.dt - s[r] * (m[r] - me[r])) * R(1.0 / flow_norm<DIM>(r));
                else
                    m[r] = src[r] * a.dt * R(1.0 / flow_norm<DIM>(r));
            }
        });
        // inverse transform on pairs: f_a += E + O, f_opp += E - O with E the
        // even rows' and O the odd rows' contribution
        R e0 = R(0);
        static_for<0, Q>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            constexpr double c0 = FL::row(r, 0);
            if constexpr (flow_row_active<DIM>(r) && c0 != 0.0) e0 += R(c0) * m[r];
        });
        fp[0] = f[0] + e0;
        static_for<0, NP>([&](auto P) {
            constexpr int p = decltype(P)::value;
            constexpr int al = pair_first<DIM>(p), op = opposite<DIM>(al);
            R ev = R(0), od = R(0);
            static_for<0, Q>([&](auto Rr) {
                constexpr int r = decltype(Rr)::value;
                constexpr double co = FL::row(r, al);
                if constexpr (flow_row_active<DIM>(r) && co != 0.0) {
                    R &acc = row_even<DIM>(r) ? ev : od;
                    if constexpr (co == 1.0) acc += m[r];
                    else if constexpr (co == -1.0) acc -= m[r];
                    else acc += R(co) * m[r];
                }
            });
            fp[al] = f[al] + (ev + od);
            fp[op] = f[op] + (ev - od);
        });
    }
}

// n = N g, relax towards n_eq, g_post = N^-1 n.  baseT = T (plain) or
// T - Tref (shifted).  (2dim_mono.f90:522-569, 3dim_mono.f90:671-730)
template <int DIM, typename R, bool EXACT, typename A>
__device__ __forceinline__ void collide_heat(const R *g, R baseT, R T, R u, R v,
                                             R w, const A &a, R *gp) {
    using HT = Heat<DIM>;
    constexpr int Q = HT::Q;
    R nm[Q], ne[Q], q[Q];
    static_for<0, Q>([&](auto Rr) {
        constexpr int r = decltype(Rr)::value;
        R acc = R(0);
        static_for<0, Q>([&](auto Aa) {
            constexpr int al = decltype(Aa)::value;
            constexpr double co = HT::row(r, al);
            if constexpr (co == 1.0) acc += g[al];
            else if constexpr (co == -1.0) acc -= g[al];
            else if constexpr (co != 0.0) acc += R(co) * g[al];
        });
        nm[r] = acc;
    });
    ne[0] = baseT;
    ne[1] = T * u;
    ne[2] = T * v;
    q[0] = R(1.0);
    q[1] = a.sig_k;
    q[2] = a.sig_k;
    if constexpr (DIM == 2) {
        ne[3] = baseT * R(0.5);
        ne[4] = R(0);
        q[3] = a.q_e;
        q[4] = a.q_e;
    } else {
        ne[3] = T * w;
        ne[4] = baseT * R(0.5);
        ne[5] = R(0);
        ne[6] = R(0);
        q[3] = a.sig_k;
        q[4] = a.q_e;
        q[5] = a.q_e;
        q[6] = a.q_e;
    }
    if constexpr (EXACT) {
        static_for<0, Q>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            nm[r] = nm[r] - q[r] * (nm[r] - ne[r]);
        });
        static_for<0, Q>([&](auto Aa) {
            constexpr int al = decltype(Aa)::value;
            R acc = R(0);
            static_for<0, Q>([&](auto Rr) {
                constexpr int r = decltype(Rr)::value;
                constexpr int co = HT::ninv[al][r];
                if constexpr (co == 1) acc += nm[r];
                else if constexpr (co == -1) acc -= nm[r];
                else if constexpr (co != 0) acc += R(co) * nm[r];
            });
            gp[al] = acc / double(HT::den);
        });
    } else {
        // increment form; row 0 (temperature) is conserved exactly because
        // n_0 and n_eq,0 are the same sum
        static_for<1, Q>([&](auto Rr) {
            constexpr int r = decltype(Rr)::value;
            nm[r] = -q[r] * (nm[r] - ne[r]);
        });
        static_for<0, Q>([&](auto Aa) {
            constexpr int al = decltype(Aa)::value;
            R acc = R(0);
            static_for<1, Q>([&](auto Rr) {
                constexpr int r = decltype(Rr)::value;
                constexpr int co = HT::ninv[al][r];
                if constexpr (co == 1) acc += nm[r];
                else if constexpr (co == -1) acc -= nm[r];
                else if constexpr (co != 0) acc += R(co) * nm[r];
            });
            gp[al] = g[al] + acc * R(1.0 / HT::den);
        });
    }
}

// ------------------------------------------------------- interpolation
// Quadratic Lagrange weights are precomputed per (grid index, sign of the
// lattice velocity) in the reference's formula (2dim_mono.f90:649-660);
// a zero velocity component gives the weights (0, 1, 0), so the reference's
// trivial interpolation stages reduce exactly to copies and are skipped.
template <typename R> struct AxisStencil {
    int o0, o2;     // offsets of stencil points 0 and 2 (point 1 is the node)
    R wp[3], wm[3]; // weights for c = +1 and c = -1
};

template <typename A>
__device__ __forceinline__ AxisStencil<typename A::real> load_stencil(const A &a, int idx) {
    using R = typename A::real;
    AxisStencil<R> s;
    int2 o = a.off[idx];
    s.o0 = o.x;
    s.o2 = o.y;
    const R *wt = a.wt + 6 * idx;
#pragma unroll
    for (int b = 0; b < 3; b++) {
        s.wp[b] = wt[b];
        s.wm[b] = wt[3 + b];
    }
    return s;
}

// FP64: the reference's sum w0 f0 + w1 f1 + w2 f2.  FP32: the same value as
// f1 + w0 (f0 - f1) + w2 (f2 - f1) (the weights sum to one), so that the
// rounding of the stored weights acts on differences between neighbours
// only and a uniform field streams unchanged.
template <bool EXACT, typename R>
__device__ __forceinline__ R lerp3(const R *w, R f0, R f1, R f2) {
    if constexpr (EXACT) return w[0] * f0 + w[1] * f1 + w[2] * f2;
    else return f1 + w[0] * (f0 - f1) + w[2] * (f2 - f1);
}

// Element idx of a population plane: a uniform 64-bit base plus a 32-bit
// byte offset, so the load is a scalar-base + vector-offset access and the
// stencil offsets common to all populations are formed once per node.  The
// caller bounds a plane to fewer than 2^32 bytes.
template <typename R>
__device__ __forceinline__ R ldp(const R *__restrict__ base, uint32_t idx) {
    return *reinterpret_cast<const R *>(reinterpret_cast<const char *>(base) +
                                        idx * (uint32_t)sizeof(R));
}
// Workgroup barrier for LDS handoffs: orders LDS accesses only.  The march
// kernels exchange data between threads through LDS alone (global loads go
// to the loading thread's registers, and each thread stores its own nodes),
// so the barrier need not wait for outstanding global loads or stores, as
// __syncthreads does; the prefetch and the stores stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// NT: a streaming (non-temporal) store, for states larger than the MALL:
// the written planes are not read again within the step, so they should not
// displace the rows that neighbouring workgroups are about to re-read.
template <bool NT = false, typename R>
__device__ __forceinline__ void stp(R *__restrict__ base, uint32_t idx, R v) {
    R *p = reinterpret_cast<R *>(reinterpret_cast<char *>(base) + idx * (uint32_t)sizeof(R));
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// One population: a 1D interpolation when one velocity component is
// non-zero, a two-stage tensor interpolation (first along the higher axis,
// then along the lower one) when two are.
template <int CX, int CY, int CZ, bool EXACT, typename S, typename R>
__device__ __forceinline__ R interp_pop(const S *__restrict__ F, uint32_t node, R fc,
                                        const AxisStencil<R> *st, const uint32_t *stride) {
    constexpr int c[3] = {CX, CY, CZ};
    constexpr int nz = (CX != 0) + (CY != 0) + (CZ != 0);
    if constexpr (nz == 1) {
        constexpr int ax = CX ? 0 : (CY ? 1 : 2);
        const AxisStencil<R> &s = st[ax];
        const R *w = c[ax] > 0 ? s.wp : s.wm;
        return lerp3<EXACT>(w, R(ldp(F, node + (uint32_t)s.o0 * stride[ax])), fc,
                            R(ldp(F, node + (uint32_t)s.o2 * stride[ax])));
    } else {
        constexpr int q = CX ? 0 : 1;          // second (outer) stage axis
        constexpr int p = CZ ? 2 : 1;          // first (inner) stage axis
        const AxisStencil<R> &sp = st[p], &sq = st[q];
        const R *wpp = c[p] > 0 ? sp.wp : sp.wm;
        const R *wqq = c[q] > 0 ? sq.wp : sq.wm;
        const uint32_t dp0 = (uint32_t)sp.o0 * stride[p], dp2 = (uint32_t)sp.o2 * stride[p];
        const uint32_t n0 = node + (uint32_t)sq.o0 * stride[q];
        const uint32_t n2 = node + (uint32_t)sq.o2 * stride[q];
        R c0 = lerp3<EXACT>(wpp, R(ldp(F, n0 + dp0)), R(ldp(F, n0)), R(ldp(F, n0 + dp2)));
        R c1 = lerp3<EXACT>(wpp, R(ldp(F, node + dp0)), fc, R(ldp(F, node + dp2)));
        R c2 = lerp3<EXACT>(wpp, R(ldp(F, n2 + dp0)), R(ldp(F, n2)), R(ldp(F, n2 + dp2)));
        return lerp3<EXACT>(wqq, c0, c1, c2);
    }
}

// ---------------------------------------------------------------- step
// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  With remap on, hardware workgroup P runs logical tile
// (P % 8) * (G/8) + P/8, so each XCD sweeps one contiguous range of tiles and
// the neighbour tiles whose rows it re-reads sit in its own L2.  The tail
// G % 8 tiles keep their identity.
__device__ __forceinline__ int logical_tile(int P, int G, int remap) {
    if (remap != 1) return P;
    const int g8 = G >> 3;
    if (P >= (g8 << 3)) return P;
    return (P & 7) * g8 + (P >> 3);
}

// Tile coordinates (tx, ty, tz) of hardware workgroup P.  remap 0 and 1 use
// the x-fastest order of logical_tile.  remap 2 (needs gy % 8 == 0) gives XCD
// P % 8 the band of gy/8 tile rows [x*gy/8, (x+1)*gy/8) and sweeps it
// x fastest, then y, then z: x and y neighbour tiles share the XCD's L2 and
// the tile above in z comes gx*gy/8 workgroups later on the same XCD.
__device__ __forceinline__ void tile_coords(int P, int G, int gx, int gy, int remap, int &tx,
                                            int &ty, int &tz) {
    if (remap == 2) {
        const int band = gy >> 3, q = P >> 3;
        tx = q % gx;
        const int r = q / gx;
        ty = (P & 7) * band + r % band;
        tz = r / band;
        return;
    }
    const int L = logical_tile(P, G, remap);
    tx = L % gx;
    const int r = L / gx;
    ty = r % gy;
    tz = r / gy;
}

// The shifted path's per-node constants along the heated axis: Tref at the
// node and the streaming corrections for c = +1 and c = -1.
template <typename R> struct HeatShift {
    R tloc, tc_plus, tc_minus;
};
template <int DIM, bool SHIFT, typename R, typename S>
__device__ __forceinline__ HeatShift<R> load_heat_shift(const StepArgs<R, S> &a,
                                                        const int (&idx)[3]) {
    HeatShift<R> h{R(0), R(0), R(0)};
    if constexpr (SHIFT) {
        const int t = idx[heated_axis<DIM>()];
        h.tloc = a.tref[t];
        h.tc_plus = a.tcorr[2 * t];
        h.tc_minus = a.tcorr[2 * t + 1];
    }
    return h;
}

// Step 2 for the flow populations: the wall rules at wall nodes
// (half-way bounce-back, 2dim_mono.f90:662-696, 3dim_mono.f90:878-943),
// with the node's own P(k) in pcf.
template <int DIM, typename R>
__device__ __forceinline__ void node_walls_flow(int n, const int (&idx)[3], const R *pcf, R *f) {
    using FL = Flow<DIM>;
    constexpr int QF = FL::Q;
    static_for<0, DIM>([&](auto AX) {
        constexpr int ax = decltype(AX)::value;
        const int side = idx[ax] == 0 ? 1 : (idx[ax] == n - 1 ? -1 : 0);
        if (side != 0) {
            static_for<1, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                constexpr int ca = FL::c[al][ax];
                if constexpr (ca != 0) {
                    if (ca * side > 0) f[al] = pcf[opposite<DIM>(al)];
                }
            });
        }
    });
}

// Step 2 for the thermal populations: the shifted path's streaming
// correction, then the wall rules (hot/cold anti-bounce-back on the heated
// axis, adiabatic bounce-back elsewhere; 2dim_mono.f90:698-724,
// 3dim_mono.f90:945-986) with the node's own P(k) in pcg.
template <int DIM, bool SHIFT, typename R, typename S>
__device__ __forceinline__ void node_walls_heat(const StepArgs<R, S> &a, const int (&idx)[3],
                                                const R *pcg, R *g, const HeatShift<R> &hs) {
    using FL = Flow<DIM>;
    using HT = Heat<DIM>;
    constexpr int QG = HT::Q;
    const int n = a.n;
    constexpr int HAX = heated_axis<DIM>();
    if constexpr (SHIFT) {
        static_for<1, QG>([&](auto A) {
            constexpr int al = decltype(A)::value;
            if constexpr (FL::c[al][HAX] != 0) {
                constexpr double wa = HT::w[al];
                g[al] += R(wa) * (FL::c[al][HAX] > 0 ? hs.tc_plus : hs.tc_minus);
            }
        });
    }
    static_for<0, DIM>([&](auto AX) {
        constexpr int ax = decltype(AX)::value;
        const int side = idx[ax] == 0 ? 1 : (idx[ax] == n - 1 ? -1 : 0);
        if (side != 0) {
            static_for<1, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                constexpr int ca = FL::c[al][ax];
                if constexpr (ca != 0) {
                    if (ca * side > 0) {
                        R gpo = pcg[opposite<DIM>(al)];
                        if constexpr (ax == heated_axis<DIM>()) {
                            constexpr double w2 = 2.0 * HT::w[al];
                            R tw = side > 0 ? a.tw_hot : a.tw_cold;
                            g[al] = -gpo + R(w2) * tw;
                        } else {
                            g[al] = gpo;
                        }
                    }
                }
            });
        }
    });
}

// The lagged force F(k) of the FP32 path from the sums of the node's own
// P(k), i.e. from rho(k-1) and T(k-1), which the collision conserves.
template <bool SHIFT, typename R>
__device__ __forceinline__ R lagged_force(R gbeta, R sp_f, R sp_g, R tloc) {
    const R rho_prev = SHIFT ? R(1) + sp_f : sp_f;
    const R T_prev = SHIFT ? tloc + sp_g : sp_g;
    return gbeta * T_prev * rho_prev;
}

// Steps 3 and 4 of a node: the macroscopic fields of this step with the
// lagged force F_prev (2dim_mono.f90:726-753), written to the FP64 fields
// on MACRO steps, and the collision with the force of this step
// (2dim_mono.f90:431-569), giving the node's P(k+1) in fp, gp.  fnode is
// the node's field index.
template <int DIM, typename S, typename R, bool SHIFT, bool MACRO, bool NT = false>
__device__ __forceinline__ void node_collide(const StepArgs<R, S> &a, uint32_t fnode, const R *f,
                                             const R *g, R F_prev, R tloc, R *fp, R *gp) {
    constexpr bool EXACT = std::is_same<S, double>::value;
    using FL = Flow<DIM>;
    using HT = Heat<DIM>;
    constexpr int QF = FL::Q, QG = HT::Q;
    R srho = R(0), jx = R(0), jy = R(0), jz = R(0), sT = R(0);
    if constexpr (EXACT) {
        static_for<0, QF>([&](auto A) { srho += f[decltype(A)::value]; });
        static_for<1, QF>([&](auto A) {
            constexpr int al = decltype(A)::value;
            if constexpr (FL::c[al][0] > 0) jx += f[al];
            if constexpr (FL::c[al][0] < 0) jx -= f[al];
            if constexpr (FL::c[al][1] > 0) jy += f[al];
            if constexpr (FL::c[al][1] < 0) jy -= f[al];
            if constexpr (FL::c[al][2] > 0) jz += f[al];
            if constexpr (FL::c[al][2] < 0) jz -= f[al];
        });
    } else {
        // on opposite pairs, shared with the collision's moment transform
        srho = f[0];
        static_for<0, npairs<DIM>()>([&](auto P) {
            constexpr int p = decltype(P)::value;
            constexpr int al = pair_first<DIM>(p), op = opposite<DIM>(al);
            const R sm = f[al] + f[op], df = f[al] - f[op];
            srho += sm;
            if constexpr (FL::c[al][0] > 0) jx += df;
            if constexpr (FL::c[al][0] < 0) jx -= df;
            if constexpr (FL::c[al][1] > 0) jy += df;
            if constexpr (FL::c[al][1] < 0) jy -= df;
            if constexpr (FL::c[al][2] > 0) jz += df;
            if constexpr (FL::c[al][2] < 0) jz -= df;
        });
    }
    static_for<0, QG>([&](auto A) { sT += g[decltype(A)::value]; });
    const R rho = SHIFT ? R(1) + srho : srho;
    const R T = SHIFT ? tloc + sT : sT;
    R ux, uy, uz;
    if constexpr (EXACT) {
        if constexpr (DIM == 2) {
            ux = jx / rho;
            uy = (jy + a.half_dt * F_prev) / rho;
            uz = R(0);
        } else {
            ux = jx / rho;
            uy = jy / rho;
            uz = (jz + a.half_dt * F_prev) / rho;
        }
    } else {
        const R irho = R(1) / rho;
        ux = jx * irho;
        if constexpr (DIM == 2) {
            uy = (jy + a.half_dt * F_prev) * irho;
            uz = R(0);
        } else {
            uy = jy * irho;
            uz = (jz + a.half_dt * F_prev) * irho;
        }
    }
    if constexpr (MACRO) {
        if constexpr (SHIFT) {
            // FP64 sums of the stored deviations for the exported fields
            double dr = 0.0, dT = 0.0;
            static_for<0, QF>([&](auto A) { dr += double(f[decltype(A)::value]); });
            static_for<0, QG>([&](auto A) { dT += double(g[decltype(A)::value]); });
            a.rho[fnode] = 1.0 + dr;
            a.T[fnode] = double(tloc) + dT;
        } else {
            a.rho[fnode] = double(rho);
            a.T[fnode] = double(T);
        }
        a.u[fnode] = double(ux);
        a.v[fnode] = double(uy);
        if constexpr (DIM == 3) a.w[fnode] = double(uz);
    }

    // 4. collision with the force of this step
    const R Fb = a.gbeta * T * rho;
    if constexpr (EXACT) stp<NT>(a.force, fnode, Fb);
    collide_flow<DIM, R, EXACT>(f, SHIFT ? srho : rho, rho, ux, uy, uz, Fb, a, fp);
    collide_heat<DIM, R, EXACT>(g, SHIFT ? sT : T, T, ux, uy, uz, a, gp);
}

// Steps 2-4 of a node once f(k), g(k) are interpolated: the thermal shift
// correction, the wall rules (with the node's own P(k) in pcf, pcg), the
// macroscopic fields with the lagged force and the collision, whose result
// P(k+1) is left in fp, gp.  sp_f, sp_g are the sums of the node's own P(k)
// (the zeroth moments that carry rho(k-1), T(k-1) on the FP32 path).  fnode:
// the node's field index.
template <int DIM, typename S, typename R, bool SHIFT, bool MACRO, bool NT = false>
__device__ __forceinline__ void node_tail_collide(const StepArgs<R, S> &a, uint32_t fnode,
                                                  const int (&idx)[3], const R *pcf,
                                                  const R *pcg, R *f, R *g, R sp_f, R sp_g,
                                                  const HeatShift<R> &hs, R *fp, R *gp) {
    constexpr bool EXACT = std::is_same<S, double>::value;
    using FL = Flow<DIM>;
    using HT = Heat<DIM>;
    constexpr int QF = FL::Q, QG = HT::Q;
    const int n = a.n;
    constexpr int HAX = heated_axis<DIM>();
    // node_walls_heat and node_walls_flow, interleaved per axis (the
    // instruction order the compiler schedules with the fewest registers)
    if constexpr (SHIFT) {
        static_for<1, QG>([&](auto A) {
            constexpr int al = decltype(A)::value;
            if constexpr (FL::c[al][HAX] != 0) {
                constexpr double wa = HT::w[al];
                g[al] += R(wa) * (FL::c[al][HAX] > 0 ? hs.tc_plus : hs.tc_minus);
            }
        });
    }
    static_for<0, DIM>([&](auto AX) {
        constexpr int ax = decltype(AX)::value;
        const int side = idx[ax] == 0 ? 1 : (idx[ax] == n - 1 ? -1 : 0);
        if (side != 0) {
            static_for<1, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                constexpr int ca = FL::c[al][ax];
                if constexpr (ca != 0) {
                    if (ca * side > 0) f[al] = pcf[opposite<DIM>(al)];
                }
            });
            static_for<1, QG>([&](auto A) {
                constexpr int al = decltype(A)::value;
                constexpr int ca = FL::c[al][ax];
                if constexpr (ca != 0) {
                    if (ca * side > 0) {
                        R gpo = pcg[opposite<DIM>(al)];
                        if constexpr (ax == heated_axis<DIM>()) {
                            constexpr double w2 = 2.0 * HT::w[al];
                            R tw = side > 0 ? a.tw_hot : a.tw_cold;
                            g[al] = -gpo + R(w2) * tw;
                        } else {
                            g[al] = gpo;
                        }
                    }
                }
            });
        }
    });
    // 3. the force of the previous step.  The FP64 path keeps F(k) per
    // node, as the reference does (Fx/Fy/Fz, 2dim_mono.f90:47); the FP32
    // path recovers it from the conserved zeroth moments of the node's own
    // P(k).
    R F_prev;
    if constexpr (EXACT) F_prev = ldp(a.force, fnode);
    else F_prev = lagged_force<SHIFT>(a.gbeta, sp_f, sp_g, hs.tloc);
    node_collide<DIM, S, R, SHIFT, MACRO, NT>(a, fnode, f, g, F_prev, hs.tloc, fp, gp);
}

// The stores of a node's P(k+1); node: its population index.
template <int DIM, typename S, bool NT, typename R>
__device__ __forceinline__ void node_store(S *pout, long plane, uint32_t node, const R *fp,
                                           const R *gp) {
    constexpr int QF = Flow<DIM>::Q, QG = Heat<DIM>::Q;
    static_for<0, QF>([&](auto A) {
        constexpr int al = decltype(A)::value;
        stp<NT>(pout + al * plane, node, S(fp[al]));
    });
    static_for<0, QG>([&](auto A) {
        constexpr int al = decltype(A)::value;
        stp<NT>(pout + (QF + al) * plane, node, S(gp[al]));
    });
}

// node_tail_collide, then the stores.
template <int DIM, typename S, typename R, bool SHIFT, bool MACRO, bool NT = false>
__device__ __forceinline__ void node_tail(const StepArgs<R, S> &a, uint32_t node, uint32_t fnode,
                                          const int (&idx)[3], const R *pcf, const R *pcg,
                                          R *f, R *g, R sp_f, R sp_g, const HeatShift<R> &hs) {
    constexpr int QF = Flow<DIM>::Q, QG = Heat<DIM>::Q;
    R fp[QF], gp[QG];
    node_tail_collide<DIM, S, R, SHIFT, MACRO, NT>(a, fnode, idx, pcf, pcg, f, g, sp_f, sp_g, hs,
                                                   fp, gp);
    static_for<0, QF>([&](auto A) {
        constexpr int al = decltype(A)::value;
        stp<NT>(a.pout + al * a.plane, node, S(fp[al]));
    });
    static_for<0, QG>([&](auto A) {
        constexpr int al = decltype(A)::value;
        stp<NT>(a.pout + (QF + al) * a.plane, node, S(gp[al]));
    });
}

// WPE > 0 asks the compiler for at least WPE resident waves per SIMD (it
// then caps the registers per lane), trading load hoisting for latency
// hiding; the autotuner picks.
// One node (i, j, k) of the fused step (the node must be inside the grid).
template <int DIM, typename S, typename R, bool SHIFT, bool MACRO, bool NT = false>
__device__ __forceinline__ void step_node(const StepArgs<R, S> &a, int i, int j, int k) {
    constexpr bool EXACT = std::is_same<S, double>::value;
    using FL = Flow<DIM>;
    using HT = Heat<DIM>;
    constexpr int QF = FL::Q, QG = HT::Q;
    const int n = a.n;
    const uint32_t sy = a.pitch, sz = (uint32_t)a.pitch * n;
    const uint32_t node = i + sy * j + sz * k;
    const uint32_t fnode = i + (uint32_t)n * j + (uint32_t)n * n * k;
    const int idx[3] = {i, j, k};
    const uint32_t stride[3] = {1, sy, sz};
    AxisStencil<R> st[3];
    st[0] = load_stencil(a, i);
    st[1] = load_stencil(a, j);
    if constexpr (DIM == 3) st[2] = load_stencil(a, k);

    R pcf[QF], pcg[QG], f[QF], g[QG];
    static_for<0, QF>([&](auto A) {
        constexpr int al = decltype(A)::value;
        pcf[al] = R(ldp(a.pin + al * a.plane, node));
    });
    static_for<0, QG>([&](auto A) {
        constexpr int al = decltype(A)::value;
        pcg[al] = R(ldp(a.pin + (QF + al) * a.plane, node));
    });
    R sp_f = R(0), sp_g = R(0);
    static_for<0, QF>([&](auto A) { sp_f += pcf[decltype(A)::value]; });
    static_for<0, QG>([&](auto A) { sp_g += pcg[decltype(A)::value]; });
    f[0] = pcf[0];
    g[0] = pcg[0];
    static_for<1, QF>([&](auto A) {
        constexpr int al = decltype(A)::value;
        f[al] = interp_pop<FL::c[al][0], FL::c[al][1], FL::c[al][2], EXACT>(
            a.pin + al * a.plane, node, pcf[al], st, stride);
    });
    static_for<1, QG>([&](auto A) {
        constexpr int al = decltype(A)::value;
        g[al] = interp_pop<FL::c[al][0], FL::c[al][1], FL::c[al][2], EXACT>(
            a.pin + (QF + al) * a.plane, node, pcg[al], st, stride);
    });
    node_tail<DIM, S, R, SHIFT, MACRO, NT>(a, node, fnode, idx, pcf, pcg, f, g, sp_f, sp_g,
                                           load_heat_shift<DIM, SHIFT>(a, idx));
}

// MAXT: the largest workgroup the kernel is launched with.  The register
// budget per lane follows from it (1024 threads: 128 VGPRs; 256 threads:
// all 512), and the FP64 3D node update needs more than 128.
template <int DIM, typename S, typename R, bool SHIFT, bool MACRO, int WPE = 0, bool NT = false,
          int MAXT = 1024>
__global__ __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void k_step(StepArgs<R, S> a) {
    int tx, ty, tz;
    tile_coords((int)blockIdx.x, a.ntiles, a.gx, a.gy, a.remap, tx, ty, tz);
    const int i = tx * blockDim.x + threadIdx.x;
    const int j = ty * blockDim.y + threadIdx.y;
    const int k = DIM == 3 ? tz : 0;
    if (i >= a.n || j >= a.n) return;
    step_node<DIM, S, R, SHIFT, MACRO, NT>(a, i, j, k);
}

// Initial collision (iteration 1's collision_U/collision_T): equilibrium
// populations from the FP64 macroscopic fields (2dim_mono.f90:227-240),
// collided in FP64 with F = g beta T rho, stored as R (shifted if SHIFT).
template <int DIM, typename S, typename R, bool SHIFT>
__global__ __launch_bounds__(256) void k_prologue(StepArgs<R, S> a, StepArgs<double> ad) {
    using FL = Flow<DIM>;
    using HT = Heat<DIM>;
    constexpr int QF = FL::Q, QG = HT::Q;
    const long nn = DIM == 2 ? (long)a.n * a.n : (long)a.n * a.n * a.n;
    const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn) return;
    // population index of the field node (rows padded to the pitch)
    const long row = node / a.n;
    const long pnode = row * a.pitch + (node - row * a.n);
    const double rho = a.rho[node], u = a.u[node], v = a.v[node];
    const double w = DIM == 3 ? a.w[node] : 0.0, T = a.T[node];
    double f[QF], g[QG], fp[QF], gp[QG];
    const double u2 = u * u + v * v + w * w;
    static_for<0, QF>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr double wa = FL::w[al];
        const double un = u * double(FL::c[al][0]) + v * double(FL::c[al][1]) +
                          w * double(FL::c[al][2]);
        f[al] = rho * wa * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * u2);
    });
    static_for<0, QG>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr double wa = HT::w[al];
        const double un = u * double(FL::c[al][0]) + v * double(FL::c[al][1]) +
                          w * double(FL::c[al][2]);
        g[al] = wa * T * (1.0 + HT::cs2inv * un);
    });
    const double Fb = ad.gbeta * T * rho;
    if constexpr (std::is_same<S, double>::value) a.force[node] = Fb;
    collide_flow<DIM, double, true>(f, rho, rho, u, v, w, Fb, ad, fp);
    collide_heat<DIM, double, true>(g, T, T, u, v, w, ad, gp);
    static_for<0, QF>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr double wa = FL::w[al];
        a.pout[al * a.plane + pnode] = S(SHIFT ? fp[al] - wa : fp[al]);
    });
    int hidx = DIM == 2 ? (int)(node % a.n) : (int)((node / a.n) % a.n);
    const double tloc = SHIFT ? double(a.tref[hidx]) : 0.0;
    static_for<0, QG>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr double wa = HT::w[al];
        a.pout[(QF + al) * a.plane + pnode] = S(SHIFT ? gp[al] - wa * tloc : gp[al]);
    });
}

}  // namespace lbm
