// lbm_step3z.h -- z-marching form of the fused 3D step (D3Q19 + D3Q7, FP32
// shifted populations).  Same arithmetic per node as k_step (lbm_step.h);
// only the way the stencil values reach the registers differs.
//
// A workgroup owns a TX x TY tile of (x, y) columns and marches a chunk of z
// planes.  Per output plane k it stages, once each, the plane tiles that an
// in-plane stencil needs into LDS (with a one-node halo on the sides that
// stencil reaches), so each population value crosses the memory system about
// once instead of up to nine times:
//   * in-plane populations at plane k: x-axis (1, 2, g1, g2; x halo),
//     y-axis (3, 4, g3, g4; y halo), xoy diagonals (7-10; both halos);
//   * the populations that also move in z (xoz 11-14, yoz 15-18) at the plane
//     entering the z window, where their in-plane stage (along x or y) is
//     done; the z stage then combines three planes from a register window.
//     z-axis populations (5, 6, g5, g6) go to the window unstaged.
// The quadratic interpolation is a tensor product, so doing the in-plane
// stage before the z stage gives the reference's value (3dim_mono.f90:767-824
// interpolates z first); the FP64 path keeps the reference's order.
//
// The z window holds planes ws..ws+2 with ws = clamp(k - 1, 0, n - 3): the
// interior stencil (k-1, k, k+1) and the one-sided wall stencils (1, 0, 2)
// and (n-2, n-1, n-3) all fall inside it.  In x and y the wall stencils reach
// two nodes inwards, which stays inside the tile plus halo when TX, TY >= 3
// and n mod TX, n mod TY != 1 (checked by the launcher).
#pragma once

#include "lbm_step.h"

namespace lbm {

template <int TX, int TY> struct Tile {
    float v[TY + 2][TX + 2];  // interior at [1..TY][1..TX]
};

// Cooperative load of one population plane tile with an optional halo.
template <int TX, int TY, int HX, int HY>
__device__ __forceinline__ void stage_tile(Tile<TX, TY> &t, const float *__restrict__ P,
                                           int x0, int y0, int n, int tid) {
    constexpr int W = TX + 2 * HX, H = TY + 2 * HY, C = W * H;
#pragma unroll
    for (int c0 = 0; c0 < C; c0 += TX * TY) {
        const int c = c0 + tid;
        if (c < C) {
            const int cx = c % W, cy = c / W;
            const int gx = x0 - HX + cx, gy = y0 - HY + cy;
            if (gx >= 0 && gx < n && gy >= 0 && gy < n)
                t.v[cy + 1 - HY][cx + 1 - HX] = P[gx + (long)n * gy];
        }
    }
}

__device__ __forceinline__ float lerp3f(const float *w, float f0, float f1, float f2) {
    return f1 + w[0] * (f0 - f1) + w[2] * (f2 - f1);
}

// Populations of the z window: xoz 11-14, yoz 15-18, z-axis 5, 6, g5, g6
// (thermal populations are numbered after the 19 flow planes).
constexpr int zpop(int z) {
    constexpr int t[12] = {11, 12, 13, 14, 15, 16, 17, 18, 5, 6, 19 + 5, 19 + 6};
    return t[z];
}

template <int TX, int TY, bool MACRO>
__global__ __launch_bounds__(TX *TY) void k_step3z(StepArgs<float> a, int kc) {
    using FL = Flow<3>;
    using HT = Heat<3>;
    constexpr int QF = 19;
    const int n = a.n;
    const long nxy = (long)n * n;
    int tx_, ty_, kcid;
    tile_coords((int)blockIdx.x, a.ntiles, a.gx, a.gy, a.remap, tx_, ty_, kcid);
    const int x0 = tx_ * TX, y0 = ty_ * TY;
    const int lx = threadIdx.x, ly = threadIdx.y;
    const int tid = lx + TX * ly;
    const int i = x0 + lx, j = y0 + ly;
    const bool active = i < n && j < n;
    const int k0 = kcid * kc;
    const int k1 = min(k0 + kc, n);

    // 12 in-plane tiles for plane k, 8 tiles for the plane entering the window
    __shared__ Tile<TX, TY> sx[4], sy[4], sxy[4], sz[8];

    AxisStencil<float> stx = load_stencil(a, active ? i : 0);
    AxisStencil<float> sty = load_stencil(a, active ? j : 0);
    if (!active) {  // lanes past the domain read only their own (unused) cell
        stx.o0 = stx.o2 = 0;
        sty.o0 = sty.o2 = 0;
    }
    const float tloc = a.tref[active ? j : 0];

    // z window: stage-1 values W (xoz: x-interpolated, yoz: y-interpolated,
    // z-axis: raw) and the raw centre values C of xoz/yoz
    float W[12][3], Cw[8][3];
    const float *P = a.pin;

    // Stage plane kk of the window populations into slot s (0..2).
    auto load_window_plane = [&](int kk, int s) {
        const long po = kk * nxy;
        static_for<0, 4>([&](auto Z) {
            constexpr int z = decltype(Z)::value;
            stage_tile<TX, TY, 1, 0>(sz[z], P + zpop(z) * a.plane + po, x0, y0, n, tid);
            stage_tile<TX, TY, 0, 1>(sz[4 + z], P + zpop(4 + z) * a.plane + po, x0, y0, n, tid);
        });
        float direct[4];
        static_for<0, 4>([&](auto Z) {
            constexpr int z = decltype(Z)::value;
            direct[z] = active ? P[zpop(8 + z) * a.plane + po + i + (long)n * j] : 0.f;
        });
        __syncthreads();
        static_for<0, 4>([&](auto Z) {
            constexpr int z = decltype(Z)::value;
            constexpr int al = zpop(z);           // xoz: x stage
            const float *wxs = FL::c[al][0] > 0 ? stx.wp : stx.wm;
            const auto &t = sz[z].v[ly + 1];
            Cw[z][s] = t[lx + 1];
            W[z][s] = lerp3f(wxs, t[lx + 1 + stx.o0], t[lx + 1], t[lx + 1 + stx.o2]);
            constexpr int be = zpop(4 + z);       // yoz: y stage
            const float *wys = FL::c[be][1] > 0 ? sty.wp : sty.wm;
            const auto &u = sz[4 + z].v;
            Cw[4 + z][s] = u[ly + 1][lx + 1];
            W[4 + z][s] = lerp3f(wys, u[ly + 1 + sty.o0][lx + 1], u[ly + 1][lx + 1],
                                 u[ly + 1 + sty.o2][lx + 1]);
            W[8 + z][s] = direct[z];
        });
        __syncthreads();
    };

    int ws = max(0, min(k0 - 1, n - 3));
    load_window_plane(ws, 0);
    load_window_plane(ws + 1, 1);
    load_window_plane(ws + 2, 2);

    for (int k = k0; k < k1; k++) {
        const int wk = max(0, min(k - 1, n - 3));
        const bool shift = wk != ws;
        const long po = k * nxy;
        // in-plane tiles of plane k
        stage_tile<TX, TY, 1, 0>(sx[0], P + 1 * a.plane + po, x0, y0, n, tid);
        stage_tile<TX, TY, 1, 0>(sx[1], P + 2 * a.plane + po, x0, y0, n, tid);
        stage_tile<TX, TY, 1, 0>(sx[2], P + (QF + 1) * a.plane + po, x0, y0, n, tid);
        stage_tile<TX, TY, 1, 0>(sx[3], P + (QF + 2) * a.plane + po, x0, y0, n, tid);
        stage_tile<TX, TY, 0, 1>(sy[0], P + 3 * a.plane + po, x0, y0, n, tid);
        stage_tile<TX, TY, 0, 1>(sy[1], P + 4 * a.plane + po, x0, y0, n, tid);
        stage_tile<TX, TY, 0, 1>(sy[2], P + (QF + 3) * a.plane + po, x0, y0, n, tid);
        stage_tile<TX, TY, 0, 1>(sy[3], P + (QF + 4) * a.plane + po, x0, y0, n, tid);
        static_for<0, 4>([&](auto Z) {
            constexpr int z = decltype(Z)::value;
            stage_tile<TX, TY, 1, 1>(sxy[z], P + (7 + z) * a.plane + po, x0, y0, n, tid);
        });
        const long node = po + i + (long)n * j;
        const float r0 = active ? P[node] : 0.f;
        const float rg0 = active ? P[QF * a.plane + node] : 0.f;
        // the window advances by one plane: stage plane wk + 2
        float direct[4];
        if (shift) {
            const long pn = (long)(wk + 2) * nxy;
            static_for<0, 4>([&](auto Z) {
                constexpr int z = decltype(Z)::value;
                stage_tile<TX, TY, 1, 0>(sz[z], P + zpop(z) * a.plane + pn, x0, y0, n, tid);
                stage_tile<TX, TY, 0, 1>(sz[4 + z], P + zpop(4 + z) * a.plane + pn, x0, y0, n,
                                         tid);
                direct[z] = active ? P[zpop(8 + z) * a.plane + pn + i + (long)n * j] : 0.f;
            });
        }
        __syncthreads();
        if (shift) {
            static_for<0, 12>([&](auto Z) {
                constexpr int z = decltype(Z)::value;
                W[z][0] = W[z][1];
                W[z][1] = W[z][2];
                if constexpr (z < 8) {
                    Cw[z][0] = Cw[z][1];
                    Cw[z][1] = Cw[z][2];
                }
            });
            static_for<0, 4>([&](auto Z) {
                constexpr int z = decltype(Z)::value;
                constexpr int al = zpop(z);
                const float *wxs = FL::c[al][0] > 0 ? stx.wp : stx.wm;
                const auto &t = sz[z].v[ly + 1];
                Cw[z][2] = t[lx + 1];
                W[z][2] = lerp3f(wxs, t[lx + 1 + stx.o0], t[lx + 1], t[lx + 1 + stx.o2]);
                constexpr int be = zpop(4 + z);
                const float *wys = FL::c[be][1] > 0 ? sty.wp : sty.wm;
                const auto &u = sz[4 + z].v;
                Cw[4 + z][2] = u[ly + 1][lx + 1];
                W[4 + z][2] = lerp3f(wys, u[ly + 1 + sty.o0][lx + 1], u[ly + 1][lx + 1],
                                     u[ly + 1 + sty.o2][lx + 1]);
                W[8 + z][2] = direct[z];
            });
            ws = wk;
        }
        if (active) {
            const AxisStencil<float> stz = load_stencil(a, k);
            // window slots of the z stencil: interior (0,1,2), bottom wall
            // (1,0,2), top wall (1,2,0)
            const int zcase = k == 0 ? 1 : (k == n - 1 ? 2 : 0);
            float pcf[QF], pcg[7], f[QF], g[7];
            auto zsel = [&](const float *v3, int which) {
                // which: 0 = stencil point 0, 1 = node, 2 = stencil point 2
                const int slot = zcase == 0 ? which : (zcase == 1 ? (which == 0 ? 1 : (which == 1 ? 0 : 2))
                                                                  : (which == 0 ? 1 : (which == 1 ? 2 : 0)));
                return slot == 0 ? v3[0] : (slot == 1 ? v3[1] : v3[2]);
            };
            // raw centre values of the node
            pcf[0] = r0;
            pcg[0] = rg0;
            const int cx = lx + 1, cy = ly + 1;
            pcf[1] = sx[0].v[cy][cx];
            pcf[2] = sx[1].v[cy][cx];
            pcg[1] = sx[2].v[cy][cx];
            pcg[2] = sx[3].v[cy][cx];
            pcf[3] = sy[0].v[cy][cx];
            pcf[4] = sy[1].v[cy][cx];
            pcg[3] = sy[2].v[cy][cx];
            pcg[4] = sy[3].v[cy][cx];
            static_for<0, 4>([&](auto Z) {
                constexpr int z = decltype(Z)::value;
                pcf[7 + z] = sxy[z].v[cy][cx];
                pcf[zpop(z)] = zsel(Cw[z], 1);
                pcf[zpop(4 + z)] = zsel(Cw[4 + z], 1);
            });
            pcf[5] = zsel(W[8], 1);
            pcf[6] = zsel(W[9], 1);
            pcg[5] = zsel(W[10], 1);
            pcg[6] = zsel(W[11], 1);

            // 1. streaming
            f[0] = pcf[0];
            g[0] = pcg[0];
            auto xline = [&](const Tile<TX, TY> &t, const float *w) {
                return lerp3f(w, t.v[cy][cx + stx.o0], t.v[cy][cx], t.v[cy][cx + stx.o2]);
            };
            auto yline = [&](const Tile<TX, TY> &t, const float *w) {
                return lerp3f(w, t.v[cy + sty.o0][cx], t.v[cy][cx], t.v[cy + sty.o2][cx]);
            };
            f[1] = xline(sx[0], stx.wp);
            f[2] = xline(sx[1], stx.wm);
            g[1] = xline(sx[2], stx.wp);
            g[2] = xline(sx[3], stx.wm);
            f[3] = yline(sy[0], sty.wp);
            f[4] = yline(sy[1], sty.wm);
            g[3] = yline(sy[2], sty.wp);
            g[4] = yline(sy[3], sty.wm);
            static_for<0, 4>([&](auto Z) {
                constexpr int z = decltype(Z)::value;
                constexpr int al = 7 + z;
                const float *wxs = FL::c[al][0] > 0 ? stx.wp : stx.wm;
                const float *wys = FL::c[al][1] > 0 ? sty.wp : sty.wm;
                const auto &t = sxy[z].v;
                float c0 = lerp3f(wys, t[cy + sty.o0][cx + stx.o0], t[cy][cx + stx.o0],
                                  t[cy + sty.o2][cx + stx.o0]);
                float c1 = lerp3f(wys, t[cy + sty.o0][cx], t[cy][cx], t[cy + sty.o2][cx]);
                float c2 = lerp3f(wys, t[cy + sty.o0][cx + stx.o2], t[cy][cx + stx.o2],
                                  t[cy + sty.o2][cx + stx.o2]);
                f[al] = lerp3f(wxs, c0, c1, c2);
            });
            static_for<0, 12>([&](auto Z) {
                constexpr int z = decltype(Z)::value;
                constexpr int al = zpop(z);
                constexpr int cz = al < QF ? FL::c[al][2] : FL::c[al - QF][2];
                const float *wzs = cz > 0 ? stz.wp : stz.wm;
                const float v = lerp3f(wzs, zsel(W[z], 0), zsel(W[z], 1), zsel(W[z], 2));
                if constexpr (al < QF) f[al] = v;
                else g[al - QF] = v;
            });
            // streaming of the conduction profile along y (heated axis)
            constexpr double wc3 = HT::w[3] * FL::c[3][1], wc4 = HT::w[4] * FL::c[4][1];
            g[3] += float(wc3) * a.tcorr;
            g[4] += float(wc4) * a.tcorr;

            // 2. walls
            const int idx[3] = {i, j, k};
            static_for<0, 3>([&](auto AX) {
                constexpr int ax = decltype(AX)::value;
                const int side = idx[ax] == 0 ? 1 : (idx[ax] == n - 1 ? -1 : 0);
                if (side != 0) {
                    static_for<1, QF>([&](auto A) {
                        constexpr int al = decltype(A)::value;
                        constexpr int ca = FL::c[al][ax];
                        if constexpr (ca != 0) {
                            if (ca * side > 0) f[al] = pcf[opposite<3>(al)];
                        }
                    });
                    static_for<1, 7>([&](auto A) {
                        constexpr int al = decltype(A)::value;
                        constexpr int ca = FL::c[al][ax];
                        if constexpr (ca != 0) {
                            if (ca * side > 0) {
                                const float gpo = pcg[opposite<3>(al)];
                                if constexpr (ax == heated_axis<3>()) {
                                    constexpr double w2 = 2.0 * HT::w[al];
                                    const float tw = side > 0 ? a.tw_hot : a.tw_cold;
                                    g[al] = -gpo + float(w2) * tw;
                                } else {
                                    g[al] = gpo;
                                }
                            }
                        }
                    });
                }
            });

            // 3. macroscopic fields and the force of the previous step
            float sp_f = 0.f, sp_g = 0.f;
            static_for<0, QF>([&](auto A) { sp_f += pcf[decltype(A)::value]; });
            static_for<0, 7>([&](auto A) { sp_g += pcg[decltype(A)::value]; });
            const float F_prev = a.gbeta * (tloc + sp_g) * (1.f + sp_f);
            float srho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f, sT = 0.f;
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
            static_for<0, 7>([&](auto A) { sT += g[decltype(A)::value]; });
            const float rho = 1.f + srho;
            const float T = tloc + sT;
            const float ux = jx / rho, uy = jy / rho;
            const float uz = (jz + a.half_dt * F_prev) / rho;
            if constexpr (MACRO) {
                double dr = 0.0, dT = 0.0;
                static_for<0, QF>([&](auto A) { dr += double(f[decltype(A)::value]); });
                static_for<0, 7>([&](auto A) { dT += double(g[decltype(A)::value]); });
                const long mn = i + (long)n * (j + (long)n * k);
                a.rho[mn] = 1.0 + dr;
                a.T[mn] = double(tloc) + dT;
                a.u[mn] = double(ux);
                a.v[mn] = double(uy);
                a.w[mn] = double(uz);
            }

            // 4. collision and store
            const float Fb = a.gbeta * T * rho;
            float fp[QF], gp[7];
            collide_flow<3, float, true>(f, srho, rho, ux, uy, uz, Fb, a, fp);
            collide_heat<3, float>(g, sT, T, ux, uy, uz, a, gp);
            static_for<0, QF>([&](auto A) {
                constexpr int al = decltype(A)::value;
                a.pout[al * a.plane + node] = fp[al];
            });
            static_for<0, 7>([&](auto A) {
                constexpr int al = decltype(A)::value;
                a.pout[(QF + al) * a.plane + node] = gp[al];
            });
        }
        __syncthreads();
    }
}

}  // namespace lbm
