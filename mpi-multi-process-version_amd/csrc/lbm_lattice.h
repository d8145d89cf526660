// lbm_lattice.h -- compile-time descriptors of the four lattices of the
// coupled solver (D2Q9 + D2Q5, D3Q19 + D3Q7), usable on host and device.
//
// Direction numbering is the reference's (natural-convection_2dim_mono.f90:58-60,
// natural-convection_3dim_mono.f90:60-63); the thermal lattices are the first
// five / seven directions.  The MRT moment bases are built from their
// polynomial definitions (Lallemand & Luo 2000 for D2Q9, d'Humieres et al.
// 2002 for D3Q19); both are orthogonal, so M^-1 = M^T diag(1/|row|^2).
#pragma once

#include <hip/hip_runtime.h>

namespace lbm {

template <int DIM> struct Flow;
template <int DIM> struct Heat;

template <> struct Flow<2> {
    static constexpr int Q = 9;
    static constexpr int c[Q][3] = {{0, 0, 0},  {1, 0, 0},   {0, 1, 0},
                                    {-1, 0, 0}, {0, -1, 0},  {1, 1, 0},
                                    {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
    static constexpr double w[Q] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                    1.0 / 9.0,  1.0 / 9.0,  1.0 / 36.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    // rho, e, eps, jx, qx, jy, qy, pxx, pxy
    static constexpr double row(int r, int a) {
        double cx = c[a][0], cy = c[a][1], q = cx * cx + cy * cy;
        switch (r) {
        case 0: return 1;
        case 1: return -4 + 3 * q;
        case 2: return 4 - 10.5 * q + 4.5 * q * q;
        case 3: return cx;
        case 4: return (-5 + 3 * q) * cx;
        case 5: return cy;
        case 6: return (-5 + 3 * q) * cy;
        case 7: return cx * cx - cy * cy;
        default: return cx * cy;
        }
    }
};

template <> struct Flow<3> {
    static constexpr int Q = 19;
    static constexpr int c[Q][3] = {
        {0, 0, 0},   {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
        {0, 0, 1},   {0, 0, -1}, {1, 1, 0},  {-1, 1, 0},  {1, -1, 0},
        {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1},  {-1, 0, -1},
        {0, 1, 1},   {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
    static constexpr double w[Q] = {
        1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
        1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
        1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
    // rho, e, eps, jx, qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz,
    // pxz, mx, my, mz
    static constexpr double row(int r, int a) {
        double cx = c[a][0], cy = c[a][1], cz = c[a][2];
        double q = cx * cx + cy * cy + cz * cz;
        switch (r) {
        case 0: return 1;
        case 1: return 19 * q - 30;
        case 2: return (21 * q * q - 53 * q + 24) / 2;
        case 3: return cx;
        case 4: return (5 * q - 9) * cx;
        case 5: return cy;
        case 6: return (5 * q - 9) * cy;
        case 7: return cz;
        case 8: return (5 * q - 9) * cz;
        case 9: return 3 * cx * cx - q;
        case 10: return (3 * q - 5) * (3 * cx * cx - q);
        case 11: return cy * cy - cz * cz;
        case 12: return (3 * q - 5) * (cy * cy - cz * cz);
        case 13: return cx * cy;
        case 14: return cy * cz;
        case 15: return cx * cz;
        case 16: return (cy * cy - cz * cz) * cx;
        case 17: return (cz * cz - cx * cx) * cy;
        default: return (cx * cx - cy * cy) * cz;
        }
    }
};

// Thermal lattices: moments 1, c_x, c_y, |c|^2, c_x^2 - c_y^2 (D2Q5) and
// 1, c_x, c_y, c_z, |c|^2, c_x^2 - c_y^2, c_x^2 - c_z^2 (D3Q7); the inverse
// is stored as integer numerators over a common denominator.
template <> struct Heat<2> {
    static constexpr int Q = 5;
    static constexpr double w[Q] = {0.5, 0.125, 0.125, 0.125, 0.125};
    static constexpr double cs2inv = 4.0;  // g_eq = w T (1 + 4 c.u)
    static constexpr int den = 4;
    static constexpr int ninv[Q][Q] = {{4, 0, 0, -4, 0},
                                       {0, 2, 0, 1, 1},
                                       {0, 0, 2, 1, -1},
                                       {0, -2, 0, 1, 1},
                                       {0, 0, -2, 1, -1}};
    static constexpr double row(int r, int a) {
        double cx = Flow<2>::c[a][0], cy = Flow<2>::c[a][1];
        switch (r) {
        case 0: return 1;
        case 1: return cx;
        case 2: return cy;
        case 3: return cx * cx + cy * cy;
        default: return cx * cx - cy * cy;
        }
    }
};

template <> struct Heat<3> {
    static constexpr int Q = 7;
    static constexpr double w[Q] = {0.5,        1.0 / 12.0, 1.0 / 12.0,
                                    1.0 / 12.0, 1.0 / 12.0, 1.0 / 12.0,
                                    1.0 / 12.0};
    static constexpr double cs2inv = 6.0;  // g_eq = w T (1 + 6 c.u)
    static constexpr int den = 6;
    static constexpr int ninv[Q][Q] = {
        {6, 0, 0, 0, -6, 0, 0}, {0, 3, 0, 0, 1, 1, 1},  {0, -3, 0, 0, 1, 1, 1},
        {0, 0, 3, 0, 1, -2, 1}, {0, 0, -3, 0, 1, -2, 1}, {0, 0, 0, 3, 1, 1, -2},
        {0, 0, 0, -3, 1, 1, -2}};
    static constexpr double row(int r, int a) {
        double cx = Flow<3>::c[a][0], cy = Flow<3>::c[a][1], cz = Flow<3>::c[a][2];
        switch (r) {
        case 0: return 1;
        case 1: return cx;
        case 2: return cy;
        case 3: return cz;
        case 4: return cx * cx + cy * cy + cz * cz;
        case 5: return cx * cx - cy * cy;
        default: return cx * cx - cz * cz;
        }
    }
};

template <int DIM> constexpr double flow_norm(int r) {
    double s = 0;
    for (int a = 0; a < Flow<DIM>::Q; a++) s += Flow<DIM>::row(r, a) * Flow<DIM>::row(r, a);
    return s;
}

template <int DIM> constexpr int opposite(int a) {
    for (int b = 0; b < Flow<DIM>::Q; b++)
        if (Flow<DIM>::c[b][0] == -Flow<DIM>::c[a][0] &&
            Flow<DIM>::c[b][1] == -Flow<DIM>::c[a][1] &&
            Flow<DIM>::c[b][2] == -Flow<DIM>::c[a][2])
            return b;
    return -1;
}

// Axis with the heated walls: x in 2D, y in 3D (reference orientation).
template <int DIM> constexpr int heated_axis() { return DIM == 2 ? 0 : 1; }

}  // namespace lbm
