// lbm_aux.hip -- FP64 helper kernels: the reference initial condition, the
// windowed residual (check) and the Nusselt reduction.  Reductions are
// deterministic: a fixed grid accumulates per-thread partial sums, reduces
// them per block with an LDS tree, and one block adds the block results in
// a fixed order.  Compiled with -ffp-contract=off.
#include <mutex>

#include "lbm_internal.h"

namespace lbm {

// initial (2dim_mono.f90:199-212, 3dim_mono.f90:256-271): rho = 1, u = 0,
// T linear across the heated direction (x in 2D, y in 3D).
__global__ void k_init_rest(double *rho, double *u, double *v, double *w, double *T,
                            int n, int dim, long nn, double Thot, double Tcold) {
    long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn) return;
    int ih = dim == 2 ? (int)(node % n) : (int)((node / n) % n);
    rho[node] = 1.0;
    u[node] = 0.0;
    v[node] = 0.0;
    if (dim == 3) w[node] = 0.0;
    T[node] = (double)ih / (double)(n - 1) * (Tcold - Thot) + Thot;
}

template <int NQ>
__device__ void block_reduce_store(double (&acc)[NQ], double *out) {
    // butterfly over the 64 lanes of each wave (a fixed pairing, so every
    // lane ends with the same, reproducible sum), then the wave sums in
    // wave order
    constexpr int NW = kRedThreads / 64;
    __shared__ double sh[NQ][NW];
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) sh[q][wv] = acc[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            double s = sh[q][0];
            for (int w = 1; w < NW; w++) s += sh[q][w];
            out[q] = s;
        }
    }
}

// check (2dim_mono.f90:771-798, 3dim_mono.f90:1043-1070): partial sums of
// |u - u_prev|^2, |u|^2, (T - T_prev)^2, T^2; the snapshot is updated.
template <int DIM>
__global__ __launch_bounds__(kRedThreads) void k_resid_part(
    const double *__restrict__ u, const double *__restrict__ v,
    const double *__restrict__ w, const double *__restrict__ T,
    double *__restrict__ up, double *__restrict__ vp, double *__restrict__ wp,
    double *__restrict__ Tp, long nn, double *__restrict__ part) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
         i += (long)gridDim.x * blockDim.x) {
        double uu = u[i], vv = v[i];
        double du = uu - up[i], dv = vv - vp[i];
        acc[0] += du * du + dv * dv;
        acc[1] += uu * uu + vv * vv;
        if constexpr (DIM == 3) {
            double ww = w[i], dw = ww - wp[i];
            acc[0] += dw * dw;
            acc[1] += ww * ww;
            wp[i] = ww;
        }
        double tt = T[i], dT = tt - Tp[i];
        acc[2] += dT * dT;
        acc[3] += tt * tt;
        up[i] = uu;
        vp[i] = vv;
        Tp[i] = tt;
    }
    block_reduce_store<4>(acc, part + 4 * blockIdx.x);
}

// Mean Nusselt numbers of the hot and cold walls: one-sided quadratic
// gradient through (0, T_wall), (d1, T1), (d2, T2), control-volume weights.
__device__ inline double wall_gradient(double Tw, double d1, double T1, double d2,
                                       double T2) {
    return Tw * (-(d1 + d2) / (d1 * d2)) + T1 * d2 / (d1 * (d2 - d1)) -
           T2 * d1 / (d2 * (d2 - d1));
}

__global__ __launch_bounds__(kRedThreads) void k_nu_part(
    const double *__restrict__ T, const double *__restrict__ x,
    const double *__restrict__ cvw, int n, int dim, double H, double Thot,
    double Tcold, double *__restrict__ part) {
    double acc[2] = {0.0, 0.0};
    const long nw = dim == 2 ? (long)n : (long)n * n;
    const double dT = Thot - Tcold;
    const double d1h = x[0], d2h = x[1], d1c = H - x[n - 1], d2c = H - x[n - 2];
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nw;
         t += (long)gridDim.x * blockDim.x) {
        long h1, h2, c1, c2;
        double wgt;
        if (dim == 2) {  // heated x-walls; wall coordinate j = t
            h1 = (long)n * t;
            h2 = 1 + (long)n * t;
            c1 = (n - 1) + (long)n * t;
            c2 = (n - 2) + (long)n * t;
            wgt = cvw[t];
        } else {  // heated y-walls; wall coordinates (i, k)
            long i = t % n, k = t / n;
            h1 = i + (long)n * ((long)n * k);
            h2 = i + (long)n * (1 + (long)n * k);
            c1 = i + (long)n * ((n - 1) + (long)n * k);
            c2 = i + (long)n * ((n - 2) + (long)n * k);
            wgt = cvw[i] * cvw[k];
        }
        double gh = wall_gradient(Thot, d1h, T[h1], d2h, T[h2]);
        double gc = wall_gradient(Tcold, d1c, T[c1], d2c, T[c2]);
        acc[0] += -H * gh / dT * wgt;
        acc[1] += H * gc / dT * wgt;
    }
    block_reduce_store<2>(acc, part + 2 * blockIdx.x);
}

// Adds nparts partial results of NQ values each, in a fixed order.
template <int NQ>
__global__ __launch_bounds__(kRedThreads) void k_reduce_final(const double *__restrict__ part,
                                                              int nparts, double *__restrict__ out) {
    double acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
#pragma unroll
        for (int q = 0; q < NQ; q++) acc[q] += part[NQ * b + q];
    }
    block_reduce_store<NQ>(acc, out);
}

// One solver state in another's storage (lbm_copy_state): populations as
// stored by src (SS; shifted: f - w_a and g - w_a Tref, Tref the FP32
// conduction profile both handles hold) converted to dst's storage (DS),
// and, for an FP64 destination, the lagged force F = g beta T rho of the
// FP64 path from the macroscopic fields of the same step (which the FP32
// path recovers from the node's own populations).
// The two storages may pad their population rows differently (spitch,
// dpitch; plane strides splane, dplane); the fields are unpadded.
template <int DIM, typename SS, bool SSHIFT, typename DS, bool DSHIFT>
__global__ __launch_bounds__(256) void k_convert_state(
    const SS *__restrict__ src, long splane, int spitch, DS *__restrict__ dst, long dplane,
    int dpitch, long nn, int n, const float *__restrict__ tref, const double *__restrict__ rho,
    const double *__restrict__ T, double *__restrict__ force, double gbeta) {
    using FL = Flow<DIM>;
    using HT = Heat<DIM>;
    const long node = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn) return;
    const int ih = DIM == 2 ? (int)(node % n) : (int)((node / n) % n);
    const long row = node / n, col = node - row * n;
    const long sn = row * spitch + col, dn = row * dpitch + col;
    const double tl = (double)tref[ih];
    static_for<0, FL::Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr double wa = FL::w[al];
        const double f = (double)src[al * splane + sn] + (SSHIFT ? wa : 0.0);
        dst[al * dplane + dn] = (DS)(DSHIFT ? f - wa : f);
    });
    static_for<0, HT::Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr double wa = HT::w[al];
        const double ws = wa * tl;
        const double g = (double)src[(FL::Q + al) * splane + sn] + (SSHIFT ? ws : 0.0);
        dst[(FL::Q + al) * dplane + dn] = (DS)(DSHIFT ? g - ws : g);
    });
    if (force) force[node] = gbeta * T[node] * rho[node];
}

template <int DIM>
static hipError_t convert_state(const Ctx &dst, const Ctx &src) {
    dim3 block(256), grid((unsigned)((src.nn + 255) / 256));
    const void *s = src.pbuf[src.cur];
    void *d = dst.pbuf[dst.cur];
    // the FP64 path's force, from the macroscopic fields (FP32 to FP64)
    double *force = dst.fp64 ? dst.force : nullptr;
    const double gb = dst.info[LBM_INFO_GBETA];
    if (src.fp64)
        hipLaunchKernelGGL((k_convert_state<DIM, double, false, float, true>), grid, block, 0,
                           dst.stream, (const double *)s, src.plane, src.pitch, (float *)d,
                           dst.plane, dst.pitch, src.nn, src.n, src.d_tref, src.rho, src.T,
                           force, gb);
    else
        hipLaunchKernelGGL((k_convert_state<DIM, float, true, double, false>), grid, block, 0,
                           dst.stream, (const float *)s, src.plane, src.pitch, (double *)d,
                           dst.plane, dst.pitch, src.nn, src.n, src.d_tref, src.rho, src.T,
                           force, gb);
    return hipGetLastError();
}

hipError_t launch_convert_state(const Ctx &dst, const Ctx &src) {
    if (src.fp64 == dst.fp64) {
        // one storage: a plain copy (FP32 (f + w) - w would round deviations
        // far below w), and the FP64 force as it is
        const size_t esz = src.fp64 ? sizeof(double) : sizeof(float);
        hipError_t e = hipMemcpyAsync(dst.pbuf[dst.cur], src.pbuf[src.cur],
                                      esz * (size_t)src.plane * (size_t)(src.qf + src.qg),
                                      hipMemcpyDeviceToDevice, dst.stream);
        if (e == hipSuccess && src.force && dst.force)
            e = hipMemcpyAsync(dst.force, src.force, sizeof(double) * (size_t)src.nn,
                               hipMemcpyDeviceToDevice, dst.stream);
        return e;
    }
    return src.dim == 2 ? convert_state<2>(dst, src) : convert_state<3>(dst, src);
}

hipError_t launch_init_rest(const Ctx &c) {
    dim3 block(256), grid((unsigned)((c.nn + 255) / 256));
    hipLaunchKernelGGL(k_init_rest, grid, block, 0, c.stream, c.rho, c.u, c.v, c.w, c.T,
                       c.n, c.dim, c.nn, c.p.Thot, c.p.Tcold);
    return hipGetLastError();
}

// The residual's two kernels with their partial sums in `part` (4 per
// block) and the four sums in `red`.
static hipError_t residual_into(const Ctx &c, double *part, double *red) {
    dim3 block(kRedThreads), grid(c.red_blocks);
    if (c.dim == 2)
        hipLaunchKernelGGL(k_resid_part<2>, grid, block, 0, c.stream, c.u, c.v, c.w, c.T,
                           c.up, c.vp, c.wp, c.Tp, c.nn, part);
    else
        hipLaunchKernelGGL(k_resid_part<3>, grid, block, 0, c.stream, c.u, c.v, c.w, c.T,
                           c.up, c.vp, c.wp, c.Tp, c.nn, part);
    hipLaunchKernelGGL(k_reduce_final<4>, dim3(1), block, 0, c.stream, part, c.red_blocks, red);
    return hipGetLastError();
}

hipError_t launch_residual(const Ctx &c) { return residual_into(c, c.d_part, c.d_red); }

// Scratch memory of lbm_check_noalloc: the residual's partial sums (4 per
// block, at most kRedMaxBlocks blocks) followed by the 4 sums, one copy per
// GPU for the life of the process (32.8 KB).  Host code uses the device
// address from hipGetSymbolAddress (scratch_address), never the array's
// name, which on the host is only a stand-in.
constexpr int kScratchDoubles = 4 * kRedMaxBlocks + 4;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};
}  // namespace

// The device address of `scratch` on the current device, looked up once
// per device (with that device current) and kept.
static hipError_t scratch_address(double **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

hipError_t launch_residual_scratch(const Ctx &c, double **sums) {
    if (4 * c.red_blocks + 4 > kScratchDoubles) return hipErrorInvalidValue;
    double *base = nullptr;
    hipError_t e = scratch_address(&base);
    if (e != hipSuccess) return e;
    // every partial the final kernel reads is written by the first kernel
    // of the same call, so nothing left by an earlier call needs clearing
    *sums = base + 4 * kRedMaxBlocks;
    return residual_into(c, base, *sums);
}

hipError_t launch_nusselt(const Ctx &c) {
    dim3 block(kRedThreads), grid(c.red_blocks);
    hipLaunchKernelGGL(k_nu_part, grid, block, 0, c.stream, c.T, c.d_x, c.d_cvw, c.n, c.dim,
                       c.info[LBM_INFO_H], c.p.Thot, c.p.Tcold, c.d_part);
    hipLaunchKernelGGL(k_reduce_final<2>, dim3(1), block, 0, c.stream, c.d_part,
                       c.red_blocks, c.d_red + 4);
    return hipGetLastError();
}

}  // namespace lbm
