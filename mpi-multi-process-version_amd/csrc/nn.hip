// nn.hip -- the learned warm start (include/lbm_nn.h): kernels for gfx950
// and the host side of the encoder-decoder.
//
// Every convolution is an im2col product on the matrix cores:
//   forward   Z[cout][N]  = W[cout][K] . col[K][N]            (K = cin k^2)
//   backward  dW[cout][K] = dZ[cout][N] . col[K][N]^T         (split over N)
//             dcol[K][N]  = W[cout][K]^T . dZ[cout][N], then col2im
// with N = batch x output pixels.  The product kernel multiplies BF16
// operands with v_mfma_f32_16x16x16_bf16 and accumulates in FP32; one wave
// computes one or four 16 x 16 tiles of the result stacked along m (sharing
// the right operand's fragment).  Activations are FP32 in
// channel-major layout [C][B][H][W], so a channel concatenation is a
// contiguous stack and a layer's product output is its activation.  All
// reductions run in a fixed order, so every call is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lbm_nn.h"
#include "lbm_internal.h"

namespace nn {

using bf16 = uint16_t;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// round to nearest even; NaN stays NaN
__device__ __forceinline__ bf16 to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (bf16)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16)(u >> 16);
}

// C(m, n) = sum_k A(m, k) B(k, n) with A(m, k) = A[m as_m + k as_k] and
// B(k, n) = B[k bs_k + n bs_n]; C row-major with leading dimension ldc.
// blockIdx.z selects the k range [z kchunk, (z + 1) kchunk); a split
// product writes its partial result to C + z M ldc.
struct GemmArgs {
    const bf16 *A;
    long as_m, as_k;
    const bf16 *B;
    long bs_k, bs_n;
    float *C;
    long ldc;
    int M, N, K, kchunk;
};

// Read in place of an element outside an operand: a masked-out load reads
// zeros from here, so every load is issued and its value used as it comes
// (a value selected after the load lets the compiler turn the load into a
// branch of its own, with a wait before the join).
__device__ __align__(16) uint32_t kZeros[4] = {0u, 0u, 0u, 0u};   // never written

// The 16 x 16 x 16 MFMA's operand fragment of one lane: lane (r, q) holds
// X(r, 4q + e), e = 0..3, of the 16-row slice starting at row `row` and the
// k range [k0, k0 + 16) clipped to [.., ke), with X(i, k) = X[i si + k sk]
// and rows >= rows zero.  VEC: k is the contiguous axis (sk == 1) of an
// 8-byte aligned operand and every k range ends on a multiple of 4, so the
// four values are one 8-byte load.  Every load is issued (from kZeros for a
// masked-out element): no branch separates the loads of a k step, so they
// are in flight together.
template <bool VEC>
__device__ __forceinline__ v4s fragment(const bf16 *X, long si, long sk, long i, bool iok,
                                        long k, long ke) {
    v4s f;
    if constexpr (VEC) {
        const bool ok = iok && k < ke;
        const uint2 w = *(ok ? (const uint2 *)(X + i * si + k) : (const uint2 *)kZeros);
        f[0] = (short)(w.x & 0xffffu);
        f[1] = (short)(w.x >> 16);
        f[2] = (short)(w.y & 0xffffu);
        f[3] = (short)(w.y >> 16);
    } else {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const bool ok = iok && k + e < ke;
            f[e] = (short)*(ok ? X + i * si + (k + e) * sk : (const bf16 *)kZeros);
        }
    }
    return f;
}

// One wave per MT*16 x 16 block of C: MT 16 x 16 MFMA tiles down the m axis
// share each B fragment (B, the large operand, is then read once per MT*16
// rows of C).  The host picks MT and the k split per product so that a
// launch has enough waves, and AV / BV (the 8-byte fragment loads of A / B)
// where the layouts allow them.  The fragments of step k0 + 16 are loaded
// while the MFMAs of step k0 run (two register sets in turn); each tile's
// MFMAs run in k order, and tiles past M multiply zero fragments and are
// not stored.
template <int MT, bool AV, bool BV>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int tn = (g.N + 15) >> 4, tg = (g.M + 16 * MT - 1) / (16 * MT);
    if (wave >= tn * tg) return;   // whole waves only
    const int mb = (wave / tn) * 16 * MT, n0 = (wave % tn) << 4;
    const int kb = blockIdx.z * g.kchunk, ke = min(g.K, kb + g.kchunk);
    const int r = lane & 15, q = lane >> 4;
    const long n = n0 + r;
    const bool nok = n < g.N;
    const int nt = min(MT, (g.M - mb + 15) >> 4);   // m tiles in this block (wave-uniform)
    v4f acc[MT];
    // two fragment sets in turn: the loads of one k step are in flight
    // while the MFMAs of the other run
    v4s a0[MT], a1[MT], b0, b1;
    auto load = [&](long k, v4s (&a)[MT], v4s &b) {
        b = fragment<BV>(g.B, g.bs_n, g.bs_k, n, nok, k, ke);
#pragma unroll
        for (int t = 0; t < MT; t++) {
            const long m = mb + 16 * t + r;
            a[t] = fragment<AV>(g.A, g.as_m, g.as_k, m, m < g.M, k, ke);
        }
    };
    auto mfma = [&](const v4s (&a)[MT], const v4s &b) {
#pragma unroll
        for (int t = 0; t < MT; t++)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], b, acc[t], 0, 0, 0);
    };
#pragma unroll
    for (int t = 0; t < MT; t++) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const long kq = 4 * q;
    load(kb + kq, a0, b0);
    for (int k0 = kb; k0 < ke; k0 += 32) {
        load(k0 + 16 + kq, a1, b1);
        mfma(a0, b0);
        if (k0 + 16 >= ke) break;   // wave-uniform
        load(k0 + 32 + kq, a0, b0);
        mfma(a1, b1);
    }
    float *C = g.C + (long)blockIdx.z * g.M * g.ldc;
#pragma unroll
    for (int t = 0; t < MT; t++) {
        if (t >= nt) break;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const long mi = mb + 16 * t + 4 * q + e, nj = n0 + r;
            if (mi < g.M && nj < g.N) C[mi * g.ldc + nj] = acc[t][e];
        }
    }
}

// The forward convolution's product without its im2col matrix: the right
// operand B(k, n) = X[c][b][yo s + ky - p][xo s + kx - p] (k = (c KS + ky) KS
// + kx, n = (b Ho + yo) Wo + xo) is gathered from the activations and
// rounded to BF16 where k_gemm would read it from col: the same operands,
// MFMAs and order, so the same results.  Pipelined like k_gemm.
struct ConvGeom {
    const float *X;
    int B, H, W, stride, pad, Ho, Wo;
};

template <int MT, int KS, bool AV>
__global__ __launch_bounds__(256) void k_gemm_conv(GemmArgs g, ConvGeom cg) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int tn = (g.N + 15) >> 4, tg = (g.M + 16 * MT - 1) / (16 * MT);
    if (wave >= tn * tg) return;   // whole waves only
    const int mb = (wave / tn) * 16 * MT, n0 = (wave % tn) << 4;
    const int kb = blockIdx.z * g.kchunk, ke = min(g.K, kb + g.kchunk);
    const int r = lane & 15, q = lane >> 4;
    const int n = n0 + r;
    const bool nok = n < g.N;
    const int nn = nok ? n : 0;
    const int t2 = nn / cg.Wo, xo = nn - t2 * cg.Wo;
    const int b = t2 / cg.Ho, yo = t2 - b * cg.Ho;
    const int y0 = yo * cg.stride - cg.pad, x0 = xo * cg.stride - cg.pad;
    const long bstride = (long)cg.B * cg.H * cg.W;   // one channel
    const float *Xb = cg.X + (long)b * cg.H * cg.W;
    const int nt = min(MT, (g.M - mb + 15) >> 4);
    // the four activations of B's fragment for the k range starting at k
    // (from kZeros for an element outside the image or past ke); they are
    // rounded to BF16 one step later, so that the loads of the next step
    // stay in flight while this step's MFMAs run
    auto gather = [&](int k, float (&v)[4]) {
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int kk = k + e;
            const int c = kk / (KS * KS), rr = kk - c * (KS * KS);
            const int ky = rr / KS, kx = rr - ky * KS;
            const int y = y0 + ky, x = x0 + kx;
            const bool ok = nok && kk < ke && y >= 0 && y < cg.H && x >= 0 && x < cg.W;
            v[e] = *(ok ? Xb + c * bstride + y * cg.W + x : (const float *)kZeros);
        }
    };
    v4f acc[MT];
    // two sets in turn, as in k_gemm
    v4s a0[MT], a1[MT];
    float v0[4], v1[4];
    auto load = [&](int k, v4s (&a)[MT], float (&v)[4]) {
        gather(k, v);
#pragma unroll
        for (int t = 0; t < MT; t++) {
            const long m = mb + 16 * t + r;
            a[t] = fragment<AV>(g.A, g.as_m, g.as_k, m, m < g.M, k, ke);
        }
    };
    auto mfma = [&](const v4s (&a)[MT], const float (&v)[4]) {
        v4s bf;
#pragma unroll
        for (int e = 0; e < 4; e++) bf[e] = (short)to_bf16(v[e]);
#pragma unroll
        for (int t = 0; t < MT; t++)
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[t], bf, acc[t], 0, 0, 0);
    };
#pragma unroll
    for (int t = 0; t < MT; t++) acc[t] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
    const int kq = 4 * q;
    load(kb + kq, a0, v0);
    for (int k0 = kb; k0 < ke; k0 += 32) {
        load(k0 + 16 + kq, a1, v1);
        mfma(a0, v0);
        if (k0 + 16 >= ke) break;   // wave-uniform
        load(k0 + 32 + kq, a0, v0);
        mfma(a1, v1);
    }
    float *C = g.C + (long)blockIdx.z * g.M * g.ldc;
#pragma unroll
    for (int t = 0; t < MT; t++) {
        if (t >= nt) break;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const long mi = mb + 16 * t + 4 * q + e, nj = n0 + r;
            if (mi < g.M && nj < g.N) C[mi * g.ldc + nj] = acc[t][e];
        }
    }
}

// C[m ldc + n] = sum_z P[z M N + m N + n], z in order
__global__ void k_reduce_splits(const float *P, float *C, long ldc, int M, int N, int splits) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)M * N) return;
    const long m = t / N, n = t % N;
    float s = 0.0f;
    for (int z = 0; z < splits; z++) s += P[(long)z * M * N + t];
    C[m * ldc + n] = s;
}

// col[(c k^2 + ky k + kx) N + n] = X[c][b][yo s + ky - p][xo s + kx - p], n = (b Ho + yo) Wo + xo.
// Block row blockIdx.y is one row kk of col (so c, ky, kx are uniform); the
// threads run over n.  Indices are 32-bit: N and C B H W stay below 2^31.
__global__ __launch_bounds__(256) void k_im2col(const float *X, bf16 *col, int B, int H, int W,
                                                int ks, int stride, int pad, int Ho, int Wo) {
    const int N = B * Ho * Wo;
    const int kk = blockIdx.y, c = kk / (ks * ks), rr = kk - c * ks * ks;
    const int ky = rr / ks, kx = rr - ky * ks;
    const float *Xc = X + (long)c * B * H * W;
    bf16 *row = col + (long)kk * N;
    for (int n = blockIdx.x * 256 + threadIdx.x; n < N; n += gridDim.x * 256) {
        const int t2 = n / Wo, xo = n - t2 * Wo;
        const int b = t2 / Ho, yo = t2 - b * Ho;
        const int y = yo * stride + ky - pad, x = xo * stride + kx - pad;
        float v = 0.0f;
        if (y >= 0 && y < H && x >= 0 && x < W) v = Xc[(b * H + y) * W + x];
        row[n] = to_bf16(v);
    }
}

// k_im2col for Wo % 8 == 0: a thread writes eight consecutive n (the same
// image row, xo = 8m .. 8m+7) as one 16-byte store; the same values.
__global__ __launch_bounds__(256) void k_im2col8(const float *X, bf16 *col, int B, int H, int W,
                                                 int ks, int stride, int pad, int Ho, int Wo) {
    const int N = B * Ho * Wo;
    const int kk = blockIdx.y, c = kk / (ks * ks), rr = kk - c * ks * ks;
    const int ky = rr / ks, kx = rr - ky * ks;
    const float *Xc = X + (long)c * B * H * W;
    bf16 *row = col + (long)kk * N;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < (N >> 3); t += gridDim.x * 256) {
        const int n = t << 3;
        const int t2 = n / Wo, xo = n - t2 * Wo;
        const int b = t2 / Ho, yo = t2 - b * Ho;
        const int y = yo * stride + ky - pad;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        if (y >= 0 && y < H) {
            const float *xr = Xc + (b * H + y) * W;
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int x = (xo + e) * stride + kx - pad;
                const float v = (x >= 0 && x < W) ? xr[x] : 0.0f;
                w[e >> 1] |= (uint32_t)to_bf16(v) << (16 * (e & 1));
            }
        }
        *reinterpret_cast<uint4 *>(row + n) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// dX[c][b][y][x] (+)= sum of the dcol entries that im2col read from (c, b, y, x)
__global__ __launch_bounds__(256) void k_col2im(const float *dcol, float *dX, int C, int B, int H,
                                                int W, int ks, int stride, int pad, int Ho, int Wo,
                                                int accumulate) {
    const int N = B * Ho * Wo;
    const int total = C * B * H * W;
    for (int t = blockIdx.x * 256 + threadIdx.x; t < total; t += gridDim.x * 256) {
        const int t2 = t / W, x = t - t2 * W;
        const int t3 = t2 / H, y = t2 - t3 * H;
        const int c = t3 / B, b = t3 - c * B;
        float s = 0.0f;
        for (int ky = 0; ky < ks; ky++) {
            const int yy = y + pad - ky;
            if (yy < 0 || yy % stride) continue;
            const int yo = yy / stride;
            if (yo >= Ho) continue;
            for (int kx = 0; kx < ks; kx++) {
                const int xx = x + pad - kx;
                if (xx < 0 || xx % stride) continue;
                const int xo = xx / stride;
                if (xo >= Wo) continue;
                s += dcol[(long)((c * ks + ky) * ks + kx) * N + (b * Ho + yo) * Wo + xo];
            }
        }
        dX[t] = accumulate ? dX[t] + s : s;
    }
}

__global__ void k_bias_act(float *Y, const float *bias, int rows, long N, int relu) {
    const long total = (long)rows * N;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        float v = Y[t] + bias[t / N];
        Y[t] = relu ? fmaxf(v, 0.0f) : v;
    }
}

// dZ = dA where the ReLU output A is positive, else 0 (in place)
__global__ void k_relu_bwd(float *dA, const float *A, long n) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        if (!(A[t] > 0.0f)) dA[t] = 0.0f;
}

__global__ void k_to_bf16(const float *x, bf16 *y, long n) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x)
        y[t] = to_bf16(x[t]);
}

// part[r P + p] = sum of row r of X over the p-th of P equal column ranges:
// blocks (r, p), fixed order within each
__global__ __launch_bounds__(256) void k_rowsum_part(const float *X, float *part, long N) {
    __shared__ float sh[256];
    const long len = (N + gridDim.y - 1) / gridDim.y;
    const long c0 = (long)blockIdx.y * len, c1 = min(N, c0 + len);
    const float *row = X + (long)blockIdx.x * N;
    float s = 0.0f;
    for (long c = c0 + threadIdx.x; c < c1; c += 256) s += row[c];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * gridDim.y + blockIdx.y] = sh[0];
}

// out[r] = sum_p part[r P + p], p in order
__global__ void k_rowsum_final(const float *part, float *out, int rows, int P) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    float s = 0.0f;
    for (int p = 0; p < P; p++) s += part[(long)r * P + p];
    out[r] = s;
}

// nearest x2 upsampling of [C][B][h][w] and its adjoint
__global__ void k_up2(const float *X, float *Y, long CB, int h, int w) {
    const long total = CB * 4 * h * w;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % (2 * w));
        const long t2 = t / (2 * w);
        const int y = (int)(t2 % (2 * h));
        const long cb = t2 / (2 * h);
        Y[t] = X[(cb * h + y / 2) * w + x / 2];
    }
}
__global__ void k_up2_bwd(const float *dY, float *dX, long CB, int h, int w) {
    const long total = CB * h * w;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int x = (int)(t % w);
        const long t2 = t / w;
        const int y = (int)(t2 % h);
        const long cb = t2 / h;
        const float *d = dY + (cb * 2 * h + 2 * y) * (2 * w) + 2 * x;
        dX[t] = (d[0] + d[1]) + (d[2 * w] + d[2 * w + 1]);
    }
}

// sample-major [B][C][P] <-> channel-major [C][B][P]
__global__ void k_to_cmajor(const float *src, float *dst, int B, int C, long P) {
    const long total = (long)B * C * P;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const long p = t % P, t2 = t / P;
        const int c = (int)(t2 % C), b = (int)(t2 / C);
        dst[((long)c * B + b) * P + p] = src[t];
    }
}
__global__ void k_to_smajor(const float *src, float *dst, int B, int C, long P) {
    const long total = (long)B * C * P;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const long p = t % P, t2 = t / P;
        const int c = (int)(t2 % C), b = (int)(t2 / C);
        dst[t] = src[((long)c * B + b) * P + p];
    }
}

// dY = 2 (Y - T) / n; part[block] = sum over the block of (Y - T)^2
__global__ __launch_bounds__(256) void k_mse(const float *Y, const float *T, float *dY,
                                             float *part, long n) {
    __shared__ float sh[256];
    float s = 0.0f;
    const float scale = 2.0f / (float)n;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const float d = Y[t] - T[t];
        dY[t] = scale * d;
        s += d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_adam(float *w, const float *g, float *m, float *v, long n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long)gridDim.x * blockDim.x) {
        const float gi = g[t];
        const float mi = b1 * m[t] + (1.0f - b1) * gi;
        const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
        m[t] = mi;
        v[t] = vi;
        w[t] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

constexpr int kMseBlocks = 256;

unsigned grid_for(long n) { return (unsigned)std::min<long>((n + 255) / 256, 65535L); }

}  // namespace nn

using nn::bf16;

namespace {

struct ConvLayer {
    int cin, cout, ks, stride, relu;
    long woff, boff;   // offsets into the flat parameter array
};

}  // namespace

struct lbm_nn_ctx {
    int G = 0, Bmax = 0, device = 0;
    int W = 16;   // base width: encoder W, 2W, 4W, 4W; decoder 4W, 2W, W channels
    hipStream_t s = nullptr;
    std::vector<ConvLayer> L;
    long np = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *wb = nullptr;
    long t = 0;   // Adam steps taken
    // activations and their gradients, [C][B][H][W]
    float *x0 = nullptr, *c7 = nullptr, *c6 = nullptr, *c5 = nullptr, *a4 = nullptr,
          *a5 = nullptr, *a6 = nullptr, *a7 = nullptr, *y = nullptr;
    float *dc7 = nullptr, *dc6 = nullptr, *dc5 = nullptr, *da4 = nullptr, *da5 = nullptr,
          *da6 = nullptr, *da7 = nullptr, *dy = nullptr;
    bf16 *col = nullptr, *dzb = nullptr;
    float *dcol = nullptr, *part = nullptr, *red = nullptr, *stage = nullptr, *tgt = nullptr;
    std::vector<void *> allocs;
};

namespace {

int32_t fail_hip(hipError_t e, const char *where) {
    return lbm::set_error(LBM_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define NN_TRY(expr)                                         \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return fail_hip(e_, #expr);    \
    } while (0)

void release(lbm_nn n) {
    for (void *p : n->allocs) (void)hipFree(p);
    if (n->s) (void)hipStreamDestroy(n->s);
    delete n;
}

template <typename T> int32_t dalloc(lbm_nn n, T *&p, long count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, sizeof(T) * (size_t)std::max(count, 1L));
    if (e != hipSuccess)
        return lbm::set_error(LBM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    n->allocs.push_back(q);
    p = (T *)q;
    return LBM_OK;
}

// splitmix64: the weight initialisation's random stream
uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// partial products of split-k products (splits x M x N floats) and the
// bias-gradient partial sums
constexpr long kPartFloats = 64L * 64 * 128 * 9;

hipError_t gemm(lbm_nn n, int M, int N, int K, const bf16 *A, long as_m, long as_k,
                const bf16 *B, long bs_k, long bs_n, float *C, long ldc,
                const nn::ConvGeom *conv = nullptr, int ks = 0) {
    // Waves per launch: one per MT*16 x 16 block of C and k range.  Split
    // the k range (at least 256 long, into the partials buffer) until a
    // launch has about kWaves waves; take four tiles per wave unless one
    // tile per wave gives more waves.
    constexpr long kWaves = 2048;
    const long tn = (N + 15) / 16, mn = (long)M * N;
    auto plan = [&](int mt, int &splits) {
        const long blocks = ((M + 16L * mt - 1) / (16L * mt)) * tn;
        const long want = (kWaves + blocks - 1) / blocks;
        splits = (int)std::max<long>(
            1, std::min<long>({want, std::max<long>(1, K / 256), 64L, kPartFloats / mn}));
        return blocks * splits;
    };
    int s4 = 1, s1 = 1, mt = 4, splits;
    const long w4 = plan(4, s4);
    splits = s4;
    if (w4 < kWaves && M > 16 && plan(1, s1) > w4) {
        mt = 1;
        splits = s1;
    }
    const int kchunk = ((K + splits - 1) / splits + 15) / 16 * 16;
    splits = (K + kchunk - 1) / kchunk;
    nn::GemmArgs g{A, as_m, as_k, B, bs_k, bs_n, splits > 1 ? n->part : C,
                   splits > 1 ? (long)N : ldc, M, N, K, kchunk};
    const long blocks = ((M + 16L * mt - 1) / (16L * mt)) * tn;
    dim3 grid((unsigned)((blocks + 3) / 4), 1, (unsigned)splits);
    if (conv && ks != 1 && ks != 3) return hipErrorInvalidValue;   // kernels for 1x1 and 3x3
    // 8-byte fragment loads where an operand's k axis is contiguous and
    // aligned and every k range ends on a multiple of 4 (the ranges are
    // multiples of 16 but the last, which ends at K)
    const bool av = as_k == 1 && (as_m & 3) == 0 && ((uintptr_t)A & 7) == 0 && (K & 3) == 0;
    const bool bv = bs_k == 1 && (bs_n & 3) == 0 && ((uintptr_t)B & 7) == 0 && (K & 3) == 0;
    const dim3 block(256);
    if (conv) {
        auto c = [&](auto kern) { hipLaunchKernelGGL(kern, grid, block, 0, n->s, g, *conv); };
        if (ks == 3 && mt == 4) av ? c(nn::k_gemm_conv<4, 3, true>) : c(nn::k_gemm_conv<4, 3, false>);
        else if (ks == 3) av ? c(nn::k_gemm_conv<1, 3, true>) : c(nn::k_gemm_conv<1, 3, false>);
        else if (mt == 4) av ? c(nn::k_gemm_conv<4, 1, true>) : c(nn::k_gemm_conv<4, 1, false>);
        else av ? c(nn::k_gemm_conv<1, 1, true>) : c(nn::k_gemm_conv<1, 1, false>);
    } else {
        auto l = [&](auto kern) { hipLaunchKernelGGL(kern, grid, block, 0, n->s, g); };
        if (mt == 4) {
            if (av && bv) l(nn::k_gemm<4, true, true>);
            else if (av) l(nn::k_gemm<4, true, false>);
            else if (bv) l(nn::k_gemm<4, false, true>);
            else l(nn::k_gemm<4, false, false>);
        } else {
            if (av && bv) l(nn::k_gemm<1, true, true>);
            else if (av) l(nn::k_gemm<1, true, false>);
            else if (bv) l(nn::k_gemm<1, false, true>);
            else l(nn::k_gemm<1, false, false>);
        }
    }
    if (splits > 1)
        hipLaunchKernelGGL(nn::k_reduce_splits, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0,
                           n->s, (const float *)n->part, C, ldc, M, N, splits);
    return hipGetLastError();
}

// col of the K x N im2col matrix of X ([C][B][H][H], Ho x Ho outputs)
void im2col(lbm_nn n, const float *X, int B, int H, int ks, int stride, int pad, int Ho, long K) {
    const long N = (long)B * Ho * Ho;
    if (Ho % 8 == 0)
        hipLaunchKernelGGL(nn::k_im2col8, dim3(nn::grid_for(N / 8), (unsigned)K), dim3(256), 0,
                           n->s, X, n->col, B, H, H, ks, stride, pad, Ho, Ho);
    else
        hipLaunchKernelGGL(nn::k_im2col, dim3(nn::grid_for(N), (unsigned)K), dim3(256), 0, n->s,
                           X, n->col, B, H, H, ks, stride, pad, Ho, Ho);
}

// forward of one convolution: Y[cout][B][Ho][Wo] = act(W * X + b)
hipError_t conv_fwd(lbm_nn n, const ConvLayer &l, int B, const float *X, int H, float *Y) {
    const int Ho = H / l.stride, pad = l.ks / 2;
    const long N = (long)B * Ho * Ho, K = (long)l.cin * l.ks * l.ks;
    // the product gathers its right operand from X (no im2col matrix)
    const nn::ConvGeom cg{X, B, H, H, l.stride, pad, Ho, Ho};
    hipError_t e = gemm(n, l.cout, (int)N, (int)K, n->wb + l.woff, K, 1, nullptr, N, 1, Y, N,
                        &cg, l.ks);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nn::k_bias_act, dim3(nn::grid_for((long)l.cout * N)), dim3(256), 0, n->s,
                       Y, (const float *)(n->w + l.boff), l.cout, N, l.relu);
    return hipGetLastError();
}

// backward of one convolution.  dA (the gradient of its output) is turned
// into dZ in place; dW, db go to the gradient array; dX is written or added
// to when given.
hipError_t conv_bwd(lbm_nn n, const ConvLayer &l, int B, const float *X, int H, const float *A,
                    float *dA, float *dX, bool accumulate) {
    const int Ho = H / l.stride, pad = l.ks / 2;
    const long N = (long)B * Ho * Ho, K = (long)l.cin * l.ks * l.ks;
    if (l.relu)
        hipLaunchKernelGGL(nn::k_relu_bwd, dim3(nn::grid_for(l.cout * N)), dim3(256), 0, n->s, dA,
                           A, (long)l.cout * N);
    // bias gradient: row sums of dZ, split into P column ranges per row
    const int P = (int)std::min<long>(64, std::max<long>(1, (N + 8191) / 8192));
    hipLaunchKernelGGL(nn::k_rowsum_part, dim3(l.cout, P), dim3(256), 0, n->s,
                       (const float *)dA, n->part, N);
    hipLaunchKernelGGL(nn::k_rowsum_final, dim3((l.cout + 63) / 64), dim3(64), 0, n->s,
                       (const float *)n->part, n->g + l.boff, l.cout, P);
    hipLaunchKernelGGL(nn::k_to_bf16, dim3(nn::grid_for(l.cout * N)), dim3(256), 0, n->s,
                       (const float *)dA, n->dzb, (long)l.cout * N);
    im2col(n, X, B, H, l.ks, l.stride, pad, Ho, K);
    // dW[cout][K] = dZ[cout][N] . col[K][N]^T
    hipError_t e = gemm(n, l.cout, (int)K, (int)N, n->dzb, N, 1, n->col, 1, N, n->g + l.woff, K);
    if (e != hipSuccess || !dX) return e;
    // dcol[K][N] = W^T . dZ
    e = gemm(n, (int)K, (int)N, l.cout, n->wb + l.woff, 1, K, n->dzb, N, 1, n->dcol, N);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(nn::k_col2im, dim3(nn::grid_for((long)l.cin * B * H * H)), dim3(256), 0,
                       n->s, (const float *)n->dcol, dX, l.cin, B, H, H, l.ks, l.stride, pad, Ho,
                       Ho, accumulate ? 1 : 0);
    return hipGetLastError();
}

int32_t check_net(lbm_nn n, int32_t batch) {
    if (!n) return lbm::set_error(LBM_ERR_ARG, "null network");
    if (batch < 1 || batch > n->Bmax) return lbm::set_error(LBM_ERR_ARG, "batch out of range");
    NN_TRY(hipSetDevice(n->device));
    return LBM_OK;
}

// x (host, sample-major) -> y (device, channel-major) through the network
int32_t forward(lbm_nn n, int B, const float *x_host) {
    const int G = n->G;
    const long P = (long)G * G, W = n->W;
    NN_TRY(hipMemcpyAsync(n->stage, x_host, sizeof(float) * B * LBM_NN_IN_CHANNELS * P,
                          hipMemcpyHostToDevice, n->s));
    hipLaunchKernelGGL(nn::k_to_cmajor, dim3(nn::grid_for(B * 4 * P)), dim3(256), 0, n->s,
                       (const float *)n->stage, n->x0, B, 4, P);
    hipLaunchKernelGGL(nn::k_to_bf16, dim3(nn::grid_for(n->np)), dim3(256), 0, n->s,
                       (const float *)n->w, n->wb, n->np);
    const int G2 = G / 2, G4 = G / 4, G8 = G / 8;
    // the encoder maps a1, a2, a3 sit behind the upsampled decoder maps in
    // the concatenations c7 = [up(a6), a1], c6 = [up(a5), a2], c5 = [up(a4), a3]
    float *A1 = n->c7 + 2 * W * B * P, *A2 = n->c6 + 4 * W * B * G2 * G2,
          *A3 = n->c5 + 4 * W * B * G4 * G4;
    hipError_t e = hipSuccess;
    auto run = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    run(conv_fwd(n, n->L[0], B, n->x0, G, A1));
    run(conv_fwd(n, n->L[1], B, A1, G, A2));
    run(conv_fwd(n, n->L[2], B, A2, G2, A3));
    run(conv_fwd(n, n->L[3], B, A3, G4, n->a4));
    hipLaunchKernelGGL(nn::k_up2, dim3(nn::grid_for(4 * W * B * G4 * G4)), dim3(256), 0, n->s,
                       (const float *)n->a4, n->c5, 4 * W * B, G8, G8);
    run(conv_fwd(n, n->L[4], B, n->c5, G4, n->a5));
    hipLaunchKernelGGL(nn::k_up2, dim3(nn::grid_for(4 * W * B * G2 * G2)), dim3(256), 0, n->s,
                       (const float *)n->a5, n->c6, 4 * W * B, G4, G4);
    run(conv_fwd(n, n->L[5], B, n->c6, G2, n->a6));
    hipLaunchKernelGGL(nn::k_up2, dim3(nn::grid_for(2 * W * B * P)), dim3(256), 0, n->s,
                       (const float *)n->a6, n->c7, 2 * W * B, G2, G2);
    run(conv_fwd(n, n->L[6], B, n->c7, G, n->a7));
    run(conv_fwd(n, n->L[7], B, n->a7, G, n->y));
    run(hipGetLastError());
    if (e != hipSuccess) return fail_hip(e, "forward");
    return LBM_OK;
}

// loss of y against target (host, sample-major); fills dy and the gradient
int32_t backward(lbm_nn n, int B, const float *target, float *loss) {
    const int G = n->G, G2 = G / 2, G4 = G / 4, G8 = G / 8;
    const long P = (long)G * G, W = n->W;
    const long ny = 3L * B * P;
    NN_TRY(hipMemcpyAsync(n->stage, target, sizeof(float) * ny, hipMemcpyHostToDevice, n->s));
    hipLaunchKernelGGL(nn::k_to_cmajor, dim3(nn::grid_for(ny)), dim3(256), 0, n->s,
                       (const float *)n->stage, n->tgt, B, 3, P);
    hipLaunchKernelGGL(nn::k_mse, dim3(nn::kMseBlocks), dim3(256), 0, n->s, (const float *)n->y,
                       (const float *)n->tgt, n->dy, n->red, ny);
    float *A1 = n->c7 + 2 * W * B * P, *A2 = n->c6 + 4 * W * B * G2 * G2,
          *A3 = n->c5 + 4 * W * B * G4 * G4;
    float *dA1 = n->dc7 + 2 * W * B * P, *dA2 = n->dc6 + 4 * W * B * G2 * G2,
          *dA3 = n->dc5 + 4 * W * B * G4 * G4;
    hipError_t e = hipSuccess;
    auto run = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    run(conv_bwd(n, n->L[7], B, n->a7, G, n->y, n->dy, n->da7, false));
    run(conv_bwd(n, n->L[6], B, n->c7, G, n->a7, n->da7, n->dc7, false));
    hipLaunchKernelGGL(nn::k_up2_bwd, dim3(nn::grid_for(2 * W * B * G2 * G2)), dim3(256), 0,
                       n->s, (const float *)n->dc7, n->da6, 2 * W * B, G2, G2);
    run(conv_bwd(n, n->L[5], B, n->c6, G2, n->a6, n->da6, n->dc6, false));
    hipLaunchKernelGGL(nn::k_up2_bwd, dim3(nn::grid_for(4 * W * B * G4 * G4)), dim3(256), 0,
                       n->s, (const float *)n->dc6, n->da5, 4 * W * B, G4, G4);
    run(conv_bwd(n, n->L[4], B, n->c5, G4, n->a5, n->da5, n->dc5, false));
    hipLaunchKernelGGL(nn::k_up2_bwd, dim3(nn::grid_for(4 * W * B * G8 * G8)), dim3(256), 0,
                       n->s, (const float *)n->dc5, n->da4, 4 * W * B, G8, G8);
    run(conv_bwd(n, n->L[3], B, A3, G4, n->a4, n->da4, dA3, true));
    run(conv_bwd(n, n->L[2], B, A2, G2, A3, dA3, dA2, true));
    run(conv_bwd(n, n->L[1], B, A1, G, A2, dA2, dA1, true));
    run(conv_bwd(n, n->L[0], B, n->x0, G, A1, dA1, nullptr, false));
    run(hipGetLastError());
    if (e != hipSuccess) return fail_hip(e, "backward");
    std::vector<float> part(nn::kMseBlocks);
    NN_TRY(hipMemcpyAsync(part.data(), n->red, sizeof(float) * part.size(),
                          hipMemcpyDeviceToHost, n->s));
    NN_TRY(hipStreamSynchronize(n->s));
    double s = 0.0;
    for (float p : part) s += p;
    if (loss) *loss = (float)(s / (double)ny);
    return LBM_OK;
}

}  // namespace

extern "C" {

int32_t lbm_nn_create(int32_t grid, int32_t batch_max, uint64_t seed, int32_t device,
                      lbm_nn *out) {
    return lbm_nn_create_width(grid, 16, batch_max, seed, device, out);
}

int32_t lbm_nn_create_width(int32_t grid, int32_t width, int32_t batch_max, uint64_t seed,
                            int32_t device, lbm_nn *out) {
    if (!out) return lbm::set_error(LBM_ERR_ARG, "null output");
    *out = nullptr;
    if (grid < 16 || grid % 8 || batch_max < 1 || (long)batch_max * grid * grid > (1L << 26))
        return lbm::set_error(LBM_ERR_ARG, "grid must be a multiple of 8 (>= 16), batch >= 1");
    if (width < 4 || width > 64 || width % 4)
        return lbm::set_error(LBM_ERR_ARG, "width must be a multiple of 4 from 4 to 64");
    lbm_nn n = new (std::nothrow) lbm_nn_ctx();
    if (!n) return lbm::set_error(LBM_ERR_NOMEM, "network");
    n->G = grid;
    n->W = width;
    n->Bmax = batch_max;
    n->device = device;
    // layers: cin, cout, k, stride, relu
    const int W = width;
    const int spec[8][5] = {{4, W, 3, 1, 1},         {W, 2 * W, 3, 2, 1},
                            {2 * W, 4 * W, 3, 2, 1}, {4 * W, 4 * W, 3, 2, 1},
                            {8 * W, 4 * W, 3, 1, 1}, {6 * W, 2 * W, 3, 1, 1},
                            {3 * W, W, 3, 1, 1},     {W, 3, 1, 1, 0}};
    long off = 0;
    for (auto &sp : spec) {
        ConvLayer l{sp[0], sp[1], sp[2], sp[3], sp[4], 0, 0};
        l.woff = off;
        off += (long)l.cout * l.cin * l.ks * l.ks;
        l.boff = off;
        off += l.cout;
        n->L.push_back(l);
    }
    n->np = off;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking);
    if (e != hipSuccess) {
        release(n);
        return fail_hip(e, "network stream");
    }
    const long B = batch_max, P = (long)grid * grid;
    const long P2 = P / 4, P4 = P / 16, P8 = P / 64;
    // largest im2col matrix: K N over the layers
    long colmax = 0, zmax = 0;
    {
        const int hin[8] = {grid, grid, grid / 2, grid / 4, grid / 4, grid / 2, grid, grid};
        for (int i = 0; i < 8; i++) {
            const ConvLayer &l = n->L[i];
            const long Ho = hin[i] / l.stride;
            const long N = B * Ho * Ho;
            colmax = std::max(colmax, (long)l.cin * l.ks * l.ks * N);
            zmax = std::max(zmax, (long)l.cout * N);
        }
    }
    int32_t rc = LBM_OK;
    auto A = [&](auto *&p, long cnt) { if (!rc) rc = dalloc(n, p, cnt); };
    A(n->w, n->np); A(n->g, n->np); A(n->m, n->np); A(n->v, n->np); A(n->wb, n->np);
    const long Wl = W;
    A(n->x0, 4 * B * P); A(n->c7, 3 * Wl * B * P); A(n->c6, 6 * Wl * B * P2);
    A(n->c5, 8 * Wl * B * P4); A(n->a4, 4 * Wl * B * P8); A(n->a5, 4 * Wl * B * P4);
    A(n->a6, 2 * Wl * B * P2); A(n->a7, Wl * B * P);
    A(n->y, 3 * B * P);
    A(n->dc7, 3 * Wl * B * P); A(n->dc6, 6 * Wl * B * P2); A(n->dc5, 8 * Wl * B * P4);
    A(n->da4, 4 * Wl * B * P8); A(n->da5, 4 * Wl * B * P4); A(n->da6, 2 * Wl * B * P2);
    A(n->da7, Wl * B * P);
    A(n->dy, 3 * B * P);
    A(n->col, colmax); A(n->dzb, zmax); A(n->dcol, colmax);
    A(n->part, kPartFloats); A(n->red, nn::kMseBlocks);
    A(n->stage, 4 * B * P); A(n->tgt, 3 * B * P);
    if (rc) {
        release(n);
        return rc;
    }
    // He-uniform weights, zero biases
    std::vector<float> w(n->np, 0.0f);
    uint64_t st = seed;
    for (const ConvLayer &l : n->L) {
        const double lim = std::sqrt(6.0 / (double)(l.cin * l.ks * l.ks));
        for (long i = 0; i < (long)l.cout * l.cin * l.ks * l.ks; i++) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            w[l.woff + i] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    *out = n;
    return lbm_nn_set_params(n, w.data());
}

int32_t lbm_nn_destroy(lbm_nn n) {
    if (!n) return LBM_OK;
    (void)hipSetDevice(n->device);
    (void)hipStreamSynchronize(n->s);
    release(n);
    return LBM_OK;
}

int32_t lbm_nn_param_count(lbm_nn n, int64_t *count) {
    if (!n || !count) return lbm::set_error(LBM_ERR_ARG, "null argument");
    *count = n->np;
    return LBM_OK;
}

int32_t lbm_nn_get_params(lbm_nn n, float *params) {
    if (!n || !params) return lbm::set_error(LBM_ERR_ARG, "null argument");
    NN_TRY(hipSetDevice(n->device));
    NN_TRY(hipStreamSynchronize(n->s));
    NN_TRY(hipMemcpy(params, n->w, sizeof(float) * n->np, hipMemcpyDeviceToHost));
    return LBM_OK;
}

int32_t lbm_nn_set_params(lbm_nn n, const float *params) {
    if (!n || !params) return lbm::set_error(LBM_ERR_ARG, "null argument");
    NN_TRY(hipSetDevice(n->device));
    NN_TRY(hipStreamSynchronize(n->s));
    NN_TRY(hipMemcpy(n->w, params, sizeof(float) * n->np, hipMemcpyHostToDevice));
    NN_TRY(hipMemset(n->m, 0, sizeof(float) * n->np));
    NN_TRY(hipMemset(n->v, 0, sizeof(float) * n->np));
    n->t = 0;
    return LBM_OK;
}

int32_t lbm_nn_forward(lbm_nn n, int32_t batch, const float *x, float *y) {
    int32_t rc = check_net(n, batch);
    if (rc) return rc;
    if (!x || !y) return lbm::set_error(LBM_ERR_ARG, "null array");
    if ((rc = forward(n, batch, x))) return rc;
    const long P = (long)n->G * n->G;
    hipLaunchKernelGGL(nn::k_to_smajor, dim3(nn::grid_for(3L * batch * P)), dim3(256), 0, n->s,
                       (const float *)n->y, n->stage, batch, 3, P);
    NN_TRY(hipMemcpyAsync(y, n->stage, sizeof(float) * 3 * batch * P, hipMemcpyDeviceToHost,
                          n->s));
    NN_TRY(hipStreamSynchronize(n->s));
    return LBM_OK;
}

int32_t lbm_nn_loss_grad(lbm_nn n, int32_t batch, const float *x, const float *target,
                         float *loss, float *grad) {
    int32_t rc = check_net(n, batch);
    if (rc) return rc;
    if (!x || !target) return lbm::set_error(LBM_ERR_ARG, "null array");
    if ((rc = forward(n, batch, x)) || (rc = backward(n, batch, target, loss))) return rc;
    if (grad) NN_TRY(hipMemcpy(grad, n->g, sizeof(float) * n->np, hipMemcpyDeviceToHost));
    return LBM_OK;
}

int32_t lbm_nn_train_step(lbm_nn n, int32_t batch, const float *x, const float *target,
                          float lr, float *loss) {
    int32_t rc = check_net(n, batch);
    if (rc) return rc;
    if (!x || !target) return lbm::set_error(LBM_ERR_ARG, "null array");
    if ((rc = forward(n, batch, x)) || (rc = backward(n, batch, target, loss))) return rc;
    n->t++;
    const float b1 = 0.9f, b2 = 0.999f;
    // bias corrections 1 - beta^t of the decay rates the kernel applies,
    // formed in FP64 and rounded once: 1.0f - (float)beta^t would lose
    // five digits to cancellation (1 - 0.999^2 = 0.001999)
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)n->t));
    const float c2 = (float)(1.0 - std::pow((double)b2, (double)n->t));
    hipLaunchKernelGGL(nn::k_adam, dim3(nn::grid_for(n->np)), dim3(256), 0, n->s, n->w,
                       (const float *)n->g, n->m, n->v, n->np, lr, b1, b2, 1e-8f, c1, c2);
    NN_TRY(hipGetLastError());
    return LBM_OK;
}

}  // extern "C"
