// lbm_step_f32.hip -- FP32 shifted-population instantiation of the fused
// step (the production path).  Built with default FMA contraction.
#include "lbm_internal.h"
#include "lbm_step3z.h"

namespace lbm {

// Tile shapes with a z-marching instantiation.
static const int kZShapes[][2] = {{64, 4}, {64, 8}, {128, 2}, {128, 4}, {32, 8}, {96, 6}};

bool zmarch_supported(int n, int bx, int by) {
    bool known = false;
    for (const auto &s : kZShapes) known = known || (s[0] == bx && s[1] == by);
    return known && n >= 5 && bx >= 3 && by >= 2 && n % bx != 1 && n % by != 1;
}

template <int TX, int TY>
static void launch_z(bool macro, StepArgs<float> &a, int kc, hipStream_t s) {
    dim3 block(TX, TY, 1);
    dim3 grid(a.ntiles);
    if (macro) hipLaunchKernelGGL((k_step3z<TX, TY, true>), grid, block, 0, s, a, kc);
    else hipLaunchKernelGGL((k_step3z<TX, TY, false>), grid, block, 0, s, a, kc);
}

hipError_t launch_step_f32(int dim, bool macro, const StepArgs<float> &a0, const LaunchShape &ls,
                           hipStream_t s) {
    StepArgs<float> a = a0;
    if (dim == 3 && ls.kind == 1 && zmarch_supported(a.n, ls.bx, ls.by)) {
        const int kc = ls.zchunk < 1 ? a.n : ls.zchunk;
        a.gx = (a.n + ls.bx - 1) / ls.bx;
        a.gy = (a.n + ls.by - 1) / ls.by;
        a.ntiles = a.gx * a.gy * ((a.n + kc - 1) / kc);
        a.remap = (ls.remap == 2 && a.gy % 8) ? 0 : ls.remap;
        if (ls.bx == 64 && ls.by == 4) launch_z<64, 4>(macro, a, kc, s);
        else if (ls.bx == 64 && ls.by == 8) launch_z<64, 8>(macro, a, kc, s);
        else if (ls.bx == 128 && ls.by == 2) launch_z<128, 2>(macro, a, kc, s);
        else if (ls.bx == 128 && ls.by == 4) launch_z<128, 4>(macro, a, kc, s);
        else if (ls.bx == 32 && ls.by == 8) launch_z<32, 8>(macro, a, kc, s);
        else launch_z<96, 6>(macro, a, kc, s);
        return hipGetLastError();
    }
    dim3 block(ls.bx, ls.by, 1);
    a.gx = (a.n + ls.bx - 1) / ls.bx;
    a.gy = (a.n + ls.by - 1) / ls.by;
    a.ntiles = a.gx * a.gy * (dim == 3 ? a.n : 1);
    a.remap = (ls.remap == 2 && a.gy % 8) ? 0 : ls.remap;
    dim3 grid(a.ntiles);
    if (dim == 2) {
        if (macro) hipLaunchKernelGGL((k_step<2, float, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step<2, float, true, false>), grid, block, 0, s, a);
    } else if (ls.occ == 5) {
        if (macro) hipLaunchKernelGGL((k_step<3, float, true, true, 5>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step<3, float, true, false, 5>), grid, block, 0, s, a);
    } else if (ls.occ == 6) {
        if (macro) hipLaunchKernelGGL((k_step<3, float, true, true, 6>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step<3, float, true, false, 6>), grid, block, 0, s, a);
    } else {
        if (macro) hipLaunchKernelGGL((k_step<3, float, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step<3, float, true, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_prologue_f32(int dim, const StepArgs<float> &a,
                               const StepArgs<double> &ad, hipStream_t s) {
    long nn = dim == 2 ? (long)a.n * a.n : (long)a.n * a.n * a.n;
    dim3 block(256);
    dim3 grid((unsigned)((nn + 255) / 256));
    if (dim == 2) hipLaunchKernelGGL((k_prologue<2, float, true>), grid, block, 0, s, a, ad);
    else hipLaunchKernelGGL((k_prologue<3, float, true>), grid, block, 0, s, a, ad);
    return hipGetLastError();
}

}  // namespace lbm
