// lbm_step_f32.hip -- FP32 shifted-population instantiation of the fused
// step (the production path).  Built with default FMA contraction.
#include "lbm_internal.h"

namespace lbm {

template <int DIM, int WPE>
static void launch_k(bool macro, dim3 grid, dim3 block, const StepArgs<float> &a,
                     hipStream_t s) {
    if (macro)
        hipLaunchKernelGGL((k_step<DIM, float, float, true, true, WPE>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((k_step<DIM, float, float, true, false, WPE>), grid, block, 0, s, a);
}

hipError_t launch_step_f32(int dim, bool macro, const StepArgs<float> &a0, const LaunchShape &ls,
                           hipStream_t s) {
    StepArgs<float> a = a0;
    dim3 block(ls.bx, ls.by, 1);
    a.gx = (a.n + ls.bx - 1) / ls.bx;
    a.gy = (a.n + ls.by - 1) / ls.by;
    a.ntiles = a.gx * a.gy * (dim == 3 ? a.n : 1);
    a.remap = (ls.remap == 2 && a.gy % 8) ? 0 : ls.remap;
    dim3 grid(a.ntiles);
    if (dim == 2) launch_k<2, 0>(macro, grid, block, a, s);
    else if (ls.occ == 5) launch_k<3, 5>(macro, grid, block, a, s);
    else if (ls.occ == 6) launch_k<3, 6>(macro, grid, block, a, s);
    else launch_k<3, 0>(macro, grid, block, a, s);
    return hipGetLastError();
}

hipError_t launch_prologue_f32(int dim, const StepArgs<float> &a,
                               const StepArgs<double> &ad, hipStream_t s) {
    long nn = dim == 2 ? (long)a.n * a.n : (long)a.n * a.n * a.n;
    dim3 block(256);
    dim3 grid((unsigned)((nn + 255) / 256));
    if (dim == 2) hipLaunchKernelGGL((k_prologue<2, float, float, true>), grid, block, 0, s, a, ad);
    else hipLaunchKernelGGL((k_prologue<3, float, float, true>), grid, block, 0, s, a, ad);
    return hipGetLastError();
}

}  // namespace lbm
