// lbm_step_f32.hip -- FP32 shifted-population instantiation of the fused
// step (the production path).  Built with default FMA contraction.
#include "lbm_internal.h"

namespace lbm {

static dim3 step_grid(int dim, int n, dim3 block) {
    return dim3((n + block.x - 1) / block.x, (n + block.y - 1) / block.y,
                dim == 3 ? n : 1);
}

hipError_t launch_step_f32(int dim, bool macro, const StepArgs<float> &a, hipStream_t s) {
    dim3 block(64, 4, 1);
    dim3 grid = step_grid(dim, a.n, block);
    if (dim == 2) {
        if (macro) hipLaunchKernelGGL((k_step<2, float, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step<2, float, true, false>), grid, block, 0, s, a);
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
