// lbm_step_f64.hip -- FP64 instantiation of the fused step with unshifted
// populations and the force stored per node.  This translation unit is
// compiled with -ffp-contract=off, so every product and sum is rounded on
// its own, in the same order as the FP64 CPU restatement.
#include "lbm_internal.h"

namespace lbm {

hipError_t launch_step_f64(int dim, bool macro, const StepArgs<double> &a0, const LaunchShape &ls,
                           hipStream_t s) {
    StepArgs<double> a = a0;
    dim3 block(ls.bx, ls.by, 1);
    a.gx = (a.n + ls.bx - 1) / ls.bx;
    a.gy = (a.n + ls.by - 1) / ls.by;
    a.ntiles = a.gx * a.gy * (dim == 3 ? a.n : 1);
    a.remap = (ls.remap == 2 && a.gy % 8) ? 0 : ls.remap;
    dim3 grid(a.ntiles);
    if (dim == 2) {
        if (macro) hipLaunchKernelGGL((k_step<2, double, double, false, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step<2, double, double, false, false>), grid, block, 0, s, a);
    } else {
        if (macro) hipLaunchKernelGGL((k_step<3, double, double, false, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((k_step<3, double, double, false, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_prologue_f64(int dim, const StepArgs<double> &a,
                               const StepArgs<double> &ad, hipStream_t s) {
    long nn = dim == 2 ? (long)a.n * a.n : (long)a.n * a.n * a.n;
    dim3 block(256);
    dim3 grid((unsigned)((nn + 255) / 256));
    if (dim == 2) hipLaunchKernelGGL((k_prologue<2, double, double, false>), grid, block, 0, s, a, ad);
    else hipLaunchKernelGGL((k_prologue<3, double, double, false>), grid, block, 0, s, a, ad);
    return hipGetLastError();
}

}  // namespace lbm
