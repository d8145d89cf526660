// lbm_step_f64.hip -- FP64 instantiation of the fused step with unshifted
// populations and the force stored per node.  This translation unit is
// compiled with -ffp-contract=off, so every product and sum is rounded on
// its own, in the same order as the FP64 CPU restatement.
#include "lbm_internal.h"

namespace lbm {

template <int DIM, bool NT, int MAXT>
static void launch_k(bool macro, dim3 grid, dim3 block, const StepArgs<double> &a, hipStream_t s) {
    if (macro)
        hipLaunchKernelGGL((k_step<DIM, double, double, false, true, 0, NT, MAXT>), grid, block, 0,
                           s, a);
    else
        hipLaunchKernelGGL((k_step<DIM, double, double, false, false, 0, NT, MAXT>), grid, block,
                           0, s, a);
}

// 3D blocks of up to 256 threads run the instance compiled for 256, whose
// register budget holds the FP64 D3Q19 + D3Q7 node update without spilling
// (the 1024-thread instance is capped at 128 VGPRs and spills about 70).
template <int DIM>
static void launch_k(bool macro, bool nt, dim3 grid, dim3 block, const StepArgs<double> &a,
                     hipStream_t s) {
    const bool small = DIM == 3 && block.x * block.y <= 256;
    if (nt && small) launch_k<DIM, true, 256>(macro, grid, block, a, s);
    else if (nt) launch_k<DIM, true, 1024>(macro, grid, block, a, s);
    else if (small) launch_k<DIM, false, 256>(macro, grid, block, a, s);
    else launch_k<DIM, false, 1024>(macro, grid, block, a, s);
}

hipError_t launch_step_f64(int dim, bool macro, const StepArgs<double> &a0, const LaunchShape &ls,
                           hipStream_t s) {
    StepArgs<double> a = a0;
    dim3 block(ls.bx, ls.by, 1);
    a.gx = (a.n + ls.bx - 1) / ls.bx;
    a.gy = (a.n + ls.by - 1) / ls.by;
    a.ntiles = a.gx * a.gy * (dim == 3 ? a.n : 1);
    a.remap = (ls.remap == 2 && a.gy % 8) ? 0 : ls.remap;
    dim3 grid(a.ntiles);
    // populations and the per-node force, read and written
    const double nodes = dim == 2 ? (double)a.n * a.n : (double)a.n * a.n * a.n;
    const bool nt = stream_stores(nodes * ((dim == 2 ? 14 : 26) + 1) * 8.0 * 2.0);
    if (dim == 2) launch_k<2>(macro, nt, grid, block, a, s);
    else launch_k<3>(macro, nt, grid, block, a, s);
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
