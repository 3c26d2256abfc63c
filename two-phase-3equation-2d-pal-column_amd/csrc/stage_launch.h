// stage_launch.h -- launch helpers of the fused stage kernels, instantiated per precision
// policy and scheme group by the wenofv_kernels_*.hip translation units.
#pragma once
#include "stage_impl.h"

#ifndef WFV_GH_SPLIT
#define WFV_GH_SPLIT 1
#endif

namespace wfv {

// one wavefront per workgroup: waves spread evenly over the CUs; the loader addresses a ring
// of fixed ghost cells only where the grid has one (boundary 3)
template <class PR, int MINW, int RECON, int SCHEME, int FLUX>
hipError_t launch_stage_t(const StageArgs &A, hipStream_t s) {
    const int nstrip = A.sp.total;
    if (A.bc == 3 || !WFV_GH_SPLIT)
        hipLaunchKernelGGL((stage_kernel<PR, RECON, SCHEME, FLUX, MINW, true>), dim3(nstrip), dim3(LANES), 0, s, A);
    else
        hipLaunchKernelGGL((stage_kernel<PR, RECON, SCHEME, FLUX, MINW, false>), dim3(nstrip), dim3(LANES), 0, s, A);
    return hipGetLastError();
}

template <class PR, int MINW, int RECON, int SCHEME>
hipError_t launch_stage_f(const StageArgs &A, hipStream_t s) {
    return A.flux == 1 ? launch_stage_t<PR, MINW, RECON, SCHEME, 1>(A, s)
                       : launch_stage_t<PR, MINW, RECON, SCHEME, 0>(A, s);
}

template <class PR, int MINW, int RECON, int SCHEME, int FLUX, bool GH>
hipError_t launch_batch_g(const BatchStageArgs &B, int grid, hipStream_t s) {
    if (grid < B.total_items)  // persistent: the wavefronts stride over the items
        hipLaunchKernelGGL((batch_stage_kernel<PR, RECON, SCHEME, FLUX, MINW, true, GH>), dim3(grid), dim3(LANES), 0, s,
                           B);
    else
        hipLaunchKernelGGL((batch_stage_kernel<PR, RECON, SCHEME, FLUX, MINW, false, GH>), dim3(B.total_items),
                           dim3(LANES), 0, s, B);
    return hipGetLastError();
}

// the ghost-ring loader only for a batch with a boundary-3 instance
template <class PR, int MINW, int RECON, int SCHEME, int FLUX>
hipError_t launch_batch_t(const BatchStageArgs &B, int grid, hipStream_t s) {
    return (B.ghost || !WFV_GH_SPLIT) ? launch_batch_g<PR, MINW, RECON, SCHEME, FLUX, true>(B, grid, s)
                   : launch_batch_g<PR, MINW, RECON, SCHEME, FLUX, false>(B, grid, s);
}

template <class PR, int MINW, int RECON, int SCHEME>
hipError_t launch_batch_f(const BatchStageArgs &B, int flux, int grid, hipStream_t s) {
    return flux == 1 ? launch_batch_t<PR, MINW, RECON, SCHEME, 1>(B, grid, s)
                     : launch_batch_t<PR, MINW, RECON, SCHEME, 0>(B, grid, s);
}

}  // namespace wfv
