// wenofv_kernels_mixed.hip -- mixed-precision stage kernels (FP32 WENO and flux arithmetic,
// FP64 equation of state, theta and RK), compiled with fused multiply-add contraction.
#include "stage_impl.h"

// minimum wavefronts per SIMD the register allocation must allow
#ifndef WFV_MIXED_MINW
#define WFV_MIXED_MINW 2
#endif

namespace wfv {

template <int RECON, int SCHEME, int FLUX>
static hipError_t launch_t(const StageArgs &A, hipStream_t s) {
    // one wavefront per workgroup: waves spread evenly over the CUs
    const int nstrip = A.strips_x * A.strips_y;
    hipLaunchKernelGGL((stage_kernel<PrecMixed, RECON, SCHEME, FLUX, WFV_MIXED_MINW>), dim3(nstrip), dim3(LANES), 0, s, A);
    return hipGetLastError();
}

template <int RECON, int SCHEME>
static hipError_t launch_f(const StageArgs &A, hipStream_t s) {
    return A.flux == 1 ? launch_t<RECON, SCHEME, 1>(A, s) : launch_t<RECON, SCHEME, 0>(A, s);
}

hipError_t launch_stage_mixed(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    switch (recon * 2 + scheme) {
    case 0: return launch_f<0, 0>(A, s);
    case 1: return launch_f<0, 1>(A, s);
    case 2: return launch_f<1, 0>(A, s);
    case 3: return launch_f<1, 1>(A, s);
    case 4: return launch_f<2, 0>(A, s);
    case 5: return launch_f<2, 1>(A, s);
    case 6: return launch_f<3, 0>(A, s);
    case 7: return launch_f<3, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

template <int RECON, int SCHEME, int FLUX>
static hipError_t launch_bt(const BatchStageArgs &B, int grid, hipStream_t s) {
    if (grid < B.total_items)  // persistent: the wavefronts stride over the items
        hipLaunchKernelGGL((batch_stage_kernel<PrecMixed, RECON, SCHEME, FLUX, WFV_MIXED_MINW, true>), dim3(grid), dim3(LANES), 0,
                           s, B);
    else
        hipLaunchKernelGGL((batch_stage_kernel<PrecMixed, RECON, SCHEME, FLUX, WFV_MIXED_MINW, false>), dim3(B.total_items),
                           dim3(LANES), 0, s, B);
    return hipGetLastError();
}

template <int RECON, int SCHEME>
static hipError_t launch_bf(const BatchStageArgs &B, int flux, int grid, hipStream_t s) {
    return flux == 1 ? launch_bt<RECON, SCHEME, 1>(B, grid, s) : launch_bt<RECON, SCHEME, 0>(B, grid, s);
}

hipError_t launch_batch_mixed(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s) {
    switch (recon * 2 + scheme) {
    case 0: return launch_bf<0, 0>(B, flux, grid, s);
    case 1: return launch_bf<0, 1>(B, flux, grid, s);
    case 2: return launch_bf<1, 0>(B, flux, grid, s);
    case 3: return launch_bf<1, 1>(B, flux, grid, s);
    case 4: return launch_bf<2, 0>(B, flux, grid, s);
    case 5: return launch_bf<2, 1>(B, flux, grid, s);
    case 6: return launch_bf<3, 0>(B, flux, grid, s);
    case 7: return launch_bf<3, 1>(B, flux, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
