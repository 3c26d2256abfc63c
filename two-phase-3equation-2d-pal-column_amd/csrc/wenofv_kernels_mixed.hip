// wenofv_kernels_mixed.hip -- mixed-precision stage kernels of the 5-point schemes (FP32 WENO
// and flux arithmetic, FP64 equation of state, theta and RK), compiled with fused
// multiply-add contraction.
#include "stage_launch.h"

// minimum wavefronts per SIMD the register allocation must allow
#ifndef WFV_MIXED_MINW
#define WFV_MIXED_MINW 2
#endif
// the characteristic form keeps its eigenvector matrices and double-precision WENO in
// registers: at two wavefronts per SIMD it spills 200-320 bytes per lane, and one wavefront
// per SIMD without spills is 5 % faster (1024^2 and the shipped 100^2 case)
#ifndef WFV_MIXED_CHAR_MINW
#define WFV_MIXED_CHAR_MINW 1
#endif

namespace wfv {

// the 7-point forms (SCHEME 2, RECONSTRUCTIONTYPE 4) live in wenofv_kernels_mixed_r3.hip
hipError_t launch_stage_mixed_r3(const StageArgs &A, int recon, int scheme, hipStream_t s);
hipError_t launch_batch_mixed_r3(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s);

hipError_t launch_stage_mixed(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    if (scheme == 2 || recon == 4) return launch_stage_mixed_r3(A, recon, scheme, s);
    switch (recon * 2 + scheme) {
    case 0: return launch_stage_f<PrecMixed, WFV_MIXED_MINW, 0, 0>(A, s);
    case 1: return launch_stage_f<PrecMixed, WFV_MIXED_MINW, 0, 1>(A, s);
    case 2: return launch_stage_f<PrecMixed, WFV_MIXED_CHAR_MINW, 1, 0>(A, s);
    case 3: return launch_stage_f<PrecMixed, WFV_MIXED_CHAR_MINW, 1, 1>(A, s);
    case 4: return launch_stage_f<PrecMixed, WFV_MIXED_MINW, 2, 0>(A, s);
    case 5: return launch_stage_f<PrecMixed, WFV_MIXED_MINW, 2, 1>(A, s);
    case 6: return launch_stage_f<PrecMixed, WFV_MIXED_MINW, 3, 0>(A, s);
    case 7: return launch_stage_f<PrecMixed, WFV_MIXED_MINW, 3, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_mixed(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s) {
    if (scheme == 2 || recon == 4) return launch_batch_mixed_r3(B, recon, scheme, flux, grid, s);
    switch (recon * 2 + scheme) {
    case 0: return launch_batch_f<PrecMixed, WFV_MIXED_MINW, 0, 0>(B, flux, grid, s);
    case 1: return launch_batch_f<PrecMixed, WFV_MIXED_MINW, 0, 1>(B, flux, grid, s);
    case 2: return launch_batch_f<PrecMixed, WFV_MIXED_CHAR_MINW, 1, 0>(B, flux, grid, s);
    case 3: return launch_batch_f<PrecMixed, WFV_MIXED_CHAR_MINW, 1, 1>(B, flux, grid, s);
    case 4: return launch_batch_f<PrecMixed, WFV_MIXED_MINW, 2, 0>(B, flux, grid, s);
    case 5: return launch_batch_f<PrecMixed, WFV_MIXED_MINW, 2, 1>(B, flux, grid, s);
    case 6: return launch_batch_f<PrecMixed, WFV_MIXED_MINW, 3, 0>(B, flux, grid, s);
    case 7: return launch_batch_f<PrecMixed, WFV_MIXED_MINW, 3, 1>(B, flux, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
