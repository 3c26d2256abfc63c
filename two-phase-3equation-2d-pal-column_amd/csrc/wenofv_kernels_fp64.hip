// wenofv_kernels_fp64.hip -- all-FP64 stage kernels of the 5-point schemes (WENO-JS, WENO-Z),
// compiled with -ffp-contract=off so that every product and sum rounds as in the reference
// build (no fused multiply-add).
#include "stage_launch.h"

// minimum wavefronts per SIMD the register allocation must allow
#ifndef WFV_FP64_MINW
#define WFV_FP64_MINW 2
#endif

namespace wfv {

// the 7-point forms (SCHEME 2, RECONSTRUCTIONTYPE 4) live in wenofv_kernels_fp64_r3.hip
hipError_t launch_stage_fp64_r3(const StageArgs &A, int recon, int scheme, hipStream_t s);
hipError_t launch_batch_fp64_r3(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s);

hipError_t launch_stage_fp64(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    if (scheme == 2 || recon == 4) return launch_stage_fp64_r3(A, recon, scheme, s);
    switch (recon * 2 + scheme) {
    case 0: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 0, 0>(A, s);
    case 1: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 0, 1>(A, s);
    case 2: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 1, 0>(A, s);
    case 3: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 1, 1>(A, s);
    case 4: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 2, 0>(A, s);
    case 5: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 2, 1>(A, s);
    case 6: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 3, 0>(A, s);
    case 7: return launch_stage_f<PrecFP64, WFV_FP64_MINW, 3, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_fp64(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s) {
    if (scheme == 2 || recon == 4) return launch_batch_fp64_r3(B, recon, scheme, flux, grid, s);
    switch (recon * 2 + scheme) {
    case 0: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 0, 0>(B, flux, grid, s);
    case 1: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 0, 1>(B, flux, grid, s);
    case 2: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 1, 0>(B, flux, grid, s);
    case 3: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 1, 1>(B, flux, grid, s);
    case 4: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 2, 0>(B, flux, grid, s);
    case 5: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 2, 1>(B, flux, grid, s);
    case 6: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 3, 0>(B, flux, grid, s);
    case 7: return launch_batch_f<PrecFP64, WFV_FP64_MINW, 3, 1>(B, flux, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
