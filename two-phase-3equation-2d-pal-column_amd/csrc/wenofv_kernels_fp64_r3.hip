// wenofv_kernels_fp64_r3.hip -- all-FP64 stage kernels of the 7-point forms: WENO-EXP
// (SCHEME 2) with every reconstruction form, and primitive3 (RECONSTRUCTIONTYPE 4, THINC on
// z1, WENO-Z on p, u, v whatever SCHEME says).  Compiled with -ffp-contract=off.
#include "stage_launch.h"

// minimum wavefronts per SIMD: one for WENO-EXP, so that the 7-row window and the 7-point
// stencils of the x-sweep keep their values in the register file instead of scratch
// memory; two for primitive3, which then spills 300-600 bytes per lane but runs 15 % faster
// on a 4096 x 2048 grid with the second wavefront hiding latency (4 % slower at 1024 x 512)
#ifndef WFV_FP64_R3_MINW
#define WFV_FP64_R3_MINW 1
#endif
#ifndef WFV_FP64_P3_MINW
#define WFV_FP64_P3_MINW 2
#endif

namespace wfv {

hipError_t launch_stage_fp64_r3(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    (void)scheme;  // primitive3 ignores SCHEME
    switch (recon) {
    case 0: return launch_stage_f<PrecFP64, WFV_FP64_R3_MINW, 0, 2>(A, s);
    case 1: return launch_stage_f<PrecFP64, WFV_FP64_R3_MINW, 1, 2>(A, s);
    case 2: return launch_stage_f<PrecFP64, WFV_FP64_R3_MINW, 2, 2>(A, s);
    case 3: return launch_stage_f<PrecFP64, WFV_FP64_R3_MINW, 3, 2>(A, s);
    case 4: return launch_stage_f<PrecFP64, WFV_FP64_P3_MINW, 4, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_fp64_r3(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s) {
    (void)scheme;
    switch (recon) {
    case 0: return launch_batch_f<PrecFP64, WFV_FP64_R3_MINW, 0, 2>(B, flux, grid, s);
    case 1: return launch_batch_f<PrecFP64, WFV_FP64_R3_MINW, 1, 2>(B, flux, grid, s);
    case 2: return launch_batch_f<PrecFP64, WFV_FP64_R3_MINW, 2, 2>(B, flux, grid, s);
    case 3: return launch_batch_f<PrecFP64, WFV_FP64_R3_MINW, 3, 2>(B, flux, grid, s);
    case 4: return launch_batch_f<PrecFP64, WFV_FP64_P3_MINW, 4, 1>(B, flux, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
