// wenofv_kernels_mixed_r3.hip -- mixed-precision stage kernels of the 7-point forms (WENO-EXP,
// primitive3): FP64 reconstruction, FP32 fluxes.  Compiled with fused multiply-add contraction.
#include "stage_launch.h"

// minimum wavefronts per SIMD: one for WENO-EXP, so that the 7-row window and the 7-point
// stencils of the x-sweep keep their values in the register file instead of scratch memory
// (at two its kernels spill over 500 bytes per lane and run 25 % slower); two for
// primitive3, whose WENO-Z on p, u, v needs fewer registers: it spills under 150 bytes per
// lane there and runs 26 % faster on a 4096 x 2048 grid with the second wavefront hiding
// latency
#ifndef WFV_MIXED_R3_MINW
#define WFV_MIXED_R3_MINW 1
#endif
#ifndef WFV_MIXED_P3_MINW
#define WFV_MIXED_P3_MINW 2
#endif

namespace wfv {

hipError_t launch_stage_mixed_r3(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    (void)scheme;  // primitive3 ignores SCHEME
    switch (recon) {
    case 0: return launch_stage_f<PrecMixed, WFV_MIXED_R3_MINW, 0, 2>(A, s);
    case 1: return launch_stage_f<PrecMixed, WFV_MIXED_R3_MINW, 1, 2>(A, s);
    case 2: return launch_stage_f<PrecMixed, WFV_MIXED_R3_MINW, 2, 2>(A, s);
    case 3: return launch_stage_f<PrecMixed, WFV_MIXED_R3_MINW, 3, 2>(A, s);
    case 4: return launch_stage_f<PrecMixed, WFV_MIXED_P3_MINW, 4, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_mixed_r3(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s) {
    (void)scheme;
    switch (recon) {
    case 0: return launch_batch_f<PrecMixed, WFV_MIXED_R3_MINW, 0, 2>(B, flux, grid, s);
    case 1: return launch_batch_f<PrecMixed, WFV_MIXED_R3_MINW, 1, 2>(B, flux, grid, s);
    case 2: return launch_batch_f<PrecMixed, WFV_MIXED_R3_MINW, 2, 2>(B, flux, grid, s);
    case 3: return launch_batch_f<PrecMixed, WFV_MIXED_R3_MINW, 3, 2>(B, flux, grid, s);
    case 4: return launch_batch_f<PrecMixed, WFV_MIXED_P3_MINW, 4, 1>(B, flux, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
