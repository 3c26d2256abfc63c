// wenofv_kernels_mixedw.hip -- stage kernels of precision 2: the mixed policy with the 5-point
// reconstruction in double (PrecMixedW64) for the conservative and primitive forms.  The
// characteristic and 7-point forms reconstruct in double under the mixed policy already, so
// precision 2 runs their mixed kernels.  Compiled with fused multiply-add contraction.
#include "stage_launch.h"

// minimum wavefronts per SIMD the register allocation must allow
#ifndef WFV_MIXEDW_MINW
#define WFV_MIXEDW_MINW 2
#endif

namespace wfv {

hipError_t launch_stage_mixedw(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    switch (recon * 2 + scheme) {
    case 0: return launch_stage_f<PrecMixedW64, WFV_MIXEDW_MINW, 0, 0>(A, s);
    case 1: return launch_stage_f<PrecMixedW64, WFV_MIXEDW_MINW, 0, 1>(A, s);
    case 4: return launch_stage_f<PrecMixedW64, WFV_MIXEDW_MINW, 2, 0>(A, s);
    case 5: return launch_stage_f<PrecMixedW64, WFV_MIXEDW_MINW, 2, 1>(A, s);
    case 6: return launch_stage_f<PrecMixedW64, WFV_MIXEDW_MINW, 3, 0>(A, s);
    case 7: return launch_stage_f<PrecMixedW64, WFV_MIXEDW_MINW, 3, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_mixedw(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s) {
    switch (recon * 2 + scheme) {
    case 0: return launch_batch_f<PrecMixedW64, WFV_MIXEDW_MINW, 0, 0>(B, flux, grid, s);
    case 1: return launch_batch_f<PrecMixedW64, WFV_MIXEDW_MINW, 0, 1>(B, flux, grid, s);
    case 4: return launch_batch_f<PrecMixedW64, WFV_MIXEDW_MINW, 2, 0>(B, flux, grid, s);
    case 5: return launch_batch_f<PrecMixedW64, WFV_MIXEDW_MINW, 2, 1>(B, flux, grid, s);
    case 6: return launch_batch_f<PrecMixedW64, WFV_MIXEDW_MINW, 3, 0>(B, flux, grid, s);
    case 7: return launch_batch_f<PrecMixedW64, WFV_MIXEDW_MINW, 3, 1>(B, flux, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
