// wenofv_kernels_fp64c.hip -- precision 3: the all-FP64 stage kernels of the 5-point schemes
// compiled with fused multiply-add contraction.  Every operation is the all-FP64 policy's
// (IEEE division and square root, the reference's association), but products feeding sums
// fuse into one rounding, so the results are FP64-accurate without being bit-identical to
// the reference: a solver that differs from the bit-faithful one by rounding alone, at every
// step.  PrecFP64C is a type of its own so that its kernels are distinct instantiations.
// a rounding control: the general kernel variant only
#define WFV_SPECIALISE 0
#include "stage_launch.h"

#ifndef WFV_FP64_MINW
#define WFV_FP64_MINW 2
#endif

namespace wfv {
namespace {
struct PrecFP64C : PrecFP64 {};
}  // namespace

hipError_t launch_stage_fp64c(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    switch (recon * 2 + scheme) {
    case 0: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 0, 0>(A, s);
    case 1: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 0, 1>(A, s);
    case 2: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 1, 0>(A, s);
    case 3: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 1, 1>(A, s);
    case 4: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 2, 0>(A, s);
    case 5: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 2, 1>(A, s);
    case 6: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 3, 0>(A, s);
    case 7: return launch_stage_f<PrecFP64C, WFV_FP64_MINW, 3, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_batch_fp64c(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s) {
    switch (recon * 2 + scheme) {
    case 0: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 0, 0>(B, flux, grid, s);
    case 1: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 0, 1>(B, flux, grid, s);
    case 2: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 1, 0>(B, flux, grid, s);
    case 3: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 1, 1>(B, flux, grid, s);
    case 4: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 2, 0>(B, flux, grid, s);
    case 5: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 2, 1>(B, flux, grid, s);
    case 6: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 3, 0>(B, flux, grid, s);
    case 7: return launch_batch_f<PrecFP64C, WFV_FP64_MINW, 3, 1>(B, flux, grid, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
