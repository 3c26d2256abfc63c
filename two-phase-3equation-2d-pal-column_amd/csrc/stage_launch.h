// stage_launch.h -- launch helpers of the fused stage kernels, instantiated per precision
// policy and scheme group by the wenofv_kernels_*.hip translation units.
#pragma once
#include "stage_impl.h"

// Kernel variants per form.  A unit that defines WFV_SPECIALISE 0 (the rounding control of
// precision 3) has the general kernel only; the others add the kernel without the loader of
// the fixed ghost ring (boundary 3), and the forms that trim their halo rows (trims_rows) one
// more with the limiter known to be on.  (For the forms that reconstruct every row that
// variant only raised spills: primitive3 at two wavefronts per SIMD 18 -> 63 registers.)
#ifndef WFV_SPECIALISE
#define WFV_SPECIALISE 1
#endif

namespace wfv {

// whether a single-domain form has the limiter-on variant: the forms that trim their halo
// rows, except under precision 2, whose conservative kernel then spills 45 registers against
// none and runs slower (the batch kernel of precision 2 spills 24 and still gains: it keeps it)
template <class PR, int RECON, int SCHEME>
constexpr bool limiter_variant() {
    return trims_rows<RECON, SCHEME>() && !std::is_same<PR, PrecMixedW64>::value;
}

// one wavefront per workgroup: waves spread evenly over the CUs
template <class PR, int MINW, int RECON, int SCHEME, int FLUX, bool GH, bool BPK>
hipError_t launch_stage_v(const StageArgs &A, hipStream_t s) {
    hipLaunchKernelGGL((stage_kernel<PR, RECON, SCHEME, FLUX, MINW, GH, BPK>), dim3(A.sp.total), dim3(LANES), 0, s, A);
    return hipGetLastError();
}

template <class PR, int MINW, int RECON, int SCHEME, int FLUX>
hipError_t launch_stage_t(const StageArgs &A, hipStream_t s) {
    if (!WFV_SPECIALISE || A.bc == 3) return launch_stage_v<PR, MINW, RECON, SCHEME, FLUX, true, false>(A, s);
    if constexpr (WFV_SPECIALISE) {
        if constexpr (limiter_variant<PR, RECON, SCHEME>())
            if (A.bp) return launch_stage_v<PR, MINW, RECON, SCHEME, FLUX, false, true>(A, s);
        return launch_stage_v<PR, MINW, RECON, SCHEME, FLUX, false, false>(A, s);
    }
    return hipErrorInvalidValue;
}

template <class PR, int MINW, int RECON, int SCHEME>
hipError_t launch_stage_f(const StageArgs &A, hipStream_t s) {
    return A.flux == 1 ? launch_stage_t<PR, MINW, RECON, SCHEME, 1>(A, s)
                       : launch_stage_t<PR, MINW, RECON, SCHEME, 0>(A, s);
}

template <class PR, int MINW, int RECON, int SCHEME, int FLUX, bool GH, bool BPK>
hipError_t launch_batch_v(const BatchStageArgs &B, int grid, hipStream_t s) {
    if (grid < B.total_items)  // persistent: the wavefronts stride over the items
        hipLaunchKernelGGL((batch_stage_kernel<PR, RECON, SCHEME, FLUX, MINW, true, GH, BPK>), dim3(grid), dim3(LANES),
                           0, s, B);
    else
        hipLaunchKernelGGL((batch_stage_kernel<PR, RECON, SCHEME, FLUX, MINW, false, GH, BPK>), dim3(B.total_items),
                           dim3(LANES), 0, s, B);
    return hipGetLastError();
}

// a batch takes the ghost-ring loader if any instance has boundary 3; its limiter flag is
// the batch's own
template <class PR, int MINW, int RECON, int SCHEME, int FLUX>
hipError_t launch_batch_t(const BatchStageArgs &B, int grid, hipStream_t s) {
    if (!WFV_SPECIALISE || B.ghost) return launch_batch_v<PR, MINW, RECON, SCHEME, FLUX, true, false>(B, grid, s);
    if constexpr (WFV_SPECIALISE) {
        if constexpr (trims_rows<RECON, SCHEME>())
            if (B.bp) return launch_batch_v<PR, MINW, RECON, SCHEME, FLUX, false, true>(B, grid, s);
        return launch_batch_v<PR, MINW, RECON, SCHEME, FLUX, false, false>(B, grid, s);
    }
    return hipErrorInvalidValue;
}

template <class PR, int MINW, int RECON, int SCHEME>
hipError_t launch_batch_f(const BatchStageArgs &B, int flux, int grid, hipStream_t s) {
    return flux == 1 ? launch_batch_t<PR, MINW, RECON, SCHEME, 1>(B, grid, s)
                     : launch_batch_t<PR, MINW, RECON, SCHEME, 0>(B, grid, s);
}

}  // namespace wfv
