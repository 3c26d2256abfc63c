// wenofv_kernels_fp64.hip -- all-FP64 stage kernels, compiled with -ffp-contract=off so
// that every product and sum rounds as in the reference build (no fused multiply-add).
#include "stage_impl.h"

namespace wfv {

template <int RECON, int SCHEME>
static hipError_t launch_t(const StageArgs &A, hipStream_t s) {
    constexpr int WAVES = 1;  // one wavefront per workgroup: waves spread evenly over the CUs
    const int nstrip = A.strips_x * A.strips_y;
    const int blocks = (nstrip + WAVES - 1) / WAVES;
    hipLaunchKernelGGL((stage_kernel<PrecFP64, RECON, SCHEME, WAVES, 2>), dim3(blocks), dim3(LANES * WAVES), 0, s,
                       A);
    return hipGetLastError();
}

hipError_t launch_stage_fp64(const StageArgs &A, int recon, int scheme, hipStream_t s) {
    switch (recon * 2 + scheme) {
    case 0: return launch_t<0, 0>(A, s);
    case 1: return launch_t<0, 1>(A, s);
    case 2: return launch_t<1, 0>(A, s);
    case 3: return launch_t<1, 1>(A, s);
    case 4: return launch_t<2, 0>(A, s);
    case 5: return launch_t<2, 1>(A, s);
    case 6: return launch_t<3, 0>(A, s);
    case 7: return launch_t<3, 1>(A, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace wfv
