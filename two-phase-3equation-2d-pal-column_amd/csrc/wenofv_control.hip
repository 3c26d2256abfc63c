// wenofv_control.hip -- CFL reduction and device-side step control of the time loop.
#include <hip/hip_runtime.h>

#include "wenofv_device.h"
#include "wenofv_kernels.h"

namespace wfv {

constexpr int DT_THREADS = 256;

// CFL time step of a state (calculateDeltaT, src/WENOFV.cpp:474-540).  Positive doubles
// order like their bit patterns, so the minimum is an integer atomic; one atomic per block.
__global__ __launch_bounds__(DT_THREADS) void dt_kernel(DtArgs A) {
    __shared__ double smin[DT_THREADS / 64];
    __shared__ unsigned serr[DT_THREADS / 64];
    const size_t nc = (size_t)A.nx * A.ny;
    double dtloc = 1.0e300;
    unsigned err = 0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (size_t)gridDim.x * blockDim.x) {
        const double U[4] = {A.U[c], A.U[nc + c], A.U[2 * nc + c], A.U[3 * nc + c]};
        Eos e;
        eos(A.P, U, e);
        const double u = U[2] / (U[0] + U[1]);
        const double v = U[3] / (U[0] + U[1]);
        const double ex = fabs(u * 1.0 + v * 0.0) + e.c;
        const double ey = fabs(u * 0.0 + v * 1.0) + e.c;
        const double tmp = A.cfl / (ex / A.dX + ey / A.dY);
        if (A.cell_dt) A.cell_dt[c] = tmp;  // the reference's m_totalDeltaT (src/WENOFV.cpp:500)
        if (isnan(tmp)) err |= WFV_FLAG_NAN_DT;
        else dtloc = fmin(dtloc, tmp);
    }
    for (int off = 32; off > 0; off >>= 1) {
        err |= __shfl_xor(err, off, 64);
        dtloc = fmin(dtloc, __shfl_xor(dtloc, off, 64));
    }
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        smin[w] = dtloc;
        serr[w] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < DT_THREADS / 64; ++k) {
            dtloc = fmin(dtloc, smin[k]);
            err |= serr[k];
        }
        if (err && A.flags) atomicOr(A.flags, err);
        if (A.dtmin && dtloc < 1.0e300) atomicMin(A.dtmin, (unsigned long long)__double_as_longlong(dtloc));
    }
}

// the step control (step_control) as a kernel of its own: before the first step of a device
// loop, and before every step where the last stage does not run it (SSP-RK1, the
// allocation-free loop)
__global__ void step_control_kernel(StepCtl *S) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    step_control(S, S->dtmin_bits, S->flags);
}

// step control of every instance of a batch (one thread per instance)
__global__ void batch_step_control_kernel(StepCtl *ctl, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    StepCtl *S = ctl + i;
    if (S->active == 0) return;
    bool go;
    if (S->fixed_steps > 0) go = S->count < S->fixed_steps;
    else go = fabs(S->now - S->t_end) > 1e-10;
    if (!go || (S->flags & WFV_FLAG_NAN_DT)) {
        S->active = 0;
        return;
    }
    double timestep = 1.0;
    const double m = __longlong_as_double((long long)S->dtmin_bits);
    if (S->dtmin_bits != ~0ull && !(m > timestep)) timestep = m;
    if (S->fixed_steps <= 0) timestep = stdMin(timestep, S->t_end - S->now);
    S->dt = timestep;
    if (S->count < S->max_record && S->dts) S->dts[S->count] = timestep;
    S->now += timestep;
    S->count += 1;
    S->dtmin_bits = ~0ull;
}

// CFL minimum of every instance's state (blockIdx.y = instance)
__global__ __launch_bounds__(DT_THREADS) void batch_dt_kernel(const BatchInst *inst) {
    __shared__ double smin[DT_THREADS / 64];
    __shared__ unsigned serr[DT_THREADS / 64];
    const BatchInst &I = inst[blockIdx.y];
    const size_t nc = (size_t)I.nx * I.ny;
    const double *U = I.reg[0];
    double dtloc = 1.0e300;
    unsigned err = 0;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (size_t)gridDim.x * blockDim.x) {
        const double u4[4] = {U[c], U[nc + c], U[2 * nc + c], U[3 * nc + c]};
        Eos e;
        eos(I.P, u4, e);
        const double u = u4[2] / (u4[0] + u4[1]);
        const double v = u4[3] / (u4[0] + u4[1]);
        const double ex = fabs(u * 1.0 + v * 0.0) + e.c;
        const double ey = fabs(u * 0.0 + v * 1.0) + e.c;
        const double tmp = I.cfl / (ex / I.dX + ey / I.dY);
        if (isnan(tmp)) err |= WFV_FLAG_NAN_DT;
        else dtloc = fmin(dtloc, tmp);
    }
    for (int off = 32; off > 0; off >>= 1) {
        err |= __shfl_xor(err, off, 64);
        dtloc = fmin(dtloc, __shfl_xor(dtloc, off, 64));
    }
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        smin[w] = dtloc;
        serr[w] = err;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < DT_THREADS / 64; ++k) {
            dtloc = fmin(dtloc, smin[k]);
            err |= serr[k];
        }
        if (err) atomicOr(&I.ctl->flags, err);
        if (dtloc < 1.0e300) atomicMin(&I.ctl->dtmin_bits, (unsigned long long)__double_as_longlong(dtloc));
    }
}

// SSP-RK1 writes its stage into register 1 and this copies it back into the state.  A stage
// enqueued after the device loop has finished writes nothing, so neither does its copy.
__global__ __launch_bounds__(DT_THREADS) void copy_state_kernel(double *dst, const double *src, size_t n,
                                                                const int *active) {
    if (active && *active == 0) return;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        dst[k] = src[k];
}

// the same for every instance of a batch (blockIdx.y = instance)
__global__ __launch_bounds__(DT_THREADS) void batch_copy_state_kernel(const BatchInst *inst) {
    const BatchInst &I = inst[blockIdx.y];
    if (I.ctl->active == 0) return;
    const size_t n = 4 * (size_t)I.nx * I.ny;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        I.reg[0][k] = I.reg[1][k];
}

hipError_t launch_copy_state(double *dst, const double *src, size_t n, const int *active, hipStream_t s) {
    size_t blocks = (n + DT_THREADS - 1) / DT_THREADS;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(copy_state_kernel, dim3((unsigned)blocks), dim3(DT_THREADS), 0, s, dst, src, n, active);
    return hipGetLastError();
}

hipError_t launch_batch_copy_state(const BatchInst *inst, int n, int max_cells, hipStream_t s) {
    int bx = (int)((4 * (size_t)max_cells + DT_THREADS - 1) / DT_THREADS);
    if (bx > 256) bx = 256;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(batch_copy_state_kernel, dim3(bx, n), dim3(DT_THREADS), 0, s, inst);
    return hipGetLastError();
}

hipError_t launch_batch_step_control(StepCtl *ctl, int n, hipStream_t s) {
    hipLaunchKernelGGL(batch_step_control_kernel, dim3((n + 63) / 64), dim3(64), 0, s, ctl, n);
    return hipGetLastError();
}

hipError_t launch_batch_dt(const BatchInst *inst, int n, int max_cells, hipStream_t s) {
    int bx = (max_cells + DT_THREADS - 1) / DT_THREADS;
    if (bx > 64) bx = 64;
    if (bx < 1) bx = 1;
    hipLaunchKernelGGL(batch_dt_kernel, dim3(bx, n), dim3(DT_THREADS), 0, s, inst);
    return hipGetLastError();
}

hipError_t launch_batch_fp64(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s);
hipError_t launch_batch_mixed(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s);
hipError_t launch_batch_mixedw(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s);

// precision 2 has kernels of its own for the 5-point conservative and primitive forms; its
// other forms reconstruct in double under the mixed policy already
static bool own_mixedw(int recon, int scheme) { return scheme != 2 && (recon == 0 || recon == 2 || recon == 3); }

hipError_t launch_batch_fp64c(const BatchStageArgs &B, int recon, int scheme, int flux, int grid, hipStream_t s);
hipError_t launch_stage_fp64c(const StageArgs &A, int recon, int scheme, hipStream_t s);

hipError_t launch_batch_stage(const BatchStageArgs &B, int precision, int recon, int scheme, int flux, int grid,
                              hipStream_t s) {
    if (precision == 3) return launch_batch_fp64c(B, recon, scheme, flux, grid, s);
    if (precision == 2 && own_mixedw(recon, scheme)) return launch_batch_mixedw(B, recon, scheme, flux, grid, s);
    if (precision >= 1) return launch_batch_mixed(B, recon, scheme, flux, grid, s);
    return launch_batch_fp64(B, recon, scheme, flux, grid, s);
}

hipError_t launch_stage_fp64(const StageArgs &A, int recon, int scheme, hipStream_t s);
hipError_t launch_stage_mixed(const StageArgs &A, int recon, int scheme, hipStream_t s);
hipError_t launch_stage_mixedw(const StageArgs &A, int recon, int scheme, hipStream_t s);

hipError_t launch_stage(const StageArgs &A, int precision, int recon, int scheme, hipStream_t s) {
    if (precision == 3) return launch_stage_fp64c(A, recon, scheme, s);
    if (precision == 2 && own_mixedw(recon, scheme)) return launch_stage_mixedw(A, recon, scheme, s);
    if (precision >= 1) return launch_stage_mixed(A, recon, scheme, s);
    return launch_stage_fp64(A, recon, scheme, s);
}

hipError_t launch_dt(const DtArgs &A, hipStream_t s) {
    const size_t nc = (size_t)A.nx * A.ny;
    int blocks = (int)((nc + DT_THREADS - 1) / DT_THREADS);
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(dt_kernel, dim3(blocks), dim3(DT_THREADS), 0, s, A);
    return hipGetLastError();
}

hipError_t launch_step_control(StepCtl *S, hipStream_t s) {
    hipLaunchKernelGGL(step_control_kernel, dim3(1), dim3(64), 0, s, S);
    return hipGetLastError();
}

int strip_width(int scheme, int recon) {
    return (scheme == 2 || recon == 4) ? Stencil<2, 0>::SW : Stencil<0, 0>::SW;
}
int strip_halo(int scheme, int recon) {
    return (scheme == 2 || recon == 4) ? Stencil<2, 0>::HALO : Stencil<0, 0>::HALO;
}

}  // namespace wfv
