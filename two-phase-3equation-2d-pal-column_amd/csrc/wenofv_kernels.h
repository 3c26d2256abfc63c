// wenofv_kernels.h -- launch interface of the fused WENOFV stage kernels (library-internal).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "wenofv.h"
#include "wenofv_device.h"

namespace wfv {

constexpr int WAVE_LANES = 64;

// cells of a fixed-ghost ring (boundary 3) per variable, WFV_GHOST layers around nx x ny
__host__ __device__ inline size_t ring_size(int nx, int ny) {
    return 2 * (size_t)WFV_GHOST * (nx + 2 * WFV_GHOST) + 2 * (size_t)WFV_GHOST * ny;
}

// Fixed ghost cells (boundary 3): a ring of WFV_GHOST layers around the nx x ny interior,
// SoA over the 4 variables with ring_size(nx, ny) cells each: the bottom rows (-G..-1) and
// top rows (ny..ny+G-1) over columns -G..nx+G-1, then the left and right columns over rows
// 0..ny-1.  Columns beyond the ring (lanes past the last strip's halo) clamp to its edge.
__host__ __device__ inline size_t ghost_offset(int nx, int ny, int i, int j) {
    constexpr int G = WFV_GHOST;
    const int px = nx + 2 * G;
    i = i < -G ? -G : (i >= nx + G ? nx + G - 1 : i);
    j = j < -G ? -G : (j >= ny + G ? ny + G - 1 : j);
    if (j < 0) return (size_t)(j + G) * px + (i + G);
    if (j >= ny) return (size_t)G * px + (size_t)(j - ny) * px + (i + G);
    const size_t sides = 2 * (size_t)G * px;
    if (i < 0) return sides + (size_t)j * G + (i + G);
    if (i >= nx) return sides + (size_t)ny * G + (size_t)j * G + (i - nx);
    return 0;  // interior: not a ghost cell
}

// Stencil of a scheme: radius R = 2 for the 5-point WENO-JS/Z, 3 for the 7-point WENO-EXP
// (SCHEME 2) and primitive3 (RECONSTRUCTIONTYPE 4).  A strip keeps R+1 halo columns on each
// side (R for the stencil, one for the neighbouring strips' edge cells), so a wavefront owns
// 58 or 56 columns.
template <int SCHEME, int RECON>
struct Stencil {
    static constexpr int R = (SCHEME == 2 || RECON == 4) ? 3 : 2;
    static constexpr int NW = 2 * R + 1;              // stencil width
    static constexpr int HALO = R + 1;
    static constexpr int SW = WAVE_LANES - 2 * HALO;  // owned columns per wavefront
};

// One RK stage: Uo = (a*Un + b*U) + (c*dt)*rhs(U).  Un == nullptr stands for Un = U.
struct StageArgs {
    const double *U;      // stage input, SoA [4][ny][nx]
    const double *Un;     // RK base state or nullptr
    double *Uo;           // stage output (may alias Un, never U)
    const double *dxC;    // [nx] column widths
    const double *dxW;    // [nx] west-neighbour widths as the owning rank sees them
    const double *dyC;    // [ny] row widths
    const double *dt_ptr; // device dt (time loop on the device) or nullptr
    const int *active;    // device loop flag or nullptr
    double dt;            // host dt when dt_ptr == nullptr
    double a, b, c;       // RK coefficients
    Phys P;
    double dX, dY, cfl, gravity;
    double dx4;           // pow(dX, 4): WENO-EXP's smoothness threshold, rounded as the host's libm
    const double *ghost;  // fixed ghost ring (boundary 3), SoA [4][ring_size]
    int nx, ny, bc, flux, bp;
    int ty;               // rows per strip
    int strips_x, strips_y;
    unsigned *flags;                 // error bits (WFV_FLAG_*)
    unsigned long long *dtmin;       // if set: CFL minimum of the output state (bit pattern)
    double *faces_out;               // parity hook: limited face values [cell][side][v][g]
    double *theta_out;               // parity hook: limiter theta per cell
    double *theta_flag;              // if set: theta < 1-1e-8 per cell, overwritten every stage
};

struct DtArgs {
    const double *U;
    Phys P;
    double dX, dY, cfl;
    int nx, ny;
    unsigned *flags;
    unsigned long long *dtmin;
};

// device-resident state of the reference time loop for one solver instance
struct StepCtl {
    unsigned long long dtmin_bits;  // CFL minimum of the current state (~0: none)
    double dt;                      // dt of the step in progress
    double now;                     // physical time
    double t_end;
    int count;                      // steps taken
    int fixed_steps;                // > 0: unclipped fixed-step mode
    int active;                     // 0 once the loop has finished
    unsigned flags;                 // error bits
    double *dts;                    // optional record of every dt (device memory)
    int max_record;
    int max_count;                  // time-driven loop: stop after this many steps (0: no cap)
};

// one instance of a batch (device memory)
struct BatchInst {
    double *reg[3];              // RK registers; reg[0] is the state
    const double *dxC, *dxW, *dyC;
    StepCtl *ctl;                // this instance's time loop
    Phys P;
    double dX, dY, cfl, gravity;
    double dx4;                  // pow(dX, 4) (WENO-EXP)
    const double *ghost;         // fixed ghost ring (boundary 3) or nullptr
    int nx, ny, bc, ty;
    int strips_x, strips_y;
    int item_begin;              // first item of this instance in the flattened list
    int pad;
};

// one stage over every instance of a batch
struct BatchStageArgs {
    const BatchInst *inst;
    const int *item_inst;  // [total_items] instance of each item
    int total_items;
    int in, base, out;     // RK register indices (base < 0: Un = U)
    double a, b, c;
    int bp, want_dtmin;
};

hipError_t launch_stage(const StageArgs &A, int precision, int recon, int scheme, hipStream_t s);
hipError_t launch_batch_stage(const BatchStageArgs &B, int precision, int recon, int scheme, int flux, int grid,
                              hipStream_t s);
hipError_t launch_batch_step_control(StepCtl *ctl, int n, hipStream_t s);
hipError_t launch_batch_dt(const BatchInst *inst, int n, int max_cells, hipStream_t s);
hipError_t launch_dt(const DtArgs &A, hipStream_t s);
hipError_t launch_step_control(StepCtl *S, hipStream_t s);
int strip_width(int scheme, int recon);  // owned columns per wavefront

}  // namespace wfv
