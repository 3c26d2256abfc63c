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

// Wavefronts of one grid.  Full strips own SW = 64 - 2*halo columns each and march `ty` rows
// (nfull x sy of them).  The remaining nx % SW columns form a narrow strip of seg = width +
// 2*halo lanes; when K >= 2 such segments fit in a wavefront, that strip is cut into row
// segments of tyr = ceil(ty / K) rows and K of them share a wavefront, so a grid a little
// wider than a multiple of SW does not leave most lanes of its last strip idle.
struct StripPlan {
    int nfull, sy;  // full strips: nfull columns x sy rows of strips
    int seg, K;     // lanes per remainder segment and segments per wavefront (K = 1, seg = 64: unpacked)
    int tyr;        // rows per remainder segment
    int total;      // wavefronts of the grid
};
__host__ __device__ inline StripPlan strip_plan(int nx, int ny, int ty, int halo) {
    StripPlan p;
    const int sw = WAVE_LANES - 2 * halo;
    p.nfull = nx / sw;
    p.sy = (ny + ty - 1) / ty;
    const int wr = nx - p.nfull * sw;
    p.seg = WAVE_LANES;
    p.K = 1;
    p.tyr = ty;
    int nrem = 0;
    if (wr > 0) {
        const int k = WAVE_LANES / (wr + 2 * halo);
        if (k >= 2) {
            p.seg = wr + 2 * halo;
            p.K = k;
            p.tyr = (ty + k - 1) / k;
            const int nseg = (ny + p.tyr - 1) / p.tyr;
            nrem = (nseg + k - 1) / k;
        } else {
            nrem = p.sy;
        }
    }
    p.total = p.nfull * p.sy + nrem;
    return p;
}

#ifndef WFV_XCD_ORDER
#define WFV_XCD_ORDER 1
#endif
// XCD-aware strip order of the single-domain stage kernels.  The hardware hands workgroup b
// to XCD b mod 8, each XCD with its own L2.  Strips are numbered along row bands, so
// workgroup b takes strip xcd_strip(b, n): each XCD runs one contiguous run of whole row
// bands, and the rows that vertically adjacent strips share stay in that XCD's L2 (with b
// itself as the strip, every XCD would see every eighth strip and fetch the shared rows from
// HBM again).  A bijection of [0, n).
__host__ __device__ inline int xcd_strip(int b, int n) {
    if (!WFV_XCD_ORDER) return b;
    constexpr int NXCD = 8;
    const int q = n / NXCD, r = n % NXCD, x = b % NXCD;
    return x * q + (x < r ? x : r) + b / NXCD;
}

// Rows-per-strip candidates that fit a grid's wavefronts to whole rounds of the device.  A
// grid of a few hundred thousand to a few million cells gives fewer wavefronts than the
// device holds at once, or a last round that is nearly empty, and a launch lasts as long as
// its most loaded SIMD.  For 1 and 2 resident wavefronts per SIMD (4 and 8 per CU) and 1-4
// rounds: the fewest rows per strip whose wavefront count fits that many rounds.  Writes at
// most 8 values (unsorted, possibly repeated) to out and returns their count.
__host__ __device__ inline int occupancy_candidates(int nx, int ny, int halo, int cus, int out[8]) {
    int n = 0;
    const int top = ny < 128 ? ny : 128;
    for (int per_cu = 4; per_cu <= 8; per_cu += 4)
        for (int rounds = 1; rounds <= 4; ++rounds) {
            const long long cap = (long long)rounds * per_cu * cus;
            for (int ty = 1; ty <= top; ++ty)
                if (strip_plan(nx, ny, ty, halo).total <= cap) {
                    out[n++] = ty;
                    break;
                }
        }
    return n;
}

// The segment of lane `lane` in wavefront `strip`: a full strip is one segment of 64 lanes;
// a wavefront of the packed remainder strip holds K segments of sp.seg lanes, each with its
// own rows (segments past the last row get y0 = y1), and the 64 - K*seg lanes after them form
// a tail segment of that width with no rows, too narrow for any stencil.  Lane j of a
// segment holds column x0 - halo + j.  nrow_w is the wavefront's row count (its longest
// segment).
struct Segment {
    int x0, y0, y1, j, segw, nrow_w;
};
__host__ __device__ inline Segment segment_of(const StripPlan &sp, int strip, int lane, int ty, int ny, int halo) {
    const int sw = WAVE_LANES - 2 * halo;
    const int nfs = sp.nfull * sp.sy;
    Segment g;
    if (strip < nfs) {
        g.x0 = (strip % sp.nfull) * sw;
        g.y0 = (strip / sp.nfull) * ty;
        g.y1 = g.y0 + ty < ny ? g.y0 + ty : ny;
        g.j = lane;
        g.segw = WAVE_LANES;
        g.nrow_w = g.y1 - g.y0;
    } else {
        g.x0 = sp.nfull * sw;
        const int k = lane / sp.seg;
        if (k < sp.K) {
            g.segw = sp.seg;
            g.j = lane - k * sp.seg;
            const int s0 = (strip - nfs) * sp.K + k;
            g.y0 = s0 * sp.tyr < ny ? s0 * sp.tyr : ny;
            g.y1 = g.y0 + sp.tyr < ny ? g.y0 + sp.tyr : ny;
        } else {  // tail
            g.segw = WAVE_LANES - sp.K * sp.seg;
            g.j = lane - sp.K * sp.seg;
            g.y0 = g.y1 = ny;
        }
        g.nrow_w = sp.tyr;
    }
    return g;
}

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
    StripPlan sp;         // wavefronts of the grid (strip_plan)
    unsigned *flags;                 // error bits (WFV_FLAG_*)
    unsigned long long *dtmin;       // if set: CFL minimum of the output state (bit pattern)
    struct StepCtl *ctl;             // if set (last stage of a device-loop step): the last
                                     // wavefront to finish runs the next step's control
    double *faces_out;               // parity hook: limited face values [cell][side][v][g]
    double *theta_out;               // parity hook: limiter theta per cell
    int faces_ring;                  // hooks over [ny+2][nx+2]: the cells and their ghost ring
    double *theta_flag;              // if set: theta < 1-1e-8 per cell, overwritten every stage
};

struct DtArgs {
    const double *U;
    Phys P;
    double dX, dY, cfl;
    int nx, ny;
    unsigned *flags;
    unsigned long long *dtmin;  // CFL minimum (bit pattern) or nullptr
    double *cell_dt;            // if set: each cell's CFL time step, [ny][nx]
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
    unsigned done;                  // wavefronts of the step's last stage that have finished
};

// Device-side step control of the reference time loop (src/WENOFV.cpp:318-332):
// dt = min(timestep, T - now) with timestep = min(1.0, CFL minimum); the loop stops when
// |now - T| <= 1e-10 (or after max_count steps, where the caller writes output), or after a
// fixed number of unclipped steps.  It writes the dt the following stages read, advances
// `now`, and re-arms the CFL minimum for the last stage.  dtmin_bits and flags are passed in
// by the caller, which reads them as its memory ordering requires.
WFV_DEV void step_control(StepCtl *S, unsigned long long dtmin_bits, unsigned flags) {
    if (S->active == 0) return;
    bool go;
    if (S->fixed_steps > 0) go = S->count < S->fixed_steps;
    else go = fabs(S->now - S->t_end) > 1e-10 && (S->max_count <= 0 || S->count < S->max_count);
    if (!go || (flags & WFV_FLAG_NAN_DT)) {
        S->active = 0;
        return;
    }
    double timestep = 1.0;
    const double m = __longlong_as_double((long long)dtmin_bits);
    if (dtmin_bits != ~0ull && !(m > timestep)) timestep = m;
    if (S->fixed_steps <= 0) timestep = stdMin(timestep, S->t_end - S->now);
    S->dt = timestep;
    if (S->count < S->max_record && S->dts) S->dts[S->count] = timestep;
    S->now += timestep;
    S->count += 1;
    S->dtmin_bits = ~0ull;
}

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
    StripPlan sp;                // wavefronts of the grid (strip_plan)
    int pad;
};

// one stage over every instance of a batch
struct BatchStageArgs {
    const BatchInst *inst;
    const int *item_inst;  // [total_items] instance of each item (-1: none, the wavefront returns)
    const int *item_strip; // [total_items] its strip within the instance
    int total_items;
    int in, base, out;     // RK register indices (base < 0: Un = U)
    double a, b, c;
    int bp, want_dtmin;
    int ghost;             // some instance has fixed ghost cells (boundary 3)
};

hipError_t launch_stage(const StageArgs &A, int precision, int recon, int scheme, hipStream_t s);
hipError_t launch_batch_stage(const BatchStageArgs &B, int precision, int recon, int scheme, int flux, int grid,
                              hipStream_t s);
hipError_t launch_batch_step_control(StepCtl *ctl, int n, hipStream_t s);
hipError_t launch_batch_dt(const BatchInst *inst, int n, int max_cells, hipStream_t s);
hipError_t launch_dt(const DtArgs &A, hipStream_t s);
hipError_t launch_step_control(StepCtl *S, hipStream_t s);
// dst[0..n) = src[0..n) unless *active == 0 (active may be null: always)
hipError_t launch_copy_state(double *dst, const double *src, size_t n, const int *active, hipStream_t s);
// register 1 -> register 0 of every batch instance whose time loop is active
hipError_t launch_batch_copy_state(const BatchInst *inst, int n, int max_cells, hipStream_t s);
int strip_width(int scheme, int recon);  // owned columns per wavefront
int strip_halo(int scheme, int recon);   // halo columns on each side of a strip

}  // namespace wfv
