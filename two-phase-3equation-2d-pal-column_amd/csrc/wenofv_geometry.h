// wenofv_geometry.h -- grid geometry and launch-shape defaults.
#pragma once
#include <hip/hip_runtime.h>

namespace wfv {

// Cell widths as CWENOFV::initializeSolver builds them for its `nranks` x-slabs
// (src/WENOFV.cpp:90-123): x0 = xL + rank*(nx/nranks)*dx - 3*dx and a cell's width is
// (x0 + (i+1)*dx) - (x0 + i*dx) in its rank's local index i.  These per-cell widths enter
// the face quadrature and the RHS divisors (src/WENOFV.cpp:1774, 1793, 1809-1824); the
// global dx, dy enter the CFL step and the limiter.  dxW of column g is the width of column
// g-1 as column g's rank computes it (its ghost column when g starts a slab).  The same
// functions serve host and device code (both without contraction), so the widths round
// alike on either.
__host__ __device__ inline void column_geometry(int g, int nx, double xL, double dX, int nranks, double *dxC,
                                                double *dxW, double *xcen) {
    const int ng = 3;
    if (nranks < 1 || nx % nranks) nranks = 1;
    const int nloc = nx / nranks;
    const int rank = g / nloc;
    const double procXLeft = xL + (double)(rank * nloc) * dX;
    const double x0 = procXLeft - dX * ng;
    const int ei = g - rank * nloc + ng;
    const double xl = x0 + ei * dX, xr = x0 + (ei + 1) * dX;
    if (dxC) *dxC = xr - xl;
    if (dxW) *dxW = (x0 + ei * dX) - (x0 + (ei - 1) * dX);
    if (xcen) *xcen = (xl + xr) / 2;
}

__host__ __device__ inline void row_geometry(int j, int ny, double yL, double dY, double *dyC, double *ycen) {
    const int ng = 3;
    const double y0 = (yL + 0 * ny * dY) - dY * ng;
    const int ej = j + ng;
    const double yl = y0 + ej * dY, yr = y0 + (ej + 1) * dY;
    if (dyC) *dyC = yr - yl;
    if (ycen) *ycen = (yl + yr) / 2;
}

inline void grid_widths(int nx, int ny, double xL, double xR, double yL, double yR, int nranks, double *dxC,
                        double *dxW, double *dyC, double *xcen, double *ycen) {
    const double dX = (xR - xL) / nx, dY = (yR - yL) / ny;
    for (int g = 0; g < nx; ++g)
        column_geometry(g, nx, xL, dX, nranks, dxC ? dxC + g : nullptr, dxW ? dxW + g : nullptr,
                        xcen ? xcen + g : nullptr);
    for (int j = 0; j < ny; ++j) row_geometry(j, ny, yL, dY, dyC ? dyC + j : nullptr, ycen ? ycen + j : nullptr);
}

// Rows per wavefront strip: long strips amortise the two recomputed halo rows, short ones
// give more wavefronts; aim for >= 8 wavefronts per CU of a 256-CU device.  Small grids
// cannot fill the device and take the shortest strips: their stage time is one strip's
// latency.
inline int default_rows_per_strip(int nx, int ny) {
    const int sx = (nx + 57) / 58;
    const int cand[] = {64, 48, 32, 24, 16, 12, 8, 4, 2};
    for (int ty : cand) {
        const long waves = (long)sx * ((ny + ty - 1) / ty);
        if (waves >= 2048) return ty;
    }
    return 1;
}

// Rows per strip of every instance of a batch: the longest strips that still give the
// launch at least two rounds of 2048 wavefronts (8 per CU of a 256-CU device).
inline int default_batch_rows_per_strip(const int *nx, const int *ny, int n) {
    const int cand[] = {64, 48, 32, 24, 16, 12, 8, 4, 2};
    for (int ty : cand) {
        long waves = 0;
        for (int i = 0; i < n; ++i) waves += (long)((nx[i] + 57) / 58) * ((ny[i] + ty - 1) / ty);
        if (waves >= 4096) return ty;
    }
    return 1;
}

}  // namespace wfv
