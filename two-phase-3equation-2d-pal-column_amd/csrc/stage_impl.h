// stage_impl.h -- fused WENOFV RK stage for gfx950 (template; instantiated per precision
// policy by wenofv_kernels_fp64.hip and wenofv_kernels_mixed.hip).
//
// One launch performs a whole SSP-RK stage of the reference's assembleRHS + RK update
// (src/WENOFV.cpp:436-471, 568-615): ghost fill, y- and x-direction WENO reconstruction
// at the faces' Gauss-Lobatto nodes, the bound-preserving limiter, LLF/HLLC fluxes,
// face quadrature, divergence and the RK combination; the last stage also reduces the
// next CFL time step.
//
// Work decomposition.  A wavefront (64 lanes) owns a strip of 58 columns and marches up
// `ty` rows.  Lane l holds column x0-3+l.  For each row r it
//   1. keeps cell averages of rows r-2..r+2 of its column in registers (the periodic or
//      wall ghost map is applied by the loader, so no ghost cells exist in memory),
//   2. reconstructs its column in y at the 4 Gauss-Lobatto nodes (one stencil per column
//      and row, shared by the 5 cells whose x-stencil uses it) and publishes the 16
//      values in LDS,
//   3. for lanes 2..61, reconstructs its cell in x from the neighbours' y-values,
//      converts back (characteristic/primitive forms) and limits it (theta),
//   4. solves the Riemann problem on the face below its cell with the previous row's
//      top face values kept in registers, and on the face east of its cell with the
//      neighbour lane's west face values exchanged through LDS,
//   5. completes the RHS of the cell of row r-1, applies the RK combination and stores.
// Lanes 3..60 own output cells; lanes 2 and 61 recompute the neighbouring strips'
// edge cells, whose face values the strip's boundary fluxes need.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "wenofv_device.h"
#include "wenofv_kernels.h"

namespace wfv {

constexpr int LANES = 64;
constexpr int STRIP_W = 58;  // owned columns per wavefront

WFV_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ghost-cell map: periodic wrap (src/WENOFV.cpp:2175-2295) or reflective wall mirror with
// the normal momentum negated; the result is clamped so that every load stays in bounds.
WFV_DEV int map_index(int i, int n, int bc, bool &mirrored) {
    mirrored = false;
    if (bc == 0) {
        i %= n;
        if (i < 0) i += n;
        return i;
    }
    if (i < 0) { i = -1 - i; mirrored = true; }
    if (i >= n) { i = 2 * n - 1 - i; mirrored = true; }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

template <class T>
WFV_DEV void load_cell(const double *__restrict__ U, size_t nc, size_t off, bool mu, bool mv, T out[4]) {
    double a = U[off], b = U[nc + off], c = U[2 * nc + off], d = U[3 * nc + off];
    out[0] = T(a); out[1] = T(b); out[2] = T(mu ? -c : c); out[3] = T(mv ? -d : d);
}

// primitive variables (src/WENOFV.cpp:937-956 PRIMITIVE, 1097-1117 PRIMITIVE2)
template <class W, int RECON>
WFV_DEV void to_prim(const Phys &P, const double U[4], W out[4]) {
    Eos e;
    eos(P, U, e);
    const double z1 = e.z1;
    out[0] = W(z1);
    if constexpr (RECON == 2) {
        out[1] = W(U[0] / z1);
    } else {
        double pr;
        if (z1 > 0.5) pr = P.preb + P.c1_2 * (U[0] / z1 - P.rho10);
        else pr = P.preb + P.c2_2 * (U[1] / (1.0 - z1) - P.rho20);
        out[1] = W(pr);
    }
    out[2] = W(U[2] / (U[0] + U[1]));
    out[3] = W(U[3] / (U[0] + U[1]));
}

// conservative variables at a node (src/WENOFV.cpp:1007-1032, 1171-1189)
template <class W, class F, int RECON>
WFV_DEV void from_prim(const Phys &P, const W w[4], F out[4]) {
    const double z1 = double(w[0]);
    double z1rho1, z2rho2;
    if constexpr (RECON == 2) {
        const double rho1 = double(w[1]);
        z2rho2 = P.c1_2 / P.c2_2 * (rho1 - P.rho10) + P.rho20;
        z2rho2 = z2rho2 * (1.0 - z1);
        z1rho1 = z1 * rho1;
    } else {
        const double pr = double(w[1]);
        z1rho1 = ((pr - P.preb) / P.c1_2 + P.rho10) * z1;
        z2rho2 = ((pr - P.preb) / P.c2_2 + P.rho20) * (1.0 - z1);
    }
    const double rho = z1rho1 + z2rho2;
    out[0] = F(z1rho1); out[1] = F(z2rho2); out[2] = F(rho * double(w[2])); out[3] = F(rho * double(w[3]));
}

template <class W, int SCHEME>
WFV_DEV void recon4(const W s[5], W out[4]) {
    Weno5<W, SCHEME> st(s[0], s[1], s[2], s[3], s[4]);
    out[0] = st.template node<0>();
    out[1] = st.template node<1>();
    out[2] = st.template node<2>();
    out[3] = st.template node<3>();
}

// LDS per wavefront.  The y-sweep values are dead once the x-sweep has read them, so the
// west face values of the same row reuse their space.
template <class PR>
struct WaveLds {
    union {
        typename PR::W y[16][LANES];     // y-sweep values [g*4+v][lane]
        typename PR::F west[16][LANES];  // limited west face values [v*4+g][lane]
    };
    double fx[4][LANES];  // integrated flux through the east face [v][lane]
};

template <class PR, int RECON, int SCHEME, int WAVES, int MINW>
__global__ __launch_bounds__(LANES *WAVES) __attribute__((amdgpu_waves_per_eu(MINW))) void stage_kernel(
    StageArgs A) {
    using W = typename PR::W;
    using F = typename PR::F;
    __shared__ WaveLds<PR> lds_all[WAVES];
    const int lane = threadIdx.x & (LANES - 1);
    const int wave = threadIdx.x / LANES;
    WaveLds<PR> &L = lds_all[wave];

    const int strip = blockIdx.x * WAVES + wave;
    if (strip >= A.strips_x * A.strips_y) return;  // whole wavefront leaves; no block barrier below
    if (A.active && *A.active == 0) return;         // device-side time loop already finished
    const int sx = strip % A.strips_x, sy = strip / A.strips_x;
    const int nx = A.nx, ny = A.ny;
    const size_t nc = (size_t)nx * ny;
    const int x0 = sx * STRIP_W;
    const int y0 = sy * A.ty;
    const int y1 = min(y0 + A.ty, ny);
    const int nrow = y1 - y0;
    const int xc = x0 - 3 + lane;
    bool mu;
    const int cx = map_index(xc, nx, A.bc, mu);
    const bool owned = lane >= 3 && lane < 3 + STRIP_W && xc < nx;
    const bool recon_lane = lane >= 2 && lane <= 61;
    const bool xface_lane = lane >= 2 && lane <= 60;
    const Phys P = A.P;
    const int flux = A.flux;
    const bool bp = A.bp != 0;

    // RK coefficients; c*dt formed as the reference does (c * deltaT, src/WENOFV.cpp:471)
    const double dt = A.dt_ptr ? *A.dt_ptr : A.dt;
    const double cdt = A.c * dt;
    const double dxc = owned ? A.dxC[cx] : 1.0;
    const double dxw = owned ? A.dxW[cx] : 1.0;
    double qwx[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) qwx[g] = GLW[g] * dxc / 2.0;
    constexpr double omega1 = 0.1;
    constexpr double omegaR = 1.0 - 2 * omega1;

    // register window: conservative averages of rows r-2..r+2 of this lane's column
    double win[5][4];
    W pwin[5][4];  // primitive variables of the same rows (RECON 2, 3)
    auto load_row = [&](int row, double out[4]) {
        bool mv;
        const int cy = map_index(row, ny, A.bc, mv);
        load_cell<double>(A.U, nc, (size_t)cy * nx + cx, mu, mv, out);
    };
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        load_row(y0 - 3 + k, win[k]);
        if constexpr (RECON >= 2) to_prim<W, RECON>(P, win[k], pwin[k]);
    }

    F top_prev[4][4];      // [v][g] top face values of row r-1
    double rhs_part[4];    // RHS of row r-1 without its north face
    unsigned err = 0;
    double dtloc = 1.0e300;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        rhs_part[v] = 0.0;
#pragma unroll
        for (int g = 0; g < 4; ++g) top_prev[v][g] = F(0);
    }

    for (int it = 0; it <= nrow + 1; ++it) {
        const int r = y0 - 1 + it;
        if (it > 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    win[k][v] = win[k + 1][v];
                    if constexpr (RECON >= 2) pwin[k][v] = pwin[k + 1][v];
                }
            }
            load_row(r + 2, win[4]);
            if constexpr (RECON >= 2) to_prim<W, RECON>(P, win[4], pwin[4]);
        }

        // ---- 1. y-direction reconstruction of this lane's column at row r
        //      (getFlux_* step 3, src/WENOFV.cpp:670-687; characteristic 796-833)
        {
            W yv[4][4];  // [g][v]
            if constexpr (RECON == 1) {
                W Lm[16], Rm[16], pc[5][4], yc[4][4];
                eigen_LR<W, false>(P, win[2], Lm, Rm);
#pragma unroll
                for (int t = 0; t < 5; ++t) {
                    W u4[4] = {W(win[t][0]), W(win[t][1]), W(win[t][2]), W(win[t][3])};
                    matvec(Lm, u4, pc[t]);
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    W s[5] = {pc[0][k], pc[1][k], pc[2][k], pc[3][k], pc[4][k]};
                    W o[4];
                    recon4<W, SCHEME>(s, o);
#pragma unroll
                    for (int g = 0; g < 4; ++g) yc[g][k] = o[g];
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) matvec(Rm, yc[g], yv[g]);
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    W s[5];
#pragma unroll
                    for (int t = 0; t < 5; ++t) s[t] = (RECON == 0) ? W(win[t][v]) : pwin[t][v];
                    W o[4];
                    recon4<W, SCHEME>(s, o);
#pragma unroll
                    for (int g = 0; g < 4; ++g) yv[g][v] = o[g];
                }
            }
#pragma unroll
            for (int g = 0; g < 4; ++g)
#pragma unroll
                for (int v = 0; v < 4; ++v) L.y[g * 4 + v][lane] = yv[g][v];
        }
        wave_sync();

        // ---- 2. x-direction reconstruction, back-conversion and limiting of cell (xc, r)
        //      (src/WENOFV.cpp:690-722, 836-880; limiter 1895-2000)
        F fw[4][4], fe[4][4], fs[4][4], fn[4][4];  // west/east/south/north [v][g]
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int g = 0; g < 4; ++g) fw[v][g] = fe[v][g] = fs[v][g] = fn[v][g] = F(0);
        double theta = 1.0;
        if (recon_lane) {
            // Ugp[gx][gy][v] at the 12 boundary nodes, in W before conversion
            W ug[4][4][4];
            if constexpr (RECON == 1) {
                W Lm[16], Rm[16];
                eigen_LR<W, true>(P, win[2], Lm, Rm);
#pragma unroll
                for (int gy = 0; gy < 4; ++gy) {
                    W pc[5][4];
#pragma unroll
                    for (int t = 0; t < 5; ++t) {
                        W y4[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) y4[v] = L.y[gy * 4 + v][lane - 2 + t];
                        matvec(Lm, y4, pc[t]);
                    }
                    W uc[4][4];  // [gx][k]
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        Weno5<W, SCHEME> st(pc[0][k], pc[1][k], pc[2][k], pc[3][k], pc[4][k]);
                        uc[0][k] = st.template node<0>();
                        uc[3][k] = st.template node<3>();
                        if (gy == 0 || gy == 3) {
                            uc[1][k] = st.template node<1>();
                            uc[2][k] = st.template node<2>();
                        }
                    }
                    matvec(Rm, uc[0], ug[0][gy]);
                    matvec(Rm, uc[3], ug[3][gy]);
                    if (gy == 0 || gy == 3) {
                        matvec(Rm, uc[1], ug[1][gy]);
                        matvec(Rm, uc[2], ug[2][gy]);
                    }
                }
            } else {
#pragma unroll
                for (int gy = 0; gy < 4; ++gy) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int row = gy * 4 + v;
                        Weno5<W, SCHEME> st(L.y[row][lane - 2], L.y[row][lane - 1], L.y[row][lane],
                                            L.y[row][lane + 1], L.y[row][lane + 2]);
                        ug[0][gy][v] = st.template node<0>();
                        ug[3][gy][v] = st.template node<3>();
                        if (gy == 0 || gy == 3) {
                            ug[1][gy][v] = st.template node<1>();
                            ug[2][gy][v] = st.template node<2>();
                        }
                    }
                }
            }
            // faces: left = Ugp[0][g], right = Ugp[3][g], bottom = Ugp[g][0], top = Ugp[g][3]
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                F cw[4], ce[4], cs[4], cn[4];
                if constexpr (RECON >= 2) {
                    from_prim<W, F, RECON>(P, ug[0][g], cw);
                    from_prim<W, F, RECON>(P, ug[3][g], ce);
                    from_prim<W, F, RECON>(P, ug[g][0], cs);
                    from_prim<W, F, RECON>(P, ug[g][3], cn);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        cw[v] = F(ug[0][g][v]); ce[v] = F(ug[3][g][v]);
                        cs[v] = F(ug[g][0][v]); cn[v] = F(ug[g][3][v]);
                    }
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    fw[v][g] = cw[v]; fe[v][g] = ce[v]; fs[v][g] = cs[v]; fn[v][g] = cn[v];
                }
            }
            if (bp) {
                // getPositivityPreservingTheta (src/WENOFV.cpp:1936-2000), double in both policies
                const double *Uc = win[2];
                Eos e;
                eos_p<!PR::exact>(P, Uc, e);
                const double u = Uc[2] / (Uc[0] + Uc[1]);
                const double v = Uc[3] / (Uc[0] + Uc[1]);
                const double a1 = fabs(u * 1.0 + v * 0.0) + e.c;
                const double a2 = fabs(u * 0.0 + v * 1.0) + e.c;
                const double mu1 = (a1 / A.dX) / (a1 / A.dX + a2 / A.dY);
                const double mu2 = (a2 / A.dY) / (a1 / A.dX + a2 / A.dY);
                double th[2];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double uave = Uc[k];
                    const double eps = stdMin(1e-12, uave);
                    double rem = uave;
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const double cg = omega1 * GLW[g] / 2;
                        rem -= cg * mu2 * double(fs[k][g]);
                        rem -= cg * mu2 * double(fn[k][g]);
                        rem -= cg * mu1 * double(fw[k][g]);
                        rem -= cg * mu1 * double(fe[k][g]);
                    }
                    rem /= omegaR;
                    double mloc = rem;
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        mloc = scMin(mloc, double(fs[k][g]));
                        mloc = scMin(mloc, double(fn[k][g]));
                        mloc = scMin(mloc, double(fw[k][g]));
                        mloc = scMin(mloc, double(fe[k][g]));
                    }
                    th[k] = scMin(fabs((uave - eps) / (uave - mloc)), 1.0);
                }
                theta = scMin(th[0], th[1]);
                if (theta >= 0.0 && theta <= 1.0) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double ub = Uc[v];
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            fe[v][g] = F(theta * (double(fe[v][g]) - ub) + ub);
                            fw[v][g] = F(theta * (double(fw[v][g]) - ub) + ub);
                            fn[v][g] = F(theta * (double(fn[v][g]) - ub) + ub);
                            fs[v][g] = F(theta * (double(fs[v][g]) - ub) + ub);
                        }
                    }
                } else if (owned && r >= y0 && r < y1) {
                    err |= WFV_FLAG_BAD_THETA;
                }
            }
            if (A.faces_out && owned && r >= y0 && r < y1) {
                double *fo = A.faces_out + ((size_t)r * nx + xc) * 64;
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        fo[0 * 16 + v * 4 + g] = double(fw[v][g]);
                        fo[1 * 16 + v * 4 + g] = double(fe[v][g]);
                        fo[2 * 16 + v * 4 + g] = double(fs[v][g]);
                        fo[3 * 16 + v * 4 + g] = double(fn[v][g]);
                    }
                if (A.theta_out) A.theta_out[(size_t)r * nx + xc] = theta;
            }
        }

        if (it > 0) {
            const bool interior_row = it <= nrow;
            // ---- 3. face below row r: Riemann problem (top of r-1 | bottom of r) and
            //      Gauss-Lobatto quadrature (src/WENOFV.cpp:1728-1746, 1779-1799)
            double G[4] = {0.0, 0.0, 0.0, 0.0};
            if (owned) {
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    F ul[4], ur[4], fl[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) { ul[v] = top_prev[v][g]; ur[v] = fs[v][g]; }
                    riemann<F, 0, 1>(P, flux, ul, ur, fl);
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        if (isnan(double(fl[v]))) err |= WFV_FLAG_NAN_FLUX;
                        G[v] += double(fl[v]) * qwx[g];
                    }
                }
            }
            double rhs_x[4] = {0.0, 0.0, 0.0, 0.0};
            double dyr = 1.0;
            if (interior_row) {
                // ---- 4. face east of cell (xc, r): right face | neighbour lane's left face
                //      (src/WENOFV.cpp:1707-1725, 1764-1778)
                wave_sync();  // every lane has read the y-sweep values the west faces overwrite
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int g = 0; g < 4; ++g) L.west[v * 4 + g][lane] = fw[v][g];
                wave_sync();
                dyr = A.dyC[r];
                if (xface_lane) {
                    double Fx[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const double qw = GLW[g] * dyr / 2.0;
                        F ul[4], ur[4], fl[4];
#pragma unroll
                        for (int v = 0; v < 4; ++v) { ul[v] = fe[v][g]; ur[v] = L.west[v * 4 + g][lane + 1]; }
                        riemann<F, 1, 0>(P, flux, ul, ur, fl);
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            if (owned && isnan(double(fl[v]))) err |= WFV_FLAG_NAN_FLUX;
                            Fx[v] += double(fl[v]) * qw;
                        }
                    }
#pragma unroll
                    for (int v = 0; v < 4; ++v) L.fx[v][lane] = Fx[v];
                }
                wave_sync();
                if (owned) {
                    // RHS order of assembleBoundaryFaceTerm: +west, -east (src/WENOFV.cpp:1802-1813)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        double acc = 0.0;
                        acc += L.fx[v][lane - 1] / dxw / dyr;
                        acc += -L.fx[v][lane] / dxc / dyr;
                        rhs_x[v] = acc;
                    }
                }
            }
            // ---- 5. complete row r-1: -north face, source, RK combination (src/WENOFV.cpp:1816-1827,
            //      1831-1872, 466-471)
            if (it >= 2 && owned) {
                const int rr = r - 1;
                const double dyp = A.dyC[rr];
                const double *Ucur = win[1];
                double Ub[4];
                if (A.Un) {
                    load_cell<double>(A.Un, nc, (size_t)rr * nx + xc, false, false, Ub);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) Ub[v] = Ucur[v];
                }
                double Unew[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double rhs = rhs_part[v];
                    rhs -= G[v] / dxc / dyp;
                    if (v == 3 && A.gravity != 0.0) rhs -= (Ucur[0] + Ucur[1]) * A.gravity;
                    Unew[v] = A.a * Ub[v] + A.b * Ucur[v] + cdt * rhs;
                    A.Uo[v * nc + (size_t)rr * nx + xc] = Unew[v];
                }
                if (A.dtmin) {
                    // calculateDeltaT on the new state (src/WENOFV.cpp:489-513)
                    Eos e;
                    eos_p<!PR::exact>(P, Unew, e);
                    const double u = Unew[2] / (Unew[0] + Unew[1]);
                    const double v = Unew[3] / (Unew[0] + Unew[1]);
                    const double ex = fabs(u * 1.0 + v * 0.0) + e.c;
                    const double ey = fabs(u * 0.0 + v * 1.0) + e.c;
                    const double tmp = A.cfl / (ex / A.dX + ey / A.dY);
                    if (isnan(tmp)) err |= WFV_FLAG_NAN_DT;
                    else dtloc = fmin(dtloc, tmp);
                }
            }
            if (interior_row && owned) {
#pragma unroll
                for (int v = 0; v < 4; ++v) rhs_part[v] = rhs_x[v] + G[v] / dxc / dyr;
            }
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int g = 0; g < 4; ++g) top_prev[v][g] = fn[v][g];
    }

    // wavefront reductions: error bits and the CFL minimum
    for (int off = 32; off > 0; off >>= 1) {
        err |= __shfl_xor(err, off, LANES);
        dtloc = fmin(dtloc, __shfl_xor(dtloc, off, LANES));
    }
    if (lane == 0) {
        if (err && A.flags) atomicOr(A.flags, err);
        if (A.dtmin && dtloc < 1.0e300) atomicMin(A.dtmin, (unsigned long long)__double_as_longlong(dtloc));
    }
}

}  // namespace wfv
