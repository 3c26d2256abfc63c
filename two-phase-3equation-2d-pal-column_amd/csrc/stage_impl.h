// stage_impl.h -- fused WENOFV RK stage for gfx950 (template; instantiated per precision
// policy by wenofv_kernels_fp64.hip and wenofv_kernels_mixed.hip).
//
// One launch performs a whole SSP-RK stage of the reference's assembleRHS + RK update
// (src/WENOFV.cpp:436-471, 568-615): ghost fill, y- and x-direction WENO reconstruction
// at the faces' Gauss-Lobatto nodes, the bound-preserving limiter, LLF/HLLC fluxes,
// face quadrature, divergence and the RK combination; the last stage also reduces the
// next CFL time step.
//
// Work decomposition.  A wavefront (64 lanes) owns a strip of 58 columns (56 for the 7-point
// forms, see Stencil) and marches up `ty` rows.  Lane l holds column x0-3+l (x0-4+l).  For
// each row r it
//   1. keeps cell averages of rows r-2..r+2 (r-3..r+3) of its column in registers (the
//      periodic or wall ghost map is applied by the loader, so no ghost cells exist in memory),
//   2. reconstructs its column in y at the 4 Gauss-Lobatto nodes (one stencil per column
//      and row, shared by the 5 cells whose x-stencil uses it) and publishes the 16
//      values in LDS,
//   3. for lanes 2..61, reconstructs its cell in x from the neighbours' y-values,
//      converts back (characteristic/primitive forms) and limits it (theta),
//   4. solves the Riemann problem on the face below its cell with the previous row's
//      top face values kept in registers, and on the face east of its cell with the
//      neighbour lane's west face values exchanged through LDS,
//   5. completes the RHS of the cell of row r-1, applies the RK combination and stores.
// Lanes 3..60 (4..59) own output cells; lanes 2 and 61 (3 and 60) recompute the neighbouring
// strips' edge cells, whose face values the strip's boundary fluxes need.  The columns left
// over after the full strips (nx mod 58) form a narrower strip, which is cut into row
// segments packed several per wavefront (strip_plan, segment_of in wenofv_kernels.h); every
// lane then works relative to its segment.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "wenofv_device.h"
#include "wenofv_kernels.h"

namespace wfv {

constexpr int LANES = WAVE_LANES;

// Row prefetch (stage_body's PF) in a kernel, single-domain or batch: where one wavefront per
// SIMD leaves nothing to hide a load's latency, and in the mixed 5-point conservative and
// characteristic kernels, which have the registers for it (+0.9 % on a batch of 256
// liquid-column instances of 64² to 512² cells)
template <class PR, int RECON, int SCHEME, int MINW>
constexpr bool prefetch_rows() {
    return MINW == 1 || (!PR::exact && SCHEME != 2 && RECON <= 1);
}

// The forms whose halo rows reconstruct only the node row they lend when the limiter is off
// (the 5-point conservative and characteristic forms); the others reconstruct every row.
template <int RECON, int SCHEME>
constexpr bool trims_rows() {
    return Stencil<SCHEME, RECON>::R == 2 && RECON <= 1;
}

WFV_DEV void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ghost-cell map of a boundary kind (include/Equations.hpp:25-32): 0 periodic wrap
// (src/WENOFV.cpp:2175-2295); 1 slip walls (NEUMANN, FREESLIP) and 2 no-slip walls (NoSlip)
// mirror the interior; 3 (NEUMANN2) reads fixed ghost cells (ghost_offset).  The result is
// clamped so that every load stays in bounds.
WFV_DEV int map_index(int i, int n, int bc, bool &mirrored) {
    mirrored = false;
    if (bc == 0) {
        // the indices asked for lie within a stencil's reach of the grid (-4 <= i < n + 4) and
        // n >= 3, so two corrections each way wrap them without an integer division
        i += i < 0 ? n : 0;
        i += i < 0 ? n : 0;
        i -= i >= n ? n : 0;
        i -= i >= n ? n : 0;
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    }
    if (i < 0) { i = -1 - i; mirrored = true; }
    if (i >= n) { i = 2 * n - 1 - i; mirrored = true; }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// A state across mirrors: slip walls negate the momentum normal to each mirror (mu: x, mv:
// y); no-slip walls negate both components once per mirror.
template <class T>
WFV_DEV void load_cell(const double *__restrict__ Up, size_t nc, size_t off, bool mu, bool mv, int bc, T out[4]) {
    const WFV_GLOBAL double *U = (const WFV_GLOBAL double *)Up;
    double a = U[off], b = U[nc + off], c = U[2 * nc + off], d = U[3 * nc + off];
    const bool nu = bc == 2 ? mu != mv : mu;
    const bool nv = bc == 2 ? mu != mv : mv;
    out[0] = T(a); out[1] = T(b); out[2] = T(nu ? -c : c); out[3] = T(nv ? -d : d);
}



// primitive variables (src/WENOFV.cpp:937-956 PRIMITIVE, 1097-1117 PRIMITIVE2).  FAST (mixed
// policy): the volume fraction and pressure of the division-free equation of state (with
// N, D as in eos_fast: z1 = N/(D+N), rho1 = z1rho1 (D+N)/N, and p the phases' common
// pressure), the velocities by one reciprocal of rho.
template <class W, int RECON, bool FAST = false>
WFV_DEV void to_prim(const Phys &P, const double U[4], W out[4]) {
    if constexpr (FAST) {
        const double a = U[0], b = U[1];
        const double t1 = a * P.c1_2, t2 = b * P.c2_2;
        const double dq = P.q - (t2 - t1);
        const double x = dq * dq + 4.0 * t1 * t2;
        const double N = dq + x * rsq_nr(fmax(x, 1.0e-300)), DN = 2.0 * t2 + N;
        out[0] = W(N * rcp_nr(DN));
        if constexpr (RECON == 2) out[1] = W(a * DN * rcp_nr(N));
        else out[1] = W(P.preb + (0.5 * DN - P.c2_2 * P.rho20));
        const double irho = rcp_nr(a + b);
        out[2] = W(U[2] * irho);
        out[3] = W(U[3] * irho);
        return;
    }
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

// conservative variables at a node (src/WENOFV.cpp:1007-1032, 1171-1189); FAST divides by
// the sound speeds' squares through their reciprocals
template <class W, class F, int RECON, bool FAST = false>
WFV_DEV void from_prim(const Phys &P, const W w[4], F out[4]) {
    const double z1 = double(w[0]);
    double z1rho1, z2rho2;
    if constexpr (RECON == 2) {
        const double rho1 = double(w[1]);
        z2rho2 = (FAST ? P.c1_2 * P.ic2_2 : P.c1_2 / P.c2_2) * (rho1 - P.rho10) + P.rho20;
        z2rho2 = z2rho2 * (1.0 - z1);
        z1rho1 = z1 * rho1;
    } else {
        const double pr = double(w[1]);
        if constexpr (FAST) {
            z1rho1 = ((pr - P.preb) * P.ic1_2 + P.rho10) * z1;
            z2rho2 = ((pr - P.preb) * P.ic2_2 + P.rho20) * (1.0 - z1);
        } else {
            z1rho1 = ((pr - P.preb) / P.c1_2 + P.rho10) * z1;
            z2rho2 = ((pr - P.preb) / P.c2_2 + P.rho20) * (1.0 - z1);
        }
    }
    const double rho = z1rho1 + z2rho2;
    out[0] = F(z1rho1); out[1] = F(z2rho2); out[2] = F(rho * double(w[2])); out[3] = F(rho * double(w[3]));
}

// the 4 Gauss-Lobatto node values of a 5-point stencil; nodes outside `nodes` (bit g,
// wave-uniform) are left at 0
template <bool EX, class W, int SCHEME, bool ZC = false>
WFV_DEV void recon4(const W s[5], W out[4], unsigned nodes = 0xFu) {
    WenoSel<EX, W, SCHEME, ZC> st(s[0], s[1], s[2], s[3], s[4]);
    out[0] = out[1] = out[2] = out[3] = W(0);
    if (nodes & 1u) out[0] = st.template node<0>();
    if (nodes & 2u) out[1] = st.template node<1>();
    if (nodes & 4u) out[2] = st.template node<2>();
    if (nodes & 8u) out[3] = st.template node<3>();
}

// a wave-uniform union of per-lane bit sets (the low 4 bits)
WFV_DEV unsigned wave_or4(unsigned bits) {
    unsigned r = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (__ballot((bits >> b) & 1u) != 0) r |= 1u << b;
    return r;
}

// Register rotation for loops kept rolled to bound the code size: slot 0 always holds the
// element of the current iteration, so every register index stays static.
template <class T>
WFV_DEV void rotate_left(T (&a)[4]) {
    const T t = a[0];
    a[0] = a[1]; a[1] = a[2]; a[2] = a[3]; a[3] = t;
}
template <class T>
WFV_DEV void rotate_push(T (&a)[4], T x) {  // after 4 pushes a[k] holds the k-th value
    a[0] = a[1]; a[1] = a[2]; a[2] = a[3]; a[3] = x;
}

// LDS per wavefront.  The y-sweep values are dead once the x-sweep has read them, so the
// west face values of the same row reuse their space.
template <class W>
struct WaveLds {
    union {
        W y[16][LANES];          // y-sweep values [g*4+v][lane] (mixed: deviations)
        struct {
            double d[8][LANES];  // mixed conservative form: partial densities [g*2+v][lane]
            W m[8][LANES];       // and momenta [g*2+v-2][lane], as deviations
        } yh;
        double west[16][LANES];  // limited west face states [v*4+g][lane]
    };
    double base[4][LANES];  // mixed: reconstructed variables of the centre row [v][lane]
    double fx[4][LANES];    // integrated flux through the east face [v][lane]
    double top[16][LANES];  // north face values of the previous row [v*4+g][lane]
};

// Face values of one cell side, [v][g].  All-FP64: the values themselves.  Mixed: the
// partial densities in FP64 (they alone enter the equation of state and the limiter's
// positivity bound, so a limited value of 1e-12 beside a cell average of 0.5 must stay
// exact) and the momenta as FP32 deviations from the cell's average (the base).
template <class PR>
struct FaceSide {
    double d[4][4];
    WFV_DEV double get(int v, int g, double) const { return d[v][g]; }
    WFV_DEV void set(int v, int g, double x, double) { d[v][g] = x; }
    template <class W>
    WFV_DEV void set_recon(int v, int g, W w, double) { d[v][g] = double(w); }
    WFV_DEV void limit(int v, int g, double theta, double ub) { d[v][g] = theta * (d[v][g] - ub) + ub; }
    WFV_DEV void rotate() {
#pragma unroll
        for (int v = 0; v < 4; ++v) rotate_left(d[v]);
    }
};
template <>
struct FaceSide<PrecMixed> {
    double r[2][4];
    float m[2][4];
    WFV_DEV double get(int v, int g, double base) const {
        return v < 2 ? r[v & 1][g] : base + double(m[v & 1][g]);
    }
    WFV_DEV void set(int v, int g, double x, double base) {
        if (v < 2) r[v & 1][g] = x;
        else m[v & 1][g] = float(x - base);
    }
    // a reconstructed deviation w from the base
    template <class W>
    WFV_DEV void set_recon(int v, int g, W w, double base) {
        if (v < 2) r[v & 1][g] = base + double(w);
        else m[v & 1][g] = float(w);
    }
    WFV_DEV void limit(int v, int g, double theta, double ub) {  // theta*(value - ub) + ub
        if (v < 2) r[v & 1][g] = theta * (r[v & 1][g] - ub) + ub;
        else m[v & 1][g] *= float(theta);  // a deviation from ub itself
    }
    WFV_DEV void rotate() {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            rotate_left(r[k]);
            rotate_left(m[k]);
        }
    }
};

// precision 2 keeps the mixed policy's face values (only its reconstruction differs)
template <>
struct FaceSide<PrecMixedW64> : FaceSide<PrecMixed> {};

// Riemann flux through one face, integrated over its 4 Gauss-Lobatto nodes
// (src/WENOFV.cpp:1707-1746 fluxes, 1764-1799 quadrature).
template <class PR, int FLUX, int NX, int NY>
WFV_DEV void face_flux(const Phys &P, FaceSide<PR> &a, const double ba[4], FaceSide<PR> &b, const double bb[4],
                       double width, double G[4], unsigned &err, bool check) {
#pragma unroll
    for (int v = 0; v < 4; ++v) G[v] = 0.0;
    if constexpr (PR::exact) {
        // node in register slot 0 with quadrature weight GLW[g] width / 2; the node values
        // rotate into that slot (unrolled, the rotation costs no instructions)
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            double ul[4], ur[4], fl[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                ul[v] = a.get(v, 0, ba[v]);
                ur[v] = b.get(v, 0, bb[v]);
            }
            riemann<FLUX, NX, NY>(P, ul, ur, fl);
            const double qw = GLW[g] * width / 2.0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (check && isnan(fl[v])) err |= WFV_FLAG_NAN_FLUX;
                G[v] += fl[v] * qw;
            }
            a.rotate();  // node g+1 moves to slot 0
            b.rotate();
        }
    } else {
        // Mixed: unrolled; momenta as float base + float deviation, the face frame's
        // (normal, tangential) momenta mapped back to (x, y); a NaN in any node's flux makes
        // its quadrature sum NaN, so the sums are checked once per face.
        constexpr int N = NX ? 2 : 3, T = NX ? 3 : 2;
        const float ban = float(ba[N]), bat = float(ba[T]), bbn = float(bb[N]), bbt = float(bb[T]);
        const float djn = float(bb[N] - ba[N]), djt = float(bb[T] - ba[T]);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const SideF L = {a.r[0][g], a.r[1][g], ban + a.m[N - 2][g], bat + a.m[T - 2][g]};
            const SideF R = {b.r[0][g], b.r[1][g], bbn + b.m[N - 2][g], bbt + b.m[T - 2][g]};
            double pref;
            const double qw = GLW[g] * width / 2.0;
            if constexpr (FLUX == 1) {
                double fd[2];
                float fm[2];
                hllc_f(P, L, R, fd, fm, pref);
                G[0] += fd[0] * qw;
                G[1] += fd[1] * qw;
                G[N] += (double(fm[0]) + pref) * qw;
                G[T] += double(fm[1]) * qw;
            } else {
                float fl[4];
                const float jump[4] = {float(R.a - L.a), float(R.b - L.b), djn + (b.m[N - 2][g] - a.m[N - 2][g]),
                                       djt + (b.m[T - 2][g] - a.m[T - 2][g])};
                llf_f(P, L, R, jump, fl, pref);
                G[0] += double(fl[0]) * qw;
                G[1] += double(fl[1]) * qw;
                G[N] += (double(fl[2]) + pref) * qw;
                G[T] += double(fl[3]) * qw;
            }
        }
        if (check) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (isnan(G[v])) err |= WFV_FLAG_NAN_FLUX;
        }
    }
}

// reconstruction arithmetic of a policy and form (see stage_body)
template <class PR, int RECON, int SCHEME>
using ReconT = typename std::conditional<RECON == 1 || Stencil<SCHEME, RECON>::R == 3, double, typename PR::W>::type;

// the 4 Gauss-Lobatto node values of variable v from a 7-point stencil: WENO-EXP (SCHEME 2),
// or for primitive3 THINC on z1 and WENO-Z on p, u, v (src/WENOFV.cpp:1287-1329)
template <bool EX, int SCHEME, int RECON>
WFV_DEV void recon7(const double s[7], int v, double dx4, double out[4]) {
    if (RECON == 4 && v != 0) {
        WenoSel<EX, double, 1> st(s[1], s[2], s[3], s[4], s[5]);
        out[0] = st.template node<0>(); out[1] = st.template node<1>();
        out[2] = st.template node<2>(); out[3] = st.template node<3>();
    } else if (RECON == 4) {
        ExpThinc7<!EX> st(s, dx4);
        out[0] = st.template node<0>(); out[1] = st.template node<1>();
        out[2] = st.template node<2>(); out[3] = st.template node<3>();
    } else {
        WenoExp7<!EX> st(s, dx4);
        out[0] = st.template node<0>(); out[1] = st.template node<1>();
        out[2] = st.template node<2>(); out[3] = st.template node<3>();
    }
}
// the same at the 12 boundary nodes' x positions: nodes 0 and 3 always, 1 and 2 if `all`
template <bool EX, int SCHEME, int RECON>
WFV_DEV void recon7_x(const double s[7], int v, double dx4, bool all, double out[4]) {
    if (RECON == 4 && v != 0) {
        WenoSel<EX, double, 1> st(s[1], s[2], s[3], s[4], s[5]);
        out[0] = st.template node<0>(); out[3] = st.template node<3>();
        if (all) { out[1] = st.template node<1>(); out[2] = st.template node<2>(); }
    } else if (RECON == 4) {
        ExpThinc7<!EX> st(s, dx4);
        out[0] = st.template node<0>(); out[3] = st.template node<3>();
        if (all) { out[1] = st.template node<1>(); out[2] = st.template node<2>(); }
    } else {
        WenoExp7<!EX> st(s, dx4);
        out[0] = st.template node<0>(); out[3] = st.template node<3>();
        if (all) { out[1] = st.template node<1>(); out[2] = st.template node<2>(); }
    }
}

// One wavefront's strip of one stage (the whole algorithm; see the file comment).  PF:
// load each row of the register window one row ahead, so that its global-memory latency
// overlaps a row of arithmetic (for kernels that run one wavefront per SIMD, where no
// second wavefront hides it).  GH: the grid may have fixed ghost cells (boundary 3); without
// it the loader carries no ring addressing (fewer instructions and scalar registers).  BPK: the
// limiter is known to be on, so every row reconstructs every node row and the row masks,
// their ballots and the node values of skipped rows compile away.
template <class PR, int RECON, int SCHEME, int FLUX, bool PF = false, bool GH = true, bool BPK = false>
WFV_DEV void stage_body(const StageArgs &A, const int strip, WaveLds<ReconT<PR, RECON, SCHEME>> &L) {
    // Reconstruction arithmetic.  The characteristic form projects with eigenvector
    // matrices whose condition number reaches ~1e8 in near-pure phases (z1 -> 0 or 1), which
    // would amplify FP32 rounding of the characteristic values to O(1); the mixed policy
    // therefore keeps that form's reconstruction in FP64 (its fluxes stay FP32), and so it
    // does for the 7-point forms.  DEV: the mixed 5-point forms reconstruct deviations from
    // the cell's mean; the 7-point forms reconstruct the values themselves (THINC tests the
    // absolute volume fraction).
    using W = ReconT<PR, RECON, SCHEME>;
    using FS = FaceSide<PR>;
    using ST = Stencil<SCHEME, RECON>;
    constexpr int R = ST::R, NW = ST::NW, HALO = ST::HALO, SW = ST::SW;
    constexpr bool EX = PR::exact;
    constexpr bool DEV = !EX && R == 2;
    // HYB (mixed conservative form): the partial densities are reconstructed in double, the
    // momenta in W.  The pressure at a face is c_k^2 (rho_k - rho_k0) of the reconstructed
    // partial densities, so it amplifies their rounding by c_k^2 rho_k / p (about 100 for a
    // liquid at c = 10, p = 1): an FP32-rounded z1rho1 deviation of O(1) at an interface would
    // put a 6e-6 relative error into the pressure of every flux there.  The primitive forms
    // reconstruct the pressure itself and convert back in double, so they keep W throughout.
    constexpr bool HYB = DEV && RECON == 0 && !std::is_same<W, double>::value;
    constexpr bool TRIM = trims_rows<RECON, SCHEME>();
    // EARLY (mixed policy): a row's height and its RK base state are loaded well before their
    // use, so that their latency overlaps arithmetic; the all-FP64 kernels, whose registers
    // are spent, load them where they are used
    constexpr bool EARLY = !EX;
    // Loops over the 4 variables or characteristic fields: unrolled (no register rotation,
    // fewer spills; measured faster in every kernel, the all-FP64 ones included, although
    // several of those then outgrow the 64 KiB instruction cache).
    constexpr int VU = 4;
    const int lane = threadIdx.x & (LANES - 1);

    const StripPlan &sp = A.sp;
    if (strip >= sp.total) return;
    // the device-side time loop's flag; the wavefront returns if the loop has finished, once
    // the loads that every strip starts with are in flight (one memory latency, not three)
    const int act = A.active ? *A.active : 1;
    const int nx = A.nx, ny = A.ny;
    const size_t nc = (size_t)nx * ny;
    // this lane's segment (segment_of); stencils and face exchanges stay inside it
    const Segment sg = segment_of(sp, strip, lane, A.ty, ny, HALO);
    const int x0 = sg.x0, y0 = sg.y0, y1 = sg.y1, j = sg.j, segw = sg.segw, nrow_w = sg.nrow_w;
    const int nrow = y1 - y0;
    const int xc = x0 - HALO + j;
    bool mu;
    const int cx = map_index(xc, nx, A.bc, mu);
    const bool owned = j >= HALO && j < HALO + SW && xc < nx && nrow > 0;
    const bool recon_lane = j >= R && j <= segw - 1 - R;
    const bool xface_lane = j >= R && j <= segw - 2 - R;
    const Phys P = A.P;
    const bool bp = BPK || A.bp != 0;

    // RK coefficients; c*dt formed as the reference does (c * deltaT, src/WENOFV.cpp:471)
    const double dt = A.dt_ptr ? *A.dt_ptr : A.dt;
    const double cdt = A.c * dt;
    const double dxc = owned ? A.dxC[cx] : 1.0;
    const double dxw = owned ? A.dxW[cx] : 1.0;
    constexpr double omega1 = 0.1;
    constexpr double omegaR = 1.0 - 2 * omega1;

    // register window: cell averages of rows r-R..r+R of this lane's column, and their
    // reconstructed (primitive) variables for RECON 2, 3, 4.  A row's cell is a fixed ghost
    // cell (boundary 3) from the ring or a cell of the state through the periodic or wall
    // map: one load per variable either way, with no branch, so that the window's loads are
    // all in flight together.
    // SLIM (primitive forms): the stencils read the primitive window pwin; of the cell
    // averages only rows r-1 (the RK completion, win[0]) and r (faces, limiter, RK, win[1])
    // are kept, and row r is loaded again when it becomes the centre (it entered the window R
    // iterations earlier, so the caches hold it): the other rows of a second window would
    // only occupy registers.
    constexpr bool SLIM = RECON >= 2;
    double win[SLIM ? 2 : NW][4];
    double pwin[NW][4];
    // load_raw: a row's four loads and whether its momenta are to be negated (slip walls
    // negate the momentum normal to each mirror, no-slip walls both components); the signs are
    // applied where the values are first used, so that PF's loads stay in flight through a row
    // of arithmetic instead of being waited for at once
    auto load_raw = [&](int row, double out[4], bool &nu, bool &nv) {
        bool mv;
        const int cy = map_index(row, ny, A.bc, mv);
        const bool g = GH && A.bc == 3 && (xc < 0 || xc >= nx || row < 0 || row >= ny);
        const WFV_GLOBAL double *src = (const WFV_GLOBAL double *)(g ? A.ghost : A.U);
        const size_t stride = g ? ring_size(nx, ny) : nc;
        const size_t off = g ? ghost_offset(nx, ny, xc, row) : (size_t)cy * nx + cx;
        out[0] = src[off]; out[1] = src[stride + off]; out[2] = src[2 * stride + off]; out[3] = src[3 * stride + off];
        nu = !g && (A.bc == 2 ? mu != mv : mu);
        nv = !g && (A.bc == 2 ? mu != mv : mv);
    };
    auto set_signs = [](double out[4], bool nu, bool nv) {
        out[2] = nu ? -out[2] : out[2];
        out[3] = nv ? -out[3] : out[3];
    };
    auto load_row = [&](int row, double out[4]) {
        bool nu, nv;
        load_raw(row, out, nu, nv);
        set_signs(out, nu, nv);
    };
    // The window of row r = y0 - 1 (the first iteration's).  PF: one row lower, in positions
    // 1..NW-1, and every iteration, the first too, shifts the next row in from `ahead`: a loop
    // body without a branch around the shift, so the loads into `ahead` are not waited for at
    // a join.
    constexpr int K0 = PF ? 1 : 0;
    const int wrow = PF ? y0 - 2 - R : y0 - 1 - R;  // row of window position 0
#pragma unroll
    for (int k = K0; k < NW; ++k) {
        if constexpr (SLIM) {
            double t[4];
            load_row(wrow + k, t);
            to_prim<double, RECON, !EX>(P, t, pwin[k]);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (!PF && k == R - 1) win[0][v] = t[v];  // rows r-1 and r (PF: r-2 and r-1,
                if (k == R) win[1][v] = t[v];             // of which the shift keeps r-1)
            }
        } else {
            load_row(wrow + k, win[k]);
        }
    }

    // The north face values of row r-1 wait in LDS (L.top) for the face below row r; their
    // base (mixed: row r-1 cell average) stays in registers.
    double top_base[4];
    double rhs_part[4];    // RHS of row r-1 without its north face
    unsigned err = 0;
    double dtloc = 1.0e300;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        rhs_part[v] = 0.0;
        top_base[v] = 0.0;
    }

    // PF: the row the next shift takes into the window (the same rows as without PF)
    double ahead[4];
    bool ahead_nu = false, ahead_nv = false;  // its signs, applied as it enters the window
    if constexpr (PF) load_raw(y0 - 1 + R, ahead, ahead_nu, ahead_nv);
    // mixed policy: reciprocals of the widths replace the reference's divisions by them
    const double idxc = EX ? 0.0 : 1.0 / dxc;
    const double idxw = EX ? 0.0 : 1.0 / dxw;
    const double idX = EX ? 0.0 : 1.0 / A.dX;
    const double idY = EX ? 0.0 : 1.0 / A.dY;
    asm volatile("" ::: "memory");  // the loads above are issued before the flag is tested
    if (act == 0) return;
    // PF: the first row's signs here, where the window is waited for anyway, so that at the
    // loop head only the loads of the previous iteration are waited for
    if constexpr (PF) {
        set_signs(ahead, ahead_nu, ahead_nv);
        ahead_nu = ahead_nv = false;
    }
    // height of the previous row and (mixed) its reciprocal, for the RK completion of that row
    double dy_prev = 1.0, idy_prev = 0.0;

    for (int it = 0; it <= nrow_w + 1; ++it) {
        const int r = y0 - 1 + it;
        const bool live = it <= nrow + 1;  // this lane's segment still has rows
        // this row's height (for the RK completion), loaded early so that its latency overlaps
        // the sweeps; PF: after the shift, whose signs would otherwise wait for this load too
        auto row_height = [&]() -> double {
            return (EARLY && it >= 1 && it <= nrow) ? ((const WFV_GLOBAL double *)A.dyC)[r] : 1.0;
        };
        double dy_row = 1.0;
        if constexpr (!PF) dy_row = row_height();
        // Node rows (bit gy) the sweeps must reconstruct: all four in an owned row; in the rows
        // below and above the segment, whose cells only lend it their north (gy = 3) or south
        // (gy = 0) face, just that row, unless the limiter needs the cells' other faces too or
        // every face value is recorded.  A union over the wavefront's lanes.
        // (The primitive and 7-point forms reconstruct every row: their register budget is spent.)
        const bool full_l = bp || A.faces_out || (it >= 1 && it <= nrow);
        const unsigned rows =
            (!TRIM || BPK) ? 0xFu
                  : wave_or4(!live ? 0u : (full_l ? 0xFu : ((it == 0 ? 0x8u : 0u) | (it == nrow + 1 ? 0x1u : 0u))));
        if (PF || it > 0) {
            if constexpr (SLIM) {
#pragma unroll
                for (int k = 0; k < NW - 1; ++k)
#pragma unroll
                    for (int v = 0; v < 4; ++v) pwin[k][v] = pwin[k + 1][v];
                double t[4];
                if constexpr (PF) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) t[v] = ahead[v];
                    set_signs(t, ahead_nu, ahead_nv);
                    // without a branch (which would make the loaded registers wait for a
                    // copy at its end): the last iteration loads its own last row again, unused
                    load_raw(it <= nrow_w ? r + R + 1 : r + R, ahead, ahead_nu, ahead_nv);
                } else {
                    load_row(r + R, t);
                }
                to_prim<double, RECON, !EX>(P, t, pwin[NW - 1]);
#pragma unroll
                for (int v = 0; v < 4; ++v) win[0][v] = win[1][v];
                load_row(r, win[1]);  // the new centre row's cell averages, again
            } else {
#pragma unroll
                for (int k = 0; k < NW - 1; ++k)
#pragma unroll
                    for (int v = 0; v < 4; ++v) win[k][v] = win[k + 1][v];
                if constexpr (PF) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) win[NW - 1][v] = ahead[v];
                    set_signs(win[NW - 1], ahead_nu, ahead_nv);
                    // without a branch (which would make the loaded registers wait for a
                    // copy at its end): the last iteration loads its own last row again, unused
                    load_raw(it <= nrow_w ? r + R + 1 : r + R, ahead, ahead_nu, ahead_nv);
                } else {
                    load_row(r + R, win[NW - 1]);
                }
            }
        }
        if constexpr (PF) dy_row = row_height();
        double Ub[4];  // the RK base state of row r-1, completed in this iteration
        const bool complete = it >= 2 && owned && live;
        if constexpr (EARLY && PF)
            if (complete && A.Un) load_cell<double>(A.Un, nc, (size_t)(r - 1) * nx + xc, false, false, 0, Ub);
        const double *Uc;  // cell (xc, r)
        const double *Bc;  // its reconstructed variables
        if constexpr (SLIM) {
            Uc = win[1];
            Bc = pwin[R];
        } else {
            Uc = win[R];
            Bc = win[R];
        }

        // ---- 1. y-direction reconstruction of this lane's column at row r
        //      (getFlux_* step 3, src/WENOFV.cpp:670-687; characteristic 796-833)
        {
            auto sv = [&](int t, int v) -> W {  // stencil value (DEV: deviation from row r)
                double x;
                if constexpr (SLIM) x = pwin[t][v];
                else x = win[t][v];
                if constexpr (!DEV) return x;
                else return W(x - Bc[v]);
            };
            if constexpr (R == 3 && RECON == 1) {
                double Lm[16], Rm[16], pc[7][4], yc[4][4], yv[4];
                eigen_LR<double, false, !EX>(P, Uc, Lm, Rm);
#pragma unroll
                for (int t = 0; t < 7; ++t) {
                    double u4[4] = {sv(t, 0), sv(t, 1), sv(t, 2), sv(t, 3)};
                    matvec(Lm, u4, pc[t]);
                }
#pragma unroll VU
                for (int k = 0; k < 4; ++k) {  // characteristic field k sits in slot 0
                    const double s7[7] = {pc[0][0], pc[1][0], pc[2][0], pc[3][0], pc[4][0], pc[5][0], pc[6][0]};
                    double o[4];
                    recon7<EX, SCHEME, RECON>(s7, k, A.dx4, o);
#pragma unroll
                    for (int g = 0; g < 4; ++g) rotate_push(yc[g], o[g]);
#pragma unroll
                    for (int t = 0; t < 7; ++t) rotate_left(pc[t]);
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    matvec(Rm, yc[g], yv);
#pragma unroll
                    for (int v = 0; v < 4; ++v) L.y[g * 4 + v][lane] = yv[v];
                }
            } else if constexpr (R == 3) {
#pragma unroll VU
                for (int v = 0; v < 4; ++v) {
                    double s7[7];
#pragma unroll
                    for (int t = 0; t < 7; ++t) s7[t] = sv(t, v);
                    double o[4];
                    recon7<EX, SCHEME, RECON>(s7, v, A.dx4, o);
#pragma unroll
                    for (int g = 0; g < 4; ++g) L.y[g * 4 + v][lane] = o[g];
                }
            } else if constexpr (RECON == 1) {
                W Lm[16], Rm[16], pc[5][4], yc[4][4], yv[4];
                eigen_LR<W, false, !EX>(P, Uc, Lm, Rm);
#pragma unroll
                for (int t = 0; t < 5; ++t) {
                    W u4[4] = {sv(t, 0), sv(t, 1), sv(t, 2), sv(t, 3)};
                    matvec(Lm, u4, pc[t]);
                }
#pragma unroll VU
                for (int k = 0; k < 4; ++k) {  // characteristic field k sits in slot 0
                    W s5[5] = {pc[0][0], pc[1][0], pc[2][0], pc[3][0], pc[4][0]};
                    W o[4];
                    recon4<EX, W, SCHEME>(s5, o, rows);
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        rotate_push(yc[g], o[g]);
                    }
#pragma unroll
                    for (int t = 0; t < 5; ++t) rotate_left(pc[t]);
                }
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    matvec(Rm, yc[g], yv);
#pragma unroll
                    for (int v = 0; v < 4; ++v) L.y[g * 4 + v][lane] = yv[v];
                }
            } else if constexpr (HYB) {
                // partial densities in double, momenta in W; deviations from row r (the
                // centre's own is the literal 0)
                double sd[5][2];
                W sm[5][2];
#pragma unroll
                for (int t = 0; t < 5; ++t)
#pragma unroll
                    for (int v = 0; v < 2; ++v) {
                        sd[t][v] = t == 2 ? 0.0 : win[t][v] - Bc[v];
                        sm[t][v] = t == 2 ? W(0) : W(win[t][v + 2] - Bc[v + 2]);
                    }
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const double s5[5] = {sd[0][v], sd[1][v], sd[2][v], sd[3][v], sd[4][v]};
                    double o[4];
                    recon4<EX, double, SCHEME, true>(s5, o, rows);
#pragma unroll
                    for (int g = 0; g < 4; ++g) L.yh.d[g * 2 + v][lane] = o[g];
                }
#pragma unroll
                for (int v = 0; v < 2; ++v) {
                    const W s5[5] = {sm[0][v], sm[1][v], sm[2][v], sm[3][v], sm[4][v]};
                    W o[4];
                    recon4<EX, W, SCHEME, true>(s5, o, rows);
#pragma unroll
                    for (int g = 0; g < 4; ++g) L.yh.m[g * 2 + v][lane] = o[g];
                }
            } else {
                W sw[5][4];
#pragma unroll
                for (int t = 0; t < 5; ++t)
#pragma unroll
                    // DEV: the centre's deviation from itself is 0 (a literal the compiler folds)
                    for (int v = 0; v < 4; ++v) sw[t][v] = (DEV && t == 2) ? W(0) : sv(t, v);
#pragma unroll VU
                for (int v = 0; v < 4; ++v) {  // variable v sits in slot 0
                    W s5[5] = {sw[0][0], sw[1][0], sw[2][0], sw[3][0], sw[4][0]};
                    W o[4];
                    recon4<EX, W, SCHEME, DEV>(s5, o, rows);
#pragma unroll
                    for (int g = 0; g < 4; ++g) L.y[g * 4 + v][lane] = o[g];
#pragma unroll
                    for (int t = 0; t < 5; ++t) rotate_left(sw[t]);
                }
            }
            if constexpr (DEV) {
#pragma unroll
                for (int v = 0; v < 4; ++v) L.base[v][lane] = Bc[v];
            }
        }
        wave_sync();

        // ---- 2. x-direction reconstruction, back-conversion and limiting of cell (xc, r)
        //      (src/WENOFV.cpp:690-722, 836-880; limiter 1895-2000)
        // West/east/south/north face values: set by the lanes that reconstruct (recon_lane) and
        // read only by them (fluxes, limiter, outputs) and, through L.west and L.top, only from
        // them; the stencil-halo lanes leave them unset and never read them.
        // (The forms that reconstruct every row, whose registers are spent, zero them: that
        // measured fewer spills there.)
        FS fw, fe, fs, fn;
        if constexpr (!TRIM) {
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    fw.set(v, g, 0.0, 0.0);
                    fe.set(v, g, 0.0, 0.0);
                    fs.set(v, g, 0.0, 0.0);
                    fn.set(v, g, 0.0, 0.0);
                }
        }
        double theta = 1.0;
        if (recon_lane) {
            // x-stencil value of node row gy, variable v, offset t (DEV: deviation from Bc)
            auto xv = [&](int gy, int v, int t) -> W {
                const int l = lane - R + t;
                if constexpr (!DEV) return L.y[gy * 4 + v][l];
                else return W(L.base[v][l] - Bc[v]) + L.y[gy * 4 + v][l];
            };
            // node values -> face values of the four sides; left = Ugp[0][g], right = Ugp[3][g],
            // bottom = Ugp[g][0], top = Ugp[g][3]
            auto to_face = [&](const W w[4], FS &f, int g) {
                if constexpr (RECON >= 2) {
                    double pv[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) pv[v] = DEV ? Bc[v] + double(w[v]) : double(w[v]);
                    double cv[4];
                    from_prim<double, double, RECON, !EX>(P, pv, cv);
#pragma unroll
                    for (int v = 0; v < 4; ++v) f.set(v, g, cv[v], Uc[v]);
                } else if constexpr (DEV || EX) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) f.set_recon(v, g, w[v], Uc[v]);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v) f.set(v, g, double(w[v]), Uc[v]);
                }
            };
            // the characteristic forms hand each node's values to its faces as soon as they are
            // formed (node (gx, gy) lies on the west face if gx = 0, the east if gx = 3, the
            // south if gy = 0 and the north if gy = 3), so that the 48 node values and the
            // faces built from them are never live together
            auto emit = [&](int gx, int gy, const W o[4]) {
                if (gx == 0) to_face(o, fw, gy);
                if (gx == 3) to_face(o, fe, gy);
                if (gy == 0) to_face(o, fs, gx);
                if (gy == 3) to_face(o, fn, gx);
            };
            W ug[4][4][4];  // Ugp[gx][gy][v] at the 12 boundary nodes
            double ugd[4][4][2];  // HYB: the partial densities' node values [gx][gy][v]
            W ugm[4][4][2];       // and the momenta's [gx][gy][v-2]
            if constexpr (R == 3 && RECON == 1) {
                double Lm[16], Rm[16];
                eigen_LR<double, true, !EX>(P, Uc, Lm, Rm);
#pragma unroll VU
                for (int gy = 0; gy < 4; ++gy) {  // node row gy sits in slot 0 of ug[.][0]
                    double pc[7][4];
#pragma unroll
                    for (int t = 0; t < 7; ++t) {
                        double y4[4] = {xv(gy, 0, t), xv(gy, 1, t), xv(gy, 2, t), xv(gy, 3, t)};
                        matvec(Lm, y4, pc[t]);
                    }
                    const bool all = gy == 0 || gy == 3;
                    double uc[4][4];  // [gx][k]
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const double s7[7] = {pc[0][k], pc[1][k], pc[2][k], pc[3][k], pc[4][k], pc[5][k], pc[6][k]};
                        double o[4] = {0.0, 0.0, 0.0, 0.0};
                        recon7_x<EX, SCHEME, RECON>(s7, k, A.dx4, all, o);
#pragma unroll
                        for (int gx = 0; gx < 4; ++gx) uc[gx][k] = o[gx];
                    }
#pragma unroll
                    for (int gx = 0; gx < 4; ++gx) {
                        if (!all && (gx == 1 || gx == 2)) continue;  // inner nodes of the side faces only
                        double o[4];
                        matvec(Rm, uc[gx], o);
                        emit(gx, gy, o);
                    }
                }
            } else if constexpr (R == 3) {
#pragma unroll VU
                for (int v = 0; v < 4; ++v) {  // variable v sits in slot 3 after the loop
#pragma unroll
                    for (int gy = 0; gy < 4; ++gy) {
                        double s7[7];
#pragma unroll
                        for (int t = 0; t < 7; ++t) s7[t] = xv(gy, v, t);
                        double o[4] = {0.0, 0.0, 0.0, 0.0};
                        recon7_x<EX, SCHEME, RECON>(s7, v, A.dx4, gy == 0 || gy == 3, o);
#pragma unroll
                        for (int gx = 0; gx < 4; ++gx) rotate_push(ug[gx][gy], o[gx]);
                    }
                }
            } else if constexpr (RECON == 1) {
                W Lm[16], Rm[16];
                eigen_LR<W, true, !EX>(P, Uc, Lm, Rm);
#pragma unroll
                for (int gy = 0; gy < 4; ++gy) {
                    const bool all = gy == 0 || gy == 3;
                    if (!(rows & (1u << gy))) {
                        const W z[4] = {W(0), W(0), W(0), W(0)};
#pragma unroll
                        for (int gx = 0; gx < 4; ++gx)
                            if (all || gx == 0 || gx == 3) emit(gx, gy, z);
                        continue;
                    }
                    W pc[5][4];
#pragma unroll
                    for (int t = 0; t < 5; ++t) {
                        W y4[4] = {xv(gy, 0, t), xv(gy, 1, t), xv(gy, 2, t), xv(gy, 3, t)};
                        matvec(Lm, y4, pc[t]);
                    }
                    W uc[4][4];  // [gx][k]
#pragma unroll VU
                    for (int k = 0; k < 4; ++k) {  // characteristic field k sits in slot 0
                        WenoSel<EX, W, SCHEME> st(pc[0][0], pc[1][0], pc[2][0], pc[3][0], pc[4][0]);
                        rotate_push(uc[0], st.template node<0>());
                        rotate_push(uc[3], st.template node<3>());
                        if (all) {
                            rotate_push(uc[1], st.template node<1>());
                            rotate_push(uc[2], st.template node<2>());
                        }
#pragma unroll
                        for (int t = 0; t < 5; ++t) rotate_left(pc[t]);
                    }
#pragma unroll
                    for (int gx = 0; gx < 4; ++gx) {
                        if (!all && (gx == 1 || gx == 2)) continue;
                        W o[4];
                        matvec(Rm, uc[gx], o);
                        emit(gx, gy, o);
                    }
                }
            } else if constexpr (EX) {
                // all-FP64: unrolled (a rolled loop would keep every node value live across it)
#pragma unroll
                for (int gy = 0; gy < 4; ++gy)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
#pragma unroll
                        for (int gx = 0; gx < 4; ++gx) ug[gx][gy][v] = W(0);
                        if (!(rows & (1u << gy))) continue;
                        WenoSel<EX, W, SCHEME> st(xv(gy, v, 0), xv(gy, v, 1), xv(gy, v, 2), xv(gy, v, 3), xv(gy, v, 4));
                        ug[0][gy][v] = st.template node<0>();
                        ug[3][gy][v] = st.template node<3>();
                        if (gy == 0 || gy == 3) {
                            ug[1][gy][v] = st.template node<1>();
                            ug[2][gy][v] = st.template node<2>();
                        }
                    }
            } else if constexpr (HYB) {
                // variable v of every node row: its y-values ys[gy*2 + v%2] and the neighbours'
                // bases relative to this cell's (shared by the node rows), in T
                auto sweep = [&](auto zero, const auto (&ys)[8][LANES], auto (&un)[4][4][2], int v) {
                    using T = decltype(zero);
                    T bd[5];
#pragma unroll
                    for (int t = 0; t < 5; ++t) bd[t] = t == 2 ? T(0) : T(L.base[v][lane - 2 + t] - Bc[v]);
#pragma unroll
                    for (int gy = 0; gy < 4; ++gy) {
                        T n[4] = {T(0), T(0), T(0), T(0)};
                        if (rows & (1u << gy)) {
                            T st5[5];
#pragma unroll
                            for (int t = 0; t < 5; ++t) {
                                // the centre's base is this cell's: its value is its y-deviation
                                const T y = ys[gy * 2 + (v & 1)][lane - 2 + t];
                                st5[t] = t == 2 ? y : bd[t] + y;
                            }
                            WenoSel<EX, T, SCHEME> st(st5[0], st5[1], st5[2], st5[3], st5[4]);
                            n[0] = st.template node<0>();
                            n[3] = st.template node<3>();
                            if (gy == 0 || gy == 3) {
                                n[1] = st.template node<1>();
                                n[2] = st.template node<2>();
                            }
                        }
                        un[0][gy][v & 1] = n[0];
                        un[3][gy][v & 1] = n[3];
                        if (gy == 0 || gy == 3) {
                            un[1][gy][v & 1] = n[1];
                            un[2][gy][v & 1] = n[2];
                        }
                    }
                };
#pragma unroll
                for (int v = 0; v < 2; ++v) sweep(0.0, L.yh.d, ugd, v);
#pragma unroll
                for (int v = 2; v < 4; ++v) sweep(W(0), L.yh.m, ugm, v);
            } else {
                double bc[4] = {Bc[0], Bc[1], Bc[2], Bc[3]};
#pragma unroll VU
                for (int v = 0; v < 4; ++v) {  // variable v's base sits in bc[0]
                    // the neighbours' bases relative to this cell's, shared by the node rows
                    W bd[5];
#pragma unroll
                    for (int t = 0; t < 5; ++t) bd[t] = t == 2 ? W(0) : W(L.base[v][lane - 2 + t] - bc[0]);
#pragma unroll
                    for (int gy = 0; gy < 4; ++gy) {
                        W n[4] = {W(0), W(0), W(0), W(0)};
                        if (rows & (1u << gy)) {
                            W st5[5];
#pragma unroll
                            for (int t = 0; t < 5; ++t) {
                                // the centre's base is bc[0] itself: its value is its y-deviation
                                const W y = L.y[gy * 4 + v][lane - 2 + t];
                                st5[t] = t == 2 ? y : bd[t] + y;
                            }
                            WenoSel<EX, W, SCHEME> st(st5[0], st5[1], st5[2], st5[3], st5[4]);
                            n[0] = st.template node<0>();
                            n[3] = st.template node<3>();
                            if (gy == 0 || gy == 3) {
                                n[1] = st.template node<1>();
                                n[2] = st.template node<2>();
                            }
                        }
                        rotate_push(ug[0][gy], n[0]);
                        rotate_push(ug[3][gy], n[3]);
                        if (gy == 0 || gy == 3) {
                            rotate_push(ug[1][gy], n[1]);
                            rotate_push(ug[2][gy], n[2]);
                        }
                    }
                    rotate_left(bc);
                }
            }
            auto to_face_h = [&](int gx, int gy, FS &f, int g) {
                f.set_recon(0, g, ugd[gx][gy][0], Uc[0]);
                f.set_recon(1, g, ugd[gx][gy][1], Uc[1]);
                f.set_recon(2, g, ugm[gx][gy][0], Uc[2]);
                f.set_recon(3, g, ugm[gx][gy][1], Uc[3]);
            };
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                if constexpr (RECON == 1) {
                    // the characteristic branches emitted their faces already
                } else if constexpr (HYB) {
                    to_face_h(0, g, fw, g);
                    to_face_h(3, g, fe, g);
                    to_face_h(g, 0, fs, g);
                    to_face_h(g, 3, fn, g);
                } else {
                    to_face(ug[0][g], fw, g);
                    to_face(ug[3][g], fe, g);
                    to_face(ug[g][0], fs, g);
                    to_face(ug[g][3], fn, g);
                }
            }
            if (bp) {
                // getPositivityPreservingTheta (src/WENOFV.cpp:1936-2000), double in both policies
                Eos e;
                eos_p<!EX>(P, Uc, e);
                double mu1, mu2;
                if constexpr (EX) {
                    const double u = Uc[2] / (Uc[0] + Uc[1]);
                    const double v = Uc[3] / (Uc[0] + Uc[1]);
                    const double a1 = fabs(u * 1.0 + v * 0.0) + e.c;
                    const double a2 = fabs(u * 0.0 + v * 1.0) + e.c;
                    mu1 = (a1 / A.dX) / (a1 / A.dX + a2 / A.dY);
                    mu2 = (a2 / A.dY) / (a1 / A.dX + a2 / A.dY);
                } else {
                    const double irho = rcp_nr(Uc[0] + Uc[1]);
                    const double a1 = fabs(Uc[2] * irho) + e.c;
                    const double a2 = fabs(Uc[3] * irho) + e.c;
                    const double t1 = a1 * idX, t2 = a2 * idY;
                    const double it = rcp_nr(t1 + t2);
                    mu1 = t1 * it;
                    mu2 = t2 * it;
                }
                double th[2];
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double uave = Uc[k];
                    const double eps = stdMin(1e-12, uave);
                    double rem = uave;
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        const double cg = omega1 * GLW[g] / 2;
                        rem -= cg * mu2 * fs.get(k, g, uave);
                        rem -= cg * mu2 * fn.get(k, g, uave);
                        rem -= cg * mu1 * fw.get(k, g, uave);
                        rem -= cg * mu1 * fe.get(k, g, uave);
                    }
                    if constexpr (EX) rem /= omegaR;
                    else rem *= 1.0 / omegaR;
                    double mloc = rem;
                    // mixed: the hardware minimum (it differs from sc_math::Min only for NaN
                    // operands, which the flux check reports) and a reciprocal
                    auto mn = [](double a, double b) { return EX ? scMin(a, b) : fmin(a, b); };
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        mloc = mn(mloc, fs.get(k, g, uave));
                        mloc = mn(mloc, fn.get(k, g, uave));
                        mloc = mn(mloc, fw.get(k, g, uave));
                        mloc = mn(mloc, fe.get(k, g, uave));
                    }
                    if constexpr (EX) {
                        th[k] = scMin(fabs((uave - eps) / (uave - mloc)), 1.0);
                    } else {
                        // the quotient is below 1 only where |uave - eps| < |uave - mloc|; elsewhere
                        // (a zero or NaN denominator included) theta_k is 1 without a reciprocal
                        const double nm = uave - eps, dn = uave - mloc;
                        th[k] = 1.0;
                        if (fabs(nm) < fabs(dn)) th[k] = scMin(fabs(nm * rcp_nr(dn)), 1.0);
                    }
                }
                theta = scMin(th[0], th[1]);
                // the limiter flag of the reference's outputs (src/WENOFV.cpp:1906-1909)
                if (A.theta_flag && owned && r >= y0 && r < y1)
                    A.theta_flag[(size_t)r * nx + xc] = theta < 1 - 1e-8 ? 1.0 : 0.0;
                // mixed: theta = 1 leaves the values as they are, so a wavefront with no
                // limited cell skips the loop
                if (theta >= 0.0 && (EX ? theta <= 1.0 : theta < 1.0)) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double ub = Uc[v];
#pragma unroll
                        for (int g = 0; g < 4; ++g) {
                            fe.limit(v, g, theta, ub);
                            fw.limit(v, g, theta, ub);
                            fn.limit(v, g, theta, ub);
                            fs.limit(v, g, theta, ub);
                        }
                    }
                } else if (!(theta >= 0.0 && theta <= 1.0) && owned && r >= y0 && r < y1) {
                    err |= WFV_FLAG_BAD_THETA;
                }
            }
            // parity hook: limited face values (and theta) of the owned cells; with faces_ring
            // also of the ring of ghost cells around the grid, in a layout padded by one cell
            bool fwrite = false;
            size_t fcell = 0;
            if (A.faces_out) {
                if (A.faces_ring) {
                    const bool col = (j >= HALO && j < HALO + SW && xc < nx) || xc == -1 || xc == nx;
                    const bool row = (r >= y0 && r < y1) || (r == -1 && y0 == 0) || (r == ny && y1 == ny);
                    fwrite = col && row && nrow > 0;
                    fcell = (size_t)(r + 1) * (nx + 2) + (xc + 1);
                } else {
                    fwrite = owned && r >= y0 && r < y1;
                    fcell = (size_t)r * nx + xc;
                }
            }
            if (fwrite) {
                double *fo = A.faces_out + fcell * 64;
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int g = 0; g < 4; ++g) {
                        fo[0 * 16 + v * 4 + g] = fw.get(v, g, Uc[v]);
                        fo[1 * 16 + v * 4 + g] = fe.get(v, g, Uc[v]);
                        fo[2 * 16 + v * 4 + g] = fs.get(v, g, Uc[v]);
                        fo[3 * 16 + v * 4 + g] = fn.get(v, g, Uc[v]);
                    }
                if (A.theta_out) A.theta_out[fcell] = theta;
            }
        }

        if (it > 0) {
            const bool interior_row = it <= nrow;
            // the RK base state of row r-1, loaded ahead of the fluxes that precede its use
            // (PF: at the top of the iteration, ahead of the sweeps too)
            if constexpr (EARLY && !PF)
                if (complete && A.Un) load_cell<double>(A.Un, nc, (size_t)(r - 1) * nx + xc, false, false, 0, Ub);
            // ---- 3. face below row r: Riemann problem (top of r-1 | bottom of r)
            double G[4] = {0.0, 0.0, 0.0, 0.0};
            if (owned && live) {
                FS top_prev;
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int g = 0; g < 4; ++g) top_prev.set(v, g, L.top[v * 4 + g][lane], top_base[v]);
                face_flux<PR, FLUX, 0, 1>(P, top_prev, top_base, fs, Uc, dxc, G, err, true);
            }
            double rhs_x[4] = {0.0, 0.0, 0.0, 0.0};
            double dyr = 1.0, idyr = 0.0;
            if (interior_row) {
                // ---- 4. face east of cell (xc, r): right face | neighbour lane's left face
                wave_sync();  // every lane has read the y-sweep values the west faces overwrite
                if (recon_lane) {  // the lanes whose west faces a face lane (lane - 1) reads
#pragma unroll
                    for (int v = 0; v < 4; ++v)
#pragma unroll
                        for (int g = 0; g < 4; ++g) L.west[v * 4 + g][lane] = fw.get(v, g, Uc[v]);
                }
                wave_sync();
                dyr = EARLY ? dy_row : ((const WFV_GLOBAL double *)A.dyC)[r];
                if constexpr (!EX) idyr = rcp_nr(dyr);
                if (xface_lane) {
                    // the neighbour's west face values, as exact FP64 states (base 0)
                    FS nb;  // the neighbour's west face values; momenta relative to its node 0
                    double nbase[4];
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        nbase[v] = EX ? 0.0 : L.west[v * 4 + 0][lane + 1];
#pragma unroll
                        for (int g = 0; g < 4; ++g) nb.set(v, g, L.west[v * 4 + g][lane + 1], nbase[v]);
                    }
                    double Fx[4];
                    face_flux<PR, FLUX, 1, 0>(P, fe, Uc, nb, nbase, dyr, Fx, err, owned);
#pragma unroll
                    for (int v = 0; v < 4; ++v) L.fx[v][lane] = Fx[v];
                }
                wave_sync();
                if (owned) {
                    // RHS order of assembleBoundaryFaceTerm: +west, -east (src/WENOFV.cpp:1802-1813)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        if constexpr (EX) {
                            double acc = 0.0;
                            acc += L.fx[v][lane - 1] / dxw / dyr;
                            acc += -L.fx[v][lane] / dxc / dyr;
                            rhs_x[v] = acc;
                        } else {
                            rhs_x[v] = (L.fx[v][lane - 1] * idxw - L.fx[v][lane] * idxc) * idyr;
                        }
                    }
                }
            }
            // ---- 5. complete row r-1: -north face, source, RK combination (src/WENOFV.cpp:1816-1827,
            //      1831-1872, 466-471)
            if (complete) {
                const int rr = r - 1;
                const double dyp = EARLY ? dy_prev : ((const WFV_GLOBAL double *)A.dyC)[rr];
                const double idyp = idy_prev;
                const double *Ucur;
                if constexpr (SLIM) Ucur = win[0];
                else Ucur = win[R - 1];
                if (!A.Un) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) Ub[v] = Ucur[v];
                } else if (!EARLY) {
                    load_cell<double>(A.Un, nc, (size_t)rr * nx + xc, false, false, 0, Ub);
                }
                double Unew[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double rhs = rhs_part[v];
                    if constexpr (EX) rhs -= G[v] / dxc / dyp;
                    else rhs -= G[v] * idxc * idyp;
                    if (v == 3 && A.gravity != 0.0) rhs -= (Ucur[0] + Ucur[1]) * A.gravity;
                    Unew[v] = A.a * Ub[v] + A.b * Ucur[v] + cdt * rhs;
                    ((WFV_GLOBAL double *)A.Uo)[v * nc + (size_t)rr * nx + xc] = Unew[v];
                }
                if (A.dtmin) {
                    // calculateDeltaT on the new state (src/WENOFV.cpp:489-513)
                    Eos e;
                    eos_p<!EX>(P, Unew, e);
                    double tmp;
                    if constexpr (EX) {
                        const double u = Unew[2] / (Unew[0] + Unew[1]);
                        const double v = Unew[3] / (Unew[0] + Unew[1]);
                        const double ex = fabs(u * 1.0 + v * 0.0) + e.c;
                        const double ey = fabs(u * 0.0 + v * 1.0) + e.c;
                        tmp = A.cfl / (ex / A.dX + ey / A.dY);
                    } else {
                        const double irho = rcp_nr(Unew[0] + Unew[1]);
                        tmp = A.cfl * rcp_nr((fabs(Unew[2] * irho) + e.c) * idX + (fabs(Unew[3] * irho) + e.c) * idY);
                    }
                    if (isnan(tmp)) err |= WFV_FLAG_NAN_DT;
                    else dtloc = fmin(dtloc, tmp);
                }
            }
            if (interior_row && owned) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if constexpr (EX) rhs_part[v] = rhs_x[v] + G[v] / dxc / dyr;
                    else rhs_part[v] = rhs_x[v] + G[v] * (idxc * idyr);
                }
            }
            dy_prev = dyr;
            idy_prev = idyr;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) top_base[v] = Uc[v];
        if (recon_lane) {  // the next row's south faces of the owned cells (recon lanes) read them
#pragma unroll
            for (int v = 0; v < 4; ++v)
#pragma unroll
                for (int g = 0; g < 4; ++g) L.top[v * 4 + g][lane] = fn.get(v, g, Uc[v]);
        }
    }

    // wavefront reductions: error bits and the CFL minimum
    for (int off = 32; off > 0; off >>= 1) {
        err |= __shfl_xor(err, off, LANES);
        dtloc = fmin(dtloc, __shfl_xor(dtloc, off, LANES));
    }
    if (lane == 0) {
        if (err && A.flags) atomicOr(A.flags, err);
        if (A.dtmin && dtloc < 1.0e300) atomicMin(A.dtmin, (unsigned long long)__double_as_longlong(dtloc));
        // Fused step control: every wavefront counts itself out once its error bits and CFL
        // minimum are in (release); the last of the launch's sp.total wavefronts (acquire) runs
        // the next step's control, in place of a launch of its own, and re-arms the count.
        // A wavefront that returned at the top (loop finished) took no part: then none did.
        if (A.ctl && __hip_atomic_fetch_add(&A.ctl->done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                         (unsigned)sp.total - 1u) {
            const unsigned long long bits =
                __hip_atomic_load(&A.ctl->dtmin_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned fl = __hip_atomic_load(&A.ctl->flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            step_control(A.ctl, bits, fl);
            A.ctl->done = 0;
        }
    }
    wave_sync();  // the next strip of a persistent wavefront reuses the LDS
}

// single domain: one wavefront per strip
template <class PR, int RECON, int SCHEME, int FLUX, int MINW, bool GH, bool BPK>
__global__ __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(MINW))) void stage_kernel(StageArgs A) {
    __shared__ WaveLds<ReconT<PR, RECON, SCHEME>> L;
    stage_body<PR, RECON, SCHEME, FLUX, prefetch_rows<PR, RECON, SCHEME, MINW>(), GH, BPK>(
        A, xcd_strip((int)blockIdx.x, (int)gridDim.x), L);
}

// Batch of independent instances with uneven grids in one persistent launch: the
// wavefronts stride over the flattened list of (instance, strip) items; each instance
// keeps its own grid, parameters, RK registers and device-side time step.
// Batch of independent instances with uneven grids in one launch: the wavefronts take the
// items of the flattened (instance, strip) list; each instance keeps its own grid,
// parameters, RK registers and device-side time step.  PERSIST: a grid smaller than the
// item list strides over it (a persistent launch); otherwise one wavefront per item.
template <class PR, int RECON, int SCHEME, int FLUX, bool PF, bool GH, bool BPK>
WFV_DEV void batch_item(const BatchStageArgs &B, int item, WaveLds<ReconT<PR, RECON, SCHEME>> &L) {
    const int i = B.item_inst[item];
    if (i < 0) return;  // a padding entry of the XCD-aware order
    const BatchInst &I = B.inst[i];
    StepCtl *ctl = I.ctl;
    StageArgs A;
    A.U = I.reg[B.in];
    A.Un = B.base < 0 ? nullptr : I.reg[B.base];
    A.Uo = I.reg[B.out];
    A.dxC = I.dxC;
    A.dxW = I.dxW;
    A.dyC = I.dyC;
    A.dt_ptr = &ctl->dt;
    A.active = &ctl->active;
    A.dt = 0.0;
    A.a = B.a;
    A.b = B.b;
    A.c = B.c;
    A.P = I.P;
    A.dX = I.dX;
    A.dY = I.dY;
    A.cfl = I.cfl;
    A.dx4 = I.dx4;
    A.ghost = I.ghost;
    A.gravity = I.gravity;
    A.nx = I.nx;
    A.ny = I.ny;
    A.bc = I.bc;
    A.flux = FLUX;
    A.bp = B.bp;
    A.ty = I.ty;
    A.sp = I.sp;
    A.flags = &ctl->flags;
    A.dtmin = B.want_dtmin ? &ctl->dtmin_bits : nullptr;
    A.ctl = nullptr;  // the batch runs its step control as a launch of its own
    A.faces_out = nullptr;
    A.faces_ring = 0;
    A.theta_out = nullptr;
    A.theta_flag = nullptr;
    stage_body<PR, RECON, SCHEME, FLUX, PF, GH, BPK>(A, B.item_strip[item], L);
}

template <class PR, int RECON, int SCHEME, int FLUX, int MINW, bool PERSIST, bool GH, bool BPK>
__global__ __launch_bounds__(LANES) __attribute__((amdgpu_waves_per_eu(MINW))) void batch_stage_kernel(
    BatchStageArgs B) {
    __shared__ WaveLds<ReconT<PR, RECON, SCHEME>> L;
    if constexpr (PERSIST) {
        for (int item = (int)blockIdx.x; item < B.total_items; item += (int)gridDim.x)
            batch_item<PR, RECON, SCHEME, FLUX, prefetch_rows<PR, RECON, SCHEME, MINW>(), GH, BPK>(B, item, L);
    } else {
        if ((int)blockIdx.x < B.total_items)
            batch_item<PR, RECON, SCHEME, FLUX, prefetch_rows<PR, RECON, SCHEME, MINW>(), GH, BPK>(B, (int)blockIdx.x, L);
    }
}

}  // namespace wfv
