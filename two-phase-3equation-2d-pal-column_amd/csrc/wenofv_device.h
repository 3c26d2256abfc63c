// wenofv_device.h -- device physics and reconstruction kernels of the WENOFV stage.
//
// Two precision policies share this code:
//   PrecFP64  : every operation in double, in the reference's association order, with
//               IEEE division and square root; compiled with -ffp-contract=off this
//               reproduces the reference CPU solver bit for bit.
//   PrecMixed : WENO smoothness indicators, nonlinear weights and flux arithmetic in
//               float with hardware reciprocals; the equation of state (volume
//               fraction, pressure, sound speed), the limiter's theta and the RK
//               accumulation stay in double, and so does the reconstruction of the
//               partial densities, the equation of state's inputs at the faces (the
//               conservative form; see stage_body).  The forms that must reconstruct in
//               double (characteristic, WENO-EXP, primitive3) do so with Newton-refined
//               reciprocals in place of IEEE divisions.
//   PrecMixedW64 : the mixed policy with every 5-point reconstruction in double.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "weno_constants.h"

#define WFV_DEV __device__ __forceinline__
// the all-FP64 policy's point functions, also compiled for the host: the CWENOFV and
// TwoPhaseEquation methods that evaluate single points (equation of state, fluxes,
// eigenvectors, WENO, the limiter's theta) call the very code the kernels run
#define WFV_HD __host__ __device__ __forceinline__
// global-memory view of a device pointer: accesses through it compile to global_* rather
// than flat_* instructions, which wait on the vector-memory counter only (a flat access
// also holds up the wavefront's next wait for its LDS operations)
#define WFV_GLOBAL __attribute__((address_space(1)))

namespace wfv {

struct PrecFP64 {
    using W = double;  // WENO arithmetic and the y-sweep values kept in LDS
    using F = double;  // Riemann-flux arithmetic and face values
    static constexpr bool exact = true;
};
struct PrecMixed {
    using W = float;
    using F = float;
    static constexpr bool exact = false;
};
// PrecMixedW64 (precision 2): the mixed policy with the 5-point reconstruction in double, as
// the mixed characteristic and 7-point forms already reconstruct; the Riemann-flux arithmetic
// stays FP32.
struct PrecMixedW64 {
    using W = double;
    using F = float;
    static constexpr bool exact = false;
};

// per-case constants; the host fills them with the reference's own expressions
// (c_k^2 = g_sound_k * g_sound_k, q = rho20*c2^2 - rho10*c1^2, include/Equations.hpp:86-92)
struct Phys {
    double rho10, rho20, c1_2, c2_2, preb, q;
    double ic1_2, ic2_2;  // 1/c1^2, 1/c2^2 (mixed policy)
};

// Gauss-Lobatto-4 weights (third_party/scTools/GaussQuadratureInRef.hpp:252-262)
constexpr double GLW[4] = {0.166666666666666666666666666667, 0.833333333333333333333333333333,
                                               0.833333333333333333333333333333, 0.166666666666666666666666666667};

WFV_HD double stdMin(double a, double b) { return (b < a) ? b : a; }  // std::min
WFV_HD double stdMax(double a, double b) { return (a < b) ? b : a; }  // std::max
WFV_HD double scMin(double a, double b) { return a < b ? a : b; }      // sc_math::Min

// ------------------------------------------------------------------ equation of state
// TwoPhaseEquation::getVolumeFrac / getPressure / getSoundSpeed (include/Equations.hpp:84-146)
// evaluated once per state and shared by every use, in double in both policies.
struct Eos {
    double rho, z1, rho1, rho2, p, c;
};
WFV_HD void eos(const Phys &P, const double U[4], Eos &e) {
    const double z1rho1 = U[0], z2rho2 = U[1];
    const double qs = z2rho2 * P.c2_2 - z1rho1 * P.c1_2;
    const double dq = P.q - qs;
    const double r = (dq + sqrt(dq * dq + 4.0 * z1rho1 * P.c1_2 * z2rho2 * P.c2_2)) / (2.0 * z2rho2 * P.c2_2);
    const double z1 = r / (1.0 + r);
    const double z2 = 1.0 - z1;
    e.rho = z1rho1 + z2rho2;
    e.z1 = z1;
    e.rho1 = z1rho1 / z1;
    e.rho2 = z2rho2 / z2;
    const double pre1 = P.c1_2 * (e.rho1 - P.rho10);
    const double pre2 = P.c2_2 * (e.rho2 - P.rho20);
    e.p = (z1 < z2) ? pre2 : pre1;
    const double sc = sqrt(e.rho * (z1 / (e.rho1 * P.c1_2) + z2 / (e.rho2 * P.c2_2)));
    e.c = 1.0 / sc;
}

// Double reciprocal and inverse square root from the hardware estimates plus one
// Newton-Raphson step (relative error ~1e-15); used only by the mixed policy.
WFV_DEV double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    return fma(fma(-x, r, 1.0), r, r);
}
WFV_DEV double rsq_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x * y;
    return fma(y, fma(-h, y, 0.5), y);
}

WFV_DEV double sqrt_nr(double x) { return x > 0.0 ? x * rsq_nr(x) : 0.0; }

// Mixed policy: pressure and sound speed (still double) with the algebra rearranged to no
// division.  With S the square root of the volume-fraction quadratic's discriminant,
// N = dq + S and D = 2 z2rho2 c2^2: z1 = N/(D+N), z2 = D/(D+N) and rho2 = (D+N)/(2 c2^2),
// so p = c2^2 (rho2 - rho20) = (D+N)/2 - c2^2 rho20 (equal to c1^2 (rho1 - rho10) by the
// pressure equilibrium the quadratic expresses), and Wood's
// 1/c^2 = rho (z1^2/(z1rho1 c1^2) + z2^2/(z2rho2 c2^2)) = rho (N^2/t1 + 2D)/(D+N)^2 with
// t1 = z1rho1 c1^2, that is c = (D+N) |t1| / sqrt(t1 rho (N^2 + 2 D t1)): one Newton-refined
// inverse square root and no reciprocal (|t1| = sqrt(t1^2) keeps the sign of the quotient
// form where an unlimited reconstruction leaves z1rho1 < 0).  2 D t1 = 4 t1 t2 (t2 =
// z2rho2 c2^2) is the discriminant's own product term, passed in as f.  Only rho, p and c
// are set.
WFV_DEV double wood_c(double rho, double t1, double N, double f, double DN) {
    return (DN * fabs(t1)) * rsq_nr(t1 * rho * (N * N + f));
}
WFV_DEV void eos_fast(const Phys &P, const double U[4], Eos &e) {
    const double a = U[0], b = U[1];
    const double t1 = a * P.c1_2, t2 = b * P.c2_2;
    const double dq = P.q - (t2 - t1);
    const double f = 4.0 * t1 * t2;
    const double S = sqrt_nr(dq * dq + f);
    const double N = dq + S;
    const double DN = 2.0 * t2 + N;
    e.rho = a + b;
    e.p = 0.5 * DN - P.c2_2 * P.rho20;
    e.c = wood_c(e.rho, t1, N, f, DN);
}

// The mixed Riemann solvers' equation of state: the pressure as eos_fast and the sound speed
// with Newton-refined reciprocal and inverse square root, both to double accuracy.  The
// discriminant's square root S = x y (1.5 - 0.5 x y^2) = g (1.5 - 0.5 g y) with y the inverse
// square root estimate and g = x y: one Newton step, applied to g; at x = 0 (y taken at a
// tiny floor) g and S are 0.
WFV_DEV void eos_flux(const Phys &P, double a, double b, Eos &e) {
    const double t1 = a * P.c1_2, t2 = b * P.c2_2;
    const double dq = P.q - (t2 - t1);
    const double f = 4.0 * t1 * t2;
    const double x = dq * dq + f;
    const double y = __builtin_amdgcn_rsq(fmax(x, 1.0e-300));
    const double g = x * y;
    const double N = g * ((g * y) * -0.5 + 1.5) + dq;
    const double DN = 2.0 * t2 + N;
    e.rho = a + b;
    e.p = 0.5 * DN - P.c2_2 * P.rho20;
    e.c = wood_c(e.rho, t1, N, f, DN);
}

template <bool FAST>
WFV_DEV void eos_p(const Phys &P, const double U[4], Eos &e) {
    if constexpr (FAST) eos_fast(P, U, e);
    else eos(P, U, e);
}

// ------------------------------------------------------------------ Riemann fluxes
// F = the physical flux along x (xdir: the reference's nx > ny) or y.
template <class T>
WFV_HD void phy_flux(const T U[4], T rho, T u, T v, T p, bool xdir, T F[4]) {
    // TwoPhaseEquation::getPhyFlux (include/Equations.hpp:158-186)
    if (xdir) {
        F[0] = U[0] * u; F[1] = U[1] * u; F[2] = rho * (u * u) + p; F[3] = rho * u * v;
    } else {
        F[0] = U[0] * v; F[1] = U[1] * v; F[2] = rho * v * u; F[3] = rho * (v * v) + p;
    }
}

// The all-FP64 policy's Riemann solvers: the reference's formulas, operation by operation,
// through the face normal (nx, ny).  The kernels pass (1, 0) or (0, 1) as constants (the
// riemann_* templates below); the literal products with the normal's components are kept,
// so that the rounding and signed zeros match the reference's.
WFV_HD void riemann_llf_n(const Phys &P, const double UL[4], const double UR[4], double nx, double ny, double F[4]) {
    // getLLFRiemannFlux (include/Equations.hpp:363-378)
    Eos el, er;
    eos(P, UL, el);
    eos(P, UR, er);
    const double rl = el.rho, rr = er.rho;
    const double ul = UL[2] / rl, vl = UL[3] / rl;
    const double ur = UR[2] / rr, vr = UR[3] / rr;
    const double wsL = fabs(ul * nx + vl * ny) + el.c;
    const double wsR = fabs(ur * nx + vr * ny) + er.c;
    const double ws = stdMax(wsL, wsR);
    double FL[4], FR[4];
    phy_flux<double>(UL, rl, ul, vl, el.p, nx > ny, FL);
    phy_flux<double>(UR, rr, ur, vr, er.p, nx > ny, FR);
#pragma unroll
    for (int r = 0; r < 4; ++r) F[r] = 0.5 * (FL[r] + FR[r] - ws * (UR[r] - UL[r]));
}

WFV_HD void riemann_hllc_n(const Phys &P, const double UL[4], const double UR[4], double nx, double ny, double F[4]) {
    // getHLLCRiemannFlux (include/Equations.hpp:381-483)
    Eos el, er;
    eos(P, UL, el);
    eos(P, UR, er);
    const double rho_L = el.rho, rho_R = er.rho;
    const double u_L = UL[2] / rho_L, v_L = UL[3] / rho_L;
    const double u_R = UR[2] / rho_R, v_R = UR[3] / rho_R;
    const double pre_L = el.p, pre_R = er.p, c_L = el.c, c_R = er.c;
    const double vn_L = u_L * nx + v_L * ny;
    const double vn_R = u_R * nx + v_R * ny;
    double FL[4], FR[4];
    FL[0] = UL[0] * vn_L; FL[1] = UL[1] * vn_L;
    FL[2] = UL[2] * vn_L + pre_L * nx; FL[3] = UL[3] * vn_L + pre_L * ny;
    FR[0] = UR[0] * vn_R; FR[1] = UR[1] * vn_R;
    FR[2] = UR[2] * vn_R + pre_R * nx; FR[3] = UR[3] * vn_R + pre_R * ny;
    const double sL = sqrt(rho_L), sR = sqrt(rho_R);
    const double vn_roe = (vn_L * sL + vn_R * sR) / (sL + sR);
    double c2_roe = (c_L * c_L * sL + c_R * c_R * sR) / (sL + sR);
    c2_roe += 0.5 * sL * sR / (sL + sR) / (sL + sR) * (vn_R - vn_L) * (vn_R - vn_L);
    const double c_roe = sqrt(c2_roe);
    const double S_L = stdMin(vn_L - c_L, vn_roe - c_roe);
    const double S_R = stdMax(vn_R + c_R, vn_roe + c_roe);
    const double S_star = ((S_L - vn_L) * rho_L * vn_L - (S_R - vn_R) * rho_R * vn_R + pre_R - pre_L) /
                          ((S_L - vn_L) * rho_L - (S_R - vn_R) * rho_R);
    // The reference's four cases (S* >= 0: F_L or the left star flux; S* < 0: F_R or the
    // right star flux) as one star-flux evaluation on the upwind side's values: the taken
    // case performs the same operations, and a wavefront whose lanes take different cases
    // evaluates one star state instead of two.
    const bool left = S_star >= 0.0;
    const bool outside = left ? (S_L >= 0.0) : (S_R <= 0.0);
    const double S_K = left ? S_L : S_R, vn_K = left ? vn_L : vn_R, rho_K = left ? rho_L : rho_R;
    const double u_K = left ? u_L : u_R, v_K = left ? v_L : v_R;
    double U_K[4], F_K[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        U_K[r] = left ? UL[r] : UR[r];
        F_K[r] = left ? FL[r] : FR[r];
    }
    const double k = (S_K - vn_K) / (S_K - S_star);
    const double Us[4] = {U_K[0], U_K[1], rho_K * (u_K + (S_star - vn_K) * nx), rho_K * (v_K + (S_star - vn_K) * ny)};
#pragma unroll
    for (int r = 0; r < 4; ++r) F[r] = outside ? F_K[r] : S_K * (k * Us[r] - U_K[r]) + F_K[r];
}

template <int NX, int NY>
WFV_HD void riemann_llf(const Phys &P, const double UL[4], const double UR[4], double F[4]) {
    riemann_llf_n(P, UL, UR, NX, NY, F);
}
template <int NX, int NY>
WFV_HD void riemann_hllc(const Phys &P, const double UL[4], const double UR[4], double F[4]) {
    riemann_hllc_n(P, UL, UR, NX, NY, F);
}

// Mixed-policy node solvers on a face whose normal is x (NX = 1) or y.  Each side arrives as
// its partial densities in double (they alone enter the equation of state) and its normal
// and tangential momenta in float.  The fluxes are returned minus pref * n with pref the left
// state's pressure: the O(1) pressure term of the momentum flux would otherwise dominate its
// FP32 rounding, while its constant part cancels in the divergence once added back in double
// (the caller adds pref * n).  The same algorithms as above, with the products by the
// normal's 0/1 components dropped, the left pressure offset (p_L - pref = 0) known, and the
// HLLC star-state flux taken in its closed form (below); only rounding differs.
struct SideF {
    double a, b;  // z1 rho1, z2 rho2
    float mn, mt; // normal and tangential momentum
};

WFV_DEV void hllc_f(const Phys &P, const SideF &L, const SideF &R, double Fd[2], float F[2], double &pref) {
    Eos el, er;
    eos_flux(P, L.a, L.b, el);
    eos_flux(P, R.a, R.b, er);
    pref = el.p;
    const float rho_L = float(el.rho), rho_R = float(er.rho);
    const float irho_L = __builtin_amdgcn_rcpf(rho_L);
    const float vn_L = L.mn * irho_L, vn_R = R.mn * __builtin_amdgcn_rcpf(rho_R);
    const float pre_R = float(er.p - el.p), c_L = float(el.c), c_R = float(er.c);
    // Roe weights sqrt(rho_K) / (sqrt(rho_L) + sqrt(rho_R)) from one square root:
    // w_L = 1 / (1 + t), w_R = t w_L with t = sqrt(rho_R / rho_L)
    const float t = __builtin_amdgcn_sqrtf(rho_R * irho_L);
    const float wL = __builtin_amdgcn_rcpf(1.0f + t), wR = t * wL;
    const float dvn = vn_R - vn_L;
    const float vn_roe = wR * dvn + vn_L;
    const float cL2 = c_L * c_L;
    const float c2_roe = wR * (c_R * c_R - cL2) + cL2 + (0.5f * wL * wR) * (dvn * dvn);
    const float c_roe = __builtin_amdgcn_sqrtf(c2_roe);
    const float S_L = __builtin_fminf(vn_L - c_L, vn_roe - c_roe);
    const float S_R = __builtin_fmaxf(vn_R + c_R, vn_roe + c_roe);
    const float qL = (S_L - vn_L) * rho_L, qR = (S_R - vn_R) * rho_R;
    const float S_star = (qL * vn_L - qR * vn_R + pre_R) * __builtin_amdgcn_rcpf(qL - qR);
    // upwind side K; the flux F_K + S'(U*_K - U_K) with S' = min(S_L, 0) (left) or
    // max(S_R, 0) (right), which is F_K outside the fan.  With U*_K = k (a, b, rho S*, m_t)_K,
    // k = (S_K - vn_K)/(S_K - S*) = 1 + e, e = (S* - vn_K)/(S_K - S*), the transported
    // components move with phi = vn_K + S' e and the normal momentum flux is
    // m_K vn_K + p_K + S' e S_K rho_K.  Where S' = 0 (outside the fan) the correction is 0
    // exactly, whatever e is (S_K = S* would make it infinite).
    const bool left = S_star >= 0.0f;
    const float S_K = left ? S_L : S_R, vn_K = left ? vn_L : vn_R, rho_K = left ? rho_L : rho_R;
    const float Sp = left ? __builtin_fminf(S_L, 0.0f) : __builtin_fmaxf(S_R, 0.0f);
    const float p_K = left ? 0.0f : pre_R, mn_K = left ? L.mn : R.mn, mt_K = left ? L.mt : R.mt;
    const float e = (S_star - vn_K) * __builtin_amdgcn_rcpf(S_K - S_star);
    const float g = Sp != 0.0f ? Sp * e : 0.0f;
    const float phi = vn_K + g;
    const double phid = double(phi);
    Fd[0] = (left ? L.a : R.a) * phid;
    Fd[1] = (left ? L.b : R.b) * phid;
    F[0] = g * (S_K * rho_K) + (mn_K * vn_K + p_K);
    F[1] = mt_K * phi;
}

// jump = right minus left state, formed by the caller from exact double differences of the
// bases plus the float deviations
WFV_DEV void llf_f(const Phys &P, const SideF &L, const SideF &R, const float jump[4], float F[4], double &pref) {
    Eos el, er;
    eos_flux(P, L.a, L.b, el);
    eos_flux(P, R.a, R.b, er);
    pref = el.p;
    const float rho_L = float(el.rho), rho_R = float(er.rho);
    const float vn_L = L.mn * __builtin_amdgcn_rcpf(rho_L), vn_R = R.mn * __builtin_amdgcn_rcpf(rho_R);
    const float pre_R = float(er.p - el.p);
    const float ws = __builtin_fmaxf(fabsf(vn_L) + float(el.c), fabsf(vn_R) + float(er.c));
    // 0.5 (F_L + F_R - ws jump) with the halves taken into the speeds
    const float hL = 0.5f * vn_L, hR = 0.5f * vn_R, hw = 0.5f * ws;
    const float UL[4] = {float(L.a), float(L.b), L.mn, L.mt}, UR[4] = {float(R.a), float(R.b), R.mn, R.mt};
#pragma unroll
    for (int r = 0; r < 4; ++r) F[r] = UL[r] * hL + (UR[r] * hR - hw * jump[r]);
    F[2] += 0.5f * pre_R;
}

// getRiemannFlux dispatch (include/Equations.hpp:329-360), resolved at compile time
template <int FLUX, int NX, int NY>
WFV_HD void riemann(const Phys &P, const double UL[4], const double UR[4], double F[4]) {
    if constexpr (FLUX == 1) riemann_hllc<NX, NY>(P, UL, UR, F);
    else riemann_llf<NX, NY>(P, UL, UR, F);
}

// ------------------------------------------------------------------ eigenvectors
// getLEigenMatrix / getREigenMatrix (include/Equations.hpp:189-327), row-major 4x4,
// built in double (they involve the equation of state) and used in W.
template <class T, bool XDIR, bool FAST = false>
WFV_HD void eigen_LR(const Phys &P, const double U[4], T L[16], T R[16]) {
    if constexpr (FAST) {
        // Mixed policy: the same matrices with the volume fraction from the division-free
        // equation of state (z1 = N/(D+N), z2 = D/(D+N)) and every quotient as a product
        // with a Newton-refined reciprocal; sc_k^2 uses 1 - zrho1/(zrho1+zrho2) =
        // zrho2/(zrho1+zrho2), which does not cancel.
        const double a = U[0], b = U[1];
        const double t1 = a * P.c1_2, t2 = b * P.c2_2;
        const double dq = P.q - (t2 - t1);
        const double f = 4.0 * t1 * t2;
        const double x = dq * dq + f;
        const double S = x * rsq_nr(fmax(x, 1.0e-300));
        const double N = dq + S, D = 2.0 * t2, DN = D + N;
        const double iDN = rcp_nr(DN);
        const double z1 = N * iDN, z2 = D * iDN;
        const double rho = a + b, irho = rcp_nr(rho);
        const double c = wood_c(rho, t1, N, f, DN);
        const double u = U[2] * irho, v = U[3] * irho;
        const double iz1 = rcp_nr(z1), iz2 = rcp_nr(z2);
        const double zrho1 = t1 * iz1 * iz1, zrho2 = t2 * iz2 * iz2;
        const double izz = rcp_nr(zrho1 + zrho2);
        const double sc1_2 = P.c1_2 * iz1 * (zrho2 * izz);
        const double sc2_2 = P.c2_2 * iz2 * (zrho1 * izz);
        const double ic = rcp_nr(c), i2c2 = 0.5 * ic * ic, irc2 = irho * ic * ic;
        const double d12 = sc1_2 - sc2_2, id12 = rcp_nr(d12);
        const double un = XDIR ? u : v;
        L[0] = T((sc1_2 + c * un) * i2c2);
        L[1] = T((sc2_2 + c * un) * i2c2);
        L[4] = T(-b * d12 * irc2);
        L[5] = T(a * d12 * irc2);
        L[6] = T(0); L[7] = T(0);
        L[12] = T((sc1_2 - c * un) * i2c2);
        L[13] = T((sc2_2 - c * un) * i2c2);
        if (XDIR) {
            L[2] = T(-0.5 * ic); L[3] = T(0);
            L[8] = T(-v); L[9] = T(-v); L[10] = T(0); L[11] = T(1);
            L[14] = T(0.5 * ic); L[15] = T(0);
        } else {
            L[2] = T(0); L[3] = T(-0.5 * ic);
            L[8] = T(-u); L[9] = T(-u); L[10] = T(1); L[11] = T(0);
            L[14] = T(0); L[15] = T(0.5 * ic);
        }
        R[0] = T(a * irho); R[4] = T(b * irho);
        R[1] = T(-sc2_2 * id12); R[5] = T(sc1_2 * id12); R[9] = T(u); R[13] = T(v);
        R[2] = T(0); R[6] = T(0);
        R[3] = T(a * irho); R[7] = T(b * irho);
        if (XDIR) {
            R[8] = T(u - c); R[12] = T(v);
            R[10] = T(0); R[14] = T(1);
            R[11] = T(u + c); R[15] = T(v);
        } else {
            R[8] = T(u); R[12] = T(v - c);
            R[10] = T(1); R[14] = T(0);
            R[11] = T(u); R[15] = T(v + c);
        }
        return;
    }
    Eos e;
    eos(P, U, e);
    const double rho = e.rho, u = U[2] / rho, v = U[3] / rho, c = e.c;
    const double z1 = e.z1, z2 = 1.0 - z1;
    const double zrho1 = U[0] * P.c1_2 / (z1 * z1);
    const double zrho2 = U[1] * P.c2_2 / (z2 * z2);
    const double sc1_2 = P.c1_2 / z1 - P.c1_2 / z1 * zrho1 / (zrho1 + zrho2);
    const double sc2_2 = P.c2_2 / z2 - P.c2_2 / z2 * zrho2 / (zrho1 + zrho2);
    const double c2 = c * c;
    const double d12 = sc1_2 - sc2_2;
    const double un = XDIR ? u : v;
    L[0] = T((sc1_2 + c * un) / (2.0 * c2));
    L[1] = T((sc2_2 + c * un) / (2.0 * c2));
    L[4] = T(-U[1] * d12 / (rho * c2));
    L[5] = T(U[0] * d12 / (rho * c2));
    L[6] = T(0); L[7] = T(0);
    L[12] = T((sc1_2 - c * un) / (2.0 * c2));
    L[13] = T((sc2_2 - c * un) / (2.0 * c2));
    if (XDIR) {
        L[2] = T(-1.0 / (2.0 * c)); L[3] = T(0);
        L[8] = T(-v); L[9] = T(-v); L[10] = T(0); L[11] = T(1);
        L[14] = T(1.0 / (2.0 * c)); L[15] = T(0);
    } else {
        L[2] = T(0); L[3] = T(-1.0 / (2.0 * c));
        L[8] = T(-u); L[9] = T(-u); L[10] = T(1); L[11] = T(0);
        L[14] = T(0); L[15] = T(1.0 / (2.0 * c));
    }
    R[0] = T(U[0] / rho); R[4] = T(U[1] / rho);
    R[1] = T(-sc2_2 / d12); R[5] = T(sc1_2 / d12); R[9] = T(u); R[13] = T(v);
    R[2] = T(0); R[6] = T(0);
    R[3] = T(U[0] / rho); R[7] = T(U[1] / rho);
    if (XDIR) {
        R[8] = T(u - c); R[12] = T(v);
        R[10] = T(0); R[14] = T(1);
        R[11] = T(u + c); R[15] = T(v);
    } else {
        R[8] = T(u); R[12] = T(v - c);
        R[10] = T(1); R[14] = T(0);
        R[11] = T(u); R[15] = T(v + c);
    }
}

// Array2D * Array1D (third_party/scTools/Array.h:604-626): C[i] = ((0 + A_i0 B0) + ...) + A_i3 B3
template <class T>
WFV_HD void matvec(const T A[16], const T B[4], T C[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T acc = T(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc += A[i * 4 + j] * B[j];
        C[i] = acc;
    }
}

// ------------------------------------------------------------------ limiter
// getPositivityPreservingTheta (src/WENOFV.cpp:1936-2000) of one cell in double, with the
// operations of the stage kernel's all-FP64 limiter: U the cell average, F[side][v][g] its
// face values (sides left, right, bottom, top), dX and dY the grid's global widths.
WFV_HD double bp_theta(const Phys &P, const double U[4], const double F[4][4][4], double dX, double dY) {
    constexpr double omega1 = 0.1;
    constexpr double omegaR = 1.0 - 2 * omega1;
    Eos e;
    eos(P, U, e);
    const double u = U[2] / (U[0] + U[1]);
    const double v = U[3] / (U[0] + U[1]);
    const double a1 = fabs(u * 1.0 + v * 0.0) + e.c;
    const double a2 = fabs(u * 0.0 + v * 1.0) + e.c;
    const double mu1 = (a1 / dX) / (a1 / dX + a2 / dY);
    const double mu2 = (a2 / dY) / (a1 / dX + a2 / dY);
    double th[2];
    for (int k = 0; k < 2; ++k) {
        const double uave = U[k];
        const double eps = stdMin(1e-12, uave);
        double rem = uave;
        for (int g = 0; g < 4; ++g) {
            const double cg = omega1 * GLW[g] / 2;
            rem -= cg * mu2 * F[2][k][g];
            rem -= cg * mu2 * F[3][k][g];
            rem -= cg * mu1 * F[0][k][g];
            rem -= cg * mu1 * F[1][k][g];
        }
        rem /= omegaR;
        double m = rem;
        for (int g = 0; g < 4; ++g) {
            m = scMin(m, F[2][k][g]);
            m = scMin(m, F[3][k][g]);
            m = scMin(m, F[0][k][g]);
            m = scMin(m, F[1][k][g]);
        }
        th[k] = scMin(fabs((uave - eps) / (uave - m)), 1.0);
    }
    return scMin(th[0], th[1]);
}

// ------------------------------------------------------------------ WENO
// One 5-point stencil {a,b,c,d,e} = {u_mm, u_m, u, u_p, u_pp}; the smoothness indicators
// are evaluated once and shared by every Gauss-Lobatto node reconstructed from the
// stencil (the reference recomputes them per node with identical operands).
template <class T>
WFV_HD void beta3(T a, T b, T c, T d, T e, T &b0, T &b1, T &b2) {
    // include/Utility.hpp:48-50 (and 270-272 for WENO-Z on the normalised stencil)
    T t;
    t = a - T(4) * b + T(3) * c; b0 = T(0.25) * (t * t);
    t = a - T(2) * b + c;        b0 = b0 + T(13.0 / 12.0) * (t * t);
    t = b - d;                   b1 = T(0.25) * (t * t);
    t = b - T(2) * c + d;        b1 = b1 + T(13.0 / 12.0) * (t * t);
    t = T(3) * c - T(4) * d + e; b2 = T(0.25) * (t * t);
    t = c - T(2) * d + e;        b2 = b2 + T(13.0 / 12.0) * (t * t);
}

// Double (both policies' FP64 reconstructions): the reference's arithmetic, operation by
// operation (include/Utility.hpp:5-95, 193-290).
template <class T, int SCHEME>
struct Weno5 {
    static_assert(sizeof(T) == 8, "the mixed policy uses WenoLin");
    T s0, s1, s2, s3, s4;
    T k0, k1, k2;  // JS: (eps+beta_k)^2; Z: tau/(beta_k+eps)

    WFV_HD Weno5(T a, T b, T c, T d, T e) : s0(a), s1(b), s2(c), s3(d), s4(e) {
        T b0, b1, b2;
        if constexpr (SCHEME == 0) {
            // WENO5threconstruction (include/Utility.hpp:48-61), eps = 1e-6
            beta3(a, b, c, d, e, b0, b1, b2);
            const T eps = T(1.0e-6);
            k0 = (eps + b0) * (eps + b0);
            k1 = (eps + b1) * (eps + b1);
            k2 = (eps + b2) * (eps + b2);
        } else {
            // WENO5Zthreconstruction (include/Utility.hpp:261-280): beta of the min-max
            // normalised stencil, tau = |beta0 - beta2|, eps = 1e-10
            const T mx = stdMax(stdMax(stdMax(a, b), stdMax(e, d)), c);
            const T mn = stdMin(stdMin(stdMin(a, b), stdMin(e, d)), c);
            const T den = mx - mn + T(1e-20);
            const T a_ = (a - mn) / den, b_ = (b - mn) / den, c_ = (c - mn) / den;
            const T d_ = (d - mn) / den, e_ = (e - mn) / den;
            beta3(a_, b_, c_, d_, e_, b0, b1, b2);
            const T eps = T(1.0e-10);
            const T tau = fabs(b0 - b2);
            k0 = tau / (b0 + eps);
            k1 = tau / (b1 + eps);
            k2 = tau / (b2 + eps);
        }
    }

    template <int GP>
    WFV_HD T node() const {
        const T h0 = T(WENO_H[GP][0][0]) * s0 + T(WENO_H[GP][0][1]) * s1 + T(WENO_H[GP][0][2]) * s2;
        const T h1 = T(WENO_H[GP][1][0]) * s1 + T(WENO_H[GP][1][1]) * s2 + T(WENO_H[GP][1][2]) * s3;
        const T h2 = T(WENO_H[GP][2][0]) * s2 + T(WENO_H[GP][2][1]) * s3 + T(WENO_H[GP][2][2]) * s4;
        T w0, w1, w2;
        if constexpr (SCHEME == 0) {
            w0 = T(WENO_D[GP][0]) / k0;
            w1 = T(WENO_D[GP][1]) / k1;
            w2 = T(WENO_D[GP][2]) / k2;
        } else {
            w0 = T(WENO_D[GP][0]) * (T(1) + k0);
            w1 = T(WENO_D[GP][1]) * (T(1) + k1);
            w2 = T(WENO_D[GP][2]) * (T(1) + k2);
        }
        const T sum = w0 + w1 + w2;
        w0 = w0 / sum; w1 = w1 / sum; w2 = w2 / sum;
        return w0 * h0 + w1 * h1 + w2 * h2;
    }
};

// Mixed policy (float, and double for the forms it reconstructs in double).  Candidate k is
// the quadratic with the stencil's cell averages,
//   p_k(x) = c + u_k x / 2 + t_k (x^2 - 1/12) / 2   (x in cell widths from the cell centre),
// whose first- and second-difference forms u_k, t_k are exactly the linear forms inside the
// smoothness indicator beta_k = u_k^2 / 4 + 13 t_k^2 / 12.  The indicators are taken as
// 4 beta_k = u_k^2 + (13/3) t_k^2 with the regularisation scaled alike (4 eps, and 4e-10 den^2
// for WENO-Z), which leaves every nonlinear weight as it is.  With the node-independent
// factors q_k (alpha_k = d_k q_k, up to a common factor) folded in once per stencil, a node
// costs
//   c + (sum_k d_k (A u_k q_k + B t_k q_k)) / (sum_k d_k q_k),  A = x/2, B = (x^2 - 1/12)/2,
// the same value as the tabulated candidates (WENO_H) up to rounding.  u_k, t_k come from
// the stencil's adjacent differences, or straight from the values when the centre is a
// literal zero (ZC: stencils of deviations from the centre).  Reciprocals: the hardware
// estimate in float, one Newton step on it in double.
template <class T>
WFV_DEV T rcp_mixed(T x) {
    if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
    else return rcp_nr(x);
}
template <class T, int SCHEME, bool ZC = false>
struct WenoLin {
    T c, pu[3], pt[3], q[3];

    WFV_DEV WenoLin(T a, T b, T cc, T d, T e) : c(cc) {
        T u0, t0, u1, t1, u2, t2;
        if constexpr (ZC) {
            u0 = a - T(4) * b; t0 = a - T(2) * b;
            u1 = d - b;        t1 = b + d;
            u2 = T(4) * d - e; t2 = e - T(2) * d;
        } else {
            const T d1 = b - a, d2 = cc - b, d3 = d - cc, d4 = e - d;
            u0 = T(3) * d2 - d1; t0 = d2 - d1;
            u1 = d2 + d3;        t1 = d3 - d2;
            u2 = T(3) * d3 - d4; t2 = d4 - d3;
        }
        constexpr T K = T(13.0 / 3.0);
        const T b0 = u0 * u0 + (K * t0) * t0;
        const T b1 = u1 * u1 + (K * t1) * t1;
        const T b2 = u2 * u2 + (K * t2) * t2;
        // In double the q_k are scaled by the product of their three denominators, which the
        // ratio of a node's q-weighted sums cancels: two products per q_k replace a
        // Newton-refined reciprocal (three instructions, one of them a transcendental).  The
        // denominators are at least 1.6e-11 (JS) and 1e-100 (Z) and, for stencil differences
        // below 1e15, at most 1e61, so their products stay normal.
        constexpr bool SCALED = sizeof(T) == 8;
        if constexpr (SCHEME == 0) {
            // WENO5threconstruction (include/Utility.hpp:48-61): alpha_k = d_k/(eps+beta_k)^2
            const T eps = T(4.0e-6);
            const T E0 = (eps + b0) * (eps + b0), E1 = (eps + b1) * (eps + b1), E2 = (eps + b2) * (eps + b2);
            if constexpr (SCALED) {
                q[0] = E1 * E2;
                q[1] = E0 * E2;
                q[2] = E0 * E1;
            } else {
                q[0] = rcp_mixed(E0);
                q[1] = rcp_mixed(E1);
                q[2] = rcp_mixed(E2);
            }
        } else if constexpr (SCALED) {
            // WENO5Zthreconstruction as below, q_k = (B_k + tau) prod_{j != k} B_j with
            // B_k = beta_k + eps den^2; the floor added to the regularisation is 1e-100 (it
            // matters only for stencils flat to 1e-38, whose node values do not depend on
            // the weights), so that a constant stencil's q_k = B^3 stays normal
            const T mx = fmax(fmax(fmax(fmax(a, b), cc), d), e);
            const T mn = fmin(fmin(fmin(fmin(a, b), cc), d), e);
            const T den = mx - mn;
            const T e2 = (T(4.0e-10) * den) * den + T(1.0e-100);
            const T tau = fabs(b0 - b2);
            const T B0 = b0 + e2, B1 = b1 + e2, B2 = b2 + e2;
            const T P01 = B0 * B1, P02 = B0 * B2, P12 = B1 * B2, PI = P01 * B2;
            q[0] = tau * P12 + PI;
            q[1] = tau * P02 + PI;
            q[2] = tau * P01 + PI;
        } else {
            // WENO5Zthreconstruction (include/Utility.hpp:261-280) on the min-max normalised
            // stencil: the indicators are shift-invariant quadratic forms, so tau/(beta_k + eps)
            // there equals tau_raw/(beta_k,raw + eps den^2) here; the smallest normal float
            // added keeps a constant stencil at q = 1 where eps den^2 underflows.  The
            // normalisation's + 1e-20 offset matters only below that floor and is left out;
            // the max/min chains become 3-operand instructions.
            const T mx = fmax(fmax(fmax(fmax(a, b), cc), d), e);
            const T mn = fmin(fmin(fmin(fmin(a, b), cc), d), e);
            const T den = mx - mn;
            const T e2 = (T(4.0e-10) * den) * den + T(1.17549435e-38f);
            const T tau = fabs(b0 - b2);
            q[0] = tau * rcp_mixed(b0 + e2) + T(1);
            q[1] = tau * rcp_mixed(b1 + e2) + T(1);
            q[2] = tau * rcp_mixed(b2 + e2) + T(1);
        }
        pu[0] = q[0] * u0; pu[1] = q[1] * u1; pu[2] = q[2] * u2;
        pt[0] = q[0] * t0; pt[1] = q[1] * t1; pt[2] = q[2] * t2;
    }

    template <int GP>
    WFV_DEV T node() const {
        // Gauss-Lobatto nodes x = -1/2, -sqrt(5)/10, sqrt(5)/10, 1/2 with A = x/2 and
        // B = (x^2 - 1/12)/2: the node is c + (A PU + B PT)/S with PU, PT, S the sums of
        // pu_k, pt_k, q_k weighted by the node's linear weights d_k.  The d_k enter only
        // through their ratios, so they are scaled to whole numbers at the end nodes
        // (d = (3, 6, 1)/10 and its mirror) and to a unit centre weight at the inner ones,
        // and numerator and denominator are divided by B: c + (PT + (A/B) PU)/(S/B) with
        // A/B = -/+3, 1/B = 12 at the end nodes, A/B = +/-3 sqrt(5), 1/B = -60 inside.  Every
        // constant is then an exact small integer but three, the inner nodes' d_0/d_1,
        // d_2/d_1 and 3 sqrt(5).
        constexpr int k0 = (GP == 0 || GP == 1) ? 0 : 2, k2 = 2 - k0;  // mirrored at GP 2, 3
        if constexpr (GP == 0 || GP == 3) {
            const T PU = T(3) * pu[k0] + (T(6) * pu[1] + pu[k2]);
            const T PT = T(3) * pt[k0] + (T(6) * pt[1] + pt[k2]);
            const T S = T(3) * q[k0] + (T(6) * q[1] + q[k2]);
            const T num = GP == 0 ? PT - T(3) * PU : PT + T(3) * PU;
            return c + num * rcp_mixed(T(12) * S);
        } else {
            constexpr double R0 = WENO_D[1][0] / WENO_D[1][1], R2 = WENO_D[1][2] / WENO_D[1][1];
            constexpr double AL = 6.7082039324993690892;  // 3 sqrt(5)
            const T PU = T(R0) * pu[k0] + (pu[1] + T(R2) * pu[k2]);
            const T PT = T(R0) * pt[k0] + (pt[1] + T(R2) * pt[k2]);
            const T S = T(R0) * q[k0] + (q[1] + T(R2) * q[k2]);
            const T num = GP == 1 ? PT + T(AL) * PU : PT - T(AL) * PU;
            return c + num * rcp_mixed(T(-60) * S);
        }
    }
};

// the 5-point reconstruction of a policy: the reference's arithmetic in the all-FP64 policy,
// the linear form in the mixed one (ZC: the centre value is a literal zero)
template <bool EXACT, class T, int SCHEME, bool ZC = false>
using WenoSel = typename std::conditional<EXACT, Weno5<T, SCHEME>, WenoLin<T, SCHEME, ZC>>::type;

// ------------------------------------------------------------------ 7-point kernels
// WENO-EXP (include/Utility.hpp:291-472) and the THINC kernel of the primitive3 form
// (include/Utility.hpp:473-663), in double.  s = {u_mmm, u_mm, u_m, u, u_p, u_pp, u_ppp}.

// x^3 rounded once (the reference's pow(x, 3)): x^2 = p + e exactly, p x = hi + lo exactly
WFV_HD double cube_rn(double x) {
    const double p = x * x;
    const double e = fma(x, x, -p);
    const double hi = p * x;
    const double lo = fma(p, x, -hi);
    return hi + (lo + e * x);
}

// FAST (mixed policy): the same formulas with each quotient a product by a Newton-refined
// reciprocal, hardware min/max and a plain cube; the all-FP64 policy keeps the reference's
// IEEE operations.

// fitted linear weights, cubic in l2h2 (include/Utility.hpp:347-380); pow(l, 2) is l*l
template <int GP, bool FAST = false>
WFV_HD void wexp_d(double l, double &d0, double &d1, double &d2) {
    const double l2 = l * l, l3 = FAST ? l2 * l : cube_rn(l);
    if constexpr (GP == 0) {
        d0 = 0.3 - 0.057142858 * l + 0.0071428571 * l2 - 0.00077200577 * l3;
        d1 = 0.6 + 0.078571429 * l - 0.0101190476 * l2 + 0.00112012987 * l3;
        d2 = 0.1 - 0.021428571 * l + 0.0029761905 * l2 - 0.00034812410 * l3;
    } else if constexpr (GP == 1) {
        d0 = 0.16108042 - 0.035187607 * l + 0.0049466098 * l2 - 0.00058211525 * l3;
        d1 = 0.58636364 + 0.087692641 * l - 0.0120401515 * l2 + 0.00139436871 * l3;
        d2 = 0.25255594 - 0.052505034 * l + 0.0070935417 * l2 - 0.00081225346 * l3;
    } else if constexpr (GP == 2) {
        d0 = 0.25255594 - 0.052505034 * l + 0.0070935417 * l2 - 0.00081225346 * l3;
        d1 = 0.58636364 + 0.087692641 * l - 0.0120401515 * l2 + 0.00139436871 * l3;
        d2 = 0.16108042 - 0.035187607 * l + 0.0049466098 * l2 - 0.00058211525 * l3;
    } else {
        d0 = 0.1 - 0.021428571 * l + 0.0029761905 * l2 - 0.00034812410 * l3;
        d1 = 0.6 + 0.078571429 * l - 0.0101190476 * l2 + 0.00112012987 * l3;
        d2 = 0.3 - 0.057142858 * l + 0.0071428571 * l2 - 0.00077200577 * l3;
    }
}

// l2h2 = dif5/dif3 of the 6 cells biased toward the node (gp < 2: u_mmm..u_pp; else
// u_mm..u_ppp), 0 where |dif3| < dx^4 (include/Utility.hpp:297-321)
template <bool FAST = false>
WFV_HD double wexp_l2h2(const double *t, double dx4, bool clamp) {
    const double dif5 = (-t[0] + 5.0 * t[1] - 10.0 * t[2] + 10.0 * t[3] - 5.0 * t[4] + t[5]);
    const double d6 = t[0] - 11.0 * t[1] + 28.0 * t[2] - 28.0 * t[3] + 11.0 * t[4] - t[5];
    const double dif3 = FAST ? d6 * (1.0 / 6.0) : d6 / 6.0;
    if (fabs(dif3) < dx4) return 0.0;
    double l;
    if constexpr (FAST) l = dif5 * rcp_nr(dif3);
    else l = dif5 / dif3;
    if (clamp && fabs(l) > 1) l = 1;
    return l;
}

// the inner five cells scaled to [0, 1] by their min and max (include/Utility.hpp:393-401)
template <bool FAST = false>
WFV_HD void wexp_norm(const double s[7], double n[5]) {
    if constexpr (FAST) {
        const double mx = fmax(fmax(fmax(fmax(s[1], s[2]), s[3]), s[4]), s[5]);
        const double mn = fmin(fmin(fmin(fmin(s[1], s[2]), s[3]), s[4]), s[5]);
        const double r = rcp_nr(mx - mn + 1e-20);
        for (int k = 0; k < 5; ++k) n[k] = (s[1 + k] - mn) * r;
    } else {
        const double mx = stdMax(stdMax(stdMax(s[1], s[2]), stdMax(s[3], s[4])), s[5]);
        const double mn = stdMin(stdMin(stdMin(s[1], s[2]), stdMin(s[3], s[4])), s[5]);
        for (int k = 0; k < 5; ++k) n[k] = (s[1 + k] - mn) / (mx - mn + 1e-20);
    }
}

// weights -> value: w_k = d_k (1 + a_k), normalised, sum w_k h_k (candidates as WENO5)
template <int GP, bool FAST = false>
WFV_HD double wexp_combine(const double s[7], double d0, double d1, double d2, double a0, double a1, double a2) {
    const double h0 = WENO_H[GP][0][0] * s[1] + WENO_H[GP][0][1] * s[2] + WENO_H[GP][0][2] * s[3];
    const double h1 = WENO_H[GP][1][0] * s[2] + WENO_H[GP][1][1] * s[3] + WENO_H[GP][1][2] * s[4];
    const double h2 = WENO_H[GP][2][0] * s[3] + WENO_H[GP][2][1] * s[4] + WENO_H[GP][2][2] * s[5];
    double w0 = d0 * a0, w1 = d1 * a1, w2 = d2 * a2;
    const double S = w0 + w1 + w2;
    if constexpr (FAST) return (w0 * h0 + w1 * h1 + w2 * h2) * rcp_nr(S);
    w0 = w0 / S; w1 = w1 / S; w2 = w2 / S;
    return w0 * h0 + w1 * h1 + w2 * h2;
}

// wenoExpthreconstruction: the node-independent parts (indicators, tau) once per stencil
template <bool FAST = false>
struct WenoExp7 {
    const double *s;
    double dx4, a0, a1, a2, l_lo, l_hi;
    WFV_HD WenoExp7(const double *st, double dx4_) : s(st), dx4(dx4_) {
        const double theta = 1;
        const double eps = 1e-16;
        if constexpr (FAST) {
            // The indicators and tau are differences, of degree 1 in the stencil: on the values
            // less their minimum (the reference's numerators) they are den times the
            // normalised ones (den = max - min + 1e-20), so a_k = 1 + 2 tau^2 / (b_k^2 + eps)
            // keeps its value with eps den^2 in place of eps, and the division by den goes.
            // The a_k are then scaled by the product of their three denominators (>= 1e-56
            // each), which the node's weight ratio cancels.
            const double mx = fmax(fmax(fmax(fmax(s[1], s[2]), s[3]), s[4]), s[5]);
            const double mn = fmin(fmin(fmin(fmin(s[1], s[2]), s[3]), s[4]), s[5]);
            const double den = mx - mn + 1e-20;
            const double e2 = (eps * den) * den;
            const double u[5] = {s[1] - mn, s[2] - mn, s[3] - mn, s[4] - mn, s[5] - mn};
            const double b0 = 0.5 * fabs(u[0] - 4 * u[1] + 3 * u[2]) + fabs(u[0] - 2 * u[1] + u[2]);
            const double b1 = 0.5 * fabs(u[1] - u[3]) + fabs(u[1] - 2 * u[2] + u[3]);
            const double b2 = 0.5 * fabs(3 * u[2] - 4 * u[3] + u[4]) + fabs(u[2] - 2 * u[3] + u[4]);
            const double tau = u[0] - 4.0 * u[1] + 6.0 * u[2] - 4.0 * u[3] + u[4];
            const double t2 = 2 * (tau * tau);
            const double E0 = b0 * b0 + e2, E1 = b1 * b1 + e2, E2 = b2 * b2 + e2;
            const double P01 = E0 * E1, PI = P01 * E2;
            a0 = t2 * (E1 * E2) + PI;
            a1 = t2 * (E0 * E2) + PI;
            a2 = t2 * P01 + PI;
        } else {
            double n[5];
            wexp_norm<FAST>(s, n);
            const double b0 = theta * 1.0 / 2.0 * fabs(n[0] - 4 * n[1] + 3 * n[2]) + fabs(n[0] - 2 * n[1] + n[2]);
            const double b1 = theta * 1.0 / 2.0 * fabs(n[1] - n[3]) + fabs(n[1] - 2 * n[2] + n[3]);
            const double b2 = theta * 1.0 / 2.0 * fabs(3 * n[2] - 4 * n[3] + n[4]) + fabs(n[2] - 2 * n[3] + n[4]);
            const double tau = n[0] - 4.0 * n[1] + 6.0 * n[2] - 4.0 * n[3] + n[4];
            a0 = 1.0 + 2 * (tau * tau) / ((b0 * b0) + eps);
            a1 = 1.0 + 2 * (tau * tau) / ((b1 * b1) + eps);
            a2 = 1.0 + 2 * (tau * tau) / ((b2 * b2) + eps);
        }
        l_lo = wexp_l2h2<FAST>(s, dx4, true);
        l_hi = wexp_l2h2<FAST>(s + 1, dx4, true);
    }
    template <int GP>
    WFV_HD double node() const {
        double d0, d1, d2;
        wexp_d<GP, FAST>(GP >= 2 ? l_hi : l_lo, d0, d1, d2);
        return wexp_combine<GP, FAST>(s, d0, d1, d2, a0, a1, a2);
    }
};

// wenoExpWithThincReconstruction: a THINC profile in cells with 1e-6 < u < 1-1e-6, else WENO-Z
// indicators (eps 1e-16) with the unclamped WENO-EXP linear weights
template <bool FAST = false>
struct ExpThinc7 {
    const double *s;
    bool thinc, flat;
    double a0, a1, a2, l_lo, l_hi;      // WENO branch
    double qmin, qmax, theta, xcc;      // THINC branch
    WFV_DEV ExpThinc7(const double *st, double dx4) : s(st) {
        const double um = s[2], uc = s[3], up = s[4];
        thinc = (1e-6 < uc) & (uc < 1 - 1e-6);
        l_lo = wexp_l2h2<FAST>(s, dx4, false);
        l_hi = wexp_l2h2<FAST>(s + 1, dx4, false);
        // FAST: the indicators on the values less their minimum (den^2 times the normalised
        // ones) with eps den^2 in place of eps, and the a_k scaled by the product of their
        // denominators (>= 1e-56 each), as in WenoExp7
        double n[5];
        double e2 = 0.0;
        if constexpr (FAST) {
            const double mx = fmax(fmax(fmax(fmax(s[1], s[2]), s[3]), s[4]), s[5]);
            const double mn = fmin(fmin(fmin(fmin(s[1], s[2]), s[3]), s[4]), s[5]);
            const double den = mx - mn + 1e-20;
            e2 = (1.0e-16 * den) * den;
            for (int k = 0; k < 5; ++k) n[k] = s[1 + k] - mn;
        } else {
            wexp_norm<FAST>(s, n);
        }
        double t;
        t = n[0] - 4 * n[1] + 3 * n[2];
        double b0 = 0.25 * (t * t);
        t = n[0] - 2 * n[1] + n[2];
        b0 = b0 + 13.0 / 12.0 * (t * t);
        t = n[1] - n[3];
        double b1 = 0.25 * (t * t);
        t = n[1] - 2 * n[2] + n[3];
        b1 = b1 + 13.0 / 12.0 * (t * t);
        t = 3 * n[2] - 4 * n[3] + n[4];
        double b2 = 0.25 * (t * t);
        t = n[2] - 2 * n[3] + n[4];
        b2 = b2 + 13.0 / 12.0 * (t * t);
        const double eps = 1.0e-16;
        const double tau = fabs(b0 - b2);
        if constexpr (FAST) {
            const double E0 = b0 + e2, E1 = b1 + e2, E2 = b2 + e2;
            const double P01 = E0 * E1, PI = P01 * E2;
            a0 = tau * (E1 * E2) + PI;
            a1 = tau * (E0 * E2) + PI;
            a2 = tau * P01 + PI;
        } else {
            a0 = 1 + tau / (b0 + eps);
            a1 = 1 + tau / (b1 + eps);
            a2 = 1 + tau / (b2 + eps);
        }
        // THINC (include/Utility.hpp:633-661); exp(+-beta) as glibc rounds them
        qmin = stdMin(um, up);
        qmax = stdMax(um, up) - qmin;
        theta = (up - uc >= 0) ? 1.0 : -1.0;  // sc_math::Sign (third_party/scTools/MyMath.h:47-50)
        const double beta = 1.5;
        const double exp_beta = 0x1.1ed3fe64fc541p+2, exp_mbeta = 0x1.c8f87724b5c1dp-3;
        if constexpr (FAST) {
            // mixed policy: the profile's exponential and logarithm in float (hardware
            // exp2/log2; the node values are volume fractions in [0, 1], where their ~1e-7
            // absolute error is float rounding)
            const double C = (uc - qmin + 1e-20) * rcp_nr(qmax + 1e-20);
            const float B = __expf(float(theta * beta * (2.0 * C - 1.0)));
            const float xc = (float(exp_beta) - B) * __builtin_amdgcn_rcpf(B - float(exp_mbeta));
            xcc = double(__logf(xc) * float(1.0 / (2.0 * beta)));
        } else {
            const double C = (uc - qmin + 1e-20) / (qmax + 1e-20);
            const double B = exp(theta * beta * (2.0 * C - 1.0));
            const double xc = (exp_beta - B) / (B - exp_mbeta);
            xcc = log(xc) / (2.0 * beta);
        }
        flat = (uc - um) * (up - uc) < 0;
    }
    template <int GP>
    WFV_DEV double node() const {
        if (thinc) {
            // Gauss-Lobatto-4 points (third_party/scTools/GaussQuadratureInRef.hpp:252-262)
            constexpr double glp[4] = {-1, -0.447213595499957939281834733746, 0.447213595499957939281834733746, 1};
            if (flat) return s[3];
            const double beta = 1.5;
            if constexpr (FAST) {
                // tanh x = 1 - 2 / (exp(2x) + 1), in float (exact limits +-1 at +-inf)
                const float x = float(beta * ((glp[GP] + 1) / 2 - xcc));
                const float th = 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * x) + 1.0f);
                return 0.5 * qmax * (1 + theta * double(th)) + qmin;
            }
            return 0.5 * qmax * (1 + theta * tanh(beta * ((glp[GP] + 1) / 2 - xcc))) + qmin;
        }
        double d0, d1, d2;
        wexp_d<GP, FAST>(GP >= 2 ? l_hi : l_lo, d0, d1, d2);
        return wexp_combine<GP, FAST>(s, d0, d1, d2, a0, a1, a2);
    }
};

}  // namespace wfv
