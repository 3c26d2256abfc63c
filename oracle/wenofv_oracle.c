/*
 * wenofv_oracle.c -- CPU FP64 restatement of the reference WENOFV time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP solver; see wenofv_oracle.h).
 * Built with -O2 -ffp-contract=off so that no multiply-add is fused, as in the
 * reference build (makefile:8 compiles for the x86-64 baseline, which has no FMA).
 *
 * Single-domain formulation: the reference splits x over 4 MPI ranks and exchanges
 * three ghost columns of cell averages (src/WENOFV.cpp:2175-2248) and one ring of
 * reconstructed face values (src/WENOFV.cpp:1558-1668).  Here every cell of the padded
 * domain (3 ghost layers, filled by the boundary map) is reconstructed and limited by
 * the same code, which yields the identical face values the neighbour rank would have
 * sent, so the result equals the 4-rank run bit for bit.
 */
#include "wenofv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NV 4
#define NG 4

/* Gauss-Lobatto-4 weights on [-1,1] (third_party/scTools/GaussQuadratureInRef.hpp:252-262) */
static const double GLW[4] = {0.166666666666666666666666666667, 0.833333333333333333333333333333,
                              0.833333333333333333333333333333, 0.166666666666666666666666666667};

/* sc_math::Min (third_party/scTools/MyMath.h:52-55) */
static inline double scMin(double a, double b) { return a < b ? a : b; }
/* std::min / std::max as used by the reference (libstdc++: (b<a)?b:a and (a<b)?b:a) */
static inline double stdMin(double a, double b) { return (b < a) ? b : a; }
static inline double stdMax(double a, double b) { return (a < b) ? b : a; }

/* ---------------------------------------------------------------- WENO point kernels */

/* linear weights d_k per Gauss-Lobatto node (include/Utility.hpp:12-39, 205-230) */
static void weno_d(int gp, double *d0, double *d1, double *d2) {
    switch (gp) {
    case 0: *d0 = 0.3; *d1 = 0.6; *d2 = 0.1; break;
    case 1: *d0 = (91.0 - 9.0 * sqrt(5.0)) / 440.0; *d1 = 129.0 / 220.0;
            *d2 = (91.0 + 9.0 * sqrt(5.0)) / 440.0; break;
    case 2: *d0 = (91.0 + 9.0 * sqrt(5.0)) / 440.0; *d1 = 129.0 / 220.0;
            *d2 = (91.0 - 9.0 * sqrt(5.0)) / 440.0; break;
    default: *d0 = 0.1; *d1 = 0.6; *d2 = 0.3; break;
    }
}

/* candidate values h_k at node gp (include/Utility.hpp:64-90, 233-259) */
static void weno_h(int gp, double a, double b, double c, double d, double e,
                   double *h0, double *h1, double *h2) {
    const double s5 = sqrt(5.0);
    switch (gp) {
    case 0:
        *h0 = -1.0 / 6.0 * a + 5.0 / 6.0 * b + 1.0 / 3.0 * c;
        *h1 = 1.0 / 3.0 * b + 5.0 / 6.0 * c - 1.0 / 6 * d;
        *h2 = 11.0 / 6.0 * c - 7.0 / 6.0 * d + 1.0 / 3.0 * e;
        break;
    case 1:
        *h0 = (-1.0 / 60.0 - s5 / 20.0) * a + (1.0 / 30.0 + s5 / 5.0) * b + (59.0 / 60.0 - 3 * s5 / 20.0) * c;
        *h1 = (-1.0 / 60.0 + s5 / 20.0) * b + (31.0 / 30.0) * c + (-1.0 / 60 - s5 / 20.0) * d;
        *h2 = (59.0 / 60.0 + 3 * s5 / 20) * c + (1.0 / 30.0 - s5 / 5.0) * d + (-1.0 / 60.0 + s5 / 20.0) * e;
        break;
    case 2:
        *h0 = (-1.0 / 60.0 + s5 / 20.0) * a + (1.0 / 30.0 - s5 / 5.0) * b + (59.0 / 60.0 + 3 * s5 / 20) * c;
        *h1 = (-1.0 / 60 - s5 / 20.0) * b + (31.0 / 30.0) * c + (-1.0 / 60.0 + s5 / 20.0) * d;
        *h2 = (59.0 / 60.0 - 3 * s5 / 20.0) * c + (1.0 / 30.0 + s5 / 5.0) * d + (-1.0 / 60.0 - s5 / 20.0) * e;
        break;
    default:
        *h0 = 1.0 / 3.0 * a - 7.0 / 6.0 * b + 11.0 / 6.0 * c;
        *h1 = -1.0 / 6 * b + 5.0 / 6.0 * c + 1.0 / 3.0 * d;
        *h2 = 1.0 / 3.0 * c + 5.0 / 6.0 * d - 1.0 / 6.0 * e;
        break;
    }
}

/* Jiang-Shu smoothness indicators (include/Utility.hpp:48-50) */
static void weno_beta(double a, double b, double c, double d, double e,
                      double *b0, double *b1, double *b2) {
    double t;
    t = a - 4 * b + 3 * c;       *b0 = 0.25 * (t * t);
    t = a - 2 * b + c;           *b0 = *b0 + 13.0 / 12.0 * (t * t);
    t = b - d;                   *b1 = 0.25 * (t * t);
    t = b - 2 * c + d;           *b1 = *b1 + 13.0 / 12.0 * (t * t);
    t = 3 * c - 4 * d + e;       *b2 = 0.25 * (t * t);
    t = c - 2 * d + e;           *b2 = *b2 + 13.0 / 12.0 * (t * t);
}

/* WENO5threconstruction, single-node form (include/Utility.hpp:5-95) */
static double weno_js(double a, double b, double c, double d, double e, int gp) {
    double d0, d1, d2, h0, h1, h2, b0, b1, b2;
    weno_d(gp, &d0, &d1, &d2);
    weno_beta(a, b, c, d, e, &b0, &b1, &b2);
    const double eps = 1.0e-6;
    double w0 = d0 / ((eps + b0) * (eps + b0));
    double w1 = d1 / ((eps + b1) * (eps + b1));
    double w2 = d2 / ((eps + b2) * (eps + b2));
    const double sum = w0 + w1 + w2;
    w0 = w0 / sum; w1 = w1 / sum; w2 = w2 / sum;
    weno_h(gp, a, b, c, d, e, &h0, &h1, &h2);
    return w0 * h0 + w1 * h1 + w2 * h2;
}

/* WENO5Zthreconstruction (include/Utility.hpp:193-290): beta on the min-max normalised
   stencil, candidates on the raw stencil, tau = |beta0-beta2|, eps = 1e-10. */
static double weno_z(double a, double b, double c, double d, double e, int gp) {
    double d0, d1, d2, p0, p1, p2, b0, b1, b2;
    weno_d(gp, &d0, &d1, &d2);
    weno_h(gp, a, b, c, d, e, &p0, &p1, &p2);
    const double mx = stdMax(stdMax(stdMax(a, b), stdMax(e, d)), c);
    const double mn = stdMin(stdMin(stdMin(a, b), stdMin(e, d)), c);
    const double a_ = (a - mn) / (mx - mn + 1e-20);
    const double b_ = (b - mn) / (mx - mn + 1e-20);
    const double c_ = (c - mn) / (mx - mn + 1e-20);
    const double d_ = (d - mn) / (mx - mn + 1e-20);
    const double e_ = (e - mn) / (mx - mn + 1e-20);
    weno_beta(a_, b_, c_, d_, e_, &b0, &b1, &b2);
    const double eps = 1.0e-10;
    const double tau = fabs(b0 - b2);
    double w0 = d0 * (1 + tau / (b0 + eps));
    double w1 = d1 * (1 + tau / (b1 + eps));
    double w2 = d2 * (1 + tau / (b2 + eps));
    const double sum = w0 + w1 + w2;
    w0 = w0 / sum; w1 = w1 / sum; w2 = w2 / sum;
    return w0 * p0 + w1 * p1 + w2 * p2;
}

/* ------------------------------------------------ 7-point kernels (WENO-EXP, THINC) */

/* fitted linear weights of WENO-EXP, cubic in l2h2 (include/Utility.hpp:347-380, 545-578) */
static void wexp_d(int gp, double l, double *d0, double *d1, double *d2) {
    switch (gp) {
    case 0:
        *d0 = 0.3 - 0.057142858 * l + 0.0071428571 * pow(l, 2) - 0.00077200577 * pow(l, 3);
        *d1 = 0.6 + 0.078571429 * l - 0.0101190476 * pow(l, 2) + 0.00112012987 * pow(l, 3);
        *d2 = 0.1 - 0.021428571 * l + 0.0029761905 * pow(l, 2) - 0.00034812410 * pow(l, 3);
        break;
    case 1:
        *d0 = 0.16108042 - 0.035187607 * l + 0.0049466098 * pow(l, 2) - 0.00058211525 * pow(l, 3);
        *d1 = 0.58636364 + 0.087692641 * l - 0.0120401515 * pow(l, 2) + 0.00139436871 * pow(l, 3);
        *d2 = 0.25255594 - 0.052505034 * l + 0.0070935417 * pow(l, 2) - 0.00081225346 * pow(l, 3);
        break;
    case 2:
        *d0 = 0.25255594 - 0.052505034 * l + 0.0070935417 * pow(l, 2) - 0.00081225346 * pow(l, 3);
        *d1 = 0.58636364 + 0.087692641 * l - 0.0120401515 * pow(l, 2) + 0.00139436871 * pow(l, 3);
        *d2 = 0.16108042 - 0.035187607 * l + 0.0049466098 * pow(l, 2) - 0.00058211525 * pow(l, 3);
        break;
    default:
        *d0 = 0.1 - 0.021428571 * l + 0.0029761905 * pow(l, 2) - 0.00034812410 * pow(l, 3);
        *d1 = 0.6 + 0.078571429 * l - 0.0101190476 * pow(l, 2) + 0.00112012987 * pow(l, 3);
        *d2 = 0.3 - 0.057142858 * l + 0.0071428571 * pow(l, 2) - 0.00077200577 * pow(l, 3);
        break;
    }
}

/* fifth and third differences of the 6 cells upwind-biased toward node gp
   (include/Utility.hpp:297-307, 486-496); s = {u_mmm, ..., u_ppp} */
static void wexp_dif(const double s[7], int gp, double *dif5, double *dif3) {
    const double *t = gp >= 2 ? s + 1 : s;
    *dif5 = (-t[0] + 5.0 * t[1] - 10.0 * t[2] + 10.0 * t[3] - 5.0 * t[4] + t[5]);
    *dif3 = (t[0] - 11.0 * t[1] + 28.0 * t[2] - 28.0 * t[3] + 11.0 * t[4] - t[5]) / 6.0;
}

/* min-max normalisation of the inner 5 cells (include/Utility.hpp:393-401) */
static void wexp_norm(const double s[7], double n[5]) {
    const double mx = stdMax(stdMax(stdMax(s[1], s[2]), stdMax(s[3], s[4])), s[5]);
    const double mn = stdMin(stdMin(stdMin(s[1], s[2]), stdMin(s[3], s[4])), s[5]);
    for (int k = 0; k < 5; ++k) n[k] = (s[1 + k] - mn) / (mx - mn + 1e-20);
}

/* wenoExpthreconstruction (include/Utility.hpp:291-472): WENO-EXP with l2h2 from the
   fifth/third difference ratio (clamped to 1 in magnitude, 0 where the third difference is
   below dx^4) and the exponential-type indicators on the normalised stencil */
double orc_weno_exp(const double s[7], int gp, double dx) {
    double dif5, dif3;
    wexp_dif(s, gp, &dif5, &dif3);
    double l2h2 = 0.0;
    if (!(fabs(dif3) < pow(dx, 4))) {
        l2h2 = dif5 / dif3;
        if (fabs(l2h2) > 1) l2h2 = 1;
    }
    double h0, h1, h2, d0, d1, d2, n[5];
    weno_h(gp, s[1], s[2], s[3], s[4], s[5], &h0, &h1, &h2);
    wexp_d(gp, l2h2, &d0, &d1, &d2);
    wexp_norm(s, n);
    const double theta = 1;
    const double beta0 = theta * 1.0 / 2.0 * fabs(n[0] - 4 * n[1] + 3 * n[2]) + fabs(n[0] - 2 * n[1] + n[2]);
    const double beta1 = theta * 1.0 / 2.0 * fabs(n[1] - n[3]) + fabs(n[1] - 2 * n[2] + n[3]);
    const double beta2 = theta * 1.0 / 2.0 * fabs(3 * n[2] - 4 * n[3] + n[4]) + fabs(n[2] - 2 * n[3] + n[4]);
    const double tau = n[0] - 4.0 * n[1] + 6.0 * n[2] - 4.0 * n[3] + n[4];
    const double epsilon = 1e-16;
    double w0 = d0 * (1.0 + 2 * pow(tau, 2.0) / (pow(beta0, 2.0) + epsilon));
    double w1 = d1 * (1.0 + 2 * pow(tau, 2.0) / (pow(beta1, 2.0) + epsilon));
    double w2 = d2 * (1.0 + 2 * pow(tau, 2.0) / (pow(beta2, 2.0) + epsilon));
    const double S = w0 + w1 + w2;
    w0 /= S; w1 /= S; w2 /= S;
    return w0 * h0 + w1 * h1 + w2 * h2;
}

/* Gauss-Lobatto-4 points on [-1,1] (third_party/scTools/GaussQuadratureInRef.hpp:252-262) */
static const double GLP[4] = {-1, -0.447213595499957939281834733746, 0.447213595499957939281834733746, 1};

/* wenoExpWithThincReconstruction (include/Utility.hpp:473-663): in cells whose value lies
   strictly inside (1e-6, 1-1e-6) a THINC tanh profile through the 3-cell stencil, elsewhere
   WENO-Z indicators (eps 1e-16) with the WENO-EXP linear weights (l2h2 not clamped) */
double orc_weno_exp_thinc(const double s[7], int gp, double dx) {
    double dif5, dif3;
    wexp_dif(s, gp, &dif5, &dif3);
    double l2h2 = 0.0;
    if (!(fabs(dif3) < pow(dx, 4))) l2h2 = dif5 / dif3;
    const double um = s[2], uc = s[3], up = s[4];
    if (!((1e-6 < uc) & (uc < 1 - 1e-6))) {
        double h0, h1, h2, d0, d1, d2, n[5];
        weno_h(gp, s[1], s[2], s[3], s[4], s[5], &h0, &h1, &h2);
        wexp_d(gp, l2h2, &d0, &d1, &d2);
        wexp_norm(s, n);
        const double beta0 = 0.25 * pow(n[0] - 4 * n[1] + 3 * n[2], 2) + 13.0 / 12.0 * pow(n[0] - 2 * n[1] + n[2], 2);
        const double beta1 = 0.25 * pow(n[1] - n[3], 2) + 13.0 / 12.0 * pow(n[1] - 2 * n[2] + n[3], 2);
        const double beta2 = 0.25 * pow(3 * n[2] - 4 * n[3] + n[4], 2) + 13.0 / 12.0 * pow(n[2] - 2 * n[3] + n[4], 2);
        const double eps = 1.0e-16;
        const double tau = fabs(beta0 - beta2);
        double w0 = d0 * (1 + tau / (beta0 + eps));
        double w1 = d1 * (1 + tau / (beta1 + eps));
        double w2 = d2 * (1 + tau / (beta2 + eps));
        const double S = w0 + w1 + w2;
        w0 /= S; w1 /= S; w2 /= S;
        return w0 * h0 + w1 * h1 + w2 * h2;
    }
    const double Qmin = stdMin(um, up);
    const double Qmax = stdMax(um, up) - Qmin;
    const double theta = (up - uc >= 0) ? 1 : -1; /* sc_math::Sign (MyMath.h:47-50) */
    const double beta = 1.5;
    const double C = (uc - Qmin + 1e-20) / (Qmax + 1e-20);
    const double B = exp(theta * beta * (2.0 * C - 1.0));
    const double xc = (exp(beta) - B) / (B - exp(-beta));
    const double xcc = log(xc) / (2.0 * beta);
    if ((uc - um) * (up - uc) < 0) return uc;
    return 0.5 * Qmax * (1 + theta * tanh(beta * ((GLP[gp] + 1) / 2 - xcc))) + Qmin;
}

/* CWENOFV::useWENO (src/WENOFV.cpp:1873-1893) on the 7-point stencil s = {u_mmm, ..., u_ppp};
   dx is the reference's m_deltaX (WENO-EXP's smoothness threshold) */
double orc_weno7(int scheme, const double s[7], int gp, double dx) {
    if (scheme == 2) return orc_weno_exp(s, gp, dx);
    return scheme == 1 ? weno_z(s[1], s[2], s[3], s[4], s[5], gp) : weno_js(s[1], s[2], s[3], s[4], s[5], gp);
}

/* CWENOFV::useWENO dispatch (src/WENOFV.cpp:1873-1893); s = {u_mm, u_m, u, u_p, u_pp} */
double orc_weno(int scheme, const double s[5], int gp) {
    return scheme == 1 ? weno_z(s[0], s[1], s[2], s[3], s[4], gp)
                       : weno_js(s[0], s[1], s[2], s[3], s[4], gp);
}

/* ---------------------------------------------------------------- equation of state */

/* TwoPhaseEquation::getVolumeFrac (include/Equations.hpp:84-99) */
double orc_volume_frac(const orc_params *p, const double U[4]) {
    const double z1rho1 = U[0], z2rho2 = U[1];
    const double c1_2 = p->sound1 * p->sound1;
    const double c2_2 = p->sound2 * p->sound2;
    const double q = p->rho20 * c2_2 - p->rho10 * c1_2;
    const double qs = z2rho2 * c2_2 - z1rho1 * c1_2;
    const double r = ((q - qs) + sqrt((q - qs) * (q - qs) + 4.0 * z1rho1 * c1_2 * z2rho2 * c2_2)) /
                     (2.0 * z2rho2 * c2_2);
    return r / (1.0 + r);
}

/* TwoPhaseEquation::getPressure (include/Equations.hpp:101-124); g_preb is not added */
double orc_pressure(const orc_params *p, const double U[4]) {
    const double z1 = orc_volume_frac(p, U);
    const double z2 = 1.0 - z1;
    const double c1_2 = p->sound1 * p->sound1;
    const double c2_2 = p->sound2 * p->sound2;
    const double pre1 = c1_2 * (U[0] / z1 - p->rho10);
    const double pre2 = c2_2 * (U[1] / z2 - p->rho20);
    return (z1 < z2) ? pre2 : pre1;
}

/* TwoPhaseEquation::getSoundSpeed, Wood's formula (include/Equations.hpp:126-146) */
double orc_sound_speed(const orc_params *p, const double U[4]) {
    const double z1 = orc_volume_frac(p, U);
    const double z2 = 1.0 - z1;
    const double rho = U[0] + U[1];
    const double rho1 = U[0] / z1;
    const double rho2 = U[1] / z2;
    const double c1_2 = p->sound1 * p->sound1;
    const double c2_2 = p->sound2 * p->sound2;
    const double sc = sqrt(rho * (z1 / (rho1 * c1_2) + z2 / (rho2 * c2_2)));
    return 1.0 / sc;
}

/* TwoPhaseEquation::getMaxEigenValue (include/Equations.hpp:149-156) */
double orc_max_eig(const orc_params *p, const double U[4], double nx, double ny) {
    const double sc = orc_sound_speed(p, U);
    const double u = U[2] / (U[0] + U[1]);
    const double v = U[3] / (U[0] + U[1]);
    return fabs(u * nx + v * ny) + sc;
}

/* TwoPhaseEquation::getPhyFlux (include/Equations.hpp:158-186); branch on nx > ny */
void orc_phy_flux(const orc_params *p, const double U[4], double F[4], double nx, double ny) {
    const double rho = U[0] + U[1];
    const double u = U[2] / rho;
    const double v = U[3] / rho;
    const double pre = orc_pressure(p, U);
    if (nx > ny) {
        F[0] = U[0] * u; F[1] = U[1] * u; F[2] = rho * (u * u) + pre; F[3] = rho * u * v;
    } else {
        F[0] = U[0] * v; F[1] = U[1] * v; F[2] = rho * v * u; F[3] = rho * (v * v) + pre;
    }
}

/* getLLFRiemannFlux (include/Equations.hpp:363-378) */
static void riemann_llf(const orc_params *p, const double UL[4], const double UR[4], double F[4],
                        double nx, double ny) {
    const double wsL = orc_max_eig(p, UL, nx, ny);
    const double wsR = orc_max_eig(p, UR, nx, ny);
    const double ws = stdMax(wsL, wsR);
    double FL[4], FR[4];
    orc_phy_flux(p, UL, FL, nx, ny);
    orc_phy_flux(p, UR, FR, nx, ny);
    for (int r = 0; r < 4; ++r) F[r] = 0.5 * (FL[r] + FR[r] - ws * (UR[r] - UL[r]));
}

/* getHLLCRiemannFlux (include/Equations.hpp:381-483) */
static void riemann_hllc(const orc_params *p, const double UL[4], const double UR[4], double F[4],
                         double nx, double ny) {
    double FL[4], FR[4], Us[4];
    const double rho_L = UL[0] + UL[1];
    const double u_L = UL[2] / rho_L, v_L = UL[3] / rho_L;
    const double pre_L = orc_pressure(p, UL);
    const double vn_L = u_L * nx + v_L * ny;
    const double c_L = orc_sound_speed(p, UL);
    const double rho_R = UR[0] + UR[1];
    const double u_R = UR[2] / rho_R, v_R = UR[3] / rho_R;
    const double pre_R = orc_pressure(p, UR);
    const double vn_R = u_R * nx + v_R * ny;
    const double c_R = orc_sound_speed(p, UR);
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
    if (S_star >= 0) {
        if (S_L >= 0.0) {
            for (int r = 0; r < 4; ++r) F[r] = FL[r];
        } else {
            const double k = (S_L - vn_L) / (S_L - S_star);
            Us[0] = UL[0]; Us[1] = UL[1];
            Us[2] = rho_L * (u_L + (S_star - vn_L) * nx);
            Us[3] = rho_L * (v_L + (S_star - vn_L) * ny);
            for (int r = 0; r < 4; ++r) F[r] = S_L * (k * Us[r] - UL[r]) + FL[r];
        }
    } else {
        if (S_R <= 0) {
            for (int r = 0; r < 4; ++r) F[r] = FR[r];
        } else {
            const double k = (S_R - vn_R) / (S_R - S_star);
            Us[0] = UR[0]; Us[1] = UR[1];
            Us[2] = rho_R * (u_R + (S_star - vn_R) * nx);
            Us[3] = rho_R * (v_R + (S_star - vn_R) * ny);
            for (int r = 0; r < 4; ++r) F[r] = S_R * (k * Us[r] - UR[r]) + FR[r];
        }
    }
}

/* getRiemannFlux dispatch (include/Equations.hpp:329-360) */
void orc_riemann(const orc_params *p, const double UL[4], const double UR[4], double F[4],
                 double nx, double ny, int type) {
    if (type == 1) riemann_hllc(p, UL, UR, F, nx, ny);
    else riemann_llf(p, UL, UR, F, nx, ny);
}

/* shared quantities of the eigenvector matrices (include/Equations.hpp:191-210, 261-281) */
typedef struct { double rho, u, v, z1, z2, c, sc1_2, sc2_2; } eig_state;
static void eig_prep(const orc_params *p, const double U[4], eig_state *s) {
    s->rho = U[0] + U[1];
    s->u = U[2] / s->rho;
    s->v = U[3] / s->rho;
    s->z1 = orc_volume_frac(p, U);
    s->z2 = 1.0 - s->z1;
    s->c = orc_sound_speed(p, U);
    const double c1_2 = p->sound1 * p->sound1;
    const double c2_2 = p->sound2 * p->sound2;
    const double zrho1 = U[0] * c1_2 / (s->z1 * s->z1);
    const double zrho2 = U[1] * c2_2 / (s->z2 * s->z2);
    s->sc1_2 = c1_2 / s->z1 - c1_2 / s->z1 * zrho1 / (zrho1 + zrho2);
    s->sc2_2 = c2_2 / s->z2 - c2_2 / s->z2 * zrho2 / (zrho1 + zrho2);
}

/* getREigenMatrix (include/Equations.hpp:189-256); R[row*4+col] */
void orc_eigen_R(const orc_params *p, const double U[4], double nx, double ny, double R[16]) {
    eig_state s;
    eig_prep(p, U, &s);
    const double u = s.u, v = s.v, c = s.c, rho = s.rho;
    R[0 * 4 + 0] = U[0] / rho; R[1 * 4 + 0] = U[1] / rho;
    R[0 * 4 + 1] = -s.sc2_2 / (s.sc1_2 - s.sc2_2); R[1 * 4 + 1] = s.sc1_2 / (s.sc1_2 - s.sc2_2);
    R[2 * 4 + 1] = u; R[3 * 4 + 1] = v;
    R[0 * 4 + 2] = 0; R[1 * 4 + 2] = 0;
    R[0 * 4 + 3] = U[0] / rho; R[1 * 4 + 3] = U[1] / rho;
    if (fabs(nx) > fabs(ny)) {
        R[2 * 4 + 0] = u - c; R[3 * 4 + 0] = v;
        R[2 * 4 + 2] = 0; R[3 * 4 + 2] = 1;
        R[2 * 4 + 3] = u + c; R[3 * 4 + 3] = v;
    } else {
        R[2 * 4 + 0] = u; R[3 * 4 + 0] = v - c;
        R[2 * 4 + 2] = 1; R[3 * 4 + 2] = 0;
        R[2 * 4 + 3] = u; R[3 * 4 + 3] = v + c;
    }
}

/* getLEigenMatrix (include/Equations.hpp:259-327); L[row*4+col] */
void orc_eigen_L(const orc_params *p, const double U[4], double nx, double ny, double L[16]) {
    eig_state s;
    eig_prep(p, U, &s);
    const double u = s.u, v = s.v, c = s.c, rho = s.rho;
    const double c2 = c * c;
    const double d12 = s.sc1_2 - s.sc2_2;
    const double un = (fabs(nx) > fabs(ny)) ? u : v;  /* normal velocity of the direction */
    L[0] = (s.sc1_2 + c * un) / (2.0 * c2);
    L[1] = (s.sc2_2 + c * un) / (2.0 * c2);
    L[4] = -U[1] * d12 / (rho * c2);
    L[5] = U[0] * d12 / (rho * c2);
    L[6] = 0; L[7] = 0;
    L[12] = (s.sc1_2 - c * un) / (2.0 * c2);
    L[13] = (s.sc2_2 - c * un) / (2.0 * c2);
    if (fabs(nx) > fabs(ny)) {
        L[2] = -1.0 / (2.0 * c); L[3] = 0;
        L[8] = -v; L[9] = -v; L[10] = 0; L[11] = 1.0;
        L[14] = 1.0 / (2.0 * c); L[15] = 0;
    } else {
        L[2] = 0; L[3] = -1.0 / (2.0 * c);
        L[8] = -u; L[9] = -u; L[10] = 1.0; L[11] = 0;
        L[14] = 0; L[15] = 1.0 / (2.0 * c);
    }
}

/* Array2D * Array1D (third_party/scTools/Array.h:604-626): C[i] = sum_j A[i][j]*B[j] from 0 */
static inline void matvec(const double A[16], const double B[4], double C[4]) {
    for (int i = 0; i < 4; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 4; ++j) acc += A[i * 4 + j] * B[j];
        C[i] = acc;
    }
}

/* ---------------------------------------------------------------- limiter */

/* getPositivityPreservingTheta (src/WENOFV.cpp:1936-2000) */
double orc_bp_theta(const orc_params *p, const double U[4], const double faces[64]) {
    const double zeps = 1e-12, omega1 = 0.1;
    const double omegaR = 1.0 - 2 * omega1;
    const double a1 = orc_max_eig(p, U, 1, 0);
    const double a2 = orc_max_eig(p, U, 0, 1);
    const double mu1 = (a1 / p->dX) / (a1 / p->dX + a2 / p->dY);
    const double mu2 = (a2 / p->dY) / (a1 / p->dX + a2 / p->dY);
    double th[2];
    for (int r = 0; r < 2; ++r) {
        th[r] = 1.0;
        const double uave = U[r];
        const double eps = stdMin(zeps, uave);
        const double *L = faces + 0 * 16 + r * 4, *R = faces + 1 * 16 + r * 4;
        const double *B = faces + 2 * 16 + r * 4, *T = faces + 3 * 16 + r * 4;
        double rem = uave;
        for (int g = 0; g < NG; ++g) {
            rem -= omega1 * GLW[g] / 2 * mu2 * B[g];
            rem -= omega1 * GLW[g] / 2 * mu2 * T[g];
            rem -= omega1 * GLW[g] / 2 * mu1 * L[g];
            rem -= omega1 * GLW[g] / 2 * mu1 * R[g];
        }
        rem /= omegaR;
        double mloc = rem;
        for (int g = 0; g < NG; ++g) {
            mloc = scMin(mloc, B[g]);
            mloc = scMin(mloc, T[g]);
            mloc = scMin(mloc, L[g]);
            mloc = scMin(mloc, R[g]);
        }
        th[r] = scMin(fabs((uave - eps) / (uave - mloc)), th[r]);
    }
    return scMin(th[0], th[1]);
}

/* ---------------------------------------------------------------- geometry */

/* per-cell widths as CWENOFV::initializeSolver builds them (src/WENOFV.cpp:90-123) */
void orc_geometry(int nx, int ny, double xL, double xR, double yL, double yR, int nranks,
                  double *dxC, double *dxW, double *dyC, double *xcen, double *ycen) {
    const double dX = (xR - xL) / nx, dY = (yR - yL) / ny;
    const int ng = 3;
    if (nranks < 1 || nx % nranks) nranks = 1;
    const int nloc = nx / nranks;
    for (int g = 0; g < nx; ++g) {
        const int rank = g / nloc;
        const double procXLeft = xL + (double)(rank * nloc) * dX;
        const double x0 = procXLeft - dX * ng;
        const int ei = g - rank * nloc + ng;
        const double xl = x0 + ei * dX, xr = x0 + (ei + 1) * dX;
        const double xlw = x0 + (ei - 1) * dX, xrw = x0 + ei * dX;
        if (dxC) dxC[g] = xr - xl;
        if (dxW) dxW[g] = xrw - xlw;
        if (xcen) xcen[g] = (xl + xr) / 2;
    }
    const double y0 = (yL + 0 * ny * dY) - dY * ng;
    for (int j = 0; j < ny; ++j) {
        const int ej = j + ng;
        const double yl = y0 + ej * dY, yr = y0 + (ej + 1) * dY;
        if (dyC) dyC[j] = yr - yl;
        if (ycen) ycen[j] = (yl + yr) / 2;
    }
}

/* ---------------------------------------------------------------- stencil sweeps */

/* ghost layers of the padded state: the 7-point stencils (WENO-EXP, primitive3) of the
   ring cells reach 4 cells out; the 5-point schemes read the inner 5 of the same stencil */
#define PAD 4

typedef struct {
    const orc_params *p;
    int nx, ny, PX, PY;   /* padded extents (PAD ghost layers) */
    double *Up;           /* [PY][PX][4] padded cell averages         */
    double *Wp;           /* [PY][PX][4] reconstructed variables (primitive modes) */
    double *Y;            /* [(ny+2)][PX][4 g][4 v]: y-sweep, rows -1..ny */
    double *Fc;           /* [(ny+2)][(nx+2)][4 side][4 v][4 g]: faces, cells -1..n */
    double *th;           /* [(ny+2)][(nx+2)] */
    int err;
} work;

static inline double *UP(work *w, int i, int j) { return w->Up + ((size_t)(j + PAD) * w->PX + (i + PAD)) * 4; }
static inline double *WP(work *w, int i, int j) { return w->Wp + ((size_t)(j + PAD) * w->PX + (i + PAD)) * 4; }
static inline double *YP(work *w, int i, int j) { return w->Y + ((size_t)(j + 1) * w->PX + (i + PAD)) * 16; }
static inline double *FP(work *w, int i, int j) { return w->Fc + ((size_t)(j + 1) * (w->nx + 2) + (i + 1)) * 64; }

/* ghost-cell fill: periodic wrap as exchangeGhostCellsValue + setBoundaryAverages2
   (src/WENOFV.cpp:2175-2295, corners wrap in both directions).  The other boundary kinds
   (include/Equations.hpp:25-32) have no reference implementation: slip walls (1) mirror the
   interior and negate the normal momentum, no-slip walls (2) negate both momenta, fixed
   ghost cells (3) come from p->ghost. */
static void fill_padded(work *w, const double *U) {
    const int nx = w->nx, ny = w->ny;
    const size_t nc = (size_t)nx * ny;
    for (int j = -PAD; j < ny + PAD; ++j) {
        for (int i = -PAD; i < nx + PAD; ++i) {
            int si = i, sj = j;
            double *d = UP(w, i, j);
            const int bc = w->p->bc;
            if (bc == 3 && (i < 0 || i >= nx || j < 0 || j >= ny)) {
                /* fixed ghost cells (NEUMANN2): the padded array the caller provides */
                const size_t pc = (size_t)(nx + 2 * PAD) * (ny + 2 * PAD);
                const size_t o = (size_t)(j + PAD) * (nx + 2 * PAD) + (i + PAD);
                for (int v = 0; v < 4; ++v) d[v] = w->p->ghost[v * pc + o];
                continue;
            }
            int mu = 0, mv = 0;
            if (bc != 0) {
                if (si < 0) { si = -1 - si; mu = 1; }
                if (si >= nx) { si = 2 * nx - 1 - si; mu = 1; }
                if (sj < 0) { sj = -1 - sj; mv = 1; }
                if (sj >= ny) { sj = 2 * ny - 1 - sj; mv = 1; }
            } else {
                si = ((si % nx) + nx) % nx;
                sj = ((sj % ny) + ny) % ny;
            }
            const size_t c = (size_t)sj * nx + si;
            d[0] = U[c];
            d[1] = U[nc + c];
            d[2] = U[2 * nc + c];
            d[3] = U[3 * nc + c];
            /* slip walls negate the momentum normal to each mirror; no-slip walls (bc 2)
               negate both components once per mirror */
            const int nu = bc == 2 ? mu != mv : mu, nv = bc == 2 ? mu != mv : mv;
            if (nu) d[2] = -d[2];
            if (nv) d[3] = -d[3];
        }
    }
}

/* primitive variables of one cell (src/WENOFV.cpp:937-956 for PRIMITIVE,
   src/WENOFV.cpp:1097-1117 for PRIMITIVE2) */
static void to_primitive(const orc_params *p, const double U[4], double W[4]) {
    const double z1 = orc_volume_frac(p, U);
    if (p->recon == 2) {
        W[0] = z1;
        W[1] = U[0] / z1;
    } else {
        double pr;
        if (z1 > 0.5) pr = p->preb + p->sound1 * p->sound1 * (U[0] / z1 - p->rho10);
        else pr = p->preb + p->sound2 * p->sound2 * (U[1] / (1.0 - z1) - p->rho20);
        W[0] = z1;
        W[1] = pr;
    }
    W[2] = U[2] / (U[0] + U[1]);
    W[3] = U[3] / (U[0] + U[1]);
}

/* conservative variables at a node (src/WENOFV.cpp:1007-1032, 1171-1189) */
static void from_primitive(const orc_params *p, const double W[4], double U[4]) {
    const double z1 = W[0];
    double z1rho1, z2rho2;
    if (p->recon == 2) {
        const double rho1 = W[1];
        z2rho2 = p->sound1 * p->sound1 / (p->sound2 * p->sound2) * (rho1 - p->rho10) + p->rho20;
        z2rho2 = z2rho2 * (1.0 - z1);
        z1rho1 = z1 * rho1;
    } else {
        const double pr = W[1];
        z1rho1 = ((pr - p->preb) / (p->sound1 * p->sound1) + p->rho10) * z1;
        z2rho2 = ((pr - p->preb) / (p->sound2 * p->sound2) + p->rho20) * (1.0 - z1);
    }
    const double rho = z1rho1 + z2rho2;
    U[0] = z1rho1; U[1] = z2rho2; U[2] = rho * W[2]; U[3] = rho * W[3];
}

/* reconstruction of variable v at node gp from a 7-point stencil: useWENO for RECON 0-3;
   for PRIMITIVE3 wenoExpWithThincReconstruction on z1 and WENO5Zthreconstruction on
   p, u, v whatever SCHEME says (src/WENOFV.cpp:1287-1329) */
static double recon7(const orc_params *p, const double s[7], int gp, int v) {
    if (p->recon == 4)
        return v == 0 ? orc_weno_exp_thinc(s, gp, p->dX) : weno_z(s[1], s[2], s[3], s[4], s[5], gp);
    return orc_weno7(p->scheme, s, gp, p->dX);
}

/* y-direction sweep of one column x-offset (steps 2-4 of getFlux_*):
   Y[g*4+v] = value at y-node g of column i, row j (src/WENOFV.cpp:670-687, 796-833). */
static void sweep_y(work *w, int i, int j, double *Yo) {
    const orc_params *p = w->p;
    double s[7];
    if (p->recon == 1) {
        double L[16], R[16], Uc[7][4], Yc[4][4];
        orc_eigen_L(p, UP(w, i, j), 0, 1, L);
        for (int t = 0; t < 7; ++t) matvec(L, UP(w, i, j - 3 + t), Uc[t]);
        for (int g = 0; g < NG; ++g)
            for (int v = 0; v < NV; ++v) {
                for (int t = 0; t < 7; ++t) s[t] = Uc[t][v];
                Yc[g][v] = recon7(p, s, g, v);
            }
        orc_eigen_R(p, UP(w, i, j), 0, 1, R);
        for (int g = 0; g < NG; ++g) matvec(R, Yc[g], Yo + g * 4);
    } else {
        for (int g = 0; g < NG; ++g)
            for (int v = 0; v < NV; ++v) {
                for (int t = 0; t < 7; ++t)
                    s[t] = (p->recon == 0) ? UP(w, i, j - 3 + t)[v] : WP(w, i, j - 3 + t)[v];
                Yo[g * 4 + v] = recon7(p, s, g, v);
            }
    }
}

/* x-direction sweep, back-conversion and limiting of one cell (steps 5-7 of getFlux_*,
   src/WENOFV.cpp:690-722, 836-880; limiter src/WENOFV.cpp:1895-1935). */
static void sweep_x(work *w, int i, int j) {
    const orc_params *p = w->p;
    double Ugp[4][4][4]; /* [gx][gy][v] */
    double s[7];
    const double *Yw[7];
    for (int t = 0; t < 7; ++t) Yw[t] = YP(w, i - 3 + t, j);
    if (p->recon == 1) {
        double L[16], R[16], Yc[7][4][4], Uc[4];
        orc_eigen_L(p, UP(w, i, j), 1, 0, L);
        orc_eigen_R(p, UP(w, i, j), 1, 0, R);
        for (int t = 0; t < 7; ++t)
            for (int g = 0; g < NG; ++g) matvec(L, Yw[t] + g * 4, Yc[t][g]);
        for (int gy = 0; gy < NG; ++gy)
            for (int gx = 0; gx < NG; ++gx) {
                if (!(gx == 0 || gx == 3 || gy == 0 || gy == 3)) continue;
                for (int v = 0; v < NV; ++v) {
                    for (int t = 0; t < 7; ++t) s[t] = Yc[t][gy][v];
                    Uc[v] = recon7(p, s, gx, v);
                }
                matvec(R, Uc, Ugp[gx][gy]);
            }
    } else {
        for (int gy = 0; gy < NG; ++gy)
            for (int gx = 0; gx < NG; ++gx) {
                if (!(gx == 0 || gx == 3 || gy == 0 || gy == 3)) continue;
                for (int v = 0; v < NV; ++v) {
                    for (int t = 0; t < 7; ++t) s[t] = Yw[t][gy * 4 + v];
                    Ugp[gx][gy][v] = recon7(p, s, gx, v);
                }
                if (p->recon >= 2) {
                    double Wn[4];
                    memcpy(Wn, Ugp[gx][gy], sizeof Wn);
                    from_primitive(p, Wn, Ugp[gx][gy]);
                }
            }
    }
    double *F = FP(w, i, j);
    for (int v = 0; v < NV; ++v)
        for (int g = 0; g < NG; ++g) {
            F[0 * 16 + v * 4 + g] = Ugp[0][g][v];
            F[1 * 16 + v * 4 + g] = Ugp[3][g][v];
            F[2 * 16 + v * 4 + g] = Ugp[g][0][v];
            F[3 * 16 + v * 4 + g] = Ugp[g][3][v];
        }
    double theta = 1.0;
    if (p->bp) {
        const double *Uc = UP(w, i, j);
        theta = orc_bp_theta(p, Uc, F);
        if (theta >= 0.0 && theta <= 1.0) {
            for (int side = 0; side < 4; ++side)
                for (int v = 0; v < NV; ++v)
                    for (int g = 0; g < NG; ++g) {
                        double *x = &F[side * 16 + v * 4 + g];
                        *x = theta * (*x - Uc[v]) + Uc[v];
                    }
        } else {
            w->err |= ORC_ERR_BAD_THETA;
        }
    }
    w->th[(size_t)(j + 1) * (w->nx + 2) + (i + 1)] = theta;
}

static int work_init(work *w, const orc_params *p) {
    memset(w, 0, sizeof *w);
    w->p = p;
    w->nx = p->nx; w->ny = p->ny;
    w->PX = p->nx + 2 * PAD; w->PY = p->ny + 2 * PAD;
    w->Up = (double *)malloc(sizeof(double) * 4 * w->PX * w->PY);
    w->Wp = (double *)malloc(sizeof(double) * 4 * w->PX * w->PY);
    w->Y = (double *)malloc(sizeof(double) * 16 * w->PX * (p->ny + 2));
    w->Fc = (double *)malloc(sizeof(double) * 64 * (p->nx + 2) * (p->ny + 2));
    w->th = (double *)malloc(sizeof(double) * (p->nx + 2) * (p->ny + 2));
    return (w->Up && w->Wp && w->Y && w->Fc && w->th) ? 0 : -1;
}
static void work_free(work *w) {
    free(w->Up); free(w->Wp); free(w->Y); free(w->Fc); free(w->th);
}

/* CWENOFV::assembleRHS (src/WENOFV.cpp:568-615) for a single domain */
static int assemble(work *w, const double *U, double *rhs, double *faces_out, double *theta_out) {
    const orc_params *p = w->p;
    const int nx = p->nx, ny = p->ny;
    const size_t nc = (size_t)nx * ny;
    w->err = 0;
    fill_padded(w, U);
    if (p->recon >= 2) {
        for (int j = -PAD; j < ny + PAD; ++j)
            for (int i = -PAD; i < nx + PAD; ++i) to_primitive(p, UP(w, i, j), WP(w, i, j));
    }
    /* y-sweep: every padded column on rows -1..ny (rows -1 and ny only feed the ring cells) */
#pragma omp parallel for schedule(static)
    for (int j = -1; j <= ny; ++j)
        for (int i = -PAD; i < nx + PAD; ++i) sweep_y(w, i, j, YP(w, i, j));
    /* x-sweep + limiter: interior cells and the one-cell ring without corners */
    int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int j = -1; j <= ny; ++j) {
        work wl = *w;
        wl.err = 0;
        for (int i = -1; i <= nx; ++i) {
            const int ring_x = (i == -1 || i == nx), ring_y = (j == -1 || j == ny);
            if (ring_x && ring_y) continue;
            sweep_x(&wl, i, j);
        }
        err |= wl.err;
    }
    w->err |= err;
    /* Riemann fluxes, Gauss-Lobatto face integrals and RHS (src/WENOFV.cpp:1670-1829) */
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            double ul[4], ur[4], fl[4];
            double Fw[4], Fe[4], Gs[4], Gn[4];
            const double dy = p->dyC[j], dxc = p->dxC[i], dxw = p->dxW[i];
            /* four faces: west (cell i-1 right | cell i left), east, south, north */
            for (int face = 0; face < 4; ++face) {
                const double *A, *B;
                double nxn, nyn, width;
                double *out = face == 0 ? Fw : face == 1 ? Fe : face == 2 ? Gs : Gn;
                if (face == 0) { A = FP(w, i - 1, j) + 16; B = FP(w, i, j) + 0; nxn = 1; nyn = 0; width = dy; }
                else if (face == 1) { A = FP(w, i, j) + 16; B = FP(w, i + 1, j) + 0; nxn = 1; nyn = 0; width = dy; }
                else if (face == 2) { A = FP(w, i, j - 1) + 48; B = FP(w, i, j) + 32; nxn = 0; nyn = 1; width = dxc; }
                else { A = FP(w, i, j) + 48; B = FP(w, i, j + 1) + 32; nxn = 0; nyn = 1; width = dxc; }
                double ave[4] = {0.0, 0.0, 0.0, 0.0};
                for (int g = 0; g < NG; ++g) {
                    for (int v = 0; v < NV; ++v) { ul[v] = A[v * 4 + g]; ur[v] = B[v * 4 + g]; }
                    orc_riemann(p, ul, ur, fl, nxn, nyn, p->flux);
                    for (int v = 0; v < NV; ++v)
                        if (isnan(fl[v])) err |= ORC_ERR_NAN_FLUX;
                    const double gw = GLW[g] * width / 2.0;
                    for (int v = 0; v < NV; ++v) ave[v] += fl[v] * gw;
                }
                for (int v = 0; v < NV; ++v) out[v] = ave[v];
            }
            for (int v = 0; v < NV; ++v) {
                acc[v] += Fw[v] / dxw / dy;
                acc[v] += -Fe[v] / dxc / dy;
                acc[v] += Gs[v] / dxc / dy;
                acc[v] -= Gn[v] / dxc / dy;
            }
            /* assembleSourceTerm (src/WENOFV.cpp:1831-1872) */
            if (p->gravity != 0.0) {
                const double *Uc = UP(w, i, j);
                const double rho = Uc[0] + Uc[1];
                acc[3] -= rho * p->gravity;
            }
            const size_t c = (size_t)j * nx + i;
            for (int v = 0; v < NV; ++v) rhs[v * nc + c] = acc[v];
            if (faces_out) memcpy(faces_out + c * 64, FP(w, i, j), 64 * sizeof(double));
            if (theta_out) theta_out[c] = w->th[(size_t)(j + 1) * (nx + 2) + (i + 1)];
        }
    }
    w->err |= err;
    return w->err;
}

int orc_assemble_rhs(const orc_params *p, const double *U, double *rhs, double *faces_out,
                     double *theta_out) {
    work w;
    if (work_init(&w, p)) { work_free(&w); return -1; }
    const int e = assemble(&w, U, rhs, faces_out, theta_out);
    work_free(&w);
    return e;
}

/* CWENOFV::calculateDeltaT (src/WENOFV.cpp:474-540) */
int orc_compute_dt(const orc_params *p, const double *U, double *dt) {
    const int nx = p->nx, ny = p->ny;
    const size_t nc = (size_t)nx * ny;
    double timestep = 1.0;
    int nan = 0;
    for (size_t c = 0; c < nc; ++c) {
        const double u[4] = {U[c], U[nc + c], U[2 * nc + c], U[3 * nc + c]};
        const double ex = orc_max_eig(p, u, 1, 0);
        const double ey = orc_max_eig(p, u, 0, 1);
        const double tmp = p->cfl / (ex / p->dX + ey / p->dY);
        if (isnan(tmp)) nan = 1;
        if (tmp > timestep) continue;
        timestep = tmp;
    }
    *dt = timestep;
    return nan ? ORC_ERR_NAN_DT : 0;
}

/* CWENOFV::RunRK1/2/3 (src/WENOFV.cpp:376-473) */
static int rk_step(work *w, double *U, double *Un, double *rhs, double dt) {
    const orc_params *p = w->p;
    const size_t n = (size_t)4 * p->nx * p->ny;
    int err = 0;
    if (p->rk == 0) {
        err |= assemble(w, U, rhs, NULL, NULL);
        for (size_t k = 0; k < n; ++k) U[k] += dt * rhs[k];
        return err;
    }
    memcpy(Un, U, n * sizeof(double));
    const int ns = p->rk == 1 ? 2 : 3;
    for (int st = 0; st < ns; ++st) {
        double a, b, c;
        if (p->rk == 1) {
            if (st == 0) { a = 1.0; b = 0.0; c = 1.0; }
            else { a = 0.5; b = 0.5; c = 0.5; }
        } else {
            if (st == 0) { a = 1.0; b = 0.0; c = 1.0; }
            else if (st == 1) { a = 3.0 / 4.0; b = 1.0 / 4.0; c = 1.0 / 4.0; }
            else { a = 1.0 / 3.0; b = 2.0 / 3.0; c = 2.0 / 3.0; }
        }
        err |= assemble(w, U, rhs, NULL, NULL);
        for (size_t k = 0; k < n; ++k) U[k] = a * Un[k] + b * U[k] + c * dt * rhs[k];
    }
    return err;
}

int orc_rk_step(const orc_params *p, double *U, double dt) {
    work w;
    const size_t n = (size_t)4 * p->nx * p->ny;
    double *Un = (double *)malloc(n * sizeof(double)), *rhs = (double *)malloc(n * sizeof(double));
    if (!Un || !rhs || work_init(&w, p)) { free(Un); free(rhs); return -1; }
    const int e = rk_step(&w, U, Un, rhs, dt);
    work_free(&w); free(Un); free(rhs);
    return e;
}

/* CWENOFV::run time loop (src/WENOFV.cpp:318-364): dt = min(CFL dt, remaining) until
   |now - T| <= 1e-10; or a fixed number of unclipped CFL steps. */
int orc_run(const orc_params *p, double *U, double t_start, double t_end, int fixed_steps,
            double *dts_out, int max_dts, int *nsteps_out, double *t_out) {
    work w;
    const size_t n = (size_t)4 * p->nx * p->ny;
    double *Un = (double *)malloc(n * sizeof(double)), *rhs = (double *)malloc(n * sizeof(double));
    if (!Un || !rhs || work_init(&w, p)) { free(Un); free(rhs); return -1; }
    double now = t_start;
    int count = 0, err = 0;
    for (;;) {
        if (fixed_steps > 0) {
            if (count >= fixed_steps) break;
        } else if (!(fabs(now - t_end) > 1e-10)) {
            break;
        }
        double dt;
        err |= orc_compute_dt(p, U, &dt);
        if (err & ORC_ERR_NAN_DT) break;
        if (fixed_steps <= 0) dt = stdMin(dt, t_end - now);
        err |= rk_step(&w, U, Un, rhs, dt);
        if (dts_out && count < max_dts) dts_out[count] = dt;
        now += dt;
        count++;
    }
    if (nsteps_out) *nsteps_out = count;
    if (t_out) *t_out = now;
    work_free(&w); free(Un); free(rhs);
    return err;
}
