// ref_point.cpp -- C entry points over the reference's own inline point kernels.
//
// TEST INFRASTRUCTURE ONLY, built in the development container by oracle/build_ref.sh
// into oracle/_ref/librefpoint.so.  It includes the reference headers where they lie
// (include/Equations.hpp, include/Utility.hpp) and exposes the reference's own
// TwoPhaseEquation methods and WENO functions through plain C signatures, so that
// tests/golden/make_golden.py can record their outputs as fixtures.  Nothing here is
// compiled into the product, and nothing of the reference's source is copied.
#include "Equations.hpp"
#include "Utility.hpp"

namespace {
TwoPhaseEquation g_eq;

Array1D<double> vec4(const double *u) {
    Array1D<double> a(4);
    for (int r = 0; r < 4; ++r) a[r] = u[r];
    return a;
}
}  // namespace

extern "C" {

// Set the TwoPhaseEquation parameters used by the EOS/flux/eigen entry points.
void ref_set_params(double rho10, double rho20, double sound1, double sound2, double preb) {
    g_eq.g_rho10 = rho10;
    g_eq.g_rho20 = rho20;
    g_eq.g_sound1 = sound1;
    g_eq.g_sound2 = sound2;
    g_eq.g_preb = preb;
}

// Load a reference test case's parameters (TwoPhaseEquation::setEquationParameters).
int ref_set_case(int testcase, double *out5) {
    g_eq.setEquationParameters((TESTCASETYPE)testcase);
    out5[0] = g_eq.g_rho10;
    out5[1] = g_eq.g_rho20;
    out5[2] = g_eq.g_sound1;
    out5[3] = g_eq.g_sound2;
    out5[4] = g_eq.g_preb;
    return 0;
}

double ref_weno_js(const double *s, int gp) { return WENO5threconstruction(s[0], s[1], s[2], s[3], s[4], gp); }
double ref_weno_z(const double *s, int gp) { return WENO5Zthreconstruction(s[0], s[1], s[2], s[3], s[4], gp); }
// 7-point kernels (s = {u_mmm, ..., u_ppp}); dx is the reference's m_deltaX
double ref_weno_exp(const double *s, int gp, double dx) {
    return wenoExpthreconstruction(s[0], s[1], s[2], s[3], s[4], s[5], s[6], gp, dx);
}
double ref_weno_exp_thinc(const double *s, int gp, double dx) {
    return wenoExpWithThincReconstruction(s[0], s[1], s[2], s[3], s[4], s[5], s[6], gp, dx);
}

double ref_volume_frac(const double *u) { return g_eq.getVolumeFrac(vec4(u)); }
double ref_pressure(const double *u) { return g_eq.getPressure(vec4(u)); }
double ref_sound_speed(const double *u) { return g_eq.getSoundSpeed(vec4(u)); }
double ref_max_eig(const double *u, double nx, double ny) {
    Array1D<double> a = vec4(u);
    return g_eq.getMaxEigenValue(a, nx, ny);
}
void ref_phy_flux(const double *u, double *f, double nx, double ny) {
    Array1D<double> F(4);
    g_eq.getPhyFlux(vec4(u), F, nx, ny);
    for (int r = 0; r < 4; ++r) f[r] = F[r];
}
void ref_riemann(const double *ul, const double *ur, double *f, double nx, double ny, int type) {
    Array1D<double> F(4);
    if (type == 1) g_eq.getHLLCRiemannFlux(vec4(ul), vec4(ur), F, nx, ny);
    else g_eq.getLLFRiemannFlux(vec4(ul), vec4(ur), F, nx, ny);
    for (int r = 0; r < 4; ++r) f[r] = F[r];
}
void ref_eigen(const double *u, double nx, double ny, double *L, double *R) {
    Array2D<double> A(4, 4), B(4, 4);
    Array1D<double> a = vec4(u);
    g_eq.getLEigenMatrix(a, nx, ny, A);
    g_eq.getREigenMatrix(a, nx, ny, B);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            L[i * 4 + j] = A[i][j];
            R[i * 4 + j] = B[i][j];
        }
}
// TwoPhaseEquation::getU0 for a test case at one point (IC pinning).
void ref_u0(double x, double y, double *u) {
    Array1D<double> U(4);
    g_eq.getU0(x, y, U);
    for (int r = 0; r < 4; ++r) u[r] = U[r];
}
}
