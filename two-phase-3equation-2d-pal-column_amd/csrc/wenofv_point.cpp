// wenofv_point.cpp -- the model's point functions behind the C ABI: the kernels' own all-FP64
// functions (wenofv_device.h) compiled for the host without contraction, so that a value
// computed here rounds as the reference's inline TwoPhaseEquation and CWENOFV functions do
// (include/Equations.hpp:84-483, include/Utility.hpp:5-472, src/WENOFV.cpp:1873-2000) and as
// the all-FP64 kernels do.
#include <cmath>

#include "wenofv.h"
#include "wenofv_device.h"
#include "wenofv_host.h"

using namespace wfv;

extern "C" {

int wfv_point_eos(const wfv_case *cs, const double *U, double *out) {
    if (!cs || !U || !out) return WFV_ERR_ARG;
    Eos e;
    eos(make_phys(*cs), U, e);
    out[0] = e.z1;
    out[1] = e.p;
    out[2] = e.c;
    out[3] = e.rho;
    out[4] = e.rho1;
    out[5] = e.rho2;
    return WFV_OK;
}

int wfv_point_flux(const wfv_case *cs, const double *U, double nx, double ny, double *F) {
    if (!cs || !U || !F) return WFV_ERR_ARG;
    Eos e;
    eos(make_phys(*cs), U, e);
    phy_flux<double>(U, e.rho, U[2] / e.rho, U[3] / e.rho, e.p, nx > ny, F);
    return WFV_OK;
}

int wfv_point_riemann(const wfv_case *cs, int riemann_flux, const double *UL, const double *UR, double nx, double ny,
                      double *F) {
    if (!cs || !UL || !UR || !F || riemann_flux < 0 || riemann_flux > 1) return WFV_ERR_ARG;
    if (riemann_flux == 1) riemann_hllc_n(make_phys(*cs), UL, UR, nx, ny, F);
    else riemann_llf_n(make_phys(*cs), UL, UR, nx, ny, F);
    return WFV_OK;
}

int wfv_point_eigen(const wfv_case *cs, const double *U, double nx, double ny, double *L, double *R) {
    if (!cs || !U || !L || !R) return WFV_ERR_ARG;
    if (std::fabs(nx) > std::fabs(ny)) eigen_LR<double, true>(make_phys(*cs), U, L, R);
    else eigen_LR<double, false>(make_phys(*cs), U, L, R);
    return WFV_OK;
}

int wfv_point_weno(int scheme, const double *s, double dx, double *out) {
    if (!s || !out || scheme < 0 || scheme > 2) return WFV_ERR_ARG;
    auto nodes = [out](const auto &st) {
        out[0] = st.template node<0>();
        out[1] = st.template node<1>();
        out[2] = st.template node<2>();
        out[3] = st.template node<3>();
    };
    if (scheme == 0) nodes(Weno5<double, 0>(s[1], s[2], s[3], s[4], s[5]));
    else if (scheme == 1) nodes(Weno5<double, 1>(s[1], s[2], s[3], s[4], s[5]));
    else nodes(WenoExp7<false>(s, std::pow(dx, 4)));  // the reference's pow(dx, 4), include/Utility.hpp:297
    return WFV_OK;
}

int wfv_point_theta(const wfv_case *cs, const double *U, const double *faces, double dX, double dY, double *theta) {
    if (!cs || !U || !faces || !theta) return WFV_ERR_ARG;
    double F[4][4][4];
    for (int k = 0; k < 64; ++k) F[k / 16][(k / 4) % 4][k % 4] = faces[k];
    *theta = bp_theta(make_phys(*cs), U, F, dX, dY);
    return WFV_OK;
}

}  // extern "C"
