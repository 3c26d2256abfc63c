// Equations.hpp -- host-side model of the two-phase three-equation system: test cases,
// initial conditions and the equation of state, with the public surface of the
// reference's TwoPhaseEquation (include/Equations.hpp:52-1044 of the reference), so that
// code written against that class compiles against this one.
//
// The time-stepping arithmetic itself runs in the HIP kernels (wenofv_device.h); this
// class only feeds them parameters and initial cell averages.
#pragma once
#include <cmath>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

enum RIEMANNFLUXTYPE { LLF = 0, HLLC = 1 };

// reference test cases (include/Equations.hpp:12-23)
enum TESTCASETYPE {
    SMOOTH = 0,
    VORTEX = 1,
    SHARP = 2,
    DAMBREAK = 3,
    DAMBREAK2 = 4,
    DAMBREAK3 = 5,
    RTI = 6,
    STANDING_WAVE = 7,
    ZALESAK = 8,
};

// boundary kinds (include/Equations.hpp:25-32).  The reference's MPI solver runs PERIOD
// only; here every non-periodic kind maps onto reflective walls.
enum BOUNDARYTYPE { PERIOD = 0, NEUMANN = 1, NoSlip = 2, NEUMANN2 = 3, FREESLIP = 4 };

inline const std::string TESTCASETYPE_STRINGS[] = {"SMOOTH",    "VORTEX", "SHARP",         "DAMBREAK", "DAMBREAK2",
                                                   "DAMBREAK3", "RTI",    "STANDING_WAVE", "ZALESAK"};
inline const std::string RIEMANNFLUXTYPE_STRINGS[] = {"LLF", "HLLC"};

class TwoPhaseEquation {
public:
    double xL = 0, xR = 1, yL = 0, yR = 1;
    std::vector<double> outputTimes;
    BOUNDARYTYPE boundaryCondition = PERIOD;
    std::function<double(const double *)> theVarUh;        // scalar checked by the accuracy output
    std::function<double(double, double, double)> theVarExact;
    bool u_exact_exist = false;

    double g_rho10 = 1, g_rho20 = 1;
    double g_sound1 = 1, g_sound2 = 1;
    double g_u0 = 0, g_v0 = 0, g_preb = 0;
    double g_epsa = 1e-8;
    double g_gravity = 0;  // source term of assembleSourceTerm (src/WENOFV.cpp:1831-1872)

    std::function<double(double, double)> z1_0, p1_0, p2_0, u_0, v_0;

    int getVarNum() const { return 4; }

    // TwoPhaseEquation::getVolumeFrac (include/Equations.hpp:84-99)
    double getVolumeFrac(const double *U) const {
        const double c1_2 = g_sound1 * g_sound1, c2_2 = g_sound2 * g_sound2;
        const double q = g_rho20 * c2_2 - g_rho10 * c1_2;
        const double qs = U[1] * c2_2 - U[0] * c1_2;
        const double r = ((q - qs) + std::sqrt((q - qs) * (q - qs) + 4.0 * U[0] * c1_2 * U[1] * c2_2)) /
                         (2.0 * U[1] * c2_2);
        return r / (1.0 + r);
    }
    // TwoPhaseEquation::getPressure (include/Equations.hpp:101-124)
    double getPressure(const double *U) const {
        const double z1 = getVolumeFrac(U), z2 = 1.0 - z1;
        const double pre1 = g_sound1 * g_sound1 * (U[0] / z1 - g_rho10);
        const double pre2 = g_sound2 * g_sound2 * (U[1] / z2 - g_rho20);
        return z1 < z2 ? pre2 : pre1;
    }

    // TwoPhaseEquation::setEquationParameters (include/Equations.hpp:486-977): domain,
    // output times, boundary, phase parameters and initial fields of each test case.
    void setEquationParameters(TESTCASETYPE type) {
        auto uniform = [this]() {
            p1_0 = [this](double, double) { return g_preb; };
            p2_0 = [this](double, double) { return g_preb; };
            u_0 = [this](double, double) { return g_u0; };
            v_0 = [this](double, double) { return g_v0; };
        };
        g_gravity = 0;
        switch (type) {
        case SMOOTH:
            xL = 0; xR = 1; yL = 0; yR = 1;
            outputTimes = {0, 0.05};
            boundaryCondition = PERIOD;
            g_rho10 = 1.2; g_rho20 = 1.0; g_sound1 = 1.0; g_sound2 = 1.0;
            g_u0 = 1.0; g_v0 = 1.0; g_preb = 0.0; g_epsa = 1.0e-8;
            z1_0 = [this](double x, double) { return g_epsa + (1 - 2 * g_epsa) * std::pow(std::sin(2.0 * M_PI * (x)), 2); };
            uniform();
            u_exact_exist = true;
            theVarExact = [this](double x, double y, double t) { return z1_0(x - g_u0 * t, y - g_v0 * t); };
            theVarUh = [this](const double *U) { return getVolumeFrac(U); };
            break;
        case VORTEX: {
            xL = -1; xR = 1; yL = -1; yR = 1;
            outputTimes = {0, 0.005, 0.01};
            boundaryCondition = PERIOD;
            g_rho10 = 1000.0; g_rho20 = 1.0; g_sound1 = 30.0; g_sound2 = 30.0;
            g_u0 = 2; g_v0 = 2; g_preb = 0; g_epsa = 1.0e-4;
            auto wrap = [](double &x, double &y) {
                while (x >= 1) x -= 2;
                while (x < -1) x += 2;
                while (y >= 1) y -= 2;
                while (y < -1) y += 2;
            };
            z1_0 = [this, wrap](double x, double y) {
                wrap(x, y);
                const double r = std::sqrt(x * x + y * y);
                return (1 - 2 * g_epsa) * std::exp(-std::pow(r, 2) / std::pow(0.2, 2)) + g_epsa;
            };
            p1_0 = [wrap](double x, double y) {
                wrap(x, y);
                return 0 * std::sqrt(x * x + y * y);
            };
            p2_0 = p1_0;
            u_0 = [this, wrap](double x, double y) {
                wrap(x, y);
                const double L = 1, theta = std::atan2(y, x), r = std::sqrt(x * x + y * y);
                return g_u0 + r * 0.0 + theta * 0.0 + L * 0.0;
            };
            v_0 = [this, wrap](double x, double y) {
                wrap(x, y);
                const double L = 1, theta = std::atan2(y, x), r = std::sqrt(x * x + y * y);
                return g_v0 + r * 0.0 + theta * 0.0 + L * 0.0;
            };
            u_exact_exist = true;
            theVarExact = [this](double x, double y, double t) { return z1_0(x - g_u0 * t, y - g_v0 * t); };
            theVarUh = [this](const double *U) { return getVolumeFrac(U); };
            break;
        }
        case SHARP:
            xL = 0; xR = 1; yL = 0; yR = 1;
            outputTimes = {0, 1.0};
            boundaryCondition = PERIOD;
            g_rho10 = 1000; g_rho20 = 1.0; g_sound1 = 70.0; g_sound2 = 70.0;
            g_u0 = 1.0; g_v0 = 1.0; g_preb = 0.0; g_epsa = 1e-8;
            z1_0 = [this](double x, double y) {
                const double r = (xR - xL) / 10.0, xc = (xL + xR) / 2.0, yc = (yL + yR) / 2.0;
                return (std::fabs(x - xc) < r && std::fabs(y - yc) < r) ? 1 - g_epsa : g_epsa;
            };
            uniform();
            u_exact_exist = true;
            theVarExact = [this](double x, double y, double t) {
                (void)t;
                const double r = (xR - xL) / 10, xc = (xL + xR) / 2, yc = (yL + yR) / 2;
                const double z1 = (std::fabs(x - xc) < r && std::fabs(y - yc) < r) ? 1 - g_epsa : g_epsa;
                return z1 * g_rho10;
            };
            theVarUh = [](const double *U) { return U[0]; };
            break;
        case ZALESAK:
            xL = 0; xR = 1; yL = 0; yR = 1;
            outputTimes = {2};
            boundaryCondition = NEUMANN2;
            g_rho10 = 1000; g_rho20 = 1.0; g_sound1 = 70.0; g_sound2 = 70.0;
            g_u0 = 1.0; g_v0 = 1.0; g_preb = 0.0; g_epsa = 1e-8;
            z1_0 = [this](double x, double y) {
                return std::pow(x - 0.5, 2) + std::pow(y - 0.5, 2) < std::pow(0.15, 2) ? 1 - g_epsa : g_epsa;
            };
            uniform();
            u_0 = [](double, double y) { return M_PI * (0.5 - y); };
            v_0 = [](double x, double) { return M_PI * (x - 0.5); };
            u_exact_exist = false;
            theVarExact = [](double, double, double) { return 0.0; };
            theVarUh = [](const double *U) { return U[0]; };
            break;
        case STANDING_WAVE:
            xL = 0; xR = 0.5; yL = 0; yR = 0.5;
            outputTimes = {3.0};
            boundaryCondition = NEUMANN;
            g_rho10 = 1000.0; g_rho20 = 1.0; g_sound1 = 100; g_sound2 = 100;
            g_u0 = 0.0; g_v0 = 0.0; g_preb = 1e4; g_epsa = 1e-6;
            g_gravity = 2 * M_PI;
            z1_0 = [this](double x, double y) {
                const double hs = 0.25, A = 0.025, kn = 2 * M_PI;
                return y <= hs + A * std::cos(kn * x) ? 1.0 - g_epsa : g_epsa;
            };
            uniform();
            u_exact_exist = false;
            theVarExact = [](double, double, double) { return 0.0; };
            theVarUh = [](const double *U) { return U[0]; };
            break;
        case RTI:
            xL = 0; xR = 1; yL = 0; yR = 4;
            outputTimes = {0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.0};
            boundaryCondition = NEUMANN;
            g_rho10 = 1.225; g_rho20 = 0.1694; g_sound1 = 100.0; g_sound2 = 100.0;
            g_u0 = 0; g_v0 = 0; g_preb = 10000.0; g_epsa = 1.0e-8;
            g_gravity = 9.81;
            z1_0 = [this](double x, double y) {
                const double yi = 2.0 + 0.05 * std::cos(2.0 * M_PI * x);
                return y >= yi ? 1.0 - g_epsa : g_epsa;
            };
            uniform();
            u_exact_exist = false;
            theVarExact = [](double, double, double) { return 0.0; };
            theVarUh = [](const double *U) { return U[0]; };
            break;
        case DAMBREAK:
        case DAMBREAK2:
        case DAMBREAK3: {
            const double xr[] = {1.61, 1.61, 0.8}, yr[] = {0.805, 0.6, 0.6}, eps[] = {1e-6, 1e-5, 1e-5};
            const double H0[] = {0.3, 0.6, 0.4}, L0[] = {0.6, 0.6, 0.2};
            const int k = type - DAMBREAK;
            xL = 0; xR = xr[k]; yL = 0; yR = yr[k];
            outputTimes = k == 0 ? std::vector<double>{0.0, 0.1599, 0.2766, 0.3733, 0.4499, 0.5733, 0.8623, 1.0233, 1.1666}
                                 : std::vector<double>{1.5};
            boundaryCondition = NEUMANN;
            g_rho10 = 1000; g_rho20 = 1.0; g_sound1 = 100.0; g_sound2 = 100.0;
            g_u0 = 0; g_v0 = 0; g_preb = 1e4; g_epsa = eps[k];
            g_gravity = 9.81;
            const double h0 = H0[k], l0 = L0[k];
            z1_0 = [this, h0, l0, k](double x, double y) {
                const bool in = k == 2 ? (x >= xL && x <= xL + l0 && y >= yL && y <= yL + h0)
                                       : (x >= xR - l0 && x <= xR && y >= yL && y <= yL + h0);
                return in ? 1.0 - g_epsa : g_epsa;
            };
            uniform();
            u_exact_exist = false;
            theVarExact = [](double, double, double) { return 0.0; };
            theVarUh = [](const double *U) { return U[0]; };
            break;
        }
        default:
            throw std::invalid_argument("Invalid example number");
        }
    }

    // TwoPhaseEquation::getU0 (include/Equations.hpp:979-1002)
    void getU0(double xP, double yP, double *U) const {
        const double z_1 = z1_0(xP, yP), z_2 = 1.0 - z_1;
        const double p_1 = p1_0(xP, yP), p_2 = p2_0(xP, yP);
        const double rho10 = (p_1 - g_preb) / std::pow(g_sound1, 2) + g_rho10;
        const double rho20 = (p_2 - g_preb) / std::pow(g_sound2, 2) + g_rho20;
        const double u = u_0(xP, yP), v = v_0(xP, yP);
        const double z1rho1 = rho10 * z_1, z2rho2 = rho20 * z_2;
        const double rho = z1rho1 + z2rho2;
        U[0] = z1rho1;
        U[1] = z2rho2;
        U[2] = rho * u;
        U[3] = rho * v;
    }
};
