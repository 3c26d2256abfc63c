// Equations.hpp -- host-side model of the two-phase three-equation system: test cases,
// initial conditions, and the point functions of the model (equation of state, physical and
// Riemann fluxes, eigenvectors, derived output variables), with the public surface of the
// reference's TwoPhaseEquation (include/Equations.hpp:52-1044 of the reference), so that
// code written against that class compiles against this one.
//
// Header-only, so it compiles with the caller's containers (wenofv_array.h).  The point
// functions call the library's wfv_point_* entry points, which evaluate the kernels' own
// all-FP64 functions: a value computed here is the value a kernel computes.
#pragma once
#include <cmath>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "wenofv.h"
#include "wenofv_array.h"

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
    std::function<double(Array1D<double>)> theVarUh;       // scalar checked by the accuracy output
    std::function<double(double, double, double)> theVarExact;
    bool u_exact_exist = false;

    double g_rho10 = 1, g_rho20 = 1;
    double g_sound1 = 1, g_sound2 = 1;
    double g_u0 = 0, g_v0 = 0, g_preb = 0;
    double g_epsa = 1e-8;
    double g_gravity = 0;  // source term of assembleSourceTerm (src/WENOFV.cpp:1831-1872)

    std::function<double(double, double)> z1_0, p1_0, p2_0, u_0, v_0;

    int getVarNum() const { return 4; }

    // ------------------------------------------------------------ point functions
    // U = (z1 rho1, z2 rho2, rho u, rho v).  A normal (nx, ny) selects the x direction where
    // the reference does: getPhyFlux where nx > ny, the eigenvector matrices where
    // |nx| > |ny|; the Riemann fluxes and the maximum eigenvalue take it as it is.

    // the phase parameters as the library's point functions take them
    wfv_case pointCase() const {
        wfv_case c{};
        c.xL = xL;
        c.xR = xR;
        c.yL = yL;
        c.yR = yR;
        c.rho10 = g_rho10;
        c.rho20 = g_rho20;
        c.sound1 = g_sound1;
        c.sound2 = g_sound2;
        c.preb = g_preb;
        c.gravity = g_gravity;
        return c;
    }
    // z1, p, c, rho, rho1, rho2 of a state
    void getEos(const double *U, double e[6]) const {
        const wfv_case c = pointCase();
        if (wfv_point_eos(&c, U, e) != WFV_OK) throw std::invalid_argument("getEos");
    }

    // getVolumeFrac (include/Equations.hpp:84-99)
    double getVolumeFrac(const Array1D<double> U) const { return getVolumeFrac(&U[0]); }
    double getVolumeFrac(const double *U) const {
        double e[6];
        getEos(U, e);
        return e[0];
    }
    // getPressure (include/Equations.hpp:101-124)
    double getPressure(const Array1D<double> U) const { return getPressure(&U[0]); }
    double getPressure(const double *U) const {
        double e[6];
        getEos(U, e);
        return e[1];
    }
    // getSoundSpeed (include/Equations.hpp:126-146): Wood's mixture sound speed
    double getSoundSpeed(const Array1D<double> U) const {
        double e[6];
        getEos(&U[0], e);
        return e[2];
    }
    // getMaxEigenValue (include/Equations.hpp:149-156): |u.n| + c
    double getMaxEigenValue(const Array1D<double> &U, double nx, double ny) const {
        const double c = getSoundSpeed(U);
        const double u = U[2] / (U[0] + U[1]);
        const double v = U[3] / (U[0] + U[1]);
        return std::fabs(u * nx + v * ny) + c;
    }
    // getPhyFlux (include/Equations.hpp:158-186)
    void getPhyFlux(const Array1D<double> Uh, Array1D<double> &Flux, double nx, double ny) const {
        if (Flux.dim() < 4) Flux.Resize(4);
        const wfv_case c = pointCase();
        wfv_point_flux(&c, &Uh[0], nx, ny, &Flux[0]);
    }
    // getREigenMatrix / getLEigenMatrix (include/Equations.hpp:189-327), 4 x 4
    void getREigenMatrix(const Array1D<double> &U, double nx, double ny, Array2D<double> &eigMatrix) const {
        eigen(U, nx, ny, eigMatrix, false);
    }
    void getLEigenMatrix(const Array1D<double> &U, double nx, double ny, Array2D<double> &eigMatrix) const {
        eigen(U, nx, ny, eigMatrix, true);
    }
    // getRiemannFlux (include/Equations.hpp:329-360): LLF (riemannFluxType 0) or HLLC (1); a
    // NaN flux is reported on std::cerr, as the reference does (without its pause on stdin)
    void getRiemannFlux(const Array1D<double> UL, const Array1D<double> UR, Array1D<double> &Flux, double nx, double ny,
                        int riemannFluxType) const {
        if (riemannFluxType != LLF && riemannFluxType != HLLC) {
            std::cerr << "Riemann flux type is not supported." << std::endl;
            return;
        }
        if (Flux.dim() < 4) Flux.Resize(4);
        const wfv_case c = pointCase();
        wfv_point_riemann(&c, riemannFluxType, &UL[0], &UR[0], nx, ny, &Flux[0]);
        for (int r = 0; r != getVarNum(); ++r)
            if (std::isnan(Flux[r])) {
                std::cerr << "Riemann flux is nan." << std::endl;
                std::cerr << "UL: " << UL << std::endl << "UR: " << UR << std::endl;
                std::cerr << "nx: " << nx << std::endl << "ny: " << ny << std::endl;
                std::cerr << "riemannFluxType: " << riemannFluxType << std::endl << "r: " << r << std::endl;
                std::cerr << "Flux: " << Flux << std::endl;
            }
    }
    // getLLFRiemannFlux (include/Equations.hpp:363-378), getHLLCRiemannFlux (381-483)
    void getLLFRiemannFlux(const Array1D<double> UL, const Array1D<double> UR, Array1D<double> &Flux, double nx,
                           double ny) const {
        if (Flux.dim() < 4) Flux.Resize(4);
        const wfv_case c = pointCase();
        wfv_point_riemann(&c, LLF, &UL[0], &UR[0], nx, ny, &Flux[0]);
    }
    void getHLLCRiemannFlux(const Array1D<double> &UL, const Array1D<double> &UR, Array1D<double> &Flux, double nx,
                            double ny) const {
        if (Flux.dim() < 4) Flux.Resize(4);
        const wfv_case c = pointCase();
        wfv_point_riemann(&c, HLLC, &UL[0], &UR[0], nx, ny, &Flux[0]);
    }
    // getVitalVarVal / Num / Name (include/Equations.hpp:1005-1043): the 9 output variables
    // (z1 rho1, z2 rho2, z1, rho1, rho2, rho, u, v, p)
    void getVitalVarVal(const Array1D<double> Uh, Array1D<double> &VitalVar) const {
        double e[6];
        getEos(&Uh[0], e);
        if (VitalVar.dim() < 9) VitalVar.Resize(9);
        VitalVar[0] = Uh[0];
        VitalVar[1] = Uh[1];
        VitalVar[2] = e[0];
        VitalVar[3] = e[4];
        VitalVar[4] = e[5];
        VitalVar[5] = Uh[0] + Uh[1];
        VitalVar[6] = Uh[2] / (Uh[0] + Uh[1]);
        VitalVar[7] = Uh[3] / (Uh[0] + Uh[1]);
        VitalVar[8] = e[1];
    }
    int getVitalVarNum() const { return 9; }
    void getVitalVarName(Array1D<std::string> &VitalVarName) const {
        static const char *names[9] = {"RHO1Z1", "RHO2Z2", "Z1", "RHO1", "RHO2", "RHO", "U", "V", "PRE"};
        if (VitalVarName.dim() < 9) VitalVarName.Resize(9);
        for (int k = 0; k < 9; ++k) VitalVarName[k] = names[k];
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
            theVarUh = [this](Array1D<double> U) { return getVolumeFrac(U); };
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
            theVarUh = [this](Array1D<double> U) { return getVolumeFrac(U); };
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
            theVarUh = [](Array1D<double> U) { return U[0]; };
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
            theVarUh = [](Array1D<double> U) { return U[0]; };
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
            theVarUh = [](Array1D<double> U) { return U[0]; };
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
            theVarUh = [](Array1D<double> U) { return U[0]; };
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
            theVarUh = [](Array1D<double> U) { return U[0]; };
            break;
        }
        default:
            throw std::invalid_argument("Invalid example number");
        }
    }

    // TwoPhaseEquation::getU0 (include/Equations.hpp:979-1002)
    void getU0(double xP, double yP, Array1D<double> &U) const {
        if (U.dim() < 4) U.Resize(4);
        getU0(xP, yP, &U[0]);
    }
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

private:
    void eigen(const Array1D<double> &U, double nx, double ny, Array2D<double> &M, bool left) const {
        double L[16], R[16];
        const wfv_case c = pointCase();
        wfv_point_eigen(&c, &U[0], nx, ny, L, R);
        if (M.row() < 4 || M.col() < 4) M.Resize(4, 4);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) M[i][j] = left ? L[i * 4 + j] : R[i * 4 + j];
    }
};
