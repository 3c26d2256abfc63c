// wenofv_cases.cpp -- the reference's test cases behind the C ABI: TwoPhaseEquation's
// setEquationParameters (include/Equations.hpp:486-977) as a wfv_case, and CWENOFV's
// initializeAve (src/WENOFV.cpp:232-279) as cell averages on the grid the kernels use.
#include <vector>

#include "Equations.hpp"
#include "wenofv.h"
#include "wenofv_geometry.h"

namespace {
// Gauss-Legendre-4 on [-1,1] (third_party/scTools/GaussQuadratureInRef.hpp:64-75)
const double GLP[4] = {-0.861136311594052575223946488893, -0.339981043584856264802665759103,
                       0.339981043584856264802665759103, 0.861136311594052575223946488893};
const double GLWT[4] = {0.347854845137453857373063949222, 0.652145154862546142626936050778,
                        0.652145154862546142626936050778, 0.347854845137453857373063949222};

// BOUNDARYTYPE (include/Equations.hpp:25-32) -> wfv_options.boundary
int boundary_kind(BOUNDARYTYPE b) {
    switch (b) {
    case PERIOD: return 0;
    case NoSlip: return 2;
    case NEUMANN2: return 3;
    default: return 1;  // NEUMANN, FREESLIP: reflective slip walls
    }
}
}  // namespace

extern "C" int wfv_reference_case(int testcase, wfv_case *cs, int *boundary, double *t_end) {
    if (!cs || testcase < 0 || testcase > 8) return WFV_ERR_ARG;
    TwoPhaseEquation e;
    e.setEquationParameters((TESTCASETYPE)testcase);
    *cs = wfv_case{e.xL, e.xR, e.yL, e.yR, e.g_rho10, e.g_rho20, e.g_sound1, e.g_sound2, e.g_preb, e.g_gravity};
    if (boundary) *boundary = boundary_kind(e.boundaryCondition);
    if (t_end) *t_end = e.outputTimes.back();
    return WFV_OK;
}

// cell averages of getU0 over one cell: 4x4 Gauss-Legendre (src/WENOFV.cpp:232-279)
static void cell_average(TwoPhaseEquation &e, double xc, double yc, double dx, double dy, double out[4]) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0}, u[4];
    for (int gi = 0; gi < 4; ++gi)
        for (int gj = 0; gj < 4; ++gj) {
            const double gx = xc + 0.5 * dx * GLP[gi];
            const double gy = yc + 0.5 * dy * GLP[gj];
            const double wx = 0.5 * dx * GLWT[gi];
            const double wy = 0.5 * dy * GLWT[gj];
            e.getU0(gx, gy, u);
            for (int r = 0; r < 4; ++r) acc[r] += u[r] * wx * wy;
        }
    const double area = dx * dy;
    for (int r = 0; r < 4; ++r) out[r] = acc[r] / area;
}

extern "C" int wfv_reference_initial_state_padded(int testcase, int nx, int ny, int geometry_ranks, double *Upad) {
    if (!Upad || testcase < 0 || testcase > 8 || nx < 1 || ny < 1) return WFV_ERR_ARG;
    const int G = WFV_GHOST, px = nx + 2 * G, py = ny + 2 * G;
    std::vector<double> U(4 * (size_t)nx * ny);
    const int rc = wfv_reference_initial_state(testcase, nx, ny, geometry_ranks, U.data());
    if (rc != WFV_OK) return rc;
    TwoPhaseEquation e;
    e.setEquationParameters((TESTCASETYPE)testcase);
    const double dX = (e.xR - e.xL) / nx, dY = (e.yR - e.yL) / ny;
    const size_t pc = (size_t)px * py, nc = (size_t)nx * ny;
    for (int j = -G; j < ny + G; ++j)
        for (int i = -G; i < nx + G; ++i) {
            const size_t o = (size_t)(j + G) * px + (i + G);
            if (i >= 0 && i < nx && j >= 0 && j < ny) {
                for (int r = 0; r < 4; ++r) Upad[r * pc + o] = U[r * nc + (size_t)j * nx + i];
                continue;
            }
            double a[4];  // ghost cells continue the uniform grid
            cell_average(e, e.xL + (i + 0.5) * dX, e.yL + (j + 0.5) * dY, dX, dY, a);
            for (int r = 0; r < 4; ++r) Upad[r * pc + o] = a[r];
        }
    return WFV_OK;
}

extern "C" int wfv_reference_initial_state(int testcase, int nx, int ny, int geometry_ranks, double *U) {
    if (!U || testcase < 0 || testcase > 8 || nx < 1 || ny < 1) return WFV_ERR_ARG;
    TwoPhaseEquation e;
    e.setEquationParameters((TESTCASETYPE)testcase);
    const size_t nc = (size_t)nx * ny;
    std::vector<double> dxC(nx), dxW(nx), dyC(ny), xc(nx), yc(ny);
    wfv::grid_widths(nx, ny, e.xL, e.xR, e.yL, e.yR, geometry_ranks, dxC.data(), dxW.data(), dyC.data(), xc.data(),
                     yc.data());
    double u[4];
    for (int ei = 0; ei < nx; ++ei)
        for (int ej = 0; ej < ny; ++ej) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int gi = 0; gi < 4; ++gi)
                for (int gj = 0; gj < 4; ++gj) {
                    const double gx = xc[ei] + 0.5 * dxC[ei] * GLP[gi];
                    const double gy = yc[ej] + 0.5 * dyC[ej] * GLP[gj];
                    const double wx = 0.5 * dxC[ei] * GLWT[gi];
                    const double wy = 0.5 * dyC[ej] * GLWT[gj];
                    e.getU0(gx, gy, u);
                    for (int r = 0; r < 4; ++r) acc[r] += u[r] * wx * wy;
                }
            const double area = dxC[ei] * dyC[ej];
            for (int r = 0; r < 4; ++r) U[r * nc + (size_t)ej * nx + ei] = acc[r] / area;
        }
    return WFV_OK;
}
