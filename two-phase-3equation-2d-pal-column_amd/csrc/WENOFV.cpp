// WENOFV.cpp -- CWENOFV facade (see WENOFV.hpp) over the C ABI.
#include "WENOFV.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "wenofv_geometry.h"

extern "C" int wfv_reference_initial_state(int testcase, int nx, int ny, int geometry_ranks, double *U);

namespace {
// Gauss-Legendre-4 on [-1,1] (third_party/scTools/GaussQuadratureInRef.hpp:64-75)
const double GLP[4] = {-0.861136311594052575223946488893, -0.339981043584856264802665759103,
                       0.339981043584856264802665759103, 0.861136311594052575223946488893};
const double GLWT[4] = {0.347854845137453857373063949222, 0.652145154862546142626936050778,
                        0.652145154862546142626936050778, 0.347854845137453857373063949222};
}  // namespace

void CWENOFV::check(int rc, const char *what) const {
    if (rc != WFV_OK) throw std::runtime_error(std::string(what) + ": " + wfv_last_error(ctx_));
}

// BOUNDARYTYPE (include/Equations.hpp:25-32) -> wfv_options.boundary
static int boundary_kind(BOUNDARYTYPE b) {
    switch (b) {
    case PERIOD: return 0;
    case NoSlip: return 2;
    case NEUMANN2: return 3;
    default: return 1;  // NEUMANN, FREESLIP: reflective slip walls
    }
}

int CWENOFV::stoi_key(const char *key) {
    // a missing key makes the reference's std::stoi("") throw (src/WENOFV.cpp:43-63)
    auto it = option.find(key);
    if (it == option.end() || it->second.empty()) throw std::invalid_argument(std::string("missing option ") + key);
    return std::stoi(it->second);
}

void CWENOFV::initializeSolver() {
    m_testcase = (TESTCASETYPE)stoi_key("TESTCASE");
    m_rkMeth = (RKMETHOD)stoi_key("RKMETH");
    m_riemannFluxType = (RIEMANNFLUXTYPE)stoi_key("RIEMANNFLUX");
    m_reconstructionType = (RECONSTRUCTIONTYPE)stoi_key("RECONSTRUCTIONTYPE");
    equation->setEquationParameters(m_testcase);
    m_timeMoments = equation->outputTimes;
    // ENDTIME (no reference counterpart): replace the case's output times by {0, ENDTIME}
    if (option.count("ENDTIME")) m_timeMoments = {0.0, std::stod(option["ENDTIME"])};
    m_scheme = (SCHEMETYPE)stoi_key("SCHEME");
    m_pplimiter = stoi_key("PPlimiter") != 0;
    auto cfl = option.find("CFL");
    if (cfl == option.end() || cfl->second.empty()) throw std::invalid_argument("missing option CFL");
    m_cfl = std::stod(cfl->second);
    m_worldElemNumX = stoi_key("ELEMNUMX");
    m_worldElemNumY = stoi_key("ELEMNUMY");
    m_precision = option.count("PRECISION") ? std::stoi(option["PRECISION"]) : 0;
    m_deltaX = (equation->xR - equation->xL) / m_worldElemNumX;
    m_deltaY = (equation->yR - equation->yL) / m_worldElemNumY;

    wfv_options o{};
    o.elem_num_x = m_worldElemNumX;
    o.elem_num_y = m_worldElemNumY;
    o.rk_method = m_rkMeth;
    o.cfl = m_cfl;
    o.scheme = m_scheme;
    o.pp_limiter = m_pplimiter ? 1 : 0;
    o.riemann_flux = m_riemannFluxType;
    o.reconstruction = m_reconstructionType;
    o.precision = m_precision;
    o.boundary = boundary_kind(equation->boundaryCondition);
    // the reference splits x over exactly 4 ranks and needs ELEMNUMX % 4 == 0
    // (src/WENOFV.cpp:24, 71-76); its per-rank cell widths are reproduced when possible
    o.geometry_ranks = (m_worldElemNumX % 4 == 0) ? 4 : 1;
    o.rows_per_strip = option.count("ROWSPERSTRIP") ? std::stoi(option["ROWSPERSTRIP"]) : 0;
    wfv_case c{};
    c.xL = equation->xL;
    c.xR = equation->xR;
    c.yL = equation->yL;
    c.yR = equation->yR;
    c.rho10 = equation->g_rho10;
    c.rho20 = equation->g_rho20;
    c.sound1 = equation->g_sound1;
    c.sound2 = equation->g_sound2;
    c.preb = equation->g_preb;
    c.gravity = equation->g_gravity;
    wfv_destroy(ctx_);
    ctx_ = nullptr;
    const int rc = wfv_create(&o, &c, &ctx_);
    if (rc != WFV_OK) throw std::runtime_error(std::string("wfv_create: ") + wfv_last_error(nullptr));
    // ./output/<timestamp>_<case>_<scheme>_<reconstruction>_<flux>_<nx>x<ny>/ (src/WENOFV.cpp:100-105)
    m_outputDir = "./output/" + timestamp() + "_" + TESTCASETYPE_STRINGS[m_testcase] + "_" + SCHEMETYPE_STRINGS[m_scheme] + "_" +
                  RECONSTRUCTIONTYPE_STRINGS[m_reconstructionType] + "_" + RIEMANNFLUXTYPE_STRINGS[m_riemannFluxType] +
                  "_" + std::to_string(m_worldElemNumX) + "x" + std::to_string(m_worldElemNumY) + "/";
    if (m_pplimiter) check(wfv_record_limiter_flags(ctx_, 1), "record limiter flags");
    initializeAve();
}

void CWENOFV::initializeAve() {
    m_Uh.assign(4 * (size_t)m_worldElemNumX * m_worldElemNumY, 0.0);
    check(wfv_reference_initial_state(m_testcase, m_worldElemNumX, m_worldElemNumY,
                                      (m_worldElemNumX % 4 == 0) ? 4 : 1, m_Uh.data()),
          "initializeAve");
    check(wfv_set_state(ctx_, m_Uh.data()), "wfv_set_state");
    if (equation->boundaryCondition == NEUMANN2) {
        // NEUMANN2 ("boundary from the initial values"): ghost cells hold the initial state
        std::vector<double> pad(4 * (size_t)(m_worldElemNumX + 2 * WFV_GHOST) * (m_worldElemNumY + 2 * WFV_GHOST));
        check(wfv_reference_initial_state_padded(m_testcase, m_worldElemNumX, m_worldElemNumY,
                                                 (m_worldElemNumX % 4 == 0) ? 4 : 1, pad.data()),
              "initializeAve");
        check(wfv_set_ghost_cells(ctx_, pad.data()), "wfv_set_ghost_cells");
    }
}

double CWENOFV::calculateDeltaT(double) {
    double dt = 0;
    check(wfv_compute_dt(ctx_, &dt), "calculateDeltaT");
    return dt;
}

void CWENOFV::performTimeIntegration(double timestep) { check(wfv_step(ctx_, timestep), "performTimeIntegration"); }

void CWENOFV::rk_checked(RKMETHOD m, double dt) {
    if (m != m_rkMeth) throw std::invalid_argument("the context was built for RKMETH " + std::to_string(m_rkMeth));
    performTimeIntegration(dt);
}

void CWENOFV::assembleRHS() {
    const size_t nc = (size_t)m_worldElemNumX * m_worldElemNumY;
    m_rhs.resize(4 * nc);
    m_faces.resize(64 * nc);
    m_theta.resize(nc);
    check(wfv_assemble_rhs(ctx_, m_rhs.data(), m_faces.data(), m_theta.data()), "assembleRHS");
}

double CWENOFV::getPositivityPreservingTheta(int ei, int ej) const {
    const int ix = ei - m_ghostCellNum, iy = ej - m_ghostCellNum;
    if (m_theta.empty() || ix < 0 || iy < 0 || ix >= m_worldElemNumX || iy >= m_worldElemNumY)
        throw std::out_of_range("getPositivityPreservingTheta: call assembleRHS first / cell outside the domain");
    return m_theta[(size_t)iy * m_worldElemNumX + ix];
}

void CWENOFV::getWorldUh() {
    m_Uh.resize(4 * (size_t)m_worldElemNumX * m_worldElemNumY);
    check(wfv_get_state(ctx_, m_Uh.data()), "getWorldUh");
}

std::vector<double> CWENOFV::outputAccuracy(const std::string &prefix, double now) {
    getWorldUh();
    const int nx = m_worldElemNumX, ny = m_worldElemNumY;
    const size_t nc = (size_t)nx * ny;
    double err1 = 0, err2 = 0, errinf = 0;
    for (int ei = 0; ei < nx; ++ei)
        for (int ej = 0; ej < ny; ++ej) {
            // world grid of rank 0 (src/WENOFV.cpp:127-141): xl = xL - 3 dx + i dx
            const int wi = ei + m_ghostCellNum, wj = ej + m_ghostCellNum;
            const double xl = equation->xL - m_ghostCellNum * m_deltaX + wi * m_deltaX;
            const double xr = equation->xL - m_ghostCellNum * m_deltaX + (wi + 1) * m_deltaX;
            const double yl = equation->yL - m_ghostCellNum * m_deltaY + wj * m_deltaY;
            const double yr = equation->yL - m_ghostCellNum * m_deltaY + (wj + 1) * m_deltaY;
            const double xc = (xl + xr) / 2, yc = (yl + yr) / 2;
            double ex = 0;
            for (int gi = 0; gi < 4; ++gi)
                for (int gj = 0; gj < 4; ++gj) {
                    const double gx = xc + 0.5 * m_deltaX * GLP[gi];
                    const double gy = yc + 0.5 * m_deltaY * GLP[gj];
                    const double wx = 0.5 * m_deltaX * GLWT[gi];
                    const double wy = 0.5 * m_deltaY * GLWT[gj];
                    ex += equation->theVarExact(gx, gy, now) * wx * wy / m_deltaX / m_deltaY;
                }
            const size_t c = (size_t)ej * nx + ei;
            const double u4[4] = {m_Uh[c], m_Uh[nc + c], m_Uh[2 * nc + c], m_Uh[3 * nc + c]};
            const double e = std::fabs(ex - equation->theVarUh(u4));
            err1 += e;
            err2 += std::pow(e, 2);
            if (e > errinf) errinf = e;
        }
    err1 = err1 / nx / ny;
    err2 = std::sqrt(err2 / nx / ny);
    std::filesystem::create_directories(m_outputDir);
    const std::string fn = m_outputDir + "accuracy_" + prefix + ".csv";
    const bool exists = std::filesystem::exists(fn);
    std::ofstream f(fn, std::ios::app);
    if (!exists) f << "Cell Number, Linf-norm, L1-norm, L2-norm" << std::endl;
    f << nx << ", " << std::setprecision(15) << std::setw(20) << std::scientific << errinf << ", " << std::setw(20)
      << err1 << ", " << std::setw(20) << err2 << std::endl;
    return {errinf, err1, err2};
}

void CWENOFV::writeState(const std::string &path) const {
    std::ofstream f(path, std::ios::binary);
    const int hdr[4] = {0x57464f56, 4, m_worldElemNumX, m_worldElemNumY};
    f.write(reinterpret_cast<const char *>(hdr), sizeof hdr);
    f.write(reinterpret_cast<const char *>(m_Uh.data()), m_Uh.size() * sizeof(double));
}

void CWENOFV::readState(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    int hdr[4];
    f.read(reinterpret_cast<char *>(hdr), sizeof hdr);
    if (!f || hdr[0] != 0x57464f56 || hdr[2] != m_worldElemNumX || hdr[3] != m_worldElemNumY)
        throw std::runtime_error("readState: " + path + " is not a state of this grid");
    m_Uh.resize(4 * (size_t)m_worldElemNumX * m_worldElemNumY);
    f.read(reinterpret_cast<char *>(m_Uh.data()), m_Uh.size() * sizeof(double));
    check(wfv_set_state(ctx_, m_Uh.data()), "readState");
}

void CWENOFV::run() {
    // CWENOFV::run (src/WENOFV.cpp:281-374): the loop over the output times with the
    // reference's outputs.  The steps between two outputs run on the device
    // (wfv_advance_steps: clipped CFL steps, up to the next step count that writes a file).
    const auto t0 = std::chrono::steady_clock::now();
    initializeSolver();
    double now = 0;
    int count = 0;
    outputResults(count, now, "initial");
    const auto t1 = std::chrono::steady_clock::now();
    double loop = 0;
    for (size_t k = 1; k < m_timeMoments.size(); ++k) {
        const double target = m_timeMoments[k];
        while (std::abs(now - target) > 1e-10) {
            const auto s0 = std::chrono::steady_clock::now();
            int n = 0;
            double t = now;
            check(wfv_advance_steps(ctx_, now, target, nextOutputCount(count) - count, &n, &t), "run");
            loop += std::chrono::duration<double>(std::chrono::steady_clock::now() - s0).count();
            if (n == 0) break;
            now = t;
            count += n;
            outputResults(count, now);
        }
        outputResults(count, now, "intermediate");
        std::cout << "t = " << now << "  steps " << count << std::endl;
    }
    unsigned flags = 0;
    check(wfv_get_flags(ctx_, &flags), "flags");
    if (flags & WFV_FLAG_NAN_FLUX) std::cerr << "Riemann flux is nan." << std::endl;
    if (flags & WFV_FLAG_BAD_THETA) std::cerr << "Error: illegal theta in PP limiting..." << std::endl;
    std::cout << "\nTotal iterations completed. Total iterations: " << count << std::endl;
    outputResults(count, now, "final");
    writeState(m_outputDir + "final_state.bin");
    const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const double cs = 3.0 * count * m_worldElemNumX * m_worldElemNumY;
    std::cout << std::defaultfloat << "time loop " << loop << " s (" << (loop > 0 ? cs / loop / 1e6 : 0.0)
              << " M cell-stage/s for RK3), total " << total << " s" << std::endl;
    (void)t1;
}

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
