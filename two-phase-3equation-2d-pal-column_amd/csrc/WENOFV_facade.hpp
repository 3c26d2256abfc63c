// WENOFV_facade.hpp -- the CWENOFV methods (declared in WENOFV.hpp), inline: the facade
// is compiled in the caller's translation unit, over the C ABI.  Host arithmetic here (the
// face term's quadrature and divergence, the accuracy norms) follows the reference's
// operation order; like the reference, it is meant to be compiled without floating-point
// contraction (GCC's default on baseline x86-64; add -ffp-contract=off where FMA is enabled).
#pragma once
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>

namespace wfv_facade {
// Gauss-Legendre-4 on [-1,1] (third_party/scTools/GaussQuadratureInRef.hpp:64-75)
inline constexpr double GLP[4] = {-0.861136311594052575223946488893, -0.339981043584856264802665759103,
                                  0.339981043584856264802665759103, 0.861136311594052575223946488893};
inline constexpr double GLWT[4] = {0.347854845137453857373063949222, 0.652145154862546142626936050778,
                                   0.652145154862546142626936050778, 0.347854845137453857373063949222};
// Gauss-Lobatto-4 nodes and weights on [-1,1] (third_party/scTools/GaussQuadratureInRef.hpp:252-262)
inline constexpr double GLB[4] = {-1, -0.447213595499957939281834733746, 0.447213595499957939281834733746, 1};
inline constexpr double GLW[4] = {0.166666666666666666666666666667, 0.833333333333333333333333333333,
                                  0.833333333333333333333333333333, 0.166666666666666666666666666667};

// sc_common::intToString / doubleToString: the default stream conversion
template <class T>
std::string to_str(T v) {
    std::stringstream ss;
    ss << v;
    return ss.str();
}

inline Array1D<double> state4(const double *u) {
    Array1D<double> a(4);
    for (int r = 0; r < 4; ++r) a[r] = u[r];
    return a;
}

// a cell's face values as the limiter takes them, [side][v][g] (left, right, bottom, top)
inline void faces_of(const CgridFlux &G, double F[64]) {
    for (int v = 0; v < 4; ++v)
        for (int g = 0; g < 4; ++g) {
            F[0 * 16 + v * 4 + g] = G.leftFlux[v][g];
            F[1 * 16 + v * 4 + g] = G.rightFlux[v][g];
            F[2 * 16 + v * 4 + g] = G.bottomFlux[v][g];
            F[3 * 16 + v * 4 + g] = G.topFlux[v][g];
        }
}

// BOUNDARYTYPE (include/Equations.hpp:25-32) -> wfv_options.boundary
inline int boundary_kind(BOUNDARYTYPE b) {
    switch (b) {
    case PERIOD: return 0;
    case NoSlip: return 2;
    case NEUMANN2: return 3;
    default: return 1;  // NEUMANN, FREESLIP: reflective slip walls
    }
}
}  // namespace wfv_facade

inline void CWENOFV::check(int rc, const char *what) const {
    if (rc != WFV_OK) throw std::runtime_error(std::string(what) + ": " + wfv_last_error(ctx_));
}

inline int CWENOFV::stoi_key(const char *key) {
    // a missing key makes the reference's std::stoi("") throw (src/WENOFV.cpp:43-63)
    auto it = option.find(key);
    if (it == option.end() || it->second.empty()) throw std::invalid_argument(std::string("missing option ") + key);
    return std::stoi(it->second);
}

inline void CWENOFV::initializeSolver() {
    m_testcase = (TESTCASETYPE)stoi_key("TESTCASE");
    m_rkMeth = (RKMETHOD)stoi_key("RKMETH");
    m_riemannFluxType = (RIEMANNFLUXTYPE)stoi_key("RIEMANNFLUX");
    m_reconstructionType = (RECONSTRUCTIONTYPE)stoi_key("RECONSTRUCTIONTYPE");
    equation->setEquationParameters(m_testcase);
    m_timeMoments = equation->outputTimes;
    // ENDTIME (no reference counterpart): replace the case's output times by {0, ENDTIME}
    if (option.count("ENDTIME")) m_timeMoments = {0.0, std::stod(option["ENDTIME"])};
    m_varNum = equation->getVarNum();
    m_scheme = (SCHEMETYPE)stoi_key("SCHEME");
    m_pplimiter = stoi_key("PPlimiter") != 0;
    auto cfl = option.find("CFL");
    if (cfl == option.end() || cfl->second.empty()) throw std::invalid_argument("missing option CFL");
    m_cfl = std::stod(cfl->second);
    m_ghostCellNum = 3;
    m_worldElemNumX = stoi_key("ELEMNUMX");
    m_worldElemNumY = stoi_key("ELEMNUMY");
    m_precision = option.count("PRECISION") ? std::stoi(option["PRECISION"]) : 0;
    // one rank holding the whole domain (src/WENOFV.cpp:64-88)
    m_rank = 0;
    m_size = 1;
    m_numProcessorsX = 1;
    m_procIndexX = m_procIndexY = 0;
    m_worldStartElemX = m_worldStartElemY = m_startElemX = m_startElemY = m_ghostCellNum;
    m_elemNumX = m_worldElemNumX;
    m_elemNumY = m_worldElemNumY;
    m_worldEndElemX = m_endElemX = m_startElemX + m_elemNumX;
    m_worldEndElemY = m_endElemY = m_startElemY + m_elemNumY;
    m_totalElemNumX = m_elemNumX + 2 * m_ghostCellNum;
    m_totalElemNumY = m_elemNumY + 2 * m_ghostCellNum;
    m_globalXL = equation->xL;
    m_globalXR = equation->xR;
    m_globalYL = equation->yL;
    m_globalYR = equation->yR;
    m_deltaX = (equation->xR - equation->xL) / m_worldElemNumX;
    m_deltaY = (equation->yR - equation->yL) / m_worldElemNumY;
    m_gpNum = 4;
    m_gpoints_ref.Resize(m_gpNum);
    m_gweights_ref.Resize(m_gpNum);
    for (int g = 0; g < m_gpNum; ++g) {
        m_gpoints_ref[g] = wfv_facade::GLB[g];
        m_gweights_ref[g] = wfv_facade::GLW[g];
    }

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
    o.boundary = bc_ = wfv_facade::boundary_kind(equation->boundaryCondition);
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
    dxC_.resize(m_worldElemNumX);
    dxW_.resize(m_worldElemNumX);
    dyC_.resize(m_worldElemNumY);
    check(wfv_grid(ctx_, nullptr, nullptr, dxC_.data(), dxW_.data(), dyC_.data()), "wfv_grid");
    // ./output/<timestamp>_<case>_<scheme>_<reconstruction>_<flux>_<nx>x<ny>/ (src/WENOFV.cpp:100-105)
    m_outputDir = "./output/" + timestamp() + "_" + TESTCASETYPE_STRINGS[m_testcase] + "_" + SCHEMETYPE_STRINGS[m_scheme] + "_" +
                  RECONSTRUCTIONTYPE_STRINGS[m_reconstructionType] + "_" + RIEMANNFLUXTYPE_STRINGS[m_riemannFluxType] +
                  "_" + std::to_string(m_worldElemNumX) + "x" + std::to_string(m_worldElemNumY) + "/";
    if (m_pplimiter) check(wfv_record_limiter_flags(ctx_, 1), "record limiter flags");
    allocate_mirrors();
    initializeAve();
}

// the arrays the reference allocates in initializeSolver that every use needs; the face
// values, right-hand side, RK base and world copies come with the calls that fill them
inline void CWENOFV::allocate_mirrors() {
    const int tx = m_totalElemNumX, ty = m_totalElemNumY;
    m_Uh.Resize(tx, ty);
    for (int ei = 0; ei < tx; ++ei)
        for (int ej = 0; ej < ty; ++ej) m_Uh[ei][ej].vector.Resize(m_varNum);
    // the world grid (src/WENOFV.cpp:125-141); the single rank's grid is the world's
    m_worldGrids.Resize(tx, ty);
    for (int ei = 0; ei < tx; ++ei)
        for (int ej = 0; ej < ty; ++ej) {
            const double xl = equation->xL - m_ghostCellNum * m_deltaX + ei * m_deltaX;
            const double xr = equation->xL - m_ghostCellNum * m_deltaX + (ei + 1) * m_deltaX;
            const double yl = equation->yL - m_ghostCellNum * m_deltaY + ej * m_deltaY;
            const double yr = equation->yL - m_ghostCellNum * m_deltaY + (ej + 1) * m_deltaY;
            m_worldGrids[ei][ej] = Cgrid(xl, xr, yl, yr);
        }
    m_grids = m_worldGrids;
    m_totalTheta.Resize(tx, ty, 0.0);
    m_totalDeltaT.Resize(tx, ty, 0.0);
    m_totalPr.Resize(tx, ty, 0.0);
}

inline void CWENOFV::ensure(Array2D<Cvector> &a) {
    if (a.row() == (unsigned)m_totalElemNumX && a.col() == (unsigned)m_totalElemNumY) return;
    a.Resize(m_totalElemNumX, m_totalElemNumY);
    for (int ei = 0; ei < m_totalElemNumX; ++ei)
        for (int ej = 0; ej < m_totalElemNumY; ++ej) a[ei][ej].vector.Resize(m_varNum, 0.0);
}

inline void CWENOFV::ensure_faces(Array2D<CgridFlux> &a) {
    if (a.row() == (unsigned)m_totalElemNumX && a.col() == (unsigned)m_totalElemNumY) return;
    a.Resize(m_totalElemNumX, m_totalElemNumY);
    for (int ei = 0; ei < m_totalElemNumX; ++ei)
        for (int ej = 0; ej < m_totalElemNumY; ++ej) {
            a[ei][ej].rightFlux.Resize(m_varNum, m_gpNum, 0.0);
            a[ei][ej].leftFlux.Resize(m_varNum, m_gpNum, 0.0);
            a[ei][ej].topFlux.Resize(m_varNum, m_gpNum, 0.0);
            a[ei][ej].bottomFlux.Resize(m_varNum, m_gpNum, 0.0);
        }
}

inline void CWENOFV::ensure(Array2D<double> &a) {
    if (a.row() != (unsigned)m_totalElemNumX || a.col() != (unsigned)m_totalElemNumY)
        a.Resize(m_totalElemNumX, m_totalElemNumY, 0.0);
}

// Ghost layers of a state mirror as the kernels read them (map_index / load_cell of the
// stage kernel): periodic wrap (the reference's exchanges, src/WENOFV.cpp:2175-2269); slip
// walls mirror the interior and negate the momentum normal to each mirror; no-slip walls
// negate both momenta once per mirror; fixed ghost cells (NEUMANN2) hold the initial state.
inline void CWENOFV::fill_ghosts(Array2D<Cvector> &U) {
    const int G = m_ghostCellNum, nx = m_worldElemNumX, ny = m_worldElemNumY;
    auto map = [&](int i, int n, bool &mirrored) {
        mirrored = false;
        if (bc_ == 0) return ((i % n) + n) % n;
        if (i < 0) { i = -1 - i; mirrored = true; }
        if (i >= n) { i = 2 * n - 1 - i; mirrored = true; }
        return i < 0 ? 0 : (i >= n ? n - 1 : i);
    };
    const size_t px = nx + 2 * WFV_GHOST, pc = px * (size_t)(ny + 2 * WFV_GHOST);
    for (int i = -G; i < nx + G; ++i)
        for (int j = -G; j < ny + G; ++j) {
            if (i >= 0 && i < nx && j >= 0 && j < ny) continue;
            double *dst = &U[i + G][j + G].vector[0];
            if (bc_ == 3) {
                const size_t o = (size_t)(j + WFV_GHOST) * px + (i + WFV_GHOST);
                for (int r = 0; r < 4; ++r) dst[r] = ghost_pad_.empty() ? 0.0 : ghost_pad_[r * pc + o];
                continue;
            }
            bool mu, mv;
            const int si = map(i, nx, mu), sj = map(j, ny, mv);
            const double *src = &U[si + G][sj + G].vector[0];
            const bool nu = bc_ == 2 ? mu != mv : mu;
            const bool nv = bc_ == 2 ? mu != mv : mv;
            dst[0] = src[0];
            dst[1] = src[1];
            dst[2] = nu ? -src[2] : src[2];
            dst[3] = nv ? -src[3] : src[3];
        }
}

inline void CWENOFV::push_state() {
    const int G = m_ghostCellNum, nx = m_worldElemNumX, ny = m_worldElemNumY;
    const size_t nc = (size_t)nx * ny;
    soa_.resize(4 * nc);
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int r = 0; r < 4; ++r) soa_[r * nc + (size_t)iy * nx + ix] = m_Uh[ix + G][iy + G].vector[r];
    check(wfv_set_state(ctx_, soa_.data()), "wfv_set_state");
}

inline void CWENOFV::pull_state() {
    const int G = m_ghostCellNum, nx = m_worldElemNumX, ny = m_worldElemNumY;
    const size_t nc = (size_t)nx * ny;
    soa_.resize(4 * nc);
    check(wfv_get_state(ctx_, soa_.data()), "wfv_get_state");
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int r = 0; r < 4; ++r) m_Uh[ix + G][iy + G].vector[r] = soa_[r * nc + (size_t)iy * nx + ix];
    fill_ghosts(m_Uh);
}

inline void CWENOFV::initializeAve() {
    const size_t nc = (size_t)m_worldElemNumX * m_worldElemNumY;
    std::vector<double> U(4 * nc);
    check(wfv_reference_initial_state(m_testcase, m_worldElemNumX, m_worldElemNumY,
                                      (m_worldElemNumX % 4 == 0) ? 4 : 1, U.data()),
          "initializeAve");
    if (equation->boundaryCondition == NEUMANN2) {
        // NEUMANN2 ("boundary from the initial values"): ghost cells hold the initial state
        ghost_pad_.resize(4 * (size_t)(m_worldElemNumX + 2 * WFV_GHOST) * (m_worldElemNumY + 2 * WFV_GHOST));
        check(wfv_reference_initial_state_padded(m_testcase, m_worldElemNumX, m_worldElemNumY,
                                                 (m_worldElemNumX % 4 == 0) ? 4 : 1, ghost_pad_.data()),
              "initializeAve");
        check(wfv_set_ghost_cells(ctx_, ghost_pad_.data()), "wfv_set_ghost_cells");
    }
    check(wfv_set_state(ctx_, U.data()), "wfv_set_state");
    pull_state();
}

// m_totalDeltaT of the device state; on a NaN time step the reference prints, writes the
// DeltaT map and stops (src/WENOFV.cpp:519-531): so does this, by an exception
inline void CWENOFV::cell_dt_map(double now) {
    const int G = m_ghostCellNum, nx = m_worldElemNumX, ny = m_worldElemNumY;
    std::vector<double> cell((size_t)nx * ny);
    check(wfv_cell_dt(ctx_, cell.data()), "cell time steps");
    bool nan = false;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            const double t = cell[(size_t)iy * nx + ix];
            m_totalDeltaT[ix + G][iy + G] = t;
            nan |= std::isnan(t);
        }
    if (nan) {
        std::cout << "Error: exist nan DeltaT" << std::endl;
        outputDeltaT(wfv_facade::to_str(now));
        throw std::runtime_error("Error: exist nan DeltaT");
    }
}

inline double CWENOFV::calculateDeltaT(double now) {
    push_state();
    cell_dt_map(now);
    double dt = 0;
    check(wfv_compute_dt(ctx_, &dt), "calculateDeltaT");
    return dt;
}

inline void CWENOFV::performTimeIntegration(double timestep) {
    // m_Un <- m_Uh (RunRK*: src/WENOFV.cpp:381, 399, 431), then the step on the device
    ensure(m_Un);
    for (int ei = 0; ei < m_totalElemNumX; ++ei)
        for (int ej = 0; ej < m_totalElemNumY; ++ej) m_Un[ei][ej].vector = m_Uh[ei][ej].vector;
    push_state();
    check(wfv_step(ctx_, timestep), "performTimeIntegration");
    pull_state();
    if (m_pplimiter) {
        // the limiter flags of the step's last stage (m_totalTheta)
        flags_.resize((size_t)m_worldElemNumX * m_worldElemNumY);
        check(wfv_get_limiter_flags(ctx_, flags_.data()), "limiter flags");
        for (int ix = 0; ix < m_worldElemNumX; ++ix)
            for (int iy = 0; iy < m_worldElemNumY; ++iy)
                m_totalTheta[ix + m_ghostCellNum][iy + m_ghostCellNum] = flags_[(size_t)iy * m_worldElemNumX + ix];
    }
}

inline void CWENOFV::rk_checked(RKMETHOD m, double dt) {
    if (m != m_rkMeth) throw std::invalid_argument("the context was built for RKMETH " + std::to_string(m_rkMeth));
    performTimeIntegration(dt);
}

inline void CWENOFV::assembleRHS() {
    push_state();
    const int G = m_ghostCellNum, nx = m_worldElemNumX, ny = m_worldElemNumY;
    const size_t nc = (size_t)nx * ny;
    std::vector<double> rhs(4 * nc), faces(64 * nc), theta(nc);
    check(wfv_assemble_rhs(ctx_, rhs.data(), faces.data(), theta.data()), "assembleRHS");
    ensure(m_rhs);
    ensure_faces(m_gridFlux);
    for (int ei = 0; ei < m_totalElemNumX; ++ei)
        for (int ej = 0; ej < m_totalElemNumY; ++ej) m_rhs[ei][ej].vector.setZero();
    m_totalPr.setZero();
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            const size_t c = (size_t)iy * nx + ix;
            for (int r = 0; r < 4; ++r) m_rhs[ix + G][iy + G].vector[r] = rhs[r * nc + c];
            CgridFlux &F = m_gridFlux[ix + G][iy + G];
            const double *f = &faces[c * 64];
            for (int v = 0; v < 4; ++v)
                for (int g = 0; g < 4; ++g) {
                    F.leftFlux[v][g] = f[0 * 16 + v * 4 + g];
                    F.rightFlux[v][g] = f[1 * 16 + v * 4 + g];
                    F.bottomFlux[v][g] = f[2 * 16 + v * 4 + g];
                    F.topFlux[v][g] = f[3 * 16 + v * 4 + g];
                }
            if (m_pplimiter) m_totalTheta[ix + G][iy + G] = theta[c] < 1 - 1e-8 ? 1 : 0;
        }
}

inline void CWENOFV::face_values(RECONSTRUCTIONTYPE recon) {
    push_state();
    const int G = m_ghostCellNum, nx = m_worldElemNumX, ny = m_worldElemNumY;
    std::vector<double> faces((size_t)(nx + 2) * (ny + 2) * 64);
    check(wfv_face_values(ctx_, recon, 0, faces.data(), nullptr), "getFlux");
    ensure_faces(m_gridFlux);
    m_totalPr.setZero();  // as every getFlux_* of the reference does (src/WENOFV.cpp:637)
    for (int ix = -1; ix <= nx; ++ix)
        for (int iy = -1; iy <= ny; ++iy) {
            const double *f = &faces[((size_t)(iy + 1) * (nx + 2) + (ix + 1)) * 64];
            CgridFlux &F = m_gridFlux[ix + G][iy + G];
            for (int v = 0; v < 4; ++v)
                for (int g = 0; g < 4; ++g) {
                    F.leftFlux[v][g] = f[0 * 16 + v * 4 + g];
                    F.rightFlux[v][g] = f[1 * 16 + v * 4 + g];
                    F.bottomFlux[v][g] = f[2 * 16 + v * 4 + g];
                    F.topFlux[v][g] = f[3 * 16 + v * 4 + g];
                }
        }
}

inline double CWENOFV::getPositivityPreservingTheta(int ei, int ej) {
    if (m_gridFlux.row() == 0) throw std::logic_error("getPositivityPreservingTheta: no face values (getFlux_* first)");
    if (ei < 0 || ej < 0 || ei >= m_totalElemNumX || ej >= m_totalElemNumY)
        throw std::out_of_range("getPositivityPreservingTheta: cell outside the grid");
    double F[64], theta = 0;
    wfv_facade::faces_of(m_gridFlux[ei][ej], F);
    const wfv_case c = equation->pointCase();
    check(wfv_point_theta(&c, &m_Uh[ei][ej].vector[0], F, m_deltaX, m_deltaY, &theta), "getPositivityPreservingTheta");
    return theta;
}

inline void CWENOFV::useBoundPreservingLimiter() {
    // src/WENOFV.cpp:1895-1935 over the interior cells, and the same for the ring of ghost
    // cells, whose face values the reference takes from its limited neighbours by the
    // exchange (the kernels limit them with their own theta: equal for periodic grids)
    ensure_faces(m_gridFlux);
    for (int ei = m_startElemX - 1; ei <= m_endElemX; ei++)
        for (int ej = m_startElemY - 1; ej <= m_endElemY; ej++) {
            const bool interior = ei >= m_startElemX && ei < m_endElemX && ej >= m_startElemY && ej < m_endElemY;
            const double theta = getPositivityPreservingTheta(ei, ej);
            if (interior) m_totalTheta[ei][ej] = theta < 1 - 1e-8 ? 1 : 0;
            if (theta >= 0.0 && theta <= 1.0) {
                CgridFlux &F = m_gridFlux[ei][ej];
                for (int r = 0; r != m_varNum; ++r) {
                    const double ub = m_Uh[ei][ej].vector[r];
                    for (int gp = 0; gp != m_gpNum; ++gp) {
                        F.rightFlux[r][gp] = theta * (F.rightFlux[r][gp] - ub) + ub;
                        F.leftFlux[r][gp] = theta * (F.leftFlux[r][gp] - ub) + ub;
                        F.topFlux[r][gp] = theta * (F.topFlux[r][gp] - ub) + ub;
                        F.bottomFlux[r][gp] = theta * (F.bottomFlux[r][gp] - ub) + ub;
                    }
                }
            } else if (interior) {
                std::cout << "Error: illegal theta in PP limiting..." << std::endl;
                std::cout << "Theta=" << theta << std::endl;
            }
        }
}

inline void CWENOFV::useFVLimiter() {
    // src/WENOFV.cpp:2002-2058: a partial density of a face value further than 0.9 times the
    // cell's average from it resets the left, top or bottom face value to the average; such
    // a right face value is only reported
    ensure_faces(m_gridFlux);
    auto far = [&](const Array2D<double> &F, const double *U, int gp) {
        bool lim = false;
        for (int r = 0; r != 2; r++)
            if (std::fabs(F[r][gp] - U[r]) > 0.9 * U[r]) lim = true;
        return lim;
    };
    for (int ei = m_startElemX; ei != m_endElemX; ei++)
        for (int ej = m_startElemY; ej != m_endElemY; ej++) {
            const double *U = &m_Uh[ei][ej].vector[0];
            CgridFlux &G = m_gridFlux[ei][ej];
            for (int gp = 0; gp != m_gpNum; gp++) {
                if (far(G.rightFlux, U, gp))
                    for (int r = 0; r != m_varNum; r++)
                        std::cout << std::scientific << "rightFlux: " << G.rightFlux[r][gp] << " Uh: " << U[r] << std::endl;
                for (Array2D<double> *F : {&G.leftFlux, &G.topFlux, &G.bottomFlux})
                    if (far(*F, U, gp))
                        for (int r = 0; r != m_varNum; r++) (*F)[r][gp] = U[r];
            }
        }
}

inline void CWENOFV::assembleBoundaryFaceTerm() {
    // src/WENOFV.cpp:1670-1829: the Riemann flux at every face node (x faces from the
    // ghost column left of the grid, y faces from the ghost row below it), Gauss-Lobatto
    // quadrature over each face, and the divergence added to m_rhs in the order west, east,
    // south, north; the cell widths are the grid's own (the 4-rank slab widths of the
    // kernels: a west face divides by the west neighbour's width as its rank computes it)
    ensure_faces(m_gridFlux);
    ensure(m_rhs);
    const int G = m_ghostCellNum, nx = m_worldElemNumX, ny = m_worldElemNumY;
    const wfv_case P = equation->pointCase();
    const int ftype = m_riemannFluxType;
    auto flux = [&](const Array2D<double> &A, const Array2D<double> &B, bool xdir, double width, double ave[4]) {
        for (int r = 0; r < 4; ++r) ave[r] = 0.0;
        for (int gp = 0; gp < m_gpNum; ++gp) {
            double UL[4], UR[4], F[4];
            for (int r = 0; r < 4; ++r) {
                UL[r] = A[r][gp];
                UR[r] = B[r][gp];
            }
            const double fnx = xdir ? 1 : 0, fny = xdir ? 0 : 1;
            wfv_point_riemann(&P, ftype, UL, UR, fnx, fny, F);
            for (int r = 0; r < 4; ++r) {
                if (std::isnan(F[r])) std::cerr << "Riemann flux is nan." << std::endl;
                ave[r] += F[r] * (wfv_facade::GLW[gp] * width / 2.0);
            }
        }
    };
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            const int ei = ix + G, ej = iy + G;
            const double dxc = dxC_[ix], dxw = dxW_[ix], dy = dyC_[iy];
            double w[4], e[4], so[4], no[4];
            flux(m_gridFlux[ei - 1][ej].rightFlux, m_gridFlux[ei][ej].leftFlux, true, dy, w);
            flux(m_gridFlux[ei][ej].rightFlux, m_gridFlux[ei + 1][ej].leftFlux, true, dy, e);
            flux(m_gridFlux[ei][ej - 1].topFlux, m_gridFlux[ei][ej].bottomFlux, false, dxc, so);
            flux(m_gridFlux[ei][ej].topFlux, m_gridFlux[ei][ej + 1].bottomFlux, false, dxc, no);
            double *rhs = &m_rhs[ei][ej].vector[0];
            for (int r = 0; r < 4; ++r) {
                rhs[r] += w[r] / dxw / dy;
                rhs[r] += -e[r] / dxc / dy;
                rhs[r] += so[r] / dxc / dy;
                rhs[r] -= no[r] / dxc / dy;
            }
        }
}

inline void CWENOFV::assembleSourceTerm() {
    // src/WENOFV.cpp:1831-1872: gravity g of the case (9.81 for DAMBREAK*, RTI; 2 pi for
    // STANDING_WAVE; none otherwise) on the vertical momentum
    ensure(m_rhs);
    const double g = equation->g_gravity;
    if (g == 0.0) return;
    for (int ei = m_startElemX; ei != m_endElemX; ei++)
        for (int ej = m_startElemY; ej != m_endElemY; ej++) {
            const double rho = m_Uh[ei][ej].vector[0] + m_Uh[ei][ej].vector[1];
            m_rhs[ei][ej].vector[3] -= rho * g;
        }
}

inline double CWENOFV::useWENO(double uavemmm, double uavemm, double uavem, double uave, double uavep, double uavepp,
                        double uaveppp, int gp) {
    if (gp < 0 || gp > 3) throw std::out_of_range("useWENO: gp is 0..3");
    if (m_scheme != WENO && m_scheme != WENOZ && m_scheme != WENOEXP) {
        std::cerr << "the scheme is not supported" << std::endl;
        return 0.0;
    }
    const double s[7] = {uavemmm, uavemm, uavem, uave, uavep, uavepp, uaveppp};
    double out[4];
    wfv_point_weno(m_scheme, s, m_deltaX, out);
    return out[gp];
}

inline void CWENOFV::getWorldUh() {
    pull_state();
    ensure(m_worldUh);
    for (int ei = 0; ei < m_totalElemNumX; ++ei)
        for (int ej = 0; ej < m_totalElemNumY; ++ej) m_worldUh[ei][ej].vector = m_Uh[ei][ej].vector;
}

inline void CWENOFV::worldToProcUh() {
    ensure(m_worldUh);
    for (int ei = m_startElemX; ei < m_endElemX; ++ei)
        for (int ej = m_startElemY; ej < m_endElemY; ++ej) m_Uh[ei][ej].vector = m_worldUh[ei][ej].vector;
    fill_ghosts(m_Uh);
    push_state();
}

inline void CWENOFV::exchangeGhostCellsValue() { fill_ghosts(m_Uh); }

inline void CWENOFV::getWorldTheta() {
    ensure(m_worldTotalTheta);
    m_worldTotalTheta = m_totalTheta;
}

inline void CWENOFV::getWorldPr() {
    ensure(m_worldTotalPr);
    m_worldTotalPr = m_totalPr;
}

inline void CWENOFV::getWorldDeltaT() {
    ensure(m_worldTotalDeltaT);
    m_worldTotalDeltaT = m_totalDeltaT;
}

inline void CWENOFV::outputDeltaT(std::string prefix) {
    // src/WENOFV.cpp:2806-2864
    getWorldDeltaT();
    std::cout << "\noutputing DeltaT..." << std::endl;
    std::filesystem::create_directories(m_outputDir);
    const std::string filename = m_outputDir + "DeltaT_" + prefix + ".plt";
    std::ofstream f(filename);
    if (!f.is_open()) {
        std::cout << "Unable to open file" << std::endl;
        return;
    }
    f << "TITLE=FvSolution" << std::endl;
    f << "VARIABLES=" << "X " << "Y " << "DeltaT" << " " << "nanDeltaT" << " " << std::endl;
    f << "ZONE T=TA, " << "I=" << m_worldElemNumY << ", " << "J=" << m_worldElemNumX << ", "
      << "DATAPACKING = POINT" << std::endl;
    for (int ei = m_ghostCellNum; ei != m_worldElemNumX + m_ghostCellNum; ei++)
        for (int ej = m_ghostCellNum; ej != m_worldElemNumY + m_ghostCellNum; ej++) {
            const double t = m_worldTotalDeltaT[ei][ej];
            f << m_worldGrids[ei][ej].m_xCenter << " " << m_worldGrids[ei][ej].m_yCenter << " ";
            if (std::isnan(t)) f << 0 << " ";
            else f << t << " ";
            f << std::isnan(t) << " " << std::endl;
        }
    std::cout << "The file " << filename << " has been output successfully..." << std::endl;
}

inline void CWENOFV::backupProjectFiles() {
    // src/WENOFV.cpp:3392-3421
    namespace fs = std::filesystem;
    const fs::path out = m_outputDir + "sourceCode/";
    fs::create_directories(out);
    fs::create_directories(out / "output/");
    const char *dirs[] = {"bin", "build", "config", "docs", "include", "src", "makefile", "third_party"};
    for (const char *d : dirs) {
        const fs::path src = fs::path(".") / d;
        if (fs::exists(src))
            fs::copy(src, out / d, fs::copy_options::recursive | fs::copy_options::overwrite_existing);
        else
            std::cerr << "warning: source directory does not exist " << src << std::endl;
    }
}

inline std::vector<double> CWENOFV::outputAccuracy(std::string prefix, double now) {
    snapshot();
    const int nx = m_worldElemNumX, ny = m_worldElemNumY;
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
                    const double gx = xc + 0.5 * m_deltaX * wfv_facade::GLP[gi];
                    const double gy = yc + 0.5 * m_deltaY * wfv_facade::GLP[gj];
                    const double wx = 0.5 * m_deltaX * wfv_facade::GLWT[gi];
                    const double wy = 0.5 * m_deltaY * wfv_facade::GLWT[gj];
                    ex += equation->theVarExact(gx, gy, now) * wx * wy / m_deltaX / m_deltaY;
                }
            double u4[4];
            cell(ei, ej, u4);
            const double e = std::fabs(ex - equation->theVarUh(wfv_facade::state4(u4)));
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

inline void CWENOFV::writeState(const std::string &path) const {
    std::ofstream f(path, std::ios::binary);
    const int hdr[4] = {0x57464f56, 4, m_worldElemNumX, m_worldElemNumY};
    f.write(reinterpret_cast<const char *>(hdr), sizeof hdr);
    f.write(reinterpret_cast<const char *>(soa_.data()), soa_.size() * sizeof(double));
}

inline void CWENOFV::readState(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    int hdr[4];
    f.read(reinterpret_cast<char *>(hdr), sizeof hdr);
    if (!f || hdr[0] != 0x57464f56 || hdr[2] != m_worldElemNumX || hdr[3] != m_worldElemNumY)
        throw std::runtime_error("readState: " + path + " is not a state of this grid");
    soa_.resize(4 * (size_t)m_worldElemNumX * m_worldElemNumY);
    f.read(reinterpret_cast<char *>(soa_.data()), soa_.size() * sizeof(double));
    check(wfv_set_state(ctx_, soa_.data()), "readState");
    pull_state();
}

inline void CWENOFV::run() {
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
            const int rc = wfv_advance_steps(ctx_, now, target, nextOutputCount(count) - count, &n, &t);
            if (rc == WFV_ERR_NONFINITE_DT) {
                pull_state();  // the state whose time step is NaN
                cell_dt_map(t);
            }
            check(rc, "run");
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

