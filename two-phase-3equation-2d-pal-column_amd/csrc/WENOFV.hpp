// WENOFV.hpp -- CWENOFV-compatible C++ facade over the C ABI (include/wenofv.h).
//
// Mirrors the public surface of the reference's solver class (include/WENOFV.hpp:108-257 of
// the reference): the same constructor, option keys, enums, helper classes (Cgrid,
// CgridFlux, Cvector), data members and methods, backed by one single-domain GPU context.
// The facade is header-only (the methods are in WENOFV_facade.hpp and WENOFV_output.hpp) and
// compiles in the caller's translation unit, with the caller's containers (wenofv_array.h);
// libwenofv.so exports the C ABI only.
//
// Where the state lives.  The device holds the solver state; the host arrays mirror it with
// the reference's layout and indices (x-major, 3 ghost layers: m_Uh[ei][ej].vector[r] with
// interior cells ei in [m_startElemX, m_endElemX)), the single domain being rank 0 of 1.
//  - The stepping methods (calculateDeltaT, performTimeIntegration, RunRK1-3) take m_Uh as
//    their input and leave the new state in it, like the reference's; run() advances on the
//    device between its outputs and leaves the final state in m_Uh.
//  - The step-by-step methods work on the mirrors as the reference's do: getFlux_*
//    reconstruct m_Uh's face values (m_gridFlux, interior cells and the ring of ghost cells
//    around them) on the GPU; useBoundPreservingLimiter, assembleBoundaryFaceTerm,
//    assembleSourceTerm, getPositivityPreservingTheta and useWENO take the Riemann fluxes,
//    theta and node values from the library's point functions (wfv_point_*, the all-FP64
//    kernels' arithmetic), so every value equals the value the all-FP64 kernels compute;
//    assembleRHS runs the fused GPU stage.
//  - m_Un, m_rhs, m_gridFlux and the world arrays are allocated by the first call that fills
//    them (a grid's face values take 1 KiB per cell).
// Methods that in the reference move data between MPI ranks become copies between the
// device and the mirrors (getWorld*, worldToProcUh) or ghost fills of the mirrors.
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "Equations.hpp"
#include "wenofv.h"

enum RKMETHOD { RK1 = 0, RK2 = 1, RK3 = 2 };
enum SCHEMETYPE { WENO = 0, WENOZ = 1, WENOEXP = 2 };
enum RECONSTRUCTIONTYPE { CONVERVATIVE = 0, CHARACTERISTIC = 1, PRIMITIVE = 2, PRIMITIVE2 = 3, PRIMITIVE3 = 4,
                          ThincOnlyForZ1 = 5 };

inline const std::string SCHEMETYPE_STRINGS[] = {"WENO", "WENOZ", "WENOEXP"};
inline const std::string RECONSTRUCTIONTYPE_STRINGS[] = {"CONVERVATIVE", "CHARACTERISTIC", "PRIMITIVE",
                                                         "PRIMITIVE2", "PRIMITIVE3"};

// a cell of the grid (include/WENOFV.hpp:54-89)
class Cgrid {
public:
    double m_xLeft = 0, m_xRight = 0, m_xCenter = 0, m_xDistance = 0;
    double m_yLeft = 0, m_yRight = 0, m_yCenter = 0, m_yDistance = 0;
    double m_area = 0;
    Cgrid() = default;
    Cgrid(double xLeft, double xRight, double yLeft, double yRight)
        : m_xLeft(xLeft), m_xRight(xRight), m_yLeft(yLeft), m_yRight(yRight) {
        m_xCenter = (m_xLeft + m_xRight) / 2;
        m_xDistance = m_xRight - m_xLeft;
        m_yCenter = (m_yLeft + m_yRight) / 2;
        m_yDistance = m_yRight - m_yLeft;
        m_area = m_xDistance * m_yDistance;
    }
};

// face values of a cell at the 4 Gauss-Lobatto nodes of each side, [var][gp]
// (include/WENOFV.hpp:91-99)
class CgridFlux {
public:
    Array2D<double> rightFlux, leftFlux, topFlux, bottomFlux;
};

// the conserved variables of a cell (include/WENOFV.hpp:101-106)
class Cvector {
public:
    Array1D<double> vector;
};

class CWENOFV {
public:
    CWENOFV() {}
    CWENOFV(TwoPhaseEquation *equation, std::map<std::string, std::string> option)
        : equation(equation), option(std::move(option)) {}
    ~CWENOFV() { wfv_destroy(ctx_); }
    CWENOFV(const CWENOFV &) = delete;
    CWENOFV &operator=(const CWENOFV &) = delete;

    TwoPhaseEquation *equation = nullptr;
    std::map<std::string, std::string> option;

    // decomposition (one rank: the whole domain) and grid sizes (src/WENOFV.cpp:16-90)
    int m_procIndexX = 0, m_procIndexY = 0;
    int m_rank = 0, m_size = 1;
    int m_worldElemNumX = 0, m_worldElemNumY = 0;
    int m_numProcessorsX = 1;
    int m_ghostCellNum = 3;
    int m_startElemX = 3, m_endElemX = 3, m_startElemY = 3, m_endElemY = 3;
    int m_elemNumX = 0, m_elemNumY = 0;
    int m_totalElemNumX = 0, m_totalElemNumY = 0;
    int m_worldStartElemX = 3, m_worldStartElemY = 3, m_worldEndElemX = 3, m_worldEndElemY = 3;
    int m_varNum = 4;
    double m_globalXL = 0, m_globalXR = 0, m_globalYL = 0, m_globalYR = 0;
    double m_cfl = 0, m_deltaX = 0, m_deltaY = 0;
    int m_gpNum = 4;
    bool m_pplimiter = false;
    std::vector<double> m_timeMoments;
    Array1D<double> m_gpoints_ref, m_gweights_ref;  // Gauss-Lobatto-4 nodes and weights on [-1, 1]

    // host mirrors in the reference's layout, [m_totalElemNumX][m_totalElemNumY]
    Array2D<Cvector> m_rhs, m_Un, m_Uh, m_worldUh;
    Array2D<Cgrid> m_grids, m_worldGrids;
    Array2D<CgridFlux> m_gridFlux, m_worldGridFlux;
    Array2D<double> m_worldTotalTheta, m_totalTheta;    // 1 where the limiter's theta < 1 - 1e-8
    Array2D<double> m_worldTotalDeltaT, m_totalDeltaT;  // CFL time step per cell
    Array2D<double> m_worldTotalPr, m_totalPr;          // zero, as in the reference

    std::string m_outputDir = "./output/";
    TESTCASETYPE m_testcase = SMOOTH;
    SCHEMETYPE m_scheme = WENO;
    RKMETHOD m_rkMeth = RK3;
    RIEMANNFLUXTYPE m_riemannFluxType = LLF;
    RECONSTRUCTIONTYPE m_reconstructionType = CONVERVATIVE;
    int m_precision = 0;  // PRECISION key: 0 all-FP64 (default), 1 mixed, 2 mixed with FP64 WENO, 3 FP64 + FMA

    // initializeSolver (src/WENOFV.cpp:16-230): options, case, grid, mirrors, device context
    void initializeSolver();
    // initializeAve (src/WENOFV.cpp:232-279): 4x4 Gauss-Legendre cell averages into m_Uh
    void initializeAve();

    // performTimeIntegration / RunRK1-3 (src/WENOFV.cpp:376-566): one SSP-RK step of m_Uh
    void performTimeIntegration(double timestep);
    void RunRK1(double deltaT) { rk_checked(RK1, deltaT); }
    void RunRK2(double deltaT) { rk_checked(RK2, deltaT); }
    void RunRK3(double deltaT) { rk_checked(RK3, deltaT); }

    // declared by the reference, not defined there: no-op
    void setBoundaryGPs() {}

    // run (src/WENOFV.cpp:281-374): the time loop over the case's output times with outputs
    void run();
    // assembleRHS (src/WENOFV.cpp:568-615) of m_Uh, on the GPU (the fused stage): m_rhs,
    // m_gridFlux (limited face values), m_totalTheta
    void assembleRHS();
    // calculateDeltaT (src/WENOFV.cpp:474-540) of m_Uh: also m_totalDeltaT; a NaN time step
    // writes the DeltaT file and throws (the reference writes it and exits)
    double calculateDeltaT(double now);
    // getFlux_* (src/WENOFV.cpp:617-1372): m_Uh's face values of the form into m_gridFlux
    void getFlux_conservative() { face_values(CONVERVATIVE); }
    void getFlux_characteristic() { face_values(CHARACTERISTIC); }
    void getFlux_primitive() { face_values(PRIMITIVE); }
    void getFlux_primitive2() { face_values(PRIMITIVE2); }
    void getFlux_primitive3() { face_values(PRIMITIVE3); }
    // assembleBoundaryFaceTerm (src/WENOFV.cpp:1670-1829): Riemann fluxes of m_gridFlux at
    // every face node, quadrature, added to m_rhs
    void assembleBoundaryFaceTerm();
    // assembleSourceTerm (src/WENOFV.cpp:1831-1872): m_rhs[3] -= rho g
    void assembleSourceTerm();
    // useWENO (src/WENOFV.cpp:1873-1893): the node value gp of the scheme from 7 averages
    double useWENO(double uavemmm, double uavemm, double uavem, double uave, double uavep, double uavepp,
                   double uaveppp, int gp);
    // useBoundPreservingLimiter / getPositivityPreservingTheta (src/WENOFV.cpp:1895-2000) on
    // m_Uh and m_gridFlux (ghost ring included, as the reference's exchanged copies are)
    void useBoundPreservingLimiter();
    double getPositivityPreservingTheta(int ei, int ej);
    // useFVLimiter (src/WENOFV.cpp:2002-2058; the reference never calls it): resets the
    // left, top and bottom face values whose partial density departs from the average by
    // more than 0.9 times it to the average, and reports such right face values
    void useFVLimiter();

    // the reference's gathers to rank 0 (src/WENOFV.cpp:2059-2117, 2297-2445)
    void getWorldUh();     // device state -> m_Uh and m_worldUh (ghost cells filled)
    void getWorldTheta();  // m_totalTheta -> m_worldTotalTheta
    void getWorldPr();
    void getWorldDeltaT();
    // m_worldUh -> m_Uh -> device state (the reference's commented-out counterpart)
    void worldToProcUh();

    // ghost fills (src/WENOFV.cpp:1542-1668, 2175-2295): the kernels read ghost cells
    // through the boundary map; on the mirrors these fill m_Uh's ghost layers from its
    // interior (exchangeGhostCellsValue, setBoundaryAverages2).  The face values of the
    // ghost ring come with getFlux_*, so the face-value exchanges have nothing to do.
    void exchangeGhostCellsValue();
    void exchangeGhostCellsGridFlux() {}
    void setBoundaryAverages2() { exchangeGhostCellsValue(); }
    void setBoundaryGPs1D() {}

    // outputs (src/WENOFV.cpp:2447-3391): the reference's files and cadence
    void outputResults(int count, double now, std::string specialCase = "");
    // backupProjectFiles (src/WENOFV.cpp:3392-3421): copies ./bin ./build ./config ./docs
    // ./include ./src ./makefile ./third_party (those present) under m_outputDir/sourceCode/
    void backupProjectFiles();
    // declared by the reference, not defined there: no-ops
    void outputInterfaceLocationY(int, double) {}
    void outputBoundError(std::string) {}
    void outputSpecialTestCasebyTime(double now);
    void outputUPoscillationByTime(double now);
    void outputPressureByTime(double now);
    void outputThincknessByTime(double now);
    void outputSymmetryErr(double now);
    void outputSpecialTestCasebyCell(std::string prefix);
    void outputUPoscillation(std::string prefix);
    void outputSliceAve(std::string prefix);
    void outputAve(std::string filename);
    // outputDeltaT (src/WENOFV.cpp:2806-2864): DeltaT_<prefix>.plt of m_totalDeltaT
    void outputDeltaT(std::string prefix);
    // outputAccuracy (src/WENOFV.cpp:3152-3251): L-inf, L1, L2 of theVarUh against the exact
    // cell averages; returns them and appends the reference's CSV line
    std::vector<double> outputAccuracy(std::string prefix, double now);
    void outputAccuracyAve(std::string prefix, double now);

    // not in the reference: state dump / restore (raw FP64 SoA with a small header) and the
    // device context
    void writeState(const std::string &path) const;
    void readState(const std::string &path);
    double Uh(int ix, int iy, int r) const { return m_Uh[ix + m_ghostCellNum][iy + m_ghostCellNum].vector[r]; }
    wfv_ctx *context() const { return ctx_; }

private:
    wfv_ctx *ctx_ = nullptr;
    int bc_ = 0;                            // wfv_options.boundary of the context
    std::vector<double> dxC_, dxW_, dyC_;   // the context's cell widths
    std::vector<double> soa_;               // state [4][ny][nx] as exchanged with the device
    std::vector<double> flags_;             // limiter flags of the last stage, [ny][nx]
    static std::string timestamp();
    double xCenter(int wi) const;  // world grid cell centres, ghost-offset index
    double yCenter(int wj) const;
    void snapshot();                                  // device state + flags -> host (outputs)
    void cell(int ix, int iy, double U[4]) const;     // interior cell of the snapshot
    int nextOutputCount(int count) const;             // next step count with an output
    void check(int rc, const char *what) const;
    void rk_checked(RKMETHOD m, double dt);
    int stoi_key(const char *key);
    void allocate_mirrors();
    void ensure(Array2D<Cvector> &a);
    void ensure_faces(Array2D<CgridFlux> &a);
    void ensure(Array2D<double> &a);
    void push_state();                          // m_Uh interior -> device
    void pull_state();                          // device -> m_Uh (ghosts filled)
    void fill_ghosts(Array2D<Cvector> &U);      // ghost layers from the interior / ghost ring
    void face_values(RECONSTRUCTIONTYPE recon); // getFlux_*
    void cell_dt_map(double now);               // m_totalDeltaT; DeltaT file and throw on NaN
    std::vector<double> ghost_pad_;             // NEUMANN2 ghost values, [4][ny+8][nx+8]
};

#include "WENOFV_facade.hpp"
#include "WENOFV_output.hpp"
