// WENOFV.hpp -- CWENOFV-compatible C++ facade over the C ABI (include/wenofv.h).
//
// Mirrors the surface of the reference's solver class (include/WENOFV.hpp:108-257 of the
// reference): the same constructor, option keys, enums and stepping methods, backed by
// one single-domain GPU context.  Methods that in the reference move data between MPI
// ranks (ghost exchanges, world gathers) are single-domain no-ops or host copies here,
// because the kernels apply the boundary map themselves.
//
// Host mirrors use the reference's layout only through accessors: Uh(ix, iy, r) reads the
// cell averages of interior cell (ix, iy) after syncState().
#pragma once
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "Equations.hpp"
#include "wenofv.h"

enum RKMETHOD { RK1 = 0, RK2 = 1, RK3 = 2 };
enum SCHEMETYPE { WENO = 0, WENOZ = 1, WENOEXP = 2 };
enum RECONSTRUCTIONTYPE { CONVERVATIVE = 0, CHARACTERISTIC = 1, PRIMITIVE = 2, PRIMITIVE2 = 3, PRIMITIVE3 = 4 };

inline const std::string SCHEMETYPE_STRINGS[] = {"WENO", "WENOZ", "WENOEXP"};
inline const std::string RECONSTRUCTIONTYPE_STRINGS[] = {"CONVERVATIVE", "CHARACTERISTIC", "PRIMITIVE",
                                                         "PRIMITIVE2", "PRIMITIVE3"};

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

    // the reference's configuration members (src/WENOFV.cpp:43-63, 90-91)
    int m_worldElemNumX = 0, m_worldElemNumY = 0;
    int m_varNum = 4, m_ghostCellNum = 3, m_gpNum = 4;
    double m_cfl = 0, m_deltaX = 0, m_deltaY = 0;
    bool m_pplimiter = false;
    std::vector<double> m_timeMoments;
    TESTCASETYPE m_testcase = SMOOTH;
    SCHEMETYPE m_scheme = WENO;
    RKMETHOD m_rkMeth = RK3;
    RIEMANNFLUXTYPE m_riemannFluxType = LLF;
    RECONSTRUCTIONTYPE m_reconstructionType = CONVERVATIVE;
    int m_precision = 0;  // PRECISION key: 0 all-FP64 (default), 1 mixed
    std::string m_outputDir = "./output/";

    // host mirrors, SoA [4][ny][nx]
    std::vector<double> m_Uh, m_rhs, m_faces, m_theta;

    // CWENOFV::initializeSolver (src/WENOFV.cpp:16-230): options, case, grid, device context
    void initializeSolver();
    // CWENOFV::initializeAve (src/WENOFV.cpp:232-279): 4x4 Gauss-Legendre cell averages
    void initializeAve();
    // CWENOFV::run (src/WENOFV.cpp:281-374): time loop over the case's output times
    void run();
    // CWENOFV::calculateDeltaT (src/WENOFV.cpp:474-540)
    double calculateDeltaT(double now);
    // CWENOFV::performTimeIntegration / RunRK1-3 (src/WENOFV.cpp:376-566)
    void performTimeIntegration(double timestep);
    void RunRK1(double dt) { rk_checked(RK1, dt); }
    void RunRK2(double dt) { rk_checked(RK2, dt); }
    void RunRK3(double dt) { rk_checked(RK3, dt); }
    // CWENOFV::assembleRHS (src/WENOFV.cpp:568-615): fills m_rhs, m_faces, m_theta
    void assembleRHS();
    // CWENOFV::getPositivityPreservingTheta (src/WENOFV.cpp:1936-2000) of the last assembleRHS;
    // ei, ej are the reference's ghost-offset world indices (interior starts at 3)
    double getPositivityPreservingTheta(int ei, int ej) const;
    // single-domain counterparts of the MPI exchanges (src/WENOFV.cpp:2175-2295, 1542-1668):
    // the kernels read ghost cells through the boundary map, so nothing is exchanged
    void exchangeGhostCellsValue() {}
    void setBoundaryAverages2() {}
    void exchangeGhostCellsGridFlux() {}
    void setBoundaryGPs1D() {}
    // CWENOFV::getWorldUh (src/WENOFV.cpp:2059-2117): device state -> m_Uh
    void getWorldUh();
    // CWENOFV::outputAccuracy (src/WENOFV.cpp:3152-3251): L-inf, L1, L2 of theVarUh against the
    // exact cell averages; returns them and appends the reference's CSV line
    std::vector<double> outputAccuracy(const std::string &prefix, double now);
    // CWENOFV::outputResults and the output files it writes (src/WENOFV.cpp:2447-3391): the
    // reference's cadence (initial, step 1, every 200 / 1000 steps, output times, final)
    void outputResults(int count, double now, const std::string &specialCase = "");
    void outputAve(const std::string &prefix);
    void outputSliceAve(const std::string &prefix);
    void outputAccuracyAve(const std::string &prefix, double now);
    void outputUPoscillation(const std::string &prefix);
    void outputUPoscillationByTime(double now);
    void outputThincknessByTime(double now);
    void outputSymmetryErr(double now);
    void outputPressureByTime(double now);
    void outputSpecialTestCasebyTime(double now);
    void outputSpecialTestCasebyCell(const std::string &prefix);
    // state dump / restore (raw FP64 SoA with a small header)
    void writeState(const std::string &path) const;
    void readState(const std::string &path);
    double Uh(int ix, int iy, int r) const { return m_Uh[(size_t)r * m_worldElemNumX * m_worldElemNumY + (size_t)iy * m_worldElemNumX + ix]; }
    wfv_ctx *context() const { return ctx_; }

private:
    wfv_ctx *ctx_ = nullptr;
    std::vector<double> m_flags;  // limiter flags of the last stage, [ny][nx] (theta column)
    static std::string timestamp();
    double xCenter(int wi) const;  // world grid cell centres, ghost-offset index
    double yCenter(int wj) const;
    void snapshot();                                  // device state + flags -> host
    void cell(int ix, int iy, double U[4]) const;     // interior cell of the snapshot
    int nextOutputCount(int count) const;             // next step count with an output
    void check(int rc, const char *what) const;
    void rk_checked(RKMETHOD m, double dt);
    int stoi_key(const char *key);
};
