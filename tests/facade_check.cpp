// facade_check.cpp -- the CWENOFV / TwoPhaseEquation facade used the way code written against
// the reference's classes uses them (include/WENOFV.hpp:108-257, include/Equations.hpp:52-1044
// of the reference): compiled with a plain host C++ compiler against the facade headers
// (csrc/WENOFV.hpp, csrc/Equations.hpp) and linked with libwenofv.so.  Every public member
// and method of the reference's two classes is named in this file with the reference's
// signatures.  Inputs and outputs are raw little-endian float64 files in a directory given
// on the command line; tests/test_facade.py (host) and tests/test_gpu_facade.py write the
// inputs from the reference's fixtures and compare the outputs.
//
//   facade_check point  DIR   point functions on the host (no GPU): the equation of state,
//                             physical and Riemann fluxes, eigenvectors and output variables
//                             of TwoPhaseEquation, CWENOFV::useWENO (WENO-JS, -Z, -EXP) and
//                             CWENOFV::getPositivityPreservingTheta on built cells
//   facade_check solver DIR   the solver on the GPU: assembleRHS, and its steps (getFlux_*,
//                             useBoundPreservingLimiter, assembleBoundaryFaceTerm,
//                             assembleSourceTerm) on the given states; the reference's time
//                             loop through calculateDeltaT / performTimeIntegration; every
//                             other public method once; a NaN time step
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "Equations.hpp"
#include "WENOFV.hpp"

namespace {

std::string g_dir;

std::vector<double> load(const std::string &name, size_t n) {
    std::vector<double> v(n);
    FILE *f = std::fopen((g_dir + "/" + name).c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) throw std::runtime_error("cannot read " + name);
    std::fclose(f);
    return v;
}

void save(const std::string &name, const std::vector<double> &v) {
    FILE *f = std::fopen((g_dir + "/" + name).c_str(), "wb");
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

Array1D<double> vec4(const double *p) {
    Array1D<double> a(4);
    for (int r = 0; r < 4; ++r) a[r] = p[r];
    return a;
}

void set_params(TwoPhaseEquation &eq, const double *p) {
    eq.g_rho10 = p[0];
    eq.g_rho20 = p[1];
    eq.g_sound1 = p[2];
    eq.g_sound2 = p[3];
    eq.g_preb = p[4];
}

// ------------------------------------------------------------------ point functions
int point() {
    const std::vector<double> sz = load("sizes.bin", 8);
    const int neos = (int)sz[0], nst = (int)sz[1], nrie = (int)sz[2], npair = (int)sz[3];
    const int nweno = (int)sz[4], nexp = (int)sz[5], ndx = (int)sz[6], ntheta = (int)sz[7];
    TwoPhaseEquation eq;
    if (eq.getVarNum() != 4 || eq.getVitalVarNum() != 9) return 1;

    // equation of state, maximum eigenvalues, physical fluxes, eigenvectors, output variables
    {
        const auto P = load("eos_params.bin", neos * 5), U = load("eos_U.bin", (size_t)neos * nst * 4);
        std::vector<double> out;
        for (int k = 0; k < neos; ++k) {
            set_params(eq, &P[k * 5]);
            for (int n = 0; n < nst; ++n) {
                const Array1D<double> u = vec4(&U[((size_t)k * nst + n) * 4]);
                out.push_back(eq.getVolumeFrac(u));
                out.push_back(eq.getPressure(u));
                out.push_back(eq.getSoundSpeed(u));
                out.push_back(eq.getMaxEigenValue(u, 1, 0));
                out.push_back(eq.getMaxEigenValue(u, 0, 1));
                Array1D<double> F(4);
                eq.getPhyFlux(u, F, 1, 0);
                for (int r = 0; r < 4; ++r) out.push_back(F[r]);
                eq.getPhyFlux(u, F, 0, 1);
                for (int r = 0; r < 4; ++r) out.push_back(F[r]);
                Array2D<double> M(4, 4);
                for (int d = 0; d < 2; ++d) {
                    const double nx = d == 0 ? 1 : 0, ny = 1 - nx;
                    eq.getLEigenMatrix(u, nx, ny, M);
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j) out.push_back(M[i][j]);
                    eq.getREigenMatrix(u, nx, ny, M);
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j) out.push_back(M(i, j));
                }
                Array1D<double> V(eq.getVitalVarNum());
                eq.getVitalVarVal(u, V);
                for (int i = 0; i < 9; ++i) out.push_back(V[i]);
            }
        }
        save("eos_out.bin", out);
        Array1D<std::string> names(9);
        eq.getVitalVarName(names);
        std::ofstream(g_dir + "/names.txt") << names << "\n";
    }

    // Riemann fluxes through getRiemannFlux (and the two solvers by name)
    {
        const auto P = load("rie_params.bin", nrie * 5);
        const auto UL = load("rie_UL.bin", (size_t)nrie * npair * 4), UR = load("rie_UR.bin", (size_t)nrie * npair * 4);
        std::vector<double> out;
        for (int k = 0; k < nrie; ++k) {
            set_params(eq, &P[k * 5]);
            for (int n = 0; n < npair; ++n) {
                const Array1D<double> l = vec4(&UL[((size_t)k * npair + n) * 4]), r = vec4(&UR[((size_t)k * npair + n) * 4]);
                for (int t = 0; t < 2; ++t)
                    for (int d = 0; d < 2; ++d) {
                        Array1D<double> F(4), G(4);
                        const double nx = d == 0 ? 1 : 0, ny = 1 - nx;
                        eq.getRiemannFlux(l, r, F, nx, ny, t);
                        if (t == 0) eq.getLLFRiemannFlux(l, r, G, nx, ny);
                        else eq.getHLLCRiemannFlux(l, r, G, nx, ny);
                        for (int q = 0; q < 4; ++q) {
                            if (!(F[q] == G[q]) && !(std::isnan(F[q]) && std::isnan(G[q]))) return 2;
                            out.push_back(F[q]);
                        }
                    }
            }
        }
        save("rie_out.bin", out);
    }

    // useWENO: WENO-JS and WENO-Z on the middle five of seven values, WENO-EXP on seven
    {
        CWENOFV s;
        const auto S = load("weno_st.bin", (size_t)nweno * 5);
        std::vector<double> out;
        for (int n = 0; n < nweno; ++n) {
            const double *a = &S[(size_t)n * 5];
            for (SCHEMETYPE sc : {WENO, WENOZ}) {
                s.m_scheme = sc;
                for (int gp = 0; gp < 4; ++gp) out.push_back(s.useWENO(0.0, a[0], a[1], a[2], a[3], a[4], 0.0, gp));
            }
        }
        save("weno_out.bin", out);
        const auto E = load("exp_st.bin", (size_t)nexp * 7), DX = load("exp_dx.bin", ndx);
        out.clear();
        s.m_scheme = WENOEXP;
        for (int d = 0; d < ndx; ++d) {
            s.m_deltaX = DX[d];
            for (int n = 0; n < nexp; ++n) {
                const double *a = &E[(size_t)n * 7];
                for (int gp = 0; gp < 4; ++gp) out.push_back(s.useWENO(a[0], a[1], a[2], a[3], a[4], a[5], a[6], gp));
            }
        }
        save("exp_out.bin", out);
    }

    // getPositivityPreservingTheta on cell (3, 3) built from records [U | faces[side][v][g]]
    {
        std::vector<double> out;
        const auto dxy = load("theta_dxy.bin", 2);
        for (const char *key : {"c1", "c70"}) {
            const auto P = load(std::string("theta_") + key + "_params.bin", 5);
            const auto R = load(std::string("theta_") + key + "_cells.bin", (size_t)ntheta * 68);
            set_params(eq, P.data());
            CWENOFV s(&eq, {});
            s.m_deltaX = dxy[0];
            s.m_deltaY = dxy[1];
            s.m_totalElemNumX = s.m_totalElemNumY = 7;
            s.m_Uh.Resize(7, 7);
            s.m_gridFlux.Resize(7, 7);
            s.m_Uh[3][3].vector.Resize(4);
            CgridFlux &G = s.m_gridFlux[3][3];
            G.leftFlux.Resize(4, 4);
            G.rightFlux.Resize(4, 4);
            G.bottomFlux.Resize(4, 4);
            G.topFlux.Resize(4, 4);
            for (int k = 0; k < ntheta; ++k) {
                const double *rec = &R[(size_t)k * 68];
                for (int r = 0; r < 4; ++r) s.m_Uh[3][3].vector[r] = rec[r];
                for (int v = 0; v < 4; ++v)
                    for (int g = 0; g < 4; ++g) {
                        G.leftFlux[v][g] = rec[4 + 0 * 16 + v * 4 + g];
                        G.rightFlux[v][g] = rec[4 + 1 * 16 + v * 4 + g];
                        G.bottomFlux[v][g] = rec[4 + 2 * 16 + v * 4 + g];
                        G.topFlux[v][g] = rec[4 + 3 * 16 + v * 4 + g];
                    }
                out.push_back(s.getPositivityPreservingTheta(3, 3));
            }
        }
        save("theta_out.bin", out);
    }
    return 0;
}

// ------------------------------------------------------------------ the solver (GPU)
std::map<std::string, std::string> options(int testcase, int n, int rk, double cfl, int scheme, int pp, int flux,
                                           int recon) {
    std::map<std::string, std::string> o;
    o["TESTCASE"] = std::to_string(testcase);
    o["ELEMNUMX"] = std::to_string(n);
    o["ELEMNUMY"] = std::to_string(n);
    o["RKMETH"] = std::to_string(rk);
    o["CFL"] = std::to_string(cfl);
    o["SCHEME"] = std::to_string(scheme);
    o["PPlimiter"] = std::to_string(pp);
    o["RIEMANNFLUX"] = std::to_string(flux);
    o["RECONSTRUCTIONTYPE"] = std::to_string(recon);
    return o;
}

// m_Uh interior <- SoA [4][ny][nx]
void put_state(CWENOFV &s, const std::vector<double> &U) {
    const int nx = s.m_worldElemNumX, ny = s.m_worldElemNumY;
    const size_t nc = (size_t)nx * ny;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int r = 0; r < 4; ++r) s.m_Uh[s.m_startElemX + i][s.m_startElemY + j].vector[r] = U[r * nc + (size_t)j * nx + i];
    s.exchangeGhostCellsValue();
}

std::vector<double> get_cells(CWENOFV &s, Array2D<Cvector> &A) {
    const int nx = s.m_worldElemNumX, ny = s.m_worldElemNumY;
    const size_t nc = (size_t)nx * ny;
    std::vector<double> o(4 * nc);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int r = 0; r < 4; ++r) o[r * nc + (size_t)j * nx + i] = A[s.m_startElemX + i][s.m_startElemY + j].vector[r];
    return o;
}

// face values [ny][nx][side][v][g] of the interior cells, or of the ring too (ring: [ny+2][nx+2])
std::vector<double> get_faces(CWENOFV &s, bool ring) {
    const int nx = s.m_worldElemNumX, ny = s.m_worldElemNumY, e = ring ? 1 : 0;
    std::vector<double> o;
    for (int j = -e; j < ny + e; ++j)
        for (int i = -e; i < nx + e; ++i) {
            const CgridFlux &G = s.m_gridFlux[s.m_startElemX + i][s.m_startElemY + j];
            for (const Array2D<double> *F : {&G.leftFlux, &G.rightFlux, &G.bottomFlux, &G.topFlux})
                for (int v = 0; v < 4; ++v)
                    for (int g = 0; g < 4; ++g) o.push_back((*F)[v][g]);
        }
    return o;
}

std::vector<double> get_flags(CWENOFV &s, Array2D<double> &A) {
    std::vector<double> o;
    for (int j = 0; j < s.m_worldElemNumY; ++j)
        for (int i = 0; i < s.m_worldElemNumX; ++i) o.push_back(A[s.m_startElemX + i][s.m_startElemY + j]);
    return o;
}

// every public method of the reference's CWENOFV not exercised elsewhere, once
void touch_all(CWENOFV &s) {
    s.setBoundaryGPs();
    s.exchangeGhostCellsGridFlux();
    s.setBoundaryAverages2();
    s.setBoundaryGPs1D();
    s.getWorldUh();
    s.worldToProcUh();
    s.getWorldTheta();
    s.getWorldPr();
    s.getWorldDeltaT();
    s.getFlux_conservative();
    s.useFVLimiter();
    s.getFlux_characteristic();
    s.getFlux_primitive();
    s.getFlux_primitive2();
    s.getFlux_primitive3();
    s.outputInterfaceLocationY(0, 0.0);
    s.outputBoundError("check");
    s.outputAve("check");
    s.outputSliceAve("check");
    s.outputAccuracy("check", 0.0);
    s.outputAccuracyAve("check", 0.0);
    s.outputUPoscillation("check");
    s.outputSpecialTestCasebyCell("check");
    s.outputUPoscillationByTime(0.0);
    s.outputPressureByTime(0.0);
    s.outputThincknessByTime(0.0);
    s.outputSymmetryErr(0.0);
    s.outputSpecialTestCasebyTime(0.0);
    s.outputResults(0, 0.0, "initial");
    s.backupProjectFiles();
    const double dt = s.calculateDeltaT(0.0);
    s.outputDeltaT("check");
    s.RunRK3(dt);
    if (s.m_rkMeth != RK3) throw std::runtime_error("RK3 expected");
    bool refused = false;
    try {
        s.RunRK1(dt);
    } catch (const std::invalid_argument &) {
        refused = true;
    }
    if (!refused) throw std::runtime_error("RunRK1 on an RK3 solver was not refused");
    s.initializeAve();
    (void)s.m_procIndexX, (void)s.m_procIndexY, (void)s.m_rank, (void)s.m_size, (void)s.m_numProcessorsX;
    (void)s.m_elemNumX, (void)s.m_elemNumY, (void)s.m_worldEndElemX, (void)s.m_worldEndElemY;
    (void)s.m_globalXL, (void)s.m_globalXR, (void)s.m_globalYL, (void)s.m_globalYR, (void)s.m_gpNum;
    (void)s.m_gpoints_ref[0], (void)s.m_gweights_ref[0], (void)s.m_grids[3][3].m_area;
    (void)s.m_worldGrids[3][3].m_xCenter, (void)s.m_worldGridFlux, (void)s.m_worldTotalPr, (void)s.m_Un;
    (void)s.m_outputDir, (void)s.m_timeMoments, (void)s.m_pplimiter, (void)s.m_cfl;
}

int solver() {
    const std::vector<double> sz = load("sizes.bin", 3);
    const int nrhs = (int)sz[0], n = (int)sz[1], nrun = (int)sz[2];
    // right-hand sides: opts per case = testcase scheme recon flux bp, state [4][n][n]
    const auto opts = load("rhs_opts.bin", (size_t)nrhs * 5);
    for (int k = 0; k < nrhs; ++k) {
        const double *o = &opts[k * 5];
        TwoPhaseEquation eq;
        CWENOFV s(&eq, options((int)o[0], n, 2, 0.4, (int)o[1], (int)o[4], (int)o[3], (int)o[2]));
        s.initializeSolver();
        const auto U = load("rhs_U_" + std::to_string(k) + ".bin", 4 * (size_t)n * n);
        put_state(s, U);
        s.assembleRHS();
        save("rhs_fused_" + std::to_string(k) + ".bin", get_cells(s, s.m_rhs));
        save("faces_fused_" + std::to_string(k) + ".bin", get_faces(s, false));
        save("flags_fused_" + std::to_string(k) + ".bin", get_flags(s, s.m_totalTheta));
        // the reference's assembleRHS, step by step (src/WENOFV.cpp:568-615)
        for (int i = 0; i < s.m_totalElemNumX; i++)
            for (int j = 0; j < s.m_totalElemNumY; j++) s.m_rhs[i][j].vector.setZero();
        switch (s.m_reconstructionType) {
        case CONVERVATIVE: s.getFlux_conservative(); break;
        case CHARACTERISTIC: s.getFlux_characteristic(); break;
        case PRIMITIVE: s.getFlux_primitive(); break;
        case PRIMITIVE2: s.getFlux_primitive2(); break;
        default: s.getFlux_primitive3(); break;
        }
        save("faces_raw_" + std::to_string(k) + ".bin", get_faces(s, true));
        if (s.m_pplimiter) s.useBoundPreservingLimiter();
        s.exchangeGhostCellsGridFlux();
        s.assembleBoundaryFaceTerm();
        s.assembleSourceTerm();
        save("rhs_steps_" + std::to_string(k) + ".bin", get_cells(s, s.m_rhs));
        save("faces_steps_" + std::to_string(k) + ".bin", get_faces(s, false));
        save("flags_steps_" + std::to_string(k) + ".bin", get_flags(s, s.m_totalTheta));
    }
    // the reference's time loop (src/WENOFV.cpp:318-332) through the per-step methods:
    // opts per run = testcase nx rk cfl scheme bp flux recon fixed t_end|steps
    const auto ropts = load("run_opts.bin", (size_t)nrun * 10);
    for (int k = 0; k < nrun; ++k) {
        const double *o = &ropts[k * 10];
        TwoPhaseEquation eq;
        CWENOFV s(&eq, options((int)o[0], (int)o[1], (int)o[2], o[3], (int)o[4], (int)o[5], (int)o[6], (int)o[7]));
        s.option["CFL"] = ([&] { char b[64]; std::snprintf(b, sizeof b, "%.17g", o[3]); return std::string(b); })();
        s.initializeSolver();
        std::vector<double> dts;
        double now = 0.0;
        const double tend = o[9];
        const bool fixed = o[8] != 0;
        for (int count = 0; fixed ? count < (int)tend : std::abs(now - tend) > 1e-10; ++count) {
            double dt = s.calculateDeltaT(now);
            if (!fixed) dt = std::min(dt, tend - now);
            switch (s.m_rkMeth) {
            case RK1: s.RunRK1(dt); break;
            case RK2: s.RunRK2(dt); break;
            default: s.RunRK3(dt); break;
            }
            now += dt;
            dts.push_back(dt);
        }
        save("run_dts_" + std::to_string(k) + ".bin", dts);
        save("run_final_" + std::to_string(k) + ".bin", get_cells(s, s.m_Uh));
        save("run_un_" + std::to_string(k) + ".bin", get_cells(s, s.m_Un));
    }
    // every other method, on the SMOOTH case
    {
        TwoPhaseEquation eq;
        CWENOFV s(&eq, options(0, 16, 2, 0.4, 1, 1, 1, 0));
        s.initializeSolver();
        touch_all(s);
        std::cout << "output dir " << s.m_outputDir << std::endl;
        // a NaN in one cell: calculateDeltaT writes the DeltaT map and stops
        s.m_Uh[s.m_startElemX + 5][s.m_startElemY + 2].vector[0] = std::nan("");
        s.exchangeGhostCellsValue();
        bool stopped = false;
        try {
            s.calculateDeltaT(0.25);
        } catch (const std::runtime_error &e) {
            stopped = std::string(e.what()).find("nan DeltaT") != std::string::npos;
        }
        if (!stopped) throw std::runtime_error("NaN time step not reported");
        std::ofstream(g_dir + "/outdir.txt") << s.m_outputDir << "\n";
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: facade_check point|solver DIR\n");
        return 2;
    }
    g_dir = argv[2];
    try {
        const std::string mode = argv[1];
        if (mode == "point") return point();
        if (mode == "solver") return solver();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 2;
}
