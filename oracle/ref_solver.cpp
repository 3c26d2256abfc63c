// ref_solver.cpp -- fixture/timing driver over the reference's unmodified 4-rank solver.
//
// TEST INFRASTRUCTURE ONLY, built in the development container by oracle/build_ref.sh
// into oracle/_ref/refsolver (linked with the reference's own src/WENOFV.cpp, compiled
// from where it lies).  It is the builder's own main(): it constructs the reference's
// TwoPhaseEquation and CWENOFV exactly as src/main.cpp:20-35 does, optionally replaces
// the initial cell averages with a state read from a file, and then either
//   rhs   -- calls CWENOFV::assembleRHS once and gathers rhs, limited face values, theta;
//   run   -- runs the reference time loop (src/WENOFV.cpp:318-332: dt = min(CFL dt,
//            T - now) until |now - T| <= 1e-10) or a fixed number of CFL steps;
//   theta -- evaluates getPositivityPreservingTheta on cells built from a file.
// Outputs are raw little-endian doubles written by rank 0.  Run under `mpirun -np 4`
// from a scratch directory: initializeSolver writes ./output/ there.
//
// An optional last argument names a test case whose source term assembleSourceTerm applies
// (src/WENOFV.cpp:1831-1872 keys on the public m_testcase alone): the solver is initialised
// as the periodic `testcase` (the gravity cases' own NEUMANN boundary stops the reference's
// ghost fill, src/WENOFV.cpp:2287-2292, already inside initializeSolver) and m_testcase is
// then set to it, so DAMBREAK, RTI or STANDING_WAVE gravity runs on a periodic grid.
#include <mpi.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Equations.hpp"
#include "WENOFV.hpp"

namespace {

std::vector<double> read_file(const char *path, size_t n) {
    std::vector<double> v(n);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(double), n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", path);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    std::fclose(f);
    return v;
}
void write_file(const char *path, const std::vector<double> &v) {
    FILE *f = std::fopen(path, "wb");
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

// world SoA [4][ny][nx] <-> per-rank m_Uh (x-major with 3 ghost layers)
void scatter_state(CWENOFV &s, const std::vector<double> &W) {
    const int nx = s.m_worldElemNumX, ny = s.m_worldElemNumY, nl = s.m_elemNumX;
    const size_t nc = (size_t)nx * ny;
    for (int li = 0; li < nl; ++li) {
        const int g = s.m_rank * nl + li;
        for (int j = 0; j < ny; ++j)
            for (int r = 0; r < 4; ++r) s.m_Uh[li + 3][j + 3].vector[r] = W[r * nc + (size_t)j * nx + g];
    }
}

// gather per-cell blocks of k doubles (interior cells) to rank 0 as [cell j*nx+i][k]
std::vector<double> gather_cells(CWENOFV &s, int k, const std::vector<double> &local) {
    const int nx = s.m_worldElemNumX, ny = s.m_worldElemNumY, nl = s.m_elemNumX;
    std::vector<double> all;
    if (s.m_rank == 0) all.resize((size_t)nx * ny * k);
    std::vector<double> buf((size_t)nl * ny * k);
    for (int src = 0; src < s.m_size; ++src) {
        if (src == 0 && s.m_rank == 0) buf = local;
        else if (s.m_rank == src) MPI_Send(local.data(), (int)local.size(), MPI_DOUBLE, 0, 7, MPI_COMM_WORLD);
        else if (s.m_rank == 0)
            MPI_Recv(buf.data(), (int)buf.size(), MPI_DOUBLE, src, 7, MPI_COMM_WORLD, MPI_STATUS_IGNORE);
        if (s.m_rank == 0)
            for (int li = 0; li < nl; ++li)
                for (int j = 0; j < ny; ++j) {
                    const size_t c = (size_t)j * nx + src * nl + li;
                    std::memcpy(&all[c * k], &buf[((size_t)li * ny + j) * k], k * sizeof(double));
                }
    }
    return all;
}

std::vector<double> local_state(CWENOFV &s, Array2D<Cvector> &A) {
    const int ny = s.m_worldElemNumY, nl = s.m_elemNumX;
    std::vector<double> v((size_t)nl * ny * 4);
    for (int li = 0; li < nl; ++li)
        for (int j = 0; j < ny; ++j)
            for (int r = 0; r < 4; ++r) v[((size_t)li * ny + j) * 4 + r] = A[li + 3][j + 3].vector[r];
    return v;
}

// [cell][4] -> SoA [4][ny][nx]
std::vector<double> to_soa(const std::vector<double> &cells, size_t nc) {
    std::vector<double> o(4 * nc);
    for (size_t c = 0; c < nc; ++c)
        for (int r = 0; r < 4; ++r) o[r * nc + c] = cells[c * 4 + r];
    return o;
}

}  // namespace

int main(int argc, char **argv) {
    MPI_Init(&argc, &argv);
    // argv: mode testcase nx ny rk cfl scheme pp flux recon t_end|nsteps state_in(or -) out_prefix
    //       [rho10 rho20 sound1 sound2 preb [source_testcase]]
    if (argc < 14) {
        std::fprintf(stderr, "usage: refsolver mode testcase nx ny rk cfl scheme pp flux recon "
                             "tend_or_steps state_in out_prefix [rho10 rho20 c1 c2 preb]\n");
        MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const std::string mode = argv[1];
    std::map<std::string, std::string> option;
    option["TESTCASE"] = argv[2];
    option["ELEMNUMX"] = argv[3];
    option["ELEMNUMY"] = argv[4];
    option["RKMETH"] = argv[5];
    option["CFL"] = argv[6];
    option["SCHEME"] = argv[7];
    option["PPlimiter"] = argv[8];
    option["RIEMANNFLUX"] = argv[9];
    option["RECONSTRUCTIONTYPE"] = argv[10];
    const double tend = std::atof(argv[11]);
    const std::string state_in = argv[12];
    const std::string out = argv[13];

    TwoPhaseEquation *equation = new TwoPhaseEquation();
    CWENOFV solver(equation, option);
    solver.initializeSolver();
    if (argc >= 19) {
        equation->g_rho10 = std::atof(argv[14]);
        equation->g_rho20 = std::atof(argv[15]);
        equation->g_sound1 = std::atof(argv[16]);
        equation->g_sound2 = std::atof(argv[17]);
        equation->g_preb = std::atof(argv[18]);
    }
    if (argc >= 20) solver.m_testcase = (TESTCASETYPE)std::atoi(argv[19]);
    const int nx = solver.m_worldElemNumX, ny = solver.m_worldElemNumY;
    const size_t nc = (size_t)nx * ny;
    if (state_in != "-" && mode != "theta") {
        scatter_state(solver, read_file(state_in.c_str(), 4 * nc));
        solver.exchangeGhostCellsValue();
    }
    {
        std::vector<double> init = gather_cells(solver, 4, local_state(solver, solver.m_Uh));
        if (solver.m_rank == 0) write_file((out + "_init.bin").c_str(), to_soa(init, nc));
    }

    if (mode == "rhs") {
        solver.assembleRHS();
        std::vector<double> r = gather_cells(solver, 4, local_state(solver, solver.m_rhs));
        const int nl = solver.m_elemNumX;
        std::vector<double> fl((size_t)nl * ny * 64), th((size_t)nl * ny);
        for (int li = 0; li < nl; ++li)
            for (int j = 0; j < ny; ++j) {
                double *f = &fl[((size_t)li * ny + j) * 64];
                CgridFlux &G = solver.m_gridFlux[li + 3][j + 3];
                for (int v = 0; v < 4; ++v)
                    for (int g = 0; g < 4; ++g) {
                        f[0 * 16 + v * 4 + g] = G.leftFlux[v][g];
                        f[1 * 16 + v * 4 + g] = G.rightFlux[v][g];
                        f[2 * 16 + v * 4 + g] = G.bottomFlux[v][g];
                        f[3 * 16 + v * 4 + g] = G.topFlux[v][g];
                    }
                th[(size_t)li * ny + j] = solver.m_totalTheta[li + 3][j + 3];
            }
        std::vector<double> F = gather_cells(solver, 64, fl);
        std::vector<double> T = gather_cells(solver, 1, th);
        if (solver.m_rank == 0) {
            write_file((out + "_rhs.bin").c_str(), to_soa(r, nc));
            write_file((out + "_faces.bin").c_str(), F);
            write_file((out + "_thetaflag.bin").c_str(), T);
        }
    } else if (mode == "run" || mode == "steps") {
        std::vector<double> dts;
        double now = 0.0;
        int count = 0;
        MPI_Barrier(MPI_COMM_WORLD);
        const auto t0 = std::chrono::steady_clock::now();
        if (mode == "run") {
            while (std::abs(now - tend) > 1e-10) {
                const double remaining = tend - now;
                double timestep = solver.calculateDeltaT(now);
                timestep = std::min(timestep, remaining);
                solver.performTimeIntegration(timestep);
                now += timestep;
                count++;
                dts.push_back(timestep);
            }
        } else {
            const int nsteps = (int)tend;
            for (int k = 0; k < nsteps; ++k) {
                const double timestep = solver.calculateDeltaT(now);
                solver.performTimeIntegration(timestep);
                now += timestep;
                count++;
                dts.push_back(timestep);
            }
        }
        MPI_Barrier(MPI_COMM_WORLD);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::vector<double> fin = gather_cells(solver, 4, local_state(solver, solver.m_Uh));
        if (solver.m_rank == 0) {
            write_file((out + "_final.bin").c_str(), to_soa(fin, nc));
            write_file((out + "_dts.bin").c_str(), dts);
            std::printf("steps %d t %.17g loop_seconds %.6f\n", count, now, secs);
        }
    } else if (mode == "outputs") {
        // the reference's run() loop (src/WENOFV.cpp:305-373) with its output calls, one output
        // time t_end: the reference writes its own .plt / .csv files under ./output/
        double now = 0.0;
        int count = 0;
        solver.outputResults(count, now, "initial");
        while (std::abs(now - tend) > 1e-10) {
            double timestep = solver.calculateDeltaT(now);
            timestep = std::min(timestep, tend - now);
            solver.performTimeIntegration(timestep);
            now += timestep;
            count++;
            solver.outputResults(count, now);
        }
        solver.outputResults(count, now, "intermediate");
        solver.outputResults(count, now, "final");
        if (solver.m_rank == 0) std::printf("steps %d t %.17g dir %s\n", count, now, solver.m_outputDir.c_str());
    } else if (mode == "theta") {
        // file: ncase records of [U(4) | faces(64) in [side][v][g]]; rank 0 evaluates them
        // one by one in interior cell (3,3).
        const int ncase = (int)tend;
        if (solver.m_rank == 0) {
            std::vector<double> in = read_file(state_in.c_str(), (size_t)ncase * 68);
            std::vector<double> th(ncase);
            for (int k = 0; k < ncase; ++k) {
                const double *rec = &in[(size_t)k * 68];
                for (int r = 0; r < 4; ++r) solver.m_Uh[3][3].vector[r] = rec[r];
                CgridFlux &G = solver.m_gridFlux[3][3];
                for (int v = 0; v < 4; ++v)
                    for (int g = 0; g < 4; ++g) {
                        G.leftFlux[v][g] = rec[4 + 0 * 16 + v * 4 + g];
                        G.rightFlux[v][g] = rec[4 + 1 * 16 + v * 4 + g];
                        G.bottomFlux[v][g] = rec[4 + 2 * 16 + v * 4 + g];
                        G.topFlux[v][g] = rec[4 + 3 * 16 + v * 4 + g];
                    }
                th[k] = solver.getPositivityPreservingTheta(3, 3);
            }
            write_file((out + "_theta.bin").c_str(), th);
        }
    }
    MPI_Finalize();
    return 0;
}
