// WENOFV_output.hpp -- the reference's on-disk outputs of a run (CWENOFV methods declared in
// WENOFV.hpp, inline), written from a host snapshot of the device state: Tecplot ASCII
// averages and slices, the accuracy averages, and the per-case time series
// (src/WENOFV.cpp:2447-3151, 3252-3391).  File names, headers, columns and the default stream
// formatting follow the reference's files, so that a run of this build and a run of the
// reference write the same text.
#pragma once
#include <algorithm>
#include <ctime>

namespace wfv_facade {
// std::max as the reference calls it: (a < b) ? b : a, so a NaN second operand is dropped
inline double stdmax(double a, double b) { return (a < b) ? b : a; }

// WENO5threconstruction (include/Utility.hpp:5-95) at node gp, for the pressure probes
inline double weno_js(const double s[5], int gp) {
    const double s7[7] = {0.0, s[0], s[1], s[2], s[3], s[4], 0.0};
    double out[4];
    wfv_point_weno(0, s7, 1.0, out);
    return out[gp];
}

// the reference's derived variables (TwoPhaseEquation::getVitalVarVal, include/Equations.hpp:1005-1027)
inline void vital(const TwoPhaseEquation &eq, const double U[4], double V[9]) {
    Array1D<double> out(9);
    eq.getVitalVarVal(state4(U), out);
    for (int k = 0; k < 9; ++k) V[k] = out[k];
}
inline const char *const VITAL_NAMES[9] = {"RHO1Z1", "RHO2Z2", "Z1", "RHO1", "RHO2", "RHO", "U", "V", "PRE"};

// open a time-series file for appending; a new file starts with its header
inline std::ofstream series(const std::string &fn, const std::string &header) {
    const bool exists = std::filesystem::exists(fn);
    std::ofstream f(fn, std::ios::app);
    if (!exists) f << header;
    return f;
}
}  // namespace wfv_facade

// ------------------------------------------------------------------ the world snapshot

inline std::string CWENOFV::timestamp() {
    // third_party/scTools/Timer.hpp:79-86
    const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    std::stringstream ss;
    ss << std::put_time(std::localtime(&t), "%Y%m%d_%H%M%S");
    return ss.str();
}

// cell centre of world column wi / row wj (ghost-offset indices, src/WENOFV.cpp:127-141)
inline double CWENOFV::xCenter(int wi) const {
    const double xl = equation->xL - m_ghostCellNum * m_deltaX + wi * m_deltaX;
    const double xr = equation->xL - m_ghostCellNum * m_deltaX + (wi + 1) * m_deltaX;
    return (xl + xr) / 2;
}
inline double CWENOFV::yCenter(int wj) const {
    const double yl = equation->yL - m_ghostCellNum * m_deltaY + wj * m_deltaY;
    const double yr = equation->yL - m_ghostCellNum * m_deltaY + (wj + 1) * m_deltaY;
    return (yl + yr) / 2;
}

inline void CWENOFV::snapshot() {
    getWorldUh();
    const size_t nc = (size_t)m_worldElemNumX * m_worldElemNumY;
    flags_.assign(nc, 0.0);  // theta column: limiter flags of the last stage (0 without the limiter)
    if (m_pplimiter) check(wfv_get_limiter_flags(context(), flags_.data()), "limiter flags");
}

inline void CWENOFV::cell(int ix, int iy, double U[4]) const {
    const size_t nc = (size_t)m_worldElemNumX * m_worldElemNumY, c = (size_t)iy * m_worldElemNumX + ix;
    for (int r = 0; r < 4; ++r) U[r] = soa_[r * nc + c];
}

// ------------------------------------------------------------------ cadence

inline void CWENOFV::outputResults(int count, double now, std::string specialCase) {
    // CWENOFV::outputResults (src/WENOFV.cpp:2447-2524)
    const bool acc = m_testcase == SMOOTH || m_testcase == VORTEX;
    if (specialCase == "initial") {
        outputAve("initial");
        if (acc) {
            outputAccuracy("initial", now);
            outputAccuracyAve("initial", now);
        }
        if (m_testcase == SHARP) outputUPoscillation(wfv_facade::to_str(count) + "_" + wfv_facade::to_str(now));
    } else if (specialCase == "final") {
        outputAve("final");
        if (acc) {
            outputAccuracy("final", now);
            outputAccuracyAve("final", now);
        }
        outputSpecialTestCasebyCell("final");
    } else if (specialCase == "intermediate") {
        outputAve(wfv_facade::to_str(now) + "_intermediate");
    } else if (count == 1) {
        outputSpecialTestCasebyTime(now);
        outputAve(wfv_facade::to_str(count) + "_" + wfv_facade::to_str(now));
        if (m_testcase == SHARP) outputUPoscillation(wfv_facade::to_str(count) + "_" + wfv_facade::to_str(now));
        if (acc) outputAccuracy(wfv_facade::to_str(count), now);
    } else {
        if (count % 200 == 0) outputSpecialTestCasebyTime(now);
        if (count % 1000 == 0) outputAve(wfv_facade::to_str(count) + "_" + wfv_facade::to_str(now));
    }
}

inline int CWENOFV::nextOutputCount(int count) const {
    if (count < 1) return 1;
    const bool by_time = m_testcase == SHARP || m_testcase == RTI || m_testcase == DAMBREAK;
    const int every = by_time ? 200 : 1000;
    return (count / every + 1) * every;
}

inline void CWENOFV::outputSpecialTestCasebyTime(double now) {
    // src/WENOFV.cpp:3363-3380
    if (m_testcase == SHARP) {
        outputUPoscillationByTime(now);
        outputThincknessByTime(now);
    } else if (m_testcase == RTI) {
        outputSymmetryErr(now);
    } else if (m_testcase == DAMBREAK) {
        outputPressureByTime(now);
    }
}

inline void CWENOFV::outputSpecialTestCasebyCell(std::string prefix) {
    if (m_testcase == SHARP) outputUPoscillation(prefix);  // src/WENOFV.cpp:3381-3391
}

// ------------------------------------------------------------------ whole-field files

inline void CWENOFV::outputAve(std::string prefix) {
    // CWENOFV::outputAve (src/WENOFV.cpp:2525-2601)
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    const std::string fn = m_outputDir + "average_" + prefix + ".plt";
    std::ofstream f(fn);
    f << "TITLE=FvSolution" << std::endl;
    f << "VARIABLES=X Y ";
    for (const char *n : wfv_facade::VITAL_NAMES) f << n << " ";
    f << "theta Pr " << std::endl;
    f << "ZONE T=TA, I=" << m_worldElemNumY << ", J=" << m_worldElemNumX << ", DATAPACKING = POINT" << std::endl;
    double U[4], V[9];
    for (int ix = 0; ix < m_worldElemNumX; ++ix)
        for (int iy = 0; iy < m_worldElemNumY; ++iy) {
            cell(ix, iy, U);
            wfv_facade::vital(*equation, U, V);
            f << xCenter(ix + m_ghostCellNum) << " " << yCenter(iy + m_ghostCellNum) << " ";
            for (double v : V) f << v << " ";
            f << flags_[(size_t)iy * m_worldElemNumX + ix] << " " << 0.0 << " " << std::endl;  // Pr: always zero
        }
}

inline void CWENOFV::outputSliceAve(std::string prefix) {
    // CWENOFV::outputSliceAve (src/WENOFV.cpp:2668-2738): the row above the first
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    std::ofstream f(m_outputDir + "slice_" + prefix + ".plt");
    f << "TITLE=FvSolution_Slice" << std::endl;
    f << "VARIABLES=X ";
    for (const char *n : wfv_facade::VITAL_NAMES) f << n << " ";
    f << "theta Pr " << std::endl;
    f << "ZONE T=TA, I=" << m_worldElemNumX << ", DATAPACKING = POINT" << std::endl;
    const int iy = 1;
    double U[4], V[9];
    for (int ix = 0; ix < m_worldElemNumX; ++ix) {
        cell(ix, iy, U);
        wfv_facade::vital(*equation, U, V);
        f << xCenter(ix + m_ghostCellNum) << " ";
        for (double v : V) f << v << " ";
        f << flags_[(size_t)iy * m_worldElemNumX + ix] << " " << 0.0 << " " << std::endl;
    }
}

inline void CWENOFV::outputAccuracyAve(std::string prefix, double now) {
    // CWENOFV::outputAccuracyAve (src/WENOFV.cpp:3252-3362): computed and exact averages of the
    // checked variable per cell
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    std::ofstream f(m_outputDir + "average_accuracy_" + prefix + ".plt");
    f << "TITLE=FvSolution" << std::endl;
    f << "VARIABLES=X Y num real err " << std::endl;
    f << "ZONE T=TA, I=" << m_worldElemNumY << ", J=" << m_worldElemNumX << ", DATAPACKING = POINT" << std::endl;
    double U[4];
    for (int ix = 0; ix < m_worldElemNumX; ++ix)
        for (int iy = 0; iy < m_worldElemNumY; ++iy) {
            const double xc = xCenter(ix + m_ghostCellNum), yc = yCenter(iy + m_ghostCellNum);
            double ex = 0.0;
            for (int gi = 0; gi < 4; ++gi)
                for (int gj = 0; gj < 4; ++gj) {
                    const double gx = xc + 0.5 * m_deltaX * wfv_facade::GLP[gi];
                    const double gy = yc + 0.5 * m_deltaY * wfv_facade::GLP[gj];
                    const double wx = 0.5 * m_deltaX * wfv_facade::GLWT[gi];
                    const double wy = 0.5 * m_deltaY * wfv_facade::GLWT[gj];
                    ex += equation->theVarExact(gx, gy, now) * wx * wy / m_deltaX / m_deltaY;
                }
            cell(ix, iy, U);
            const double num = equation->theVarUh(wfv_facade::state4(U));
            f << xc << " " << yc << " " << num << " " << ex << " " << std::fabs(num - ex) << " " << std::endl;
        }
}

inline void CWENOFV::outputUPoscillation(std::string prefix) {
    // CWENOFV::outputUPoscillation (src/WENOFV.cpp:2739-2805): departures from the SHARP case's
    // uniform (u, v) = (1, 1) and p = 0, scaled by the phase-1 sound speed
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    std::ofstream f(m_outputDir + "UPoscillation_" + prefix + ".plt");
    f << "TITLE=FvSolution" << std::endl;
    f << "VARIABLES=X Y Speed_oscillation P_oscillation U_oscillation V_oscillation ";
    f << "\nZONE T=TA, I=" << m_worldElemNumY << ", J=" << m_worldElemNumX << ", DATAPACKING = POINT" << std::endl;
    const double c1 = equation->g_sound1;
    double U[4];
    for (int ix = 0; ix < m_worldElemNumX; ++ix)
        for (int iy = 0; iy < m_worldElemNumY; ++iy) {
            cell(ix, iy, U);
            const double u = U[2] / (U[0] + U[1]), v = U[3] / (U[0] + U[1]);
            f << xCenter(ix + m_ghostCellNum) << " " << yCenter(iy + m_ghostCellNum) << " ";
            f << std::fabs(std::sqrt(std::pow(u - 1, 2) + std::pow(v - 1, 2))) / c1 << " ";
            f << std::fabs(equation->getPressure(U)) / equation->g_rho10 / std::pow(c1, 2) << " ";
            f << std::fabs(u - 1) / c1 << " ";
            f << std::fabs(v - 1) / c1 << " ";
            f << std::endl;
        }
}

// ------------------------------------------------------------------ time series

inline void CWENOFV::outputUPoscillationByTime(double now) {
    // src/WENOFV.cpp:2865-2941: largest departures over the cells
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    std::ofstream f = wfv_facade::series(m_outputDir + "oscillationByTime_" + wfv_facade::to_str(m_worldElemNumX) + ".plt",
                             "TITLE=FvSolution\nVARIABLES=Time Speed_oscillation P_oscillation U_oscillation "
                             "V_oscillation \nZONE T=TA\n");
    const double c1 = equation->g_sound1;
    double pmax = 0, smax = 0, umax = 0, vmax = 0, U[4];
    for (int ix = 0; ix < m_worldElemNumX; ++ix)
        for (int iy = 0; iy < m_worldElemNumY; ++iy) {
            cell(ix, iy, U);
            const double pre = std::fabs(equation->getPressure(U)) / equation->g_rho10 / std::pow(c1, 2);
            const double spd = std::sqrt(std::pow(U[2] / (U[0] + U[1]) - 1, 2) + std::pow(U[3] / (U[0] + U[1]) - 1, 2)) / c1;
            const double uo = (U[2] / (U[0] + U[1]) - 1) / c1;
            const double vo = (U[3] / (U[0] + U[1]) - 1) / c1;
            pmax = wfv_facade::stdmax(pmax, pre);
            smax = wfv_facade::stdmax(smax, spd);
            umax = wfv_facade::stdmax(umax, uo);
            vmax = wfv_facade::stdmax(vmax, vo);
        }
    f << now << " " << pmax << " " << smax << " " << umax << " " << vmax << std::endl;
}

inline void CWENOFV::outputThincknessByTime(double now) {
    // src/WENOFV.cpp:3049-3108: 1 / largest jump of z1 to the east or north neighbour.  The
    // reference reads the cells past the last column and row from its never-filled world
    // ghost cells (z1 = NaN there), which std::max drops, together with the north jump of the
    // last column; the same skips are made here.
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    std::ofstream f = wfv_facade::series(m_outputDir + "thincknessByTime.plt", "TITLE=FvSolution\nVARIABLES=time thinckness\n\nZONE T=TA\n");
    const int nx = m_worldElemNumX, ny = m_worldElemNumY;
    auto z1 = [&](int ix, int iy) {
        if (ix >= nx || iy >= ny) return std::numeric_limits<double>::quiet_NaN();
        double U[4];
        cell(ix, iy, U);
        return equation->getVolumeFrac(U);
    };
    double maxDiff = 0;
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy) {
            const double d1 = std::fabs(z1(ix, iy) - z1(ix + 1, iy));
            const double d2 = std::fabs(z1(ix, iy) - z1(ix, iy + 1));
            maxDiff = wfv_facade::stdmax(maxDiff, wfv_facade::stdmax(d1, d2));
        }
    f << now << " " << 1.0 / maxDiff << std::endl;
}

inline void CWENOFV::outputSymmetryErr(double now) {
    // src/WENOFV.cpp:3109-3151: largest |z1 - z1 of the mirror cell about x = (xL+xR)/2| over
    // the left half (the reference starts the maximum from an uninitialised value; 0 here)
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    std::ofstream f(m_outputDir + "symmetry_err.plt", std::ios::app);
    double maxErr = 0, U[4], Us[4];
    for (int ix = 0; ix < m_worldElemNumX / 2; ++ix)
        for (int iy = 0; iy < m_worldElemNumY; ++iy) {
            cell(ix, iy, U);
            cell(m_worldElemNumX - 1 - ix, iy, Us);
            maxErr = wfv_facade::stdmax(std::fabs(equation->getVolumeFrac(U) - equation->getVolumeFrac(Us)), maxErr);
        }
    f << now << " " << maxErr << std::endl;
}

inline void CWENOFV::outputPressureByTime(double now) {
    // src/WENOFV.cpp:2942-3048: scaled pressure at the left wall at heights 3, 15, 30 and 80 mm,
    // WENO-reconstructed at x = 0 and at the Gauss-Lobatto node nearest each height.  Cells
    // beyond the wall are the wall's mirror images (the reference reads unfilled ghost cells).
    snapshot();
    std::filesystem::create_directories(m_outputDir);
    std::ofstream f = wfv_facade::series(m_outputDir + "pressureByTime.plt",
                             "TITLE=FvSolution\nVARIABLES=Time P_y3mm P_y15mm P_y30mm P_y80mm\n\nZONE T=TA\n");
    const int nx = m_worldElemNumX, ny = m_worldElemNumY;
    auto pres = [&](int ix, int iy) {
        bool mx = false;
        if (ix < 0) { ix = -1 - ix; mx = true; }
        if (iy < 0) iy = -1 - iy;
        if (iy >= ny) iy = 2 * ny - 1 - iy;
        ix = std::min(ix, nx - 1);
        double U[4];
        cell(ix, iy, U);
        if (mx) U[2] = -U[2];
        return equation->getPressure(U) / equation->g_rho10 / 9.81 / 0.3;
    };
    f << now * std::sqrt(9.81 / 0.3);
    for (double y_mm : {3.0, 15.0, 30.0, 80.0}) {
        const double y_m = y_mm / 1000.0;
        const int iy = (int)std::floor(y_m / m_deltaY);
        double atx0[5];
        for (int j = 0; j < 5; ++j) {
            double s[5];
            for (int i = 0; i < 5; ++i) s[i] = pres(i - 2, iy + j - 2);
            atx0[j] = wfv_facade::weno_js(s, 0);
        }
        double best = std::numeric_limits<double>::max(), pre = 0;
        for (int g = 0; g < 4; ++g) {
            const double py = yCenter(iy + m_ghostCellNum) + 0.5 * m_deltaY * wfv_facade::GLB[g];
            const double diff = std::fabs(py - y_m);
            if (diff < best) {
                best = diff;
                pre = wfv_facade::weno_js(atx0, g);
            }
        }
        f << " " << pre;
    }
    f << std::endl;
}
