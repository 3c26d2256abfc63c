// wenofv_host.h -- host-side state of solver contexts and batches (library-internal).
//
// Shared by the C ABI (wenofv_capi.cpp) and the launch-shape autotuner (wenofv_tune.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "wenofv.h"
#include "wenofv_kernels.h"

struct wfv_ctx {
    wfv_options opt{};
    wfv_case cs{};
    int nx = 0, ny = 0;
    size_t nc = 0;
    double *reg[3] = {nullptr, nullptr, nullptr};  // RK registers; reg[0] is the state
    double *d_dxC = nullptr, *d_dxW = nullptr, *d_dyC = nullptr;
    double *d_faces = nullptr, *d_theta = nullptr;
    double *d_faces_ring = nullptr, *d_theta_ring = nullptr;  // wfv_face_values, [ny+2][nx+2]
    double *d_cell_dt = nullptr;                               // wfv_cell_dt, [ny][nx]
    double *d_ghost = nullptr;        // fixed ghost ring (boundary 3)
    double *d_theta_flag = nullptr;   // limiter flags of the last stage (when recorded)
    wfv::StepCtl *d_ctl = nullptr;
    double *d_dts = nullptr;
    int dts_cap = 0;
    hipStream_t stream = nullptr;     // stream in use
    hipStream_t own_stream = nullptr; // stream created with the context
    wfv::Phys P{};
    double dX = 0, dY = 0;
    int ty = 16;                      // rows per wavefront strip
    std::vector<double> dxC, dxW, dyC;
    std::string err;
    hipGraphExec_t graph = nullptr;   // captured steps; rebuilt when the launch shape changes
    int graph_steps = 0;
};

struct wfv_batch {
    wfv_options opt{};
    int n = 0;
    std::vector<wfv_instance> inst;
    std::vector<wfv::BatchInst> hinst;  // host copy of the instance table
    std::vector<size_t> cell_off;       // first cell of each instance in the registers
    std::vector<int> ty;                // rows per strip per instance
    int default_ty = 16;                // the batch default of rows per strip
    size_t total_cells = 0;
    int total_items = 0;
    int max_cells = 0;
    double *reg[3] = {nullptr, nullptr, nullptr};
    double *geo = nullptr;              // all instances' dxC, dxW, dyC
    double *ghost = nullptr;            // fixed ghost rings of the boundary-3 instances
    wfv::StepCtl *d_ctl = nullptr;
    wfv::BatchInst *d_inst = nullptr;
    int *d_item_inst = nullptr;         // [total_items] instance (-1: empty), then [total_items] strip
    hipStream_t stream = nullptr, own_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int persistent = 0;                 // schedule: 0 one wavefront per item, 1 persistent grid
    int grid = 0;                       // wavefronts per stage launch
    std::string err;
};

namespace wfv {

#ifndef WFV_BATCH_XCD
#define WFV_BATCH_XCD 1
#endif
#ifndef WFV_BATCH_RUN
#define WFV_BATCH_RUN 16  // strips per run of the batch's XCD-aware order
#endif
// Dispatch order of a batch's (instance, strip) items, XCD-aware.  The hardware hands
// workgroup p to XCD p mod 8, each XCD with its own L2.  prio lists the instances in
// priority order (rows per strip descending: the longest items start first), nstrip and rows
// give each instance's strip count and rows per strip.  Each instance's strips, numbered along
// row bands, are cut into runs of `chunk`; the runs go to the XCDs longest-processing-time
// first by their work (strips x (rows + 1)); each XCD orders its items by instance priority
// and strip, and its k-th item takes position 8k + x.  Vertically adjacent strips thus share
// an L2 except at the ends of runs, and each XCD still starts its longest items first.
// Positions past an XCD's last item hold instance -1 (a wavefront that returns at once).
// With WFV_BATCH_XCD 0: the items in priority order, no padding.  Returns the item count.
inline int batch_item_order(const std::vector<int> &prio, const std::vector<int> &nstrip, const std::vector<int> &rows,
                            int chunk, std::vector<int> &item_inst, std::vector<int> &item_strip) {
    item_inst.clear();
    item_strip.clear();
    if (!WFV_BATCH_XCD) {
        for (int i : prio)
            for (int s = 0; s < nstrip[i]; ++s) {
                item_inst.push_back(i);
                item_strip.push_back(s);
            }
        return (int)item_inst.size();
    }
    constexpr int NXCD = 8;
    struct Run {
        int inst, s0, s1, rank;
        double work;
    };
    std::vector<Run> runs;
    for (int r = 0; r < (int)prio.size(); ++r) {
        const int i = prio[r];
        for (int s = 0; s < nstrip[i]; s += chunk) {
            const int e = s + chunk < nstrip[i] ? s + chunk : nstrip[i];
            runs.push_back({i, s, e, r, (e - s) * (rows[i] + 1.0)});
        }
    }
    std::vector<int> order(runs.size());
    for (size_t k = 0; k < order.size(); ++k) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return runs[a].work > runs[b].work; });
    double load[NXCD] = {};
    std::vector<int> mine[NXCD];
    for (int k : order) {
        int x = 0;
        for (int y = 1; y < NXCD; ++y)
            if (load[y] < load[x]) x = y;
        load[x] += runs[k].work;
        mine[x].push_back(k);
    }
    size_t len = 0;
    for (int x = 0; x < NXCD; ++x) {
        std::sort(mine[x].begin(), mine[x].end(), [&](int a, int b) {
            return runs[a].rank != runs[b].rank ? runs[a].rank < runs[b].rank : runs[a].s0 < runs[b].s0;
        });
        size_t n = 0;
        for (int k : mine[x]) n += runs[k].s1 - runs[k].s0;
        len = n > len ? n : len;
    }
    item_inst.assign(NXCD * len, -1);
    item_strip.assign(NXCD * len, -1);
    for (int x = 0; x < NXCD; ++x) {
        size_t k = 0;
        for (int q : mine[x])
            for (int s = runs[q].s0; s < runs[q].s1; ++s, ++k) {
                item_inst[NXCD * k + x] = runs[q].inst;
                item_strip[NXCD * k + x] = s;
            }
    }
    return (int)item_inst.size();
}

// RK schedule of SSP-RK1/2/3 (src/WENOFV.cpp:376-473): per stage (a, b, c) and registers
struct StagePlan {
    double a, b, c;
    int in, base, out;  // register indices; base < 0: Un = U
};
int rk_plan(int rk, StagePlan *s);

int fail(wfv_ctx *c, int code, const std::string &msg);
// the message wfv_last_error(NULL) returns on this thread (copied into a fixed buffer)
void set_thread_error(const char *msg);
// nullptr if the options and case are valid, else why not
const char *check_options(const wfv_options *opt, const wfv_case *cs);
// equation-of-state parameters of a case
Phys make_phys(const wfv_case &cs);
// stage arguments of a grid without registers, RK coefficients or dt: widths dxC, dxW
// [nx] and dyC [ny] (device), the fixed ghost ring or nullptr, error bits in ctl->flags
StageArgs stage_base(const wfv_options &opt, const wfv_case &cs, int nx, int ny, int ty, const double *dxC,
                     const double *dxW, const double *dyC, const double *ghost, StepCtl *ctl);
// enqueue the stages of one RK step from reg[0] over reg[1], reg[2] (SSP-RK1: reg[1] copied
// back into reg[0]); dt from the host (dt_dev == nullptr) or the device loop (dt_dev,
// active); dtmin (optional) receives the CFL minimum of the new state
// fuse_ctl (optional, SSP-RK2/3 with dtmin): the last stage's last wavefront runs the next
// step's control on it (fused step control)
hipError_t enqueue_rk_step(const StageArgs &base, double *const reg[3], const wfv_options &opt, double dt,
                           const double *dt_dev, const int *active, unsigned long long *dtmin, hipStream_t s,
                           StepCtl *fuse_ctl = nullptr);

// a padded state [4][ny+2G][nx+2G] -> its ghost ring, SoA [4][ring_size(nx, ny)]
inline void pack_ghost_ring(int nx, int ny, const double *Upad, double *ring) {
    const int G = WFV_GHOST, px = nx + 2 * G, py = ny + 2 * G;
    const size_t rs = ring_size(nx, ny), pc = (size_t)px * py;
    for (int v = 0; v < 4; ++v)
        for (int j = -G; j < ny + G; ++j)
            for (int i = -G; i < nx + G; ++i) {
                if (i >= 0 && i < nx && j >= 0 && j < ny) continue;
                ring[v * rs + ghost_offset(nx, ny, i, j)] = Upad[v * pc + (size_t)(j + G) * px + (i + G)];
            }
}
int bfail(wfv_batch *b, int code, const std::string &msg);

// stage arguments of a context without registers, RK coefficients or dt
StageArgs base_args(wfv_ctx *c);
// (re)build a batch's instance table, item list and launch grid from its rows per strip
// and schedule; drops the captured step graph
int batch_layout(wfv_batch *b);

}  // namespace wfv

#define HIPCHK(ctx, call)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return wfv::fail(ctx, WFV_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define BHIPCHK(b, call)                                                                         \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return wfv::bfail(b, WFV_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)
