// wenofv_tune.cpp -- launch-shape autotuner of the stage kernels.
//
// A stage kernel has one free launch parameter: the number of rows a wavefront marches
// (rows per strip).  Long strips amortise the two halo rows that every strip reconstructs
// again; short strips give more wavefronts to fill the device.  A batch also has a
// schedule: one wavefront per (instance, strip) item, or a persistent grid of eight
// wavefronts per CU striding over the items.  The tuner times the first RK stage on the
// live state (it reads register 0 and writes register 1, so the state is left unchanged)
// for every candidate with HIP events on the solver's stream and keeps the fastest.  A
// batch is tuned by two passes of coordinate descent over its grid-size classes (instances
// of equal nx x ny share rows per strip), largest class first, on the whole one-launch stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "wenofv.h"
#include "wenofv_geometry.h"
#include "wenofv_host.h"
#include "wenofv_kernels.h"

using namespace wfv;

namespace {

constexpr int CANDIDATES[] = {1, 2, 4, 8, 12, 16, 24, 32, 48, 64, 96, 128};
constexpr int REPEATS = 3;

// candidate rows per strip for a grid of ny rows
std::vector<int> row_candidates(int ny) {
    std::vector<int> c;
    for (int r : CANDIDATES)
        if (r < ny) c.push_back(r);
    c.push_back(std::min(ny, 128));
    c.erase(std::unique(c.begin(), c.end()), c.end());
    return c;
}

// Best-of-REPEATS mean duration (ms) of `trials` launches issued by `launch` on `s`.
template <class L>
hipError_t time_launches(hipStream_t s, int trials, L launch, float *ms) {
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return e;
    }
    float best = 1e30f;
    e = launch();  // warm-up
    for (int r = 0; r < REPEATS && e == hipSuccess; ++r) {
        e = hipEventRecord(e0, s);
        for (int t = 0; t < trials && e == hipSuccess; ++t) e = launch();
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float el = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&el, e0, e1);
        if (e == hipSuccess) best = std::min(best, el / trials);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = best;
    return e;
}

hipError_t time_ctx_stage(wfv_ctx *c, int trials, float *ms) {
    StageArgs A = base_args(c);
    // the timed stages leave the recorded limiter flags and the error bits alone
    A.theta_flag = nullptr;
    A.flags = nullptr;
    A.U = c->reg[0];
    A.Un = nullptr;
    A.Uo = c->reg[1];
    A.a = 1.0;
    A.b = 0.0;
    A.c = 1.0;
    A.dt = 0.0;
    return time_launches(c->stream, trials,
                         [&] { return launch_stage(A, c->opt.precision, c->opt.reconstruction, c->opt.scheme, c->stream); },
                         ms);
}

hipError_t time_batch_stage(wfv_batch *b, int trials, float *ms) {
    BatchStageArgs B{};
    B.inst = b->d_inst;
    B.item_inst = b->d_item_inst;
    B.item_strip = b->d_item_inst + b->total_items;
    B.total_items = b->total_items;
    B.in = 0;
    B.base = -1;
    B.out = 1;
    B.a = 1.0;
    B.b = 0.0;
    B.c = 1.0;
    B.bp = b->opt.pp_limiter;
    B.ghost = b->ghost != nullptr;
    B.want_dtmin = 0;
    return time_launches(b->stream, trials, [&] {
        return launch_batch_stage(B, b->opt.precision, b->opt.reconstruction, b->opt.scheme, b->opt.riemann_flux,
                                  b->grid, b->stream);
    }, ms);
}

}  // namespace

extern "C" {

int wfv_set_rows_per_strip(wfv_ctx *c, int rows) {
    if (!c || rows < 0) return WFV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // 0: the default; strips longer than the grid are the whole grid
    c->ty = rows > 0 ? std::min(rows, c->ny) : default_rows_per_strip(c->nx, c->ny);
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    return WFV_OK;
}

int wfv_get_rows_per_strip(const wfv_ctx *c, int *rows) {
    if (!c || !rows) return WFV_ERR_ARG;
    *rows = c->ty;
    return WFV_OK;
}

int wfv_autotune(wfv_ctx *c, int trials, wfv_tuning *best) {
    if (!c || trials < 1) return WFV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    StepCtl saved{};
    HIPCHK(c, hipMemcpy(&saved, c->d_ctl, sizeof saved, hipMemcpyDeviceToHost));
    const int start = c->ty;
    int best_ty = start;
    float best_ms = 1e30f;
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    std::vector<int> cand = row_candidates(c->ny);
    int occ[8];
    const int nocc = occupancy_candidates(c->nx, c->ny, strip_halo(c->opt.scheme, c->opt.reconstruction), cus, occ);
    cand.insert(cand.end(), occ, occ + nocc);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    for (int ty : cand) {
        c->ty = ty;
        float ms;
        hipError_t e = time_ctx_stage(c, trials, &ms);
        if (e != hipSuccess) {
            c->ty = start;
            return fail(c, WFV_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        }
        if (ms < best_ms) {
            best_ms = ms;
            best_ty = ty;
        }
    }
    c->ty = best_ty;
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    HIPCHK(c, hipMemcpy(c->d_ctl, &saved, sizeof saved, hipMemcpyHostToDevice));
    if (best) {
        best->rows_per_strip = best_ty;
        best->persistent = 0;
        best->stage_ms = best_ms;
    }
    return WFV_OK;
}

int wfv_batch_get_rows_per_strip(const wfv_batch *b, int i, int *rows) {
    if (!b || !rows || i < 0 || i >= b->n) return WFV_ERR_ARG;
    *rows = b->ty[i];
    return WFV_OK;
}

int wfv_batch_set_schedule(wfv_batch *b, int persistent) {
    if (!b || persistent < 0 || persistent > 1) return WFV_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    b->persistent = persistent;
    return batch_layout(b);
}

int wfv_batch_get_schedule(const wfv_batch *b, int *persistent) {
    if (!b || !persistent) return WFV_ERR_ARG;
    *persistent = b->persistent;
    return WFV_OK;
}

int wfv_batch_autotune(wfv_batch *b, int trials, wfv_tuning *best) {
    if (!b || trials < 1) return WFV_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    // every instance takes part in the timed launches, whatever its loop state
    std::vector<StepCtl> saved(b->n), armed;
    BHIPCHK(b, hipMemcpy(saved.data(), b->d_ctl, b->n * sizeof(StepCtl), hipMemcpyDeviceToHost));
    armed = saved;
    for (auto &s : armed) s.active = 1;
    BHIPCHK(b, hipMemcpy(b->d_ctl, armed.data(), b->n * sizeof(StepCtl), hipMemcpyHostToDevice));

    // grid-size classes, largest total first
    std::map<std::pair<int, int>, std::vector<int>> cls;
    for (int i = 0; i < b->n; ++i) cls[{b->inst[i].nx, b->inst[i].ny}].push_back(i);
    std::vector<std::pair<long long, std::vector<int>>> order;
    for (auto &kv : cls) order.push_back({(long long)kv.first.first * kv.first.second * kv.second.size(), kv.second});
    std::stable_sort(order.begin(), order.end(), [](const auto &p, const auto &q) { return p.first > q.first; });

    auto measure = [&](float *ms) -> int {
        int rc = batch_layout(b);
        if (rc != WFV_OK) return rc;
        hipError_t e = time_batch_stage(b, trials, ms);
        if (e != hipSuccess) return bfail(b, WFV_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
        return WFV_OK;
    };
    float best_ms;
    int rc = measure(&best_ms);
    // two passes: the first tunes each class with the later ones still at the default
    for (size_t k = 0; k < 2 * order.size() && rc == WFV_OK; ++k) {
        const std::vector<int> &members = order[k % order.size()].second;
        const int ny = b->inst[members[0]].ny;
        int keep = b->ty[members[0]];
        for (int ty : row_candidates(ny)) {
            if (ty == keep || rc != WFV_OK) continue;
            for (int i : members) b->ty[i] = ty;
            float ms;
            rc = measure(&ms);
            if (rc == WFV_OK && ms < best_ms) {
                best_ms = ms;
                keep = ty;
            }
        }
        for (int i : members) b->ty[i] = keep;
    }
    // schedule
    for (int p = 0; p < 2 && rc == WFV_OK; ++p) {
        if (p == b->persistent) continue;
        const int keep = b->persistent;
        b->persistent = p;
        float ms;
        rc = measure(&ms);
        if (rc == WFV_OK && ms < best_ms) best_ms = ms;
        else b->persistent = keep;
    }
    if (rc == WFV_OK) rc = batch_layout(b);
    BHIPCHK(b, hipMemcpy(b->d_ctl, saved.data(), b->n * sizeof(StepCtl), hipMemcpyHostToDevice));
    if (rc != WFV_OK) return rc;
    if (best) {
        best->rows_per_strip = 0;  // per instance: wfv_batch_get_rows_per_strip
        best->persistent = b->persistent;
        best->stage_ms = best_ms;
    }
    return WFV_OK;
}

}  // extern "C"
