// wenofv_capi.cpp -- C ABI of the single-domain WENOFV solver (include/wenofv.h).
//
// Host side of the time-stepping path: option validation, grid geometry, the three RK
// registers in HBM, stage scheduling and the device-resident time loop replayed from a
// HIP graph.  All arithmetic runs in the kernels of wenofv_kernels.hip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "wenofv.h"
#include "wenofv_geometry.h"
#include "wenofv_host.h"
#include "wenofv_kernels.h"

using namespace wfv;


namespace {
// failure of the last call without a context on this thread (wfv_last_error(NULL)); a fixed
// buffer, so that recording it allocates nothing
thread_local char g_create_error[256];
}  // namespace

namespace wfv {

void set_thread_error(const char *msg) {
    std::snprintf(g_create_error, sizeof g_create_error, "%s", msg ? msg : "");
}

const char *check_options(const wfv_options *opt, const wfv_case *cs) {
    if (opt->elem_num_x < 1 || opt->elem_num_y < 1) return "ELEMNUMX/ELEMNUMY must be positive";
    if (opt->elem_num_x < 3 || opt->elem_num_y < 3) return "grids need at least 3 cells per direction";
    if (opt->rk_method < 0 || opt->rk_method > 2) return "RKMETH must be 0, 1 or 2";
    if (opt->scheme < 0 || opt->scheme > 2) return "SCHEME must be 0 (WENO-JS), 1 (WENO-Z) or 2 (WENO-EXP)";
    if (opt->riemann_flux < 0 || opt->riemann_flux > 1) return "RIEMANNFLUX must be 0 or 1";
    if (opt->reconstruction < 0 || opt->reconstruction > 4) return "RECONSTRUCTIONTYPE must be 0..4";
    if (opt->precision < 0 || opt->precision > 3)
        return "precision must be 0 (fp64), 1 (mixed), 2 (mixed, reconstruction in fp64) or 3 (fp64 with fused "
               "multiply-add)";
    if (opt->precision == 3 && (opt->scheme == 2 || opt->reconstruction == 4))
        return "precision 3 has the 5-point schemes only (SCHEME 0-1, RECONSTRUCTIONTYPE 0-3)";
    if (opt->boundary < 0 || opt->boundary > 3)
        return "boundary must be 0 (periodic), 1 (slip walls), 2 (no-slip walls) or 3 (fixed ghost cells)";
    if (!(opt->cfl > 0)) return "CFL must be positive";
    if (opt->geometry_ranks < 0 || (opt->geometry_ranks > 1 && opt->elem_num_x % opt->geometry_ranks))
        return "ELEMNUMX must be divisible by the number of ranks of the reference grid";
    if (!(cs->xR > cs->xL) || !(cs->yR > cs->yL)) return "empty domain";
    return nullptr;
}

Phys make_phys(const wfv_case &cs) {
    Phys P{};
    P.rho10 = cs.rho10;
    P.rho20 = cs.rho20;
    P.c1_2 = cs.sound1 * cs.sound1;
    P.c2_2 = cs.sound2 * cs.sound2;
    P.preb = cs.preb;
    P.q = cs.rho20 * P.c2_2 - cs.rho10 * P.c1_2;
    P.ic1_2 = 1.0 / P.c1_2;
    P.ic2_2 = 1.0 / P.c2_2;
    return P;
}

StageArgs stage_base(const wfv_options &opt, const wfv_case &cs, int nx, int ny, int ty, const double *dxC,
                     const double *dxW, const double *dyC, const double *ghost, StepCtl *ctl) {
    StageArgs A{};
    A.dxC = dxC;
    A.dxW = dxW;
    A.dyC = dyC;
    A.P = make_phys(cs);
    A.dX = (cs.xR - cs.xL) / nx;
    A.dY = (cs.yR - cs.yL) / ny;
    A.cfl = opt.cfl;
    A.gravity = cs.gravity;
    A.nx = nx;
    A.ny = ny;
    A.bc = opt.boundary;
    A.ghost = ghost;
    A.theta_flag = nullptr;
    A.flux = opt.riemann_flux;
    A.bp = opt.pp_limiter;
    A.ty = ty;
    A.sp = strip_plan(nx, ny, ty, strip_halo(opt.scheme, opt.reconstruction));
    A.dx4 = std::pow(A.dX, 4);  // WENO-EXP's threshold, as the reference's pow(m_deltaX, 4)
    A.flags = &ctl->flags;
    return A;
}

hipError_t enqueue_rk_step(const StageArgs &base, double *const reg[3], const wfv_options &opt, double dt,
                           const double *dt_dev, const int *active, unsigned long long *dtmin, hipStream_t s,
                           StepCtl *fuse_ctl) {
    StagePlan pl[3];
    const int ns = rk_plan(opt.rk_method, pl);
    for (int st = 0; st < ns; ++st) {
        StageArgs A = base;
        A.U = reg[pl[st].in];
        A.Un = pl[st].base < 0 ? nullptr : reg[pl[st].base];
        A.Uo = reg[pl[st].out];
        A.a = pl[st].a;
        A.b = pl[st].b;
        A.c = pl[st].c;
        A.dt = dt;
        A.dt_ptr = dt_dev;
        A.active = active;
        A.dtmin = (st == ns - 1 && opt.rk_method != 0) ? dtmin : nullptr;
        A.ctl = (st == ns - 1 && opt.rk_method != 0 && dtmin) ? fuse_ctl : nullptr;
        hipError_t e = launch_stage(A, opt.precision, opt.reconstruction, opt.scheme, s);
        if (e != hipSuccess) return e;
    }
    if (opt.rk_method == 0) {
        // RK1 writes register 1; copy back so that register 0 stays the state (skipped, like
        // the stage, once the device loop has finished)
        hipError_t e = launch_copy_state(reg[0], reg[1], 4 * (size_t)base.nx * base.ny, active, s);
        if (e != hipSuccess) return e;
        if (dtmin) {
            DtArgs D{reg[0], base.P, base.dX, base.dY, base.cfl, base.nx, base.ny, base.flags, dtmin};
            e = launch_dt(D, s);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

int fail(wfv_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int bfail(wfv_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    return code;
}

StageArgs base_args(wfv_ctx *c) {
    StageArgs A = stage_base(c->opt, c->cs, c->nx, c->ny, c->ty, c->d_dxC, c->d_dxW, c->d_dyC, c->d_ghost, c->d_ctl);
    A.theta_flag = c->d_theta_flag;
    return A;
}

int rk_plan(int rk, StagePlan *s) {
    if (rk == 0) {
        // U += dt*rhs, written as (0*U + 1*U) + (1*dt)*rhs, which rounds identically
        s[0] = {0.0, 1.0, 1.0, 0, -1, 1};
        return 1;
    }
    if (rk == 1) {
        s[0] = {1.0, 0.0, 1.0, 0, -1, 1};
        s[1] = {0.5, 0.5, 0.5, 1, 0, 0};
        return 2;
    }
    s[0] = {1.0, 0.0, 1.0, 0, -1, 1};
    s[1] = {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0, 1, 0, 2};
    s[2] = {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 2, 0, 0};
    return 3;
}

}  // namespace wfv

namespace {

// enqueue one RK step; dt from the host (dt_dev == nullptr) or from the device loop; fuse: the
// step's last stage also runs the next step's control (fused_control)
int enqueue_step(wfv_ctx *c, double dt, const double *dt_dev, const int *active, bool want_dtmin, bool fuse = false) {
    HIPCHK(c, enqueue_rk_step(base_args(c), c->reg, c->opt, dt, dt_dev, active,
                              want_dtmin ? &c->d_ctl->dtmin_bits : nullptr, c->stream, fuse ? c->d_ctl : nullptr));
    return WFV_OK;
}

// whether a device loop's steps run their control in the last stage (every scheme whose last
// stage reduces the CFL minimum: SSP-RK2 and RK3) rather than as a launch before each step
bool fused_control(const wfv_ctx *c) { return c->opt.rk_method != 0; }

}  // namespace

extern "C" {

const char *wfv_version(void) {
    return "wenofv-gfx950 1.2 (fp64, mixed, mixed-w64; js, z, exp; cons, char, prim, prim2, prim3)";
}

int wfv_create(const wfv_options *opt, const wfv_case *cs, wfv_ctx **out) {
    if (!opt || !cs || !out) return WFV_ERR_ARG;
    *out = nullptr;
    wfv_ctx *c = new wfv_ctx();
    c->opt = *opt;
    c->cs = *cs;
    if (const char *why = check_options(opt, cs)) {
        delete c;
        set_thread_error(why);
        return WFV_ERR_CONFIG;
    }
    c->nx = opt->elem_num_x;
    c->ny = opt->elem_num_y;
    c->nc = (size_t)c->nx * c->ny;
    c->ty = opt->rows_per_strip > 0 ? std::min(opt->rows_per_strip, c->ny) : default_rows_per_strip(c->nx, c->ny);
    c->P = make_phys(*cs);
    c->dX = (cs->xR - cs->xL) / c->nx;
    c->dY = (cs->yR - cs->yL) / c->ny;
    c->dxC.resize(c->nx);
    c->dxW.resize(c->nx);
    c->dyC.resize(c->ny);
    grid_widths(c->nx, c->ny, cs->xL, cs->xR, cs->yL, cs->yR, opt->geometry_ranks < 1 ? 1 : opt->geometry_ranks,
                c->dxC.data(), c->dxW.data(), c->dyC.data(), nullptr, nullptr);
    auto bad = [&](hipError_t e) {
        if (e == hipSuccess) return false;
        set_thread_error(hipGetErrorString(e));
        wfv_destroy(c);
        return true;
    };
    if (bad(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking))) return WFV_ERR_HIP;
    c->stream = c->own_stream;
    for (int k = 0; k < 3; ++k)
        if (bad(hipMalloc(&c->reg[k], 4 * c->nc * sizeof(double)))) return WFV_ERR_HIP;
    if (bad(hipMalloc(&c->d_dxC, c->nx * sizeof(double))) || bad(hipMalloc(&c->d_dxW, c->nx * sizeof(double))) ||
        bad(hipMalloc(&c->d_dyC, c->ny * sizeof(double))) || bad(hipMalloc(&c->d_ctl, sizeof(StepCtl))))
        return WFV_ERR_HIP;
    if (bad(hipMemcpy(c->d_dxC, c->dxC.data(), c->nx * sizeof(double), hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(c->d_dxW, c->dxW.data(), c->nx * sizeof(double), hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(c->d_dyC, c->dyC.data(), c->ny * sizeof(double), hipMemcpyHostToDevice)))
        return WFV_ERR_HIP;
    StepCtl init{};
    init.dtmin_bits = ~0ull;
    init.active = 1;
    if (bad(hipMemcpy(c->d_ctl, &init, sizeof init, hipMemcpyHostToDevice))) return WFV_ERR_HIP;
    if (bad(hipMemset(c->reg[0], 0, 4 * c->nc * sizeof(double)))) return WFV_ERR_HIP;
    if (opt->boundary == 3) {
        const size_t bytes = 4 * ring_size(c->nx, c->ny) * sizeof(double);
        if (bad(hipMalloc(&c->d_ghost, bytes)) || bad(hipMemset(c->d_ghost, 0, bytes))) return WFV_ERR_HIP;
    }
    *out = c;
    return WFV_OK;
}

void wfv_destroy(wfv_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (auto &p : c->reg) (void)hipFree(p);
    (void)hipFree(c->d_dxC);
    (void)hipFree(c->d_dxW);
    (void)hipFree(c->d_dyC);
    (void)hipFree(c->d_faces);
    (void)hipFree(c->d_theta);
    (void)hipFree(c->d_faces_ring);
    (void)hipFree(c->d_theta_ring);
    (void)hipFree(c->d_cell_dt);
    (void)hipFree(c->d_ghost);
    (void)hipFree(c->d_theta_flag);
    (void)hipFree(c->d_ctl);
    (void)hipFree(c->d_dts);
    if (c->own_stream) {
        (void)hipStreamSynchronize(c->own_stream);
        (void)hipStreamDestroy(c->own_stream);
    }
    delete c;
}

const char *wfv_last_error(const wfv_ctx *c) { return c ? c->err.c_str() : g_create_error; }

int wfv_set_state(wfv_ctx *c, const double *U) {
    if (!c || !U) return WFV_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(c->reg[0], U, 4 * c->nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WFV_OK;
}
int wfv_get_state(wfv_ctx *c, double *U) {
    if (!c || !U) return WFV_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(U, c->reg[0], 4 * c->nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WFV_OK;
}
int wfv_set_state_device(wfv_ctx *c, const double *dU) {
    if (!c || !dU) return WFV_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(c->reg[0], dU, 4 * c->nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WFV_OK;
}
int wfv_get_state_device(wfv_ctx *c, double *dU) {
    if (!c || !dU) return WFV_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(dU, c->reg[0], 4 * c->nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WFV_OK;
}

int wfv_compute_dt(wfv_ctx *c, double *dt) {
    if (!c || !dt) return WFV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    StepCtl h{};
    HIPCHK(c, hipMemcpy(&h, c->d_ctl, sizeof h, hipMemcpyDeviceToHost));
    h.flags &= ~WFV_FLAG_NAN_DT;  // this call's own NaN time step only
    h.dtmin_bits = ~0ull;
    HIPCHK(c, hipMemcpy(c->d_ctl, &h, sizeof h, hipMemcpyHostToDevice));
    DtArgs D{c->reg[0], c->P, c->dX, c->dY, c->opt.cfl, c->nx, c->ny, &c->d_ctl->flags, &c->d_ctl->dtmin_bits};
    HIPCHK(c, launch_dt(D, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&h, c->d_ctl, sizeof h, hipMemcpyDeviceToHost));
    if (h.flags & WFV_FLAG_NAN_DT) return fail(c, WFV_ERR_NONFINITE_DT, "Error: exist nan DeltaT");
    double m;
    std::memcpy(&m, &h.dtmin_bits, sizeof m);
    double timestep = 1.0;  // src/WENOFV.cpp:477, 510-513
    if (h.dtmin_bits != ~0ull && !(m > timestep)) timestep = m;
    *dt = timestep;
    return WFV_OK;
}

int wfv_step(wfv_ctx *c, double dt) {
    if (!c) return WFV_ERR_ARG;
    int rc = enqueue_step(c, dt, nullptr, nullptr, false);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WFV_OK;
}

int wfv_assemble_rhs(wfv_ctx *c, double *rhs, double *faces, double *theta) {
    if (!c || !rhs) return WFV_ERR_ARG;
    if ((faces || theta) && !c->d_faces) {
        HIPCHK(c, hipMalloc(&c->d_faces, c->nc * 64 * sizeof(double)));
        HIPCHK(c, hipMalloc(&c->d_theta, c->nc * sizeof(double)));
    }
    StageArgs A = base_args(c);
    A.U = c->reg[0];
    A.Un = nullptr;
    A.Uo = c->reg[2];
    A.a = 0.0;  // Uo = (0*U + 0*U) + (1*1)*rhs = rhs
    A.b = 0.0;
    A.c = 1.0;
    A.dt = 1.0;
    A.theta_flag = nullptr;
    A.faces_out = (faces || theta) ? c->d_faces : nullptr;
    A.theta_out = (faces || theta) ? c->d_theta : nullptr;
    HIPCHK(c, launch_stage(A, c->opt.precision, c->opt.reconstruction, c->opt.scheme, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(rhs, c->reg[2], 4 * c->nc * sizeof(double), hipMemcpyDeviceToHost));
    if (faces) HIPCHK(c, hipMemcpy(faces, c->d_faces, c->nc * 64 * sizeof(double), hipMemcpyDeviceToHost));
    if (theta) HIPCHK(c, hipMemcpy(theta, c->d_theta, c->nc * sizeof(double), hipMemcpyDeviceToHost));
    return WFV_OK;
}

int wfv_face_values(wfv_ctx *c, int reconstruction, int limit, double *faces, double *theta) {
    if (!c || !faces) return WFV_ERR_ARG;
    if (reconstruction < 0 || reconstruction > 4 || limit < 0 || limit > 1)
        return fail(c, WFV_ERR_ARG, "wfv_face_values: reconstruction 0..4, limit 0 or 1");
    const size_t np = (size_t)(c->nx + 2) * (c->ny + 2);
    if (!c->d_faces_ring) {
        HIPCHK(c, hipMalloc(&c->d_faces_ring, np * 64 * sizeof(double)));
        HIPCHK(c, hipMalloc(&c->d_theta_ring, np * sizeof(double)));
    }
    wfv_options o = c->opt;
    o.reconstruction = reconstruction;
    o.pp_limiter = limit;
    StageArgs A = stage_base(o, c->cs, c->nx, c->ny, c->ty, c->d_dxC, c->d_dxW, c->d_dyC, c->d_ghost, c->d_ctl);
    A.U = c->reg[0];
    A.Un = nullptr;
    A.Uo = c->reg[2];  // the right-hand side, unused
    A.a = 0.0;
    A.b = 0.0;
    A.c = 1.0;
    A.dt = 1.0;
    A.faces_out = c->d_faces_ring;
    A.theta_out = c->d_theta_ring;
    A.faces_ring = 1;
    HIPCHK(c, launch_stage(A, c->opt.precision, reconstruction, c->opt.scheme, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(faces, c->d_faces_ring, np * 64 * sizeof(double), hipMemcpyDeviceToHost));
    if (theta) HIPCHK(c, hipMemcpy(theta, c->d_theta_ring, np * sizeof(double), hipMemcpyDeviceToHost));
    return WFV_OK;
}

int wfv_cell_dt(wfv_ctx *c, double *dt) {
    if (!c || !dt) return WFV_ERR_ARG;
    if (!c->d_cell_dt) HIPCHK(c, hipMalloc(&c->d_cell_dt, c->nc * sizeof(double)));
    DtArgs D{c->reg[0], c->P, c->dX, c->dY, c->opt.cfl, c->nx, c->ny, nullptr, nullptr, c->d_cell_dt};
    HIPCHK(c, launch_dt(D, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(dt, c->d_cell_dt, c->nc * sizeof(double), hipMemcpyDeviceToHost));
    return WFV_OK;
}

}  // extern "C"

namespace {

// the device-resident time loop: fixed unclipped steps (fixed_steps > 0), or toward t_end with
// the last step clipped, stopping early after max_count steps when max_count > 0
int advance_impl(wfv_ctx *c, double t_start, double t_end, int fixed_steps, int max_count, int *nsteps, double *t_out,
                 double *dts, int max_dts) {
    if (!c) return WFV_ERR_ARG;
    if (dts && max_dts > c->dts_cap) {
        (void)hipFree(c->d_dts);
        c->d_dts = nullptr;
        HIPCHK(c, hipMalloc(&c->d_dts, (size_t)max_dts * sizeof(double)));
        c->dts_cap = max_dts;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    StepCtl h{};
    HIPCHK(c, hipMemcpy(&h, c->d_ctl, sizeof h, hipMemcpyDeviceToHost));
    // a NaN time step ends the call that met it (and was reported by it); the other flags
    // stay until wfv_get_flags reads them
    h.flags &= ~WFV_FLAG_NAN_DT;
    h.dtmin_bits = ~0ull;
    h.now = t_start;
    h.t_end = t_end;
    h.count = 0;
    h.fixed_steps = fixed_steps;
    h.active = 1;
    h.dts = dts ? c->d_dts : nullptr;
    h.max_record = dts ? max_dts : 0;
    h.max_count = fixed_steps > 0 ? 0 : max_count;
    h.done = 0;
    HIPCHK(c, hipMemcpy(c->d_ctl, &h, sizeof h, hipMemcpyHostToDevice));
    // CFL minimum of the initial state
    DtArgs D{c->reg[0], c->P, c->dX, c->dY, c->opt.cfl, c->nx, c->ny, &c->d_ctl->flags, &c->d_ctl->dtmin_bits};
    HIPCHK(c, launch_dt(D, c->stream));
    // The control of the first step; every later one runs in the last wavefront of the
    // previous step's last stage (fused_control), or as a launch before its step.
    const bool fuse = fused_control(c);
    if (fuse) HIPCHK(c, launch_step_control(c->d_ctl, c->stream));
    // Steps are replayed from a graph of GRAPH_STEPS steps captured once per context; a
    // step enqueued after the loop has finished returns at once (device flag `active`).
    // A fixed-step run enqueues its remainder as individual launches, so it never runs past
    // its last step.
    constexpr int GRAPH_STEPS = 32;
    if (!c->graph) {
        hipGraph_t g;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = WFV_OK;
        for (int k = 0; k < GRAPH_STEPS && rc == WFV_OK; ++k) {
            if (!fuse && launch_step_control(c->d_ctl, c->stream) != hipSuccess) rc = WFV_ERR_HIP;
            if (rc == WFV_OK) rc = enqueue_step(c, 0.0, &c->d_ctl->dt, &c->d_ctl->active, true, fuse);
        }
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != WFV_OK) return rc;
        if (e != hipSuccess) return fail(c, WFV_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(c, WFV_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(e));
        c->graph_steps = GRAPH_STEPS;
    }
    // a short capped run is enqueued step by step rather than as graph replays of no-op steps
    const int capped = (fixed_steps <= 0 && max_count > 0 && max_count < GRAPH_STEPS) ? max_count : 0;
    if (fixed_steps > 0 || capped) {
        const int nfix = fixed_steps > 0 ? fixed_steps : capped;
        const int full = nfix / GRAPH_STEPS, rest = nfix % GRAPH_STEPS;
        for (int k = 0; k < full; ++k) HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
        for (int k = 0; k < rest; ++k) {
            if (!fuse) HIPCHK(c, launch_step_control(c->d_ctl, c->stream));
            int rc = enqueue_step(c, 0.0, &c->d_ctl->dt, &c->d_ctl->active, true, fuse);
            if (rc) return rc;
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(&h, c->d_ctl, sizeof h, hipMemcpyDeviceToHost));
    } else {
        for (;;) {
            // a few graphs in flight between host checks of the loop flag
            for (int k = 0; k < 4; ++k) HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            HIPCHK(c, hipMemcpy(&h, c->d_ctl, sizeof h, hipMemcpyDeviceToHost));
            if (!h.active) break;
        }
    }
    if (nsteps) *nsteps = h.count;
    if (t_out) *t_out = h.now;
    if (dts && h.count > 0) {
        const int n = h.count < max_dts ? h.count : max_dts;
        HIPCHK(c, hipMemcpy(dts, c->d_dts, n * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (h.flags & WFV_FLAG_NAN_DT) return fail(c, WFV_ERR_NONFINITE_DT, "Error: exist nan DeltaT");
    return WFV_OK;
}

}  // namespace

extern "C" {

int wfv_advance(wfv_ctx *c, double t_start, double t_end, int fixed_steps, int *nsteps, double *t_out,
                double *dts, int max_dts) {
    return advance_impl(c, t_start, t_end, fixed_steps, 0, nsteps, t_out, dts, max_dts);
}

int wfv_advance_steps(wfv_ctx *c, double t_start, double t_end, int max_steps, int *nsteps, double *t_out) {
    if (max_steps < 1) return WFV_ERR_ARG;
    return advance_impl(c, t_start, t_end, 0, max_steps, nsteps, t_out, nullptr, 0);
}

int wfv_set_ghost_cells(wfv_ctx *c, const double *Upad) {
    if (!c || !Upad) return WFV_ERR_ARG;
    if (!c->d_ghost) return fail(c, WFV_ERR_CONFIG, "ghost cells need boundary 3");
    std::vector<double> ring(4 * ring_size(c->nx, c->ny));
    pack_ghost_ring(c->nx, c->ny, Upad, ring.data());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_ghost, ring.data(), ring.size() * sizeof(double), hipMemcpyHostToDevice));
    return WFV_OK;
}

int wfv_record_limiter_flags(wfv_ctx *c, int on) {
    if (!c) return WFV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (on && !c->d_theta_flag) {
        HIPCHK(c, hipMalloc(&c->d_theta_flag, c->nc * sizeof(double)));
        HIPCHK(c, hipMemset(c->d_theta_flag, 0, c->nc * sizeof(double)));
    } else if (!on && c->d_theta_flag) {
        (void)hipFree(c->d_theta_flag);
        c->d_theta_flag = nullptr;
    }
    if (c->graph) {  // the captured steps carry the flag array
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    return WFV_OK;
}

int wfv_get_limiter_flags(wfv_ctx *c, double *flags) {
    if (!c || !flags) return WFV_ERR_ARG;
    if (!c->d_theta_flag) return fail(c, WFV_ERR_CONFIG, "limiter flags are not recorded");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(flags, c->d_theta_flag, c->nc * sizeof(double), hipMemcpyDeviceToHost));
    return WFV_OK;
}

int wfv_get_flags(wfv_ctx *c, unsigned *flags) {
    if (!c || !flags) return WFV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    StepCtl h{};
    HIPCHK(c, hipMemcpy(&h, c->d_ctl, sizeof h, hipMemcpyDeviceToHost));
    *flags = h.flags;
    h.flags = 0;
    HIPCHK(c, hipMemcpy(c->d_ctl, &h, sizeof h, hipMemcpyHostToDevice));
    return WFV_OK;
}

int wfv_set_stream(wfv_ctx *c, void *hip_stream) {
    if (!c) return WFV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->stream = hip_stream ? (hipStream_t)hip_stream : c->own_stream;
    return WFV_OK;
}

int wfv_synchronize(wfv_ctx *c) {
    if (!c) return WFV_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return WFV_OK;
}

int wfv_grid(const wfv_ctx *c, int *nx, int *ny, double *dxC, double *dxW, double *dyC) {
    if (!c) return WFV_ERR_ARG;
    if (nx) *nx = c->nx;
    if (ny) *ny = c->ny;
    if (dxC) std::memcpy(dxC, c->dxC.data(), c->nx * sizeof(double));
    if (dxW) std::memcpy(dxW, c->dxW.data(), c->nx * sizeof(double));
    if (dyC) std::memcpy(dyC, c->dyC.data(), c->ny * sizeof(double));
    return WFV_OK;
}

}  // extern "C"

// ======================================================================= batches

namespace wfv {

// (re)build the instance table and the item -> instance map from b->ty.  Items are issued
// longest first (instances with more rows per strip first), so the short items fill the
// tail of the launch.
int batch_layout(wfv_batch *b) {
    std::vector<int> order(b->n);
    for (int i = 0; i < b->n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return b->ty[p] > b->ty[q]; });
    std::vector<int> nstrip(b->n), rows(b->n);
    for (int i = 0; i < b->n; ++i) {
        BatchInst &I = b->hinst[i];
        I.ty = b->ty[i];
        I.sp = strip_plan(I.nx, I.ny, I.ty, strip_halo(b->opt.scheme, b->opt.reconstruction));
        nstrip[i] = I.sp.total;
        rows[i] = I.ty;
    }
    // runs of 16 strips: a few row bands of one instance per XCD, while the XCDs' loads and
    // each one's longest-first start stay as balanced as the one global order's
    std::vector<int> item_inst, item_strip;
    b->total_items = batch_item_order(order, nstrip, rows, WFV_BATCH_RUN, item_inst, item_strip);
    item_inst.insert(item_inst.end(), item_strip.begin(), item_strip.end());
    (void)hipFree(b->d_item_inst);
    b->d_item_inst = nullptr;
    BHIPCHK(b, hipMalloc(&b->d_item_inst, item_inst.size() * sizeof(int)));
    BHIPCHK(b, hipMemcpy(b->d_item_inst, item_inst.data(), item_inst.size() * sizeof(int), hipMemcpyHostToDevice));
    BHIPCHK(b, hipMemcpy(b->d_inst, b->hinst.data(), b->n * sizeof(BatchInst), hipMemcpyHostToDevice));
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // schedule 0: one wavefront per item; 1: persistent, 8 wavefronts per CU striding over
    // the items
    b->grid = b->persistent ? std::min(b->total_items, cus * 8) : b->total_items;
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return WFV_OK;
}

}  // namespace wfv

namespace {

int batch_enqueue_step(wfv_batch *b) {
    StagePlan pl[3];
    const int ns = rk_plan(b->opt.rk_method, pl);
    BHIPCHK(b, launch_batch_step_control(b->d_ctl, b->n, b->stream));
    for (int s = 0; s < ns; ++s) {
        BatchStageArgs B{};
        B.inst = b->d_inst;
        B.item_inst = b->d_item_inst;
        B.item_strip = b->d_item_inst + b->total_items;
        B.total_items = b->total_items;
        B.in = pl[s].in;
        B.base = pl[s].base;
        B.out = pl[s].out;
        B.a = pl[s].a;
        B.b = pl[s].b;
        B.c = pl[s].c;
        B.bp = b->opt.pp_limiter;
        B.ghost = b->ghost != nullptr;
        B.want_dtmin = (s == ns - 1 && b->opt.rk_method != 0) ? 1 : 0;
        BHIPCHK(b, launch_batch_stage(B, b->opt.precision, b->opt.reconstruction, b->opt.scheme,
                                      b->opt.riemann_flux, b->grid, b->stream));
    }
    if (b->opt.rk_method == 0) {
        // register 1 back into the state of every instance whose loop is still running
        BHIPCHK(b, launch_batch_copy_state(b->d_inst, b->n, b->max_cells, b->stream));
        BHIPCHK(b, launch_batch_dt(b->d_inst, b->n, b->max_cells, b->stream));
    }
    return WFV_OK;
}

}  // namespace

extern "C" {

int wfv_batch_create(const wfv_options *opt, const wfv_instance *inst, int n, wfv_batch **out) {
    if (!opt || !inst || n < 1 || !out) return WFV_ERR_ARG;
    *out = nullptr;
    std::string why;
    if (opt->rk_method < 0 || opt->rk_method > 2) why = "RKMETH must be 0, 1 or 2";
    else if (opt->scheme < 0 || opt->scheme > 2) why = "SCHEME must be 0, 1 or 2";
    else if (opt->riemann_flux < 0 || opt->riemann_flux > 1) why = "RIEMANNFLUX must be 0 or 1";
    else if (opt->reconstruction < 0 || opt->reconstruction > 4) why = "RECONSTRUCTIONTYPE must be 0..4";
    else if (opt->precision < 0 || opt->precision > 3) why = "precision must be 0, 1, 2 or 3";
    else if (opt->precision == 3 && (opt->scheme == 2 || opt->reconstruction == 4))
        why = "precision 3 has the 5-point schemes only (SCHEME 0-1, RECONSTRUCTIONTYPE 0-3)";
    else if (!(opt->cfl > 0)) why = "CFL must be positive";
    for (int i = 0; i < n && why.empty(); ++i) {
        const wfv_instance &I = inst[i];
        if (I.nx < 3 || I.ny < 3) why = "instance grids need at least 3 cells per direction";
        else if (I.boundary < 0 || I.boundary > 3) why = "instance boundary must be 0..3";
        else if (!(I.cs.xR > I.cs.xL) || !(I.cs.yR > I.cs.yL)) why = "empty instance domain";
        else if (opt->geometry_ranks > 1 && I.nx % opt->geometry_ranks) why = "nx not divisible by geometry_ranks";
    }
    if (!why.empty()) {
        set_thread_error(why.c_str());
        return WFV_ERR_CONFIG;
    }
    wfv_batch *b = new wfv_batch();
    b->opt = *opt;
    b->n = n;
    b->inst.assign(inst, inst + n);
    b->hinst.resize(n);
    b->cell_off.resize(n);
    b->ty.resize(n);
    size_t geo_total = 0, ghost_total = 0;
    std::vector<size_t> ghost_off(n, 0);
    for (int i = 0; i < n; ++i)
        if (inst[i].boundary == 3) {
            ghost_off[i] = ghost_total;
            ghost_total += 4 * ring_size(inst[i].nx, inst[i].ny);
        }
    std::vector<int> nxs(n), nys(n);
    for (int i = 0; i < n; ++i) {
        nxs[i] = inst[i].nx;
        nys[i] = inst[i].ny;
    }
    const int ty0 = opt->rows_per_strip > 0 ? opt->rows_per_strip : default_batch_rows_per_strip(nxs.data(), nys.data(), n);
    for (int i = 0; i < n; ++i) {
        b->cell_off[i] = b->total_cells;
        b->total_cells += (size_t)inst[i].nx * inst[i].ny;
        geo_total += 2 * (size_t)inst[i].nx + inst[i].ny;
        b->max_cells = std::max(b->max_cells, inst[i].nx * inst[i].ny);
        b->ty[i] = std::min(ty0, inst[i].ny);  // strips longer than the grid are the whole grid
    }
    b->default_ty = ty0;
    auto bad = [&](hipError_t e) {
        if (e == hipSuccess) return false;
        set_thread_error(hipGetErrorString(e));
        wfv_batch_destroy(b);
        return true;
    };
    if (bad(hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking))) return WFV_ERR_HIP;
    b->stream = b->own_stream;
    for (int k = 0; k < 3; ++k)
        if (bad(hipMalloc(&b->reg[k], 4 * b->total_cells * sizeof(double)))) return WFV_ERR_HIP;
    if (bad(hipMalloc(&b->geo, geo_total * sizeof(double))) || bad(hipMalloc(&b->d_ctl, n * sizeof(StepCtl))) ||
        bad(hipMalloc(&b->d_inst, n * sizeof(BatchInst))))
        return WFV_ERR_HIP;
    if (bad(hipMemset(b->reg[0], 0, 4 * b->total_cells * sizeof(double)))) return WFV_ERR_HIP;
    if (ghost_total &&
        (bad(hipMalloc(&b->ghost, ghost_total * sizeof(double))) || bad(hipMemset(b->ghost, 0, ghost_total * sizeof(double)))))
        return WFV_ERR_HIP;
    std::vector<double> geo(geo_total);
    size_t go = 0;
    for (int i = 0; i < n; ++i) {
        const wfv_instance &in = inst[i];
        BatchInst &I = b->hinst[i];
        std::memset(&I, 0, sizeof I);
        for (int k = 0; k < 3; ++k) I.reg[k] = b->reg[k] + 4 * b->cell_off[i];
        double *dxC = geo.data() + go, *dxW = dxC + in.nx, *dyC = dxW + in.nx;
        grid_widths(in.nx, in.ny, in.cs.xL, in.cs.xR, in.cs.yL, in.cs.yR,
                    opt->geometry_ranks < 1 ? 1 : opt->geometry_ranks, dxC, dxW, dyC, nullptr, nullptr);
        I.dxC = b->geo + go;
        I.dxW = I.dxC + in.nx;
        I.dyC = I.dxW + in.nx;
        go += 2 * (size_t)in.nx + in.ny;
        I.ctl = b->d_ctl + i;
        I.P = make_phys(in.cs);
        I.dX = (in.cs.xR - in.cs.xL) / in.nx;
        I.dY = (in.cs.yR - in.cs.yL) / in.ny;
        I.dx4 = std::pow(I.dX, 4);
        I.cfl = opt->cfl;
        I.gravity = in.cs.gravity;
        I.nx = in.nx;
        I.ny = in.ny;
        I.bc = in.boundary;
        I.ghost = in.boundary == 3 ? b->ghost + ghost_off[i] : nullptr;
    }
    if (bad(hipMemcpy(b->geo, geo.data(), geo_total * sizeof(double), hipMemcpyHostToDevice))) return WFV_ERR_HIP;
    std::vector<StepCtl> ctl(n);
    for (auto &c : ctl) {
        std::memset(&c, 0, sizeof c);
        c.dtmin_bits = ~0ull;
    }
    if (bad(hipMemcpy(b->d_ctl, ctl.data(), n * sizeof(StepCtl), hipMemcpyHostToDevice))) return WFV_ERR_HIP;
    if (batch_layout(b) != WFV_OK) {
        set_thread_error(b->err.c_str());
        wfv_batch_destroy(b);
        return WFV_ERR_HIP;
    }
    *out = b;
    return WFV_OK;
}

void wfv_batch_destroy(wfv_batch *b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (auto &p : b->reg) (void)hipFree(p);
    (void)hipFree(b->geo);
    (void)hipFree(b->ghost);
    (void)hipFree(b->d_ctl);
    (void)hipFree(b->d_inst);
    (void)hipFree(b->d_item_inst);
    if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
    delete b;
}

const char *wfv_batch_last_error(const wfv_batch *b) { return b ? b->err.c_str() : g_create_error; }

int wfv_batch_set_state(wfv_batch *b, int i, const double *U) {
    if (!b || !U || i < 0 || i >= b->n) return WFV_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    BHIPCHK(b, hipMemcpy(b->hinst[i].reg[0], U, 4 * (size_t)b->inst[i].nx * b->inst[i].ny * sizeof(double),
                         hipMemcpyHostToDevice));
    return WFV_OK;
}

int wfv_batch_get_state(wfv_batch *b, int i, double *U) {
    if (!b || !U || i < 0 || i >= b->n) return WFV_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    BHIPCHK(b, hipMemcpy(U, b->hinst[i].reg[0], 4 * (size_t)b->inst[i].nx * b->inst[i].ny * sizeof(double),
                         hipMemcpyDeviceToHost));
    return WFV_OK;
}

int wfv_batch_set_ghost_cells(wfv_batch *b, int i, const double *Upad) {
    if (!b || !Upad || i < 0 || i >= b->n) return WFV_ERR_ARG;
    if (!b->hinst[i].ghost) return bfail(b, WFV_ERR_CONFIG, "ghost cells need boundary 3");
    const int nx = b->inst[i].nx, ny = b->inst[i].ny;
    std::vector<double> ring(4 * ring_size(nx, ny));
    pack_ghost_ring(nx, ny, Upad, ring.data());
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    BHIPCHK(b, hipMemcpy((void *)b->hinst[i].ghost, ring.data(), ring.size() * sizeof(double), hipMemcpyHostToDevice));
    return WFV_OK;
}

int wfv_batch_set_stream(wfv_batch *b, void *hip_stream) {
    if (!b) return WFV_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    b->stream = hip_stream ? (hipStream_t)hip_stream : b->own_stream;
    return WFV_OK;
}

int wfv_batch_set_rows_per_strip(wfv_batch *b, int i, int rows) {
    if (!b || i < 0 || i >= b->n || rows < 0) return WFV_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    // 0: the batch default; strips longer than the grid are the whole grid
    b->ty[i] = std::min(rows > 0 ? rows : b->default_ty, b->inst[i].ny);
    return batch_layout(b);
}

int wfv_batch_cells(const wfv_batch *b, long long *cells) {
    if (!b || !cells) return WFV_ERR_ARG;
    *cells = (long long)b->total_cells;
    return WFV_OK;
}

int wfv_batch_advance(wfv_batch *b, double t_end, int fixed_steps, int *nsteps, double *t_out, unsigned *flags) {
    if (!b) return WFV_ERR_ARG;
    BHIPCHK(b, hipStreamSynchronize(b->stream));
    std::vector<StepCtl> ctl(b->n);
    BHIPCHK(b, hipMemcpy(ctl.data(), b->d_ctl, b->n * sizeof(StepCtl), hipMemcpyDeviceToHost));
    for (auto &c : ctl) {
        c.flags &= ~WFV_FLAG_NAN_DT;  // as in wfv_advance
        c.dtmin_bits = ~0ull;
        c.now = 0.0;
        c.t_end = t_end;
        c.count = 0;
        c.fixed_steps = fixed_steps;
        c.active = 1;
        c.dts = nullptr;
        c.max_record = 0;
    }
    BHIPCHK(b, hipMemcpy(b->d_ctl, ctl.data(), b->n * sizeof(StepCtl), hipMemcpyHostToDevice));
    BHIPCHK(b, launch_batch_dt(b->d_inst, b->n, b->max_cells, b->stream));
    constexpr int GRAPH_STEPS = 16;
    if (!b->graph) {
        hipGraph_t g;
        BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = WFV_OK;
        for (int k = 0; k < GRAPH_STEPS && rc == WFV_OK; ++k) rc = batch_enqueue_step(b);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != WFV_OK) return rc;
        if (e != hipSuccess) return bfail(b, WFV_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return bfail(b, WFV_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(e));
    }
    if (fixed_steps > 0) {
        const int full = fixed_steps / GRAPH_STEPS, rest = fixed_steps % GRAPH_STEPS;
        for (int k = 0; k < full; ++k) BHIPCHK(b, hipGraphLaunch(b->graph, b->stream));
        for (int k = 0; k < rest; ++k) {
            int rc = batch_enqueue_step(b);
            if (rc) return rc;
        }
        BHIPCHK(b, hipStreamSynchronize(b->stream));
        BHIPCHK(b, hipMemcpy(ctl.data(), b->d_ctl, b->n * sizeof(StepCtl), hipMemcpyDeviceToHost));
    } else {
        for (;;) {
            for (int k = 0; k < 4; ++k) BHIPCHK(b, hipGraphLaunch(b->graph, b->stream));
            BHIPCHK(b, hipStreamSynchronize(b->stream));
            BHIPCHK(b, hipMemcpy(ctl.data(), b->d_ctl, b->n * sizeof(StepCtl), hipMemcpyDeviceToHost));
            bool any = false;
            for (auto &c : ctl) any |= c.active != 0;
            if (!any) break;
        }
    }
    bool nan_dt = false;
    for (int i = 0; i < b->n; ++i) {
        if (nsteps) nsteps[i] = ctl[i].count;
        if (t_out) t_out[i] = ctl[i].now;
        if (flags) flags[i] = ctl[i].flags;
        nan_dt |= (ctl[i].flags & WFV_FLAG_NAN_DT) != 0;
        ctl[i].flags = 0;
    }
    BHIPCHK(b, hipMemcpy(b->d_ctl, ctl.data(), b->n * sizeof(StepCtl), hipMemcpyHostToDevice));
    if (nan_dt) return bfail(b, WFV_ERR_NONFINITE_DT, "Error: exist nan DeltaT");
    return WFV_OK;
}

}  // extern "C"
