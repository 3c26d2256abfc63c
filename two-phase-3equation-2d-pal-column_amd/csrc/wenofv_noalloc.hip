// wenofv_noalloc.hip -- wfv_advance_fixed_noalloc: the fixed-step time loop of wfv_advance on
// a caller-owned device state, allocating no memory.
//
// The RK registers 1 and 2, the grid widths and the loop control live in `scratch`, a static
// device array of WFV_NOALLOC_SCRATCH_BYTES.  Every GPU the library runs on holds its own copy
// of it, at its own address, until the process ends.  On the host the array's name is only a
// stand-in, not a device address: the address is looked up once per device with
// hipGetSymbolAddress (with that device current) and kept, and host code passes only that
// address to kernels.  Every call on a GPU shares that GPU's array, which keeps the last
// call's contents: the loop control is written anew by every call, and the registers are
// written before they are read.  Calls that overlap on one GPU race.
//
// The work is the same sequence of kernels wfv_advance enqueues for a fixed-step run (the CFL
// minimum of the initial state, then per step the step control and the RK stages), so the
// results are bit-identical to it.  The grid widths are computed on the device by the
// functions the host uses for a context (wenofv_geometry.h).
#include <hip/hip_runtime.h>

#include <mutex>

#include "wenofv.h"
#include "wenofv_geometry.h"
#include "wenofv_host.h"
#include "wenofv_kernels.h"

static __device__ __align__(256) unsigned char scratch[WFV_NOALLOC_SCRATCH_BYTES];

namespace {

using namespace wfv;

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[MAX_DEVICES];  // per device: the device address of `scratch`

// the device address of `scratch` on device dev, which is current; looked up on the first
// call on that device
hipError_t scratch_of(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        g_scratch[dev] = ptr;
    }
    *out = (unsigned char *)g_scratch[dev];
    return hipSuccess;
}

constexpr size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// layout of one call's scratch: loop control, dxC, dxW [nx], dyC [ny], registers 1 and 2
struct Layout {
    size_t ctl, dxC, dxW, dyC, reg1, reg2, total;
};
Layout layout(int nx, int ny) {
    Layout L;
    const size_t nc = (size_t)nx * ny;
    L.ctl = 0;
    L.dxC = align256(sizeof(StepCtl));
    L.dxW = L.dxC + align256(8 * (size_t)nx);
    L.dyC = L.dxW + align256(8 * (size_t)nx);
    L.reg1 = L.dyC + align256(8 * (size_t)ny);
    L.reg2 = L.reg1 + align256(32 * nc);
    L.total = L.reg2 + align256(32 * nc);
    return L;
}

// loop control of a fixed-step run from t_start (as wfv_advance arms it) and the grid widths
__global__ __launch_bounds__(256) void noalloc_setup_kernel(StepCtl *ctl, double *dxC, double *dxW, double *dyC,
                                                            int nx, int ny, double xL, double dX, double yL, double dY,
                                                            int nranks, double t_start, int fixed_steps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        ctl->dtmin_bits = ~0ull;
        ctl->dt = 0.0;
        ctl->now = t_start;
        ctl->t_end = 0.0;
        ctl->count = 0;
        ctl->fixed_steps = fixed_steps;
        ctl->active = 1;
        ctl->flags = 0;
        ctl->dts = nullptr;
        ctl->max_record = 0;
        ctl->max_count = 0;
    }
    if (k < nx) column_geometry(k, nx, xL, dX, nranks, dxC + k, dxW + k, nullptr);
    if (k < ny) row_geometry(k, ny, yL, dY, dyC + k, nullptr);
}

int refuse(int code, const char *msg) {
    set_thread_error(msg);
    return code;
}

#define NACHK(call)                                                     \
    do {                                                                \
        const hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return refuse(WFV_ERR_HIP, hipGetErrorString(e_)); \
    } while (0)

}  // namespace

extern "C" int wfv_advance_fixed_noalloc(const wfv_options *opt, const wfv_case *cs, double *dU, double t_start,
                                         int fixed_steps, void *hip_stream, int *nsteps, double *t_out) {
    if (!opt || !cs || !dU) return refuse(WFV_ERR_ARG, "null argument");
    if (fixed_steps < 1) return refuse(WFV_ERR_ARG, "fixed_steps must be positive");
    if (const char *why = check_options(opt, cs)) return refuse(WFV_ERR_CONFIG, why);
    if (opt->boundary == 3) return refuse(WFV_ERR_CONFIG, "fixed ghost cells (boundary 3) need a context");
    const int nx = opt->elem_num_x, ny = opt->elem_num_y;
    const Layout lay = layout(nx, ny);
    if (lay.total > (size_t)WFV_NOALLOC_SCRATCH_BYTES)
        return refuse(WFV_ERR_ARG, "grid too large for the allocation-free scratch (WFV_NOALLOC_SCRATCH_BYTES)");
    int dev = 0;
    NACHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return refuse(WFV_ERR_ARG, "device index beyond the scratch table");
    unsigned char *base = nullptr;
    NACHK(scratch_of(dev, &base));
    const hipStream_t s = (hipStream_t)hip_stream;

    StepCtl *ctl = (StepCtl *)(base + lay.ctl);
    double *dxC = (double *)(base + lay.dxC), *dxW = (double *)(base + lay.dxW), *dyC = (double *)(base + lay.dyC);
    double *const reg[3] = {dU, (double *)(base + lay.reg1), (double *)(base + lay.reg2)};
    const double dX = (cs->xR - cs->xL) / nx, dY = (cs->yR - cs->yL) / ny;
    const int nranks = opt->geometry_ranks < 1 ? 1 : opt->geometry_ranks;
    const int nthreads = nx > ny ? nx : ny;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3((nthreads + 255) / 256), dim3(256), 0, s, ctl, dxC, dxW, dyC, nx, ny,
                       cs->xL, dX, cs->yL, dY, nranks, t_start, fixed_steps);
    NACHK(hipGetLastError());

    const int ty = opt->rows_per_strip > 0 ? (opt->rows_per_strip < ny ? opt->rows_per_strip : ny)
                                           : default_rows_per_strip(nx, ny);
    const StageArgs A = stage_base(*opt, *cs, nx, ny, ty, dxC, dxW, dyC, nullptr, ctl);
    // CFL minimum of the initial state, then the steps
    const DtArgs D{dU, A.P, A.dX, A.dY, A.cfl, nx, ny, &ctl->flags, &ctl->dtmin_bits};
    NACHK(launch_dt(D, s));
    for (int k = 0; k < fixed_steps; ++k) {
        NACHK(launch_step_control(ctl, s));
        NACHK(enqueue_rk_step(A, reg, *opt, 0.0, &ctl->dt, &ctl->active, &ctl->dtmin_bits, s));
    }
    if (!nsteps && !t_out) return WFV_OK;
    StepCtl h;
    NACHK(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, s));
    NACHK(hipStreamSynchronize(s));
    if (nsteps) *nsteps = h.count;
    if (t_out) *t_out = h.now;
    if (h.flags & WFV_FLAG_NAN_DT) return refuse(WFV_ERR_NONFINITE_DT, "Error: exist nan DeltaT");
    return WFV_OK;
}
