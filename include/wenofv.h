/*
 * wenofv.h -- C ABI of the MI355X (gfx950) WENOFV time-stepping library (libwenofv.so).
 *
 * The reference has no plugin or FFI layer: its boundary is the C++ class CWENOFV
 * (include/WENOFV.hpp:108-257 of the reference) configured by the key = value options of
 * config/settings.cfg.  Each entry point below replaces one CWENOFV method (cited per
 * function) for a single domain on one GPU; the C++ facade (csrc/WENOFV.hpp) restores the
 * class interface on top of it.
 *
 * Conventions.  Host arrays are caller-owned; the library copies them.  A state array is
 * double[4][ny][nx] (SoA, x fastest) holding (z1*rho1, z2*rho2, rho*u, rho*v).  Every
 * function returns a WFV_* status; wfv_last_error() describes the last failure.  A context
 * is not thread-safe; distinct contexts are independent.  Nothing here exits or reads stdin.
 */
#ifndef WENOFV_H
#define WENOFV_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define WFV_OK 0
#define WFV_ERR_CONFIG 1        /* invalid option (reference: std::stoi throw, MPI_Abort) */
#define WFV_ERR_HIP 2           /* HIP runtime failure                                     */
#define WFV_ERR_NONFINITE_DT 3  /* NaN CFL time step (reference: exit(1), WENOFV.cpp:502-531) */
#define WFV_ERR_ARG 4           /* null pointer / bad size                                 */

/* ghost layers of a fixed-ghost boundary (boundary 3): the deepest stencil halo */
#define WFV_GHOST 4

/* error bits accumulated on the device (wfv_get_flags) */
#define WFV_FLAG_NAN_FLUX 1u   /* a Riemann flux was NaN (reference prints: Equations.hpp:344-359) */
#define WFV_FLAG_BAD_THETA 2u  /* limiter theta outside [0,1] (reference prints: WENOFV.cpp:1927)  */
#define WFV_FLAG_NAN_DT 4u     /* a NaN CFL time step: ends the wfv_advance / wfv_compute_dt call
                                  that meets it (WFV_ERR_NONFINITE_DT) and is cleared when the
                                  next such call starts, so a context given a valid state runs */

/* Solver options: the config/settings.cfg keys (reference config/settings.cfg:11-24, read in
   src/WENOFV.cpp:43-63) plus the options of this implementation (defaults reproduce the
   reference). */
typedef struct wfv_options {
    int elem_num_x;      /* ELEMNUMX */
    int elem_num_y;      /* ELEMNUMY */
    int rk_method;       /* RKMETH: 0 SSP-RK1, 1 SSP-RK2, 2 SSP-RK3 */
    double cfl;          /* CFL */
    int scheme;          /* SCHEME: 0 WENO-JS, 1 WENO-Z (5-point stencils); 2 WENO-EXP (7-point) */
    int pp_limiter;      /* PPlimiter: bound-preserving limiter on the partial densities */
    int riemann_flux;    /* RIEMANNFLUX: 0 LLF, 1 HLLC */
    int reconstruction;  /* RECONSTRUCTIONTYPE: 0 conservative, 1 characteristic,
                            2 primitive (z1, rho1, u, v), 3 primitive2 (z1, p, u, v),
                            4 primitive3 (THINC on z1 and WENO-Z on p, u, v; 7-point) */
    int precision;       /* 0: all-FP64, bit-faithful to the reference; 1: mixed precision (FP32
                            WENO of the momenta and FP32 Riemann fluxes; FP64 reconstruction of
                            the partial densities, equation of state, limiter theta and RK
                            accumulation); 2: mixed with every reconstruction in FP64;
                            3: all-FP64 with fused multiply-add (FP64-accurate, not
                            bit-faithful: a rounding-sensitivity control; 5-point schemes) */
    int boundary;        /* BOUNDARYTYPE (reference include/Equations.hpp:25-32; its MPI solver
                            runs only 0): 0 periodic; 1 slip walls (NEUMANN, FREESLIP: mirror,
                            normal momentum negated); 2 no-slip walls (NoSlip: mirror, both
                            momenta negated); 3 fixed ghost cells (NEUMANN2, "from the
                            initial values": wfv_set_ghost_cells) */
    int geometry_ranks;  /* cell widths as the reference's N-rank slab grid builds them (4),
                            or 1 for a single-domain grid */
    int rows_per_strip;  /* rows one wavefront marches (0: tuned default) */
} wfv_options;

/* Physical case: domain and TwoPhaseEquation parameters (reference include/Equations.hpp:61-70). */
typedef struct wfv_case {
    double xL, xR, yL, yR;
    double rho10, rho20;   /* g_rho10, g_rho20 */
    double sound1, sound2; /* g_sound1, g_sound2 */
    double preb;           /* g_preb */
    double gravity;        /* source term rhs[3] -= rho*g (assembleSourceTerm); 0: none */
} wfv_case;

typedef struct wfv_ctx wfv_ctx;

/* CWENOFV::CWENOFV + initializeSolver (include/WENOFV.hpp:114-118, src/WENOFV.cpp:16-230):
   validates the options, builds the grid and allocates the three RK registers on the
   current HIP device. */
int wfv_create(const wfv_options *opt, const wfv_case *cs, wfv_ctx **out);
void wfv_destroy(wfv_ctx *ctx);
const char *wfv_last_error(const wfv_ctx *ctx);

/* Cell averages in / out (the reference's m_Uh over interior cells, getWorldUh:
   src/WENOFV.cpp:2059-2117).  The _device forms take device pointers. */
int wfv_set_state(wfv_ctx *ctx, const double *U);
int wfv_get_state(wfv_ctx *ctx, double *U);
int wfv_set_state_device(wfv_ctx *ctx, const double *dU);
int wfv_get_state_device(wfv_ctx *ctx, double *dU);

/* CWENOFV::calculateDeltaT (src/WENOFV.cpp:474-540): min(1, min_cells CFL/(ax/dx+ay/dy)). */
int wfv_compute_dt(wfv_ctx *ctx, double *dt);

/* CWENOFV::performTimeIntegration (src/WENOFV.cpp:541-566): one SSP-RK step of size dt. */
int wfv_step(wfv_ctx *ctx, double dt);

/* CWENOFV::assembleRHS (src/WENOFV.cpp:568-615) of the current state.  rhs: [4][ny][nx];
   faces (optional): limited face values per cell [ny][nx][side][var][gp] with side order
   left, right, bottom, top (the reference's m_gridFlux); theta (optional): [ny][nx]. */
int wfv_assemble_rhs(wfv_ctx *ctx, double *rhs, double *faces, double *theta);

/* CWENOFV::getFlux_conservative / _characteristic / _primitive / _primitive2 / _primitive3
   (src/WENOFV.cpp:617-1372) of the current state: the face values (the reference's
   m_gridFlux) of reconstruction form `reconstruction` (RECONSTRUCTIONTYPE 0..4) with the
   context's scheme, boundary and precision, limited by the bound-preserving limiter
   (useBoundPreservingLimiter, src/WENOFV.cpp:1895-1935) when limit is 1, for the cells and
   the ring of ghost cells around them: faces [ny+2][nx+2][side][var][gp] (sides left, right,
   bottom, top; padded row 0 / column 0 hold the ghost cells below / left of the grid);
   theta (optional) [ny+2][nx+2] the limiter's theta (1 where limit is 0). */
int wfv_face_values(wfv_ctx *ctx, int reconstruction, int limit, double *faces, double *theta);

/* CWENOFV::calculateDeltaT's time step of every cell (m_totalDeltaT, src/WENOFV.cpp:489-500)
   of the current state: dt [ny][nx] = CFL/(ax/dx + ay/dy), NaN where the state gives none
   (the map the reference writes with outputDeltaT before it stops on a NaN time step). */
int wfv_cell_dt(wfv_ctx *ctx, double *dt);

/* CWENOFV::run's time loop (src/WENOFV.cpp:318-332) on the device: from t_start, steps of
   dt = min(CFL dt, t_end - now) until |now - t_end| <= 1e-10, or exactly fixed_steps
   unclipped CFL steps when fixed_steps > 0.  Steps are replayed from a HIP graph without
   host round trips.  dts (optional, host, max_dts entries) receives each dt. */
int wfv_advance(wfv_ctx *ctx, double t_start, double t_end, int fixed_steps, int *nsteps,
                double *t_out, double *dts, int max_dts);

/* Ghost cells of a boundary-3 context: Upad is [4][ny+2*WFV_GHOST][nx+2*WFV_GHOST] (the
   interior part is ignored).  They stay fixed while the solver runs.  No reference
   counterpart (the reference's setBoundaryAverages2 stops on NEUMANN2,
   src/WENOFV.cpp:2287-2292). */
int wfv_set_ghost_cells(wfv_ctx *ctx, const double *Upad);

/* Limiter flags of the reference's outputs (m_totalTheta, src/WENOFV.cpp:1906-1909): with
   recording on, every stage stores per cell 1.0 where the limiter's theta < 1-1e-8, else
   0.0; wfv_get_limiter_flags returns those of the last stage, [ny][nx]. */
int wfv_record_limiter_flags(wfv_ctx *ctx, int on);
int wfv_get_limiter_flags(wfv_ctx *ctx, double *flags);

/* The same time loop toward t_end, stopping early after max_steps steps (>= 1): lets a
   caller write the reference's step-count outputs (src/WENOFV.cpp:2483-2520) between runs. */
int wfv_advance_steps(wfv_ctx *ctx, double t_start, double t_end, int max_steps, int *nsteps, double *t_out);

/* Error bits raised by the kernels since the last call (WFV_FLAG_*); clears them. */
int wfv_get_flags(wfv_ctx *ctx, unsigned *flags);
int wfv_synchronize(wfv_ctx *ctx);

/* Run the context's work on a caller-owned HIP stream (hipStream_t passed as void*; NULL
   restores the context's own stream).  Lets a caller order the solver with its own work
   and time it with events on that stream.  No reference counterpart (the reference has no
   device). */
int wfv_set_stream(wfv_ctx *ctx, void *hip_stream);

/* ---------------------------------------------------------------------------------------
   Batches: many independent solver instances with uneven grids advanced together, one
   persistent launch per RK stage over all of them, each instance with its own time step
   and step count on the device.  No reference counterpart (the reference runs one case
   per MPI job); every instance behaves exactly as a single-domain context with the same
   options would.  Options shared by the batch come from wfv_options (elem_num_x/y are
   ignored); grid size, boundary and physical case are per instance. */
typedef struct wfv_instance {
    int nx, ny;       /* interior cells */
    int boundary;     /* as wfv_options.boundary */
    int pad;
    wfv_case cs;
} wfv_instance;

typedef struct wfv_batch wfv_batch;

int wfv_batch_create(const wfv_options *opt, const wfv_instance *inst, int n, wfv_batch **out);
void wfv_batch_destroy(wfv_batch *b);
const char *wfv_batch_last_error(const wfv_batch *b);
int wfv_batch_set_state(wfv_batch *b, int i, const double *U);  /* instance i, [4][ny][nx] */
int wfv_batch_get_state(wfv_batch *b, int i, double *U);
int wfv_batch_set_ghost_cells(wfv_batch *b, int i, const double *Upad);  /* see wfv_set_ghost_cells */
/* Every instance runs the reference time loop from t = 0 (see wfv_advance): to t_end with
   the last step clipped, or fixed_steps unclipped CFL steps.  nsteps and t_out (optional)
   receive each instance's step count and final time; flags (optional) its error bits. */
int wfv_batch_advance(wfv_batch *b, double t_end, int fixed_steps, int *nsteps, double *t_out,
                      unsigned *flags);
int wfv_batch_set_stream(wfv_batch *b, void *hip_stream);
/* rows per wavefront strip of instance i (0: the batch default) */
int wfv_batch_set_rows_per_strip(wfv_batch *b, int i, int rows);
int wfv_batch_get_rows_per_strip(const wfv_batch *b, int i, int *rows);
/* schedule of the stage launches: 0 one wavefront per (instance, strip) item, 1 a
   persistent grid of 8 wavefronts per CU striding over the items */
int wfv_batch_set_schedule(wfv_batch *b, int persistent);
int wfv_batch_get_schedule(const wfv_batch *b, int *persistent);
int wfv_batch_cells(const wfv_batch *b, long long *cells);  /* total interior cells */

/* ---------------------------------------------------------------------------------------
   Launch-shape autotuner.  No reference counterpart (the reference's CPU loops have no
   tile shape).  Times `trials` launches of the first RK stage on the current state (which
   it leaves unchanged) for each candidate rows-per-strip (1..128) and, for a batch, each
   schedule, and keeps the fastest; a batch is tuned per grid-size class (instances of equal
   nx x ny) by coordinate descent.  Results are also settable directly (a cache of earlier
   tunings): wfv_set_rows_per_strip, wfv_batch_set_rows_per_strip, wfv_batch_set_schedule. */
typedef struct wfv_tuning {
    int rows_per_strip;  /* single context: the choice (batches: per instance, see above) */
    int persistent;      /* batch schedule chosen */
    double stage_ms;     /* mean duration of one stage launch with the choice */
} wfv_tuning;
int wfv_autotune(wfv_ctx *ctx, int trials, wfv_tuning *best);
int wfv_batch_autotune(wfv_batch *b, int trials, wfv_tuning *best);
int wfv_set_rows_per_strip(wfv_ctx *ctx, int rows);
int wfv_get_rows_per_strip(const wfv_ctx *ctx, int *rows);

/* Grid of the context: interior cells and the per-column / per-row widths. */
int wfv_grid(const wfv_ctx *ctx, int *nx, int *ny, double *dxC, double *dxW, double *dyC);

/* TwoPhaseEquation::setEquationParameters (reference include/Equations.hpp:486-977) of a
   reference test case (TESTCASE 0..8): its domain and phase parameters, boundary (0
   periodic, 1 walls for the reference's non-periodic kinds) and final output time.  Host
   only, no device work. */
int wfv_reference_case(int testcase, wfv_case *cs, int *boundary, double *t_end);
/* CWENOFV::initializeAve (reference src/WENOFV.cpp:232-279) of a reference test case: 4x4
   Gauss-Legendre cell averages of TwoPhaseEquation::getU0 on the grid of geometry_ranks
   x-slabs; U: [4][ny][nx].  Host only, no device work. */
int wfv_reference_initial_state(int testcase, int nx, int ny, int geometry_ranks, double *U);
/* The same on the grid padded by WFV_GHOST cells on every side (cells extended with the
   same spacing): Upad [4][ny+2*WFV_GHOST][nx+2*WFV_GHOST], the ghost cells of boundary 3. */
int wfv_reference_initial_state_padded(int testcase, int nx, int ny, int geometry_ranks, double *Upad);

/* ---------------------------------------------------------------------------------------
   Point functions of the model, for one state U[4] = (z1 rho1, z2 rho2, rho u, rho v) of a
   case cs (its rho10, rho20, sound1, sound2, preb): the reference's inline TwoPhaseEquation
   and CWENOFV functions, evaluated by the all-FP64 kernels' own arithmetic (so they round
   as the reference's do).  Host only, no device work.  A normal (nx, ny) selects the x
   direction as the reference does: the physical flux where nx > ny, the eigenvectors where
   |nx| > |ny|; the Riemann fluxes take it as it is. */
/* TwoPhaseEquation::getVolumeFrac / getPressure / getSoundSpeed (include/Equations.hpp:84-146):
   out = z1, p, c, rho, rho1, rho2 */
int wfv_point_eos(const wfv_case *cs, const double *U, double *out);
/* TwoPhaseEquation::getPhyFlux (include/Equations.hpp:158-186) */
int wfv_point_flux(const wfv_case *cs, const double *U, double nx, double ny, double *F);
/* TwoPhaseEquation::getLLFRiemannFlux / getHLLCRiemannFlux (include/Equations.hpp:363-483):
   riemann_flux 0 LLF, 1 HLLC */
int wfv_point_riemann(const wfv_case *cs, int riemann_flux, const double *UL, const double *UR, double nx,
                      double ny, double *F);
/* TwoPhaseEquation::getLEigenMatrix / getREigenMatrix (include/Equations.hpp:189-327),
   row-major 4 x 4 */
int wfv_point_eigen(const wfv_case *cs, const double *U, double nx, double ny, double *L, double *R);
/* CWENOFV::useWENO (src/WENOFV.cpp:1873-1893) at the 4 Gauss-Lobatto nodes: s[7] the averages
   u_mmm..u_ppp, scheme 0 WENO-JS and 1 WENO-Z (on s[1..5]), 2 WENO-EXP (dx the cell width) */
int wfv_point_weno(int scheme, const double *s, double dx, double *out);
/* CWENOFV::getPositivityPreservingTheta (src/WENOFV.cpp:1936-2000) of a cell: faces[64] its
   face values [side][var][gp] (sides left, right, bottom, top), dX, dY the global widths */
int wfv_point_theta(const wfv_case *cs, const double *U, const double *faces, double dX, double dY, double *theta);

/* ---------------------------------------------------------------------------------------
   Allocation-free entry point.  No counterpart in the reference: it serves callers that
   cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
   forbid hipMalloc).  The fixed-step form of wfv_advance (fixed_steps unclipped CFL steps
   from t_start) on a caller-owned device state dU, [4][ny][nx], which is input and output;
   the grid (opt->elem_num_x x opt->elem_num_y, geometry_ranks), options and case are those
   wfv_create takes.  The RK registers 1 and 2, the grid widths and the loop control live in
   a static scratch array of WFV_NOALLOC_SCRATCH_BYTES on each GPU the library runs on (held
   until the process ends): a grid needs 64 nx ny + 16 nx + 8 ny bytes plus at most 1 KiB, so
   grids up to 1024 x 1024 cells fit; larger ones are refused (WFV_ERR_ARG), as is a
   fixed-ghost boundary (3: WFV_ERR_CONFIG).  Every call on a GPU shares that GPU's scratch,
   so calls that overlap on one GPU race.  The work is enqueued on hip_stream (a hipStream_t;
   NULL: the null stream).  With nsteps and t_out both NULL the call only enqueues kernels,
   with no allocation, copy or synchronization (so it can be captured into a graph); else it
   waits for them, returns the step count and final time, and WFV_ERR_NONFINITE_DT if a CFL
   time step was NaN.  Failures are described by wfv_last_error(NULL).  Results are
   bit-identical to wfv_advance(ctx, t_start, 0, fixed_steps, ...) on a context created with
   the same options and case and given the same state. */
#define WFV_NOALLOC_SCRATCH_BYTES (65u << 20)
int wfv_advance_fixed_noalloc(const wfv_options *opt, const wfv_case *cs, double *dU, double t_start,
                              int fixed_steps, void *hip_stream, int *nsteps, double *t_out);

/* Library build identification (kernel variants compiled for gfx950). */
const char *wfv_version(void);

#ifdef __cplusplus
}
#endif
#endif
