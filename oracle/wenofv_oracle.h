/*
 * wenofv_oracle.h -- CPU FP64 restatement of the reference WENOFV time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver: only the
 * tests, the smoke check and the CPU-baseline leg of bench.py load it.  The product
 * library never links or calls it, and there is no CPU fallback built on it.
 *
 * Every function follows the reference's operation order (no FMA contraction, same
 * association, same divisions) so that it reproduces the reference CPU solver to the
 * last bit on a single domain; the citation on each definition names the reference
 * file:line it restates (paths relative to the reference repository root).
 *
 * Pinning (how each function is checked, tests/test_oracle_*.py):
 *   orc_weno          golden vectors from the compiled reference (tests/golden/point_*.npz),
 *                     polynomial exactness, constant-stencil, JS/Z equal-beta identity
 *   orc_weno_exp(_thinc)  golden vectors from the reference's 7-point kernels
 *                     (tests/golden/point_wenoexp.npz)
 *   orc_eos / flux    golden vectors from the reference, p1(z1)=p2(z1) equilibrium,
 *                     Wood sound speed closed form, F(U,U)=F_phys(U) consistency
 *   orc_eigen_L/R     golden vectors, L*R = I
 *   orc_bp_theta      golden vectors, theta in [0,1], limited face values >= eps
 *   orc_assemble_rhs  golden right-hand sides and face values from the 4-rank reference
 *   orc_run           golden end states and dt sequences from the 4-rank reference
 *
 * State layout used by every array argument: SoA, U[v*ny*nx + j*nx + i], v = 0..3 holds
 * (z1*rho1, z2*rho2, rho*u, rho*v); i runs along x (fastest), j along y.
 */
#ifndef WENOFV_ORACLE_H
#define WENOFV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int nx, ny;        /* interior cells */
    int scheme;        /* 0 WENO-JS, 1 WENO-Z, 2 WENO-EXP                       */
    int recon;         /* 0 conservative, 1 characteristic, 2 primitive, 3 primitive2,
                          4 primitive3 (THINC on z1) */
    int flux;          /* 0 LLF, 1 HLLC                                          */
    int bp;            /* bound-preserving limiter on partial densities          */
    int bc;            /* 0 periodic, 1 slip walls, 2 no-slip walls, 3 fixed ghost cells */
    int rk;            /* 0 SSP-RK1, 1 SSP-RK2, 2 SSP-RK3                         */
    double cfl;
    double dX, dY;     /* global cell sizes (reference m_deltaX, m_deltaY)        */
    double rho10, rho20, sound1, sound2, preb;
    double gravity;    /* source term rhs[3] -= rho*gravity (0: no source)        */
    const double *dxC; /* [nx] x-width of each column                            */
    const double *dxW; /* [nx] x-width of the west neighbour column, as seen by
                          the rank that owns the column                          */
    const double *dyC; /* [ny] y-width of each row                               */
    const double *ghost; /* bc 3: padded state [4][ny+8][nx+8] holding the ghost cells */
} orc_params;

/* error bits returned by the stepping functions */
#define ORC_ERR_NAN_FLUX   1  /* a Riemann flux was NaN   (reference Equations.hpp:344-359) */
#define ORC_ERR_BAD_THETA  2  /* theta outside [0,1]      (reference WENOFV.cpp:1911-1932)  */
#define ORC_ERR_NAN_DT     4  /* non-finite CFL time step (reference WENOFV.cpp:502-531)    */

/* point kernels */
double orc_weno(int scheme, const double s[5], int gp);
/* 7-point stencils s = {u_mmm, ..., u_ppp}; dx is the reference's m_deltaX */
double orc_weno7(int scheme, const double s[7], int gp, double dx);       /* useWENO, SCHEME 0-2 */
double orc_weno_exp(const double s[7], int gp, double dx);               /* WENO-EXP            */
double orc_weno_exp_thinc(const double s[7], int gp, double dx);         /* z1 of PRIMITIVE3    */
double orc_volume_frac(const orc_params *p, const double U[4]);
double orc_pressure(const orc_params *p, const double U[4]);
double orc_sound_speed(const orc_params *p, const double U[4]);
double orc_max_eig(const orc_params *p, const double U[4], double nx, double ny);
void orc_phy_flux(const orc_params *p, const double U[4], double F[4], double nx, double ny);
void orc_riemann(const orc_params *p, const double UL[4], const double UR[4], double F[4],
                 double nx, double ny, int type);
void orc_eigen_L(const orc_params *p, const double U[4], double nx, double ny, double L[16]);
void orc_eigen_R(const orc_params *p, const double U[4], double nx, double ny, double R[16]);
/* faces[side*16 + v*4 + g], side 0 left, 1 right, 2 bottom, 3 top */
double orc_bp_theta(const orc_params *p, const double U[4], const double faces[64]);

/* geometry of the reference's 4-rank slab decomposition (nranks=1: single domain) */
void orc_geometry(int nx, int ny, double xL, double xR, double yL, double yR, int nranks,
                  double *dxC, double *dxW, double *dyC, double *xcen, double *ycen);

/* one right-hand side evaluation; faces_out (nx*ny*64, per cell [side][v][g], limited)
   and theta_out (nx*ny) may be NULL.  Returns error bits. */
int orc_assemble_rhs(const orc_params *p, const double *U, double *rhs, double *faces_out,
                     double *theta_out);
/* CFL time step min over cells of cfl/(ax/dX + ay/dY), capped at 1.0 */
int orc_compute_dt(const orc_params *p, const double *U, double *dt);
/* one SSP-RK step of size dt, in place on U */
int orc_rk_step(const orc_params *p, double *U, double dt);
/* reference run loop from t_start: to t_end with the last step clipped (fixed_steps<=0) or exactly
   fixed_steps CFL steps.  dts_out (optional) receives each dt; *nsteps_out the count. */
int orc_run(const orc_params *p, double *U, double t_start, double t_end, int fixed_steps,
            double *dts_out, int max_dts, int *nsteps_out, double *t_out);

#ifdef __cplusplus
}
#endif
#endif
