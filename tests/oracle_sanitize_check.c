/* Memory and undefined-behaviour check of the oracle (oracle/wenofv_oracle.c), built with
 * AddressSanitizer and UndefinedBehaviorSanitizer (tests/test_oracle_sanitize.py): every
 * scheme, reconstruction form, flux, limiter setting and boundary kind, on grids with
 * ragged strips and the smallest sizes the boundaries allow, through one right-hand side
 * with face and theta output and a few RK3 steps.  Prints "ok <cases>"; a sanitizer report
 * ends the program with a nonzero status. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "wenofv_oracle.h"

static double uni(unsigned long long *s) {
    *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(*s >> 11) / 9007199254740992.0;
}

static void fill(double *U, int nx, int ny, unsigned long long seed) {
    const size_t n = (size_t)nx * ny;
    for (size_t c = 0; c < n; ++c) {
        const double z1 = 0.2 + 0.6 * uni(&seed), p = 1.0 + 0.05 * uni(&seed);
        const double u = 0.3 * (uni(&seed) - 0.5), v = 0.3 * (uni(&seed) - 0.5);
        const double a = z1 * (1.0 + p / 100.0), b = (1.0 - z1) * (0.1 + p / 100.0);
        U[c] = a;
        U[n + c] = b;
        U[2 * n + c] = (a + b) * u;
        U[3 * n + c] = (a + b) * v;
    }
}

int main(void) {
    const int shapes[][2] = {{4, 4}, {5, 9}, {12, 7}, {17, 20}, {33, 8}};
    int cases = 0;
    for (size_t si = 0; si < sizeof shapes / sizeof shapes[0]; ++si) {
        const int nx = shapes[si][0], ny = shapes[si][1];
        const size_t n = (size_t)nx * ny, npad = (size_t)(nx + 8) * (ny + 8);
        double *dxC = malloc(nx * sizeof(double)), *dxW = malloc(nx * sizeof(double));
        double *dyC = malloc(ny * sizeof(double));
        double *U = malloc(4 * n * sizeof(double)), *rhs = malloc(4 * n * sizeof(double));
        double *faces = malloc(64 * n * sizeof(double)), *theta = malloc(n * sizeof(double));
        double *ghost = malloc(4 * npad * sizeof(double));
        orc_geometry(nx, ny, 0.0, 1.0, 0.0, 1.0, 1, dxC, dxW, dyC, NULL, NULL);
        fill(ghost, nx + 8, ny + 8, 99);
        for (int scheme = 0; scheme < 3; ++scheme)
            for (int recon = 0; recon < 5; ++recon)
                for (int flux = 0; flux < 2; ++flux)
                    for (int bp = 0; bp < 2; ++bp)
                        for (int bc = 0; bc < 4; ++bc) {
                            /* the walls mirror once: 4 cells per direction at least */
                            if (bc == 1 || bc == 2) {
                                if (nx < 4 || ny < 4) continue;
                            }
                            orc_params p = {0};
                            p.nx = nx; p.ny = ny; p.scheme = scheme; p.recon = recon; p.flux = flux; p.bp = bp;
                            p.bc = bc; p.rk = 2; p.cfl = 0.4; p.dX = 1.0 / nx; p.dY = 1.0 / ny;
                            p.rho10 = 1.0; p.rho20 = 0.1; p.sound1 = 10.0; p.sound2 = 10.0; p.preb = 0.0;
                            p.gravity = bc == 1 ? 9.81 : 0.0;
                            p.dxC = dxC; p.dxW = dxW; p.dyC = dyC; p.ghost = bc == 3 ? ghost : NULL;
                            fill(U, nx, ny, 1000u + 7u * (unsigned)cases);
                            orc_assemble_rhs(&p, U, rhs, faces, theta);
                            double dts[4];
                            int nsteps = 0;
                            double t = 0.0;
                            orc_run(&p, U, 0.0, 1.0, 3, dts, 4, &nsteps, &t);
                            ++cases;
                        }
        free(dxC); free(dxW); free(dyC); free(U); free(rhs); free(faces); free(theta); free(ghost);
    }
    printf("ok %d\n", cases);
    return 0;
}
