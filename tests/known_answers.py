"""Accuracy norms the reference itself wrote (accuracy_final.csv of CWENOFV::outputAccuracy,
src/WENOFV.cpp:3152-3251) for the SMOOTH case (TESTCASE 0) at t = 0.05, SSP-RK3.
Columns: cells per side, SCHEME, RECONSTRUCTIONTYPE, RIEMANNFLUX, PPlimiter, CFL,
L-inf, L1, L2 of the volume fraction z1 against the exact cell averages."""
import numpy as np

KNOWN = [
    (64, 0, 0, 0, 0, 0.4, 2.723367646728647e-05, 1.430925290169891e-05, 1.669623161789131e-05),
    (64, 1, 1, 1, 1, 0.4, 1.291158859073605e-05, 2.710294813085365e-06, 4.437817289545201e-06),
    (100, 1, 1, 0, 0, 0.5, 3.841719624329831e-07, 2.400853110142917e-07, 2.689826478343496e-07),
    (128, 0, 0, 0, 0, 0.4, 8.959543789759106e-07, 4.385185658570656e-07, 5.073015442039611e-07),
    (128, 1, 0, 1, 1, 0.4, 5.260008001162220e-06, 5.544307332086379e-07, 1.272185228754960e-06),
    (256, 1, 0, 1, 1, 0.4, 1.745234240346782e-06, 9.679611370939215e-08, 3.039245770986761e-07),
]

GLP = np.array([-0.861136311594052575223946488893, -0.339981043584856264802665759103,
                0.339981043584856264802665759103, 0.861136311594052575223946488893])
GLW = np.array([0.347854845137453857373063949222, 0.652145154862546142626936050778,
                0.652145154862546142626936050778, 0.347854845137453857373063949222])


def smooth_accuracy(U, n, t):
    """L-inf, L1, L2 of z1 against the exact SMOOTH solution, as outputAccuracy computes them."""
    dX = dY = 1.0 / n
    idx = np.arange(n) + 3
    xl = 0.0 - 3 * dX + idx * dX
    xr = 0.0 - 3 * dX + (idx + 1) * dX
    xc = (xl + xr) / 2
    eps = 1e-8
    ex = np.zeros((n, n))  # [iy][ix]
    for gi in range(4):
        for gj in range(4):
            gx = xc + 0.5 * dX * GLP[gi]
            wx, wy = 0.5 * dX * GLW[gi], 0.5 * dY * GLW[gj]
            z = eps + (1 - 2 * eps) * np.sin(2.0 * np.pi * (gx - 1.0 * t)) ** 2
            ex += (z[None, :] * wx * wy / dX / dY) * np.ones((n, 1))
    c1 = c2 = 1.0
    q = 1.0 * c2 - 1.2 * c1
    qs = U[1] * c2 - U[0] * c1
    r = ((q - qs) + np.sqrt((q - qs) * (q - qs) + 4.0 * U[0] * c1 * U[1] * c2)) / (2.0 * U[1] * c2)
    z1 = r / (1.0 + r)
    e = np.abs(ex - z1)
    return e.max(), e.sum() / n / n, np.sqrt((e ** 2).sum() / n / n)
