"""The oracle's right-hand side, limiter and time loop against the reference's unmodified
4-rank solver (tests/golden/solver_*.npz from tests/golden/make_golden.py)."""
import numpy as np
import pytest

import oracle as orc
from conftest import golden

# SCHEME 0/1 x RECONSTRUCTIONTYPE 0-3, and SCHEME 2 (WENO-EXP) with RECONSTRUCTIONTYPE 4
# (primitive3, THINC on z1) in their own files
RHS_FILES = {"base": golden("solver_rhs.npz"), "exp": golden("solver_rhs_exp.npz")}
RUN_FILES = {"base": golden("solver_runs.npz"), "exp": golden("solver_runs_exp.npz")}

# reference test-case parameters (include/Equations.hpp:490-628): domain, rho10, rho20, c1, c2, preb
CASE = {
    0: ((0.0, 1.0, 0.0, 1.0), 1.2, 1.0, 1.0, 1.0, 0.0),
    1: ((-1.0, 1.0, -1.0, 1.0), 1000.0, 1.0, 30.0, 30.0, 0.0),
    2: ((0.0, 1.0, 0.0, 1.0), 1000.0, 1.0, 70.0, 70.0, 0.0),
}


def rhs_cases():
    for src, RHS in RHS_FILES.items():
        for key in sorted(RHS.files):
            if key.endswith("_rhs"):
                yield src, key[: -len("_rhs")]


@pytest.mark.parametrize("src,tag", list(rhs_cases()))
def test_assemble_rhs_matches_reference(src, tag):
    RHS = RHS_FILES[src]
    kind, s, r, f, b = tag.split("_")
    params = RHS[kind + "_params"]
    U = RHS[kind + "_U"]
    ny, nx = U.shape[1:]
    o = orc.Oracle(nx, ny, scheme=int(s[1]), recon=int(r[1]), flux=int(f[1]), bp=int(b[1]), bc=0,
                   rho10=params[0], rho20=params[1], sound1=params[2], sound2=params[3], preb=params[4])
    rhs, faces, theta, err = o.assemble_rhs(U, want_faces=True)
    np.testing.assert_array_equal(rhs, RHS[tag + "_rhs"])
    if tag + "_faces" in RHS.files:
        np.testing.assert_array_equal(faces, RHS[tag + "_faces"])
    if int(b[1]):
        flag = (theta < 1 - 1e-8).astype(float)
        np.testing.assert_array_equal(flag, RHS[tag + "_thflag"])
        if kind == "sharp":
            assert flag.sum() > 0  # the limiter is active on this state


def run_names():
    return [(src, k[: -len("_opts")]) for src, R in RUN_FILES.items() for k in sorted(R.files) if k.endswith("_opts")]


@pytest.mark.parametrize("src,name", run_names())
def test_time_loop_matches_reference(src, name):
    RUNS = RUN_FILES[src]
    tc, nx, ny, rk, cfl, scheme, bp, flux, recon, fixed, tend = RUNS[name + "_opts"]
    dom, rho10, rho20, c1, c2, preb = CASE[int(tc)]
    o = orc.Oracle(int(nx), int(ny), domain=dom, scheme=int(scheme), recon=int(recon), flux=int(flux),
                   bp=int(bp), rk=int(rk), cfl=cfl, rho10=rho10, rho20=rho20, sound1=c1, sound2=c2, preb=preb)
    U0 = RUNS[name + "_init"]
    if fixed:
        U, dts, t, err = o.run(U0, fixed_steps=int(tend))
    else:
        U, dts, t, err = o.run(U0, t_end=tend)
    assert err == 0
    np.testing.assert_array_equal(dts, RUNS[name + "_dts"])
    ref = RUNS[name + "_final"]
    np.testing.assert_array_equal(U, ref)
