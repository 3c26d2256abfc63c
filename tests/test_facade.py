"""The CWENOFV / TwoPhaseEquation facade as reference code uses it, on the host.

tests/facade_check.cpp names every public member and method of the reference's CWENOFV
(include/WENOFV.hpp:108-257) and TwoPhaseEquation (include/Equations.hpp:52-1044) with the
reference's signatures; it is compiled here with g++ against the facade headers and linked
with libwenofv.so, which shows that code written against the reference's classes compiles
and links against this build.  Its point mode then runs, without a GPU, the facade's point
functions, which are the kernels' all-FP64 functions compiled for the host, on the inputs of
the reference's own fixtures, and every output must equal the reference's bit for bit: the
equation of state, maximum eigenvalues, physical fluxes and eigenvector matrices
(point_eos.npz), LLF and HLLC fluxes through getRiemannFlux (point_riemann.npz), useWENO for
WENO-JS, WENO-Z (point_weno.npz) and WENO-EXP (point_wenoexp.npz), and
getPositivityPreservingTheta on built cells (point_theta.npz).  The output variables of
getVitalVarVal are checked against their definitions (include/Equations.hpp:1005-1027)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import PKG, REPO, golden

LIB = os.path.join(PKG, "libwenofv.so")
GXX = shutil.which("g++")


def compile_check(out):
    cmd = [GXX, "-std=c++17", "-O1", "-Wall", "-Werror", "-I" + os.path.join(REPO, "include"),
           "-I" + os.path.join(PKG, "csrc"), os.path.join(REPO, "tests", "facade_check.cpp"), "-L" + PKG, "-lwenofv",
           "-Wl,-rpath," + PKG, "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr[-3000:]


@pytest.fixture(scope="module")
def point_outputs(tmp_path_factory):
    if not (GXX and os.path.exists(LIB)):
        pytest.skip("g++ or libwenofv.so missing")
    d = tmp_path_factory.mktemp("facade")
    exe = d / "facade_check"
    compile_check(exe)
    eos, rie, wen, wex, th = (golden(f) for f in ("point_eos.npz", "point_riemann.npz", "point_weno.npz",
                                                   "point_wenoexp.npz", "point_theta.npz"))
    w = lambda name, a: np.ascontiguousarray(a, dtype=np.float64).tofile(d / name)
    nt = th["c1_cells"].shape[0]
    assert th["c70_cells"].shape[0] == nt
    w("sizes.bin", [eos["U"].shape[0], eos["U"].shape[1], rie["UL"].shape[0], rie["UL"].shape[1],
                    wen["stencils"].shape[0], wex["stencils"].shape[0], wex["dx"].shape[0], nt])
    w("eos_params.bin", eos["params"])
    w("eos_U.bin", eos["U"])
    w("rie_params.bin", rie["params"])
    w("rie_UL.bin", rie["UL"])
    w("rie_UR.bin", rie["UR"])
    w("weno_st.bin", wen["stencils"])
    w("exp_st.bin", wex["stencils"])
    w("exp_dx.bin", wex["dx"])
    w("theta_dxy.bin", [th["dX"], th["dY"]])
    for key in ("c1", "c70"):
        w(f"theta_{key}_params.bin", th[key + "_params"])
        w(f"theta_{key}_cells.bin", th[key + "_cells"])
    res = subprocess.run([str(exe), "point", str(d)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    r = lambda name: np.fromfile(d / name, dtype=np.float64)
    return dict(eos=r("eos_out.bin").reshape(eos["U"].shape[:2] + (86,)), rie=r("rie_out.bin"),
                weno=r("weno_out.bin").reshape(-1, 2, 4), exp=r("exp_out.bin").reshape(wex["exp"].shape),
                theta=r("theta_out.bin"), names=open(d / "names.txt").read().split())


def test_equation_of_state_fluxes_and_eigenvectors_match_reference(point_outputs):
    g = golden("point_eos.npz")
    o = point_outputs["eos"]
    np.testing.assert_array_equal(o[..., 0:3], g["eos"])
    np.testing.assert_array_equal(o[..., 3:5], g["eig"])
    np.testing.assert_array_equal(o[..., 5:13], g["flux"])
    np.testing.assert_array_equal(o[..., 13:77].reshape(g["LR"].shape), g["LR"])


def test_vital_variables(point_outputs):
    g = golden("point_eos.npz")
    o = point_outputs["eos"]
    V, U = o[..., 77:86], g["U"]
    z1, p = g["eos"][..., 0], g["eos"][..., 1]
    np.testing.assert_array_equal(V[..., 0], U[..., 0])
    np.testing.assert_array_equal(V[..., 1], U[..., 1])
    np.testing.assert_array_equal(V[..., 2], z1)
    np.testing.assert_array_equal(V[..., 3], U[..., 0] / z1)
    np.testing.assert_array_equal(V[..., 4], U[..., 1] / (1.0 - z1))
    np.testing.assert_array_equal(V[..., 5], U[..., 0] + U[..., 1])
    np.testing.assert_array_equal(V[..., 6], U[..., 2] / (U[..., 0] + U[..., 1]))
    np.testing.assert_array_equal(V[..., 7], U[..., 3] / (U[..., 0] + U[..., 1]))
    np.testing.assert_array_equal(V[..., 8], p)
    assert point_outputs["names"] == ["RHO1Z1", "RHO2Z2", "Z1", "RHO1", "RHO2", "RHO", "U", "V", "PRE"]


def test_riemann_fluxes_match_reference(point_outputs):
    g = golden("point_riemann.npz")
    np.testing.assert_array_equal(point_outputs["rie"].reshape(g["F"].shape), g["F"])


def test_use_weno_matches_reference(point_outputs):
    g = golden("point_weno.npz")
    np.testing.assert_array_equal(point_outputs["weno"][:, 0], g["js"])
    np.testing.assert_array_equal(point_outputs["weno"][:, 1], g["z"])


def test_use_weno_exp_matches_reference(point_outputs):
    np.testing.assert_array_equal(point_outputs["exp"], golden("point_wenoexp.npz")["exp"])


def test_positivity_theta_matches_reference(point_outputs):
    g = golden("point_theta.npz")
    np.testing.assert_array_equal(point_outputs["theta"], np.concatenate([g["c1_theta"], g["c70_theta"]]))
