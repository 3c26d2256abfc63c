"""The CWENOFV facade's solver methods on the GPU, driven as reference code drives the
reference's class (tests/facade_check.cpp, built by build() with the host compiler against
the facade headers).

- assembleRHS, and the reference's assembleRHS taken step by step through the public
  methods (getFlux_* for the form, useBoundPreservingLimiter, assembleBoundaryFaceTerm,
  assembleSourceTerm), on the 40 states of tests/golden/solver_rhs.npz: m_rhs, the limited
  face values of m_gridFlux and the limiter flags of m_totalTheta equal the reference's
  bit for bit, both ways; getFlux_* also fills the ring of ghost cells, which on these
  periodic grids holds exactly the face values of the wrapped interior cells;
- the reference's time loop through calculateDeltaT and RunRK1/2/3 (the 8 runs of
  solver_runs.npz): every dt and the final m_Uh bit for bit, m_Un the last step's base state;
- every other public method once, and a NaN cell: calculateDeltaT writes the reference's
  DeltaT file, which flags that cell, and stops."""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, golden

pytestmark = pytest.mark.gpu
EXE = os.path.join(REPO, "tests", "facade_check")
RHS = golden("solver_rhs.npz")
RUNS = golden("solver_runs.npz")
TESTCASE = {"smooth": 0, "sharp": 2}


def rhs_tags():
    return sorted(k[:-4] for k in RHS.files if k.endswith("_rhs"))


def run_names():
    return sorted(k[:-5] for k in RUNS.files if k.endswith("_opts"))


@pytest.fixture(scope="module")
def outputs(tmp_path_factory):
    assert os.path.exists(EXE), "tests/facade_check is built by __graft_entry__.build()"
    d = tmp_path_factory.mktemp("facade_gpu")
    w = lambda name, a: np.ascontiguousarray(a, dtype=np.float64).tofile(d / name)
    tags, runs = rhs_tags(), run_names()
    n = RHS["smooth_U"].shape[1]
    w("sizes.bin", [len(tags), n, len(runs)])
    opts = []
    for k, tag in enumerate(tags):
        kind, s, r, f, b = tag.split("_")
        opts.append([TESTCASE[kind], int(s[1]), int(r[1]), int(f[1]), int(b[1])])
        w(f"rhs_U_{k}.bin", RHS[kind + "_U"])
    w("rhs_opts.bin", opts)
    ropts = []
    for name in runs:
        tc, nx, ny, rk, cfl, scheme, bp, flux, recon, fixed, tend = RUNS[name + "_opts"]
        assert nx == ny
        ropts.append([tc, nx, rk, cfl, scheme, bp, flux, recon, fixed, tend])
    w("run_opts.bin", ropts)
    print(f"facade_check solver: {len(tags)} right-hand sides, {len(runs)} runs", flush=True)
    res = subprocess.run([EXE, "solver", str(d)], cwd=d, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-3000:]
    return d, tags, runs


def out(d, name, shape):
    return np.fromfile(d / name, dtype=np.float64).reshape(shape)


def test_rhs_fused_and_step_by_step_match_reference(outputs):
    d, tags, _ = outputs
    n = RHS["smooth_U"].shape[1]
    for k, tag in enumerate(tags):
        bp = int(tag.split("_")[4][1])
        for way in ("fused", "steps"):
            np.testing.assert_array_equal(out(d, f"rhs_{way}_{k}.bin", (4, n, n)), RHS[tag + "_rhs"], err_msg=f"{tag} {way}")
            if tag + "_faces" in RHS.files:
                np.testing.assert_array_equal(out(d, f"faces_{way}_{k}.bin", (n, n, 64)), RHS[tag + "_faces"],
                                              err_msg=f"{tag} {way} faces")
            if bp:
                np.testing.assert_array_equal(out(d, f"flags_{way}_{k}.bin", (n, n)), RHS[tag + "_thflag"],
                                              err_msg=f"{tag} {way} flags")


def test_ghost_ring_face_values_wrap_periodically(outputs):
    d, tags, _ = outputs
    n = RHS["smooth_U"].shape[1]
    for k, tag in enumerate(tags):
        F = out(d, f"faces_raw_{k}.bin", (n + 2, n + 2, 64))
        np.testing.assert_array_equal(F[0, 1:-1], F[n, 1:-1], err_msg=tag)      # row -1 = row n-1
        np.testing.assert_array_equal(F[n + 1, 1:-1], F[1, 1:-1], err_msg=tag)  # row n = row 0
        np.testing.assert_array_equal(F[1:-1, 0], F[1:-1, n], err_msg=tag)      # column -1 = column n-1
        np.testing.assert_array_equal(F[1:-1, n + 1], F[1:-1, 1], err_msg=tag)
        np.testing.assert_array_equal(F[0, 0], F[n, n], err_msg=tag)            # corner


def test_time_loop_through_the_step_methods_matches_reference(outputs):
    d, _, runs = outputs
    for k, name in enumerate(runs):
        ref = RUNS[name + "_final"]
        np.testing.assert_array_equal(np.fromfile(d / f"run_dts_{k}.bin", dtype=np.float64), RUNS[name + "_dts"],
                                      err_msg=name)
        np.testing.assert_array_equal(out(d, f"run_final_{k}.bin", ref.shape), ref, err_msg=name)
        Un = out(d, f"run_un_{k}.bin", ref.shape)
        assert np.isfinite(Un).all() and not np.array_equal(Un, ref)


def test_every_method_and_the_nan_time_step_map(outputs):
    d, _, _ = outputs
    outdir = (d / open(d / "outdir.txt").read().strip()).resolve()
    for name in ("DeltaT_check.plt", "DeltaT_0.25.plt", "average_check.plt", "accuracy_check.csv"):
        assert (outdir / name).exists(), (name, sorted(os.listdir(outdir)))
    lines = open(outdir / "DeltaT_0.25.plt").read().splitlines()
    assert lines[0] == "TITLE=FvSolution" and lines[2].startswith("ZONE T=TA, I=16, J=16")
    rows = np.array([[float(x) for x in ln.split()] for ln in lines[3:]])
    assert rows.shape == (256, 4)
    nan_rows = np.flatnonzero(rows[:, 3] == 1)
    # x-major order: cell (5, 2) is row 5 * 16 + 2
    assert list(nan_rows) == [5 * 16 + 2] and rows[nan_rows[0], 2] == 0
    assert (rows[rows[:, 3] == 0, 2] > 0).all()
    ok = [ln for ln in open(outdir / "DeltaT_check.plt").read().splitlines()[3:]]
    assert len(ok) == 256 and all(ln.split()[3] == "0" for ln in ok)
