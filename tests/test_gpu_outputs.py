"""The files `wenofv_run` writes, against the files the reference's own run loop wrote for the
same settings (tests/golden/reference_outputs.npz, recorded by make_golden.py's `outputs`
job), written at the reference's cadence (src/WENOFV.cpp:2447-2523): Tecplot cell averages
(src/WENOFV.cpp:2525-2600), accuracy CSV and averages (src/WENOFV.cpp:3152-3361), SHARP's
oscillation snapshots and time series (src/WENOFV.cpp:2739-2804, 2865-2940, 3049-3107).
FP64 runs the reference's arithmetic bit for bit and both sides format with the default
ostream settings, so every file must be byte-identical."""
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG, golden

pytestmark = pytest.mark.gpu
CLI = os.path.join(PKG, "wenofv_run")
OUT = golden("reference_outputs.npz")

SETTINGS = """TESTCASE = {tc}
ELEMNUMX = {nx}
ELEMNUMY = {ny}
RKMETH = {rk}
CFL = {cfl}
SCHEME = {scheme}
PPlimiter = {bp}
RIEMANNFLUX = {flux}
RECONSTRUCTIONTYPE = {recon}
ENDTIME = {tend}
"""


def runs():
    return sorted(k[:-len("__opts")] for k in OUT.files if k.endswith("__opts"))


def first_difference(a, b):
    la, lb = a.splitlines(), b.splitlines()
    for i, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return f"line {i + 1}: ours {x!r} reference {y!r}"
    return f"line counts differ: ours {len(la)} reference {len(lb)}"


@pytest.mark.parametrize("name", runs())
def test_output_files_match_reference(name, tmp_path):
    tc, nx, ny, rk, cfl, scheme, bp, flux, recon, tend = OUT[name + "__opts"]
    os.makedirs(tmp_path / "config")
    (tmp_path / "config" / "settings.cfg").write_text(SETTINGS.format(
        tc=int(tc), nx=int(nx), ny=int(ny), rk=int(rk), cfl=repr(float(cfl)), scheme=int(scheme), bp=int(bp),
        flux=int(flux), recon=int(recon), tend=repr(float(tend))))
    print(f"outputs {name}", flush=True)
    res = subprocess.run([CLI], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert res.returncode == 0, res.stderr
    (d,) = os.listdir(tmp_path / "output")
    ours = sorted(f for f in os.listdir(tmp_path / "output" / d) if f.endswith((".plt", ".csv")))
    expected = [str(f) for f in OUT[name + "__files"]]
    assert ours == expected
    for f in expected:
        ref = bytes(OUT[name + "__" + f]).decode()
        got = open(tmp_path / "output" / d / f).read()
        assert got == ref, f"{f}: {first_difference(got, ref)}"
