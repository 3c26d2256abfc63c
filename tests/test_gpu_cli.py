"""The command-line driver (config/settings.cfg, the reference's CLI behaviour) on the GPU:
its accuracy output reproduces the norms the reference itself wrote for the SMOOTH case."""
import glob
import os
import subprocess

import numpy as np
import pytest

from conftest import PKG
from known_answers import KNOWN

pytestmark = pytest.mark.gpu
CLI = os.path.join(PKG, "wenofv_run")

SETTINGS = """# SMOOTH = 0
TESTCASE = 0
ELEMNUMX = {n}
ELEMNUMY = {n}
# 0:RK1 1:RK2 2:RK3
RKMETH = 2
CFL = {cfl}
SCHEME = {scheme}
PPlimiter = {bp}
RIEMANNFLUX = {flux}
RECONSTRUCTIONTYPE = {recon}
"""


@pytest.mark.parametrize("row", KNOWN, ids=lambda k: f"{k[0]}-s{k[1]}r{k[2]}f{k[3]}b{k[4]}")
def test_cli_reproduces_reference_accuracy(row, tmp_path):
    n, scheme, recon, flux, bp, cfl, linf, l1, l2 = row
    os.makedirs(tmp_path / "config")
    (tmp_path / "config" / "settings.cfg").write_text(
        SETTINGS.format(n=n, cfl=cfl, scheme=scheme, bp=bp, flux=flux, recon=recon))
    print(f"cli smooth {n} s{scheme} r{recon} f{flux} b{bp}", flush=True)
    res = subprocess.run([CLI], cwd=tmp_path, capture_output=True, text=True, timeout=100)
    assert res.returncode == 0, res.stderr
    (csv,) = glob.glob(str(tmp_path / "output" / "*" / "accuracy_final.csv"))
    vals = [float(x) for x in open(csv).read().splitlines()[1].split(",")[1:]]
    np.testing.assert_allclose(vals, (linf, l1, l2), rtol=1e-12)


def test_cli_reports_missing_key(tmp_path):
    os.makedirs(tmp_path / "config")
    (tmp_path / "config" / "settings.cfg").write_text("TESTCASE = 0\nELEMNUMX = 16\n")
    res = subprocess.run([CLI], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert res.returncode == 1 and "missing option" in res.stderr
