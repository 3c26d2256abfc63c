"""The algorithm's operation count per cell-stage (oracle/flop_count.cpp): the counting type
against known answers, and the figures bench.py reports (ALGO_FLOPS) against what the counter
measures on the oracle for every configuration they cover."""
import os
import subprocess
import sys

import pytest

from conftest import REPO

sys.path.insert(0, REPO)
import bench  # noqa: E402

ORACLE = os.path.join(REPO, "oracle")
NAMES = {(0, 0, 0, 0): "js_cons_llf", (1, 0, 1, 1): "z_cons_hllc_bp", (1, 0, 0, 0): "z_cons_llf",
         (1, 1, 0, 0): "z_char_llf", (1, 1, 1, 1): "z_char_hllc_bp", (1, 2, 1, 1): "z_prim_hllc_bp",
         (1, 3, 1, 1): "z_prim2_hllc_bp", (1, 4, 1, 1): "z_prim3_hllc_bp", (2, 0, 1, 1): "exp_cons_hllc_bp"}


@pytest.fixture(scope="module")
def counter():
    subprocess.check_call(["make", "-s", "-C", ORACLE, "flop_count"])
    return os.path.join(ORACLE, "flop_count")


def test_counting_known_answers(counter):
    res = subprocess.run([counter, "--selftest"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "selftest ok" in res.stdout, res.stdout


def test_bench_figures_match_the_count(counter):
    assert set(bench.ALGO_FLOPS) == set(NAMES)
    res = subprocess.run([counter] + list(NAMES.values()), capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stderr
    rows = {}
    for line in res.stdout.splitlines():
        if line.startswith("#"):
            continue
        f = line.split("|")[0].split()
        rows[f[0]] = [float(x) for x in f[1:]]
    for key, name in NAMES.items():
        distinct, literal, rhs_d, rhs_l = rows[name]
        assert abs(bench.ALGO_FLOPS[key] - distinct) < 0.06, (name, distinct)
        # the reference recomputes (volume fraction per EOS call, indicators per node)
        assert literal > 2 * distinct and rhs_l > rhs_d
        # the right-hand side is nearly all of a stage: RK combination + a third of the CFL step
        assert 0 < distinct - rhs_d < 60
    # WENO-Z costs more than WENO-JS, the characteristic form more than the conservative one,
    # HLLC with the limiter more than LLF without it
    f = bench.ALGO_FLOPS
    assert f[(0, 0, 0, 0)] < f[(1, 0, 0, 0)] < f[(1, 0, 1, 1)] < f[(1, 1, 1, 1)]
    assert f[(1, 0, 0, 0)] < f[(1, 1, 0, 0)]
