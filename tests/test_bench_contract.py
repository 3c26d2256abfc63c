"""bench.py's contract pieces that need no GPU: one GPU only (a request for more exits with a
message); the committed rocprofv3 summaries carry every machine-readable figure the bench
line reads from them (roofline traffic, VALU busy and issue fractions, instructions per
cell-stage, the profiled stage duration); the roofline block is the VALU bound with the HBM
one beside it; the accuracy block counts instances against the tolerance and the FP64
spread and locates the ones beyond both."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import REPO

sys.path.insert(0, REPO)
import bench  # noqa: E402

KEYS = ("stage_kernel_avg_ns", "traffic_bytes_per_launch", "algorithmic_bytes_per_launch", "valu_issue_frac",
        "valu_busy_frac", "lane_instructions_per_cell_stage")


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 2 and "one GPU only" in res.stderr and res.stdout == ""


@pytest.mark.parametrize("name", ["batch_stage_mixed_batch256.md", "batch_stage_fp64_batch256.md"])
def test_profile_summaries_hold_the_bench_figures(name):
    for key in KEYS:
        v = bench.profiled(name, key)
        assert v is not None and v > 0, key
    # quad-cycles with a VALU instruction executing, summed over the waves: it passes 1 when
    # two waves' instructions overlap, up to the two resident waves per SIMD
    assert 0.5 < bench.profiled(name, "valu_busy_frac") <= 2.0
    assert 0.5 < bench.profiled(name, "valu_issue_frac") <= 1.0
    # the summary's algorithmic bytes are the roofline unit (85.33 B per cell-stage) over the
    # batch's cells, and its measured traffic is at least that
    cells = 14315520
    assert abs(bench.profiled(name, "algorithmic_bytes_per_launch") - bench.BYTES_PER_CELL_STAGE * cells) < 2
    assert bench.profiled(name, "traffic_bytes_per_launch") >= bench.profiled(name, "algorithmic_bytes_per_launch")


def test_roofline_is_the_valu_bound_with_the_hbm_one_beside_it():
    opts = dict(scheme=1, recon=0, flux=1, bp=1)
    cells, stage_ms = 1_000_000, 0.1
    r = bench.roofline(cells, stage_ms, opts, traffic=1.0e8)
    assert r["bound"] == "valu" and r["unit"] == "TFLOP/s" and r["peak"] == 78.6
    assert r["achieved"] == pytest.approx(3868.3 * cells / (stage_ms * 1e-3) / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / 78.6)
    h = r["hbm"]
    assert h["bound"] == "hbm" and h["unit"] == "GB/s" and h["traffic"] == 1.0e8
    assert h["achieved"] == pytest.approx(256.0 / 3.0 * cells / (stage_ms * 1e-3) / 1e9)
    # an option set without an operation count falls back to the HBM roofline
    assert bench.roofline(cells, stage_ms, dict(opts, scheme=0))["bound"] == "hbm"


def test_accuracy_block_counts_instances_against_tolerance_and_spread():
    import cases
    wls = [cases.batch_instance(sp) for sp in cases.batch_specs(3, seed=4)]
    Sd = [w.U0.copy() for w in wls]
    rng = np.random.default_rng(1)
    # instance 0 exact, 1 off by 1e-9 relative, 2 off by 1e-3 in the momenta at one cell
    S = [Sd[0].copy(), Sd[1] * (1 + 1e-9 * rng.standard_normal(Sd[1].shape)), Sd[2].copy()]
    w2 = wls[2]
    j, i = w2.ny // 2, w2.nx // 2
    S[2][2:, j, i] += 1e-3 * np.abs(Sd[2][:2, j, i].sum()) + 1.0
    spread = [s * (1 + 1e-12) for s in Sd]
    acc = bench.accuracy_block(wls, S, Sd, spread, [spread, spread], 500)
    pi = acc["per_instance_max_field_rel_l1"]
    assert acc["steps"] == 500 and pi["fraction_within_tolerance"] == pytest.approx(2 / 3)
    assert acc["fraction_within_tolerance_or_fp64_6e-8_spread"] == pytest.approx(2 / 3)
    (b,) = acc["beyond_tolerance_and_spread"]
    assert b["instance"] == 2 and b["field"] in ("u", "v") and (b["where"]["i"], b["where"]["j"]) == (i, j)
    assert len(b["fp64_6e-8_spread_per_direction"]) == 2
    assert acc["rel_l1_per_field_batch"]["z1rho1"] < 1e-8
    # an instance's spread is the largest over the perturbation directions: a second direction
    # that moves instance 2 as far as the mixed state covers it, the first direction alone does not
    wide = [spread[0], spread[1], S[2].copy()]
    acc = bench.accuracy_block(wls, S, Sd, spread, [spread, wide], 500)
    assert acc["fraction_within_tolerance_or_fp64_6e-8_spread"] == pytest.approx(1.0)
    assert acc["fraction_within_tolerance_or_fp64_6e-8_spread_first_direction"] == pytest.approx(2 / 3)
    assert acc["beyond_tolerance_and_spread"] == []
