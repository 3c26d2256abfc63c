"""bench.py's contract pieces that need no GPU: one GPU only (a request for more exits with a
message), and the committed rocprofv3 summaries carry every machine-readable figure the
bench line reads from them (roofline traffic, VALU busy and issue fractions, instructions
per cell-stage, the profiled stage duration)."""
import os
import subprocess
import sys

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
    assert 0.5 < bench.profiled(name, "valu_busy_frac") <= 1.0
    # the summary's algorithmic bytes are the roofline unit (85.33 B per cell-stage) over the
    # batch's cells, and its measured traffic is at least that
    cells = 14315520
    assert abs(bench.profiled(name, "algorithmic_bytes_per_launch") - bench.BYTES_PER_CELL_STAGE * cells) < 2
    assert bench.profiled(name, "traffic_bytes_per_launch") >= bench.profiled(name, "algorithmic_bytes_per_launch")
