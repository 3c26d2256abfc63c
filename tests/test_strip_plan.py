"""The stage kernels' work decomposition, checked on the host: tests/strip_plan_check.cpp
compiles the kernels' own strip_plan / segment_of (hipcc, host code only) and verifies, over
35,640 combinations of grid shape, rows per strip (every value up to 40 and the larger ones
the autotuner tries) and stencil halo, that every cell is owned exactly once and that every
lane a cell's reconstruction and face fluxes read stays in its segment, on the adjacent
column, with the same rows, under the host sanitizers; that the autotuner's occupancy
candidates for a 1024 x 512 grid on 256 CUs are the rows per strip that fill whole rounds; and
that the XCD-aware strip order (xcd_strip) is a bijection giving each XCD one contiguous run
of strips; and that the batch's XCD-aware item order (batch_item_order) lists every item of
200 seeded random instance lists once, in priority order within each XCD, with balanced
loads."""
import os
import shutil
import subprocess

import pytest

from conftest import PKG, REPO

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_every_cell_owned_once_and_stencils_stay_in_segment(tmp_path):
    exe = tmp_path / "strip_plan_check"
    # host code only, under AddressSanitizer and UndefinedBehaviorSanitizer
    cmd = [HIPCC, "-std=c++17", "-O1", "--cuda-host-only", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
           "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize-recover=all", "-I" + os.path.join(REPO, "include"),
           "-I" + os.path.join(PKG, "csrc"), os.path.join(REPO, "tests", "strip_plan_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=300)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr
