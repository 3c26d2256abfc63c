"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: tests/oracle_sanitize_check.c
drives every scheme, reconstruction form, flux, limiter setting and boundary kind of the C
restatement on small and ragged grids (right-hand side with face and theta output, then
RK3 steps); any out-of-bounds access, use of freed memory or undefined operation ends it
with a report and a nonzero status."""
import os
import shutil
import subprocess

import pytest

from conftest import REPO


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc is not installed")
def test_oracle_is_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_sanitize_check"
    cmd = ["gcc", "-std=gnu99", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-Wno-unknown-pragmas",
           "-I" + os.path.join(REPO, "oracle"), os.path.join(REPO, "tests", "oracle_sanitize_check.c"),
           os.path.join(REPO, "oracle", "wenofv_oracle.c"), "-lm", "-o", str(exe)]
    built = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if built.returncode != 0 and "sanitizer" in built.stderr.lower():
        pytest.skip("this gcc has no sanitizer runtime")
    assert built.returncode == 0, built.stderr
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0 and res.stdout.startswith("ok"), res.stdout + res.stderr[-3000:]
