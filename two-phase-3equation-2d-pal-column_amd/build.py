"""Build libwenofv.so (HIP kernels + C ABI) in place for gfx950.

    python build.py            # from any directory

Flags: -ffp-contract=off keeps the all-FP64 kernels free of fused multiply-adds so they
round exactly like the reference CPU build; packed FP32 instructions are disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwenofv.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("WENOFV_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(REPO, "include"), "-I" + CSRC, "-Wno-unused-command-line-argument"]
# translation units and their contraction policy
UNITS = [
    ("wenofv_kernels_fp64.hip", ["-ffp-contract=off"]),
    ("wenofv_kernels_mixed.hip", ["-ffp-contract=fast"]),
    ("wenofv_control.hip", ["-ffp-contract=off"]),
    ("wenofv_capi.cpp", ["-ffp-contract=off"]),
]


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src, extra in UNITS:  # compiled one after another (each unit takes ~10-30 s)
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
