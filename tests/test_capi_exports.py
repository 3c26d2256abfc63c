"""The C-ABI library loads and exports every entry point include/wenofv.h declares
(no device work: runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import REPO, PKG

HEADER = os.path.join(REPO, "include", "wenofv.h")
LIB = os.path.join(PKG, "libwenofv.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wfv_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("wfv_create", "wfv_destroy", "wfv_step", "wfv_assemble_rhs", "wfv_advance", "wfv_compute_dt"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwenofv.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
