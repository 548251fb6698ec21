"""The C-ABI library loads on a host without a GPU and exports every symbol that
include/dugks.h declares (no compute calls)."""
import ctypes as ct
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dugks.h")
LIB = os.path.join(ROOT, "psp-vst_amd", "dugks_amd", "libdugks_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dugks_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("dugks_create", "dugks_evolution", "dugks_co_num", "dugks_destroy", "dugks_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libdugks_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"symbols declared in include/dugks.h but not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libdugks_hip.so not built")
def test_abi_version_and_errors_without_gpu():
    from dugks_amd import hip
    L = hip.load_library()
    assert L.dugks_abi_version() == 3
    # invalid arguments are rejected before any device call
    h = ct.c_void_p()
    assert L.dugks_create(None, None, 0, None, None, None, ct.byref(h)) != 0
    assert b"required" in L.dugks_last_error()
