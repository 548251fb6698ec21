"""The C-ABI libraries load on a host without a GPU and export every symbol that their
headers declare (no compute calls): include/dugks.h -> libdugks_hip.so, include/dugks_foam.h
-> libdugks_foam.so."""
import ctypes as ct
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dugks.h")
LIB = os.path.join(ROOT, "psp-vst_amd", "dugks_amd", "libdugks_hip.so")
FOAM_HEADER = os.path.join(ROOT, "include", "dugks_foam.h")
FOAM_LIB = os.path.join(ROOT, "psp-vst_amd", "dugks_amd", "libdugks_foam.so")


def declared_functions(header=HEADER, prefix="dugks_"):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(FOAM_LIB), reason="libdugks_foam.so not built")
def test_foam_library_exports_every_declared_symbol():
    names = declared_functions(FOAM_HEADER, "dugks_foam_")
    assert "dugks_foam_read" in names and "dugks_foam_read_vmesh" in names and "dugks_foam_write" in names
    lib = ct.CDLL(FOAM_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"symbols declared in include/dugks_foam.h but not exported: {missing}"


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
