"""Build libdugks_hip.so for gfx950 in-tree (psp-vst_amd/dugks_amd/libdugks_hip.so), the
streaming probe libdugks_probe.so beside it (HBM counter calibration for the benchmark), the
OpenFOAM case library libdugks_foam.so (host C++, g++) and the dugksFoam driver executable
that runs a case directory through both.

Usage: python psp-vst_amd/build.py [--jobs N] [--verbose]

Every call recompiles.  No cmake/ninja: one hipcc invocation per translation unit, then
one link.  All HIP code is built with packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "dugks_amd")
LIB = os.path.join(OUT_DIR, "libdugks_hip.so")
PROBE_LIB = os.path.join(OUT_DIR, "libdugks_probe.so")
PROBE_SOURCES = ["dugks_probe.hip"]
FOAM_DIR = os.path.join(HERE, "foam")
FOAM_LIB = os.path.join(OUT_DIR, "libdugks_foam.so")
FOAM_SOURCES = ["foam_dict.cpp", "foam_case.cpp", "foam_capi.cpp"]
DRIVER = os.path.join(OUT_DIR, "dugksFoam")
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(HERE, "..", "include")]
ROCM_LIB = os.environ.get("ROCM_PATH", "/opt/rocm") + "/lib"
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("DUGKS_OFFLOAD_ARCH", "gfx950")

SOURCES = ["dugks_setup.cpp", "dugks_fused_plan.cpp", "dugks_fused_rc_plan.cpp", "dugks_kernels.hip",
           "dugks_fused.hip", "dugks_fused_rc.hip", "dugks_adjoint.hip",
           "dugks_api.hip"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(HERE, "..", "include")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_one(src: str, verbose: bool) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + ["-c", "-x", "hip", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    return obj


def link(lib: str, objs: list, verbose: bool) -> None:
    tmp = lib + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {lib}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)


def run(cmd: list, what: str, verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed\n{r.stdout}\n{r.stderr}")


def gxx() -> str:
    return os.environ.get("CXX") or shutil.which("g++") or "g++"


def build_foam(verbose: bool = False) -> None:
    """libdugks_foam.so (OpenFOAM case I/O, host only) and the dugksFoam driver, which links
    it and libdugks_hip.so from the same directory (rpath $ORIGIN)."""
    srcs = [os.path.join(FOAM_DIR, f) for f in FOAM_SOURCES]
    tmp = FOAM_LIB + ".tmp"
    run([gxx()] + HOST_FLAGS + ["-shared", "-o", tmp] + srcs + ["-lz"], "libdugks_foam.so build", verbose)
    os.replace(tmp, FOAM_LIB)
    tmp = DRIVER + ".tmp"
    run([gxx()] + HOST_FLAGS + [os.path.join(FOAM_DIR, "dugksFoam.C"), "-o", tmp, "-L", OUT_DIR, "-ldugks_foam",
                                "-ldugks_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link," + ROCM_LIB],
        "dugksFoam build", verbose)
    os.replace(tmp, DRIVER)


def build(jobs: int = 4, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), SOURCES + PROBE_SOURCES))
    link(LIB, objs[:len(SOURCES)], verbose)
    link(PROBE_LIB, objs[len(SOURCES):], verbose)
    build_foam(verbose)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
