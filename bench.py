"""Benchmark of the DUGKS step on one MI355X.

Workload: the 256x256 lid-driven cavity with the 64x64 trapezoid velocity set on
[-5, 5] sqrt(R T0) (65,536 cells x 4,096 velocities = 2.68e8 cell-velocity updates per
step), Kn = 0.1, synthetic nondimensional initial state (uniform rest equilibrium).  A step
is one DUGKS time step including the device-side Courant number / time-step update.

Prints one JSON line.  ``value`` is the mixed-precision path's throughput in
cell-velocity updates per second with all data resident in HBM; the FP64 path, the
relative error of the mixed path against the FP64 path, per-kernel HIP-event timings, the
HBM roofline of the dominant kernel and a CPU baseline (the FP64 oracle on the host
cores, a bounded sample) ride along.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-steps S] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "psp-vst_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel: str, nC: int, nF: int, Nv: int, s: int) -> float | None:
    """Minimal HBM bytes one launch of a kernel must move (values of storage size s)."""
    cv = float(nC) * Nv
    fv = float(nF) * Nv
    table = {
        # staged path
        "bar": 4 * cv * s,                       # read g~,h~; write g+,h+
        "face_recon": 2 * cv * s + 2 * fv * s,   # read g+,h+; write face g,h
        "face_moments": 2 * fv * s,              # read face g,h
        "face_relax": 4 * fv * s,                # read + write face g,h
        "update": 6 * cv * s + 2 * fv * s,       # read g~,h~,g+,h+, face g,h; write g~,h~
        "cell_moments": 2 * cv * s,              # read g~,h~
        # fused path
        "fused_face": 2 * cv * s,                # read g~,h~
        "fused_update": 4 * cv * s,              # read g~,h~; write g~,h~
    }
    return table.get(kernel)


def run_gpu(case, precision, steps, warmup, fused):
    import numpy as np
    from dugks_amd import hip

    solver = hip.FvDVM(case, precision=precision, fused=fused)
    solver.run(warmup)
    solver.synchronize()
    solver.profile(True)
    t0 = time.perf_counter()
    solver.run(steps)
    solver.synchronize()
    wall = time.perf_counter() - t0
    ktimes = solver.kernel_times()
    solver.profile(False)
    macros = solver.cell_macros()
    info = dict(wall=wall, ktimes=ktimes, macros=macros, fused="fused_update" in ktimes)
    del solver
    return info


def rel(a, b):
    import numpy as np
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def cpu_baseline(case, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, lib
    threads = min(16, os.cpu_count() or 1)
    lib().orc_set_threads(threads)
    o = Oracle(case, threads=threads)
    o.step(1)  # first touch of the big arrays outside the timed sample
    t0 = time.perf_counter()
    o.step(steps)
    wall = time.perf_counter() - t0
    return dict(value=case.n_cells * case.nv * steps / wall, cores=threads, wall=wall)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--kn", type=float, default=0.1)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--ndv", type=int, default=64)
    ap.add_argument("--fused", type=int, default=-1)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs the DUGKS step on one GPU; --gpus must be 1 "
                                   "(velocity-space sharding across GPUs is a design, not a build)"}))
        sys.exit(2)

    from dugks_amd import cases
    case = cases.config1(a.kn, N=a.cells, n_dv=a.ndv)
    nC, Nv = case.n_cells, case.nv
    nF = case.mesh.n_internal_faces + sum(p.size for p in case.mesh.patches if p.geometric_type != "empty")
    upd_per_step = nC * Nv

    mixed = run_gpu(case, 1, a.steps, a.warmup, a.fused)
    ms_per_step = mixed["wall"] / a.steps * 1e3
    value = upd_per_step * a.steps / mixed["wall"]

    # dominant kernel of the mixed path
    kt = mixed["ktimes"]
    dom = max(kt, key=lambda k: kt[k][0])
    dom_avg = kt[dom][0] / kt[dom][1]
    alg = algorithmic_bytes(dom, nC, nF, Nv, 4)
    roofline = None
    if alg is not None:
        ach = alg / dom_avg / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_launch_ms": round(dom_avg * 1e3, 4)}
    step_frac = value * 24.0 / (HBM_PEAK_GBS * 1e9)

    out = {
        "metric": "cell-velocity updates/s (mixed FP32-storage path) + HBM roofline %",
        "value": value,
        "unit": "cell*velocity updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage / f64 moments+equilibrium+update",
        "data": "synthetic (uniform rest equilibrium, nondimensional)",
        "config": {"workload": f"lid-driven cavity {a.cells}x{a.cells}, {a.ndv}x{a.ndv} trapezoid on [-5,5], "
                               f"Kn={a.kn}", "cells": nC, "velocities": Nv, "faces": nF,
                   "path": "fused" if mixed["fused"] else "staged"},
        "step_roofline_frac": round(step_frac, 4),
        "kernels_ms": {k: round(v[0] / v[1] * 1e3, 4) for k, v in kt.items()},
        "roofline": roofline,
    }
    if not a.no_fp64:
        fp64 = run_gpu(case, 0, a.steps, a.warmup, a.fused)
        v64 = upd_per_step * a.steps / fp64["wall"]
        mo, mf = mixed["macros"], fp64["macros"]
        out["fp64"] = {"value": v64, "ms_per_step": fp64["wall"] / a.steps * 1e3,
                       "step_roofline_frac": round(v64 * 48.0 / (HBM_PEAK_GBS * 1e9), 4),
                       "kernels_ms": {k: round(v[0] / v[1] * 1e3, 4) for k, v in fp64["ktimes"].items()}}
        out["mixed_over_fp64"] = value / v64
        out["mixed_rel_err_vs_fp64"] = {k: rel(mo[k], mf[k]) for k in ("rho", "U", "T", "q")}
        out["mixed_rel_err_note"] = f"after {a.warmup + a.steps} steps, same dt control"
    if not a.no_cpu:
        cb = cpu_baseline(case, a.cpu_steps)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "cell*velocity updates/s", "cores": cb["cores"],
                               "kind": "port",
                               "sample": f"{a.cpu_steps} FP64 steps of the same workload (after 1 untimed step), "
                                         f"OpenMP over velocities, {cb['wall']:.1f} s"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
