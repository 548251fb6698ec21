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

HBM traffic of the dominant kernel (``roofline.traffic``) comes from two rocprofv3 counter
passes of this script with ``--copy-probe-mib M`` (FETCH_SIZE, then WRITE_SIZE); the probe
is a plain device-to-device copy of M MiB whose counter readings calibrate the counters
(bytes moved / bytes counted).  ``--pmc-summary`` turns the two passes into
profiles/hbm_traffic_<workload>.json, which later runs read:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- python3 bench.py \
        --steps 2 --warmup 1 --no-cpu --no-fp64 --copy-probe-mib 1024
    rocprofv3 --pmc WRITE_SIZE ... (same)
    python bench.py --pmc-summary OUT/fetch/run_counter_collection.csv OUT/write/run_counter_collection.csv
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
KERNEL_SYMBOLS = {"fused_update": "k_pass_b", "fused_face": "k_pass_a", "face_recon": "face_recon",
                  "update": "k_update", "bar": "k_bar", "cell_moments": "k_cell_moments"}
COPY_SYMBOL = "__amd_rocclr_copyBuffer"


def workload_tag(cells: int, ndv: int, kn: float) -> str:
    return f"cavity{cells}_v{ndv}_kn{kn:g}"


def traffic_file(tag: str) -> str:
    return os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}.json")


def copy_probe(mib: int) -> None:
    """A plain device-to-device copy of mib MiB (counter calibration for the traffic figure),
    issued through the HIP runtime libdugks_hip.so is linked against (no second runtime)."""
    import ctypes as ct
    from dugks_amd import hip
    hip.load_library()
    rt = ct.CDLL("libamdhip64.so.7")
    n = ct.c_size_t(mib << 20)
    src, dst = ct.c_void_p(), ct.c_void_p()
    if rt.hipMalloc(ct.byref(src), n) or rt.hipMalloc(ct.byref(dst), n):
        raise RuntimeError("copy probe: hipMalloc failed")
    try:
        if rt.hipMemset(src, 1, n) or rt.hipDeviceSynchronize():
            raise RuntimeError("copy probe: hipMemset failed")
        if rt.hipMemcpy(dst, src, n, 3) or rt.hipDeviceSynchronize():  # 3 = device to device
            raise RuntimeError("copy probe: hipMemcpy failed")
    finally:
        rt.hipFree(src)
        rt.hipFree(dst)


def pmc_summary(fetch_csv: str, write_csv: str, tag: str, mib: int) -> dict:
    """Average FETCH_SIZE / WRITE_SIZE (KiB per dispatch) per kernel, scaled by the copy probe."""
    import csv
    import collections

    def read(path, counter):
        vals = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                vals[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024.0)
        return vals

    f, w = read(fetch_csv, "FETCH_SIZE"), read(write_csv, "WRITE_SIZE")
    probe = float(mib << 20)
    # the probe's copy dispatches (the copy may be split); the step itself copies only small arrays
    fcopy = sum(v for k, vs in f.items() if COPY_SYMBOL in k for v in vs if v > probe / 64)
    wcopy = sum(v for k, vs in w.items() if COPY_SYMBOL in k for v in vs if v > probe / 64)
    # FETCH_SIZE reads about half of a device-to-device copy's bytes on this hardware, so the
    # accepted window around the probe size is wide (a factor of 4 either way)
    if not (0.25 * probe < fcopy < 4 * probe and 0.25 * probe < wcopy < 4 * probe):
        raise SystemExit(f"copy probe dispatches not found in the counter files ({fcopy:.3e}, {wcopy:.3e} B)")
    kf, kw = probe / fcopy, probe / wcopy
    out = {"workload": tag, "copy_probe_bytes": probe, "fetch_scale": kf, "write_scale": kw, "kernels": {}}
    for short, sym in KERNEL_SYMBOLS.items():
        names = [k for k in f if sym in k and k in w]
        for k in names:
            fb = sum(f[k]) / len(f[k]) * kf
            wb = sum(w[k]) / len(w[k]) * kw
            out["kernels"][short] = {"symbol": k, "fetch_bytes": fb, "write_bytes": wb,
                                     "traffic_bytes": fb + wb, "dispatches": len(f[k])}
    return out


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
    pinfo = solver.path_info()
    info = dict(wall=wall, ktimes=ktimes, macros=macros, fused=bool(pinfo["fused"]), path_info=pinfo)
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


def rel2(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def compare_precisions(case, steps, every, fused=-1, label=None):
    """Mixed (FP32 storage) and FP64 runs of the same case on the GPU, compared every `every`
    steps up to `steps` (max-norm relative error max|a-b|/max|b| and 2-norm relative error).
    The two runs go one after the other, each keeping host copies of its cell macros at the
    comparison steps, so only one solver's distributions are resident at a time; a run whose
    fields stop being finite ends the comparison (the step is reported)."""
    import numpy as np
    from dugks_amd import hip

    def run(precision):
        s = hip.FvDVM(case, precision=precision, fused=fused)
        snaps, done, t = [], 0, 0.0
        while done < steps:
            n = min(every, steps - done)
            t0 = time.perf_counter()
            s.run(n)
            s.synchronize()
            t += time.perf_counter() - t0
            done += n
            m = s.cell_macros()
            snaps.append((done, m))
            print(json.dumps({"progress": label, "precision": "mixed" if precision == hip.MIXED else "fp64",
                              "step": done}), flush=True)
            if not np.all(np.isfinite(m["rho"])):
                break
        s.close()
        return snaps, t

    mixed, tm = run(hip.MIXED)
    fp64, tf = run(hip.FP64)
    rows, diverged = [], None
    for (done, mo), (_, mf) in zip(mixed, fp64):
        finite = {k: bool(np.all(np.isfinite(v))) for k, v in (("mixed", mo["rho"]), ("fp64", mf["rho"]))}
        if not all(finite.values()):
            diverged = {"step": done, "finite": finite}
            break
        row = {"step": done, **{k: rel(mo[k], mf[k]) for k in ("rho", "U", "T", "q")},
               "l2": {k: rel2(mo[k], mf[k]) for k in ("rho", "U", "T", "q")},
               "rho_range": [float(mf["rho"].min()), float(mf["rho"].max())],
               "T_range": [float(mf["T"].min()), float(mf["T"].max())]}
        rows.append(row)
        print(json.dumps({"progress": label, **row}), flush=True)
    if diverged is None and len(mixed) != len(fp64):
        short = mixed if len(mixed) < len(fp64) else fp64
        diverged = {"step": short[-1][0], "finite": {"mixed": len(mixed) >= len(fp64), "fp64": len(fp64) >= len(mixed)}}
    upd = case.n_cells * case.nv
    nm, nf = mixed[-1][0], fp64[-1][0]
    return {"mixed_rel_err_vs_fp64": rows, "diverged": diverged, "mixed_value": upd * nm / tm if tm else None,
            "fp64_value": upd * nf / tf if tf else None,
            "mixed_over_fp64": (tf / nf) / (tm / nm) if tm and tf else None}


def kn_sweep(a):
    """Mixed against FP64 on the cavity for each Kn of a.kn_sweep (a.steps steps)."""
    from dugks_amd import cases
    for kn in [float(x) for x in a.kn_sweep.split(",")]:
        case = cases.config1(kn, N=a.cells, n_dv=a.ndv)
        r = compare_precisions(case, a.steps, a.sweep_every, a.fused, label=f"kn={kn}")
        print(json.dumps({"kn": kn, "steps": a.steps, "max_co": case.controls.maxCo,
                          "lid_speed": case.meta.get("lid_speed"), "lid_jitter": case.meta.get("lid_jitter"),
                          **r}), flush=True)


def adjoint_check(a):
    """Adjoint gradient against central differences of the FP64 GPU forward (fixed time
    steps) on the adjoint cavity (100x100, 32x32 trapezoid on [-4,4], Kn 0.1, 500 steps,
    checkpoints every 50) or the adjoint sphere (~51k cells, 16^3 Gauss-Hermite, Kn 0.1,
    Ma 0.2, 300 steps, checkpoints every 30, mixed precision, 64 wall patches)."""
    import numpy as np
    from dugks_amd import adjoint, hip
    eps = 1e-6
    if a.adjoint == "cavity":
        prob = adjoint.CavityHeatFlux(steps=a.adjoint_steps or 500, checkpoint_every=50, precision=hip.FP64)
        th = prob.nominal()
        fd_idx = [0, 1, 2 + 10, 2 + 50, 2 + 90]
        names = ["lid_speed", "Kn", "T_w[10]", "T_w[50]", "T_w[90]"]
    else:
        prob = adjoint.SphereDrag(steps=a.adjoint_steps or 300, checkpoint_every=30, precision=hip.MIXED)
        th = prob.nominal()
        th[1:] = 1.0 + 0.05 * np.sin(np.arange(prob.n_patch))
        fd_idx = [0, 17]
        names = ["Kn", "T_patch[16]"]
    t0 = time.perf_counter()
    prob.time_steps(th)
    t1 = time.perf_counter()
    res = prob.evaluate(th)
    t2 = time.perf_counter()
    out = {"problem": a.adjoint, "steps": prob.steps, "checkpoint_every": prob.ckpt, "J": res.J,
           "n_params": len(th), "adjoint_s": t2 - t1, "forward_dt_run_s": t1 - t0, "checks": []}
    print(json.dumps({"progress": "adjoint done", "J": res.J, "s": t2 - t1}), flush=True)
    for i, nm in zip(fd_idx, names):
        h = eps * max(abs(th[i]), 1.0) if i > 1 or a.adjoint == "sphere" else eps * th[i]
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (prob.value(tp, precision=hip.FP64) - prob.value(tm, precision=hip.FP64)) / (2 * h)
        out["checks"].append({"param": nm, "adjoint": float(res.grad[i]), "fd_fp64": fd,
                              "rel_err": abs(res.grad[i] - fd) / max(abs(fd), 1e-300)})
        print(json.dumps({"progress": nm, **out["checks"][-1]}), flush=True)
    if a.adjoint == "sphere":
        p64 = adjoint.SphereDrag(steps=prob.steps, checkpoint_every=30, precision=hip.FP64, dts=prob.dts)
        r64 = p64.evaluate(th)
        out["fp64_adjoint_J"] = r64.J
        out["mixed_vs_fp64_grad_rel_err"] = float(np.max(np.abs(res.grad - r64.grad)) / np.max(np.abs(r64.grad)))
        out["mixed_vs_fp64_J_rel_err"] = abs(res.J - r64.J) / abs(r64.J)
    print(json.dumps(out), flush=True)


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
    ap.add_argument("--copy-probe-mib", type=int, default=0)
    ap.add_argument("--workload", choices=("cavity", "sphere"), default="cavity")
    ap.add_argument("--kn-sweep", type=str, default="",
                    help="comma-separated Kn values: mixed vs FP64 error table on the cavity, one JSON line each")
    ap.add_argument("--sweep-every", type=int, default=100)
    ap.add_argument("--compare", action="store_true",
                    help="mixed against FP64 side by side for --steps steps (error every --sweep-every)")
    ap.add_argument("--adjoint", choices=("cavity", "sphere"), default=None,
                    help="adjoint gradient check against central differences instead of the benchmark")
    ap.add_argument("--adjoint-steps", type=int, default=0)
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    a = ap.parse_args()
    tag = workload_tag(a.cells, a.ndv, a.kn)
    if a.pmc_summary:
        summ = pmc_summary(a.pmc_summary[0], a.pmc_summary[1], tag, a.copy_probe_mib or 1024)
        os.makedirs(os.path.dirname(traffic_file(tag)), exist_ok=True)
        with open(traffic_file(tag), "w") as fh:
            json.dump(summ, fh, indent=1)
        print(json.dumps(summ))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs the DUGKS step on one GPU; --gpus must be 1 "
                                   "(velocity-space sharding across GPUs is a design, not a build)"}))
        sys.exit(2)

    from dugks_amd import cases
    if a.workload == "sphere":
        case = cases.config2(n_dv=a.ndv if a.ndv != 64 else 24)
        tag = f"sphere{case.n_cells}_v{case.nv}"
        wl = (f"flow past a sphere, {case.n_cells} hex cells, {round(case.nv ** (1 / 3))}^3 half-range "
              f"Gauss-Hermite velocities, Kn=0.1, Ma=0.2")
    else:
        case = cases.config1(a.kn, N=a.cells, n_dv=a.ndv)
        wl = f"lid-driven cavity {a.cells}x{a.cells}, {a.ndv}x{a.ndv} trapezoid on [-5,5], Kn={a.kn}"
    nC, Nv = case.n_cells, case.nv
    nF = case.mesh.n_internal_faces + sum(p.size for p in case.mesh.patches if p.geometric_type != "empty")
    upd_per_step = nC * Nv
    if a.kn_sweep:
        kn_sweep(a)
        return
    if a.adjoint:
        adjoint_check(a)
        return
    if a.compare:
        r = compare_precisions(case, a.steps, a.sweep_every, a.fused, label=wl)
        print(json.dumps({"workload": wl, "steps": a.steps, "max_co": case.controls.maxCo, **r}), flush=True)
        return

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
                    "algorithmic_bytes": alg, "avg_launch_ms": round(dom_avg * 1e3, 4)}
        tf = traffic_file(tag)
        if os.path.exists(tf):
            rec = json.load(open(tf)).get("kernels", {}).get(dom)
            if rec:
                roofline["traffic"] = round(rec["traffic_bytes"])
                roofline["traffic_over_algorithmic"] = round(rec["traffic_bytes"] / alg, 3)
                roofline["traffic_source"] = os.path.relpath(tf, ROOT)
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
        "config": {"workload": wl, "cells": nC, "velocities": Nv, "faces": nF,
                   "path": {1: "fused, stored face values", 2: "fused, recomputed halo rings"}.get(
                       mixed["path_info"]["fused"], "staged"),
                   "block_cells": mixed["path_info"]["block_cells"],
                   "chunk_velocities": mixed["path_info"]["chunk_velocities"],
                   "velocity_groups": mixed["path_info"]["velocity_groups"]},
        "step_roofline_frac": round(step_frac, 4),
        "kernels_ms": {k: round(v[0] / v[1] * 1e3, 4) for k, v in kt.items()},
        "kernels_gbs": {k: round(algorithmic_bytes(k, nC, nF, Nv, 4) / (v[0] / v[1]) / 1e9, 1)
                        for k, v in kt.items() if algorithmic_bytes(k, nC, nF, Nv, 4)},
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
    if not a.no_cpu and a.workload == "cavity":
        cb = cpu_baseline(case, a.cpu_steps)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "cell*velocity updates/s", "cores": cb["cores"],
                               "kind": "port",
                               "sample": f"{a.cpu_steps} FP64 steps of the same workload (after 1 untimed step), "
                                         f"OpenMP over velocities, {cb['wall']:.1f} s"}
    print(json.dumps(out), flush=True)
    if a.copy_probe_mib:
        copy_probe(a.copy_probe_mib)


if __name__ == "__main__":
    main()
