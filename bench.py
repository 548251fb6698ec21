"""Benchmark of the DUGKS step on one MI355X.

Workload (default): flow past a sphere, 211,520 hex cells, 24^3 half-range Gauss-Hermite
velocities (13,824), Kn = 0.1, Ma = 0.2, diffuse isothermal sphere wall, freestream
inlet / outlet / side patches; 2.92e9 cell-velocity updates per step.  A step is one DUGKS
time step including the device-side Courant number / time-step update.  ``--workload
cavity`` runs the 256x256 lid-driven cavity (64x64 trapezoid on [-5, 5], Kn 0.1) instead;
the default run also times the cavity and reports it under ``cavity``.

Prints one JSON line.  ``value`` is the mixed-precision path's throughput in
cell-velocity updates per second with all data resident in HBM; the FP64 path, the
relative error of the mixed path against the FP64 path, per-kernel HIP-event timings, the
HBM roofline of the dominant kernel, the measured streaming bandwidth and a CPU baseline
(the FP64 oracle on the host cores over a bounded sample) ride along.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sphere|cavity]
                    [--cpu-steps S] [--no-cpu] [--no-fp64] [--no-cavity]

HBM traffic of the dominant kernel (``roofline.traffic``) comes from two rocprofv3 counter
passes of this script with ``--probe-mib M`` (FETCH_SIZE, then WRITE_SIZE).  The probe runs
streaming read and write kernels over M MiB with 4-, 8- and 16-byte accesses per lane; the
counter readings of each give a counter scale per access width (bytes moved / bytes
counted), and each pass is scaled by the scale of the width that carries most of its bytes.
``--pmc-summary`` turns the two passes into profiles/hbm_traffic_<workload>.json, which
later runs read:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- python3 bench.py \
        --steps 1 --warmup 1 --no-cpu --no-fp64 --no-cavity --probe-mib 1024
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
KERNEL_SYMBOLS = {"fused_update": "k_pass_b", "fused_face": "k_pass_a", "fused_bar": "k_bar3",
                  "face_recon": "face_recon", "update": "k_update", "bar": "k_bar<",
                  "cell_moments": "k_cell_moments"}
PROBE_WIDTHS = (4, 8, 16)
# access width (bytes per lane) that carries most of each pass's HBM bytes: the face-value
# pass B reads (g, h) FP32 face-value pairs (8 B) for ~3.7 face slots per cell-velocity of
# the sphere, and so does pass A write them; fused_bar reads FP32 g~ and h~ one value per
# lane (4 B), as do the recompute schedule's passes
DOMINANT_WIDTH = {"fused_update": 8, "fused_face": 8, "fused_bar": 4}


def workload_tag(cells: int, ndv: int, kn: float) -> str:
    return f"cavity{cells}_v{ndv}_kn{kn:g}"


def traffic_file(tag: str) -> str:
    return os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}.json")


def probe_lib():
    import ctypes as ct
    from dugks_amd import hip
    hip.load_library()  # the probe shares the HIP runtime libdugks_hip.so is linked against
    path = os.path.join(ROOT, "psp-vst_amd", "dugks_amd", "libdugks_probe.so")
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python psp-vst_amd/build.py`")
    L = ct.CDLL(path)
    L.dugks_probe_stream.argtypes = [ct.c_size_t, ct.c_int, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
    return L


def stream_probe(mib: int) -> dict:
    """Streaming read and write of mib MiB with 4-, 8- and 16-byte accesses per lane: the
    counter calibration of the traffic figure and the measured HBM bandwidth."""
    import ctypes as ct
    L = probe_lib()
    out = {"bytes": mib << 20}
    for w in PROBE_WIDTHS:
        r, wr = ct.c_double(), ct.c_double()
        if L.dugks_probe_stream(ct.c_size_t(mib << 20), w, ct.byref(r), ct.byref(wr)) != 0:
            raise RuntimeError("stream probe failed")
        out[f"read{w}_gbs"] = round((mib << 20) / (r.value * 1e-3) / 1e9, 1)
        out[f"write{w}_gbs"] = round((mib << 20) / (wr.value * 1e-3) / 1e9, 1)
    return out


def pmc_summary(fetch_csv: str, write_csv: str, tag: str, mib: int) -> dict:
    """Average FETCH_SIZE / WRITE_SIZE (KiB per dispatch) per kernel, each scaled by the
    stream probe's counter scale for the access width that carries most of its bytes."""
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

    def probe_avg(vals, sym):
        xs = [v for k, vs in vals.items() if sym in k for v in vs]
        if not xs:
            raise SystemExit(f"probe kernel {sym} not found in the counter files")
        return sum(xs) / len(xs)

    fscale = {wd: probe / probe_avg(f, f"k_probe_read<{wd}>") for wd in PROBE_WIDTHS}
    wscale = {wd: probe / probe_avg(w, f"k_probe_write<{wd}>") for wd in PROBE_WIDTHS}
    out = {"workload": tag, "probe_bytes": probe,
           "fetch_scale_by_width": {str(k): round(v, 4) for k, v in fscale.items()},
           "write_scale_by_width": {str(k): round(v, 4) for k, v in wscale.items()},
           "kernels": {}}
    for short, sym in KERNEL_SYMBOLS.items():
        names = [k for k in f if sym in k and k in w and "probe" not in k]
        if not names:
            continue
        k = max(names, key=lambda n: sum(f[n]))
        # the recompute schedule's passes ("rc::") read g~, h~ only: 4-byte values
        wd = 4 if "rc::" in k else DOMINANT_WIDTH.get(short, 4)
        fraw = sum(f[k]) / len(f[k])
        wraw = sum(w[k]) / len(w[k])
        fb, wb = fraw * fscale[wd], wraw * wscale[wd]
        out["kernels"][short] = {"symbol": k, "fetch_bytes": fb, "write_bytes": wb,
                                 "traffic_bytes": fb + wb, "dispatches": len(f[k]), "scale_width": wd,
                                 "fetch_scale": fscale[wd], "write_scale": wscale[wd],
                                 "fetch_counted_bytes": fraw, "write_counted_bytes": wraw}
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
        # fused path (3-D FP32 storage: fused_bar forms f-bar-plus and u of every cell, pass A
        # reads f-bar-plus instead of g~,h~ and pass B reads u instead of g~,h~)
        "fused_bar": 6 * cv * s,                 # read g~,h~; write f-bar-plus and u pairs
        "fused_face": 2 * cv * s,                # read g~,h~ (or f-bar-plus pairs)
        "fused_update": 4 * cv * s,              # read g~,h~ (or u pairs); write g~,h~
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


def host_cores() -> int:
    """Threads for the CPU baseline: the process's CPU share (OMP_NUM_THREADS when the
    environment sets it, as the GPU hosts do, else the CPUs this process may run on)."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, steps):
    """The FP64 oracle (OpenMP over velocities, as the reference's omp loops over dvi) on
    this host: one untimed step, then `steps` timed steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, lib
    threads = host_cores()
    lib().orc_set_threads(threads)
    o = Oracle(case, threads=threads)
    o.step(1)  # first touch of the big arrays outside the timed sample
    t0 = time.perf_counter()
    o.step(steps)
    wall = time.perf_counter() - t0
    del o
    return dict(value=case.n_cells * case.nv * steps / wall, cores=threads, wall=wall,
                nproc=os.cpu_count(), cpu_model=cpu_model())


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


def measure(case, a, with_fp64=True):
    """Mixed (and FP64) runs of one case: throughput, per-kernel times, dominant-kernel
    roofline (algorithmic bytes over the average launch) and the mixed-vs-FP64 error."""
    nC, Nv = case.n_cells, case.nv
    nF = case.mesh.n_internal_faces + sum(p.size for p in case.mesh.patches if p.geometric_type != "empty")
    upd_per_step = nC * Nv
    mixed = run_gpu(case, 1, a.steps, a.warmup, a.fused)
    value = upd_per_step * a.steps / mixed["wall"]
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
    r = {"value": value, "ms_per_step": mixed["wall"] / a.steps * 1e3, "nC": nC, "Nv": Nv, "nF": nF,
         "path_info": mixed["path_info"], "roofline": roofline,
         "step_roofline_frac": round(value * 24.0 / (HBM_PEAK_GBS * 1e9), 4),
         "kernels_ms": {k: round(v[0] / v[1] * 1e3, 4) for k, v in kt.items()},
         "kernels_gbs": {k: round(algorithmic_bytes(k, nC, nF, Nv, 4) / (v[0] / v[1]) / 1e9, 1)
                         for k, v in kt.items() if algorithmic_bytes(k, nC, nF, Nv, 4)}}
    if with_fp64:
        fp64 = run_gpu(case, 0, a.steps, a.warmup, a.fused)
        v64 = upd_per_step * a.steps / fp64["wall"]
        mo, mf = mixed["macros"], fp64["macros"]
        r["fp64"] = {"value": v64, "ms_per_step": fp64["wall"] / a.steps * 1e3,
                     "step_roofline_frac": round(v64 * 48.0 / (HBM_PEAK_GBS * 1e9), 4),
                     "kernels_ms": {k: round(v[0] / v[1] * 1e3, 4) for k, v in fp64["ktimes"].items()}}
        r["mixed_over_fp64"] = value / v64
        r["mixed_rel_err_vs_fp64"] = {k: rel(mo[k], mf[k]) for k in ("rho", "U", "T", "q")}
        r["mixed_rel_err_note"] = f"max-norm relative, after {a.warmup + a.steps} steps, same dt control"
    return r


def path_name(pinfo):
    return {1: "fused, stored face values", 2: "fused, recomputed halo rings"}.get(pinfo["fused"], "staged")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--kn", type=float, default=0.1)
    ap.add_argument("--cells", type=int, default=256)
    ap.add_argument("--ndv", type=int, default=0, help="velocities per dimension (default 24 sphere, 64 cavity)")
    ap.add_argument("--fused", type=int, default=-1)
    ap.add_argument("--cpu-steps", type=int, default=0, help="timed oracle steps (default 8 sphere, 2 cavity)")
    ap.add_argument("--cpu-ndv", type=int, default=6,
                    help="velocities per dimension of the sphere CPU sample (same mesh; host memory)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cavity", action="store_true", help="skip the extra cavity line of the sphere run")
    ap.add_argument("--probe-mib", type=int, default=0,
                    help="run the stream probe over this many MiB at the end (counter calibration)")
    ap.add_argument("--workload", choices=("cavity", "sphere"), default="sphere")
    ap.add_argument("--kn-sweep", type=str, default="",
                    help="comma-separated Kn values: mixed vs FP64 error table on the cavity, one JSON line each")
    ap.add_argument("--sweep-every", type=int, default=100)
    ap.add_argument("--compare", action="store_true",
                    help="mixed against FP64 side by side for --steps steps (error every --sweep-every)")
    ap.add_argument("--adjoint", choices=("cavity", "sphere"), default=None,
                    help="adjoint gradient check against central differences instead of the benchmark")
    ap.add_argument("--adjoint-steps", type=int, default=0)
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--tag", type=str, default="", help="workload tag of --pmc-summary (default: the workload's)")
    a = ap.parse_args()
    from dugks_amd import cases
    if a.workload == "sphere":
        ndv = a.ndv or 24
        tag = f"sphere211520_v{ndv ** 3}"
    else:
        ndv = a.ndv or 64
        tag = workload_tag(a.cells, ndv, a.kn)
    if a.pmc_summary:
        summ = pmc_summary(a.pmc_summary[0], a.pmc_summary[1], a.tag or tag, a.probe_mib or 1024)
        os.makedirs(os.path.dirname(traffic_file(a.tag or tag)), exist_ok=True)
        with open(traffic_file(a.tag or tag), "w") as fh:
            json.dump(summ, fh, indent=1)
        print(json.dumps(summ))
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs the DUGKS step on one GPU; --gpus must be 1 "
                                   "(velocity-space sharding across GPUs is a design, not a build)"}))
        sys.exit(2)
    if a.kn_sweep:
        a.ndv = a.ndv or 64
        kn_sweep(a)
        return
    if a.adjoint:
        adjoint_check(a)
        return

    if a.workload == "sphere":
        case = cases.config2(n_dv=ndv)
        tag = f"sphere{case.n_cells}_v{case.nv}"
        wl = (f"flow past a sphere, {case.n_cells} hex cells, {ndv}^3 half-range Gauss-Hermite velocities, "
              f"Kn=0.1, Ma=0.2")
    else:
        case = cases.config1(a.kn, N=a.cells, n_dv=ndv)
        wl = f"lid-driven cavity {a.cells}x{a.cells}, {ndv}x{ndv} trapezoid on [-5,5], Kn={a.kn}"
    if a.compare:
        r = compare_precisions(case, a.steps, a.sweep_every, a.fused, label=wl)
        print(json.dumps({"workload": wl, "steps": a.steps, "max_co": case.controls.maxCo, **r}), flush=True)
        return

    m = measure(case, a, with_fp64=not a.no_fp64)
    roofline = m["roofline"]
    if roofline is not None:
        tf = traffic_file(tag)
        if os.path.exists(tf):
            rec = json.load(open(tf)).get("kernels", {}).get(roofline["kernel"])
            if rec:
                roofline["traffic"] = round(rec["traffic_bytes"])
                roofline["traffic_over_algorithmic"] = round(rec["traffic_bytes"] / roofline["algorithmic_bytes"], 3)
                roofline["traffic_source"] = os.path.relpath(tf, ROOT)
    pinfo = m["path_info"]
    out = {
        "metric": "cell-velocity updates/s (mixed FP32-storage path) + HBM roofline %",
        "value": m["value"],
        "unit": "cell*velocity updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 storage / f64 moments+equilibrium+update",
        "data": ("synthetic (uniform freestream equilibrium, nondimensional)" if a.workload == "sphere"
                 else "synthetic (uniform rest equilibrium, nondimensional)"),
        "config": {"workload": wl, "cells": m["nC"], "velocities": m["Nv"], "faces": m["nF"],
                   "path": path_name(pinfo), "block_cells": pinfo["block_cells"],
                   "chunk_velocities": pinfo["chunk_velocities"], "velocity_groups": pinfo["velocity_groups"]},
        "step_roofline_frac": m["step_roofline_frac"],
        "kernels_ms": m["kernels_ms"],
        "kernels_gbs": m["kernels_gbs"],
        "roofline": roofline,
    }
    for k in ("fp64", "mixed_over_fp64", "mixed_rel_err_vs_fp64", "mixed_rel_err_note"):
        if k in m:
            out[k] = m[k]
    out["hbm_stream_measured"] = stream_probe(1024)
    if a.workload == "sphere" and not a.no_cavity:
        cav_case = cases.config1(0.1, N=256, n_dv=64)
        c = measure(cav_case, a, with_fp64=not a.no_fp64)
        out["cavity"] = {"workload": "lid-driven cavity 256x256, 64x64 trapezoid on [-5,5], Kn=0.1",
                         "value": c["value"], "ms_per_step": c["ms_per_step"], "path": path_name(c["path_info"]),
                         "step_roofline_frac": c["step_roofline_frac"], "kernels_ms": c["kernels_ms"],
                         "roofline_frac": c["roofline"]["frac"] if c["roofline"] else None}
        for k in ("mixed_over_fp64", "mixed_rel_err_vs_fp64"):
            if k in c:
                out["cavity"][k] = c[k]
        if "fp64" in c:
            out["cavity"]["fp64_value"] = c["fp64"]["value"]
    if not a.no_cpu:
        if a.workload == "sphere":
            cpu_case = cases.config2(n_dv=a.cpu_ndv)
            steps = a.cpu_steps or 8
            sample = (f"{steps} FP64 steps (after 1 untimed step) of the same sphere mesh with a "
                      f"{a.cpu_ndv}^3 half-range Gauss-Hermite set ({cpu_case.nv} velocities; the full 24^3 "
                      f"set needs ~240 GB of host memory in FP64), OpenMP over velocities")
        else:
            cpu_case, steps = case, a.cpu_steps or 2
            sample = f"{steps} FP64 steps of the same workload (after 1 untimed step), OpenMP over velocities"
        cb = cpu_baseline(cpu_case, steps)
        out["cpu_baseline"] = {"value": cb["value"], "unit": "cell*velocity updates/s", "cores": cb["cores"],
                               "kind": "port", "sample": f"{sample}, {cb['wall']:.1f} s",
                               "nproc": cb["nproc"], "cpu_model": cb["cpu_model"]}
    print(json.dumps(out), flush=True)
    if a.probe_mib:
        stream_probe(a.probe_mib)


if __name__ == "__main__":
    main()
