"""Host-side arithmetic of bench.py: the HBM-traffic summary of two rocprofv3 counter passes
(per-dispatch averages scaled by the stream probe of the pass's dominant access width) and
the algorithmic bytes per launch.
Synthetic counter files only; nothing here touches a GPU."""
import csv
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _write_counters(path, counter, rows):
    with open(path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for name, kib in rows:
            w.writerow({"Kernel_Name": name, "Counter_Name": counter, "Counter_Value": kib})


PASS_B_RC = "void dugks::rc::(anonymous namespace)::k_pass_b<float, 2, 4>(dugks::Geo, FArgs, ...)"
PASS_B_FV = "void dugks::(anonymous namespace)::k_pass_b<float, 3, 4>(dugks::Geo, FArgs, ...)"
PASS_A = "void dugks::(anonymous namespace)::k_pass_a<float, 3, 4>(dugks::Geo, FArgs, ...)"


def probe(kind, w):
    return f"void (anonymous namespace)::k_probe_{kind}<{w}>(float const*, unsigned long, float*)"


def _probe_rows(probe_kib, fetch_div, write_div):
    """probe dispatches reading probe_kib / fetch_div[w] in FETCH_SIZE and probe_kib /
    write_div[w] in WRITE_SIZE (the write kernel fetches nothing, the read kernel writes
    nothing)"""
    f = [(probe("read", w), probe_kib / fetch_div[w]) for w in (4, 8, 16)]
    f += [(probe("write", w), 0.0) for w in (4, 8, 16)]
    wr = [(probe("write", w), probe_kib / write_div[w]) for w in (4, 8, 16)]
    wr += [(probe("read", w), 0.0) for w in (4, 8, 16)]
    return f, wr


def test_traffic_summary_scales_by_the_probe_of_the_dominant_width(tmp_path):
    """Per-width probe scales: the face-value passes (8-byte face-value pairs) take the
    8-byte scale, the recompute schedule's pass B (4-byte values) the 4-byte one; kernel
    figures are per-dispatch means times those scales."""
    b = _bench()
    mib = 64
    probe_kib = (mib << 20) / 1024.0
    pf, pw = _probe_rows(probe_kib, {4: 2.0, 8: 1.6, 16: 1.0}, {4: 1.0, 8: 1.0, 16: 0.5})
    fetch = tmp_path / "fetch.csv"
    write = tmp_path / "write.csv"
    _write_counters(fetch, "FETCH_SIZE", pf + [(PASS_B_FV, 1000.0), (PASS_B_FV, 1100.0), (PASS_B_FV, 1200.0),
                                               (PASS_A, 500.0)])
    _write_counters(write, "WRITE_SIZE", pw + [(PASS_B_FV, 400.0), (PASS_B_FV, 400.0), (PASS_B_FV, 400.0),
                                               (PASS_A, 10.0)])
    out = b.pmc_summary(str(fetch), str(write), "synthetic", mib)
    assert out["fetch_scale_by_width"] == {"4": 2.0, "8": 1.6, "16": 1.0}
    assert out["write_scale_by_width"] == {"4": 1.0, "8": 1.0, "16": 0.5}
    kb = out["kernels"]["fused_update"]
    assert kb["dispatches"] == 3 and kb["scale_width"] == 8
    assert kb["fetch_bytes"] == pytest.approx(1100.0 * 1024 * 1.6)
    assert kb["write_bytes"] == pytest.approx(400.0 * 1024)
    assert kb["traffic_bytes"] == pytest.approx(kb["fetch_bytes"] + kb["write_bytes"])
    ka = out["kernels"]["fused_face"]
    assert ka["scale_width"] == 8
    assert ka["traffic_bytes"] == pytest.approx(500.0 * 1024 * 1.6 + 10.0 * 1024)


def test_recompute_schedule_pass_b_takes_the_4_byte_scale(tmp_path):
    b = _bench()
    mib = 16
    probe_kib = (mib << 20) / 1024.0
    pf, pw = _probe_rows(probe_kib, {4: 2.0, 8: 1.5, 16: 1.0}, {4: 1.0, 8: 1.0, 16: 1.0})
    fetch = tmp_path / "fetch.csv"
    write = tmp_path / "write.csv"
    _write_counters(fetch, "FETCH_SIZE", pf + [(PASS_B_RC, 300.0)])
    _write_counters(write, "WRITE_SIZE", pw + [(PASS_B_RC, 100.0)])
    kb = b.pmc_summary(str(fetch), str(write), "synthetic", mib)["kernels"]["fused_update"]
    assert kb["scale_width"] == 4
    assert kb["fetch_bytes"] == pytest.approx(300.0 * 1024 * 2.0)


def test_traffic_summary_refuses_files_without_the_probe(tmp_path):
    b = _bench()
    fetch = tmp_path / "fetch.csv"
    write = tmp_path / "write.csv"
    _write_counters(fetch, "FETCH_SIZE", [(PASS_B_FV, 1000.0)])
    _write_counters(write, "WRITE_SIZE", [(PASS_B_FV, 400.0)])
    with pytest.raises(SystemExit):
        b.pmc_summary(str(fetch), str(write), "synthetic", 64)


def test_algorithmic_bytes_of_the_fused_passes():
    """Three compulsory array passes of two distributions per step: pass A reads g~, h~
    (2 s per cell-velocity), pass B reads and writes them (4 s); 24 B per update in the
    mixed path, 48 B in FP64."""
    b = _bench()
    nC, nF, Nv = 65536, 131584, 4096
    a4 = b.algorithmic_bytes("fused_face", nC, nF, Nv, 4)
    b4 = b.algorithmic_bytes("fused_update", nC, nF, Nv, 4)
    assert a4 == 2 * 4 * nC * Nv
    assert b4 == 4 * 4 * nC * Nv == 4294967296
    assert (a4 + b4) / (nC * Nv) == 24
    assert (b.algorithmic_bytes("fused_face", nC, nF, Nv, 8) + b.algorithmic_bytes("fused_update", nC, nF, Nv, 8)) \
        / (nC * Nv) == 48
    assert b.algorithmic_bytes("no_such_kernel", nC, nF, Nv, 4) is None
