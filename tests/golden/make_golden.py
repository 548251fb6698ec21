"""Writes tests/golden/cavity4_nc38x7_5steps.npz: the FP64 oracle's state after 5 steps of
a 4x4 cavity with a 7-point Newton-Cotes (Simpson 3/8) set per direction, Kn = 0.1.

Before writing, the oracle's result is checked against the independent numpy restatement
(tests/numpy_ref.py) to 1e-13; the script refuses to write a fixture that disagrees.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "psp-vst_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]

from dugks_amd import cases  # noqa: E402
from dugks_amd.timecontrol import next_delta_t  # noqa: E402
from numpy_ref import NumpyDUGKS  # noqa: E402
from oracle import Oracle  # noqa: E402


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def main():
    c = cases.cavity_case(4, "NC38", 7, 4.0, 0.1)
    o = Oracle(c)
    n = NumpyDUGKS(c)
    n.init()
    dt = c.controls.deltaT
    for _ in range(5):
        co = n.co_num(dt)
        dt = next_delta_t(dt, co, c.controls)
        n.step(dt)
        o.step()
    m = o.cell_macros()
    g, h = o.df()
    errs = [rel(m["rho"], n.rho), rel(m["U"], n.U), rel(m["T"], n.T), rel(m["q"], n.q), rel(g, n.g), rel(h, n.h)]
    print("oracle vs numpy restatement:", errs)
    if max(errs) > 1e-13:
        raise SystemExit("oracle and numpy restatement disagree; fixture not written")
    np.savez_compressed(os.path.join(HERE, "cavity4_nc38x7_5steps.npz"), rho=m["rho"], U=m["U"], T=m["T"], q=m["q"],
                        g=g, h=h, dt=o.dt)


if __name__ == "__main__":
    main()
