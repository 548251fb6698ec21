"""Copies the velocity set that ships with the reference's cavity demo
(demo/cavity/constant/Xis and demo/cavity/constant/weights: the 28-point half-range
Gauss-Hermite set for C = sqrt(2 R T0), R = 208.244343891, T0 = 273) and the cavity's
DVMProperties / 0/rho numbers into tests/golden/demo_cavity.json.  Data only.

    python tests/golden/extract_demo_vset.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def read_list(path):
    text = open(path).read()
    m = re.search(r"\n(\d+)\s*\n\(", text)
    n = int(m.group(1))
    vals = re.findall(r"[-+]?\d+\.\d+e[-+]\d+", text[m.end():])
    assert len(vals) == n, (path, len(vals), n)
    return [float(v) for v in vals]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    d = os.path.join(ref, "demo", "cavity")
    xis = read_list(os.path.join(d, "constant", "Xis"))
    wts = read_list(os.path.join(d, "constant", "weights"))
    props = open(os.path.join(d, "constant", "DVMProperties")).read()
    rho0 = open(os.path.join(d, "0", "rho")).read()
    grab = lambda key, text: float(re.search(key + r"[^;]*?\s([-+]?\d+\.?\d*(?:e[-+]?\d+)?)\s*;", text).group(1))
    out = {
        "source": "demo/cavity/constant/{Xis,weights,DVMProperties}, demo/cavity/0/rho",
        "xis": xis, "weights": wts,
        "R": grab("R", props), "Tref": grab("Tref", props), "muRef": grab("muRef", props),
        "omega": grab("omega", props), "xiMax": grab("xiMax", props),
        "rho0": float(re.search(r"internalField\s+uniform\s+([-+\d.e]+);", rho0).group(1)),
        "Kn": 0.075, "T0": 273.0, "L": 1.0,
    }
    with open(os.path.join(HERE, "demo_cavity.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(xis), "nodes")


if __name__ == "__main__":
    main()
