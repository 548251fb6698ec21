"""Writes the fixture tests/golden/demo_cavity/: the input files of the reference's
demo/cavity case (60x60 lid-driven cavity, half-range Gauss-Hermite 28-point set, SI units)
as data, with the polyMesh files gzip-compressed (OpenFOAM's own compressed form, which the
case reader takes as name.gz).  The blockMeshDict, decomposeParDict and scripts are not
copied: the reader needs the polyMesh, dictionaries, velocity set and initial fields only.

    python tests/golden/make_demo_case.py [/root/reference/demo/cavity]
"""
import gzip
import os
import shutil
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/demo/cavity"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "demo_cavity")

FILES = ["constant/DVMProperties", "constant/Xis", "constant/weights", "constant/polyMesh/boundary",
         "system/controlDict", "system/fvSchemes", "system/fvSolution", "0/rho", "0/U", "0/T"]
MESH_GZ = ["constant/polyMesh/points", "constant/polyMesh/faces", "constant/polyMesh/owner",
           "constant/polyMesh/neighbour"]


def main():
    for rel in FILES:
        os.makedirs(os.path.dirname(os.path.join(DST, rel)), exist_ok=True)
        shutil.copyfile(os.path.join(SRC, rel), os.path.join(DST, rel))
    for rel in MESH_GZ:
        with open(os.path.join(SRC, rel), "rb") as fi, gzip.GzipFile(os.path.join(DST, rel + ".gz"), "wb",
                                                                      mtime=0) as fo:
            fo.write(fi.read())
    print("wrote", DST)


if __name__ == "__main__":
    main()
