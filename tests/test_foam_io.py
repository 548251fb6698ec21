"""OpenFOAM case I/O (libdugks_foam.so through dugks_amd.foam) and the C++ facade and
driver, on the CPU.

* The reference's demo/cavity case (tests/golden/demo_cavity, made by
  tests/golden/make_demo_case.py; polyMesh gzip-compressed) parses into the mesh, gas,
  velocity set, controls and fields it describes: a 60x60 one-cell-thick cavity of 1 x 1 x
  0.1 m, 28-point half-range Gauss-Hermite set, R = 208.244343891, ...
* The C++ geometry equals the independent Python restatement (dugks_amd.mesh.geometry) on
  the same polyMesh, and closed forms on the uniform mesh.
* Generated cases survive write_case -> read_case (2-D cavity, 3-D sphere, compressed or
  plain mesh files); written fields read back.
* Error paths: missing files, unsupported boundary types, binary format, directives,
  inconsistent lists; the facade header compiles and the driver refuses bad options and
  reports a missing GPU without crashing.
"""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

DEMO = os.path.join(ROOT, "tests", "golden", "demo_cavity")


def _foam():
    from dugks_amd import foam
    return foam


def test_demo_cavity_parses_into_the_case_it_describes():
    c = _foam().read_case(DEMO)
    m = c.mesh
    assert c.n_cells == 3600 and m.n_internal_faces == 7080 and m.n_faces == 14520
    assert c.D == 2
    names = [(p.name, p.geometric_type, p.start, p.size) for p in m.patches]
    assert names == [("movingWall", "wall", 7080, 60), ("fixedWalls", "wall", 7140, 180),
                     ("frontAndBack", "empty", 7320, 7200)]
    assert [(b.name, b.rho_type) for b in c.bcs] == [("movingWall", "calculatedMaxwell"),
                                                     ("fixedWalls", "calculatedMaxwell")]
    # geometry: 1 x 1 x 0.1 box of 60 x 60 equal cells
    V = c.geo.V
    assert abs(V.sum() - 0.1) < 1e-15 and np.max(np.abs(V - 0.1 / 3600)) < 1e-18
    h = 1.0 / 60
    C = c.geo.C
    ij = np.round((C[:, :2] - h / 2) / h)
    assert np.max(np.abs(C[:, :2] - (ij * h + h / 2))) < 1e-14 and np.max(np.abs(C[:, 2] - 0.05)) < 1e-15
    assert sorted(set(map(tuple, ij.astype(int).tolist()))) == [(i, j) for i in range(60) for j in range(60)]
    magSf = np.linalg.norm(c.geo.Sf[:7080], axis=1)
    assert np.max(np.abs(magSf - h * 0.1)) < 1e-16
    assert np.max(np.abs(c.geo.weights - 0.5)) < 1e-12 and np.max(np.abs(c.geo.delta_coeffs - 60.0)) < 1e-9
    # DVMProperties and the velocity set
    g = c.gas
    assert (g.R, g.omega, g.Tref, g.muRef, g.Pr, g.KInner) == (208.244343891, 0.81, 273.0, 1.60281882485e-05,
                                                               0.666666666666667, 0)
    assert c.xi_max == 1766.485662735688 and c.xi_min == -1766.485662735688
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "demo_cavity.json")))
    assert np.array_equal(c.xis, np.array(gold["xis"])) and np.array_equal(c.wts, np.array(gold["weights"]))
    # controlDict and initial fields
    k = c.controls
    assert (k.deltaT, k.adjustTimeStep, k.maxCo, k.maxDeltaT) == (4.0e-4, True, 0.8, 1.0)
    assert c.meta["end_time"] == 200.0 and c.meta["time"] == 0.0
    assert np.all(c.rho0 == 1.14423514927e-06) and np.all(c.T0 == 273.0) and np.all(c.U0 == 0.0)
    lid, walls = c.bcs
    assert np.all(lid.U == [50.0, 0.0, 0.0]) and np.all(walls.U == 0.0)
    assert np.all(lid.rho == 8.5817636195e-07) and np.all(lid.T == 273.0)


def test_cpp_geometry_equals_python_restatement():
    from dugks_amd import cases, mesh as meshmod
    for c in (_foam().read_case(DEMO),
              cases.sphere_case(dict(nc=4, nr=2, n_up=2, n_down=3, n_side=2), "GH", 6, 1.0, 0.1)):
        gp = meshmod.geometry(c.mesh)
        if c.name.startswith("sphere"):
            import tempfile
            d = tempfile.mkdtemp()
            _foam().write_case(c, d)
            c = _foam().read_case(d)
            shutil.rmtree(d)
        for k in ("Sf", "Cf", "C", "V", "weights", "delta_coeffs"):
            a, b = getattr(c.geo, k), getattr(gp, k)
            assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b)), k


@pytest.mark.parametrize("compress", [False, True])
def test_generated_cases_round_trip(tmp_path, compress):
    from dugks_amd import cases
    for c in (cases.config0(), cases.sphere_case(dict(nc=4, nr=2, n_up=2, n_down=3, n_side=2), "GH", 6, 1.0, 0.1)):
        d = tmp_path / c.name
        _foam().write_case(c, str(d), compress_mesh=compress)
        assert os.path.exists(d / "constant" / "polyMesh" / ("points.gz" if compress else "points"))
        r = _foam().read_case(str(d))
        a, b = c.arrays(), r.arrays()
        for k, v in a.items():
            if isinstance(v, np.ndarray) and v.dtype.kind == "f" and v.size:
                assert np.max(np.abs(v - b[k])) <= 1e-13 * np.max(np.abs(v)), (c.name, k)
            elif isinstance(v, np.ndarray):
                assert np.array_equal(v, b[k]), (c.name, k)
            else:
                assert v == b[k], (c.name, k)


def test_written_fields_read_back(tmp_path):
    from dugks_amd import cases
    c = cases.cavity_case(6, "NC38", 7, 4.0, 0.1)
    d = str(tmp_path / "cav")
    _foam().write_case(c, d)
    rng = np.random.default_rng(1)
    n, nB = c.n_cells, c.arrays()["nB"]
    rho, T = rng.uniform(0.9, 1.1, n), rng.uniform(0.9, 1.1, n)
    U, q = rng.uniform(-0.1, 0.1, (n, 3)), rng.uniform(-0.1, 0.1, (n, 3))
    nF = c.mesh.n_internal_faces + nB
    _foam().write_fields(d, 0.125, 42, 0.01, 0.009, rho, U, T, q, rho_b=rng.uniform(0.9, 1.1, nB),
                         q_surf=rng.uniform(-1, 1, (nF, 3)))
    for f in ("rho", "U", "T", "q", "qSurf", "stressSurf", "qWall", "stressWall", "uniform/time"):
        assert os.path.exists(os.path.join(d, "0.125", f)), f
    back = _foam().read_fields(d, "0.125")
    assert np.array_equal(back["rho"], rho) and np.array_equal(back["U"], U) and np.array_equal(back["T"], T)
    txt = open(os.path.join(d, "0.125", "uniform", "time")).read()
    assert "index           42;" in txt and 'name            "0.125";' in txt


def test_boundary_types_round_trip(tmp_path):
    """Every rho boundary type the path carries survives write_case -> read_case with its
    DF code and pressure (discreteVelocity.C:188-195)."""
    from dugks_amd import cases
    U = (0.1, 0.0, 0.0)
    sides = dict(left=("pressureIn", 1.0, U, 1.1, 1.05), right=("pressureOut", 1.0, U, 1.0, 0.95),
                 bottom=("symmetryPlane", 1.0, U, 1.0), top=("symmetryMod", 1.0, U, 1.0))
    c = cases.box_case(5, 4, sides, rule="NC38", n_dv=7, vel_extent=4.0)
    d = str(tmp_path / "box")
    _foam().write_case(c, d)
    r = _foam().read_case(d)
    a, b = c.arrays(), r.arrays()
    assert list(b["patch_type"]) == [5, 6, 8, 7]     # pressureIn, pressureOut, symmetryPlane, symmetryMod
    assert list(b["patch_pressure"]) == [1.05, 0.95, 0.0, 0.0]
    assert np.array_equal(a["bf_T"][:4], b["bf_T"][:4])
    assert [p.geometric_type for p in r.mesh.patches] == ["patch", "patch", "symmetryPlane", "patch", "empty"]
    zg = ("zeroGradient", 1.0, U, 1.0)
    c2 = cases.box_case(4, 4, dict(left=("farField", 1.2, U, 0.9), right=zg, bottom=zg, top=zg), rule="NC38",
                        n_dv=7, vel_extent=4.0)
    d2 = str(tmp_path / "box2")
    _foam().write_case(c2, d2)
    b2 = _foam().read_case(d2).arrays()
    assert list(b2["patch_type"]) == [4, 3, 3, 3]
    assert np.all(b2["bf_rho"][:4] == 1.2) and np.all(b2["bf_T"][:4] == 0.9)


def _copy_demo(tmp_path):
    d = tmp_path / "demo"
    shutil.copytree(DEMO, d)
    return d


def test_reader_error_paths(tmp_path):
    foam = _foam()
    with pytest.raises(foam.FoamError, match="cannot open"):
        foam.read_case(str(tmp_path / "nothing"))
    d = _copy_demo(tmp_path)
    rho = (d / "0" / "rho").read_text()
    (d / "0" / "rho").write_text(rho.replace("calculatedMaxwell", "totalPressure", 1))
    with pytest.raises(foam.FoamError, match="totalPressure.*not supported"):
        foam.read_case(str(d))
    (d / "0" / "rho").write_text(rho.replace("calculatedMaxwell", "pressureIn", 1))
    with pytest.raises(foam.FoamError, match="positive `pressureIn`"):
        foam.read_case(str(d))
    (d / "0" / "rho").write_text(rho)
    ctl = (d / "system" / "controlDict").read_text()
    (d / "system" / "controlDict").write_text("#include \"other\"\n" + ctl)
    with pytest.raises(foam.FoamError, match="not supported"):
        foam.read_case(str(d))
    (d / "system" / "controlDict").write_text(ctl)
    xis = (d / "constant" / "Xis").read_text()
    (d / "constant" / "Xis").write_text(xis.replace("format      ascii", "format      binary"))
    with pytest.raises(foam.FoamError, match="binary"):
        foam.read_case(str(d))
    (d / "constant" / "Xis").write_text(xis.replace("28\n(", "27\n("))
    with pytest.raises(foam.FoamError, match="announces 27"):
        foam.read_case(str(d))
    (d / "constant" / "Xis").write_text(xis)
    dv = (d / "constant" / "DVMProperties").read_text()
    (d / "constant" / "DVMProperties").write_text(dv.replace("nDV", "nXX"))
    with pytest.raises(foam.FoamError, match="nDV"):
        foam.read_case(str(d))
    (d / "constant" / "DVMProperties").write_text(dv)
    assert foam.read_case(str(d)).n_cells == 3600


def test_facade_header_compiles_and_driver_error_paths(tmp_path):
    src = tmp_path / "use_facade.C"
    src.write_text('#include "fvDVM.H"\nint main() {\n  try { dugks::fvDVM dvm("/nonexistent");\n'
                   '    double a, b; dvm.getCoNum(a, b); dvm.evolution(); dvm.write(); }\n'
                   '  catch (const dugks::Error &e) { return 3; }\n  return 0;\n}\n')
    foam_dir = os.path.join(ROOT, "psp-vst_amd", "foam")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", foam_dir, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    foam = _foam()
    rc, out, err = foam.run_driver(str(tmp_path), "-bogus")
    assert rc == 1 and "unknown option -bogus" in err
    rc, out, err = foam.run_driver(str(tmp_path), "-dvParallel")
    assert rc == 1 and "not part of this one-GPU build" in err
    rc, out, err = foam.run_driver(str(tmp_path / "nothing"))
    assert rc == 1 and "cannot open" in err
    import torch
    if not torch.cuda.is_available():  # a readable case on a host without a GPU: a clean error
        d = _copy_demo(tmp_path)
        rc, out, err = foam.run_driver(str(d), "-maxSteps", "1")
        assert rc == 1 and "dugks_create" in err
