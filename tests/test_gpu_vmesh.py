"""The velocity mesh of the VMESH build and the macroFlux cell update on the GPU (staged HIP
path) against the FP64 oracle, through the C-ABI (src/fvDVM/fvDVM/fvDVM.C:129-253,
1521-1682).

* 2-D graded cavity with a jittered 12 x 12 velocity mesh (general hexahedra: the velocities
  are no tensor product), macroFlux no / yes (face-macro fluxes) / yes (velocity-moment
  fluxes): FP64 within 1e-12 of the oracle after 30 steps, FP32 storage within the mixed
  budget (rho, U, T 1e-5; q 1e-3);
* the two flux forms on a tensor-product set (the fused path steps aside for them);
* 3-D small sphere with a jittered 7^3 velocity mesh and the velocity-moment flux update;
* a case directory with constant/vMesh read by dugks_foam_read_vmesh and run through the
  dugksFoam driver (-vmesh): written fields within 1e-10 of the oracle;
* the adjoint with the tensor set given as a velocity list equals the tensor-set adjoint;
* refusals: the fused path for a velocity list, the adjoint for the flux update.
"""
import dataclasses
import os

import numpy as np
import pytest

from conftest import field_rel_err

pytestmark = pytest.mark.gpu

WALL = ("calculatedMaxwell", 1.0, (0.0, 0.0, 0.0), 1.0)
LID = ("calculatedMaxwell", 1.0, (0.2, 0.0, 0.0), 1.0)
SMALL_SPHERE = dict(nc=4, nr=2, n_up=2, n_down=3, n_side=2)


def _vcase(macro_flux):
    from dugks_amd import cases, mesh
    c = cases.box_case(10, 8, dict(left=WALL, right=WALL, bottom=WALL, top=LID), rule="NC38", n_dv=10,
                       vel_extent=4.5, grading=2.5, T0=lambda C: 1.0 + 0.1 * C[:, 0],
                       rho0=lambda C: 1.0 + 0.05 * np.sin(3 * C[:, 1]))
    vm = mesh.velocity_mesh(np.linspace(-4.5, 4.5, 13), np.linspace(-4.5, 4.5, 13), [-0.5, 0.5], jitter=0.3)
    g = mesh.geometry(vm)
    xi = g.C.copy()
    xi[:, 2] = 0.0
    return dataclasses.replace(c, vlist=(xi, g.V.copy()), xi_max=float(np.max(np.linalg.norm(xi, axis=1))),
                               macro_flux=macro_flux)


def _pair(case, steps, precision, fused=-1):
    from dugks_amd import hip
    from oracle import Oracle
    o = Oracle(case)
    g = hip.FvDVM(case, precision=precision, fused=fused)
    ho = []
    for _ in range(steps):
        o.step()
        ho.append(o.dt)
    hg = g.run(steps, record_dt=True)
    mo, mg = o.cell_macros(), g.cell_macros()
    return {k: field_rel_err(mg[k], mo[k]) for k in ("rho", "U", "T", "q")}, np.array(ho), hg, g


def _check(e, precision):
    tol = 1e-12 if precision == 0 else 1e-5
    for k in ("rho", "U", "T"):
        assert e[k] < tol, (k, e[k])
    assert e["q"] < (1e-12 if precision == 0 else 1e-3)


@pytest.mark.parametrize("macro_flux,precision", [(0, 0), (1, 0), (2, 0), (0, 1), (2, 1)])
def test_velocity_mesh_cavity_matches_oracle(macro_flux, precision):
    c = _vcase(macro_flux)
    print("velocity mesh cavity", macro_flux, precision, flush=True)
    e, ho, hg, g = _pair(c, 30, precision)
    print(e)
    assert g.path_info()["fused"] == 0 and g.Nv == 144
    assert np.max(np.abs(hg - ho) / ho) < (1e-12 if precision == 0 else 1e-5)
    _check(e, precision)


@pytest.mark.parametrize("macro_flux", [1, 2])
def test_macro_flux_tensor_set_matches_oracle(macro_flux):
    from dugks_amd import cases
    c = dataclasses.replace(cases.cavity_case(8, "NC38", 10, 4.0, 0.1), macro_flux=macro_flux)
    print("macro flux tensor set", macro_flux, flush=True)
    e, _, _, g = _pair(c, 30, 0)
    print(e)
    assert g.path_info()["fused"] == 0
    _check(e, 0)


def test_velocity_mesh_sphere_matches_oracle():
    from dugks_amd import cases, mesh
    c = cases.sphere_case(SMALL_SPHERE, "GH", 6, 1.0, 0.1)
    e = np.linspace(-3.5, 3.5, 8)
    vm = mesh.velocity_mesh(e, e, e, jitter=0.25, seed=11)
    g = mesh.geometry(vm)
    c = dataclasses.replace(c, vlist=(g.C.copy(), g.V.copy()),
                            xi_max=float(np.max(np.linalg.norm(g.C, axis=1))), macro_flux=2)
    print("velocity mesh sphere", c.n_cells, len(g.V), flush=True)
    err, _, _, _ = _pair(c, 10, 0)
    print(err)
    _check(err, 0)


def test_vmesh_case_directory_through_driver(tmp_path):
    from dugks_amd import cases, foam, mesh
    from oracle import Oracle
    c = dataclasses.replace(cases.cavity_case(10, "NC38", 10, 4.0, 0.1), macro_flux=1)
    vm = mesh.velocity_mesh(np.linspace(-4.0, 4.0, 11), np.linspace(-4.0, 4.0, 11), [0.0, 1.0], jitter=0.2)
    d = tmp_path / "vmesh"
    foam.write_case(c, str(d), vmesh=vm)
    print("dugksFoam -vmesh", d, flush=True)
    rc, out, err = foam.run_driver(str(d), "-maxSteps", "40", "-vmesh")
    assert rc == 0, (rc, out[-2000:], err[-2000:])
    r = foam.read_case(str(d), "0", vmesh=True)
    assert r.macro_flux == 2 and r.nv == 100
    o = Oracle(r)
    for _ in range(40):
        o.step()
    times = []
    for n in os.listdir(d):
        try:
            times.append((float(n), n))
        except ValueError:
            pass
    got = foam.read_fields(str(d), max(times)[1])
    m = o.cell_macros()
    for k in ("rho", "U", "T"):
        e = field_rel_err(got[k], m[k])
        print("vmesh via driver", k, e)
        assert e < 1e-10, k


def test_refusals():
    from dugks_amd import cases, hip
    with pytest.raises(hip.DugksError, match="tensor-product velocity set"):
        hip.FvDVM(_vcase(0), fused=1)
    c = dataclasses.replace(cases.cavity_case(6, "NC38", 7, 4.0, 0.1), macro_flux=2)
    with pytest.raises(hip.DugksError, match="macro_flux"):
        hip.FvDVM(c, fused=2)
    s = hip.FvDVM(c)
    with pytest.raises(hip.DugksError, match="macro_flux"):
        s.adjoint([s.dt, s.dt], 1)
    sym = ("symmetryPlane", 1.0, (0.0, 0.0, 0.0), 1.0)
    b = cases.box_case(4, 4, dict(left=WALL, right=sym, bottom=WALL, top=LID), rule="NC38", n_dv=7, vel_extent=4.0)
    xi, w = _vcase(0).vlist
    with pytest.raises(hip.DugksError, match="symmetry"):
        hip.FvDVM(dataclasses.replace(b, vlist=(xi, w)))


def test_adjoint_with_a_velocity_list_equals_the_tensor_set():
    """The reverse kernels take the velocity set as a list: the tensor set given as an explicit
    list (same order, ix fastest) gives the tensor set's objective and gradients."""
    from dugks_amd import adjoint, cases, hip, quadrature
    c = cases.cavity_case(6, "NC38", 10, 4.0, 0.1, wall_T=1.0 + 0.05 * np.sin(np.arange(6)))
    xi, w = quadrature.tensor_product(c.xis, c.wts, c.D)
    lst = dataclasses.replace(c, vlist=(xi, w))
    s = hip.FvDVM(c, precision=hip.FP64)
    dts = s.run(4, record_dt=True)
    s.close()
    qw = adjoint.bottom_wall_q_weight(c)
    out = []
    for case in (c, lst):
        s = hip.FvDVM(case, precision=hip.FP64)
        out.append(s.adjoint(dts, 2, q_weight=qw))
        s.close()
    a, b = out
    print("adjoint tensor vs list", a["J"], b["J"], a["dmuRef"], b["dmuRef"])
    assert abs(a["J"] - b["J"]) <= 1e-13 * abs(a["J"])
    assert abs(a["dmuRef"] - b["dmuRef"]) <= 1e-12 * abs(a["dmuRef"])
    assert np.max(np.abs(a["dTw"] - b["dTw"])) <= 1e-12 * np.max(np.abs(a["dTw"]))
    assert np.max(np.abs(a["dUw"] - b["dUw"])) <= 1e-12 * np.max(np.abs(a["dUw"]))
