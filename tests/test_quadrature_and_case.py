"""Velocity sets, the Kn -> muRef formula, the adaptive time step and the mesh geometry.

Pinned by data that ships with the reference (the cavity demo's Xis/weights and muRef,
copied into tests/golden/demo_cavity.json by tests/golden/extract_demo_vset.py), by closed
forms (Maxwellian moments, exact polynomial integration, uniform-mesh geometry) and by
invariants of the geometry (closed cells, divergence theorem).
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import ROOT

DEMO = json.load(open(os.path.join(ROOT, "tests", "golden", "demo_cavity.json")))


def test_half_range_gauss_hermite_reproduces_demo_set():
    from dugks_amd import quadrature
    C = math.sqrt(2.0 * DEMO["R"] * DEMO["T0"])
    x, w = quadrature.half_range_gauss_hermite(C, 28)
    # the demo's scale C came from R = k/m carried to more digits (2e-12 apart) and a
    # non-symmetric eigensolver; agreement is ~1e-11, a wrong recurrence is off by O(1)
    assert np.max(np.abs(x - np.array(DEMO["xis"])) / C) < 1e-10
    assert np.max(np.abs(w - np.array(DEMO["weights"])) / np.max(DEMO["weights"])) < 1e-10
    assert abs(DEMO["xiMax"] - x.max()) < 1e-10 * DEMO["xiMax"]


@pytest.mark.parametrize("source", ["demo", "generated"])
def test_gauss_hermite_maxwellian_moments(source):
    from dugks_amd import quadrature
    C = math.sqrt(2.0 * DEMO["R"] * DEMO["T0"])
    if source == "demo":
        x, w = np.array(DEMO["xis"]), np.array(DEMO["weights"])
    else:
        x, w = quadrature.half_range_gauss_hermite(C, 28)
    M = np.exp(-(x / C) ** 2) / (C * math.sqrt(math.pi))
    for k, exact in enumerate([1.0, 0.5, 0.75, 15.0 / 8.0, 105.0 / 16.0]):
        assert abs(np.sum(w * M * (x / C) ** (2 * k)) - exact) < 1e-13 * exact


@pytest.mark.parametrize("rule,N,degree", [("NC", 9, 5), ("NC", 17, 5), ("NC38", 28, 3), ("NC38", 7, 3),
                                           ("trapezoid", 64, 1)])
def test_newton_cotes_polynomial_exactness(rule, N, degree):
    from dugks_amd import quadrature
    L = 3.0
    x, w = quadrature.make_rule(rule, L, N)
    for k in range(degree + 1):
        exact = 0.0 if k % 2 else 2 * L ** (k + 1) / (k + 1)
        assert abs(np.sum(w * x ** k) - exact) < 1e-12 * max(1.0, exact)
    k = degree + 1 if (degree + 1) % 2 == 0 else degree + 2
    exact = 2 * L ** (k + 1) / (k + 1)
    assert abs(np.sum(w * x ** k) - exact) > 1e-10 * exact  # the next even power is not exact


def test_rule_validation():
    from dugks_amd import quadrature
    with pytest.raises(ValueError):
        quadrature.make_rule("NC", 4.0, 28)       # composite Boole needs 4Z+1 (setDV.py:153-156)
    with pytest.raises(ValueError):
        quadrature.make_rule("GH", 1.0, 7)        # half-range GH needs an even count
    with pytest.raises(ValueError):
        quadrature.make_rule("nope", 1.0, 7)


def test_tensor_product_order_ix_fastest():
    from dugks_amd import quadrature
    x = np.array([-1.0, 0.5, 2.0])
    w = np.array([0.1, 0.2, 0.3])
    xi, wv = quadrature.tensor_product(x, w, 3)
    assert xi.shape == (27, 3)
    i = 1 + 3 * (2 + 3 * 0)   # ix=1, iy=2, iz=0
    assert np.allclose(xi[i], [0.5, 2.0, -1.0]) and abs(wv[i] - 0.1 * 0.3 * 0.2) < 1e-17
    xi2, w2 = quadrature.tensor_product(x, w, 2)
    assert np.allclose(xi2[1 + 3 * 2], [0.5, 2.0, 0.0])


def test_mu_ref_formula_reproduces_demo_value():
    """demo/cavity/para.py:33-46 with its hard-sphere exponent gives the demo's muRef."""
    from dugks_amd.cases import mu_ref_from_kn
    mu = mu_ref_from_kn(DEMO["Kn"], rho0=DEMO["rho0"], R=DEMO["R"], T0=DEMO["T0"], L=1.0, omega=0.5)
    assert abs(mu - DEMO["muRef"]) < 1e-10 * DEMO["muRef"]


def test_next_delta_t_rule():
    from dugks_amd.cases import Controls
    from dugks_amd.timecontrol import next_delta_t
    c = Controls(deltaT=1.0, adjustTimeStep=True, maxCo=0.8, maxDeltaT=1.0)
    assert abs(next_delta_t(1.0, 8.0, c) - 0.1) < 1e-15            # immediate reduction to maxCo
    assert abs(next_delta_t(0.1, 0.08, c) - 0.1 * 1.2) < 1e-15      # growth capped at 1.2
    assert abs(next_delta_t(0.1, 0.5, c) - 0.1 * (1 + 0.1 * 1.6)) < 1e-15  # damped growth
    assert next_delta_t(0.9, 1e-6, c) == 1.0                       # maxDeltaT cap
    assert next_delta_t(0.3, 8.0, Controls(adjustTimeStep=False)) == 0.3


def test_cavity_geometry_closed_forms():
    from dugks_amd import mesh
    N, dz = 8, 0.1
    m = mesh.cavity(N, dz=dz)
    g = mesh.geometry(m)
    h = 1.0 / N
    assert m.n_solution_d() == 2
    assert np.allclose(g.V, h * h * dz, rtol=1e-14)
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="xy")
    assert np.allclose(g.C[:, 0], ((ii.ravel() + 0.5) * h), atol=1e-15)
    assert np.allclose(g.C[:, 1], ((jj.ravel() + 0.5) * h), atol=1e-15)
    assert np.allclose(g.weights, 0.5, atol=1e-14)
    assert np.allclose(g.delta_coeffs, N, rtol=1e-13)
    side = np.arange(m.n_internal_faces)
    assert np.allclose(g.magSf[side], h * dz, rtol=1e-14)
    # owner < neighbour, normals out of the owner
    assert np.all(m.owner[:m.n_internal_faces] < m.neighbour)
    d = g.C[m.neighbour] - g.C[m.owner[:m.n_internal_faces]]
    assert np.all(np.sum(d * g.Sf[side], 1) > 0)


def test_distorted_mesh_geometry_invariants():
    from test_oracle_pins import _distorted_cavity
    from dugks_amd import mesh
    m = _distorted_cavity(7)
    g = mesh.geometry(m)
    nC, nIF = m.n_cells, m.n_internal_faces
    close = np.zeros((nC, 3))
    np.add.at(close, m.owner, g.Sf)
    np.add.at(close, m.neighbour, -g.Sf[:nIF])
    assert np.max(np.abs(close)) < 1e-15
    assert abs(g.V.sum() - 0.1) < 1e-14
    # divergence theorem: sum_f Sf (x) Cf over a closed planar-faced cell = V I
    T = np.zeros((nC, 3, 3))
    np.add.at(T, m.owner, np.einsum("fi,fj->fij", g.Sf, g.Cf))
    np.add.at(T, m.neighbour, -np.einsum("fi,fj->fij", g.Sf[:nIF], g.Cf[:nIF]))
    assert np.max(np.abs(T - g.V[:, None, None] * np.eye(3)[None])) < 1e-14


def test_case_arrays_skip_empty_patches_and_map_bc_types():
    from dugks_amd import cases
    c = cases.config0()
    a = c.arrays()
    assert a["nB"] == 4 * 40 and a["nIF"] == 2 * 40 * 39
    assert list(a["patch_type"]) == [1, 1] and list(a["patch_start"]) == [0, 40]
    assert np.allclose(a["bf_U"][:40, 0], 0.16 * math.sqrt(2.0))
    c.bcs[0].rho_type = "farField"
    assert list(c.arrays()["patch_type"]) == [4, 1]
    c.bcs[0].rho_type = "totalPressure"
    with pytest.raises(ValueError, match="not supported"):
        c.arrays()


def test_sphere_mesh_invariants():
    from dugks_amd import mesh, sphere
    m = sphere.sphere_box_mesh(nc=6, nr=3, n_up=3, n_down=6, n_side=3)
    g = mesh.geometry(m)
    nC, nIF = m.n_cells, m.n_internal_faces
    assert m.n_solution_d() == 3
    close = np.zeros((nC, 3))
    np.add.at(close, m.owner, g.Sf)
    np.add.at(close, m.neighbour, -g.Sf[:nIF])
    assert np.max(np.abs(close)) < 1e-13
    assert np.all(g.V > 0)
    d = g.C[m.neighbour] - g.C[m.owner[:nIF]]
    assert np.all(np.sum(d * g.Sf[:nIF], 1) > 0)
    sp = m.patch("sphere")
    area = g.magSf[sp.start:sp.start + sp.size].sum()
    assert 0.9 * math.pi < area < math.pi          # inscribed facets of the unit-diameter sphere
    box = 10 * 6 * 6
    assert box - math.pi / 6 < g.V.sum() < box      # box minus the faceted sphere
    names = [p.name for p in m.patches]
    assert names == ["sphere", "inlet", "outlet", "sides"]


def test_sphere_config_sizes():
    from dugks_amd import cases, sphere
    m = sphere.sphere_box_mesh(**cases.SPHERE_200K)
    assert 195_000 < m.n_cells < 215_000
    m4 = sphere.sphere_box_mesh(**cases.SPHERE_50K)
    assert 45_000 < m4.n_cells < 55_000
