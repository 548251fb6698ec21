"""Gradient schemes of system/fvSchemes on the path: leastSquares (the library's own vectors)
and Gauss linear (demo/ball/system/fvSchemes:23-27), carried as vectors of the form
grad_P += own . (phi_N - phi_P), grad_N -= nei . (phi_N - phi_P).

The Gauss linear vectors are checked against the gradient computed from the scheme's
definition, grad_P = (1/V_P) sum over every face of the cell of S_f phi_f with
phi_f = w phi_P + (1 - w) phi_N on internal faces and the cell value on boundary faces
(the value the fixedGradient feedback of fvDVM.C:775-809 keeps), on a randomly distorted
mesh; and through the oracle's reconstruction of a linear field on a uniform mesh, where
Gauss linear is exact on interior cells.
"""
import numpy as np

from conftest import ROOT


def test_gauss_linear_vectors_equal_the_gauss_gradient_on_a_distorted_mesh():
    from dugks_amd import cases, mesh as meshmod
    from test_oracle_pins import _distorted_cavity
    c = cases.cavity_case(9, "NC38", 7, 4.0, 0.1)
    c.mesh = _distorted_cavity(9, seed=21)
    c.geo = meshmod.geometry(c.mesh)
    c.grad_scheme = "Gauss linear"
    a = c.arrays()
    m, g = c.mesh, c.geo
    rng = np.random.default_rng(4)
    phi = rng.uniform(-1, 1, m.n_cells)
    nIF = m.n_internal_faces
    own, nei = m.owner, m.neighbour
    # definition: every face of the cell, boundary and empty faces at the cell value
    direct = np.zeros((m.n_cells, 3))
    w = g.weights
    pf = w * phi[own[:nIF]] + (1 - w) * phi[nei]
    np.add.at(direct, own[:nIF], g.Sf[:nIF] * pf[:, None])
    np.add.at(direct, nei, -g.Sf[:nIF] * pf[:, None])
    np.add.at(direct, own[nIF:], g.Sf[nIF:] * phi[own[nIF:]][:, None])
    direct /= g.V[:, None]
    vec = np.zeros((m.n_cells, 3))
    d = phi[nei] - phi[own[:nIF]]
    np.add.at(vec, own[:nIF], a["ls_own"] * d[:, None])
    np.add.at(vec, nei, -a["ls_nei"] * d[:, None])
    assert np.max(np.abs(vec - direct)) < 1e-12 * np.max(np.abs(direct))


def test_oracle_reconstruction_with_gauss_linear_is_exact_for_linear_fields():
    from dugks_amd import cases
    from oracle import Oracle
    c = cases.cavity_case(8, "NC38", 7, 4.0, 0.1)
    c.grad_scheme = "Gauss linear"
    o = Oracle(c)
    a = c.arrays()
    nC, nIF = a["nC"], a["nIF"]
    o.set_cell_macros(np.full(nC, 1.1), np.tile((0.1, -0.05, 0.0), (nC, 1)), np.full(nC, 1.2),
                      np.tile((0.02, 0.01, 0.0), (nC, 1)))
    xi, wv = o.vset()
    Nv = len(wv)
    C = a["C"][:, :2]
    rng = np.random.default_rng(9)
    al, be = rng.uniform(0.2, 1, Nv), rng.uniform(-1, 1, (Nv, 2))
    gt = al[:, None] + be @ C.T
    o.set_df(gt, 2 * gt)
    dt = 0.02
    o.stages(dt, 0, 2)
    gB, _ = o.bar()
    gS, _ = o.surf()
    tau = c.gas.muRef * (1.2 / c.gas.Tref) ** c.gas.omega / (1.1 * c.gas.R * 1.2)
    keep = (1 - 0.25 * dt / tau) / (1 + 0.5 * dt / tau)     # f~ -> f-bar+ (see test_oracle_identities)
    const = gB[:, 0] - keep * gt[:, 0]
    foot = a["Cf"][None, :nIF, :2] - 0.5 * dt * xi[:, None, :2]
    exact = const[:, None] + keep * (al[:, None] + np.einsum("vk,vfk->vf", be, foot))
    bcells = set(a["owner"][nIF:].tolist())
    own, nei = a["owner"][:nIF], a["neighbour"]
    sig = xi @ a["Sf"][:nIF].T
    ok_o = np.array([x not in bcells for x in own])
    ok_n = np.array([x not in bcells for x in nei])
    mask = np.where(sig > 1e-300, ok_o[None, :], np.where(sig < -1e-300, ok_n[None, :], (ok_o & ok_n)[None, :]))
    assert mask.sum() > 500
    assert np.max(np.abs(gS[:, :nIF] - exact)[mask]) < 1e-12 * np.max(np.abs(exact))
    # and the least-squares gradient differs from Gauss linear near the walls
    lsq = Oracle(cases.cavity_case(8, "NC38", 7, 4.0, 0.1)).ls_vectors()[0]
    assert np.max(np.abs(lsq - a["ls_own"])) > 1e-3


def test_case_reader_takes_gauss_linear_from_fv_schemes(tmp_path):
    import shutil
    from dugks_amd import foam
    d = tmp_path / "demo"
    shutil.copytree(f"{ROOT}/tests/golden/demo_cavity", d)
    assert foam.read_case(str(d)).grad_scheme == "leastSquares"
    fs = (d / "system" / "fvSchemes").read_text()
    (d / "system" / "fvSchemes").write_text(fs.replace("default         leastSquares;", "default         Gauss linear;"))
    c = foam.read_case(str(d))
    assert c.grad_scheme == "Gauss linear" and c.arrays()["ls_own"].shape == (7080, 3)
    (d / "system" / "fvSchemes").write_text(fs.replace("default         leastSquares;",
                                                       "default         cellLimited Gauss linear 1;"))
    import pytest
    with pytest.raises(foam.FoamError, match="cellLimited Gauss linear 1"):
        foam.read_case(str(d))
