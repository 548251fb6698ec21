"""Pins of the FP64 CPU oracle (oracle/dugks_oracle.c) that do not re-type its formulas.

* an independent numpy restatement (tests/numpy_ref.py) of whole steps, <= 1e-12;
* least-squares gradient exact for linear fields on interior cells of a distorted mesh,
  and equal to a brute-force numpy lstsq of the weighted system;
* a uniform rest equilibrium with walls at T0 stays fixed (round-off) when the velocity
  quadrature integrates the Maxwellian exactly (half-range Gauss-Hermite);
* the Shakhov equilibrium reproduces rho, rho U, rho E and a heat flux of (1 - Pr) q;
* the internal face fluxes telescope: sum_c V_c (g~_new - 4/3 g+ + 1/3 g~) equals
  -dt sum_boundary (xi.Sf) g_f for every velocity;
* (tests/test_oracle_identities.py) closed forms and identities of the DUGKS definitions for
  the reconstruction, the f-bar+ blend, the face relaxation, both heat-flux corrections, the
  diffuse wall and the q1_compat read;
* golden vectors committed under tests/golden/ reproduce.
"""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, field_rel_err

GOLDEN = os.path.join(ROOT, "tests", "golden")


def _cases():
    from dugks_amd import cases
    return cases


def test_oracle_matches_independent_numpy_restatement():
    from numpy_ref import NumpyDUGKS
    from oracle import Oracle
    from dugks_amd.timecontrol import next_delta_t
    c = _cases().cavity_case(5, "NC38", 7, 4.0, 0.2)
    o = Oracle(c)
    n = NumpyDUGKS(c)
    n.init()
    dt = c.controls.deltaT
    for _ in range(6):
        co_o, _ = o.co_num()
        co_n = n.co_num(dt)
        assert abs(co_o - co_n) <= 1e-13 * co_n
        o.step()
        dt = next_delta_t(dt, co_n, c.controls)
        n.step(dt)
    m = o.cell_macros()
    assert field_rel_err(m["rho"], n.rho) < 1e-12
    assert field_rel_err(m["U"], n.U) < 1e-12
    assert field_rel_err(m["T"], n.T) < 1e-12
    assert field_rel_err(m["q"], n.q) < 1e-11
    g, h = o.df()
    assert field_rel_err(g, n.g) < 1e-12 and field_rel_err(h, n.h) < 1e-12
    fm = o.face_macros()
    assert field_rel_err(fm["q"], n.face[3]) < 1e-11
    assert field_rel_err(o.wall_rho(), n.rho_w) < 1e-12


def test_oracle_matches_numpy_on_small_sphere():
    """3-D hex mesh, diffuse sphere wall, frozen-incoming inlet/outlet/side patches, and
    velocities tangent to shell faces (xi.Sf = +-1e-19 round-off decides the upwind side)."""
    from numpy_ref import NumpyDUGKS
    from oracle import Oracle
    from dugks_amd.timecontrol import next_delta_t
    c = _cases().sphere_case(dict(nc=4, nr=2, n_up=2, n_down=3, n_side=2), "GH", 6, 1.0, 0.1)
    o = Oracle(c)
    n = NumpyDUGKS(c)
    n.init()
    dt = c.controls.deltaT
    for _ in range(5):
        dt = next_delta_t(dt, n.co_num(dt), c.controls)
        n.step(dt)
        o.step()
    m = o.cell_macros()
    for k, v in (("rho", n.rho), ("U", n.U), ("T", n.T)):
        assert field_rel_err(m[k], v) < 1e-13, k
    assert field_rel_err(m["q"], n.q) < 1e-12
    wall = n.btype == 1
    assert field_rel_err(o.wall_rho()[wall], n.rho_w[wall]) < 1e-13


def _distorted_cavity(N, seed=3):
    from dugks_amd import mesh as meshmod
    m = meshmod.cavity(N)
    rng = np.random.default_rng(seed)
    h = 1.0 / N
    pts = m.points.copy()
    interior = (pts[:, 0] > 1e-12) & (pts[:, 0] < 1 - 1e-12) & (pts[:, 1] > 1e-12) & (pts[:, 1] < 1 - 1e-12)
    # same in-plane displacement on both z layers keeps the faces planar
    n1 = N + 1
    disp = rng.uniform(-0.25 * h, 0.25 * h, size=(n1 * n1, 2))
    for k in range(2):
        sl = slice(k * n1 * n1, (k + 1) * n1 * n1)
        pts[sl, :2] += np.where(interior[sl, None], disp, 0.0)
    m.points = pts
    return m


def test_ls_gradient_exact_for_linear_fields_and_equals_lstsq():
    from dugks_amd import mesh as meshmod
    from oracle import Oracle
    c = _cases().cavity_case(6, "NC38", 4, 4.0, 0.1)
    c.mesh = _distorted_cavity(6)
    c.geo = meshmod.geometry(c.mesh)
    o = Oracle(c)
    own_ls, nei_ls, _ = o.ls_vectors()
    a = c.arrays()
    own, nei, C = a["owner"], a["neighbour"], a["C"]
    nIF = a["nIF"]
    grad_true = np.array([0.7, -1.3, 0.0])
    phi = C @ grad_true + 0.25
    G = np.zeros((a["nC"], 3))
    delta = phi[nei] - phi[own[:nIF]]
    np.add.at(G, own[:nIF], own_ls * delta[:, None])
    np.add.at(G, nei, -nei_ls * delta[:, None])
    bcells = set(own[nIF:].tolist())
    interior = np.array([i for i in range(a["nC"]) if i not in bcells])
    assert np.max(np.abs(G[interior] - grad_true)) < 1e-12
    # brute force: weighted least squares of each cell's differences, boundary rows with
    # zero right-hand side, via numpy lstsq
    magSf = np.linalg.norm(a["Sf"], axis=1)
    for cell in range(a["nC"]):
        rows, rhs = [], []
        for f in range(nIF):
            if own[f] == cell or nei[f] == cell:
                d = C[nei[f]] - C[own[f]]
                wgt = magSf[f] / (d @ d) * ((1 - a["weights"][f]) if own[f] == cell else a["weights"][f])
                s = 1.0 if own[f] == cell else -1.0
                rows.append(math.sqrt(wgt) * s * d)
                rhs.append(math.sqrt(wgt) * s * (phi[nei[f]] - phi[own[f]]))
        for b in range(a["nB"]):
            f = nIF + b
            if own[f] == cell:
                n = a["Sf"][f] / magSf[f]
                d = n * (n @ (a["Cf"][f] - C[cell]))
                rows.append(math.sqrt(magSf[f] / (d @ d)) * d)
                rhs.append(0.0)
        Aw = np.array(rows)[:, :2]
        sol = np.linalg.lstsq(Aw, np.array(rhs), rcond=None)[0]
        assert np.max(np.abs(sol - G[cell, :2])) < 1e-11 * max(1.0, np.max(np.abs(sol)))


def test_rest_state_is_a_fixed_point():
    from oracle import Oracle
    c = _cases().cavity_case(6, "GH", 28, 1.0 * math.sqrt(2.0), 0.1, lid_speed=0.0)
    o = Oracle(c)
    m0 = o.cell_macros()
    for _ in range(10):
        o.step()
    m = o.cell_macros()
    assert np.max(np.abs(m["rho"] - m0["rho"])) < 1e-13
    assert np.max(np.abs(m["T"] - m0["T"])) < 1e-13
    assert np.max(np.abs(m["U"])) < 1e-13
    assert np.max(np.abs(m["q"])) < 1e-13
    assert np.max(np.abs(o.wall_rho() - 1.0)) < 1e-13


def test_shakhov_equilibrium_moments():
    """f-bar-plus of a zero distribution is a * f^S: its moments give rho, rho U, T and a
    heat flux (1 - Pr) q (fvDVM.C:673-679), up to the quadrature error of a fine set."""
    from oracle import Oracle
    c = _cases().cavity_case(3, "trapezoid", 96, 9.0, 0.1)
    o = Oracle(c)
    nC = o.nC
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.8, 1.2, nC)
    U = np.zeros((nC, 3))
    U[:, :2] = rng.uniform(-0.3, 0.3, (nC, 2))
    T = rng.uniform(0.8, 1.3, nC)
    q = np.zeros((nC, 3))
    q[:, :2] = rng.uniform(-0.05, 0.05, (nC, 2))
    o.set_cell_macros(rho, U, T, q)
    o.set_df(np.zeros((o.Nv, nC)), np.zeros((o.Nv, nC)))
    dt = 0.01
    o.evolution(dt)
    gB, hB = o.bar()
    tau = c.gas.muRef * (T / c.gas.Tref) ** c.gas.omega / (rho * T * c.gas.R)
    a = 1.5 * dt / (2 * tau + dt)
    gS, hS = gB / a, hB / a
    xi, w = o.vset()
    r = w @ gS
    ru = np.einsum("v,vc,vk->ck", w, gS, xi)
    E = 0.5 * (w @ (gS * np.sum(xi ** 2, 1)[:, None] + hS))
    assert np.max(np.abs(r - rho)) < 1e-11
    assert np.max(np.abs(ru / r[:, None] - U)) < 1e-11
    Tm = (E - 0.5 * r * np.sum((ru / r[:, None]) ** 2, 1)) / (1.5 * c.gas.R * r)
    assert np.max(np.abs(Tm - T)) < 1e-11
    cc = xi[:, None, :] - U[None, :, :]
    qm = 0.5 * np.einsum("v,vck,vc->ck", w, cc, np.sum(cc * cc, 2) * gS + hS)
    assert np.max(np.abs(qm - (1 - c.gas.Pr) * q)) < 1e-11


def test_flux_telescoping():
    """Internal face fluxes cancel in the volume sum: what is left is the boundary flux."""
    from oracle import Oracle
    c = _cases().cavity_case(5, "NC38", 10, 4.0, 0.1)
    o = Oracle(c)
    o.step(3)
    g0, _ = o.df()
    o.step(1)
    g1, _ = o.df()
    gB, _ = o.bar()
    gS, _ = o.surf()
    xi, w = o.vset()
    a = c.arrays()
    sig = xi @ a["Sf"].T
    V = a["V"]
    lhs = (g1 - (4.0 / 3 * gB - 1.0 / 3 * g0)) @ V          # [Nv]
    rhs = -o.dt * np.sum(sig[:, o.nIF:] * gS[:, o.nIF:], 1)
    assert np.max(np.abs(lhs - rhs)) < 1e-13 * np.max(np.abs(g1 @ V))


def test_golden_vectors_reproduce():
    from oracle import Oracle
    path = os.path.join(GOLDEN, "cavity4_nc38x7_5steps.npz")
    gold = np.load(path)
    c = _cases().cavity_case(4, "NC38", 7, 4.0, 0.1)
    o = Oracle(c)
    o.step(5)
    m = o.cell_macros()
    for k in ("rho", "U", "T", "q"):
        assert field_rel_err(m[k], gold[k]) < 1e-13, k
    g, h = o.df()
    assert field_rel_err(g, gold["g"]) < 1e-13
    assert field_rel_err(h, gold["h"]) < 1e-13
