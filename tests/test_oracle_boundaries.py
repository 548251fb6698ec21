"""Boundary types beyond the diffuse wall and the frozen inlet, in the FP64 oracle:
zeroGradient, farField, pressureIn, pressureOut (fvDVM.C:883-940, 1745-1821), symmetryMod /
symmetryPlane (fvDVM.C:1009-1025, 1100-1179, 1449-1456; discreteVelocity.C:691-775).

* A uniform flow with farField and zeroGradient sides, or pressureIn / pressureOut sides at
  the flow's pressure, is a fixed point to round-off with a velocity set that integrates
  shifted Maxwellians exactly (40-point trapezoid on [-10, 10]).  The farField patch U is
  set far off the flow U: the incoming Maxwellian takes the owner cell's U (fvDVM.C:927-933),
  so the fixed point holds only if the patch U is not read.
* zeroGradient faces hold the owner cell's f-bar+ for every velocity before relaxation.
* A pressure difference drives the flow from the pressureIn side to the pressureOut side,
  and the inlet density is the one of the equation of state p = rho R T at the patch T.
* Symmetry faces carry no mass and no energy before the face relaxation: every incoming
  entry holds the value of its mirror velocity (discreteVelocity.C:725-770), whose flux is
  the opposite.
"""
import numpy as np
import pytest


def _cases():
    from dugks_amd import cases
    return cases


def _oracle(case):
    from oracle import Oracle
    return Oracle(case)


FINE = dict(rule="trapezoid", n_dv=40, vel_extent=10.0)
U = (0.2, 0.05, 0.0)


@pytest.mark.parametrize("kind", ["farField_zeroGradient", "pressure"])
def test_uniform_flow_is_a_fixed_point(kind):
    cases = _cases()
    if kind == "farField_zeroGradient":
        ff = ("farField", 1.0, (9.0, -9.0, 0.0), 1.0)   # patch U is not read by the DF boundary
        zg = ("zeroGradient", 1.0, (9.0, -9.0, 0.0), 1.0)
        sides = dict(left=ff, right=ff, bottom=zg, top=zg)
    else:
        sides = dict(left=("pressureIn", 0.0, (9.0, 0.0, 0.0), 1.0, 1.0), right=("pressureOut", 0.0, U, 0.5, 1.0),
                     bottom=("zeroGradient", 1.0, U, 1.0), top=("farField", 1.0, (0.0, 9.0, 0.0), 1.0))
    c = cases.box_case(8, 6, sides, U0=U, **FINE)
    o = _oracle(c)
    o.step(10)
    m = o.cell_macros()
    assert np.max(np.abs(m["rho"] - 1.0)) < 1e-14
    assert np.max(np.abs(m["U"] - np.array(U))) < 1e-14
    assert np.max(np.abs(m["T"] - 1.0)) < 1e-14
    assert np.max(np.abs(m["q"])) < 1e-14


def test_zero_gradient_faces_hold_the_owner_f_bar_plus():
    cases = _cases()
    zg = ("zeroGradient", 1.0, U, 1.0)
    c = cases.box_case(6, 5, dict(left=zg, right=zg, bottom=zg, top=zg),
                       rho0=lambda C: 1.0 + 0.1 * np.sin(3 * C[:, 0]) * np.cos(2 * C[:, 1]),
                       U0=lambda C: np.stack([0.1 * C[:, 1], -0.05 * C[:, 0], 0 * C[:, 0]], 1),
                       T0=lambda C: 1.0 + 0.05 * C[:, 0] * C[:, 1], rule="NC38", n_dv=7, vel_extent=4.0)
    o = _oracle(c)
    o.step(2)
    o.stages(o.dt, 0, 4)
    gB, hB = o.bar()
    gS, hS = o.surf()
    a = c.arrays()
    own = a["owner"][a["nIF"]:]
    assert np.array_equal(gS[:, a["nIF"]:], gB[:, own]) and np.array_equal(hS[:, a["nIF"]:], hB[:, own])


def test_pressure_difference_drives_flow_and_inlet_obeys_equation_of_state():
    cases = _cases()
    Tin = 1.1
    sides = dict(left=("pressureIn", 0.0, (0.0, 0.0, 0.0), Tin, 1.05),
                 right=("pressureOut", 0.0, (0.0, 0.0, 0.0), 1.0, 0.95),
                 bottom=("calculatedMaxwell", 1.0, (0.0, 0.0, 0.0), 1.0),
                 top=("calculatedMaxwell", 1.0, (0.0, 0.0, 0.0), 1.0))
    c = cases.box_case(12, 6, sides, extent=(0.0, 2.0, 0.0, 1.0), **FINE)
    o = _oracle(c)
    ux = []
    for _ in range(4):
        o.step(10)
        ux.append(float(np.mean(o.cell_macros()["U"][:, 0])))
    assert ux[0] > 0 and all(b > a for a, b in zip(ux, ux[1:]))   # accelerating from left to right
    a = c.arrays()
    left = slice(int(a["patch_start"][0]), int(a["patch_start"][0] + a["patch_size"][0]))
    rho_in = o.wall_rho()[left]
    assert np.max(np.abs(rho_in * c.gas.R * Tin - 1.05)) < 1e-15


@pytest.mark.parametrize("sym", ["symmetryMod", "symmetryPlane"])
def test_symmetry_faces_carry_no_mass_or_energy_before_relaxation(sym):
    cases = _cases()
    wall = ("calculatedMaxwell", 1.0, (0.0, 0.0, 0.0), 1.0)
    sides = dict(left=wall, right=(sym, 1.0, (0.0, 0.0, 0.0), 1.0), bottom=wall,
                 top=("calculatedMaxwell", 1.0, (0.0, 0.2, 0.0), 1.0))
    c = cases.box_case(6, 8, sides, extent=(-1.0, 0.0, 0.0, 1.0), rule="NC38", n_dv=10, vel_extent=4.0,
                       T0=lambda C: 1.0 + 0.1 * C[:, 1])
    o = _oracle(c)
    o.step(3)
    o.stages(o.dt, 0, 4)
    gS, hS = o.surf()
    xi, w = o.vset()
    a = c.arrays()
    k = int(np.nonzero(np.array(a["patch_names"]) == "right")[0][0])
    fs = a["nIF"] + int(a["patch_start"][k]) + np.arange(int(a["patch_size"][k]))
    Sf = a["Sf"][fs]
    sig = xi @ Sf.T
    n = len(c.xis)
    ix, iy = np.arange(len(w)) % n, np.arange(len(w)) // n
    mirror = (n - 1 - ix) + n * iy
    incoming = sig <= 0
    for v in np.nonzero(incoming.any(1))[0]:
        assert np.array_equal(gS[v, fs][incoming[v]], gS[mirror[v], fs][incoming[v]])
    g, h = gS[:, fs], hS[:, fs]
    mass = w @ (sig * g)
    energy = w @ (sig * (np.sum(xi * xi, 1)[:, None] * g + h))
    out = w @ (np.where(sig > 0, sig, 0.0) * g)
    assert np.all(out > 0)
    assert np.max(np.abs(mass) / out) < 1e-14 and np.max(np.abs(energy) / out) < 1e-13
