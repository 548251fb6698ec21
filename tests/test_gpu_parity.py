"""GPU parity of the DUGKS step against the FP64 CPU oracle (oracle/dugks_oracle.c).

Tolerances (relative to the field's max magnitude, see conftest.field_rel_err):
  * FP64 GPU path vs oracle: 1e-12 on the tiny cavity (30 steps), 1e-10 on the 40x40
    cavity with the 28x28 Newton-Cotes set (1000 steps, the configuration the FP64 parity
    requirement names).
  * mixed path (FP32 g/h storage) vs oracle: 1e-5 in rho, U, T (the stated mixed budget);
    q is reported and bounded at 1e-3 of its scale.

`fused` values follow dugks_options.fused: 0 staged, 1 fused with the default schedule
(2-D: recompute for FP32 storage, stored face values for FP64 storage; 3-D: stored face
values; 3-D FP32 storage in its three-kernel form), 2 stored-face-value schedule, 3
recompute schedule, 4 stored face values in the two-pass form.
"""
import numpy as np
import pytest

from conftest import field_rel_err

pytestmark = pytest.mark.gpu


def _cases():
    from dugks_amd import cases
    return cases


def _run_pair(case, steps, precision, fused=-1):
    from dugks_amd import hip
    from oracle import Oracle
    o = Oracle(case)
    g = hip.FvDVM(case, precision=precision, fused=fused)
    hist_o = []
    for _ in range(steps):
        o.step()
        hist_o.append(o.dt)
    hist_g = g.run(steps, record_dt=True)
    return o, g, np.array(hist_o), hist_g


def _errs(o, g):
    mo, mg = o.cell_macros(), g.cell_macros()
    return {k: field_rel_err(mg[k], mo[k]) for k in ("rho", "U", "T", "q")}


@pytest.mark.parametrize("fused", [0, 1, 2, 3])
def test_tiny_cavity_fp64_matches_oracle(fused):
    c = _cases().cavity_case(6, "NC38", 10, 4.0, 0.1)
    o, g, ho, hg = _run_pair(c, 30, 0, fused)
    e = _errs(o, g)
    print("tiny fp64", fused, e)
    assert np.max(np.abs(hg - ho) / ho) < 1e-12
    for k, v in e.items():
        assert v < 1e-12, (k, v)
    fo, fg = o.face_macros(), g.face_macros()
    for k in ("rho", "U", "T", "q"):
        assert field_rel_err(fg[k], fo[k]) < 1e-12, k
    assert field_rel_err(g.boundary_rho(), o.wall_rho()) < 1e-12
    go, ho_ = o.df()
    gg, hg_ = g.df()
    assert field_rel_err(gg, go) < 1e-12
    assert field_rel_err(hg_, ho_) < 1e-12


@pytest.mark.parametrize("fused,precision", [(0, 0), (2, 0), (3, 0), (2, 1), (3, 1)])
def test_tiny_cavity_odd_grid_zero_normal_velocity(fused, precision):
    """Odd Newton-Cotes grid: xi_d = 0 exists, so xi.Sf = 0 on axis-aligned faces and the
    averaging branch of the internal reconstruction (fvDVM.C:847-863) is exercised; with
    FP32 storage (precision 1) the tangent branch of each fused schedule's FP32 pass A runs
    too (halved moments, no stored value), checked against the mixed budget."""
    c = _cases().cavity_case(5, "NC", 9, 4.0, 0.3)
    o, g, ho, hg = _run_pair(c, 20, precision, fused)
    e = _errs(o, g)
    print("odd grid", fused, precision, e)
    tol = 1e-12 if precision == 0 else 1e-5
    for k in ("rho", "U", "T"):
        assert e[k] < tol, (k, e[k])
    assert e["q"] < (1e-12 if precision == 0 else 1e-3)


@pytest.mark.parametrize("fused", [0, 2, 3])
def test_config0_fp64_1000_steps(fused):
    c = _cases().config0()
    o, g, ho, hg = _run_pair(c, 1000, 0, fused)
    e = _errs(o, g)
    print("config0 fp64 1000 steps", e, "dt rel", np.max(np.abs(hg - ho) / ho))
    assert np.max(np.abs(hg - ho) / ho) < 1e-10
    for k, v in e.items():
        assert v < 1e-10, (k, v)


@pytest.mark.parametrize("fused", [0, 2, 3])
def test_config0_mixed_200_steps(fused):
    c = _cases().config0()
    o, g, ho, hg = _run_pair(c, 200, 1, fused)
    e = _errs(o, g)
    print("config0 mixed 200 steps", e)
    for k in ("rho", "U", "T"):
        assert e[k] < 1e-5, (k, e[k])
    assert e["q"] < 1e-3


@pytest.mark.parametrize("fused", [0, 2, 3])
def test_determinism_and_checkpoint_resume(tmp_path, fused):
    from dugks_amd import hip
    c = _cases().cavity_case(20, "NC38", 10, 4.0, 0.1)
    a = hip.FvDVM(c, fused=fused)
    a.run(12)
    b = hip.FvDVM(c, fused=fused)
    b.run(5)
    path = tmp_path / "ck.bin"
    b.checkpoint_save(path)
    dt_b = b.dt
    d = hip.FvDVM(c, fused=fused)
    d.checkpoint_load(path)
    d.dt = dt_b
    d.run(7)
    ga, ha = a.df()
    gd, hd = d.df()
    assert np.array_equal(ga, gd) and np.array_equal(ha, hd)
    ma, md = a.cell_macros(), d.cell_macros()
    for k in ma:
        assert np.array_equal(ma[k], md[k]), k


@pytest.mark.parametrize("fused", [2, 3])
def test_fused_and_staged_agree_on_multiblock_mesh(fused):
    """A 48x48 cavity gives nine 16x16 RCB blocks, so block halos, ring-2 cells, faces shared
    between blocks (a slot in each block, one face-value column) and, in the face-value
    schedule, the 32 face slots
    beyond 512 per block that pass A spreads over velocity lanes (13 velocities per axis:
    a short last stripe) are all exercised; FP64 fused vs staged after 50 steps."""
    from dugks_amd import hip
    c = _cases().cavity_case(48, "NC38", 13, 4.0, 0.05)
    a = hip.FvDVM(c, fused=0)
    b = hip.FvDVM(c, fused=fused)
    a.run(50)
    b.run(50)
    ma, mb = a.cell_macros(), b.cell_macros()
    for k in ("rho", "U", "T", "q"):
        assert field_rel_err(mb[k], ma[k]) < 1e-12, k
    fa, fb = a.face_macros(), b.face_macros()
    for k in ("rho", "U", "T", "q"):
        assert field_rel_err(fb[k], fa[k]) < 1e-12, k


@pytest.mark.parametrize("fused", [2, 3])
def test_fused_fp64_is_bitwise_deterministic(fused):
    from dugks_amd import hip
    c = _cases().cavity_case(40, "NC38", 16, 4.0, 0.1)
    runs = []
    for _ in range(2):
        s = hip.FvDVM(c, fused=fused)
        s.run(20)
        runs.append(s.df())
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])


def test_default_schedule_choice():
    """Default fused schedule: 2-D with FP32 storage recomputes the halo rings (path 2),
    2-D with FP64 storage and 3-D store face values (path 1); options 2 and 3 force one."""
    from dugks_amd import hip
    c2 = _cases().cavity_case(20, "NC38", 10, 4.0, 0.1)
    c3 = _cases().sphere_case(SMALL_SPHERE, "GH", 6, 1.0, 0.1)
    expect = [(c2, hip.MIXED, 1, 2), (c2, hip.FP64, 1, 1), (c3, hip.MIXED, 1, 1), (c3, hip.FP64, 1, 1),
              (c2, hip.MIXED, 2, 1), (c2, hip.FP64, 3, 2)]
    for case, prec, fused, path in expect:
        s = hip.FvDVM(case, precision=prec, fused=fused)
        assert s.path_info()["fused"] == path, (prec, fused)
        s.close()


SMALL_SPHERE = dict(nc=4, nr=2, n_up=2, n_down=3, n_side=2)


@pytest.mark.parametrize("fused", [0, 1])
def test_small_sphere_fp64_matches_oracle(fused):
    """3-D hex mesh with a diffuse sphere wall and mixed inlet/outlet/side patches."""
    c = _cases().sphere_case(SMALL_SPHERE, "GH", 6, 1.0, 0.1)
    o, g, ho, hg = _run_pair(c, 15, 0, fused)
    e = _errs(o, g)
    print("small sphere fp64", fused, c.n_cells, e)
    for k, v in e.items():
        assert v < 1e-12, (k, v)
    assert field_rel_err(g.boundary_rho(), o.wall_rho()) < 1e-12


def _three_kernel_form(g):
    """True when the handle's step ran the f-bar-plus kernel (the 3-D FP32 form)."""
    g.profile(True)
    g.run(1)
    names = set(g.kernel_times())
    g.profile(False)
    return "fused_bar" in names


# fused 1: 3-D FP32 storage takes the three-kernel form (k_bar3, pass A gathering f-bar-plus,
# the 512-thread pass B); fused 4: the two-pass form it falls back to without memory
@pytest.mark.parametrize("fused", [1, 4])
def test_small_sphere_mixed_matches_oracle(fused):
    c = _cases().sphere_case(SMALL_SPHERE, "GH", 6, 1.0, 0.1)
    o, g, ho, hg = _run_pair(c, 15, 1, fused)
    e = _errs(o, g)
    print("small sphere mixed", fused, e)
    for k in ("rho", "U", "T"):
        assert e[k] < 1e-5, (k, e[k])
    assert _three_kernel_form(g) == (fused == 1)


@pytest.mark.parametrize("fused", [1, 4])
def test_small_sphere_odd_velocity_grid_mixed_matches_oracle(fused):
    """3-D with an odd Simpson 3/8 set (xi_d = 0 on every axis): velocities tangent to the
    axis-aligned box faces and to the O-grid shell faces reach the FP32 tangent branch of
    the face-value pass A; 7 abscissae make a short last stripe in every chunked array."""
    c = _cases().sphere_case(SMALL_SPHERE, "NC38", 7, 4.0, 0.1)
    o, g, ho, hg = _run_pair(c, 10, 1, fused)
    e = _errs(o, g)
    print("small sphere odd grid mixed", e)
    for k in ("rho", "U", "T"):
        assert e[k] < 1e-5, (k, e[k])
    assert e["q"] < 1e-3


@pytest.mark.parametrize("fused", [0, 1])
def test_high_mach_sphere_fp64_matches_oracle(fused):
    """Ma = 5 freestream (U = 6.45 sqrt(R T0)) past the small sphere, 22^3 trapezoid set on
    [-14, 14]: a bow shock forms at the wall and q reaches ~20 while rho U^3 ~ 270, so the
    raw power sums the passes accumulate (sum w g xi^3) cancel strongly when q is formed
    about U.  FP64 GPU vs oracle after 5 steps: <= 1e-10 (max-norm relative) in rho, U, T, q,
    cell and face fields, and the wall density."""
    c = _cases().sphere_case(SMALL_SPHERE, "trapezoid", 22, 14.0, 0.1, Ma=5.0)
    o, g, ho, hg = _run_pair(c, 5, 0, fused)
    e = _errs(o, g)
    fo, fg = o.face_macros(), g.face_macros()
    ef = {k: field_rel_err(fg[k], fo[k]) for k in ("rho", "U", "T", "q")}
    print("high-Ma sphere fp64", fused, e, ef, flush=True)
    assert np.max(np.abs(o.cell_macros()["q"])) > 10.0
    for k in ("rho", "U", "T", "q"):
        assert e[k] < 1e-10, (k, e[k])
        assert ef[k] < 1e-10, (k, ef[k])
    assert field_rel_err(g.boundary_rho(), o.wall_rho()) < 1e-10


def test_full_size_cavity_matches_oracle():
    """The 256^2 cavity with the 64^2 trapezoid set on [-5, 5] (Kn 0.1) at full size, 3 steps
    from the same state: both default schedules (FP64: stored face values, 1e-10; FP32
    storage: recomputed rings, 1e-5 in rho, U, T) against one oracle run, time steps
    included."""
    from dugks_amd import hip
    from oracle import Oracle
    c = _cases().config1(0.1)
    o = Oracle(c)
    ho = []
    for _ in range(3):
        o.step()
        ho.append(o.dt)
    ho = np.array(ho)
    mo, fo = o.cell_macros(), o.face_macros()
    wo = o.wall_rho()
    del o
    for prec, tol in ((hip.FP64, 1e-10), (hip.MIXED, 1e-5)):
        g = hip.FvDVM(c, precision=prec)
        hg = g.run(3, record_dt=True)
        mg, fg = g.cell_macros(), g.face_macros()
        e = {k: field_rel_err(mg[k], mo[k]) for k in ("rho", "U", "T", "q")}
        ef = {k: field_rel_err(fg[k], fo[k]) for k in ("rho", "U", "T", "q")}
        print("full-size cavity", prec, g.path_info()["fused"], e, ef, flush=True)
        assert np.max(np.abs(hg - ho) / ho) < tol
        for k in ("rho", "U", "T"):
            assert e[k] < tol, (prec, k, e[k])
            assert ef[k] < tol, (prec, k, ef[k])
        if prec == hip.FP64:
            assert e["q"] < tol and ef["q"] < tol
            assert field_rel_err(g.boundary_rho(), wo) < tol
        g.close()


def test_courant_limit_for_long_runs():
    """dugksFoam's Courant rule with maxCo 0.8 lets diagonal velocities exceed the unsplit
    reconstruction's stability bound (sum_d |xi_d| dt/dx <= 1): a 128^2 cavity with the
    64^2 trapezoid set on [-5, 5] diverges within 300 steps, the same run at maxCo 0.65
    (the limit used for the long synthetic runs) stays bounded.  Both precisions agree."""
    from dugks_amd import cases, hip
    out = {}
    for co in (0.8, cases.STABLE_MAX_CO[2]):
        c = cases.cavity_case(128, "trapezoid", 64, 5.0, 0.1, controls=cases.Controls(maxCo=co))
        s = hip.FvDVM(c, precision=hip.FP64)
        s.run(300)
        m = s.cell_macros()
        s.close()
        out[co] = m
    assert not np.all(np.isfinite(out[0.8]["rho"]))
    m = out[cases.STABLE_MAX_CO[2]]
    assert np.all(np.isfinite(m["rho"])) and 0.8 < m["rho"].min() and m["rho"].max() < 1.25
    assert 0.95 < m["T"].min() and m["T"].max() < 1.05


@pytest.mark.parametrize("fused", [0, 2, 3])
def test_mixed_storage_drift_stays_small(fused):
    """FP32 storage holds deviations from a per-velocity reference Maxwellian, so rounding
    does not accumulate into a drift of the slowly varying fields: a 64^2 cavity (32^2
    trapezoid set on [-5, 5], Kn 0.1) run for 1000 steps with the FP64 run's time steps
    stays within 1e-6 of FP64 in rho, U, T on the staged path and both fused schedules.
    (Rounding the full values gave 5e-5 in U after 2000 steps of the 256^2 cavity.)"""
    from dugks_amd import cases, hip
    c = cases.cavity_case(64, "trapezoid", 32, 5.0, 0.1, controls=cases.Controls(maxCo=cases.STABLE_MAX_CO[2]))
    f64 = hip.FvDVM(c, precision=hip.FP64, fused=fused)
    dts = f64.run(1000, record_dt=True)
    m32 = hip.FvDVM(c, precision=hip.MIXED, fused=fused)
    m32.run_fixed(dts)
    a, b = m32.cell_macros(), f64.cell_macros()
    e = {k: field_rel_err(a[k], b[k]) for k in ("rho", "U", "T", "q")}
    print("mixed drift", fused, e)
    for k in ("rho", "U", "T"):
        assert e[k] < 1e-6, (k, e[k])


def test_checkpoint_is_in_caller_cell_order_and_moves_between_paths(tmp_path):
    """A checkpoint holds g~ in the caller's cell order (the RCB block numbering inside a
    handle depends on its path and schedule), so a file saved by the staged path loads into
    both fused schedules with the same fields, and the runs continue alike; a mesh with the
    same counts but other cells is refused."""
    from dugks_amd import hip, mesh as meshmod
    from test_oracle_pins import _distorted_cavity
    c = _cases().cavity_case(36, "NC38", 10, 4.0, 0.1)
    a = hip.FvDVM(c, fused=0)
    a.run(6)
    path = tmp_path / "ck.bin"
    a.checkpoint_save(path)
    g0, h0 = a.df()
    m0 = a.cell_macros()
    nC, Nv = a.nC, a.Nv
    raw = np.fromfile(path, dtype=np.float64, count=2 * Nv * nC, offset=64).reshape(2, Nv, nC)
    assert np.array_equal(raw[0], g0) and np.array_equal(raw[1], h0)
    a.dt_saved = a.dt
    a.run(4)
    ma = a.cell_macros()
    for fused in (2, 3):
        b = hip.FvDVM(c, fused=fused)
        b.checkpoint_load(path)
        gb, hb = b.df()
        assert np.array_equal(gb, g0) and np.array_equal(hb, h0)
        mb = b.cell_macros()
        for k in m0:
            assert np.array_equal(mb[k], m0[k]), k
        b.dt = a.dt_saved
        b.run(4)
        mb = b.cell_macros()
        for k in ("rho", "U", "T", "q"):
            assert field_rel_err(mb[k], ma[k]) < 1e-12, (fused, k)
        b.close()
    c2 = _cases().cavity_case(36, "NC38", 10, 4.0, 0.1)
    c2.mesh = _distorted_cavity(36, seed=2)
    c2.geo = meshmod.geometry(c2.mesh)
    other = hip.FvDVM(c2, fused=0)
    with pytest.raises(hip.DugksError, match="mesh"):
        other.checkpoint_load(path)


def test_checkpoint_refuses_other_storage_reference(tmp_path):
    """A mixed-precision checkpoint holds deviations from the handle's reference Maxwellian
    (T = Tref); a handle with another Tref must refuse it."""
    from dugks_amd import cases, hip
    c1 = cases.cavity_case(16, "NC38", 10, 4.0, 0.1)
    gas2 = cases.Gas(R=c1.gas.R, omega=c1.gas.omega, Tref=1.1, muRef=c1.gas.muRef, Pr=c1.gas.Pr,
                     KInner=c1.gas.KInner)
    c2 = cases.cavity_case(16, "NC38", 10, 4.0, 0.1, gas=gas2)
    a = hip.FvDVM(c1, precision=hip.MIXED)
    a.run(3)
    path = tmp_path / "ck.bin"
    a.checkpoint_save(path)
    same = hip.FvDVM(c1, precision=hip.MIXED)
    same.checkpoint_load(path)
    other = hip.FvDVM(c2, precision=hip.MIXED)
    with pytest.raises(hip.DugksError, match="reference"):
        other.checkpoint_load(path)


def _boundary_cases():
    from dugks_amd import cases
    U = (0.15, 0.05, 0.0)
    wall = ("calculatedMaxwell", 1.0, (0.0, 0.0, 0.0), 1.0)
    zg = ("zeroGradient", 1.0, U, 1.0)
    bumpy = dict(rho0=lambda C: 1.0 + 0.1 * np.sin(3 * C[:, 0]) * np.cos(2 * C[:, 1]),
                 U0=lambda C: np.stack([0.1 + 0.05 * C[:, 1], -0.05 * C[:, 0], 0 * C[:, 0]], 1),
                 T0=lambda C: 1.0 + 0.05 * C[:, 0] * C[:, 1], rule="NC38", n_dv=13, vel_extent=4.0)
    return {
        "farField_zeroGradient": cases.box_case(10, 8, dict(left=("farField", 1.0, (9.0, 0, 0), 1.0),
                                                            right=("farField", 0.9, U, 1.1), bottom=zg, top=zg),
                                                **bumpy),
        "pressure": cases.box_case(12, 6, dict(left=("pressureIn", 1.0, U, 1.1, 1.05),
                                               right=("pressureOut", 1.0, U, 1.0, 0.95), bottom=wall, top=wall),
                                   extent=(0.0, 2.0, 0.0, 1.0), **bumpy),
        "symmetryMod": cases.box_case(6, 8, dict(left=wall, right=("symmetryMod", 1.0, U, 1.0), bottom=wall,
                                                 top=("calculatedMaxwell", 1.0, (0.0, 0.2, 0.0), 1.0)),
                                      extent=(-1.0, 0.0, 0.0, 1.0), **bumpy),
        "symmetryPlane": cases.box_case(6, 8, dict(left=wall, right=("symmetryPlane", 1.0, U, 1.0), bottom=wall,
                                                   top=("calculatedMaxwell", 1.0, (0.0, 0.2, 0.0), 1.0)),
                                        extent=(-1.0, 0.0, 0.0, 1.0), **bumpy),
    }


@pytest.mark.parametrize("kind", ["farField_zeroGradient", "pressure", "symmetryMod", "symmetryPlane"])
def test_boundary_types_match_oracle(kind):
    """zeroGradient, farField, pressureIn / pressureOut and both symmetry types on the staged
    path (the fused schedules carry walls and fixedValue patches only and refuse these; with
    fused = -1 the handle falls back to the staged path): FP64 within 1e-12 of the oracle
    after 20 steps in cell and face fields and boundary rho, mixed storage within 1e-5."""
    from dugks_amd import hip
    c = _boundary_cases()[kind]
    with pytest.raises(hip.DugksError, match="fused path carries"):
        hip.FvDVM(c, fused=1)
    o, g, ho, hg = _run_pair(c, 20, 0, -1)
    assert g.path_info()["fused"] == 0
    e = _errs(o, g)
    fo, fg = o.face_macros(), g.face_macros()
    ef = {k: field_rel_err(fg[k], fo[k]) for k in ("rho", "U", "T", "q")}
    print("boundary", kind, e, ef, flush=True)
    assert np.max(np.abs(hg - ho) / ho) < 1e-12
    for k in ("rho", "U", "T", "q"):
        assert e[k] < 1e-12 and ef[k] < 1e-12, k
    assert field_rel_err(g.boundary_rho(), o.wall_rho()) < 1e-12
    o2, g2, _, _ = _run_pair(c, 20, 1, -1)
    e2 = _errs(o2, g2)
    for k in ("rho", "U", "T"):
        assert e2[k] < 1e-5, (k, e2[k])


@pytest.mark.parametrize("fused", [0, 2, 3])
def test_gauss_linear_gradient_matches_oracle(fused):
    """Gauss linear gradients (demo/ball's scheme) reach every path as vectors of the
    least-squares form: FP64 within 1e-12 of the oracle after 30 steps on a 20x20 cavity."""
    c = _cases().cavity_case(20, "NC38", 10, 4.0, 0.1)
    c.grad_scheme = "Gauss linear"
    o, g, ho, hg = _run_pair(c, 30, 0, fused)
    e = _errs(o, g)
    print("gauss linear", fused, e)
    for k, v in e.items():
        assert v < 1e-12, (k, v)
