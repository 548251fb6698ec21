"""Adjoint of N DUGKS steps (dugks_adjoint) against central finite differences.

The time steps are held fixed (recorded from an adaptive run), so J(theta) is smooth and
central differences of the same forward map (the GPU FP64 forward, and the CPU oracle
forward) check the hand-written reverse kernels:
  * gradients in muRef, the lid velocity and bottom-wall temperatures, 1e-6 relative;
  * the derivative along a random direction of the initial distributions (dot-product
    test through every reverse kernel, including the cell-moment path);
  * checkpoint spacing does not change the gradient;
  * 3-D sphere (mixed patches, diffuse wall, drag objective): FP64 adjoint against FD,
    and the mixed (FP32 storage) adjoint against the FP64 adjoint.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = 1e-6


def _mods():
    from dugks_amd import adjoint, cases, hip
    return adjoint, cases, hip


def _tiny_case(lid=0.16 * 2 ** 0.5, kn=0.1, Tb=None, N=6):
    _, cases, _ = _mods()
    if Tb is None:
        Tb = 1.0 + 0.05 * np.sin(np.arange(N))
    return cases.cavity_case(N, "NC38", 10, 4.0, kn, lid_speed=lid, wall_T=Tb)


def _dts(case, n):
    _, _, hip = _mods()
    s = hip.FvDVM(case, precision=hip.FP64)
    d = s.run(n, record_dt=True)
    s.close()
    return d


def _weights(case):
    adjoint, _, _ = _mods()
    qw = adjoint.bottom_wall_q_weight(case)
    mw = np.zeros_like(qw)
    mw[adjoint.face_ids_of_patch(case, "movingWall")] = [1.0, 0.3, 0.0]
    return qw, mw


def _J(case, dts, precision=0, g0=None):
    _, _, hip = _mods()
    qw, mw = _weights(case)
    s = hip.FvDVM(case, precision=precision)
    if g0 is not None:
        s.set_df(*g0)
    J = s.objective(dts, q_weight=qw, mom_weight=mw)
    s.close()
    return J


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_adjoint_matches_fd_tiny_cavity_fp64():
    adjoint, cases, hip = _mods()
    N = 6
    Tb = 1.0 + 0.05 * np.sin(np.arange(N))
    lid, kn = 0.16 * 2 ** 0.5, 0.1
    c = _tiny_case(lid, kn, Tb)
    dts = _dts(c, 5)
    qw, mw = _weights(c)
    s = hip.FvDVM(c, precision=hip.FP64)
    r = s.adjoint(dts, 2, q_weight=qw, mom_weight=mw)
    s.close()
    assert np.isfinite(r["J"]) and r["J"] != 0.0
    # muRef through Kn (muRef is linear in Kn)
    fd = (_J(_tiny_case(lid, kn * (1 + EPS), Tb), dts) - _J(_tiny_case(lid, kn * (1 - EPS), Tb), dts)) / (2 * EPS * kn)
    ad = r["dmuRef"] * c.gas.muRef / kn
    print("dJ/dKn", ad, fd)
    assert _rel(ad, fd) < 1e-6
    # lid speed: sum of dJ/dU_x over the lid faces
    lid_b = adjoint.face_ids_of_patch(c, "movingWall") - c.mesh.n_internal_faces
    h = EPS * lid
    fd = (_J(_tiny_case(lid + h, kn, Tb), dts) - _J(_tiny_case(lid - h, kn, Tb), dts)) / (2 * h)
    ad = r["dUw"][lid_b, 0].sum()
    print("dJ/dlid", ad, fd)
    assert _rel(ad, fd) < 1e-6
    # bottom wall temperatures
    fw_b = adjoint.face_ids_of_patch(c, "fixedWalls") - c.mesh.n_internal_faces
    for i in (0, 3):
        e = np.zeros(N)
        e[i] = EPS
        fd = (_J(_tiny_case(lid, kn, Tb + e), dts) - _J(_tiny_case(lid, kn, Tb - e), dts)) / (2 * EPS)
        ad = r["dTw"][fw_b[N + i]]
        print("dJ/dTw", i, ad, fd)
        assert _rel(ad, fd) < 1e-6


def test_adjoint_initial_state_direction_and_checkpoints():
    _, _, hip = _mods()
    c = _tiny_case()
    dts = _dts(c, 4)
    qw, mw = _weights(c)
    s = hip.FvDVM(c, precision=hip.FP64)
    g0, h0 = s.df()
    r1 = s.adjoint(dts, 1, q_weight=qw, mom_weight=mw, want_df0=True)
    s.close()
    assert r1["J"] == _J(c, dts)            # forward-only evaluation gives the same objective
    s = hip.FvDVM(c, precision=hip.FP64)
    r4 = s.adjoint(dts, 4, q_weight=qw, mom_weight=mw, want_df0=True)
    s.close()
    # checkpoint spacing only changes what is recomputed
    for k in ("J", "dmuRef"):
        assert r1[k] == r4[k], k
    assert np.array_equal(r1["dTw"], r4["dTw"]) and np.array_equal(r1["dg0"], r4["dg0"])
    rng = np.random.default_rng(7)
    dg = rng.standard_normal(g0.shape) * g0 * 1e-2
    dh = rng.standard_normal(h0.shape) * h0 * 1e-2
    t = 1e-4
    fd = (_J(c, dts, g0=(g0 + t * dg, h0 + t * dh)) - _J(c, dts, g0=(g0 - t * dg, h0 - t * dh))) / (2 * t)
    ad = float(np.sum(r1["dg0"] * dg) + np.sum(r1["dh0"] * dh))
    print("dJ/df0 . d", ad, fd)
    assert _rel(ad, fd) < 1e-6


def test_adjoint_matches_oracle_fd():
    """Central differences of the CPU oracle's forward (fixed dt) against the GPU adjoint."""
    from oracle import Oracle
    adjoint, cases, hip = _mods()
    N = 6
    Tb = 1.0 + 0.05 * np.sin(np.arange(N))
    lid, kn = 0.16 * 2 ** 0.5, 0.1
    c = _tiny_case(lid, kn, Tb)
    dts = _dts(c, 5)
    qw, mw = _weights(c)

    def J_oracle(case):
        o = Oracle(case)
        for dt in dts:
            o.evolution(float(dt))
        q = o.face_macros()["q"]
        gS, _ = o.surf()
        xi, w = o.vset()
        a = case.arrays()
        sig = xi @ a["Sf"].T                     # [Nv, nF]
        mom = np.einsum("v,vf,vf->f", w, xi @ mw.T, sig * gS)
        return float(np.sum(q * qw) + np.sum(mom))

    s = hip.FvDVM(c, precision=hip.FP64)
    r = s.adjoint(dts, 5, q_weight=qw, mom_weight=mw)
    s.close()
    assert _rel(r["J"], J_oracle(c)) < 1e-10
    fd = (J_oracle(_tiny_case(lid, kn * (1 + EPS), Tb)) - J_oracle(_tiny_case(lid, kn * (1 - EPS), Tb))) / (2 * EPS * kn)
    assert _rel(r["dmuRef"] * c.gas.muRef / kn, fd) < 1e-6
    fw_b = adjoint.face_ids_of_patch(c, "fixedWalls") - c.mesh.n_internal_faces
    e = np.zeros(N)
    e[2] = EPS
    fd = (J_oracle(_tiny_case(lid, kn, Tb + e)) - J_oracle(_tiny_case(lid, kn, Tb - e))) / (2 * EPS)
    assert _rel(r["dTw"][fw_b[N + 2]], fd) < 1e-6


def test_adjoint_small_sphere_drag():
    adjoint, cases, hip = _mods()
    kw = dict(nc=4, nr=2, n_up=2, n_down=3, n_side=2)
    prob64 = adjoint.SphereDrag(mesh_kw=kw, n_dv=6, steps=4, checkpoint_every=2, precision=hip.FP64, n_patch=4)
    th = prob64.nominal()
    th[1:] = 1.0 + 0.1 * np.arange(4)
    r64 = prob64.evaluate(th)
    assert np.isfinite(r64.J) and r64.J != 0.0
    # FD of the FP64 forward (same fixed time steps) in Kn and one patch temperature
    for i, h in ((0, EPS * th[0]), (2, EPS)):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (prob64.value(tp) - prob64.value(tm)) / (2 * h)
        print("sphere dJ/dtheta", i, r64.grad[i], fd)
        assert _rel(r64.grad[i], fd) < 1e-6
    # mixed (FP32 storage, FP64 adjoint arithmetic) against the FP64 adjoint
    probm = adjoint.SphereDrag(mesh_kw=kw, n_dv=6, steps=4, checkpoint_every=2, precision=hip.MIXED, n_patch=4,
                               dts=prob64.dts)
    rm = probm.evaluate(th)
    err = np.max(np.abs(rm.grad - r64.grad)) / np.max(np.abs(r64.grad))
    print("sphere mixed vs fp64 gradient", err)
    assert err < 1e-4


def test_torch_autograd_binding():
    torch = pytest.importorskip("torch")
    adjoint, _, hip = _mods()
    prob = adjoint.CavityHeatFlux(N=6, n_dv=10, steps=3, checkpoint_every=2, rule="NC38")
    th = torch.tensor(prob.nominal(), dtype=torch.float64, requires_grad=True)
    J = adjoint.objective(prob, th)
    J.backward()
    ref = prob.evaluate(th.detach().numpy())
    assert float(J) == ref.J
    assert np.array_equal(th.grad.numpy(), ref.grad)
