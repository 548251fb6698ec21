"""Gradients of wall heat flux and drag through N DUGKS steps (dugks_adjoint of the C-ABI).

Two problems, each a map theta -> J with its gradient from one forward + reverse sweep:

* :class:`CavityHeatFlux`: lid-driven cavity; J = integrated heat flux q.S over the bottom
  wall faces after N steps; theta = [lid speed, Kn, T_w of each bottom-wall face].
* :class:`SphereDrag`: flow past a sphere; J = drag coefficient from the momentum flux
  through the sphere faces, F_x / (rho U^2 A / 2), after N steps; theta = [Kn, T_w of each
  of n_patch seeded surface patches].

The time steps are fixed (taken from an adaptive run at the nominal theta unless given),
so J is a smooth function of theta and central finite differences of the same forward
check the gradient.  :func:`objective` wraps a problem as a ``torch.autograd.Function``
on CPU tensors (the DUGKS state lives in the library, not in torch).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import cases, hip


def bottom_wall_q_weight(case) -> np.ndarray:
    """[nF, 3]: S_f on the bottom wall faces of a cavity case (heat flux q.S), 0 elsewhere."""
    a = case.arrays()
    nIF, nB = a["nIF"], a["nB"]
    w = np.zeros((nIF + nB, 3))
    Cf, Sf = a["Cf"], a["Sf"]
    bot = nIF + np.nonzero(np.abs(Cf[nIF:, 1]) < 1e-12)[0]
    w[bot] = Sf[bot]
    return w


def face_ids_of_patch(case, name: str) -> np.ndarray:
    """Face numbers (C-ABI numbering) of a patch."""
    a = case.arrays()
    k = a["patch_names"].index(name)
    s, n = int(a["patch_start"][k]), int(a["patch_size"][k])
    return a["nIF"] + s + np.arange(n)


def drag_mom_weight(case, patch: str = "sphere", direction=(1.0, 0.0, 0.0), D: float = 1.0) -> np.ndarray:
    """[nF, 3]: e / (rho_inf U_inf^2 A / 2) on the patch faces, A = pi D^2 / 4.

    The momentum flux sum_v w xi (xi.S) g through a wall face (S pointing out of the gas)
    is the force of the gas on that face, so the objective term is the drag coefficient."""
    a = case.arrays()
    U = case.meta["U_inf"]
    qdyn = 0.5 * 1.0 * U * U * math.pi * D * D / 4.0
    w = np.zeros((a["nIF"] + a["nB"], 3))
    w[face_ids_of_patch(case, patch)] = np.asarray(direction, float) / qdyn
    return w


def seeded_surface_patches(case, n_patch: int, patch: str = "sphere", seed: int = 2024) -> np.ndarray:
    """Label in [0, n_patch) of each face of a wall patch: nearest of n_patch seeded points
    on the unit sphere around the patch centroid (directions of the face centres)."""
    a = case.arrays()
    ids = face_ids_of_patch(case, patch)
    Cf = a["Cf"][ids]
    d = Cf - Cf.mean(0)
    d /= np.linalg.norm(d, axis=1)[:, None]
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n_patch, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    return np.argmax(d @ p.T, axis=1)


@dataclasses.dataclass
class AdjointResult:
    J: float
    grad: np.ndarray
    raw: dict


class CavityHeatFlux:
    """theta = [lid_speed, Kn, T_bottom[N]] for an N x N cavity (trapezoid set on [-4, 4])."""

    def __init__(self, N: int = 100, n_dv: int = 32, steps: int = 500, checkpoint_every: int = 50,
                 precision: int = hip.FP64, rule: str = "trapezoid", vel_extent: float = 4.0, dts=None):
        self.N, self.n_dv, self.steps, self.ckpt = N, n_dv, steps, checkpoint_every
        self.precision, self.rule, self.vel_extent = precision, rule, vel_extent
        self.dts = None if dts is None else np.asarray(dts, float)

    def nominal(self) -> np.ndarray:
        return np.concatenate([[0.16 * math.sqrt(2.0)], [0.1], cases.smoothed_wall_profile(self.N)])

    def case(self, theta):
        theta = np.asarray(theta, float)
        return cases.cavity_case(self.N, self.rule, self.n_dv, self.vel_extent, float(theta[1]),
                                 lid_speed=float(theta[0]), wall_T=theta[2:2 + self.N],
                                 controls=cases.Controls(maxCo=cases.STABLE_MAX_CO[2]))

    def time_steps(self, theta):
        if self.dts is None:
            s = hip.FvDVM(self.case(theta), precision=hip.FP64)
            self.dts = s.run(self.steps, record_dt=True)
            s.close()
        return self.dts

    def value(self, theta, precision=None) -> float:
        """J(theta) from the forward only (same fixed time steps)."""
        theta = np.asarray(theta, float)
        dts = self.time_steps(theta)
        c = self.case(theta)
        s = hip.FvDVM(c, precision=self.precision if precision is None else precision)
        J = s.objective(dts, q_weight=bottom_wall_q_weight(c))
        s.close()
        return J

    def evaluate(self, theta) -> AdjointResult:
        theta = np.asarray(theta, float)
        dts = self.time_steps(theta)
        c = self.case(theta)
        s = hip.FvDVM(c, precision=self.precision)
        r = s.adjoint(dts, self.ckpt, q_weight=bottom_wall_q_weight(c))
        s.close()
        lid = face_ids_of_patch(c, "movingWall") - c.mesh.n_internal_faces
        fw = face_ids_of_patch(c, "fixedWalls") - c.mesh.n_internal_faces
        bottom = fw[self.N:2 * self.N]          # cavity_case puts the bottom wall in the middle third
        grad = np.concatenate([[r["dUw"][lid, 0].sum()],
                               [r["dmuRef"] * c.gas.muRef / float(theta[1])],   # muRef is linear in Kn
                               r["dTw"][bottom]])
        return AdjointResult(r["J"], grad, r)


class SphereDrag:
    """theta = [Kn, T_patch[n_patch]] for the sphere (half-range Gauss-Hermite set)."""

    def __init__(self, mesh_kw=None, n_dv: int = 16, steps: int = 300, checkpoint_every: int = 30,
                 precision: int = hip.MIXED, n_patch: int = 64, dts=None):
        self.mesh_kw = mesh_kw or cases.SPHERE_50K
        self.n_dv, self.steps, self.ckpt, self.precision = n_dv, steps, checkpoint_every, precision
        self.n_patch = n_patch
        self.dts = None if dts is None else np.asarray(dts, float)
        self._labels = None

    def nominal(self) -> np.ndarray:
        return np.concatenate([[0.1], np.ones(self.n_patch)])

    def case(self, theta):
        theta = np.asarray(theta, float)
        base = cases.sphere_case(self.mesh_kw, "GH", self.n_dv, 1.0, float(theta[0]))
        if self._labels is None:
            self._labels = seeded_surface_patches(base, self.n_patch)
        return cases.sphere_case(self.mesh_kw, "GH", self.n_dv, 1.0, float(theta[0]),
                                 wall_T=theta[1:][self._labels])

    def time_steps(self, theta):
        if self.dts is None:
            s = hip.FvDVM(self.case(theta), precision=hip.FP64)
            self.dts = s.run(self.steps, record_dt=True)
            s.close()
        return self.dts

    def value(self, theta, precision=None) -> float:
        """J(theta) from the forward only (same fixed time steps)."""
        theta = np.asarray(theta, float)
        dts = self.time_steps(theta)
        c = self.case(theta)
        s = hip.FvDVM(c, precision=self.precision if precision is None else precision)
        J = s.objective(dts, mom_weight=drag_mom_weight(c))
        s.close()
        return J

    def evaluate(self, theta) -> AdjointResult:
        theta = np.asarray(theta, float)
        dts = self.time_steps(theta)
        c = self.case(theta)
        s = hip.FvDVM(c, precision=self.precision)
        r = s.adjoint(dts, self.ckpt, mom_weight=drag_mom_weight(c))
        s.close()
        wall = face_ids_of_patch(c, "sphere") - c.mesh.n_internal_faces
        gT = np.bincount(self._labels, weights=r["dTw"][wall], minlength=self.n_patch)
        grad = np.concatenate([[r["dmuRef"] * c.gas.muRef / float(theta[0])], gT])
        return AdjointResult(r["J"], grad, r)


def objective(problem, theta):
    """J(theta) as a differentiable torch scalar (CPU, float64)."""
    import torch

    class _Objective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, th):
            res = problem.evaluate(th.detach().cpu().numpy())
            ctx.save_for_backward(torch.as_tensor(res.grad, dtype=th.dtype))
            return th.new_tensor(res.J)

        @staticmethod
        def backward(ctx, gJ):
            (grad,) = ctx.saved_tensors
            return gJ * grad

    return _Objective.apply(theta)
