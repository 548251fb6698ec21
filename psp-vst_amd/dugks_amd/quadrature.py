"""One-dimensional velocity quadratures (the `constant/Xis` and `constant/weights` lists).

The DUGKS step takes a 1-D list of abscissae and weights and forms the tensor-product
velocity set from it (src/fvDVM/fvDVM/fvDVM.C:293-373, ix fastest).  The weights are full
quadrature measures: the reference multiplies them by `dXiCellSize_ = 1`
(src/fvDVM/fvDVM/fvDVM.C:1934-1939).

Rules provided here:

* ``half_range_gauss_hermite(C, N)`` -- the half-range Gauss-Hermite rule of
  src/scripts/setDV.py:83-112 (weight exp(-x^2) on [0, inf), mirrored), derived by
  Golub-Welsch from the three-term recurrence for that weight; weights are returned
  multiplied by exp(x^2)*C so that sum w f(xi) approximates the integral of f.
* ``boole_newton_cotes(xi_max, N)`` -- composite Boole rule, N = 4Z+1
  (src/scripts/setDV.py:114-138).
* ``simpson38_newton_cotes(xi_max, N)`` -- composite closed 4-point Newton-Cotes rule
  (Simpson 3/8), N = 3Z+1.  This is the 28-point Newton-Cotes set of the 40x40 cavity
  case (27 intervals = 9 panels).
* ``trapezoid(xi_max, N)`` -- uniform trapezoid rule on [-xi_max, xi_max].
* ``gauss_hermite(C, N)`` -- full-range Gauss-Hermite (weight exp(-x^2) on R), scaled
  the same way.

All return ``(xis, weights)`` as float64 numpy arrays in ascending order of xis.
"""
from __future__ import annotations

import math

import numpy as np


def _sorted(x: np.ndarray, w: np.ndarray):
    order = np.argsort(x, kind="stable")
    return np.ascontiguousarray(x[order]), np.ascontiguousarray(w[order])


def half_range_gauss_hermite(C: float, N: int):
    """Half-range Gauss-Hermite set of src/scripts/setDV.py:83-112.

    N must be even (N/2 nodes on each half line).  The recurrence coefficients of the
    polynomials orthogonal to exp(-x^2) on [0, inf) follow setDV.py:86-95; the nodes are
    the eigenvalues of the symmetric Jacobi matrix and the weights are
    (first eigenvector component)^2 * sqrt(pi)/2 (Golub-Welsch).
    """
    if N % 2 != 0 or N < 6:
        raise ValueError("half-range Gauss-Hermite needs an even N >= 6")
    n = N // 2
    a = np.zeros(n)
    b = np.zeros(n)
    a[0] = 1.0 / math.sqrt(math.pi)
    a[1] = 2.0 / math.sqrt(math.pi) / (math.pi - 2.0)
    b[1] = a[0] / (a[0] + a[1]) / 2.0
    for i in range(2, n):
        b[i] = (i - 1) + 0.5 - b[i - 1] - a[i - 1] ** 2
        a[i] = (i * i / 4.0 / b[i] - b[i - 1] - 0.5) / a[i - 1] - a[i - 1]
    J = np.diag(a) + np.diag(np.sqrt(b[1:]), 1) + np.diag(np.sqrt(b[1:]), -1)
    nodes, vecs = np.linalg.eigh(J)
    wts = vecs[0, :] ** 2 * math.sqrt(math.pi) / 2.0
    order = np.argsort(nodes)
    v = nodes[order]
    w = wts[order]
    xis = np.concatenate([-v[::-1], v])
    weights = np.concatenate([w[::-1], w])
    weights = weights * np.exp(xis ** 2) * C
    return _sorted(xis * C, weights)


def gauss_hermite(C: float, N: int):
    """Full-range Gauss-Hermite rule for weight exp(-x^2), scaled like setDV.py."""
    x, w = np.polynomial.hermite.hermgauss(N)
    return _sorted(x * C, w * np.exp(x ** 2) * C)


def boole_newton_cotes(xi_max: float, N: int):
    """Composite Boole rule on [-xi_max, xi_max], N = 4Z+1 (setDV.py:114-138)."""
    if N % 4 != 1 or N < 5:
        raise ValueError("composite Boole rule needs N = 4Z+1")
    h = 2.0 * xi_max / (N - 1)
    xis = -xi_max + h * np.arange(N)
    w = np.zeros(N)
    for p in range((N - 1) // 4):
        w[4 * p: 4 * p + 5] += np.array([7.0, 32.0, 12.0, 32.0, 7.0]) * (2.0 * h / 45.0)
    return xis, w


def simpson38_newton_cotes(xi_max: float, N: int):
    """Composite Simpson 3/8 (closed 4-point Newton-Cotes) on [-xi_max, xi_max], N = 3Z+1."""
    if N % 3 != 1 or N < 4:
        raise ValueError("composite Simpson 3/8 rule needs N = 3Z+1")
    h = 2.0 * xi_max / (N - 1)
    xis = -xi_max + h * np.arange(N)
    w = np.zeros(N)
    for p in range((N - 1) // 3):
        w[3 * p: 3 * p + 4] += np.array([1.0, 3.0, 3.0, 1.0]) * (3.0 * h / 8.0)
    return xis, w


def trapezoid(xi_max: float, N: int):
    """Uniform trapezoid rule with N nodes on [-xi_max, xi_max]."""
    if N < 2:
        raise ValueError("trapezoid rule needs N >= 2")
    h = 2.0 * xi_max / (N - 1)
    xis = -xi_max + h * np.arange(N)
    w = np.full(N, h)
    w[0] = w[-1] = 0.5 * h
    return xis, w


RULES = {
    "GH": half_range_gauss_hermite,
    "GHfull": gauss_hermite,
    "NC": boole_newton_cotes,
    "NC38": simpson38_newton_cotes,
    "trapezoid": trapezoid,
}


def make_rule(kind: str, scale: float, N: int):
    """Dispatch by name; ``scale`` is C for the Gauss-Hermite rules and xi_max otherwise."""
    try:
        fn = RULES[kind]
    except KeyError as exc:
        raise ValueError(f"unknown velocity quadrature {kind!r}; known: {sorted(RULES)}") from exc
    return fn(scale, N)


def tensor_product(xis: np.ndarray, weights: np.ndarray, D: int):
    """Tensor-product velocity set with ix fastest (src/fvDVM/fvDVM/fvDVM.C:293-373).

    Returns ``(xi[Nv,3], w[Nv])``; components beyond D are zero.
    """
    n = len(xis)
    if D == 3:
        iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
        xi = np.stack([xis[ix], xis[iy], xis[iz]], axis=1)
        w = weights[iz] * weights[iy] * weights[ix]
    elif D == 2:
        iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        ix, iy = ix.ravel(), iy.ravel()
        xi = np.stack([xis[ix], xis[iy], np.zeros(n * n)], axis=1)
        w = weights[iy] * weights[ix] * 1.0
    elif D == 1:
        xi = np.stack([xis, np.zeros(n), np.zeros(n)], axis=1)
        w = weights * 1.0 * 1.0
    else:
        raise ValueError("D must be 1, 2 or 3")
    return np.ascontiguousarray(xi), np.ascontiguousarray(w)
