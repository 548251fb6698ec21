"""Synthetic 3-D mesh for flow past a sphere: a body-fitted hexahedral O-grid shell around a
sphere of diameter D, inside a graded Cartesian hex box 10D x 6D x 6D (sphere centre 3D
behind the inlet).  All cells are hexahedra; the mesh is unstructured in the polyMesh sense
(owner/neighbour addressing, seven zones glued along non-matching block orientations).

Layout: a cube of half width ``a`` around the sphere is cut out of the Cartesian box; its
six faces carry an ``nc`` x ``nc`` grid shared with the box; the shell between the cube and
the sphere has ``nr`` layers clustered geometrically toward the sphere.  Box lines grow
geometrically away from the cube.  Patches: ``sphere`` (wall), ``inlet`` (x = xmin),
``outlet`` (x = xmax), ``sides`` (y, z = +-6D/2).
"""
from __future__ import annotations

import numpy as np

from . import mesh as meshmod


def _graded(L: float, h0: float, n: int) -> np.ndarray:
    """n cell widths summing to L, geometric, first width as close to h0 as allowed."""
    if n * h0 >= L:
        return np.full(n, L / n)
    lo, hi = 1.0, 2.0
    while h0 * (hi ** n - 1) / (hi - 1) < L:
        hi *= 2
    for _ in range(200):
        r = 0.5 * (lo + hi)
        if h0 * (r ** n - 1) / (r - 1) < L:
            lo = r
        else:
            hi = r
    r = 0.5 * (lo + hi)
    w = h0 * r ** np.arange(n)
    return w * (L / w.sum())


def _lines(lo: float, hi: float, a: float, nc: int, n_lo: int, n_hi: int) -> np.ndarray:
    core = np.linspace(-a, a, nc + 1)
    h = 2 * a / nc
    left = -a - np.cumsum(_graded(-a - lo, h, n_lo))[::-1]
    right = a + np.cumsum(_graded(hi - a, h, n_hi))
    left[0] = lo
    right[-1] = hi
    return np.concatenate([left, core, right])


def sphere_box_mesh(D: float = 1.0, nc: int = 24, nr: int = 12, a_over_D: float = 1.0,
                    n_up: int = 14, n_down: int = 30, n_side: int = 14, shell_ratio: float = 1.12,
                    box=(-3.0, 7.0, -3.0, 3.0, -3.0, 3.0)) -> meshmod.PolyMesh:
    r0 = 0.5 * D
    a = a_over_D * D
    x0, x1, y0, y1, z0, z1 = [b * D for b in box]
    X = _lines(x0, x1, a, nc, n_up, n_down)
    Y = _lines(y0, y1, a, nc, n_side, n_side)
    Z = _lines(z0, z1, a, nc, n_side, n_side)
    core = np.linspace(-a, a, nc + 1)

    pts = []
    # Cartesian box points
    zz, yy, xx = np.meshgrid(Z, Y, X, indexing="ij")
    box_pts = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
    nx, ny, nz = len(X) - 1, len(Y) - 1, len(Z) - 1
    pts.append(box_pts)

    def bid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    # Cartesian cells outside the cube
    k_, j_, i_ = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    i_, j_, k_ = i_.ravel(), j_.ravel(), k_.ravel()
    cx = 0.5 * (X[i_] + X[i_ + 1])
    cy = 0.5 * (Y[j_] + Y[j_ + 1])
    cz = 0.5 * (Z[k_] + Z[k_ + 1])
    keep = ~((np.abs(cx) < a) & (np.abs(cy) < a) & (np.abs(cz) < a))
    i_, j_, k_ = i_[keep], j_[keep], k_[keep]
    box_cells = np.stack([bid(i_, j_, k_), bid(i_ + 1, j_, k_), bid(i_ + 1, j_ + 1, k_), bid(i_, j_ + 1, k_),
                          bid(i_, j_, k_ + 1), bid(i_ + 1, j_, k_ + 1), bid(i_ + 1, j_ + 1, k_ + 1),
                          bid(i_, j_ + 1, k_ + 1)], 1)
    cells = [box_cells]
    offset = len(box_pts)

    # shell: six cube faces, (nc+1)^2 points per layer, nr+1 layers
    s = (shell_ratio ** np.arange(nr + 1) - 1) / (shell_ratio ** nr - 1)
    vv, uu = np.meshgrid(core, core, indexing="ij")
    u, v = uu.ravel(), vv.ravel()
    faces = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            P = np.zeros((len(u), 3))
            P[:, axis] = sign * a
            P[:, (axis + 1) % 3] = u
            P[:, (axis + 2) % 3] = v
            faces.append(P)
    m = nc + 1
    for P in faces:
        Ps = r0 * P / np.linalg.norm(P, axis=1)[:, None]
        layer = Ps[None, :, :] + s[:, None, None] * (P - Ps)[None, :, :]     # [nr+1, m*m, 3]
        pts.append(layer.reshape(-1, 3))
        kk, jj, ii = np.meshgrid(np.arange(nr), np.arange(nc), np.arange(nc), indexing="ij")
        ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
        sid = lambda i, j, k: offset + k * m * m + j * m + i
        cells.append(np.stack([sid(ii, jj, kk), sid(ii + 1, jj, kk), sid(ii + 1, jj + 1, kk), sid(ii, jj + 1, kk),
                               sid(ii, jj, kk + 1), sid(ii + 1, jj, kk + 1), sid(ii + 1, jj + 1, kk + 1),
                               sid(ii, jj + 1, kk + 1)], 1))
        offset += (nr + 1) * m * m
    allp = np.concatenate(pts)
    cells = np.concatenate(cells)
    # merge coincident points (cube surface shared by box and shell, shell block edges)
    scale = max(abs(x0), abs(x1), abs(y0), abs(y1), abs(z0), abs(z1))
    q = np.round(allp / scale * 1e11).astype(np.int64)
    _, first, inv = np.unique(q, axis=0, return_index=True, return_inverse=True)
    inv = inv.ravel()
    points = allp[first]
    cells = inv[cells]

    tol = 1e-9 * scale

    def classify(fx):
        r = np.linalg.norm(fx, axis=2)
        out = np.full(len(fx), -1)
        out[np.all(np.abs(r - r0) < tol, axis=1)] = 0
        out[np.all(np.abs(fx[:, :, 0] - x0) < tol, axis=1)] = 1
        out[np.all(np.abs(fx[:, :, 0] - x1) < tol, axis=1)] = 2
        side = (np.all(np.abs(fx[:, :, 1] - y0) < tol, axis=1) | np.all(np.abs(fx[:, :, 1] - y1) < tol, axis=1) |
                np.all(np.abs(fx[:, :, 2] - z0) < tol, axis=1) | np.all(np.abs(fx[:, :, 2] - z1) < tol, axis=1))
        out[side & (out < 0)] = 3
        return out

    return meshmod.hex_mesh(points, cells, ["sphere", "inlet", "outlet", "sides"], ["wall", "patch", "patch", "patch"],
                            classify)
