"""Finite-volume mesh in OpenFOAM polyMesh form, its geometry, and the cavity generator.

The DUGKS step reads the mesh through OpenFOAM's fvMesh: owner/neighbour addressing,
face area vectors Sf, face centres Cf, cell centres C, cell volumes V, interpolation
weights and deltaCoeffs (src/fvDVM/fvDVM/fvDVM.C:757-762, 1473-1476, 2180-2182).  OpenFOAM
itself is not part of this build, so :func:`geometry` restates the primitiveMesh /
surfaceInterpolation formulas (face centres and areas by triangle fans about the point
average, cell centres and volumes by face pyramids about the face-centre average,
w = |Sf.(C_N - Cf)| / (|Sf.(Cf - C_P)| + |Sf.(C_N - Cf)|), deltaCoeffs = 1/|C_N - C_P|).

Face ordering follows the polyMesh convention: internal faces first in upper-triangular
order (sorted by owner, then neighbour), then boundary faces patch by patch.  Face normals
point out of the owner cell.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ROOTVSMALL = 1.0e-150
VSMALL = 1.0e-300


@dataclass
class Patch:
    name: str
    geometric_type: str  # "wall", "patch" or "empty"
    start: int
    size: int


@dataclass
class PolyMesh:
    points: np.ndarray                 # [nPoints, 3]
    face_offsets: np.ndarray           # [nFaces+1] CSR offsets into face_points
    face_points: np.ndarray            # point labels of all faces
    owner: np.ndarray                  # [nFaces]
    neighbour: np.ndarray              # [nInternalFaces]
    patches: list = field(default_factory=list)
    n_cells: int = 0

    @property
    def n_faces(self) -> int:
        return len(self.owner)

    @property
    def n_internal_faces(self) -> int:
        return len(self.neighbour)

    def face(self, i: int) -> np.ndarray:
        return self.face_points[self.face_offsets[i]:self.face_offsets[i + 1]]

    def patch(self, name: str) -> Patch:
        for p in self.patches:
            if p.name == name:
                return p
        raise KeyError(name)

    def n_solution_d(self) -> int:
        """Number of solved directions: 3 minus the directions spanned only by empty patches.

        Mirrors polyMesh::nSolutionD for meshes whose empty patches are planes normal to
        a coordinate axis (the cases handled here).
        """
        empty_dirs = set()
        for p in self.patches:
            if p.geometric_type != "empty" or p.size == 0:
                continue
            areas, _ = _face_centres_areas(self, range(p.start, p.start + p.size))
            n = np.abs(areas).sum(axis=0)
            empty_dirs.add(int(np.argmax(n)))
        return 3 - len(empty_dirs)


def _face_centres_areas(mesh: PolyMesh, face_ids):
    pts = mesh.points
    ctrs, areas = [], []
    for f in face_ids:
        fp = pts[mesh.face(f)]
        n = len(fp)
        if n == 3:
            ctrs.append((1.0 / 3.0) * (fp[0] + fp[1] + fp[2]))
            areas.append(0.5 * np.cross(fp[1] - fp[0], fp[2] - fp[0]))
            continue
        fc = fp.sum(axis=0) / n
        nxt = np.roll(fp, -1, axis=0)
        c = fp + nxt + fc
        nv = np.cross(nxt - fp, fc - fp)
        a = np.sqrt((nv * nv).sum(axis=1))
        sumA = a.sum()
        if sumA < ROOTVSMALL:
            ctrs.append(fc)
            areas.append(np.zeros(3))
        else:
            ctrs.append((1.0 / 3.0) * (a[:, None] * c).sum(axis=0) / sumA)
            areas.append(0.5 * nv.sum(axis=0))
    return np.array(areas).reshape(-1, 3), np.array(ctrs).reshape(-1, 3)


def _face_centres_areas_vectorised(mesh: PolyMesh):
    """Vectorised form of the triangle-fan face geometry, grouped by face size."""
    nF = mesh.n_faces
    sizes = np.diff(mesh.face_offsets)
    Sf = np.zeros((nF, 3))
    Cf = np.zeros((nF, 3))
    pts = mesh.points
    for n in np.unique(sizes):
        ids = np.nonzero(sizes == n)[0]
        idx = mesh.face_offsets[ids][:, None] + np.arange(n)[None, :]
        fp = pts[mesh.face_points[idx]]                       # [k, n, 3]
        if n == 3:
            Cf[ids] = (1.0 / 3.0) * (fp[:, 0] + fp[:, 1] + fp[:, 2])
            Sf[ids] = 0.5 * np.cross(fp[:, 1] - fp[:, 0], fp[:, 2] - fp[:, 0])
            continue
        fc = fp[:, 0].copy()
        for k in range(1, n):
            fc += fp[:, k]
        fc /= n
        sumN = np.zeros((len(ids), 3))
        sumA = np.zeros(len(ids))
        sumAc = np.zeros((len(ids), 3))
        for k in range(n):
            p = fp[:, k]
            q = fp[:, (k + 1) % n]
            c = p + q + fc
            nv = np.cross(q - p, fc - p)
            a = np.sqrt((nv * nv).sum(axis=1))
            sumN += nv
            sumA += a
            sumAc += a[:, None] * c
        small = sumA < ROOTVSMALL
        Cf[ids] = np.where(small[:, None], fc, (1.0 / 3.0) * sumAc / np.where(small, 1.0, sumA)[:, None])
        Sf[ids] = np.where(small[:, None], 0.0, 0.5 * sumN)
    return Sf, Cf


@dataclass
class Geometry:
    Sf: np.ndarray        # [nFaces, 3]
    Cf: np.ndarray        # [nFaces, 3]
    magSf: np.ndarray     # [nFaces]
    C: np.ndarray         # [nCells, 3]
    V: np.ndarray         # [nCells]
    weights: np.ndarray   # [nInternalFaces]
    delta_coeffs: np.ndarray  # [nInternalFaces]


def geometry(mesh: PolyMesh) -> Geometry:
    """Face/cell geometry with the primitiveMesh and surfaceInterpolation formulas."""
    Sf, Cf = _face_centres_areas_vectorised(mesh)
    own = mesh.owner
    nei = mesh.neighbour
    nC = mesh.n_cells
    nIF = mesh.n_internal_faces
    cEst = np.zeros((nC, 3))
    nCellFaces = np.zeros(nC)
    np.add.at(cEst, own, Cf)
    np.add.at(nCellFaces, own, 1.0)
    np.add.at(cEst, nei, Cf[:nIF])
    np.add.at(nCellFaces, nei, 1.0)
    cEst /= nCellFaces[:, None]

    pyr_own = (Sf * (Cf - cEst[own])).sum(axis=1)
    pc_own = (3.0 / 4.0) * Cf + (1.0 / 4.0) * cEst[own]
    pyr_nei = (Sf[:nIF] * (cEst[nei] - Cf[:nIF])).sum(axis=1)
    pc_nei = (3.0 / 4.0) * Cf[:nIF] + (1.0 / 4.0) * cEst[nei]
    cellCtrs = np.zeros((nC, 3))
    cellVols = np.zeros(nC)
    np.add.at(cellCtrs, own, pyr_own[:, None] * pc_own)
    np.add.at(cellVols, own, pyr_own)
    np.add.at(cellCtrs, nei, pyr_nei[:, None] * pc_nei)
    np.add.at(cellVols, nei, pyr_nei)
    ok = np.abs(cellVols) > VSMALL
    C = np.where(ok[:, None], cellCtrs / np.where(ok, cellVols, 1.0)[:, None], cEst)
    V = cellVols * (1.0 / 3.0)

    SfI = Sf[:nIF]
    CfI = Cf[:nIF]
    SfdOwn = np.abs((SfI * (CfI - C[own[:nIF]])).sum(axis=1))
    SfdNei = np.abs((SfI * (C[nei] - CfI)).sum(axis=1))
    w = SfdNei / (SfdOwn + SfdNei)
    d = C[nei] - C[own[:nIF]]
    delta_coeffs = 1.0 / np.sqrt((d * d).sum(axis=1))
    magSf = np.sqrt((Sf * Sf).sum(axis=1))
    return Geometry(Sf=Sf, Cf=Cf, magSf=magSf, C=C, V=V, weights=w, delta_coeffs=delta_coeffs)


# --------------------------------------------------------------------------- builders

def build_from_cell_faces(points, cell_faces_fn, n_cells, patch_specs):
    """Assemble a polyMesh from per-cell face lists.

    ``cell_faces_fn(c)`` yields point-label tuples of the faces of cell ``c``.  Shared
    faces are matched by their sorted point labels; internal faces are ordered upper-
    triangularly; unmatched faces are boundary faces assigned to patches by
    ``patch_specs`` = list of (name, geometric_type, predicate(face_points_xyz) -> bool),
    tested in order.  Face point order is fixed so that normals point out of the owner.
    """
    table = {}
    internal = []
    boundary = []
    for c in range(n_cells):
        for fpts in cell_faces_fn(c):
            key = tuple(sorted(fpts))
            hit = table.pop(key, None)
            if hit is None:
                table[key] = (c, tuple(fpts))
            else:
                c0, f0 = hit
                internal.append((min(c0, c), max(c0, c), f0 if c0 < c else tuple(fpts)))
    for key, (c, fpts) in table.items():
        boundary.append((c, fpts))
    internal.sort(key=lambda t: (t[0], t[1]))
    pts = np.asarray(points, dtype=np.float64)
    return _assemble(pts, internal, boundary, patch_specs, n_cells, cell_faces_fn)


def _assemble(pts, internal, boundary, patch_specs, n_cells, cell_faces_fn):
    # point-average cell centres, used only to orient faces
    cavg = np.zeros((n_cells, 3))
    for c in range(n_cells):
        labels = sorted({p for f in cell_faces_fn(c) for p in f})
        cavg[c] = pts[labels].mean(axis=0)

    def normal(fpts):
        fp = pts[list(fpts)]
        fc = fp.mean(axis=0)
        n = np.zeros(3)
        for k in range(len(fp)):
            n += np.cross(fp[(k + 1) % len(fp)] - fp[k], fc - fp[k])
        return n, fc

    face_list, owner, neighbour = [], [], []
    for own, nei, fpts in internal:
        n, fc = normal(fpts)
        if np.dot(n, cavg[nei] - cavg[own]) < 0:
            fpts = tuple(reversed(fpts))
        face_list.append(fpts)
        owner.append(own)
        neighbour.append(nei)
    patches = []
    groups = [[] for _ in patch_specs]
    for c, fpts in boundary:
        xyz = pts[list(fpts)]
        for k, (_, _, pred) in enumerate(patch_specs):
            if pred(xyz):
                groups[k].append((c, fpts))
                break
        else:
            raise ValueError(f"boundary face {fpts} of cell {c} matches no patch")
    start = len(face_list)
    for (name, gtype, _), grp in zip(patch_specs, groups):
        grp.sort(key=lambda t: (t[0], tuple(sorted(t[1]))))
        for c, fpts in grp:
            n, fc = normal(fpts)
            if np.dot(n, fc - cavg[c]) < 0:
                fpts = tuple(reversed(fpts))
            face_list.append(fpts)
            owner.append(c)
        patches.append(Patch(name, gtype, start, len(grp)))
        start += len(grp)
    offsets = np.zeros(len(face_list) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([len(f) for f in face_list])
    fpoints = np.array([p for f in face_list for p in f], dtype=np.int64)
    return PolyMesh(points=pts, face_offsets=offsets, face_points=fpoints,
                    owner=np.array(owner, dtype=np.int64),
                    neighbour=np.array(neighbour, dtype=np.int64),
                    patches=patches, n_cells=n_cells)


def cavity(N: int, L: float = 1.0, dz: float = 0.1) -> PolyMesh:
    """Uniform N x N x 1 hex mesh of [0,L]^2 x [0,dz] (the lid-driven cavity).

    Cells are numbered x fastest.  Patches, in order: ``movingWall`` (y = L),
    ``fixedWalls`` (x = 0, then y = 0, then x = L, mirroring the face order of
    demo/cavity/constant/polyMesh/blockMeshDict:53-92) and ``frontAndBack`` (empty).
    Built directly (not through the generic face matcher) so large N is fast.
    """
    n1 = N + 1
    h = L / N
    xs = np.arange(n1) * h
    xs[-1] = L
    k_, j_, i_ = np.meshgrid(np.arange(2), np.arange(n1), np.arange(n1), indexing="ij")
    pts = np.stack([xs[i_.ravel()], xs[j_.ravel()], np.where(k_.ravel() == 0, 0.0, dz)], axis=1)

    def pid(i, j, k):
        return i + n1 * (j + n1 * k)

    faces, owner, neighbour = [], [], []
    # internal faces, upper-triangular: for each cell, +x neighbour then +y neighbour
    for j in range(N):
        for i in range(N):
            c = i + N * j
            if i + 1 < N:   # x-face at i+1, normal +x
                faces.append((pid(i + 1, j, 0), pid(i + 1, j + 1, 0), pid(i + 1, j + 1, 1), pid(i + 1, j, 1)))
                owner.append(c)
                neighbour.append(c + 1)
            if j + 1 < N:   # y-face at j+1, normal +y
                faces.append((pid(i, j + 1, 0), pid(i, j + 1, 1), pid(i + 1, j + 1, 1), pid(i + 1, j + 1, 0)))
                owner.append(c)
                neighbour.append(c + N)
    patches = []
    start = len(faces)
    # movingWall: y = L, normal +y
    for i in range(N):
        c = i + N * (N - 1)
        faces.append((pid(i, N, 0), pid(i, N, 1), pid(i + 1, N, 1), pid(i + 1, N, 0)))
        owner.append(c)
    patches.append(Patch("movingWall", "wall", start, N))
    start += N
    # fixedWalls: x = 0 (normal -x), y = 0 (normal -y), x = L (normal +x)
    for j in range(N):
        faces.append((pid(0, j, 0), pid(0, j, 1), pid(0, j + 1, 1), pid(0, j + 1, 0)))
        owner.append(N * j)
    for i in range(N):
        faces.append((pid(i, 0, 0), pid(i + 1, 0, 0), pid(i + 1, 0, 1), pid(i, 0, 1)))
        owner.append(i)
    for j in range(N):
        faces.append((pid(N, j, 0), pid(N, j + 1, 0), pid(N, j + 1, 1), pid(N, j, 1)))
        owner.append(N - 1 + N * j)
    patches.append(Patch("fixedWalls", "wall", start, 3 * N))
    start += 3 * N
    # frontAndBack: z = 0 (normal -z), z = dz (normal +z)
    for j in range(N):
        for i in range(N):
            faces.append((pid(i, j, 0), pid(i, j + 1, 0), pid(i + 1, j + 1, 0), pid(i + 1, j, 0)))
            owner.append(i + N * j)
    for j in range(N):
        for i in range(N):
            faces.append((pid(i, j, 1), pid(i + 1, j, 1), pid(i + 1, j + 1, 1), pid(i, j + 1, 1)))
            owner.append(i + N * j)
    patches.append(Patch("frontAndBack", "empty", start, 2 * N * N))
    offsets = np.arange(len(faces) + 1, dtype=np.int64) * 4
    fpoints = np.array(faces, dtype=np.int64).ravel()
    return PolyMesh(points=pts, face_offsets=offsets, face_points=fpoints,
                    owner=np.array(owner, dtype=np.int64),
                    neighbour=np.array(neighbour, dtype=np.int64),
                    patches=patches, n_cells=N * N)


def box2d(nx: int, ny: int, x0: float = 0.0, x1: float = 1.0, y0: float = 0.0, y1: float = 1.0,
          dz: float = 0.1, types=None, grading: float = 1.0) -> PolyMesh:
    """nx x ny x 1 hex mesh of [x0,x1] x [y0,y1] x [0,dz] with the patches ``left``,
    ``right``, ``bottom``, ``top`` (geometric types from ``types``, default ``patch``) and
    ``frontAndBack`` (empty).  ``grading``: ratio of the last to the first cell width in x
    and in y (geometric progression; 1 = uniform)."""
    def edges(a, b, n):
        if grading == 1.0 or n < 2:
            return np.linspace(a, b, n + 1)
        r = grading ** (1.0 / (n - 1))
        w = r ** np.arange(n)
        return a + (b - a) * np.concatenate([[0.0], np.cumsum(w) / np.sum(w)])
    xs = edges(x0, x1, nx)
    ys = edges(y0, y1, ny)
    xs[-1], ys[-1] = x1, y1
    k_, j_, i_ = np.meshgrid(np.arange(2), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    pts = np.stack([xs[i_.ravel()], ys[j_.ravel()], np.where(k_.ravel() == 0, 0.0, dz)], axis=1)
    pid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    cells = []
    for j in range(ny):
        for i in range(nx):
            cells.append([pid(i, j, 0), pid(i + 1, j, 0), pid(i + 1, j + 1, 0), pid(i, j + 1, 0),
                          pid(i, j, 1), pid(i + 1, j, 1), pid(i + 1, j + 1, 1), pid(i, j + 1, 1)])
    names = ["left", "right", "bottom", "top", "frontAndBack"]
    types = dict(types or {})
    gtypes = [types.get(n, "patch") for n in names[:4]] + ["empty"]
    tol = 1e-12 * max(abs(x1 - x0), abs(y1 - y0))

    def classify(q):
        c = q.mean(axis=1)
        out = np.full(len(q), -1)
        out[np.abs(c[:, 0] - x0) < tol] = 0
        out[np.abs(c[:, 0] - x1) < tol] = 1
        out[np.abs(c[:, 1] - y0) < tol] = 2
        out[np.abs(c[:, 1] - y1) < tol] = 3
        out[(np.abs(c[:, 2]) < tol) | (np.abs(c[:, 2] - dz) < tol)] = 4
        return out

    return hex_mesh(pts, np.array(cells), names, gtypes, classify)


def velocity_mesh(ex, ey, ez, jitter: float = 0.0, seed: int = 7) -> PolyMesh:
    """Hex mesh of velocity space with node coordinates ex x ey x ez (one patch ``boundary``),
    as dugksFoam's VMESH build reads from constant/vMesh: its cell centres are the discrete
    velocities and its cell volumes the weights (src/fvDVM/fvDVM/fvDVM.C:170-249).  ``jitter``
    moves every interior node by up to that fraction of the local spacing (seeded), so the
    cells are general hexahedra with non-planar faces and the velocities no tensor product; the
    outer box and its volume stay as they are."""
    ex, ey, ez = (np.asarray(e, dtype=np.float64) for e in (ex, ey, ez))
    nx, ny, nz = len(ex) - 1, len(ey) - 1, len(ez) - 1
    k_, j_, i_ = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    i_, j_, k_ = i_.ravel(), j_.ravel(), k_.ravel()
    pts = np.stack([ex[i_], ey[j_], ez[k_]], axis=1)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        for ax, (idx, e, n) in enumerate(((i_, ex, nx), (j_, ey, ny), (k_, ez, nz))):
            inner = (idx > 0) & (idx < n)
            h = np.minimum(np.diff(e)[np.clip(idx - 1, 0, n - 1)], np.diff(e)[np.clip(idx, 0, n - 1)])
            pts[inner, ax] += jitter * h[inner] * rng.uniform(-1.0, 1.0, int(inner.sum()))
    pid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    cells = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                cells.append([pid(i, j, k), pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i, j + 1, k),
                              pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j + 1, k + 1), pid(i, j + 1, k + 1)])
    return hex_mesh(pts, np.array(cells), ["boundary"], ["patch"], lambda q: np.zeros(len(q), dtype=np.int64))


# OpenFOAM hex model faces (vertex order 0-3 bottom, 4-7 top), outward for a right-handed hex
_HEX_FACES = np.array([[0, 4, 7, 3], [1, 2, 6, 5], [0, 1, 5, 4], [3, 7, 6, 2], [0, 3, 2, 1], [4, 5, 6, 7]])


def hex_mesh(points: np.ndarray, cells: np.ndarray, patch_names, patch_types, classify) -> PolyMesh:
    """Vectorised polyMesh assembly from hexahedra.

    ``cells`` [nC, 8] are point labels in OpenFOAM hex order.  Shared faces are matched by
    their sorted point labels.  ``classify(face_xyz [k, 4, 3]) -> int array`` assigns every
    boundary face to a patch index (into ``patch_names``), or -1 to raise.  Internal faces
    are ordered upper-triangularly, boundary faces patch by patch and by owner; every face
    is oriented out of its owner.
    """
    points = np.asarray(points, dtype=np.float64)
    cells = np.asarray(cells, dtype=np.int64)
    nC = len(cells)
    fp = cells[:, _HEX_FACES].reshape(-1, 4)                 # [6nC, 4]
    fc = np.repeat(np.arange(nC), 6)
    key = np.sort(fp, axis=1)
    order = np.lexsort(key.T[::-1])
    ks = key[order]
    same = np.all(ks[1:] == ks[:-1], axis=1)
    first = np.nonzero(same)[0]
    if np.any(same[1:] & same[:-1]):
        raise ValueError("a face is shared by more than two cells")
    a, b = order[first], order[first + 1]
    internal_mask = np.zeros(len(fp), dtype=bool)
    internal_mask[a] = True
    internal_mask[b] = True
    ca, cb = fc[a], fc[b]
    own = np.minimum(ca, cb)
    nei = np.maximum(ca, cb)
    fpts = np.where((ca < cb)[:, None], fp[a], fp[b])
    cavg = points[cells].mean(axis=1)

    def normals(q):
        xyz = points[q]
        ctr = xyz.mean(axis=1)
        n = np.zeros((len(q), 3))
        for k in range(4):
            n += np.cross(xyz[:, (k + 1) % 4] - xyz[:, k], ctr - xyz[:, k])
        return n, ctr

    n, _ = normals(fpts)
    flip = np.sum(n * (cavg[nei] - cavg[own]), axis=1) < 0
    fpts[flip] = fpts[flip][:, ::-1]
    io = np.lexsort((nei, own))
    own, nei, fpts = own[io], nei[io], fpts[io]

    bidx = np.nonzero(~internal_mask)[0]
    bpts = fp[bidx]
    bown = fc[bidx]
    n, ctr = normals(bpts)
    flip = np.sum(n * (ctr - cavg[bown]), axis=1) < 0
    bpts[flip] = bpts[flip][:, ::-1]
    pid = np.asarray(classify(points[bpts]))
    if np.any(pid < 0):
        raise ValueError(f"{int(np.sum(pid < 0))} boundary faces match no patch")
    face_list = [fpts]
    owner = [own]
    patches = []
    start = len(own)
    for k, (name, typ) in enumerate(zip(patch_names, patch_types)):
        sel = np.nonzero(pid == k)[0]
        sel = sel[np.argsort(bown[sel], kind="stable")]
        face_list.append(bpts[sel])
        owner.append(bown[sel])
        patches.append(Patch(name, typ, start, len(sel)))
        start += len(sel)
    allf = np.concatenate(face_list)
    return PolyMesh(points=points, face_offsets=np.arange(len(allf) + 1, dtype=np.int64) * 4,
                    face_points=allf.ravel().astype(np.int64), owner=np.concatenate(owner).astype(np.int64),
                    neighbour=nei.astype(np.int64), patches=patches, n_cells=nC)


def cell_faces(mesh: PolyMesh):
    """CSR list of the faces of each cell, faces in ascending label order."""
    nC = mesh.n_cells
    nIF = mesh.n_internal_faces
    pairs_c = np.concatenate([mesh.owner, mesh.neighbour])
    pairs_f = np.concatenate([np.arange(mesh.n_faces), np.arange(nIF)])
    order = np.lexsort((pairs_f, pairs_c))
    cells = pairs_c[order]
    faces = pairs_f[order]
    offsets = np.zeros(nC + 1, dtype=np.int64)
    np.add.at(offsets, cells + 1, 1)
    offsets = np.cumsum(offsets)
    return offsets, faces
