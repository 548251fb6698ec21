"""Case descriptions: mesh + boundary conditions + velocity set + gas + controls.

A :class:`Case` carries what dugksFoam reads from an OpenFOAM case directory:
``constant/DVMProperties`` (fvDVMparas, gasProperties: src/fvDVM/fvDVM/fvDVM.C:1928-1948),
``constant/Xis`` and ``constant/weights`` (fvDVM.C:62-123), the patch types and values of
``0/rho``, ``0/U``, ``0/T`` (createFields.H; the rho patch type selects the DF face type,
src/fvDVM/discreteVelocity/discreteVelocity.C:188-195) and the ``system/controlDict`` time
controls (src/readTimeControlsExplicit.H:39-52).

The flattened arrays (:meth:`Case.arrays`) are the plain-pointer form that crosses the
C-ABI: internal faces first, then the boundary faces of the non-empty patches.
"""
from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field

import numpy as np

from . import mesh as meshmod
from . import quadrature

# rho patch types accepted by the path, with the DF face type they select
# (src/fvDVM/discreteVelocity/discreteVelocity.C:188-195; symmetryPlane is a geometric
# constraint type whose DF faces OpenFOAM sets by itself)
RHO_BC_TO_DF = {
    "calculatedMaxwell": "maxwellWall",
    "fixedValue": "mixed",
    "zeroGradient": "zeroGradient",
    "farField": "farField",
    "pressureIn": "farField",
    "pressureOut": "farField",
    "symmetryMod": "DVMsymmetry",
    "symmetryPlane": "symmetryPlane",
}
PATCH_CODE = {"calculatedMaxwell": 1, "fixedValue": 2, "zeroGradient": 3, "farField": 4, "pressureIn": 5,
              "pressureOut": 6, "symmetryMod": 7, "symmetryPlane": 8}


@dataclass
class PatchBC:
    name: str
    rho_type: str
    rho: np.ndarray   # [size]
    U: np.ndarray     # [size, 3]
    T: np.ndarray     # [size]
    pressure: float = 0.0  # pressureIn / pressureOut patches


@dataclass
class Gas:
    R: float = 1.0
    omega: float = 0.81
    Tref: float = 1.0
    muRef: float = 1.0
    Pr: float = 2.0 / 3.0
    KInner: int = 0


@dataclass
class Controls:
    deltaT: float = 1.0
    adjustTimeStep: bool = True
    maxCo: float = 0.8
    maxDeltaT: float = 1.0


@dataclass
class Case:
    name: str
    mesh: meshmod.PolyMesh
    geo: meshmod.Geometry
    D: int
    xis: np.ndarray
    wts: np.ndarray
    xi_max: float
    xi_min: float
    gas: Gas
    rho0: np.ndarray
    U0: np.ndarray
    T0: np.ndarray
    bcs: list = field(default_factory=list)   # PatchBC per non-empty patch, mesh order
    controls: Controls = field(default_factory=Controls)
    q1_compat: bool = True
    meta: dict = field(default_factory=dict)
    # gradient scheme of system/fvSchemes: "leastSquares" (dugksFoam's demos) or "Gauss linear"
    grad_scheme: str = "leastSquares"
    # velocity list of the VMESH build (cell centres [n, 3] and volumes [n] of constant/vMesh,
    # src/fvDVM/fvDVM/fvDVM.C:129-253); None: the tensor product of xis / wts
    vlist: tuple | None = None
    # updateMacroVol (fvDVM.C:1521-1682): 0 macroFlux no, 1 macroFlux yes (face-macro fluxes),
    # 2 macroFlux yes in the VMESH build (velocity-moment fluxes of the face DFs)
    macro_flux: int = 0

    @property
    def n_cells(self) -> int:
        return self.mesh.n_cells

    @property
    def nv(self) -> int:
        if self.vlist is not None:
            return len(self.vlist[1])
        return len(self.xis) ** self.D

    def arrays(self) -> dict:
        """Flattened C-ABI arrays (int32 labels, float64 geometry)."""
        m, g = self.mesh, self.geo
        nIF = m.n_internal_faces
        bc_by_name = {b.name: b for b in self.bcs}
        bfaces, ptype, pstart, psize, names, ppress = [], [], [], [], [], []
        b_rho, b_U, b_T = [], [], []
        for p in m.patches:
            if p.geometric_type == "empty" or p.size == 0:
                continue
            if p.name not in bc_by_name:
                raise ValueError(f"no boundary condition for patch {p.name!r}")
            bc = bc_by_name[p.name]
            if bc.rho_type not in RHO_BC_TO_DF:
                raise ValueError(f"rho patch type {bc.rho_type!r} on {p.name!r} is not supported")
            ptype.append(PATCH_CODE[bc.rho_type])
            ppress.append(float(bc.pressure))
            pstart.append(len(bfaces))
            psize.append(p.size)
            names.append(p.name)
            bfaces.extend(range(p.start, p.start + p.size))
            b_rho.append(np.broadcast_to(np.asarray(bc.rho, float), (p.size,)))
            b_U.append(np.broadcast_to(np.asarray(bc.U, float), (p.size, 3)))
            b_T.append(np.broadcast_to(np.asarray(bc.T, float), (p.size,)))
        faces = np.concatenate([np.arange(nIF), np.asarray(bfaces, dtype=np.int64)])
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        out = dict(
            nC=m.n_cells, nIF=nIF, nB=len(bfaces), D=self.D,
            owner=i32(m.owner[faces]), neighbour=i32(m.neighbour),
            Sf=f64(g.Sf[faces]), Cf=f64(g.Cf[faces]), C=f64(g.C), V=f64(g.V),
            weights=f64(g.weights), delta_coeffs=f64(g.delta_coeffs),
            patch_names=names, patch_type=i32(ptype), patch_start=i32(pstart), patch_size=i32(psize),
            patch_pressure=f64(ppress),
            bf_rho=f64(np.concatenate(b_rho) if b_rho else np.zeros(0)),
            bf_U=f64(np.concatenate(b_U) if b_U else np.zeros((0, 3))),
            bf_T=f64(np.concatenate(b_T) if b_T else np.zeros(0)),
            xis=f64(self.xis), wts=f64(self.wts), xi_max=float(self.xi_max),
            R=self.gas.R, omega=self.gas.omega, Tref=self.gas.Tref, muRef=self.gas.muRef,
            Pr=self.gas.Pr, K=int(self.gas.KInner), q1_compat=int(self.q1_compat),
            rho0=f64(self.rho0), U0=f64(self.U0), T0=f64(self.T0), ls_own=None, ls_nei=None,
            n_list=0 if self.vlist is None else len(self.vlist[1]),
            list_xi=None if self.vlist is None else f64(np.reshape(self.vlist[0], (-1, 3))),
            list_w=None if self.vlist is None else f64(self.vlist[1]),
            macro_flux=int(self.macro_flux),
        )
        out.update(self._gradient_vectors())
        return out

    def _gradient_vectors(self) -> dict:
        """ls_own / ls_nei [nIF, 3] for the gradient scheme: None for leastSquares (the
        library and the oracle form the least-squares vectors themselves); for Gauss linear,
        grad_P = (1/V_P) sum_f S_f phi_f with phi_f = w phi_P + (1 - w) phi_N and boundary values
        equal to the cell value, which by sum_f S_f = 0 over a closed cell is
        sum_f S_f (1 - w_f)/V_P (phi_N - phi_P) at the owner and sum_f S_f w_f / V_N (phi_N - phi_P)
        at the neighbour: the vectors S_f (1 - w_f) / V_P and -S_f w_f / V_N of the form
        grad_P += own . (phi_N - phi_P), grad_N -= nei . (phi_N - phi_P)."""
        if self.grad_scheme in ("leastSquares", "", None):
            return {}
        if self.grad_scheme.replace(" ", "") != "Gausslinear":
            raise ValueError(f"gradient scheme {self.grad_scheme!r} is not supported "
                             "(leastSquares, Gauss linear)")
        m, g = self.mesh, self.geo
        nIF = m.n_internal_faces
        Sf, w = g.Sf[:nIF], g.weights
        own, nei = m.owner[:nIF], m.neighbour
        return dict(ls_own=np.ascontiguousarray(Sf * ((1.0 - w) / g.V[own])[:, None]),
                    ls_nei=np.ascontiguousarray(-Sf * (w / g.V[nei])[:, None]))


def mu_ref_from_kn(Kn: float, rho0: float = 1.0, R: float = 1.0, T0: float = 1.0, L: float = 1.0,
                   omega: float = 0.81, alpha: float = 1.0) -> float:
    """Reference viscosity for a Knudsen number (demo/cavity/para.py:33-46, VHS form).

    mu_bar = 5 (alpha+1)(alpha+2) sqrt(pi) Kn / (4 alpha (5 - 2 omega)(7 - 2 omega)),
    mu = rho0 * sqrt(2 R T0) * L * mu_bar.
    """
    mu_bar = 5.0 * (alpha + 1.0) * (alpha + 2.0) * math.sqrt(math.pi) * Kn / (
        4.0 * alpha * (5.0 - 2.0 * omega) * (7.0 - 2.0 * omega))
    return rho0 * math.sqrt(2.0 * R * T0) * L * mu_bar


def seeded_uniform(tag: str, lo: float, hi: float) -> float:
    """Deterministic uniform draw keyed by a string (used for the per-run lid jitter)."""
    h = hashlib.sha256(tag.encode()).digest()
    u = int.from_bytes(h[:8], "little") / 2.0 ** 64
    return lo + (hi - lo) * u


def cavity_case(N: int, rule: str, n_dv: int, vel_extent: float, Kn: float,
                lid_speed: float | None = None, name: str | None = None,
                wall_T=None, seed_tag: str | None = None, gas: Gas | None = None,
                controls: Controls | None = None) -> Case:
    """Lid-driven cavity, nondimensional (R = T0 = rho0 = L = 1).

    ``vel_extent`` is the half width of the velocity domain in units of sqrt(R T0) for the
    Newton-Cotes / trapezoid rules, or the scale C (in sqrt(R T0)) for Gauss-Hermite.
    ``wall_T`` optionally gives per-face temperatures of the bottom wall (y = 0).
    """
    m = meshmod.cavity(N)
    geo = meshmod.geometry(m)
    R = 1.0
    T0 = 1.0
    if gas is None:
        gas = Gas(R=R, omega=0.81, Tref=T0, muRef=mu_ref_from_kn(Kn, omega=0.81), Pr=2.0 / 3.0, KInner=0)
    scale = vel_extent * math.sqrt(R * T0)
    xis, wts = quadrature.make_rule(rule, scale, n_dv)
    if lid_speed is None:
        lid_speed = 0.16 * math.sqrt(2.0 * R * T0)
    nC = m.n_cells
    rho0 = np.ones(nC)
    U0 = np.zeros((nC, 3))
    T0c = np.full(nC, T0)
    mw = m.patch("movingWall")
    fw = m.patch("fixedWalls")
    Tfw = np.full(fw.size, T0)
    if wall_T is not None:
        Tfw[N:2 * N] = np.asarray(wall_T, float)   # bottom wall faces sit in the middle third
    bcs = [
        PatchBC("movingWall", "calculatedMaxwell", np.ones(mw.size),
                np.tile([lid_speed, 0.0, 0.0], (mw.size, 1)), np.full(mw.size, T0)),
        PatchBC("fixedWalls", "calculatedMaxwell", np.ones(fw.size), np.zeros((fw.size, 3)), Tfw),
    ]
    return Case(name=name or f"cavity{N}", mesh=m, geo=geo, D=2, xis=xis, wts=wts,
                xi_max=float(np.max(np.abs(xis))), xi_min=float(np.min(xis)), gas=gas,
                rho0=rho0, U0=U0, T0=T0c, bcs=bcs, controls=controls or Controls(),
                meta=dict(Kn=Kn, lid_speed=lid_speed, rule=rule, n_dv=n_dv, seed_tag=seed_tag))


def box_case(nx: int, ny: int, sides: dict, rule: str = "GH", n_dv: int = 8, vel_extent: float = 1.0,
             Kn: float = 0.1, extent=(0.0, 1.0, 0.0, 1.0), rho0=1.0, U0=(0.0, 0.0, 0.0), T0=1.0,
             controls: Controls | None = None, name: str | None = None, grading: float = 1.0) -> Case:
    """A 2-D box (nondimensional, R = Tref = 1) with one boundary condition per side.

    ``sides`` maps ``left``, ``right``, ``bottom``, ``top`` to ``(rho_type, rho, U, T)`` or
    ``(rho_type, rho, U, T, pressure)``; rho, U, T may be scalars / 3-vectors or per-face
    arrays.  ``rho0``, ``U0``, ``T0`` are the initial cell fields (constants or callables of
    the cell centres).  ``symmetryPlane`` sides get that geometric type, walls ``wall``.
    ``grading`` > 1 grows the cell widths geometrically along x and y (mesh.box2d)."""
    gtypes = {k: ("symmetryPlane" if v[0] == "symmetryPlane" else "wall" if v[0] == "calculatedMaxwell" else "patch")
              for k, v in sides.items()}
    x0, x1, y0, y1 = extent
    m = meshmod.box2d(nx, ny, x0, x1, y0, y1, types=gtypes, grading=grading)
    geo = meshmod.geometry(m)
    gas = Gas(R=1.0, omega=0.81, Tref=1.0, muRef=mu_ref_from_kn(Kn, omega=0.81), Pr=2.0 / 3.0, KInner=0)
    scale = vel_extent * math.sqrt((2.0 if rule.startswith("GH") else 1.0))
    xis, wts = quadrature.make_rule(rule, scale, n_dv)
    bcs = []
    for p in m.patches:
        if p.geometric_type == "empty":
            continue
        spec = sides[p.name]
        rt, r, u, t = spec[:4]
        bcs.append(PatchBC(p.name, rt, np.broadcast_to(np.asarray(r, float), (p.size,)).copy(),
                           np.broadcast_to(np.asarray(u, float), (p.size, 3)).copy(),
                           np.broadcast_to(np.asarray(t, float), (p.size,)).copy(),
                           float(spec[4]) if len(spec) > 4 else 0.0))
    C = geo.C
    f = lambda v, width: (np.asarray(v(C), float) if callable(v) else np.broadcast_to(np.asarray(v, float),
                                                                                       (m.n_cells,) + width)).copy()
    return Case(name=name or f"box{nx}x{ny}", mesh=m, geo=geo, D=2, xis=xis, wts=wts,
                xi_max=float(np.max(np.abs(xis))), xi_min=float(np.min(xis)), gas=gas,
                rho0=f(rho0, ()), U0=f(U0, (3,)), T0=f(T0, ()), bcs=bcs,
                controls=controls or Controls(maxCo=STABLE_MAX_CO[2]), meta=dict(Kn=Kn, rule=rule, n_dv=n_dv))


def config0() -> Case:
    """40x40 cavity, 28x28 Newton-Cotes (composite Simpson 3/8) on [-4, 4], Kn = 0.075."""
    return cavity_case(40, "NC38", 28, 4.0, 0.075, name="cavity40_nc28_kn0.075")


# Courant limit for long runs.  The step reconstructs at x_f - xi dt/2 from the upwind cell
# with its full gradient (transverse terms included), an unsplit scheme whose von Neumann
# factor for free transport exceeds 1 once sum_d |xi_d| dt / dx > 1 (checkerboard mode;
# 1.20 per step at |xi_d| dt / dx = 0.55 in 2-D).  dugksFoam's Courant number bounds
# (|U| + sqrt(D) xi_max) dt / dx, so maxCo 0.8 allows sum_d |xi_d| dt / dx up to 0.8 sqrt(D):
# corner velocities of the set grow and, on meshes fine enough for the growth to outrun
# the walls (256^2, 100^2 with 2000 / 500 steps), the run diverges.  maxCo <= 1 / sqrt(D)
# keeps every velocity stable.
STABLE_MAX_CO = {2: 0.65, 3: 0.55}


def config1(Kn: float, N: int = 256, n_dv: int = 64) -> Case:
    """256x256 cavity, 64x64 trapezoid on [-5, 5], lid jittered by +-5 % per (config, Kn),
    Courant limit STABLE_MAX_CO[2]."""
    tag = f"cavity{N}_trap{n_dv}_Kn{Kn!r}"
    jitter = seeded_uniform(tag, -0.05, 0.05)
    lid = 0.16 * math.sqrt(2.0) * (1.0 + jitter)
    c = cavity_case(N, "trapezoid", n_dv, 5.0, Kn, lid_speed=lid, name=tag, seed_tag=tag,
                    controls=Controls(maxCo=STABLE_MAX_CO[2]))
    c.meta["lid_jitter"] = jitter
    return c


def smoothed_wall_profile(n: int, seed: int = 12345, amp: float = 0.05, width: float = 4.0):
    """T0 (1 + amp * smooth(N(0,1))) on n faces: Gaussian-kernel smoothed white noise,
    renormalised to unit standard deviation before scaling."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(n)
    x = np.arange(-3 * int(width), 3 * int(width) + 1)
    k = np.exp(-0.5 * (x / width) ** 2)
    zs = np.convolve(np.pad(z, len(x) // 2, mode="wrap"), k / k.sum(), mode="valid")
    zs = (zs - zs.mean()) / zs.std()
    return 1.0 + amp * zs


def sphere_case(mesh_kw: dict, rule: str, n_dv: int, vel_scale: float, Kn: float, Ma: float = 0.2,
                wall_T=1.0, name: str | None = None, controls: Controls | None = None) -> Case:
    """Flow past a sphere (diameter 1), nondimensional (R = T0 = rho0 = 1, L = D).

    Freestream Ma = U / sqrt(gamma R T0), gamma = 5/3, on the inlet / outlet / side patches
    (rho fixedValue -> DF mixed: incoming half frozen at the freestream Maxwellian), and a
    diffuse wall (rho calculatedMaxwell) at temperature ``wall_T`` (scalar or per face) on
    the sphere.  Uniform freestream initial state.
    """
    from . import sphere
    m = sphere.sphere_box_mesh(**mesh_kw)
    geo = meshmod.geometry(m)
    R = T0 = 1.0
    U_inf = Ma * math.sqrt(5.0 / 3.0 * R * T0)
    gas = Gas(R=R, omega=0.81, Tref=T0, muRef=mu_ref_from_kn(Kn, omega=0.81, L=1.0), Pr=2.0 / 3.0, KInner=0)
    scale = vel_scale * math.sqrt((2.0 if rule.startswith("GH") else 1.0) * R * T0)
    xis, wts = quadrature.make_rule(rule, scale, n_dv)
    nC = m.n_cells
    bcs = []
    for p in m.patches:
        if p.name == "sphere":
            bcs.append(PatchBC(p.name, "calculatedMaxwell", np.ones(p.size), np.zeros((p.size, 3)),
                               np.broadcast_to(np.asarray(wall_T, float), (p.size,)).copy()))
        else:
            bcs.append(PatchBC(p.name, "fixedValue", np.ones(p.size), np.tile([U_inf, 0.0, 0.0], (p.size, 1)),
                               np.full(p.size, T0)))
    return Case(name=name or f"sphere{nC}", mesh=m, geo=geo, D=3, xis=xis, wts=wts,
                xi_max=float(np.max(np.abs(xis))), xi_min=float(np.min(xis)), gas=gas,
                rho0=np.ones(nC), U0=np.tile([U_inf, 0.0, 0.0], (nC, 1)), T0=np.full(nC, T0), bcs=bcs,
                controls=controls or Controls(maxCo=STABLE_MAX_CO[3]),
                meta=dict(Kn=Kn, Ma=Ma, U_inf=U_inf, rule=rule, n_dv=n_dv))


SPHERE_200K = dict(nc=24, nr=12, n_up=14, n_down=30, n_side=14)
SPHERE_50K = dict(nc=16, nr=8, n_up=8, n_down=18, n_side=8)


def config2(n_dv: int = 24) -> Case:
    """Sphere, ~211k hex cells, 24^3 half-range Gauss-Hermite set, Kn = 0.1, Ma = 0.2."""
    return sphere_case(SPHERE_200K, "GH", n_dv, 1.0, 0.1, name=f"sphere211k_gh{n_dv}_kn0.1")


def config3() -> Case:
    """100x100 cavity, 32x32 trapezoid on [-4, 4], Kn = 0.1, seeded bottom-wall profile."""
    Tw = smoothed_wall_profile(100)
    return cavity_case(100, "trapezoid", 32, 4.0, 0.1, wall_T=Tw, name="cavity100_trap32_kn0.1",
                       controls=Controls(maxCo=STABLE_MAX_CO[2]))
