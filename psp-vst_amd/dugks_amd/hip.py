"""ctypes binding of libdugks_hip.so (include/dugks.h) with an fvDVM-shaped facade.

:class:`FvDVM` mirrors the public interface of Foam::fvDVM (src/fvDVM/fvDVM/fvDVM.H:301-408):
``evolution()``, ``getCoNum()``, the macroscopic field accessors (``rhoVol``, ``Uvol``,
``Tvol``, ``qVol``, ``tauVol``, ``rhoSurf``, ``Usurf``, ``Tsurf``, ``qSurf``, ``tauSurf``) and
the velocity-set / gas accessors.  The time loop of dugksFoam.C:63-76 is :meth:`FvDVM.run`.

There is no CPU fallback: if the HIP library is missing or fails to load, construction
raises.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from .timecontrol import next_delta_t

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdugks_hip.so")

FP64 = 0
MIXED = 1

_D = ct.POINTER(ct.c_double)
_I = ct.POINTER(ct.c_int)


class Mesh(ct.Structure):
    _fields_ = [("n_cells", ct.c_int), ("n_internal_faces", ct.c_int), ("n_boundary_faces", ct.c_int),
                ("n_solution_d", ct.c_int), ("owner", _I), ("neighbour", _I), ("Sf", _D), ("Cf", _D),
                ("C", _D), ("V", _D), ("weights", _D), ("delta_coeffs", _D), ("ls_own", _D), ("ls_nei", _D)]


class Patch(ct.Structure):
    _fields_ = [("type", ct.c_int), ("start", ct.c_int), ("size", ct.c_int),
                ("rho", _D), ("U", _D), ("T", _D), ("pressure", ct.c_double)]


class VSet(ct.Structure):
    _fields_ = [("n_per_dim", ct.c_int), ("xis", _D), ("weights", _D), ("xi_max", ct.c_double),
                ("xi_min", ct.c_double), ("n_list", ct.c_int), ("list_xi", _D), ("list_w", _D)]


class Gas(ct.Structure):
    _fields_ = [("R", ct.c_double), ("omega", ct.c_double), ("Tref", ct.c_double), ("muRef", ct.c_double),
                ("Pr", ct.c_double), ("KInner", ct.c_int)]


class Options(ct.Structure):
    _fields_ = [("precision", ct.c_int), ("q1_compat", ct.c_int), ("device", ct.c_int), ("fused", ct.c_int),
                ("macro_flux", ct.c_int)]


EXPORTS = [
    "dugks_create", "dugks_init_from_macros", "dugks_co_num", "dugks_evolution", "dugks_advance",
    "dugks_advance_noalloc",
    "dugks_synchronize", "dugks_get_cell_macros", "dugks_get_face_macros", "dugks_get_boundary_rho",
    "dugks_get_df", "dugks_set_df", "dugks_set_cell_macros", "dugks_checkpoint_save", "dugks_checkpoint_load",
    "dugks_num_velocities", "dugks_velocity_set", "dugks_kernel_count", "dugks_kernel_name", "dugks_profile",
    "dugks_kernel_times", "dugks_last_error", "dugks_abi_version", "dugks_destroy",
]

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libdugks_hip.so and declare its signatures.  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python psp-vst_amd/build.py` "
                           "(there is no CPU fallback for the DUGKS step)")
    L = ct.CDLL(path)
    vp = ct.c_void_p
    L.dugks_create.argtypes = [ct.POINTER(Mesh), ct.POINTER(Patch), ct.c_int, ct.POINTER(VSet), ct.POINTER(Gas),
                               ct.POINTER(Options), ct.POINTER(vp)]
    L.dugks_init_from_macros.argtypes = [vp, _D, _D, _D]
    L.dugks_co_num.argtypes = [vp, ct.c_double, _D, _D]
    L.dugks_evolution.argtypes = [vp, ct.c_double]
    L.dugks_advance.argtypes = [vp, ct.c_int, ct.c_int, ct.c_double, ct.c_double, _D, _D]
    L.dugks_advance_noalloc.argtypes = [vp, ct.c_int, ct.c_int, ct.c_double, ct.c_double, _D, _D]
    L.dugks_synchronize.argtypes = [vp]
    L.dugks_get_cell_macros.argtypes = [vp, _D, _D, _D, _D, _D]
    L.dugks_get_face_macros.argtypes = [vp, _D, _D, _D, _D, _D]
    L.dugks_get_boundary_rho.argtypes = [vp, _D]
    L.dugks_get_df.argtypes = [vp, _D, _D]
    L.dugks_set_df.argtypes = [vp, _D, _D]
    L.dugks_set_cell_macros.argtypes = [vp, _D, _D, _D, _D]
    L.dugks_checkpoint_save.argtypes = [vp, ct.c_char_p]
    L.dugks_checkpoint_load.argtypes = [vp, ct.c_char_p]
    L.dugks_num_velocities.argtypes = [vp]
    L.dugks_velocity_set.argtypes = [vp, _D, _D]
    L.dugks_kernel_count.argtypes = [vp]
    L.dugks_kernel_name.argtypes = [vp, ct.c_int]
    L.dugks_kernel_name.restype = ct.c_char_p
    L.dugks_profile.argtypes = [vp, ct.c_int]
    L.dugks_kernel_times.argtypes = [vp, _D, _I]
    L.dugks_last_error.restype = ct.c_char_p
    L.dugks_destroy.argtypes = [vp]
    L.dugks_destroy.restype = None
    L.dugks_adjoint.argtypes = [vp, ct.c_int, _D, ct.c_int, ct.POINTER(Objective), _D, _D, _D, _D, _D, _D]
    L.dugks_path_info.argtypes = [vp, _I, _I, _I, _I]
    _lib = L
    return L


class Objective(ct.Structure):
    _fields_ = [("q_weight", _D), ("mom_weight", _D)]


def _p(a, typ=_D):
    return a.ctypes.data_as(typ) if a is not None else None


class DugksError(RuntimeError):
    pass


class FvDVM:
    """One DUGKS solver on one GPU, built from a :class:`dugks_amd.cases.Case`.

    ``fused`` is ``dugks_options.fused``: -1 fused when available, else staged; 0 staged;
    1 fused with the default schedule; 2 fused with stored face values; 3 fused recomputing
    the halo rings; 4 stored face values in the two-pass form (each pass forms f-bar-plus
    itself; the form 3-D FP32 storage falls back to when its f-bar-plus / u arrays do not
    fit)."""

    def __init__(self, case, precision: int = FP64, fused: int = -1, device: int = 0, q1_compat=None,
                 ls_vectors=None):
        L = load_library()
        self.L = L
        self.case = case
        a = case.arrays()
        self._a = a
        self.nC, self.nIF, self.nB = a["nC"], a["nIF"], a["nB"]
        self.nF = self.nIF + self.nB
        self.D = a["D"]
        ls_own = ls_nei = None
        if ls_vectors is None and a.get("ls_own") is not None:   # the case's gradient scheme
            ls_vectors = (a["ls_own"], a["ls_nei"])
        if ls_vectors is not None:
            ls_own = np.ascontiguousarray(ls_vectors[0], dtype=np.float64)
            ls_nei = np.ascontiguousarray(ls_vectors[1], dtype=np.float64)
            self._ls = (ls_own, ls_nei)
        mesh = Mesh(n_cells=self.nC, n_internal_faces=self.nIF, n_boundary_faces=self.nB, n_solution_d=self.D,
                    owner=_p(a["owner"], _I), neighbour=_p(a["neighbour"], _I), Sf=_p(a["Sf"]), Cf=_p(a["Cf"]),
                    C=_p(a["C"]), V=_p(a["V"]), weights=_p(a["weights"]), delta_coeffs=_p(a["delta_coeffs"]),
                    ls_own=_p(ls_own), ls_nei=_p(ls_nei))
        npatch = len(a["patch_type"])
        patches = (Patch * max(npatch, 1))()
        self._pv = []
        for i in range(npatch):
            s, n = int(a["patch_start"][i]), int(a["patch_size"][i])
            rho = np.ascontiguousarray(a["bf_rho"][s:s + n])
            U = np.ascontiguousarray(a["bf_U"][s:s + n])
            T = np.ascontiguousarray(a["bf_T"][s:s + n])
            self._pv += [rho, U, T]
            patches[i] = Patch(type=int(a["patch_type"][i]), start=s, size=n, rho=_p(rho), U=_p(U), T=_p(T),
                               pressure=float(a["patch_pressure"][i]))
        vset = VSet(n_per_dim=len(a["xis"]), xis=_p(a["xis"]), weights=_p(a["wts"]), xi_max=a["xi_max"],
                    xi_min=float(case.xi_min), n_list=a["n_list"], list_xi=_p(a["list_xi"]),
                    list_w=_p(a["list_w"]))
        gas = Gas(R=a["R"], omega=a["omega"], Tref=a["Tref"], muRef=a["muRef"], Pr=a["Pr"], KInner=a["K"])
        q1 = a["q1_compat"] if q1_compat is None else int(q1_compat)
        opt = Options(precision=int(precision), q1_compat=q1, device=int(device), fused=int(fused),
                      macro_flux=int(a["macro_flux"]))
        h = ct.c_void_p()
        if L.dugks_create(ct.byref(mesh), patches, npatch, ct.byref(vset), ct.byref(gas), ct.byref(opt),
                          ct.byref(h)) != 0:
            raise DugksError(L.dugks_last_error().decode())
        self._h = h
        self.precision = precision
        self.Nv = L.dugks_num_velocities(h)
        self.controls = case.controls
        self.dt = case.controls.deltaT
        self.time = 0.0
        self.step_index = 0
        self._check(L.dugks_init_from_macros(h, _p(a["rho0"]), _p(a["U0"]), _p(a["T0"])))

    # -- lifetime -------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self.L.dugks_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise DugksError(self.L.dugks_last_error().decode())

    # -- fvDVM interface --------------------------------------------------------------
    def evolution(self, dt=None):
        """fvDVM::evolution (fvDVM.C:2128-2174) with time step dt (default: current)."""
        if dt is not None:
            self.dt = float(dt)
        self._check(self.L.dugks_evolution(self._h, self.dt))
        self.time += self.dt
        self.step_index += 1

    def getCoNum(self):
        """fvDVM::getCoNum (fvDVM.C:2177-2185): (max, mean) Courant number for the current dt."""
        mx, mean = ct.c_double(), ct.c_double()
        self._check(self.L.dugks_co_num(self._h, self.dt, ct.byref(mx), ct.byref(mean)))
        return mx.value, mean.value

    def run(self, n_steps: int, record_dt: bool = False, device_loop: bool = True):
        """n iterations of the dugksFoam loop (CourantNo.H, setDeltaTvar.H, evolution)."""
        c = self.controls
        if device_loop:
            dt = ct.c_double(self.dt)
            hist = np.empty(n_steps) if record_dt else None
            self._check(self.L.dugks_advance(self._h, int(n_steps), int(bool(c.adjustTimeStep)), c.maxCo,
                                             c.maxDeltaT, ct.byref(dt), _p(hist)))
            self.dt = dt.value
            self.step_index += n_steps
            if hist is not None:
                self.time += float(hist.sum())
            return hist
        hist = []
        for _ in range(n_steps):
            co, _ = self.getCoNum()
            self.dt = next_delta_t(self.dt, co, c)
            hist.append(self.dt)
            self.evolution()
        return np.array(hist)

    def run_noalloc(self, n_steps: int, record_dt: bool = False):
        """run() through dugks_advance_noalloc: no device allocation in the call; the time-step
        history (record_dt) is limited to 65536 steps per call."""
        c = self.controls
        dt = ct.c_double(self.dt)
        hist = np.empty(n_steps) if record_dt else None
        self._check(self.L.dugks_advance_noalloc(self._h, int(n_steps), int(bool(c.adjustTimeStep)), c.maxCo,
                                                 c.maxDeltaT, ct.byref(dt), _p(hist)))
        self.dt = dt.value
        self.step_index += n_steps
        if hist is not None:
            self.time += float(hist.sum())
        return hist

    def run_fixed(self, dts):
        """evolution() with the given time steps (no Courant-number control)."""
        for dt in np.asarray(dts, dtype=np.float64):
            self.evolution(float(dt))

    def adjoint(self, dts, checkpoint_every: int = 1, q_weight=None, mom_weight=None, want_df0: bool = False):
        """Objective and gradients of len(dts) steps with fixed time steps (dugks_adjoint).

        q_weight / mom_weight: [nF, 3] face weights of the heat-flux and momentum-flux terms
        of the objective on the last step.  Returns dict(J, dmuRef, dTw [nB], dUw [nB, 3]
        and, with want_df0, dg0 / dh0 [Nv, nC]); the solver is left after the steps.
        """
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        qw = None if q_weight is None else np.ascontiguousarray(q_weight, dtype=np.float64).reshape(self.nF, 3)
        mw = None if mom_weight is None else np.ascontiguousarray(mom_weight, dtype=np.float64).reshape(self.nF, 3)
        obj = Objective(q_weight=_p(qw), mom_weight=_p(mw))
        J, dmu = ct.c_double(), ct.c_double()
        dTw = np.zeros(self.nB)
        dUw = np.zeros((self.nB, 3))
        dg0 = np.empty((self.Nv, self.nC)) if want_df0 else None
        dh0 = np.empty((self.Nv, self.nC)) if want_df0 else None
        self._check(self.L.dugks_adjoint(self._h, len(dts), _p(dts), int(checkpoint_every), ct.byref(obj),
                                         ct.byref(J), ct.byref(dmu), _p(dTw), _p(dUw), _p(dg0), _p(dh0)))
        self.step_index += len(dts)
        self.time += float(dts.sum())
        r = dict(J=J.value, dmuRef=dmu.value, dTw=dTw, dUw=dUw)
        if want_df0:
            r["dg0"], r["dh0"] = dg0, dh0
        return r

    def objective(self, dts, q_weight=None, mom_weight=None):
        """Objective of the adjoint after len(dts) fixed steps, forward only."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        qw = None if q_weight is None else np.ascontiguousarray(q_weight, dtype=np.float64).reshape(self.nF, 3)
        mw = None if mom_weight is None else np.ascontiguousarray(mom_weight, dtype=np.float64).reshape(self.nF, 3)
        obj = Objective(q_weight=_p(qw), mom_weight=_p(mw))
        J = ct.c_double()
        self._check(self.L.dugks_adjoint(self._h, len(dts), _p(dts), len(dts), ct.byref(obj), ct.byref(J),
                                         None, None, None, None, None))
        self.step_index += len(dts)
        self.time += float(dts.sum())
        return J.value

    def synchronize(self):
        self._check(self.L.dugks_synchronize(self._h))

    def cell_macros(self):
        n = self.nC
        r = dict(rho=np.empty(n), U=np.empty((n, 3)), T=np.empty(n), q=np.empty((n, 3)), tau=np.empty(n))
        self._check(self.L.dugks_get_cell_macros(self._h, _p(r["rho"]), _p(r["U"]), _p(r["T"]), _p(r["q"]),
                                                 _p(r["tau"])))
        return r

    def face_macros(self):
        n = self.nF
        r = dict(rho=np.empty(n), U=np.empty((n, 3)), T=np.empty(n), q=np.empty((n, 3)), tau=np.empty(n))
        self._check(self.L.dugks_get_face_macros(self._h, _p(r["rho"]), _p(r["U"]), _p(r["T"]), _p(r["q"]),
                                                 _p(r["tau"])))
        return r

    def rhoVol(self): return self.cell_macros()["rho"]
    def Uvol(self): return self.cell_macros()["U"]
    def Tvol(self): return self.cell_macros()["T"]
    def qVol(self): return self.cell_macros()["q"]
    def tauVol(self): return self.cell_macros()["tau"]
    def rhoSurf(self): return self.face_macros()["rho"]
    def Usurf(self): return self.face_macros()["U"]
    def Tsurf(self): return self.face_macros()["T"]
    def qSurf(self): return self.face_macros()["q"]
    def tauSurf(self): return self.face_macros()["tau"]
    def nXi(self): return self.Nv
    def nXiPerDim(self): return len(self.case.xis)
    def xiMax(self): return self.case.xi_max
    def xiMin(self): return self.case.xi_min
    def R(self): return self.case.gas.R
    def omega(self): return self.case.gas.omega
    def Tref(self): return self.case.gas.Tref
    def muRef(self): return self.case.gas.muRef
    def Pr(self): return self.case.gas.Pr
    def KInner(self): return self.case.gas.KInner

    def boundary_rho(self):
        r = np.empty(self.nB)
        self._check(self.L.dugks_get_boundary_rho(self._h, _p(r)))
        return r

    def df(self):
        g = np.empty((self.Nv, self.nC))
        h = np.empty((self.Nv, self.nC))
        self._check(self.L.dugks_get_df(self._h, _p(g), _p(h)))
        return g, h

    def set_df(self, g, h):
        g = np.ascontiguousarray(g, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        self._check(self.L.dugks_set_df(self._h, _p(g), _p(h)))

    def set_cell_macros(self, rho, U, T, q):
        f = lambda x: np.ascontiguousarray(x, dtype=np.float64)
        rho, U, T, q = f(rho), f(U), f(T), f(q)
        self._check(self.L.dugks_set_cell_macros(self._h, _p(rho), _p(U), _p(T), _p(q)))

    def checkpoint_save(self, path):
        self._check(self.L.dugks_checkpoint_save(self._h, str(path).encode()))

    def checkpoint_load(self, path):
        self._check(self.L.dugks_checkpoint_load(self._h, str(path).encode()))

    def path_info(self):
        """dict(fused, block_cells, chunk_velocities, velocity_groups) of this handle."""
        v = [ct.c_int() for _ in range(4)]
        self._check(self.L.dugks_path_info(self._h, *[ct.byref(x) for x in v]))
        return dict(zip(("fused", "block_cells", "chunk_velocities", "velocity_groups"), [x.value for x in v]))

    def velocity_set(self):
        xi = np.empty((self.Nv, 3))
        w = np.empty(self.Nv)
        self._check(self.L.dugks_velocity_set(self._h, _p(xi), _p(w)))
        return xi, w

    def profile(self, enable=True):
        self._check(self.L.dugks_profile(self._h, int(enable)))

    def kernel_times(self):
        n = self.L.dugks_kernel_count(self._h)
        secs = np.zeros(n)
        calls = np.zeros(n, dtype=np.int32)
        self._check(self.L.dugks_kernel_times(self._h, _p(secs), calls.ctypes.data_as(_I)))
        names = [self.L.dugks_kernel_name(self._h, i).decode() for i in range(n)]
        return {nm: (float(s), int(c)) for nm, s, c in zip(names, secs, calls) if c > 0}
