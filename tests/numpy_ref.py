"""Independent numpy restatement of one DUGKS step, used only to cross-check the C oracle.

Written separately from oracle/dugks_oracle.c: vectorised over the velocity axis, the
least-squares gradient obtained by solving each cell's weighted normal equations
directly, fluxes summed with np.add.at.  Same reference algorithm
(src/fvDVM/fvDVM/fvDVM.C:648-1743), different code.  Small cases only.
"""
from __future__ import annotations

import math

import numpy as np

VSMALL = 1e-300


class NumpyDUGKS:
    def __init__(self, case):
        a = case.arrays()
        self.a = a
        self.nC, self.nIF, self.nB, self.D = a["nC"], a["nIF"], a["nB"], a["D"]
        self.nF = self.nIF + self.nB
        n = len(a["xis"])
        x, w = a["xis"], a["wts"]
        if self.D == 2:
            iy, ix = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
            ix, iy = ix.ravel(), iy.ravel()
            self.xi = np.stack([x[ix], x[iy], np.zeros(n * n)], 1)
            self.w = w[iy] * w[ix]
        else:
            iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
            ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
            self.xi = np.stack([x[ix], x[iy], x[iz]], 1)
            self.w = w[iz] * w[iy] * w[ix]
        self.R, self.Pr, self.K = a["R"], a["Pr"], a["K"]
        self.own, self.nei = a["owner"].astype(int), a["neighbour"].astype(int)
        self.Sf, self.Cf, self.C, self.V = a["Sf"], a["Cf"], a["C"], a["V"]
        self.btype = np.zeros(self.nB, dtype=int)
        self.blocal = np.zeros(self.nB, dtype=int)
        for t, s, z in zip(a["patch_type"], a["patch_start"], a["patch_size"]):
            self.btype[s:s + z] = t
            self.blocal[s:s + z] = np.arange(z)
        # xi . Sf as (x x' + y y') + z z', each operation rounded on its own (the sign of
        # round-off-sized values decides the upwind side for velocities tangent to a face)
        X, S = self.xi, self.Sf
        self.sigma = (X[:, None, 0] * S[None, :, 0] + X[:, None, 1] * S[None, :, 1]) + X[:, None, 2] * S[None, :, 2]
        self._ls_matrix()

    # ---- least squares: grad_c = A_c^-1 sum_f wt_f d_f (phi_nb - phi_c)
    def _ls_matrix(self):
        nC, nIF, nB = self.nC, self.nIF, self.nB
        C = self.C
        d = C[self.nei] - C[self.own[:nIF]]
        magSf = np.linalg.norm(self.Sf, axis=1)
        wt = magSf[:nIF] / np.sum(d * d, axis=1)
        lam = self.a["weights"]
        A = np.zeros((nC, 3, 3))
        dd = np.einsum("fi,fj->fij", d, d)
        np.add.at(A, self.own[:nIF], ((1 - lam) * wt)[:, None, None] * dd)
        np.add.at(A, self.nei, (lam * wt)[:, None, None] * dd)
        n = self.Sf[nIF:] / magSf[nIF:, None]
        db = n * np.sum(n * (self.Cf[nIF:] - C[self.own[nIF:]]), axis=1)[:, None]
        wb = magSf[nIF:] / np.sum(db * db, axis=1)
        np.add.at(A, self.own[nIF:], wb[:, None, None] * np.einsum("fi,fj->fij", db, db))
        if self.D == 2:
            A[:, 2, 2] += 1.0
        Ainv = np.linalg.inv(A)
        if self.D == 2:
            Ainv[:, 2, 2] -= 1.0
        # per internal face: contribution vectors for owner and neighbour rows
        self.Lown = ((1 - lam) * wt)[:, None] * np.einsum("fij,fj->fi", Ainv[self.own[:nIF]], d)
        self.Lnei = (lam * wt)[:, None] * np.einsum("fij,fj->fi", Ainv[self.nei], d)

    def grad(self, phi):
        """phi [Nv, nC] -> grad [Nv, nC, 3]"""
        nIF = self.nIF
        delta = phi[:, self.nei] - phi[:, self.own[:nIF]]          # [Nv, nIF]
        g = np.zeros(phi.shape + (3,))
        for k in range(3):
            gk = np.zeros_like(phi)
            np.add.at(gk.T, self.own[:nIF], (self.Lown[:, k][None, :] * delta).T)
            np.add.at(gk.T, self.nei, (self.Lnei[:, k][None, :] * delta).T)
            g[..., k] = gk
        return g

    def shakhov(self, rho, U, T, q, Th=None):
        """f^S for every (v, x): rho [n], U [n,3], T [n], q [n,3] -> (g, h) [Nv, n]"""
        R, D, K, Pr = self.R, self.D, self.K, self.Pr
        c = self.xi[:, None, :] - U[None, :, :]
        c2 = np.sum(c * c, axis=2) / (R * T)[None, :]
        cq = np.sum(c * q[None, :, :], axis=2) / (5 * rho * (R * T) ** 2)[None, :]
        gM = (rho / (2 * math.pi * R * T) ** (D / 2))[None, :] * np.exp(-c2 / 2)
        g = (1 + (1 - Pr) * cq * (c2 - D - 2)) * gM
        Th = T if Th is None else Th
        h = ((K + 3 - D) + (1 - Pr) * cq * ((c2 - D) * (K + 3 - D) - 2 * K)) * gM * (R * Th)[None, :]
        return g, h

    def maxwell(self, rho, U, T):
        R, D = self.R, self.D
        c = self.xi[:, None, :] - U[None, :, :]
        return (rho / (2 * math.pi * R * T) ** (D / 2))[None, :] * np.exp(-np.sum(c * c, 2) / (2 * R * T)[None, :])

    def tau(self, rho, T):
        a = self.a
        return a["muRef"] * (T / a["Tref"]) ** a["omega"] / (rho * T * self.R)

    def moments(self, g, h):
        w = self.w[:, None]
        rho = np.sum(w * g, 0)
        U = np.einsum("v,vx,vk->xk", self.w, g, self.xi) / rho[:, None]
        E = 0.5 * np.sum(w * (g * np.sum(self.xi ** 2, 1)[:, None] + h), 0)
        T = (E - 0.5 * rho * np.sum(U * U, 1)) / ((self.K + 3) / 2 * self.R * rho)
        c = self.xi[:, None, :] - U[None, :, :]
        q = 0.5 * np.einsum("v,vxk,vx->xk", self.w, c, np.sum(c * c, 2) * g + h)
        return rho, U, T, q

    def init(self):
        a = self.a
        rho, U, T = a["rho0"], a["U0"], a["T0"]
        g = self.maxwell(rho, U, T)
        self.g, self.h = g, (self.K + 3 - self.D) * g * (self.R * T)[None, :]
        self.rho, self.U, self.T, self.q = rho.copy(), U.copy(), T.copy(), np.zeros_like(U)
        bU, bT, brho = a["bf_U"], a["bf_T"], a["bf_rho"]
        sig_b = self.sigma[:, self.nIF:]
        Mw = self.maxwell(np.ones(self.nB), bU, bT)
        self.in_by_rho = np.sum(np.where(sig_b < 0, -self.w[:, None] * sig_b * Mw, 0.0), 0)
        gm = self.maxwell(brho, bU, bT)
        self.frozen_g = gm
        self.frozen_h = (self.K + 3 - self.D) * self.R * bT[None, :] * gm
        lam = a["weights"]
        self.Usurf = np.concatenate([lam[:, None] * U[self.own[:self.nIF]] + (1 - lam)[:, None] * U[self.nei], bU])

    def co_num(self, dt):
        u = self.a["delta_coeffs"] * (np.linalg.norm(self.Usurf[:self.nIF], axis=1) +
                                      math.sqrt(self.D) * self.a["xi_max"])
        return u.max() * dt

    def step(self, dt):
        nIF, nB = self.nIF, self.nB
        R, D, K, Pr = self.R, self.D, self.K, self.Pr
        tau = self.tau(self.rho, self.T)
        a = 1.5 * dt / (2 * tau + dt)
        gS, hS = self.shakhov(self.rho, self.U, self.T, self.q)
        gB = (1 - a) * self.g + a * gS
        hB = (1 - a) * self.h + a * hS
        self.gB, self.hB = gB, hB
        Gg, Gh = self.grad(gB), self.grad(hB)

        def rec(phi, G, cells, faces):
            r = self.Cf[faces][None, :, :] - self.C[cells][None, :, :] - 0.5 * dt * self.xi[:, None, :]
            return phi[:, cells] + np.sum(G[:, cells, :] * r, 2)

        fi = np.arange(nIF)
        sig = self.sigma[:, :nIF]
        go, ho = rec(gB, Gg, self.own[:nIF], fi), rec(hB, Gh, self.own[:nIF], fi)
        gn, hn = rec(gB, Gg, self.nei, fi), rec(hB, Gh, self.nei, fi)
        gf = np.where(sig >= VSMALL, go, np.where(sig < -VSMALL, gn, 0.5 * (gn + go)))
        hf = np.where(sig >= VSMALL, ho, np.where(sig < -VSMALL, hn, 0.5 * (hn + ho)))
        bi = np.arange(nIF, nIF + nB)
        sb = self.sigma[:, nIF:]
        gbo, hbo = rec(gB, Gg, self.own[nIF:], bi), rec(hB, Gh, self.own[nIF:], bi)
        out = np.sum(np.where(sb > 0, self.w[:, None] * sb * gbo, 0.0), 0)
        wall = self.btype == 1
        rho_w = out / np.abs(np.where(wall, self.in_by_rho, 1.0))
        bU, bT = self.a["bf_U"], self.a["bf_T"]
        Mw = self.maxwell(np.ones(nB), bU, bT)
        g_in = np.where(wall[None, :], rho_w[None, :] * Mw, self.frozen_g)
        h_in = np.where(wall[None, :], rho_w[None, :] * Mw * (R * bT)[None, :] * (K + 3 - D), self.frozen_h)
        gbf = np.where(sb > 0, gbo, g_in)
        hbf = np.where(sb > 0, hbo, h_in)
        gF = np.concatenate([gf, gbf], 1)
        hF = np.concatenate([hf, hbf], 1)
        self.rho_w = rho_w
        # face macros
        fr, fU, fT, fq = self.moments(gF, hF)
        ftau = self.tau(fr, fT)
        fq = (2 * ftau / (2 * ftau + 0.5 * dt * Pr))[:, None] * fq
        self.face = (fr, fU, fT, fq, ftau)
        self.Usurf = fU
        # relaxation
        b = 0.5 * dt / (2 * ftau + 0.5 * dt)
        Th = fT.copy()
        if self.a["q1_compat"]:
            Th[nIF:] = fT[self.blocal]
        gE, hE = self.shakhov(fr, fU, fT, fq, Th)
        relax = np.concatenate([np.ones((len(self.w), nIF), bool), sb > 0], 1)
        gF = np.where(relax, (1 - b) * gF + b * gE, gF)
        hF = np.where(relax, (1 - b) * hF + b * hE, hF)
        self.gF, self.hF = gF, hF
        # update
        gN = -1.0 / 3 * self.g + 4.0 / 3 * gB
        hN = -1.0 / 3 * self.h + 4.0 / 3 * hB
        fluxg = self.sigma * gF * dt
        fluxh = self.sigma * hF * dt
        accg = np.zeros_like(gN)
        acch = np.zeros_like(hN)
        np.add.at(accg.T, self.own, -fluxg.T)
        np.add.at(accg.T, self.nei, fluxg[:, :nIF].T)
        np.add.at(acch.T, self.own, -fluxh.T)
        np.add.at(acch.T, self.nei, fluxh[:, :nIF].T)
        self.g = gN + accg / self.V[None, :]
        self.h = hN + acch / self.V[None, :]
        rho, U, T, q = self.moments(self.g, self.h)
        tau = self.tau(rho, T)
        self.rho, self.U, self.T = rho, U, T
        self.q = (2 * tau / (2 * tau + dt * Pr))[:, None] * q
