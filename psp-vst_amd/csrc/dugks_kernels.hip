// Staged HIP kernels of the DUGKS step: one kernel per sub-step of fvDVM::evolution
// (src/fvDVM/fvDVM/fvDVM.C:2128-2174), distributions velocity-major [v][x], face
// distributions materialised.  This path keeps the reference's data flow one-to-one and
// is the cross-check for the fused path; it also covers velocity sets without tensor
// structure.  Storage type S is double (FP64 path) or float (mixed path); all arithmetic
// is FP64.
#include <hip/hip_runtime.h>

#include "dugks_internal.h"

namespace dugks {

namespace {

constexpr int kBlock = 256;

// Shakhov equilibrium (fvDVM.C:673-679; the same expression at faces, fvDVM.C:1383-1389).
// Th is the temperature multiplying h (differs from T only for the Q1 wall read).
__device__ inline void shakhov(const Geo &g, double rho, const d3 &U, double T, const d3 &q, const d3 &xi, double Th,
                               double &gEq, double &hEq)
{
    const d3 um{U.x - xi.x, U.y - xi.y, U.z - xi.z};
    const double cSqrByRT = magsqr(um) / (g.R * T);
    const d3 xu{xi.x - U.x, xi.y - U.y, xi.z - U.z};
    const double cqBy5pRT = dot(xu, q) / (5.0 * rho * g.R * T * g.R * T);
    const double gEqBGK = rho / pow(sqrt(2.0 * kPi * g.R * T), (double)g.D) * exp(-cSqrByRT / 2.0);
    gEq = (1.0 + (1.0 - g.Pr) * cqBy5pRT * (cSqrByRT - g.D - 2.0)) * gEqBGK;
    hEq = ((g.K + 3.0 - g.D) + (1.0 - g.Pr) * cqBy5pRT * ((cSqrByRT - g.D) * (g.K + 3.0 - g.D) - 2 * g.K)) * gEqBGK *
          g.R * Th;
}

// discreteVelocity::equilibriumMaxwellByRho (discreteVelocity.C:1027-1039)
__device__ inline double maxwell_by_rho(const Geo &g, const d3 &xi, const d3 &U, double T)
{
    const d3 d{U.x - xi.x, U.y - xi.y, U.z - xi.z};
    return 1.0 / pow(sqrt(2.0 * kPi * g.R * T), (double)g.D) * exp(-magsqr(d) / (2.0 * g.R * T));
}

__device__ inline double tau_of(const Geo &g, double T, double rho)
{
    return g.muRef * exp(g.omega * log(T / g.Tref)) / rho / T / g.R;
}

// phi_c + grad(phi)_c . r with the least-squares gradient gathered through the cell's
// internal faces in ascending face order.
template <class S>
__device__ inline double recon(const Geo &g, const S *phi, int c, const d3 &r)
{
    const double pc = (double)phi[c];
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int k = g.cg_off[c]; k < g.cg_off[c + 1]; ++k) {
        const double d = (double)phi[g.cg_nb[k]] - pc;
        const d3 L = g.cg_ls[k];
        gx += L.x * d;
        gy += L.y * d;
        gz += L.z * d;
    }
    return pc + (gx * r.x + gy * r.y + gz * r.z);
}

__device__ inline d3 recon_arm(const d3 &Cf, const d3 &C, const d3 &xi, double dt)
{
    return d3{Cf.x - C.x - 0.5 * xi.x * dt, Cf.y - C.y - 0.5 * xi.y * dt, Cf.z - C.z - 0.5 * xi.z * dt};
}

// ----------------------------------------------------------------------- init kernels

template <class S>
__global__ void __launch_bounds__(kBlock) k_init_df(Geo g, Macros cm, S *gT, S *hT)
{
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (c >= g.nC) return;
    const d3 xi = g.xi[v];
    const d3 U = cm.U[c];
    const double T = cm.T[c];
    const d3 d{U.x - xi.x, U.y - xi.y, U.z - xi.z};
    const double cSqrByRT = magsqr(d) / (g.R * T);
    const double gEq = cm.rho[c] / pow(sqrt(2.0 * kPi * g.R * T), (double)g.D) * exp(-cSqrByRT / 2.0);
    const size_t k = (size_t)v * g.nC + c;
    gT[k] = (S)(gEq - g.ref_g[v]);
    hT[k] = (S)((g.K + 3.0 - g.D) * gEq * g.R * T - g.ref_h[v]);
}

template <class S>
__global__ void __launch_bounds__(kBlock) k_init_mixed_faces(Geo g, S *gS, S *hS)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (b >= g.nB || g.bf_type[b] != kMixed) return;
    const d3 xi = g.xi[v];
    const d3 Ub = g.bf_U[b];
    const double Tb = g.bf_T[b];
    const d3 d{Ub.x - xi.x, Ub.y - xi.y, Ub.z - xi.z};
    const double geq = g.bf_rho[b] / pow(sqrt(2.0 * kPi * g.R * Tb), (double)g.D) * exp(-magsqr(d) / (2.0 * g.R * Tb));
    const size_t k = (size_t)v * g.nF + g.nIF + b;
    gS[k] = (S)(geq - g.ref_g[v]);
    hS[k] = (S)((g.K + 3 - g.D) * g.R * Tb * geq - g.ref_h[v]);
}

__global__ void __launch_bounds__(kBlock) k_init_incoming(Geo g, double *in_by_rho)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= g.nB) return;
    double acc = 0.0;
    if (g.bf_type[b] == kMaxwellWall) {
        const d3 Sf = g.Sf[g.nIF + b];
        const d3 Ub = g.bf_U[b];
        const double Tb = g.bf_T[b];
        for (int v = 0; v < g.Nv; ++v) {
            const d3 xi = g.xi[v];
            const double xs = sigma_dot(xi, Sf);
            if (xs < 0) acc += -g.w[v] * xs * maxwell_by_rho(g, xi, Ub, Tb);
        }
    }
    in_by_rho[b] = acc;
}

// Usurf = linear interpolate(U) for the first Courant number (fvDVM.C:2116):
// surfaceInterpolationScheme::interpolate forms lambda*(U_P - U_N) + U_N.
__global__ void __launch_bounds__(kBlock) k_init_usurf(Geo g, Macros cm, Macros fm)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= g.nF) return;
    d3 u;
    if (f < g.nIF) {
        const d3 P = cm.U[g.owner[f]], N = cm.U[g.neighbour[f]];
        const double l = g.weights[f];
        u = d3{l * (P.x - N.x) + N.x, l * (P.y - N.y) + N.y, l * (P.z - N.z) + N.z};
    } else {
        u = g.bf_U[f - g.nIF];
    }
    fm.U[f] = u;
}

__global__ void __launch_bounds__(kBlock) k_tau(Geo g, Macros m, int n)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    m.tau[i] = tau_of(g, m.T[i], m.rho[i]);
}

// ----------------------------------------------------------------------- step kernels

// updateGHbarPvol (fvDVM.C:648-683)
template <class S>
__global__ void __launch_bounds__(kBlock)
    k_bar(Geo g, Macros cm, const S *__restrict__ gT, const S *__restrict__ hT, S *__restrict__ gB,
          S *__restrict__ hB, const double *dtp)
{
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (c >= g.nC) return;
    const double dt = *dtp;
    const double T = cm.T[c];
    double gEq, hEq;
    shakhov(g, cm.rho[c], cm.U[c], T, cm.q[c], g.xi[v], T, gEq, hEq);
    const double a = 1.5 * dt / (2.0 * cm.tau[c] + dt);
    const size_t k = (size_t)v * g.nC + c;
    // f-bar-plus kept as a deviation from the storage reference, like g~
    gB[k] = (S)((1.0 - a) * (double)gT[k] + a * (gEq - g.ref_g[v]));
    hB[k] = (S)((1.0 - a) * (double)hT[k] + a * (hEq - g.ref_h[v]));
}

// updateGrad + updateGHbarSurf (fvDVM.C:749-1029), outgoing boundary values included.
// f-bar-plus and the face values are deviations from the storage reference, which is the
// same in every cell, so the reconstruction of the deviation is the deviation of the
// reconstruction.
template <class S>
__global__ void __launch_bounds__(kBlock)
    k_face_recon(Geo g, const S *__restrict__ gB, const S *__restrict__ hB, S *__restrict__ gS, S *__restrict__ hS,
                 const double *dtp)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (f >= g.nF) return;
    const double dt = *dtp;
    const d3 xi = g.xi[v];
    const d3 Cf = g.Cf[f];
    const double xs = sigma_dot(xi, g.Sf[f]);
    const S *gv = gB + (size_t)v * g.nC;
    const S *hv = hB + (size_t)v * g.nC;
    const size_t k = (size_t)v * g.nF + f;
    if (f < g.nIF) {
        const int own = g.owner[f], nei = g.neighbour[f];
        double gf, hf;
        if (xs >= kVSmall) {
            const d3 r = recon_arm(Cf, g.C[own], xi, dt);
            gf = recon(g, gv, own, r);
            hf = recon(g, hv, own, r);
        } else if (xs < -kVSmall) {
            const d3 r = recon_arm(Cf, g.C[nei], xi, dt);
            gf = recon(g, gv, nei, r);
            hf = recon(g, hv, nei, r);
        } else {
            const d3 ro = recon_arm(Cf, g.C[own], xi, dt), rn = recon_arm(Cf, g.C[nei], xi, dt);
            gf = 0.5 * (recon(g, gv, nei, rn) + recon(g, gv, own, ro));
            hf = 0.5 * (recon(g, hv, nei, rn) + recon(g, hv, own, ro));
        }
        gS[k] = (S)gf;
        hS[k] = (S)hf;
    } else {
        const int b = f - g.nIF, t = g.bf_type[b], own = g.owner[f];
        if (t == kZeroGradient) {  // every entry takes the owner's f-bar-plus (fvDVM.C:887-891)
            const size_t kc = (size_t)v * g.nC + own;
            gS[k] = gB[kc];
            hS[k] = hB[kc];
        } else if (is_symmetry(t) ? xs > -kVSmall : xs > 0) {
            // outgoing entries (symmetry faces: tangent ones too, fvDVM.C:1013) reconstructed
            // from the owner; incoming ones are set by the wall / far-field / symmetry kernels
            const d3 r = recon_arm(Cf, g.C[own], xi, dt);
            gS[k] = (S)recon(g, gv, own, r);
            hS[k] = (S)recon(g, hv, own, r);
        }
    }
}

// incoming entries of farField / pressureIn / pressureOut faces: patch rho times the
// Maxwellian of the owner cell's U and the patch T (fvDVM.C:923-937)
template <class S>
__global__ void __launch_bounds__(kBlock) k_far_field_in(Geo g, Macros cm, const double *rho_b, S *gS, S *hS)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (b >= g.nB || !is_far_field(g.bf_type[b])) return;
    const int f = g.nIF + b;
    const d3 xi = g.xi[v];
    if (sigma_dot(xi, g.Sf[f]) > 0) return;
    const double Tp = g.bf_T[b];
    const double gp = rho_b[b] * maxwell_by_rho(g, xi, cm.U[g.owner[f]], Tp);
    const size_t k = (size_t)v * g.nF + f;
    gS[k] = (S)(gp - g.ref_g[v]);
    hS[k] = (S)(gp * (g.R * Tp) * (g.K + 3 - g.D) - g.ref_h[v]);
}

// updateGHbarSurfSymmetryIn (fvDVM.C:1100-1179, discreteVelocity.C:691-775): an incoming
// entry (xi . Sf0 <= 0, Sf0 the patch's first face) takes the value of the velocity mirrored
// about the patch axis.  The mirror of an incoming velocity is outgoing (or itself when
// tangent), so no entry read here is written here.
template <class S>
__global__ void __launch_bounds__(kBlock) k_symmetry_in(Geo g, S *gS, S *hS)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (b >= g.nB) return;
    const int axis = g.bf_axis[b];
    if (axis < 0 || sigma_dot(g.xi[v], g.bf_sf0[b]) > 0) return;
    const int f = g.nIF + b, vm = g.mirror[(size_t)axis * g.Nv + v];
    // deviations: the reference value is the same for v and its mirror
    gS[(size_t)v * g.nF + f] = gS[(size_t)vm * g.nF + f];
    hS[(size_t)v * g.nF + f] = hS[(size_t)vm * g.nF + f];
}

// outGoing accumulation (fvDVM.C:959-960, velocity order) and the wall density
// rho_w = outGoing / |inComingByRho| (calculatedMaxwellFvPatchField::evaluate, CM:158)
template <class S>
__global__ void __launch_bounds__(kBlock) k_wall_rho(Geo g, const S *__restrict__ gS, double *rho_b)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= g.nB || g.bf_type[b] != kMaxwellWall) return;
    const int f = g.nIF + b;
    const d3 Sf = g.Sf[f];
    double out = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const double xs = sigma_dot(g.xi[v], Sf);
        if (xs > 0) out += g.w[v] * xs * ((double)gS[(size_t)v * g.nF + f] + g.ref_g[v]);
    }
    rho_b[b] = out / fabs(g.in_by_rho[b]);
}

// updateGHbarSurfMaxwellWallIn (fvDVM.C:1054-1098)
template <class S>
__global__ void __launch_bounds__(kBlock) k_wall_in(Geo g, const double *rho_b, S *gS, S *hS)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (b >= g.nB || g.bf_type[b] != kMaxwellWall) return;
    const int f = g.nIF + b;
    const d3 xi = g.xi[v];
    if (sigma_dot(xi, g.Sf[f]) <= 0) {
        const double Tb = g.bf_T[b];
        const double gw = rho_b[b] * maxwell_by_rho(g, xi, g.bf_U[b], Tb);
        const size_t k = (size_t)v * g.nF + f;
        gS[k] = (S)(gw - g.ref_g[v]);
        hS[k] = (S)(gw * (g.R * Tb) * (g.K + 3 - g.D) - g.ref_h[v]);
    }
}

// updateMacroSurf (fvDVM.C:1181-1300)
template <class S>
__global__ void __launch_bounds__(kBlock)
    k_face_moments(Geo g, const S *__restrict__ gS, const S *__restrict__ hS, Macros fm, const double *dtp)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= g.nF) return;
    const double dt = *dtp;
    double rho = 0.0, rx = 0.0, ry = 0.0, rz = 0.0, rhoE = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const size_t k = (size_t)v * g.nF + f;
        const double gv = (double)gS[k] + g.ref_g[v], hv = (double)hS[k] + g.ref_h[v];
        const d3 xi = g.xi[v];
        const double w = g.w[v];
        rho += w * gv;
        const double wg = w * gv;
        rx += wg * xi.x;
        ry += wg * xi.y;
        rz += wg * xi.z;
        rhoE += 0.5 * w * (gv * magsqr(xi) + hv);
    }
    const d3 U{rx / rho, ry / rho, rz / rho};
    const double T = (rhoE - 0.5 * rho * magsqr(U)) / ((g.K + 3) / 2.0 * g.R * rho);
    const double tau = tau_of(g, T, rho);
    double qx = 0.0, qy = 0.0, qz = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const size_t k = (size_t)v * g.nF + f;
        const double gv = (double)gS[k] + g.ref_g[v], hv = (double)hS[k] + g.ref_h[v];
        const d3 xi = g.xi[v];
        const d3 c{xi.x - U.x, xi.y - U.y, xi.z - U.z};
        const double m = magsqr(c) * gv + hv;
        const double hw = 0.5 * g.w[v];
        qx += hw * c.x * m;
        qy += hw * c.y * m;
        qz += hw * c.z * m;
    }
    const double corr = 2.0 * tau / (2.0 * tau + 0.5 * dt * g.Pr);
    fm.rho[f] = rho;
    fm.U[f] = U;
    fm.T[f] = T;
    fm.tau[f] = tau;
    fm.q[f] = d3{corr * qx, corr * qy, corr * qz};
}

// updateGHsurf (fvDVM.C:1362-1460)
template <class S>
__global__ void __launch_bounds__(kBlock) k_face_relax(Geo g, Macros fm, S *gS, S *hS, const double *dtp)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (f >= g.nF) return;
    const d3 xi = g.xi[v];
    const bool bnd = f >= g.nIF;
    // patches: outgoing entries (fvDVM.C:1431); DVMsymmetry faces: then every entry once
    // more, outgoing ones included (fvDVM.C:1449-1456)
    const bool out = !bnd || sigma_dot(xi, g.Sf[f]) > 0;
    const bool sym_all = bnd && g.bf_type[f - g.nIF] == kSymmetry;
    if (!out && !sym_all) return;
    const double dt = *dtp;
    const double T = fm.T[f];
    const double Th = (bnd && g.q1) ? fm.T[g.bf_local[f - g.nIF]] : T;
    double gEq, hEq;
    shakhov(g, fm.rho[f], fm.U[f], T, fm.q[f], xi, Th, gEq, hEq);
    const double b = 0.5 * dt / (2.0 * fm.tau[f] + 0.5 * dt);
    const size_t k = (size_t)v * g.nF + f;
    double gv = (double)gS[k], hv = (double)hS[k];
    for (int pass = (out ? 0 : 1); pass < (sym_all ? 2 : 1); ++pass) {
        gv = (1.0 - b) * gv + b * (gEq - g.ref_g[v]);
        hv = (1.0 - b) * hv + b * (hEq - g.ref_h[v]);
    }
    gS[k] = (S)gv;
    hS[k] = (S)hv;
}

// updateGHtildeVol (fvDVM.C:1462-1519) as a cell-wise gather in ascending face order
template <class S>
__global__ void __launch_bounds__(kBlock)
    k_update(Geo g, const S *__restrict__ gB, const S *__restrict__ hB, const S *__restrict__ gS,
             const S *__restrict__ hS, S *gT, S *hT, const double *dtp)
{
    const int c = blockIdx.x * kBlock + threadIdx.x;
    const int v = blockIdx.y;
    if (c >= g.nC) return;
    const double dt = *dtp;
    const d3 xi = g.xi[v];
    const size_t kc = (size_t)v * g.nC + c;
    double ag = -1.0 / 3 * (double)gT[kc] + 4.0 / 3 * (double)gB[kc];  // deviations
    double ah = -1.0 / 3 * (double)hT[kc] + 4.0 / 3 * (double)hB[kc];
    const double V = g.V[c];
    const S *gv = gS + (size_t)v * g.nF;
    const S *hv = hS + (size_t)v * g.nF;
    for (int k = g.cf_off[c]; k < g.cf_off[c + 1]; ++k) {
        const int f = g.cf_face[k];
        const double xs = sigma_dot(xi, g.Sf[f]);
        const double fg = xs * ((double)gv[f] + g.ref_g[v]) * dt / V;  // fluxes of values
        const double fh = xs * ((double)hv[f] + g.ref_h[v]) * dt / V;
        if (f < g.nIF && g.owner[f] != c) {
            ag += fg;
            ah += fh;
        } else {
            ag -= fg;
            ah -= fh;
        }
    }
    gT[kc] = (S)ag;
    hT[kc] = (S)ah;
}

// Fluxes of the velocity moments 1, xi, (|xi|^2 g + h)/2 of the relaxed face DFs through
// each face: flux[f] = sum_v w_v (xi_v . S_f) m(v) (updateMacroVol, macroFlux yes in the VMESH
// build, fvDVM.C:1562-1612).  One thread per face, velocities in sequence: coalesced reads of
// the [v][face] arrays.
template <class S>
__global__ void __launch_bounds__(kBlock)
    k_face_flux(Geo g, const S *__restrict__ gS, const S *__restrict__ hS, double *flux)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= g.nF) return;
    const d3 Sf = g.Sf[f];
    double m0 = 0.0, mx = 0.0, my = 0.0, mz = 0.0, me = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const size_t k = (size_t)v * g.nF + f;
        const double gv = (double)gS[k] + g.ref_g[v], hv = (double)hS[k] + g.ref_h[v];
        const d3 xi = g.xi[v];
        const double xs = sigma_dot(xi, Sf);
        const double wg = g.w[v] * gv;
        m0 += xs * wg;
        mx += xs * wg * xi.x;
        my += xs * wg * xi.y;
        mz += xs * wg * xi.z;
        me += xs * 0.5 * g.w[v] * (gv * magsqr(xi) + hv);
    }
    double *o = flux + 5 * (size_t)f;
    o[0] = m0;
    o[1] = mx;
    o[2] = my;
    o[3] = mz;
    o[4] = me;
}

// updateMacroVol (fvDVM.C:1521-1743).  rho, rhoU, rhoE: moments of g~, h~ (macroFlux no,
// F:1528-1554), or the previous conserved variables minus the face fluxes (macroFlux yes):
//   macro_flux 1 (F:1620-1679): fluxes of the face macros with zero face stress (stressSurf_ is
//     never accumulated), rho_f (|U_f|^2 + (K+3)/2 R T_f) U_f + q_f as the energy flux, and the
//     boundary momentum flux dotted with the internal-face area vector at the patch-local
//     index (Sf there is the internal field), subtracted face by face in ascending order;
//   macro_flux 2 (F:1562-1618): the velocity-moment fluxes of k_face_flux, summed apart and
//     added at the end.
// Then U, T (F:1682-1684), tau, and q from the moments of g~, h~ about U with the
// 2 tau / (2 tau + dt Pr) correction (F:1700-1743).
template <class S>
__global__ void __launch_bounds__(kBlock)
    k_cell_moments(Geo g, const S *__restrict__ gT, const S *__restrict__ hT, Macros cm, Macros fm,
                   const double *dtp)
{
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= g.nC) return;
    const double dt = *dtp;
    double rho = 0.0, rx = 0.0, ry = 0.0, rz = 0.0, rhoE = 0.0;
    if (g.macro_flux == 0) {
        for (int v = 0; v < g.Nv; ++v) {
            const size_t k = (size_t)v * g.nC + c;
            const double gv = (double)gT[k] + g.ref_g[v], hv = (double)hT[k] + g.ref_h[v];
            const d3 xi = g.xi[v];
            const double w = g.w[v];
            rho += w * gv;
            const double wg = w * gv;
            rx += wg * xi.x;
            ry += wg * xi.y;
            rz += wg * xi.z;
            rhoE += 0.5 * w * (magsqr(xi) * gv + hv);
        }
    } else {
        const double cv = (g.K + 3) / 2.0 * g.R;
        const double V = g.V[c];
        rho = cm.rho[c];
        const d3 U0 = cm.U[c];
        rx = rho * U0.x;
        ry = rho * U0.y;
        rz = rho * U0.z;
        rhoE = rho * (0.5 * magsqr(U0) + cv * cm.T[c]);
        double s0 = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, se = 0.0;  // macro_flux 2: summed apart
        for (int k = g.cf_off[c]; k < g.cf_off[c + 1]; ++k) {
            const int f = g.cf_face[k];
            const double sgn = (f < g.nIF && g.owner[f] != c) ? 1.0 : -1.0;
            if (g.macro_flux == 1) {
                const d3 Sf = g.Sf[f];
                const d3 Sm = f < g.nIF ? Sf : g.Sf[g.bf_local[f - g.nIF]];
                const double rf = fm.rho[f];
                const d3 Uf = fm.U[f], qf = fm.q[f];
                const double un = dot(Uf, Sm);
                const double eU = rf * (magsqr(Uf) + cv * fm.T[f]);
                rho += sgn * (rf * dot(Uf, Sf) * dt / V);
                rx += sgn * (rf * Uf.x * un * dt / V);
                ry += sgn * (rf * Uf.y * un * dt / V);
                rz += sgn * (rf * Uf.z * un * dt / V);
                rhoE += sgn * ((eU * dot(Uf, Sf) + dot(qf, Sf)) * dt / V);
            } else {
                const double *F = g.face_flux + 5 * (size_t)f;
                s0 += sgn * (F[0] * dt / V);
                sx += sgn * (F[1] * dt / V);
                sy += sgn * (F[2] * dt / V);
                sz += sgn * (F[3] * dt / V);
                se += sgn * (F[4] * dt / V);
            }
        }
        rho += s0;
        rx += sx;
        ry += sy;
        rz += sz;
        rhoE += se;
    }
    const d3 U{rx / rho, ry / rho, rz / rho};
    const double T = (rhoE - 0.5 * rho * magsqr(U)) / ((g.K + 3) / 2.0 * g.R * rho);
    const double tau = tau_of(g, T, rho);
    double qx = 0.0, qy = 0.0, qz = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const size_t k = (size_t)v * g.nC + c;
        const double gv = (double)gT[k] + g.ref_g[v], hv = (double)hT[k] + g.ref_h[v];
        const d3 xi = g.xi[v];
        const d3 cc{xi.x - U.x, xi.y - U.y, xi.z - U.z};
        const double m = magsqr(cc) * gv + hv;
        const double hw = 0.5 * g.w[v];
        qx += hw * cc.x * m;
        qy += hw * cc.y * m;
        qz += hw * cc.z * m;
    }
    const double corr = 2.0 * tau / (2.0 * tau + dt * g.Pr);
    cm.rho[c] = rho;
    cm.U[c] = U;
    cm.T[c] = T;
    cm.tau[c] = tau;
    cm.q[c] = d3{corr * qx, corr * qy, corr * qz};
}

// Courant number partials over internal faces: deltaCoeffs*(|Usurf| + sqrt(D) xiMax)
// (fvDVM.C:2180-2182), block max and block sum.
__global__ void __launch_bounds__(kCflBlock) k_cfl_partials(Geo g, const d3 *Usurf, double *part)
{
    __shared__ double smax[kCflBlock], ssum[kCflBlock];
    const int f = blockIdx.x * kCflBlock + threadIdx.x;
    double u = 0.0, m = -1e300;
    if (f < g.nIF) {
        u = g.delta_coeffs[f] * (sqrt(magsqr(Usurf[f])) + g.sqrtD * g.xi_max);
        m = u;
    }
    smax[threadIdx.x] = m;
    ssum[threadIdx.x] = u;
    __syncthreads();
    for (int s = kCflBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
            ssum[threadIdx.x] += ssum[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smax[0];
        part[2 * blockIdx.x + 1] = ssum[0];
    }
}

// getCoNum + setDeltaTvar.H:34-50 on the device: dt[0] = new dt, dt[1] = Co max,
// dt[2] = Co mean (both with the old dt), hist[step] = new dt.
__global__ void __launch_bounds__(kCflBlock)
    k_dt_update(Geo g, const double *part, int nblk, double *dt, int adjust, double max_co, double max_delta_t,
                double *hist, int step)
{
    __shared__ double smax[kCflBlock], ssum[kCflBlock];
    double m = -1e300, s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kCflBlock) {
        m = fmax(m, part[2 * b]);
        s += part[2 * b + 1];
    }
    smax[threadIdx.x] = m;
    ssum[threadIdx.x] = s;
    __syncthreads();
    for (int k = kCflBlock / 2; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + k]);
            ssum[threadIdx.x] += ssum[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double dt0 = dt[0];
        const double co = smax[0] * dt0;
        const double mean = g.nIF > 0 ? ssum[0] / g.nIF * dt0 : 0.0;
        double dtn = dt0;
        if (adjust) {
            const double maxDeltaTFact = max_co / (co + kSmall);
            const double deltaTFact = fmin(fmin(maxDeltaTFact, 1.0 + 0.1 * maxDeltaTFact), 1.2);
            dtn = fmin(deltaTFact * dt0, max_delta_t);
        }
        dt[0] = dtn;
        dt[1] = co;
        dt[2] = mean;
        if (hist) hist[step] = dtn;
    }
}

// updatePressureInOutBC (fvDVM.C:1745-1821): pressureIn faces rho_p = p / (R T_p); pressureOut
// faces rho_p = rho_i + (p - rho_i R T_i) / a_i^2 and T_p = p / (R rho_i), with the owner
// cell's state and a_i^2 = R T_i (K + 5) / (K + 3)
__global__ void __launch_bounds__(kBlock) k_pressure_bc(Geo g, Macros cm, double *rho_b, double *bf_T)
{
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= g.nB) return;
    const int t = g.bf_type[b];
    if (t == kPressureIn) {
        rho_b[b] = g.bf_p[b] / g.R / bf_T[b];
    } else if (t == kPressureOut) {
        const int own = g.owner[g.nIF + b];
        const double Ti = cm.T[own], rhoi = cm.rho[own];
        const double ai = sqrt(g.R * Ti * (g.K + 5) / (g.K + 3));
        rho_b[b] = rhoi + (g.bf_p[b] - rhoi * g.R * Ti) / ai / ai;
        bf_T[b] = g.bf_p[b] / (g.R * rhoi);
    }
}

inline dim3 grid2(int n, int nv) { return dim3((n + kBlock - 1) / kBlock, nv); }
inline dim3 grid1(int n) { return dim3((n + kBlock - 1) / kBlock); }

}  // namespace

int cfl_blocks(int nIF) { return nIF > 0 ? (nIF + kCflBlock - 1) / kCflBlock : 1; }

template <class S>
void launch_init_df(const Geo &g, const Macros &cm, S *gT, S *hT, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_df<S>, grid2(g.nC, g.Nv), dim3(kBlock), 0, st, g, cm, gT, hT);
}
template <class S>
void launch_init_mixed_faces(const Geo &g, S *gS, S *hS, hipStream_t st)
{
    if (g.nB > 0) hipLaunchKernelGGL(k_init_mixed_faces<S>, grid2(g.nB, g.Nv), dim3(kBlock), 0, st, g, gS, hS);
}
void launch_init_incoming(const Geo &g, double *in_by_rho, hipStream_t st)
{
    if (g.nB > 0) hipLaunchKernelGGL(k_init_incoming, grid1(g.nB), dim3(kBlock), 0, st, g, in_by_rho);
}
void launch_init_usurf(const Geo &g, const Macros &cm, Macros fm, hipStream_t st)
{
    hipLaunchKernelGGL(k_init_usurf, grid1(g.nF), dim3(kBlock), 0, st, g, cm, fm);
}
void launch_tau(const Geo &g, Macros m, int n, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_tau, grid1(n), dim3(kBlock), 0, st, g, m, n);
}
void launch_pressure_bc(const Geo &g, const Macros &cm, double *rho_b, double *bf_T, hipStream_t st)
{
    if (g.nB > 0) hipLaunchKernelGGL(k_pressure_bc, grid1(g.nB), dim3(kBlock), 0, st, g, cm, rho_b, bf_T);
}
void launch_cfl_partials(const Geo &g, const Macros &fm, double *part, hipStream_t st)
{
    hipLaunchKernelGGL(k_cfl_partials, dim3(cfl_blocks(g.nIF)), dim3(kCflBlock), 0, st, g, fm.U, part);
}
void launch_dt_update(const Geo &g, const double *part, double *dt, int adjust, double max_co, double max_delta_t,
                      double *hist, int step, hipStream_t st)
{
    hipLaunchKernelGGL(k_dt_update, dim3(1), dim3(kCflBlock), 0, st, g, part, cfl_blocks(g.nIF), dt, adjust, max_co,
                       max_delta_t, hist, step);
}

template <class S>
void Staged<S>::bar(const Geo &g, const Macros &cm, const S *gT, const S *hT, S *gB, S *hB, const double *dt,
                    hipStream_t st)
{
    hipLaunchKernelGGL(k_bar<S>, grid2(g.nC, g.Nv), dim3(kBlock), 0, st, g, cm, gT, hT, gB, hB, dt);
}
template <class S>
void Staged<S>::face_recon(const Geo &g, const S *gB, const S *hB, S *gS, S *hS, const double *dt, hipStream_t st)
{
    hipLaunchKernelGGL(k_face_recon<S>, grid2(g.nF, g.Nv), dim3(kBlock), 0, st, g, gB, hB, gS, hS, dt);
}
template <class S>
void Staged<S>::wall_rho(const Geo &g, const S *gS, double *rho_b, hipStream_t st)
{
    if (g.nB > 0) hipLaunchKernelGGL(k_wall_rho<S>, grid1(g.nB), dim3(kBlock), 0, st, g, gS, rho_b);
}
template <class S>
void Staged<S>::wall_in(const Geo &g, const double *rho_b, S *gS, S *hS, hipStream_t st)
{
    if (g.nB > 0) hipLaunchKernelGGL(k_wall_in<S>, grid2(g.nB, g.Nv), dim3(kBlock), 0, st, g, rho_b, gS, hS);
}
template <class S>
void Staged<S>::far_field_in(const Geo &g, const Macros &cm, const double *rho_b, S *gS, S *hS, hipStream_t st)
{
    if (g.nB > 0) hipLaunchKernelGGL(k_far_field_in<S>, grid2(g.nB, g.Nv), dim3(kBlock), 0, st, g, cm, rho_b, gS, hS);
}
template <class S>
void Staged<S>::symmetry_in(const Geo &g, S *gS, S *hS, hipStream_t st)
{
    if (g.nB > 0) hipLaunchKernelGGL(k_symmetry_in<S>, grid2(g.nB, g.Nv), dim3(kBlock), 0, st, g, gS, hS);
}
template <class S>
void Staged<S>::face_moments(const Geo &g, const S *gS, const S *hS, Macros fm, const double *dt, double *cfl_part,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_face_moments<S>, grid1(g.nF), dim3(kBlock), 0, st, g, gS, hS, fm, dt);
    launch_cfl_partials(g, fm, cfl_part, st);
}
template <class S>
void Staged<S>::face_relax(const Geo &g, const Macros &fm, S *gS, S *hS, const double *dt, hipStream_t st)
{
    hipLaunchKernelGGL(k_face_relax<S>, grid2(g.nF, g.Nv), dim3(kBlock), 0, st, g, fm, gS, hS, dt);
}
template <class S>
void Staged<S>::update(const Geo &g, const S *gB, const S *hB, const S *gS, const S *hS, S *gT, S *hT,
                       const double *dt, hipStream_t st)
{
    hipLaunchKernelGGL(k_update<S>, grid2(g.nC, g.Nv), dim3(kBlock), 0, st, g, gB, hB, gS, hS, gT, hT, dt);
}
template <class S>
void Staged<S>::face_flux(const Geo &g, const S *gS, const S *hS, hipStream_t st)
{
    hipLaunchKernelGGL(k_face_flux<S>, grid1(g.nF), dim3(kBlock), 0, st, g, gS, hS, g.face_flux);
}
template <class S>
void Staged<S>::cell_moments(const Geo &g, const S *gT, const S *hT, Macros cm, const Macros &fm, const double *dt,
                             hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_moments<S>, grid1(g.nC), dim3(kBlock), 0, st, g, gT, hT, cm, fm, dt);
}

template void launch_init_df<float>(const Geo &, const Macros &, float *, float *, hipStream_t);
template void launch_init_df<double>(const Geo &, const Macros &, double *, double *, hipStream_t);
template void launch_init_mixed_faces<float>(const Geo &, float *, float *, hipStream_t);
template void launch_init_mixed_faces<double>(const Geo &, double *, double *, hipStream_t);
template struct Staged<float>;
template struct Staged<double>;

}  // namespace dugks
