// Hand-written reverse (adjoint) of N DUGKS steps.  No counterpart in the reference; the
// forward map it differentiates is the staged step of dugks_kernels.hip (the sub-steps of
// fvDVM::evolution, src/fvDVM/fvDVM/fvDVM.C:2128-2174) with the time steps held fixed.
//
// Forward of one step (cell macros M = rho, U, T, q, tau of the step's start):
//   f+   = (1 - a) f~ + a fS(M),  a = 1.5 dt / (2 tau + dt)                  (fvDVM.C:648-683)
//   fR   = upwind reconstruction of f+ with its least-squares gradient at x_f - xi dt/2
//          (fvDVM.C:824-963); incoming wall values rho_w M_w, rho_w = out / inComingByRho
//          (fvDVM.C:959-960, 1054-1098); frozen incoming mixed-patch values
//   Mf   = macros of the face moments of fR, q corrected with h = dt/2          (fvDVM.C:1181-1300)
//   fF   = (1 - b) fR + b fS(Mf) on internal faces and outgoing patch entries   (fvDVM.C:1362-1460)
//   f~'  = 4/3 f+ - 1/3 f~ - dt/V sum(+-(xi.S) fF)                            (fvDVM.C:1462-1519)
//   M'   = macros of the moments of f~', q corrected with dt                    (fvDVM.C:1521-1743)
// Reverse of one step, with lambda = dJ/df~' (FP64):
//   k_adj_flux    fF-bar = dt (xi.S)(lambda_nei / V_nei - lambda_own / V_own)
//   k_adj_face    per face: relaxation and equilibrium VJP -> Mf-bar, macro VJP -> moment
//                 cotangents, fR-bar; incoming wall values -> rho_w-bar, T_w-bar, U_w-bar;
//                 rho_w = out / inc -> outgoing fR-bar, inc(T_w, U_w)-bar
//   k_adj_recon   per cell: f+-bar and grad-bar gathered from the faces it is upwind of
//   k_adj_grad    per cell: least-squares gradient transposed
//   k_adj_bar     per cell: f~-bar, a-bar, equilibrium VJP -> M-bar, macro VJP -> moment
//                 cotangents -> lambda of the previous step
// muRef enters every tau (cells and faces, and the initial cells); its cotangent is
// accumulated per cell and per face over all steps and summed in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "dugks_adjoint.h"

namespace dugks {

namespace {

constexpr int kB = 256;

inline dim3 grid_x(int n) { return dim3((n + kB - 1) / kB); }
inline dim3 grid_xv(int n, int nv) { return dim3((n + kB - 1) / kB, nv); }

struct MacroBar {
    double rho, T, tau;
    d3 U, q;
};

// Shakhov equilibrium (fvDVM.C:673-679) and its vector-Jacobian product: given the
// cotangents (ag, ah) of (gEq, hEq), accumulate the cotangents of rho, U, T, q and of the
// temperature Th that multiplies h.
__device__ inline void shakhov_vjp(const Geo &g, double rho, const d3 &U, double T, const d3 &q, double Th,
                                   const d3 &xi, double ag, double ah, double &gEq, double &hEq, MacroBar &acc,
                                   double &Th_bar)
{
    const double RT = g.R * T;
    const double A = rho / pow(sqrt(2.0 * kPi * RT), (double)g.D);
    const d3 c{xi.x - U.x, xi.y - U.y, xi.z - U.z};
    const double s = magsqr(c) / RT;
    const double gM = A * exp(-s / 2.0);
    const double Q5 = 1.0 / (5.0 * rho * RT * RT);
    const double cq = dot(c, q) * Q5;
    const double al = 1.0 - g.Pr;
    const double kd = g.K + 3.0 - g.D;
    const double Pg = 1.0 + al * cq * (s - g.D - 2.0);
    const double Ph = kd + al * cq * ((s - g.D) * kd - 2.0 * g.K);
    const double RTh = g.R * Th;
    gEq = Pg * gM;
    hEq = Ph * gM * RTh;
    const double Gm = ag * Pg + ah * Ph * RTh;                                             // on gM
    const double Cq = al * gM * (ag * (s - g.D - 2.0) + ah * RTh * ((s - g.D) * kd - 2.0 * g.K));  // on cq
    const double Cs = al * gM * cq * (ag + ah * RTh * kd);                                 // on s
    acc.rho += Gm * gM / rho - Cq * cq / rho;
    acc.T += Gm * gM * (s - g.D) / (2.0 * T) - 2.0 * Cq * cq / T - Cs * s / T;
    const double fU = Gm * gM / RT - 2.0 * Cs / RT;
    acc.U.x += fU * c.x - Cq * Q5 * q.x;
    acc.U.y += fU * c.y - Cq * Q5 * q.y;
    acc.U.z += fU * c.z - Cq * Q5 * q.z;
    acc.q.x += Cq * Q5 * c.x;
    acc.q.y += Cq * Q5 * c.y;
    acc.q.z += Cq * Q5 * c.z;
    Th_bar += ah * Ph * gM * g.R;
}

// raw moments of one (value, velocity) pair, layout of the fused kernels:
// 2-D: G0, G1x, G1y, G2xx, G2xy, G2yy, G3x, G3y, H0, H1x, H1y
// 3-D: G0, G1xyz, G2xx, G2xy, G2xz, G2yy, G2yz, G2zz, G3xyz, H0, H1xyz
template <int D>
__device__ inline void add_mom(double *m, double wg, double wh, const d3 &xi)
{
    const double xx = xi.x * xi.x, yy = xi.y * xi.y;
    if (D == 2) {
        const double s2 = xx + yy;
        m[0] += wg; m[1] += wg * xi.x; m[2] += wg * xi.y;
        m[3] += wg * xx; m[4] += wg * (xi.x * xi.y); m[5] += wg * yy;
        m[6] += wg * (s2 * xi.x); m[7] += wg * (s2 * xi.y);
        m[8] += wh; m[9] += wh * xi.x; m[10] += wh * xi.y;
    } else {
        const double zz = xi.z * xi.z, s2 = xx + yy + zz;
        m[0] += wg; m[1] += wg * xi.x; m[2] += wg * xi.y; m[3] += wg * xi.z;
        m[4] += wg * xx; m[5] += wg * (xi.x * xi.y); m[6] += wg * (xi.x * xi.z);
        m[7] += wg * yy; m[8] += wg * (xi.y * xi.z); m[9] += wg * zz;
        m[10] += wg * (s2 * xi.x); m[11] += wg * (s2 * xi.y); m[12] += wg * (s2 * xi.z);
        m[13] += wh; m[14] += wh * xi.x; m[15] += wh * xi.y; m[16] += wh * xi.z;
    }
}

// cotangent of a distribution value from moment cotangents mb: (sum_k mb_k phi_k^g, sum_k mb_k phi_k^h)
template <int D>
__device__ inline void mom_dot(const double *mb, const d3 &xi, double &cg, double &ch)
{
    const double xx = xi.x * xi.x, yy = xi.y * xi.y;
    if (D == 2) {
        const double s2 = xx + yy;
        cg = mb[0] + mb[1] * xi.x + mb[2] * xi.y + mb[3] * xx + mb[4] * (xi.x * xi.y) + mb[5] * yy +
             mb[6] * (s2 * xi.x) + mb[7] * (s2 * xi.y);
        ch = mb[8] + mb[9] * xi.x + mb[10] * xi.y;
    } else {
        const double zz = xi.z * xi.z, s2 = xx + yy + zz;
        cg = mb[0] + mb[1] * xi.x + mb[2] * xi.y + mb[3] * xi.z + mb[4] * xx + mb[5] * (xi.x * xi.y) +
             mb[6] * (xi.x * xi.z) + mb[7] * yy + mb[8] * (xi.y * xi.z) + mb[9] * zz + mb[10] * (s2 * xi.x) +
             mb[11] * (s2 * xi.y) + mb[12] * (s2 * xi.z);
        ch = mb[13] + mb[14] * xi.x + mb[15] * xi.y + mb[16] * xi.z;
    }
}

// Reverse of (raw moments, muRef) -> (rho, U, T, tau, q) with q corrected by
// 2 tau / (2 tau + hc Pr) (fvDVM.C:1249-1300 at faces with hc = dt/2, fvDVM.C:1682-1743 at
// cells with hc = dt; q = 1/2 sum w c (|c|^2 g + h) expanded in raw moments).
template <int D>
__device__ inline void macros_vjp(const Geo &g, const double *m, double hc, const MacroBar &in, double *mb,
                                  double &mu_bar)
{
    double G0, G1[3], G2[6], G3[3], H0, H1[3];
    if (D == 2) {
        G0 = m[0]; G1[0] = m[1]; G1[1] = m[2]; G1[2] = 0.0;
        G2[0] = m[3]; G2[1] = m[4]; G2[2] = 0.0; G2[3] = m[5]; G2[4] = 0.0; G2[5] = 0.0;
        G3[0] = m[6]; G3[1] = m[7]; G3[2] = 0.0;
        H0 = m[8]; H1[0] = m[9]; H1[1] = m[10]; H1[2] = 0.0;
    } else {
        G0 = m[0]; G1[0] = m[1]; G1[1] = m[2]; G1[2] = m[3];
        G2[0] = m[4]; G2[1] = m[5]; G2[2] = m[6]; G2[3] = m[7]; G2[4] = m[8]; G2[5] = m[9];
        G3[0] = m[10]; G3[1] = m[11]; G3[2] = m[12];
        H0 = m[13]; H1[0] = m[14]; H1[1] = m[15]; H1[2] = m[16];
    }
    // forward
    const double rho = G0;
    const double u[3] = {G1[0] / rho, G1[1] / rho, G1[2] / rho};
    const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double trG2 = G2[0] + G2[3] + G2[5];
    const double E = 0.5 * (trG2 + H0);
    const double cK = (g.K + 3) / 2.0;
    const double num = E - 0.5 * rho * uu;
    const double den = cK * g.R * rho;
    const double T = num / den;
    const double tau = g.muRef * exp(g.omega * log(T / g.Tref)) / rho / T / g.R;
    const int ix[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    double G1u = 0.0, G2u[3], qq[3];
    for (int j = 0; j < 3; ++j) G1u += G1[j] * u[j];
    for (int k = 0; k < 3; ++k) {
        G2u[k] = 0.0;
        for (int j = 0; j < 3; ++j) G2u[k] += G2[ix[k][j]] * u[j];
    }
    for (int k = 0; k < 3; ++k)
        qq[k] = 0.5 * (G3[k] - 2.0 * G2u[k] + uu * G1[k] - trG2 * u[k] + 2.0 * u[k] * G1u - uu * u[k] * G0 + H1[k] -
                       H0 * u[k]);
    const double dd = 2.0 * tau + hc * g.Pr;
    const double corr = 2.0 * tau / dd;
    // reverse
    const double qb[3] = {in.q.x, in.q.y, in.q.z};
    double bG0 = 0, bG1[3] = {0, 0, 0}, bG2[6] = {0, 0, 0, 0, 0, 0}, bG3[3] = {0, 0, 0}, bH0 = 0,
           bH1[3] = {0, 0, 0};
    double bU[3] = {in.U.x, in.U.y, in.U.z};
    double brho = in.rho, bT = in.T, btau = in.tau, buu = 0.0, btr = 0.0, bG1u = 0.0, bG2u[3] = {0, 0, 0};
    double bcorr = 0.0, bqq[3];
    for (int k = 0; k < 3; ++k) {
        bcorr += qb[k] * qq[k];
        bqq[k] = corr * qb[k];
    }
    btau += bcorr * 2.0 * hc * g.Pr / (dd * dd);
    for (int k = 0; k < 3; ++k) {
        const double b = bqq[k];
        bG3[k] += 0.5 * b;
        bG2u[k] += -b;
        buu += 0.5 * b * (G1[k] - u[k] * G0);
        bG1[k] += 0.5 * b * uu;
        btr += -0.5 * b * u[k];
        bU[k] += 0.5 * b * (-trG2 + 2.0 * G1u - uu * G0 - H0);
        bG1u += b * u[k];
        bG0 += -0.5 * b * uu * u[k];
        bH1[k] += 0.5 * b;
        bH0 += -0.5 * b * u[k];
    }
    for (int j = 0; j < 3; ++j) {
        bG1[j] += bG1u * u[j];
        bU[j] += bG1u * G1[j];
    }
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j) {
            bG2[ix[k][j]] += bG2u[k] * u[j];
            bU[j] += bG2u[k] * G2[ix[k][j]];
        }
    // tau = muRef (T/Tref)^omega / (rho T R)
    mu_bar += btau * tau / g.muRef;
    brho += -btau * tau / rho;
    bT += btau * tau * (g.omega - 1.0) / T;
    // T = num / den, den = cK R rho
    const double bnum = bT / den;
    brho += -bT * T / den * cK * g.R;
    // num = E - rho uu / 2
    const double bE = bnum;
    brho += -0.5 * bnum * uu;
    buu += -0.5 * bnum * rho;
    btr += 0.5 * bE;
    bH0 += 0.5 * bE;
    bG2[0] += btr;
    bG2[3] += btr;
    bG2[5] += btr;
    for (int k = 0; k < 3; ++k) bU[k] += 2.0 * buu * u[k];
    for (int k = 0; k < 3; ++k) {
        bG1[k] += bU[k] / rho;
        brho += -bU[k] * u[k] / rho;
    }
    bG0 += brho;
    if (D == 2) {
        mb[0] = bG0; mb[1] = bG1[0]; mb[2] = bG1[1];
        mb[3] = bG2[0]; mb[4] = bG2[1]; mb[5] = bG2[3];
        mb[6] = bG3[0]; mb[7] = bG3[1];
        mb[8] = bH0; mb[9] = bH1[0]; mb[10] = bH1[1];
    } else {
        mb[0] = bG0; mb[1] = bG1[0]; mb[2] = bG1[1]; mb[3] = bG1[2];
        mb[4] = bG2[0]; mb[5] = bG2[1]; mb[6] = bG2[2]; mb[7] = bG2[3]; mb[8] = bG2[4]; mb[9] = bG2[5];
        mb[10] = bG3[0]; mb[11] = bG3[1]; mb[12] = bG3[2];
        mb[13] = bH0; mb[14] = bH1[0]; mb[15] = bH1[1]; mb[16] = bH1[2];
    }
}

// equilibriumMaxwellByRho (discreteVelocity.C:1027-1039) and its derivatives in Tw, Uw
__device__ inline double maxwell_d(const Geo &g, const d3 &xi, const d3 &Uw, double Tw, double &dT, d3 &dU)
{
    const double RT = g.R * Tw;
    const d3 c{xi.x - Uw.x, xi.y - Uw.y, xi.z - Uw.z};
    const double s = magsqr(c) / RT;
    const double M = 1.0 / pow(sqrt(2.0 * kPi * RT), (double)g.D) * exp(-s / 2.0);
    dT = M * (s - g.D) / (2.0 * Tw);
    dU = d3{M * c.x / RT, M * c.y / RT, M * c.z / RT};
    return M;
}

// ------------------------------------------------------------------------ kernels

// fF-bar from lambda (the flux scatter of fvDVM.C:1489-1517 transposed) plus the momentum
// flux term of the objective on the last step
__global__ void __launch_bounds__(kB)
    k_adj_flux(Geo g, const double *lg, const double *lh, double *pg, double *ph, const double *dtp,
               const d3 *mom_w)
{
    const int f = blockIdx.x * kB + threadIdx.x;
    const int v = blockIdx.y;
    if (f >= g.nF) return;
    const d3 xi = g.xi[v];
    const double xs = sigma_dot(xi, g.Sf[f]);
    const double dt = *dtp;
    const int own = g.owner[f];
    const size_t ko = (size_t)v * g.nC + own;
    double sg = -lg[ko] / g.V[own], sh = -lh[ko] / g.V[own];
    if (f < g.nIF) {
        const int nei = g.neighbour[f];
        const size_t kn = (size_t)v * g.nC + nei;
        sg += lg[kn] / g.V[nei];
        sh += lh[kn] / g.V[nei];
    }
    double ag = dt * xs * sg;
    if (mom_w) ag += g.w[v] * dot(xi, mom_w[f]) * xs;
    const size_t k = (size_t)v * g.nF + f;
    pg[k] = ag;
    ph[k] = dt * xs * sh;
}

// q1_compat: the relaxation of outgoing patch entries uses the temperature of the internal
// face at the patch-local index (fvDVM.C:1367, 1413); its cotangent per boundary face
template <class S>
__global__ void __launch_bounds__(kB)
    k_adj_th(Geo g, Macros fm, const double *pg, const double *ph, const double *dtp, double *thb)
{
    const int b = blockIdx.x * kB + threadIdx.x;
    if (b >= g.nB) return;
    const int f = g.nIF + b;
    double th_bar = 0.0;
    if (g.q1) {
        const double dt = *dtp;
        const double bb = 0.5 * dt / (2.0 * fm.tau[f] + 0.5 * dt);
        const d3 Sf = g.Sf[f];
        const double Th = fm.T[g.bf_local[b]];
        for (int v = 0; v < g.Nv; ++v) {
            const d3 xi = g.xi[v];
            if (!(sigma_dot(xi, Sf) > 0)) continue;
            const size_t k = (size_t)v * g.nF + f;
            double gE, hE;
            MacroBar dummy{};
            shakhov_vjp(g, fm.rho[f], fm.U[f], fm.T[f], fm.q[f], Th, xi, bb * pg[k], bb * ph[k], gE, hE, dummy,
                        th_bar);
        }
    }
    thb[b] = th_bar;
}

struct FaceAdjArgs {
    const double *thb;       // [nB]
    const int *p_start, *p_size;  // patches (boundary-face ranges) for the q1 gather
    int n_patch;
    const d3 *q_w;           // [nF] objective heat-flux weights (last step) or null
    double *mu_f;            // [nF] accumulated muRef cotangent
    double *Tw_bar;          // [nB]
    d3 *Uw_bar;              // [nB]
};

template <class S, int D>
__global__ void __launch_bounds__(kB)
    k_adj_face(Geo g, Macros fm, const S *__restrict__ gR, const S *__restrict__ hR, const double *pg,
               const double *ph, double *rg, double *rh, const double *rho_b, const double *dtp, FaceAdjArgs a)
{
    constexpr int NM = D == 3 ? 17 : 11;
    const int f = blockIdx.x * kB + threadIdx.x;
    if (f >= g.nF) return;
    const bool bnd = f >= g.nIF;
    const int b = f - g.nIF;
    const int type = bnd ? g.bf_type[b] : 0;
    const double dt = *dtp;
    const d3 Sf = g.Sf[f];
    const double rho = fm.rho[f], T = fm.T[f], tau = fm.tau[f];
    const d3 U = fm.U[f], q = fm.q[f];
    const bool q1b = bnd && g.q1;
    const double Th = q1b ? fm.T[g.bf_local[b]] : T;
    const double bb = 0.5 * dt / (2.0 * tau + 0.5 * dt);
    // pass 1: relaxation / equilibrium cotangents, raw moments of the face distribution
    MacroBar mf{};
    double th_bar = 0.0, b_bar = 0.0;
    double m[NM];
    for (int k = 0; k < NM; ++k) m[k] = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const d3 xi = g.xi[v];
        const size_t k = (size_t)v * g.nF + f;
        // recorded face values are deviations from the storage reference
        const double gr = (double)gR[k] + g.ref_g[v], hr = (double)hR[k] + g.ref_h[v];
        const double w = g.w[v];
        add_mom<D>(m, w * gr, w * hr, xi);
        if (bnd && !(sigma_dot(xi, Sf) > 0)) continue;
        const double ag = pg[k], ah = ph[k];
        double gE, hE;
        shakhov_vjp(g, rho, U, T, q, Th, xi, bb * ag, bb * ah, gE, hE, mf, th_bar);
        b_bar += (gE - gr) * ag + (hE - hr) * ah;
    }
    mf.tau += b_bar * (-dt) / ((2.0 * tau + 0.5 * dt) * (2.0 * tau + 0.5 * dt));
    if (!q1b) mf.T += th_bar;
    if (!bnd && g.q1)
        for (int p = 0; p < a.n_patch; ++p)
            if (f < a.p_size[p]) mf.T += a.thb[a.p_start[p] + f];
    if (a.q_w) {
        const d3 qw = a.q_w[f];
        mf.q.x += qw.x;
        mf.q.y += qw.y;
        mf.q.z += qw.z;
    }
    double mb[NM];
    double mu_bar = 0.0;
    macros_vjp<D>(g, m, 0.5 * dt, mf, mb, mu_bar);
    a.mu_f[f] += mu_bar;
    // pass 2: cotangents of the reconstructed / incoming face values
    const bool wall = type == kMaxwellWall;
    d3 Uw{0, 0, 0};
    double Tw = 1.0, rw = 0.0, kdRT = 0.0;
    if (wall) {
        Uw = g.bf_U[b];
        Tw = g.bf_T[b];
        rw = rho_b[b];
        kdRT = g.R * Tw * (g.K + 3 - g.D);
    }
    double rw_bar = 0.0, Tw_bar = 0.0;
    d3 Uw_bar{0, 0, 0};
    for (int v = 0; v < g.Nv; ++v) {
        const d3 xi = g.xi[v];
        const size_t k = (size_t)v * g.nF + f;
        double cg, ch;
        mom_dot<D>(mb, xi, cg, ch);
        const double w = g.w[v];
        cg *= w;
        ch *= w;
        const bool relaxed = !bnd || sigma_dot(xi, Sf) > 0;
        if (relaxed) {
            rg[k] = (1.0 - bb) * pg[k] + cg;
            rh[k] = (1.0 - bb) * ph[k] + ch;
        } else if (wall) {
            // incoming g = rho_w M(Uw, Tw), h = g R Tw (K + 3 - D)
            const double ag = pg[k] + cg, ah = ph[k] + ch;
            double dT;
            d3 dU;
            const double M = maxwell_d(g, xi, Uw, Tw, dT, dU);
            const double c = ag + kdRT * ah;
            rw_bar += M * c;
            Tw_bar += rw * (dT * c + M * g.R * (g.K + 3 - g.D) * ah);
            Uw_bar.x += rw * dU.x * c;
            Uw_bar.y += rw * dU.y * c;
            Uw_bar.z += rw * dU.z * c;
        }
    }
    if (wall) {
        // rho_w = out / |inc|, out = sum_{xi.S > 0} w (xi.S) fR, inc = sum_{xi.S < 0} -w (xi.S) M
        const double inc = fabs(g.in_by_rho[b]);
        const double inc_bar = -rw_bar * rw / inc;
        double dinc_T = 0.0;
        d3 dinc_U{0, 0, 0};
        for (int v = 0; v < g.Nv; ++v) {
            const d3 xi = g.xi[v];
            const double xs = sigma_dot(xi, Sf);
            if (xs > 0) {
                rg[(size_t)v * g.nF + f] += rw_bar * g.w[v] * xs / inc;
            } else if (xs < 0) {
                double dT;
                d3 dU;
                (void)maxwell_d(g, xi, Uw, Tw, dT, dU);
                const double c = -g.w[v] * xs;
                dinc_T += c * dT;
                dinc_U.x += c * dU.x;
                dinc_U.y += c * dU.y;
                dinc_U.z += c * dU.z;
            }
        }
        Tw_bar += inc_bar * dinc_T;
        Uw_bar.x += inc_bar * dinc_U.x;
        Uw_bar.y += inc_bar * dinc_U.y;
        Uw_bar.z += inc_bar * dinc_U.z;
        a.Tw_bar[b] += Tw_bar;
        d3 u = a.Uw_bar[b];
        a.Uw_bar[b] = d3{u.x + Uw_bar.x, u.y + Uw_bar.y, u.z + Uw_bar.z};
    }
}

// f+-bar = 4/3 lambda + sum over the faces the cell reconstructs, grad-bar likewise with the
// reconstruction arm (the upwind choice of fvDVM.C:829-863 and the outgoing patch entries)
template <int D>
__global__ void __launch_bounds__(kB)
    k_adj_recon(Geo g, const double *lg, const double *lh, const double *rg, const double *rh, double *bg,
                double *bh, double *gg, double *gh, const double *dtp)
{
    const int c = blockIdx.x * kB + threadIdx.x;
    const int v = blockIdx.y;
    if (c >= g.nC) return;
    const double dt = *dtp;
    const d3 xi = g.xi[v];
    const d3 C = g.C[c];
    const size_t kc = (size_t)v * g.nC + c;
    double ag = 4.0 / 3 * lg[kc], ah = 4.0 / 3 * lh[kc];
    double grg[3] = {0, 0, 0}, grh[3] = {0, 0, 0};
    for (int k = g.cf_off[c]; k < g.cf_off[c + 1]; ++k) {
        const int f = g.cf_face[k];
        const double xs = sigma_dot(xi, g.Sf[f]);
        double wc;
        if (f < g.nIF) {
            if (g.owner[f] == c)
                wc = xs >= kVSmall ? 1.0 : (xs < -kVSmall ? 0.0 : 0.5);
            else
                wc = xs < -kVSmall ? 1.0 : (xs >= kVSmall ? 0.0 : 0.5);
        } else {
            wc = xs > 0 ? 1.0 : 0.0;
        }
        if (wc == 0.0) continue;
        const size_t kf = (size_t)v * g.nF + f;
        const double tg = wc * rg[kf], th = wc * rh[kf];
        const d3 Cf = g.Cf[f];
        const double r[3] = {Cf.x - C.x - 0.5 * xi.x * dt, Cf.y - C.y - 0.5 * xi.y * dt, Cf.z - C.z - 0.5 * xi.z * dt};
        ag += tg;
        ah += th;
        for (int j = 0; j < 3; ++j) {
            grg[j] += tg * r[j];
            grh[j] += th * r[j];
        }
    }
    bg[kc] = ag;
    bh[kc] = ah;
    for (int j = 0; j < 3; ++j) {
        gg[kc * 3 + j] = grg[j];
        gh[kc * 3 + j] = grh[j];
    }
}

// least-squares gradient grad_c = sum_k L_ck (f_nb(k) - f_c) transposed: cell c collects
// -L_ck . gbar_c and L_{nb->c} . gbar_nb
__global__ void __launch_bounds__(kB)
    k_adj_grad(Geo g, const int *cg_rev, const double *gg, const double *gh, double *bg, double *bh)
{
    const int c = blockIdx.x * kB + threadIdx.x;
    const int v = blockIdx.y;
    if (c >= g.nC) return;
    const size_t kc = (size_t)v * g.nC + c;
    const double *Gc = gg + kc * 3, *Hc = gh + kc * 3;
    double ag = 0.0, ah = 0.0;
    for (int k = g.cg_off[c]; k < g.cg_off[c + 1]; ++k) {
        const int nb = g.cg_nb[k];
        const d3 L = g.cg_ls[k], Lr = g.cg_ls[cg_rev[k]];
        const size_t kn = ((size_t)v * g.nC + nb) * 3;
        ag += -(L.x * Gc[0] + L.y * Gc[1] + L.z * Gc[2]) + (Lr.x * gg[kn] + Lr.y * gg[kn + 1] + Lr.z * gg[kn + 2]);
        ah += -(L.x * Hc[0] + L.y * Hc[1] + L.z * Hc[2]) + (Lr.x * gh[kn] + Lr.y * gh[kn + 1] + Lr.z * gh[kn + 2]);
    }
    bg[kc] += ag;
    bh[kc] += ah;
}

// f~-bar through f+ = (1 - a) f~ + a fS(M), the -1/3 f~ term of the update and (n >= 1) the
// macros of f~; lambda is updated in place to dJ/df~ of the step's start
template <class S, int D>
__global__ void __launch_bounds__(kB)
    k_adj_bar(Geo g, Macros cm, const S *__restrict__ gT, const S *__restrict__ hT, const double *bg,
              const double *bh, double *lg, double *lh, const double *dtp, double hc_prev, int first, double *mu_c)
{
    constexpr int NM = D == 3 ? 17 : 11;
    const int c = blockIdx.x * kB + threadIdx.x;
    if (c >= g.nC) return;
    const double dt = *dtp;
    const double rho = cm.rho[c], T = cm.T[c], tau = cm.tau[c];
    const d3 U = cm.U[c], q = cm.q[c];
    const double a = 1.5 * dt / (2.0 * tau + dt);
    MacroBar mc{};
    double th_bar = 0.0, a_bar = 0.0;
    double m[NM];
    for (int k = 0; k < NM; ++k) m[k] = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const d3 xi = g.xi[v];
        const size_t k = (size_t)v * g.nC + c;
        // stored tilde values are deviations from the per-velocity reference
        const double gt = (double)gT[k] + g.ref_g[v], ht = (double)hT[k] + g.ref_h[v];
        const double ag = bg[k], ah = bh[k];
        double gE, hE;
        shakhov_vjp(g, rho, U, T, q, T, xi, a * ag, a * ah, gE, hE, mc, th_bar);
        a_bar += (gE - gt) * ag + (hE - ht) * ah;
        lg[k] = (1.0 - a) * ag - lg[k] / 3.0;
        lh[k] = (1.0 - a) * ah - lh[k] / 3.0;
        if (!first) add_mom<D>(m, g.w[v] * gt, g.w[v] * ht, xi);
    }
    mc.T += th_bar;
    mc.tau += a_bar * (-3.0 * dt) / ((2.0 * tau + dt) * (2.0 * tau + dt));
    double mu_bar = 0.0;
    if (first) {
        // initial macros are inputs; tau = tau(rho0, T0; muRef)
        mu_bar += mc.tau * tau / g.muRef;
    } else {
        double mb[NM];
        macros_vjp<D>(g, m, hc_prev, mc, mb, mu_bar);
        for (int v = 0; v < g.Nv; ++v) {
            const size_t k = (size_t)v * g.nC + c;
            double cg, ch;
            mom_dot<D>(mb, g.xi[v], cg, ch);
            lg[k] += g.w[v] * cg;
            lh[k] += g.w[v] * ch;
        }
    }
    mu_c[c] += mu_bar;
}

// objective per face on the last step's face values
template <class S>
__global__ void __launch_bounds__(kB)
    k_objective(Geo g, Macros fm, const S *__restrict__ gS, const d3 *q_w, const d3 *mom_w, double *jf)
{
    const int f = blockIdx.x * kB + threadIdx.x;
    if (f >= g.nF) return;
    double j = 0.0;
    if (q_w) j += dot(fm.q[f], q_w[f]);
    if (mom_w) {
        const d3 e = mom_w[f], Sf = g.Sf[f];
        double s = 0.0;
        for (int v = 0; v < g.Nv; ++v) {
            const d3 xi = g.xi[v];
            s += g.w[v] * dot(xi, e) * sigma_dot(xi, Sf) * ((double)gS[(size_t)v * g.nF + f] + g.ref_g[v]);
        }
        j += s;
    }
    jf[f] = j;
}

__global__ void __launch_bounds__(kB) k_zero(double *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kB + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------------ host driver

struct Buf {
    std::vector<void *> ptrs;
    ~Buf()
    {
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
    template <class T>
    bool get(T **p, size_t n, std::string &err)
    {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) {
            err = "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed (adjoint buffers)";
            return false;
        }
        ptrs.push_back(q);
        *p = (T *)q;
        return true;
    }
};

struct MacroBufs {
    double *rho, *T, *tau;
    d3 *U, *q;
    Macros view() const { return Macros{rho, U, T, q, tau}; }
};

bool alloc_macros(Buf &B, MacroBufs &m, int n, std::string &err)
{
    return B.get(&m.rho, n, err) && B.get(&m.T, n, err) && B.get(&m.tau, n, err) && B.get(&m.U, n, err) &&
           B.get(&m.q, n, err);
}

bool copy_macros(const Macros &dst, const Macros &src, int n, hipStream_t st)
{
    return hipMemcpyAsync(dst.rho, src.rho, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
           hipMemcpyAsync(dst.T, src.T, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
           hipMemcpyAsync(dst.tau, src.tau, sizeof(double) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
           hipMemcpyAsync(dst.U, src.U, sizeof(d3) * n, hipMemcpyDeviceToDevice, st) == hipSuccess &&
           hipMemcpyAsync(dst.q, src.q, sizeof(d3) * n, hipMemcpyDeviceToDevice, st) == hipSuccess;
}

template <class S>
struct Slot {
    S *g, *h;
    MacroBufs m;
};

}  // namespace

template <class S>
int run_adjoint(const Geo &g, DevState &d, hipStream_t st, const AdjointIO &io, std::string &err)
{
    const int N = io.n_steps, K = io.ckpt;
    const int nC = g.nC, nF = g.nF, nB = g.nB, Nv = g.Nv;
    const size_t ncell = (size_t)Nv * nC, nface = (size_t)Nv * nF;
    Buf B;
#define ADJ_CHK(x)                                              \
    do {                                                        \
        if (!(x)) {                                             \
            if (err.empty()) err = "HIP call failed: " #x;      \
            return 1;                                           \
        }                                                       \
    } while (0)

    // constants and reverse-stencil table
    double *d_dts = nullptr;
    ADJ_CHK(B.get(&d_dts, N, err));
    ADJ_CHK(hipMemcpyAsync(d_dts, io.dts, sizeof(double) * N, hipMemcpyHostToDevice, st) == hipSuccess);
    std::vector<int> cg_off(nC + 1), bf_local(nB);
    ADJ_CHK(hipMemcpyAsync(cg_off.data(), g.cg_off, sizeof(int) * (nC + 1), hipMemcpyDeviceToHost, st) == hipSuccess);
    if (nB) ADJ_CHK(hipMemcpyAsync(bf_local.data(), g.bf_local, sizeof(int) * nB, hipMemcpyDeviceToHost, st) == hipSuccess);
    ADJ_CHK(hipStreamSynchronize(st) == hipSuccess);
    std::vector<int> cg_nb(cg_off[nC]);
    if (!cg_nb.empty())
        ADJ_CHK(hipMemcpy(cg_nb.data(), g.cg_nb, sizeof(int) * cg_nb.size(), hipMemcpyDeviceToHost) == hipSuccess);
    std::vector<int> rev(cg_nb.size(), -1);
    for (int c = 0; c < nC; ++c)
        for (int k = cg_off[c]; k < cg_off[c + 1]; ++k) {
            const int nb = cg_nb[k];
            for (int j = cg_off[nb]; j < cg_off[nb + 1]; ++j)
                if (cg_nb[j] == c) {
                    rev[k] = j;
                    break;
                }
            if (rev[k] < 0) {
                err = "least-squares stencil is not symmetric";
                return 1;
            }
        }
    int *d_rev = nullptr;
    ADJ_CHK(B.get(&d_rev, rev.size(), err));
    if (!rev.empty())
        ADJ_CHK(hipMemcpy(d_rev, rev.data(), sizeof(int) * rev.size(), hipMemcpyHostToDevice) == hipSuccess);
    std::vector<int> p_start, p_size;
    for (int b = 0; b < nB; ++b)
        if (bf_local[b] == 0) {
            p_start.push_back(b);
            p_size.push_back(0);
        }
    for (int b = 0; b < nB; ++b) {
        int p = (int)(std::upper_bound(p_start.begin(), p_start.end(), b) - p_start.begin()) - 1;
        p_size[p] = std::max(p_size[p], bf_local[b] + 1);
    }
    int *d_pstart = nullptr, *d_psize = nullptr;
    ADJ_CHK(B.get(&d_pstart, p_start.size(), err) && B.get(&d_psize, p_size.size(), err));
    if (!p_start.empty()) {
        ADJ_CHK(hipMemcpy(d_pstart, p_start.data(), sizeof(int) * p_start.size(), hipMemcpyHostToDevice) == hipSuccess);
        ADJ_CHK(hipMemcpy(d_psize, p_size.data(), sizeof(int) * p_size.size(), hipMemcpyHostToDevice) == hipSuccess);
    }
    d3 *d_qw = nullptr, *d_mw = nullptr;
    if (io.q_weight) {
        ADJ_CHK(B.get(&d_qw, nF, err));
        ADJ_CHK(hipMemcpy(d_qw, io.q_weight, sizeof(d3) * nF, hipMemcpyHostToDevice) == hipSuccess);
    }
    if (io.mom_weight) {
        ADJ_CHK(B.get(&d_mw, nF, err));
        ADJ_CHK(hipMemcpy(d_mw, io.mom_weight, sizeof(d3) * nF, hipMemcpyHostToDevice) == hipSuccess);
    }

    // forward work buffers (face values: gS after relaxation, gR before)
    S *gB, *hB, *gS, *hS, *gR, *hR;
    ADJ_CHK(B.get(&gB, ncell, err) && B.get(&hB, ncell, err) && B.get(&gS, nface, err) && B.get(&hS, nface, err) &&
            B.get(&gR, nface, err) && B.get(&hR, nface, err));
    launch_init_mixed_faces<S>(g, gS, hS, st);
    // states: checkpoints every K steps and one segment of K consecutive steps (none when
    // only the objective is asked for)
    const bool want_grad = io.dmu || io.dTw || io.dUw || io.dg0 || io.dh0;
    const int nseg = want_grad ? (N + K - 1) / K : 0;
    std::vector<Slot<S>> ck(nseg), seg(want_grad ? K : 0);
    for (auto *v : {&ck, &seg})
        for (auto &s : *v) ADJ_CHK(B.get(&s.g, ncell, err) && B.get(&s.h, ncell, err) && alloc_macros(B, s.m, nC, err));
    // adjoint buffers
    double *lg = nullptr, *lh = nullptr, *pg = nullptr, *ph = nullptr, *rg = nullptr, *rh = nullptr, *bg = nullptr,
           *bh = nullptr, *gg = nullptr, *gh = nullptr, *thb = nullptr, *mu_f = nullptr, *mu_c = nullptr,
           *Tw_bar = nullptr, *jf = nullptr;
    d3 *Uw_bar = nullptr;
    ADJ_CHK(B.get(&jf, nF, err));
    if (want_grad) {
        ADJ_CHK(B.get(&lg, ncell, err) && B.get(&lh, ncell, err) && B.get(&pg, nface, err) &&
                B.get(&ph, nface, err) && B.get(&rg, nface, err) && B.get(&rh, nface, err) && B.get(&bg, ncell, err) &&
                B.get(&bh, ncell, err) && B.get(&gg, 3 * ncell, err) && B.get(&gh, 3 * ncell, err) &&
                B.get(&thb, nB, err) && B.get(&mu_f, nF, err) && B.get(&mu_c, nC, err) && B.get(&Tw_bar, nB, err) &&
                B.get(&Uw_bar, nB, err));
        auto zero = [&](double *p, size_t n) {
            if (n) hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + kB - 1) / kB)), dim3(kB), 0, st, p, n);
        };
        zero(lg, ncell);
        zero(lh, ncell);
        zero(mu_f, nF);
        zero(mu_c, nC);
        zero(Tw_bar, nB);
        zero((double *)Uw_bar, 3 * (size_t)nB);
        zero(rg, nface);
        zero(rh, nface);
    }

    Macros fm = d.face;
    // one forward step n on the state (gt, ht, cm): cm holds the step's start macros on entry
    // and the end macros on exit; gR/hR keep the face values before relaxation
    auto step = [&](int n, S *gt, S *ht, const Macros &cm) {
        const double *dt = d_dts + n;
        Staged<S>::bar(g, cm, gt, ht, gB, hB, dt, st);
        Staged<S>::face_recon(g, gB, hB, gS, hS, dt, st);
        Staged<S>::wall_rho(g, gS, d.rho_b, st);
        Staged<S>::wall_in(g, d.rho_b, gS, hS, st);
        (void)hipMemcpyAsync(gR, gS, nface * sizeof(S), hipMemcpyDeviceToDevice, st);
        (void)hipMemcpyAsync(hR, hS, nface * sizeof(S), hipMemcpyDeviceToDevice, st);
        Staged<S>::face_moments(g, gS, hS, fm, dt, d.cfl_part, st);
        Staged<S>::face_relax(g, fm, gS, hS, dt, st);
        Staged<S>::update(g, gB, hB, gS, hS, gt, ht, dt, st);
        Staged<S>::cell_moments(g, gt, ht, cm, fm, dt, st);
    };
    auto copy_state = [&](S *g2, S *h2, const Macros &m2, const S *g1, const S *h1, const Macros &m1) {
        return hipMemcpyAsync(g2, g1, ncell * sizeof(S), hipMemcpyDeviceToDevice, st) == hipSuccess &&
               hipMemcpyAsync(h2, h1, ncell * sizeof(S), hipMemcpyDeviceToDevice, st) == hipSuccess &&
               copy_macros(m2, m1, nC, st);
    };

    // forward sweep with checkpoints, on the handle's own state
    S *gT = (S *)d.gT, *hT = (S *)d.hT;
    for (int n = 0; n < N; ++n) {
        if (want_grad && n % K == 0)
            ADJ_CHK(copy_state(ck[n / K].g, ck[n / K].h, ck[n / K].m.view(), gT, hT, d.cell));
        step(n, gT, hT, d.cell);
    }
    hipLaunchKernelGGL(k_objective<S>, grid_x(nF), dim3(kB), 0, st, g, fm, gS, d_qw, d_mw, jf);
    std::vector<double> jh(nF);
    ADJ_CHK(hipMemcpyAsync(jh.data(), jf, sizeof(double) * nF, hipMemcpyDeviceToHost, st) == hipSuccess);
    ADJ_CHK(hipStreamSynchronize(st) == hipSuccess);
    ADJ_CHK(hipGetLastError() == hipSuccess);
    double J = 0.0;
    for (double x : jh) J += x;
    if (io.J) *io.J = J;
    if (!want_grad) return 0;   // objective only: the handle already holds the state after N steps
    // the reverse sweep uses the handle's buffers as scratch: keep the state after N steps
    // (distributions, cell and face macros, wall density, CFL partials) and restore it at the end
    Slot<S> fin;
    MacroBufs fsave;
    double *rho_b_save = nullptr, *cfl_save = nullptr;
    const int ncfl = 2 * cfl_blocks(g.nIF);
    ADJ_CHK(B.get(&fin.g, ncell, err) && B.get(&fin.h, ncell, err) && alloc_macros(B, fin.m, nC, err));
    ADJ_CHK(alloc_macros(B, fsave, nF, err) && B.get(&rho_b_save, nB, err) && B.get(&cfl_save, ncfl, err));
    ADJ_CHK(copy_state(fin.g, fin.h, fin.m.view(), gT, hT, d.cell));
    ADJ_CHK(copy_macros(fsave.view(), fm, nF, st));
    ADJ_CHK(hipMemcpyAsync(rho_b_save, d.rho_b, sizeof(double) * nB, hipMemcpyDeviceToDevice, st) == hipSuccess);
    ADJ_CHK(hipMemcpyAsync(cfl_save, d.cfl_part, sizeof(double) * ncfl, hipMemcpyDeviceToDevice, st) == hipSuccess);

    // reverse sweep
    FaceAdjArgs fa{thb, d_pstart, d_psize, (int)p_start.size(), nullptr, mu_f, Tw_bar, Uw_bar};
    for (int s = nseg - 1; s >= 0; --s) {
        const int n0 = s * K, n1 = std::min(N, n0 + K);
        ADJ_CHK(copy_state(seg[0].g, seg[0].h, seg[0].m.view(), ck[s].g, ck[s].h, ck[s].m.view()));
        for (int n = n0 + 1; n < n1; ++n) {
            Slot<S> &a = seg[n - 1 - n0], &b = seg[n - n0];
            ADJ_CHK(copy_state(b.g, b.h, b.m.view(), a.g, a.h, a.m.view()));
            step(n - 1, b.g, b.h, b.m.view());
        }
        for (int n = n1 - 1; n >= n0; --n) {
            Slot<S> &sl = seg[n - n0];
            // step n's intermediates (gR, fm, rho_b) from a scratch copy of its start state in
            // the handle's own buffers
            ADJ_CHK(copy_state(gT, hT, d.cell, sl.g, sl.h, sl.m.view()));
            step(n, gT, hT, d.cell);
            const double *dt = d_dts + n;
            const bool last = n == N - 1;
            hipLaunchKernelGGL(k_adj_flux, grid_xv(nF, Nv), dim3(kB), 0, st, g, lg, lh, pg, ph, dt,
                               last ? d_mw : nullptr);
            if (nB) hipLaunchKernelGGL(k_adj_th<S>, grid_x(nB), dim3(kB), 0, st, g, fm, pg, ph, dt, thb);
            fa.q_w = last ? d_qw : nullptr;
            if (g.D == 3) {
                hipLaunchKernelGGL((k_adj_face<S, 3>), grid_x(nF), dim3(kB), 0, st, g, fm, gR, hR, pg, ph, rg, rh,
                                   d.rho_b, dt, fa);
                hipLaunchKernelGGL(k_adj_recon<3>, grid_xv(nC, Nv), dim3(kB), 0, st, g, lg, lh, rg, rh, bg, bh, gg,
                                   gh, dt);
            } else {
                hipLaunchKernelGGL((k_adj_face<S, 2>), grid_x(nF), dim3(kB), 0, st, g, fm, gR, hR, pg, ph, rg, rh,
                                   d.rho_b, dt, fa);
                hipLaunchKernelGGL(k_adj_recon<2>, grid_xv(nC, Nv), dim3(kB), 0, st, g, lg, lh, rg, rh, bg, bh, gg,
                                   gh, dt);
            }
            hipLaunchKernelGGL(k_adj_grad, grid_xv(nC, Nv), dim3(kB), 0, st, g, d_rev, gg, gh, bg, bh);
            const double hc_prev = n > 0 ? io.dts[n - 1] : 0.0;
            if (g.D == 3)
                hipLaunchKernelGGL((k_adj_bar<S, 3>), grid_x(nC), dim3(kB), 0, st, g, sl.m.view(), sl.g, sl.h, bg, bh,
                                   lg, lh, dt, hc_prev, n == 0 ? 1 : 0, mu_c);
            else
                hipLaunchKernelGGL((k_adj_bar<S, 2>), grid_x(nC), dim3(kB), 0, st, g, sl.m.view(), sl.g, sl.h, bg, bh,
                                   lg, lh, dt, hc_prev, n == 0 ? 1 : 0, mu_c);
            ADJ_CHK(hipGetLastError() == hipSuccess);
        }
    }
    ADJ_CHK(hipStreamSynchronize(st) == hipSuccess);

    // outputs
    if (io.dmu) {
        std::vector<double> a(nF), b(nC);
        ADJ_CHK(hipMemcpy(a.data(), mu_f, sizeof(double) * nF, hipMemcpyDeviceToHost) == hipSuccess);
        ADJ_CHK(hipMemcpy(b.data(), mu_c, sizeof(double) * nC, hipMemcpyDeviceToHost) == hipSuccess);
        double s = 0.0;
        for (double x : a) s += x;
        for (double x : b) s += x;
        *io.dmu = s;
    }
    if (io.dTw && nB) ADJ_CHK(hipMemcpy(io.dTw, Tw_bar, sizeof(double) * nB, hipMemcpyDeviceToHost) == hipSuccess);
    if (io.dUw && nB) ADJ_CHK(hipMemcpy(io.dUw, Uw_bar, sizeof(d3) * nB, hipMemcpyDeviceToHost) == hipSuccess);
    if (io.dg0) ADJ_CHK(hipMemcpy(io.dg0, lg, sizeof(double) * ncell, hipMemcpyDeviceToHost) == hipSuccess);
    if (io.dh0) ADJ_CHK(hipMemcpy(io.dh0, lh, sizeof(double) * ncell, hipMemcpyDeviceToHost) == hipSuccess);

    // leave the handle at the state after N steps
    ADJ_CHK(copy_state(gT, hT, d.cell, fin.g, fin.h, fin.m.view()));
    ADJ_CHK(copy_macros(fm, fsave.view(), nF, st));
    ADJ_CHK(hipMemcpyAsync(d.rho_b, rho_b_save, sizeof(double) * nB, hipMemcpyDeviceToDevice, st) == hipSuccess);
    ADJ_CHK(hipMemcpyAsync(d.cfl_part, cfl_save, sizeof(double) * ncfl, hipMemcpyDeviceToDevice, st) == hipSuccess);
    ADJ_CHK(hipStreamSynchronize(st) == hipSuccess);
    return 0;
#undef ADJ_CHK
}

template int run_adjoint<float>(const Geo &, DevState &, hipStream_t, const AdjointIO &, std::string &);
template int run_adjoint<double>(const Geo &, DevState &, hipStream_t, const AdjointIO &, std::string &);

}  // namespace dugks
