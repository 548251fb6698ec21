// Fused DUGKS step kernels (see dugks_fused.h for the data flow).
//
// Work-group = (cell block, velocity group).  Per chunk of VC velocities:
//   stage 1  f-bar-plus of every block / ring-1 / ring-2 cell -> LDS
//            (fvDVM::updateGHbarPvol, fvDVM.C:648-683; equilibrium in FP64)
//   stage 2  least-squares gradient of every block / ring-1 cell -> LDS
//            (leastSquaresGrad through discreteVelocity::updateGrad, discreteVelocity.C:366-369)
//   stage 3A reconstruction of owned faces at x_f - xi dt/2 (fvDVM.C:824-963), FP64 raw
//            moments accumulated per face in registers (fvDVM.C:1181-1297)
//   stage 3B face values relaxed to the face equilibrium (fvDVM.C:1362-1460), incoming
//            wall / inlet values (fvDVM.C:1054-1098, discreteVelocity.C:244-257), flux in LDS
//   stage 4B g~ <- 4/3 f-bar-plus - 1/3 g~ - dt/V sum(+-flux) in ascending face order
//            (fvDVM.C:1462-1519), FP64 raw moments of the new g~, h~ per cell
// Reconstruction arithmetic runs in the storage type S (FP32 for the mixed path).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dugks_fused.h"

namespace dugks {

namespace {

constexpr int kNTA = kFusedThreadsA;
constexpr int kSlotsA = kFusedSlotsA;
constexpr int kNTB = kFusedThreadsB;
constexpr int kBlk = 256;

template <class S>
struct V3 {
    S x, y, z;
};

struct FArgs {
    const BlockInfo *blk;
    const int *lcell, *gs_beg, *gs_nb, *fa_face, *fa_own, *fa_nei, *fb_face, *fb_own, *fb_nei, *fbc_beg, *fbc_idx;
    const signed char *fbc_sgn;
    const void *gs_ls;
    double *fpart, *cpart;
    const double *bin, *cA, *fA;
    int nvg, vper, max_nl, max_ng, max_nfb, max_nc, nm, nma;
};

FArgs fargs(const FusedPlan &p)
{
    FArgs a;
    a.blk = p.d_blk; a.lcell = p.d_lcell; a.gs_beg = p.d_gs_beg; a.gs_nb = p.d_gs_nb;
    a.fa_face = p.d_fa_face; a.fa_own = p.d_fa_own; a.fa_nei = p.d_fa_nei;
    a.fb_face = p.d_fb_face; a.fb_own = p.d_fb_own; a.fb_nei = p.d_fb_nei;
    a.fbc_beg = p.d_fbc_beg; a.fbc_idx = p.d_fbc_idx; a.fbc_sgn = p.d_fbc_sgn; a.gs_ls = p.d_gs_ls;
    a.fpart = p.fpart; a.cpart = p.cpart; a.bin = p.bin; a.cA = p.cA; a.fA = p.fA;
    a.nvg = p.nvg; a.vper = p.vper; a.max_nl = p.max_nl; a.max_ng = p.max_ng; a.max_nfb = p.max_nfb;
    a.max_nc = p.max_nc; a.nm = p.nm; a.nma = p.nma;
    return a;
}

// Shakhov equilibrium with a precomputed rho/(2 pi R T)^(D/2) (fvDVM.C:673-679)
__device__ inline void shakhov_a(const Geo &g, double A, double rho, const d3 &U, double T, const d3 &q,
                                 const d3 &xi, double Th, double &gEq, double &hEq)
{
    const d3 um{U.x - xi.x, U.y - xi.y, U.z - xi.z};
    const double cSqrByRT = magsqr(um) / (g.R * T);
    const d3 xu{xi.x - U.x, xi.y - U.y, xi.z - U.z};
    const double cqBy5pRT = dot(xu, q) / (5.0 * rho * g.R * T * g.R * T);
    const double gEqBGK = A * exp(-cSqrByRT / 2.0);
    gEq = (1.0 + (1.0 - g.Pr) * cqBy5pRT * (cSqrByRT - g.D - 2.0)) * gEqBGK;
    hEq = ((g.K + 3.0 - g.D) + (1.0 - g.Pr) * cqBy5pRT * ((cSqrByRT - g.D) * (g.K + 3.0 - g.D) - 2 * g.K)) * gEqBGK *
          g.R * Th;
}

__device__ inline double eq_prefactor(const Geo &g, double rho, double T)
{
    return rho / pow(sqrt(2.0 * kPi * g.R * T), (double)g.D);
}

// raw moments of one (value, velocity) pair: m[k] += w * g * phi_k(xi), h likewise
template <int D>
__device__ inline void add_moments(double *m, double wg, double wh, const d3 &xi)
{
    const double xx = xi.x * xi.x, yy = xi.y * xi.y;
    if (D == 2) {
        const double s2 = xx + yy;
        m[0] += wg;
        m[1] += wg * xi.x;
        m[2] += wg * xi.y;
        m[3] += wg * xx;
        m[4] += wg * (xi.x * xi.y);
        m[5] += wg * yy;
        m[6] += wg * (s2 * xi.x);
        m[7] += wg * (s2 * xi.y);
        m[8] += wh;
        m[9] += wh * xi.x;
        m[10] += wh * xi.y;
    } else {
        const double zz = xi.z * xi.z;
        const double s2 = xx + yy + zz;
        m[0] += wg;
        m[1] += wg * xi.x;
        m[2] += wg * xi.y;
        m[3] += wg * xi.z;
        m[4] += wg * xx;
        m[5] += wg * (xi.x * xi.y);
        m[6] += wg * (xi.x * xi.z);
        m[7] += wg * yy;
        m[8] += wg * (xi.y * xi.z);
        m[9] += wg * zz;
        m[10] += wg * (s2 * xi.x);
        m[11] += wg * (s2 * xi.y);
        m[12] += wg * (s2 * xi.z);
        m[13] += wh;
        m[14] += wh * xi.x;
        m[15] += wh * xi.y;
        m[16] += wh * xi.z;
    }
}

// rho, U, T, tau, q from raw moments (the conversions of fvDVM.C:1249-1253, 1683-1696;
// q = 1/2 sum w c (|c|^2 g + h) with c = xi - U expanded in raw moments), then the
// 2 tau / (2 tau + h Pr) correction with h = corr_dt.
template <int D>
__device__ inline void macros_from_moments(const Geo &g, const double *m, double corr_dt, double &rho, d3 &U,
                                           double &T, double &tau, d3 &q)
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
    rho = G0;
    U = d3{G1[0] / rho, G1[1] / rho, G1[2] / rho};
    const double trG2 = G2[0] + G2[3] + G2[5];
    const double rhoE = 0.5 * (trG2 + H0);
    T = (rhoE - 0.5 * rho * magsqr(U)) / ((g.K + 3) / 2.0 * g.R * rho);
    tau = g.muRef * exp(g.omega * log(T / g.Tref)) / rho / T / g.R;
    const double u[3] = {U.x, U.y, U.z};
    const double uu = magsqr(U);
    const double G1u = G1[0] * u[0] + G1[1] * u[1] + G1[2] * u[2];
    const double G2u[3] = {G2[0] * u[0] + G2[1] * u[1] + G2[2] * u[2], G2[1] * u[0] + G2[3] * u[1] + G2[4] * u[2],
                           G2[2] * u[0] + G2[4] * u[1] + G2[5] * u[2]};
    double qq[3];
    for (int k = 0; k < 3; ++k) {
        const double qg = G3[k] - 2.0 * G2u[k] + uu * G1[k] - trG2 * u[k] + 2.0 * u[k] * G1u - uu * u[k] * G0;
        qq[k] = 0.5 * (qg + H1[k] - H0 * u[k]);
    }
    const double corr = 2.0 * tau / (2.0 * tau + corr_dt * g.Pr);
    q = d3{corr * qq[0], corr * qq[1], corr * qq[2]};
}

// ----------------------------------------------------------------- shared stages

// stage 1: f-bar-plus of local cells for the velocity chunk [vb, vb+VC)
template <class S, int VC, int NT>
__device__ inline void stage_bar(const Geo &g, const FArgs &a, const BlockInfo &bi, const S *gT, const S *hT,
                                 const Macros &cm, double dt, int vb, int v1, S *fbg, S *fbh, S *ftg, S *fth)
{
    const int nl = bi.nl;
    for (int i = threadIdx.x; i < nl * VC; i += NT) {
        const int vv = i / nl;
        const int l = i - vv * nl;
        const int v = vb + vv;
        if (v >= v1) continue;
        const int c = a.lcell[bi.l_off + l];
        const size_t k = (size_t)v * g.nC + c;
        const double gt = (double)gT[k], ht = (double)hT[k];
        const double T = cm.T[c];
        double gEq, hEq;
        shakhov_a(g, a.cA[c], cm.rho[c], cm.U[c], T, cm.q[c], g.xi[v], T, gEq, hEq);
        const double af = 1.5 * dt / (2.0 * cm.tau[c] + dt);
        fbg[vv * a.max_nl + l] = (S)((1.0 - af) * gt + af * gEq);
        fbh[vv * a.max_nl + l] = (S)((1.0 - af) * ht + af * hEq);
        if (ftg && l < bi.nc) {
            ftg[vv * a.max_nc + l] = (S)gt;
            fth[vv * a.max_nc + l] = (S)ht;
        }
    }
}

// stage 2: gradients of block + ring-1 cells
template <class S, int D, int VC, int NT>
__device__ inline void stage_grad(const FArgs &a, const BlockInfo &bi, int vb, int v1, const S *fbg, const S *fbh,
                                  S *grg, S *grh)
{
    const int ng = bi.ng;
    const V3<S> *ls = (const V3<S> *)a.gs_ls;
    for (int i = threadIdx.x; i < ng * VC; i += NT) {
        const int vv = i / ng;
        const int l = i - vv * ng;
        if (vb + vv >= v1) continue;
        const S *pg = fbg + vv * a.max_nl, *ph = fbh + vv * a.max_nl;
        const S cg = pg[l], ch = ph[l];
        S gx = 0, gy = 0, gz = 0, hx = 0, hy = 0, hz = 0;
        const int k1 = a.gs_beg[bi.gsb_off + l + 1];
        for (int k = a.gs_beg[bi.gsb_off + l]; k < k1; ++k) {
            const int n = a.gs_nb[k];
            const V3<S> L = ls[k];
            const S dg = pg[n] - cg, dh = ph[n] - ch;
            gx += L.x * dg; gy += L.y * dg; gz += L.z * dg;
            hx += L.x * dh; hy += L.y * dh; hz += L.z * dh;
        }
        S *og = grg + (vv * a.max_ng + l) * D, *oh = grh + (vv * a.max_ng + l) * D;
        og[0] = gx; og[1] = gy; oh[0] = hx; oh[1] = hy;
        if (D == 3) {
            og[2] = gz;
            oh[2] = hz;
        }
    }
}

template <class S, int D>
__device__ inline S recon_l(const S *fb, const S *gr, int l, const d3 &arm)
{
    const S *gl = gr + l * D;
    S r = fb[l] + (gl[0] * (S)arm.x + gl[1] * (S)arm.y);
    if (D == 3) r = fb[l] + (gl[0] * (S)arm.x + gl[1] * (S)arm.y + gl[2] * (S)arm.z);
    return r;
}

// ------------------------------------------------------------------------- pass A
template <class S, int D, int VC>
__global__ void __launch_bounds__(kNTA)
    k_pass_a(Geo g, FArgs a, const S *__restrict__ gT, const S *__restrict__ hT, Macros cm, const double *dtp)
{
    constexpr int NMA = (D == 3 ? 17 : 11) + 1;
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockInfo bi = a.blk[blockIdx.x];
    const int vg = blockIdx.y;
    const int v0 = vg * a.vper, v1 = min(g.Nv, v0 + a.vper);
    S *fbg = (S *)smem;
    S *fbh = fbg + VC * a.max_nl;
    S *grg = fbh + VC * a.max_nl;
    S *grh = grg + VC * a.max_ng * D;
    const double dt = *dtp;

    double acc[kSlotsA][NMA];
    int fid[kSlotsA], fown[kSlotsA], fnei[kSlotsA], ftyp[kSlotsA];
    d3 fS[kSlotsA], armO[kSlotsA], armN[kSlotsA];
#pragma unroll
    for (int s = 0; s < kSlotsA; ++s) {
#pragma unroll
        for (int k = 0; k < NMA; ++k) acc[s][k] = 0.0;
        const int j = threadIdx.x + s * kNTA;
        fid[s] = -1;
        if (j < bi.nfa) {
            const int f = a.fa_face[bi.fa_off + j];
            fid[s] = f;
            fown[s] = a.fa_own[bi.fa_off + j];
            fnei[s] = a.fa_nei[bi.fa_off + j];
            ftyp[s] = f < g.nIF ? 0 : g.bf_type[f - g.nIF];
            fS[s] = g.Sf[f];
            const d3 Cf = g.Cf[f];
            const d3 Co = g.C[a.lcell[bi.l_off + fown[s]]];
            armO[s] = d3{Cf.x - Co.x, Cf.y - Co.y, Cf.z - Co.z};
            if (fnei[s] >= 0) {
                const d3 Cn = g.C[a.lcell[bi.l_off + fnei[s]]];
                armN[s] = d3{Cf.x - Cn.x, Cf.y - Cn.y, Cf.z - Cn.z};
            } else {
                armN[s] = armO[s];
            }
        }
    }

    for (int vb = v0; vb < v1; vb += VC) {
        stage_bar<S, VC, kNTA>(g, a, bi, gT, hT, cm, dt, vb, v1, fbg, fbh, (S *)nullptr, (S *)nullptr);
        __syncthreads();
        stage_grad<S, D, VC, kNTA>(a, bi, vb, v1, fbg, fbh, grg, grh);
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kSlotsA; ++s) {
            if (fid[s] < 0) continue;
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                const int v = vb + vv;
                if (v >= v1) break;
                const d3 xi = g.xi[v];
                const double w = g.w[v];
                const double xs = dot(xi, fS[s]);
                const S *pg = fbg + vv * a.max_nl, *ph = fbh + vv * a.max_nl;
                const S *qg = grg + vv * a.max_ng * D, *qh = grh + vv * a.max_ng * D;
                const double hx = 0.5 * dt;
                double gf, hf;
                if (ftyp[s] == 0) {
                    if (xs >= kVSmall) {
                        const d3 r{armO[s].x - xi.x * hx, armO[s].y - xi.y * hx, armO[s].z - xi.z * hx};
                        gf = (double)recon_l<S, D>(pg, qg, fown[s], r);
                        hf = (double)recon_l<S, D>(ph, qh, fown[s], r);
                    } else if (xs < -kVSmall) {
                        const d3 r{armN[s].x - xi.x * hx, armN[s].y - xi.y * hx, armN[s].z - xi.z * hx};
                        gf = (double)recon_l<S, D>(pg, qg, fnei[s], r);
                        hf = (double)recon_l<S, D>(ph, qh, fnei[s], r);
                    } else {
                        const d3 ro{armO[s].x - xi.x * hx, armO[s].y - xi.y * hx, armO[s].z - xi.z * hx};
                        const d3 rn{armN[s].x - xi.x * hx, armN[s].y - xi.y * hx, armN[s].z - xi.z * hx};
                        gf = 0.5 * ((double)recon_l<S, D>(pg, qg, fnei[s], rn) + (double)recon_l<S, D>(pg, qg, fown[s], ro));
                        hf = 0.5 * ((double)recon_l<S, D>(ph, qh, fnei[s], rn) + (double)recon_l<S, D>(ph, qh, fown[s], ro));
                    }
                } else {
                    if (!(xs > 0)) continue;  // incoming half added from constants
                    const d3 r{armO[s].x - xi.x * hx, armO[s].y - xi.y * hx, armO[s].z - xi.z * hx};
                    gf = (double)recon_l<S, D>(pg, qg, fown[s], r);
                    hf = (double)recon_l<S, D>(ph, qh, fown[s], r);
                    if (ftyp[s] == kMaxwellWall) acc[s][NMA - 1] += w * xs * gf;
                }
                add_moments<D>(acc[s], w * gf, w * hf, xi);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < kSlotsA; ++s) {
        if (fid[s] < 0) continue;
        double *o = a.fpart + ((size_t)vg * g.nF + fid[s]) * NMA;
#pragma unroll
        for (int k = 0; k < NMA; ++k) o[k] = acc[s][k];
    }
}

// ------------------------------------------------------------------- face macros
template <int D>
__global__ void __launch_bounds__(kBlk)
    k_face_macros(Geo g, FArgs a, Macros fm, double *rho_b, const double *dtp)
{
    constexpr int NM = D == 3 ? 17 : 11;
    constexpr int NMA = NM + 1;
    const int f = blockIdx.x * kBlk + threadIdx.x;
    if (f >= g.nF) return;
    double m[NMA];
#pragma unroll
    for (int k = 0; k < NMA; ++k) m[k] = 0.0;
    for (int vg = 0; vg < a.nvg; ++vg) {
        const double *p = a.fpart + ((size_t)vg * g.nF + f) * NMA;
#pragma unroll
        for (int k = 0; k < NMA; ++k) m[k] += p[k];
    }
    if (f >= g.nIF) {
        const int b = f - g.nIF;
        double scale = 1.0;
        if (g.bf_type[b] == kMaxwellWall) {
            // calculatedMaxwellFvPatchField::evaluate: rho_w = outGoing / |inComingByRho|
            scale = m[NMA - 1] / fabs(g.in_by_rho[b]);
            rho_b[b] = scale;
        }
        const double *in = a.bin + (size_t)b * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) m[k] += scale * in[k];
    }
    double rho, T, tau;
    d3 U, q;
    macros_from_moments<D>(g, m, 0.5 * (*dtp), rho, U, T, tau, q);
    fm.rho[f] = rho;
    fm.U[f] = U;
    fm.T[f] = T;
    fm.tau[f] = tau;
    fm.q[f] = q;
    ((double *)a.fA)[f] = eq_prefactor(g, rho, T);
}

// ------------------------------------------------------------------------- pass B
template <class S, int D, int VC>
__global__ void __launch_bounds__(kNTB)
    k_pass_b(Geo g, FArgs a, const S *__restrict__ gT, const S *__restrict__ hT, S *__restrict__ gN,
             S *__restrict__ hN, Macros cm, Macros fm, const double *rho_b, const double *dtp)
{
    constexpr int NM = D == 3 ? 17 : 11;
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockInfo bi = a.blk[blockIdx.x];
    const int vg = blockIdx.y;
    const int v0 = vg * a.vper, v1 = min(g.Nv, v0 + a.vper);
    double *flg = (double *)smem;                   // [VC][max_nfb]
    double *flh = flg + VC * a.max_nfb;
    S *fbg = (S *)(flh + VC * a.max_nfb);
    S *fbh = fbg + VC * a.max_nl;
    S *grg = fbh + VC * a.max_nl;
    S *grh = grg + VC * a.max_ng * D;
    S *ftg = grh + VC * a.max_ng * D;
    S *fth = ftg + VC * a.max_nc;
    const double dt = *dtp;
    const double hx = 0.5 * dt;
    const double bf = 0.5 * dt;

    double mom[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) mom[k] = 0.0;
    const int my = threadIdx.x;
    const bool has_cell = my < bi.nc;
    const int cell = bi.c0 + my;
    const double Vc = has_cell ? g.V[cell] : 1.0;
    int k0 = 0, k1 = 0;
    if (has_cell) {
        k0 = a.fbc_beg[bi.fbc_off + my];
        k1 = a.fbc_beg[bi.fbc_off + my + 1];
    }

    for (int vb = v0; vb < v1; vb += VC) {
        stage_bar<S, VC, kNTB>(g, a, bi, gT, hT, cm, dt, vb, v1, fbg, fbh, ftg, fth);
        __syncthreads();
        stage_grad<S, D, VC, kNTB>(a, bi, vb, v1, fbg, fbh, grg, grh);
        __syncthreads();
        // stage 3: face values -> flux sigma * f_f in LDS
        const int nfb = bi.nfb;
        for (int i = threadIdx.x; i < nfb * VC; i += kNTB) {
            const int vv = i / nfb;
            const int j = i - vv * nfb;
            const int v = vb + vv;
            if (v >= v1) continue;
            const int f = a.fb_face[bi.fb_off + j];
            const int lo = a.fb_own[bi.fb_off + j], ln = a.fb_nei[bi.fb_off + j];
            const d3 xi = g.xi[v];
            const d3 Sf = g.Sf[f];
            const double xs = dot(xi, Sf);
            const d3 Cf = g.Cf[f];
            const S *pg = fbg + vv * a.max_nl, *ph = fbh + vv * a.max_nl;
            const S *qg = grg + vv * a.max_ng * D, *qh = grh + vv * a.max_ng * D;
            double gv, hv;
            bool relax = true;
            if (f < g.nIF) {
                if (xs >= kVSmall) {
                    const d3 C = g.C[a.lcell[bi.l_off + lo]];
                    const d3 r{Cf.x - C.x - xi.x * hx, Cf.y - C.y - xi.y * hx, Cf.z - C.z - xi.z * hx};
                    gv = (double)recon_l<S, D>(pg, qg, lo, r);
                    hv = (double)recon_l<S, D>(ph, qh, lo, r);
                } else if (xs < -kVSmall) {
                    const d3 C = g.C[a.lcell[bi.l_off + ln]];
                    const d3 r{Cf.x - C.x - xi.x * hx, Cf.y - C.y - xi.y * hx, Cf.z - C.z - xi.z * hx};
                    gv = (double)recon_l<S, D>(pg, qg, ln, r);
                    hv = (double)recon_l<S, D>(ph, qh, ln, r);
                } else {
                    const d3 Co = g.C[a.lcell[bi.l_off + lo]], Cn = g.C[a.lcell[bi.l_off + ln]];
                    const d3 ro{Cf.x - Co.x - xi.x * hx, Cf.y - Co.y - xi.y * hx, Cf.z - Co.z - xi.z * hx};
                    const d3 rn{Cf.x - Cn.x - xi.x * hx, Cf.y - Cn.y - xi.y * hx, Cf.z - Cn.z - xi.z * hx};
                    gv = 0.5 * ((double)recon_l<S, D>(pg, qg, ln, rn) + (double)recon_l<S, D>(pg, qg, lo, ro));
                    hv = 0.5 * ((double)recon_l<S, D>(ph, qh, ln, rn) + (double)recon_l<S, D>(ph, qh, lo, ro));
                }
            } else {
                const int b = f - g.nIF;
                if (xs > 0) {
                    const d3 C = g.C[a.lcell[bi.l_off + lo]];
                    const d3 r{Cf.x - C.x - xi.x * hx, Cf.y - C.y - xi.y * hx, Cf.z - C.z - xi.z * hx};
                    gv = (double)recon_l<S, D>(pg, qg, lo, r);
                    hv = (double)recon_l<S, D>(ph, qh, lo, r);
                } else {
                    relax = false;
                    const d3 Ub = g.bf_U[b];
                    const double Tb = g.bf_T[b];
                    const d3 d{Ub.x - xi.x, Ub.y - xi.y, Ub.z - xi.z};
                    if (g.bf_type[b] == kMaxwellWall) {
                        gv = rho_b[b] * (1.0 / pow(sqrt(2.0 * kPi * g.R * Tb), (double)g.D) *
                                         exp(-magsqr(d) / (2.0 * g.R * Tb)));
                        hv = gv * (g.R * Tb) * (g.K + 3 - g.D);
                    } else {
                        gv = g.bf_rho[b] / pow(sqrt(2.0 * kPi * g.R * Tb), (double)g.D) *
                             exp(-magsqr(d) / (2.0 * g.R * Tb));
                        hv = (g.K + 3 - g.D) * g.R * Tb * gv;
                    }
                }
            }
            if (relax) {
                const double Tf = fm.T[f];
                const double Th = (f >= g.nIF && g.q1) ? fm.T[g.bf_local[f - g.nIF]] : Tf;
                double gEq, hEq;
                shakhov_a(g, a.fA[f], fm.rho[f], fm.U[f], Tf, fm.q[f], xi, Th, gEq, hEq);
                const double b = bf / (2.0 * fm.tau[f] + bf);
                gv = (1.0 - b) * gv + b * gEq;
                hv = (1.0 - b) * hv + b * hEq;
            }
            flg[vv * a.max_nfb + j] = xs * gv;
            flh[vv * a.max_nfb + j] = xs * hv;
        }
        __syncthreads();
        // stage 4: cell update in ascending face order, new values to the other buffer
        if (has_cell) {
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                const int v = vb + vv;
                if (v >= v1) break;
                double ag = -1.0 / 3 * (double)ftg[vv * a.max_nc + my] + 4.0 / 3 * (double)fbg[vv * a.max_nl + my];
                double ah = -1.0 / 3 * (double)fth[vv * a.max_nc + my] + 4.0 / 3 * (double)fbh[vv * a.max_nl + my];
                for (int k = k0; k < k1; ++k) {
                    const int j = a.fbc_idx[k];
                    const double tg = flg[vv * a.max_nfb + j] * dt / Vc;
                    const double th = flh[vv * a.max_nfb + j] * dt / Vc;
                    if (a.fbc_sgn[k] > 0) {
                        ag += tg;
                        ah += th;
                    } else {
                        ag -= tg;
                        ah -= th;
                    }
                }
                const S sg = (S)ag, sh = (S)ah;
                const size_t kk = (size_t)v * g.nC + cell;
                gN[kk] = sg;
                hN[kk] = sh;
                const double w = g.w[v];
                add_moments<D>(mom, w * (double)sg, w * (double)sh, g.xi[v]);
            }
        }
        __syncthreads();
    }
    if (has_cell) {
        double *o = a.cpart + ((size_t)vg * g.nC + cell) * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) o[k] = mom[k];
    }
}

// ------------------------------------------------------------------- cell macros
template <int D>
__global__ void __launch_bounds__(kBlk) k_cell_macros(Geo g, FArgs a, Macros cm, const double *dtp)
{
    constexpr int NM = D == 3 ? 17 : 11;
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= g.nC) return;
    double m[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) m[k] = 0.0;
    for (int vg = 0; vg < a.nvg; ++vg) {
        const double *p = a.cpart + ((size_t)vg * g.nC + c) * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) m[k] += p[k];
    }
    double rho, T, tau;
    d3 U, q;
    macros_from_moments<D>(g, m, *dtp, rho, U, T, tau, q);
    cm.rho[c] = rho;
    cm.U[c] = U;
    cm.T[c] = T;
    cm.tau[c] = tau;
    cm.q[c] = q;
    ((double *)a.cA)[c] = eq_prefactor(g, rho, T);
}

__global__ void __launch_bounds__(kBlk) k_cell_coef(Geo g, Macros cm, double *cA)
{
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= g.nC) return;
    cA[c] = eq_prefactor(g, cm.rho[c], cm.T[c]);
}

// incoming-half moments of the wall (per unit wall density) and inlet faces, fixed for
// the run (the wall and inlet values are constant, fvDVM.C:578-624, discreteVelocity.C:244-257)
template <int D>
__global__ void __launch_bounds__(kBlk) k_incoming_moments(Geo g, double *bin)
{
    constexpr int NM = D == 3 ? 17 : 11;
    const int b = blockIdx.x * kBlk + threadIdx.x;
    if (b >= g.nB) return;
    double m[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) m[k] = 0.0;
    const d3 Sf = g.Sf[g.nIF + b];
    const d3 Ub = g.bf_U[b];
    const double Tb = g.bf_T[b];
    const bool wall = g.bf_type[b] == kMaxwellWall;
    for (int v = 0; v < g.Nv; ++v) {
        const d3 xi = g.xi[v];
        if (dot(xi, Sf) > 0) continue;
        const d3 d{Ub.x - xi.x, Ub.y - xi.y, Ub.z - xi.z};
        double gv, hv;
        if (wall) {
            gv = 1.0 / pow(sqrt(2.0 * kPi * g.R * Tb), (double)g.D) * exp(-magsqr(d) / (2.0 * g.R * Tb));
            hv = gv * (g.R * Tb) * (g.K + 3 - g.D);
        } else {
            gv = g.bf_rho[b] / pow(sqrt(2.0 * kPi * g.R * Tb), (double)g.D) * exp(-magsqr(d) / (2.0 * g.R * Tb));
            hv = (g.K + 3 - g.D) * g.R * Tb * gv;
        }
        add_moments<D>(m, g.w[v] * gv, g.w[v] * hv, xi);
    }
    for (int k = 0; k < NM; ++k) bin[(size_t)b * NM + k] = m[k];
}

template <class S>
size_t lds_pass_a(const FusedPlan &p, int D, int VC)
{
    return (size_t)VC * (2 * p.max_nl + 2 * p.max_ng * D) * sizeof(S);
}
template <class S>
size_t lds_pass_b(const FusedPlan &p, int D, int VC)
{
    return (size_t)VC * 2 * p.max_nfb * sizeof(double) + lds_pass_a<S>(p, D, VC) +
           (size_t)VC * 2 * p.max_nc * sizeof(S);
}

constexpr size_t kLdsBudget = 80 * 1024;

template <class S>
int choose_vc(const FusedPlan &p, int D)
{
    for (int vc : {4, 2, 1})
        if (lds_pass_b<S>(p, D, vc) <= kLdsBudget) return vc;
    return 0;
}

template <class S, int D, int VC>
void launch_a(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const Macros &cm, const double *dt, hipStream_t st)
{
    const size_t lds = lds_pass_a<S>(p, D, VC);
    hipLaunchKernelGGL((k_pass_a<S, D, VC>), dim3(p.nblk, p.nvg), dim3(kNTA), lds, st, g, fargs(p), gT, hT, cm, dt);
}

template <class S, int D, int VC>
void launch_b(const Geo &g, FusedPlan &p, const S *gT, const S *hT, S *gN, S *hN, const Macros &cm, const Macros &fm,
              const double *rho_b, const double *dt, hipStream_t st)
{
    const size_t lds = lds_pass_b<S>(p, D, VC);
    hipLaunchKernelGGL((k_pass_b<S, D, VC>), dim3(p.nblk, p.nvg), dim3(kNTB), lds, st, g, fargs(p), gT, hT, gN, hN,
                       cm, fm, rho_b, dt);
}

template <class S>
bool set_lds_attrs(int D, int vc, const FusedPlan &p)
{
    auto set = [](const void *fn, size_t bytes) {
        return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
    };
    bool ok = true;
#define DUGKS_SET(DD, VV)                                                                              \
    if (D == DD && vc == VV) {                                                                         \
        ok &= set((const void *)k_pass_a<S, DD, VV>, lds_pass_a<S>(p, DD, VV));                        \
        ok &= set((const void *)k_pass_b<S, DD, VV>, lds_pass_b<S>(p, DD, VV));                        \
    }
    DUGKS_SET(2, 1) DUGKS_SET(2, 2) DUGKS_SET(2, 4) DUGKS_SET(3, 1) DUGKS_SET(3, 2) DUGKS_SET(3, 4)
#undef DUGKS_SET
    return ok;
}

inline dim3 grid1(int n) { return dim3((n + kBlk - 1) / kBlk); }

}  // namespace

// ----------------------------------------------------------------- plan upload / init

template <class T>
static bool up(FusedPlan &p, T **dst, const std::vector<T> &src, std::string &err)
{
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(1, src.size()) * sizeof(T);
    if (hipMalloc(&q, bytes) != hipSuccess) {
        err = "hipMalloc failed for fused tables";
        return false;
    }
    p.allocs.push_back(q);
    if (!src.empty() && hipMemcpy(q, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        err = "hipMemcpy failed for fused tables";
        return false;
    }
    *dst = (T *)q;
    return true;
}

static bool dbuf(FusedPlan &p, void **dst, size_t bytes, std::string &err)
{
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(1, bytes)) != hipSuccess) {
        err = "hipMalloc of " + std::to_string(bytes) + " bytes failed (fused buffers)";
        return false;
    }
    p.allocs.push_back(q);
    *dst = q;
    return true;
}

bool fused_plan_upload(FusedPlan &p, const Geo &g, size_t es, std::string &err)
{
    const int vc = es == 4 ? choose_vc<float>(p, g.D) : choose_vc<double>(p, g.D);
    if (vc == 0) {
        err = "block tables exceed the LDS budget";
        return false;
    }
    p.vc = vc;
    if (!(es == 4 ? set_lds_attrs<float>(g.D, vc, p) : set_lds_attrs<double>(g.D, vc, p))) {
        err = "cannot raise the dynamic LDS limit";
        return false;
    }
    if (!up(p, &p.d_blk, p.blk, err) || !up(p, &p.d_lcell, p.lcell, err) || !up(p, &p.d_gs_beg, p.gs_beg, err) ||
        !up(p, &p.d_gs_nb, p.gs_nb, err) || !up(p, &p.d_fa_face, p.fa_face, err) ||
        !up(p, &p.d_fa_own, p.fa_own, err) || !up(p, &p.d_fa_nei, p.fa_nei, err) ||
        !up(p, &p.d_fb_face, p.fb_face, err) || !up(p, &p.d_fb_own, p.fb_own, err) ||
        !up(p, &p.d_fb_nei, p.fb_nei, err) || !up(p, &p.d_fbc_beg, p.fbc_beg, err) ||
        !up(p, &p.d_fbc_idx, p.fbc_idx, err) || !up(p, &p.d_fbc_sgn, p.fbc_sgn, err))
        return false;
    if (es == 4) {
        std::vector<V3<float>> ls(p.gs_ls.size());
        for (size_t i = 0; i < ls.size(); ++i) ls[i] = V3<float>{(float)p.gs_ls[i].x, (float)p.gs_ls[i].y, (float)p.gs_ls[i].z};
        V3<float> *d = nullptr;
        if (!up(p, &d, ls, err)) return false;
        p.d_gs_ls = d;
    } else {
        d3 *d = nullptr;
        if (!up(p, &d, p.gs_ls, err)) return false;
        p.d_gs_ls = d;
    }
    void *fp = nullptr, *cp = nullptr, *bi = nullptr, *ca = nullptr, *fa = nullptr;
    if (!dbuf(p, &fp, (size_t)p.nvg * g.nF * p.nma * 8, err) || !dbuf(p, &cp, (size_t)p.nvg * g.nC * p.nm * 8, err) ||
        !dbuf(p, &bi, (size_t)g.nB * p.nm * 8, err) || !dbuf(p, &ca, (size_t)g.nC * 8, err) ||
        !dbuf(p, &fa, (size_t)g.nF * 8, err) || !dbuf(p, &p.gT2, (size_t)g.Nv * g.nC * es, err) ||
        !dbuf(p, &p.hT2, (size_t)g.Nv * g.nC * es, err))
        return false;
    p.fpart = (double *)fp;
    p.cpart = (double *)cp;
    p.bin = (double *)bi;
    p.cA = (double *)ca;
    p.fA = (double *)fa;
    p.lds_a = es == 4 ? lds_pass_a<float>(p, g.D, vc) : lds_pass_a<double>(p, g.D, vc);
    p.lds_b = es == 4 ? lds_pass_b<float>(p, g.D, vc) : lds_pass_b<double>(p, g.D, vc);
    return true;
}

bool fused_init_boundary(FusedPlan &p, const Geo &g, const Macros &cm, hipStream_t st, std::string &err)
{
    if (g.nB > 0) {
        if (g.D == 3)
            hipLaunchKernelGGL(k_incoming_moments<3>, grid1(g.nB), dim3(kBlk), 0, st, g, p.bin);
        else
            hipLaunchKernelGGL(k_incoming_moments<2>, grid1(g.nB), dim3(kBlk), 0, st, g, p.bin);
    }
    hipLaunchKernelGGL(k_cell_coef, grid1(g.nC), dim3(kBlk), 0, st, g, cm, p.cA);
    if (hipGetLastError() != hipSuccess) {
        err = "fused init kernels failed to launch";
        return false;
    }
    return true;
}

void fused_plan_free(FusedPlan &p)
{
    for (void *q : p.allocs)
        if (q) (void)hipFree(q);
    p.allocs.clear();
    p.ready = false;
}

// ------------------------------------------------------------------- launchers

template <class S>
void Fused<S>::face_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const Macros &cm, const double *dt,
                         hipStream_t st)
{
#define DUGKS_A(DD, VV) \
    if (g.D == DD && p.vc == VV) return launch_a<S, DD, VV>(g, p, gT, hT, cm, dt, st);
    DUGKS_A(2, 1) DUGKS_A(2, 2) DUGKS_A(2, 4) DUGKS_A(3, 1) DUGKS_A(3, 2) DUGKS_A(3, 4)
#undef DUGKS_A
}

template <class S>
void Fused<S>::face_macros(const Geo &g, FusedPlan &p, Macros fm, double *rho_b, const double *dt, double *cfl_part,
                           hipStream_t st)
{
    if (g.D == 3)
        hipLaunchKernelGGL(k_face_macros<3>, grid1(g.nF), dim3(kBlk), 0, st, g, fargs(p), fm, rho_b, dt);
    else
        hipLaunchKernelGGL(k_face_macros<2>, grid1(g.nF), dim3(kBlk), 0, st, g, fargs(p), fm, rho_b, dt);
    launch_cfl_partials(g, fm, cfl_part, st);
}

template <class S>
void Fused<S>::update_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, S *gN, S *hN, const Macros &cm,
                           const Macros &fm, const double *rho_b, const double *dt, hipStream_t st)
{
#define DUGKS_B(DD, VV) \
    if (g.D == DD && p.vc == VV) return launch_b<S, DD, VV>(g, p, gT, hT, gN, hN, cm, fm, rho_b, dt, st);
    DUGKS_B(2, 1) DUGKS_B(2, 2) DUGKS_B(2, 4) DUGKS_B(3, 1) DUGKS_B(3, 2) DUGKS_B(3, 4)
#undef DUGKS_B
}

template <class S>
void Fused<S>::cell_macros(const Geo &g, FusedPlan &p, Macros cm, const double *dt, hipStream_t st)
{
    if (g.D == 3)
        hipLaunchKernelGGL(k_cell_macros<3>, grid1(g.nC), dim3(kBlk), 0, st, g, fargs(p), cm, dt);
    else
        hipLaunchKernelGGL(k_cell_macros<2>, grid1(g.nC), dim3(kBlk), 0, st, g, fargs(p), cm, dt);
}

template <class S>
void Fused<S>::cell_coef(const Geo &g, FusedPlan &p, const Macros &cm, hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_coef, grid1(g.nC), dim3(kBlk), 0, st, g, cm, p.cA);
}

template struct Fused<float>;
template struct Fused<double>;

}  // namespace dugks
