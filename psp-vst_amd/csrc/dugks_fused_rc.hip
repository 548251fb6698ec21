// Fused DUGKS step kernels, recompute schedule (see dugks_fused_rc.h for the data flow).
//
// Work-group = (cell block, velocity group).  Per chunk of up to VC velocities (one x-stripe
// of a velocity row):
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
//
// Each thread owns fixed cells / faces for the whole work-group.  The velocity set is a
// tensor product, so the Maxwellian factor A exp(-(xi_x - U_x)^2 / 2RT) of an entity is
// formed once per x-stripe and kept in registers; a chunk then needs one exp for the y/z
// part per entity.  Work-group ids are mapped so that the velocity groups of a block run on
// the same accelerator die and share its L2.  Reconstruction arithmetic runs in the storage
// type S (FP32 for the mixed path); equilibria and moments in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dugks_fused_rc.h"

namespace dugks {
namespace rc {

namespace {

constexpr int kNTA = kFusedThreadsA;
constexpr int kSlotsA = kFusedSlotsA;
constexpr int kNTB = kFusedThreadsB;
constexpr int kNb = kFusedMaxNb;
constexpr int kBlk = 256;

// Tuning knobs of the recompute schedule: LDS budget per work-group (caps the chunk size)
// and the minimum waves per SIMD the passes are compiled for (caps their registers).
#ifndef DUGKS_RC_LDS_KB
#define DUGKS_RC_LDS_KB 80
#endif
#ifndef DUGKS_RC_WAVES_A
#define DUGKS_RC_WAVES_A 1
#endif
#ifndef DUGKS_RC_WAVES_B
#define DUGKS_RC_WAVES_B 1
#endif
constexpr int kWavesA = DUGKS_RC_WAVES_A;
constexpr int kWavesB = DUGKS_RC_WAVES_B;

template <class S>
struct V3 {
    S x, y, z;
};

// g and h of one (cell or face, velocity) side by side: one LDS access moves both
template <class S>
struct alignas(2 * sizeof(S)) P2 {
    S g, h;
};

struct FArgs {
    const BlockInfo *blk;
    const int *lcell, *gs_cnt, *gs_nb, *fa_face, *fa_own, *fa_nei, *fb_face, *fb_own, *fb_nei, *fbc_beg, *fbc_idx;
    const FaceGeo *fb_geo;
    EqCoef *fb_coef;
    const signed char *fbc_sgn;
    const void *gs_ls;
    double *fpart, *cpart;
    const double *bin, *bref, *cref;
    EqCoef *cc, *fc;
    const BndCoef *bc;
    int nblk, b8, nvg, n1, nrow, nchunk, cpg;
    int max_nl, max_ng, max_nfb, max_nc, nm, nma;
    int nbuf_a;        // pass-A LDS tile buffers (2: alternate per chunk, one barrier fewer)
};

FArgs fargs(const FusedPlan &p)
{
    FArgs a;
    a.blk = p.d_blk; a.lcell = p.d_lcell; a.gs_cnt = p.d_gs_cnt; a.gs_nb = p.d_gs_nb;
    a.fa_face = p.d_fa_face; a.fa_own = p.d_fa_own; a.fa_nei = p.d_fa_nei;
    a.fb_face = p.d_fb_face; a.fb_own = p.d_fb_own; a.fb_nei = p.d_fb_nei; a.fb_geo = p.fb_geo; a.fb_coef = p.fb_coef;
    a.fbc_beg = p.d_fbc_beg; a.fbc_idx = p.d_fbc_idx; a.fbc_sgn = p.d_fbc_sgn; a.gs_ls = p.d_gs_ls;
    a.fpart = p.fpart; a.cpart = p.cpart; a.bin = p.bin; a.bref = p.bref; a.cref = p.cref; a.cc = p.cc; a.fc = p.fc; a.bc = p.bc;
    a.nblk = p.nblk; a.b8 = p.b8; a.nvg = p.nvg; a.n1 = p.n1; a.nrow = p.nrow; a.nchunk = p.nchunk; a.cpg = p.cpg;
    a.max_nl = p.max_nl; a.max_ng = p.max_ng; a.max_nfb = p.max_nfb; a.max_nc = p.max_nc;
    a.nm = p.nm; a.nma = p.nma;
    a.nbuf_a = p.nbuf_a;
    return a;
}

// Work-group id -> (block, velocity group).  Ids are dealt round-robin over the dies, so die
// k receives ids k, k+8, ...; those are mapped to the velocity groups of blocks
// [k*b8, (k+1)*b8), block-major, and a die's resident work-groups share a few blocks' data.
__device__ inline bool wg_map(const FArgs &a, int &blk, int &vg)
{
    const int id = blockIdx.x;
    const int die = id % kFusedXcds, s = id / kFusedXcds;
    blk = die * a.b8 + s / a.nvg;
    vg = s % a.nvg;
    return blk < a.nblk;
}

__device__ inline double eq_prefactor(const Geo &g, double rho, double T)
{
    return rho / pow(sqrt(2.0 * kPi * g.R * T), (double)g.D);
}

// Coefficients of the Shakhov equilibrium (fvDVM.C:673-679) for rho, U, T, q.
__device__ inline EqCoef make_coef(const Geo &g, double rho, const d3 &U, double T, const d3 &q, double RTh, double tr)
{
    EqCoef k;
    const double RT = g.R * T;
    k.A = eq_prefactor(g, rho, T);
    k.ux = U.x; k.uy = U.y; k.uz = U.z;
    k.invRT = 1.0 / RT;
    k.RTh = RTh;
    const double q5 = 1.0 / (5.0 * rho * RT * RT);
    k.qx = q.x * q5; k.qy = q.y * q5; k.qz = q.z * q5;
    k.tr = tr;
    return k;
}

// Shakhov equilibrium (fvDVM.C:673-679) given gM = A exp(-|xi - U|^2 / 2RT):
// c = xi - U, s = |c|^2/(R T), cq = c.q/(5 rho (R T)^2)
__device__ inline void shakhov_poly(const Geo &g, const EqCoef &k, const d3 &xi, double gM, double &gEq, double &hEq)
{
    const double cx = xi.x - k.ux, cy = xi.y - k.uy, cz = xi.z - k.uz;
    const double s = (cx * cx + cy * cy + cz * cz) * k.invRT;
    const double cq = cx * k.qx + cy * k.qy + cz * k.qz;
    const double omp = 1.0 - g.Pr;
    const double kd = g.K + 3.0 - g.D;
    gEq = (1.0 + omp * cq * (s - g.D - 2.0)) * gM;
    hEq = (kd + omp * cq * ((s - g.D) * kd - 2 * g.K)) * gM * k.RTh;
}

// exp(-((xi_y - U_y)^2 + (xi_z - U_z)^2) / 2RT) of the chunk's row (xr = xi of the row)
__device__ inline double row_factor(const EqCoef &k, const d3 &xr)
{
    const double dy = xr.y - k.uy, dz = xr.z - k.uz;
    return exp(-((dy * dy + dz * dz) * k.invRT) / 2.0);
}

// A exp(-(xi_x - U_x)^2 / 2RT) for the abscissae ix0 .. ix0+VC-1 of a stripe; xi[ix] is the
// velocity (x[ix], x[0], x[0]) of the tensor set (ix fastest)
template <int VC>
__device__ inline void stripe_factors(const Geo &g, const EqCoef &k, int ix0, int n1, double *ex)
{
#pragma unroll
    for (int vv = 0; vv < VC; ++vv) {
        const int ix = min(ix0 + vv, n1 - 1);
        const double dx = cload(g.xi, ix).x - k.ux;
        ex[vv] = k.A * exp(-((dx * dx) * k.invRT) / 2.0);
    }
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

// Power sums over the velocities of one chunk (one x-stripe of a velocity row: xi.y and
// xi.z are shared): g0..g3 = sum w g xi_x^k, h0, h1 = sum w h xi_x^k.  fold_row adds the raw
// moments they stand for (the order of add_moments) once per chunk, so a face-velocity pair
// costs 10 operations instead of the 20-30 of add_moments.
struct RowSums {
    double g0, g1, g2, g3, h0, h1;
};

__device__ inline void row_add(RowSums &r, double wg, double wh, double x)
{
    r.g0 += wg;
    const double t1 = wg * x;
    r.g1 += t1;
    const double t2 = t1 * x;
    r.g2 += t2;
    r.g3 += t2 * x;
    r.h0 += wh;
    r.h1 += wh * x;
}

template <int D>
__device__ inline void fold_row(double *m, const RowSums &r, double y, double z)
{
    if (D == 2) {
        const double yy = y * y;
        m[0] += r.g0;
        m[1] += r.g1;
        m[2] += y * r.g0;
        m[3] += r.g2;
        m[4] += y * r.g1;
        m[5] += yy * r.g0;
        m[6] += r.g3 + yy * r.g1;          // sum w g (x^2 + y^2) x
        m[7] += y * r.g2 + yy * y * r.g0;  // sum w g (x^2 + y^2) y
        m[8] += r.h0;
        m[9] += r.h1;
        m[10] += y * r.h0;
    } else {
        const double yy = y * y, zz = z * z, r2 = yy + zz;
        m[0] += r.g0;
        m[1] += r.g1;
        m[2] += y * r.g0;
        m[3] += z * r.g0;
        m[4] += r.g2;
        m[5] += y * r.g1;
        m[6] += z * r.g1;
        m[7] += yy * r.g0;
        m[8] += (y * z) * r.g0;
        m[9] += zz * r.g0;
        m[10] += r.g3 + r2 * r.g1;         // sum w g |xi|^2 x
        m[11] += y * r.g2 + (y * r2) * r.g0;
        m[12] += z * r.g2 + (z * r2) * r.g0;
        m[13] += r.h0;
        m[14] += r.h1;
        m[15] += y * r.h0;
        m[16] += z * r.h0;
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

// Velocity chunk ch of the (x-stripe, row) order: stripe = ch / nrow, row = ch % nrow.
struct Chunk {
    int ix0, nvv, vrow;  // first abscissa of the stripe, valid velocities, v of (ix = 0, row)
    bool new_stripe;     // stripe differs from the previous chunk's: x-factors must be formed
};

__device__ inline Chunk chunk_of(const FArgs &a, int ch, int vc, int &stripe)
{
    const int st = ch / a.nrow;
    const int row = ch - st * a.nrow;
    Chunk c;
    c.ix0 = st * vc;
    c.nvv = min(vc, a.n1 - c.ix0);
    c.vrow = row * a.n1;
    c.new_stripe = st != stripe;
    stripe = st;
    return c;
}

// g~, h~ of the thread's local cells for one chunk (lc[s]: cell of the slot-s local cell);
// velocities past the end of a short stripe repeat its last one
template <class S, int VC, int NT, int CPT>
__device__ inline void load_tilde(const Geo &g, const BlockInfo &bi, const S *__restrict__ gT,
                                  const S *__restrict__ hT, const Chunk &ck, const int (&lc)[CPT],
                                  S (&gt)[CPT][VC], S (&ht)[CPT][VC])
{
#pragma unroll
    for (int s = 0; s < CPT; ++s) {
        const int l = threadIdx.x + s * NT;
        if (l >= bi.nl) break;
#pragma unroll
        for (int vv = 0; vv < VC; ++vv) {
            const size_t kk = (size_t)(ck.vrow + ck.ix0 + min(vv, ck.nvv - 1)) * g.nC + lc[s];
            gt[s][vv] = gT[kk];
            ht[s][vv] = hT[kk];
        }
    }
}

// stage 1: f-bar-plus of local cells for one chunk; cex[s] holds the slot-s cell's
// x-factors for the current stripe.  PRE: g~, h~ were loaded into pg, ph beforehand;
// otherwise each slot loads its own just before use (fewer live registers)
template <class S, int VC, int NT, int CPT, bool PRE>
__device__ inline void stage_bar(const Geo &g, const FArgs &a, const BlockInfo &bi, const S *__restrict__ gT,
                                 const S *__restrict__ hT, double dt, const Chunk &ck, const int (&lc)[CPT],
                                 const S (&pg)[CPT][VC], const S (&ph)[CPT][VC], double (&cex)[CPT][VC],
                                 P2<S> *fb, P2<S> *ft)
{
    const d3 xr = cload(g.xi, ck.vrow);
#pragma unroll
    for (int s = 0; s < CPT; ++s) {
        const int l = threadIdx.x + s * NT;
        if (l >= bi.nl) break;
        const int c = lc[s];
        const EqCoef k = a.cc[c];
        S gt[VC], ht[VC];
#pragma unroll
        for (int vv = 0; vv < VC; ++vv) {
            if (PRE) {
                gt[vv] = pg[s][vv];
                ht[vv] = ph[s][vv];
            } else {
                const size_t kk = (size_t)(ck.vrow + ck.ix0 + min(vv, ck.nvv - 1)) * g.nC + c;
                gt[vv] = gT[kk];
                ht[vv] = hT[kk];
            }
        }
        if (ck.new_stripe) stripe_factors<VC>(g, k, ck.ix0, a.n1, cex[s]);
        const double eyz = row_factor(k, xr);
        const double af = 1.5 * dt / (2.0 * k.tr + dt);
#pragma unroll
        for (int vv = 0; vv < VC; ++vv) {
            if (vv >= ck.nvv) break;
            const int v = ck.vrow + ck.ix0 + vv;
            double gEq, hEq;
            shakhov_poly(g, k, cload(g.xi, v), cex[s][vv] * eyz, gEq, hEq);
            // LDS holds f-bar-plus as deviations from the storage reference, like g~ itself
            // (FP32 storage only; FP64 storage keeps values, its reference being zero)
            if (sizeof(S) == 4) {
                gEq -= cload(g.ref_g, v);
                hEq -= cload(g.ref_h, v);
            }
            fb[vv * a.max_nl + l] = P2<S>{(S)((1.0 - af) * (double)gt[vv] + af * gEq),
                                          (S)((1.0 - af) * (double)ht[vv] + af * hEq)};
            if (ft && l < bi.nc) {
                ft[vv * a.max_nc + l] = P2<S>{gt[vv], ht[vv]};
            }
        }
    }
}

// gradient of local cell l for the chunk velocities [v0, v1)
template <class S, int D>
__device__ inline void grad_cell(const FArgs &a, const BlockInfo &bi, int l, int v0, int v1, const P2<S> *fb,
                                 P2<S> *gr)
{
    const V3<S> *ls = (const V3<S> *)a.gs_ls;
    const int k0 = (bi.gsb_off + l) * kNb;
    const int n = a.gs_cnt[bi.gsb_off + l];
    // the first NU entries (faces of a quad / hex) are read without waiting for the count;
    // padded entries are local cell 0 with a zero vector and are not used
    constexpr int NU = D == 2 ? 4 : 6;
    int nb[kNb];
    V3<S> L[kNb];
#pragma unroll
    for (int j = 0; j < kNb; ++j) {
        if (j < NU || j < n) {
            nb[j] = a.gs_nb[k0 + j];
            L[j] = ls[k0 + j];
        }
    }
#pragma unroll 1
    for (int vv = v0; vv < v1; ++vv) {
        const P2<S> *p = fb + vv * a.max_nl;
        const P2<S> c = p[l];
        S gx = 0, gy = 0, gz = 0, hx = 0, hy = 0, hz = 0;
#pragma unroll
        for (int j = 0; j < kNb; ++j) {
            if (j < n) {
                const P2<S> q = p[nb[j]];
                const S dg = q.g - c.g, dh = q.h - c.h;
                gx += L[j].x * dg; gy += L[j].y * dg;
                hx += L[j].x * dh; hy += L[j].y * dh;
                if (D == 3) {
                    gz += L[j].z * dg;
                    hz += L[j].z * dh;
                }
            }
        }
        P2<S> *o = gr + (vv * a.max_ng + l) * D;
        o[0] = P2<S>{gx, hx};
        o[1] = P2<S>{gy, hy};
        if (D == 3) o[2] = P2<S>{gz, hz};
    }
}

// stage 2: gradients of block + ring-1 cells.  Cells l < NT take one thread each for all
// chunk velocities; the (cell, velocity) pairs of the cells beyond are dealt over all
// threads, so no thread does two whole cells while others wait at the barrier
template <class S, int D, int VC, int NT>
__device__ inline void stage_grad(const FArgs &a, const BlockInfo &bi, int nvv, const P2<S> *fb, P2<S> *gr)
{
    if ((int)threadIdx.x < bi.ng) grad_cell<S, D>(a, bi, threadIdx.x, 0, nvv, fb, gr);
    const int rest = (bi.ng - NT) * nvv;
    for (int i = threadIdx.x; i < rest; i += NT) {
        const int vv = i % nvv;
        grad_cell<S, D>(a, bi, NT + i / nvv, vv, vv + 1, fb, gr);
    }
}

// g and h at arm (ax, ay, az) from local cell l: f + grad f . arm
template <class S, int D>
__device__ inline P2<S> recon_l(const P2<S> *fb, const P2<S> *gr, int l, double ax, double ay, double az)
{
    const P2<S> *gl = gr + l * D;
    const P2<S> f = fb[l];
    const S sx = (S)ax, sy = (S)ay;
    if (D == 3) {
        const S sz = (S)az;
        return P2<S>{f.g + (gl[0].g * sx + gl[1].g * sy + gl[2].g * sz),
                     f.h + (gl[0].h * sx + gl[1].h * sy + gl[2].h * sz)};
    }
    return P2<S>{f.g + (gl[0].g * sx + gl[1].g * sy), f.h + (gl[0].h * sx + gl[1].h * sy)};
}

// ------------------------------------------------------------------------- pass A
template <class S, int D, int VC>
__global__ void __launch_bounds__(kNTA, kWavesA)
    k_pass_a(Geo g, FArgs a, const S *__restrict__ gT,
             const S *__restrict__ hT, const double *dtp)
{
    constexpr int NMA = (D == 3 ? 17 : 11) + 1;
    constexpr int CPT = fused_cells_per_thread_a(D);
    int blkid, vg;
    if (!wg_map(a, blkid, vg)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockInfo bi = cload(a.blk, blkid);
    const int ch0 = vg * a.cpg, ch1 = min(a.nchunk, ch0 + a.cpg);
    // LDS: nbuf_a buffers of {f-bar-plus [VC][max_nl], gradients [VC][max_ng][D]}.  With two,
    // chunk c uses buffer c % 2: the next chunk's stage 1 writes the other buffer while slow
    // threads still reconstruct from this one, so the barrier closing each chunk is not needed
    // (a thread reaches the next chunk's first barrier only after its own faces are done).
    P2<S> *const tiles = (P2<S> *)smem;
    const int tile = VC * (a.max_nl + a.max_ng * D);
    const double dt = *dtp;
    const double hx = 0.5 * dt;

    double acc[kSlotsA][NMA];
    int fid[kSlotsA], fown[kSlotsA], fnei[kSlotsA], ftyp[kSlotsA];
    d3 fS[kSlotsA], armO[kSlotsA], armN[kSlotsA];
#pragma unroll
    for (int s = 0; s < kSlotsA; ++s) {
#pragma unroll
        for (int k = 0; k < NMA; ++k) acc[s][k] = 0.0;
        const int j = threadIdx.x + s * kNTA;
        fid[s] = -1;
        fown[s] = fnei[s] = ftyp[s] = 0;
        fS[s] = armO[s] = armN[s] = d3{0, 0, 0};
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

    double cex[CPT][VC];
    const S unused[CPT][VC] = {};
    int lc[CPT];
#pragma unroll
    for (int s = 0; s < CPT; ++s) {
        const int l = threadIdx.x + s * kNTA;
        lc[s] = l < bi.nl ? a.lcell[bi.l_off + l] : 0;
    }
    int stripe = -1;
    for (int ch = ch0; ch < ch1; ++ch) {
        const Chunk ck = chunk_of(a, ch, VC, stripe);
        P2<S> *fb = tiles + (a.nbuf_a == 2 ? ((ch - ch0) & 1) * tile : 0);  // [VC][max_nl]
        P2<S> *gr = fb + VC * a.max_nl;                                      // [VC][max_ng][D]
        stage_bar<S, VC, kNTA, CPT, false>(g, a, bi, gT, hT, dt, ck, lc, unused, unused, cex, fb,
                                           (P2<S> *)nullptr);
        __syncthreads();
        stage_grad<S, D, VC, kNTA>(a, bi, ck.nvv, fb, gr);
        __syncthreads();
        const d3 xr = cload(g.xi, ck.vrow);
#pragma unroll
        for (int s = 0; s < kSlotsA; ++s) {
            if (fid[s] < 0) continue;
            RowSums rs{};
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                if (vv >= ck.nvv) break;
                const int v = ck.vrow + ck.ix0 + vv;
                const d3 xi = cload(g.xi, v);
                const double w = cload(g.w, v);
                const double xs = sigma_dot(xi, fS[s]);
                const P2<S> *pf = fb + vv * a.max_nl, *pq = gr + vv * a.max_ng * D;
                if (ftyp[s] != 0 && !(xs > 0)) continue;  // incoming half added from constants
                // upwind side: the owner for xs >= VSMALL (and outgoing boundary halves), the
                // neighbour for xs < -VSMALL, chosen before one reconstruction (the lanes of a
                // wave hold different faces, so two branches would both run); tangent internal
                // faces average both sides
                const bool tangent = ftyp[s] == 0 && xs < kVSmall && xs >= -kVSmall;
                double gf, hf;
                if (!tangent) {
                    const bool own = ftyp[s] != 0 || xs >= kVSmall;
                    const int l = own ? fown[s] : fnei[s];
                    const double ax = (own ? armO[s].x : armN[s].x) - xi.x * hx;
                    const double ay = (own ? armO[s].y : armN[s].y) - xi.y * hx;
                    const double az = (own ? armO[s].z : armN[s].z) - xi.z * hx;
                    const P2<S> r = recon_l<S, D>(pf, pq, l, ax, ay, az);
                    gf = (double)r.g;
                    hf = (double)r.h;
                } else {
                    const double ox = armO[s].x - xi.x * hx, oy = armO[s].y - xi.y * hx, oz = armO[s].z - xi.z * hx;
                    const double nx = armN[s].x - xi.x * hx, ny = armN[s].y - xi.y * hx, nz = armN[s].z - xi.z * hx;
                    const P2<S> rn = recon_l<S, D>(pf, pq, fnei[s], nx, ny, nz);
                    const P2<S> ro = recon_l<S, D>(pf, pq, fown[s], ox, oy, oz);
                    gf = 0.5 * ((double)rn.g + (double)ro.g);
                    hf = 0.5 * ((double)rn.h + (double)ro.h);
                }
                if (ftyp[s] == kMaxwellWall) acc[s][NMA - 1] += w * xs * gf;
                row_add(rs, w * gf, w * hf, xi.x);
            }
            fold_row<D>(acc[s], rs, xr.y, xr.z);
        }
        if (a.nbuf_a == 1) __syncthreads();
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
    {
        // the storage reference's share: every velocity on internal faces, the outgoing half
        // on boundary faces (the passes accumulate deviations from it)
        const double *r = f < g.nIF ? a.cref : a.bref + (size_t)(f - g.nIF) * NMA;
#pragma unroll
        for (int k = 0; k < NMA; ++k) m[k] += r[k];
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
}

// face equilibrium coefficients, with the relaxation weight h/(2 tau + h) (fvDVM.C:1380)
// and, on patches with q1_compat, R times the internal-face temperature at the
// patch-local index (fvDVM.C:1367, 1413)
// pass-B face slot geometry (one work-group per block)
__global__ void __launch_bounds__(kBlk) k_fb_geo(Geo g, FArgs a)
{
    const BlockInfo bi = cload(a.blk, blockIdx.x);
    for (int j = threadIdx.x; j < bi.nfb; j += kBlk) {
        const int f = a.fb_face[bi.fb_off + j];
        const int lo = a.fb_own[bi.fb_off + j], ln = a.fb_nei[bi.fb_off + j];
        const d3 Cf = g.Cf[f];
        const d3 Co = g.C[a.lcell[bi.l_off + lo]];
        FaceGeo r;
        r.Sf = g.Sf[f];
        r.ao = d3{Cf.x - Co.x, Cf.y - Co.y, Cf.z - Co.z};
        r.an = r.ao;
        if (ln >= 0) {
            const d3 Cn = g.C[a.lcell[bi.l_off + ln]];
            r.an = d3{Cf.x - Cn.x, Cf.y - Cn.y, Cf.z - Cn.z};
        }
        ((FaceGeo *)a.fb_geo)[bi.fb_off + j] = r;
    }
}

__global__ void __launch_bounds__(kBlk) k_face_coef(Geo g, Macros fm, EqCoef *fc, const double *dtp)
{
    const int f = blockIdx.x * kBlk + threadIdx.x;
    if (f >= g.nF) return;
    const double h = 0.5 * (*dtp);
    const double T = fm.T[f];
    double Th = T;
    if (f >= g.nIF && g.q1) Th = fm.T[g.bf_local[f - g.nIF]];
    fc[f] = make_coef(g, fm.rho[f], fm.U[f], T, fm.q[f], g.R * Th, h / (2.0 * fm.tau[f] + h));
}

// face records gathered into pass B's face-slot order, so pass B loads them by slot
__global__ void __launch_bounds__(kBlk) k_slot_coef(FArgs a, int n_slots)
{
    const int j = blockIdx.x * kBlk + threadIdx.x;
    if (j < n_slots) a.fb_coef[j] = a.fc[a.fb_face[j]];
}

// ------------------------------------------------------------------------- pass B
// flux sigma * f of velocity v through block face f (local owner lo, neighbour ln or -1):
// upwind reconstruction at x_f - xi dt/2 relaxed towards the face equilibrium
// (fvDVM.C:1362-1460), or the incoming half of a wall / inlet face; gM is the Maxwellian
// factor A exp(-|xi - U|^2 / 2RT) of the face record k; pf, pq: the velocity's LDS rows
template <class S, int D>
__device__ inline P2<double> face_flux(const Geo &g, const FArgs &a, const P2<S> *pf, const P2<S> *pq, int v,
                                       int f, int lo, int ln, const FaceGeo &fgeo, const EqCoef &k, double gM,
                                       double hx, const double *rho_b)
{
    const d3 xi = cload(g.xi, v);
    const double xs = sigma_dot(xi, fgeo.Sf);
    const int b = f - g.nIF;
    if (b >= 0 && !(xs > 0)) {  // incoming half of a wall or inlet face
        const BndCoef kb = a.bc[b];
        const double scale = g.bf_type[b] == kMaxwellWall ? rho_b[b] : 1.0;
        const double cx = xi.x - kb.ux, cy = xi.y - kb.uy, cz = xi.z - kb.uz;
        const double gv = scale * (kb.A * exp(-((cx * cx + cy * cy + cz * cz) * kb.invRT) / 2.0));
        return P2<double>{xs * gv, xs * (gv * kb.hfac)};
    }
    // upwind side: the owner for xs >= VSMALL (and outgoing boundary halves), the neighbour
    // for xs < -VSMALL, chosen before one reconstruction (the lanes of a wave hold different
    // faces, so two branches would both run); tangent internal faces average both sides
    const bool tangent = b < 0 && xs < kVSmall && xs >= -kVSmall;
    double gv, hv;
    if (!tangent) {
        const bool own = b >= 0 || xs >= kVSmall;
        const double ax = (own ? fgeo.ao.x : fgeo.an.x) - xi.x * hx;
        const double ay = (own ? fgeo.ao.y : fgeo.an.y) - xi.y * hx;
        const double az = (own ? fgeo.ao.z : fgeo.an.z) - xi.z * hx;
        const P2<S> r = recon_l<S, D>(pf, pq, own ? lo : ln, ax, ay, az);
        gv = (double)r.g;
        hv = (double)r.h;
    } else {
        const double ox = fgeo.ao.x - xi.x * hx, oy = fgeo.ao.y - xi.y * hx, oz = fgeo.ao.z - xi.z * hx;
        const double nx = fgeo.an.x - xi.x * hx, ny = fgeo.an.y - xi.y * hx, nz = fgeo.an.z - xi.z * hx;
        const P2<S> rn = recon_l<S, D>(pf, pq, ln, nx, ny, nz);
        const P2<S> ro = recon_l<S, D>(pf, pq, lo, ox, oy, oz);
        gv = 0.5 * ((double)rn.g + (double)ro.g);
        hv = 0.5 * ((double)rn.h + (double)ro.h);
    }
    if (sizeof(S) == 4) {  // reconstructed deviation -> value
        gv += cload(g.ref_g, v);
        hv += cload(g.ref_h, v);
    }
    double gEq, hEq;
    shakhov_poly(g, k, xi, gM, gEq, hEq);
    gv = (1.0 - k.tr) * gv + k.tr * gEq;
    hv = (1.0 - k.tr) * hv + k.tr * hEq;
    return P2<double>{xs * gv, xs * hv};
}

template <class S, int D, int VC>
__global__ void __launch_bounds__(kNTB, kWavesB)
    k_pass_b(Geo g, FArgs a, const S *__restrict__ gT, const S *__restrict__ hT, S *__restrict__ gN,
             S *__restrict__ hN, const double *rho_b, const double *dtp)
{
    constexpr int NM = D == 3 ? 17 : 11;
    constexpr int CPT = fused_cells_per_thread_b(D);
    constexpr int FPT = fused_faces_per_thread_b(D);
    int blkid, vg;
    if (!wg_map(a, blkid, vg)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockInfo bi = cload(a.blk, blkid);
    const int ch0 = vg * a.cpg, ch1 = min(a.nchunk, ch0 + a.cpg);
    P2<double> *fl = (P2<double> *)smem;            // [VC][max_nfb] fluxes sigma * f_f
    P2<S> *fb = (P2<S> *)(fl + VC * a.max_nfb);   // [VC][max_nl]
    P2<S> *gr = fb + VC * a.max_nl;                // [VC][max_ng][D]
    P2<S> *ft = gr + VC * a.max_ng * D;            // [VC][max_nc] stored g~, h~ of the block
    const double dt = *dtp;
    const double hx = 0.5 * dt;

    double mom[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) mom[k] = 0.0;
    const int my = threadIdx.x;
    const bool has_cell = my < bi.nc;
    const int cell = bi.c0 + my;
    double dtV = 0.0;
    int nf = 0;
    int fidx[kNb];
    signed char fsg[kNb];
    if (has_cell) {
        dtV = dt / g.V[cell];
        const int k0 = a.fbc_beg[bi.fbc_off + my];
        nf = a.fbc_beg[bi.fbc_off + my + 1] - k0;
#pragma unroll
        for (int j = 0; j < kNb; ++j) {
            if (j < nf) {
                fidx[j] = a.fbc_idx[k0 + j];
                fsg[j] = a.fbc_sgn[k0 + j];
            }
        }
    }

    constexpr int FW = FPT - 1;  // whole faces per thread in stage 3
    double cex[CPT][VC], fex[FW][VC];
    int lc[CPT];
#pragma unroll
    for (int s = 0; s < CPT; ++s) {
        const int l = threadIdx.x + s * kNTB;
        lc[s] = l < bi.nl ? a.lcell[bi.l_off + l] : 0;
    }
    // 2-D FP32 storage: the next chunk's g~, h~ are loaded before the chunk's last barrier,
    // so their latency overlaps the barrier wait (16 registers; elsewhere the registers
    // are not there to spare)
    constexpr bool kAhead = D == 2 && sizeof(S) == 4;
    S gt[CPT][VC] = {}, ht[CPT][VC] = {};
    int stripe = -1;
    if (kAhead && ch0 < ch1) {
        int st = stripe;
        load_tilde<S, VC, kNTB, CPT>(g, bi, gT, hT, chunk_of(a, ch0, VC, st), lc, gt, ht);
    }
    for (int ch = ch0; ch < ch1; ++ch) {
        const Chunk ck = chunk_of(a, ch, VC, stripe);
        stage_bar<S, VC, kNTB, CPT, kAhead>(g, a, bi, gT, hT, dt, ck, lc, gt, ht, cex, fb,
                                            kAhead ? (P2<S> *)nullptr : ft);
        __syncthreads();
        stage_grad<S, D, VC, kNTB>(a, bi, ck.nvv, fb, gr);
        __syncthreads();
        // stage 3: face values -> flux sigma * f_f in LDS.  Faces j < FW * kNTB: whole faces,
        // FW per thread, x-factors kept for the stripe; the (face, velocity) pairs of the
        // faces beyond are dealt over all threads, so no thread takes a third whole face
        // while the others wait at the barrier
        const d3 xr = cload(g.xi, ck.vrow);
#pragma unroll
        for (int s = 0; s < FW; ++s) {
            const int j = threadIdx.x + s * kNTB;
            if (j >= bi.nfb) break;
            const FaceGeo fgeo = a.fb_geo[bi.fb_off + j];
            const EqCoef k = a.fb_coef[bi.fb_off + j];
            const int f = a.fb_face[bi.fb_off + j];
            const int lo = a.fb_own[bi.fb_off + j], ln = a.fb_nei[bi.fb_off + j];
            if (ck.new_stripe) stripe_factors<VC>(g, k, ck.ix0, a.n1, fex[s]);
            const double eyz = row_factor(k, xr);
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                if (vv >= ck.nvv) break;
                fl[vv * a.max_nfb + j] = face_flux<S, D>(g, a, fb + vv * a.max_nl, gr + vv * a.max_ng * D,
                                                        ck.vrow + ck.ix0 + vv, f, lo, ln, fgeo, k,
                                                        fex[s][vv] * eyz, hx, rho_b);
            }
        }
        {
            const int rest = (bi.nfb - FW * kNTB) * ck.nvv;
            for (int i = threadIdx.x; i < rest; i += kNTB) {
                const int q = ck.nvv == VC ? i / VC : i / ck.nvv;
                const int vv = i - q * ck.nvv;
                const int j = FW * kNTB + q;
                const FaceGeo fgeo = a.fb_geo[bi.fb_off + j];
                const EqCoef k = a.fb_coef[bi.fb_off + j];
                const int f = a.fb_face[bi.fb_off + j];
                const int lo = a.fb_own[bi.fb_off + j], ln = a.fb_nei[bi.fb_off + j];
                // the x-factor of this velocity alone, formed as stripe_factors forms it
                const double dx = cload(g.xi, ck.ix0 + vv).x - k.ux;
                const double ex = k.A * exp(-((dx * dx) * k.invRT) / 2.0);
                fl[vv * a.max_nfb + j] = face_flux<S, D>(g, a, fb + vv * a.max_nl, gr + vv * a.max_ng * D,
                                                        ck.vrow + ck.ix0 + vv, f, lo, ln, fgeo, k,
                                                        ex * row_factor(k, xr), hx, rho_b);
            }
        }
        __syncthreads();
        // stage 4: cell update in ascending face order, new values to the other buffer
        if (has_cell) {
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                if (vv >= ck.nvv) break;
                const int v = ck.vrow + ck.ix0 + vv;
                // g~ of the thread's own cell: still in the registers stage 1 took it from
                // (slot 0 is local cell l = thread, the block cell it updates)
                const P2<S> t = kAhead ? P2<S>{gt[0][vv], ht[0][vv]} : ft[vv * a.max_nc + my];
                const P2<S> bb = fb[vv * a.max_nl + my];
                double ag = -1.0 / 3 * (double)t.g + 4.0 / 3 * (double)bb.g;
                double ah = -1.0 / 3 * (double)t.h + 4.0 / 3 * (double)bb.h;
#pragma unroll
                for (int j = 0; j < kNb; ++j) {
                    if (j < nf) {
                        const P2<double> q = fl[vv * a.max_nfb + fidx[j]];
                        const double tg = q.g * dtV;
                        const double th = q.h * dtV;
                        if (fsg[j] > 0) {
                            ag += tg;
                            ah += th;
                        } else {
                            ag -= tg;
                            ah -= th;
                        }
                    }
                }
                const S sg = (S)ag, sh = (S)ah;
                const size_t kk = (size_t)v * g.nC + cell;
                gN[kk] = sg;
                hN[kk] = sh;
                const double w = cload(g.w, v);
                add_moments<D>(mom, w * (double)sg, w * (double)sh, cload(g.xi, v));  // deviations
            }
        }
        if (kAhead && ch + 1 < ch1) {
            int st = stripe;
            load_tilde<S, VC, kNTB, CPT>(g, bi, gT, hT, chunk_of(a, ch + 1, VC, st), lc, gt, ht);
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
#pragma unroll
    for (int k = 0; k < NM; ++k) m[k] += cload(a.cref, k);  // the storage reference's share
    double rho, T, tau;
    d3 U, q;
    macros_from_moments<D>(g, m, *dtp, rho, U, T, tau, q);
    cm.rho[c] = rho;
    cm.U[c] = U;
    cm.T[c] = T;
    cm.tau[c] = tau;
    cm.q[c] = q;
    a.cc[c] = make_coef(g, rho, U, T, q, g.R * T, tau);
}

__global__ void __launch_bounds__(kBlk) k_cell_coef(Geo g, Macros cm, EqCoef *cc)
{
    const int c = blockIdx.x * kBlk + threadIdx.x;
    if (c >= g.nC) return;
    cc[c] = make_coef(g, cm.rho[c], cm.U[c], cm.T[c], cm.q[c], g.R * cm.T[c], cm.tau[c]);
}

// incoming-half moments of the wall (per unit wall density) and inlet faces, fixed for
// the run (the wall and inlet values are constant, fvDVM.C:578-624, discreteVelocity.C:244-257),
// and the coefficients pass B uses to form the incoming values
template <int D>
__global__ void __launch_bounds__(kBlk) k_incoming(Geo g, double *bin, BndCoef *bc, double *bref)
{
    constexpr int NM = D == 3 ? 17 : 11;
    constexpr int NMA = NM + 1;
    const int b = blockIdx.x * kBlk + threadIdx.x;
    if (b >= g.nB) return;
    const d3 Sf = g.Sf[g.nIF + b];
    const d3 Ub = g.bf_U[b];
    const double Tb = g.bf_T[b];
    const bool wall = g.bf_type[b] == kMaxwellWall;
    BndCoef k{};
    k.A = (wall ? 1.0 : g.bf_rho[b]) / pow(sqrt(2.0 * kPi * g.R * Tb), (double)g.D);
    k.ux = Ub.x; k.uy = Ub.y; k.uz = Ub.z;
    k.invRT = 1.0 / (g.R * Tb);
    k.hfac = (g.K + 3 - g.D) * g.R * Tb;
    bc[b] = k;
    double m[NM];
#pragma unroll
    for (int j = 0; j < NM; ++j) m[j] = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const d3 xi = cload(g.xi, v);
        if (sigma_dot(xi, Sf) > 0) continue;
        const double cx = xi.x - k.ux, cy = xi.y - k.uy, cz = xi.z - k.uz;
        const double gv = k.A * exp(-((cx * cx + cy * cy + cz * cz) * k.invRT) / 2.0);
        add_moments<D>(m, cload(g.w, v) * gv, cload(g.w, v) * (gv * k.hfac), xi);
    }
    for (int j = 0; j < NM; ++j) bin[(size_t)b * NM + j] = m[j];
    // outgoing half of the storage reference (the passes accumulate deviations from it),
    // with its share of the wall's outgoing flux
    double r[NMA];
#pragma unroll
    for (int j = 0; j < NMA; ++j) r[j] = 0.0;
    for (int v = 0; v < g.Nv; ++v) {
        const d3 xi = cload(g.xi, v);
        const double xs = sigma_dot(xi, Sf);
        if (!(xs > 0)) continue;
        const double wg = cload(g.w, v) * cload(g.ref_g, v);
        add_moments<D>(r, wg, cload(g.w, v) * cload(g.ref_h, v), xi);
        if (wall) r[NMA - 1] += wg * xs;
    }
    for (int j = 0; j < NMA; ++j) bref[(size_t)b * NMA + j] = r[j];
}

// moments of the storage reference over all velocities (one thread; run once per init)
template <int D>
__global__ void k_ref_moments(Geo g, double *cref)
{
    constexpr int NMA = (D == 3 ? 17 : 11) + 1;
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double r[NMA];
#pragma unroll
    for (int j = 0; j < NMA; ++j) r[j] = 0.0;
    for (int v = 0; v < g.Nv; ++v) add_moments<D>(r, cload(g.w, v) * cload(g.ref_g, v), cload(g.w, v) * cload(g.ref_h, v), cload(g.xi, v));
    for (int j = 0; j < NMA; ++j) cref[j] = r[j];
}

// f-bar-plus and gradient tiles of one chunk
template <class S>
size_t lds_tiles(const FusedPlan &p, int D, int VC)
{
    return (size_t)VC * (2 * p.max_nl + 2 * p.max_ng * D) * sizeof(S);
}
template <class S>
size_t lds_pass_a(const FusedPlan &p, int D, int VC)
{
    return (size_t)p.nbuf_a * lds_tiles<S>(p, D, VC);
}
template <class S>
size_t lds_pass_b(const FusedPlan &p, int D, int VC)
{
    return (size_t)VC * 2 * p.max_nfb * sizeof(double) + lds_tiles<S>(p, D, VC) +
           (size_t)VC * 2 * p.max_nc * sizeof(S);
}

constexpr size_t kLdsBudget = DUGKS_RC_LDS_KB * 1024;

template <class S>
int choose_vc(const FusedPlan &p, int D)
{
    for (int vc : {4, 2, 1})
        if (lds_pass_b<S>(p, D, vc) <= kLdsBudget) return vc;
    return 0;
}

inline int grid_ab(const FusedPlan &p) { return kFusedXcds * p.b8 * p.nvg; }

template <class S, int D, int VC>
void launch_a(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const double *dt, hipStream_t st)
{
    const size_t lds = lds_pass_a<S>(p, D, VC);
    hipLaunchKernelGGL((k_pass_a<S, D, VC>), dim3(grid_ab(p)), dim3(kNTA), lds, st, g, fargs(p), gT, hT, dt);
}

template <class S, int D, int VC>
void launch_b(const Geo &g, FusedPlan &p, const S *gT, const S *hT, S *gN, S *hN, const double *rho_b,
              const double *dt, hipStream_t st)
{
    const size_t lds = lds_pass_b<S>(p, D, VC);
    hipLaunchKernelGGL((k_pass_b<S, D, VC>), dim3(grid_ab(p)), dim3(kNTB), lds, st, g, fargs(p), gT, hT, gN, hN,
                       rho_b, dt);
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

void fused_velocity_groups(FusedPlan &p, int D, int n1, int vc)
{
    p.vc = vc;
    p.n1 = n1;
    p.nrow = 1;
    for (int k = 1; k < D; ++k) p.nrow *= n1;
    p.nstripe = (n1 + vc - 1) / vc;
    p.nchunk = p.nstripe * p.nrow;
    const int want = std::max(1, std::min(p.nchunk, (p.wg_target + p.nblk - 1) / p.nblk));
    p.cpg = (p.nchunk + want - 1) / want;
    p.nvg = (p.nchunk + p.cpg - 1) / p.cpg;
    p.b8 = (p.nblk + kFusedXcds - 1) / kFusedXcds;
}

bool fused_plan_upload(FusedPlan &p, const Geo &g, size_t es, std::string &err)
{
    const int vc = es == 4 ? choose_vc<float>(p, g.D) : choose_vc<double>(p, g.D);
    if (vc == 0) {
        err = "block tables exceed the LDS budget";
        return false;
    }
    // two pass-A tile buffers where they fit the budget that keeps two work-groups per CU
    {
        const size_t t = es == 4 ? lds_tiles<float>(p, g.D, vc) : lds_tiles<double>(p, g.D, vc);
        p.nbuf_a = 2 * t <= kLdsBudget ? 2 : 1;
    }
    if (g.n1d < 1 || g.Nv != (g.D == 3 ? g.n1d * g.n1d * g.n1d : g.n1d * g.n1d)) {
        err = "fused path needs a tensor-product velocity set";
        return false;
    }
    fused_velocity_groups(p, g.D, g.n1d, vc);
    if (!(es == 4 ? set_lds_attrs<float>(g.D, vc, p) : set_lds_attrs<double>(g.D, vc, p))) {
        err = "cannot raise the dynamic LDS limit";
        return false;
    }
    if (!up(p, &p.d_blk, p.blk, err) || !up(p, &p.d_lcell, p.lcell, err) || !up(p, &p.d_gs_cnt, p.gs_cnt, err) ||
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
    void *fp = nullptr, *cp = nullptr, *bi = nullptr, *cc = nullptr, *fc = nullptr, *bc = nullptr;
    void *br = nullptr, *cr = nullptr;
    if (!dbuf(p, &fp, (size_t)p.nvg * g.nF * p.nma * 8, err) || !dbuf(p, &cp, (size_t)p.nvg * g.nC * p.nm * 8, err) ||
        !dbuf(p, &bi, (size_t)g.nB * p.nm * 8, err) || !dbuf(p, &br, (size_t)g.nB * p.nma * 8, err) ||
        !dbuf(p, &cr, (size_t)p.nma * 8, err) || !dbuf(p, &cc, (size_t)g.nC * sizeof(EqCoef), err) ||
        !dbuf(p, &fc, (size_t)g.nF * sizeof(EqCoef), err) || !dbuf(p, &bc, (size_t)g.nB * sizeof(BndCoef), err) ||
        !dbuf(p, &p.gT2, (size_t)g.Nv * g.nC * es, err) || !dbuf(p, &p.hT2, (size_t)g.Nv * g.nC * es, err))
        return false;
    void *geo = nullptr, *fbc = nullptr;
    if (!dbuf(p, &geo, p.fb_face.size() * sizeof(FaceGeo), err) ||
        !dbuf(p, &fbc, p.fb_face.size() * sizeof(EqCoef), err))
        return false;
    p.fb_geo = (FaceGeo *)geo;
    p.fb_coef = (EqCoef *)fbc;
    hipLaunchKernelGGL(k_fb_geo, dim3(p.nblk), dim3(kBlk), 0, 0, g, fargs(p));
    if (hipDeviceSynchronize() != hipSuccess) {
        err = "forming the face geometry failed";
        return false;
    }
    p.fpart = (double *)fp;
    p.cpart = (double *)cp;
    p.bin = (double *)bi;
    p.bref = (double *)br;
    p.cref = (double *)cr;
    p.cc = (EqCoef *)cc;
    p.fc = (EqCoef *)fc;
    p.bc = (BndCoef *)bc;
    p.lds_a = es == 4 ? lds_pass_a<float>(p, g.D, vc) : lds_pass_a<double>(p, g.D, vc);
    p.lds_b = es == 4 ? lds_pass_b<float>(p, g.D, vc) : lds_pass_b<double>(p, g.D, vc);
    return true;
}

bool fused_init_boundary(FusedPlan &p, const Geo &g, const Macros &cm, hipStream_t st, std::string &err)
{
    if (g.nB > 0) {
        if (g.D == 3)
            hipLaunchKernelGGL(k_incoming<3>, grid1(g.nB), dim3(kBlk), 0, st, g, p.bin, p.bc, p.bref);
        else
            hipLaunchKernelGGL(k_incoming<2>, grid1(g.nB), dim3(kBlk), 0, st, g, p.bin, p.bc, p.bref);
    }
    if (g.D == 3)
        hipLaunchKernelGGL(k_ref_moments<3>, dim3(1), dim3(64), 0, st, g, p.cref);
    else
        hipLaunchKernelGGL(k_ref_moments<2>, dim3(1), dim3(64), 0, st, g, p.cref);
    hipLaunchKernelGGL(k_cell_coef, grid1(g.nC), dim3(kBlk), 0, st, g, cm, p.cc);
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
void Fused<S>::face_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const Macros &, const double *dt,
                         hipStream_t st)
{
#define DUGKS_A(DD, VV) \
    if (g.D == DD && p.vc == VV) return launch_a<S, DD, VV>(g, p, gT, hT, dt, st);
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
    hipLaunchKernelGGL(k_face_coef, grid1(g.nF), dim3(kBlk), 0, st, g, fm, p.fc, dt);
    const int n_slots = (int)p.fb_face.size();
    hipLaunchKernelGGL(k_slot_coef, grid1(n_slots), dim3(kBlk), 0, st, fargs(p), n_slots);
    launch_cfl_partials(g, fm, cfl_part, st);
}

template <class S>
void Fused<S>::update_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, S *gN, S *hN, const Macros &,
                           const Macros &, const double *rho_b, const double *dt, hipStream_t st)
{
#define DUGKS_B(DD, VV) \
    if (g.D == DD && p.vc == VV) return launch_b<S, DD, VV>(g, p, gT, hT, gN, hN, rho_b, dt, st);
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
    hipLaunchKernelGGL(k_cell_coef, grid1(g.nC), dim3(kBlk), 0, st, g, cm, p.cc);
}

template struct Fused<float>;
template struct Fused<double>;

}  // namespace rc
}  // namespace dugks
