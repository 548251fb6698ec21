// Fused DUGKS step kernels (see dugks_fused.h for the data flow).
//
// Work-group = (cell block, velocity group).  Per chunk of up to VC velocities (one x-stripe
// of a velocity row):
//   pass A
//     stage 1  f-bar-plus of the block cells and their halo ring -> LDS
//              (fvDVM::updateGHbarPvol, fvDVM.C:648-683; equilibrium in FP64)
//     stage 2  least-squares gradient of the block cells -> LDS
//              (leastSquaresGrad through discreteVelocity::updateGrad, discreteVelocity.C:366-369)
//     stage 3  faces of the block cells: reconstruction at x_f - xi dt/2 from the upwind cell
//              when it is a block cell (fvDVM.C:824-963), value written for pass B, FP64 raw
//              moments accumulated per (face, side) in registers (fvDVM.C:1181-1297)
//   pass B
//     stage 1  f-bar-plus of the block cells in registers
//     stage 2  faces of the block cells: stored value relaxed to the face equilibrium
//              (fvDVM.C:1362-1460) or the incoming wall / inlet value (fvDVM.C:1054-1098,
//              discreteVelocity.C:244-257); flux sigma f in LDS (FP64)
//     stage 3  g~ <- 4/3 f-bar-plus - 1/3 g~ - dt/V sum(+-flux) in ascending face order
//              (fvDVM.C:1462-1519), in place; FP64 raw moments of the new g~, h~ per cell
//
// 3-D with FP32 storage runs a three-kernel form of the same step:
//   k_bar3     f-bar-plus and u = 4/3 f-bar-plus - 1/3 g~ of every cell, once (streaming)
//   pass A     (BARIN) no stage 1: the next chunk's f-bar-plus of the block and halo goes
//              into LDS by DMA during the face loop; stage 2 also forms
//              f* = f-bar-plus - (dt/2) xi.grad, so stage 3's face value
//              f* + grad.(x_f - x_cell) has a velocity-independent arm; stage 3 picks the
//              upwind side by sigma thresholds with no branch but the face-value stores
//   pass B     (k_pass_b3) 512 threads, records in registers, u instead of stage 1:
//              g~ <- u - dt/V sum(+-flux)
// FP64 storage and 2-D keep the two-pass form above.
//
// Each thread owns fixed cells / faces for the whole work-group.  The passes wait on memory
// latency more than on arithmetic, so the loads of a stage are issued together before their
// values are used (or a chunk ahead), the chunk index is wave-uniform (per-velocity tables
// are scalar loads), and exponentials use a short exp (exp_cw).  Cell Maxwellians use the
// tensor-set factorisation A exp(-(xi_x - U_x)^2 / 2RT) per x-stripe times one exp per chunk
// for the y/z part (pass A in LDS, k_bar3 in registers, pass B3 for its face slots in LDS);
// the two-pass pass B forms each face and cell velocity's factor on its own, which frees the
// registers the cached factors held for the batched loads.  Work-group ids are mapped so
// that the velocity groups of a block run on the same accelerator die and share its L2.
// Reconstruction arithmetic runs in the storage type S (FP32 for the mixed path);
// equilibria, relaxation, update and moments in FP64.  Entries with |xi.S| < VSMALL carry no
// flux (sigma f = 0 in double precision); their face value is the average of both sides'
// reconstructions, which only the face moments use.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dugks_fused.h"

namespace dugks {

namespace {

constexpr int kNTA = kFusedThreadsA;
constexpr int kSlotsA = kFusedSlotsA;
static_assert(2 * kFusedThreadsB <= kFusedThreadsA,
              "pass A takes each gradient (block) cell on two threads, one per half of the chunk");
constexpr int kNTB = kFusedThreadsB;
constexpr int kNb = kFusedMaxNb;
constexpr int kBlk = 256;

template <class S>
struct V3 {
    S x, y, z;
};

// g and h of one (cell, velocity) side by side: one LDS access moves both
template <class S>
struct alignas(2 * sizeof(S)) P2 {
    S g, h;
};


// Streamed (g, h) pairs: face values and u are written once per step and read once by the
// other pass, and the new g~, h~ once by the next step, so their loads and stores carry the
// non-temporal hint and leave the L2 to the halo, record and table lines that are re-read
// (sphere, FP32 storage: pass A 94.3 -> 90.1 ms, pass B 64.4 -> 59.9 ms)
template <class S>
__device__ inline P2<S> ld_nt(const P2<S> *p)
{
    typedef S v2 __attribute__((ext_vector_type(2)));
    const v2 x = __builtin_nontemporal_load((const v2 *)p);
    return P2<S>{x.x, x.y};
}
template <class S>
__device__ inline void st_nt(P2<S> *p, const P2<S> v)
{
    typedef S v2 __attribute__((ext_vector_type(2)));
    v2 x;
    x.x = v.g;
    x.y = v.h;
    __builtin_nontemporal_store(x, (v2 *)p);
}

// the power-sum weights of one velocity, m_k = w xi_x^k (k = 0..3), formed once per plan
struct alignas(16) RowW {
    double m0, m1, m2, m3;
};

struct FArgs {
    const BlockInfo *blk;
    const int *lcell, *gs_beg, *gs_nb, *fs_face, *fs_col, *fs_lo, *fs_ln, *fbc_beg, *fbc_idx;
    d3 *fs_sf;
    EqCoef *fs_coef;
    const signed char *fs_flag, *fbc_sgn;
    const void *gs_ls;
    double *fpart, *cpart;
    const double *bin, *bref, *cref;
    const RowW *rw;  // [Nv] power-sum weights
    EqCoef *cc, *fc;
    const BndCoef *bc;
    int nblk, b8, nvg, n1, nrow, nchunk, cpg, n_slots, n_cols;
    int max_nl, max_nfs, max_nc, nm, nma;
};

FArgs fargs(const FusedPlan &p)
{
    FArgs a;
    a.blk = p.d_blk; a.lcell = p.d_lcell; a.gs_beg = p.d_gs_beg; a.gs_nb = p.d_gs_nb;
    a.fs_face = p.d_fs_face; a.fs_col = p.d_fs_col; a.fs_lo = p.d_fs_lo; a.fs_ln = p.d_fs_ln;
    a.fs_sf = p.fs_sf; a.fs_coef = p.fs_coef;
    a.n_slots = p.n_slots;
    a.n_cols = p.n_cols;
    a.fbc_beg = p.d_fbc_beg; a.fbc_idx = p.d_fbc_idx; a.fs_flag = p.d_fs_flag; a.fbc_sgn = p.d_fbc_sgn;
    a.gs_ls = p.d_gs_ls;
    a.fpart = p.fpart; a.cpart = p.cpart; a.bin = p.bin; a.bref = p.bref; a.cref = p.cref; a.cc = p.cc; a.fc = p.fc; a.bc = p.bc;
    a.rw = (const RowW *)p.rw;
    a.nblk = p.nblk; a.b8 = p.b8; a.nvg = p.nvg; a.n1 = p.n1; a.nrow = p.nrow; a.nchunk = p.nchunk; a.cpg = p.cpg;
    a.max_nl = p.max_nl; a.max_nfs = p.max_nfs; a.max_nc = p.max_nc;
    a.nm = p.nm; a.nma = p.nma;
    return a;
}

// pass B (and the cell partials it writes) use their own velocity chunking
FArgs fargs_b(const FusedPlan &p)
{
    FArgs a = fargs(p);
    a.nvg = p.nvg_b;
    a.cpg = p.cpg_b;
    a.nchunk = p.nchunk_b;
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

// exp(x) for Maxwellian factors (x <= 0, above the underflow limit -708): Cody and Waite's
// argument reduction (Software Manual for the Elementary Functions, 1980) x = n ln2 + r,
// |r| <= ln2/2, with ln2 split into a high part of 32 significant bits (n ln2_hi is exact)
// and a low part, then exp(r) by its Taylor series to degree 13 in Horner form (truncation
// below 5e-18 relative) and the scaling by 2^n.  19 instructions, against about 45 for the
// general-purpose exp; a few units in the last place from exp.
__device__ inline double exp_cw(double x)
{
    const double n = __builtin_rint(x * 1.4426950408889634074);
    double r = fma(-n, 6.93147180369123816490e-01, x);
    r = fma(-n, 1.90821492927058770002e-10, r);
    double p = 1.0 / 6227020800.0;          // 1/13!
    p = fma(p, r, 1.0 / 479001600.0);       // 1/12!
    p = fma(p, r, 1.0 / 39916800.0);
    p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0);
    p = fma(p, r, 1.0 / 40320.0);
    p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0);
    p = fma(p, r, 1.0 / 120.0);
    p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}

// Shakhov equilibrium (fvDVM.C:673-679) of record k at velocity xi, Maxwellian factor
// A exp(-|xi - U|^2 / 2RT) formed here with exp_cw
__device__ inline void shakhov_eq(const Geo &g, const EqCoef &k, const d3 &xi, double &gEq, double &hEq)
{
    const double cx = xi.x - k.ux, cy = xi.y - k.uy, cz = xi.z - k.uz;
    const double s = (cx * cx + cy * cy + cz * cz) * k.invRT;
    const double gM = k.A * exp_cw(-0.5 * s);
    const double cq = cx * k.qx + cy * k.qy + cz * k.qz;
    const double omp = 1.0 - g.Pr;
    const double kd = g.K + 3.0 - g.D;
    gEq = (1.0 + omp * cq * (s - g.D - 2.0)) * gM;
    hEq = (kd + omp * cq * ((s - g.D) * kd - 2 * g.K)) * gM * k.RTh;
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

// sigma_dot(xi, Sf) for a velocity of the chunk's row: the row's y and z products py =
// xi_y Sf_y, pz = xi_z Sf_z are formed once per face and chunk (row_products), then
// (xi_x Sf_x + py) + pz -- the same roundings in the same order, so the upwind side and the
// outgoing / incoming split match sigma_dot bit for bit in both passes
__device__ inline void row_products(const d3 &xr, const d3 &Sf, double &py, double &pz)
{
#pragma clang fp contract(off)
    py = xr.y * Sf.y;
    pz = xr.z * Sf.z;
}
__device__ inline double sigma_row(double xix, double sfx, double py, double pz)
{
#pragma clang fp contract(off)
    const double px = xix * sfx;
    return (px + py) + pz;
}

// exp(-((xi_y - U_y)^2 + (xi_z - U_z)^2) / 2RT) of the chunk's row (xr = xi of the row)
__device__ inline double row_factor(const EqCoef &k, const d3 &xr)
{
    const double dy = xr.y - k.uy, dz = xr.z - k.uz;
    return exp_cw(-0.5 * ((dy * dy + dz * dz) * k.invRT));
}

// A exp(-(xi_x - U_x)^2 / 2RT) for the abscissae ix0 .. ix0+VC-1 of a stripe; xi[ix] is the
// velocity (x[ix], x[0], x[0]) of the tensor set (ix fastest)
template <int VC>
__device__ inline void stripe_factors(const Geo &g, const EqCoef &k, int ix0, int n1, double *ex, int es = 1)
{
#pragma unroll
    for (int vv = 0; vv < VC; ++vv) {
        const int ix = min(ix0 + vv, n1 - 1);
        const double dx = cload(g.xi, ix).x - k.ux;
        ex[vv * es] = k.A * exp_cw(-0.5 * ((dx * dx) * k.invRT));
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

// row_add with the weights of the velocity: six independent FMAs instead of a chain of
// products (g0..g3 = sum m_k g, h0, h1 = sum m_k h); 3-D pass B 53.9 -> 53.2 ms
__device__ inline void row_add_w(RowSums &r, const RowW &m, double gv, double hv)
{
    r.g0 = fma(m.m0, gv, r.g0);
    r.g1 = fma(m.m1, gv, r.g1);
    r.g2 = fma(m.m2, gv, r.g2);
    r.g3 = fma(m.m3, gv, r.g3);
    r.h0 = fma(m.m0, hv, r.h0);
    r.h1 = fma(m.m1, hv, r.h1);
}

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
    bool new_stripe;     // stripe differs from the previous chunk's
};

// The chunk index is the same for every lane of a wave; readfirstlane tells the compiler,
// so the velocity indices derived from it stay in scalar registers and the per-velocity
// tables are read with scalar loads.
__device__ inline Chunk chunk_of(const FArgs &a, int ch, int vc, int &stripe)
{
    const int st = __builtin_amdgcn_readfirstlane(ch / a.nrow);
    const int row = __builtin_amdgcn_readfirstlane(ch - st * a.nrow);
    Chunk c;
    c.ix0 = st * vc;
    c.nvv = min(vc, a.n1 - c.ix0);
    c.vrow = row * a.n1;
    c.new_stripe = st != stripe;
    stripe = st;
    return c;
}

// g~, h~ of cell c for the chunk's velocities (velocities past the end of a short stripe
// repeat its last one)
template <class S, int VC>
__device__ inline void load_cell(const Geo &g, int c, const S *__restrict__ gT, const S *__restrict__ hT,
                                 const Chunk &ck, S *gt, S *ht)
{
#pragma unroll
    for (int vv = 0; vv < VC; ++vv) {
        const size_t kk = (size_t)(ck.vrow + ck.ix0 + min(vv, ck.nvv - 1)) * g.nC + c;
        gt[vv] = gT[kk];
        ht[vv] = hT[kk];
    }
}

// f-bar-plus of cell c (equilibrium record k) for the chunk's velocities (FP64), x-factors
// cex[vv * cs] of the stripe kept by the caller (registers, or LDS with a stride); LOAD:
// g~, h~ are loaded here, else gt, ht hold them
template <class S, int VC, bool LOAD = true>
__device__ inline void cell_bar(const Geo &g, const FArgs &a, const EqCoef &k, int c, const S *__restrict__ gT,
                                const S *__restrict__ hT, double dt, const Chunk &ck, const d3 &xr, double *cex,
                                S *gt, S *ht, double *gb, double *hb, int cs = 1)
{
    if (LOAD) load_cell<S, VC>(g, c, gT, hT, ck, gt, ht);
    if (ck.new_stripe) stripe_factors<VC>(g, k, ck.ix0, a.n1, cex, cs);
    const double eyz = row_factor(k, xr);
    const double af = 1.5 * dt / (2.0 * k.tr + dt);
#pragma unroll
    for (int vv = 0; vv < VC; ++vv) {
        const int v = ck.vrow + ck.ix0 + min(vv, ck.nvv - 1);
        double gEq, hEq;
        shakhov_poly(g, k, cload(g.xi, v), cex[vv * cs] * eyz, gEq, hEq);
        // with FP32 storage, stored values and gb, hb are deviations from the per-velocity
        // reference; FP64 storage keeps values (its reference is zero)
        if (sizeof(S) == 4) {
            gEq -= cload(g.ref_g, v);
            hEq -= cload(g.ref_h, v);
        }
        gb[vv] = (1.0 - af) * (double)gt[vv] + af * gEq;
        hb[vv] = (1.0 - af) * (double)ht[vv] + af * hEq;
    }
}


// f-bar-plus of cell c (equilibrium record k) for the chunk's velocities (FP64), each
// velocity's Maxwellian factor formed on its own (shakhov_eq), so no factors are held in
// registers across chunks; LOAD: g~, h~ are loaded here, else gt, ht hold them
template <class S, int VC, bool LOAD = true>
__device__ inline void cell_bar_eq(const Geo &g, const EqCoef &k, int c, const S *__restrict__ gT,
                                   const S *__restrict__ hT, double dt, const Chunk &ck, S *gt, S *ht, double *gb,
                                   double *hb)
{
    if (LOAD) load_cell<S, VC>(g, c, gT, hT, ck, gt, ht);
    const double af = 1.5 * dt / (2.0 * k.tr + dt);
#pragma unroll
    for (int vv = 0; vv < VC; ++vv) {
        const int v = ck.vrow + ck.ix0 + min(vv, ck.nvv - 1);
        double gEq, hEq;
        shakhov_eq(g, k, cload(g.xi, v), gEq, hEq);
        if (sizeof(S) == 4) {
            gEq -= cload(g.ref_g, v);
            hEq -= cload(g.ref_h, v);
        }
        gb[vv] = (1.0 - af) * (double)gt[vv] + af * gEq;
        hb[vv] = (1.0 - af) * (double)ht[vv] + af * hEq;
    }
}

// pass A stage 1: f-bar-plus of the local cells (block + ring) into LDS.  PRE: g~, h~ of
// the thread's cells were loaded into pg, ph beforehand, else each cell loads its own.
// HK: the cells' equilibrium records are held in kc for the kernel, else loaded per chunk
// CL: the x-factors live in LDS, cexL[vv * max_nl + l] (each thread touches only its own
// cells' entries), else in cex
template <class S, int VC, int NT, int CPT, bool D3, bool PRE, bool HK, bool CL>
__device__ inline void stage_bar(const Geo &g, const FArgs &a, const BlockInfo &bi, const S *__restrict__ gT,
                                 const S *__restrict__ hT, double dt, const Chunk &ck, const int (&lc)[CPT],
                                 const S (&pg)[CPT][VC], const S (&ph)[CPT][VC], const EqCoef (&kc)[CPT],
                                 double (&cex)[CPT][VC], double *cexL, P2<S> *fb)
{
    const d3 xr = cload(g.xi, ck.vrow);
#pragma unroll
    for (int s = 0; s < CPT; ++s) {
        const int l = threadIdx.x + s * NT;
        if (l >= bi.nl) break;
        S gt[VC], ht[VC];
        double gb[VC], hb[VC];
        // the hoisted cell ids cost FP64 3-D pass A registers it does not have (spills)
        const int c = (sizeof(S) == 8 && D3) ? a.lcell[bi.l_off + l] : lc[s];
        if (PRE) {
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                gt[vv] = pg[s][vv];
                ht[vv] = ph[s][vv];
            }
        }
        double *cx = CL ? cexL + l : cex[s];
        const int cs = CL ? a.max_nl : 1;
        if (HK) {
            cell_bar<S, VC, !PRE>(g, a, kc[s], c, gT, hT, dt, ck, xr, cx, gt, ht, gb, hb, cs);
        } else {
            const EqCoef k = a.cc[c];
            cell_bar<S, VC, !PRE>(g, a, k, c, gT, hT, dt, ck, xr, cx, gt, ht, gb, hb, cs);
        }
#pragma unroll
        for (int vv = 0; vv < VC; ++vv) {
            if (vv >= ck.nvv) break;
            fb[vv * a.max_nl + l] = P2<S>{(S)gb[vv], (S)hb[vv]};
        }
    }
}

// pass A stage 2: gradients of the block cells.  Block cell l is taken by threads l and
// l + NT/2, which split the chunk's velocities (even / odd), so the whole work-group works.
// Stencil range (first entry, count) of the thread's gradient cell, fixed for the kernel
template <int NT>
__device__ inline int2 grad_range(const FArgs &a, const BlockInfo &bi)
{
    const int l = threadIdx.x % (NT / 2);
    if (l >= bi.nc) return int2{0, 0};
    const int k0 = a.gs_beg[bi.gsb_off + l];
    return int2{k0, a.gs_beg[bi.gsb_off + l + 1] - k0};
}

// Entries per (cell, velocity) of the gradient tile: the gradient's D components and, with
// FP32 storage, f* = f-bar-plus - (dt/2) xi . grad f-bar-plus, so that a face value is
// f* + grad . (x_f - x_cell) with an arm that does not depend on the velocity.
template <class S, int D>
constexpr int grad_entries() { return sizeof(S) == 4 ? D + 1 : D; }

// The block cells' gradient stencils as an LDS table, [kNb][kNTB] entries (least-squares
// vector, local neighbour index as the fourth word), written once per kernel by the first
// thread of each cell (FP32 storage): the gradient stage of every chunk reads it instead of
// global memory, whose latency followed every chunk's first barrier.
__device__ inline void stencil_table(const FArgs &a, const BlockInfo &bi, int2 gk, float4 *gsl)
{
    const int l = threadIdx.x;
    if (l >= bi.nc) return;
    const V3<float> *ls = (const V3<float> *)a.gs_ls;
#pragma unroll
    for (int j = 0; j < kNb; ++j) {
        if (j < gk.y) {
            const V3<float> L = ls[gk.x + j];
            gsl[j * kNTB + l] = float4{L.x, L.y, L.z, __int_as_float(a.gs_nb[gk.x + j])};
        }
    }
}

template <class S, int D, int VC, int NT, bool GSL = false>
__device__ inline void stage_grad(const Geo &g, const FArgs &a, const BlockInfo &bi, const Chunk &ck, double hdt,
                                  int2 gk, const P2<S> *fb, P2<S> *gr, const float4 *gsl = nullptr)
{
    constexpr int GE = grad_entries<S, D>();
    const int nvv = ck.nvv;
    const V3<S> *ls = (const V3<S> *)a.gs_ls;
    const int l = threadIdx.x % (NT / 2);
    const int v0 = threadIdx.x / (NT / 2);
    if (l < bi.nc) {
        const int k0 = gk.x;
        const int n = gk.y;
        int nb[kNb];
        V3<S> L[kNb];
#pragma unroll
        for (int j = 0; j < kNb; ++j) {
            if (j < n) {
                if constexpr (GSL) {
                    const float4 e = gsl[j * kNTB + l];
                    nb[j] = __float_as_int(e.w);
                    L[j] = V3<S>{e.x, e.y, e.z};
                } else {
                    nb[j] = a.gs_nb[k0 + j];
                    L[j] = ls[k0 + j];
                }
            }
        }
#pragma unroll 1
        for (int vv = v0; vv < nvv; vv += 2) {
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
            // component-major tile, [vv][component][kNTB]: consecutive cells in consecutive
            // words, so neither these writes nor the gathers of stage 3 pile up on a few LDS
            // banks; the fixed stride keeps the component offsets in the instructions
            P2<S> *o = gr + vv * GE * kNTB + l;
            o[0] = P2<S>{gx, hx};
            o[kNTB] = P2<S>{gy, hy};
            if (D == 3) o[2 * kNTB] = P2<S>{gz, hz};
            if (GE > D) {
                const d3 xi = cload(g.xi, ck.vrow + ck.ix0 + vv);
                const S sx = (S)(xi.x * hdt), sy = (S)(xi.y * hdt), sz = (S)(xi.z * hdt);
                S dg = sx * gx + sy * gy, dh = sx * hx + sy * hy;
                if (D == 3) {
                    dg += sz * gz;
                    dh += sz * hz;
                }
                o[D * kNTB] = P2<S>{c.g - dg, c.h - dh};
            }
        }
    }
}

// Gradient stage of the 3-D FP32 pass A that gathers f-bar-plus: the chunk's f-bar-plus as
// (g, h) of velocity pairs, fq[p][nlp] = {g(2p), h(2p), g(2p+1), h(2p+1)}, one 16-byte gather
// per neighbour and pair; each velocity's gradient and f* take the operations of stage_grad
// in the same order.  The two threads of a cell take alternate pairs.
template <int VC, int NT>
__device__ inline void stage_grad_q(const Geo &g, const BlockInfo &bi, const Chunk &ck, double hdt, int2 gk,
                                    const float4 *fq, int nlp, P2<float> *gr, const float4 *gsl)
{
    constexpr int GE = grad_entries<float, 3>();
    const int l = threadIdx.x % (NT / 2);
    // the same for every lane of a wave: the velocity tables are read with scalar loads
    const int h = __builtin_amdgcn_readfirstlane(threadIdx.x / (NT / 2));
    if (l >= bi.nc) return;
    const int n = gk.y;
    int nb[kNb];
    V3<float> L[kNb];
#pragma unroll
    for (int j = 0; j < kNb; ++j) {
        if (j < n) {
            const float4 e = gsl[j * kNTB + l];
            nb[j] = __float_as_int(e.w);
            L[j] = V3<float>{e.x, e.y, e.z};
        }
    }
    auto put = [&](int vv, float cg, float ch, float gx, float gy, float gz, float hx, float hy, float hz) {
        P2<float> *o = gr + vv * GE * kNTB + l;
        o[0] = P2<float>{gx, hx};
        o[kNTB] = P2<float>{gy, hy};
        o[2 * kNTB] = P2<float>{gz, hz};
        const d3 xi = cload(g.xi, ck.vrow + ck.ix0 + vv);
        const float sx = (float)(xi.x * hdt), sy = (float)(xi.y * hdt), sz = (float)(xi.z * hdt);
        float dg = sx * gx + sy * gy, dh = sx * hx + sy * hy;
        dg += sz * gz;
        dh += sz * hz;
        o[3 * kNTB] = P2<float>{cg - dg, ch - dh};
    };
#pragma unroll 1
    for (int p = h; 2 * p < ck.nvv; p += 2) {
        const float4 *q = fq + p * nlp;
        const float4 c = q[l];
        float ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;  // velocity 2p: grad g, grad h
        float cx = 0, cy = 0, cz = 0, dx = 0, dy = 0, dz = 0;  // velocity 2p + 1
#pragma unroll
        for (int j = 0; j < kNb; ++j) {
            if (j < n) {
                const float4 e = q[nb[j]];
                const float dg0 = e.x - c.x, dh0 = e.y - c.y, dg1 = e.z - c.z, dh1 = e.w - c.w;
                ax += L[j].x * dg0; ay += L[j].y * dg0;
                bx += L[j].x * dh0; by += L[j].y * dh0;
                az += L[j].z * dg0; bz += L[j].z * dh0;
                cx += L[j].x * dg1; cy += L[j].y * dg1;
                dx += L[j].x * dh1; dy += L[j].y * dh1;
                cz += L[j].z * dg1; dz += L[j].z * dh1;
            }
        }
        put(2 * p, c.x, c.y, ax, ay, az, bx, by, bz);
        if (2 * p + 1 < ck.nvv) put(2 * p + 1, c.z, c.w, cx, cy, cz, dx, dy, dz);
    }
}

// g and h of local cell l at the face with arm a = x_f - x_cell: f* + grad f . a, from the
// gradient tile with f* entries (FP32 storage)
template <class S, int D>
__device__ inline P2<S> recon_s(const P2<S> *gr, int l, const V3<S> &ar)
{
    constexpr int nc = kNTB;
    const P2<S> *gl = gr + l;
    const P2<S> f = gl[D * nc], g0 = gl[0], g1 = gl[nc];
    if (D == 3) {
        const P2<S> g2 = gl[2 * nc];
        return P2<S>{f.g + (g0.g * ar.x + g1.g * ar.y + g2.g * ar.z), f.h + (g0.h * ar.x + g1.h * ar.y + g2.h * ar.z)};
    }
    return P2<S>{f.g + (g0.g * ar.x + g1.g * ar.y), f.h + (g0.h * ar.x + g1.h * ar.y)};
}

// g and h at arm (ax, ay, az) from local cell l: f + grad f . arm
template <class S, int D>
__device__ inline P2<S> recon_l(const P2<S> *fb, const P2<S> *gr, int l, double ax, double ay, double az)
{
    constexpr int nc = kNTB;
    const P2<S> *gl = gr + l;
    const P2<S> f = fb[l];
    const S sx = (S)ax, sy = (S)ay;
    if (D == 3) {
        const S sz = (S)az;
        return P2<S>{f.g + (gl[0].g * sx + gl[nc].g * sy + gl[2 * nc].g * sz),
                     f.h + (gl[0].h * sx + gl[nc].h * sy + gl[2 * nc].h * sz)};
    }
    return P2<S>{f.g + (gl[0].g * sx + gl[nc].g * sy), f.h + (gl[0].h * sx + gl[nc].h * sy)};
}

// ------------------------------------------------------------------------- pass A
// BARIN: f-bar-plus of every cell comes from Fb (k_bar3), each cell's VC values of a chunk
// one contiguous run that LDS DMA moves into the chunk's tile (no stage 1); otherwise stage 1
// forms them from g~, h~ and the cell records.
template <class S, int D, int VC, bool BARIN>
__global__ void __launch_bounds__(kNTA)
    k_pass_a(Geo g, FArgs a, const S *__restrict__ gT, const S *__restrict__ hT, P2<S> *__restrict__ F,
             const P2<S> *__restrict__ Fb, const double *dtp)
{
    static_assert(!BARIN || (sizeof(S) == 4 && D == 3), "f-bar-plus is gathered in 3-D with FP32 storage only");
    constexpr int NMA = (D == 3 ? 17 : 11) + 1;
    constexpr int CPT = fused_cells_per_thread_a(D);
    int blkid, vg;
    if (!wg_map(a, blkid, vg)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockInfo bi = cload(a.blk, blkid);
    const int ch0 = vg * a.cpg, ch1 = min(a.nchunk, ch0 + a.cpg);
    // BARIN (FP32 storage): the gradient stencils' LDS table first, then the tiles
    float4 *gsl = (float4 *)smem;      // [kNb][kNTB]
    P2<S> *fb = (P2<S> *)(smem + (BARIN ? kNb * kNTB * sizeof(float4) : 0));  // [VC][max_nl]
    // BARIN: f-bar-plus of the local cells as velocity pairs instead, fq[VC / 2][nlp] (the local
    // cells padded to whole waves), filled by LDS DMA
    const int nlp = (a.max_nl + 63) & ~63;
    float4 *fq = (float4 *)fb;
    constexpr int GE = grad_entries<S, D>();
    P2<S> *gr = fb + VC * (BARIN ? nlp : a.max_nl);    // [VC][GE][kNTB]
    // 3-D: the local cells' x-factors in LDS, [VC][max_nl] (registers are short there; the
    // 512-thread work-group is alone on its CU, so the LDS is free)
    constexpr bool kCexLds = D == 3;
    double *cexL = (double *)(gr + VC * kNTB * GE);
    const double dt = *dtp;
    const double hx = 0.5 * dt;

    // Face slots: slot 0 of each thread takes block slot j = thread.  Slot 1 takes j = kNTA +
    // thread in general; in 2-D blocks with at most kNTA / VC slots beyond kNTA (a 16 x 16
    // block has 544) it takes velocity offset vo = thread % VC of slot kNTA + thread / VC
    // instead, so that the few slots beyond kNTA spread over VC lanes each rather than
    // giving a whole second face to some threads while the rest wait at the barrier.  The
    // VC lanes of such a slot add their partial moments at the end, in lane order.
    const bool items = D == 2 && bi.nfs > kNTA && bi.nfs <= kNTA + kNTA / VC;
    const int vo = threadIdx.x % VC;
    auto slot_of = [&](int s) { return (s == 1 && items) ? kNTA + (int)threadIdx.x / VC : (int)threadIdx.x + s * kNTA; };
    double acc[kSlotsA][NMA];
    // fmeta: slot flags; fcol: the face's column of the face-value array
    int fid[kSlotsA], fmeta[kSlotsA], fcol[kSlotsA], flo[kSlotsA], fln[kSlotsA], ftyp[kSlotsA];
    d3 fS[kSlotsA], armO[kSlotsA], armN[kSlotsA];
    V3<S> aO[kSlotsA], aN[kSlotsA];  // the arms in the storage type (FP32 storage: recon_s)
#pragma unroll
    for (int s = 0; s < kSlotsA; ++s) {
#pragma unroll
        for (int k = 0; k < NMA; ++k) acc[s][k] = 0.0;
        const int j = slot_of(s);
        fid[s] = -1;
        flo[s] = fln[s] = fmeta[s] = fcol[s] = ftyp[s] = 0;
        fS[s] = armO[s] = armN[s] = d3{0, 0, 0};
        if (j < bi.nfs) {
            const int f = a.fs_face[bi.fs_off + j];
            fid[s] = f;
            flo[s] = a.fs_lo[bi.fs_off + j];
            fln[s] = a.fs_ln[bi.fs_off + j];
            fmeta[s] = a.fs_flag[bi.fs_off + j];
            fcol[s] = a.fs_col[bi.fs_off + j];
            ftyp[s] = f < g.nIF ? 0 : g.bf_type[f - g.nIF];
            fS[s] = g.Sf[f];
            const d3 Cf = g.Cf[f];
            const d3 Co = g.C[a.lcell[bi.l_off + flo[s]]];
            armO[s] = d3{Cf.x - Co.x, Cf.y - Co.y, Cf.z - Co.z};
            if (fln[s] >= 0) {
                const d3 Cn = g.C[a.lcell[bi.l_off + fln[s]]];
                armN[s] = d3{Cf.x - Cn.x, Cf.y - Cn.y, Cf.z - Cn.z};
            } else {
                armN[s] = armO[s];
            }
        }
        aO[s] = V3<S>{(S)armO[s].x, (S)armO[s].y, (S)armO[s].z};
        aN[s] = V3<S>{(S)armN[s].x, (S)armN[s].y, (S)armN[s].z};
    }

    double cex[CPT][VC];
    int lc[CPT];  // the thread's local cells (fixed for the kernel)
#pragma unroll
    for (int s = 0; s < CPT; ++s) {
        const int l = threadIdx.x + s * kNTA;
        lc[s] = l < bi.nl ? a.lcell[bi.l_off + l] : 0;
    }
    const int2 gk = grad_range<kNTA>(a, bi);
    if constexpr (BARIN) stencil_table(a, bi, gk, gsl);  // read after the first chunk's barrier
    // BARIN face loop: sigma > tO picks the owner side (in the block), sigma < tN the
    // neighbour side, tT <= sigma < VSMALL a tangent internal face; neither: nothing (the
    // incoming half of a boundary face, the upwind cell outside the block, an empty slot).
    // wf = 1 on walls (outgoing mass flux).  flO, flN: the sides' local cells, 0 where the
    // side is not in the block, so every lane reads inside the gradient tile.
    double tO[kSlotsA], tN[kSlotsA], tT[kSlotsA], wf[kSlotsA];
    int flO[kSlotsA], flN[kSlotsA];
    if constexpr (BARIN) {
        const double inf = __builtin_huge_val();
        const double vsmall_below = __longlong_as_double(__double_as_longlong(kVSmall) - 1);  // x > it: x >= VSMALL
#pragma unroll
        for (int s = 0; s < kSlotsA; ++s) {
            const bool valid = fid[s] >= 0, bnd = ftyp[s] != 0;
            const bool oin = fmeta[s] & kSlotOwnIn, nin = fmeta[s] & kSlotNeiIn;
            tO[s] = (!valid || !oin) ? inf : bnd ? 0.0 : vsmall_below;
            tN[s] = (!valid || bnd || !nin) ? -inf : -kVSmall;
            tT[s] = (!valid || bnd) ? inf : -kVSmall;
            wf[s] = (valid && ftyp[s] == kMaxwellWall) ? 1.0 : 0.0;
            flO[s] = oin ? flo[s] : 0;
            flN[s] = nin ? fln[s] : 0;
        }
    }
    // 2-D: the next chunk's g~, h~ are loaded before the chunk's last barrier, so their
    // latency overlaps the barrier wait (3-D has no registers to spare)
    constexpr bool kAhead = D == 2 && !BARIN;
    S gt[CPT][VC] = {}, ht[CPT][VC] = {};
    // 2-D: the local cells' equilibrium records are read once per kernel, not per chunk
    constexpr bool kHoldK = D == 2 && !BARIN;
    EqCoef kc[CPT] = {};
    if (kHoldK) {
#pragma unroll
        for (int s = 0; s < CPT; ++s)
            if (threadIdx.x + s * kNTA < bi.nl) kc[s] = a.cc[lc[s]];
    }
    // BARIN: chunk ch of f-bar-plus of the local cells into fq by LDS DMA, issued after the
    // gradient stage of the chunk before, so the gathers run while its faces are formed and
    // need no registers (the first barrier of the next chunk waits for them); 16 bytes, one
    // velocity pair of one cell, per lane: a wave fills 64 consecutive local cells of a pair,
    // lanes past the last cell read cell 0's run into the padding
    auto fill_fq = [&](int ch) {
        if constexpr (BARIN) {
#pragma unroll
            for (int s = 0; s < CPT; ++s) {
                const int l0 = (int)(threadIdx.x & ~63u) + s * kNTA;
                if (l0 >= bi.nl) break;
                const P2<S> *src = Fb + ((size_t)ch * g.nC + lc[s]) * VC;
#pragma unroll
                for (int p = 0; p < VC / 2; ++p)
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(src + 2 * p),
                                                     (__attribute__((address_space(3))) void *)(fq + p * nlp + l0),
                                                     16, 0, 0);
            }
        }
    };
    if (BARIN && ch0 < ch1) fill_fq(ch0);
    int stripe = -1;
    if (kAhead && ch0 < ch1) {
        int st = stripe;
        const Chunk c0 = chunk_of(a, ch0, VC, st);
#pragma unroll
        for (int s = 0; s < CPT; ++s)
            if (threadIdx.x + s * kNTA < bi.nl) load_cell<S, VC>(g, lc[s], gT, hT, c0, gt[s], ht[s]);
    }
    for (int ch = ch0; ch < ch1; ++ch) {
        const Chunk ck = chunk_of(a, ch, VC, stripe);
        if constexpr (!BARIN)
            stage_bar<S, VC, kNTA, CPT, D == 3, kAhead, kHoldK, kCexLds>(g, a, bi, gT, hT, dt, ck, lc, gt, ht, kc,
                                                                         cex, cexL, fb);
        __syncthreads();
        if constexpr (BARIN)
            stage_grad_q<VC, kNTA>(g, bi, ck, hx, gk, fq, nlp, (P2<float> *)gr, gsl);
        else
            stage_grad<S, D, VC, kNTA>(g, a, bi, ck, hx, gk, fb, gr);
        __syncthreads();
        if (ch + 1 < ch1) fill_fq(ch + 1);
        const d3 xr = cload(g.xi, ck.vrow);
        if constexpr (BARIN) {
            // (FP32 storage, f-bar-plus gathered) velocities outer, the thread's slots inner: the
            // slots share each velocity's scalar loads and their independent work overlaps;
            // unrolled over the chunk, so the velocities' scalar loads are issued together
            // (sphere pass A 61.1 -> 59.2 -> 58.0 ms).  The face logic of the general loop below
            // as compares of sigma with the slot's thresholds and selects: every lane runs the
            // one reconstruction (inactive lanes read a block cell and add zeros), so only the
            // stores branch (58.2 -> 53.8 ms); tangent faces are added after the chunk's loop,
            // a wall's outgoing mass flux sum w sigma g is formed from the row sums and the
            // power sums take the weight table (53.8 -> 50.2 ms).
            RowSums rs[kSlotsA];
            double spy[kSlotsA], spz[kSlotsA];
            bool tg[kSlotsA];  // some velocity of the chunk is tangent to the slot's face
#pragma unroll
            for (int s = 0; s < kSlotsA; ++s) {
                rs[s] = RowSums{};
                row_products(xr, fS[s], spy[s], spz[s]);
                tg[s] = false;
            }
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                if (vv >= ck.nvv) break;
                const int v = ck.vrow + ck.ix0 + vv;
                const double xix = cload((const double *)g.xi, 3 * v);
                const P2<S> *pq = gr + vv * kNTB * GE;
#pragma unroll
                for (int s = 0; s < kSlotsA; ++s) {
                    const double xs = sigma_row(xix, fS[s].x, spy[s], spz[s]);
                    const bool po = xs > tO[s], pn = xs < tN[s];  // owner / neighbour upwind, in the block
                    const bool act = po || pn;
                    tg[s] = tg[s] | ((xs >= tT[s]) & (xs < kVSmall));  // no short-circuit branch
                    const P2<S> r = recon_s<S, D>(pq, po ? flO[s] : flN[s], po ? aO[s] : aN[s]);
                    if (act) st_nt(F + (size_t)v * a.n_cols + fcol[s], r);
                    // scaled by 0 or 1 rather than selected, so that the reconstruction is not
                    // moved into the store's branch (inactive lanes read a block cell: finite)
                    const S am = act ? (S)1 : (S)0;
                    const double gf = (double)(r.g * am), hf = (double)(r.h * am);
                    row_add_w(rs[s], cload(a.rw, v), gf, hf);
                }
            }
#pragma unroll
            for (int s = 0; s < kSlotsA; ++s) {
                if (tg[s]) {  // tangent internal faces: no flux; each side in the block adds half
                    const bool own_in = fmeta[s] & kSlotOwnIn, nei_in = fmeta[s] & kSlotNeiIn;
                    for (int vv = 0; vv < ck.nvv; ++vv) {
                        const int v = ck.vrow + ck.ix0 + vv;
                        const double xix = cload((const double *)g.xi, 3 * v);
                        const double xs = sigma_row(xix, fS[s].x, spy[s], spz[s]);
                        if (!(xs >= tT[s] && xs < kVSmall)) continue;
                        const P2<S> *pq = gr + vv * kNTB * GE;
                        double gf, hf;
                        if (own_in && nei_in) {
                            const P2<S> rn = recon_s<S, D>(pq, fln[s], aN[s]);
                            const P2<S> ro = recon_s<S, D>(pq, flo[s], aO[s]);
                            gf = 0.5 * ((double)rn.g + (double)ro.g);
                            hf = 0.5 * ((double)rn.h + (double)ro.h);
                        } else {
                            const P2<S> rt = own_in ? recon_s<S, D>(pq, flo[s], aO[s]) : recon_s<S, D>(pq, fln[s], aN[s]);
                            gf = 0.5 * (double)rt.g;
                            hf = 0.5 * (double)rt.h;
                        }
                        row_add_w(rs[s], cload(a.rw, v), gf, hf);
                    }
                }
                // walls: sum w sigma g = S_x sum w xi_x g + (xi_y S_y + xi_z S_z) sum w g
                acc[s][NMA - 1] += wf[s] * (fS[s].x * rs[s].g1 + (spy[s] + spz[s]) * rs[s].g0);
            }
#pragma unroll
            for (int s = 0; s < kSlotsA; ++s)
                if (fid[s] >= 0) fold_row<D>(acc[s], rs[s], xr.y, xr.z);
        } else
#pragma unroll
        for (int s = 0; s < kSlotsA; ++s) {
            if (fid[s] < 0) continue;
            RowSums rs{};
            const bool own_in = fmeta[s] & kSlotOwnIn, nei_in = fmeta[s] & kSlotNeiIn;
            const bool split = s == 1 && items;
            const int vb = split ? vo : 0, ve = split ? min(vo + 1, ck.nvv) : ck.nvv;
            // BARIN: two velocities per iteration, so the scalar and LDS loads of both overlap
            // (the forms that build f-bar-plus themselves spill with it)
            constexpr int kUnroll = BARIN ? 2 : 1;
            double spy, spz;
            row_products(xr, fS[s], spy, spz);
#pragma unroll kUnroll
            for (int vv = vb; vv < ve; ++vv) {
                const int v = ck.vrow + ck.ix0 + vv;
                const d3 xi{cload((const double *)g.xi, 3 * v), xr.y, xr.z};
                const double w = cload(g.w, v);
                const double xs = sigma_row(xi.x, fS[s].x, spy, spz);
                const P2<S> *pf = fb + vv * a.max_nl, *pq = gr + vv * kNTB * GE;
                double gf, hf;
                bool store = true;
                // FP32 storage: the upwind side (the owner for xs >= VSMALL and outgoing boundary
                // halves, the neighbour for xs < -VSMALL) is chosen before one reconstruction;
                // the lanes of a wave hold different faces, so two branches would both run
                // (sphere pass A 115.1 -> 105.5 ms).  With FP64 storage the chosen form costs
                // registers the pass does not have (cavity +2.6 %, sphere +11 %), so it keeps
                // the branches.  The block holding the upwind cell writes the value; tangent
                // internal faces carry no flux and each side adds half of its reconstruction
                // to the moments.
                constexpr bool kPick = sizeof(S) == 4;
                if (kPick) {
                    if (ftyp[s] != 0 && !(xs > 0)) continue;  // incoming half added from constants
                    const bool tangent = ftyp[s] == 0 && xs < kVSmall && xs >= -kVSmall;
                    if (!tangent) {
                        const bool own = ftyp[s] != 0 || xs >= kVSmall;
                        if (!(own ? own_in : nei_in)) continue;
                        const V3<S> ar = own ? aO[s] : aN[s];
                        const P2<S> r = recon_s<S, D>(pq, own ? flo[s] : fln[s], ar);
                        gf = (double)r.g;
                        hf = (double)r.h;
                    } else {
                        store = false;
                        if (own_in && nei_in) {
                            const P2<S> rn = recon_s<S, D>(pq, fln[s], aN[s]);
                            const P2<S> ro = recon_s<S, D>(pq, flo[s], aO[s]);
                            gf = 0.5 * ((double)rn.g + (double)ro.g);
                            hf = 0.5 * ((double)rn.h + (double)ro.h);
                        } else if (own_in) {
                            const P2<S> r = recon_s<S, D>(pq, flo[s], aO[s]);
                            gf = 0.5 * (double)r.g;
                            hf = 0.5 * (double)r.h;
                        } else {
                            const P2<S> r = recon_s<S, D>(pq, fln[s], aN[s]);
                            gf = 0.5 * (double)r.g;
                            hf = 0.5 * (double)r.h;
                        }
                    }
                } else if (ftyp[s] == 0) {
                    if (xs >= kVSmall) {
                        if (!own_in) continue;
                        const double ox = armO[s].x - xi.x * hx, oy = armO[s].y - xi.y * hx, oz = armO[s].z - xi.z * hx;
                        const P2<S> r = recon_l<S, D>(pf, pq, flo[s], ox, oy, oz);
                        gf = (double)r.g;
                        hf = (double)r.h;
                    } else if (xs < -kVSmall) {
                        if (!nei_in) continue;
                        const double nx = armN[s].x - xi.x * hx, ny = armN[s].y - xi.y * hx, nz = armN[s].z - xi.z * hx;
                        const P2<S> r = recon_l<S, D>(pf, pq, fln[s], nx, ny, nz);
                        gf = (double)r.g;
                        hf = (double)r.h;
                    } else {
                        store = false;
                        const double ox = armO[s].x - xi.x * hx, oy = armO[s].y - xi.y * hx, oz = armO[s].z - xi.z * hx;
                        const double nx = armN[s].x - xi.x * hx, ny = armN[s].y - xi.y * hx, nz = armN[s].z - xi.z * hx;
                        if (own_in && nei_in) {
                            const P2<S> rn = recon_l<S, D>(pf, pq, fln[s], nx, ny, nz);
                            const P2<S> ro = recon_l<S, D>(pf, pq, flo[s], ox, oy, oz);
                            gf = 0.5 * ((double)rn.g + (double)ro.g);
                            hf = 0.5 * ((double)rn.h + (double)ro.h);
                        } else if (own_in) {
                            const P2<S> r = recon_l<S, D>(pf, pq, flo[s], ox, oy, oz);
                            gf = 0.5 * (double)r.g;
                            hf = 0.5 * (double)r.h;
                        } else {
                            const P2<S> r = recon_l<S, D>(pf, pq, fln[s], nx, ny, nz);
                            gf = 0.5 * (double)r.g;
                            hf = 0.5 * (double)r.h;
                        }
                    }
                } else {
                    if (!(xs > 0)) continue;  // incoming half added from constants
                    const double ox = armO[s].x - xi.x * hx, oy = armO[s].y - xi.y * hx, oz = armO[s].z - xi.z * hx;
                    const P2<S> r = recon_l<S, D>(pf, pq, flo[s], ox, oy, oz);
                    gf = (double)r.g;
                    hf = (double)r.h;
                }
                if (store)  // face values are stored as deviations, one column per face
                    st_nt(F + (size_t)v * a.n_cols + fcol[s], P2<S>{(S)gf, (S)hf});
                if (ftyp[s] == kMaxwellWall) acc[s][NMA - 1] += w * xs * gf;
                // (the weight table of pass B measured slower here: its scalar load sits on
                // the loop's path)
                row_add(rs, w * gf, w * hf, xi.x);
            }
            fold_row<D>(acc[s], rs, xr.y, xr.z);
        }
        if (kAhead && ch + 1 < ch1) {
            int st = stripe;
            const Chunk nx = chunk_of(a, ch + 1, VC, st);
#pragma unroll
            for (int s = 0; s < CPT; ++s)
                if (threadIdx.x + s * kNTA < bi.nl) load_cell<S, VC>(g, lc[s], gT, hT, nx, gt[s], ht[s]);
        }
        // BARIN: the next chunk's first barrier ends the chunk (no stage 1 writes before it)
        if constexpr (!BARIN) __syncthreads();
    }
    if (items) {  // uniform over the work-group: every lane takes part in the shuffles
#pragma unroll
        for (int k = 0; k < NMA; ++k) {
            const double t = acc[1][k];
            double sum = t;
#pragma unroll
            for (int d = 1; d < VC; ++d) sum += __shfl_down(t, d, VC);
            acc[1][k] = sum;
        }
        if (vo != 0) fid[1] = -1;
    }
#pragma unroll
    for (int s = 0; s < kSlotsA; ++s) {
        if (fid[s] < 0) continue;
        const int side = (fmeta[s] & kSlotOwnIn) ? 0 : 1;
        double *o = a.fpart + (((size_t)vg * g.nF + fid[s]) * 2 + side) * NMA;
#pragma unroll
        for (int k = 0; k < NMA; ++k) o[k] = acc[s][k];
    }
}

// ---------------------------------------------------------- f-bar-plus of every cell
// 3-D with FP32 storage: f-bar-plus = (1 - a) g~ + a g^S (fvDVM::updateGHbarPvol,
// fvDVM.C:648-683, FP64) and u = 4/3 f-bar-plus - 1/3 g~ of every cell, both stored as
// deviations in FP32 for the two passes.  One thread per cell, a range of velocity rows per
// work-group row, stripe by stripe; the cell's Maxwellian is its x-factor (per abscissa of
// the stripe, in registers) times one row factor per row, and the record stays in registers.  u is stored [v][cell] like g~ (pass B
// reads it over cells); f-bar-plus in pass A's chunks, [chunk][cell][VC] with chunk = stripe
// * nrow + row: a cell's VC values of a chunk are one contiguous 8 VC-byte run, so pass A's
// gathers of halo cells, which lie scattered in memory, move whole runs (sphere pass A HBM
// fetch 116 GB per launch with the [v][cell] layout).
constexpr int kBarRowGroups = 8;

template <class S, int VC>
__global__ void __launch_bounds__(kBlk)
    k_bar3(Geo g, const EqCoef *__restrict__ cc, int n1, const S *__restrict__ gT, const S *__restrict__ hT,
           P2<S> *__restrict__ Fb, P2<S> *__restrict__ U, const double *dtp)
{
    static_assert(16 % sizeof(P2<S>) == 0 && (VC * sizeof(P2<S>)) % 16 == 0, "16-byte pieces of whole pairs");
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x % 64;
    // the wave's store patch, [64][VC] pairs
    P2<S> *patch = (P2<S> *)smem + (threadIdx.x / 64) * 64 * VC;
    const int c = blockIdx.x * kBlk + threadIdx.x;
    const int wave_c0 = c - lane;
    const bool live = c < g.nC;
    const int cs = live ? c : g.nC - 1;
    const EqCoef k = cc[cs];
    const double dt = *dtp;
    const double af = 1.5 * dt / (2.0 * k.tr + dt);
    const int nrow = n1 * n1;
    const int nstripe = (n1 + VC - 1) / VC;
    const int rpg = (nrow + kBarRowGroups - 1) / kBarRowGroups;
    const int r0 = blockIdx.y * rpg, r1 = min(nrow, r0 + rpg);
    for (int st = 0; st < nstripe; ++st) {
        double ex[VC];  // the stripe's x-factors of the cell
#pragma unroll
        for (int vv = 0; vv < VC; ++vv) {
            const double dx = cload(g.xi, min(st * VC + vv, n1 - 1)).x - k.ux;
            ex[vv] = k.A * exp_cw(-0.5 * ((dx * dx) * k.invRT));
        }
        for (int row = r0; row < r1; ++row) {
            const int vrow = row * n1;
            const double eyz = row_factor(k, cload(g.xi, vrow));
            P2<S> fbv[VC];
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                const int ix = min(st * VC + vv, n1 - 1);  // a short last stripe repeats its last abscissa
                const int v = vrow + ix;
                const size_t kk = (size_t)v * g.nC + cs;
                const S gt = gT[kk], ht = hT[kk];
                double gEq, hEq;
                shakhov_poly(g, k, cload(g.xi, v), ex[vv] * eyz, gEq, hEq);
                gEq -= cload(g.ref_g, v);
                hEq -= cload(g.ref_h, v);
                const double gb = (1.0 - af) * (double)gt + af * gEq;
                const double hb = (1.0 - af) * (double)ht + af * hEq;
                fbv[vv] = P2<S>{(S)gb, (S)hb};
                if (live && st * VC + vv < n1)
                    st_nt(U + kk, P2<S>{(S)(-1.0 / 3 * (double)gt + 4.0 / 3 * gb),
                                        (S)(-1.0 / 3 * (double)ht + 4.0 / 3 * hb)});
            }
            // the wave's 64 cells hold one contiguous run of 64 VC pairs: staged through the
            // wave's LDS patch so that each store instruction writes whole consecutive
            // 16-byte pieces (per-lane runs written straight out leave every instruction with
            // a quarter of each segment)
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) patch[lane * VC + vv] = fbv[vv];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            P2<S> *o = Fb + ((size_t)(st * nrow + row) * g.nC + wave_c0) * VC;
            constexpr int PP = 16 / sizeof(P2<S>);  // pairs per 16-byte piece
#pragma unroll
            for (int q = 0; q < VC / PP; ++q) {
                const int e = (q * 64 + lane) * PP;  // first pair of the lane's piece
                if (wave_c0 + e / VC < g.nC) {
                    typedef unsigned u4 __attribute__((ext_vector_type(4)));
                    __builtin_nontemporal_store(*(const u4 *)(patch + e), (u4 *)(o + e));
                }
            }
            __builtin_amdgcn_wave_barrier();
        }
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
        const double *p = a.fpart + ((size_t)vg * g.nF + f) * 2 * NMA;
#pragma unroll
        for (int k = 0; k < NMA; ++k) m[k] += p[k];
#pragma unroll
        for (int k = 0; k < NMA; ++k) m[k] += p[NMA + k];
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

// face records and area vectors in face-slot order, so pass B loads them by slot
__global__ void __launch_bounds__(kBlk) k_slot_coef(FArgs a, int n_slots)
{
    const int j = blockIdx.x * kBlk + threadIdx.x;
    if (j < n_slots) a.fs_coef[j] = a.fc[a.fs_face[j]];
}

__global__ void __launch_bounds__(kBlk) k_slot_sf(Geo g, FArgs a, int n_slots)
{
    const int j = blockIdx.x * kBlk + threadIdx.x;
    if (j < n_slots) a.fs_sf[j] = g.Sf[a.fs_face[j]];
}

// ------------------------------------------------------------------------- pass B
// what pass B reads of one velocity: xi and (FP32 storage) the storage reference.  Loaded
// outside the branches of slot_flux, where a wave-uniform velocity makes them scalar loads.
template <class S>
struct VelK {
    d3 xi;
    double rg, rh;
};
template <class S>
__device__ inline VelK<S> vel_k(const Geo &g, int v)
{
    VelK<S> q;
    q.xi = cload(g.xi, v);
    q.rg = sizeof(S) == 4 ? cload(g.ref_g, v) : 0.0;
    q.rh = sizeof(S) == 4 ? cload(g.ref_h, v) : 0.0;
    return q;
}

// flux sigma * f of velocity v through a face slot (face f, area vector Sf, equilibrium
// record k) whose stored value is fv: the stored face value relaxed towards the face
// equilibrium (fvDVM.C:1362-1460), or the incoming half of a wall / inlet face.  Tangent
// velocities of internal faces carry no flux.  The caller loads fv for every slot and
// velocity, incoming entries included (they are not read), and the velocity's constants vk,
// so that the loads of a slot are in flight together instead of one behind the other under
// this branch.  FP32 slots hold deviations from the storage reference.
template <class S>
__device__ inline P2<double> slot_flux(const Geo &g, const FArgs &a, const P2<S> fv, int f, const VelK<S> &vk,
                                       const d3 &Sf, const EqCoef &k, const double *rho_b)
{
    const d3 xi = vk.xi;
    const double xs = sigma_dot(xi, Sf);
    const int b = f - g.nIF;
    if (b < 0 ? (xs >= kVSmall || xs < -kVSmall) : xs > 0) {
        double gEq, hEq;
        shakhov_eq(g, k, xi, gEq, hEq);
        return P2<double>{xs * ((1.0 - k.tr) * ((double)fv.g + vk.rg) + k.tr * gEq),
                          xs * ((1.0 - k.tr) * ((double)fv.h + vk.rh) + k.tr * hEq)};
    }
    if (b >= 0) {  // incoming half of a wall or inlet face
        const BndCoef kb = a.bc[b];
        const double scale = g.bf_type[b] == kMaxwellWall ? rho_b[b] : 1.0;
        const double cx = xi.x - kb.ux, cy = xi.y - kb.uy, cz = xi.z - kb.uz;
        const double gv = scale * (kb.A * exp_cw(-0.5 * ((cx * cx + cy * cy + cz * cz) * kb.invRT)));
        return P2<double>{xs * gv, xs * (gv * kb.hfac)};
    }
    return P2<double>{0.0, 0.0};
}

// Shakhov equilibrium (fvDVM.C:673-679) for a velocity of the chunk's row given the row's
// parts of s = |c|^2 / RT and c.q (c = xi - U): sy = (c_y^2 + c_z^2) / RT and qy = c_y q_y +
// c_z q_z (RowEq, formed once per face and chunk), and gM = A exp(-|c|^2 / 2RT); three
// operations per velocity for s and c.q instead of ten (3-D pass B 52.9 -> 51.6 ms)
struct RowEq {
    double sy, qy;
};
__device__ inline RowEq row_eq(const EqCoef &k, const d3 &xr)
{
    const double cy = xr.y - k.uy, cz = xr.z - k.uz;
    return RowEq{(cy * cy + cz * cz) * k.invRT, cy * k.qy + cz * k.qz};
}
__device__ inline void shakhov_row(const Geo &g, const EqCoef &k, const RowEq &r, double xix, double gM, double &gEq,
                                   double &hEq)
{
    const double cx = xix - k.ux;
    const double s = fma(cx * cx, k.invRT, r.sy);
    const double cq = fma(cx, k.qx, r.qy);
    const double omp = 1.0 - g.Pr;
    const double kd = g.K + 3.0 - g.D;
    gEq = (1.0 + omp * cq * (s - g.D - 2.0)) * gM;
    hEq = (kd + omp * cq * ((s - g.D) * kd - 2 * g.K)) * gM * k.RTh;
}

// slot_flux with the face Maxwellian factor gM = A exp(-|xi - U|^2 / 2RT) given by the caller
template <class S>
__device__ inline P2<double> slot_flux_m(const Geo &g, const FArgs &a, const P2<S> fv, int f, const VelK<S> &vk,
                                         const d3 &Sf, double spy, double spz, const EqCoef &k, const RowEq &re,
                                         double gM, const double *rho_b)
{
    const d3 xi = vk.xi;
    const double xs = sigma_row(xi.x, Sf.x, spy, spz);
    const int b = f - g.nIF;
    if (b < 0 ? (xs >= kVSmall || xs < -kVSmall) : xs > 0) {
        double gEq, hEq;
        shakhov_row(g, k, re, xi.x, gM, gEq, hEq);
        return P2<double>{xs * ((1.0 - k.tr) * ((double)fv.g + vk.rg) + k.tr * gEq),
                          xs * ((1.0 - k.tr) * ((double)fv.h + vk.rh) + k.tr * hEq)};
    }
    if (b >= 0) {  // incoming half of a wall or inlet face
        const BndCoef kb = a.bc[b];
        const double scale = g.bf_type[b] == kMaxwellWall ? rho_b[b] : 1.0;
        const double cx = xi.x - kb.ux, cy = xi.y - kb.uy, cz = xi.z - kb.uz;
        const double gv = scale * (kb.A * exp_cw(-0.5 * ((cx * cx + cy * cy + cz * cz) * kb.invRT)));
        return P2<double>{xs * gv, xs * (gv * kb.hfac)};
    }
    return P2<double>{0.0, 0.0};
}

template <class S, int D, int VC>
__global__ void __launch_bounds__(kNTB)
    k_pass_b(Geo g, FArgs a, S *gT, S *hT, const P2<S> *__restrict__ F, const double *rho_b,
             const double *dtp)
{
    constexpr int NM = D == 3 ? 17 : 11;
    constexpr int FPT = fused_faces_per_thread_b(D);
    int blkid, vg;
    if (!wg_map(a, blkid, vg)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockInfo bi = cload(a.blk, blkid);
    const int ch0 = vg * a.cpg, ch1 = min(a.nchunk, ch0 + a.cpg);
    P2<double> *fl = (P2<double> *)smem;            // [VC][max_nfs] fluxes sigma * f_f
    const double dt = *dtp;

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

    // whole faces per thread in stage 2: every slot the plan allows (FPT per thread), so that
    // a thread issues each of its slots' loads together; the dealt (slot, velocity) items
    // below take only what a larger block would leave (sphere pass B 101.9 -> 97.3 ms with
    // FP32 storage against FPT - 1 whole faces plus about three dealt rounds)
    constexpr int FW = FPT;
    // 2-D FP64: the next chunk's g~, h~ of the cell are loaded before the chunk's last
    // barrier (7.52 -> 7.42 ms on the 256^2 cavity; with FP32 storage it measured slower, and
    // 3-D has no registers to spare).  The update below writes only this chunk's velocities
    // of the cell, so the early loads read values no one writes first.
    constexpr bool kAhead = D == 2 && sizeof(S) == 8;
    S pg[VC] = {}, ph[VC] = {};
    int stripe = -1;
    if (kAhead && has_cell && ch0 < ch1) {
        int st = stripe;
        load_cell<S, VC>(g, cell, gT, hT, chunk_of(a, ch0, VC, st), pg, ph);
    }
    for (int ch = ch0; ch < ch1; ++ch) {
        const Chunk ck = chunk_of(a, ch, VC, stripe);
        const d3 xr = cload(g.xi, ck.vrow);
        // stage 1: f-bar-plus of the thread's cell
        S gt[VC], ht[VC];
        double gb[VC], hb[VC];
        if (kAhead) {
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                gt[vv] = pg[vv];
                ht[vv] = ph[vv];
            }
        }
        if (has_cell) {
            const EqCoef k = a.cc[cell];
            cell_bar_eq<S, VC, !kAhead>(g, k, cell, gT, hT, dt, ck, gt, ht, gb, hb);
        }
        // stage 2: face values -> flux sigma * f_f in LDS.  Slots j < FW * kNTB: whole faces,
        // FW per thread; the (slot, velocity) pairs of any slots beyond are dealt over all
        // threads.  Each slot's loads (stored values of the chunk's velocities, records) are
        // issued together before any is used.
        // Per-slot copies are loads that do not wait on the face label: the face records
        // always, the area vectors except in 3-D with FP64 storage (there they cost the
        // registers of the second wave per SIMD).  A shared face's value is in the face's one
        // column, written by whichever block held the upwind cell.
#pragma unroll
        for (int s = 0; s < FW; ++s) {
            const int j = threadIdx.x + s * kNTB;
            if (j >= bi.nfs) break;
            const int own_slot = bi.fs_off + j;
            P2<S> fv[VC];
#pragma unroll
            for (int vv = 0; vv < VC; ++vv)
                fv[vv] = ld_nt(F + (size_t)(ck.vrow + ck.ix0 + min(vv, ck.nvv - 1)) * a.n_cols + a.fs_col[own_slot]);
            const int f = a.fs_face[own_slot];
            const d3 Sf = (D == 2 || sizeof(S) == 4) ? a.fs_sf[own_slot] : g.Sf[f];
            const EqCoef k = a.fs_coef[own_slot];
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                if (vv >= ck.nvv) break;
                const VelK<S> vk = vel_k<S>(g, ck.vrow + ck.ix0 + vv);
                fl[vv * a.max_nfs + j] = slot_flux<S>(g, a, fv[vv], f, vk, Sf, k, rho_b);
            }
        }
        {
            const int rest = (bi.nfs - FW * kNTB) * ck.nvv;
            for (int i = threadIdx.x; i < rest; i += kNTB) {
                const int q = ck.nvv == VC ? i / VC : i / ck.nvv;
                const int vv = i - q * ck.nvv;
                const int j = FW * kNTB + q;
                const int own_slot = bi.fs_off + j;
                const P2<S> fv = ld_nt(F + (size_t)(ck.vrow + ck.ix0 + vv) * a.n_cols + a.fs_col[own_slot]);
                const int f = a.fs_face[own_slot];
                const d3 Sf = (D == 2 || sizeof(S) == 4) ? a.fs_sf[own_slot] : g.Sf[f];
                const EqCoef k = a.fs_coef[own_slot];
                fl[vv * a.max_nfs + j] = slot_flux<S>(g, a, fv, f, vel_k<S>(g, ck.vrow + ck.ix0 + vv), Sf, k, rho_b);
            }
        }
        __syncthreads();
        // stage 3: cell update in ascending face order, in place
        if (has_cell) {
            RowSums rs{};
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                if (vv >= ck.nvv) break;
                const int v = ck.vrow + ck.ix0 + vv;
                double ag = -1.0 / 3 * (double)gt[vv] + 4.0 / 3 * gb[vv];
                double ah = -1.0 / 3 * (double)ht[vv] + 4.0 / 3 * hb[vv];
#pragma unroll
                for (int j = 0; j < kNb; ++j) {
                    if (j < nf) {
                        const P2<double> q = fl[vv * a.max_nfs + fidx[j]];
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
                __builtin_nontemporal_store(sg, gT + kk);
                __builtin_nontemporal_store(sh, hT + kk);
                const double w = cload(g.w, v);
                row_add(rs, w * (double)sg, w * (double)sh, cload(g.xi, v).x);  // deviations
            }
            fold_row<D>(mom, rs, xr.y, xr.z);
            if (kAhead && ch + 1 < ch1) {
                int st = stripe;
                load_cell<S, VC>(g, cell, gT, hT, chunk_of(a, ch + 1, VC, st), pg, ph);
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

// ------------------------------------------------------------------ pass B, 3-D
// 512 threads per work-group: two threads per block cell, each taking half of the chunk's
// velocities in the update, and two face slots per thread for the fluxes.  k_bar3 wrote
// u = 4/3 f-bar-plus - 1/3 g~ of every cell (U), so the update is g~ <- u - dt/V
// sum(+-flux) and this pass needs no cell equilibrium.  What does not change over the
// velocities is loaded once per kernel and held in registers: the cell's face table, each
// slot's face record and area vector.  The next chunk's stored face values and u are loaded
// while the current chunk is computed.  The cell's raw moments are summed per half and the
// halves added in a fixed order at the end.
constexpr int kNTB3 = 2 * kFusedThreadsB;
constexpr int kSlotsB3 = 2;

template <class S, int VC>
__global__ void __launch_bounds__(kNTB3)
    k_pass_b3(Geo g, FArgs a, S *gT, S *hT, const P2<S> *__restrict__ F, const P2<S> *__restrict__ U,
              const double *rho_b, const double *dtp)
{
    constexpr int NM = 17;
    constexpr int HV = VC / 2;
    static_assert(VC % 2 == 0, "the two threads of a cell split the chunk's velocities");
    int blkid, vg;
    if (!wg_map(a, blkid, vg)) return;
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockInfo bi = cload(a.blk, blkid);
    const int ch0 = vg * a.cpg, ch1 = min(a.nchunk, ch0 + a.cpg);
    // two flux tiles [VC][FS] (sigma * f_f), one per chunk parity: stage 2 of the next chunk
    // writes the other tile, so one barrier per chunk suffices.  Entry max_nfs of every row
    // stays zero: the cells' face lists are padded to kNb entries with it, so the gather has
    // neither a face-count nor a sign branch (those conditions, hoisted out of the chunk
    // loop as lane masks, overflowed the scalar registers)
    const int FS = a.max_nfs + 1;
    P2<double> *fl2 = (P2<double> *)smem;
    // the slots' Maxwellian x-factors A exp(-(xi_x - U_x)^2 / 2RT) of the current stripe,
    // [VC][max_nfs]; each thread reads only the entries it wrote (pass B 59.8 -> 57.9 ms)
    double *exs = (double *)(fl2 + 2 * VC * FS);
    if (threadIdx.x < 2 * VC) fl2[threadIdx.x * FS + a.max_nfs] = P2<double>{0.0, 0.0};
    const double dt = *dtp;

    const int my = threadIdx.x % kNTB;
    // the same for every lane of a wave (kNTB is a multiple of 64): the velocity tables are
    // read with scalar loads, whose waits leave the next chunk's vector loads in flight
    const int half = __builtin_amdgcn_readfirstlane(threadIdx.x / kNTB);
    const bool has_cell = my < bi.nc;
    const int cell = bi.c0 + my;
    double mom[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) mom[k] = 0.0;
    double dtV = 0.0;
    // the cell's faces as block slot indices (< 2^15) with the owner flag in bit 15, two per
    // word, padded with the zero entry
    unsigned fpk[kNb / 2];
#pragma unroll
    for (int j = 0; j < kNb / 2; ++j) fpk[j] = (unsigned)a.max_nfs | ((unsigned)a.max_nfs << 16);
    if (has_cell) {
        dtV = dt / g.V[cell];
        const int k0 = a.fbc_beg[bi.fbc_off + my];
        const int nf = a.fbc_beg[bi.fbc_off + my + 1] - k0;
#pragma unroll
        for (int j = 0; j < kNb; ++j) {
            if (j < nf) {
                const unsigned e = (unsigned)a.fbc_idx[k0 + j] | (a.fbc_sgn[k0 + j] > 0 ? 0x8000u : 0u);
                fpk[j / 2] = (fpk[j / 2] & ~(0xffffu << (16 * (j % 2)))) | (e << (16 * (j % 2)));
            }
        }
    }
    int sfc[kSlotsB3], scol[kSlotsB3];  // the slots' faces and face-value columns
    d3 sS[kSlotsB3];
    EqCoef sk[kSlotsB3];
#pragma unroll
    for (int s = 0; s < kSlotsB3; ++s) {
        const int j = threadIdx.x + s * kNTB3;
        sfc[s] = -1;
        scol[s] = 0;
        sS[s] = d3{0, 0, 0};
        sk[s] = EqCoef{};
        if (j < bi.nfs) {
            const int own_slot = bi.fs_off + j;
            sfc[s] = a.fs_face[own_slot];
            scol[s] = a.fs_col[own_slot];
            sS[s] = a.fs_sf[own_slot];
            sk[s] = a.fs_coef[own_slot];
        }
    }

    // the flux cases as compares of sigma = xi . S_f: relaxed face value where sigma > tP or
    // sigma < tM (|sigma| >= VSMALL on internal faces, sigma > 0 on boundary faces), the
    // incoming half where sigma <= tI (boundary faces), zero otherwise (tangent internal faces)
    double tP[kSlotsB3], tM[kSlotsB3], tI[kSlotsB3];
    {
        const double inf = __builtin_huge_val();
        const double vsmall_below = __longlong_as_double(__double_as_longlong(kVSmall) - 1);  // x > it: x >= VSMALL
#pragma unroll
        for (int s = 0; s < kSlotsB3; ++s) {
            const bool bnd = sfc[s] >= g.nIF;
            tP[s] = bnd ? 0.0 : vsmall_below;
            tM[s] = bnd ? -inf : -kVSmall;
            tI[s] = bnd ? 0.0 : -inf;
        }
    }

    // next chunk's stored face values of the thread's slots and u of its cell
    P2<S> fn[kSlotsB3][VC];
    P2<S> un[HV];
    auto load_next = [&](int ch) {
        int st = -1;
        const Chunk c = chunk_of(a, ch, VC, st);
        if (has_cell) {
#pragma unroll
            for (int i = 0; i < HV; ++i)
                un[i] = ld_nt(U + (size_t)(c.vrow + c.ix0 + min(half * HV + i, c.nvv - 1)) * g.nC + cell);
        }
#pragma unroll
        for (int s = 0; s < kSlotsB3; ++s) {
            if (sfc[s] < 0) continue;
#pragma unroll
            for (int vv = 0; vv < VC; ++vv)
                fn[s][vv] = ld_nt(F + (size_t)(c.vrow + c.ix0 + min(vv, c.nvv - 1)) * a.n_cols + scol[s]);
        }
    };
    if (ch0 < ch1) load_next(ch0);
    int stripe = -1;
    for (int ch = ch0; ch < ch1; ++ch) {
        const Chunk ck = chunk_of(a, ch, VC, stripe);
        const d3 xr = cload(g.xi, ck.vrow);
        P2<S> fv[kSlotsB3][VC];
        P2<S> ut[HV];
#pragma unroll
        for (int s = 0; s < kSlotsB3; ++s)
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) fv[s][vv] = fn[s][vv];
#pragma unroll
        for (int i = 0; i < HV; ++i) ut[i] = un[i];
        if (ch + 1 < ch1) load_next(ch + 1);
        P2<double> *fl = fl2 + ((ch - ch0) & 1) * VC * FS;
        // the thread's face slots -> flux sigma * f_f in LDS, each face Maxwellian the slot's
        // stripe x-factor times one row factor per chunk.  Every velocity runs the relaxation
        // and keeps it where sigma > tP or sigma < tM (no branch); the incoming halves of
        // boundary faces (sigma <= tI) are written after the slot's loop
        // the chunk's velocity data and each slot's x-factors read before the relaxation loop:
        // scalar and LDS loads share one counter, so a scalar load inside the loop would also
        // wait for the flux writes before it (pass B 45.0 -> 43.7 ms)
        double vx[VC], vrg[VC], vrh[VC];
#pragma unroll
        for (int vv = 0; vv < VC; ++vv) {
            const int v = ck.vrow + ck.ix0 + min(vv, ck.nvv - 1);
            vx[vv] = cload((const double *)g.xi, 3 * v);
            vrg[vv] = cload(g.ref_g, v);
            vrh[vv] = cload(g.ref_h, v);
        }
#pragma unroll
        for (int s = 0; s < kSlotsB3; ++s) {
            if (sfc[s] < 0) continue;
            const int j = threadIdx.x + s * kNTB3;
            if (ck.new_stripe) stripe_factors<VC>(g, sk[s], ck.ix0, a.n1, exs + j, a.max_nfs);
            double ex[VC];
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) ex[vv] = exs[vv * a.max_nfs + j];
            const double eyz = row_factor(sk[s], xr);
            double spy, spz;
            row_products(xr, sS[s], spy, spz);
            const RowEq re = row_eq(sk[s], xr);
            const EqCoef &k = sk[s];
            bool inc = false;
#pragma unroll
            for (int vv = 0; vv < VC; ++vv) {
                if (vv >= ck.nvv) break;
                const double xs = sigma_row(vx[vv], sS[s].x, spy, spz);
                const bool rel = xs > tP[s] || xs < tM[s];
                inc = inc || xs <= tI[s];
                double gEq, hEq;
                shakhov_row(g, k, re, vx[vv], ex[vv] * eyz, gEq, hEq);
                const double fg = xs * ((1.0 - k.tr) * ((double)fv[s][vv].g + vrg[vv]) + k.tr * gEq);
                const double fh = xs * ((1.0 - k.tr) * ((double)fv[s][vv].h + vrh[vv]) + k.tr * hEq);
                fl[vv * FS + j] = P2<double>{rel ? fg : 0.0, rel ? fh : 0.0};
            }
            if (inc) {  // incoming half of a wall or inlet face
                const int b = sfc[s] - g.nIF;
                const BndCoef kb = a.bc[b];
                const double scale = g.bf_type[b] == kMaxwellWall ? rho_b[b] : 1.0;
                for (int vv = 0; vv < ck.nvv; ++vv) {
                    const d3 xi = cload(g.xi, ck.vrow + ck.ix0 + vv);
                    const double xs = sigma_row(xi.x, sS[s].x, spy, spz);
                    if (!(xs <= tI[s])) continue;
                    const double cx = xi.x - kb.ux, cy = xi.y - kb.uy, cz = xi.z - kb.uz;
                    const double gv = scale * (kb.A * exp_cw(-0.5 * ((cx * cx + cy * cy + cz * cz) * kb.invRT)));
                    fl[vv * FS + j] = P2<double>{xs * gv, xs * (gv * kb.hfac)};
                }
            }
        }
        __syncthreads();
        // cell update of the thread's half of the chunk, in ascending face order, in place
        if (has_cell) {
            RowSums rs{};
#pragma unroll
            for (int i = 0; i < HV; ++i) {
                const int vv = half * HV + i;
                if (vv >= ck.nvv) break;
                const int v = ck.vrow + ck.ix0 + vv;
                double ag = (double)ut[i].g;
                double ah = (double)ut[i].h;
#pragma unroll
                for (int j = 0; j < kNb; ++j) {
                    const unsigned e = (fpk[j / 2] >> (16 * (j % 2))) & 0xffffu;
                    const P2<double> q = fl[vv * FS + (int)(e & 0x7fffu)];
                    // owner faces add, neighbour faces subtract: the sign bit of the scaled
                    // flux is flipped unless bit 15 is set
                    const unsigned long long flip = (unsigned long long)(~e & 0x8000u) << 48;
                    ag += __longlong_as_double(__double_as_longlong(q.g * dtV) ^ flip);
                    ah += __longlong_as_double(__double_as_longlong(q.h * dtV) ^ flip);
                }
                const S sg = (S)ag, sh = (S)ah;
                const size_t kk = (size_t)v * g.nC + cell;
                __builtin_nontemporal_store(sg, gT + kk);
                __builtin_nontemporal_store(sh, hT + kk);
                row_add_w(rs, cload(a.rw, v), (double)sg, (double)sh);
            }
            fold_row<3>(mom, rs, xr.y, xr.z);
        }
    }
    __syncthreads();  // every thread is past its last flux-tile read
    // the second half's moments through LDS, added to the first half's in a fixed order
    double *mx = (double *)smem;  // [NM][kNTB]
    if (half == 1 && has_cell) {
#pragma unroll
        for (int k = 0; k < NM; ++k) mx[k * kNTB + my] = mom[k];
    }
    __syncthreads();
    if (half == 0 && has_cell) {
        double *o = a.cpart + ((size_t)vg * g.nC + cell) * NM;
#pragma unroll
        for (int k = 0; k < NM; ++k) o[k] = mom[k] + mx[k * kNTB + my];
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

// power-sum weights w xi_x^k of every velocity (row_add_w)
__global__ void __launch_bounds__(kBlk) k_row_weights(Geo g, RowW *rw)
{
    const int v = blockIdx.x * kBlk + threadIdx.x;
    if (v >= g.Nv) return;
    const double w = g.w[v], x = g.xi[v].x;
    const double m1 = w * x, m2 = m1 * x;
    rw[v] = RowW{w, m1, m2, m2 * x};
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

// pass-A LDS: f-bar-plus and gradient tiles (what the budget limits), plus the x-factors
// in 3-D
template <class S>
size_t lds_tiles_a(const FusedPlan &p, int D, int VC)
{
    const int ge = sizeof(S) == 4 ? D + 1 : D;
    return (size_t)VC * (2 * p.max_nl + 2 * kFusedThreadsB * ge) * sizeof(S);
}
template <class S>
size_t lds_pass_a(const FusedPlan &p, int D, int VC, bool barin = false)
{
    if (barin) {  // the gradient stencils' table, f-bar-plus pairs (local cells padded to whole waves), gradient tile
        const size_t nlp = (size_t)((p.max_nl + 63) & ~63);
        return (size_t)kNb * kNTB * 16 + (size_t)VC * nlp * 2 * sizeof(S) +
               (size_t)VC * kFusedThreadsB * (D + 1) * 2 * sizeof(S);
    }
    const size_t cex = D == 3 ? (size_t)VC * p.max_nl * sizeof(double) : 0;
    return lds_tiles_a<S>(p, D, VC) + cex;
}
size_t lds_pass_b(const FusedPlan &p, int VC) { return (size_t)VC * 2 * p.max_nfs * sizeof(double); }
// k_bar3: the waves' store patches
template <class S>
size_t lds_bar3(const FusedPlan &, int VC)
{
    return (size_t)kBlk * VC * 2 * sizeof(S);
}
// 3-D pass B: two flux tiles, which the two halves' moment exchange reuses at the end, and
// the slots' x-factors
size_t lds_pass_b3(const FusedPlan &p, int VC)
{
    return std::max((size_t)2 * VC * (p.max_nfs + 1) * 2 * sizeof(double), (size_t)17 * kFusedThreadsB * sizeof(double)) +
           (size_t)VC * p.max_nfs * sizeof(double);
}
inline int vc_for_b(int vc) { return std::min(vc, kFusedVcB); }

constexpr size_t kLdsBudget = 80 * 1024;
// In 3-D the 512-thread pass-A work-group is alone on its CU (2 waves per SIMD), so its
// chunks may use most of the CU's LDS: 8 or 6 velocities per chunk (pass B keeps 4) cut the
// per-chunk barriers and loads (211k-cell sphere pass A, FP32 storage 99.3 -> 93.8 ms with
// 8; FP64 148.0 -> 145.0 ms with 6, the largest whose tiles fit)
constexpr size_t kLdsBudgetA3 = 160 * 1024;

template <class S>
int choose_vc(const FusedPlan &p, int D, bool barin = false)
{
    if (D == 3)
        for (int vc : {8, 6})
            if (lds_pass_a<S>(p, D, vc, barin) <= kLdsBudgetA3 && lds_pass_b(p, vc_for_b(vc)) <= kLdsBudget)
                return vc;
    for (int vc : {4, 2, 1})
        if (lds_tiles_a<S>(p, D, vc) <= kLdsBudget && lds_pass_b(p, vc_for_b(vc)) <= kLdsBudget) return vc;
    return 0;
}

inline int grid_a(const FusedPlan &p) { return kFusedXcds * p.b8 * p.nvg; }
inline int grid_b(const FusedPlan &p) { return kFusedXcds * p.b8 * p.nvg_b; }

template <class S, int D, int VC>
void launch_a(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const double *dt, hipStream_t st)
{
    if constexpr (sizeof(S) == 4 && D == 3) {
        if (p.Fb != nullptr) {
            const size_t lds = lds_pass_a<S>(p, D, VC, true);
            hipLaunchKernelGGL((k_pass_a<S, D, VC, true>), dim3(grid_a(p)), dim3(kNTA), lds, st, g, fargs(p), gT, hT,
                               (P2<S> *)p.F, (const P2<S> *)p.Fb, dt);
            return;
        }
    }
    const size_t lds = lds_pass_a<S>(p, D, VC);
    hipLaunchKernelGGL((k_pass_a<S, D, VC, false>), dim3(grid_a(p)), dim3(kNTA), lds, st, g, fargs(p), gT, hT,
                       (P2<S> *)p.F, (const P2<S> *)nullptr, dt);
}

template <class S, int D, int VC>
void launch_b(const Geo &g, FusedPlan &p, S *gT, S *hT, const double *rho_b, const double *dt, hipStream_t st)
{
    const size_t lds = lds_pass_b(p, VC);
    hipLaunchKernelGGL((k_pass_b<S, D, VC>), dim3(grid_b(p)), dim3(kNTB), lds, st, g, fargs_b(p), gT, hT,
                       (const P2<S> *)p.F, rho_b, dt);
}

template <class S, int VC>
void launch_b3(const Geo &g, FusedPlan &p, S *gT, S *hT, const double *rho_b, const double *dt, hipStream_t st)
{
    const size_t lds = lds_pass_b3(p, VC);
    hipLaunchKernelGGL((k_pass_b3<S, VC>), dim3(grid_b(p)), dim3(kNTB3), lds, st, g, fargs_b(p), gT, hT,
                       (const P2<S> *)p.F, (const P2<S> *)p.U, rho_b, dt);
}

template <class S>
bool set_lds_attrs(int D, int vc, const FusedPlan &p)
{
    auto set = [](const void *fn, size_t bytes) {
        return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
    };
    bool ok = true;
#define DUGKS_SET(DD, VV)                                                                              \
    if (D == DD && vc == VV) ok &= set((const void *)k_pass_a<S, DD, VV, false>, lds_pass_a<S>(p, DD, VV));   \
    if (D == DD && vc_for_b(vc) == VV) ok &= set((const void *)k_pass_b<S, DD, VV>, lds_pass_b(p, VV));
    DUGKS_SET(2, 1) DUGKS_SET(2, 2) DUGKS_SET(2, 4) DUGKS_SET(3, 1) DUGKS_SET(3, 2) DUGKS_SET(3, 4)
#undef DUGKS_SET
    // 3-D pass A alone takes 6 or 8 velocities per chunk (pass B keeps 4)
    if (D == 3 && vc == 6) ok &= set((const void *)k_pass_a<S, 3, 6, false>, lds_pass_a<S>(p, 3, 6));
    if (D == 3 && vc == 8) ok &= set((const void *)k_pass_a<S, 3, 8, false>, lds_pass_a<S>(p, 3, 8));
    if constexpr (sizeof(S) == 4) {
        if (D == 3 && vc == 6) ok &= set((const void *)k_pass_a<S, 3, 6, true>, lds_pass_a<S>(p, 3, 6, true));
        if (D == 3 && vc == 8) ok &= set((const void *)k_pass_a<S, 3, 8, true>, lds_pass_a<S>(p, 3, 8, true));
        if (D == 3 && vc == 6) ok &= set((const void *)k_bar3<S, 6>, lds_bar3<S>(p, 6));
        if (D == 3 && vc == 8) ok &= set((const void *)k_bar3<S, 8>, lds_bar3<S>(p, 8));
    }
    if (D == 3 && vc_for_b(vc) == 4) ok &= set((const void *)k_pass_b3<S, 4>, lds_pass_b3(p, 4));
    return ok;
}

inline dim3 grid1(int n) { return dim3((n + kBlk - 1) / kBlk); }

__global__ void __launch_bounds__(kBlk) k_zero_d(double *p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * kBlk + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

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
    p.vc_b = vc_for_b(vc);
    p.nchunk_b = ((n1 + p.vc_b - 1) / p.vc_b) * p.nrow;
    const int want_b = std::max(1, std::min(p.nchunk_b, (p.wg_target + p.nblk - 1) / p.nblk));
    p.cpg_b = (p.nchunk_b + want_b - 1) / want_b;
    p.nvg_b = (p.nchunk_b + p.cpg_b - 1) / p.cpg_b;
    p.b8 = (p.nblk + kFusedXcds - 1) / kFusedXcds;
}

bool fused_plan_upload(FusedPlan &p, const Geo &g, size_t es, std::string &err)
{
    // 3-D with FP32 storage: f-bar-plus and u of every cell from k_bar3 (Fb, U), pass B in
    // the 512-thread form (k_pass_b3)
    bool fbar = !p.two_pass && es == 4 && g.D == 3 && p.max_nfs <= kSlotsB3 * kNTB3 &&
                lds_pass_b3(p, 4) <= kLdsBudgetA3;
    int vc = es == 4 ? choose_vc<float>(p, g.D, fbar) : choose_vc<double>(p, g.D);
    if (fbar && (vc_for_b(vc) != 4 || (vc != 6 && vc != 8))) {  // the k_bar3 / BARIN instances
        fbar = false;
        vc = choose_vc<float>(p, g.D);
    }
    if (vc == 0) {
        err = "block tables exceed the LDS budget";
        return false;
    }
    if (g.n1d < 1 || g.Nv != (g.D == 3 ? g.n1d * g.n1d * g.n1d : g.n1d * g.n1d)) {
        err = "fused path needs a tensor-product velocity set";
        return false;
    }
    fused_velocity_groups(p, g.D, g.n1d, vc);
    // the face values, then the k_bar3 form's two more [Nv][nC] pair arrays: without the
    // memory for those the two-pass form runs (it needs 2 Nv nC es bytes less than that)
    if (!dbuf(p, &p.F, (size_t)g.Nv * p.n_cols * 2 * es, err)) return false;
    p.U = p.Fb = nullptr;
    if (fbar) {
        void *u = nullptr, *fb = nullptr;
        if (hipMalloc(&u, (size_t)g.Nv * g.nC * 2 * es) == hipSuccess &&
            hipMalloc(&fb, (size_t)p.nchunk * vc * g.nC * 2 * es) == hipSuccess) {
            p.allocs.push_back(u);
            p.allocs.push_back(fb);
            p.U = u;
            p.Fb = fb;
        } else {
            if (u) (void)hipFree(u);
            (void)hipGetLastError();
            fbar = false;
            vc = choose_vc<float>(p, g.D);
            if (vc == 0) {
                err = "block tables exceed the LDS budget";
                return false;
            }
            fused_velocity_groups(p, g.D, g.n1d, vc);
        }
    }
    if (!(es == 4 ? set_lds_attrs<float>(g.D, vc, p) : set_lds_attrs<double>(g.D, vc, p))) {
        err = "cannot raise the dynamic LDS limit";
        return false;
    }
    if (!up(p, &p.d_blk, p.blk, err) || !up(p, &p.d_lcell, p.lcell, err) || !up(p, &p.d_gs_beg, p.gs_beg, err) ||
        !up(p, &p.d_gs_nb, p.gs_nb, err) || !up(p, &p.d_fs_face, p.fs_face, err) ||
        !up(p, &p.d_fs_col, p.fs_col, err) || !up(p, &p.d_fs_lo, p.fs_lo, err) || !up(p, &p.d_fs_ln, p.fs_ln, err) ||
        !up(p, &p.d_fs_flag, p.fs_flag, err) || !up(p, &p.d_fbc_beg, p.fbc_beg, err) ||
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
    const size_t nfpart = (size_t)p.nvg * g.nF * 2 * p.nma;
    if (!dbuf(p, &fp, nfpart * 8, err) || !dbuf(p, &cp, (size_t)p.nvg_b * g.nC * p.nm * 8, err) ||
        !dbuf(p, &bi, (size_t)g.nB * p.nm * 8, err) || !dbuf(p, &br, (size_t)g.nB * p.nma * 8, err) ||
        !dbuf(p, &cr, (size_t)p.nma * 8, err) || !dbuf(p, &cc, (size_t)g.nC * sizeof(EqCoef), err) ||
        !dbuf(p, &fc, (size_t)g.nF * sizeof(EqCoef), err) || !dbuf(p, &bc, (size_t)g.nB * sizeof(BndCoef), err))
        return false;
    p.fpart = (double *)fp;
    p.cpart = (double *)cp;
    p.bin = (double *)bi;
    p.bref = (double *)br;
    p.cref = (double *)cr;
    p.cc = (EqCoef *)cc;
    p.fc = (EqCoef *)fc;
    void *sf = nullptr, *sc = nullptr;
    if (!dbuf(p, &sf, (size_t)p.n_slots * sizeof(d3), err) || !dbuf(p, &sc, (size_t)p.n_slots * sizeof(EqCoef), err))
        return false;
    p.fs_sf = (d3 *)sf;
    p.fs_coef = (EqCoef *)sc;
    if (!dbuf(p, &p.rw, (size_t)g.Nv * sizeof(RowW), err)) return false;
    hipLaunchKernelGGL(k_row_weights, dim3((unsigned)((g.Nv + kBlk - 1) / kBlk)), dim3(kBlk), 0, 0, g, (RowW *)p.rw);
    p.bc = (BndCoef *)bc;
    // (face, side) slots no block writes (the second side of a face inside one block) stay 0
    hipLaunchKernelGGL(k_zero_d, dim3((unsigned)((nfpart + kBlk - 1) / kBlk)), dim3(kBlk), 0, 0, p.fpart, nfpart);
    hipLaunchKernelGGL(k_slot_sf, dim3((unsigned)((p.n_slots + kBlk - 1) / kBlk)), dim3(kBlk), 0, 0, g, fargs(p),
                       p.n_slots);
    if (hipDeviceSynchronize() != hipSuccess) {
        err = "clearing the face partials failed";
        return false;
    }
    p.lds_a = es == 4 ? lds_pass_a<float>(p, g.D, vc, fbar) : lds_pass_a<double>(p, g.D, vc);
    p.lds_b = lds_pass_b(p, p.vc_b);
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
void Fused<S>::face_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const double *dt, hipStream_t st)
{
#define DUGKS_A(DD, VV) \
    if (g.D == DD && p.vc == VV) return launch_a<S, DD, VV>(g, p, gT, hT, dt, st);
    DUGKS_A(2, 1) DUGKS_A(2, 2) DUGKS_A(2, 4) DUGKS_A(3, 1) DUGKS_A(3, 2) DUGKS_A(3, 4) DUGKS_A(3, 6) DUGKS_A(3, 8)
#undef DUGKS_A
}

template <class S>
void Fused<S>::bar_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const double *dt, hipStream_t st)
{
    if constexpr (sizeof(S) == 4) {
        if (p.Fb == nullptr) return;
        const dim3 grid((g.nC + kBlk - 1) / kBlk, kBarRowGroups);
        const size_t lds = lds_bar3<S>(p, p.vc);
        if (p.vc == 8)
            hipLaunchKernelGGL((k_bar3<S, 8>), grid, dim3(kBlk), lds, st, g, p.cc, p.n1, gT, hT, (P2<S> *)p.Fb,
                               (P2<S> *)p.U, dt);
        else if (p.vc == 6)
            hipLaunchKernelGGL((k_bar3<S, 6>), grid, dim3(kBlk), lds, st, g, p.cc, p.n1, gT, hT, (P2<S> *)p.Fb,
                               (P2<S> *)p.U, dt);
    }
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
    // pass B reads the records per face slot
    hipLaunchKernelGGL(k_slot_coef, grid1(p.n_slots), dim3(kBlk), 0, st, fargs(p), p.n_slots);
    launch_cfl_partials(g, fm, cfl_part, st);
}

template <class S>
void Fused<S>::update_pass(const Geo &g, FusedPlan &p, S *gT, S *hT, const double *rho_b, const double *dt,
                           hipStream_t st)
{
    if (p.U != nullptr) return launch_b3<S, 4>(g, p, gT, hT, rho_b, dt, st);
#define DUGKS_B(DD, VV) \
    if (g.D == DD && p.vc_b == VV) return launch_b<S, DD, VV>(g, p, gT, hT, rho_b, dt, st);
    DUGKS_B(2, 1) DUGKS_B(2, 2) DUGKS_B(2, 4) DUGKS_B(3, 1) DUGKS_B(3, 2) DUGKS_B(3, 4)
#undef DUGKS_B
}

template <class S>
void Fused<S>::cell_macros(const Geo &g, FusedPlan &p, Macros cm, const double *dt, hipStream_t st)
{
    if (g.D == 3)
        hipLaunchKernelGGL(k_cell_macros<3>, grid1(g.nC), dim3(kBlk), 0, st, g, fargs_b(p), cm, dt);
    else
        hipLaunchKernelGGL(k_cell_macros<2>, grid1(g.nC), dim3(kBlk), 0, st, g, fargs_b(p), cm, dt);
}

template <class S>
void Fused<S>::cell_coef(const Geo &g, FusedPlan &p, const Macros &cm, hipStream_t st)
{
    hipLaunchKernelGGL(k_cell_coef, grid1(g.nC), dim3(kBlk), 0, st, g, cm, p.cc);
}

template struct Fused<float>;
template struct Fused<double>;

}  // namespace dugks
