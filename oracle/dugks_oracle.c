/*
 * FP64 CPU oracle for the DUGKS time step -- TEST INFRASTRUCTURE ONLY (see header).
 *
 * Each function restates one member of Foam::fvDVM (src/fvDVM/fvDVM/fvDVM.C, "F" below)
 * or of the OpenFOAM library code it calls, keeping the reference's loop structure, its
 * expression association and its summation order where that order is defined:
 *   - velocity moments are summed in ascending velocity order (F:1201-1297, F:1536-1546;
 *     the TBB reductions run on one thread, F:41, F:2107);
 *   - the flux scatter visits internal faces, then boundary faces, in face order
 *     (F:1489-1517).
 * Velocity loops run under OpenMP where the reference has `omp parallel for` over dvi
 * (F:659, F:764, F:1374, F:1464); per-velocity results do not depend on thread count.
 */
#include "dugks_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846
#define ORC_VSMALL 1.0e-300
#define ORC_SMALL 1.0e-15

static char g_err[512];
static void set_err(const char *m) { snprintf(g_err, sizeof g_err, "%s", m); }
const char *orc_last_error(void) { return g_err; }

int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

struct orc_state {
    orc_case cs;
    int Nv, nF;
    double *xi, *w;                 /* [Nv][3], [Nv] */
    double *gT, *hT, *gB, *hB;      /* [Nv][nC] */
    double *gS, *hS;                /* [Nv][nF] */
    double *rho, *U, *T, *q, *tau;  /* cell macros */
    double *rhoS, *US, *TS, *qS, *tauS; /* face macros [nF] */
    double *rho_b;                  /* boundary rho patch values [nB] */
    double *in_by_rho, *out_going;  /* [nB] */
    double *own_ls, *nei_ls, *patch_ls;
    int *bf_type, *bf_local;        /* [nB] */
    double *bf_p;                   /* [nB] pressure of pressureIn / pressureOut faces */
    int *bf_mirror;                 /* [nB] symmetry faces: patch axis (0, 1, 2), else -1 */
    double *bf_sf0;                 /* [nB][3] symmetry faces: Sf of the patch's first face */
    int *mirror;                    /* [3][Nv] velocity mirrored about each axis (symmetric sets) */
    /* owned copies of the case arrays */
    int *owner, *neighbour, *ptype, *pstart, *psize;
    double *Sf, *Cf, *C, *V, *wts_if, *dc, *bf_rho, *bf_U, *bf_T, *xis, *wts;
};

static void *dup(const void *p, size_t n)
{
    void *d = malloc(n ? n : 1);
    if (d && n) memcpy(d, p, n);
    return d;
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double magsqr3(const double *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

/* ---------------------------------------------------------------------------------
 * Least-squares vectors: OpenFOAM leastSquaresVectors as recorded in-tree by
 * src/fvDVM/zeroBoundaryGrad/zeroBoundaryVectors.C:115-122 (internal rows weighted by
 * magSf/|d|^2 split with the interpolation weights), with the non-coupled boundary rows
 * of the standard scheme restored (zeroBoundaryVectors.C:156-165, commented out there):
 * dd += magSf/|d|^2 d d, d = patch delta = n (n.(Cf - C)).  The inverse uses OpenFOAM's
 * inv(symmTensorField), which adds 1 to diagonal components that are negligible in the
 * first cell's tensor (2-D and 1-D meshes), inverts, and subtracts the 1 again.
 * ------------------------------------------------------------------------------- */
typedef struct { double xx, xy, xz, yy, yz, zz; } symm;

static symm symm_inv(symm s)
{
    double det = s.xx * s.yy * s.zz + s.xy * s.yz * s.xz + s.xz * s.xy * s.yz
               - s.xx * s.yz * s.yz - s.xy * s.xy * s.zz - s.xz * s.yy * s.xz;
    symm r;
    r.xx = (s.yy * s.zz - s.yz * s.yz) / det;
    r.xy = (s.xz * s.yz - s.xy * s.zz) / det;
    r.xz = (s.xy * s.yz - s.xz * s.yy) / det;
    r.yy = (s.xx * s.zz - s.xz * s.xz) / det;
    r.yz = (s.xy * s.xz - s.xx * s.yz) / det;
    r.zz = (s.xx * s.yy - s.xy * s.xy) / det;
    return r;
}

static void symm_add_dd(symm *a, double f, const double *d)
{
    a->xx += f * (d[0] * d[0]); a->xy += f * (d[0] * d[1]); a->xz += f * (d[0] * d[2]);
    a->yy += f * (d[1] * d[1]); a->yz += f * (d[1] * d[2]); a->zz += f * (d[2] * d[2]);
}

static void symm_dot(const symm *a, const double *d, double *out)
{
    out[0] = a->xx * d[0] + a->xy * d[1] + a->xz * d[2];
    out[1] = a->xy * d[0] + a->yy * d[1] + a->yz * d[2];
    out[2] = a->xz * d[0] + a->yz * d[1] + a->zz * d[2];
}

static void boundary_delta(const orc_state *s, int bf, double *d)
{
    const int f = s->cs.nIF + bf;
    const double *Sf = s->Sf + 3 * f;
    double mag = sqrt(magsqr3(Sf));
    double n[3] = {Sf[0] / mag, Sf[1] / mag, Sf[2] / mag};
    const double *c = s->C + 3 * s->owner[f];
    double r[3] = {s->Cf[3 * f] - c[0], s->Cf[3 * f + 1] - c[1], s->Cf[3 * f + 2] - c[2]};
    double nd = dot3(n, r);
    d[0] = n[0] * nd; d[1] = n[1] * nd; d[2] = n[2] * nd;
}

static void compute_ls_vectors(orc_state *s)
{
    const int nC = s->cs.nC, nIF = s->cs.nIF, nB = s->cs.nB;
    symm *dd = calloc((size_t)nC, sizeof(symm));
    for (int f = 0; f < nIF; ++f) {
        int own = s->owner[f], nei = s->neighbour[f];
        double d[3] = {s->C[3 * nei] - s->C[3 * own], s->C[3 * nei + 1] - s->C[3 * own + 1],
                       s->C[3 * nei + 2] - s->C[3 * own + 2]};
        double magSf = sqrt(magsqr3(s->Sf + 3 * f));
        double fac = magSf / magsqr3(d);
        symm wdd = {0, 0, 0, 0, 0, 0};
        symm_add_dd(&wdd, fac, d);
        double w = s->wts_if[f];
        dd[own].xx += (1 - w) * wdd.xx; dd[own].xy += (1 - w) * wdd.xy; dd[own].xz += (1 - w) * wdd.xz;
        dd[own].yy += (1 - w) * wdd.yy; dd[own].yz += (1 - w) * wdd.yz; dd[own].zz += (1 - w) * wdd.zz;
        dd[nei].xx += w * wdd.xx; dd[nei].xy += w * wdd.xy; dd[nei].xz += w * wdd.xz;
        dd[nei].yy += w * wdd.yy; dd[nei].yz += w * wdd.yz; dd[nei].zz += w * wdd.zz;
    }
    for (int b = 0; b < nB; ++b) {
        double d[3];
        boundary_delta(s, b, d);
        double magSf = sqrt(magsqr3(s->Sf + 3 * (nIF + b)));
        symm_add_dd(&dd[s->owner[nIF + b]], magSf / magsqr3(d), d);
    }
    /* inv(symmTensorField) with removal of negligible directions (first tensor decides) */
    int rm[3] = {0, 0, 0};
    if (nC > 0) {
        symm t = dd[0];
        double scale = t.xx * t.xx + 2 * t.xy * t.xy + 2 * t.xz * t.xz + t.yy * t.yy
                     + 2 * t.yz * t.yz + t.zz * t.zz;
        rm[0] = (t.xx * t.xx) / scale < ORC_SMALL;
        rm[1] = (t.yy * t.yy) / scale < ORC_SMALL;
        rm[2] = (t.zz * t.zz) / scale < ORC_SMALL;
    }
    for (int c = 0; c < nC; ++c) {
        symm t = dd[c];
        if (rm[0]) t.xx += 1; if (rm[1]) t.yy += 1; if (rm[2]) t.zz += 1;
        t = symm_inv(t);
        if (rm[0]) t.xx -= 1; if (rm[1]) t.yy -= 1; if (rm[2]) t.zz -= 1;
        dd[c] = t;
    }
    for (int f = 0; f < nIF; ++f) {
        int own = s->owner[f], nei = s->neighbour[f];
        double d[3] = {s->C[3 * nei] - s->C[3 * own], s->C[3 * nei + 1] - s->C[3 * own + 1],
                       s->C[3 * nei + 2] - s->C[3 * own + 2]};
        double magSf = sqrt(magsqr3(s->Sf + 3 * f));
        double fac = magSf / magsqr3(d);
        double w = s->wts_if[f], a[3], b[3];
        symm_dot(&dd[own], d, a);
        symm_dot(&dd[nei], d, b);
        for (int k = 0; k < 3; ++k) {
            s->own_ls[3 * f + k] = (1 - w) * fac * a[k];
            s->nei_ls[3 * f + k] = -w * fac * b[k];
        }
    }
    for (int b = 0; b < nB; ++b) {
        double d[3], a[3];
        boundary_delta(s, b, d);
        double magSf = sqrt(magsqr3(s->Sf + 3 * (nIF + b)));
        symm_dot(&dd[s->owner[nIF + b]], d, a);
        for (int k = 0; k < 3; ++k) s->patch_ls[3 * b + k] = magSf * (1.0 / magsqr3(d)) * a[k];
    }
    free(dd);
}

/* ------------------------------------------------------------------------- create */
orc_state *orc_create(const orc_case *cs)
{
    if (!cs || cs->nC <= 0 || cs->nIF < 0 || cs->nB < 0 || cs->D < 1 || cs->D > 3 ||
        (cs->n_list > 0 ? !cs->list_xi || !cs->list_w : cs->n1d <= 0)) {
        set_err("orc_create: invalid sizes");
        return NULL;
    }
    if (cs->macro_flux < 0 || cs->macro_flux > 2) {
        set_err("orc_create: macro_flux must be 0, 1 or 2");
        return NULL;
    }
    orc_state *s = calloc(1, sizeof *s);
    s->cs = *cs;
    const int nC = cs->nC, nIF = cs->nIF, nB = cs->nB, nF = nIF + nB;
    s->nF = nF;
    s->owner = dup(cs->owner, sizeof(int) * nF);
    s->neighbour = dup(cs->neighbour, sizeof(int) * nIF);
    s->Sf = dup(cs->Sf, sizeof(double) * 3 * nF);
    s->Cf = dup(cs->Cf, sizeof(double) * 3 * nF);
    s->C = dup(cs->C, sizeof(double) * 3 * nC);
    s->V = dup(cs->V, sizeof(double) * nC);
    s->wts_if = dup(cs->weights, sizeof(double) * nIF);
    s->dc = dup(cs->delta_coeffs, sizeof(double) * nIF);
    s->ptype = dup(cs->patch_type, sizeof(int) * cs->n_patches);
    s->pstart = dup(cs->patch_start, sizeof(int) * cs->n_patches);
    s->psize = dup(cs->patch_size, sizeof(int) * cs->n_patches);
    s->bf_rho = dup(cs->bf_rho, sizeof(double) * nB);
    s->bf_U = dup(cs->bf_U, sizeof(double) * 3 * nB);
    s->bf_T = dup(cs->bf_T, sizeof(double) * nB);
    s->xis = cs->n_list > 0 ? NULL : dup(cs->xis, sizeof(double) * cs->n1d);
    s->wts = cs->n_list > 0 ? NULL : dup(cs->wts, sizeof(double) * cs->n1d);

    s->bf_type = calloc((size_t)nB + 1, sizeof(int));
    s->bf_local = calloc((size_t)nB + 1, sizeof(int));
    s->bf_p = calloc((size_t)nB + 1, sizeof(double));
    s->bf_mirror = calloc((size_t)nB + 1, sizeof(int));
    s->bf_sf0 = calloc(3 * (size_t)nB + 3, sizeof(double));
    for (int b = 0; b < nB; ++b) {
        s->bf_type[b] = -1;
        s->bf_mirror[b] = -1;
    }
    int any_sym = 0;
    for (int p = 0; p < cs->n_patches; ++p) {
        int t = s->ptype[p];
        if (t < ORC_PATCH_MAXWELL_WALL || t > ORC_PATCH_SYMMETRY_PLANE) {
            set_err("orc_create: unsupported patch type");
            orc_destroy(s);
            return NULL;
        }
        if (s->pstart[p] < 0 || s->pstart[p] + s->psize[p] > nB) {
            set_err("orc_create: patch range outside boundary faces");
            orc_destroy(s);
            return NULL;
        }
        if ((cs->q1_compat || cs->macro_flux == 1) && s->psize[p] > nIF) {
            set_err("orc_create: patch larger than the internal-face count (Q1 read, macro-flux Sf read)");
            orc_destroy(s);
            return NULL;
        }
        int axis = -1;
        if ((t == ORC_PATCH_SYMMETRY || t == ORC_PATCH_SYMMETRY_PLANE) && s->psize[p] > 0) {
            /* the mirror velocity is taken about the normal of the patch's first face, which
             * must be a coordinate axis (discreteVelocity.C:731-737, 760-764) */
            const double *S0 = s->Sf + 3 * (nIF + s->pstart[p]);
            const double m = sqrt(magsqr3(S0));
            for (int k = 0; k < 3; ++k)
                if (fabs(fabs(S0[k]) / m - 1.0) < 1e-12) axis = k;
            if (axis < 0) {
                set_err("orc_create: a symmetry patch must be normal to a coordinate axis");
                orc_destroy(s);
                return NULL;
            }
            any_sym = 1;
        }
        for (int i = 0; i < s->psize[p]; ++i) {
            const int b = s->pstart[p] + i;
            s->bf_type[b] = t;
            s->bf_local[b] = i;
            s->bf_p[b] = cs->patch_pressure ? cs->patch_pressure[p] : 0.0;
            s->bf_mirror[b] = axis;
            if (axis >= 0)
                for (int k = 0; k < 3; ++k) s->bf_sf0[3 * b + k] = s->Sf[3 * (nIF + s->pstart[p]) + k];
        }
    }
    for (int b = 0; b < nB; ++b)
        if (s->bf_type[b] < 0) {
            set_err("orc_create: boundary face not covered by any patch");
            orc_destroy(s);
            return NULL;
        }

    /* tensor-product velocity set, ix fastest (F:293-373), or the explicit list of the
     * VMESH build: xi = cell centre with the components beyond D set to 0, weight = cell
     * volume (F:170-249) */
    const int n = cs->n1d, D = cs->D;
    int Nv = cs->n_list > 0 ? cs->n_list : ((D == 3) ? n * n * n : (D == 2 ? n * n : n));
    s->Nv = Nv;
    s->xi = calloc((size_t)Nv * 3, sizeof(double));
    s->w = calloc((size_t)Nv, sizeof(double));
    int i = 0;
    if (cs->n_list > 0) {
        if (any_sym) {
            set_err("orc_create: symmetry patches need a tensor-product velocity set "
                    "(the VMESH build sets no mirror map)");
            orc_destroy(s);
            return NULL;
        }
        for (int v = 0; v < Nv; ++v) {
            s->w[v] = cs->list_w[v];
            for (int k = 0; k < D; ++k) s->xi[3 * v + k] = cs->list_xi[3 * v + k];
        }
    } else if (D == 3) {
        for (int iz = 0; iz < n; ++iz)
            for (int iy = 0; iy < n; ++iy)
                for (int ix = 0; ix < n; ++ix, ++i) {
                    s->w[i] = s->wts[iz] * s->wts[iy] * s->wts[ix];
                    s->xi[3 * i] = s->xis[ix]; s->xi[3 * i + 1] = s->xis[iy]; s->xi[3 * i + 2] = s->xis[iz];
                }
    } else if (D == 2) {
        for (int iy = 0; iy < n; ++iy)
            for (int ix = 0; ix < n; ++ix, ++i) {
                s->w[i] = s->wts[iy] * s->wts[ix] * 1;
                s->xi[3 * i] = s->xis[ix]; s->xi[3 * i + 1] = s->xis[iy]; s->xi[3 * i + 2] = 0.0;
            }
    } else {
        for (int ix = 0; ix < n; ++ix, ++i) {
            s->w[i] = s->wts[ix] * 1 * 1;
            s->xi[3 * i] = s->xis[ix]; s->xi[3 * i + 1] = 0.0; s->xi[3 * i + 2] = 0.0;
        }
    }

    /* mirror velocities of a symmetric 1-D set (x[n-1-i] = -x[i], equal weights) */
    s->mirror = calloc(3 * (size_t)Nv, sizeof(int));
    if (any_sym && cs->n_list <= 0) {
        for (int k = 0; k < n; ++k)
            if (fabs(s->xis[n - 1 - k] + s->xis[k]) > 1e-12 * fabs(s->xis[0]) || s->wts[n - 1 - k] != s->wts[k]) {
                set_err("orc_create: symmetry patches need a velocity set symmetric about 0");
                orc_destroy(s);
                return NULL;
            }
    }
    for (int v = 0; v < Nv && cs->n_list <= 0; ++v) {
        const int ix = v % n, iy = (v / n) % n, iz = v / (n * n);
        s->mirror[v] = (n - 1 - ix) + n * (iy + n * iz);
        s->mirror[Nv + v] = D >= 2 ? ix + n * ((n - 1 - iy) + n * iz) : v;
        s->mirror[2 * Nv + v] = D == 3 ? ix + n * (iy + n * (n - 1 - iz)) : v;
    }

    size_t nvc = (size_t)Nv * nC, nvf = (size_t)Nv * nF;
    s->gT = calloc(nvc, sizeof(double)); s->hT = calloc(nvc, sizeof(double));
    s->gB = calloc(nvc, sizeof(double)); s->hB = calloc(nvc, sizeof(double));
    s->gS = calloc(nvf, sizeof(double)); s->hS = calloc(nvf, sizeof(double));
    s->rho = calloc(nC, sizeof(double)); s->U = calloc(3 * (size_t)nC, sizeof(double));
    s->T = calloc(nC, sizeof(double)); s->q = calloc(3 * (size_t)nC, sizeof(double));
    s->tau = calloc(nC, sizeof(double));
    s->rhoS = calloc(nF + 1, sizeof(double)); s->US = calloc(3 * (size_t)nF + 1, sizeof(double));
    s->TS = calloc(nF + 1, sizeof(double)); s->qS = calloc(3 * (size_t)nF + 1, sizeof(double));
    s->tauS = calloc(nF + 1, sizeof(double));
    s->rho_b = dup(s->bf_rho, sizeof(double) * nB);
    s->in_by_rho = calloc(nB + 1, sizeof(double));
    s->out_going = calloc(nB + 1, sizeof(double));
    s->own_ls = calloc(3 * (size_t)nIF + 1, sizeof(double));
    s->nei_ls = calloc(3 * (size_t)nIF + 1, sizeof(double));
    s->patch_ls = calloc(3 * (size_t)nB + 1, sizeof(double));
    if (!s->gT || !s->hT || !s->gB || !s->hB || !s->gS || !s->hS) {
        set_err("orc_create: out of memory");
        orc_destroy(s);
        return NULL;
    }
    compute_ls_vectors(s);
    if (cs->ls_own && cs->ls_nei) {  /* another gradient scheme in the same form (Gauss linear) */
        memcpy(s->own_ls, cs->ls_own, sizeof(double) * 3 * (size_t)nIF);
        memcpy(s->nei_ls, cs->ls_nei, sizeof(double) * 3 * (size_t)nIF);
    }
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    void *p[] = {s->xi, s->w, s->gT, s->hT, s->gB, s->hB, s->gS, s->hS, s->rho, s->U, s->T, s->q,
                 s->tau, s->rhoS, s->US, s->TS, s->qS, s->tauS, s->rho_b, s->in_by_rho,
                 s->out_going, s->own_ls, s->nei_ls, s->patch_ls, s->bf_type, s->bf_local,
                 s->bf_p, s->bf_mirror, s->bf_sf0, s->mirror,
                 s->owner, s->neighbour, s->ptype, s->pstart, s->psize, s->Sf, s->Cf, s->C, s->V,
                 s->wts_if, s->dc, s->bf_rho, s->bf_U, s->bf_T, s->xis, s->wts};
    for (size_t i = 0; i < sizeof p / sizeof p[0]; ++i) free(p[i]);
    free(s);
}

/* updateTau (F:1823-1832): tau = muRef*exp(omega*log(T/Tref))/rho/T/R */
static double tau_of(const orc_case *cs, double T, double rho)
{
    return cs->muRef * exp(cs->omega * log(T / cs->Tref)) / rho / T / cs->R;
}

/* discreteVelocity::equilibriumMaxwellByRho (discreteVelocity.C:1027-1039) */
static double maxwell_by_rho(const orc_state *s, const double *xi, const double *U, double T)
{
    double R = s->cs.R;
    double c[3] = {U[0] - xi[0], U[1] - xi[1], U[2] - xi[2]};
    return 1.0 / pow(sqrt(2.0 * ORC_PI * R * T), (double)s->cs.D) * exp(-magsqr3(c) / (2.0 * R * T));
}

/* updatePressureInOutBC (fvDVM.C:1745-1821): patch rho (and T at outlets) from the cell
 * state by the characteristic relations; the patch U the reference also sets there is not
 * read by the DF boundary (farField incoming takes the owner cell's U, fvDVM.C:927-933) */
static void update_pressure_bc(orc_state *s)
{
    const orc_case *cs = &s->cs;
    const double R = cs->R;
    for (int b = 0; b < cs->nB; ++b) {
        const int t = s->bf_type[b];
        if (t == ORC_PATCH_PRESSURE_IN) {
            s->rho_b[b] = s->bf_p[b] / R / s->bf_T[b];
        } else if (t == ORC_PATCH_PRESSURE_OUT) {
            const int own = s->owner[cs->nIF + b];
            const double Ti = s->T[own], rhoi = s->rho[own];
            const double ai = sqrt(R * Ti * (cs->K + 5) / (cs->K + 3));
            s->rho_b[b] = rhoi + (s->bf_p[b] - rhoi * R * Ti) / ai / ai;
            s->bf_T[b] = s->bf_p[b] / (R * rhoi);
        }
    }
}

/* ------------------------------------------------------------------------- init */
int orc_init(orc_state *s, const double *rho, const double *U, const double *T)
{
    const orc_case *cs = &s->cs;
    const int nC = cs->nC, nIF = cs->nIF, nB = cs->nB, nF = s->nF, Nv = s->Nv, D = cs->D;
    const double R = cs->R;
    memcpy(s->rho, rho, sizeof(double) * nC);
    memcpy(s->U, U, sizeof(double) * 3 * nC);
    memcpy(s->T, T, sizeof(double) * nC);
    memset(s->q, 0, sizeof(double) * 3 * nC);   /* Q5: initial q = 0 (F:2042-2054) */
    /* initial BGK Maxwellians (F:414-426) */
#pragma omp parallel for schedule(static)
    for (int v = 0; v < Nv; ++v) {
        const double *xi = s->xi + 3 * v;
        for (int c = 0; c < nC; ++c) {
            double d[3] = {s->U[3 * c] - xi[0], s->U[3 * c + 1] - xi[1], s->U[3 * c + 2] - xi[2]};
            double cSqrByRT = magsqr3(d) / (R * s->T[c]);
            double gEq = s->rho[c] / pow(sqrt(2.0 * ORC_PI * R * s->T[c]), (double)D) * exp(-cSqrByRT / 2.0);
            s->gT[(size_t)v * nC + c] = gEq;
            s->hT[(size_t)v * nC + c] = (cs->K + 3.0 - D) * gEq * R * s->T[c];
        }
    }
    /* setCalculatedMaxwellRhoBC (F:578-624): inComingByRho from initial wall U, T (Q4) */
    for (int b = 0; b < nB; ++b) {
        s->in_by_rho[b] = 0.0;
        s->out_going[b] = 0.0;
        if (s->bf_type[b] != ORC_PATCH_MAXWELL_WALL) continue;
        const double *Sf = s->Sf + 3 * (nIF + b);
        for (int v = 0; v < Nv; ++v) {
            double xs = dot3(s->xi + 3 * v, Sf);
            if (xs < 0)
                s->in_by_rho[b] += -s->w[v] * xs * maxwell_by_rho(s, s->xi + 3 * v, s->bf_U + 3 * b, s->bf_T[b]);
        }
    }
    /* initBoundaryField (discreteVelocity.C:226-258, 1010-1024): mixed patches hold the
     * Maxwellian of the fixed patch values; incoming entries are never touched again. */
    memcpy(s->rho_b, s->bf_rho, sizeof(double) * nB);
    for (int v = 0; v < Nv; ++v)
        for (int b = 0; b < nB; ++b) {
            size_t k = (size_t)v * nF + nIF + b;
            if (s->bf_type[b] == ORC_PATCH_MIXED) {
                const double *xi = s->xi + 3 * v, *Ub = s->bf_U + 3 * b;
                double Tb = s->bf_T[b];
                double d[3] = {Ub[0] - xi[0], Ub[1] - xi[1], Ub[2] - xi[2]};
                double geq = s->bf_rho[b] / pow(sqrt(2.0 * ORC_PI * R * Tb), (double)D) * exp(-magsqr3(d) / (2.0 * R * Tb));
                s->gS[k] = geq;
                s->hS[k] = (cs->K + 3 - D) * R * Tb * geq;
            } else {
                s->gS[k] = 0.0;
                s->hS[k] = 0.0;
            }
        }
    for (int c = 0; c < nC; ++c) s->tau[c] = tau_of(cs, s->T[c], s->rho[c]);
    update_pressure_bc(s);  /* "set initial rho in pressureIn/Out BC" (F:2112-2113) */
    /* Usurf = linear interpolate(U) for the first CFL (F:2116) */
    for (int f = 0; f < nIF; ++f) {
        double lam = s->wts_if[f];
        int P = s->owner[f], N = s->neighbour[f];
        for (int k = 0; k < 3; ++k)
            s->US[3 * f + k] = lam * (s->U[3 * P + k] - s->U[3 * N + k]) + s->U[3 * N + k];
    }
    for (int b = 0; b < nB; ++b)
        for (int k = 0; k < 3; ++k) s->US[3 * (nIF + b) + k] = s->bf_U[3 * b + k];
    return 0;
}

/* ------------------------------------------------------------------- sub-steps */

/* updateGHbarPvol (F:648-683) */
static void update_gh_bar_pvol(orc_state *s, double dt)
{
    const orc_case *cs = &s->cs;
    const int nC = cs->nC, D = cs->D, K = cs->K;
    const double R = cs->R, Pr = cs->Pr;
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < s->Nv; ++v) {
        const double *xi = s->xi + 3 * v;
        for (int c = 0; c < nC; ++c) {
            const double rho = s->rho[c], T = s->T[c], *U = s->U + 3 * c, *q = s->q + 3 * c;
            double relaxFactor = 1.5 * dt / (2.0 * s->tau[c] + dt);
            double um[3] = {U[0] - xi[0], U[1] - xi[1], U[2] - xi[2]};
            double cSqrByRT = magsqr3(um) / (R * T);
            double xu[3] = {xi[0] - U[0], xi[1] - U[1], xi[2] - U[2]};
            double cqBy5pRT = dot3(xu, q) / (5.0 * rho * R * T * R * T);
            double gEqBGK = rho / pow(sqrt(2.0 * ORC_PI * R * T), (double)D) * exp(-cSqrByRT / 2.0);
            size_t k = (size_t)v * nC + c;
            s->gB[k] = (1.0 - relaxFactor) * s->gT[k]
                     + relaxFactor * (1.0 + (1.0 - Pr) * cqBy5pRT * (cSqrByRT - D - 2.0)) * gEqBGK;
            s->hB[k] = (1.0 - relaxFactor) * s->hT[k]
                     + relaxFactor * ((K + 3.0 - D) + (1.0 - Pr) * cqBy5pRT * ((cSqrByRT - D) * (K + 3.0 - D) - 2 * K)) * gEqBGK * R * T;
        }
    }
}

/* face value reconstructed from cell c: phi_c + grad_c . (Cf - C_c - 0.5 xi dt) (F:832-836) */
static double recon(const double *phi, const double *grad, int c, const double *Cf, const double *C,
                    const double *xi, double dt)
{
    double r[3];
    for (int k = 0; k < 3; ++k) r[k] = Cf[k] - C[3 * c + k] - 0.5 * xi[k] * dt;
    return phi[c] + dot3(grad + 3 * c, r);
}

/* updateGrad (discreteVelocity.C:366-369 -> leastSquaresGrad) + updateGHbarSurf
 * (F:755-1029) for one velocity.  Boundary rows of the gradient have zero right-hand
 * side: the fixedGradient patch values of f-bar-plus equal the cell values because the
 * patch gradient is fed back from the corrected boundary gradient (F:801-807), a fixed
 * point at its initial value 0. */
static void grad_of(const orc_state *s, const double *phi, double *grad)
{
    const int nC = s->cs.nC, nIF = s->cs.nIF;
    memset(grad, 0, sizeof(double) * 3 * nC);
    for (int f = 0; f < nIF; ++f) {
        int own = s->owner[f], nei = s->neighbour[f];
        double delta = phi[nei] - phi[own];
        for (int k = 0; k < 3; ++k) {
            grad[3 * own + k] += s->own_ls[3 * f + k] * delta;
            grad[3 * nei + k] -= s->nei_ls[3 * f + k] * delta;
        }
    }
}

static void update_gh_bar_surf(orc_state *s, double dt)
{
    const orc_case *cs = &s->cs;
    const int nC = cs->nC, nIF = cs->nIF, nB = cs->nB, nF = s->nF;
#pragma omp parallel
    {
        double *gg = malloc(sizeof(double) * 3 * nC);
        double *gh = malloc(sizeof(double) * 3 * nC);
#pragma omp for schedule(dynamic)
        for (int v = 0; v < s->Nv; ++v) {
            const double *xi = s->xi + 3 * v;
            const double *gB = s->gB + (size_t)v * nC, *hB = s->hB + (size_t)v * nC;
            double *gS = s->gS + (size_t)v * nF, *hS = s->hS + (size_t)v * nF;
            grad_of(s, gB, gg);
            grad_of(s, hB, gh);
            for (int f = 0; f < nIF; ++f) {
                int own = s->owner[f], nei = s->neighbour[f];
                const double *Cf = s->Cf + 3 * f;
                double xs = dot3(xi, s->Sf + 3 * f);
                if (xs >= ORC_VSMALL) {
                    gS[f] = recon(gB, gg, own, Cf, s->C, xi, dt);
                    hS[f] = recon(hB, gh, own, Cf, s->C, xi, dt);
                } else if (xs < -ORC_VSMALL) {
                    gS[f] = recon(gB, gg, nei, Cf, s->C, xi, dt);
                    hS[f] = recon(hB, gh, nei, Cf, s->C, xi, dt);
                } else {
                    gS[f] = 0.5 * (recon(gB, gg, nei, Cf, s->C, xi, dt) + recon(gB, gg, own, Cf, s->C, xi, dt));
                    hS[f] = 0.5 * (recon(hB, gh, nei, Cf, s->C, xi, dt) + recon(hB, gh, own, Cf, s->C, xi, dt));
                }
            }
            /* boundary faces (F:883-1027) */
            for (int b = 0; b < nB; ++b) {
                int f = nIF + b, own = s->owner[f];
                const int t = s->bf_type[b];
                double xs = dot3(xi, s->Sf + 3 * f);
                if (t == ORC_PATCH_ZERO_GRADIENT) {
                    /* every entry takes the owner's f-bar-plus, no reconstruction (F:887-891) */
                    gS[f] = gB[own];
                    hS[f] = hB[own];
                } else if (t == ORC_PATCH_SYMMETRY || t == ORC_PATCH_SYMMETRY_PLANE) {
                    /* outgoing and tangent entries reconstructed; incoming ones copied below (F:1009-1025) */
                    if (xs > -ORC_VSMALL) {
                        gS[f] = recon(gB, gg, own, s->Cf + 3 * f, s->C, xi, dt);
                        hS[f] = recon(hB, gh, own, s->Cf + 3 * f, s->C, xi, dt);
                    }
                } else if (xs > 0) {  /* maxwellWall, mixed, farField: outgoing reconstructed */
                    gS[f] = recon(gB, gg, own, s->Cf + 3 * f, s->C, xi, dt);
                    hS[f] = recon(hB, gh, own, s->Cf + 3 * f, s->C, xi, dt);
                } else if (t == ORC_PATCH_FAR_FIELD || t == ORC_PATCH_PRESSURE_IN || t == ORC_PATCH_PRESSURE_OUT) {
                    /* incoming: patch rho times the Maxwellian of the owner cell's U and the patch T (F:923-937) */
                    const double Tp = s->bf_T[b];
                    gS[f] = s->rho_b[b] * maxwell_by_rho(s, xi, s->U + 3 * own, Tp);
                    hS[f] = gS[f] * (cs->R * Tp) * (cs->K + 3 - cs->D);
                }
            }
        }
        free(gg);
        free(gh);
    }
    /* outGoing accumulation in velocity order (F:959-960) */
    for (int b = 0; b < nB; ++b) {
        if (s->bf_type[b] != ORC_PATCH_MAXWELL_WALL) continue;
        const double *Sf = s->Sf + 3 * (nIF + b);
        for (int v = 0; v < s->Nv; ++v) {
            double xs = dot3(s->xi + 3 * v, Sf);
            if (xs > 0) s->out_going[b] += s->w[v] * xs * s->gS[(size_t)v * nF + nIF + b];
        }
    }
}

/* updateMaxwellWallRho (F:1032-1052) -> calculatedMaxwellFvPatchField::evaluate (CM:140-164) */
static void update_maxwell_wall_rho(orc_state *s)
{
    for (int b = 0; b < s->cs.nB; ++b) {
        if (s->bf_type[b] != ORC_PATCH_MAXWELL_WALL) continue;
        s->rho_b[b] = s->out_going[b] / fabs(s->in_by_rho[b]);
        s->out_going[b] = 0.0;
    }
}

/* updateGHbarSurfMaxwellWallIn (F:1054-1098) */
static void update_gh_bar_surf_maxwell_wall_in(orc_state *s)
{
    const orc_case *cs = &s->cs;
    const int nIF = cs->nIF, nB = cs->nB, nF = s->nF, D = cs->D;
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < s->Nv; ++v) {
        const double *xi = s->xi + 3 * v;
        for (int b = 0; b < nB; ++b) {
            if (s->bf_type[b] != ORC_PATCH_MAXWELL_WALL) continue;
            int f = nIF + b;
            if (dot3(xi, s->Sf + 3 * f) <= 0) {
                size_t k = (size_t)v * nF + f;
                s->gS[k] = s->rho_b[b] * maxwell_by_rho(s, xi, s->bf_U + 3 * b, s->bf_T[b]);
                s->hS[k] = s->gS[k] * (cs->R * s->bf_T[b]) * (cs->K + 3 - D);
            }
        }
    }
}

/* updateGHbarSurfSymmetryIn (F:1100-1179, discreteVelocity.C:691-775): on symmetry
 * faces every incoming entry (xi . Sf0 <= 0, Sf0 the patch's first face) takes the face
 * value of the velocity mirrored about the patch axis */
static void update_gh_bar_surf_symmetry_in(orc_state *s)
{
    const int nIF = s->cs.nIF, nB = s->cs.nB, nF = s->nF, Nv = s->Nv;
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < Nv; ++v) {
        const double *xi = s->xi + 3 * v;
        for (int b = 0; b < nB; ++b) {
            const int axis = s->bf_mirror[b];
            if (axis < 0 || dot3(xi, s->bf_sf0 + 3 * b) > 0) continue;
            const int f = nIF + b, vm = s->mirror[(size_t)axis * Nv + v];
            s->gS[(size_t)v * nF + f] = s->gS[(size_t)vm * nF + f];
            s->hS[(size_t)v * nF + f] = s->hS[(size_t)vm * nF + f];
        }
    }
}

/* updateMacroSurf (F:1181-1300) */
static void update_macro_surf(orc_state *s, double dt)
{
    const orc_case *cs = &s->cs;
    const int nF = s->nF, Nv = s->Nv;
    const double R = cs->R, Pr = cs->Pr;
#pragma omp parallel for schedule(static)
    for (int f = 0; f < nF; ++f) {
        double rho = 0.0, rhoU[3] = {0, 0, 0}, rhoE = 0.0;
        for (int v = 0; v < Nv; ++v) {
            const double *xi = s->xi + 3 * v;
            double g = s->gS[(size_t)v * nF + f], h = s->hS[(size_t)v * nF + f];
            rho += s->w[v] * g;
            double wg = s->w[v] * g;
            rhoU[0] += wg * xi[0]; rhoU[1] += wg * xi[1]; rhoU[2] += wg * xi[2];
            rhoE += 0.5 * s->w[v] * (g * magsqr3(xi) + h);
        }
        double U[3] = {rhoU[0] / rho, rhoU[1] / rho, rhoU[2] / rho};
        double T = (rhoE - 0.5 * rho * magsqr3(U)) / ((cs->K + 3) / 2.0 * R * rho);
        double tau = tau_of(cs, T, rho);
        double q[3] = {0, 0, 0};
        for (int v = 0; v < Nv; ++v) {
            const double *xi = s->xi + 3 * v;
            double g = s->gS[(size_t)v * nF + f], h = s->hS[(size_t)v * nF + f];
            double c[3] = {xi[0] - U[0], xi[1] - U[1], xi[2] - U[2]};
            double m = magsqr3(c) * g + h;
            for (int k = 0; k < 3; ++k) q[k] += 0.5 * s->w[v] * c[k] * m;
        }
        double corr = 2.0 * tau / (2.0 * tau + 0.5 * dt * Pr);
        s->rhoS[f] = rho;
        s->TS[f] = T;
        s->tauS[f] = tau;
        for (int k = 0; k < 3; ++k) {
            s->US[3 * f + k] = U[k];
            s->qS[3 * f + k] = corr * q[k];
        }
    }
}

/* updateGHsurf (F:1362-1460) */
static void update_gh_surf(orc_state *s, double dt)
{
    const orc_case *cs = &s->cs;
    const int nIF = cs->nIF, nF = s->nF, D = cs->D, K = cs->K;
    const double R = cs->R, Pr = cs->Pr;
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < s->Nv; ++v) {
        const double *xi = s->xi + 3 * v;
        double *gS = s->gS + (size_t)v * nF, *hS = s->hS + (size_t)v * nF;
        for (int f = 0; f < nF; ++f) {
            int bnd = f >= nIF;
            if (bnd && !(dot3(xi, s->Sf + 3 * f) > 0)) continue;  /* patches: outgoing only */
            const double rho = s->rhoS[f], T = s->TS[f], *U = s->US + 3 * f, *q = s->qS + 3 * f;
            double relaxFactor = 0.5 * dt / (2.0 * s->tauS[f] + 0.5 * dt);
            double um[3] = {U[0] - xi[0], U[1] - xi[1], U[2] - xi[2]};
            double cSqrByRT = magsqr3(um) / (R * T);
            double xu[3] = {xi[0] - U[0], xi[1] - U[1], xi[2] - U[2]};
            double cqBy5pRT = dot3(xu, q) / (5.0 * rho * R * T * R * T);
            double gEqBGK = rho / pow(sqrt(2.0 * ORC_PI * R * T), (double)D) * exp(-cSqrByRT / 2.0);
            double gEq = (1.0 + (1.0 - Pr) * cqBy5pRT * (cSqrByRT - D - 2.0)) * gEqBGK;
            /* Q1: on patches hEq takes T from the internal-face Tsurf at the patch-local index */
            double Th = (bnd && cs->q1_compat) ? s->TS[s->bf_local[f - nIF]] : T;
            double hEq = ((K + 3.0 - D) + (1.0 - Pr) * cqBy5pRT * ((cSqrByRT - D) * (K + 3.0 - D) - 2 * K)) * gEqBGK * R * Th;
            gS[f] = (1.0 - relaxFactor) * gS[f] + relaxFactor * gEq;
            hS[f] = (1.0 - relaxFactor) * hS[f] + relaxFactor * hEq;
        }
        /* DVMsymmetry patches: the whole face is relaxed once more, outgoing entries
         * included (F:1449-1456) */
        for (int b = 0; b < cs->nB; ++b) {
            if (s->bf_type[b] != ORC_PATCH_SYMMETRY) continue;
            const int f = nIF + b;
            const double rho = s->rhoS[f], T = s->TS[f], *U = s->US + 3 * f, *q = s->qS + 3 * f;
            double relaxFactor = 0.5 * dt / (2.0 * s->tauS[f] + 0.5 * dt);
            double um[3] = {U[0] - xi[0], U[1] - xi[1], U[2] - xi[2]};
            double cSqrByRT = magsqr3(um) / (R * T);
            double xu[3] = {xi[0] - U[0], xi[1] - U[1], xi[2] - U[2]};
            double cqBy5pRT = dot3(xu, q) / (5.0 * rho * R * T * R * T);
            double gEqBGK = rho / pow(sqrt(2.0 * ORC_PI * R * T), (double)D) * exp(-cSqrByRT / 2.0);
            double gEq = (1.0 + (1.0 - Pr) * cqBy5pRT * (cSqrByRT - D - 2.0)) * gEqBGK;
            double Th = cs->q1_compat ? s->TS[s->bf_local[b]] : T;
            double hEq = ((K + 3.0 - D) + (1.0 - Pr) * cqBy5pRT * ((cSqrByRT - D) * (K + 3.0 - D) - 2 * K)) * gEqBGK * R * Th;
            gS[f] = (1.0 - relaxFactor) * gS[f] + relaxFactor * gEq;
            hS[f] = (1.0 - relaxFactor) * hS[f] + relaxFactor * hEq;
        }
    }
}

/* updateGHtildeVol (F:1462-1519) */
static void update_gh_tilde_vol(orc_state *s, double dt)
{
    const orc_case *cs = &s->cs;
    const int nC = cs->nC, nIF = cs->nIF, nB = cs->nB, nF = s->nF;
#pragma omp parallel for schedule(dynamic)
    for (int v = 0; v < s->Nv; ++v) {
        const double *xi = s->xi + 3 * v;
        double *gT = s->gT + (size_t)v * nC, *hT = s->hT + (size_t)v * nC;
        const double *gB = s->gB + (size_t)v * nC, *hB = s->hB + (size_t)v * nC;
        const double *gS = s->gS + (size_t)v * nF, *hS = s->hS + (size_t)v * nF;
        for (int c = 0; c < nC; ++c) {
            gT[c] = -1.0 / 3 * gT[c] + 4.0 / 3 * gB[c];
            hT[c] = -1.0 / 3 * hT[c] + 4.0 / 3 * hB[c];
        }
        for (int f = 0; f < nIF; ++f) {
            int own = s->owner[f], nei = s->neighbour[f];
            double xs = dot3(xi, s->Sf + 3 * f);
            gT[own] -= (xs * gS[f] * dt / s->V[own]);
            gT[nei] += (xs * gS[f] * dt / s->V[nei]);
            hT[own] -= (xs * hS[f] * dt / s->V[own]);
            hT[nei] += (xs * hS[f] * dt / s->V[nei]);
        }
        for (int b = 0; b < nB; ++b) {
            int f = nIF + b, own = s->owner[f];
            double xs = dot3(xi, s->Sf + 3 * f);
            gT[own] -= xs * gS[f] * dt / s->V[own];
            hT[own] -= xs * hS[f] * dt / s->V[own];
        }
    }
}

/* Conserved variables rho, rhoU, rhoE [nC][5] after the face fluxes of updateMacroVol's
 * macroFlux branch (F:1556-1679), starting from the previous step's cell macros (F:1523-1525).
 *  form 1 (default build, F:1620-1679): fluxes of the face macros.  The face stress is zero
 *    (stressSurf_ is reset and never accumulated, F:1187-1192, 1304-1306), the energy flux
 *    carries rho_f (|U_f|^2 + (K+3)/2 R T_f) U_f + q_f (no 1/2 on |U_f|^2, F:1630), and the
 *    boundary momentum flux is dotted with Sf[pFacei] of the internal-face field, i.e. the
 *    area vector of internal face number <patch-local index> (F:1562, 1667-1668).
 *  form 2 (VMESH build, F:1562-1618): fluxes of the moments 1, xi, (|xi|^2 g + h)/2 of the
 *    relaxed face DFs, summed into separate flux fields and added at the end (F:1615-1617). */
static void macro_flux_conserved(const orc_state *s, double dt, double *cons)
{
    const orc_case *cs = &s->cs;
    const int nC = cs->nC, nIF = cs->nIF, nF = s->nF, Nv = s->Nv;
    const double cv = (cs->K + 3) / 2.0 * cs->R;
    for (int c = 0; c < nC; ++c) {
        const double rho = s->rho[c], *U = s->U + 3 * c;
        cons[5 * c] = rho;
        for (int k = 0; k < 3; ++k) cons[5 * c + 1 + k] = rho * U[k];
        cons[5 * c + 4] = rho * (0.5 * magsqr3(U) + cv * s->T[c]);
    }
    if (cs->macro_flux == 1) {
        for (int f = 0; f < nF; ++f) {
            const int own = s->owner[f];
            const double *Sf = s->Sf + 3 * f, *U = s->US + 3 * f, *q = s->qS + 3 * f;
            const double rho = s->rhoS[f];
            /* momentum: (rho U U + 0) . S with S the internal-face field at the patch-local index */
            const double *Sm = f < nIF ? Sf : s->Sf + 3 * s->bf_local[f - nIF];
            const double mass = rho * dot3(U, Sf);
            const double un = dot3(U, Sm);
            const double rhoE = rho * (magsqr3(U) + cv * s->TS[f]);
            const double energy = rhoE * dot3(U, Sf) + dot3(q, Sf);
            double F[5] = {mass, rho * U[0] * un, rho * U[1] * un, rho * U[2] * un, energy};
            for (int k = 0; k < 5; ++k) cons[5 * own + k] -= F[k] * dt / s->V[own];
            if (f < nIF) {
                const int nei = s->neighbour[f];
                for (int k = 0; k < 5; ++k) cons[5 * nei + k] += F[k] * dt / s->V[nei];
            }
        }
        return;
    }
    double *flux = calloc(5 * (size_t)nC, sizeof(double));
    double *F = malloc(sizeof(double) * 5 * (size_t)nF);
#pragma omp parallel for schedule(static)
    for (int f = 0; f < nF; ++f) {
        double a[5] = {0, 0, 0, 0, 0};
        for (int v = 0; v < Nv; ++v) {
            const double *xi = s->xi + 3 * v;
            const double g = s->gS[(size_t)v * nF + f], h = s->hS[(size_t)v * nF + f];
            const double xs = dot3(xi, s->Sf + 3 * f);
            const double wg = s->w[v] * g;
            a[0] += xs * wg;
            for (int k = 0; k < 3; ++k) a[1 + k] += xs * wg * xi[k];
            a[4] += xs * 0.5 * s->w[v] * (g * magsqr3(xi) + h);
        }
        for (int k = 0; k < 5; ++k) F[5 * (size_t)f + k] = a[k];
    }
    for (int f = 0; f < nF; ++f) {
        const int own = s->owner[f];
        for (int k = 0; k < 5; ++k) flux[5 * own + k] -= F[5 * (size_t)f + k] * dt / s->V[own];
        if (f < nIF) {
            const int nei = s->neighbour[f];
            for (int k = 0; k < 5; ++k) flux[5 * nei + k] += F[5 * (size_t)f + k] * dt / s->V[nei];
        }
    }
    for (size_t k = 0; k < 5 * (size_t)nC; ++k) cons[k] += flux[k];
    free(F);
    free(flux);
}

/* updateMacroVol (F:1521-1743): rho, rhoU, rhoE from the moments of g~, h~ (macroFlux no,
 * F:1528-1554) or from the face fluxes (macroFlux yes, macro_flux_conserved); U, T from them
 * (F:1682-1684); tau; q from the moments of g~, h~ about U with the 2 tau / (2 tau + dt Pr)
 * correction (F:1700-1743). */
static void update_macro_vol(orc_state *s, double dt)
{
    const orc_case *cs = &s->cs;
    const int nC = cs->nC, Nv = s->Nv;
    const double R = cs->R, Pr = cs->Pr;
    double *cons = NULL;
    if (cs->macro_flux) {
        cons = malloc(sizeof(double) * 5 * (size_t)nC);
        macro_flux_conserved(s, dt, cons);
    }
#pragma omp parallel for schedule(static)
    for (int c = 0; c < nC; ++c) {
        double rho = 0.0, rhoU[3] = {0, 0, 0}, rhoE = 0.0;
        if (cons) {
            rho = cons[5 * c];
            for (int k = 0; k < 3; ++k) rhoU[k] = cons[5 * c + 1 + k];
            rhoE = cons[5 * c + 4];
        } else {
            for (int v = 0; v < Nv; ++v) {
                const double *xi = s->xi + 3 * v;
                double g = s->gT[(size_t)v * nC + c], h = s->hT[(size_t)v * nC + c];
                rho += s->w[v] * g;
                double wg = s->w[v] * g;
                rhoU[0] += wg * xi[0]; rhoU[1] += wg * xi[1]; rhoU[2] += wg * xi[2];
                rhoE += 0.5 * s->w[v] * (magsqr3(xi) * g + h);
            }
        }
        double U[3] = {rhoU[0] / rho, rhoU[1] / rho, rhoU[2] / rho};
        double T = (rhoE - 0.5 * rho * magsqr3(U)) / ((cs->K + 3) / 2.0 * R * rho);
        double tau = tau_of(cs, T, rho);
        double q[3] = {0, 0, 0};
        for (int v = 0; v < Nv; ++v) {
            const double *xi = s->xi + 3 * v;
            double g = s->gT[(size_t)v * nC + c], h = s->hT[(size_t)v * nC + c];
            double cc[3] = {xi[0] - U[0], xi[1] - U[1], xi[2] - U[2]};
            double m = magsqr3(cc) * g + h;
            for (int k = 0; k < 3; ++k) q[k] += 0.5 * s->w[v] * cc[k] * m;
        }
        double corr = 2.0 * tau / (2.0 * tau + dt * Pr);
        s->rho[c] = rho;
        s->T[c] = T;
        s->tau[c] = tau;
        for (int k = 0; k < 3; ++k) {
            s->U[3 * c + k] = U[k];
            s->q[3 * c + k] = corr * q[k];
        }
    }
    free(cons);
}

/* fvDVM::evolution (F:2128-2174).  GHbarPvolComm only exchanges processor halos and
 * evaluates fixedGradient patches whose values equal the cell values. */
int orc_evolution_stages(orc_state *s, double dt, int first, int last)
{
    if (!(dt > 0)) {
        set_err("orc_evolution: dt must be positive");
        return 1;
    }
    if (first < 0 || last > ORC_N_STAGES || first > last) {
        set_err("orc_evolution_stages: stage range outside [0, 8]");
        return 1;
    }
    for (int k = first; k < last; ++k) {
        switch (k) {
        case 0: update_gh_bar_pvol(s, dt); break;
        case 1: update_gh_bar_surf(s, dt); break;
        case 2: update_maxwell_wall_rho(s); break;
        case 3:
            update_gh_bar_surf_maxwell_wall_in(s);
            update_gh_bar_surf_symmetry_in(s);
            break;
        case 4: update_macro_surf(s, dt); break;
        case 5: update_gh_surf(s, dt); break;
        case 6: update_gh_tilde_vol(s, dt); break;
        case 7:
            update_macro_vol(s, dt);
            update_pressure_bc(s);  /* F:2172 */
            break;
        }
    }
    return 0;
}

int orc_evolution(orc_state *s, double dt) { return orc_evolution_stages(s, dt, 0, ORC_N_STAGES); }

/* getCoNum (F:2177-2185): internal faces, previous Usurf (Q10) */
int orc_co_num(const orc_state *s, double dt, double *co_max, double *co_mean)
{
    const int nIF = s->cs.nIF;
    double mx = -1e300, sum = 0.0;
    for (int f = 0; f < nIF; ++f) {
        double u = s->dc[f] * (sqrt(magsqr3(s->US + 3 * f)) + sqrt((double)s->cs.D) * s->cs.xi_max);
        if (u > mx) mx = u;
        sum += u;
    }
    *co_max = mx * dt;
    *co_mean = nIF ? sum / nIF * dt : 0.0;
    return 0;
}

/* ----------------------------------------------------------------------- getters */
int orc_nv(const orc_state *s) { return s->Nv; }

int orc_get_cell_macros(const orc_state *s, double *rho, double *U, double *T, double *q, double *tau)
{
    const int nC = s->cs.nC;
    if (rho) memcpy(rho, s->rho, sizeof(double) * nC);
    if (U) memcpy(U, s->U, sizeof(double) * 3 * nC);
    if (T) memcpy(T, s->T, sizeof(double) * nC);
    if (q) memcpy(q, s->q, sizeof(double) * 3 * nC);
    if (tau) memcpy(tau, s->tau, sizeof(double) * nC);
    return 0;
}

int orc_set_cell_macros(orc_state *s, const double *rho, const double *U, const double *T, const double *q)
{
    const int nC = s->cs.nC;
    memcpy(s->rho, rho, sizeof(double) * nC);
    memcpy(s->U, U, sizeof(double) * 3 * nC);
    memcpy(s->T, T, sizeof(double) * nC);
    memcpy(s->q, q, sizeof(double) * 3 * nC);
    for (int c = 0; c < nC; ++c) s->tau[c] = tau_of(&s->cs, s->T[c], s->rho[c]);
    return 0;
}

int orc_get_face_macros(const orc_state *s, double *rho, double *U, double *T, double *q, double *tau)
{
    const int nF = s->nF;
    if (rho) memcpy(rho, s->rhoS, sizeof(double) * nF);
    if (U) memcpy(U, s->US, sizeof(double) * 3 * nF);
    if (T) memcpy(T, s->TS, sizeof(double) * nF);
    if (q) memcpy(q, s->qS, sizeof(double) * 3 * nF);
    if (tau) memcpy(tau, s->tauS, sizeof(double) * nF);
    return 0;
}

int orc_get_df(const orc_state *s, double *g, double *h)
{
    size_t n = (size_t)s->Nv * s->cs.nC;
    if (g) memcpy(g, s->gT, sizeof(double) * n);
    if (h) memcpy(h, s->hT, sizeof(double) * n);
    return 0;
}

int orc_set_df(orc_state *s, const double *g, const double *h)
{
    size_t n = (size_t)s->Nv * s->cs.nC;
    memcpy(s->gT, g, sizeof(double) * n);
    memcpy(s->hT, h, sizeof(double) * n);
    return 0;
}

int orc_get_bar(const orc_state *s, double *g, double *h)
{
    size_t n = (size_t)s->Nv * s->cs.nC;
    if (g) memcpy(g, s->gB, sizeof(double) * n);
    if (h) memcpy(h, s->hB, sizeof(double) * n);
    return 0;
}

int orc_get_surf(const orc_state *s, double *g, double *h)
{
    size_t n = (size_t)s->Nv * s->nF;
    if (g) memcpy(g, s->gS, sizeof(double) * n);
    if (h) memcpy(h, s->hS, sizeof(double) * n);
    return 0;
}

int orc_get_wall_rho(const orc_state *s, double *rho_b)
{
    memcpy(rho_b, s->rho_b, sizeof(double) * s->cs.nB);
    return 0;
}

int orc_get_incoming_by_rho(const orc_state *s, double *in_b)
{
    memcpy(in_b, s->in_by_rho, sizeof(double) * s->cs.nB);
    return 0;
}

int orc_get_ls_vectors(const orc_state *s, double *own_ls, double *nei_ls, double *patch_ls)
{
    if (own_ls) memcpy(own_ls, s->own_ls, sizeof(double) * 3 * s->cs.nIF);
    if (nei_ls) memcpy(nei_ls, s->nei_ls, sizeof(double) * 3 * s->cs.nIF);
    if (patch_ls) memcpy(patch_ls, s->patch_ls, sizeof(double) * 3 * s->cs.nB);
    return 0;
}

int orc_get_vset(const orc_state *s, double *xi, double *w)
{
    if (xi) memcpy(xi, s->xi, sizeof(double) * 3 * s->Nv);
    if (w) memcpy(w, s->w, sizeof(double) * s->Nv);
    return 0;
}
