// Host-side mesh preparation for libdugks_hip.so: least-squares gradient vectors and the
// cell-wise CSR stencils the kernels gather through.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "dugks_internal.h"

namespace dugks {

namespace {

struct Symm {
    double xx = 0, xy = 0, xz = 0, yy = 0, yz = 0, zz = 0;
    void add(double f, const double *d)
    {
        xx += f * (d[0] * d[0]); xy += f * (d[0] * d[1]); xz += f * (d[0] * d[2]);
        yy += f * (d[1] * d[1]); yz += f * (d[1] * d[2]); zz += f * (d[2] * d[2]);
    }
    void apply(const double *d, double *o) const
    {
        o[0] = xx * d[0] + xy * d[1] + xz * d[2];
        o[1] = xy * d[0] + yy * d[1] + yz * d[2];
        o[2] = xz * d[0] + yz * d[1] + zz * d[2];
    }
};

Symm inverse(const Symm &s)
{
    // cofactor inverse of a symmetric 3x3 tensor (OpenFOAM inv(SymmTensor))
    const double det = s.xx * s.yy * s.zz + s.xy * s.yz * s.xz + s.xz * s.xy * s.yz - s.xx * s.yz * s.yz -
                       s.xy * s.xy * s.zz - s.xz * s.yy * s.xz;
    Symm r;
    r.xx = (s.yy * s.zz - s.yz * s.yz) / det;
    r.xy = (s.xz * s.yz - s.xy * s.zz) / det;
    r.xz = (s.xy * s.yz - s.xz * s.yy) / det;
    r.yy = (s.xx * s.zz - s.xz * s.xz) / det;
    r.yz = (s.xy * s.xz - s.xx * s.yz) / det;
    r.zz = (s.xx * s.yy - s.xy * s.xy) / det;
    return r;
}

double msq(const double *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

}  // namespace

// Least-squares vectors of OpenFOAM's leastSquaresVectors, in the form recorded in the
// reference tree (src/fvDVM/zeroBoundaryGrad/zeroBoundaryVectors.C:115-122, 181-184):
//   dd_P += (1-w) |Sf|/|d|^2 d d,  dd_N += w |Sf|/|d|^2 d d     (internal faces)
//   dd_P += |Sf|/|d_b|^2 d_b d_b,  d_b = n (n.(Cf - C_P))        (non-coupled boundary faces)
//   ownLs = (1-w) |Sf|/|d|^2 dd_P^-1 d,  neiLs = -w |Sf|/|d|^2 dd_N^-1 d
// dd^-1 follows OpenFOAM inv(symmTensorField): directions negligible in the first cell's
// tensor (one-cell-thick meshes) are regularised with +1 before and -1 after inversion;
// first_cell is the internal label of the caller's cell 0.
bool compute_ls_vectors(int nC, int nIF, int nB, const int *owner, const int *neighbour, const double *Sf,
                        const double *Cf, const double *C, const double *weights, int first_cell,
                        std::vector<double> &own_ls, std::vector<double> &nei_ls, std::string &err)
{
    std::vector<Symm> dd(nC);
    for (int f = 0; f < nIF; ++f) {
        const int P = owner[f], N = neighbour[f];
        const double d[3] = {C[3 * N] - C[3 * P], C[3 * N + 1] - C[3 * P + 1], C[3 * N + 2] - C[3 * P + 2]};
        const double dsq = msq(d);
        if (!(dsq > 0)) {
            err = "coincident cell centres across internal face " + std::to_string(f);
            return false;
        }
        Symm wdd;
        wdd.add(std::sqrt(msq(Sf + 3 * f)) / dsq, d);
        const double w = weights[f];
        Symm &a = dd[P], &b = dd[N];
        a.xx += (1 - w) * wdd.xx; a.xy += (1 - w) * wdd.xy; a.xz += (1 - w) * wdd.xz;
        a.yy += (1 - w) * wdd.yy; a.yz += (1 - w) * wdd.yz; a.zz += (1 - w) * wdd.zz;
        b.xx += w * wdd.xx; b.xy += w * wdd.xy; b.xz += w * wdd.xz;
        b.yy += w * wdd.yy; b.yz += w * wdd.yz; b.zz += w * wdd.zz;
    }
    for (int b = 0; b < nB; ++b) {
        const int f = nIF + b, P = owner[f];
        const double *s = Sf + 3 * f;
        const double mag = std::sqrt(msq(s));
        const double n[3] = {s[0] / mag, s[1] / mag, s[2] / mag};
        const double r[3] = {Cf[3 * f] - C[3 * P], Cf[3 * f + 1] - C[3 * P + 1], Cf[3 * f + 2] - C[3 * P + 2]};
        const double nd = n[0] * r[0] + n[1] * r[1] + n[2] * r[2];
        const double d[3] = {n[0] * nd, n[1] * nd, n[2] * nd};
        const double dsq = msq(d);
        if (!(dsq > 0)) {
            err = "boundary face centre lies in its cell centre plane, face " + std::to_string(f);
            return false;
        }
        dd[P].add(mag / dsq, d);
    }
    bool rm[3] = {false, false, false};
    if (nC > 0) {
        const Symm &t = dd[first_cell];
        const double scale = t.xx * t.xx + 2 * t.xy * t.xy + 2 * t.xz * t.xz + t.yy * t.yy + 2 * t.yz * t.yz + t.zz * t.zz;
        rm[0] = (t.xx * t.xx) / scale < kSmall;
        rm[1] = (t.yy * t.yy) / scale < kSmall;
        rm[2] = (t.zz * t.zz) / scale < kSmall;
    }
    for (int c = 0; c < nC; ++c) {
        Symm t = dd[c];
        if (rm[0]) t.xx += 1;
        if (rm[1]) t.yy += 1;
        if (rm[2]) t.zz += 1;
        t = inverse(t);
        if (rm[0]) t.xx -= 1;
        if (rm[1]) t.yy -= 1;
        if (rm[2]) t.zz -= 1;
        if (!std::isfinite(t.xx) || !std::isfinite(t.yy) || !std::isfinite(t.zz)) {
            err = "singular least-squares matrix in cell " + std::to_string(c);
            return false;
        }
        dd[c] = t;
    }
    own_ls.assign(3 * (size_t)nIF, 0.0);
    nei_ls.assign(3 * (size_t)nIF, 0.0);
    for (int f = 0; f < nIF; ++f) {
        const int P = owner[f], N = neighbour[f];
        const double d[3] = {C[3 * N] - C[3 * P], C[3 * N + 1] - C[3 * P + 1], C[3 * N + 2] - C[3 * P + 2]};
        const double fac = std::sqrt(msq(Sf + 3 * f)) / msq(d);
        const double w = weights[f];
        double a[3], b[3];
        dd[P].apply(d, a);
        dd[N].apply(d, b);
        for (int k = 0; k < 3; ++k) {
            own_ls[3 * f + k] = (1 - w) * fac * a[k];
            nei_ls[3 * f + k] = -w * fac * b[k];
        }
    }
    return true;
}

// Cell-wise CSR lists.  cf: every face of a cell in ascending face label (internal faces
// precede boundary faces), which is the order in which the reference's face-order scatter
// (src/fvDVM/fvDVM/fvDVM.C:1489-1517) updates that cell.  cg: the internal faces of a cell
// with the other cell and the LS vector seen from this cell; a term of the gradient is
// cg_ls . (phi_nb - phi_c), equal to ownLs*(phi_N - phi_P) for the owner and to
// -neiLs*(phi_N - phi_P) for the neighbour (leastSquaresGrad face loop).
bool build_stencils(int nC, int nIF, int nB, const int *owner, const int *neighbour, const double *own_ls,
                    const double *nei_ls, HostPrep &p, std::string &err)
{
    const int nF = nIF + nB;
    std::vector<int> cnt(nC + 1, 0), gcnt(nC + 1, 0);
    for (int f = 0; f < nF; ++f) {
        const int P = owner[f];
        if (P < 0 || P >= nC) {
            err = "owner label out of range at face " + std::to_string(f);
            return false;
        }
        cnt[P + 1]++;
        if (f < nIF) {
            const int N = neighbour[f];
            if (N < 0 || N >= nC || N == P) {
                err = "neighbour label invalid at face " + std::to_string(f);
                return false;
            }
            cnt[N + 1]++;
            gcnt[P + 1]++;
            gcnt[N + 1]++;
        }
    }
    for (int c = 0; c < nC; ++c) {
        cnt[c + 1] += cnt[c];
        gcnt[c + 1] += gcnt[c];
    }
    p.cf_off = cnt;
    p.cg_off = gcnt;
    p.cf_face.assign(cnt[nC], 0);
    p.cg_nb.assign(gcnt[nC], 0);
    p.cg_ls.assign(gcnt[nC], d3{0, 0, 0});
    std::vector<int> fill(cnt.begin(), cnt.end() - 1), gfill(gcnt.begin(), gcnt.end() - 1);
    // ascending face label per cell: visit faces in order
    for (int f = 0; f < nF; ++f) {
        const int P = owner[f];
        p.cf_face[fill[P]++] = f;
        if (f < nIF) {
            const int N = neighbour[f];
            p.cf_face[fill[N]++] = f;
            int k = gfill[P]++;
            p.cg_nb[k] = N;
            p.cg_ls[k] = d3{own_ls[3 * f], own_ls[3 * f + 1], own_ls[3 * f + 2]};
            k = gfill[N]++;
            p.cg_nb[k] = P;
            p.cg_ls[k] = d3{nei_ls[3 * f], nei_ls[3 * f + 1], nei_ls[3 * f + 2]};
        }
    }
    return true;
}

}  // namespace dugks
