// Host-side planning of the fused path: cell blocks by recursive coordinate bisection, the
// halo ring of each block, and the block-local tables (gradient stencils of the block
// cells, the faces of the block cells with their gather order, the face-value column of
// every face slot, one per face even where two blocks share the face) that the kernels follow.
#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "dugks_fused.h"

namespace dugks {

namespace {

void rcb(std::vector<int> &ids, int lo, int hi, const double *C, int max_block, std::vector<int> &starts)
{
    const int n = hi - lo;
    if (n <= max_block) {
        starts.push_back(lo);
        return;
    }
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = lo; i < hi; ++i)
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], C[3 * ids[i] + k]);
            mx[k] = std::max(mx[k], C[3 * ids[i] + k]);
        }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (mx[k] - mn[k] > mx[ax] - mn[ax]) ax = k;
    // split into two parts whose sizes are multiples of max_block where possible, so that
    // regular meshes give square / cubic blocks
    const int nblocks = (n + max_block - 1) / max_block;
    const int left_blocks = nblocks / 2;
    const int mid = lo + std::min(n - 1, std::max(1, left_blocks * max_block));
    std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi, [&](int a, int b) {
        const double ca = C[3 * a + ax], cb = C[3 * b + ax];
        return ca < cb || (ca == cb && a < b);
    });
    // keep each half in ascending original order for locality inside a block
    std::sort(ids.begin() + lo, ids.begin() + mid);
    std::sort(ids.begin() + mid, ids.begin() + hi);
    rcb(ids, lo, mid, C, max_block, starts);
    rcb(ids, mid, hi, C, max_block, starts);
}

int max_local_cells(int D)
{
    return kFusedThreadsA * fused_cells_per_thread_a(D);
}

int max_face_slots(int D)
{
    return std::min(kFusedThreadsA * kFusedSlotsA, kFusedThreadsB * fused_faces_per_thread_b(D));
}

}  // namespace

void rcb_partition(int nC, const double *C, int max_block, std::vector<int> &perm, std::vector<int> &block_start)
{
    perm.resize(nC);
    std::iota(perm.begin(), perm.end(), 0);
    block_start.clear();
    rcb(perm, 0, nC, C, max_block, block_start);
    block_start.push_back(nC);
}

void rcb_refine(const double *C, const std::vector<int> &blocks, std::vector<int> &perm,
                std::vector<int> &block_start)
{
    std::vector<char> split(block_start.size(), 0);
    for (int b : blocks) split[b] = 1;
    std::vector<int> starts;
    for (size_t b = 0; b + 1 < block_start.size(); ++b) {
        const int lo = block_start[b], hi = block_start[b + 1];
        if (split[b] && hi - lo > 1)
            rcb(perm, lo, hi, C, (hi - lo + 1) / 2, starts);
        else
            starts.push_back(lo);
    }
    starts.push_back(block_start.back());
    block_start.swap(starts);
}

bool block_tables_fit(int nC, int nIF, int nF, const int *owner, const int *neighbour, int D,
                      const std::vector<int> &perm, const std::vector<int> &block_start, std::vector<int> *bad)
{
    if (bad) bad->clear();
    // cell -> internal-face neighbours and cell -> faces, in the caller's numbering
    std::vector<int> deg(nC + 1, 0), fdeg(nC + 1, 0);
    for (int f = 0; f < nIF; ++f) {
        ++deg[owner[f] + 1];
        ++deg[neighbour[f] + 1];
    }
    for (int f = 0; f < nF; ++f) {
        ++fdeg[owner[f] + 1];
        if (f < nIF) ++fdeg[neighbour[f] + 1];
    }
    for (int c = 0; c < nC; ++c) {
        deg[c + 1] += deg[c];
        fdeg[c + 1] += fdeg[c];
    }
    std::vector<int> nb(deg[nC]), fc(fdeg[nC]), pos(deg.begin(), deg.end() - 1), fpos(fdeg.begin(), fdeg.end() - 1);
    for (int f = 0; f < nIF; ++f) {
        nb[pos[owner[f]]++] = neighbour[f];
        nb[pos[neighbour[f]]++] = owner[f];
    }
    for (int f = 0; f < nF; ++f) {
        fc[fpos[owner[f]]++] = f;
        if (f < nIF) fc[fpos[neighbour[f]]++] = f;
    }
    std::vector<int> mark(nC, -1), fmark(nF, -1);
    const int nblk = (int)block_start.size() - 1;
    for (int b = 0; b < nblk; ++b) {
        const int c0 = block_start[b], c1 = block_start[b + 1];
        if (c1 - c0 > kFusedThreadsB) {
            if (!bad) return false;
            bad->push_back(b);
            continue;
        }
        int nl = c1 - c0, nfs = 0;
        for (int i = c0; i < c1; ++i) mark[perm[i]] = b;
        for (int i = c0; i < c1; ++i) {
            const int c = perm[i];
            for (int k = fdeg[c]; k < fdeg[c + 1]; ++k)
                if (fmark[fc[k]] != b) {
                    fmark[fc[k]] = b;
                    ++nfs;
                }
            for (int k = deg[c]; k < deg[c + 1]; ++k)
                if (mark[nb[k]] != b) {
                    mark[nb[k]] = b;
                    ++nl;
                }
        }
        if (nl > max_local_cells(D) || nfs > max_face_slots(D)) {
            if (!bad) return false;
            bad->push_back(b);
        }
    }
    return !bad || bad->empty();
}

bool fused_plan_build(FusedPlan &p, const Geo &g, const HostPrep &prep, const std::vector<int> &block_start,
                      size_t es, int nvg_hint, std::string &err)
{
    p = FusedPlan();
    if (g.D < 2) {
        err = "fused path needs a 2-D or 3-D mesh";
        return false;
    }
    const int nC = g.nC, nIF = g.nIF, nF = g.nF;
    const int nblk = (int)block_start.size() - 1;
    const std::vector<int> &cg_off = prep.cg_off, &cg_nb = prep.cg_nb;
    const std::vector<int> &cf_off = prep.cf_off, &cf_face = prep.cf_face;
    const std::vector<int> &owner = prep.owner_h, &neighbour = prep.neighbour_h;

    std::vector<int> block_of(nC);
    for (int b = 0; b < nblk; ++b)
        for (int c = block_start[b]; c < block_start[b + 1]; ++c) block_of[c] = b;
    // a face shared by two blocks has a slot in each; both slots name the face's one column
    // of the face-value array (the column of the face's first slot, numbered in slot order)
    std::vector<int> first_slot(nF, -1);
    int ncol = 0;

    p.nblk = nblk;
    p.blk.resize(nblk);
    std::vector<int> mark(nC, -1), local(nC, -1);
    for (int b = 0; b < nblk; ++b) {
        BlockInfo bi{};
        const int c0 = block_start[b], c1 = block_start[b + 1];
        bi.c0 = c0;
        bi.nc = c1 - c0;
        bi.l_off = (int)p.lcell.size();
        std::vector<int> L;
        for (int c = c0; c < c1; ++c) {
            L.push_back(c);
            mark[c] = b;
        }
        std::vector<int> h1;
        for (int c = c0; c < c1; ++c)
            for (int k = cg_off[c]; k < cg_off[c + 1]; ++k) {
                const int n = cg_nb[k];
                if (mark[n] != b) {
                    mark[n] = b;
                    h1.push_back(n);
                }
            }
        std::sort(h1.begin(), h1.end());
        L.insert(L.end(), h1.begin(), h1.end());
        bi.nl = (int)L.size();
        for (int l = 0; l < bi.nl; ++l) local[L[l]] = l;
        p.lcell.insert(p.lcell.end(), L.begin(), L.end());
        // gradient stencils of the block cells, ascending face label (the order of cg)
        bi.gsb_off = (int)p.gs_beg.size();
        for (int l = 0; l < bi.nc; ++l) {
            const int c = L[l];
            p.gs_beg.push_back((int)p.gs_nb.size());
            for (int k = cg_off[c]; k < cg_off[c + 1]; ++k) {
                const int n = cg_nb[k];
                if (local[n] < 0) {
                    err = "internal error: stencil cell outside the halo";
                    return false;
                }
                p.gs_nb.push_back(local[n]);
                p.gs_ls.push_back(prep.cg_ls[k]);
            }
        }
        p.gs_beg.push_back((int)p.gs_nb.size());
        // faces of the block cells, one slot each: faces inside the block first, then the
        // faces shared with each adjacent block (grouped by that block, ascending face
        // label), then boundary faces.  Both blocks of a shared face order their group the
        // same way, so the columns of a run of lanes' shared faces are a contiguous run
        // whichever block numbered them, and the passes' stores and loads stay coalesced.
        bi.fs_off = (int)p.fs_face.size();
        std::unordered_map<int, int> slot_of;
        std::vector<std::array<int, 3>> keyed;  // (group, adjacent block, face)
        for (int c = c0; c < c1; ++c)
            for (int k = cf_off[c]; k < cf_off[c + 1]; ++k) {
                const int f = cf_face[k];
                if (!slot_of.emplace(f, 0).second) continue;
                if (f >= nIF) {
                    keyed.push_back({2, 0, f});
                } else {
                    const int bo = block_of[owner[f]], bn = block_of[neighbour[f]];
                    if (bo == b && bn == b)
                        keyed.push_back({0, 0, f});
                    else
                        keyed.push_back({1, bo == b ? bn : bo, f});
                }
            }
        std::sort(keyed.begin(), keyed.end());
        for (const auto &kf : keyed) {
            const int f = kf[2];
            slot_of[f] = (int)p.fs_face.size() - bi.fs_off;
            const int o = owner[f], n = f < nIF ? neighbour[f] : -1;
            int flag = 0;
            if (block_of[o] == b) flag |= kSlotOwnIn;
            if (n >= 0 && block_of[n] == b) flag |= kSlotNeiIn;
            const int gslot = (int)p.fs_face.size();
            p.fs_face.push_back(f);
            if (first_slot[f] < 0) {
                first_slot[f] = gslot;
                p.fs_col.push_back(ncol++);
            } else {
                p.fs_col.push_back(p.fs_col[first_slot[f]]);
            }
            p.fs_lo.push_back(local[o]);
            p.fs_ln.push_back(n >= 0 ? local[n] : -1);
            p.fs_flag.push_back((signed char)flag);
        }
        // per-cell gather order: the cell's faces in ascending face label
        bi.fbc_off = (int)p.fbc_beg.size();
        for (int c = c0; c < c1; ++c) {
            p.fbc_beg.push_back((int)p.fbc_idx.size());
            for (int k = cf_off[c]; k < cf_off[c + 1]; ++k) {
                const int f = cf_face[k];
                p.fbc_idx.push_back(slot_of[f]);
                // the owner subtracts the flux, the neighbour adds it (fvDVM.C:1493-1496)
                p.fbc_sgn.push_back((signed char)((f < nIF && owner[f] != c) ? 1 : -1));
            }
        }
        p.fbc_beg.push_back((int)p.fbc_idx.size());
        bi.nfs = (int)p.fs_face.size() - bi.fs_off;
        for (int l = 0; l < bi.nl; ++l) local[L[l]] = -1;
        p.blk[b] = bi;
        p.max_nl = std::max(p.max_nl, bi.nl);
        p.max_nfs = std::max(p.max_nfs, bi.nfs);
        p.max_nc = std::max(p.max_nc, bi.nc);
    }
    int max_st = 0, max_cf = 0;
    for (size_t i = 0; i + 1 < p.gs_beg.size(); ++i) max_st = std::max(max_st, p.gs_beg[i + 1] - p.gs_beg[i]);
    for (size_t i = 0; i + 1 < p.fbc_beg.size(); ++i) max_cf = std::max(max_cf, p.fbc_beg[i + 1] - p.fbc_beg[i]);
    if (max_st > kFusedMaxNb || max_cf > kFusedMaxNb) {
        err = "cells with more than " + std::to_string(kFusedMaxNb) + " faces";
        return false;
    }
    if (p.max_nl > max_local_cells(g.D) || p.max_nfs > max_face_slots(g.D) || p.max_nc > kFusedThreadsB) {
        err = "block tables exceed the work-group shape (cells " + std::to_string(p.max_nc) + ", local cells " +
              std::to_string(p.max_nl) + ", faces " + std::to_string(p.max_nfs) + ")";
        return false;
    }
    p.n_slots = (int)p.fs_face.size();
    p.n_cols = ncol;
    p.nm = g.D == 3 ? 17 : 11;
    p.nma = p.nm + 1;
    // about 6k work-groups per pass: the last round of work-groups on each die then holds a
    // small share of the pass (211k-cell sphere, 826 blocks: 8 velocity groups instead of 5,
    // pass A 103.7 -> 99.8 ms, pass B 96.3 -> 92.5 ms with FP32 storage; 12 and 16 groups
    // gain nothing more while the face-moment partials grow)
    p.wg_target = nvg_hint > 0 ? nvg_hint * nblk : 6144;
    (void)es;
    p.ready = true;
    return true;
}

}  // namespace dugks
