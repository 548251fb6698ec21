// Host-side planning of the fused path, recompute schedule: the two halo rings of each
// block and the block-local tables (gradient stencils, owned faces, faces of the block
// cells with their gather order) that the kernels follow.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "dugks_fused_rc.h"

namespace dugks {
namespace rc {

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
    const int max_nl = std::min(kFusedThreadsA * fused_cells_per_thread_a(D), kFusedThreadsB * fused_cells_per_thread_b(D));
    const int max_nfb = kFusedThreadsB * fused_faces_per_thread_b(D);
    std::vector<int> mark(nC, -1), fmark(nF, -1);
    const int nblk = (int)block_start.size() - 1;
    for (int b = 0; b < nblk; ++b) {
        std::vector<int> ring(perm.begin() + block_start[b], perm.begin() + block_start[b + 1]);
        if ((int)ring.size() > kFusedThreadsB) {
            if (!bad) return false;
            bad->push_back(b);
            continue;
        }
        int nl = (int)ring.size(), nfb = 0;
        for (int c : ring) {
            mark[c] = b;
            for (int k = fdeg[c]; k < fdeg[c + 1]; ++k)
                if (fmark[fc[k]] != b) {
                    fmark[fc[k]] = b;
                    ++nfb;
                }
        }
        for (int r = 0; r < 2; ++r) {
            std::vector<int> next;
            for (int c : ring)
                for (int k = deg[c]; k < deg[c + 1]; ++k)
                    if (mark[nb[k]] != b) {
                        mark[nb[k]] = b;
                        next.push_back(nb[k]);
                    }
            nl += (int)next.size();
            ring.swap(next);
        }
        if (nl > max_nl || nfb > max_nfb) {
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
    const int nC = g.nC, nIF = g.nIF;
    const int nblk = (int)block_start.size() - 1;
    // cell adjacency comes from the gradient stencil (internal faces)
    const std::vector<int> &cg_off = prep.cg_off, &cg_nb = prep.cg_nb;
    const std::vector<int> &cf_off = prep.cf_off, &cf_face = prep.cf_face;
    const std::vector<int> &owner = prep.owner_h, &neighbour = prep.neighbour_h;

    p.nblk = nblk;
    p.blk.resize(nblk);
    std::vector<int> mark(nC, -1);
    std::vector<int> local(nC, -1);
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
        std::vector<int> h1, h2;
        for (int c = c0; c < c1; ++c)
            for (int k = cg_off[c]; k < cg_off[c + 1]; ++k) {
                const int n = cg_nb[k];
                if (mark[n] != b) {
                    mark[n] = b;
                    h1.push_back(n);
                }
            }
        std::sort(h1.begin(), h1.end());
        for (int c : h1)
            for (int k = cg_off[c]; k < cg_off[c + 1]; ++k) {
                const int n = cg_nb[k];
                if (mark[n] != b) {
                    mark[n] = b;
                    h2.push_back(n);
                }
            }
        std::sort(h2.begin(), h2.end());
        L.insert(L.end(), h1.begin(), h1.end());
        L.insert(L.end(), h2.begin(), h2.end());
        bi.nl = (int)L.size();
        bi.ng = bi.nc + (int)h1.size();
        for (int l = 0; l < bi.nl; ++l) local[L[l]] = l;
        p.lcell.insert(p.lcell.end(), L.begin(), L.end());
        // gradient stencils of block + ring 1, ascending face label (the order of cg)
        // fixed stride of kFusedMaxNb entries per gradient cell (padded with local cell 0 and
        // zero vectors), so a neighbour's address does not wait on a loaded range
        bi.gsb_off = (int)p.gs_cnt.size();
        for (int l = 0; l < bi.ng; ++l) {
            const int c = L[l];
            const int n_st = cg_off[c + 1] - cg_off[c];
            if (n_st > kFusedMaxNb) {
                err = "cells with more than " + std::to_string(kFusedMaxNb) + " faces";
                return false;
            }
            p.gs_cnt.push_back(n_st);
            for (int k = cg_off[c]; k < cg_off[c + 1]; ++k) {
                const int n = cg_nb[k];
                if (mark[n] != b || local[n] < 0 || L[local[n]] != n) {
                    err = "internal error: stencil cell outside the halo";
                    return false;
                }
                p.gs_nb.push_back(local[n]);
                p.gs_ls.push_back(prep.cg_ls[k]);
            }
            for (int j = n_st; j < kFusedMaxNb; ++j) {
                p.gs_nb.push_back(0);
                p.gs_ls.push_back(d3{0, 0, 0});
            }
        }
        // owned faces (owner in block), ascending face label
        bi.fa_off = (int)p.fa_face.size();
        for (int c = c0; c < c1; ++c)
            for (int k = cf_off[c]; k < cf_off[c + 1]; ++k) {
                const int f = cf_face[k];
                if (owner[f] != c) continue;
                p.fa_face.push_back(f);
                p.fa_own.push_back(local[c]);
                p.fa_nei.push_back(f < nIF ? local[neighbour[f]] : -1);
            }
        bi.nfa = (int)p.fa_face.size() - bi.fa_off;
        // faces of the block cells: unique list + per-cell gather order
        bi.fb_off = (int)p.fb_face.size();
        std::unordered_map<int, int> fb_index;
        bi.fbc_off = (int)p.fbc_beg.size();
        for (int c = c0; c < c1; ++c) {
            p.fbc_beg.push_back((int)p.fbc_idx.size());
            for (int k = cf_off[c]; k < cf_off[c + 1]; ++k) {
                const int f = cf_face[k];
                auto it = fb_index.find(f);
                int idx;
                if (it == fb_index.end()) {
                    idx = (int)p.fb_face.size() - bi.fb_off;
                    fb_index.emplace(f, idx);
                    p.fb_face.push_back(f);
                    p.fb_own.push_back(local[owner[f]]);
                    p.fb_nei.push_back(f < nIF ? local[neighbour[f]] : -1);
                } else {
                    idx = it->second;
                }
                p.fbc_idx.push_back(idx);
                // the owner subtracts the flux, the neighbour adds it (fvDVM.C:1493-1496)
                p.fbc_sgn.push_back((signed char)((f < nIF && owner[f] != c) ? 1 : -1));
            }
        }
        p.fbc_beg.push_back((int)p.fbc_idx.size());
        bi.nfb = (int)p.fb_face.size() - bi.fb_off;
        for (int l = 0; l < bi.nl; ++l) local[L[l]] = -1;
        p.blk[b] = bi;
        p.max_nl = std::max(p.max_nl, bi.nl);
        p.max_ng = std::max(p.max_ng, bi.ng);
        p.max_nfa = std::max(p.max_nfa, bi.nfa);
        p.max_nfb = std::max(p.max_nfb, bi.nfb);
        p.max_nc = std::max(p.max_nc, bi.nc);
    }
    int max_st = 0, max_cf = 0;
    for (int n_st : p.gs_cnt) max_st = std::max(max_st, n_st);
    for (size_t i = 0; i + 1 < p.fbc_beg.size(); ++i) max_cf = std::max(max_cf, p.fbc_beg[i + 1] - p.fbc_beg[i]);
    if (max_st > kFusedMaxNb || max_cf > kFusedMaxNb) {
        err = "cells with more than " + std::to_string(kFusedMaxNb) + " faces";
        return false;
    }
    if (p.max_nl > kFusedThreadsA * fused_cells_per_thread_a(g.D) ||
        p.max_nl > kFusedThreadsB * fused_cells_per_thread_b(g.D) ||
        p.max_nfb > kFusedThreadsB * fused_faces_per_thread_b(g.D)) {
        err = "block halos exceed the work-group shape (local cells " + std::to_string(p.max_nl) + ", faces " +
              std::to_string(p.max_nfb) + ")";
        return false;
    }
    if (p.max_nc > kFusedThreadsB || p.max_nfa > kFusedThreadsA * kFusedSlotsA) {
        err = "block tables exceed the work-group shape (cells " + std::to_string(p.max_nc) + ", owned faces " +
              std::to_string(p.max_nfa) + ")";
        return false;
    }
    p.nm = g.D == 3 ? 17 : 11;
    p.nma = p.nm + 1;
    p.wg_target = nvg_hint > 0 ? nvg_hint * nblk : 4096;
    (void)es;
    p.ready = true;
    return true;
}

}  // namespace rc
}  // namespace dugks
