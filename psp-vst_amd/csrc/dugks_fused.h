// Fused DUGKS step: two LDS-blocked gather passes over the distributions per step.
//
//   pass A (face moments): per (cell block, velocity group) work-group, stage f-bar-plus
//     of the block and its two halo rings in LDS, form least-squares gradients of the block
//     and the first ring, reconstruct every face the block owns at x_f - xi dt/2 and
//     accumulate FP64 raw velocity moments per face in registers;
//   face macros: sum the per-group partials in a fixed order, add the incoming half of the
//     wall / inlet faces, form rho, U, T, tau, q of every face;
//   pass B (update): same staging, face fluxes relaxed towards the face equilibrium in
//     LDS, cell-wise gather of the fluxes, new g~, h~ written to the second buffer, FP64
//     raw moments of the new distributions per cell;
//   cell macros: sum the cell partials, form rho, U, T, tau, q of every cell.
//
// No atomics: every sum has a fixed association, so the FP64 path is bitwise
// reproducible run to run.
#pragma once

#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "dugks_internal.h"

namespace dugks {

// work-group shapes of the two passes (the planner checks the block tables against them)
constexpr int kFusedThreadsA = 512;   // pass A threads
constexpr int kFusedSlotsA = 2;       // owned faces per pass-A thread
constexpr int kFusedThreadsB = 256;   // pass B threads = max cells per block

struct BlockInfo {
    int c0, nc;        // block cells: permuted ids [c0, c0 + nc)
    int l_off, nl;     // local cell table: lcell[l_off + l], l < nl (block, ring 1, ring 2)
    int ng;            // local cells l < ng (block + ring 1) get a gradient
    int gsb_off;       // gs_beg[gsb_off + l], l <= ng: stencil ranges (absolute)
    int fa_off, nfa;   // faces owned by the block (owner in block), pass A
    int fb_off, nfb;   // faces of the block cells, pass B
    int fbc_off;       // fbc_beg[fbc_off + i], i <= nc: per block cell ranges into fbc_*
    int pad;
};

struct FusedPlan {
    bool ready = false;
    int nblk = 0, nvg = 0, vper = 0;   // blocks, velocity groups, velocities per group
    int vc = 0;                        // velocities per LDS chunk
    int max_nl = 0, max_ng = 0, max_nfa = 0, max_nfb = 0, max_nc = 0;
    int nm = 0, nma = 0;               // raw moments per cell / per face (+ outgoing wall flux)
    size_t lds_a = 0, lds_b = 0;       // dynamic LDS bytes of the two passes
    // host tables
    std::vector<BlockInfo> blk;
    std::vector<int> lcell, gs_beg, gs_nb, fa_face, fa_own, fa_nei, fb_face, fb_own, fb_nei, fbc_beg, fbc_idx;
    std::vector<signed char> fbc_sgn;
    std::vector<d3> gs_ls;
    // device tables
    BlockInfo *d_blk = nullptr;
    int *d_lcell = nullptr, *d_gs_beg = nullptr, *d_gs_nb = nullptr, *d_fa_face = nullptr, *d_fa_own = nullptr,
        *d_fa_nei = nullptr, *d_fb_face = nullptr, *d_fb_own = nullptr, *d_fb_nei = nullptr, *d_fbc_beg = nullptr,
        *d_fbc_idx = nullptr;
    signed char *d_fbc_sgn = nullptr;
    void *d_gs_ls = nullptr;           // d3 (FP64 path) or float3-as-3-floats (mixed)
    double *fpart = nullptr;           // [nvg][nF][nma]
    double *cpart = nullptr;           // [nvg][nC][nm]
    double *bin = nullptr;             // [nB][nm] incoming moments (walls: per unit density)
    double *cA = nullptr, *fA = nullptr;  // rho / (2 pi R T)^(D/2) of cells / faces
    void *gT2 = nullptr, *hT2 = nullptr;  // second distribution buffer
    std::vector<void *> allocs;
};

// Recursive coordinate bisection of the cells into blocks of at most max_block cells.
// Returns the permutation (new -> old) and the block starts in the new numbering.
void rcb_partition(int nC, const double *C, int max_block, std::vector<int> &perm, std::vector<int> &block_start);

bool fused_plan_build(FusedPlan &p, const Geo &g, const HostPrep &prep, const std::vector<int> &block_start,
                      size_t es, int nvg_hint, std::string &err);
bool fused_plan_upload(FusedPlan &p, const Geo &g, size_t es, std::string &err);
bool fused_init_boundary(FusedPlan &p, const Geo &g, const Macros &cm, hipStream_t st, std::string &err);
void fused_plan_free(FusedPlan &p);

template <class S>
struct Fused {
    static void face_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const Macros &cm, const double *dt,
                          hipStream_t st);
    static void face_macros(const Geo &g, FusedPlan &p, Macros fm, double *rho_b, const double *dt, double *cfl_part,
                            hipStream_t st);
    static void update_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, S *gN, S *hN, const Macros &cm,
                            const Macros &fm, const double *rho_b, const double *dt, hipStream_t st);
    static void cell_macros(const Geo &g, FusedPlan &p, Macros cm, const double *dt, hipStream_t st);
    static void cell_coef(const Geo &g, FusedPlan &p, const Macros &cm, hipStream_t st);
};

}  // namespace dugks
