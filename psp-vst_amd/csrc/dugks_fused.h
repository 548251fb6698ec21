// Fused DUGKS step: two LDS-blocked passes over the distributions per step.
//
//   pass A (faces): per (cell block, velocity group) work-group, stage f-bar-plus of the
//     block and its halo ring in LDS, form least-squares gradients of the block cells,
//     reconstruct every face of the block's cells from the side whose upwind cell is in
//     the block, write the face value (storage precision) and accumulate FP64 raw velocity
//     moments per (face, side) in registers;
//   face macros: sum the per-group, per-side partials in a fixed order, add the incoming
//     half of the wall / inlet faces, form rho, U, T, tau, q of every face;
//   pass B (update): per (cell block, velocity group): f-bar-plus of the block cells in
//     registers, face values relaxed towards the face equilibrium and turned into fluxes in
//     LDS, cell-wise gather of the fluxes, new g~, h~ written in place, FP64 raw moments of
//     the new distributions per cell;
//   cell macros: sum the cell partials, form rho, U, T, tau, q of every cell.
//
// A face value is written once per velocity, by the block holding its upwind cell; pass B
// reads it for both adjacent cells.  No atomics: every sum has a fixed association, so the
// FP64 path is bitwise reproducible run to run.
//
// 3-D with FP32 storage (bar_pass + the passes): f-bar-plus and u = 4/3 f-bar-plus - 1/3 g~
// of every cell are formed once per step by bar_pass; pass A gathers f-bar-plus instead of
// forming it, pass B updates g~ <- u - dt/V sum(+-flux) with no cell equilibrium.
#pragma once

#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "dugks_internal.h"

namespace dugks {

// work-group shapes of the two passes (the planner checks the block tables against them)
constexpr int kFusedThreadsA = 512;   // pass A threads
constexpr int kFusedSlotsA = 2;       // face slots per pass-A thread
constexpr int kFusedThreadsB = 256;   // pass B threads = max cells per block
constexpr int kFusedMaxNb = 8;        // max internal faces per gradient cell / faces per block cell
constexpr int kFusedXcds = 8;         // accelerator dies; work-group ids are dealt round-robin over them
#ifndef DUGKS_PASS_B_VC
#define DUGKS_PASS_B_VC 4
#endif
constexpr int kFusedVcB = DUGKS_PASS_B_VC;  // max velocities per pass-B chunk (register budget)
// local cells (block + halo ring) per pass-A thread in stage 1; face slots per pass-B
// thread, the capacity the planner checks (pass B takes all but one of them as whole faces
// and deals the (slot, velocity) pairs of the rest over the work-group)
constexpr int fused_cells_per_thread_a(int D) { return D == 2 ? 1 : 2; }
constexpr int fused_faces_per_thread_b(int D) { return D == 2 ? 3 : 4; }

// Equilibrium coefficients of one cell or face, formed once per step by the macro kernels
// so that the velocity loops need no division:
//   A = rho / (2 pi R T)^(D/2), U, invRT = 1 / (R T), RTh = R T (faces on patches: the
//   temperature that multiplies h, see q1_compat), q5 = q / (5 rho (R T)^2),
//   tr = tau (cells) or the face relaxation weight h/(2 tau + h) (faces).
struct alignas(16) EqCoef {
    double A, ux, uy, uz, invRT, RTh, qx, qy, qz, tr;
};

// Incoming half of wall / inlet faces: g = s * A * exp(-|U - xi|^2 invRT / 2), h = g * hfac,
// s = rho_w for diffuse walls and 1 for inlets (A then includes the fixed patch rho).
struct alignas(16) BndCoef {
    double A, ux, uy, uz, invRT, hfac, pad0, pad1;
};

struct BlockInfo {
    int c0, nc;        // block cells: permuted ids [c0, c0 + nc)
    int l_off, nl;     // local cell table: lcell[l_off + l], l < nl (block cells first, then ring 1)
    int gsb_off;       // gs_beg[gsb_off + l], l <= nc: gradient stencil ranges of the block cells
    int fs_off, nfs;   // face slots: the faces of the block cells
    int fbc_off;       // fbc_beg[fbc_off + i], i <= nc: per block cell ranges into fbc_*
    int pad;
};

// face slot flags
constexpr int kSlotOwnIn = 1;   // owner cell in the block
constexpr int kSlotNeiIn = 2;   // neighbour cell in the block (internal faces)

struct FusedPlan {
    bool ready = false;
    bool two_pass = false;             // set between build and upload: no k_bar3 form even where it applies
    int nblk = 0, nvg = 0;             // blocks, velocity groups
    int vc = 0;                        // velocities per LDS chunk
    // Velocities are visited in chunks (x-stripe of vc abscissae, row = (iy, iz)); chunk id
    // ch -> stripe = ch / nrow, row = ch % nrow; a velocity group is a contiguous chunk range.
    int n1 = 0, nrow = 0, nstripe = 0, nchunk = 0, cpg = 0;
    int vc_b = 0, nchunk_b = 0, cpg_b = 0, nvg_b = 0;  // the same for pass B (own chunk size)
    int b8 = 0;                        // blocks per XCD in the work-group mapping
    int wg_target = 6144;              // work-groups per pass the velocity grouping aims at
    int max_nl = 0, max_nfs = 0, max_nc = 0;
    int nm = 0, nma = 0;               // raw moments per cell / per face (+ outgoing wall flux)
    size_t lds_a = 0, lds_b = 0;       // dynamic LDS bytes of the two passes
    // host tables
    std::vector<BlockInfo> blk;
    std::vector<int> lcell, gs_beg, gs_nb, fs_face, fs_col, fs_lo, fs_ln, fbc_beg, fbc_idx;
    std::vector<signed char> fs_flag, fbc_sgn;
    std::vector<d3> gs_ls;
    // device tables
    BlockInfo *d_blk = nullptr;
    int *d_lcell = nullptr, *d_gs_beg = nullptr, *d_gs_nb = nullptr, *d_fs_face = nullptr, *d_fs_col = nullptr,
        *d_fs_lo = nullptr, *d_fs_ln = nullptr, *d_fbc_beg = nullptr, *d_fbc_idx = nullptr;
    signed char *d_fs_flag = nullptr, *d_fbc_sgn = nullptr;
    void *d_gs_ls = nullptr;           // d3 (FP64 path) or float3-as-3-floats (mixed)
    double *fpart = nullptr;           // [nvg][nF][2][nma]: per (face, side) partial moments
    double *cpart = nullptr;           // [nvg][nC][nm]
    double *bin = nullptr;             // [nB][nm] incoming moments (walls: per unit density)
    double *bref = nullptr;            // [nB][nma] outgoing-half moments of the storage reference
    double *cref = nullptr;            // [nma] all-velocity moments of the storage reference
    EqCoef *cc = nullptr, *fc = nullptr;  // equilibrium coefficients of cells / faces
    BndCoef *bc = nullptr;                // incoming-half coefficients of boundary faces
    // face values [Nv][n_cols], (g, h) pairs, one column per face: written by the block holding
    // the upwind cell, read by the passes B of both blocks of a shared face
    void *F = nullptr;
    // 3-D with FP32 storage: f-bar-plus (Fb) and u = 4/3 f-bar-plus - 1/3 g~ (U) of every
    // cell, [Nv][nC] (g, h) pairs, written by bar_pass; pass A reads Fb, pass B reads U instead
    // of g~ (null on the other paths)
    void *U = nullptr, *Fb = nullptr;
    int n_slots = 0, n_cols = 0;         // face slots; face-value columns (one per face with a slot)
    d3 *fs_sf = nullptr;                  // [slots] face area vectors, formed at upload
    EqCoef *fs_coef = nullptr;            // [slots] face records, copied each step
    void *rw = nullptr;                   // [Nv] power-sum weights w xi_x^k, k = 0..3
    std::vector<void *> allocs;
};

// Recursive coordinate bisection of the cells into blocks of at most max_block cells.
// Returns the permutation (new -> old) and the block starts in the new numbering.
void rcb_partition(int nC, const double *C, int max_block, std::vector<int> &perm, std::vector<int> &block_start);
// Bisects each listed block (indices into block_start) once more, in place: the cells of a
// block stay a contiguous range of perm and the block is replaced by its two halves.
void rcb_refine(const double *C, const std::vector<int> &blocks, std::vector<int> &perm,
                std::vector<int> &block_start);
// True when every block with its halo ring fits the work-group shapes of the passes
// (owner / neighbour in the caller's numbering, perm new -> old).  With `bad`, every block
// that does not fit is listed there.
bool block_tables_fit(int nC, int nIF, int nF, const int *owner, const int *neighbour, int D,
                      const std::vector<int> &perm, const std::vector<int> &block_start,
                      std::vector<int> *bad = nullptr);

bool fused_plan_build(FusedPlan &p, const Geo &g, const HostPrep &prep, const std::vector<int> &block_start,
                      size_t es, int nvg_hint, std::string &err);
// velocity chunking: vc velocities per chunk, chunk groups sized for about wg_target work-groups
void fused_velocity_groups(FusedPlan &p, int D, int n1, int vc);
bool fused_plan_upload(FusedPlan &p, const Geo &g, size_t es, std::string &err);
bool fused_init_boundary(FusedPlan &p, const Geo &g, const Macros &cm, hipStream_t st, std::string &err);
void fused_plan_free(FusedPlan &p);

template <class S>
struct Fused {
    // f-bar-plus and u of every cell (3-D with FP32 storage; nothing on the other paths)
    static void bar_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const double *dt, hipStream_t st);
    static void face_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const double *dt, hipStream_t st);
    static void face_macros(const Geo &g, FusedPlan &p, Macros fm, double *rho_b, const double *dt, double *cfl_part,
                            hipStream_t st);
    // g~, h~ are updated in place
    static void update_pass(const Geo &g, FusedPlan &p, S *gT, S *hT, const double *rho_b, const double *dt,
                            hipStream_t st);
    static void cell_macros(const Geo &g, FusedPlan &p, Macros cm, const double *dt, hipStream_t st);
    static void cell_coef(const Geo &g, FusedPlan &p, const Macros &cm, hipStream_t st);
};

}  // namespace dugks
