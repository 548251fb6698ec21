// Fused DUGKS step, recompute schedule: two LDS-blocked passes per step that both rebuild
// f-bar-plus of the block and two halo rings (no face values stored).  Chosen where the
// halo is small against the block (2-D meshes); dugks_fused.h is the schedule for 3-D.
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
namespace rc {

// work-group shapes of the two passes (the planner checks the block tables against them)
// pass A threads: 256 (4 waves) lets two work-groups share a CU at pass A's ~240 VGPRs, so
// one work-group's barriers overlap the other's work (512 threads: one per CU, 12 % slower)
constexpr int kFusedThreadsA = 256;
constexpr int kFusedSlotsA = 3;       // owned faces per pass-A thread
constexpr int kFusedThreadsB = 256;   // pass B threads = max cells per block
constexpr int kFusedMaxNb = 8;        // max internal faces per gradient cell / faces per block cell
constexpr int kFusedXcds = 8;         // accelerator dies; work-group ids are dealt round-robin over them
// local cells per thread in stage 1 (pass A / pass B) and block faces per thread in pass B,
// the capacity the planner checks (pass B takes all but one of them as whole faces and
// deals the (face, velocity) pairs of the rest over the work-group)
constexpr int fused_cells_per_thread_a(int D) { return D == 2 ? 2 : 4; }
constexpr int fused_cells_per_thread_b(int D) { return D == 2 ? 2 : 4; }
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

// Geometry of one pass-B face slot: area vector and the arms from the owner and neighbour
// centres to the face centre (neighbour arm = owner arm on boundary faces).
struct FaceGeo {
    d3 Sf, ao, an;
};

struct BlockInfo {
    int c0, nc;        // block cells: permuted ids [c0, c0 + nc)
    int l_off, nl;     // local cell table: lcell[l_off + l], l < nl (block, ring 1, ring 2)
    int ng;            // local cells l < ng (block + ring 1) get a gradient
    int gsb_off;       // gradient cell gsb_off + l, l < ng: gs_cnt[.] neighbours at gs_nb / gs_ls
                       // [(gsb_off + l) * kFusedMaxNb + j]
    int fa_off, nfa;   // faces owned by the block (owner in block), pass A
    int fb_off, nfb;   // faces of the block cells, pass B
    int fbc_off;       // fbc_beg[fbc_off + i], i <= nc: per block cell ranges into fbc_*
    int pad;
};

struct FusedPlan {
    bool ready = false;
    int nblk = 0, nvg = 0;             // blocks, velocity groups
    int vc = 0;                        // velocities per LDS chunk
    // Velocities are visited in chunks (x-stripe of vc abscissae, row = (iy, iz)); chunk id
    // ch -> stripe = ch / nrow, row = ch % nrow; a velocity group is a contiguous chunk range.
    int n1 = 0, nrow = 0, nstripe = 0, nchunk = 0, cpg = 0;
    int b8 = 0;                        // blocks per XCD in the work-group mapping
    int wg_target = 4096;              // work-groups per pass the velocity grouping aims at
    int max_nl = 0, max_ng = 0, max_nfa = 0, max_nfb = 0, max_nc = 0;
    int nm = 0, nma = 0;               // raw moments per cell / per face (+ outgoing wall flux)
    int nbuf_a = 1;                    // pass-A LDS tile buffers (1 or 2)
    size_t lds_a = 0, lds_b = 0;       // dynamic LDS bytes of the two passes
    // host tables
    std::vector<BlockInfo> blk;
    std::vector<int> lcell, gs_cnt, gs_nb, fa_face, fa_own, fa_nei, fb_face, fb_own, fb_nei, fbc_beg, fbc_idx;
    std::vector<signed char> fbc_sgn;
    std::vector<d3> gs_ls;
    // device tables
    BlockInfo *d_blk = nullptr;
    int *d_lcell = nullptr, *d_gs_cnt = nullptr, *d_gs_nb = nullptr, *d_fa_face = nullptr, *d_fa_own = nullptr,
        *d_fa_nei = nullptr, *d_fb_face = nullptr, *d_fb_own = nullptr, *d_fb_nei = nullptr, *d_fbc_beg = nullptr,
        *d_fbc_idx = nullptr;
    signed char *d_fbc_sgn = nullptr;
    void *d_gs_ls = nullptr;           // d3 (FP64 path) or float3-as-3-floats (mixed)
    double *fpart = nullptr;           // [nvg][nF][nma]
    double *cpart = nullptr;           // [nvg][nC][nm]
    double *bin = nullptr;             // [nB][nm] incoming moments (walls: per unit density)
    double *bref = nullptr;            // [nB][nma] outgoing-half moments of the storage reference
    double *cref = nullptr;            // [nma] all-velocity moments of the storage reference
    EqCoef *cc = nullptr, *fc = nullptr;  // equilibrium coefficients of cells / faces
    BndCoef *bc = nullptr;                // incoming-half coefficients of boundary faces
    void *gT2 = nullptr, *hT2 = nullptr;  // second distribution buffer
    FaceGeo *fb_geo = nullptr;            // [pass-B face slots] geometry, formed at upload
    EqCoef *fb_coef = nullptr;            // [pass-B face slots] face records, copied each step
    std::vector<void *> allocs;
};

// True when every block with its two halo rings fits the work-group shapes of the passes
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
    static void face_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, const Macros &cm, const double *dt,
                          hipStream_t st);
    static void face_macros(const Geo &g, FusedPlan &p, Macros fm, double *rho_b, const double *dt, double *cfl_part,
                            hipStream_t st);
    static void update_pass(const Geo &g, FusedPlan &p, const S *gT, const S *hT, S *gN, S *hN, const Macros &cm,
                            const Macros &fm, const double *rho_b, const double *dt, hipStream_t st);
    static void cell_macros(const Geo &g, FusedPlan &p, Macros cm, const double *dt, hipStream_t st);
    static void cell_coef(const Geo &g, FusedPlan &p, const Macros &cm, hipStream_t st);
};

}  // namespace rc
}  // namespace dugks
