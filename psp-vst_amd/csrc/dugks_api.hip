// C-ABI of libdugks_hip.so (declared in include/dugks.h).  Owns the device buffers of
// one solver instance, sequences the kernels of one DUGKS step on the handle's stream
// and keeps the time step on the device so that n steps need no host round trip.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "dugks_adjoint.h"
#include "dugks_fused.h"
#include "dugks_fused_rc.h"
#include "dugks_internal.h"

using namespace dugks;

namespace {

thread_local std::string g_err;

int fail(const std::string &m)
{
    g_err = m;
    return 1;
}

#define HIPCHK(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

__global__ void k_set_scalar(double *p, double v) { *p = v; }

// Fused schedules.  Face values (dugks_fused.h): one halo ring, face values stored once
// per velocity by the upwind block, in-place update.  Recompute (dugks_fused_rc.h): two
// halo rings rebuilt by both passes, no face values.  Recomputing the rings costs
// (ring cells / block cells) equilibria per update: about 1.5x in 2-D and 5x in 3-D.
// Measured per step: 256^2 cavity with 64^2 velocities, FP32 storage 15.3 ms recompute vs
// 16.1 ms face values, FP64 storage 21.6 vs 17.4 ms (the recompute schedule's FP64 LDS
// tiles hold half the velocities); 211k-cell sphere 356 vs 506 ms (face values) before
// later tuning.  Face values are used in 3-D and for FP64 storage, recompute otherwise.
enum Schedule { kSchedFaceValues = 1, kSchedRecompute = 2 };

enum Phase {
    kPhBar, kPhFaceRecon, kPhWallRho, kPhWallIn, kPhFaceMom, kPhFaceRelax, kPhUpdate, kPhCellMom, kPhDt,
    kPhFusedFace, kPhFaceMacro, kPhFusedUpdate, kPhCellMacro, kPhFusedBar, kNumPhases
};
const char *kPhaseNames[kNumPhases] = {
    "bar", "face_recon", "wall_rho", "wall_in", "face_moments", "face_relax", "update", "cell_moments", "dt_update",
    "fused_face", "face_macros", "fused_update", "cell_macros", "fused_bar"};

}  // namespace

struct dugks_handle {
    dugks_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    DevState d;
    Geo g{};
    size_t es = 8;                 // bytes per stored distribution value
    bool initialised = false;
    bool fused = false;
    // fused schedule: kSchedFaceValues (dugks_fused.h) or kSchedRecompute (dugks_fused_rc.h)
    int sched = 0;
    FusedPlan plan;                // face-value schedule data
    rc::FusedPlan plan_rc;         // recompute schedule data
    bool has_far_field = false, has_symmetry = false, has_pressure = false;  // boundary types present
    std::vector<double> bf_T0;     // patch temperatures as given (pressure outlets change theirs)
    std::vector<int> perm;         // internal cell -> caller's cell label
    double geo_sum = 0.0;          // order-sensitive checksum of the caller's cell centres / volumes
    std::vector<double> ref_g, ref_h;  // host copies of the storage reference per velocity
    int hist_cap = 0;
    std::vector<void *> allocs;
    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> ev_phase;
    int ev_used = 0;
    double ptime[kNumPhases] = {0};
    int pcalls[kNumPhases] = {0};
};

namespace {

template <class T>
int dalloc(dugks_handle *h, T **p, size_t n)
{
    void *q = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess)
        return fail("hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = static_cast<T *>(q);
    return 0;
}

int dalloc_bytes(dugks_handle *h, void **p, size_t bytes)
{
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 1);
    if (e != hipSuccess)
        return fail("hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    h->allocs.push_back(q);
    *p = q;
    return 0;
}

template <class T>
int upload(T *dst, const T *src, size_t n, hipStream_t st)
{
    if (n == 0) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// ---------------------------------------------------------------- profiling helpers
int prof_begin(dugks_handle *h, int phase, hipEvent_t *stop)
{
    if (!h->prof) return 0;
    if (h->ev_used + 2 > (int)h->ev_pool.size()) {
        // fold what has been recorded so far
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int i = 0; i + 1 < h->ev_used; i += 2) {
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, h->ev_pool[i], h->ev_pool[i + 1]));
            h->ptime[h->ev_phase[i / 2]] += ms * 1e-3;
            h->pcalls[h->ev_phase[i / 2]] += 1;
        }
        h->ev_used = 0;
    }
    hipEvent_t a = h->ev_pool[h->ev_used], b = h->ev_pool[h->ev_used + 1];
    h->ev_phase[h->ev_used / 2] = phase;
    h->ev_used += 2;
    HIPCHK(hipEventRecord(a, h->stream));
    *stop = b;
    return 0;
}

int prof_end(dugks_handle *h, hipEvent_t stop)
{
    if (!h->prof) return 0;
    HIPCHK(hipEventRecord(stop, h->stream));
    return 0;
}

#define PHASE(ph, launch)                                                                              \
    do {                                                                                               \
        hipEvent_t stop_ = nullptr;                                                                    \
        if (prof_begin(h, ph, &stop_)) return 1;                                                       \
        launch;                                                                                        \
        if (prof_end(h, stop_)) return 1;                                                              \
    } while (0)

template <class S>
int staged_step(dugks_handle *h)
{
    const Geo &g = h->g;
    DevState &d = h->d;
    hipStream_t st = h->stream;
    S *gT = (S *)d.gT, *hT = (S *)d.hT, *gB = (S *)d.gB, *hB = (S *)d.hB, *gS = (S *)d.gS, *hS = (S *)d.hS;
    PHASE(kPhBar, Staged<S>::bar(g, d.cell, gT, hT, gB, hB, d.dt, st));
    PHASE(kPhFaceRecon, Staged<S>::face_recon(g, gB, hB, gS, hS, d.dt, st));
    if (h->has_far_field) PHASE(kPhFaceRecon, Staged<S>::far_field_in(g, d.cell, d.rho_b, gS, hS, st));
    PHASE(kPhWallRho, Staged<S>::wall_rho(g, gS, d.rho_b, st));
    PHASE(kPhWallIn, Staged<S>::wall_in(g, d.rho_b, gS, hS, st));
    if (h->has_symmetry) PHASE(kPhWallIn, Staged<S>::symmetry_in(g, gS, hS, st));
    PHASE(kPhFaceMom, Staged<S>::face_moments(g, gS, hS, d.face, d.dt, d.cfl_part, st));
    PHASE(kPhFaceRelax, Staged<S>::face_relax(g, d.face, gS, hS, d.dt, st));
    PHASE(kPhUpdate, Staged<S>::update(g, gB, hB, gS, hS, gT, hT, d.dt, st));
    if (g.macro_flux == 2) PHASE(kPhCellMom, Staged<S>::face_flux(g, gS, hS, st));
    PHASE(kPhCellMom, Staged<S>::cell_moments(g, gT, hT, d.cell, d.face, d.dt, st));
    if (h->has_pressure) launch_pressure_bc(g, d.cell, d.rho_b, d.bf_T, st);  // fvDVM.C:2172
    HIPCHK(hipGetLastError());
    return 0;
}

template <class S>
int fused_step(dugks_handle *h)
{
    const Geo &g = h->g;
    DevState &d = h->d;
    hipStream_t st = h->stream;
    if (h->sched == kSchedRecompute) {
        rc::FusedPlan &p = h->plan_rc;
        PHASE(kPhFusedFace, rc::Fused<S>::face_pass(g, p, (const S *)d.gT, (const S *)d.hT, d.cell, d.dt, st));
        PHASE(kPhFaceMacro, rc::Fused<S>::face_macros(g, p, d.face, d.rho_b, d.dt, d.cfl_part, st));
        PHASE(kPhFusedUpdate, rc::Fused<S>::update_pass(g, p, (const S *)d.gT, (const S *)d.hT, (S *)p.gT2,
                                                        (S *)p.hT2, d.cell, d.face, d.rho_b, d.dt, st));
        std::swap(d.gT, p.gT2);
        std::swap(d.hT, p.hT2);
        PHASE(kPhCellMacro, rc::Fused<S>::cell_macros(g, p, d.cell, d.dt, st));
    } else {
        FusedPlan &p = h->plan;
        if (p.Fb != nullptr) PHASE(kPhFusedBar, Fused<S>::bar_pass(g, p, (const S *)d.gT, (const S *)d.hT, d.dt, st));
        PHASE(kPhFusedFace, Fused<S>::face_pass(g, p, (const S *)d.gT, (const S *)d.hT, d.dt, st));
        PHASE(kPhFaceMacro, Fused<S>::face_macros(g, p, d.face, d.rho_b, d.dt, d.cfl_part, st));
        PHASE(kPhFusedUpdate, Fused<S>::update_pass(g, p, (S *)d.gT, (S *)d.hT, d.rho_b, d.dt, st));
        PHASE(kPhCellMacro, Fused<S>::cell_macros(g, p, d.cell, d.dt, st));
    }
    HIPCHK(hipGetLastError());
    return 0;
}

// cell equilibrium records of the fused schedules after the cell macros changed outside a step
void refresh_cell_coef(dugks_handle *h)
{
    if (!h->fused) return;
    if (h->sched == kSchedRecompute)
        rc::Fused<double>::cell_coef(h->g, h->plan_rc, h->d.cell, h->stream);
    else
        Fused<double>::cell_coef(h->g, h->plan, h->d.cell, h->stream);
}

int do_step(dugks_handle *h)
{
    if (h->fused) return h->es == 4 ? fused_step<float>(h) : fused_step<double>(h);
    return h->es == 4 ? staged_step<float>(h) : staged_step<double>(h);
}

// resolves the per-device address of the allocation-free entry point's step-history array
int resolve_scratch(int device);

int check_handle(dugks_handle *h, bool need_init = true)
{
    if (!h) return fail("null handle");
    if (need_init && !h->initialised) return fail("handle not initialised: call dugks_init_from_macros first");
    HIPCHK(hipSetDevice(h->device));
    return 0;
}

}  // namespace

extern "C" {

const char *dugks_last_error(void) { return g_err.c_str(); }
int dugks_abi_version(void) { return DUGKS_ABI_VERSION; }

int dugks_create(const dugks_mesh *mesh, const dugks_patch *patches, int n_patches, const dugks_vset *vset,
                 const dugks_gas *gas, const dugks_options *opt, dugks_handle **out)
{
    if (!out) return fail("dugks_create: out is NULL");
    *out = nullptr;
    if (!mesh || !vset || !gas) return fail("dugks_create: mesh, vset and gas are required");
    const int nC = mesh->n_cells, nIF = mesh->n_internal_faces, nB = mesh->n_boundary_faces;
    const bool vlist = vset->n_list > 0;  // velocity mesh of the VMESH build
    const int D = mesh->n_solution_d, n1 = vlist ? 0 : vset->n_per_dim;
    if (nC <= 0 || nIF < 0 || nB < 0) return fail("dugks_create: invalid mesh sizes");
    if (D < 1 || D > 3) return fail("dugks_create: n_solution_d must be 1, 2 or 3");
    if (vlist ? (!vset->list_xi || !vset->list_w) : (n1 <= 0 || !vset->xis || !vset->weights))
        return fail("dugks_create: empty velocity set");
    if (!mesh->owner || (nIF > 0 && !mesh->neighbour) || !mesh->Sf || !mesh->Cf || !mesh->C || !mesh->V ||
        (nIF > 0 && (!mesh->weights || !mesh->delta_coeffs)))
        return fail("dugks_create: missing mesh arrays");
    if (n_patches < 0 || (n_patches > 0 && !patches)) return fail("dugks_create: invalid patch list");
    if (!(gas->R > 0) || !(gas->Tref > 0) || !(gas->muRef > 0)) return fail("dugks_create: invalid gas properties");
    dugks_options o{};
    o.precision = DUGKS_FP64;
    o.q1_compat = 1;
    o.device = 0;
    o.fused = -1;
    if (opt) o = *opt;
    if (o.precision != DUGKS_FP64 && o.precision != DUGKS_MIXED) return fail("dugks_create: unknown precision");
    if (o.fused < -1 || o.fused > 4) return fail("dugks_create: options.fused must be -1, 0, 1, 2, 3 or 4");
    if (o.macro_flux < 0 || o.macro_flux > 2) return fail("dugks_create: options.macro_flux must be 0, 1 or 2");
    // velocity lists and the macroFlux update run on the staged path
    if (vlist || o.macro_flux != 0) {
        if (o.fused >= 1)
            return fail("dugks_create: the fused path needs a tensor-product velocity set and macro_flux = 0; "
                        "use fused = 0 or -1 for this case");
        o.fused = 0;
    }

    const int nF = nIF + nB;
    long long nv_ll = vlist ? vset->n_list : 1;
    for (int k = 0; k < D && !vlist; ++k) nv_ll *= n1;
    if (nv_ll > 65535) return fail("dugks_create: more than 65535 discrete velocities");
    const int Nv = (int)nv_ll;

    // boundary-face patch map
    std::vector<int> bf_type(nB, -1), bf_local(nB, 0), bf_axis(nB, -1);
    std::vector<double> bf_rho(nB), bf_T(nB), bf_p(nB, 0.0);
    std::vector<d3> bf_U(nB), bf_sf0(nB, d3{0, 0, 0});
    bool any_sym = false, all_fusable = true;
    for (int p = 0; p < n_patches; ++p) {
        const dugks_patch &pp = patches[p];
        if (pp.type < DUGKS_PATCH_MAXWELL_WALL || pp.type > DUGKS_PATCH_SYMMETRY_PLANE)
            return fail("dugks_create: patch " + std::to_string(p) + " has unknown type " + std::to_string(pp.type));
        if (pp.size < 0 || pp.start < 0 || pp.start + pp.size > nB)
            return fail("dugks_create: patch " + std::to_string(p) + " range outside the boundary faces");
        if (pp.size > 0 && (!pp.rho || !pp.U || !pp.T))
            return fail("dugks_create: patch " + std::to_string(p) + " lacks rho/U/T values");
        if ((pp.type == DUGKS_PATCH_PRESSURE_IN || pp.type == DUGKS_PATCH_PRESSURE_OUT) && !(pp.pressure > 0))
            return fail("dugks_create: pressure patch " + std::to_string(p) + " needs a positive pressure");
        all_fusable = all_fusable && fused_patch_type(pp.type);
        int axis = -1;
        if (is_symmetry(pp.type) && pp.size > 0) {
            // the mirror velocity is taken about the patch's first face normal, which must be a
            // coordinate axis (discreteVelocity.C:731-737, 760-764)
            const double *S0 = mesh->Sf + 3 * (size_t)(nIF + pp.start);
            const double m = std::sqrt(S0[0] * S0[0] + S0[1] * S0[1] + S0[2] * S0[2]);
            for (int k = 0; k < 3; ++k)
                if (std::fabs(std::fabs(S0[k]) / m - 1.0) < 1e-12) axis = k;
            if (axis < 0)
                return fail("dugks_create: symmetry patch " + std::to_string(p) + " is not normal to a coordinate axis");
            any_sym = true;
        }
        for (int i = 0; i < pp.size; ++i) {
            const int b = pp.start + i;
            bf_axis[b] = axis;
            bf_p[b] = pp.pressure;
            if (axis >= 0) {
                const double *S0 = mesh->Sf + 3 * (size_t)(nIF + pp.start);
                bf_sf0[b] = d3{S0[0], S0[1], S0[2]};
            }
        }
        if (o.q1_compat && pp.size > nIF)
            return fail("dugks_create: q1_compat reads the internal-face temperature at patch-local indices; patch " +
                        std::to_string(p) + " is larger than the internal-face count");
        if (o.macro_flux == 1 && pp.size > nIF)
            return fail("dugks_create: macro_flux 1 reads internal-face area vectors at patch-local indices; patch " +
                        std::to_string(p) + " is larger than the internal-face count");
        for (int i = 0; i < pp.size; ++i) {
            const int b = pp.start + i;
            if (bf_type[b] != -1) return fail("dugks_create: boundary face " + std::to_string(b) + " in two patches");
            bf_type[b] = pp.type;
            bf_local[b] = i;
            bf_rho[b] = pp.rho[i];
            bf_U[b] = d3{pp.U[3 * i], pp.U[3 * i + 1], pp.U[3 * i + 2]};
            bf_T[b] = pp.T[i];
            if (!(bf_T[b] > 0)) return fail("dugks_create: non-positive patch temperature");
        }
    }
    for (int b = 0; b < nB; ++b)
        if (bf_type[b] == -1) return fail("dugks_create: boundary face " + std::to_string(b) + " is in no patch");
    if (any_sym && vlist)
        return fail("dugks_create: symmetry patches need a tensor-product velocity set (the velocity mesh of "
                    "the VMESH build has no mirror map)");
    if (any_sym)
        for (int k = 0; k < n1; ++k)
            if (std::fabs(vset->xis[n1 - 1 - k] + vset->xis[k]) > 1e-12 * std::fabs(vset->xis[0]) ||
                vset->weights[n1 - 1 - k] != vset->weights[k])
                return fail("dugks_create: symmetry patches need a velocity set symmetric about 0");
    // zeroGradient, farField, pressure and symmetry faces run on the staged path
    if (!all_fusable) {
        if (o.fused >= 1)
            return fail("dugks_create: the fused path carries calculatedMaxwell and fixedValue patches only; "
                        "use fused = 0 or -1 for this case");
        o.fused = 0;
    }

    for (int f = 0; f < nF; ++f)
        if (mesh->owner[f] < 0 || mesh->owner[f] >= nC) return fail("dugks_create: owner label out of range");
    for (int f = 0; f < nIF; ++f)
        if (mesh->neighbour[f] < 0 || mesh->neighbour[f] >= nC) return fail("dugks_create: neighbour label out of range");

    // Internal cell numbering: recursive coordinate bisection into blocks (contiguous in the
    // internal numbering) whose halo fits the chosen fused schedule's work-group shapes.
    // Faces keep the caller's numbering; every cell-indexed array crossing the C-ABI is
    // permuted on the way in and out.
    const int sched = (o.fused == 2 || o.fused == 4) ? kSchedFaceValues
                      : o.fused == 3                   ? kSchedRecompute
                      : (D == 2 && o.precision == DUGKS_MIXED) ? kSchedRecompute
                                                                : kSchedFaceValues;
    // Blocks of 256 cells (192 for 3-D meshes with FP64 storage, whose LDS tiles then hold
    // twice the velocities per chunk: 311 vs 351 ms per step on the 211k-cell sphere, while
    // FP32 storage gains from 256: 259 vs 279 ms); a block whose halo or face count exceeds
    // the work-group shapes is bisected again (only that block), until every block fits or
    // the rounds run out (the plan build then reports the misfit and the staged path takes
    // over).
    std::vector<int> perm, block_start, inv(nC);
    rcb_partition(nC, mesh->C, (D == 3 && o.precision == DUGKS_FP64) ? 192 : 256, perm, block_start);
    for (int round = 0; o.fused != 0 && round < 8; ++round) {
        std::vector<int> bad;
        const bool fit =
            sched == kSchedRecompute
                ? rc::block_tables_fit(nC, nIF, nF, mesh->owner, mesh->neighbour, D, perm, block_start, &bad)
                : block_tables_fit(nC, nIF, nF, mesh->owner, mesh->neighbour, D, perm, block_start, &bad);
        if (fit) break;
        rcb_refine(mesh->C, bad, perm, block_start);
    }
    for (int i = 0; i < nC; ++i) inv[perm[i]] = i;
    std::vector<int> own_p(nF), nei_p(nIF);
    for (int f = 0; f < nF; ++f) own_p[f] = inv[mesh->owner[f]];
    for (int f = 0; f < nIF; ++f) nei_p[f] = inv[mesh->neighbour[f]];
    std::vector<double> C_p(3 * (size_t)nC), V_p(nC);
    for (int i = 0; i < nC; ++i) {
        for (int k = 0; k < 3; ++k) C_p[3 * i + k] = mesh->C[3 * (size_t)perm[i] + k];
        V_p[i] = mesh->V[perm[i]];
        if (!(V_p[i] > 0)) return fail("dugks_create: non-positive cell volume");
    }

    // least-squares vectors and stencils
    std::string err;
    std::vector<double> own_ls, nei_ls;
    if (mesh->ls_own && mesh->ls_nei) {
        own_ls.assign(mesh->ls_own, mesh->ls_own + 3 * (size_t)nIF);
        nei_ls.assign(mesh->ls_nei, mesh->ls_nei + 3 * (size_t)nIF);
    } else if (!compute_ls_vectors(nC, nIF, nB, own_p.data(), nei_p.data(), mesh->Sf, mesh->Cf, C_p.data(),
                                   mesh->weights, inv[0], own_ls, nei_ls, err)) {
        return fail("dugks_create: " + err);
    }
    HostPrep prep;
    if (!build_stencils(nC, nIF, nB, own_p.data(), nei_p.data(), own_ls.data(), nei_ls.data(), prep, err))
        return fail("dugks_create: " + err);
    prep.owner_h = own_p;
    prep.neighbour_h = nei_p;

    // tensor-product velocity set, ix fastest (fvDVM.C:293-373), or the velocity mesh's cell
    // centres (components beyond D dropped) and volumes (fvDVM.C:170-249)
    std::vector<d3> xi(Nv);
    std::vector<double> w(Nv);
    {
        const double *X = vset->xis, *W = vset->weights;
        int i = 0;
        if (vlist) {
            for (int v = 0; v < Nv; ++v) {
                const double *L = vset->list_xi + 3 * (size_t)v;
                xi[v] = d3{L[0], D >= 2 ? L[1] : 0.0, D == 3 ? L[2] : 0.0};
                w[v] = vset->list_w[v];
                if (!(w[v] > 0) || !std::isfinite(L[0] + L[1] + L[2]))
                    return fail("dugks_create: velocity " + std::to_string(v) + " of the list has a non-positive "
                                "weight or a non-finite component");
            }
        } else if (D == 3) {
            for (int iz = 0; iz < n1; ++iz)
                for (int iy = 0; iy < n1; ++iy)
                    for (int ix = 0; ix < n1; ++ix, ++i) {
                        w[i] = W[iz] * W[iy] * W[ix];
                        xi[i] = d3{X[ix], X[iy], X[iz]};
                    }
        } else if (D == 2) {
            for (int iy = 0; iy < n1; ++iy)
                for (int ix = 0; ix < n1; ++ix, ++i) {
                    w[i] = W[iy] * W[ix] * 1;
                    xi[i] = d3{X[ix], X[iy], 0.0};
                }
        } else {
            for (int ix = 0; ix < n1; ++ix, ++i) {
                w[i] = W[ix] * 1 * 1;
                xi[i] = d3{X[ix], 0.0, 0.0};
            }
        }
    }

    dugks_handle *h = new dugks_handle();
    h->opt = o;
    h->device = o.device;
    h->es = o.precision == DUGKS_MIXED ? 4 : 8;
    h->perm = perm;
    for (int c = 0; c < nC; ++c)  // caller's numbering: a mesh with renumbered cells differs
        h->geo_sum += (1.0 + (c % 97)) * (mesh->V[c] + mesh->C[3 * (size_t)c] + 0.5 * mesh->C[3 * (size_t)c + 1] +
                                          0.25 * mesh->C[3 * (size_t)c + 2]);
    auto bail = [&](int rc) {
        std::string m = g_err;
        dugks_destroy(h);
        g_err = m;
        return rc;
    };
    if (hipSetDevice(h->device) != hipSuccess) return bail(fail("dugks_create: hipSetDevice failed"));
    if (resolve_scratch(h->device)) return bail(1);
    if (hipStreamCreate(&h->stream) != hipSuccess) return bail(fail("dugks_create: hipStreamCreate failed"));
    DevState &d = h->d;
    hipStream_t st = h->stream;
    // geometry
    if (dalloc(h, &d.owner, nF) || dalloc(h, &d.neighbour, nIF) || dalloc(h, &d.Sf, nF) || dalloc(h, &d.Cf, nF) ||
        dalloc(h, &d.C, nC) || dalloc(h, &d.V, nC) || dalloc(h, &d.cf_off, nC + 1) ||
        dalloc(h, &d.cf_face, prep.cf_face.size()) || dalloc(h, &d.cg_off, nC + 1) ||
        dalloc(h, &d.cg_nb, prep.cg_nb.size()) || dalloc(h, &d.cg_ls, prep.cg_ls.size()) ||
        dalloc(h, &d.bf_type, nB) || dalloc(h, &d.bf_local, nB) || dalloc(h, &d.bf_rho, nB) ||
        dalloc(h, &d.bf_T, nB) || dalloc(h, &d.bf_U, nB) || dalloc(h, &d.in_by_rho, nB) || dalloc(h, &d.dc, nIF) ||
        dalloc(h, &d.weights, nIF) || dalloc(h, &d.xi, Nv) || dalloc(h, &d.w, Nv) || dalloc(h, &d.rho_b, nB) ||
        dalloc(h, &d.cfl_part, 2 * (size_t)cfl_blocks(nIF)) || dalloc(h, &d.dt, 4) || dalloc(h, &d.bf_p, nB) ||
        dalloc(h, &d.bf_axis, nB) || dalloc(h, &d.bf_sf0, nB) || dalloc(h, &d.mirror, 3 * (size_t)Nv))
        return bail(1);
    {
        // velocity mirrored about each axis (ix fastest: v = ix + n (iy + n iz))
        std::vector<int> mir(3 * (size_t)Nv);
        for (int v = 0; v < Nv && vlist; ++v) mir[v] = mir[Nv + v] = mir[2 * (size_t)Nv + v] = v;
        for (int v = 0; v < Nv && !vlist; ++v) {
            const int ix = v % n1, iy = (v / n1) % n1, iz = v / (n1 * n1);
            mir[v] = (n1 - 1 - ix) + n1 * (iy + n1 * iz);
            mir[Nv + v] = D >= 2 ? ix + n1 * ((n1 - 1 - iy) + n1 * iz) : v;
            mir[2 * (size_t)Nv + v] = D == 3 ? ix + n1 * (iy + n1 * (n1 - 1 - iz)) : v;
        }
        if (upload(d.mirror, mir.data(), mir.size(), st) || upload(d.bf_p, bf_p.data(), nB, st) ||
            upload(d.bf_axis, bf_axis.data(), nB, st) || upload(d.bf_sf0, bf_sf0.data(), nB, st))
            return bail(1);
    }
    for (int b = 0; b < nB; ++b) {
        h->has_far_field = h->has_far_field || is_far_field(bf_type[b]);
        h->has_symmetry = h->has_symmetry || is_symmetry(bf_type[b]);
        h->has_pressure = h->has_pressure || bf_type[b] == kPressureIn || bf_type[b] == kPressureOut;
    }
    h->bf_T0 = bf_T;
    if (dalloc(h, &d.cell.rho, nC) || dalloc(h, &d.cell.U, nC) || dalloc(h, &d.cell.T, nC) ||
        dalloc(h, &d.cell.q, nC) || dalloc(h, &d.cell.tau, nC) || dalloc(h, &d.face.rho, nF) ||
        dalloc(h, &d.face.U, nF) || dalloc(h, &d.face.T, nF) || dalloc(h, &d.face.q, nF) ||
        dalloc(h, &d.face.tau, nF))
        return bail(1);
    std::vector<d3> Sf(nF), Cf(nF), C(nC);
    std::memcpy(Sf.data(), mesh->Sf, sizeof(d3) * nF);
    std::memcpy(Cf.data(), mesh->Cf, sizeof(d3) * nF);
    std::memcpy(C.data(), C_p.data(), sizeof(d3) * nC);
    if (upload(d.owner, own_p.data(), nF, st) || upload(d.neighbour, nei_p.data(), nIF, st) ||
        upload(d.Sf, Sf.data(), nF, st) || upload(d.Cf, Cf.data(), nF, st) || upload(d.C, C.data(), nC, st) ||
        upload(d.V, V_p.data(), nC, st) || upload(d.cf_off, prep.cf_off.data(), nC + 1, st) ||
        upload(d.cf_face, prep.cf_face.data(), prep.cf_face.size(), st) ||
        upload(d.cg_off, prep.cg_off.data(), nC + 1, st) || upload(d.cg_nb, prep.cg_nb.data(), prep.cg_nb.size(), st) ||
        upload(d.cg_ls, prep.cg_ls.data(), prep.cg_ls.size(), st) || upload(d.bf_type, bf_type.data(), nB, st) ||
        upload(d.bf_local, bf_local.data(), nB, st) || upload(d.bf_rho, bf_rho.data(), nB, st) ||
        upload(d.bf_T, bf_T.data(), nB, st) || upload(d.bf_U, bf_U.data(), nB, st) ||
        upload(d.dc, mesh->delta_coeffs, nIF, st) || upload(d.weights, mesh->weights, nIF, st) ||
        upload(d.xi, xi.data(), Nv, st) || upload(d.w, w.data(), Nv, st) || upload(d.rho_b, bf_rho.data(), nB, st))
        return bail(1);
    // reference Maxwellian of the stored deviations (FP32 storage only)
    {
        std::vector<double> rg(Nv, 0.0), rh(Nv, 0.0);
        if (o.precision == DUGKS_MIXED) {
            const double RT = gas->R * gas->Tref;
            const double A = 1.0 / std::pow(std::sqrt(2.0 * kPi * RT), (double)D);
            for (int v = 0; v < Nv; ++v) {
                rg[v] = A * std::exp(-magsqr(xi[v]) / (2.0 * RT));
                rh[v] = (gas->KInner + 3.0 - D) * RT * rg[v];
            }
        }
        if (dalloc(h, &d.ref_g, Nv) || dalloc(h, &d.ref_h, Nv) || upload(d.ref_g, rg.data(), Nv, st) ||
            upload(d.ref_h, rh.data(), Nv, st))
            return bail(1);
        h->ref_g = rg;
        h->ref_h = rh;
    }

    Geo &g = h->g;
    g.nC = nC; g.nIF = nIF; g.nB = nB; g.nF = nF; g.D = D; g.K = gas->KInner; g.Nv = Nv; g.n1d = n1;
    g.R = gas->R; g.omega = gas->omega; g.Tref = gas->Tref; g.muRef = gas->muRef; g.Pr = gas->Pr;
    g.xi_max = vset->xi_max; g.sqrtD = std::sqrt((double)D); g.q1 = o.q1_compat ? 1 : 0;
    g.owner = d.owner; g.neighbour = d.neighbour; g.Sf = d.Sf; g.Cf = d.Cf; g.C = d.C; g.V = d.V;
    g.cf_off = d.cf_off; g.cf_face = d.cf_face; g.cg_off = d.cg_off; g.cg_nb = d.cg_nb; g.cg_ls = d.cg_ls;
    g.bf_type = d.bf_type; g.bf_local = d.bf_local; g.bf_rho = d.bf_rho; g.bf_T = d.bf_T; g.bf_U = d.bf_U;
    g.in_by_rho = d.in_by_rho; g.delta_coeffs = d.dc; g.weights = d.weights; g.xi = d.xi; g.w = d.w;
    g.ref_g = d.ref_g; g.ref_h = d.ref_h;
    g.bf_p = d.bf_p; g.bf_axis = d.bf_axis; g.bf_sf0 = d.bf_sf0; g.mirror = d.mirror;
    g.macro_flux = o.macro_flux;
    if (o.macro_flux == 2 && dalloc(h, &d.face_flux, 5 * (size_t)nF)) return bail(1);
    g.face_flux = d.face_flux;

    // distributions
    const size_t ncell = (size_t)Nv * nC, nface = (size_t)Nv * nF;
    if (dalloc_bytes(h, &d.gT, ncell * h->es) || dalloc_bytes(h, &d.hT, ncell * h->es)) return bail(1);
    h->fused = o.fused != 0;
    h->sched = sched;
    if (h->fused) {
        auto face_values = [&]() {
            if (!fused_plan_build(h->plan, g, prep, block_start, h->es, 0, err)) return false;
            h->plan.two_pass = o.fused == 4;
            return fused_plan_upload(h->plan, g, h->es, err);
        };
        const bool ok = sched == kSchedRecompute
                            ? rc::fused_plan_build(h->plan_rc, g, prep, block_start, h->es, 0, err) &&
                                  rc::fused_plan_upload(h->plan_rc, g, h->es, err)
                            : face_values();
        if (!ok) {
            if (o.fused >= 1) return bail(fail("dugks_create: fused path unavailable: " + err));
            fused_plan_free(h->plan);
            rc::fused_plan_free(h->plan_rc);
            h->fused = false;
        }
    }
    if (!h->fused) {
        if (dalloc_bytes(h, &d.gB, ncell * h->es) || dalloc_bytes(h, &d.hB, ncell * h->es) ||
            dalloc_bytes(h, &d.gS, nface * h->es) || dalloc_bytes(h, &d.hS, nface * h->es))
            return bail(1);
    }
    // profiling event pool
    h->ev_pool.resize(2 * 1024);
    h->ev_phase.resize(1024);
    for (auto &e : h->ev_pool)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail("dugks_create: hipEventCreate failed"));
    *out = h;
    return 0;
}

int dugks_init_from_macros(dugks_handle *h, const double *rho, const double *U, const double *T)
{
    if (check_handle(h, false)) return 1;
    if (!rho || !U || !T) return fail("dugks_init_from_macros: NULL field");
    const Geo &g = h->g;
    DevState &d = h->d;
    hipStream_t st = h->stream;
    for (int c = 0; c < g.nC; ++c)
        if (!(rho[c] > 0) || !(T[c] > 0)) return fail("dugks_init_from_macros: non-positive rho or T at cell " +
                                                        std::to_string(c));
    std::vector<double> rp(g.nC), Tp(g.nC);
    std::vector<d3> Up(g.nC);
    for (int i = 0; i < g.nC; ++i) {
        const int c = h->perm[i];
        rp[i] = rho[c];
        Tp[i] = T[c];
        Up[i] = d3{U[3 * (size_t)c], U[3 * (size_t)c + 1], U[3 * (size_t)c + 2]};
    }
    if (upload(d.cell.rho, rp.data(), g.nC, st) || upload(d.cell.U, Up.data(), g.nC, st) ||
        upload(d.cell.T, Tp.data(), g.nC, st))
        return 1;
    HIPCHK(hipMemsetAsync(d.cell.q, 0, sizeof(d3) * g.nC, st));
    HIPCHK(hipMemsetAsync(d.face.rho, 0, sizeof(double) * g.nF, st));
    HIPCHK(hipMemsetAsync(d.face.T, 0, sizeof(double) * g.nF, st));
    HIPCHK(hipMemsetAsync(d.face.q, 0, sizeof(d3) * g.nF, st));
    HIPCHK(hipMemsetAsync(d.face.tau, 0, sizeof(double) * g.nF, st));
    launch_tau(g, d.cell, g.nC, st);
    if (h->es == 4)
        launch_init_df<float>(g, d.cell, (float *)d.gT, (float *)d.hT, st);
    else
        launch_init_df<double>(g, d.cell, (double *)d.gT, (double *)d.hT, st);
    launch_init_incoming(g, d.in_by_rho, st);
    if (!h->fused) {
        if (h->es == 4)
            launch_init_mixed_faces<float>(g, (float *)d.gS, (float *)d.hS, st);
        else
            launch_init_mixed_faces<double>(g, (double *)d.gS, (double *)d.hS, st);
    } else {
        std::string err;
        const bool ok = h->sched == kSchedRecompute ? rc::fused_init_boundary(h->plan_rc, g, d.cell, st, err)
                                                    : fused_init_boundary(h->plan, g, d.cell, st, err);
        if (!ok) return fail("dugks_init_from_macros: " + err);
    }
    HIPCHK(hipMemcpyAsync(d.rho_b, g.bf_rho, sizeof(double) * g.nB, hipMemcpyDeviceToDevice, st));
    if (g.nB > 0 && upload(d.bf_T, h->bf_T0.data(), g.nB, st)) return 1;
    // "set initial rho in pressureIn/Out BC" (fvDVM.C:2112-2113)
    if (h->has_pressure) launch_pressure_bc(g, d.cell, d.rho_b, d.bf_T, st);
    launch_init_usurf(g, d.cell, d.face, st);
    launch_cfl_partials(g, d.face, d.cfl_part, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    h->initialised = true;
    return 0;
}

int dugks_co_num(dugks_handle *h, double dt, double *co_max, double *co_mean)
{
    if (check_handle(h)) return 1;
    double res[4];
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->stream, h->d.dt, dt);
    launch_dt_update(h->g, h->d.cfl_part, h->d.dt, 0, 0.0, 0.0, nullptr, 0, h->stream);
    HIPCHK(hipMemcpyAsync(res, h->d.dt, sizeof res, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (co_max) *co_max = res[1];
    if (co_mean) *co_mean = res[2];
    return 0;
}

int dugks_evolution(dugks_handle *h, double dt)
{
    if (check_handle(h)) return 1;
    if (!(dt > 0) || !std::isfinite(dt)) return fail("dugks_evolution: dt must be positive and finite");
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->stream, h->d.dt, dt);
    return do_step(h);
}

}  // extern "C"

namespace {

int advance_check(int n_steps, int adjust, double max_co, double max_delta_t, const double *dt, const char *who)
{
    if (n_steps < 0) return fail(std::string(who) + ": negative step count");
    if (!dt || !(*dt > 0)) return fail(std::string(who) + ": *dt must be a positive time step");
    if (adjust && (!(max_co > 0) || !(max_delta_t > 0)))
        return fail(std::string(who) + ": maxCo and maxDeltaT must be > 0");
    return 0;
}

// The loop of dugks_advance with the device time-step history in `hist` (NULL: none)
int advance_loop(dugks_handle *h, int n_steps, int adjust, double max_co, double max_delta_t, double *dt,
                 double *dt_history, double *hist)
{
    DevState &d = h->d;
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, h->stream, d.dt, *dt);
    for (int i = 0; i < n_steps; ++i) {
        PHASE(kPhDt, launch_dt_update(h->g, d.cfl_part, d.dt, adjust, max_co, max_delta_t, hist, i, h->stream));
        if (do_step(h)) return 1;
    }
    HIPCHK(hipMemcpyAsync(dt, d.dt, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (dt_history && n_steps > 0)
        HIPCHK(hipMemcpyAsync(dt_history, hist, sizeof(double) * n_steps, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// Time-step history of dugks_advance_noalloc: a file-scope device array that every GPU the
// library runs on holds for the life of the process (512 KiB).  In host code its name is
// only a stand-in for the symbol: the address each device holds it at comes from
// hipGetSymbolAddress, called once per device while dugks_create has that device current,
// and only that address is passed to kernels or copies.
constexpr int kNoAllocMaxSteps = 65536;
static __device__ __align__(256) double scratch[kNoAllocMaxSteps];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail("device ordinal beyond the library's table of 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return 0;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(std::string("hipGetSymbolAddress of the step-history array failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)ptr;
    return 0;
}

double *scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace

extern "C" {

int dugks_advance(dugks_handle *h, int n_steps, int adjust, double max_co, double max_delta_t, double *dt,
                  double *dt_history)
{
    if (check_handle(h)) return 1;
    if (advance_check(n_steps, adjust, max_co, max_delta_t, dt, "dugks_advance")) return 1;
    DevState &d = h->d;
    if (dt_history && n_steps > h->hist_cap) {
        if (d.dt_hist) {
            (void)hipFree(d.dt_hist);
            for (auto &p : h->allocs)
                if (p == d.dt_hist) p = nullptr;
        }
        d.dt_hist = nullptr;
        if (dalloc(h, &d.dt_hist, n_steps)) return 1;
        h->hist_cap = n_steps;
    }
    return advance_loop(h, n_steps, adjust, max_co, max_delta_t, dt, dt_history, dt_history ? d.dt_hist : nullptr);
}

int dugks_advance_noalloc(dugks_handle *h, int n_steps, int adjust, double max_co, double max_delta_t, double *dt,
                          double *dt_history)
{
    if (check_handle(h)) return 1;
    if (advance_check(n_steps, adjust, max_co, max_delta_t, dt, "dugks_advance_noalloc")) return 1;
    double *hist = nullptr;
    if (dt_history) {
        if (n_steps > kNoAllocMaxSteps)
            return fail("dugks_advance_noalloc: " + std::to_string(n_steps) + " steps with a time-step history exceed "
                        "its fixed buffer of " + std::to_string(kNoAllocMaxSteps) + " steps; split the call");
        hist = scratch_of(h->device);
        if (!hist) return fail("dugks_advance_noalloc: the step-history array of this device was not resolved");
    }
    return advance_loop(h, n_steps, adjust, max_co, max_delta_t, dt, dt_history, hist);
}

int dugks_synchronize(dugks_handle *h)
{
    if (!h) return fail("null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int get_macros(dugks_handle *h, const Macros &m, int n, double *rho, double *U, double *T, double *q,
                      double *tau, const std::vector<int> *perm)
{
    hipStream_t st = h->stream;
    double *outs[5] = {rho, U, T, q, tau};
    const void *src[5] = {m.rho, m.U, m.T, m.q, m.tau};
    const int width[5] = {1, 3, 1, 3, 1};
    std::vector<double> tmp;
    for (int k = 0; k < 5; ++k) {
        if (!outs[k]) continue;
        const size_t bytes = sizeof(double) * width[k] * (size_t)n;
        if (!perm) {
            HIPCHK(hipMemcpyAsync(outs[k], src[k], bytes, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            continue;
        }
        tmp.resize(width[k] * (size_t)n);
        HIPCHK(hipMemcpyAsync(tmp.data(), src[k], bytes, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < width[k]; ++j) outs[k][(size_t)width[k] * (*perm)[i] + j] = tmp[(size_t)width[k] * i + j];
    }
    return 0;
}

int dugks_get_cell_macros(dugks_handle *h, double *rho, double *U, double *T, double *q, double *tau)
{
    if (check_handle(h)) return 1;
    return get_macros(h, h->d.cell, h->g.nC, rho, U, T, q, tau, &h->perm);
}

int dugks_get_face_macros(dugks_handle *h, double *rho, double *U, double *T, double *q, double *tau)
{
    if (check_handle(h)) return 1;
    return get_macros(h, h->d.face, h->g.nF, rho, U, T, q, tau, nullptr);
}

int dugks_get_boundary_rho(dugks_handle *h, double *rho_b)
{
    if (check_handle(h)) return 1;
    if (h->g.nB == 0) return 0;
    HIPCHK(hipMemcpyAsync(rho_b, h->d.rho_b, sizeof(double) * h->g.nB, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int dugks_get_df(dugks_handle *h, double *g, double *hh)
{
    if (check_handle(h)) return 1;
    const int nC = h->g.nC, Nv = h->g.Nv;
    const size_t n = (size_t)Nv * nC;
    double *outs[2] = {g, hh};
    void *src[2] = {h->d.gT, h->d.hT};
    for (int k = 0; k < 2; ++k) {
        if (!outs[k]) continue;
        std::vector<char> tmp(n * h->es);
        HIPCHK(hipMemcpyAsync(tmp.data(), src[k], n * h->es, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        const std::vector<double> &ref = k ? h->ref_h : h->ref_g;
        for (int v = 0; v < Nv; ++v)
            for (int i = 0; i < nC; ++i) {
                const size_t s = (size_t)v * nC + i, t = (size_t)v * nC + h->perm[i];
                outs[k][t] = h->es == 8 ? ((const double *)tmp.data())[s]
                                        : (double)((const float *)tmp.data())[s] + ref[v];
            }
    }
    return 0;
}

int dugks_set_df(dugks_handle *h, const double *g, const double *hh)
{
    if (check_handle(h)) return 1;
    if (!g || !hh) return fail("dugks_set_df: NULL array");
    const int nC = h->g.nC, Nv = h->g.Nv;
    const size_t n = (size_t)Nv * nC;
    const double *ins[2] = {g, hh};
    void *dst[2] = {h->d.gT, h->d.hT};
    for (int k = 0; k < 2; ++k) {
        std::vector<char> tmp(n * h->es);
        const std::vector<double> &ref = k ? h->ref_h : h->ref_g;
        for (int v = 0; v < Nv; ++v)
            for (int i = 0; i < nC; ++i) {
                const size_t s = (size_t)v * nC + h->perm[i], t = (size_t)v * nC + i;
                if (h->es == 8)
                    ((double *)tmp.data())[t] = ins[k][s];
                else
                    ((float *)tmp.data())[t] = (float)(ins[k][s] - ref[v]);
            }
        HIPCHK(hipMemcpyAsync(dst[k], tmp.data(), n * h->es, hipMemcpyHostToDevice, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

int dugks_set_cell_macros(dugks_handle *h, const double *rho, const double *U, const double *T, const double *q)
{
    if (check_handle(h)) return 1;
    if (!rho || !U || !T || !q) return fail("dugks_set_cell_macros: NULL field");
    const Geo &g = h->g;
    DevState &d = h->d;
    std::vector<double> rp(g.nC), Tp(g.nC);
    std::vector<d3> Up(g.nC), qp(g.nC);
    for (int i = 0; i < g.nC; ++i) {
        const size_t c = h->perm[i];
        rp[i] = rho[c];
        Tp[i] = T[c];
        Up[i] = d3{U[3 * c], U[3 * c + 1], U[3 * c + 2]};
        qp[i] = d3{q[3 * c], q[3 * c + 1], q[3 * c + 2]};
    }
    if (upload(d.cell.rho, rp.data(), g.nC, h->stream) || upload(d.cell.U, Up.data(), g.nC, h->stream) ||
        upload(d.cell.T, Tp.data(), g.nC, h->stream) || upload(d.cell.q, qp.data(), g.nC, h->stream))
        return 1;
    launch_tau(g, d.cell, g.nC, h->stream);
    refresh_cell_coef(h);
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// ------------------------------------------------------------------- checkpoints
namespace {
// Version 3.  The stored distributions are deviations from the per-velocity reference (the
// header carries the sums of that reference), and every cell-indexed array is written in
// the caller's cell order, so a file does not depend on the internal block numbering (which
// differs between the staged path and the two fused schedules); the header's geometry
// checksum, taken over the caller's cell centres and volumes in their order, refuses a mesh
// with the same counts but other cells.
struct CkptHeader {
    char magic[8];
    int version, nC, nF, nB, Nv, es;
    double ref_sum_g, ref_sum_h, geo_sum, reserved;
};
constexpr int kCkptVersion = 3;

int ckpt_io(dugks_handle *h, const char *path, bool save)
{
    if (check_handle(h)) return 1;
    if (!path) return fail("checkpoint: NULL path");
    const Geo &g = h->g;
    DevState &d = h->d;
    CkptHeader hd{};
    std::memcpy(hd.magic, "DUGKSCK", 8);
    hd.version = kCkptVersion; hd.nC = g.nC; hd.nF = g.nF; hd.nB = g.nB; hd.Nv = g.Nv; hd.es = (int)h->es;
    for (int v = 0; v < g.Nv; ++v) {
        hd.ref_sum_g += h->ref_g[v];
        hd.ref_sum_h += h->ref_h[v];
    }
    hd.geo_sum = h->geo_sum;
    if (save) HIPCHK(hipStreamSynchronize(h->stream));
    std::FILE *fp = std::fopen(path, save ? "wb" : "rb");
    if (!fp) return fail(std::string("checkpoint: cannot open ") + path);
    // rows x nC elements of `elem` bytes, cell-indexed (permuted), or plain bytes
    struct Blk { void *p; size_t rows, elem; bool cells; };
    const size_t nC = g.nC;
    std::vector<Blk> blocks = {
        {d.gT, (size_t)g.Nv, h->es, true}, {d.hT, (size_t)g.Nv, h->es, true},
        {d.cell.rho, 1, 8, true}, {d.cell.U, 1, 24, true}, {d.cell.T, 1, 8, true}, {d.cell.q, 1, 24, true},
        {d.cell.tau, 1, 8, true},
        {d.face.rho, 1, 8ull * g.nF, false}, {d.face.U, 1, 24ull * g.nF, false}, {d.face.T, 1, 8ull * g.nF, false},
        {d.face.q, 1, 24ull * g.nF, false}, {d.face.tau, 1, 8ull * g.nF, false}, {d.rho_b, 1, 8ull * g.nB, false},
        {d.cfl_part, 1, 16ull * cfl_blocks(g.nIF), false}, {d.dt, 1, 32, false}};
    int rc = 0;
    if (save) {
        if (std::fwrite(&hd, sizeof hd, 1, fp) != 1) rc = fail("checkpoint: write failed");
    } else {
        CkptHeader in{};
        if (std::fread(&in, sizeof in, 1, fp) != 1 || std::memcmp(in.magic, "DUGKSCK", 8) != 0)
            rc = fail("checkpoint: not a checkpoint file");
        else if (in.version != kCkptVersion)
            rc = fail("checkpoint: file format version " + std::to_string(in.version) + ", this library reads " +
                      std::to_string(kCkptVersion));
        else if (in.nC != hd.nC || in.nF != hd.nF || in.nB != hd.nB || in.Nv != hd.Nv || in.es != hd.es)
            rc = fail("checkpoint: sizes or precision differ from this handle");
        else if (in.ref_sum_g != hd.ref_sum_g || in.ref_sum_h != hd.ref_sum_h)
            rc = fail("checkpoint: storage reference (Tref, R, K) differs from this handle");
        else if (in.geo_sum != hd.geo_sum)
            rc = fail("checkpoint: the mesh (cell centres / volumes in the caller's order) differs from this handle's");
    }
    std::vector<char> buf, perm_buf;
    for (size_t bi = 0; rc == 0 && bi < blocks.size(); ++bi) {
        const Blk &b = blocks[bi];
        if (!b.p) continue;
        if (!b.cells) {  // one plain byte range
            buf.resize(b.elem);
            if (save) {
                if (hipMemcpy(buf.data(), b.p, b.elem, hipMemcpyDeviceToHost) != hipSuccess) rc = fail("checkpoint: copy");
                else if (std::fwrite(buf.data(), 1, b.elem, fp) != b.elem) rc = fail("checkpoint: write failed");
            } else {
                if (std::fread(buf.data(), 1, b.elem, fp) != b.elem) rc = fail("checkpoint: truncated file");
                else if (hipMemcpy(b.p, buf.data(), b.elem, hipMemcpyHostToDevice) != hipSuccess) rc = fail("checkpoint: copy");
            }
            continue;
        }
        // cell-indexed rows, a few rows per transfer (about 64 MiB)
        const size_t row = nC * b.elem;
        const size_t rpc = std::max<size_t>(1, (64ull << 20) / row);
        for (size_t r0 = 0; r0 < b.rows && rc == 0; r0 += rpc) {
            const size_t nr = std::min(rpc, b.rows - r0);
            const size_t n = nr * row;
            char *dev = (char *)b.p + r0 * row;
            buf.resize(n);
            perm_buf.resize(n);
            if (save) {
                if (hipMemcpy(buf.data(), dev, n, hipMemcpyDeviceToHost) != hipSuccess) {
                    rc = fail("checkpoint: copy");
                    break;
                }
                for (size_t r = 0; r < nr; ++r)
                    for (size_t i = 0; i < nC; ++i)
                        std::memcpy(&perm_buf[r * row + (size_t)h->perm[i] * b.elem], &buf[r * row + i * b.elem], b.elem);
                if (std::fwrite(perm_buf.data(), 1, n, fp) != n) rc = fail("checkpoint: write failed");
            } else {
                if (std::fread(perm_buf.data(), 1, n, fp) != n) {
                    rc = fail("checkpoint: truncated file");
                    break;
                }
                for (size_t r = 0; r < nr; ++r)
                    for (size_t i = 0; i < nC; ++i)
                        std::memcpy(&buf[r * row + i * b.elem], &perm_buf[r * row + (size_t)h->perm[i] * b.elem], b.elem);
                if (hipMemcpy(dev, buf.data(), n, hipMemcpyHostToDevice) != hipSuccess) rc = fail("checkpoint: copy");
            }
        }
    }
    std::fclose(fp);
    return rc;
}
}  // namespace

int dugks_adjoint(dugks_handle *h, int n_steps, const double *dts, int checkpoint_every,
                  const dugks_objective *obj, double *J, double *dJ_dmuRef, double *dJ_dTw, double *dJ_dUw,
                  double *dJ_dg0, double *dJ_dh0)
{
    if (check_handle(h)) return 1;
    if (n_steps < 1 || !dts) return fail("dugks_adjoint: need n_steps >= 1 and dts");
    if (checkpoint_every < 1) return fail("dugks_adjoint: checkpoint_every must be >= 1");
    if (h->has_far_field || h->has_symmetry)
        return fail("dugks_adjoint: the reverse kernels carry diffuse-wall and fixedValue patches only");
    if (h->g.macro_flux != 0) return fail("dugks_adjoint: the reverse kernels carry macro_flux = 0 only");
    {
        std::vector<int> bt(h->g.nB);
        if (h->g.nB > 0)
            HIPCHK(hipMemcpy(bt.data(), h->g.bf_type, sizeof(int) * h->g.nB, hipMemcpyDeviceToHost));
        for (int t : bt)
            if (!fused_patch_type(t))
                return fail("dugks_adjoint: the reverse kernels carry diffuse-wall and fixedValue patches only");
    }
    for (int n = 0; n < n_steps; ++n)
        if (!(dts[n] > 0)) return fail("dugks_adjoint: non-positive time step " + std::to_string(n));
    const Geo &g = h->g;
    AdjointIO io;
    io.n_steps = n_steps;
    io.dts = dts;
    io.ckpt = std::min(checkpoint_every, n_steps);
    io.q_weight = obj ? obj->q_weight : nullptr;
    io.mom_weight = obj ? obj->mom_weight : nullptr;
    io.J = J;
    io.dmu = dJ_dmuRef;
    io.dTw = dJ_dTw;
    io.dUw = dJ_dUw;
    std::vector<double> lg, lh;
    if (dJ_dg0) {
        lg.resize((size_t)g.Nv * g.nC);
        io.dg0 = lg.data();
    }
    if (dJ_dh0) {
        lh.resize((size_t)g.Nv * g.nC);
        io.dh0 = lh.data();
    }
    HIPCHK(hipSetDevice(h->device));
    std::string err;
    const int rc = h->es == 4 ? run_adjoint<float>(g, h->d, h->stream, io, err)
                              : run_adjoint<double>(g, h->d, h->stream, io, err);
    if (rc) return fail("dugks_adjoint: " + err);
    refresh_cell_coef(h);
    HIPCHK(hipStreamSynchronize(h->stream));
    // distributions back to the caller's cell order
    for (int which = 0; which < 2; ++which) {
        double *out = which ? dJ_dh0 : dJ_dg0;
        const std::vector<double> &src = which ? lh : lg;
        if (!out) continue;
        for (int v = 0; v < g.Nv; ++v)
            for (int i = 0; i < g.nC; ++i) out[(size_t)v * g.nC + h->perm[i]] = src[(size_t)v * g.nC + i];
    }
    return 0;
}

int dugks_checkpoint_save(dugks_handle *h, const char *path) { return ckpt_io(h, path, true); }
int dugks_checkpoint_load(dugks_handle *h, const char *path)
{
    if (ckpt_io(h, path, false)) return 1;
    refresh_cell_coef(h);
    if (h->has_pressure) launch_pressure_bc(h->g, h->d.cell, h->d.rho_b, h->d.bf_T, h->stream);
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int dugks_num_velocities(const dugks_handle *h) { return h ? h->g.Nv : 0; }

int dugks_path_info(const dugks_handle *h, int *fused, int *block_cells, int *chunk_velocities,
                    int *velocity_groups)
{
    if (!h) return fail("null handle");
    const bool rc_s = h->sched == kSchedRecompute;
    if (fused) *fused = h->fused ? h->sched : 0;
    if (block_cells) *block_cells = !h->fused ? 0 : rc_s ? h->plan_rc.max_nc : h->plan.max_nc;
    if (chunk_velocities) *chunk_velocities = !h->fused ? 0 : rc_s ? h->plan_rc.vc : h->plan.vc;
    if (velocity_groups) *velocity_groups = !h->fused ? 0 : rc_s ? h->plan_rc.nvg : h->plan.nvg;
    return 0;
}

int dugks_velocity_set(const dugks_handle *h, double *xi, double *w)
{
    if (!h) return fail("null handle");
    if (xi && hipMemcpy(xi, h->d.xi, sizeof(d3) * h->g.Nv, hipMemcpyDeviceToHost) != hipSuccess)
        return fail("dugks_velocity_set: copy failed");
    if (w && hipMemcpy(w, h->d.w, sizeof(double) * h->g.Nv, hipMemcpyDeviceToHost) != hipSuccess)
        return fail("dugks_velocity_set: copy failed");
    return 0;
}

int dugks_kernel_count(const dugks_handle *) { return kNumPhases; }
const char *dugks_kernel_name(const dugks_handle *, int i) { return (i >= 0 && i < kNumPhases) ? kPhaseNames[i] : ""; }

int dugks_profile(dugks_handle *h, int enable)
{
    if (!h) return fail("null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    h->prof = enable != 0;
    h->ev_used = 0;
    for (int i = 0; i < kNumPhases; ++i) {
        h->ptime[i] = 0;
        h->pcalls[i] = 0;
    }
    return 0;
}

int dugks_kernel_times(dugks_handle *h, double *seconds, int *calls)
{
    if (!h) return fail("null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int i = 0; i + 1 < h->ev_used; i += 2) {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, h->ev_pool[i], h->ev_pool[i + 1]));
        h->ptime[h->ev_phase[i / 2]] += ms * 1e-3;
        h->pcalls[h->ev_phase[i / 2]] += 1;
    }
    h->ev_used = 0;
    for (int i = 0; i < kNumPhases; ++i) {
        if (seconds) seconds[i] = h->ptime[i];
        if (calls) calls[i] = h->pcalls[i];
    }
    return 0;
}

void dugks_destroy(dugks_handle *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto &e : h->ev_pool)
        if (e) (void)hipEventDestroy(e);
    fused_plan_free(h->plan);
    rc::fused_plan_free(h->plan_rc);
    for (void *p : h->allocs)
        if (p) (void)hipFree(p);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
