// Internal declarations shared by the HIP kernels and the C-ABI layer of libdugks_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace dugks {

constexpr double kPi = 3.14159265358979323846;
constexpr double kVSmall = 1.0e-300;   // OpenFOAM VSMALL (double)
constexpr double kSmall = 1.0e-15;     // OpenFOAM SMALL (double)

struct d3 {
    double x, y, z;
};

__host__ __device__ inline double dot(const d3 &a, const d3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// xi . Sf with each product and sum rounded on its own, (x x' + y y') + z z', exactly as
// OpenFOAM's vector inner product evaluates it on the host.  The upwind tests branch on its
// sign (fvDVM.C:829, 840, 950, 1079, 1431), and velocities tangent to a face give
// sigma = +-1e-19 round-off: a fused multiply-add there would flip the branch.
__device__ inline double sigma_dot(const d3 &xi, const d3 &Sf)
{
#pragma clang fp contract(off)
    const double px = xi.x * Sf.x;
    const double py = xi.y * Sf.y;
    const double pz = xi.z * Sf.z;
    return (px + py) + pz;
}
__host__ __device__ inline double magsqr(const d3 &a) { return a.x * a.x + a.y * a.y + a.z * a.z; }

// p[i] read through the constant address space.  For tables no kernel writes while it runs
// (velocity set, storage reference, block records): a wave-uniform index then becomes a
// scalar load (one per wave, held in SGPRs) instead of a vector load per lane, which the
// compiler cannot otherwise prove safe in kernels that also store to global memory.
template <class T>
__device__ inline T cload(const T *p, int i)
{
    static_assert(sizeof(T) % 4 == 0, "cload moves 4-byte words");
    const __attribute__((address_space(4))) unsigned *q = (const __attribute__((address_space(4))) unsigned *)(p + i);
    T r;
    unsigned *d = (unsigned *)&r;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) d[k] = q[k];
    return r;
}

// DF boundary types selected by the rho patch type (discreteVelocity.C:188-195); the
// numbers are the dugks_patch_type codes of include/dugks.h
enum PatchType : int {
    kMaxwellWall = 1,   // calculatedMaxwell: diffuse wall
    kMixed = 2,         // fixedValue: incoming half frozen at the patch Maxwellian
    kZeroGradient = 3,  // every entry takes the owner cell's f-bar-plus
    kFarField = 4,      // incoming entries: patch rho x Maxwellian(owner cell U, patch T)
    kPressureIn = 5,    // farField with rho = p / (R T) set after every step
    kPressureOut = 6,   // farField with rho, T from the outlet characteristic
    kSymmetry = 7,      // DVMsymmetry: incoming entries copy the mirror velocity, whole face relaxed again
    kSymmetryPlane = 8  // symmetryPlane: incoming entries copy the mirror velocity
};
__host__ __device__ inline bool is_far_field(int t) { return t == kFarField || t == kPressureIn || t == kPressureOut; }
__host__ __device__ inline bool is_symmetry(int t) { return t == kSymmetry || t == kSymmetryPlane; }
// boundary types the fused schedules carry (the others run on the staged path)
inline bool fused_patch_type(int t) { return t == kMaxwellWall || t == kMixed; }

// Read-only geometry and constants, passed by value to every kernel.
struct Geo {
    int nC, nIF, nB, nF, D, K, Nv, n1d;
    double R, omega, Tref, muRef, Pr, xi_max, sqrtD;
    int q1;
    const int *owner;        // [nF]
    const int *neighbour;    // [nIF]
    const d3 *Sf, *Cf, *C;   // [nF], [nF], [nC]
    const double *V;         // [nC]
    // cell -> faces (all faces of the cell, ascending face label); sign +1 owner, -1 neighbour
    const int *cf_off, *cf_face;
    // cell -> gradient stencil (internal faces, ascending face label)
    const int *cg_off, *cg_nb;
    const d3 *cg_ls;
    // boundary faces
    const int *bf_type, *bf_local;
    const double *bf_rho, *bf_T;
    const d3 *bf_U;
    const double *in_by_rho; // [nB] inComingByRho (walls)
    const double *bf_p;      // [nB] patch pressure (pressureIn / pressureOut)
    const int *bf_axis;      // [nB] symmetry faces: the axis of the patch normal, else -1
    const d3 *bf_sf0;        // [nB] symmetry faces: Sf of the patch's first face
    const int *mirror;       // [3][Nv] velocity mirrored about each axis
    // internal-face CFL data
    const double *delta_coeffs; // [nIF]
    const double *weights;      // [nIF] linear interpolation weights
    // velocity set
    const d3 *xi;            // [Nv]
    const double *w;         // [Nv]
    // Stored g~, h~ are deviations from a reference Maxwellian per velocity
    // (rho = 1, U = 0, T = Tref; h = (K+3-D) R Tref g): value = stored + ref.  FP32 storage
    // then rounds the deviation, not the full value.  Zero for FP64 storage.
    const double *ref_g, *ref_h;  // [Nv]
    // updateMacroVol: 0 moments of g~, h~; 1 face-macro fluxes; 2 velocity-moment fluxes of the
    // relaxed face DFs, held per face in face_flux [nF][5] (rho, rhoU, rhoE)
    int macro_flux;
    double *face_flux;
};

struct Macros {
    double *rho;
    d3 *U;
    double *T;
    d3 *q;
    double *tau;
};

// Dense device buffers of one solver instance.
struct DevState {
    // geometry storage (owned)
    int *owner = nullptr, *neighbour = nullptr, *cf_off = nullptr, *cf_face = nullptr;
    int *cg_off = nullptr, *cg_nb = nullptr, *bf_type = nullptr, *bf_local = nullptr;
    d3 *Sf = nullptr, *Cf = nullptr, *C = nullptr, *cg_ls = nullptr, *bf_U = nullptr, *xi = nullptr;
    double *V = nullptr, *bf_rho = nullptr, *bf_T = nullptr, *in_by_rho = nullptr, *dc = nullptr, *w = nullptr,
           *weights = nullptr, *ref_g = nullptr, *ref_h = nullptr, *bf_p = nullptr;
    int *bf_axis = nullptr, *mirror = nullptr;
    d3 *bf_sf0 = nullptr;
    double *face_flux = nullptr;  // [nF][5] (macro_flux 2)
    // distributions (storage type chosen at run time; void* + element size)
    void *gT = nullptr, *hT = nullptr, *gB = nullptr, *hB = nullptr, *gS = nullptr, *hS = nullptr;
    // macros
    Macros cell{}, face{};
    double *rho_b = nullptr;      // boundary rho (walls: rho_w)
    double *cfl_part = nullptr;   // [2 * n_cfl_blocks] partial max / sum of deltaCoeffs*(|U|+sqrt(D)xiMax)
    double *dt = nullptr;         // [4]: dt, co_max, co_mean, spare
    double *dt_hist = nullptr;    // [hist_cap]
};

// Host-side geometry preparation (dugks_setup.cpp)
struct HostPrep {
    std::vector<int> cf_off, cf_face, cg_off, cg_nb, bf_type, bf_local;
    std::vector<d3> cg_ls;
    std::vector<double> own_ls, nei_ls; // [nIF*3]
    std::vector<int> owner_h, neighbour_h; // host copies of owner / neighbour (internal numbering)
};

bool compute_ls_vectors(int nC, int nIF, int nB, const int *owner, const int *neighbour, const double *Sf,
                        const double *Cf, const double *C, const double *weights, int first_cell,
                        std::vector<double> &own_ls, std::vector<double> &nei_ls, std::string &err);
bool build_stencils(int nC, int nIF, int nB, const int *owner, const int *neighbour, const double *own_ls,
                    const double *nei_ls, HostPrep &p, std::string &err);

// kernel launchers (dugks_kernels.hip); S = float or double storage
constexpr int kCflBlock = 256;
int cfl_blocks(int nIF);

template <class S>
void launch_init_df(const Geo &g, const Macros &cm, S *gT, S *hT, hipStream_t st);
template <class S>
void launch_init_mixed_faces(const Geo &g, S *gS, S *hS, hipStream_t st);
void launch_init_incoming(const Geo &g, double *in_by_rho, hipStream_t st);
// updatePressureInOutBC (fvDVM.C:1745-1821): patch rho and (outlets) T from the cell state
void launch_pressure_bc(const Geo &g, const Macros &cm, double *rho_b, double *bf_T, hipStream_t st);
void launch_init_usurf(const Geo &g, const Macros &cm, Macros fm, hipStream_t st);
void launch_tau(const Geo &g, Macros m, int n, hipStream_t st);
void launch_cfl_partials(const Geo &g, const Macros &fm, double *part, hipStream_t st);
void launch_dt_update(const Geo &g, const double *part, double *dt, int adjust, double max_co,
                      double max_delta_t, double *hist, int step, hipStream_t st);

template <class S>
struct Staged {
    static void bar(const Geo &g, const Macros &cm, const S *gT, const S *hT, S *gB, S *hB, const double *dt,
                    hipStream_t st);
    static void face_recon(const Geo &g, const S *gB, const S *hB, S *gS, S *hS, const double *dt, hipStream_t st);
    static void wall_rho(const Geo &g, const S *gS, double *rho_b, hipStream_t st);
    static void wall_in(const Geo &g, const double *rho_b, S *gS, S *hS, hipStream_t st);
    // incoming entries of farField faces (fvDVM.C:923-937) and of symmetry faces
    // (discreteVelocity.C:691-775)
    static void far_field_in(const Geo &g, const Macros &cm, const double *rho_b, S *gS, S *hS, hipStream_t st);
    static void symmetry_in(const Geo &g, S *gS, S *hS, hipStream_t st);
    static void face_moments(const Geo &g, const S *gS, const S *hS, Macros fm, const double *dt, double *cfl_part,
                             hipStream_t st);
    static void face_relax(const Geo &g, const Macros &fm, S *gS, S *hS, const double *dt, hipStream_t st);
    static void update(const Geo &g, const S *gB, const S *hB, const S *gS, const S *hS, S *gT, S *hT,
                       const double *dt, hipStream_t st);
    // fluxes of the moments 1, xi, (|xi|^2 g + h)/2 of the relaxed face DFs (macro_flux 2)
    static void face_flux(const Geo &g, const S *gS, const S *hS, hipStream_t st);
    // cell macros; macro_flux 1 reads the face macros fm, 2 the face fluxes of face_flux
    static void cell_moments(const Geo &g, const S *gT, const S *hT, Macros cm, const Macros &fm, const double *dt,
                             hipStream_t st);
};

}  // namespace dugks
