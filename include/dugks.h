/*
 * dugks.h -- C-ABI of libdugks_hip.so, the MI355X implementation of the DUGKS time step
 * of dugksFoam (Foam::fvDVM).  Plain C types only: no OpenFOAM, HIP or torch types cross
 * this boundary.  Every call returns 0 on success and a non-zero status otherwise; the
 * message is available from dugks_last_error().  Nothing inside the library aborts.
 *
 * One handle = one solver instance on one GPU with its own HIP stream.  Calls on one
 * handle are not thread-safe; separate handles are independent.  Host arrays passed in
 * are copied; the library never keeps a host pointer.
 *
 * Face numbering follows an OpenFOAM surface field: internal faces [0, nIF) in polyMesh
 * order, then the faces of the non-empty boundary patches, patch by patch.  Vectors are
 * stored xyz-interleaved ([n][3]).
 */
#ifndef DUGKS_H
#define DUGKS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DUGKS_ABI_VERSION 3

/* storage precision of the distributions */
enum dugks_precision {
    DUGKS_FP64 = 0,   /* g/h stored and reconstructed in FP64 */
    DUGKS_MIXED = 1   /* g/h stored and reconstructed in FP32; moments, equilibrium,
                         cell update and all macroscopic fields in FP64 */
};

/* rho boundary types (0/rho "type" entries) understood by the path.  The DF face type
 * follows the map of src/fvDVM/discreteVelocity/discreteVelocity.C:188-195. */
enum dugks_patch_type {
    DUGKS_PATCH_MAXWELL_WALL = 1,   /* rho calculatedMaxwell -> DF maxwellWall (diffuse wall) */
    DUGKS_PATCH_MIXED = 2,          /* rho fixedValue        -> DF mixed (frozen incoming) */
    DUGKS_PATCH_ZERO_GRADIENT = 3,  /* rho zeroGradient      -> DF zeroGradient: every entry of
                                       the face takes the owner cell's f-bar-plus (fvDVM.C:887-891) */
    DUGKS_PATCH_FAR_FIELD = 4,      /* rho farField          -> DF farField: incoming entries are
                                       rho_p M(U_owner, T_p) (fvDVM.C:910-940) */
    DUGKS_PATCH_PRESSURE_IN = 5,    /* rho pressureIn        -> DF farField, rho_p = p/(R T_p) set
                                       after every step (fvDVM.C:1758-1784) */
    DUGKS_PATCH_PRESSURE_OUT = 6,   /* rho pressureOut       -> DF farField, rho_p and T_p from the
                                       outlet characteristic after every step (fvDVM.C:1785-1819) */
    DUGKS_PATCH_SYMMETRY = 7,       /* rho symmetryMod       -> DF DVMsymmetry: incoming entries copy
                                       the mirror velocity's face value; the face relaxation then
                                       relaxes outgoing entries and the whole face once more
                                       (discreteVelocity.C:725-749, fvDVM.C:1431-1456) */
    DUGKS_PATCH_SYMMETRY_PLANE = 8  /* geometric symmetryPlane -> DF symmetryPlane: incoming entries
                                       copy the mirror velocity's face value, outgoing entries are
                                       relaxed (discreteVelocity.C:750-770) */
};

/* Mesh geometry as OpenFOAM's fvMesh provides it (mesh.owner(), neighbour(), Sf(), Cf(),
 * C(), V(), weights(), deltaCoeffs()); replaces the fvMesh reference held by
 * Foam::fvDVM (src/fvDVM/fvDVM/fvDVM.H:80-409). */
typedef struct dugks_mesh {
    int n_cells;
    int n_internal_faces;
    int n_boundary_faces;      /* faces of the non-empty patches */
    int n_solution_d;          /* mesh.nSolutionD(): 2 for one-cell-thick meshes, else 3 */
    const int *owner;          /* [nIF + nB] */
    const int *neighbour;      /* [nIF] */
    const double *Sf;          /* [nIF + nB][3] face area vectors (out of the owner) */
    const double *Cf;          /* [nIF + nB][3] face centres */
    const double *C;           /* [n_cells][3] cell centres */
    const double *V;           /* [n_cells] cell volumes */
    const double *weights;     /* [nIF] linear interpolation weights */
    const double *delta_coeffs;/* [nIF] 1/|C_N - C_P| */
    /* optional least-squares vectors (leastSquaresVectors::pVectors()/nVectors());
     * NULL -> computed by the library with the formula of
     * src/fvDVM/zeroBoundaryGrad/zeroBoundaryVectors.C:115-122 plus non-coupled boundary rows */
    const double *ls_own;      /* [nIF][3] or NULL */
    const double *ls_nei;      /* [nIF][3] or NULL */
} dugks_mesh;

/* One non-empty boundary patch (replaces the per-patch types of 0/rho, 0/U, 0/T). */
typedef struct dugks_patch {
    int type;                  /* enum dugks_patch_type */
    int start;                 /* first face, counted within the boundary faces */
    int size;
    const double *rho;         /* [size] patch rho (mixed, farField: fixed value; wall: initial;
                                  pressure patches: initial, then updated) */
    const double *U;           /* [size][3] fixed U (walls, mixed) */
    const double *T;           /* [size] fixed T (walls, mixed, farField, pressureIn; pressureOut:
                                  initial, then updated) */
    double pressure;           /* pressureIn / pressureOut patches: the patch pressure */
} dugks_patch;

/* constant/Xis, constant/weights and the fvDVMparas entries (fvDVM.C:1928-1933), or the
 * velocity mesh of the VMESH build (fvDVM.C:129-253). */
typedef struct dugks_vset {
    int n_per_dim;             /* nDV: tensor-product set of xis x xis (x xis), ix fastest */
    const double *xis;         /* [nDV] */
    const double *weights;     /* [nDV] full quadrature measures */
    double xi_max;             /* xiMax (used by the Courant number only; the VMESH build sets it
                                  to the largest |xi| of the list) */
    double xi_min;             /* xiMin (unused by the step, kept for the record) */
    int n_list;                /* > 0: the velocity set is list_xi / list_w instead (the cell
                                  centres and volumes of constant/vMesh); components of list_xi
                                  beyond n_solution_d are dropped as fvDVM.C:213, 238 do.  Runs
                                  on the staged path; symmetry patches need a tensor set. */
    const double *list_xi;     /* [n_list][3] */
    const double *list_w;      /* [n_list] */
} dugks_vset;

/* gasProperties (fvDVM.C:1943-1948). */
typedef struct dugks_gas {
    double R, omega, Tref, muRef, Pr;
    int KInner;
} dugks_gas;

typedef struct dugks_options {
    int precision;             /* enum dugks_precision */
    int q1_compat;             /* 1: wall-face relaxation of h takes T from the
                                  internal-face Tsurf at the patch-local index, as
                                  fvDVM.C:1367,1413 does; 0: the patch temperature */
    int device;                /* HIP device ordinal */
    int fused;                 /* 1: fused two-pass path (error if unavailable), its
                                     schedule chosen by dimension and precision;
                                  2: fused, stored-face-value schedule required;
                                  3: fused, recompute schedule required;
                                  4: fused, stored-face-value schedule in its two-pass
                                     form, each pass forming f-bar-plus itself (what FP64
                                     storage and 2-D meshes run, and what 3-D FP32 storage
                                     falls back to when the f-bar-plus / u arrays of its
                                     three-kernel form do not fit in device memory);
                                  0: staged path with materialised face distributions;
                                  -1: fused when available, else staged */
    int macro_flux;            /* updateMacroVol (fvDVM.C:1521-1682): 0 macroFlux no: cell
                                  rho, rhoU, rhoE are moments of g~, h~;
                                  1 macroFlux yes: the previous conserved variables minus the
                                    face-macro fluxes, as the default build computes them
                                    (F:1620-1679: no face stress, rho_f|U_f|^2 in the energy
                                    flux, boundary momentum dotted with the internal-face Sf at
                                    the patch-local index);
                                  2 macroFlux yes as the VMESH build computes it: minus the
                                    fluxes of the velocity moments of the relaxed face DFs
                                    (F:1562-1618).
                                  1 and 2 run on the staged path. */
} dugks_options;

typedef struct dugks_handle dugks_handle;

/* Replaces fvDVM::fvDVM (fvDVM.C:1901-2117) up to, but excluding, the initial state.
 * Unknown or unsupported patch types fail here. */
int dugks_create(const dugks_mesh *mesh, const dugks_patch *patches, int n_patches,
                 const dugks_vset *vset, const dugks_gas *gas, const dugks_options *opt,
                 dugks_handle **out);

/* Initial state from cell macros: BGK Maxwellians (fvDVM::initialiseDV, fvDVM.C:414-426),
 * q = 0, tau (fvDVM.C:2115), inComingByRho of the diffuse walls
 * (setCalculatedMaxwellRhoBC, fvDVM.C:578-624), frozen mixed-patch DFs
 * (discreteVelocity::initBoundaryField, discreteVelocity.C:226-258) and
 * Usurf = linear interpolate(U) (fvDVM.C:2116). */
int dugks_init_from_macros(dugks_handle *h, const double *rho, const double *U, const double *T);

/* fvDVM::getCoNum (fvDVM.C:2177-2185) for time step dt; synchronises the stream. */
int dugks_co_num(dugks_handle *h, double dt, double *co_max, double *co_mean);

/* fvDVM::evolution (fvDVM.C:2128-2174): one step with time step dt.  Asynchronous. */
int dugks_evolution(dugks_handle *h, double dt);

/* n iterations of the dugksFoam loop body (dugksFoam.C:63-76): Courant number,
 * setDeltaTvar.H:34-50 time-step update (when adjust != 0), evolution.  The time step
 * lives on the device, so the loop needs no host round trip.  *dt is read as the current
 * time step and receives the last one used.  Asynchronous unless dt_history != NULL. */
int dugks_advance(dugks_handle *h, int n_steps, int adjust, double max_co, double max_delta_t,
                  double *dt, double *dt_history);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * dugks_advance with the same arguments and results, which calls no hipMalloc, hipFree or
 * allocating library routine.  dugks_advance keeps the device time-step history in a
 * handle buffer that it reallocates when a call asks for more steps than before; this
 * variant keeps it in a fixed per-GPU array of 65536 steps (512 KiB, held by each GPU
 * the library runs on until the process ends) and fails with an error when dt_history is
 * requested for more steps.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two handles, two host threads) race on it.  All other device memory the
 * step uses is the handle's, allocated by dugks_create. */
int dugks_advance_noalloc(dugks_handle *h, int n_steps, int adjust, double max_co, double max_delta_t,
                          double *dt, double *dt_history);

int dugks_synchronize(dugks_handle *h);

/* field output (host copies; blocking).  Cell fields [n_cells], face fields [nIF + nB]. */
int dugks_get_cell_macros(dugks_handle *h, double *rho, double *U, double *T, double *q, double *tau);
int dugks_get_face_macros(dugks_handle *h, double *rho, double *U, double *T, double *q, double *tau);
/* boundary rho of the patches (diffuse walls: the wall density of the last step) [nB] */
int dugks_get_boundary_rho(dugks_handle *h, double *rho_b);

/* distributions g~, h~ as FP64 [Nv][n_cells] (writeDFonCells, fvDVM.C:1835-1897,
 * generalised to the full field); set_df also serves checkpoint loads. */
int dugks_get_df(dugks_handle *h, double *g, double *hh);
int dugks_set_df(dugks_handle *h, const double *g, const double *hh);
int dugks_set_cell_macros(dugks_handle *h, const double *rho, const double *U, const double *T,
                          const double *q);

/* Binary checkpoint of DFs + macros + wall state (exact resume).  Cell-indexed arrays are
 * written in the caller's cell order, so a file saved by one handle loads into any handle of
 * the same mesh, velocity set, storage precision and gas reference, whatever path or
 * schedule it runs.  Layout: a 64-byte header (magic "DUGKSCK", version 3, nC, nF, nB, Nv,
 * bytes per stored value, the reference sums and a checksum of the caller's cell centres
 * and volumes), then g~ and h~ [Nv][nC] (FP64 values, or FP32 deviations from the
 * reference Maxwellian with MIXED storage), cell rho, U, T, q, tau, face rho, U, T, q, tau,
 * boundary rho, Courant partials and the device time-step record.  A load refuses a file of
 * another version, size, precision, reference or mesh. */
int dugks_checkpoint_save(dugks_handle *h, const char *path);
int dugks_checkpoint_load(dugks_handle *h, const char *path);

/* Objective of the adjoint, evaluated on the last step's face values:
 *   J = sum_f q_f . q_weight_f  +  sum_f sum_v w_v (xi_v . mom_weight_f) (xi_v . S_f) g_f,v
 * with q_f the face heat flux (qSurf, fvDVM.C:1300) and g_f,v the face distribution after
 * the face relaxation (the value the update uses, fvDVM.C:1362-1460): the first term gives
 * wall heat flux, the second the momentum flux through faces (force on a wall = drag).
 * Arrays are [nIF + nB][3] in the caller's face numbering; either may be NULL. */
typedef struct dugks_objective {
    const double *q_weight;
    const double *mom_weight;
} dugks_objective;

/* No counterpart in the reference (it has no adjoint).  Hand-written reverse of n_steps
 * DUGKS steps with the given time steps dts[n_steps] (held fixed: the gradient is that of
 * the fixed-dt map), starting from the handle's current state and leaving it advanced by
 * n_steps.  The forward keeps g~, h~ every checkpoint_every steps and recomputes the steps
 * in between during the reverse sweep.  Outputs (any may be NULL):
 *   J                 objective value;
 *   dJ_dmuRef         derivative with respect to the reference viscosity muRef (Kn);
 *   dJ_dTw, dJ_dUw    [nB], [nB][3]: with respect to each diffuse-wall face's temperature
 *                     and velocity (they enter the incoming Maxwellian and inComingByRho);
 *   dJ_dg0, dJ_dh0    [Nv][n_cells]: with respect to the initial g~, h~ (initial cell macros
 *                     held fixed).
 * With every derivative output NULL only the forward runs (objective evaluation).
 * Adjoint arithmetic is FP64; the forward storage follows the handle's precision. */
int dugks_adjoint(dugks_handle *h, int n_steps, const double *dts, int checkpoint_every,
                  const dugks_objective *obj, double *J, double *dJ_dmuRef, double *dJ_dTw,
                  double *dJ_dUw, double *dJ_dg0, double *dJ_dh0);

/* sizes */
int dugks_num_velocities(const dugks_handle *h);
int dugks_velocity_set(const dugks_handle *h, double *xi, double *w);  /* [Nv][3], [Nv] */

/* Diagnostics (no reference counterpart): the path a handle runs -- 0 staged, 1 fused with
 * stored face values (3-D meshes, FP64 storage), 2 fused recomputing two halo rings (2-D
 * meshes with FP32 storage) -- and,
 * for the fused paths, the largest cell block, velocities per LDS chunk and velocity
 * groups per block.  Any pointer may be NULL. */
int dugks_path_info(const dugks_handle *h, int *fused, int *block_cells, int *chunk_velocities,
                    int *velocity_groups);

/* per-kernel timing of the last evolution (HIP events), seconds; names via index */
int dugks_kernel_count(const dugks_handle *h);
const char *dugks_kernel_name(const dugks_handle *h, int i);
int dugks_profile(dugks_handle *h, int enable);
int dugks_kernel_times(dugks_handle *h, double *seconds, int *calls);

const char *dugks_last_error(void);
int dugks_abi_version(void);
void dugks_destroy(dugks_handle *h);

#ifdef __cplusplus
}
#endif
#endif
