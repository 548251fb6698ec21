/*
 * dugks_foam.h -- C-ABI of libdugks_foam.so: OpenFOAM case directories as dugksFoam reads
 * and writes them, for callers without OpenFOAM.  CPU only; no HIP, OpenFOAM or torch types.
 *
 * Reading replaces what dugksFoam gets from OpenFOAM before `fvDVM dvm(rho, U, T, ...)`
 * (src/dugksFoam.C:48-54: setRootCase.H, createTime.H, createMesh.H, createFields.H,
 * readTimeControlsExplicit.H) and the dictionary reads of the fvDVM constructor
 * (src/fvDVM/fvDVM/fvDVM.C:1928-1948, 62-123):
 *   system/controlDict, constant/polyMesh/{points,faces,owner,neighbour,boundary},
 *   constant/DVMProperties, constant/Xis, constant/weights, <time>/{rho,U,T}
 * in OpenFOAM's ASCII format, each file plain or gzip-compressed (name or name.gz).  Mesh
 * geometry (face area vectors and centres, cell centres and volumes, interpolation weights,
 * deltaCoeffs, solved directions) follows OpenFOAM's primitiveMesh and surfaceInterpolation.
 *
 * Writing replaces the AUTO_WRITE fields of createFields.H and fvDVM (fvDVM.C:2016-2106)
 * and Time's uniform/time, and writeDFonCell (fvDVM.C:1835-1897).
 *
 * Every call returns 0 on success, non-zero otherwise with a message from
 * dugks_foam_last_error(); nothing aborts.  Cells use the polyMesh numbering.
 */
#ifndef DUGKS_FOAM_H
#define DUGKS_FOAM_H

#include <stddef.h>

#include "dugks.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dugks_foam_case dugks_foam_case;

/* Reads the case in case_dir.  time_name NULL or "": the time directory named by
 * controlDict's startFrom (startTime, firstTime, latestTime). */
int dugks_foam_read(const char *case_dir, const char *time_name, dugks_foam_case **out);
/* As dugks_foam_read for dugksFoam built with -DVMESH (src/fvDVM/Make/options:13): the
 * velocity set is the cell centres and volumes of the polyMesh in constant/vMesh, components
 * beyond the mesh's nSolutionD set to 0, and xiMax the largest |xi| (fvDVM.C:129-253);
 * constant/Xis and constant/weights are not read.  macroFlux yes then selects the
 * velocity-moment flux update (dugks_options.macro_flux 2, fvDVM.C:1562-1618). */
int dugks_foam_read_vmesh(const char *case_dir, const char *time_name, dugks_foam_case **out);
void dugks_foam_free(dugks_foam_case *c);
const char *dugks_foam_last_error(void);

/* The case as dugks_create takes it.  Patches: the non-empty patches in polyMesh order
 * (at most patch_cap are written; *n_patches receives the count).  Pointers inside the
 * structs stay valid until dugks_foam_free.  Boundary faces of the DUGKS C-ABI are the faces
 * of these patches in that order.  dugks_options.macro_flux is the scalar "macro_flux". */
int dugks_foam_solver_inputs(const dugks_foam_case *c, dugks_mesh *mesh, dugks_patch *patches, int patch_cap,
                             int *n_patches, dugks_vset *vset, dugks_gas *gas);
/* cell fields of the time directory read: [n_cells], [n_cells][3], [n_cells] */
int dugks_foam_initial_fields(const dugks_foam_case *c, const double **rho, const double **U, const double **T);

/* system/controlDict entries that drive the dugksFoam loop (defaults as in
 * readTimeControlsExplicit.H:39-52) and the time read */
typedef struct dugks_foam_controls {
    double start_time, end_time, delta_t, write_interval, max_co, max_delta_t, converge_tol;
    int adjust_time_step, converge_check_steps, write_precision, time_precision;
    int write_control; /* 0 timeStep, 1 runTime / adjustableRunTime */
    double time;       /* value of the time directory read */
} dugks_foam_controls;
int dugks_foam_controls_get(const dugks_foam_case *c, dugks_foam_controls *ctl);

/* Introspection (used by bindings): named arrays and scalars of the case.
 * double arrays: "points" [nP][3], "Sf", "Cf" [nF][3] (polyMesh faces), "C" [nC][3], "V",
 *   "weights", "delta_coeffs" [nIF], "xis", "xi_weights" [nDV], "rho", "T" [nC], "U" [nC][3],
 *   "bf_rho", "bf_T" [nB], "bf_U" [nB][3] (C-ABI boundary faces), "patch_pressure" [non-empty
 *   patches] (pressureIn / pressureOut), "xi_list" [n][3], "xi_list_weights" [n] (VMESH);
 * int arrays: "face_offsets" [nF+1], "face_points", "owner" [nF], "neighbour" [nIF],
 *   "abi_faces" [nIF + nB] (C-ABI face -> polyMesh face), "df_write_cells";
 * scalars: "n_cells", "n_faces", "n_internal_faces", "n_boundary_faces", "n_solution_d",
 *   "n_patches" (all polyMesh patches), "nDV", "xiMax", "xiMin", "R", "omega", "Tref",
 *   "muRef", "Pr", "KInner", "time", "macro_flux" (dugks_options.macro_flux), "vmesh". */
int dugks_foam_array_d(const dugks_foam_case *c, const char *name, const double **data, size_t *count);
int dugks_foam_array_i(const dugks_foam_case *c, const char *name, const int **data, size_t *count);
int dugks_foam_scalar(const dugks_foam_case *c, const char *name, double *value);
/* gradSchemes default of system/fvSchemes ("leastSquares" or "Gauss linear"; the solver
 * inputs carry the Gauss linear gradient as ls_own / ls_nei vectors) */
const char *dugks_foam_grad_scheme(const dugks_foam_case *c);
/* polyMesh patch i (all patches, empty ones included): name, geometric type, the rho
 * boundary type of the time read, face range; strings live until dugks_foam_free */
int dugks_foam_patch(const dugks_foam_case *c, int i, const char **name, const char **geom_type,
                     const char **rho_type, int *start, int *size);

/* Writes <case>/<time>/{rho, U, T, q, qSurf, stressSurf, qWall, stressWall} and
 * <time>/uniform/time.  Cell fields in the polyMesh numbering; rho_b [nB] the boundary rho
 * of the C-ABI boundary faces (diffuse walls: the wall density); q_surf [nIF + nB][3] in the
 * C-ABI face numbering, or NULL (written as zero).  The time directory is named with
 * controlDict's timeFormat / timePrecision. */
typedef struct dugks_foam_output {
    double time, delta_t, delta_t0;
    int index;
    const double *rho, *U, *T, *q;
    const double *rho_b;
    const double *q_surf;
} dugks_foam_output;
int dugks_foam_write(const dugks_foam_case *c, const dugks_foam_output *o);

/* <time>/DF<cell>: g~ of one cell over all nv velocities (writeDFonCell, fvDVM.C:1835-1897,
 * written into the current time directory under the cell's own label) */
int dugks_foam_write_df(const dugks_foam_case *c, double time, int cell, const double *g, int nv);

#ifdef __cplusplus
}
#endif
#endif
