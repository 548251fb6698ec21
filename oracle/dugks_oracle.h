/*
 * FP64 CPU oracle for the DUGKS time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of dugksFoam's fvDVM::evolution()
 * (src/fvDVM/fvDVM/fvDVM.C:2128-2174) and the members it calls, written from the
 * reference sources.  It is the checker for the HIP path and the CPU baseline.  Only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg may load it; the
 * product library never links or calls it.
 *
 * Layout mirrors the reference: distributions are velocity-major [v][cell] and
 * [v][face] (src/fvDVM/fvDVM/fvDVM.C:403-408).  Faces are numbered internal faces first,
 * then the non-empty boundary faces patch by patch, as in an OpenFOAM surface field.
 */
#ifndef DUGKS_ORACLE_H
#define DUGKS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* rho boundary types (0/rho) that the path supports; the DF face type follows the map
 * of src/fvDVM/discreteVelocity/discreteVelocity.C:188-195. */
enum { ORC_PATCH_MAXWELL_WALL = 1,   /* rho calculatedMaxwell -> DF maxwellWall  */
       ORC_PATCH_MIXED = 2,          /* rho fixedValue        -> DF mixed        */
       ORC_PATCH_ZERO_GRADIENT = 3,  /* rho zeroGradient      -> DF zeroGradient */
       ORC_PATCH_FAR_FIELD = 4,      /* rho farField          -> DF farField     */
       ORC_PATCH_PRESSURE_IN = 5,    /* rho pressureIn        -> DF farField     */
       ORC_PATCH_PRESSURE_OUT = 6,   /* rho pressureOut       -> DF farField     */
       ORC_PATCH_SYMMETRY = 7,       /* rho symmetryMod       -> DF DVMsymmetry  */
       ORC_PATCH_SYMMETRY_PLANE = 8  /* geometric symmetryPlane -> DF symmetryPlane */
};

typedef struct orc_case {
    int nC, nIF, nB, D;
    const int *owner;        /* [nIF+nB] */
    const int *neighbour;    /* [nIF] */
    const double *Sf, *Cf;   /* [nIF+nB][3] */
    const double *C, *V;     /* [nC][3], [nC] */
    const double *weights;   /* [nIF] surface interpolation weights */
    const double *delta_coeffs; /* [nIF] */
    int n_patches;
    const int *patch_type, *patch_start, *patch_size; /* start/size within boundary faces */
    const double *bf_rho, *bf_U, *bf_T;  /* boundary-face values [nB], [nB][3], [nB] */
    int n1d;
    const double *xis, *wts; /* 1-D set */
    double xi_max;           /* DVMProperties xiMax (CFL only) */
    double R, omega, Tref, muRef, Pr;
    int K;                   /* KInner */
    int q1_compat;           /* replicate the Tsurf internal-field read of fvDVM.C:1413 */
    const double *patch_pressure; /* [n_patches] pressureIn / pressureOut values (may be NULL) */
    const double *ls_own, *ls_nei; /* [nIF][3] gradient vectors (NULL: least squares) */
    /* n_list > 0: the velocity set is this list instead of the tensor product of xis
     * (the VMESH build: cell centres and volumes of constant/vMesh, fvDVM.C:129-253);
     * components beyond D are dropped as there.  Symmetry patches need the tensor set. */
    int n_list;
    const double *list_xi;   /* [n_list][3] */
    const double *list_w;    /* [n_list] */
    /* updateMacroVol (fvDVM.C:1521-1682): 0 macroFlux no (moments of g~, h~); 1 macroFlux yes
     * in the default build (fluxes of the face macros, F:1620-1679); 2 macroFlux yes in the
     * VMESH build (fluxes of the velocity moments of the relaxed face DFs, F:1562-1618) */
    int macro_flux;
} orc_case;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_case *cs);
void orc_destroy(orc_state *s);
const char *orc_last_error(void);
int orc_set_threads(int n);

/* initial state: BGK Maxwellians from cell macros (fvDVM.C:414-426), q = 0, tau,
 * inComingByRho (fvDVM.C:578-624), frozen mixed-patch DFs (discreteVelocity.C:244-257),
 * Usurf = linear interpolate(U) (fvDVM.C:2116). */
int orc_init(orc_state *s, const double *rho, const double *U, const double *T);

int orc_co_num(const orc_state *s, double dt, double *co_max, double *co_mean);
int orc_evolution(orc_state *s, double dt);

/* The sub-steps of evolution in order (F:2128-2174): 0 updateGHbarPvol, 1 updateGHbarSurf
 * (gradients, reconstruction, wall outGoing), 2 updateMaxwellWallRho, 3
 * updateGHbarSurfMaxwellWallIn, 4 updateMacroSurf, 5 updateGHsurf, 6 updateGHtildeVol,
 * 7 updateMacroVol.  orc_evolution_stages runs sub-steps [first, last) so that tests can
 * read intermediate states (face values before relaxation, f-bar-plus, ...);
 * orc_evolution(s, dt) is orc_evolution_stages(s, dt, 0, 8). */
#define ORC_N_STAGES 8
int orc_evolution_stages(orc_state *s, double dt, int first, int last);

/* getters (host copies) */
int orc_nv(const orc_state *s);
int orc_get_cell_macros(const orc_state *s, double *rho, double *U, double *T, double *q, double *tau);
int orc_get_face_macros(const orc_state *s, double *rho, double *U, double *T, double *q, double *tau);
int orc_get_df(const orc_state *s, double *g, double *h);        /* [Nv][nC] */
int orc_set_df(orc_state *s, const double *g, const double *h);
int orc_set_cell_macros(orc_state *s, const double *rho, const double *U, const double *T, const double *q);
int orc_get_bar(const orc_state *s, double *g, double *h);       /* last f-bar-plus [Nv][nC] */
int orc_get_surf(const orc_state *s, double *g, double *h);      /* last face DFs [Nv][nIF+nB] */
int orc_get_wall_rho(const orc_state *s, double *rho_b);         /* [nB] (walls: rho_w) */
int orc_get_incoming_by_rho(const orc_state *s, double *in_b);   /* [nB] */
int orc_get_ls_vectors(const orc_state *s, double *own_ls, double *nei_ls, double *patch_ls);
/* velocity set as built (tensor product, ix fastest) */
int orc_get_vset(const orc_state *s, double *xi, double *w);

#ifdef __cplusplus
}
#endif
#endif
