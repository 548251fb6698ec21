// An OpenFOAM case directory as dugksFoam reads and writes it, without OpenFOAM.
//
// Reading (what createTime.H, createMesh.H, createFields.H, readTimeControlsExplicit.H and
// the fvDVM constructor read; src/dugksFoam.C:48-54, src/fvDVM/fvDVM/fvDVM.C:1901-1948,
// 62-123):
//   constant/polyMesh/{points,faces,owner,neighbour,boundary}  (ASCII, plain or .gz)
//   constant/DVMProperties   fvDVMparas {nDV, xiMax, xiMin, macroFlux}, gasProperties
//                            {R, omega, Tref, muRef, Pr, KInner}, DFwriteCellList
//   constant/Xis, constant/weights, or (VMESH build) constant/vMesh's cell centres and volumes
//   system/controlDict       startFrom/startTime, endTime, deltaT, writeControl,
//                            writeInterval, writePrecision, timePrecision, adjustTimeStep,
//                            maxCo, maxDeltaT, convergeTol, convergeCheckSteps
//   system/fvSchemes         gradSchemes default: leastSquares or Gauss linear
//   <time>/{rho,U,T}         internalField and boundaryField (patch type and value)
// Geometry follows OpenFOAM's primitiveMesh (face centres and area vectors by triangle
// fans about the face's point average; cell centres and volumes by face pyramids about the
// average of the cell's face centres) and surfaceInterpolation (weights, deltaCoeffs).
//
// Writing (fvDVM.C:2016-2106, the AUTO_WRITE fields, and Time::write): <time>/{rho, U, T,
// q, qSurf, stressSurf, qWall, stressWall}, <time>/uniform/time, and DF dumps
// <time>/DF<cell> of writeDFonCells (fvDVM.C:1835-1897).
#pragma once

#include <string>
#include <vector>

#include "foam_dict.h"

namespace dugks_foam {

struct PatchInfo {
    std::string name, geom_type;  // polyMesh boundary: type (wall, patch, empty, symmetryPlane ...)
    int start = 0, size = 0;      // face range in the polyMesh numbering
};

struct FieldBC {
    std::string type;             // boundaryField type
    std::vector<double> value;    // patch values [size * width] (empty when the patch has none)
    double pressure = 0.0;        // pressureIn / pressureOut entry of the patch (rho)
};

struct VolField {
    int width = 1;
    std::string dims;             // dimension set as written, e.g. "[1 -3 0 0 0 0 0]"
    std::vector<double> internal; // [nC * width]
    std::vector<FieldBC> bc;      // one per polyMesh patch
};

struct Controls {
    std::string start_from = "startTime", write_control = "timeStep", time_format = "general";
    double start_time = 0.0, end_time = 0.0, delta_t = 0.0, write_interval = 1.0;
    double max_co = 1.0, max_delta_t = 1.0e300, converge_tol = 1.0e-9;
    bool adjust_time_step = false;
    int converge_check_steps = 1, write_precision = 6, time_precision = 6;
};

// rho boundary type -> DF face type code of the DUGKS C-ABI (discreteVelocity.C:188-195);
// 0 for an empty patch, -1 for a type the path does not carry
int df_patch_code(const std::string &rho_type);

struct Case {
    std::string dir, time_name;   // case directory and the time directory read
    double time = 0.0;
    // polyMesh
    std::vector<double> points;               // [nP][3]
    std::vector<int> face_off, face_pts;      // CSR faces
    std::vector<int> owner, neighbour;
    int nC = 0, nIF = 0, nF = 0;
    std::vector<PatchInfo> patches;
    // geometry (polyMesh numbering)
    std::vector<double> Sf, Cf, C, V, weights, delta_coeffs;
    int n_solution_d = 3;
    // C-ABI arrays: internal faces, then the faces of the non-empty patches patch by patch
    int nB = 0;
    std::vector<int> abi_face;                // C-ABI face -> polyMesh face
    std::vector<int> abi_owner;
    std::vector<double> abi_Sf, abi_Cf;
    std::vector<int> solver_patch;            // non-empty patch -> polyMesh patch index
    std::vector<int> patch_code, patch_start, patch_size;  // per non-empty patch (start within boundary faces)
    std::vector<double> patch_pressure;       // per non-empty patch (pressureIn / pressureOut)
    std::vector<double> bf_rho, bf_U, bf_T;   // [nB], [nB][3], [nB]
    // DVMProperties, velocity set
    int nDV = 0;
    double xi_max = 0.0, xi_min = 0.0;
    double R = 0.0, omega = 0.0, Tref = 0.0, muRef = 0.0, Pr = 0.0;
    int KInner = 0;
    std::string macro_flux = "no";
    int macro_flux_form = 0;                    // dugks_options.macro_flux: 0 no, 1 yes, 2 yes (VMESH)
    bool vmesh = false;                         // velocity set from constant/vMesh (VMESH build)
    std::vector<double> vlist_xi, vlist_w;      // VMESH velocity list [n][3], [n]
    std::string grad_scheme = "leastSquares";   // system/fvSchemes gradSchemes default
    std::vector<double> ls_own, ls_nei;         // Gauss linear: gradient vectors [nIF][3]
    std::vector<int> df_write_cells;
    std::vector<double> xis, wts;
    // fields of the time directory read
    VolField rho, U, T;
    Controls ctl;
};

// Reads a case; `time_name` empty: the time named by controlDict's startFrom (startTime,
// firstTime or latestTime).  vmesh: the velocity set is the cell centres and volumes of
// constant/vMesh, as dugksFoam built with -DVMESH reads it (fvDVM.C:129-253).
bool read_case(const std::string &dir, const std::string &time_name, Case &c, std::string &err, bool vmesh = false);

// points, faces, owner, neighbour and boundary of a polyMesh directory (no geometry)
bool read_poly_mesh(const std::string &pm, Case &c, std::string &err);

// Geometry of a polyMesh already in `c` (points, faces, owner, neighbour, patches).
bool compute_geometry(Case &c, std::string &err);

// OpenFOAM's time name of a value under controlDict's timeFormat/timePrecision.
std::string time_name(const Controls &ctl, double t);

// Output of one time (polyMesh numbering for cells; C-ABI face numbering for faces).
struct Output {
    double time = 0.0, delta_t = 0.0, delta_t0 = 0.0;
    int index = 0;
    const double *rho = nullptr, *U = nullptr, *T = nullptr, *q = nullptr;  // cells [nC], [nC][3]
    const double *rho_b = nullptr;   // boundary rho [nB] (diffuse walls: wall density)
    const double *q_surf = nullptr;  // faces [nIF + nB][3] (qSurf)
};
bool write_time(const Case &c, const Output &o, std::string &err);
// DF dump of one cell: <time>/DF<label> as a scalarIOList of Nv values
bool write_df(const Case &c, double t, int label, const double *g, int nv, std::string &err);

}  // namespace dugks_foam
