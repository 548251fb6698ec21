// C-ABI of libdugks_foam.so (include/dugks_foam.h) over the case reader / writer.
#include <cstring>
#include <string>

#include "../../include/dugks_foam.h"
#include "foam_case.h"

struct dugks_foam_case {
    dugks_foam::Case c;
    std::vector<double> ls_none;
    std::vector<std::string> rho_types;
};

namespace {
thread_local std::string g_err;
int fail(const std::string &m)
{
    g_err = m;
    return 1;
}
}  // namespace

extern "C" {

const char *dugks_foam_last_error(void) { return g_err.c_str(); }

namespace {
int read_any(const char *case_dir, const char *time_name, dugks_foam_case **out, bool vmesh)
{
    if (!out) return fail("dugks_foam_read: out is NULL");
    *out = nullptr;
    if (!case_dir) return fail("dugks_foam_read: case_dir is NULL");
    auto *h = new dugks_foam_case();
    std::string err;
    if (!dugks_foam::read_case(case_dir, time_name ? time_name : "", h->c, err, vmesh)) {
        delete h;
        return fail(err);
    }
    for (const auto &b : h->c.rho.bc) h->rho_types.push_back(b.type);
    *out = h;
    return 0;
}
}  // namespace

int dugks_foam_read(const char *case_dir, const char *time_name, dugks_foam_case **out)
{
    return read_any(case_dir, time_name, out, false);
}

int dugks_foam_read_vmesh(const char *case_dir, const char *time_name, dugks_foam_case **out)
{
    return read_any(case_dir, time_name, out, true);
}

void dugks_foam_free(dugks_foam_case *c) { delete c; }

int dugks_foam_solver_inputs(const dugks_foam_case *h, dugks_mesh *mesh, dugks_patch *patches, int patch_cap,
                             int *n_patches, dugks_vset *vset, dugks_gas *gas)
{
    if (!h) return fail("dugks_foam_solver_inputs: NULL case");
    const dugks_foam::Case &c = h->c;
    if (mesh) {
        mesh->n_cells = c.nC;
        mesh->n_internal_faces = c.nIF;
        mesh->n_boundary_faces = c.nB;
        mesh->n_solution_d = c.n_solution_d;
        mesh->owner = c.abi_owner.data();
        mesh->neighbour = c.neighbour.data();
        mesh->Sf = c.abi_Sf.data();
        mesh->Cf = c.abi_Cf.data();
        mesh->C = c.C.data();
        mesh->V = c.V.data();
        mesh->weights = c.weights.data();
        mesh->delta_coeffs = c.delta_coeffs.data();
        mesh->ls_own = c.ls_own.empty() ? nullptr : c.ls_own.data();
        mesh->ls_nei = c.ls_nei.empty() ? nullptr : c.ls_nei.data();
    }
    const int np = (int)c.solver_patch.size();
    if (n_patches) *n_patches = np;
    if (patches) {
        for (int i = 0; i < np && i < patch_cap; ++i) {
            const int s = c.patch_start[i];
            patches[i].type = c.patch_code[i];
            patches[i].start = s;
            patches[i].size = c.patch_size[i];
            patches[i].rho = c.bf_rho.data() + s;
            patches[i].U = c.bf_U.data() + 3 * (size_t)s;
            patches[i].T = c.bf_T.data() + s;
            patches[i].pressure = c.patch_pressure[i];
        }
    }
    if (vset) {
        vset->n_per_dim = c.vmesh ? 0 : c.nDV;
        vset->xis = c.vmesh ? nullptr : c.xis.data();
        vset->weights = c.vmesh ? nullptr : c.wts.data();
        vset->xi_max = c.xi_max;
        vset->xi_min = c.xi_min;
        vset->n_list = c.vmesh ? (int)c.vlist_w.size() : 0;
        vset->list_xi = c.vmesh ? c.vlist_xi.data() : nullptr;
        vset->list_w = c.vmesh ? c.vlist_w.data() : nullptr;
    }
    if (gas) {
        gas->R = c.R;
        gas->omega = c.omega;
        gas->Tref = c.Tref;
        gas->muRef = c.muRef;
        gas->Pr = c.Pr;
        gas->KInner = c.KInner;
    }
    return 0;
}

int dugks_foam_initial_fields(const dugks_foam_case *h, const double **rho, const double **U, const double **T)
{
    if (!h) return fail("dugks_foam_initial_fields: NULL case");
    if (rho) *rho = h->c.rho.internal.data();
    if (U) *U = h->c.U.internal.data();
    if (T) *T = h->c.T.internal.data();
    return 0;
}

int dugks_foam_controls_get(const dugks_foam_case *h, dugks_foam_controls *k)
{
    if (!h || !k) return fail("dugks_foam_controls_get: NULL argument");
    const dugks_foam::Controls &c = h->c.ctl;
    k->start_time = c.start_time;
    k->end_time = c.end_time;
    k->delta_t = c.delta_t;
    k->write_interval = c.write_interval;
    k->max_co = c.max_co;
    k->max_delta_t = c.max_delta_t;
    k->converge_tol = c.converge_tol;
    k->adjust_time_step = c.adjust_time_step ? 1 : 0;
    k->converge_check_steps = c.converge_check_steps;
    k->write_precision = c.write_precision;
    k->time_precision = c.time_precision;
    k->write_control = c.write_control == "timeStep" ? 0 : 1;
    k->time = h->c.time;
    return 0;
}

int dugks_foam_array_d(const dugks_foam_case *h, const char *name, const double **data, size_t *count)
{
    if (!h || !name || !data || !count) return fail("dugks_foam_array_d: NULL argument");
    const dugks_foam::Case &c = h->c;
    const std::vector<double> *v = nullptr;
    const std::string n = name;
    if (n == "points") v = &c.points;
    else if (n == "Sf") v = &c.Sf;
    else if (n == "Cf") v = &c.Cf;
    else if (n == "C") v = &c.C;
    else if (n == "V") v = &c.V;
    else if (n == "weights") v = &c.weights;
    else if (n == "delta_coeffs") v = &c.delta_coeffs;
    else if (n == "xis") v = &c.xis;
    else if (n == "xi_weights") v = &c.wts;
    else if (n == "rho") v = &c.rho.internal;
    else if (n == "U") v = &c.U.internal;
    else if (n == "T") v = &c.T.internal;
    else if (n == "bf_rho") v = &c.bf_rho;
    else if (n == "bf_U") v = &c.bf_U;
    else if (n == "bf_T") v = &c.bf_T;
    else if (n == "patch_pressure") v = &c.patch_pressure;
    else if (n == "xi_list") v = &c.vlist_xi;
    else if (n == "xi_list_weights") v = &c.vlist_w;
    else return fail("dugks_foam_array_d: no double array named " + n);
    *data = v->data();
    *count = v->size();
    return 0;
}

int dugks_foam_array_i(const dugks_foam_case *h, const char *name, const int **data, size_t *count)
{
    if (!h || !name || !data || !count) return fail("dugks_foam_array_i: NULL argument");
    const dugks_foam::Case &c = h->c;
    const std::vector<int> *v = nullptr;
    const std::string n = name;
    if (n == "face_offsets") v = &c.face_off;
    else if (n == "face_points") v = &c.face_pts;
    else if (n == "owner") v = &c.owner;
    else if (n == "neighbour") v = &c.neighbour;
    else if (n == "abi_faces") v = &c.abi_face;
    else if (n == "df_write_cells") v = &c.df_write_cells;
    else return fail("dugks_foam_array_i: no int array named " + n);
    *data = v->data();
    *count = v->size();
    return 0;
}

int dugks_foam_scalar(const dugks_foam_case *h, const char *name, double *value)
{
    if (!h || !name || !value) return fail("dugks_foam_scalar: NULL argument");
    const dugks_foam::Case &c = h->c;
    const std::string n = name;
    if (n == "n_cells") *value = c.nC;
    else if (n == "n_faces") *value = c.nF;
    else if (n == "n_internal_faces") *value = c.nIF;
    else if (n == "n_boundary_faces") *value = c.nB;
    else if (n == "n_solution_d") *value = c.n_solution_d;
    else if (n == "n_patches") *value = (double)c.patches.size();
    else if (n == "nDV") *value = c.nDV;
    else if (n == "xiMax") *value = c.xi_max;
    else if (n == "xiMin") *value = c.xi_min;
    else if (n == "R") *value = c.R;
    else if (n == "omega") *value = c.omega;
    else if (n == "Tref") *value = c.Tref;
    else if (n == "muRef") *value = c.muRef;
    else if (n == "Pr") *value = c.Pr;
    else if (n == "KInner") *value = c.KInner;
    else if (n == "time") *value = c.time;
    else if (n == "macro_flux") *value = c.macro_flux_form;
    else if (n == "vmesh") *value = c.vmesh ? 1 : 0;
    else return fail("dugks_foam_scalar: no scalar named " + n);
    return 0;
}

const char *dugks_foam_grad_scheme(const dugks_foam_case *h) { return h ? h->c.grad_scheme.c_str() : ""; }

int dugks_foam_patch(const dugks_foam_case *h, int i, const char **name, const char **geom_type, const char **rho_type,
                     int *start, int *size)
{
    if (!h) return fail("dugks_foam_patch: NULL case");
    const dugks_foam::Case &c = h->c;
    if (i < 0 || i >= (int)c.patches.size()) return fail("dugks_foam_patch: patch index out of range");
    const dugks_foam::PatchInfo &p = c.patches[i];
    if (name) *name = p.name.c_str();
    if (geom_type) *geom_type = p.geom_type.c_str();
    if (rho_type) *rho_type = h->rho_types[i].c_str();
    if (start) *start = p.start;
    if (size) *size = p.size;
    return 0;
}

int dugks_foam_write(const dugks_foam_case *h, const dugks_foam_output *o)
{
    if (!h || !o) return fail("dugks_foam_write: NULL argument");
    if (!o->rho || !o->U || !o->T || !o->q) return fail("dugks_foam_write: rho, U, T and q are required");
    dugks_foam::Output out;
    out.time = o->time;
    out.delta_t = o->delta_t;
    out.delta_t0 = o->delta_t0;
    out.index = o->index;
    out.rho = o->rho;
    out.U = o->U;
    out.T = o->T;
    out.q = o->q;
    out.rho_b = o->rho_b;
    out.q_surf = o->q_surf;
    std::string err;
    if (!dugks_foam::write_time(h->c, out, err)) return fail(err);
    return 0;
}

int dugks_foam_write_df(const dugks_foam_case *h, double time, int cell, const double *g, int nv)
{
    if (!h || !g || nv <= 0) return fail("dugks_foam_write_df: bad arguments");
    if (cell < 0 || cell >= h->c.nC) return fail("dugks_foam_write_df: cell label out of range");
    std::string err;
    if (!dugks_foam::write_df(h->c, time, cell, g, nv, err)) return fail(err);
    return 0;
}

}  // extern "C"
