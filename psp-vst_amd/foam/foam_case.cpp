// OpenFOAM case reading, mesh geometry and field writing (see foam_case.h).
#include "foam_case.h"

#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>

namespace dugks_foam {

namespace {

constexpr double kRootVSmall = 1.0e-150;
constexpr double kVSmall = 1.0e-300;

std::string join(const std::string &a, const std::string &b) { return a.empty() ? b : a + "/" + b; }

bool is_number_name(const std::string &s, double &v)
{
    if (s.empty()) return false;
    char *end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0';
}

// numeric subdirectories of the case (time directories), sorted by value
std::vector<std::pair<double, std::string>> time_dirs(const std::string &dir)
{
    std::vector<std::pair<double, std::string>> out;
    DIR *d = opendir(dir.c_str());
    if (!d) return out;
    while (dirent *e = readdir(d)) {
        double v;
        const std::string n = e->d_name;
        if (is_number_name(n, v)) {
            struct stat st;
            if (stat(join(dir, n).c_str(), &st) == 0 && S_ISDIR(st.st_mode)) out.emplace_back(v, n);
        }
    }
    closedir(d);
    std::sort(out.begin(), out.end());
    return out;
}

bool mkdirs(const std::string &path, std::string &err)
{
    std::string cur;
    size_t i = 0;
    while (i <= path.size()) {
        const size_t j = path.find('/', i);
        const std::string part = path.substr(i, j == std::string::npos ? std::string::npos : j - i);
        cur = cur.empty() ? (path[0] == '/' ? "/" + part : part) : join(cur, part);
        if (!part.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) {
            err = "cannot create directory " + cur;
            return false;
        }
        if (j == std::string::npos) break;
        i = j + 1;
    }
    return true;
}

const Entry *need(const Dict &d, const std::string &key, const std::string &where, std::string &err)
{
    const Entry *e = d.find(key);
    if (!e) err = where + ": keyword '" + key + "' is missing";
    return e;
}

bool read_scalar_list_file(const std::string &path, std::vector<double> &v, std::string &err)
{
    FoamFile f;
    std::unique_ptr<Lexer> lx;
    if (!open_foam_list(path, f, lx, err)) return false;
    if (!lex_scalar_list(*lx, v, err)) {
        err = path + ": " + err;
        return false;
    }
    return true;
}

// `uniform v`, `nonuniform List<T> N (...)`: n entries of `width` numbers
bool field_value(const std::vector<Token> &t, int width, size_t n, std::vector<double> &out, const std::string &what,
                 std::string &err)
{
    if (t.empty() || t[0].kind != Token::kWord) {
        err = what + ": value is neither uniform nor nonuniform";
        return false;
    }
    if (t[0].text == "uniform") {
        double v[3] = {0, 0, 0};
        if (width == 1) {
            if (t.size() != 2 || t[1].kind != Token::kNumber) {
                err = what + ": bad uniform scalar";
                return false;
            }
            v[0] = t[1].num;
        } else {
            if (t.size() != 6 || t[1].kind != Token::kPunct || t[1].punct != '(' || t[2].kind != Token::kNumber ||
                t[3].kind != Token::kNumber || t[4].kind != Token::kNumber || t[5].kind != Token::kPunct ||
                t[5].punct != ')') {
                err = what + ": bad uniform vector";
                return false;
            }
            v[0] = t[2].num;
            v[1] = t[3].num;
            v[2] = t[4].num;
        }
        out.resize(n * width);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < width; ++k) out[i * width + k] = v[k];
        return true;
    }
    if (t[0].text == "nonuniform") {
        size_t k = 1;
        if (k < t.size() && t[k].kind == Token::kWord) ++k;  // List<scalar> / List<vector>
        if (!(width == 1 ? tok_scalar_list(t, k, out, err) : tok_vector_list(t, k, out, err))) {
            err = what + ": " + err;
            return false;
        }
        if (out.size() != n * width) {
            err = what + ": holds " + std::to_string(out.size() / width) + " values, expected " + std::to_string(n);
            return false;
        }
        return true;
    }
    err = what + ": value is neither uniform nor nonuniform";
    return false;
}

bool read_vol_field(const Case &c, const std::string &name, int width, VolField &fld, std::string &err)
{
    const std::string path = join(join(c.dir, c.time_name), name);
    FoamFile f;
    if (!read_foam_dict(path, f, err)) return false;
    fld.width = width;
    if (const Entry *dm = f.body.find("dimensions")) {
        fld.dims = "[";
        for (const Token &t : dm->tokens)
            if (t.kind == Token::kNumber) fld.dims += (fld.dims.size() > 1 ? " " : "") + t.text;
        fld.dims += "]";
    }
    const Entry *in = need(f.body, "internalField", path, err);
    if (!in || !field_value(in->tokens, width, (size_t)c.nC, fld.internal, path + " internalField", err)) return false;
    const Dict *bf = f.body.sub("boundaryField");
    if (!bf) {
        err = path + ": boundaryField is missing";
        return false;
    }
    fld.bc.assign(c.patches.size(), FieldBC{});
    for (size_t p = 0; p < c.patches.size(); ++p) {
        const PatchInfo &pi = c.patches[p];
        const Dict *pd = bf->sub(pi.name);
        if (!pd) {
            if (pi.geom_type == "empty") {
                fld.bc[p].type = "empty";
                continue;
            }
            err = path + ": no boundaryField entry for patch " + pi.name;
            return false;
        }
        const Entry *ty = need(*pd, "type", path + " patch " + pi.name, err);
        if (!ty || !entry_word(*ty, fld.bc[p].type, err, "type")) return false;
        if (const Entry *val = pd->find("value"))
            if (!field_value(val->tokens, width, (size_t)pi.size, fld.bc[p].value, path + " patch " + pi.name, err))
                return false;
        for (const char *key : {"pressureIn", "pressureOut"})
            if (const Entry *pe = pd->find(key))
                if (!entry_scalar(*pe, fld.bc[p].pressure, err, key)) return false;
    }
    return true;
}

void cross(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
double dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

}  // namespace

int df_patch_code(const std::string &t)
{
    if (t == "calculatedMaxwell") return 1;  // maxwellWall
    if (t == "fixedValue") return 2;         // mixed: incoming half frozen at the patch Maxwellian
    if (t == "zeroGradient") return 3;
    if (t == "farField") return 4;
    if (t == "pressureIn") return 5;         // farField with the inlet pressure
    if (t == "pressureOut") return 6;        // farField with the outlet pressure
    if (t == "symmetryMod") return 7;        // DVMsymmetry
    if (t == "symmetryPlane") return 8;      // geometric symmetry plane
    if (t == "empty") return 0;
    return -1;
}

std::string time_name(const Controls &ctl, double t)
{
    char buf[64];
    const int p = std::max(1, ctl.time_precision);
    if (ctl.time_format == "fixed")
        std::snprintf(buf, sizeof buf, "%.*f", p, t);
    else if (ctl.time_format == "scientific")
        std::snprintf(buf, sizeof buf, "%.*e", p, t);
    else
        std::snprintf(buf, sizeof buf, "%.*g", p, t);
    return buf;
}

bool compute_geometry(Case &c, std::string &err)
{
    const int nF = c.nF, nC = c.nC, nIF = c.nIF;
    c.Sf.assign(3 * (size_t)nF, 0.0);
    c.Cf.assign(3 * (size_t)nF, 0.0);
    const double *P = c.points.data();
    const int nP = (int)(c.points.size() / 3);
    for (int f = 0; f < nF; ++f) {
        const int b = c.face_off[f], n = c.face_off[f + 1] - b;
        for (int j = 0; j < n; ++j)
            if (c.face_pts[b + j] >= nP) {
                err = "face " + std::to_string(f) + " references point " + std::to_string(c.face_pts[b + j]) +
                      " beyond the " + std::to_string(nP) + " points";
                return false;
            }
        double *sf = &c.Sf[3 * (size_t)f], *cf = &c.Cf[3 * (size_t)f];
        if (n == 3) {
            const double *p0 = P + 3 * c.face_pts[b], *p1 = P + 3 * c.face_pts[b + 1], *p2 = P + 3 * c.face_pts[b + 2];
            double e1[3], e2[3], nn[3];
            for (int k = 0; k < 3; ++k) {
                cf[k] = (1.0 / 3.0) * (p0[k] + p1[k] + p2[k]);
                e1[k] = p1[k] - p0[k];
                e2[k] = p2[k] - p0[k];
            }
            cross(e1, e2, nn);
            for (int k = 0; k < 3; ++k) sf[k] = 0.5 * nn[k];
            continue;
        }
        double fc[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j)
            for (int k = 0; k < 3; ++k) fc[k] += P[3 * c.face_pts[b + j] + k];
        for (int k = 0; k < 3; ++k) fc[k] /= n;
        double sumN[3] = {0, 0, 0}, sumA = 0.0, sumAc[3] = {0, 0, 0};
        for (int j = 0; j < n; ++j) {
            const double *p = P + 3 * c.face_pts[b + j], *q = P + 3 * c.face_pts[b + (j + 1) % n];
            double cc[3], e1[3], e2[3], nn[3];
            for (int k = 0; k < 3; ++k) {
                cc[k] = p[k] + q[k] + fc[k];
                e1[k] = q[k] - p[k];
                e2[k] = fc[k] - p[k];
            }
            cross(e1, e2, nn);
            const double a = std::sqrt(dot(nn, nn));
            for (int k = 0; k < 3; ++k) {
                sumN[k] += nn[k];
                sumAc[k] += a * cc[k];
            }
            sumA += a;
        }
        if (sumA < kRootVSmall) {
            for (int k = 0; k < 3; ++k) {
                cf[k] = fc[k];
                sf[k] = 0.0;
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                cf[k] = (1.0 / 3.0) * sumAc[k] / sumA;
                sf[k] = 0.5 * sumN[k];
            }
        }
    }
    std::vector<double> cEst(3 * (size_t)nC, 0.0), nfc(nC, 0.0);
    for (int f = 0; f < nF; ++f) {
        const int o = c.owner[f];
        for (int k = 0; k < 3; ++k) cEst[3 * o + k] += c.Cf[3 * f + k];
        nfc[o] += 1.0;
        if (f < nIF) {
            const int ne = c.neighbour[f];
            for (int k = 0; k < 3; ++k) cEst[3 * ne + k] += c.Cf[3 * f + k];
            nfc[ne] += 1.0;
        }
    }
    for (int i = 0; i < nC; ++i) {
        if (nfc[i] == 0) {
            err = "cell " + std::to_string(i) + " has no faces";
            return false;
        }
        for (int k = 0; k < 3; ++k) cEst[3 * i + k] /= nfc[i];
    }
    std::vector<double> ctr(3 * (size_t)nC, 0.0), vol(nC, 0.0);
    for (int f = 0; f < nF; ++f) {
        const double *sf = &c.Sf[3 * (size_t)f], *cf = &c.Cf[3 * (size_t)f];
        const int o = c.owner[f];
        double d[3];
        for (int k = 0; k < 3; ++k) d[k] = cf[k] - cEst[3 * o + k];
        const double py = dot(sf, d);
        for (int k = 0; k < 3; ++k) ctr[3 * o + k] += py * ((3.0 / 4.0) * cf[k] + (1.0 / 4.0) * cEst[3 * o + k]);
        vol[o] += py;
        if (f < nIF) {
            const int ne = c.neighbour[f];
            for (int k = 0; k < 3; ++k) d[k] = cEst[3 * ne + k] - cf[k];
            const double pn = dot(sf, d);
            for (int k = 0; k < 3; ++k) ctr[3 * ne + k] += pn * ((3.0 / 4.0) * cf[k] + (1.0 / 4.0) * cEst[3 * ne + k]);
            vol[ne] += pn;
        }
    }
    c.C.assign(3 * (size_t)nC, 0.0);
    c.V.assign(nC, 0.0);
    for (int i = 0; i < nC; ++i) {
        const bool ok = std::fabs(vol[i]) > kVSmall;
        for (int k = 0; k < 3; ++k) c.C[3 * i + k] = ok ? ctr[3 * i + k] / vol[i] : cEst[3 * i + k];
        c.V[i] = vol[i] * (1.0 / 3.0);
        if (!(c.V[i] > 0)) {
            err = "cell " + std::to_string(i) + " has a non-positive volume (faces oriented wrongly?)";
            return false;
        }
    }
    c.weights.assign(nIF, 0.0);
    c.delta_coeffs.assign(nIF, 0.0);
    for (int f = 0; f < nIF; ++f) {
        const int o = c.owner[f], ne = c.neighbour[f];
        const double *sf = &c.Sf[3 * (size_t)f], *cf = &c.Cf[3 * (size_t)f];
        double a[3], b[3], d[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = cf[k] - c.C[3 * o + k];
            b[k] = c.C[3 * ne + k] - cf[k];
            d[k] = c.C[3 * ne + k] - c.C[3 * o + k];
        }
        const double so = std::fabs(dot(sf, a)), sn = std::fabs(dot(sf, b));
        c.weights[f] = sn / (so + sn);
        c.delta_coeffs[f] = 1.0 / std::sqrt(dot(d, d));
    }
    // solved directions: directions spanned only by empty patches are not solved
    // (polyMesh::solutionD for empty patches normal to a coordinate axis)
    std::set<int> empty_dirs;
    for (const PatchInfo &p : c.patches) {
        if (p.geom_type != "empty" || p.size == 0) continue;
        double s[3] = {0, 0, 0};
        for (int f = p.start; f < p.start + p.size; ++f)
            for (int k = 0; k < 3; ++k) s[k] += std::fabs(c.Sf[3 * (size_t)f + k]);
        empty_dirs.insert((int)(std::max_element(s, s + 3) - s));
    }
    c.n_solution_d = 3 - (int)empty_dirs.size();
    return true;
}

// points, faces, owner, neighbour and boundary of the polyMesh directory pm into c
bool read_poly_mesh(const std::string &pm, Case &c, std::string &err)
{
    {
        FoamFile f;
        std::unique_ptr<Lexer> lx;
        if (!open_foam_list(join(pm, "points"), f, lx, err) || !lex_vector_list(*lx, c.points, err)) {
            err = "points: " + err;
            return false;
        }
    }
    {
        FoamFile f;
        std::unique_ptr<Lexer> lx;
        if (!open_foam_list(join(pm, "faces"), f, lx, err)) return false;
        if (f.cls == "faceCompactList") {
            std::vector<int> off;
            if (!lex_label_list(*lx, off, err) || !lex_label_list(*lx, c.face_pts, err)) {
                err = "faces: " + err;
                return false;
            }
            c.face_off = off;
        } else if (!lex_face_list(*lx, c.face_off, c.face_pts, err)) {
            err = "faces: " + err;
            return false;
        }
    }
    {
        FoamFile f;
        std::unique_ptr<Lexer> lx;
        if (!open_foam_list(join(pm, "owner"), f, lx, err) || !lex_label_list(*lx, c.owner, err)) {
            err = "owner: " + err;
            return false;
        }
        FoamFile g;
        std::unique_ptr<Lexer> ly;
        if (!open_foam_list(join(pm, "neighbour"), g, ly, err) || !lex_label_list(*ly, c.neighbour, err)) {
            err = "neighbour: " + err;
            return false;
        }
    }
    c.nF = (int)c.owner.size();
    c.nIF = (int)c.neighbour.size();
    if ((int)c.face_off.size() != c.nF + 1) {
        err = "polyMesh: " + std::to_string(c.face_off.size() - 1) + " faces but " + std::to_string(c.nF) + " owners";
        return false;
    }
    if (c.nIF > c.nF) {
        err = "polyMesh: more neighbours than faces";
        return false;
    }
    for (int f = 0; f < c.nF; ++f) c.nC = std::max(c.nC, c.owner[f] + 1);
    for (int f = 0; f < c.nIF; ++f) c.nC = std::max(c.nC, c.neighbour[f] + 1);
    for (int f = 0; f < c.nF; ++f)
        if (c.owner[f] < 0 || (f < c.nIF && c.neighbour[f] < 0)) {
            err = "polyMesh: negative cell label at face " + std::to_string(f);
            return false;
        }
    {
        FoamFile f;
        std::unique_ptr<Lexer> lx;
        if (!open_foam_list(join(pm, "boundary"), f, lx, err)) return false;
        Token n = lx->next();
        Token b = lx->next();
        if (!n.is_int() || !(b.kind == Token::kPunct && b.punct == '(')) {
            err = "boundary: expected `N (`";
            return false;
        }
        for (int p = 0; p < (int)n.num; ++p) {
            Token name = lx->next();
            Token br = lx->next();
            if (name.kind != Token::kWord || !(br.kind == Token::kPunct && br.punct == '{')) {
                err = "boundary: patch " + std::to_string(p) + " is not `name { ... }`";
                return false;
            }
            Dict d;
            if (!parse_dict(*lx, d, err, true)) {
                err = "boundary: " + err;
                return false;
            }
            PatchInfo pi;
            pi.name = name.text;
            double v;
            const Entry *e;
            if (!(e = need(d, "type", "boundary " + pi.name, err)) || !entry_word(*e, pi.geom_type, err, "type"))
                return false;
            if (!(e = need(d, "nFaces", "boundary " + pi.name, err)) || !entry_scalar(*e, v, err, "nFaces")) return false;
            pi.size = (int)v;
            if (!(e = need(d, "startFace", "boundary " + pi.name, err)) || !entry_scalar(*e, v, err, "startFace"))
                return false;
            pi.start = (int)v;
            if (pi.start < c.nIF || pi.start + pi.size > c.nF) {
                err = "boundary: patch " + pi.name + " faces outside the boundary range";
                return false;
            }
            c.patches.push_back(pi);
        }
    }
    return true;
}

bool read_case(const std::string &dir, const std::string &tname, Case &c, std::string &err, bool vmesh)
{
    c = Case();
    c.dir = dir;
    // ---- controlDict (readTimeControlsExplicit.H:39-52, createTime.H)
    {
        const std::string path = join(dir, "system/controlDict");
        FoamFile f;
        if (!read_foam_dict(path, f, err)) return false;
        const Dict &d = f.body;
        Controls &k = c.ctl;
        const Entry *e;
        if ((e = d.find("startFrom")) && !entry_word(*e, k.start_from, err, "startFrom")) return false;
        if ((e = d.find("startTime")) && !entry_scalar(*e, k.start_time, err, "startTime")) return false;
        if (!(e = need(d, "endTime", path, err)) || !entry_scalar(*e, k.end_time, err, "endTime")) return false;
        if (!(e = need(d, "deltaT", path, err)) || !entry_scalar(*e, k.delta_t, err, "deltaT")) return false;
        if ((e = d.find("writeControl")) && !entry_word(*e, k.write_control, err, "writeControl")) return false;
        if ((e = d.find("writeInterval")) && !entry_scalar(*e, k.write_interval, err, "writeInterval")) return false;
        if ((e = d.find("timeFormat")) && !entry_word(*e, k.time_format, err, "timeFormat")) return false;
        double v;
        if ((e = d.find("writePrecision"))) {
            if (!entry_scalar(*e, v, err, "writePrecision")) return false;
            k.write_precision = (int)v;
        }
        if ((e = d.find("timePrecision"))) {
            if (!entry_scalar(*e, v, err, "timePrecision")) return false;
            k.time_precision = (int)v;
        }
        if ((e = d.find("adjustTimeStep")) && !entry_bool(*e, k.adjust_time_step, err, "adjustTimeStep")) return false;
        if ((e = d.find("maxCo")) && !entry_scalar(*e, k.max_co, err, "maxCo")) return false;
        if ((e = d.find("maxDeltaT")) && !entry_scalar(*e, k.max_delta_t, err, "maxDeltaT")) return false;
        if ((e = d.find("convergeTol")) && !entry_scalar(*e, k.converge_tol, err, "convergeTol")) return false;
        if ((e = d.find("convergeCheckSteps"))) {
            if (!entry_scalar(*e, v, err, "convergeCheckSteps")) return false;
            k.converge_check_steps = (int)v;
        }
        if (k.converge_check_steps < 1) {
            err = path + ": convergeCheckSteps must be >= 1";
            return false;
        }
        if (!(k.delta_t > 0)) {
            err = path + ": deltaT must be positive";
            return false;
        }
    }
    // ---- the time directory to start from
    const auto times = time_dirs(dir);
    if (!tname.empty()) {
        c.time_name = tname;
        if (!is_number_name(tname, c.time)) {
            err = "time name '" + tname + "' is not a number";
            return false;
        }
    } else if (times.empty()) {
        err = dir + ": no time directory";
        return false;
    } else if (c.ctl.start_from == "latestTime") {
        c.time = times.back().first;
        c.time_name = times.back().second;
    } else if (c.ctl.start_from == "firstTime") {
        c.time = times.front().first;
        c.time_name = times.front().second;
    } else {
        c.time = c.ctl.start_time;
        c.time_name = time_name(c.ctl, c.ctl.start_time);
        for (const auto &t : times)
            if (std::fabs(t.first - c.ctl.start_time) <= 1e-12 * std::max(1.0, std::fabs(t.first))) c.time_name = t.second;
    }
    // ---- polyMesh
    if (!read_poly_mesh(join(dir, "constant/polyMesh"), c, err)) return false;
    if (!compute_geometry(c, err)) return false;
    // ---- gradient scheme (system/fvSchemes; absent: leastSquares)
    {
        FoamFile f;
        std::string e2;
        if (read_foam_dict(join(dir, "system/fvSchemes"), f, e2)) {
            if (const Dict *gs = f.body.sub("gradSchemes"))
                if (const Entry *d = gs->find("default")) {
                    std::string sch;
                    for (const Token &t : d->tokens) sch += (sch.empty() ? "" : " ") + t.text;
                    c.grad_scheme = sch;
                }
        }
        if (c.grad_scheme == "Gauss linear") {
            // grad_P = (1/V_P) sum_f S_f (w phi_P + (1-w) phi_N) with boundary values equal to
            // the cell value is, as sum_f S_f = 0 over a closed cell, the form
            // grad_P += own.(phi_N - phi_P), grad_N -= nei.(phi_N - phi_P) with these vectors
            c.ls_own.resize(3 * (size_t)c.nIF);
            c.ls_nei.resize(3 * (size_t)c.nIF);
            for (int fi = 0; fi < c.nIF; ++fi) {
                const double w = c.weights[fi];
                for (int k = 0; k < 3; ++k) {
                    c.ls_own[3 * (size_t)fi + k] = c.Sf[3 * (size_t)fi + k] * (1.0 - w) / c.V[c.owner[fi]];
                    c.ls_nei[3 * (size_t)fi + k] = -c.Sf[3 * (size_t)fi + k] * w / c.V[c.neighbour[fi]];
                }
            }
        } else if (c.grad_scheme != "leastSquares") {
            err = "gradSchemes default '" + c.grad_scheme + "' is not supported (leastSquares, Gauss linear)";
            return false;
        }
    }
    // ---- DVMProperties (fvDVM.C:1928-1948) and the velocity set (fvDVM.C:62-123)
    {
        const std::string path = join(dir, "constant/DVMProperties");
        FoamFile f;
        if (!read_foam_dict(path, f, err)) return false;
        const Dict *pa = f.body.sub("fvDVMparas"), *gp = f.body.sub("gasProperties");
        if (!pa || !gp) {
            err = path + ": fvDVMparas and gasProperties are required";
            return false;
        }
        const Entry *e;
        double v;
        if (!(e = need(*pa, "nDV", path, err)) || !entry_scalar(*e, v, err, "nDV")) return false;
        c.nDV = (int)v;
        if (!(e = need(*pa, "xiMax", path, err)) || !entry_dimensioned(*e, c.xi_max, err, "xiMax")) return false;
        if (!(e = need(*pa, "xiMin", path, err)) || !entry_dimensioned(*e, c.xi_min, err, "xiMin")) return false;
        if ((e = pa->find("macroFlux")) && !entry_word(*e, c.macro_flux, err, "macroFlux")) return false;
        if (!(e = need(*gp, "R", path, err)) || !entry_dimensioned(*e, c.R, err, "R")) return false;
        if (!(e = need(*gp, "omega", path, err)) || !entry_dimensioned(*e, c.omega, err, "omega")) return false;
        if (!(e = need(*gp, "Tref", path, err)) || !entry_dimensioned(*e, c.Tref, err, "Tref")) return false;
        if (!(e = need(*gp, "muRef", path, err)) || !entry_dimensioned(*e, c.muRef, err, "muRef")) return false;
        if (!(e = need(*gp, "Pr", path, err)) || !entry_dimensioned(*e, c.Pr, err, "Pr")) return false;
        if ((e = gp->find("KInner"))) {
            if (!entry_dimensioned(*e, v, err, "KInner")) return false;
            c.KInner = (int)v;
        }
        if ((e = f.body.find("DFwriteCellList"))) {
            std::vector<double> cl;
            size_t k = 0;
            if (!tok_scalar_list(e->tokens, k, cl, err)) {
                err = path + " DFwriteCellList: " + err;
                return false;
            }
            for (double x : cl) c.df_write_cells.push_back((int)x);
        }
        // updateMacroVol takes the moments of g~ only for macroFlux "no" (fvDVM.C:1528); any
        // other word selects the flux update, whose form the VMESH build changes (F:1562)
        c.macro_flux_form = c.macro_flux == "no" ? 0 : (vmesh ? 2 : 1);
    }
    c.vmesh = vmesh;
    if (vmesh) {
        // initialiseDV of the VMESH build (fvDVM.C:129-253): the velocity set is the cell
        // centres and volumes of constant/vMesh, components beyond nSolutionD set to 0, and
        // xiMax the largest |xi|; constant/Xis and constant/weights are not read
        Case vm;
        if (!read_poly_mesh(join(dir, "constant/vMesh"), vm, err) || !compute_geometry(vm, err)) {
            err = "constant/vMesh: " + err;
            return false;
        }
        c.vlist_xi.assign(3 * (size_t)vm.nC, 0.0);
        c.vlist_w = vm.V;
        c.xi_max = 0.0;
        for (int i = 0; i < vm.nC; ++i) {
            double m2 = 0.0;
            for (int k = 0; k < c.n_solution_d; ++k) {
                const double x = vm.C[3 * (size_t)i + k];
                c.vlist_xi[3 * (size_t)i + k] = x;
                m2 += x * x;
            }
            c.xi_max = std::max(c.xi_max, std::sqrt(m2));
            if (!(vm.V[i] > 0)) {
                err = "constant/vMesh: cell " + std::to_string(i) + " has a non-positive volume";
                return false;
            }
        }
    } else {
        if (!read_scalar_list_file(join(dir, "constant/Xis"), c.xis, err) ||
            !read_scalar_list_file(join(dir, "constant/weights"), c.wts, err))
            return false;
        if ((int)c.xis.size() != c.nDV || (int)c.wts.size() != c.nDV) {
            err = "constant/Xis and constant/weights must hold nDV = " + std::to_string(c.nDV) + " values";
            return false;
        }
    }
    // ---- initial fields (createFields.H)
    if (!read_vol_field(c, "rho", 1, c.rho, err) || !read_vol_field(c, "U", 3, c.U, err) ||
        !read_vol_field(c, "T", 1, c.T, err))
        return false;
    // ---- the C-ABI view: internal faces, then the non-empty patches
    c.abi_face.resize(c.nIF);
    for (int f = 0; f < c.nIF; ++f) c.abi_face[f] = f;
    for (size_t p = 0; p < c.patches.size(); ++p) {
        const PatchInfo &pi = c.patches[p];
        if (pi.geom_type == "empty" || pi.size == 0) continue;
        const std::string &rt = c.rho.bc[p].type;
        const int code = df_patch_code(rt);
        if (code <= 0) {
            err = "rho boundary type '" + rt + "' on patch " + pi.name + " is not supported";
            return false;
        }
        const FieldBC &bu = c.U.bc[p], &bt = c.T.bc[p], &br = c.rho.bc[p];
        // values each type reads: walls and fixedValue patches rho, U, T; farField rho, T;
        // pressureIn T and the pressure; pressureOut the pressure (rho and T follow from the
        // cells); zeroGradient and symmetry patches none (missing values are taken as 1 / 0)
        const bool need_rho = code <= 2 || code == 4, need_U = code <= 2, need_T = code <= 2 || code == 4 || code == 5;
        auto size_ok = [&](const FieldBC &f, size_t w) { return f.value.size() == w * (size_t)pi.size; };
        if ((need_rho && !size_ok(br, 1)) || (need_U && !size_ok(bu, 3)) || (need_T && !size_ok(bt, 1))) {
            err = "patch " + pi.name + ": the " + rt + " boundary needs patch values of " +
                  (code <= 2 ? "rho, U and T" : code == 4 ? "rho and T" : "T");
            return false;
        }
        if ((code == 5 || code == 6) && !(br.pressure > 0)) {
            err = "patch " + pi.name + ": " + rt + " needs a positive `" + rt + "` entry";
            return false;
        }
        c.solver_patch.push_back((int)p);
        c.patch_code.push_back(code);
        c.patch_start.push_back(c.nB);
        c.patch_size.push_back(pi.size);
        c.patch_pressure.push_back(br.pressure);
        for (int i = 0; i < pi.size; ++i) {
            c.abi_face.push_back(pi.start + i);
            c.bf_rho.push_back(size_ok(br, 1) ? br.value[i] : 1.0);
            for (int k = 0; k < 3; ++k) c.bf_U.push_back(size_ok(bu, 3) ? bu.value[3 * (size_t)i + k] : 0.0);
            c.bf_T.push_back(size_ok(bt, 1) ? bt.value[i] : 1.0);
        }
        c.nB += pi.size;
    }
    const size_t nabi = c.abi_face.size();
    c.abi_owner.resize(nabi);
    c.abi_Sf.resize(3 * nabi);
    c.abi_Cf.resize(3 * nabi);
    for (size_t i = 0; i < nabi; ++i) {
        const int f = c.abi_face[i];
        c.abi_owner[i] = c.owner[f];
        for (int k = 0; k < 3; ++k) {
            c.abi_Sf[3 * i + k] = c.Sf[3 * (size_t)f + k];
            c.abi_Cf[3 * i + k] = c.Cf[3 * (size_t)f + k];
        }
    }
    return true;
}

// ----------------------------------------------------------------------------- writing
namespace {

struct Writer {
    std::FILE *f = nullptr;
    int prec = 6;
    void num(double v) { std::fprintf(f, "%.*g", prec, v); }
    void vec(const double *v)
    {
        std::fputc('(', f);
        num(v[0]);
        std::fputc(' ', f);
        num(v[1]);
        std::fputc(' ', f);
        num(v[2]);
        std::fputc(')', f);
    }
    void header(const char *cls, const std::string &location, const std::string &object)
    {
        std::fprintf(f, "// DUGKS field output (OpenFOAM ASCII format)\n");
        std::fprintf(f, "FoamFile\n{\n    version     2.0;\n    format      ascii;\n    class       %s;\n", cls);
        std::fprintf(f, "    location    \"%s\";\n    object      %s;\n}\n\n", location.c_str(), object.c_str());
    }
    // `nonuniform List<T> N ( ... )` or `uniform v` when every value is the same
    void values(const char *type, const double *v, size_t n, int width)
    {
        bool same = n > 0;
        for (size_t i = 1; same && i < n; ++i)
            for (int k = 0; k < width; ++k) same = same && v[i * width + k] == v[k];
        if (same) {
            std::fprintf(f, "uniform ");
            width == 1 ? num(v[0]) : vec(v);
            return;
        }
        std::fprintf(f, "nonuniform List<%s> \n%zu\n(\n", type, n);
        for (size_t i = 0; i < n; ++i) {
            width == 1 ? num(v[i]) : vec(v + i * width);
            std::fputc('\n', f);
        }
        std::fputc(')', f);
    }
};

bool open_out(Writer &w, const std::string &path, int prec, std::string &err)
{
    w.f = std::fopen(path.c_str(), "w");
    w.prec = prec;
    if (!w.f) err = "cannot write " + path;
    return w.f != nullptr;
}

// one volume field: internal values, then per patch (empty, or type + value)
bool write_vol(const Case &c, const std::string &tdir, const std::string &tname, const std::string &name,
               const char *cls, const char *vtype, int width, const std::string &dims, const double *cells,
               const std::vector<std::string> &types, const std::vector<std::vector<double>> &pvals, std::string &err)
{
    Writer w;
    if (!open_out(w, join(tdir, name), c.ctl.write_precision, err)) return false;
    w.header(cls, tname, name);
    std::fprintf(w.f, "dimensions      %s;\n\ninternalField   ", dims.c_str());
    w.values(vtype, cells, (size_t)c.nC, width);
    std::fprintf(w.f, ";\n\nboundaryField\n{\n");
    for (size_t p = 0; p < c.patches.size(); ++p) {
        std::fprintf(w.f, "    %s\n    {\n        type            %s;\n", c.patches[p].name.c_str(), types[p].c_str());
        if (types[p] != "empty" && !pvals[p].empty()) {
            std::fprintf(w.f, "        value           ");
            w.values(vtype, pvals[p].data(), pvals[p].size() / width, width);
            std::fprintf(w.f, ";\n");
        }
        std::fprintf(w.f, "    }\n");
    }
    std::fprintf(w.f, "}\n");
    const bool ok = std::fclose(w.f) == 0;
    if (!ok) err = "write failed: " + join(tdir, name);
    return ok;
}

}  // namespace

bool write_time(const Case &c, const Output &o, std::string &err)
{
    const std::string tname = time_name(c.ctl, o.time);
    const std::string tdir = join(c.dir, tname);
    if (!mkdirs(join(tdir, "uniform"), err)) return false;
    const size_t np = c.patches.size();
    // polyMesh patch -> C-ABI boundary start (-1: empty / not carried)
    std::vector<int> bstart(np, -1);
    for (size_t i = 0; i < c.solver_patch.size(); ++i) bstart[c.solver_patch[i]] = c.patch_start[i];
    auto per_patch = [&](const VolField *in, int width, const double *cells, const double *bvals,
                         std::vector<std::string> &types, std::vector<std::vector<double>> &vals, const char *calc) {
        types.assign(np, "");
        vals.assign(np, {});
        for (size_t p = 0; p < np; ++p) {
            const PatchInfo &pi = c.patches[p];
            if (pi.geom_type == "empty") {
                types[p] = "empty";
                continue;
            }
            types[p] = in ? in->bc[p].type : calc;
            if (bvals && bstart[p] >= 0) {
                vals[p].assign(bvals + (size_t)bstart[p] * width, bvals + (size_t)(bstart[p] + pi.size) * width);
            } else if (in && !in->bc[p].value.empty()) {
                vals[p] = in->bc[p].value;
            } else if (cells) {  // owner-cell values
                for (int i = 0; i < pi.size; ++i) {
                    const int cell = c.owner[pi.start + i];
                    vals[p].insert(vals[p].end(), cells + (size_t)cell * width, cells + (size_t)(cell + 1) * width);
                }
            } else {
                vals[p].assign((size_t)pi.size * width, 0.0);
            }
        }
    };
    std::vector<std::string> types;
    std::vector<std::vector<double>> vals;
    // rho: diffuse walls hold the wall density of the step (calculatedMaxwell), fixedValue
    // patches their fixed value
    per_patch(&c.rho, 1, o.rho, o.rho_b, types, vals, "calculated");
    if (!write_vol(c, tdir, tname, "rho", "volScalarField", "scalar", 1, c.rho.dims.empty() ? "[1 -3 0 0 0 0 0]" : c.rho.dims,
                   o.rho, types, vals, err))
        return false;
    per_patch(&c.U, 3, o.U, nullptr, types, vals, "calculated");
    if (!write_vol(c, tdir, tname, "U", "volVectorField", "vector", 3, c.U.dims.empty() ? "[0 1 -1 0 0 0 0]" : c.U.dims,
                   o.U, types, vals, err))
        return false;
    per_patch(&c.T, 1, o.T, nullptr, types, vals, "calculated");
    if (!write_vol(c, tdir, tname, "T", "volScalarField", "scalar", 1, c.T.dims.empty() ? "[0 0 0 1 0 0 0]" : c.T.dims,
                   o.T, types, vals, err))
        return false;
    // q: the reference's patch values of q are not meaningful (fvDVM.C:1687-1690, 1717-1727);
    // written as the owner-cell values
    per_patch(nullptr, 3, o.q, nullptr, types, vals, "calculated");
    if (!write_vol(c, tdir, tname, "q", "volVectorField", "vector", 3, "[1 0 -3 0 0 0 0]", o.q, types, vals, err))
        return false;
    // qWall and stressWall: zero in this version of the reference (their evaluation is
    // commented out, fvDVM.C:1317-1345)
    {
        std::vector<double> zc(3 * (size_t)c.nC, 0.0);
        per_patch(nullptr, 3, zc.data(), nullptr, types, vals, "calculated");
        if (!write_vol(c, tdir, tname, "qWall", "volVectorField", "vector", 3, "[1 0 -3 0 0 0 0]", zc.data(), types,
                       vals, err))
            return false;
        Writer w;
        if (!open_out(w, join(tdir, "stressWall"), c.ctl.write_precision, err)) return false;
        w.header("volTensorField", tname, "stressWall");
        std::fprintf(w.f, "dimensions      [1 -1 -2 0 0 0 0];\n\ninternalField   uniform (0 0 0 0 0 0 0 0 0);\n\n"
                          "boundaryField\n{\n");
        for (const PatchInfo &pi : c.patches)
            std::fprintf(w.f, "    %s\n    {\n        type            %s;\n%s    }\n", pi.name.c_str(),
                         pi.geom_type == "empty" ? "empty" : "calculated",
                         pi.geom_type == "empty" ? "" : "        value           uniform (0 0 0 0 0 0 0 0 0);\n");
        std::fprintf(w.f, "}\n");
        std::fclose(w.f);
    }
    // qSurf (surfaceVectorField over internal faces and patches) and stressSurf (zero:
    // set to zero each step and never accumulated, fvDVM.C:1188-1193, 1305-1306)
    {
        Writer w;
        if (!open_out(w, join(tdir, "qSurf"), c.ctl.write_precision, err)) return false;
        w.header("surfaceVectorField", tname, "qSurf");
        std::fprintf(w.f, "dimensions      [1 0 -3 0 0 0 0];\n\ninternalField   ");
        std::vector<double> zero3(3, 0.0);
        if (o.q_surf)
            w.values("vector", o.q_surf, (size_t)c.nIF, 3);
        else
            w.values("vector", zero3.data(), 1, 3);
        std::fprintf(w.f, ";\n\nboundaryField\n{\n");
        for (size_t p = 0; p < np; ++p) {
            const PatchInfo &pi = c.patches[p];
            if (pi.geom_type == "empty") {
                std::fprintf(w.f, "    %s\n    {\n        type            empty;\n        value           "
                                  "nonuniform List<vector> 0();\n    }\n", pi.name.c_str());
                continue;
            }
            std::fprintf(w.f, "    %s\n    {\n        type            calculated;\n        value           ",
                         pi.name.c_str());
            if (o.q_surf && bstart[p] >= 0)
                w.values("vector", o.q_surf + 3 * (size_t)(c.nIF + bstart[p]), (size_t)pi.size, 3);
            else
                w.values("vector", zero3.data(), 1, 3);
            std::fprintf(w.f, ";\n    }\n");
        }
        std::fprintf(w.f, "}\n");
        std::fclose(w.f);
        Writer s;
        if (!open_out(s, join(tdir, "stressSurf"), c.ctl.write_precision, err)) return false;
        s.header("surfaceTensorField", tname, "stressSurf");
        std::fprintf(s.f, "dimensions      [1 -1 -2 0 0 0 0];\n\ninternalField   uniform (0 0 0 0 0 0 0 0 0);\n\n"
                          "boundaryField\n{\n");
        for (const PatchInfo &pi : c.patches)
            std::fprintf(s.f, "    %s\n    {\n        type            %s;\n        value           %s;\n    }\n",
                         pi.name.c_str(), pi.geom_type == "empty" ? "empty" : "calculated",
                         pi.geom_type == "empty" ? "nonuniform List<tensor> 0()" : "uniform (0 0 0 0 0 0 0 0 0)");
        std::fprintf(s.f, "}\n");
        std::fclose(s.f);
    }
    // uniform/time (Time::writeTimeDict)
    {
        Writer w;
        if (!open_out(w, join(tdir, "uniform/time"), c.ctl.write_precision, err)) return false;
        w.header("dictionary", tname + "/uniform", "time");
        std::fprintf(w.f, "value           %.21g;\n\nname            \"%s\";\n\nindex           %d;\n\n", o.time,
                     tname.c_str(), o.index);
        std::fprintf(w.f, "deltaT          %.12g;\n\ndeltaT0         %.12g;\n", o.delta_t, o.delta_t0);
        std::fclose(w.f);
    }
    return true;
}

bool write_df(const Case &c, double t, int label, const double *g, int nv, std::string &err)
{
    const std::string tname = time_name(c.ctl, t);
    const std::string tdir = join(c.dir, tname);
    if (!mkdirs(tdir, err)) return false;
    Writer w;
    const std::string name = "DF" + std::to_string(label);
    if (!open_out(w, join(tdir, name), c.ctl.write_precision, err)) return false;
    w.header("scalarList", tname, name);
    std::fprintf(w.f, "%d\n(\n", nv);
    for (int i = 0; i < nv; ++i) {
        w.num(g[i]);
        std::fputc('\n', w.f);
    }
    std::fprintf(w.f, ")\n");
    return std::fclose(w.f) == 0;
}

}  // namespace dugks_foam
