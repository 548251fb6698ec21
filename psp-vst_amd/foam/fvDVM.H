// C++ facade of the DUGKS solver with the interface of dugksFoam's Foam::fvDVM and
// Foam::discreteVelocity (src/fvDVM/fvDVM/fvDVM.H:301-408,
// src/fvDVM/discreteVelocity/discreteVelocity.H:228-272), header-only, over the C-ABIs of
// libdugks_hip.so (include/dugks.h, the GPU step) and libdugks_foam.so
// (include/dugks_foam.h, the case directory).  Without OpenFOAM the fields are plain host
// arrays in the polyMesh numbering (scalars [n], vectors [n][3]), refreshed from the GPU
// on request; a solver loop written against Foam::fvDVM (dugksFoam.C:54-110) maps onto it
// one call to one call (see dugksFoam.C in this directory).
//
// Errors raise dugks::Error (std::runtime_error) with the library's message.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dugks.h"
#include "../../include/dugks_foam.h"

namespace dugks {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char *what)
{
    if (rc != 0) throw Error(std::string(what) + ": " + dugks_last_error());
}
inline void check_foam(int rc, const char *what)
{
    if (rc != 0) throw Error(std::string(what) + ": " + dugks_foam_last_error());
}

class fvDVM;

// One discrete velocity of the set (discreteVelocity.H:228-272): xi, weight and host copies
// of g~, h~ over the cells.
class discreteVelocity {
public:
    discreteVelocity(const fvDVM &dvm, int index) : dvm_(&dvm), i_(index) {}
    int index() const { return i_; }
    inline const double *xi() const;          // [3]
    inline double weight() const;
    inline std::vector<double> gTildeVol() const;
    inline std::vector<double> hTildeVol() const;

private:
    const fvDVM *dvm_;
    int i_;
};

struct fvDVMOptions {
    int precision = DUGKS_FP64;  // DUGKS_FP64 or DUGKS_MIXED
    int device = 0;
    int fused = -1;              // dugks_options.fused
    int q1_compat = 1;
    bool vmesh = false;          // velocity set from constant/vMesh, as dugksFoam built with -DVMESH
};

class fvDVM {
public:
    using Options = fvDVMOptions;

    // Reads the case directory (the fields of createFields.H and DVMProperties) and builds
    // the solver: fvDVM::fvDVM (fvDVM.C:1901-2117).  time_name empty: controlDict startFrom.
    explicit fvDVM(const std::string &case_dir, const Options &opt = Options(), const std::string &time_name = "")
    {
        check_foam(opt.vmesh ? dugks_foam_read_vmesh(case_dir.c_str(), time_name.c_str(), &case_)
                             : dugks_foam_read(case_dir.c_str(), time_name.c_str(), &case_),
                   "reading the case");
        try {
            int np = 0;
            check_foam(dugks_foam_solver_inputs(case_, nullptr, nullptr, 0, &np, nullptr, nullptr), "case inputs");
            dugks_mesh mesh;
            dugks_vset vset;
            dugks_gas gas;
            patches_.resize(np > 0 ? np : 1);
            check_foam(dugks_foam_solver_inputs(case_, &mesh, patches_.data(), np, &np, &vset, &gas), "case inputs");
            check_foam(dugks_foam_controls_get(case_, &ctl_), "controlDict");
            nC_ = mesh.n_cells;
            nIF_ = mesh.n_internal_faces;
            nB_ = mesh.n_boundary_faces;
            nXiPerDim_ = vset.n_per_dim;
            xiMax_ = vset.xi_max;
            xiMin_ = vset.xi_min;
            gas_ = gas;
            dugks_options o{};
            o.precision = opt.precision;
            o.q1_compat = opt.q1_compat;
            o.device = opt.device;
            o.fused = opt.fused;
            double mf = 0.0;
            check_foam(dugks_foam_scalar(case_, "macro_flux", &mf), "macroFlux");
            o.macro_flux = (int)mf;
            check(dugks_create(&mesh, patches_.data(), np, &vset, &gas, &o, &h_), "dugks_create");
            const double *rho, *U, *T;
            check_foam(dugks_foam_initial_fields(case_, &rho, &U, &T), "initial fields");
            check(dugks_init_from_macros(h_, rho, U, T), "dugks_init_from_macros");
            nXi_ = dugks_num_velocities(h_);
            xi_.resize(3 * (size_t)nXi_);
            w_.resize(nXi_);
            check(dugks_velocity_set(h_, xi_.data(), w_.data()), "dugks_velocity_set");
            deltaT_ = ctl_.delta_t;
            time_ = ctl_.time;
            refresh_cells();
        } catch (...) {
            release();
            throw;
        }
    }
    fvDVM(const fvDVM &) = delete;
    fvDVM &operator=(const fvDVM &) = delete;
    virtual ~fvDVM() { release(); }

    // fvDVM::evolution (fvDVM.C:2128-2174) with the current deltaT(); cell fields are stale
    // until refresh_cells() (the convergence check and write do it)
    void evolution()
    {
        check(dugks_evolution(h_, deltaT_), "dugks_evolution");
        time_ += deltaT_;
        ++timeIndex_;
        stale_ = true;
    }

    // fvDVM::getCoNum (fvDVM.C:2177-2185) for the current deltaT()
    void getCoNum(double &maxCoNum, double &meanCoNum) const
    {
        check(dugks_co_num(h_, deltaT_, &maxCoNum, &meanCoNum), "dugks_co_num");
    }

    // n iterations of the dugksFoam loop body (CourantNo.H, setDeltaTvar.H, evolution) on
    // the device without host round trips; dt_history [n] receives the time steps (may be
    // NULL)
    void advance(int n, double *dt_history = nullptr)
    {
        std::vector<double> hist;
        if (!dt_history && n > 0) {
            hist.resize(n);
            dt_history = hist.data();
        }
        double dt = deltaT_;
        check(dugks_advance(h_, n, ctl_.adjust_time_step, ctl_.max_co, ctl_.max_delta_t, &dt, dt_history), "dugks_advance");
        for (int i = 0; i < n; ++i) time_ += dt_history[i];
        if (n > 0) deltaT0_ = n > 1 ? dt_history[n - 2] : deltaT_;
        deltaT_ = dt;
        timeIndex_ += n;
        stale_ = true;
    }

    // time control (the parts of Foam::Time the loop uses)
    double deltaTValue() const { return deltaT_; }
    void setDeltaT(double dt)
    {
        deltaT0_ = deltaT_;
        deltaT_ = dt;
    }
    double timeValue() const { return time_; }
    int timeIndex() const { return timeIndex_; }
    const dugks_foam_controls &controls() const { return ctl_; }

    // macroscopic cell fields (host copies, polyMesh cell order)
    void refresh_cells()
    {
        rhoVol_.resize(nC_);
        Uvol_.resize(3 * (size_t)nC_);
        Tvol_.resize(nC_);
        qVol_.resize(3 * (size_t)nC_);
        tauVol_.resize(nC_);
        check(dugks_get_cell_macros(h_, rhoVol_.data(), Uvol_.data(), Tvol_.data(), qVol_.data(), tauVol_.data()),
              "dugks_get_cell_macros");
        stale_ = false;
    }
    const std::vector<double> &rhoVol() { return cells(rhoVol_); }
    const std::vector<double> &Uvol() { return cells(Uvol_); }
    const std::vector<double> &Tvol() { return cells(Tvol_); }
    const std::vector<double> &qVol() { return cells(qVol_); }
    const std::vector<double> &tauVol() { return cells(tauVol_); }
    // face fields of the last step ([nIF + nB], C-ABI face numbering)
    std::vector<double> rhoSurf() const { return face(0); }
    std::vector<double> Usurf() const { return face(1); }
    std::vector<double> Tsurf() const { return face(2); }
    std::vector<double> qSurf() const { return face(3); }
    std::vector<double> tauSurf() const { return face(4); }
    std::vector<double> boundaryRho() const
    {
        std::vector<double> r(nB_);
        check(dugks_get_boundary_rho(h_, r.data()), "dugks_get_boundary_rho");
        return r;
    }

    // velocity set and gas (fvDVMI.H)
    int nXi() const { return nXi_; }
    int nXiPerDim() const { return nXiPerDim_; }
    double xiMax() const { return xiMax_; }
    double xiMin() const { return xiMin_; }
    double R() const { return gas_.R; }
    double omega() const { return gas_.omega; }
    double Tref() const { return gas_.Tref; }
    double muRef() const { return gas_.muRef; }
    double Pr() const { return gas_.Pr; }
    int KInner() const { return gas_.KInner; }
    discreteVelocity DVi(int i) const
    {
        if (i < 0 || i >= nXi_) throw Error("DVi: velocity index out of range");
        return discreteVelocity(*this, i);
    }
    discreteVelocity DVxyz(int ix, int iy, int iz) const
    {
        const int n = nXiPerDim_;
        return DVi(ix + n * (iy + n * iz));   // ix fastest (fvDVM.C:293-373)
    }
    const double *xi(int i) const { return &xi_[3 * (size_t)i]; }
    double weight(int i) const { return w_[i]; }
    // g~, h~ of every velocity and cell [nXi][nC] (one transfer)
    void distributions(std::vector<double> &g, std::vector<double> &h) const
    {
        g.resize((size_t)nXi_ * nC_);
        h.resize((size_t)nXi_ * nC_);
        check(dugks_get_df(h_, g.data(), h.data()), "dugks_get_df");
    }

    // writeDFonCells (fvDVM.C:1892-1897): DF<cell> of every cell of DFwriteCellList into
    // the current time directory (the reference writes list positions into "0"; here the
    // listed cells into the time being written)
    void writeDFonCells() const
    {
        const int *cells = nullptr;
        size_t n = 0;
        check_foam(dugks_foam_array_i(case_, "df_write_cells", &cells, &n), "DFwriteCellList");
        if (n == 0) return;
        std::vector<double> g, h, one(nXi_);
        distributions(g, h);
        for (size_t k = 0; k < n; ++k) {
            const int c = cells[k];
            if (c < 0 || c >= nC_) throw Error("DFwriteCellList: cell " + std::to_string(c) + " out of range");
            for (int v = 0; v < nXi_; ++v) one[v] = g[(size_t)v * nC_ + c];
            check_foam(dugks_foam_write_df(case_, time_, c, one.data(), nXi_), "writing a DF dump");
        }
    }

    // the AUTO_WRITE fields and uniform/time of the current time (Time::write)
    void write()
    {
        refresh_cells();
        std::vector<double> rb = boundaryRho();
        std::vector<double> qs = qSurf();
        dugks_foam_output o{};
        o.time = time_;
        o.delta_t = deltaT_;
        o.delta_t0 = deltaT0_;
        o.index = timeIndex_;
        o.rho = rhoVol_.data();
        o.U = Uvol_.data();
        o.T = Tvol_.data();
        o.q = qVol_.data();
        o.rho_b = rb.data();
        o.q_surf = qs.data();
        check_foam(dugks_foam_write(case_, &o), "writing fields");
    }

    int nCells() const { return nC_; }
    int nFaces() const { return nIF_ + nB_; }
    dugks_handle *handle() const { return h_; }

private:
    const std::vector<double> &cells(std::vector<double> &v)
    {
        if (stale_) refresh_cells();
        return v;
    }
    std::vector<double> face(int which) const
    {
        const int n = nIF_ + nB_;
        std::vector<double> out((which == 1 || which == 3) ? 3 * (size_t)n : (size_t)n);
        double *p[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
        p[which] = out.data();
        check(dugks_get_face_macros(h_, p[0], p[1], p[2], p[3], p[4]), "dugks_get_face_macros");
        return out;
    }
    void release()
    {
        if (h_) dugks_destroy(h_);
        h_ = nullptr;
        if (case_) dugks_foam_free(case_);
        case_ = nullptr;
    }

    dugks_foam_case *case_ = nullptr;
    dugks_handle *h_ = nullptr;
    std::vector<dugks_patch> patches_;
    dugks_foam_controls ctl_{};
    dugks_gas gas_{};
    int nC_ = 0, nIF_ = 0, nB_ = 0, nXi_ = 0, nXiPerDim_ = 0;
    double xiMax_ = 0, xiMin_ = 0;
    double deltaT_ = 0, deltaT0_ = 0, time_ = 0;
    int timeIndex_ = 0;
    bool stale_ = true;
    std::vector<double> xi_, w_;
    std::vector<double> rhoVol_, Uvol_, Tvol_, qVol_, tauVol_;
};

inline const double *discreteVelocity::xi() const { return dvm_->xi(i_); }
inline double discreteVelocity::weight() const { return dvm_->weight(i_); }
inline std::vector<double> discreteVelocity::gTildeVol() const
{
    std::vector<double> g, h;
    dvm_->distributions(g, h);
    const int n = dvm_->nCells();
    return std::vector<double>(g.begin() + (size_t)i_ * n, g.begin() + (size_t)(i_ + 1) * n);
}
inline std::vector<double> discreteVelocity::hTildeVol() const
{
    std::vector<double> g, h;
    dvm_->distributions(g, h);
    const int n = dvm_->nCells();
    return std::vector<double>(h.begin() + (size_t)i_ * n, h.begin() + (size_t)(i_ + 1) * n);
}

}  // namespace dugks
