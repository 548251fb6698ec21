// dugksFoam without OpenFOAM: the solver loop of src/dugksFoam.C:38-116 over the fvDVM
// facade (fvDVM.H) on one GPU.
//
//   dugksFoam [-case DIR] [-precision fp64|mixed] [-device N] [-fused N] [-maxSteps N]
//             [-hostLoop] [-vmesh]
//
// Per iteration, as dugksFoam.C:63-108: Courant number (CourantNo.H), setDeltaTvar.H
// (dt = min(min(maxCo/(Co + SMALL), 1 + 0.1 maxCo/(Co + SMALL), 1.2) dt, maxDeltaT) when
// adjustTimeStep), time += dt, evolution, DF dumps and field output at write times, and
// every convergeCheckSteps steps the relative changes of T, rho and U against the last
// check (sum|T - Told| / sum T, sum|rho - rhoOld| / sum rho, sum|U - Uold| / sum|U|), the
// loop running while time < endTime - dt/2 and the U change exceeds convergeTol.  The cell
// fields are copied back from the GPU at every check and write, so the monitor sees the
// current state.
//
// By default the steps between two events (write, convergence check, end) run on the GPU
// in one dugks_advance call: the Courant number and setDeltaTvar live on the device, so no
// host round trip per step; the time steps come back with it and are printed afterwards.
// A chunk is cut short so that no step inside it could reach a write or end time (a
// bound on dt growth: at most x1.2 per step, capped by maxDeltaT).  -hostLoop runs the
// reference's per-step host loop instead (getCoNum, setDeltaT, evolution).  -maxSteps N
// stops after N steps (a test aid); the last state is written when the loop ends.  The
// reference's -dvParallel velocity decomposition across processes is not part of this
// one-GPU build and is refused.  -vmesh reads the velocity set from constant/vMesh, what the
// reference does when built with -DVMESH (src/fvDVM/Make/options:13; fvDVM.C:129-253); the
// macroFlux entry of DVMProperties then selects that build's flux update.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "fvDVM.H"

namespace {

constexpr double kSmall = 1.0e-15;  // OpenFOAM SMALL

struct Args {
    std::string case_dir = ".";
    dugks::fvDVM::Options opt;
    long max_steps = 0;
    bool host_loop = false;
};

int usage(const char *msg)
{
    std::fprintf(stderr, "dugksFoam: %s\nusage: dugksFoam [-case DIR] [-precision fp64|mixed] [-device N] "
                         "[-fused N] [-maxSteps N] [-hostLoop] [-vmesh]\n", msg);
    return 1;
}

bool parse(int argc, char **argv, Args &a, std::string &err)
{
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        auto val = [&](std::string &out) {
            if (i + 1 >= argc) {
                err = k + " needs a value";
                return false;
            }
            out = argv[++i];
            return true;
        };
        std::string v;
        if (k == "-case") {
            if (!val(a.case_dir)) return false;
        } else if (k == "-precision") {
            if (!val(v)) return false;
            if (v == "fp64") a.opt.precision = DUGKS_FP64;
            else if (v == "mixed") a.opt.precision = DUGKS_MIXED;
            else {
                err = "-precision takes fp64 or mixed";
                return false;
            }
        } else if (k == "-device") {
            if (!val(v)) return false;
            a.opt.device = std::atoi(v.c_str());
        } else if (k == "-fused") {
            if (!val(v)) return false;
            a.opt.fused = std::atoi(v.c_str());
        } else if (k == "-maxSteps") {
            if (!val(v)) return false;
            a.max_steps = std::atol(v.c_str());
        } else if (k == "-hostLoop") {
            a.host_loop = true;
        } else if (k == "-vmesh") {
            a.opt.vmesh = true;
        } else if (k == "-dvParallel") {
            err = "-dvParallel (velocity decomposition over processes) is not part of this one-GPU build";
            return false;
        } else {
            err = "unknown option " + k;
            return false;
        }
    }
    return true;
}

double sum_abs_diff(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += std::fabs(a[i] - b[i]);
    return s;
}

double sum_mag_vec_diff(const std::vector<double> &a, const std::vector<double> &b)
{
    double s = 0.0;
    for (size_t i = 0; i + 2 < a.size(); i += 3) {
        const double dx = a[i] - b[i], dy = a[i + 1] - b[i + 1], dz = a[i + 2] - b[i + 2];
        s += std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    return s;
}

double sum_mag_vec(const std::vector<double> &a)
{
    double s = 0.0;
    for (size_t i = 0; i + 2 < a.size(); i += 3) s += std::sqrt(a[i] * a[i] + a[i + 1] * a[i + 1] + a[i + 2] * a[i + 2]);
    return s;
}

double sum(const std::vector<double> &a)
{
    double s = 0.0;
    for (double x : a) s += x;
    return s;
}

// Largest number of steps k >= 1 such that no step of the chunk can start at or past
// `bound - dt/2` (time t, current step dt): the times and steps inside the chunk are
// bounded by dt growing x1.2 per step up to maxDeltaT (or staying fixed).
long safe_chunk(double t, double dt, double bound, const dugks_foam_controls &k, long kmax)
{
    long n = 1;
    double tt = t, d = dt;
    for (;;) {
        // step n+1 starts at tt + d (after step n) with a step of at most dn
        const double dn = k.adjust_time_step ? std::fmin(1.2 * d, k.max_delta_t) : d;
        tt += d;
        if (n >= kmax || !(tt + 0.5 * dn < bound)) return n;
        d = dn;
        ++n;
    }
}

}  // namespace

int main(int argc, char **argv)
{
    Args a;
    std::string err;
    if (!parse(argc, argv, a, err)) return usage(err.c_str());
    try {
        // convergence monitor (dugksFoam.C:43-46)
        double TemperatureChange = 1.0, rhoChange = 1.0, Uchange = 1.0;
        dugks::fvDVM dvm(a.case_dir, a.opt);
        const dugks_foam_controls &ctl = dvm.controls();
        std::printf("Case %s: %d cells, %d velocities, start time %g, deltaT %g, endTime %g\n", a.case_dir.c_str(),
                    dvm.nCells(), dvm.nXi(), dvm.timeValue(), dvm.deltaTValue(), ctl.end_time);
        // createFields.H:80-82
        std::vector<double> Told = dvm.Tvol(), rhoOld = dvm.rhoVol(), Uold = dvm.Uvol();
        std::printf("\nStarting time loop\n\n");
        const auto wall0 = std::chrono::steady_clock::now();
        const std::clock_t cpu0 = std::clock();
        long It = 0;
        long last_written = -1;
        const double t0 = dvm.timeValue();
        long write_index = 0;  // runTime write control: last output-time index
        auto is_write_time = [&](double t, double dt) {
            if (ctl.write_control == 0) return It % std::max(1L, (long)std::llround(ctl.write_interval)) == 0;
            const long idx = (long)std::floor(((t - t0) + 0.5 * dt) / ctl.write_interval);
            if (idx > write_index) {
                write_index = idx;
                return true;
            }
            return false;
        };
        auto run = [&]() { return dvm.timeValue() < ctl.end_time - 0.5 * dvm.deltaTValue(); };
        auto report = [&](long step, double t, double dt) {
            const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
            const double cpu = double(std::clock() - cpu0) / CLOCKS_PER_SEC;
            if (ctl.adjust_time_step) std::printf("deltaT = %.12g\n", dt);
            std::printf("Time = %.*g\n\nStep =%ld  ExecutionTime = %g s  ClockTime = %.0f s\n\n", ctl.time_precision, t,
                        step, cpu, wall);
        };
        while (run() && Uchange > ctl.converge_tol) {
            if (a.max_steps > 0 && It >= a.max_steps) break;
            // steps up to the next event
            const long wi = ctl.write_control == 0 ? std::max(1L, (long)std::llround(ctl.write_interval)) : 0;
            long k = 1;
            if (!a.host_loop) {
                long to_event = ctl.converge_check_steps - It % ctl.converge_check_steps;
                if (wi > 0) to_event = std::min(to_event, wi - It % wi);
                if (a.max_steps > 0) to_event = std::min(to_event, a.max_steps - It);
                double bound = ctl.end_time;
                if (ctl.write_control != 0) bound = std::fmin(bound, t0 + (write_index + 1) * ctl.write_interval);
                k = safe_chunk(dvm.timeValue(), dvm.deltaTValue(), bound, ctl, std::max(1L, to_event));
            }
            if (a.host_loop || k == 1) {
                // CourantNo.H, readTimeControlsExplicit.H, setDeltaTvar.H (dugksFoam.C:66-68)
                double CoNum = 0.0, meanCoNum = 0.0;
                dvm.getCoNum(CoNum, meanCoNum);
                if (ctl.adjust_time_step) {
                    const double maxDeltaTFact = ctl.max_co / (CoNum + kSmall);
                    const double deltaTFact = std::fmin(std::fmin(maxDeltaTFact, 1.0 + 0.1 * maxDeltaTFact), 1.2);
                    dvm.setDeltaT(std::fmin(deltaTFact * dvm.deltaTValue(), ctl.max_delta_t));
                }
                ++It;
                dvm.evolution();
                report(It, dvm.timeValue(), dvm.deltaTValue());
            } else {
                std::vector<double> hist(k);
                const double tstart = dvm.timeValue();
                dvm.advance((int)k, hist.data());
                double t = tstart;
                for (long i = 0; i < k; ++i) {
                    t += hist[i];
                    report(It + i + 1, t, hist[i]);
                }
                It += k;
            }
            if (is_write_time(dvm.timeValue(), dvm.deltaTValue())) {
                dvm.writeDFonCells();
                dvm.write();
                last_written = It;
            }
            if (It % ctl.converge_check_steps == 0 && It >= ctl.converge_check_steps) {
                dvm.refresh_cells();
                const std::vector<double> &T = dvm.Tvol(), &rho = dvm.rhoVol(), &U = dvm.Uvol();
                TemperatureChange = sum_abs_diff(T, Told) / sum(T);
                rhoChange = sum_abs_diff(rho, rhoOld) / sum(rho);
                Uchange = sum_mag_vec_diff(U, Uold) / sum_mag_vec(U);
                std::printf("Temperature changes = %g\nDensity     changes = %g\nVelocity    changes = %g\n\n",
                            TemperatureChange, rhoChange, Uchange);
                Told = T;
                rhoOld = rho;
                Uold = U;
            }
        }
        if (last_written != It && It > 0) dvm.write();
        std::printf("End\n\n");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "dugksFoam: %s\n", e.what());
        return 1;
    }
    return 0;
}
