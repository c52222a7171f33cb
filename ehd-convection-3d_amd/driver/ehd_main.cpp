// ehd_main.cpp -- host driver with the reference main.cu's surface.
//
// Same run structure as main.cu:23-315: derive the parameters, initialise
// (or restart from data_end.dat), write the first Tecplot zone, run the time
// loop with the reference's output cadence (i % N == 1), print the collector
// current, and finish with the MLUPS report, the final zone, umax.dat and
// data_end.dat.  The grid, physics and run length are command-line options
// instead of compile-time constants (LBM.h:31-86).  All GPU work goes
// through the C-ABI of include/ehd.h.
//
// Output formats follow LBM.cu:1427-1594 and 1673-1713.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ehd.h"
#include "ehd_io.h"

namespace {

struct Options {
    ehd_params p;
    long long nsteps = 6000000;      // LBM.h:82
    long long nsave = -1;            // LBM.h:83 (NSTEPS/5)
    long long ndmd = 50000;          // LBM.h:85
    long long print_current = 1000;  // LBM.h:86
    bool restart = false;            // LBM.h:30 flag
    bool quiet_files = false;
    std::string outdir = ".";
};

void usage()
{
    std::printf(
        "usage: ehd_main [options]\n"
        "  --nx N --ny N --nz N        grid (default 122 122 101)\n"
        "  --T v --C v --M v --Fe v    nondimensional groups (defaults of LBM.h)\n"
        "  --Pr v                      Prandtl-type metadata (does not change nu or D)\n"
        "  --lambda v --lambda-c v     TRT magic parameters, flow / charge (1/12, 1e-6)\n"
        "  --uw v --exf v              moving top wall, x body force (0)\n"
        "  --steps N                   time steps (6000000)\n"
        "  --nsave N --ndmd N --print-current N   output cadence (NSTEPS/5, 50000, 1000)\n"
        "  --precision fp64|fp32s      population storage (fp64)\n"
        "  --init reference|hydrostatic  initial state (reference)\n"
        "  --seed N --amp v            perturbation of the hydrostatic start (1, 1e-3)\n"
        "  --restart                   start from data_end.dat (read_data)\n"
        "  --no-files                  skip the Tecplot/DMD/end files\n"
        "  --outdir DIR                directory of the output files (.)\n");
}

bool parse(int argc, char** argv, Options& o)
{
    ehd_params_default(&o.p);
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](double& v) {
            if (i + 1 >= argc) return false;
            v = std::atof(argv[++i]);
            return true;
        };
        auto nexti = [&](long long& v) {
            if (i + 1 >= argc) return false;
            v = std::atoll(argv[++i]);
            return true;
        };
        double dv;
        long long lv;
        if (a == "--nx" && nexti(lv)) o.p.nx = (int)lv;
        else if (a == "--ny" && nexti(lv)) o.p.ny = (int)lv;
        else if (a == "--nz" && nexti(lv)) o.p.nz = (int)lv;
        else if (a == "--T" && next(dv)) o.p.T = dv;
        else if (a == "--C" && next(dv)) o.p.C = dv;
        else if (a == "--M" && next(dv)) o.p.M = dv;
        else if (a == "--Fe" && next(dv)) o.p.Fe = dv;
        else if (a == "--Pr" && next(dv)) o.p.Pr = dv;
        else if (a == "--lambda" && next(dv)) o.p.lambda_f = dv;
        else if (a == "--lambda-c" && next(dv)) o.p.lambda_c = dv;
        else if (a == "--uw" && next(dv)) o.p.uw = dv;
        else if (a == "--exf" && next(dv)) o.p.exf = dv;
        else if (a == "--steps" && nexti(lv)) o.nsteps = lv;
        else if (a == "--nsave" && nexti(lv)) o.nsave = lv;
        else if (a == "--ndmd" && nexti(lv)) o.ndmd = lv;
        else if (a == "--print-current" && nexti(lv)) o.print_current = lv;
        else if (a == "--seed" && nexti(lv)) o.p.seed = (uint64_t)lv;
        else if (a == "--amp" && next(dv)) o.p.perturb_amp = dv;
        else if (a == "--restart") o.restart = true;
        else if (a == "--no-files") o.quiet_files = true;
        else if (a == "--outdir" && i + 1 < argc) o.outdir = argv[++i];
        else if (a == "--precision" && i + 1 < argc) {
            std::string v = argv[++i];
            if (v == "fp64") o.p.precision = EHD_FP64;
            else if (v == "fp32s") o.p.precision = EHD_FP32S;
            else return false;
        } else if (a == "--init" && i + 1 < argc) {
            std::string v = argv[++i];
            if (v == "reference") o.p.init_mode = EHD_INIT_REFERENCE;
            else if (v == "hydrostatic") o.p.init_mode = EHD_INIT_HYDROSTATIC;
            else return false;
        } else {
            return false;
        }
    }
    if (o.nsave <= 0) o.nsave = o.nsteps / 5 > 0 ? o.nsteps / 5 : 1;
    if (o.ndmd <= 0) o.ndmd = 1;
    if (o.print_current <= 0) o.print_current = 1;
    return true;
}

void die(int rc, const char* what)
{
    std::fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc, ehd_last_error());
    std::exit(1);
}
#define CALL(x)                     \
    do {                            \
        int rc_ = (x);              \
        if (rc_ != EHD_OK) die(rc_, #x); \
    } while (0)

// the nine fields of the box on the host, in the order of ehd_io.h
struct Fields {
    size_t S;
    std::vector<double> f9;
    explicit Fields(const ehd_params& p) : S((size_t)p.nx * p.ny * p.nz), f9(9 * S, 0.0) {}
    double* at(int k) { return f9.data() + (size_t)k * S; }
    void read(ehd_box* b)
    {
        CALL(ehd_read_fields(b, at(0), at(1), at(2), at(3), at(4), at(5), at(6), at(7), at(8)));
    }
};

void io(int rc, const char* what)
{
    if (rc != EHD_OK) die(rc, what);
}

}  // namespace

int main(int argc, char** argv)
{
    Options o;
    if (!parse(argc, argv, o)) {
        usage();
        return 2;
    }
    ehd_params& p = o.p;
    ehd_derived d;
    std::vector<double> f9;
    double t = 0.0;
    if (o.restart) {
        f9.assign(9 * (size_t)p.nx * p.ny * p.nz, 0.0);
        io(ehd_io_read_end((o.outdir + "/data_end.dat").c_str(), p.nx, p.ny, p.nz, f9.data(), &t), "read_data");
        p.init_mode = EHD_INIT_FIELDS;
    }
    CALL(ehd_derive(&p, &d));
    double T, M, Cn, Fe;
    CALL(ehd_compute_parameters(&d, &p, &T, &M, &Cn, &Fe));

    std::printf("Simulating 3D electrohydrodynamics vortices\n");
    std::printf("      domain size: %dx%dx%d\n", p.nx, p.ny, p.nz);
    std::printf("                T: %g\n", T);
    std::printf("                M: %g\n", M);
    std::printf("                C: %g\n", Cn);
    std::printf("               Fe: %g\n", Fe);
    std::printf("               Pr: %g (metadata)\n", p.Pr);
    std::printf("           lambda: %g (charge %g)\n", p.lambda_f, p.lambda_c);
    std::printf("            uwall: %g\n", p.uw);
    std::printf("   External force: %g\n", p.exf);
    std::printf("        precision: %s\n", p.precision == EHD_FP64 ? "fp64" : "fp32s");
    std::printf("        timesteps: %lld\n", o.nsteps);
    std::printf("       save every: %lld\n", o.nsave);
    std::printf("    message every: %lld\n", o.nsave);
    std::printf("\n");

    ehd_box* b = nullptr;
    CALL(ehd_create(&p, &b));
    CALL(ehd_init(b, o.restart ? f9.data() : nullptr));

    const std::string f_data = o.outdir + "/data.dat", f_umax = o.outdir + "/umax.dat";
    const std::string f_end = o.outdir + "/data_end.dat";
    Fields F(p);
    bool umax_open = false;
    if (!o.quiet_files) {
        F.read(b);
        io(ehd_io_tecplot(f_data.c_str(), 0, 1, t, p.nx, p.ny, p.nz, d.dx, d.dy, d.dz, F.f9.data()), "data.dat");
    }

    const auto begin = std::chrono::steady_clock::now();
    // time loop (main.cu:181-239); cadence test i % N == 1 as in the reference
    long long i = 0;
    while (i < o.nsteps) {
        // advance to the next step index at which some output fires
        long long j = i;
        auto fires = [&](long long k) {
            return (k % o.nsave == 1) || (k % o.ndmd == 1) || (k % o.print_current == 1);
        };
        while (j < o.nsteps - 1 && !fires(j)) ++j;
        const long long n = j - i + 1;
        CALL(ehd_step(b, (int)n));
        for (long long k = 0; k < n; ++k) t = t + d.dt;   // one addition per step, as main.cu:208
        i = j;
        const bool save = !o.quiet_files && i % o.nsave == 1, dmd = i % o.ndmd == 1;
        const bool cur = i % o.print_current == 1;
        if (save || (dmd && !o.quiet_files) || cur) F.read(b);
        if (save) io(ehd_io_tecplot(f_data.c_str(), 1, 0, t, p.nx, p.ny, p.nz, d.dx, d.dy, d.dz, F.f9.data()),
                     "data.dat");
        if (dmd) {
            std::printf("Iteration: %lld, physical time: %g.\n", i, t);
            if (!o.quiet_files) {
                char name[128];
                std::snprintf(name, sizeof(name), "%g.dat", t);
                io(ehd_io_dmd((o.outdir + "/" + name).c_str(), t, p.nx, p.ny, p.nz, F.at(5)), name);
            }
        }
        if (cur) {
            std::printf("Current = %g\n", ehd_io_current(p.nx, p.ny, p.nz, d.K, d.dy, F.at(1), F.at(8)));
            io(ehd_io_umax(f_umax.c_str(), umax_open ? 1 : 0, t, p.nx, p.ny, p.nz, F.at(5)), "umax.dat");
            umax_open = true;
        }
        ++i;
    }
    CALL(ehd_sync(b));
    const double runtime =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - begin).count();
    const double nodes = (double)o.nsteps * (double)p.nx * p.ny * p.nz;
    std::printf(" ----- performance information -----\n");
    std::printf("               timesteps: %lld\n", o.nsteps);
    std::printf("           clock runtime: %.3f (s)\n", runtime);
    std::printf("                   speed: %.2f (Mlups)\n", nodes / (1e6 * runtime));

    F.read(b);
    io(ehd_io_umax(f_umax.c_str(), umax_open ? 1 : 0, t, p.nx, p.ny, p.nz, F.at(5)), "umax.dat");
    if (!o.quiet_files) {
        io(ehd_io_tecplot(f_data.c_str(), 1, 0, t, p.nx, p.ny, p.nz, d.dx, d.dy, d.dz, F.f9.data()), "data.dat");
        io(ehd_io_data_end(f_end.c_str(), t, p.nx, p.ny, p.nz, F.f9.data()), "data_end.dat");
    }
    ehd_destroy(b);
    return 0;
}
