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

struct Fields {
    int nx, ny, nz;
    std::vector<double> rho, q, phi, ux, uy, uz, ex, ey, ez;
    explicit Fields(const ehd_params& p) : nx(p.nx), ny(p.ny), nz(p.nz)
    {
        const size_t S = (size_t)nx * ny * nz;
        for (auto* v : {&rho, &q, &phi, &ux, &uy, &uz, &ex, &ey, &ez}) v->assign(S, 0.0);
    }
    size_t idx(int x, int y, int z) const { return (size_t)nx * ((size_t)ny * z + y) + x; }
    void read(ehd_box* b)
    {
        CALL(ehd_read_fields(b, rho.data(), q.data(), phi.data(), ux.data(), uy.data(), uz.data(), ex.data(),
                             ey.data(), ez.data()));
    }
    // linear extrapolation of the wall rows (LBM.cu:1457-1472)
    void extrapolate(std::vector<double>& v, bool bottom, bool top) const
    {
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                if (bottom) v[idx(x, y, 0)] = 2.0 * v[idx(x, y, 1)] - v[idx(x, y, 2)];
                if (top) v[idx(x, y, nz - 1)] = 2.0 * v[idx(x, y, nz - 2)] - v[idx(x, y, nz - 3)];
            }
    }
};

// save_data_tecplot (LBM.cu:1427-1503)
void save_tecplot(FILE* f, double t, ehd_box* b, const ehd_params& p, const ehd_derived& d, bool first)
{
    Fields F(p);
    F.read(b);
    for (auto* v : {&F.rho, &F.q, &F.ux, &F.uy, &F.uz}) F.extrapolate(*v, true, true);
    if (first) std::fprintf(f, "%s\n", "VARIABLES=\"x\",\"y\",\"z\",\"u\",\"v\",\"w\",\"p\",\"charge\",\"phi\",\"Ex\",\"Ey\",\"Ez\"");
    std::fprintf(f, "\n");
    std::fprintf(f, "ZONE T=\"t=%g\", F=POINT, I = %d, J = %d, K = %d\n", t, p.nx, p.ny, p.nz);
    for (int z = 0; z < p.nz; ++z)
        for (int y = 0; y < p.ny; ++y)
            for (int x = 0; x < p.nx; ++x) {
                const size_t i = F.idx(x, y, z);
                std::fprintf(f, "%g %g %g %g %g %g %g %g %10.6f %10.6f %10.6f %10.6f\n", d.dx * x, d.dy * y, d.dz * z,
                             F.ux[i], F.uy[i], F.uz[i], F.rho[i], F.q[i], F.phi[i], F.ex[i], F.ey[i], F.ez[i]);
            }
}

// save_data_end (LBM.cu:1505-1566)
void save_end(FILE* f, double t, ehd_box* b, const ehd_params& p)
{
    Fields F(p);
    F.read(b);
    for (auto* v : {&F.rho, &F.q, &F.ux, &F.uy, &F.uz}) F.extrapolate(*v, true, true);
    for (int z = 0; z < p.nz; ++z)
        for (int y = 0; y < p.ny; ++y)
            for (int x = 0; x < p.nx; ++x) {
                const size_t i = F.idx(x, y, z);
                std::fprintf(f, "%10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f\n", t, F.ux[i],
                             F.uy[i], F.uz[i], F.rho[i], F.q[i], F.phi[i], F.ex[i], F.ey[i], F.ez[i]);
            }
}

// save_data_dmd (LBM.cu:1568-1594)
void save_dmd(FILE* f, double t, ehd_box* b, const ehd_params& p)
{
    Fields F(p);
    F.read(b);
    F.extrapolate(F.uz, true, true);
    for (size_t i = 0; i < F.uz.size(); ++i) std::fprintf(f, "%10.6f %10.6f\n", t, F.uz[i]);
}

// record_umax (LBM.cu:1673-1713)
void record_umax(FILE* f, double t, ehd_box* b, const ehd_params& p)
{
    Fields F(p);
    F.read(b);
    F.extrapolate(F.uz, false, true);
    double umax = 0;
    for (double v : F.uz) umax = (umax > v) ? umax : v;
    std::fprintf(f, "%10.6f %10.6f\n", t, umax);
}

// read_data (LBM.cu:1596-1643): data_end.dat -> fields, returns time
bool read_restart(const std::string& path, const ehd_params& p, std::vector<double>& f9, double& t)
{
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return false;
    const size_t S = (size_t)p.nx * p.ny * p.nz;
    f9.assign(9 * S, 0.0);
    for (size_t i = 0; i < S; ++i) {
        double v[10];
        if (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                        &v[7], &v[8], &v[9]) != 10) {
            std::fclose(f);
            return false;
        }
        t = v[0];
        // file columns: t, u, v, w, rho, charge, phi, Ex, Ey, Ez -> rho, q, phi, u, v, w, Ex, Ey, Ez
        f9[0 * S + i] = v[4];
        f9[1 * S + i] = v[5];
        f9[2 * S + i] = v[6];
        f9[3 * S + i] = v[1];
        f9[4 * S + i] = v[2];
        f9[5 * S + i] = v[3];
        f9[6 * S + i] = v[7];
        f9[7 * S + i] = v[8];
        f9[8 * S + i] = v[9];
    }
    std::fclose(f);
    return true;
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
        if (!read_restart(o.outdir + "/data_end.dat", p, f9, t)) {
            std::fprintf(stderr, "Error: cannot read %s/data_end.dat\n", o.outdir.c_str());
            return 1;
        }
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

    FILE* fout = nullptr;
    FILE* fumax = std::fopen((o.outdir + "/umax.dat").c_str(), "wb+");
    if (!o.quiet_files) {
        fout = std::fopen((o.outdir + "/data.dat").c_str(), "wb+");
        if (fout) save_tecplot(fout, t, b, p, d, true);
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
        t += d.dt * (double)n;
        i = j;
        if (fout && i % o.nsave == 1) save_tecplot(fout, t, b, p, d, false);
        if (i % o.ndmd == 1) {
            std::printf("Iteration: %lld, physical time: %g.\n", i, t);
            if (!o.quiet_files) {
                char name[128];
                std::snprintf(name, sizeof(name), "%g.dat", t);
                FILE* f2 = std::fopen((o.outdir + "/" + name).c_str(), "wb+");
                if (f2) {
                    save_dmd(f2, t, b, p);
                    std::fclose(f2);
                }
            }
        }
        if (i % o.print_current == 1) {
            ehd_diag dg;
            CALL(ehd_diagnostics(b, &dg));
            std::printf("Current = %g\n", dg.current);
            if (fumax) record_umax(fumax, t, b, p);
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

    if (fout) {
        save_tecplot(fout, t, b, p, d, false);
        std::fclose(fout);
    }
    if (fumax) {
        record_umax(fumax, t, b, p);
        std::fclose(fumax);
    }
    if (!o.quiet_files) {
        FILE* fend = std::fopen((o.outdir + "/data_end.dat").c_str(), "wb+");
        if (fend) {
            save_end(fend, t, b, p);
            std::fclose(fend);
        }
    }
    ehd_destroy(b);
    return 0;
}
