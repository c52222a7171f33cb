// sanitize_host.cpp -- drives the FP64 oracle (oracle/ehd_oracle.c) and the
// host file layer (ehd-convection-3d_amd/csrc/ehd_io.cpp) through every entry
// point on small grids, for a build with AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitizers.py compiles and runs it).
// Any out-of-bounds access, use after free, leak or undefined operation ends
// the program with the sanitizer's report; the checks below catch wrong
// results of the file round trip.  Prints "ok" at the end.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ehd.h"
#include "../../include/ehd_io.h"
#include "../../oracle/ehd_oracle.h"

#define CHECK(c)                                                                \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(3);                                                       \
        }                                                                       \
    } while (0)

static bool all_finite(const double* v, size_t n)
{
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

static std::vector<double> fields9(oref_box* b, size_t S)
{
    static const int which[9] = {OREF_F_RHO, OREF_F_Q, OREF_F_PHI, OREF_F_UX, OREF_F_UY,
                                 OREF_F_UZ,  OREF_F_EX, OREF_F_EY, OREF_F_EZ};
    std::vector<double> f(9 * S);
    for (int k = 0; k < 9; ++k) std::memcpy(f.data() + k * S, oref_field(b, which[k]), S * sizeof(double));
    return f;
}

static void oracle_box(int nx, int ny, int nz, double lam)
{
    oref_params p;
    CHECK(oref_params_init(&p, nx, ny, nz, 300.0, 10.0, 10.0, 4000.0, lam, 1e-6, 0.01, 1.0e4, 1.0e-4, 1600.0, 1.0,
                           0.0, 0.0) == 0);
    double T, M, C, Fe;
    oref_compute_parameters(&p, &T, &M, &C, &Fe);
    CHECK(std::fabs(T / 300.0 - 1) < 1e-12 && std::fabs(Fe / 4000.0 - 1) < 1e-12);
    const size_t S = (size_t)nx * ny * nz;
    oref_box* b = oref_create(&p);
    CHECK(b && oref_get_params(b)->nz == nz);

    // the reference start, its equilibrium and coupled steps with both Poisson solves
    oref_initialization(b);
    oref_init_equilibrium(b);
    oref_step(b, 2);
    oref_step_tridiag(b, 2);
    CHECK(all_finite(oref_field(b, OREF_F_Q), S) && all_finite(oref_field(b, OREF_F_UZ), S));
    CHECK(std::isfinite(oref_current(b)));

    // the phases one by one, as the step composes them
    oref_collide_save(b);
    oref_boundary(b);
    oref_stream(b);
    oref_bc_charge(b);
    oref_stream_collide_save(b);
    oref_extension(b);
    oref_fft3d(b, oref_field(b, OREF_F_SEXT), oref_field(b, OREF_F_FREQ), -1);
    oref_derivative(b);
    oref_fft3d(b, oref_field(b, OREF_F_FREQ), oref_field(b, OREF_F_PHIEXT), 1);
    oref_extract(b);
    oref_efield(b);
    oref_poisson(b);
    oref_poisson_tridiag(b);
    CHECK(all_finite(oref_field(b, OREF_F_PHI), S) && all_finite(oref_field(b, OREF_F_EZ), S));

    // every perturbation pattern, with and without the hydrostatic density
    for (int pat = 0; pat <= 4; ++pat) {
        oref_init_hydrostatic_pattern(b, 7 + pat, 1e-3, pat & 1, pat);
        oref_init_equilibrium(b);
        oref_step(b, 1);
    }
    oref_init_hydrostatic(b, 3, 1e-3, 1);

    // restart from fields (read_data semantics) and the rate override
    const std::vector<double> f = fields9(b, S);
    oref_set_fields(b, f.data());
    double r[4];
    oref_rates(b, r);
    oref_set_rates(b, r);
    oref_init_equilibrium(b);
    oref_step(b, 1);
    CHECK(all_finite(oref_field(b, OREF_F_F0), S) && all_finite(oref_field(b, OREF_F_F1), 26 * S));
    CHECK(all_finite(oref_field(b, OREF_F_H1), 26 * S) && all_finite(oref_field(b, OREF_F_F0BC), 2 * (size_t)nx * ny));
    oref_destroy(b);
}

static void oracle_tables()
{
    int c[81], opp[27], eqf[108], eqh[108], mom[54];
    double w[27];
    oref_tables(c, opp, w, eqf, eqh, mom);
    double sw = 0;
    for (int i = 0; i < 27; ++i) sw += w[i];
    CHECK(std::fabs(sw - 1.0) < 1e-15);
    oref_params p;
    CHECK(oref_params_init(&p, 8, 8, 9, 300.0, 10.0, 10.0, 4000.0, 0.25, 1e-6, 0.01, 1.0e4, 1.0e-4, 1600.0, 1.0, 0.0,
                           0.0) == 0);
    double fe[27], he[27], fp[27];
    oref_node_terms(&p, 1600.0, 1e-3, -2e-3, 3e-3, 0.5, 1.0, -2.0, 3.0, 0.1, 0.2, 0.3, fe, he, fp);
    CHECK(all_finite(fe, 27) && all_finite(he, 27) && all_finite(fp, 27));
    for (int n : {1, 2, 3, 5, 8, 12, 16, 31}) {
        std::vector<double> d(2 * n), e(2 * n);
        for (int i = 0; i < 2 * n; ++i) d[i] = e[i] = oref_hash_uniform(5, i, n, 0, 0);
        oref_fft1d(d.data(), n, -1);
        oref_fft1d(d.data(), n, 1);
        for (int i = 0; i < 2 * n; ++i) CHECK(std::fabs(d[i] / n - e[i]) < 1e-12);
    }
    double u[3];
    oref_pattern_velocity(4, 1e-3, 0.1, 0.2, 0.3, 1.0, 1.0, 1.0, u);
    CHECK(all_finite(u, 3) && std::isfinite(oref_hydrostatic_e0(10.0)));
}

static std::string read_all(const std::string& path)
{
    std::string s;
    FILE* f = std::fopen(path.c_str(), "rb");
    CHECK(f);
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return s;
}

static void file_layer(const std::string& dir)
{
    const int nx = 6, ny = 4, nz = 5;
    const size_t S = (size_t)nx * ny * nz;
    oref_params p;
    CHECK(oref_params_init(&p, nx, ny, nz, 300.0, 10.0, 10.0, 4000.0, 0.25, 1e-6, 0.01, 1.0e4, 1.0e-4, 1600.0, 1.0,
                           0.0, 0.0) == 0);
    oref_box* b = oref_create(&p);
    oref_init_hydrostatic(b, 11, 1e-2, 1);
    oref_init_equilibrium(b);
    oref_step(b, 2);
    const std::vector<double> f = fields9(b, S);
    const std::string data = dir + "/data.dat", end = dir + "/data_end.dat", dmd = dir + "/0.1.dat",
                      umax = dir + "/umax.dat";
    CHECK(ehd_io_tecplot(data.c_str(), 0, 1, 0.0, nx, ny, nz, p.dx, p.dy, p.dz, f.data()) == 0);
    CHECK(ehd_io_tecplot(data.c_str(), 1, 0, 0.1, nx, ny, nz, p.dx, p.dy, p.dz, f.data()) == 0);
    CHECK(ehd_io_data_end(end.c_str(), 0.25, nx, ny, nz, f.data()) == 0);
    CHECK(ehd_io_dmd(dmd.c_str(), 0.1, nx, ny, nz, f.data() + 5 * S) == 0);
    CHECK(ehd_io_umax(umax.c_str(), 0, 0.1, nx, ny, nz, f.data() + 5 * S) == 0);
    CHECK(ehd_io_umax(umax.c_str(), 1, 0.2, nx, ny, nz, f.data() + 5 * S) == 0);
    CHECK(std::isfinite(ehd_io_current(nx, ny, nz, p.K, p.dy, f.data() + S, f.data() + 8 * S)));

    // data_end.dat read back: the interior rows (no wall extrapolation) to the
    // six printed decimals of %10.6f
    std::vector<double> g(9 * S);
    double t = -1;
    CHECK(ehd_io_read_end(end.c_str(), nx, ny, nz, g.data(), &t) == 0);
    CHECK(std::fabs(t - 0.25) < 1e-9);
    for (int k = 0; k < 9; ++k)
        for (size_t i = (size_t)nx * ny; i < S - (size_t)nx * ny; ++i)
            CHECK(std::fabs(g[k * S + i] - f[k * S + i]) <= 5.000001e-7 * (1 + std::fabs(f[k * S + i])));

    // a truncated file, a missing file and a malformed token are errors
    const std::string whole = read_all(end);
    const std::string cut = dir + "/cut.dat";
    FILE* o = std::fopen(cut.c_str(), "wb");
    CHECK(o);
    std::fwrite(whole.data(), 1, whole.size() / 2, o);
    std::fclose(o);
    CHECK(ehd_io_read_end(cut.c_str(), nx, ny, nz, g.data(), &t) == EHD_EINVAL);
    CHECK(std::strlen(ehd_last_error()) > 0);
    CHECK(ehd_io_read_end((dir + "/missing.dat").c_str(), nx, ny, nz, g.data(), &t) == EHD_EINVAL);
    std::string bad = whole;
    bad.replace(bad.find('.') - 1, 1, "x");
    o = std::fopen(cut.c_str(), "wb");
    std::fwrite(bad.data(), 1, bad.size(), o);
    std::fclose(o);
    CHECK(ehd_io_read_end(cut.c_str(), nx, ny, nz, g.data(), &t) == EHD_EINVAL);
    // bad grids and paths are refused before any access
    CHECK(ehd_io_data_end(end.c_str(), 0.0, nx, ny, 2, f.data()) == EHD_EINVAL);
    CHECK(ehd_io_tecplot((dir + "/no/such/dir/x.dat").c_str(), 0, 1, 0.0, nx, ny, nz, p.dx, p.dy, p.dz, f.data()) ==
          EHD_EINVAL);
    oref_destroy(b);
}

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <scratch dir>\n", argv[0]);
        return 2;
    }
    oref_set_threads(1);
    oracle_tables();
    oracle_box(8, 6, 7, 1.0 / 12.0);
    oracle_box(4, 4, 5, 0.25);
    oracle_box(6, 10, 4, 0.25);
    file_layer(argv[1]);
    std::printf("ok\n");
    return 0;
}
