// ehd_io.cpp -- host-only writers and reader of the driver's files
// (LBM.cu:1427-1713, main.cu:166-277).  No HIP here: the fields arrive as
// host arrays, so the formats are testable without a GPU.
#include <cstdio>
#include <vector>

#include "../../include/ehd.h"
#include "../../include/ehd_io.h"
#include "ehd_error.h"

using ehd::set_error;

namespace {

enum { RHO, Q, PHI, UX, UY, UZ, EX, EY, EZ };

struct Grid {
    int nx, ny, nz;
    size_t S() const { return (size_t)nx * ny * nz; }
    size_t idx(int x, int y, int z) const { return (size_t)nx * ((size_t)ny * z + y) + x; }
};

bool grid_ok(int nx, int ny, int nz)
{
    return nx >= 1 && ny >= 1 && nz >= 3;
}

// 2 v[1] - v[2] at z = 0 and 2 v[nz-2] - v[nz-3] at z = nz-1 (LBM.cu:1458-1471)
void extrapolate(const Grid& g, double* v, bool bottom, bool top)
{
    for (int y = 0; y < g.ny; ++y)
        for (int x = 0; x < g.nx; ++x) {
            if (bottom) v[g.idx(x, y, 0)] = 2.0 * v[g.idx(x, y, 1)] - v[g.idx(x, y, 2)];
            if (top) v[g.idx(x, y, g.nz - 1)] = 2.0 * v[g.idx(x, y, g.nz - 2)] - v[g.idx(x, y, g.nz - 3)];
        }
}

// private copy of fields9 with rho, q, u, v, w extrapolated at both walls
std::vector<double> walls_extrapolated(const Grid& g, const double* f9)
{
    std::vector<double> c(f9, f9 + 9 * g.S());
    for (int k : {RHO, Q, UX, UY, UZ}) extrapolate(g, c.data() + k * g.S(), true, true);
    return c;
}

FILE* open_out(const char* path, bool append)
{
    return path ? std::fopen(path, append ? "ab" : "wb+") : nullptr;
}

int close_out(FILE* f, const char* path)
{
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) return set_error(EHD_EINVAL, "write to %s failed", path);
    return EHD_OK;
}

}  // namespace

extern "C" int ehd_io_tecplot(const char* path, int append, int first, double t, int nx, int ny, int nz, double dx,
                              double dy, double dz, const double* fields9)
{
    const Grid g{nx, ny, nz};
    if (!fields9 || !grid_ok(nx, ny, nz)) return set_error(EHD_EINVAL, "ehd_io_tecplot: bad grid or null fields");
    FILE* f = open_out(path, append != 0);
    if (!f) return set_error(EHD_EINVAL, "cannot open %s", path ? path : "(null)");
    const std::vector<double> c = walls_extrapolated(g, fields9);
    const size_t S = g.S();
    if (first) std::fprintf(f, "%s\n", "VARIABLES=\"x\",\"y\",\"z\",\"u\",\"v\",\"w\",\"p\",\"charge\",\"phi\",\"Ex\",\"Ey\",\"Ez\"");
    std::fprintf(f, "\n");
    std::fprintf(f, "ZONE T=\"t=%g\", F=POINT, I = %d, J = %d, K = %d\n", t, nx, ny, nz);
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t i = g.idx(x, y, z);
                std::fprintf(f, "%g %g %g %g %g %g %g %g %10.6f %10.6f %10.6f %10.6f\n", dx * x, dy * y, dz * z,
                             c[UX * S + i], c[UY * S + i], c[UZ * S + i], c[RHO * S + i], c[Q * S + i],
                             c[PHI * S + i], c[EX * S + i], c[EY * S + i], c[EZ * S + i]);
            }
    return close_out(f, path);
}

extern "C" int ehd_io_data_end(const char* path, double t, int nx, int ny, int nz, const double* fields9)
{
    const Grid g{nx, ny, nz};
    if (!fields9 || !grid_ok(nx, ny, nz)) return set_error(EHD_EINVAL, "ehd_io_data_end: bad grid or null fields");
    FILE* f = open_out(path, false);
    if (!f) return set_error(EHD_EINVAL, "cannot open %s", path ? path : "(null)");
    const std::vector<double> c = walls_extrapolated(g, fields9);
    const size_t S = g.S();
    for (size_t i = 0; i < S; ++i)
        std::fprintf(f, "%10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f %10.6f\n", t, c[UX * S + i],
                     c[UY * S + i], c[UZ * S + i], c[RHO * S + i], c[Q * S + i], c[PHI * S + i], c[EX * S + i],
                     c[EY * S + i], c[EZ * S + i]);
    return close_out(f, path);
}

extern "C" int ehd_io_read_end(const char* path, int nx, int ny, int nz, double* fields9, double* t)
{
    const Grid g{nx, ny, nz};
    if (!fields9 || !t || !grid_ok(nx, ny, nz)) return set_error(EHD_EINVAL, "ehd_io_read_end: bad arguments");
    FILE* f = path ? std::fopen(path, "r") : nullptr;
    if (!f) return set_error(EHD_EINVAL, "cannot open %s", path ? path : "(null)");
    const size_t S = g.S();
    // file columns t, u, v, w, rho, charge, phi, Ex, Ey, Ez
    static const int col_to_field[9] = {UX, UY, UZ, RHO, Q, PHI, EX, EY, EZ};
    for (size_t i = 0; i < S; ++i) {
        double v[10];
        if (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6],
                        &v[7], &v[8], &v[9]) != 10) {
            std::fclose(f);
            return set_error(EHD_EINVAL, "%s: row %zu of %zu missing or malformed", path, i, S);
        }
        *t = v[0];
        for (int k = 0; k < 9; ++k) fields9[col_to_field[k] * S + i] = v[k + 1];
    }
    std::fclose(f);
    return EHD_OK;
}

extern "C" int ehd_io_dmd(const char* path, double t, int nx, int ny, int nz, const double* uz)
{
    const Grid g{nx, ny, nz};
    if (!uz || !grid_ok(nx, ny, nz)) return set_error(EHD_EINVAL, "ehd_io_dmd: bad grid or null field");
    FILE* f = open_out(path, false);
    if (!f) return set_error(EHD_EINVAL, "cannot open %s", path ? path : "(null)");
    std::vector<double> w(uz, uz + g.S());
    extrapolate(g, w.data(), true, true);
    for (double v : w) std::fprintf(f, "%10.6f %10.6f\n", t, v);
    return close_out(f, path);
}

extern "C" int ehd_io_umax(const char* path, int append, double t, int nx, int ny, int nz, const double* uz)
{
    const Grid g{nx, ny, nz};
    if (!uz || !grid_ok(nx, ny, nz)) return set_error(EHD_EINVAL, "ehd_io_umax: bad grid or null field");
    FILE* f = open_out(path, append != 0);
    if (!f) return set_error(EHD_EINVAL, "cannot open %s", path ? path : "(null)");
    std::vector<double> w(uz, uz + g.S());
    extrapolate(g, w.data(), false, true);
    double umax = 0.0;
    for (double v : w) umax = (umax > v) ? umax : v;
    std::fprintf(f, "%10.6f %10.6f\n", t, umax);
    return close_out(f, path);
}

extern "C" double ehd_io_current(int nx, int ny, int nz, double K, double dy, const double* q, const double* ez)
{
    const Grid g{nx, ny, nz};
    if (!q || !ez || !grid_ok(nx, ny, nz)) return 0.0;
    double I = 0.0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const double qt = 2.0 * q[g.idx(x, y, nz - 2)] - q[g.idx(x, y, nz - 3)];
            I += qt * ez[g.idx(x, y, nz - 1)];
        }
    return I * K * dy * dy;
}
