// ehd_api.cpp -- the C-ABI (include/ehd.h): box lifetime, parameter
// derivation, initial states, the coupled step loop, field I/O and
// diagnostics.  Host C++ over the HIP runtime; one HIP stream per box.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "box.h"
#include "fft_plan.h"

using namespace ehd;

namespace {



constexpr double kPi = 3.14159265358979323846;

}  // namespace


namespace {

hipEvent_t take_event(ehd_box* b)
{
    if (!b->ev_pool.empty()) {
        hipEvent_t e = b->ev_pool.back();
        b->ev_pool.pop_back();
        return e;
    }
    // timing-only events: no system-scope fence (the cache writeback and
    // invalidation it costs measured 2.4% of a 256x256x128 step with four
    // events per step); they are read after the stream has been synchronised
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

int resolve_events(ehd_box* b)
{
    if (b->lbm_pairs.empty() && b->poi_pairs.empty() && b->step_pairs.empty()) return EHD_OK;
    HIPCHK(hipStreamSynchronize(b->st));
    auto drain = [&](std::vector<std::pair<hipEvent_t, hipEvent_t>>& v, double& acc, long long& n) -> int {
        for (auto& pr : v) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, pr.first, pr.second));
            acc += ms;
            n += 1;
            b->ev_pool.push_back(pr.first);
            b->ev_pool.push_back(pr.second);
        }
        v.clear();
        return EHD_OK;
    };
    int rc;
    if ((rc = drain(b->lbm_pairs, b->lbm_ms, b->lbm_n)) != EHD_OK) return rc;
    if ((rc = drain(b->poi_pairs, b->poi_ms, b->poi_n)) != EHD_OK) return rc;
    if ((rc = drain(b->step_pairs, b->step_ms, b->step_n)) != EHD_OK) return rc;
    return EHD_OK;
}

int validate(const ehd_params* p)
{
    if (!p) return set_error(EHD_EINVAL, "null params");
    if (p->nx < 2 || p->ny < 2 || (p->nx & 1) || (p->ny & 1))
        return set_error(EHD_EINVAL, "nx and ny must be even and >= 2 (got %d, %d)", p->nx, p->ny);
    if (p->nz < 4) return set_error(EHD_EINVAL, "nz must be >= 4 (got %d)", p->nz);
    const double S = (double)p->nx * p->ny * p->nz;
    if (S * 27.0 >= 4294967295.0) return set_error(EHD_EINVAL, "grid too large: 27*nx*ny*nz must fit 32 bits");
    if (!(p->T > 0) || !(p->C > 0) || !(p->M > 0) || !(p->Fe > 0))
        return set_error(EHD_EINVAL, "T, C, M, Fe must be positive");
    if (!(p->cfl > 0) || !(p->voltage > 0) || !(p->eps > 0) || !(p->rho0 > 0) || !(p->lz > 0))
        return set_error(EHD_EINVAL, "cfl, voltage, eps, rho0, lz must be positive");
    if (!(p->lambda_f > 0) || !(p->lambda_c > 0)) return set_error(EHD_EINVAL, "TRT magic parameters must be positive");
    if (p->precision != EHD_FP64 && p->precision != EHD_FP32S) return set_error(EHD_EINVAL, "unknown precision");
    if (p->init_mode < EHD_INIT_REFERENCE || p->init_mode > EHD_INIT_FIELDS) return set_error(EHD_EINVAL, "unknown init mode");
    if (p->perturb_pattern < EHD_PERTURB_NOISE || p->perturb_pattern > EHD_PERTURB_HEXAGONS)
        return set_error(EHD_EINVAL, "unknown perturbation pattern");
    return EHD_OK;
}

void fill_lbm_params(ehd_box* b)
{
    const ehd_params& p = b->prm;
    const ehd_derived& d = b->der;
    LbmParams& L = b->lp;
    L.nx = p.nx; L.ny = p.ny; L.nz = p.nz;
    L.nxy = (unsigned)(p.nx * p.ny);
    L.S = (unsigned)((size_t)p.nx * p.ny * p.nz);
    L.dt = d.dt; L.cfl = p.cfl; L.cs_square = d.cs_square; L.K = d.K; L.exf = p.exf;
    L.charge0 = d.charge0; L.rho0 = p.rho0;
    L.w[0] = 8.0 / 27.0; L.w[1] = 2.0 / 27.0; L.w[2] = 1.0 / 54.0; L.w[3] = 1.0 / 216.0;
    // LBM.cu:299-302
    const double nu = d.nu, cs_square = d.cs_square, dt = d.dt, diffu = d.diffu;
    L.omega_plus = 1.0 / (nu / cs_square / dt + 1.0 / 2.0) / dt;
    L.omega_minus = 1.0 / (p.lambda_f / (nu / cs_square / dt) + 1.0 / 2.0) / dt;
    L.omega_c_minus = 1.0 / (diffu / cs_square / dt + 1.0 / 2.0) / dt;
    L.omega_c_plus = 1.0 / (p.lambda_c / (diffu / cs_square / dt) + 1.0 / 2.0) / dt;
    // LBM.cu:1137-1139 and 1288-1291
    for (int c = 0; c < 4; ++c) {
        L.multi[c] = 2.0 * p.rho0 * p.uw / cs_square * L.w[c] / p.cfl;
        L.inj[c] = 2.0 * d.charge0 * L.w[c];
    }
    L.dx = d.dx; L.dy = d.dy; L.dz = d.dz;
    L.rs = nullptr; L.qs = nullptr; L.shf = nullptr; L.qs0 = 0.0;
}

template <typename T>
void make_twiddles(int n, std::vector<T>& out)
{
    out.resize(2 * (size_t)n);
    for (int j = 0; j < n; ++j) {
        const double a = 2.0 * kPi * (double)j / (double)n;
        out[2 * j] = (T)std::cos(a);
        out[2 * j + 1] = (T)(-std::sin(a));
    }
}

int alloc(void** p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        return set_error(EHD_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    return EHD_OK;
}

void free_box(ehd_box* b)
{
    void* ptrs[] = {b->f[0], b->f[1], b->f0[0], b->f0[1], b->h[0], b->h[1], b->h0[0], b->h0[1],
                    b->rho, b->ux, b->uy, b->uz, b->q, b->phi, b->phibar, b->ex, b->ey, b->ez,
                    b->spec, b->twx, b->twy, b->diag_scr, b->diag_out,
                    b->shift_rs, b->shift_qs, b->shift_f, b->pjob, b->lbm_items[0], b->lbm_items[1]};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& pr : b->lbm_pairs) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (auto& pr : b->poi_pairs) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (auto& pr : b->step_pairs) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (hipEvent_t e : b->ev_pool) (void)hipEventDestroy(e);
    if (b->st) (void)hipStreamDestroy(b->st);
}

}  // namespace

namespace ehd {

PopBufs pop_bufs(ehd_box* b)
{
    const int A = b->cur, B = 1 - b->cur;
    return PopBufs{b->f[A], b->f0[A], b->h[A], b->h0[A], b->f[B], b->f0[B], b->h[B], b->h0[B]};
}

MacroBufs macro_bufs(ehd_box* b)
{
    return MacroBufs{b->rho, b->ux, b->uy, b->uz, b->q, b->phi, b->phibar, b->ex, b->ey, b->ez};
}

// launches of this step are bracketed by timing events
bool profiled_step(const ehd_box* b) { return b->prof_every > 0 && b->step % b->prof_every == 0; }

int run_lbm(ehd_box* b, bool macro, void* const* target_items)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool prof = profiled_step(b);
    if (prof) {
        e0 = take_event(b);
        e1 = take_event(b);
        if (!e0 || !e1) return set_error(EHD_EHIP, "hipEventCreate failed");
        HIPCHK(hipEventRecord(e0, b->st));
    }
    hipError_t e;
    if (b->prec == Prec::F64) e = launch_lbm_f64(b->lp, pop_bufs(b), macro_bufs(b), b->direct, b->ebuf, macro, b->st);
    else e = launch_lbm_f32s_batch(target_items ? target_items[b->cur] : b->lbm_items[b->cur], 1,
                                   (unsigned)((b->S + 255) / 256), 0, b->lbm_chunk, b->direct, b->ebuf, macro, b->st,
                                   target_items != nullptr);
    if (e != hipSuccess) return set_error(EHD_EHIP, "LBM launch failed: %s", hipGetErrorString(e));
    if (b->debug_sync) {
        e = hipStreamSynchronize(b->st);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess)
            return set_error(EHD_EHIP, "LBM launch of step %lld failed: %s", b->step, hipGetErrorString(e));
    }
    if (prof) {
        HIPCHK(hipEventRecord(e1, b->st));
        b->lbm_pairs.emplace_back(e0, e1);
    }
    b->cur = 1 - b->cur;
    b->direct = false;
    return EHD_OK;
}

PoissonGroup poisson_group_of(const ehd_box* b)
{
    PoissonGroup g{};
    g.nx = b->nx;
    g.ny = b->ny;
    g.max_n = b->nz - 2;
    g.njobs = 1;
    g.split_mean = b->prec == Prec::F32S ? 1 : 0;
    g.lpb = b->tune_lpb;
    g.ylines = b->tune_ylines;
    g.zs_partition = poisson_shape(b->prec, b->nx, b->ny, b->nz).zs_partition;
    g.twx = b->twx;
    g.twy = b->twy;
    g.jobs = b->pjob;
    g.debug_sync = b->debug_sync ? 1 : 0;
    return g;
}

int run_poisson(ehd_box* b)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool prof = profiled_step(b);
    if (prof) {
        e0 = take_event(b);
        e1 = take_event(b);
        if (!e0 || !e1) return set_error(EHD_EHIP, "hipEventCreate failed");
        HIPCHK(hipEventRecord(e0, b->st));
    }
    hipError_t e = launch_poisson_group(b->prec, poisson_group_of(b), b->st);
    if (e != hipSuccess) {
        if (b->debug_sync)
            return set_error(EHD_EHIP, "Poisson solve of step %lld failed in the %s: %s", b->step,
                             poisson_failed_pass(), hipGetErrorString(e));
        return set_error(EHD_EHIP, "Poisson launch failed: %s", hipGetErrorString(e));
    }
    if (prof) {
        HIPCHK(hipEventRecord(e1, b->st));
        b->poi_pairs.emplace_back(e0, e1);
    }
    b->ebuf = false;
    return EHD_OK;
}

int ConsistentTargets::make(ehd_box* b, const double* fields9)
{
    const size_t S = b->S;
    st = b->st;
    int rc = alloc((void**)&tgt, 4 * S * sizeof(double));
    if (rc) return rc;
    // fields9: rho, q, phi, ux, uy, uz, ex, ey, ez; targets: ux, uy, uz, q
    const int src[4] = {3, 4, 5, 1};
    for (int k = 0; k < 4; ++k)
        HIPCHK(hipMemcpyAsync(tgt + k * S, fields9 + src[k] * S, S * sizeof(double), hipMemcpyHostToDevice, st));
    const size_t isz = lbm_f32s_batch_item_size();
    std::vector<unsigned char> item(isz);
    for (int c = 0; c < 2; ++c) {
        const int A = c, B = 1 - c;
        const PopBufs pb{b->f[A], b->f0[A], b->h[A], b->h0[A], b->f[B], b->f0[B], b->h[B], b->h0[B]};
        lbm_f32s_batch_item(item.data(), b->lp, pb, macro_bufs(b), 0, tgt);
        if ((rc = alloc(&items[c], isz))) return rc;
        HIPCHK(hipMemcpyAsync(items[c], item.data(), isz, hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    return EHD_OK;
}

ConsistentTargets::~ConsistentTargets()
{
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : items)
        if (p) (void)hipFree(p);
    if (tgt) (void)hipFree(tgt);
}

}  // namespace ehd

namespace {

// Hydrostatic unipolar-injection profile (D = 0, u = 0), nondimensional:
// E^2 = E0^2 + 2 C E0 zeta, q = C E0/E, phi = 1 - (E^3 - E0^3)/(3 C E0),
// with E0 fixed by phi(1) = 0 (bisection to the last bit).
double hydrostatic_e0(double C)
{
    double lo = 0.0, hi = 1.0;
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        const double e1 = std::sqrt(mid * mid + 2.0 * C * mid);
        const double g = (e1 * e1 * e1 - mid * mid * mid) / (3.0 * C * mid) - 1.0;
        if (g > 0.0) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

}  // namespace

namespace {
// Analytic unipolar-injection profile sampled on the grid: q(z), phi(z) and
// rho(z) in one table of 3*nz doubles.  Closed form of the 1-D hydrostatic
// state: E(zeta)^2 = E0^2 + 2 C E0 zeta (nondimensional), q = C E0/E, phi from
// integrating E, with E0 fixed by phi(1) = 0.
std::vector<double> hydrostatic_table(const ehd_params& p, const ehd_derived& d)
{
    const int nz = p.nz;
    std::vector<double> tab(3 * (size_t)nz);
    const double C = d.charge0 * p.lz * p.lz / (p.voltage * p.eps);
    const double e0 = hydrostatic_e0(C);
    const double qs = p.eps * p.voltage / (p.lz * p.lz);
    // uniform current density J = q Ez balances the pressure: dp/dz = J
    const double J = C * e0 * p.eps * p.voltage * p.voltage / (p.lz * p.lz * p.lz);
    for (int z = 0; z < nz; ++z) {
        const double zeta = d.dz * z / p.lz;
        const double e = std::sqrt(e0 * e0 + 2.0 * C * e0 * zeta);
        tab[z] = C * e0 / e * qs;
        double ph = p.voltage * (1.0 - (e * e * e - e0 * e0 * e0) / (3.0 * C * e0));
        if (z == 0) ph = p.voltage;
        if (z == nz - 1) ph = 0.0;
        tab[nz + z] = ph;
        tab[2 * nz + z] = p.hydro_density ? p.rho0 + J * (d.dz * z - 0.5 * p.lz) / d.cs_square : p.rho0;
    }
    return tab;
}

// current() of LBM.cu:1646-1671 on the sampled hydrostatic profile: q
// extrapolated to the top wall, Ez of the wall copy rule (poisson.cu:62-73).
double hydrostatic_current(const ehd_params& p, const ehd_derived& d)
{
    const std::vector<double> tab = hydrostatic_table(p, d);
    const int nz = p.nz;
    const double qt = 2.0 * tab[nz - 2] - tab[nz - 3];
    const double ez = 0.5 * (tab[nz + nz - 3] - tab[nz + nz - 1]) / d.dz;
    return (double)p.nx * (double)p.ny * (qt * ez) * d.K * d.dy * d.dy;
}
}  // namespace

namespace {
struct CkptHeader {
    char magic[8];
    int abi, precision, nx, ny, nz, direct, ebuf, pad;
    long long step;
    double t;
    double T, C, M, Fe, lambda_f, lambda_c;
    double cfl, voltage, eps, rho0, lz, uw, exf;
};

void ckpt_physics(const ehd_params& p, CkptHeader& h)
{
    h.T = p.T; h.C = p.C; h.M = p.M; h.Fe = p.Fe;
    h.lambda_f = p.lambda_f; h.lambda_c = p.lambda_c;
    h.cfl = p.cfl; h.voltage = p.voltage; h.eps = p.eps; h.rho0 = p.rho0; h.lz = p.lz;
    h.uw = p.uw; h.exf = p.exf;
}

// first physical parameter in which a checkpoint differs from the box, or null
const char* ckpt_mismatch(const CkptHeader& h, const ehd_params& p)
{
    CkptHeader w{};
    ckpt_physics(p, w);
    const struct {
        const char* name;
        double a, b;
    } f[] = {{"T", h.T, w.T},       {"C", h.C, w.C},           {"M", h.M, w.M},
             {"Fe", h.Fe, w.Fe},    {"lambda_f", h.lambda_f, w.lambda_f},
             {"lambda_c", h.lambda_c, w.lambda_c}, {"cfl", h.cfl, w.cfl}, {"voltage", h.voltage, w.voltage},
             {"eps", h.eps, w.eps}, {"rho0", h.rho0, w.rho0},  {"lz", h.lz, w.lz},
             {"uw", h.uw, w.uw},    {"exf", h.exf, w.exf}};
    for (const auto& x : f)
        if (!(x.a == x.b)) return x.name;
    return nullptr;
}

int ckpt_io(ehd_box* b, FILE* f, bool save)
{
    const size_t S = b->S, es = b->esz;
    const int A = b->cur, B = 1 - b->cur;
    struct Item {
        void* dev;
        size_t bytes;
    } items[] = {
        {b->f[A], 26 * S * es}, {b->f0[A], S * es}, {b->h[A], 26 * S * es}, {b->h0[A], S * es},
        {b->f0[B], S * es},    // frozen wall rest populations live in both buffers
        {b->rho, S * es}, {b->ux, S * es}, {b->uy, S * es}, {b->uz, S * es}, {b->q, S * es}, {b->phi, S * es},
        {b->phibar, (size_t)b->nz * sizeof(double)},
        {b->ex, S * sizeof(double)}, {b->ey, S * sizeof(double)}, {b->ez, S * sizeof(double)},
    };
    std::vector<unsigned char> buf;
    for (auto& it : items) {
        buf.resize(it.bytes);
        if (save) {
            HIPCHK(hipMemcpy(buf.data(), it.dev, it.bytes, hipMemcpyDeviceToHost));
            if (std::fwrite(buf.data(), 1, it.bytes, f) != it.bytes) return set_error(EHD_EINVAL, "checkpoint write failed");
        } else {
            if (std::fread(buf.data(), 1, it.bytes, f) != it.bytes) return set_error(EHD_EINVAL, "checkpoint truncated");
            HIPCHK(hipMemcpy(it.dev, buf.data(), it.bytes, hipMemcpyHostToDevice));
        }
    }
    return EHD_OK;
}
}  // namespace

extern "C" {

int ehd_abi_version(void) { return EHD_ABI_VERSION; }

int ehd_params_default(ehd_params* p)
{
    if (!p) return set_error(EHD_EINVAL, "null params");
    std::memset(p, 0, sizeof(*p));
    p->nx = 122; p->ny = 122; p->nz = 101;                     // LBM.h:34-36
    const double K = 2.5e-5, eps = 1.0e-4, V = 1.0e4, nu = 0.0251004, rho0 = 1600.0, q0 = 10.0,
                 D = 6.25e-5, Lz = 1.0;                        // LBM.h:54-62
    p->M = std::sqrt(eps / rho0) / K;                          // LBM.cu:1375-1378
    p->T = eps * V / K / nu / rho0;
    p->C = q0 * Lz * Lz / (V * eps);
    p->Fe = K * V / D;
    p->Pr = 10.0;
    p->lambda_f = 1.0 / 12.0; p->lambda_c = 1.0e-6;            // LBM.h:79-80
    p->cfl = 0.01; p->voltage = V; p->eps = eps; p->rho0 = rho0; p->lz = Lz;
    p->uw = 0.0; p->exf = 0.0;
    p->precision = EHD_FP64;
    p->init_mode = EHD_INIT_REFERENCE;
    p->seed = 1;
    p->perturb_amp = 1.0e-3;
    return EHD_OK;
}

int ehd_derive(const ehd_params* p, ehd_derived* d)
{
    int rc = validate(p);
    if (rc) return rc;
    if (!d) return set_error(EHD_EINVAL, "null output");
    std::memset(d, 0, sizeof(*d));
    d->dz = p->lz / (double)(p->nz - 1);
    d->dx = d->dz;
    d->dy = d->dz;
    d->lx = p->nx * d->dx;
    d->ly = p->ny * d->dy;
    d->dt = p->cfl * d->dx;                       // LBM.h:52
    d->cs_square = 1.0 / 3.0 / (p->cfl * p->cfl);  // LBM.h:53
    d->K = std::sqrt(p->eps / p->rho0) / p->M;
    d->nu = p->eps * p->voltage / (d->K * p->rho0 * p->T);
    d->charge0 = p->C * p->voltage * p->eps / (p->lz * p->lz);
    d->diffu = d->K * p->voltage / p->Fe;
    const double tp = d->nu / d->cs_square / d->dt;      // tau+ - 1/2
    const double tcm = d->diffu / d->cs_square / d->dt;  // tau_c- - 1/2
    d->tau_plus = tp + 0.5;
    d->tau_minus = p->lambda_f / tp + 0.5;
    d->tau_c_minus = tcm + 0.5;
    d->tau_c_plus = p->lambda_c / tcm + 0.5;
    return EHD_OK;
}

int ehd_compute_parameters(const ehd_derived* d, const ehd_params* p, double* T, double* M, double* C, double* Fe)
{
    if (!d || !p || !T || !M || !C || !Fe) return set_error(EHD_EINVAL, "null argument");
    *M = std::sqrt(p->eps / p->rho0) / d->K;     // LBM.cu:1375-1378
    *T = p->eps * p->voltage / d->K / d->nu / p->rho0;
    *C = d->charge0 * p->lz * p->lz / (p->voltage * p->eps);
    *Fe = d->K * p->voltage / d->diffu;
    return EHD_OK;
}

int ehd_create(const ehd_params* p, ehd_box** out)
{
    if (!out) return set_error(EHD_EINVAL, "null output");
    *out = nullptr;
    int rc = validate(p);
    if (rc) return rc;
    ehd_box* b = new (std::nothrow) ehd_box();
    if (!b) return set_error(EHD_ENOMEM, "host allocation failed");
    b->prm = *p;
    ehd_derive(p, &b->der);
    b->prec = (p->precision == EHD_FP64) ? Prec::F64 : Prec::F32S;
    b->nx = p->nx; b->ny = p->ny; b->nz = p->nz; b->nxh = p->nx / 2 + 1;
    b->nxy = (size_t)p->nx * p->ny;
    b->S = b->nxy * p->nz;
    b->esz = (b->prec == Prec::F64) ? 8 : 4;
    fill_lbm_params(b);

    auto bail = [&](int code) {
        free_box(b);
        delete b;
        return code;
    };
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
        return bail(set_error(EHD_EHIP, "hipStreamCreate failed: %s", hipGetErrorString(hipGetLastError())));
    const size_t S = b->S, es = b->esz;
    for (int i = 0; i < 2; ++i) {
        if ((rc = alloc(&b->f[i], 26 * S * es)) || (rc = alloc(&b->f0[i], S * es)) ||
            (rc = alloc(&b->h[i], 26 * S * es)) || (rc = alloc(&b->h0[i], S * es)))
            return bail(rc);
    }
    if ((rc = alloc(&b->rho, S * es)) || (rc = alloc(&b->ux, S * es)) || (rc = alloc(&b->uy, S * es)) ||
        (rc = alloc(&b->uz, S * es)) || (rc = alloc(&b->q, S * es)) || (rc = alloc(&b->phi, S * es)) ||
        (rc = alloc((void**)&b->phibar, sizeof(double) * b->nz)) || (rc = alloc((void**)&b->ex, sizeof(double) * S)) ||
        (rc = alloc((void**)&b->ey, sizeof(double) * S)) || (rc = alloc((void**)&b->ez, sizeof(double) * S)))
        return bail(rc);
    const size_t nmodes = (size_t)b->nxh * b->ny, nint = (size_t)(b->nz - 2);
    if ((rc = alloc(&b->spec, nmodes * nint * 2 * es)) || (rc = alloc(&b->twx, 2 * (size_t)b->nx * es)) ||
        (rc = alloc(&b->twy, 2 * (size_t)b->ny * es)) ||
        (rc = alloc((void**)&b->diag_scr, sizeof(double) * 8 * (size_t)diag_blocks())) ||
        (rc = alloc((void**)&b->diag_out, sizeof(double) * 8)))
        return bail(rc);

    if (b->prec == Prec::F32S) {
        // Storage shifts: the hydrostatic state of unipolar injection, q(z) of
        // the analytic profile and rho(z) = rho0 + (p(z) - mean p)/cs^2 with
        // dp/dz = q Ez = const (the uniform current density).
        const int nz = b->nz;
        const double C = b->der.charge0 * p->lz * p->lz / (p->voltage * p->eps);
        const double e0 = hydrostatic_e0(C);
        const double qsc = p->eps * p->voltage / (p->lz * p->lz);
        const double J = C * e0 * p->eps * p->voltage * p->voltage / (p->lz * p->lz * p->lz);
        std::vector<double> rs(nz), qs(nz);
        for (int z = 0; z < nz; ++z) {
            const double zeta = b->der.dz * z / p->lz;
            qs[z] = C * e0 / std::sqrt(e0 * e0 + 2.0 * C * e0 * zeta) * qsc;
            rs[z] = p->rho0 + J * (b->der.dz * z - 0.5 * p->lz) / b->der.cs_square;
        }
        qs[nz - 1] = qs[nz - 2];   // the top plane holds copies of plane nz-2
        // The shifts only have to be near the hydrostatic state, so take them
        // representable in float: the kernel's rho = rs + sum g and
        // q = qs + sum k then add an exact float shift (a rounded one is off
        // by up to half an ulp of ~rho0 on a whole plane, the same on every
        // node and every step -- a systematic density and force error that
        // drives the plane-average velocity), and the differences between
        // neighbouring planes, which re-shift pulled populations, are exact
        // in float as well (Sterbenz).
        for (int z = 0; z < nz; ++z) {
            rs[z] = (double)(float)rs[z];
            qs[z] = (double)(float)qs[z];
        }
        std::vector<float> tf(6 * (size_t)nz);
        for (int z = 0; z < nz; ++z) {
            const int zm = (z == 0) ? nz - 1 : z - 1, zp = (z == nz - 1) ? 0 : z + 1;
            tf[z] = (float)rs[z];
            tf[nz + z] = (float)(rs[zm] - rs[z]);
            tf[2 * nz + z] = (float)(rs[zp] - rs[z]);
            tf[3 * nz + z] = (float)qs[z];
            tf[4 * nz + z] = (float)(qs[zm] - qs[z]);
            tf[5 * nz + z] = (float)(qs[zp] - qs[z]);
        }
        if ((rc = alloc((void**)&b->shift_rs, nz * sizeof(double))) || (rc = alloc((void**)&b->shift_qs, nz * sizeof(double))) ||
            (rc = alloc((void**)&b->shift_f, tf.size() * sizeof(float))))
            return bail(rc);
        if (hipMemcpy(b->shift_rs, rs.data(), nz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->shift_qs, qs.data(), nz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->shift_f, tf.data(), tf.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return bail(set_error(EHD_EHIP, "shift table upload failed"));
        b->lp.rs = b->shift_rs;
        b->lp.qs = b->shift_qs;
        b->lp.shf = b->shift_f;
        b->lp.qs0 = qs[0];
    }
    if (b->prec == Prec::F64) {
        std::vector<double> tx, ty;
        make_twiddles(b->nx, tx);
        make_twiddles(b->ny, ty);
        if (hipMemcpy(b->twx, tx.data(), tx.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->twy, ty.data(), ty.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
            return bail(set_error(EHD_EHIP, "twiddle upload failed"));
    } else {
        std::vector<float> tx, ty;
        make_twiddles(b->nx, tx);
        make_twiddles(b->ny, ty);
        if (hipMemcpy(b->twx, tx.data(), tx.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(b->twy, ty.data(), ty.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
            return bail(set_error(EHD_EHIP, "twiddle upload failed"));
    }
    {
        const PoissonShape ps = poisson_shape(b->prec, b->nx, b->ny, b->nz);
        if (ps.max_x_lines == 0 || ps.max_y_lines == 0 || ps.zs_partition == 0)
            return bail(set_error(EHD_EINVAL, "grid %dx%dx%d exceeds the Poisson solver's LDS tiles", b->nx, b->ny,
                                  b->nz));
        PoissonBufs pb{b->nx, b->ny, b->nz, b->nxh, b->spec, b->twx, b->twy, b->phibar};
        const PoissonJob job = make_poisson_job(b->lp, pb, b->q, b->phi, p->voltage, p->eps);
        if ((rc = alloc((void**)&b->pjob, sizeof(PoissonJob)))) return bail(rc);
        b->pjob_h = job;
        if (hipMemcpy(b->pjob, &job, sizeof(job), hipMemcpyHostToDevice) != hipSuccess)
            return bail(set_error(EHD_EHIP, "Poisson job upload failed"));
    }
    if (b->prec == Prec::F32S) {
        // one-item tables of the FP32S kernel, one per buffer parity
        const size_t isz = lbm_f32s_batch_item_size();
        std::vector<unsigned char> item(isz);
        for (int c = 0; c < 2; ++c) {
            const int A = c, B = 1 - c;
            const PopBufs pb{b->f[A], b->f0[A], b->h[A], b->h0[A], b->f[B], b->f0[B], b->h[B], b->h0[B]};
            lbm_f32s_batch_item(item.data(), b->lp, pb, macro_bufs(b), 0);
            if ((rc = alloc(&b->lbm_items[c], isz))) return bail(rc);
            if (hipMemcpy(b->lbm_items[c], item.data(), isz, hipMemcpyHostToDevice) != hipSuccess)
                return bail(set_error(EHD_EHIP, "LBM item upload failed"));
        }
        b->lbm_chunk = lbm_xcd_chunk((unsigned)(p->nx * p->ny));
    }
    *out = b;
    return EHD_OK;
}

void ehd_destroy(ehd_box* b)
{
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    free_box(b);
    delete b;
}


int ehd_hydrostatic_current(const ehd_params* p, double* current)
{
    if (!p || !current) return set_error(EHD_EINVAL, "null argument");
    int rc = validate(p);
    if (rc) return rc;
    ehd_derived d;
    if ((rc = ehd_derive(p, &d))) return rc;
    *current = hydrostatic_current(*p, d);
    return EHD_OK;
}

int ehd_init(ehd_box* b, const double* fields9)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    const ehd_params& p = b->prm;
    const size_t S = b->S;
    if (p.init_mode == EHD_INIT_FIELDS && !fields9) return set_error(EHD_EINVAL, "EHD_INIT_FIELDS needs fields");
    // double macroscopic fields rho,q,phi,u,v,w on the device
    double* tmp = nullptr;
    int rc = alloc((void**)&tmp, 6 * S * sizeof(double));
    if (rc) return rc;
    double *rho = tmp, *q = tmp + S, *phi = tmp + 2 * S, *ux = tmp + 3 * S, *uy = tmp + 4 * S, *uz = tmp + 5 * S;
    struct Guard {
        double* p;
        ~Guard() { if (p) (void)hipFree(p); }
    } guard{tmp};
    double* dqz = nullptr;
    if (p.init_mode == EHD_INIT_REFERENCE) {
        HIPCHK(launch_init_fields(b->lp, 0, nullptr, nullptr, nullptr, 0, 0.0, 0, b->der.lx, b->der.ly, p.voltage, p.lz,
                                  rho, q, phi, ux, uy, uz, b->st));
        HIPCHK(hipMemsetAsync(b->ex, 0, S * sizeof(double), b->st));
        HIPCHK(hipMemsetAsync(b->ey, 0, S * sizeof(double), b->st));
        HIPCHK(hipMemsetAsync(b->ez, 0, S * sizeof(double), b->st));
    } else if (p.init_mode == EHD_INIT_HYDROSTATIC) {
        const int nz = b->nz;
        const std::vector<double> tab = hydrostatic_table(p, b->der);
        rc = alloc((void**)&dqz, tab.size() * sizeof(double));
        if (rc) return rc;
        hipError_t e1 = hipMemcpyAsync(dqz, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, b->st);
        hipError_t e2 = launch_init_fields(b->lp, 1, dqz, dqz + nz, dqz + 2 * nz, p.seed, p.perturb_amp,
                                           p.perturb_pattern, b->der.lx, b->der.ly, p.voltage, p.lz, rho, q, phi, ux,
                                           uy, uz, b->st);
        hipError_t e3 = launch_efield(Prec::F64, b->lp, phi, nullptr, b->ex, b->ey, b->ez, b->st);
        hipError_t e4 = hipStreamSynchronize(b->st);
        (void)hipFree(dqz);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess)
            return set_error(EHD_EHIP, "hydrostatic init failed");
    } else {
        HIPCHK(hipMemcpyAsync(rho, fields9 + 0 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(q, fields9 + 1 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(phi, fields9 + 2 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(ux, fields9 + 3 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(uy, fields9 + 4 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(uz, fields9 + 5 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(b->ex, fields9 + 6 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(b->ey, fields9 + 7 * S, S * 8, hipMemcpyHostToDevice, b->st));
        HIPCHK(hipMemcpyAsync(b->ez, fields9 + 8 * S, S * 8, hipMemcpyHostToDevice, b->st));
    }
    b->cur = 0;
    b->current_hs = hydrostatic_current(p, b->der);
    HIPCHK(launch_init_equilibrium(b->prec, b->lp, rho, q, ux, uy, uz, b->ex, b->ey, b->ez, b->f[0], b->f0[0], b->h[0],
                                   b->h0[0], b->st));
    // the frozen wall rest populations must be present in both buffers
    HIPCHK(hipMemcpyAsync(b->f0[1], b->f0[0], S * b->esz, hipMemcpyDeviceToDevice, b->st));
    HIPCHK(launch_from_double(b->prec, rho, b->rho, S, b->st));
    HIPCHK(launch_from_double(b->prec, q, b->q, S, b->st));
    HIPCHK(launch_from_double(b->prec, ux, b->ux, S, b->st));
    HIPCHK(launch_from_double(b->prec, uy, b->uy, S, b->st));
    HIPCHK(launch_from_double(b->prec, uz, b->uz, S, b->st));
    HIPCHK(launch_phi_split(b->prec, b->lp, phi, b->phi, b->phibar, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    b->direct = true;
    b->ebuf = true;
    b->initialised = true;
    b->step = 0;
    b->t = 0.0;
    return EHD_OK;
}

int ehd_step(ehd_box* b, int nsteps)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    if (nsteps < 0) return set_error(EHD_EINVAL, "nsteps < 0");
    hipEvent_t s0 = nullptr, s1 = nullptr;
    if (b->prof_every > 0 && nsteps > 0) {
        s0 = take_event(b);
        s1 = take_event(b);
        if (!s0 || !s1) return set_error(EHD_EHIP, "hipEventCreate failed");
        HIPCHK(hipEventRecord(s0, b->st));
    }
    for (int i = 0; i < nsteps; ++i) {
        int rc = run_lbm(b, i == nsteps - 1);
        if (rc) return rc;
        rc = run_poisson(b);
        if (rc) return rc;
        b->step += 1;
        b->t = b->t + b->der.dt;   // main.cu:208
    }
    if (b->prof_every > 0 && nsteps > 0) {
        HIPCHK(hipEventRecord(s1, b->st));
        b->step_pairs.emplace_back(s0, s1);
        int rc = resolve_events(b);
        if (rc) return rc;
    }
    return EHD_OK;
}

int ehd_init_consistent(ehd_box* b, const double* fields9, int iterations)
{
    if (!b || !fields9) return set_error(EHD_EINVAL, "null argument");
    if (b->prec != Prec::F32S) return set_error(EHD_EINVAL, "the consistent start is implemented for FP32S boxes");
    if (b->prm.init_mode != EHD_INIT_FIELDS) return set_error(EHD_EINVAL, "the consistent start needs EHD_INIT_FIELDS");
    if (iterations < 0) return set_error(EHD_EINVAL, "iterations < 0");
    int rc = ehd_init(b, fields9);
    if (rc || iterations == 0) return rc;
    const size_t S = b->S;
    ConsistentTargets ct;
    if ((rc = ct.make(b, fields9))) return rc;
    for (int i = 0; i < iterations; ++i) {
        if ((rc = run_lbm(b, i == iterations - 1, ct.items))) return rc;
        if ((rc = run_poisson(b))) return rc;
        b->step += 1;
        b->t = b->t + b->der.dt;
    }
    (void)S;
    HIPCHK(hipStreamSynchronize(b->st));
    return EHD_OK;
}

int ehd_lbm_step(ehd_box* b)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    return run_lbm(b, true);
}

int ehd_poisson(ehd_box* b)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    return run_poisson(b);
}

int ehd_read_fields(ehd_box* b, double* rho, double* q, double* phi, double* ux, double* uy, double* uz, double* ex,
                    double* ey, double* ez)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    const size_t S = b->S;
    double* tmp = nullptr;
    int rc = alloc((void**)&tmp, S * sizeof(double));
    if (rc) return rc;
    struct Guard {
        double* p;
        ~Guard() { if (p) (void)hipFree(p); }
    } guard{tmp};
    struct Item {
        double* host;
        void* dev;
    } items[] = {{rho, b->rho}, {q, b->q}, {ux, b->ux}, {uy, b->uy}, {uz, b->uz}};
    for (auto& it : items) {
        if (!it.host) continue;
        HIPCHK(launch_to_double(b->prec, it.dev, tmp, S, b->st));
        HIPCHK(hipMemcpyAsync(it.host, tmp, S * 8, hipMemcpyDeviceToHost, b->st));
        HIPCHK(hipStreamSynchronize(b->st));
    }
    if (phi) {
        HIPCHK(launch_phi_full(b->prec, b->lp, b->phi, b->phibar, tmp, b->st));
        HIPCHK(hipMemcpyAsync(phi, tmp, S * 8, hipMemcpyDeviceToHost, b->st));
        HIPCHK(hipStreamSynchronize(b->st));
    }
    if (ex || ey || ez) {
        if (!b->ebuf) HIPCHK(launch_efield(b->prec, b->lp, b->phi, b->phibar, b->ex, b->ey, b->ez, b->st));
        if (ex) HIPCHK(hipMemcpyAsync(ex, b->ex, S * 8, hipMemcpyDeviceToHost, b->st));
        if (ey) HIPCHK(hipMemcpyAsync(ey, b->ey, S * 8, hipMemcpyDeviceToHost, b->st));
        if (ez) HIPCHK(hipMemcpyAsync(ez, b->ez, S * 8, hipMemcpyDeviceToHost, b->st));
        HIPCHK(hipStreamSynchronize(b->st));
    }
    return EHD_OK;
}

int ehd_write_phi(ehd_box* b, const double* phi)
{
    if (!b || !phi) return set_error(EHD_EINVAL, "null argument");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    double* tmp = nullptr;
    int rc = alloc((void**)&tmp, b->S * sizeof(double));
    if (rc) return rc;
    struct Guard {
        double* p;
        ~Guard() { if (p) (void)hipFree(p); }
    } guard{tmp};
    HIPCHK(hipMemcpyAsync(tmp, phi, b->S * 8, hipMemcpyHostToDevice, b->st));
    HIPCHK(launch_phi_split(b->prec, b->lp, tmp, b->phi, b->phibar, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    b->ebuf = false;
    return EHD_OK;
}

int ehd_write_charge(ehd_box* b, const double* q)
{
    if (!b || !q) return set_error(EHD_EINVAL, "null argument");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    double* tmp = nullptr;
    int rc = alloc((void**)&tmp, b->S * sizeof(double));
    if (rc) return rc;
    struct Guard {
        double* p;
        ~Guard() { if (p) (void)hipFree(p); }
    } guard{tmp};
    HIPCHK(hipMemcpyAsync(tmp, q, b->S * 8, hipMemcpyHostToDevice, b->st));
    HIPCHK(launch_from_double(b->prec, tmp, b->q, b->S, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    return EHD_OK;
}

int ehd_read_populations(ehd_box* b, double* f0, double* f1, double* h0, double* h1)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    const size_t S = b->S;
    double* tmp = nullptr;
    int rc = alloc((void**)&tmp, 54 * S * sizeof(double));
    if (rc) return rc;
    struct Guard {
        double* p;
        ~Guard() { if (p) (void)hipFree(p); }
    } guard{tmp};
    double *d_f1 = tmp, *d_h1 = tmp + 26 * S, *d_f0 = tmp + 52 * S, *d_h0 = tmp + 53 * S;
    const int A = b->cur;
    HIPCHK(launch_read_pre(b->prec, b->lp, b->direct, b->f[A], b->f0[A], b->h[A], b->h0[A], d_f1, d_f0, d_h1, d_h0,
                           b->st));
    if (f0) HIPCHK(hipMemcpyAsync(f0, d_f0, S * 8, hipMemcpyDeviceToHost, b->st));
    if (f1) HIPCHK(hipMemcpyAsync(f1, d_f1, 26 * S * 8, hipMemcpyDeviceToHost, b->st));
    if (h0) HIPCHK(hipMemcpyAsync(h0, d_h0, S * 8, hipMemcpyDeviceToHost, b->st));
    if (h1) HIPCHK(hipMemcpyAsync(h1, d_h1, 26 * S * 8, hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    return EHD_OK;
}

int ehd_write_populations(ehd_box* b, const double* f0, const double* f1, const double* h0, const double* h1)
{
    if (!b || !f0 || !f1 || !h0 || !h1) return set_error(EHD_EINVAL, "null argument");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    const size_t S = b->S;
    double* tmp = nullptr;
    int rc = alloc((void**)&tmp, 54 * S * sizeof(double));
    if (rc) return rc;
    struct Guard {
        double* p;
        ~Guard() { if (p) (void)hipFree(p); }
    } guard{tmp};
    double *d_f1 = tmp, *d_h1 = tmp + 26 * S, *d_f0 = tmp + 52 * S, *d_h0 = tmp + 53 * S;
    HIPCHK(hipMemcpyAsync(d_f0, f0, S * 8, hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(d_f1, f1, 26 * S * 8, hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(d_h0, h0, S * 8, hipMemcpyHostToDevice, b->st));
    HIPCHK(hipMemcpyAsync(d_h1, h1, 26 * S * 8, hipMemcpyHostToDevice, b->st));
    const int A = b->cur, B = 1 - b->cur;
    HIPCHK(launch_write_pre(b->prec, b->lp, d_f1, d_f0, d_h1, d_h0, b->f[A], b->f0[A], b->h[A], b->h0[A], b->st));
    // frozen wall rest populations live in both buffers
    HIPCHK(hipMemcpyAsync(b->f0[B], b->f0[A], S * b->esz, hipMemcpyDeviceToDevice, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    b->direct = true;
    return EHD_OK;
}


int ehd_save_state(ehd_box* b, const char* path)
{
    if (!b || !path) return set_error(EHD_EINVAL, "null argument");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    HIPCHK(hipStreamSynchronize(b->st));
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(EHD_EINVAL, "cannot open %s for writing", path);
    CkptHeader h{};
    std::memcpy(h.magic, "EHDCKPT2", 8);
    h.abi = EHD_ABI_VERSION;
    h.precision = b->prm.precision;
    h.nx = b->nx; h.ny = b->ny; h.nz = b->nz;
    h.direct = b->direct ? 1 : 0;
    h.ebuf = b->ebuf ? 1 : 0;
    h.step = b->step;
    h.t = b->t;
    ckpt_physics(b->prm, h);
    int rc = (std::fwrite(&h, sizeof(h), 1, f) == 1) ? ckpt_io(b, f, true) : set_error(EHD_EINVAL, "checkpoint write failed");
    if (std::fclose(f) != 0 && rc == EHD_OK) rc = set_error(EHD_EINVAL, "checkpoint close failed");
    return rc;
}

int ehd_load_state(ehd_box* b, const char* path)
{
    if (!b || !path) return set_error(EHD_EINVAL, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(EHD_EINVAL, "cannot open %s", path);
    CkptHeader h{};
    if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, "EHDCKPT2", 8) != 0) {
        std::fclose(f);
        return set_error(EHD_EINVAL, "%s is not a checkpoint of this format", path);
    }
    if (h.abi != EHD_ABI_VERSION) {
        std::fclose(f);
        return set_error(EHD_EINVAL, "checkpoint of ABI version %d, library %d", h.abi, EHD_ABI_VERSION);
    }
    if (h.precision != b->prm.precision || h.nx != b->nx || h.ny != b->ny || h.nz != b->nz) {
        std::fclose(f);
        return set_error(EHD_EINVAL, "checkpoint grid/precision (%dx%dx%d, %d) differ from the box", h.nx, h.ny, h.nz,
                    h.precision);
    }
    if (const char* what = ckpt_mismatch(h, b->prm)) {
        std::fclose(f);
        return set_error(EHD_EINVAL, "checkpoint parameter %s differs from the box's", what);
    }
    HIPCHK(hipStreamSynchronize(b->st));
    const int rc = ckpt_io(b, f, false);
    std::fclose(f);
    if (rc) return rc;
    b->direct = h.direct != 0;
    b->ebuf = h.ebuf != 0;
    b->step = h.step;
    b->t = h.t;
    b->initialised = true;
    return EHD_OK;
}

int ehd_set_nusselt_reference(ehd_box* b, double current_hs)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    b->current_hs = current_hs;
    return EHD_OK;
}

int ehd_diagnostics(ehd_box* b, ehd_diag* out)
{
    if (!b || !out) return set_error(EHD_EINVAL, "null argument");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    HIPCHK(launch_diag(b->prec, b->lp, b->q, b->ux, b->uy, b->uz, b->rho, b->phi, b->phibar, b->diag_scr,
                       b->diag_out, b->st));
    double r[8] = {0};
    HIPCHK(hipMemcpyAsync(r, b->diag_out, 5 * sizeof(double), hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    const ehd_derived& d = b->der;
    out->current = r[0] * d.K * d.dy * d.dy;   // LBM.cu:1669
    out->total_charge = r[1] * d.dx * d.dy * d.dz;
    out->umax = r[2];
    out->mean_rho = r[3] / ((double)b->nxy * (double)(b->nz - 2));
    out->nusselt = (b->current_hs != 0.0) ? out->current / b->current_hs : 0.0;
    out->step = b->step;
    out->time = b->t;
    out->nonfinite = (long long)r[4];
    if (r[4] != 0.0)
        return set_error(EHD_ENONFINITE, "non-finite charge, density or velocity at %.0f nodes (step %lld)", r[4],
                         b->step);
    return EHD_OK;
}

int ehd_hydrostatic_reference(ehd_box* b, int check_every, double tol, long long max_steps, double* current)
{
    if (!b || !current) return set_error(EHD_EINVAL, "null argument");
    if (check_every < 1 || !(tol >= 0.0) || max_steps < 1) return set_error(EHD_EINVAL, "bad convergence criteria");
    ehd_params p = b->prm;
    p.nx = 2;
    p.ny = 2;
    p.precision = EHD_FP64;
    p.init_mode = EHD_INIT_HYDROSTATIC;
    p.perturb_amp = 0.0;
    ehd_box* c = nullptr;
    int rc = ehd_create(&p, &c);
    if (rc) return rc;
    struct Destroy {
        ehd_box* c;
        ~Destroy() { ehd_destroy(c); }
    } guard{c};
    if ((rc = ehd_init(c, nullptr))) return rc;
    ehd_diag dg;
    if ((rc = ehd_diagnostics(c, &dg))) return rc;
    double prev = dg.current;
    long long taken = 0;
    bool converged = false;
    while (taken < max_steps && !converged) {
        const int n = (int)std::min<long long>(check_every, max_steps - taken);
        if ((rc = ehd_step(c, n))) return rc;
        taken += n;
        if ((rc = ehd_diagnostics(c, &dg))) return rc;
        converged = std::fabs(dg.current - prev) <= tol * std::fabs(dg.current);
        prev = dg.current;
    }
    if (!converged)
        return set_error(EHD_ESTATE, "hydrostatic reference not converged in %lld steps", max_steps);
    // a laterally uniform state: the current scales with the number of columns
    *current = dg.current * ((double)b->nx * (double)b->ny / 4.0);
    b->current_hs = *current;
    return EHD_OK;
}

int ehd_run_to_steady(ehd_box* b, int check_every, double tol, long long max_steps, ehd_steady* out)
{
    if (!b || !out) return set_error(EHD_EINVAL, "null argument");
    if (check_every < 1 || !(tol >= 0.0) || max_steps < 0) return set_error(EHD_EINVAL, "bad steady-state criteria");
    if (!b->initialised) return set_error(EHD_ESTATE, "box not initialised");
    ehd_diag dg;
    out->steps = 0;
    out->converged = 0;
    out->failed = 0;
    out->delta = 0.0;
    int rc = ehd_diagnostics(b, &dg);
    out->nusselt = dg.nusselt;
    out->current = dg.current;
    if (rc == EHD_ENONFINITE) out->failed = 1;
    if (rc) return rc;
    double prev = dg.nusselt;
    long long taken = 0;
    while (taken < max_steps) {
        const int n = (int)std::min<long long>(check_every, max_steps - taken);
        if ((rc = ehd_step(b, n))) return rc;
        taken += n;
        if ((rc = ehd_diagnostics(b, &dg))) {
            if (rc == EHD_ENONFINITE) {
                out->failed = 1;
                out->steps = taken;
                out->nusselt = dg.nusselt;
                out->current = dg.current;
            }
            return rc;
        }
        out->delta = std::fabs(dg.nusselt - prev);
        prev = dg.nusselt;
        if (n == check_every && out->delta < tol) {
            out->converged = 1;
            break;
        }
    }
    out->steps = taken;
    out->nusselt = dg.nusselt;
    out->current = dg.current;
    return EHD_OK;
}

int ehd_set_debug_sync(ehd_box* b, int on)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    b->debug_sync = on != 0;
    return EHD_OK;
}

int ehd_profile_enable(ehd_box* b, int on)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    if (on < 0) return set_error(EHD_EINVAL, "profile sampling interval < 0 (got %d)", on);
    b->prof_every = on;
    return EHD_OK;
}

int ehd_profile_read(ehd_box* b, ehd_profile* out)
{
    if (!b || !out) return set_error(EHD_EINVAL, "null argument");
    int rc = resolve_events(b);
    if (rc) return rc;
    out->lbm_ms = b->lbm_ms;
    out->lbm_launches = b->lbm_n;
    out->poisson_ms = b->poi_ms;
    out->poisson_calls = b->poi_n;
    out->step_ms = b->step_ms;
    out->steps = b->step_n;
    return EHD_OK;
}

int ehd_profile_reset(ehd_box* b)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    int rc = resolve_events(b);
    if (rc) return rc;
    b->lbm_ms = b->poi_ms = b->step_ms = 0.0;
    b->lbm_n = b->poi_n = b->step_n = 0;
    return EHD_OK;
}

int ehd_sync(ehd_box* b)
{
    if (!b) return set_error(EHD_EINVAL, "null box");
    HIPCHK(hipStreamSynchronize(b->st));
    return EHD_OK;
}

}  // extern "C"
