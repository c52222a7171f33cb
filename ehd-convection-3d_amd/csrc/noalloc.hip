// noalloc.hip -- ehd_poisson_noalloc: the FP64 Poisson solve of ehd_poisson on
// caller-owned device buffers, with every workspace (job table, twiddles,
// spectrum) in one file-scope device array instead of memory the library
// allocates, for callers that cannot call hipMalloc (code under HIP graph
// capture, tight serving loops, harnesses that forbid allocation).
//
// The array lives on every GPU the library runs on until the process ends.
// Host code never uses its name as an address: on the first call on a
// device, with that device current, hipGetSymbolAddress gives that device's
// copy, and that pointer (kept per device) is what the copies and kernels
// get.  Every call on one GPU shares its array, so calls that overlap on one
// GPU (from two threads, or on two streams without ordering) race; the call
// returns after its stream has finished the solve.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "box.h"

namespace {

constexpr size_t kScratchBytes = EHD_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// this device's copy of scratch (looked up once per device)
int scratch_ptr(void** out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return ehd::set_error(EHD_EINVAL, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return ehd::set_error(EHD_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                                  hipGetErrorString(e));
        g_ptr[dev] = p;
    }
    *out = g_ptr[dev];
    return EHD_OK;
}

size_t up256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

extern "C" int ehd_poisson_noalloc(const ehd_params* p, const double* q, double* phi, void* stream)
{
    if (!p || !q || !phi) return ehd::set_error(EHD_EINVAL, "null argument");
    ehd_derived d;
    int rc = ehd_derive(p, &d);
    if (rc) return rc;
    if (p->precision != EHD_FP64) return ehd::set_error(EHD_EINVAL, "ehd_poisson_noalloc solves in FP64 only");
    const int nx = p->nx, ny = p->ny, nz = p->nz, nxh = nx / 2 + 1;
    const ehd::PoissonShape ps = ehd::poisson_shape(ehd::Prec::F64, nx, ny, nz);
    if (ps.max_x_lines == 0 || ps.max_y_lines == 0 || ps.zs_partition == 0)
        return ehd::set_error(EHD_EINVAL, "grid %dx%dx%d exceeds the Poisson solver's LDS tiles", nx, ny, nz);
    // scratch layout: job | x twiddles | y twiddles | spectrum [nz-2][nx/2+1][ny] complex double
    const size_t o_job = 0, o_twx = up256(sizeof(ehd::PoissonJob)), o_twy = o_twx + up256(16 * (size_t)nx),
                 o_spec = o_twy + up256(16 * (size_t)ny);
    const size_t need = o_spec + 16 * (size_t)nxh * ny * (size_t)(nz - 2);
    if (need > kScratchBytes)
        return ehd::set_error(EHD_EINVAL,
                              "grid %dx%dx%d needs %zu bytes of workspace; ehd_poisson_noalloc holds %zu "
                              "(EHD_NOALLOC_SCRATCH_BYTES)",
                              nx, ny, nz, need, kScratchBytes);
    void* base = nullptr;
    if ((rc = scratch_ptr(&base))) return rc;
    unsigned char* s = (unsigned char*)base;
    hipStream_t st = (hipStream_t)stream;

    // twiddles and the job, as ehd_create builds them for a box
    std::vector<double> tw(2 * (size_t)(nx > ny ? nx : ny));
    auto twiddles = [&](int n) {
        for (int j = 0; j < n; ++j) {
            const double a = 2.0 * 3.14159265358979323846 * (double)j / (double)n;
            tw[2 * j] = std::cos(a);
            tw[2 * j + 1] = -std::sin(a);
        }
    };
    ehd::LbmParams L{};
    L.dx = d.dx;
    L.dy = d.dy;
    L.dz = d.dz;
    const ehd::PoissonBufs pb{nx, ny, nz, nxh, s + o_spec, s + o_twx, s + o_twy, nullptr};
    const ehd::PoissonJob job = ehd::make_poisson_job(L, pb, q, phi, p->voltage, p->eps);
    HIPCHK(hipMemcpyAsync(s + o_job, &job, sizeof(job), hipMemcpyHostToDevice, st));
    twiddles(nx);
    HIPCHK(hipMemcpyAsync(s + o_twx, tw.data(), 16 * (size_t)nx, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));   // the host tables are reused below
    twiddles(ny);
    HIPCHK(hipMemcpyAsync(s + o_twy, tw.data(), 16 * (size_t)ny, hipMemcpyHostToDevice, st));

    ehd::PoissonGroup g{};
    g.nx = nx;
    g.ny = ny;
    g.max_n = nz - 2;
    g.njobs = 1;
    g.split_mean = 0;
    g.zs_partition = ps.zs_partition;
    g.twx = s + o_twx;
    g.twy = s + o_twy;
    g.jobs = (const ehd::PoissonJob*)(s + o_job);
    const hipError_t e = ehd::launch_poisson_group(ehd::Prec::F64, g, st);
    if (e != hipSuccess) return ehd::set_error(EHD_EHIP, "Poisson launch failed: %s", hipGetErrorString(e));
    HIPCHK(hipStreamSynchronize(st));
    return EHD_OK;
}
