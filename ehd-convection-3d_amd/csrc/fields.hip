// fields.hip -- macroscopic-field kernels: initial states, E = -grad(phi),
// precision conversions, and the FP64 global reductions.
// Compiled with -ffp-contract=off (the FP64 E field and initial state match
// the reference's expressions bit for bit).
#include <hip/hip_runtime.h>

#include "ehd_internal.h"

namespace ehd {
namespace {

inline dim3 grid_for(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// U(-1,1) from (seed, x, y, z, component), identical on host and device.
__device__ __forceinline__ double hash_uniform(uint64_t seed, uint32_t x, uint32_t y, uint32_t z, uint32_t comp)
{
    const uint64_t key = ((uint64_t)comp << 60) ^ ((uint64_t)z << 40) ^ ((uint64_t)y << 20) ^ (uint64_t)x;
    const uint64_t h = mix64(seed ^ mix64(key));
    const double u = (double)(h >> 11) * 0x1.0p-53;
    return 2.0 * u - 1.0;
}

// Velocity perturbation patterns (EHD_PERTURB_*), after the roll, square,
// crossed-roll and hexagon fields of LBM.cu:389-411 with amplitude amp and
// coordinates zeta = z/Lz, xi = x/Lx, eta = y/Ly.
__device__ void pattern_velocity(int pattern, double amp, double zeta, double xi, double eta, double xx, double yy,
                                 double lx, double ly, double lz, double* u)
{
    const double tp = 2.0 * 3.14159265358979323846;
    const double cz = cos(tp * zeta) - 1.0, sz = sin(tp * zeta);
    u[0] = u[1] = u[2] = 0.0;
    if (pattern == EHD_PERTURB_ROLLS) {
        u[2] = amp * cz * cos(tp * xi);
        u[0] = amp * (lx / lz) * sz * sin(tp * xi);
    } else if (pattern == EHD_PERTURB_SQUARES) {
        u[2] = amp * cz * cos(tp * xi) * cos(tp * eta);
        u[0] = 0.5 * amp * (lx / lz) * sz * sin(tp * xi) * cos(tp * eta);
        u[1] = 0.5 * amp * (ly / lz) * sz * sin(tp * eta) * cos(tp * xi);
    } else if (pattern == EHD_PERTURB_CROSS) {
        u[2] = amp * cz * (cos(tp * xi) + cos(tp * eta));
        u[0] = amp * (lx / lz) * sz * sin(tp * xi);
        u[1] = amp * (ly / lz) * sz * sin(tp * eta);
    } else if (pattern == EHD_PERTURB_HEXAGONS) {
        const double L = lx / sqrt(3.0);
        const double a = 4.0 * 3.14159265358979323846 / 3.0 / L;
        const double cx = cos(tp / (sqrt(3.0) * L) * xx), sx = sin(tp / (sqrt(3.0) * L) * xx);
        const double cy = cos(tp / (3.0 * L) * yy), sy = sin(tp / (3.0 * L) * yy);
        u[2] = -amp * cz / 3.0 * (2.0 * cx * cy + cos(2.0 * tp / (3.0 * L) * yy));
        u[0] = -amp * tp * sz * (4.0 * 3.14159265358979323846) / (3.0 * sqrt(3.0) * L * a * a) * sx * cy;
        u[1] = -amp * tp * sz * (4.0 * 3.14159265358979323846) / (9.0 * L * a * a) * (cx + 2.0 * cy) * sy;
    }
}

// mode 0: gpu_initialization (LBM.cu:80-96); mode 1: hydrostatic profile
// (tables qz, phiz, rhoz per plane) plus the perturbation of u.
__global__ void __launch_bounds__(256) init_fields_kernel(LbmParams P, int mode, const double* qz,
                                                          const double* phiz, const double* rhoz, uint64_t seed,
                                                          double amp, int pattern, double lx, double ly,
                                                          double voltage, double lz, double* rho, double* q,
                                                          double* phi, double* ux, double* uy, double* uz)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.S) return;
    const unsigned x = s % P.nx, t = s / P.nx, y = t % P.ny, z = t / P.ny;
    if (mode == 0) {
        rho[s] = P.rho0;
        q[s] = 0.0;
        phi[s] = voltage * (lz - P.dz * z) / lz;
        ux[s] = 0.0;
        uy[s] = 0.0;
        uz[s] = 0.0;
    } else {
        rho[s] = rhoz[z];
        q[s] = qz[z];
        phi[s] = phiz[z];
        const bool wall = (z == 0) || (z == (unsigned)P.nz - 1);
        if (pattern == EHD_PERTURB_NOISE) {
            ux[s] = wall ? 0.0 : amp * hash_uniform(seed, x, y, z, 0);
            uy[s] = wall ? 0.0 : amp * hash_uniform(seed, x, y, z, 1);
            uz[s] = wall ? 0.0 : amp * hash_uniform(seed, x, y, z, 2);
        } else {
            double u[3];
            pattern_velocity(pattern, amp, P.dz * z / lz, P.dx * x / lx, P.dy * y / ly, P.dx * x, P.dy * y, lx, ly,
                             lz, u);
            ux[s] = wall ? 0.0 : u[0];
            uy[s] = wall ? 0.0 : u[1];
            uz[s] = wall ? 0.0 : u[2];
        }
    }
}

// gpu_efield + gpu_bc (poisson.cu:44-73).  phibar != nullptr: phi holds the
// float fluctuation about the per-plane mean phibar (FP32S).
template <typename T>
__global__ void __launch_bounds__(256) efield_kernel(LbmParams P, const T* phi, const double* phibar, double* ex,
                                                     double* ey, double* ez)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.S) return;
    const unsigned nx = P.nx, ny = P.ny, nz = P.nz, nxy = P.nxy;
    const unsigned x = s % nx, t = s / nx, y = t % ny, z = t / ny;
    const unsigned xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
    const unsigned ym = (y == 0) ? ny - 1 : y - 1, yp = (y == ny - 1) ? 0 : y + 1;
    const unsigned yc = y * nx, zc = z * nxy;
    const unsigned ze = (z == 0) ? 1u : ((z == nz - 1) ? nz - 2 : z);
    if (phibar == nullptr) {
        ex[s] = 0.5 * ((double)phi[xm + yc + zc] - (double)phi[xp + yc + zc]) / P.dx;
        ey[s] = 0.5 * ((double)phi[x + ym * nx + zc] - (double)phi[x + yp * nx + zc]) / P.dy;
        ez[s] = 0.5 * ((double)phi[x + yc + (ze - 1) * nxy] - (double)phi[x + yc + (ze + 1) * nxy]) / P.dz;
    } else {
        ex[s] = 0.5 * ((double)phi[xm + yc + zc] - (double)phi[xp + yc + zc]) / P.dx;
        ey[s] = 0.5 * ((double)phi[x + ym * nx + zc] - (double)phi[x + yp * nx + zc]) / P.dy;
        const double dfl = (double)phi[x + yc + (ze - 1) * nxy] - (double)phi[x + yc + (ze + 1) * nxy];
        const double dmn = phibar[ze - 1] - phibar[ze + 1];
        ez[s] = 0.5 * (dfl + dmn) / P.dz;
    }
}

__global__ void __launch_bounds__(256) f2d_kernel(const float* src, double* dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}
__global__ void __launch_bounds__(256) d2f_kernel(const double* src, float* dst, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (float)src[i];
}

__global__ void __launch_bounds__(256) phi_full_kernel(LbmParams P, const float* phi, const double* phibar,
                                                       double* out)
{
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.S) return;
    const unsigned z = s / P.nxy;
    out[s] = (double)phi[s] + phibar[z];
}

// per-plane mean (one block per plane, fixed reduction order) and fluctuation
__global__ void __launch_bounds__(256) phi_split_kernel(LbmParams P, const double* phi, float* phio,
                                                        double* phibar)
{
    __shared__ double red[256];
    const unsigned z = blockIdx.x;
    const unsigned nxy = P.nxy;
    const double* pz = phi + (size_t)z * nxy;
    double acc = 0.0;
    for (unsigned i = threadIdx.x; i < nxy; i += blockDim.x) acc += pz[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (unsigned w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double mean = red[0] / (double)nxy;
    for (unsigned i = threadIdx.x; i < nxy; i += blockDim.x) phio[(size_t)z * nxy + i] = (float)(pz[i] - mean);
    if (threadIdx.x == 0) phibar[z] = mean;
}

// FP64 reductions over the box: one partial per block, then one block.
constexpr int DIAG_N = 5;

template <typename T>
__global__ void __launch_bounds__(256) diag_partial_kernel(LbmParams P, const T* q, const T* ux, const T* uy,
                                                           const T* uz, const T* rho, const T* phi,
                                                           const double* phibar, double* part)
{
    __shared__ double red[DIAG_N][256];
    const unsigned nx = P.nx, ny = P.ny, nz = P.nz, nxy = P.nxy;
    double cur = 0.0, qs = 0.0, um = 0.0, rs = 0.0, bad = 0.0;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t s = tid; s < P.S; s += stride) {
        const unsigned z = (unsigned)(s / nxy);
        const double qv = (double)q[s];
        const double uv = (double)uz[s];
        const double rv = (double)rho[s];
        // watchdog: a node counts once if any of its fields is NaN or infinite
        // (|v| <= DBL_MAX is false for both)
        const double big = 1.7976931348623157e308;
        if (!(fabs(qv) <= big) || !(fabs(uv) <= big) || !(fabs(rv) <= big) || !(fabs((double)ux[s]) <= big) ||
            !(fabs((double)uy[s]) <= big))
            bad += 1.0;
        if (z >= 1 && z <= nz - 2) {
            qs += qv;
            rs += rv;
        }
        double uvv = uv;
        if (z == nz - 1) uvv = 2.0 * (double)uz[s - nxy] - (double)uz[s - 2 * nxy];
        um = fmax(um, uvv);
        if (z == nz - 1) {
            // current(): q extrapolated to the top wall, Ez(nz-1) = Ez(nz-2)
            const double qt = 2.0 * (double)q[s - nxy] - (double)q[s - 2 * nxy];
            const size_t i = s - (size_t)z * nxy;
            double dphi = (double)phi[i + (size_t)(nz - 3) * nxy] - (double)phi[i + (size_t)(nz - 1) * nxy];
            if (phibar) dphi += phibar[nz - 3] - phibar[nz - 1];
            const double ezt = 0.5 * dphi / P.dz;
            cur += qt * ezt;
        }
    }
    (void)nx; (void)ny;
    red[0][threadIdx.x] = cur;
    red[1][threadIdx.x] = qs;
    red[2][threadIdx.x] = um;
    red[3][threadIdx.x] = rs;
    red[4][threadIdx.x] = bad;
    __syncthreads();
    for (unsigned w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
            red[3][threadIdx.x] += red[3][threadIdx.x + w];
            red[4][threadIdx.x] += red[4][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < DIAG_N) part[(size_t)blockIdx.x * DIAG_N + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) diag_final_kernel(const double* part, int nblk, double* out)
{
    __shared__ double red[DIAG_N][256];
    double a[DIAG_N] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        a[0] += part[b * DIAG_N + 0];
        a[1] += part[b * DIAG_N + 1];
        a[2] = fmax(a[2], part[b * DIAG_N + 2]);
        a[3] += part[b * DIAG_N + 3];
        a[4] += part[b * DIAG_N + 4];
    }
    for (int k = 0; k < DIAG_N; ++k) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (unsigned w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
            red[3][threadIdx.x] += red[3][threadIdx.x + w];
            red[4][threadIdx.x] += red[4][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < DIAG_N) out[threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

hipError_t launch_init_fields(const LbmParams& P, int mode, const double* qz, const double* phiz, const double* rhoz,
                              uint64_t seed, double amp, int pattern, double lx, double ly, double voltage, double lz,
                              double* rho, double* q, double* phi, double* ux, double* uy, double* uz, hipStream_t st)
{
    init_fields_kernel<<<grid_for(P.S), 256, 0, st>>>(P, mode, qz, phiz, rhoz, seed, amp, pattern, lx, ly, voltage, lz,
                                                      rho, q, phi, ux, uy, uz);
    return hipGetLastError();
}

hipError_t launch_efield(Prec prec, const LbmParams& P, const void* phi, const double* phibar, double* ex,
                         double* ey, double* ez, hipStream_t st)
{
    if (prec == Prec::F64)
        efield_kernel<double><<<grid_for(P.S), 256, 0, st>>>(P, (const double*)phi, nullptr, ex, ey, ez);
    else
        efield_kernel<float><<<grid_for(P.S), 256, 0, st>>>(P, (const float*)phi, phibar, ex, ey, ez);
    return hipGetLastError();
}

hipError_t launch_to_double(Prec prec, const void* src, double* dst, size_t n, hipStream_t st)
{
    if (prec == Prec::F64) return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st);
    f2d_kernel<<<grid_for(n), 256, 0, st>>>((const float*)src, dst, n);
    return hipGetLastError();
}

hipError_t launch_from_double(Prec prec, const double* src, void* dst, size_t n, hipStream_t st)
{
    if (prec == Prec::F64) return hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st);
    d2f_kernel<<<grid_for(n), 256, 0, st>>>(src, (float*)dst, n);
    return hipGetLastError();
}

hipError_t launch_phi_full(Prec prec, const LbmParams& P, const void* phi, const double* phibar, double* out,
                           hipStream_t st)
{
    if (prec == Prec::F64) return hipMemcpyAsync(out, phi, (size_t)P.S * sizeof(double), hipMemcpyDeviceToDevice, st);
    phi_full_kernel<<<grid_for(P.S), 256, 0, st>>>(P, (const float*)phi, phibar, out);
    return hipGetLastError();
}

hipError_t launch_phi_split(Prec prec, const LbmParams& P, const double* phi, void* phio, double* phibar,
                            hipStream_t st)
{
    if (prec == Prec::F64) return hipMemcpyAsync(phio, phi, (size_t)P.S * sizeof(double), hipMemcpyDeviceToDevice, st);
    phi_split_kernel<<<P.nz, 256, 0, st>>>(P, phi, (float*)phio, phibar);
    return hipGetLastError();
}

int diag_blocks() { return 512; }

hipError_t launch_diag(Prec prec, const LbmParams& P, const void* q, const void* ux, const void* uy, const void* uz,
                       const void* rho, const void* phi, const double* phibar, double* scratch, double* out,
                       hipStream_t st)
{
    const int nb = diag_blocks();
    if (prec == Prec::F64)
        diag_partial_kernel<double><<<nb, 256, 0, st>>>(P, (const double*)q, (const double*)ux, (const double*)uy,
                                                        (const double*)uz, (const double*)rho, (const double*)phi,
                                                        nullptr, scratch);
    else
        diag_partial_kernel<float><<<nb, 256, 0, st>>>(P, (const float*)q, (const float*)ux, (const float*)uy,
                                                       (const float*)uz, (const float*)rho, (const float*)phi, phibar,
                                                       scratch);
    diag_final_kernel<<<1, 256, 0, st>>>(scratch, nb, out);
    return hipGetLastError();
}

}  // namespace ehd
