// poisson.hip -- spectral Poisson solve for the potential.
//
// The reference (poisson.cu:78-167, main.cu:191-203) odd-extends the charge
// to nz' = 2(nz-1) planes, runs a 3-D complex DFT, divides by
//   mu = 4/dz^2 sin^2(kz dz/2) + kx^2 + ky^2
// and transforms back.  Along z that is the eigen-decomposition (a DST) of the
// second-difference operator with Dirichlet walls, so the same potential
// follows from
//   (1) a real-to-complex DFT along x and a complex DFT along y of
//       s = -q/eps on each interior plane z = 1..nz-2,
//   (2) per (kx,ky) mode, the tridiagonal solve
//       phi[z-1] - (2 + k^2 dz^2) phi[z] + phi[z+1] = dz^2 s[z],
//       phi = 0 on both walls, in FP64,
//   (3) the inverse DFTs,
// with the wall potential carried by phi_L(z) = V (1 - z/(nz-1)), which the
// second-difference operator annihilates (the lifting terms V/dz^2 of
// poisson.cu:99,119 only ever reach mode (0,0)).  The DFTs are Stockham
// mixed-radix FFTs in LDS (radix 4, 2, 3, 5 butterflies; any other prime
// factor is done as a direct DFT stage).
//
// Spectral layout: [z-1][ky][kx], kx in [0, nx/2], pitch nxh = nx/2+1.
// FP32S path: the spectra are float; mode (0,0) is solved in FP64 and kept
// as the per-plane mean phibar[z] (FP64), and phi holds phi - phibar in
// float, so that E = -grad(phi) loses no digits to the large wall potential.
#include <hip/hip_runtime.h>

#include "ehd_internal.h"
#include "fft_plan.h"

namespace ehd {
namespace {

template <typename T>
struct cplx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
// i*s*a for s = +-1
template <typename T>
__device__ __forceinline__ cplx<T> cis(cplx<T> a, T s) { return {-s * a.y, s * a.x}; }

// root W_n^(sign*m) from the forward table tw[m] = exp(-2 pi i m/n)
template <typename T>
__device__ __forceinline__ cplx<T> root(const cplx<T>* tw, int m, int sign)
{
    cplx<T> w = tw[m];
    if (sign > 0) w.y = -w.y;
    return w;
}

// Stockham FFT of nlines lines (line l at buf[l*ldl]) held in LDS, ping-pong
// between A and B.  sign -1 forward, +1 inverse, unnormalised.  Returns the
// buffer that holds the result.  Every thread of the block must call it.
template <typename T>
__device__ cplx<T>* lds_fft(cplx<T>* A, cplx<T>* B, int n, int nlines, int ldl, const FftPlan& pl,
                            const cplx<T>* __restrict__ tw, int sign)
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    const T sg = (T)sign;
    cplx<T>* in = A;
    cplx<T>* out = B;
    int Ns = 1;
    for (int st = 0; st < pl.nst; ++st) {
        const int R = pl.rad[st];
        const int nb = n / R;
        const int tstep = n / (Ns * R);
        if (R == 2 || R == 4) {
            for (int w = tid; w < nlines * nb; w += nthr) {
                const int l = w / nb, j = w - l * nb, k = j % Ns;
                const cplx<T>* src = in + l * ldl + j;
                cplx<T>* dst = out + l * ldl + (j / Ns) * Ns * R + k;
                if (R == 2) {
                    cplx<T> a = src[0];
                    cplx<T> b = cmul(src[nb], root(tw, k * tstep, sign));
                    dst[0] = cadd(a, b);
                    dst[Ns] = csub(a, b);
                } else {
                    cplx<T> a = src[0];
                    cplx<T> b = cmul(src[nb], root(tw, k * tstep, sign));
                    cplx<T> c = cmul(src[2 * nb], root(tw, 2 * k * tstep, sign));
                    cplx<T> d = cmul(src[3 * nb], root(tw, 3 * k * tstep, sign));
                    cplx<T> t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = csub(b, d);
                    // W4 = exp(sign*i*pi/2) = i*sign
                    cplx<T> it3 = cis(t3, sg);
                    dst[0] = cadd(t0, t2);
                    dst[Ns] = cadd(t1, it3);
                    dst[2 * Ns] = csub(t0, t2);
                    dst[3 * Ns] = csub(t1, it3);
                }
            }
        } else if (R == 3) {
            const T s3 = (T)0.86602540378443864676;  // sin(2 pi/3)
            for (int w = tid; w < nlines * nb; w += nthr) {
                const int l = w / nb, j = w - l * nb, k = j % Ns;
                const cplx<T>* src = in + l * ldl + j;
                cplx<T>* dst = out + l * ldl + (j / Ns) * Ns * R + k;
                cplx<T> a = src[0];
                cplx<T> b = cmul(src[nb], root(tw, k * tstep, sign));
                cplx<T> c = cmul(src[2 * nb], root(tw, 2 * k * tstep, sign));
                cplx<T> t1 = cadd(b, c);
                cplx<T> t2 = {a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y};
                cplx<T> t3 = csub(b, c);
                cplx<T> r = cis(cplx<T>{s3 * t3.x, s3 * t3.y}, sg);
                dst[0] = cadd(a, t1);
                dst[Ns] = cadd(t2, r);
                dst[2 * Ns] = csub(t2, r);
            }
        } else if (R == 5) {
            const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
            const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
            for (int w = tid; w < nlines * nb; w += nthr) {
                const int l = w / nb, j = w - l * nb, k = j % Ns;
                const cplx<T>* src = in + l * ldl + j;
                cplx<T>* dst = out + l * ldl + (j / Ns) * Ns * R + k;
                cplx<T> a = src[0];
                cplx<T> b = cmul(src[nb], root(tw, k * tstep, sign));
                cplx<T> c = cmul(src[2 * nb], root(tw, 2 * k * tstep, sign));
                cplx<T> d = cmul(src[3 * nb], root(tw, 3 * k * tstep, sign));
                cplx<T> e = cmul(src[4 * nb], root(tw, 4 * k * tstep, sign));
                cplx<T> t1 = cadd(b, e), t2 = cadd(c, d), t3 = csub(b, e), t4 = csub(c, d);
                cplx<T> m1 = {a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
                cplx<T> m2 = {a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
                cplx<T> n1 = cis(cplx<T>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, sg);
                cplx<T> n2 = cis(cplx<T>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, sg);
                dst[0] = cadd(a, cadd(t1, t2));
                dst[Ns] = cadd(m1, n1);
                dst[2 * Ns] = cadd(m2, n2);
                dst[3 * Ns] = csub(m2, n2);
                dst[4 * Ns] = csub(m1, n1);
            }
        } else {
            // direct DFT stage for any other prime: one thread per output
            for (int w = tid; w < nlines * n; w += nthr) {
                const int l = w / n, o = w - l * n;
                const int k = o % Ns, q = (o / Ns) % R, jb = o / (Ns * R);
                const int j = jb * Ns + k;
                const cplx<T>* src = in + l * ldl + j;
                cplx<T> acc = {(T)0, (T)0};
                for (int r = 0; r < R; ++r) {
                    cplx<T> v = cmul(src[r * nb], root(tw, (r * k * tstep) % n, sign));
                    acc = cadd(acc, cmul(v, root(tw, ((r * q) % R) * nb, sign)));
                }
                out[l * ldl + o] = acc;
            }
        }
        __syncthreads();
        cplx<T>* tmp = in;
        in = out;
        out = tmp;
        Ns *= R;
    }
    return in;
}

// (1a) R2C along x of s = -q/eps on interior rows, two real rows per complex
// line.  Rows are numbered r = (z-1)*ny + y; line l holds rows 2l and 2l+1.
template <typename T>
__global__ void __launch_bounds__(256) fftx_fwd_kernel(const T* __restrict__ q, cplx<T>* __restrict__ spec,
                                                       const cplx<T>* __restrict__ tw, FftPlan pl, int nx,
                                                       int ny, int nxh, int npairs, int lpb, double eps)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* A = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* B = A + lpb * nx;
    const int l0 = blockIdx.x * lpb;
    const int nl = min(lpb, npairs - l0);
    const int nxy = nx * ny;
    for (int w = threadIdx.x; w < nl * nx; w += blockDim.x) {
        const int l = w / nx, i = w - l * nx;
        const int r0 = 2 * (l0 + l), r1 = r0 + 1;
        const size_t o0 = (size_t)(r0 / ny + 1) * nxy + (size_t)(r0 % ny) * nx + i;
        const size_t o1 = (size_t)(r1 / ny + 1) * nxy + (size_t)(r1 % ny) * nx + i;
        const double a = -(double)q[o0] / eps;
        const double b = -(double)q[o1] / eps;
        A[l * nx + i] = {(T)a, (T)b};
    }
    __syncthreads();
    cplx<T>* res = lds_fft<T>(A, B, nx, nl, nx, pl, tw, -1);
    for (int w = threadIdx.x; w < nl * nxh; w += blockDim.x) {
        const int l = w / nxh, k = w - l * nxh;
        const cplx<T> zk = res[l * nx + k];
        const cplx<T> zn = res[l * nx + ((nx - k) % nx)];
        // A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i)
        const cplx<T> ak = {(T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y)};
        const cplx<T> bk = {(T)0.5 * (zk.y + zn.y), (T)-0.5 * (zk.x - zn.x)};
        const int r0 = 2 * (l0 + l);
        spec[(size_t)r0 * nxh + k] = ak;
        spec[(size_t)(r0 + 1) * nxh + k] = bk;
    }
}

// (1b)/(3a) complex DFT along y, in place, for strips of w columns.
template <typename T>
__global__ void __launch_bounds__(256) ffty_kernel(cplx<T>* __restrict__ spec, const cplx<T>* __restrict__ tw,
                                                   FftPlan pl, int ny, int nxh, int wstrip, int nstrips,
                                                   int sign)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* A = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* B = A + wstrip * ny;
    const int plane = blockIdx.x / nstrips;
    const int c0 = (blockIdx.x - plane * nstrips) * wstrip;
    const int nc = min(wstrip, nxh - c0);
    cplx<T>* base = spec + (size_t)plane * ny * nxh + c0;
    for (int w = threadIdx.x; w < nc * ny; w += blockDim.x) {
        const int yy = w / nc, c = w - yy * nc;
        A[c * ny + yy] = base[(size_t)yy * nxh + c];
    }
    __syncthreads();
    cplx<T>* res = lds_fft<T>(A, B, ny, nc, ny, pl, tw, sign);
    for (int w = threadIdx.x; w < nc * ny; w += blockDim.x) {
        const int yy = w / nc, c = w - yy * nc;
        base[(size_t)yy * nxh + c] = res[c * ny + yy];
    }
}

// (2) per-mode tridiagonal solve along z (FP64 Thomas with tabulated pivots).
template <typename T>
__global__ void __launch_bounds__(256) zsolve_kernel(cplx<T>* __restrict__ spec, double* __restrict__ dscr,
                                                     const double* __restrict__ cpt, int nmodes, int n,
                                                     double dz2, int split_mean, double* __restrict__ phibar,
                                                     double inv_nxy, double voltage)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmodes) return;
    double pr = 0.0, pi = 0.0;
    for (int k = 0; k < n; ++k) {
        const size_t o = (size_t)k * nmodes + m;
        const cplx<T> sv = spec[o];
        const double c = cpt[o];
        const double rr = dz2 * (double)sv.x, ri = dz2 * (double)sv.y;
        pr = (rr - pr) * c;
        pi = (ri - pi) * c;
        dscr[2 * o] = pr;
        dscr[2 * o + 1] = pi;
    }
    double xr = 0.0, xi = 0.0;
    for (int k = n - 1; k >= 0; --k) {
        const size_t o = (size_t)k * nmodes + m;
        if (k == n - 1) {
            xr = dscr[2 * o];
            xi = dscr[2 * o + 1];
        } else {
            const double c = cpt[o];
            xr = dscr[2 * o] - c * xr;
            xi = dscr[2 * o + 1] - c * xi;
        }
        if (split_mean && m == 0) {
            const int z = k + 1;
            phibar[z] = voltage * (1.0 - (double)z / (double)(n + 1)) + xr * inv_nxy;
            spec[o] = {(T)0, (T)0};
        } else {
            spec[o] = {(T)xr, (T)xi};
        }
    }
    if (split_mean && m == 0) {
        phibar[0] = voltage;
        phibar[n + 1] = 0.0;
    }
}

// (3b) C2R along x, two real rows per complex line, scaled by 1/(nx*ny), plus
// the wall potential (F64) or nothing (FP32S fluctuation).
template <typename T>
__global__ void __launch_bounds__(256) fftx_inv_kernel(const cplx<T>* __restrict__ spec, T* __restrict__ phi,
                                                       const cplx<T>* __restrict__ tw, FftPlan pl, int nx,
                                                       int ny, int nz, int nxh, int npairs, int lpb,
                                                       double voltage, int add_linear)
{
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* A = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* B = A + lpb * nx;
    const int l0 = blockIdx.x * lpb;
    const int nl = min(lpb, npairs - l0);
    for (int w = threadIdx.x; w < nl * nx; w += blockDim.x) {
        const int l = w / nx, k = w - l * nx;
        const int r0 = 2 * (l0 + l);
        cplx<T> a, b;
        if (k < nxh) {
            a = spec[(size_t)r0 * nxh + k];
            b = spec[(size_t)(r0 + 1) * nxh + k];
            if (k == 0 || 2 * k == nx) { a.y = (T)0; b.y = (T)0; }
        } else {
            a = spec[(size_t)r0 * nxh + (nx - k)];
            b = spec[(size_t)(r0 + 1) * nxh + (nx - k)];
            a.y = -a.y;
            b.y = -b.y;
        }
        // Z = A + i B
        A[l * nx + k] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    cplx<T>* res = lds_fft<T>(A, B, nx, nl, nx, pl, tw, +1);
    const double inv = 1.0 / ((double)nx * (double)ny);
    const int nxy = nx * ny;
    for (int w = threadIdx.x; w < nl * nx; w += blockDim.x) {
        const int l = w / nx, i = w - l * nx;
        const int r0 = 2 * (l0 + l), r1 = r0 + 1;
        const int z0 = r0 / ny + 1, z1 = r1 / ny + 1;
        const cplx<T> v = res[l * nx + i];
        double p0 = (double)v.x * inv, p1 = (double)v.y * inv;
        if (add_linear) {
            p0 += voltage * (1.0 - (double)z0 / (double)(nz - 1));
            p1 += voltage * (1.0 - (double)z1 / (double)(nz - 1));
        }
        phi[(size_t)z0 * nxy + (size_t)(r0 % ny) * nx + i] = (T)p0;
        phi[(size_t)z1 * nxy + (size_t)(r1 % ny) * nx + i] = (T)p1;
    }
}

// Dirichlet values of the potential on the walls (odd_extract, poisson.cu:158-165).
__global__ void __launch_bounds__(256) phi_walls_kernel(double* phi, int nxy, int nz, double voltage)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    phi[i] = voltage;
    phi[(size_t)(nz - 1) * nxy + i] = 0.0;
}

template <typename T>
hipError_t poisson_t(const LbmParams& P, const PoissonBufs& pb, const T* q, T* phi, double voltage, double eps,
                     bool split_mean, hipStream_t st)
{
    const int nx = pb.nx, ny = pb.ny, nz = pb.nz, nxh = pb.nxh;
    const int rows = (nz - 2) * ny;
    const int npairs = rows / 2;
    FftPlan plx = make_fft_plan(nx), ply = make_fft_plan(ny);
    const size_t cs = sizeof(cplx<T>);
    // x passes: lines of nx, two LDS buffers
    int lpb = (int)(32768 / (2 * nx * cs));
    if (lpb < 1) lpb = 1;
    if (lpb > npairs) lpb = npairs;
    size_t shx = 2 * (size_t)lpb * nx * cs;
    int gx = (npairs + lpb - 1) / lpb;
    // y passes: strips of w columns
    int wstrip = (int)(65536 / (2 * ny * cs));
    if (wstrip > 16) wstrip = 16;
    if (wstrip < 1) wstrip = 1;
    const int nstrips = (nxh + wstrip - 1) / wstrip;
    size_t shy = 2 * (size_t)wstrip * ny * cs;
    const int gy = nstrips * (nz - 2);

    cplx<T>* spec = (cplx<T>*)pb.spec;
    const cplx<T>* twx = (const cplx<T>*)pb.twx;
    const cplx<T>* twy = (const cplx<T>*)pb.twy;

    fftx_fwd_kernel<T><<<gx, 256, shx, st>>>(q, spec, twx, plx, nx, ny, nxh, npairs, lpb, eps);
    ffty_kernel<T><<<gy, 256, shy, st>>>(spec, twy, ply, ny, nxh, wstrip, nstrips, -1);
    const int nmodes = nxh * ny;
    const double dz2 = P.dz * P.dz;
    zsolve_kernel<T><<<(nmodes + 255) / 256, 256, 0, st>>>(spec, pb.dscr, pb.cpt, nmodes, nz - 2, dz2,
                                                          split_mean ? 1 : 0, pb.phibar,
                                                          1.0 / ((double)nx * (double)ny), voltage);
    ffty_kernel<T><<<gy, 256, shy, st>>>(spec, twy, ply, ny, nxh, wstrip, nstrips, +1);
    fftx_inv_kernel<T><<<gx, 256, shx, st>>>(spec, phi, twx, plx, nx, ny, nz, nxh, npairs, lpb, voltage,
                                             split_mean ? 0 : 1);
    if (!split_mean) phi_walls_kernel<<<(nx * ny + 255) / 256, 256, 0, st>>>((double*)phi, nx * ny, nz, voltage);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_poisson(Prec prec, const LbmParams& P, const PoissonBufs& pb, const void* q, void* phi,
                          double voltage, double eps, hipStream_t st)
{
    if (prec == Prec::F64)
        return poisson_t<double>(P, pb, (const double*)q, (double*)phi, voltage, eps, false, st);
    return poisson_t<float>(P, pb, (const float*)q, (float*)phi, voltage, eps, true, st);
}

}  // namespace ehd
