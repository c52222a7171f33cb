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

// ---------------------------------------------------------------------------
// Compile-time planned FFT (the grid sizes the configurations use): the radix
// schedule, the strides and every index division are constants.
constexpr int ct_radix(int n, int st)
{
    int r = n, s = 0;
    const int pref[4] = {4, 2, 3, 5};
    for (int i = 0; i < 4; ++i)
        while (r % pref[i] == 0) {
            if (s == st) return pref[i];
            r /= pref[i];
            ++s;
        }
    return 0;
}
constexpr int ct_nst(int n)
{
    int s = 0;
    while (ct_radix(n, s) != 0) ++s;
    return s;
}
template <typename T, int R>
__device__ __forceinline__ void bfly(cplx<T>* v, T sg)
{
    if constexpr (R == 2) {
        const cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        const cplx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
        const cplx<T> it3 = cis(t3, sg);
        v[0] = cadd(t0, t2);
        v[1] = cadd(t1, it3);
        v[2] = csub(t0, t2);
        v[3] = csub(t1, it3);
    } else if constexpr (R == 3) {
        const T s3 = (T)0.86602540378443864676;
        const cplx<T> a = v[0], t1 = cadd(v[1], v[2]), t3 = csub(v[1], v[2]);
        const cplx<T> t2 = {a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y};
        const cplx<T> r = cis(cplx<T>{s3 * t3.x, s3 * t3.y}, sg);
        v[0] = cadd(a, t1);
        v[1] = cadd(t2, r);
        v[2] = csub(t2, r);
    } else {
        static_assert(R == 5, "radix");
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        const cplx<T> a = v[0];
        const cplx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
        const cplx<T> m1 = {a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
        const cplx<T> m2 = {a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
        const cplx<T> n1 = cis(cplx<T>{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, sg);
        const cplx<T> n2 = cis(cplx<T>{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, sg);
        v[0] = cadd(a, cadd(t1, t2));
        v[1] = cadd(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
        v[4] = csub(m1, n1);
    }
}

// LDS index with one pad element per 16: the stride-R stores of a Stockham
// stage then spread over all banks instead of a quarter of them.
__host__ __device__ constexpr int padi(int i) { return i + (i >> 4); }
// padded LDS line length of the compile-time path, the plain length otherwise
__host__ __device__ constexpr int lds_line(int N, int n) { return N > 0 ? padi(N - 1) + 1 : n; }

template <typename T, int N, int ST, int NS>
__device__ __forceinline__ cplx<T>* ct_stages(cplx<T>* in, cplx<T>* out, int nlines, int ldl,
                                              const cplx<T>* __restrict__ tw, int sign)
{
    if constexpr (ST >= ct_nst(N)) {
        return in;
    } else {
        constexpr int R = ct_radix(N, ST);
        constexpr int NB = N / R;
        constexpr int TSTEP = N / (NS * R);
        const T sg = (T)sign;
        for (int w = threadIdx.x; w < nlines * NB; w += blockDim.x) {
            const int l = w / NB, j = w - l * NB, k = j % NS;
            const cplx<T>* src = in + l * ldl;
            cplx<T>* dst = out + l * ldl;
            const int o = (j / NS) * NS * R + k;
            cplx<T> v[R];
            v[0] = src[padi(j)];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                if constexpr (NS == 1) v[r] = src[padi(j + r * NB)];
                else v[r] = cmul(src[padi(j + r * NB)], root(tw, r * k * TSTEP, sign));
            }
            bfly<T, R>(v, sg);
#pragma unroll
            for (int r = 0; r < R; ++r) dst[padi(o + r * NS)] = v[r];
        }
        __syncthreads();
        return ct_stages<T, N, ST + 1, NS * R>(out, in, nlines, ldl, tw, sign);
    }
}

// N > 0: compile-time size; N == 0: runtime plan pl.
template <typename T, int N>
__device__ __forceinline__ cplx<T>* run_fft(cplx<T>* A, cplx<T>* B, int n, int nlines, int ldl, const FftPlan& pl,
                                            const cplx<T>* tw, int sign)
{
    if constexpr (N > 0) {
        (void)n;
        (void)pl;
        return ct_stages<T, N, 0, 1>(A, B, nlines, ldl, tw, sign);
    } else {
        return lds_fft<T>(A, B, n, nlines, ldl, pl, tw, sign);
    }
}

// (1a) R2C along x of s = -q/eps on interior rows, two real rows per complex
// line.  Rows are numbered r = (z-1)*ny + y; line l holds rows 2l and 2l+1.
// LDS: twiddles (nx), then two line buffers of lpb*nx.
template <typename T, int N>
__global__ void __launch_bounds__(256) fftx_fwd_kernel(const T* __restrict__ q, cplx<T>* __restrict__ spec,
                                                       const cplx<T>* __restrict__ tw, FftPlan pl, int nx_rt,
                                                       int ny, int npairs, int lpb, double eps)
{
    const int nx = N > 0 ? N : nx_rt;
    const int nxh = nx / 2 + 1;
    const int ldl = lds_line(N, nx);
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* tws = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tws + nx;
    cplx<T>* B = A + lpb * ldl;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tws[i] = tw[i];
    const int l0 = blockIdx.x * lpb;
    const int nl = min(lpb, npairs - l0);
    const int nxy = nx * ny;
    const double neps = -1.0 / eps;
    for (int w = threadIdx.x; w < nl * nx; w += blockDim.x) {
        const int l = w / nx, i = w - l * nx;
        const int r0 = 2 * (l0 + l);
        const int z = r0 / ny + 1, y = r0 - (z - 1) * ny;
        const size_t o0 = (size_t)z * nxy + (size_t)y * nx + i;
        A[l * ldl + (N > 0 ? padi(i) : i)] = {(T)((double)q[o0] * neps), (T)((double)q[o0 + nx] * neps)};
    }
    __syncthreads();
    cplx<T>* res = run_fft<T, N>(A, B, nx, nl, ldl, pl, tws, -1);
    for (int w = threadIdx.x; w < nl * nxh; w += blockDim.x) {
        const int l = w / nxh, k = w - l * nxh;
        const int kn = (nx - k) % nx;
        const cplx<T> zk = res[l * ldl + (N > 0 ? padi(k) : k)];
        const cplx<T> zn = res[l * ldl + (N > 0 ? padi(kn) : kn)];
        // A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i)
        const cplx<T> ak = {(T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y)};
        const cplx<T> bk = {(T)0.5 * (zk.y + zn.y), (T)-0.5 * (zk.x - zn.x)};
        const int r0 = 2 * (l0 + l);
        spec[(size_t)r0 * nxh + k] = ak;
        spec[(size_t)(r0 + 1) * nxh + k] = bk;
    }
}

// (1b)/(3a) complex DFT along y, in place, for strips of W columns.
// LDS: twiddles (ny), then two buffers of W*ny.
template <typename T, int N, int W>
__global__ void __launch_bounds__(256) ffty_kernel(cplx<T>* __restrict__ spec, const cplx<T>* __restrict__ tw,
                                                   FftPlan pl, int ny_rt, int nxh, int nstrips, int sign)
{
    const int ny = N > 0 ? N : ny_rt;
    const int ldl = lds_line(N, ny);
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* tws = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tws + ny;
    cplx<T>* B = A + W * ldl;
    for (int i = threadIdx.x; i < ny; i += blockDim.x) tws[i] = tw[i];
    const int plane = blockIdx.x / nstrips;
    const int c0 = (blockIdx.x - plane * nstrips) * W;
    const int nc = min(W, nxh - c0);
    cplx<T>* base = spec + (size_t)plane * ny * nxh + c0;
    for (int w = threadIdx.x; w < W * ny; w += blockDim.x) {
        const int yy = w / W, c = w - yy * W;
        if (c < nc) A[c * ldl + (N > 0 ? padi(yy) : yy)] = base[(size_t)yy * nxh + c];
    }
    __syncthreads();
    cplx<T>* res = run_fft<T, N>(A, B, ny, nc, ldl, pl, tws, sign);
    for (int w = threadIdx.x; w < W * ny; w += blockDim.x) {
        const int yy = w / W, c = w - yy * W;
        if (c < nc) base[(size_t)yy * nxh + c] = res[c * ldl + (N > 0 ? padi(yy) : yy)];
    }
}

// (2) per-mode tridiagonal solve along z, FP64, by the partition method.
// For each (kx,ky) mode the system is x[k-1] + b x[k] + x[k+1] = dz^2 s[k],
// k = 0..n-1, x[-1] = x[n] = 0, b = -(2 + k^2 dz^2).  A block holds MB modes
// (a tile [n][MB] staged in LDS for coalesced access) and P = 256/MB threads
// per mode, each owning a chunk of at most MAXM consecutive k.  Each thread
// solves with the chunk's own matrix T (Thomas, in registers)
//   T y = r,  T v = e_first,  (T w = e_last, w = reverse of v for Toeplitz T),
// so that x = y - x_left v - x_right w.  The 2P interface unknowns (first and
// last x of every chunk) satisfy a banded system that one thread per mode
// eliminates in order; the chunks then finish independently.
constexpr int ZS_MAXM = 16;

template <typename T, int P>
__global__ void __launch_bounds__(256) zsolve_kernel(cplx<T>* __restrict__ spec, int nmodes, int n, int nxh, int ny,
                                                     double kxs, double kys, double dz2, int split_mean,
                                                     double* __restrict__ phibar, double inv_nxy, double voltage)
{
    constexpr int MB = 256 / P;
    constexpr int MAXM = ZS_MAXM;
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);                 // [n][MB]
    double* xch = reinterpret_cast<double*>(tile + (size_t)n * MB);       // [MB][P][6]
    double* res = xch + MB * P * 6;                                        // [MB][P][4]
    const int m0 = blockIdx.x * MB;
    const int nm = min(MB, nmodes - m0);
    for (int w = threadIdx.x; w < n * MB; w += 256) {
        const int k = w / MB, c = w - k * MB;
        if (c < nm) tile[w] = spec[(size_t)k * nmodes + m0 + c];
    }
    __syncthreads();
    const int c = threadIdx.x % MB, p = threadIdx.x / MB;
    const int m = m0 + c;
    const int Peff = min(P, n);
    const bool active = (c < nm) && (p < Peff);
    const int k0 = (p * n) / Peff, len = ((p + 1) * n) / Peff - k0;
    const int kxi = m % nxh, kyi = m / nxh;
    const double kxv = kxi * kxs, kyv = (kyi <= ny / 2 ? kyi : kyi - ny) * kys;
    const double bd = -(2.0 + (kxv * kxv + kyv * kyv) * dz2);

    double cp[MAXM], yr[MAXM], yi[MAXM], v[MAXM];
    if (active) {
        // forward elimination
#pragma unroll
        for (int i = 0; i < MAXM; ++i) {
            if (i < len) {
                const double den = (i == 0) ? 1.0 / bd : 1.0 / (bd - cp[i - 1]);
                cp[i] = den;
                const cplx<T> sv = tile[(k0 + i) * MB + c];
                const double rr = dz2 * (double)sv.x, ri = dz2 * (double)sv.y;
                yr[i] = (i == 0 ? rr : rr - yr[i - 1]) * den;
                yi[i] = (i == 0 ? ri : ri - yi[i - 1]) * den;
                v[i] = (i == 0 ? 1.0 : -v[i - 1]) * den;
            }
        }
        // back substitution
#pragma unroll
        for (int i = MAXM - 2; i >= 0; --i) {
            if (i < len - 1) {
                yr[i] -= cp[i] * yr[i + 1];
                yi[i] -= cp[i] * yi[i + 1];
                v[i] -= cp[i] * v[i + 1];
            }
        }
        double ylr = 0.0, yli = 0.0, vl = 0.0;
#pragma unroll
        for (int i = 0; i < MAXM; ++i)
            if (i == len - 1) {
                ylr = yr[i];
                yli = yi[i];
                vl = v[i];
            }
        double* e = xch + (c * P + p) * 6;
        e[0] = yr[0]; e[1] = yi[0]; e[2] = ylr; e[3] = yli; e[4] = v[0]; e[5] = vl;
    }
    __syncthreads();
    if (p == 0 && c < nm) {
        // interface system: x_l(q-1) = al + be * x_f(q) carried along the chunks
        double gr[P], gi[P], de[P], alr[P], ali[P], bet[P];
        double ar = 0.0, ai = 0.0, be = 0.0;
#pragma unroll
        for (int qq = 0; qq < P; ++qq) {
            if (qq < Peff) {
                const double* e = xch + (c * P + qq) * 6;
                const double vf = e[4], vlq = e[5];
                const double wf = vlq, wl = vf;   // persymmetric chunk matrix
                const double den = 1.0 / (1.0 + vf * be);
                gr[qq] = (e[0] - vf * ar) * den;
                gi[qq] = (e[1] - vf * ai) * den;
                de[qq] = -wf * den;
                const double nar = e[2] - vlq * ar - vlq * be * gr[qq];
                const double nai = e[3] - vlq * ai - vlq * be * gi[qq];
                const double nbe = -vlq * be * de[qq] - wl;
                ar = nar; ai = nai; be = nbe;
                alr[qq] = ar; ali[qq] = ai; bet[qq] = be;
            }
        }
        double xfr = 0.0, xfi = 0.0;   // x_f(q+1)
#pragma unroll
        for (int qq = P - 1; qq >= 0; --qq) {
            if (qq < Peff) {
                double* r = res + (c * P + qq) * 4;
                r[2] = xfr;            // right neighbour x_f(q+1)
                r[3] = xfi;
                const double xlr = alr[qq] + bet[qq] * xfr, xli = ali[qq] + bet[qq] * xfi;
                if (qq + 1 < Peff) {
                    double* rn = res + (c * P + qq + 1) * 4;
                    rn[0] = xlr;       // left neighbour of chunk q+1
                    rn[1] = xli;
                }
                const double nxr = gr[qq] + de[qq] * xfr, nxi = gi[qq] + de[qq] * xfi;
                xfr = nxr;
                xfi = nxi;
            }
        }
        res[(c * P) * 4 + 0] = 0.0;
        res[(c * P) * 4 + 1] = 0.0;
    }
    __syncthreads();
    if (active) {
        const double* r = res + (c * P + p) * 4;
        const double Lr = r[0], Li = r[1], Rr = r[2], Ri = r[3];
        double wn = 0.0;
#pragma unroll
        for (int i = MAXM - 1; i >= 0; --i) {
            if (i < len) {
                const double wi = (i == len - 1) ? cp[i] : -cp[i] * wn;
                wn = wi;
                const double xr = yr[i] - Lr * v[i] - Rr * wi;
                const double xi = yi[i] - Li * v[i] - Ri * wi;
                if (split_mean && m == 0) {
                    const int z = k0 + i + 1;
                    phibar[z] = voltage * (1.0 - (double)z / (double)(n + 1)) + xr * inv_nxy;
                    tile[(k0 + i) * MB + c] = {(T)0, (T)0};
                } else {
                    tile[(k0 + i) * MB + c] = {(T)xr, (T)xi};
                }
            }
        }
    }
    if (split_mean && m0 == 0 && threadIdx.x == 0) {
        phibar[0] = voltage;
        phibar[n + 1] = 0.0;
    }
    __syncthreads();
    for (int w = threadIdx.x; w < n * MB; w += 256) {
        const int k = w / MB, cc = w - k * MB;
        if (cc < nm) spec[(size_t)k * nmodes + m0 + cc] = tile[w];
    }
}

// (3b) C2R along x, two real rows per complex line, scaled by 1/(nx*ny), plus
// the wall potential (F64) or nothing (FP32S fluctuation).
template <typename T, int N>
__global__ void __launch_bounds__(256) fftx_inv_kernel(const cplx<T>* __restrict__ spec, T* __restrict__ phi,
                                                       const cplx<T>* __restrict__ tw, FftPlan pl, int nx_rt,
                                                       int ny, int nz, int npairs, int lpb, double voltage,
                                                       int add_linear)
{
    const int nx = N > 0 ? N : nx_rt;
    const int nxh = nx / 2 + 1;
    const int ldl = lds_line(N, nx);
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* tws = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tws + nx;
    cplx<T>* B = A + lpb * ldl;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tws[i] = tw[i];
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
        A[l * ldl + (N > 0 ? padi(k) : k)] = {a.x - b.y, a.y + b.x};
    }
    __syncthreads();
    cplx<T>* res = run_fft<T, N>(A, B, nx, nl, ldl, pl, tws, +1);
    const double inv = 1.0 / ((double)nx * (double)ny);
    const int nxy = nx * ny;
    for (int w = threadIdx.x; w < nl * nx; w += blockDim.x) {
        const int l = w / nx, i = w - l * nx;
        const int r0 = 2 * (l0 + l);
        const int z = r0 / ny + 1, y = r0 - (z - 1) * ny;
        const cplx<T> v = res[l * ldl + (N > 0 ? padi(i) : i)];
        double p0 = (double)v.x * inv, p1 = (double)v.y * inv;
        if (add_linear) {
            const double lin = voltage * (1.0 - (double)z / (double)(nz - 1));
            p0 += lin;
            p1 += lin;
        }
        const size_t o = (size_t)z * nxy + (size_t)y * nx + i;
        phi[o] = (T)p0;
        phi[o + nx] = (T)p1;
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

constexpr size_t kMaxLds = 65536;

template <typename T, int N>
hipError_t launch_ffty(int W, int grid, size_t shm, hipStream_t st, cplx<T>* spec, const cplx<T>* tw,
                       const FftPlan& pl, int ny, int nxh, int nstrips, int sign)
{
    switch (W) {
    case 16: ffty_kernel<T, N, 16><<<grid, 256, shm, st>>>(spec, tw, pl, ny, nxh, nstrips, sign); break;
    case 8: ffty_kernel<T, N, 8><<<grid, 256, shm, st>>>(spec, tw, pl, ny, nxh, nstrips, sign); break;
    case 4: ffty_kernel<T, N, 4><<<grid, 256, shm, st>>>(spec, tw, pl, ny, nxh, nstrips, sign); break;
    case 2: ffty_kernel<T, N, 2><<<grid, 256, shm, st>>>(spec, tw, pl, ny, nxh, nstrips, sign); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename T, int NX, int NY>
hipError_t poisson_sized(const LbmParams& P, const PoissonBufs& pb, const T* q, T* phi, double voltage, double eps,
                         bool split_mean, hipStream_t st)
{
    const int nx = pb.nx, ny = pb.ny, nz = pb.nz, nxh = pb.nxh;
    const int n = nz - 2;
    const int npairs = n * ny / 2;
    const FftPlan plx = make_fft_plan(nx), ply = make_fft_plan(ny);
    const size_t cs = sizeof(cplx<T>);
    // x passes: lpb complex lines (row pairs) per block, two buffers, ~32 KB
    const int ldx = lds_line(NX, nx), ldy = lds_line(NY, ny);
    int lpb = (int)(32768 / (2 * ldx * cs));
    if (lpb < 1) lpb = 1;
    if (lpb > npairs) lpb = npairs;
    const size_t shx = (size_t)nx * cs + 2 * (size_t)lpb * ldx * cs;
    const int gx = (npairs + lpb - 1) / lpb;
    // y passes: strips of W columns, the widest (up to 16) whose buffers fit 64 KB
    int W = 16;
    while (W > 2 && (size_t)ny * cs + 2 * (size_t)W * ldy * cs > kMaxLds) W /= 2;
    const int nstrips = (nxh + W - 1) / W;
    const size_t shy = (size_t)ny * cs + 2 * (size_t)W * ldy * cs;
    if (shx > kMaxLds || shy > kMaxLds) return hipErrorInvalidValue;
    const int gy = nstrips * n;

    cplx<T>* spec = (cplx<T>*)pb.spec;
    const cplx<T>* twx = (const cplx<T>*)pb.twx;
    const cplx<T>* twy = (const cplx<T>*)pb.twy;

    fftx_fwd_kernel<T, NX><<<gx, 256, shx, st>>>(q, spec, twx, plx, nx, ny, npairs, lpb, eps);
    hipError_t e = launch_ffty<T, NY>(W, gy, shy, st, spec, twy, ply, ny, nxh, nstrips, -1);
    if (e != hipSuccess) return e;

    // z solve: chunks per mode P (8, 16 or 32) with chunk length <= ZS_MAXM and the tile under 48 KB
    const int nmodes = nxh * ny;
    int Pz = 0;
    for (int cand : {8, 16, 32}) {
        const int mb = 256 / cand;
        const size_t lds = (size_t)n * mb * cs + (size_t)mb * cand * 10 * sizeof(double);
        if ((n + cand - 1) / cand <= ZS_MAXM && lds <= kMaxLds) {
            Pz = cand;
            break;
        }
    }
    if (Pz == 0) return hipErrorInvalidValue;
    const int mb = 256 / Pz;
    const size_t shz = (size_t)n * mb * cs + (size_t)mb * Pz * 10 * sizeof(double);
    const int gz = (nmodes + mb - 1) / mb;
    const double kxs = 2.0 * 3.14159265358979323846 / (nx * P.dx);
    const double kys = 2.0 * 3.14159265358979323846 / (ny * P.dy);
    const double dz2 = P.dz * P.dz, inv_nxy = 1.0 / ((double)nx * (double)ny);
    const int sm = split_mean ? 1 : 0;
    if (Pz == 8)
        zsolve_kernel<T, 8><<<gz, 256, shz, st>>>(spec, nmodes, n, nxh, ny, kxs, kys, dz2, sm, pb.phibar, inv_nxy, voltage);
    else if (Pz == 16)
        zsolve_kernel<T, 16><<<gz, 256, shz, st>>>(spec, nmodes, n, nxh, ny, kxs, kys, dz2, sm, pb.phibar, inv_nxy, voltage);
    else
        zsolve_kernel<T, 32><<<gz, 256, shz, st>>>(spec, nmodes, n, nxh, ny, kxs, kys, dz2, sm, pb.phibar, inv_nxy, voltage);

    e = launch_ffty<T, NY>(W, gy, shy, st, spec, twy, ply, ny, nxh, nstrips, +1);
    if (e != hipSuccess) return e;
    fftx_inv_kernel<T, NX><<<gx, 256, shx, st>>>(spec, phi, twx, plx, nx, ny, nz, npairs, lpb, voltage, split_mean ? 0 : 1);
    if (!split_mean) phi_walls_kernel<<<(nx * ny + 255) / 256, 256, 0, st>>>((double*)phi, nx * ny, nz, voltage);
    return hipGetLastError();
}

// Compile-time FFT sizes for the grids of the configurations; anything else
// takes the runtime-planned path.
template <typename T, int NX>
hipError_t poisson_ny(const LbmParams& P, const PoissonBufs& pb, const T* q, T* phi, double voltage, double eps,
                      bool split_mean, hipStream_t st)
{
    if (pb.ny == NX) return poisson_sized<T, NX, NX>(P, pb, q, phi, voltage, eps, split_mean, st);
    return poisson_sized<T, NX, 0>(P, pb, q, phi, voltage, eps, split_mean, st);
}

template <typename T>
hipError_t poisson_t(const LbmParams& P, const PoissonBufs& pb, const T* q, T* phi, double voltage, double eps,
                     bool split_mean, hipStream_t st)
{
    switch (pb.nx) {
    case 48: return poisson_ny<T, 48>(P, pb, q, phi, voltage, eps, split_mean, st);
    case 64: return poisson_ny<T, 64>(P, pb, q, phi, voltage, eps, split_mean, st);
    case 80: return poisson_ny<T, 80>(P, pb, q, phi, voltage, eps, split_mean, st);
    case 96: return poisson_ny<T, 96>(P, pb, q, phi, voltage, eps, split_mean, st);
    case 128: return poisson_ny<T, 128>(P, pb, q, phi, voltage, eps, split_mean, st);
    case 256: return poisson_ny<T, 256>(P, pb, q, phi, voltage, eps, split_mean, st);
    case 512: return poisson_ny<T, 512>(P, pb, q, phi, voltage, eps, split_mean, st);
    default: return poisson_sized<T, 0, 0>(P, pb, q, phi, voltage, eps, split_mean, st);
    }
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
