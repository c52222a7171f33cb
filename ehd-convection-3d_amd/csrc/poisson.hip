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
// Spectral layout: [z-1][kx][ky], kx in [0, nx/2], ky fastest.
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

// Buffer pointers of a solve come from its job table in memory, so the
// compiler cannot see that they are global: without the address space it
// emits flat loads and stores.  Kernels access them through gptr().
#define EHD_GLOBAL __attribute__((address_space(1)))
template <typename X>
__device__ __forceinline__ EHD_GLOBAL X* gptr(X* p)
{
    return (EHD_GLOBAL X*)p;
}
template <typename T>
__device__ __forceinline__ cplx<T> ldc(EHD_GLOBAL const cplx<T>* p)
{
    return cplx<T>{p->x, p->y};
}
template <typename T>
__device__ __forceinline__ void stc(EHD_GLOBAL cplx<T>* p, cplx<T> v)
{
    p->x = v.x;
    p->y = v.y;
}

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
// Compile-time planned FFT (the grid sizes the configurations use).  Radix 16
// and 8 butterflies run in registers (as 4x4 and 4x2 with constant inner
// twiddles), so a 256-point line takes two stages and one LDS exchange.  The
// radix schedule, the strides and every index division are constants.
// Radix of stage st: the power-of-two part 2^p is split into ceil(p/4)
// balanced stages (each 2..16, larger first: 512 = 8*8*8, 256 = 16*16,
// 128 = 16*8, 64 = 8*8), then the factors 3 and 5.  Balanced stages keep the
// butterflies per line (N / radix) small in every stage.
constexpr int ct_radix(int n, int st)
{
    if (n <= 0) return 0;
    int p = 0, r = n;
    while (r % 2 == 0) {
        r /= 2;
        ++p;
    }
    const int s2 = (p + 3) / 4;
    if (st < s2) {
        const int bits = p / s2 + (st < p % s2 ? 1 : 0);
        return 1 << bits;
    }
    int s = s2;
    const int odd[2] = {3, 5};
    for (int i = 0; i < 2; ++i)
        while (r % odd[i] == 0) {
            if (s == st) return odd[i];
            r /= odd[i];
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
// largest number of butterflies of one line over the stages (N / smallest radix)
constexpr int ct_max_nb(int n)
{
    int m = 0;
    for (int s = 0; s < ct_nst(n); ++s) m = (n / ct_radix(n, s) > m) ? n / ct_radix(n, s) : m;
    return m;
}
// the plan covers n completely (no factor left for a runtime stage)
constexpr bool ct_complete(int n)
{
    int r = n;
    for (int s = 0; s < ct_nst(n); ++s) r /= ct_radix(n, s);
    return r == 1;
}

// a * (c + i s)
template <typename T>
__device__ __forceinline__ cplx<T> crot(cplx<T> a, T c, T s)
{
    return {a.x * c - a.y * s, a.x * s + a.y * c};
}

template <typename T>
__device__ __forceinline__ void dft4(cplx<T>& a, cplx<T>& b, cplx<T>& c, cplx<T>& d, T sg)
{
    const cplx<T> t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = csub(b, d);
    const cplx<T> it3 = cis(t3, sg);
    a = cadd(t0, t2);
    b = cadd(t1, it3);
    c = csub(t0, t2);
    d = csub(t1, it3);
}

template <typename T, int R>
__device__ __forceinline__ void bfly(cplx<T>* v, T sg)
{
    if constexpr (R == 2) {
        const cplx<T> a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], sg);
    } else if constexpr (R == 8) {
        // n = 2 n1 + n2, k = k1 + 4 k2: DFT4 over n1, inner twiddle W8^(n2 k1), DFT2 over n2
        const T h = (T)0.70710678118654752440;
        cplx<T> a0 = v[0], a1 = v[2], a2 = v[4], a3 = v[6];
        cplx<T> b0 = v[1], b1 = v[3], b2 = v[5], b3 = v[7];
        dft4(a0, a1, a2, a3, sg);
        dft4(b0, b1, b2, b3, sg);
        b1 = crot(b1, h, sg * h);      // W8^1
        b2 = cis(b2, sg);              // W8^2 = i sg
        b3 = crot(b3, -h, sg * h);     // W8^3
        v[0] = cadd(a0, b0); v[4] = csub(a0, b0);
        v[1] = cadd(a1, b1); v[5] = csub(a1, b1);
        v[2] = cadd(a2, b2); v[6] = csub(a2, b2);
        v[3] = cadd(a3, b3); v[7] = csub(a3, b3);
    } else if constexpr (R == 16) {
        // n = 4 n1 + n2, k = k1 + 4 k2: DFT4 over n1, inner twiddle W16^(n2 k1), DFT4 over n2
        const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
        cplx<T> a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            a[n2][0] = v[n2]; a[n2][1] = v[4 + n2]; a[n2][2] = v[8 + n2]; a[n2][3] = v[12 + n2];
            dft4(a[n2][0], a[n2][1], a[n2][2], a[n2][3], sg);
        }
        a[1][1] = crot(a[1][1], c1, sg * s1);     // W16^1
        a[1][2] = crot(a[1][2], h, sg * h);       // W16^2
        a[1][3] = crot(a[1][3], s1, sg * c1);     // W16^3
        a[2][1] = crot(a[2][1], h, sg * h);       // W16^2
        a[2][2] = cis(a[2][2], sg);               // W16^4
        a[2][3] = crot(a[2][3], -h, sg * h);      // W16^6
        a[3][1] = crot(a[3][1], s1, sg * c1);     // W16^3
        a[3][2] = crot(a[3][2], -h, sg * h);      // W16^6
        a[3][3] = crot(a[3][3], -c1, -sg * s1);   // W16^9
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            cplx<T> p = a[0][k1], q = a[1][k1], r = a[2][k1], s = a[3][k1];
            dft4(p, q, r, s, sg);
            v[k1] = p; v[k1 + 4] = q; v[k1 + 8] = r; v[k1 + 12] = s;
        }
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

// LDS index with one pad element per 16: the stride-R accesses of a Stockham
// stage then spread over the banks instead of a few of them.
__host__ __device__ constexpr int padi(int i) { return i + (i >> 4); }
// LDS line pitch (complex elements): long enough for the padded line and
// = 1 mod 32, so that threads walking across lines hit distinct banks.
__host__ __device__ constexpr int lds_pitch(int N, int n)
{
    const int len = N > 0 ? padi(N - 1) + 1 : n;
    return ((len + 30) / 32) * 32 + 1;
}

// Runs the compile-time plan of N on nlines lines with one butterfly per
// thread per stage (nlines * ct_max_nb(N) <= blockDim.x).  Stage 0 reads its
// inputs through ld(l, i) -- global memory, or the LDS buffer itself when
// SRC_BUF; the other stages read buf.  Stages write back into buf in place
// (all reads of a stage complete before its writes), except that the last
// stage writes through sf(l, i, v) when DST_GLOBAL.  Every thread of the
// block must call it; the result is in buf (padded index) unless DST_GLOBAL.
template <typename T, int N, int ST, int NS, bool SRC_BUF, bool DST_GLOBAL, class LD, class SF>
__device__ __forceinline__ void ct_run(const LD& ld, const SF& sf, cplx<T>* buf, int nlines, int ldl,
                                       const cplx<T>* __restrict__ tw, int sign)
{
    constexpr int NST = ct_nst(N);
    if constexpr (ST < NST) {
        constexpr int R = ct_radix(N, ST);
        constexpr int NB = N / R;
        constexpr int TSTEP = N / (NS * R);
        constexpr bool from_buf = (ST > 0) || SRC_BUF;
        constexpr bool to_global = DST_GLOBAL && (ST == NST - 1);
        const T sg = (T)sign;
        const int w = threadIdx.x;
        const bool act = w < nlines * NB;
        const int l = w / NB, j = w - l * NB, k = j % NS;
        cplx<T> v[R];
        if (act) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cplx<T> x;
                if constexpr (ST == 0) x = ld(l, j + r * NB);
                else x = buf[l * ldl + padi(j + r * NB)];
                if constexpr (NS > 1) {
                    if (r > 0) x = cmul(x, root(tw, r * k * TSTEP, sign));
                }
                v[r] = x;
            }
        }
        if constexpr (from_buf && !to_global) __syncthreads();
        if (act) {
            bfly<T, R>(v, sg);
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if constexpr (to_global) sf(l, o + r * NS, v[r]);
                else buf[l * ldl + padi(o + r * NS)] = v[r];
            }
        }
        if constexpr (!to_global) __syncthreads();
        ct_run<T, N, ST + 1, NS * R, SRC_BUF, DST_GLOBAL>(ld, sf, buf, nlines, ldl, tw, sign);
    }
}

// Spectral layout (both precisions): [z-1][kx][ky], ky fastest, kx in
// [0, nx/2]: the y transforms run over contiguous lines and the z solve over
// contiguous modes m = kx*ny + ky.

// (1a) R2C along x of s = -q/eps on interior rows, two real rows per complex
// line.  A block takes 2*lpb consecutive rows y0.. of one plane; lpb divides
// ny/2.  The spectrum is written transposed, 2*lpb consecutive ky per kx.
// LDS: twiddles (nx), line buffers (lpb lines; two for the runtime plan).
template <typename T, int N>
__global__ void __launch_bounds__(256) fftx_fwd_kernel(const PoissonJob* __restrict__ jobs,
                                                       const cplx<T>* __restrict__ tw, FftPlan pl, int nx_rt,
                                                       int ny, int lpb)
{
    const PoissonJob& J = jobs[blockIdx.y];
    const int nx = N > 0 ? N : nx_rt;
    const int nxh = nx / 2 + 1;
    const int ldl = lds_pitch(N, nx);
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* tws = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tws + nx;
    const int groups = ny / (2 * lpb);
    const int zi = blockIdx.x / groups;              // interior plane index z-1
    if (zi >= J.n) return;                           // a shorter box of the batch
    const int y0 = (blockIdx.x - zi * groups) * 2 * lpb;
    EHD_GLOBAL const T* qp = gptr((const T*)J.q) + (size_t)(zi + 1) * nx * ny + (size_t)y0 * nx;
    EHD_GLOBAL cplx<T>* spec = gptr((cplx<T>*)J.spec);
    const double neps = -1.0 / J.eps;
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tws[i] = tw[i];
    __syncthreads();
    cplx<T>* res;
    if constexpr (N > 0) {
        auto ld = [&](int l, int i) -> cplx<T> {
            return {(T)((double)qp[(size_t)(2 * l) * N + i] * neps),
                    (T)((double)qp[(size_t)(2 * l + 1) * N + i] * neps)};
        };
        auto sf = [](int, int, cplx<T>) {};
        ct_run<T, N, 0, 1, false, false>(ld, sf, A, lpb, ldl, tws, -1);
        res = A;
    } else {
        cplx<T>* B = A + lpb * ldl;
        for (int w = threadIdx.x; w < lpb * nx; w += blockDim.x) {
            const int l = w / nx, i = w - l * nx;
            A[l * ldl + i] = {(T)((double)qp[(size_t)(2 * l) * nx + i] * neps),
                              (T)((double)qp[(size_t)(2 * l + 1) * nx + i] * neps)};
        }
        __syncthreads();
        res = lds_fft<T>(A, B, nx, lpb, ldl, pl, tws, -1);
    }
    const int rows = 2 * lpb;
    EHD_GLOBAL cplx<T>* out = spec + (size_t)zi * nxh * ny + y0;
    for (int w = threadIdx.x; w < nxh * rows; w += blockDim.x) {
        const int k = w / rows, r = w - k * rows;
        const int l = r >> 1;
        const int kn = (nx - k) % nx;
        const cplx<T> zk = res[l * ldl + (N > 0 ? padi(k) : k)];
        const cplx<T> zn = res[l * ldl + (N > 0 ? padi(kn) : kn)];
        // A_k = (Z_k + conj Z_{n-k})/2 (row 2l), B_k = (Z_k - conj Z_{n-k})/(2i) (row 2l+1)
        const cplx<T> v = (r & 1) ? cplx<T>{(T)0.5 * (zk.y + zn.y), (T)-0.5 * (zk.x - zn.x)}
                                  : cplx<T>{(T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y)};
        stc(out + (size_t)k * ny + r, v);
    }
}

// (1b)/(3a) complex DFT along y, in place, over contiguous lines of ny.
// A block takes G consecutive lines.
template <typename T, int N>
__global__ void __launch_bounds__(256) ffty_kernel(const PoissonJob* __restrict__ jobs, const cplx<T>* __restrict__ tw,
                                                   FftPlan pl, int ny_rt, int nxh, int G, int sign)
{
    const PoissonJob& J = jobs[blockIdx.y];
    const int ny = N > 0 ? N : ny_rt;
    const int ldl = lds_pitch(N, ny);
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* tws = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tws + ny;
    const int nlines = J.n * nxh;
    const int L0 = blockIdx.x * G;
    if (L0 >= nlines) return;                        // a shorter box of the batch
    const int nl = min(G, nlines - L0);
    EHD_GLOBAL cplx<T>* base = gptr((cplx<T>*)J.spec) + (size_t)L0 * ny;
    for (int i = threadIdx.x; i < ny; i += blockDim.x) tws[i] = tw[i];
    __syncthreads();
    if constexpr (N > 0) {
        auto ld = [&](int l, int i) -> cplx<T> { return ldc(base + (size_t)l * N + i); };
        auto sf = [&](int l, int i, cplx<T> v) { stc(base + (size_t)l * N + i, v); };
        ct_run<T, N, 0, 1, false, true>(ld, sf, A, nl, ldl, tws, sign);
    } else {
        cplx<T>* B = A + G * ldl;
        for (int w = threadIdx.x; w < nl * ny; w += blockDim.x) {
            const int l = w / ny, i = w - l * ny;
            A[l * ldl + i] = ldc(base + (size_t)l * ny + i);
        }
        __syncthreads();
        cplx<T>* res = lds_fft<T>(A, B, ny, nl, ldl, pl, tws, sign);
        for (int w = threadIdx.x; w < nl * ny; w += blockDim.x) {
            const int l = w / ny, i = w - l * ny;
            stc(base + (size_t)l * ny + i, res[l * ldl + i]);
        }
    }
}

// (2) per-mode tridiagonal solve along z, FP64, by the partition method.
// For each (kx,ky) mode the system is x[k-1] + b x[k] + x[k+1] = dz^2 s[k],
// k = 0..n-1, x[-1] = x[n] = 0, b = -(2 + k^2 dz^2).  A block holds MB =
// NT/P modes and P threads per mode, each owning a chunk of at most MAXM
// (<= ZS_MAXM) consecutive k; the threads of a wave hold consecutive modes,
// so every load and store of a chunk row is a contiguous run of modes.  Each
// thread loads its chunk straight into registers and solves with the
// chunk's own matrix T (Thomas)
//   T y = r,  T v = e_first,  (T w = e_last, w = reverse of v for Toeplitz T),
// so that x = y - x_left v - x_right w.  The interface unknowns (first and
// last x of every chunk) obey three recurrences along the chunks:
//   be_q = -vf_q - vl_q^2 be_{q-1} / (1 + vf_q be_{q-1})  (pivots),
//   ar_q = al_q + bt_q ar_{q-1},   xf_q = gr_q + de_q xf_{q+1}
// (vf, vl the first and last entries of the chunk's v), after which every
// chunk has its neighbours x_l(q-1) = ar_{q-1} + be_{q-1} xf_q and x_f(q+1)
// and finishes independently.  Per-chunk records go through LDS, fields
// outermost and modes innermost, so the threads of a wave hit consecutive
// banks.  Two ways to solve the recurrences:
// - FP64 spectrum (the parity path): one thread per mode runs them in order.
// - FP32 spectrum: three doubling (Hillis-Steele) scans over the P threads of
//   the mode: prefix products of the Moebius matrices
//   M_q = [[vl^2 - vf^2, -vf], [vf, 1]], be_q = N_q / D_q with
//   [N_q; D_q] = M_q [N_{q-1}; D_{q-1}], rescaled by powers of two; an affine
//   prefix for ar; an affine suffix for xf.  25 -> 21.5 us on the 256x256x128
//   FP32 spectrum.  The composed matrices lose more to rounding than the
//   ordered recurrence: against the oracle at 64x48x200 the FP64 solve was
//   7e-13 with scans and 3e-14 in order, so the FP64 spectrum keeps the order.
//   Rounded to float the two agree.
// LDS: 6 P MB doubles.
constexpr int ZS_MAXM = 16;

thread_local const char* g_failed_pass = "";

template <typename T, int P, int MAXM, int NT = 256>
__global__ void __launch_bounds__(NT) zsolve_kernel(const PoissonJob* __restrict__ jobs, int nmodes, int ny,
                                                      int split_mean, double inv_nxy)
{
    constexpr int MB = NT / P;
    const PoissonJob& J = jobs[blockIdx.y];
    EHD_GLOBAL cplx<T>* __restrict__ spec = gptr((cplx<T>*)J.spec);
    EHD_GLOBAL double* __restrict__ phibar = gptr(J.phibar);
    const int n = J.n;
    const double kxs = J.kxs, kys = J.kys, dz2 = J.dz2, voltage = J.voltage;
    extern __shared__ unsigned char smem_raw[];
    double* sc = reinterpret_cast<double*>(smem_raw);   // [6][P][MB] chunk records
    auto Y = [&](int f, int q, int cc) -> double& { return sc[(f * P + q) * MB + cc]; };
    const int m0 = blockIdx.x * MB;
    const int nm = min(MB, nmodes - m0);
    const int c = threadIdx.x % MB, p = threadIdx.x / MB;
    const int m = m0 + c;
    const int Peff = min(P, n);
    const bool active = (c < nm) && (p < Peff);
    const int k0 = (p * n) / Peff, len = ((p + 1) * n) / Peff - k0;
    T sx[MAXM], sy[MAXM];
#pragma unroll
    for (int i = 0; i < MAXM; ++i) {
        sx[i] = (T)0;
        sy[i] = (T)0;
        if (active && i < len) {
            const cplx<T> v = ldc(spec + (size_t)(k0 + i) * nmodes + m);
            sx[i] = v.x;
            sy[i] = v.y;
        }
    }
    const int kxi = m / ny, kyi = m - kxi * ny;
    const double kxv = kxi * kxs, kyv = (kyi <= ny / 2 ? kyi : kyi - ny) * kys;
    const double bd = -(2.0 + (kxv * kxv + kyv * kyv) * dz2);
    double cp[MAXM], yr[MAXM], yi[MAXM], v[MAXM];
    double e0 = 0.0, e1 = 0.0, e2 = 0.0, e3 = 0.0, vf = 0.0, vlq = 0.0;
    if (active) {
#pragma unroll
        for (int i = 0; i < MAXM; ++i) {
            if (i < len) {
                const double den = (i == 0) ? 1.0 / bd : 1.0 / (bd - cp[i - 1]);
                cp[i] = den;
                const double rr = dz2 * (double)sx[i], ri = dz2 * (double)sy[i];
                yr[i] = (i == 0 ? rr : rr - yr[i - 1]) * den;
                yi[i] = (i == 0 ? ri : ri - yi[i - 1]) * den;
                v[i] = (i == 0 ? 1.0 : -v[i - 1]) * den;
            }
        }
#pragma unroll
        for (int i = MAXM - 2; i >= 0; --i) {
            if (i < len - 1) {
                yr[i] -= cp[i] * yr[i + 1];
                yi[i] -= cp[i] * yi[i + 1];
                v[i] -= cp[i] * v[i + 1];
            }
        }
#pragma unroll
        for (int i = 0; i < MAXM; ++i)
            if (i == len - 1) {
                e2 = yr[i];
                e3 = yi[i];
                vlq = v[i];
            }
        e0 = yr[0]; e1 = yi[0]; vf = v[0];
    }
    double Lr = 0.0, Li = 0.0, Rr = 0.0, Ri = 0.0;
    if constexpr (sizeof(T) == sizeof(float)) {
        // (1) be_q = N_q / D_q, [N_q; D_q] = M_q [N_{q-1}; D_{q-1}], M_q = [[vl^2 - vf^2, -vf], [vf, 1]]:
        //     inclusive prefix products of the M by a doubling scan (identity for inactive chunks)
        double a = 1.0, b = 0.0, cc2 = 0.0, d = 1.0;
        if (active) { a = vlq * vlq - vf * vf; b = -vf; cc2 = vf; d = 1.0; }
#pragma unroll
        for (int off = 1; off < P; off <<= 1) {
            Y(0, p, c) = a; Y(1, p, c) = b; Y(2, p, c) = cc2; Y(3, p, c) = d;
            __syncthreads();
            if (p >= off) {
                const double a2 = Y(0, p - off, c), b2 = Y(1, p - off, c), c2 = Y(2, p - off, c), d2 = Y(3, p - off, c);
                const double na = a * a2 + b * c2, nb = a * b2 + b * d2, nc = cc2 * a2 + d * c2, nd = cc2 * b2 + d * d2;
                const double mx = fmax(fmax(fabs(na), fabs(nb)), fmax(fabs(nc), fabs(nd)));
                const int e = __builtin_amdgcn_frexp_exp(mx);
                a = ldexp(na, -e); b = ldexp(nb, -e); cc2 = ldexp(nc, -e); d = ldexp(nd, -e);
            }
            __syncthreads();
        }
        // exclusive value (N_{p-1}, D_{p-1}) = P_{p-1} [0; 1]
        Y(0, p, c) = b; Y(1, p, c) = d;
        __syncthreads();
        const double Nq = (p > 0) ? Y(0, p - 1, c) : 0.0, Dq = (p > 0) ? Y(1, p - 1, c) : 1.0;
        __syncthreads();
        double den = 0.0, de = 0.0, be = 0.0, al = 0.0, ali = 0.0, bt = 1.0;
        if (active) {
            const double inv = 1.0 / (Dq + vf * Nq);
            den = Dq * inv;
            const double bden = Nq * inv;
            de = -vlq * den;
            be = vlq * vlq * bden - vf;
            al = e2 - vlq * bden * e0;
            ali = e3 - vlq * bden * e1;
            bt = -vlq * den;
        } else {
            bt = 1.0; al = 0.0; ali = 0.0;
        }
        // (2) ar_q = al_q + bt_q ar_{q-1}: doubling scan of affine maps (identity for inactive chunks)
        double sa = al, sai = ali, sb = bt;
#pragma unroll
        for (int off = 1; off < P; off <<= 1) {
            Y(0, p, c) = sa; Y(1, p, c) = sai; Y(2, p, c) = sb;
            __syncthreads();
            if (p >= off) {
                const double pa = Y(0, p - off, c), pai = Y(1, p - off, c), pb = Y(2, p - off, c);
                sa = sa + sb * pa;
                sai = sai + sb * pai;
                sb = sb * pb;
            }
            __syncthreads();
        }
        // sa = ar_p; exclusive ar_{p-1} from the left neighbour
        Y(0, p, c) = sa; Y(1, p, c) = sai;
        __syncthreads();
        const double arp = (p > 0) ? Y(0, p - 1, c) : 0.0, aip = (p > 0) ? Y(1, p - 1, c) : 0.0;
        __syncthreads();
        const double gr = active ? (e0 - vf * arp) * den : 0.0, gi = active ? (e1 - vf * aip) * den : 0.0;
        // (3) xf_q = gr_q + de_q xf_{q+1}, xf_Peff = 0: doubling scan from the right
        double xa = gr, xai = gi, xb = active ? de : 1.0;
#pragma unroll
        for (int off = 1; off < P; off <<= 1) {
            Y(0, p, c) = xa; Y(1, p, c) = xai; Y(2, p, c) = xb;
            __syncthreads();
            if (p + off < P) {
                const double na = Y(0, p + off, c), nai = Y(1, p + off, c), nb = Y(2, p + off, c);
                xa = xa + xb * na;
                xai = xai + xb * nai;
                xb = xb * nb;
            }
            __syncthreads();
        }
        // xa = x_f(p); neighbours: x_f(p+1) from the right, x_l(p-1) = ar_{p-1} + be_{p-1} x_f(p)
        Y(0, p, c) = xa; Y(1, p, c) = xai; Y(2, p, c) = be;
        __syncthreads();
        Rr = (p + 1 < Peff) ? Y(0, p + 1, c) : 0.0;
        Ri = (p + 1 < Peff) ? Y(1, p + 1, c) : 0.0;
        const double bem = (p > 0) ? Y(2, p - 1, c) : 0.0;
        Lr = (p > 0) ? arp + bem * xa : 0.0;
        Li = (p > 0) ? aip + bem * xai : 0.0;
    } else {
        if (active) {
            Y(0, p, c) = e0; Y(1, p, c) = e1; Y(2, p, c) = e2; Y(3, p, c) = e3; Y(4, p, c) = vf;
            Y(5, p, c) = vlq;
        }
        __syncthreads();
        if (p == 0 && c < nm) {
            double ar = 0.0, ai = 0.0, be = 0.0;
#pragma unroll
            for (int qq = 0; qq < P; ++qq) {
                if (qq >= Peff) break;
                const double f0 = Y(0, qq, c), f1 = Y(1, qq, c), f2 = Y(2, qq, c), f3 = Y(3, qq, c);
                const double wl = Y(4, qq, c), wf = Y(5, qq, c);   // w reversed v: w_l = v_f, w_f = v_l
                const double dn = 1.0 / (1.0 + wl * be);
                const double gr = (f0 - wl * ar) * dn;
                const double gi = (f1 - wl * ai) * dn;
                const double de = -wf * dn;
                const double nar = f2 - wf * ar - wf * be * gr;
                const double nai = f3 - wf * ai - wf * be * gi;
                const double nbe = -wf * be * de - wl;
                ar = nar; ai = nai; be = nbe;
                Y(0, qq, c) = gr; Y(1, qq, c) = gi; Y(2, qq, c) = de; Y(3, qq, c) = ar; Y(4, qq, c) = ai;
                Y(5, qq, c) = be;
            }
            // back: record q gets (x_l(q-1), x_f(q+1)) in fields 0..3, written
            // after the record's own coefficients are read
            double xfr = 0.0, xfi = 0.0;
#pragma unroll
            for (int qq = P - 1; qq >= 0; --qq) {
                if (qq >= Peff) continue;
                const double gr = Y(0, qq, c), gi = Y(1, qq, c), de = Y(2, qq, c);
                const double ar = Y(3, qq, c), ai = Y(4, qq, c), be = Y(5, qq, c);
                const double xlr = ar + be * xfr, xli = ai + be * xfi;
                Y(2, qq, c) = xfr;
                Y(3, qq, c) = xfi;
                if (qq + 1 < Peff) {
                    Y(0, qq + 1, c) = xlr;
                    Y(1, qq + 1, c) = xli;
                }
                const double nxr = gr + de * xfr, nxi = gi + de * xfi;
                xfr = nxr;
                xfi = nxi;
            }
            Y(0, 0, c) = 0.0;
            Y(1, 0, c) = 0.0;
        }
        __syncthreads();
        if (active) {
            Lr = Y(0, p, c); Li = Y(1, p, c); Rr = Y(2, p, c); Ri = Y(3, p, c);
        }
    }
    if (active) {
        double wn = 0.0;
#pragma unroll
        for (int i = MAXM - 1; i >= 0; --i) {
            if (i < len) {
                const double wi = (i == len - 1) ? cp[i] : -cp[i] * wn;
                wn = wi;
                const double xr = yr[i] - Lr * v[i] - Rr * wi;
                const double xi = yi[i] - Li * v[i] - Ri * wi;
                EHD_GLOBAL cplx<T>* dst = spec + (size_t)(k0 + i) * nmodes + m;
                if (split_mean && m == 0) {
                    const int z = k0 + i + 1;
                    phibar[z] = voltage * (1.0 - (double)z / (double)(n + 1)) + xr * inv_nxy;
                    stc(dst, cplx<T>{(T)0, (T)0});
                } else {
                    stc(dst, cplx<T>{(T)xr, (T)xi});
                }
            }
        }
    }
    if (split_mean && m0 == 0 && threadIdx.x == 0) {
        phibar[0] = voltage;
        phibar[n + 1] = 0.0;
    }
}

// (3b) C2R along x, two real rows per complex line, scaled by 1/(nx*ny), plus
// the wall potential (F64) or nothing (FP32S fluctuation).  A block takes
// 2*lpb consecutive rows y0.. of one plane; the transposed spectrum is read
// as (row 2l, row 2l+1) pairs, coalesced along ky, and rebuilt in LDS as
// Z = A + i B with the Hermitian half.
template <typename T, int N>
__global__ void __launch_bounds__(256) fftx_inv_kernel(const PoissonJob* __restrict__ jobs,
                                                       const cplx<T>* __restrict__ tw, FftPlan pl, int nx_rt,
                                                       int ny, int lpb, int add_linear)
{
    const PoissonJob& J = jobs[blockIdx.y];
    const int nx = N > 0 ? N : nx_rt;
    const int nxh = nx / 2 + 1;
    const int ldl = lds_pitch(N, nx);
    extern __shared__ unsigned char smem_raw[];
    cplx<T>* tws = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* A = tws + nx;
    const int groups = ny / (2 * lpb);
    const int zi = blockIdx.x / groups;
    if (zi >= J.n) return;                           // a shorter box of the batch
    const int y0 = (blockIdx.x - zi * groups) * 2 * lpb;
    const int z = zi + 1;
    const int nz = J.n + 2;
    const double voltage = J.voltage;
    EHD_GLOBAL const cplx<T>* spec = gptr((const cplx<T>*)J.spec);
    EHD_GLOBAL T* phi = gptr((T*)J.phi);
    for (int i = threadIdx.x; i < nx; i += blockDim.x) tws[i] = tw[i];
    EHD_GLOBAL const cplx<T>* in = spec + (size_t)zi * nxh * ny + y0;
    // XI_LOADS row pairs in flight per thread before the LDS stores
    constexpr int XI_LOADS = 4;
    for (int w0 = threadIdx.x; w0 < nxh * lpb; w0 += blockDim.x * XI_LOADS) {
        cplx<T> av[XI_LOADS], bv[XI_LOADS];
#pragma unroll
        for (int u = 0; u < XI_LOADS; ++u) {
            const int w = w0 + u * blockDim.x;
            if (w < nxh * lpb) {
                const int k = w / lpb, l = w - k * lpb;
                av[u] = ldc(in + (size_t)k * ny + 2 * l);
                bv[u] = ldc(in + (size_t)k * ny + 2 * l + 1);
            }
        }
#pragma unroll
        for (int u = 0; u < XI_LOADS; ++u) {
            const int w = w0 + u * blockDim.x;
            if (w < nxh * lpb) {
                const int k = w / lpb, l = w - k * lpb;
                cplx<T> a = av[u], b = bv[u];
                if (k == 0 || 2 * k == nx) {
                    a.y = (T)0;
                    b.y = (T)0;
                }
                // Z_k = A_k + i B_k; Z_{n-k} = conj(A_k) + i conj(B_k)
                A[l * ldl + (N > 0 ? padi(k) : k)] = {a.x - b.y, a.y + b.x};
                if (k > 0 && 2 * k < nx) {
                    const int kn = nx - k;
                    A[l * ldl + (N > 0 ? padi(kn) : kn)] = {a.x + b.y, b.x - a.y};
                }
            }
        }
    }
    __syncthreads();
    const double inv = 1.0 / ((double)nx * (double)ny);
    const double lin = add_linear ? voltage * (1.0 - (double)z / (double)(nz - 1)) : 0.0;
    EHD_GLOBAL T* row0 = phi + (size_t)z * nx * ny + (size_t)y0 * nx;
    if constexpr (N > 0) {
        auto ld = [&](int l, int i) -> cplx<T> { return A[l * ldl + padi(i)]; };
        auto sf = [&](int l, int i, cplx<T> v) {
            row0[(size_t)(2 * l) * N + i] = (T)((double)v.x * inv + lin);
            row0[(size_t)(2 * l + 1) * N + i] = (T)((double)v.y * inv + lin);
        };
        ct_run<T, N, 0, 1, true, true>(ld, sf, A, lpb, ldl, tws, +1);
    } else {
        cplx<T>* B = A + lpb * ldl;
        cplx<T>* res = lds_fft<T>(A, B, nx, lpb, ldl, pl, tws, +1);
        for (int w = threadIdx.x; w < lpb * nx; w += blockDim.x) {
            const int l = w / nx, i = w - l * nx;
            const cplx<T> v = res[l * ldl + i];
            row0[(size_t)(2 * l) * nx + i] = (T)((double)v.x * inv + lin);
            row0[(size_t)(2 * l + 1) * nx + i] = (T)((double)v.y * inv + lin);
        }
    }
}

// Dirichlet values of the potential on the walls (odd_extract, poisson.cu:158-165).
__global__ void __launch_bounds__(256) phi_walls_kernel(const PoissonJob* __restrict__ jobs, int nxy)
{
    const PoissonJob& J = jobs[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nxy) return;
    EHD_GLOBAL double* phi = gptr((double*)J.phi);
    phi[i] = J.voltage;
    phi[(size_t)(J.n + 1) * nxy + i] = 0.0;
}

constexpr size_t kMaxLds = 65536;

// Largest line count per block for an x pass: divides ny/2, one butterfly per
// thread per stage (compile-time plan), LDS within kMaxLds.
template <typename T, int N>
int x_lines_per_block(int nx, int ny)
{
    const size_t cs = sizeof(cplx<T>);
    const int ldl = lds_pitch(N, nx);
    const int bufs = N > 0 ? 1 : 2;
    const int cap = N > 0 ? 256 / ct_max_nb(N) : 16;
    int best = 0;
    for (int l = 1; l <= ny / 2 && l <= cap; ++l) {
        if ((ny / 2) % l) continue;
        if ((size_t)nx * cs + (size_t)bufs * l * ldl * cs <= kMaxLds) best = l;
    }
    return best;
}

template <typename T, int N>
int y_lines_per_block(int ny)
{
    const size_t cs = sizeof(cplx<T>);
    const int ldl = lds_pitch(N, ny);
    const int bufs = N > 0 ? 1 : 2;
    int g = N > 0 ? 256 / ct_max_nb(N) : 16;
    while (g > 1 && (size_t)ny * cs + (size_t)bufs * g * ldl * cs > kMaxLds) --g;
    return ((size_t)ny * cs + (size_t)bufs * g * ldl * cs <= kMaxLds) ? g : 0;
}

// z-solve partition: P (16, 8 or 32) threads per mode with chunks of at most
// ZS_MAXM; fixed by n so that a box gives the same numbers alone and in a
// batch.
// LDS of a z-solve block: six doubles per (chunk, mode), P chunks x 256/P modes
size_t zsolve_lds(int P) { return (size_t)6 * P * (256 / P) * sizeof(double); }

int zsolve_partition(int n, size_t cs)
{
    (void)cs;
    // 16 threads per mode first (measured fastest at n = 126 on the 256x256
    // FP32 spectrum: 21.5 us with scans against 28 us for 32; in order 25 us
    // against 42 us for 8 and 51 us for 32)
    for (int cand : {16, 8, 32})
        if ((n + cand - 1) / cand <= ZS_MAXM && zsolve_lds(cand) <= kMaxLds) return cand;
    return 0;
}

// One Poisson solve for every job of a group (same nx, ny, precision; the
// jobs' plane counts may differ): each kernel runs once over the group, the
// job index in blockIdx.y.
template <typename T, int NX, int NY>
hipError_t poisson_group(const PoissonGroup& g, hipStream_t st)
{
    const int nx = g.nx, ny = g.ny, nxh = nx / 2 + 1;
    const int nmax = g.max_n;
    const FftPlan plx = make_fft_plan(nx), ply = make_fft_plan(ny);
    const size_t cs = sizeof(cplx<T>);
    const int lpb = g.lpb > 0 ? g.lpb : x_lines_per_block<T, NX>(nx, ny);
    const int G = g.ylines > 0 ? g.ylines : y_lines_per_block<T, NY>(ny);
    if (lpb == 0 || G == 0 || (ny / 2) % lpb != 0 || lpb > x_lines_per_block<T, NX>(nx, ny) ||
        G > y_lines_per_block<T, NY>(ny))
        return hipErrorInvalidValue;
    const size_t shx = (size_t)nx * cs + (size_t)(NX > 0 ? 1 : 2) * lpb * lds_pitch(NX, nx) * cs;
    const size_t shy = (size_t)ny * cs + (size_t)(NY > 0 ? 1 : 2) * G * lds_pitch(NY, ny) * cs;
    const dim3 gx(nmax * (ny / (2 * lpb)), g.njobs);
    const dim3 gy((nmax * nxh + G - 1) / G, g.njobs);
    const cplx<T>* twx = (const cplx<T>*)g.twx;
    const cplx<T>* twy = (const cplx<T>*)g.twy;

    // debug_sync: each pass synchronised and checked on its own
    auto check = [&](const char* pass) -> hipError_t {
        if (!g.debug_sync) return hipSuccess;
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) g_failed_pass = pass;
        return e;
    };
    hipError_t ec;
    fftx_fwd_kernel<T, NX><<<gx, 256, shx, st>>>(g.jobs, twx, plx, nx, ny, lpb);
    if ((ec = check("forward x transform")) != hipSuccess) return ec;
    ffty_kernel<T, NY><<<gy, 256, shy, st>>>(g.jobs, twy, ply, ny, nxh, G, -1);
    if ((ec = check("forward y transform")) != hipSuccess) return ec;

    const int nmodes = nxh * ny;
    const int Pz = zsolve_partition(nmax, cs);
    if (Pz == 0 || Pz != g.zs_partition) return hipErrorInvalidValue;
    const int mb = 256 / Pz;
    const size_t shz = zsolve_lds(Pz);
    const dim3 gz((nmodes + mb - 1) / mb, g.njobs);
    const double inv_nxy = 1.0 / ((double)nx * (double)ny);
    const int sm = g.split_mean;
    // chunk registers sized to the longest chunk: MAXM = 16, 8 or 4
    const int chunk = (nmax + Pz - 1) / Pz;
    hipError_t ez = hipErrorInvalidValue;
    if (Pz == 8) {
        zsolve_kernel<T, 8, 16><<<gz, 256, shz, st>>>(g.jobs, nmodes, ny, sm, inv_nxy);
        ez = hipSuccess;
    } else if (Pz == 16) {
        if (chunk <= 8) zsolve_kernel<T, 16, 8><<<gz, 256, shz, st>>>(g.jobs, nmodes, ny, sm, inv_nxy);
        else zsolve_kernel<T, 16, 16><<<gz, 256, shz, st>>>(g.jobs, nmodes, ny, sm, inv_nxy);
        ez = hipSuccess;
    } else if (Pz == 32) {
        if (chunk <= 4) zsolve_kernel<T, 32, 4><<<gz, 256, shz, st>>>(g.jobs, nmodes, ny, sm, inv_nxy);
        else if (chunk <= 8) zsolve_kernel<T, 32, 8><<<gz, 256, shz, st>>>(g.jobs, nmodes, ny, sm, inv_nxy);
        else zsolve_kernel<T, 32, 16><<<gz, 256, shz, st>>>(g.jobs, nmodes, ny, sm, inv_nxy);
        ez = hipSuccess;
    }
    if (ez != hipSuccess) return ez;
    if ((ec = check("z solve")) != hipSuccess) return ec;

    ffty_kernel<T, NY><<<gy, 256, shy, st>>>(g.jobs, twy, ply, ny, nxh, G, +1);
    if ((ec = check("inverse y transform")) != hipSuccess) return ec;
    fftx_inv_kernel<T, NX><<<gx, 256, shx, st>>>(g.jobs, twx, plx, nx, ny, lpb, sm ? 0 : 1);
    if ((ec = check("inverse x transform")) != hipSuccess) return ec;
    if (!sm) {
        phi_walls_kernel<<<dim3((nx * ny + 255) / 256, g.njobs), 256, 0, st>>>(g.jobs, nx * ny);
        if ((ec = check("wall potential")) != hipSuccess) return ec;
    }
    return hipGetLastError();
}

// Compile-time FFT sizes for the grids of the configurations; anything else
// takes the runtime-planned path.
template <typename T, int NX>
hipError_t poisson_ny(const PoissonGroup& g, hipStream_t st)
{
    static_assert(ct_complete(NX), "compile-time sizes must factor into powers of two, 3 and 5");
    if (g.ny == NX) return poisson_group<T, NX, NX>(g, st);
    return poisson_group<T, NX, 0>(g, st);
}

template <typename T>
hipError_t poisson_t(const PoissonGroup& g, hipStream_t st)
{
    switch (g.nx) {
    case 48: return poisson_ny<T, 48>(g, st);
    case 64: return poisson_ny<T, 64>(g, st);
    case 80: return poisson_ny<T, 80>(g, st);
    case 96: return poisson_ny<T, 96>(g, st);
    case 128: return poisson_ny<T, 128>(g, st);
    case 256: return poisson_ny<T, 256>(g, st);
    case 512: return poisson_ny<T, 512>(g, st);
    default: return poisson_group<T, 0, 0>(g, st);
    }
}

template <typename T>
int max_x_lines(int nx, int ny)
{
    switch (nx) {
    case 48: return x_lines_per_block<T, 48>(nx, ny);
    case 64: return x_lines_per_block<T, 64>(nx, ny);
    case 80: return x_lines_per_block<T, 80>(nx, ny);
    case 96: return x_lines_per_block<T, 96>(nx, ny);
    case 128: return x_lines_per_block<T, 128>(nx, ny);
    case 256: return x_lines_per_block<T, 256>(nx, ny);
    case 512: return x_lines_per_block<T, 512>(nx, ny);
    default: return x_lines_per_block<T, 0>(nx, ny);
    }
}

template <typename T>
int max_y_lines(int nx, int ny)
{
    if (ny != nx) return y_lines_per_block<T, 0>(ny);
    switch (ny) {
    case 48: return y_lines_per_block<T, 48>(ny);
    case 64: return y_lines_per_block<T, 64>(ny);
    case 80: return y_lines_per_block<T, 80>(ny);
    case 96: return y_lines_per_block<T, 96>(ny);
    case 128: return y_lines_per_block<T, 128>(ny);
    case 256: return y_lines_per_block<T, 256>(ny);
    case 512: return y_lines_per_block<T, 512>(ny);
    default: return y_lines_per_block<T, 0>(ny);
    }
}

}  // namespace

PoissonJob make_poisson_job(const LbmParams& P, const PoissonBufs& pb, const void* q, void* phi, double voltage,
                            double eps)
{
    PoissonJob j{};
    j.q = q;
    j.spec = pb.spec;
    j.phi = phi;
    j.phibar = pb.phibar;
    j.n = pb.nz - 2;
    j.kxs = 2.0 * 3.14159265358979323846 / (pb.nx * P.dx);
    j.kys = 2.0 * 3.14159265358979323846 / (pb.ny * P.dy);
    j.dz2 = P.dz * P.dz;
    j.eps = eps;
    j.voltage = voltage;
    return j;
}

PoissonShape poisson_shape(Prec prec, int nx, int ny, int nz)
{
    PoissonShape s{};
    const bool f64 = prec == Prec::F64;
    s.max_x_lines = f64 ? max_x_lines<double>(nx, ny) : max_x_lines<float>(nx, ny);
    s.max_y_lines = f64 ? max_y_lines<double>(nx, ny) : max_y_lines<float>(nx, ny);
    s.zs_partition = zsolve_partition(nz - 2, f64 ? 16 : 8);
    return s;
}

hipError_t launch_poisson_group(Prec prec, const PoissonGroup& g, hipStream_t st)
{
    if (prec == Prec::F64) return poisson_t<double>(g, st);
    return poisson_t<float>(g, st);
}

const char* poisson_failed_pass() { return g_failed_pass; }

}  // namespace ehd
