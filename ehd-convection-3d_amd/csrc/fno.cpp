// fno.cpp -- the Fourier-neural-operator surrogate on the GPU (include/ehd_fno.h).
//
// Every contraction is one call of the strided bf16 matrix-core GEMM
// (fno_kernels.hip): the truncated DFTs are dense matrices applied along one
// axis (z: real -> complex, y and x: complex -> complex), a complex product is
// one launch over separate real and imaginary planes (each plane read once,
// each result plane written once), and transposes or conjugates are stride
// and sign choices.  The backward pass applies the
// adjoint of each stage in reverse order (the conjugate-transposed DFT
// matrices, the conjugated mixing weights) and forms the weight gradients as
// GEMMs summed over the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/ehd.h"
#include "../../include/ehd_fno.h"
#include "box.h"
#include "fno_internal.h"

using namespace ehd;

namespace {

struct Geo {
    int M = 0, N = 0, K = 0;
    long long am = 0, ak = 0, bk = 0, bn = 0, cm = 0, cn = 0;
    int nb1 = 1, nb2 = 1;
    long long a1 = 0, a2 = 0, b1 = 0, b2 = 0, c1 = 0, c2 = 0;
    int ksplit = 0;
};

struct Cbuf {
    float* r = nullptr;
    float* i = nullptr;
};

}  // namespace

struct ehd_fno {
    ehd_fno_config c{};
    int KX = 0, KY = 0, KZ = 0, NM = 0;
    long long P = 0, ZP = 0;       // Y*X and Z*Y*X
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;
    // parameters, gradients, Adam moments (flat)
    long long nparam = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    int adam_t = 0;
    // offsets of the tensors in the flat vector
    long long oPw = 0, oPb = 0, oQ1w = 0, oQ1b = 0, oQ2w = 0, oQ2b = 0;
    std::vector<long long> oW, ob, oRr, oRi;
    // DFT matrices
    Cbuf AzT, ByT, Cx, Cxi, Byi, Azi;
    // activations
    float* in = nullptr;
    std::vector<float*> v;        // v[0..L]
    std::vector<float*> u;        // pre-activations of layers 0..L-2
    float *z1 = nullptr, *h = nullptr, *out = nullptr, *tgt = nullptr, *dout = nullptr;
    std::vector<Cbuf> S;          // spectra per layer (kept for the weight gradient)
    Cbuf Vz, Vy, Ym, Wx, Wy;      // spectral temporaries (reused by the backward pass)
    float *dv0 = nullptr, *dv1 = nullptr, *dh = nullptr;
    double* part = nullptr;
    float* wscr = nullptr;        // split-K partial sums of the weight gradients
    float* red = nullptr;         // partial sums of sum_slices and channel_sum
    long long wscr_n = 0;
    int ksplit = 0, nsplit = 1;   // K chunk of the weight-gradient GEMMs over the grid points
    ehd_fno_work work{};          // GEMM work since the last reset
    int resident = 0;             // samples of the last inputs/targets copied to the device
    float* ds_in = nullptr;       // resident training set (ehd_fno_set_dataset)
    float* ds_tgt = nullptr;
    int ds_n = 0;
};

namespace {

int falloc(ehd_fno* f, float** p, long long n)
{
    void* q = nullptr;
    if (hipMalloc(&q, (size_t)n * sizeof(float)) != hipSuccess)
        return set_error(EHD_ENOMEM, "FNO allocation of %lld floats failed", n);
    f->allocs.push_back(q);
    f->alloc_bytes.push_back((size_t)n * sizeof(float));
    *p = (float*)q;
    return EHD_OK;
}

// Every GEMM operand is checked on the host before the launch: the largest
// element it touches (all strides are non-negative) must lie in the
// allocation that holds its base pointer.
bool in_range(const ehd_fno* f, const float* base, long long last)
{
    if (!base || last < 0) return false;
    const char* b = (const char*)base;
    for (size_t i = 0; i < f->allocs.size(); ++i) {
        const char* a = (const char*)f->allocs[i];
        if (b >= a && b < a + f->alloc_bytes[i])
            return b + (size_t)(last + 1) * sizeof(float) <= a + f->alloc_bytes[i];
    }
    return false;
}

// Elementwise tail of a GEMM's epilogue: + bias[m], then GELU (the value
// before it kept in pre) or the product with gelu'(dact); pre and dact at
// C's offsets.
struct Epi {
    const float* bias = nullptr;
    int act = kActNone;
    float* pre = nullptr;
    const float* dact = nullptr;
};

hipError_t rgemm(ehd_fno* f, const Geo& g, const float* A, const float* B, float* C, float alpha, float beta,
                 hipStream_t st, const Epi& ep = Epi{})
{
    const long long kmax = g.K - 1;
    const long long la = (g.nb1 - 1) * g.a1 + (g.nb2 - 1) * g.a2 + (g.M - 1) * g.am + kmax * g.ak;
    const long long lb = (g.nb1 - 1) * g.b1 + (g.nb2 - 1) * g.b2 + kmax * g.bk + (g.N - 1) * g.bn;
    const long long lc = (g.nb1 - 1) * g.c1 + (g.nb2 - 1) * g.c2 + (g.M - 1) * g.cm + (g.N - 1) * g.cn;
    const bool neg = g.am < 0 || g.ak < 0 || g.bk < 0 || g.bn < 0 || g.cm < 0 || g.cn < 0 || g.a1 < 0 ||
                     g.a2 < 0 || g.b1 < 0 || g.b2 < 0 || g.c1 < 0 || g.c2 < 0;
    if (neg || !in_range(f, A, la) || !in_range(f, B, lb) || !in_range(f, C, lc)) return hipErrorInvalidValue;
    if ((ep.bias && !in_range(f, ep.bias, g.M - 1)) || (ep.act == kActGelu && !in_range(f, ep.pre, lc)) ||
        (ep.act == kActGeluGrad && !in_range(f, ep.dact, lc)) || (ep.act != kActNone && g.ksplit > 0))
        return hipErrorInvalidValue;
    GemmArgs a{g.M, g.N, g.K, A, g.am, g.ak, B, g.bk, g.bn, C, g.cm, g.cn,
               g.nb1, g.nb2, g.a1, g.a2, g.b1, g.b2, g.c1, g.c2, alpha, beta, g.ksplit};
    a.bias = ep.bias;
    a.act = ep.act;
    a.pre = ep.pre;
    a.dact = ep.dact;
    // with split-K the nb2 slices share the K products of one batch entry
    const double nb = (double)g.nb1 * (g.ksplit > 0 ? 1 : g.nb2), nc = (double)g.nb1 * g.nb2;
    f->work.flops += 2.0 * g.M * (double)g.N * g.K * nb;
    f->work.bytes += 4.0 * (((double)g.M * g.K + (double)g.K * g.N) * nb +
                            (double)g.M * g.N * nc * ((beta != 0.f ? 2 : 1) + (ep.act != kActNone ? 1 : 0)));
    f->work.launches += 1;
    return gemm_bf16(a, st);
}

// One launch of the complex product (Ar + i sa Ai)(Br + i sb Bi) (sa, sb = -1
// conjugate): Cr = alpha Re(.) + beta Cr, Ci likewise with Im(.).  Ai or Bi
// null: that operand is real; Ci null: only the real part is written.
hipError_t xgemm(ehd_fno* f, const Geo& g, const float* Ar, const float* Ai, const float* Br, const float* Bi,
                 float* Cr, float* Ci, int sa, int sb, float alpha, float beta, hipStream_t st)
{
    const long long kmax = g.K - 1;
    const long long la = (g.nb1 - 1) * g.a1 + (g.nb2 - 1) * g.a2 + (g.M - 1) * g.am + kmax * g.ak;
    const long long lb = (g.nb1 - 1) * g.b1 + (g.nb2 - 1) * g.b2 + kmax * g.bk + (g.N - 1) * g.bn;
    const long long lc = (g.nb1 - 1) * g.c1 + (g.nb2 - 1) * g.c2 + (g.M - 1) * g.cm + (g.N - 1) * g.cn;
    const bool neg = g.am < 0 || g.ak < 0 || g.bk < 0 || g.bn < 0 || g.cm < 0 || g.cn < 0 || g.a1 < 0 ||
                     g.a2 < 0 || g.b1 < 0 || g.b2 < 0 || g.c1 < 0 || g.c2 < 0;
    if (neg || g.ksplit != 0 || !in_range(f, Ar, la) || (Ai && !in_range(f, Ai, la)) || !in_range(f, Br, lb) ||
        (Bi && !in_range(f, Bi, lb)) || !in_range(f, Cr, lc) || (Ci && !in_range(f, Ci, lc)))
        return hipErrorInvalidValue;
    CGemmArgs x{GemmArgs{g.M, g.N, g.K, Ar, g.am, g.ak, Br, g.bk, g.bn, Cr, g.cm, g.cn, g.nb1, g.nb2, g.a1, g.a2,
                         g.b1, g.b2, g.c1, g.c2, alpha, beta, 0},
                Ai, Bi, Ci, (float)sa, (float)sb};
    // real products: Ar Br, Ai Bi (both imaginary parts), Ar Bi and Ai Br (for Ci)
    const int nprod = 1 + (Ai && Bi ? 1 : 0) + (Ci ? (Ai ? 1 : 0) + (Bi ? 1 : 0) : 0);
    const double nb = (double)g.nb1 * g.nb2;
    f->work.flops += 2.0 * g.M * (double)g.N * g.K * nb * nprod;
    f->work.bytes += 4.0 * nb * ((double)g.M * g.K * (Ai ? 2 : 1) + (double)g.K * g.N * (Bi ? 2 : 1) +
                                 (double)g.M * g.N * (Ci ? 2 : 1) * (beta != 0.f ? 2 : 1));
    f->work.launches += 1;
    return cgemm_bf16(x, st);
}

// C = op(A) op(B) (+ beta C) for complex operands in separate planes; sa/sb =
// -1 conjugates A/B:  Cr = Ar Br - sa sb Ai Bi,  Ci = sb Ar Bi + sa Ai Br.
hipError_t cgemm(ehd_fno* f, const Geo& g, const float* Ar, const float* Ai, const float* Br, const float* Bi,
                 float* Cr, float* Ci, int sa, int sb, float beta, hipStream_t st)
{
    return xgemm(f, g, Ar, Ai, Br, Bi, Cr, Ci, sa, sb, 1.f, beta, st);
}

int calloc_c(ehd_fno* f, Cbuf* b, long long n)
{
    int rc = falloc(f, &b->r, n);
    return rc ? rc : falloc(f, &b->i, n);
}

void free_dataset(ehd_fno* f)
{
    if (f->st) (void)hipStreamSynchronize(f->st);
    if (f->ds_in) (void)hipFree(f->ds_in);
    if (f->ds_tgt) (void)hipFree(f->ds_tgt);
    f->ds_in = f->ds_tgt = nullptr;
    f->ds_n = 0;
}

void free_fno(ehd_fno* f)
{
    free_dataset(f);
    if (f->st) (void)hipStreamSynchronize(f->st);
    for (void* q : f->allocs) (void)hipFree(q);
    f->allocs.clear();
    f->alloc_bytes.clear();
    if (f->st) (void)hipStreamDestroy(f->st);
}

int upload(const std::vector<float>& h, float* d)
{
    HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    return EHD_OK;
}

std::vector<int> freqs(int n, int m)
{
    std::vector<int> f;
    for (int i = 0; i < m; ++i) f.push_back(i);
    for (int i = n - m; i < n; ++i) f.push_back(i);
    return f;
}

int make_dft(ehd_fno* f)
{
    const int X = f->c.X, Y = f->c.Y, Z = f->c.Z;
    const std::vector<int> fx = freqs(X, f->c.mx), fy = freqs(Y, f->c.my);
    const double tp = 2.0 * 3.14159265358979323846;
    std::vector<float> r, im;
    auto put = [&](Cbuf& b) -> int {
        int rc = upload(r, b.r);
        return rc ? rc : upload(im, b.i);
    };
    int rc;
    // AzT[k][z] = exp(-2 pi i k z / Z)
    r.assign((size_t)f->KZ * Z, 0.f); im = r;
    for (int k = 0; k < f->KZ; ++k)
        for (int z = 0; z < Z; ++z) {
            const double a = -tp * k * z / Z;
            r[(size_t)k * Z + z] = (float)std::cos(a); im[(size_t)k * Z + z] = (float)std::sin(a);
        }
    if ((rc = put(f->AzT))) return rc;
    // Azi[k][z] = c_k/Z exp(+2 pi i k z / Z), c_0 = 1, else 2 (real part of the sum taken)
    for (int k = 0; k < f->KZ; ++k)
        for (int z = 0; z < Z; ++z) {
            const double a = tp * k * z / Z, s = (k == 0 ? 1.0 : 2.0) / Z;
            r[(size_t)k * Z + z] = (float)(s * std::cos(a)); im[(size_t)k * Z + z] = (float)(s * std::sin(a));
        }
    if ((rc = put(f->Azi))) return rc;
    // ByT[j][y] = exp(-2 pi i fy_j y / Y); Byi[j][y] = exp(+...)/Y
    r.assign((size_t)f->KY * Y, 0.f); im = r;
    for (int j = 0; j < f->KY; ++j)
        for (int y = 0; y < Y; ++y) {
            const double a = -tp * fy[j] * y / Y;
            r[(size_t)j * Y + y] = (float)std::cos(a); im[(size_t)j * Y + y] = (float)std::sin(a);
        }
    if ((rc = put(f->ByT))) return rc;
    for (int j = 0; j < f->KY; ++j)
        for (int y = 0; y < Y; ++y) {
            const double a = tp * fy[j] * y / Y;
            r[(size_t)j * Y + y] = (float)(std::cos(a) / Y); im[(size_t)j * Y + y] = (float)(std::sin(a) / Y);
        }
    if ((rc = put(f->Byi))) return rc;
    // Cx[x][i] = exp(-2 pi i fx_i x / X); Cxi[i][x] = exp(+...)/X
    r.assign((size_t)X * f->KX, 0.f); im = r;
    for (int x = 0; x < X; ++x)
        for (int i = 0; i < f->KX; ++i) {
            const double a = -tp * fx[i] * x / X;
            r[(size_t)x * f->KX + i] = (float)std::cos(a); im[(size_t)x * f->KX + i] = (float)std::sin(a);
        }
    if ((rc = put(f->Cx))) return rc;
    for (int i = 0; i < f->KX; ++i)
        for (int x = 0; x < X; ++x) {
            const double a = tp * fx[i] * x / X;
            r[(size_t)i * X + x] = (float)(std::cos(a) / X); im[(size_t)i * X + x] = (float)(std::sin(a) / X);
        }
    return put(f->Cxi);
}

#define FCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return set_error(EHD_EHIP, "FNO %s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// ---- stages --------------------------------------------------------------

// C[b][o][s] (+)= sum_c Wt[c][o] X[b][c][s]  (pointwise layer, weight [cin][cout]),
// then the epilogue's bias and activation
int pointwise(ehd_fno* f, int B, int cin, int cout, const float* Wt, const float* X, float* C, float beta,
              const Epi& ep)
{
    Geo g;
    g.M = cout; g.N = (int)f->ZP; g.K = cin;
    g.am = 1; g.ak = cout;
    g.bk = f->ZP; g.bn = 1; g.b1 = (long long)cin * f->ZP;
    g.cm = f->ZP; g.cn = 1; g.c1 = (long long)cout * f->ZP;
    g.nb1 = B;
    FCHK(rgemm(f, g, Wt, X, C, 1.f, beta, f->st, ep));
    return EHD_OK;
}

// dX[b][c][s] (+)= sum_o Wt[c][o] dC[b][o][s], then (u given) times gelu'(u)
int pointwise_dx(ehd_fno* f, int B, int cin, int cout, const float* Wt, const float* dC, float* dX, float beta,
                 const float* u = nullptr)
{
    Geo g;
    g.M = cin; g.N = (int)f->ZP; g.K = cout;
    g.am = cout; g.ak = 1;
    g.bk = f->ZP; g.bn = 1; g.b1 = (long long)cout * f->ZP;
    g.cm = f->ZP; g.cn = 1; g.c1 = (long long)cin * f->ZP;
    g.nb1 = B;
    Epi ep;
    if (u) {
        ep.act = kActGeluGrad;
        ep.dact = u;
    }
    FCHK(rgemm(f, g, Wt, dC, dX, 1.f, beta, f->st, ep));
    return EHD_OK;
}

// dW[c][o] = sum_b sum_s X[b][c][s] dC[b][o][s]: split-K over the grid points
// (batch index (b, chunk)), partial products summed by sum_slices
int pointwise_dw(ehd_fno* f, int B, int cin, int cout, const float* X, const float* dC, float* dW)
{
    const long long mn = (long long)cin * cout;
    if ((long long)B * f->nsplit * mn > f->wscr_n) return set_error(EHD_EINVAL, "weight-gradient scratch too small");
    Geo g;
    g.M = cin; g.N = cout; g.K = (int)f->ZP;
    g.am = f->ZP; g.ak = 1; g.a1 = (long long)cin * f->ZP;
    g.bk = 1; g.bn = f->ZP; g.b1 = (long long)cout * f->ZP;
    g.cm = cout; g.cn = 1; g.c1 = f->nsplit * mn; g.c2 = mn;
    g.nb1 = B; g.nb2 = f->nsplit;
    g.ksplit = f->ksplit;
    FCHK(rgemm(f, g, X, dC, f->wscr, 1.f, 0.f, f->st));
    FCHK(sum_slices(f->wscr, dW, mn, B * f->nsplit, 0.f, f->red, f->st));
    return EHD_OK;
}

// out[b][o] (+)= K_l(v)  (the spectral convolution of layer l); keeps the spectrum in S[l]
int spectral_fwd(ehd_fno* f, int l, int B, const float* v, float* out, float beta)
{
    const int w = f->c.width, X = f->c.X, Y = f->c.Y, Z = f->c.Z;
    const int KX = f->KX, KY = f->KY, KZ = f->KZ, NM = f->NM;
    const long long P = f->P, BC = (long long)B * w;
    const float* Rr = f->p + f->oRr[l];
    const float* Ri = f->p + f->oRi[l];
    Cbuf& S = f->S[l];
    {   // 1. Vz = AzT v (real input)
        Geo g;
        g.M = KZ; g.N = (int)P; g.K = Z;
        g.am = Z; g.ak = 1;
        g.bk = P; g.bn = 1; g.b1 = Z * P;
        g.cm = P; g.cn = 1; g.c1 = KZ * P;
        g.nb1 = (int)BC;
        FCHK(xgemm(f, g, f->AzT.r, f->AzT.i, v, nullptr, f->Vz.r, f->Vz.i, 1, 1, 1.f, 0.f, f->st));
    }
    {   // 2. Vy = ByT Vz along y, per (bc, kz)
        Geo g;
        g.M = KY; g.N = X; g.K = Y;
        g.am = Y; g.ak = 1;
        g.bk = X; g.bn = 1; g.b1 = (long long)Y * X;
        g.cm = X; g.cn = 1; g.c1 = (long long)KY * X;
        g.nb1 = (int)(BC * KZ);
        FCHK(cgemm(f, g, f->ByT.r, f->ByT.i, f->Vz.r, f->Vz.i, f->Vy.r, f->Vy.i, 1, 1, 0.f, f->st));
    }
    {   // 3. S = Vy Cx along x
        Geo g;
        g.M = (int)(BC * KZ * KY); g.N = KX; g.K = X;
        g.am = X; g.ak = 1;
        g.bk = KX; g.bn = 1;
        g.cm = KX; g.cn = 1;
        FCHK(cgemm(f, g, f->Vy.r, f->Vy.i, f->Cx.r, f->Cx.i, S.r, S.i, 1, 1, 0.f, f->st));
    }
    {   // 4. Ym[b][o][md] = sum_c S[b][c][md] R[md][c][o]
        Geo g;
        g.M = B; g.N = w; g.K = w;
        g.am = (long long)w * NM; g.ak = NM; g.a1 = 1;
        g.bk = w; g.bn = 1; g.b1 = (long long)w * w;
        g.cm = (long long)w * NM; g.cn = NM; g.c1 = 1;
        g.nb1 = NM;
        FCHK(cgemm(f, g, S.r, S.i, Rr, Ri, f->Ym.r, f->Ym.i, 1, 1, 0.f, f->st));
    }
    {   // 5. Wx = Ym Cxi along x
        Geo g;
        g.M = (int)(BC * KZ * KY); g.N = X; g.K = KX;
        g.am = KX; g.ak = 1;
        g.bk = X; g.bn = 1;
        g.cm = X; g.cn = 1;
        FCHK(cgemm(f, g, f->Ym.r, f->Ym.i, f->Cxi.r, f->Cxi.i, f->Wx.r, f->Wx.i, 1, 1, 0.f, f->st));
    }
    {   // 6. Wy = Byi^T Wx along y, per (bo, kz)
        Geo g;
        g.M = Y; g.N = X; g.K = KY;
        g.am = 1; g.ak = Y;
        g.bk = X; g.bn = 1; g.b1 = (long long)KY * X;
        g.cm = X; g.cn = 1; g.c1 = (long long)Y * X;
        g.nb1 = (int)(BC * KZ);
        FCHK(cgemm(f, g, f->Byi.r, f->Byi.i, f->Wx.r, f->Wx.i, f->Wy.r, f->Wy.i, 1, 1, 0.f, f->st));
    }
    {   // 7. out = Re(Azi^T Wy) along z
        Geo g;
        g.M = Z; g.N = (int)P; g.K = KZ;
        g.am = 1; g.ak = Z;
        g.bk = P; g.bn = 1; g.b1 = KZ * P;
        g.cm = P; g.cn = 1; g.c1 = Z * P;
        g.nb1 = (int)BC;
        FCHK(xgemm(f, g, f->Azi.r, f->Azi.i, f->Wy.r, f->Wy.i, out, nullptr, 1, 1, 1.f, beta, f->st));
    }
    return EHD_OK;
}

// dv (+)= K_l^T(du); dRr, dRi = gradient of the mixing weights
int spectral_bwd(ehd_fno* f, int l, int B, const float* du, float* dv, float beta)
{
    const int w = f->c.width, X = f->c.X, Y = f->c.Y, Z = f->c.Z;
    const int KX = f->KX, KY = f->KY, KZ = f->KZ, NM = f->NM;
    const long long P = f->P, BC = (long long)B * w;
    const float* Rr = f->p + f->oRr[l];
    const float* Ri = f->p + f->oRi[l];
    float* dRr = f->g + f->oRr[l];
    float* dRi = f->g + f->oRi[l];
    Cbuf& S = f->S[l];
    {   // 7b. G_Wy = conj(Azi) du  (real du)
        Geo g;
        g.M = KZ; g.N = (int)P; g.K = Z;
        g.am = Z; g.ak = 1;
        g.bk = P; g.bn = 1; g.b1 = Z * P;
        g.cm = P; g.cn = 1; g.c1 = KZ * P;
        g.nb1 = (int)BC;
        FCHK(xgemm(f, g, f->Azi.r, f->Azi.i, du, nullptr, f->Wy.r, f->Wy.i, -1, 1, 1.f, 0.f, f->st));
    }
    {   // 6b. G_Wx = conj(Byi) G_Wy along y
        Geo g;
        g.M = KY; g.N = X; g.K = Y;
        g.am = Y; g.ak = 1;
        g.bk = X; g.bn = 1; g.b1 = (long long)Y * X;
        g.cm = X; g.cn = 1; g.c1 = (long long)KY * X;
        g.nb1 = (int)(BC * KZ);
        FCHK(cgemm(f, g, f->Byi.r, f->Byi.i, f->Wy.r, f->Wy.i, f->Wx.r, f->Wx.i, -1, 1, 0.f, f->st));
    }
    {   // 5b. G_Ym = G_Wx Cxi^H along x
        Geo g;
        g.M = (int)(BC * KZ * KY); g.N = KX; g.K = X;
        g.am = X; g.ak = 1;
        g.bk = 1; g.bn = X;
        g.cm = KX; g.cn = 1;
        FCHK(cgemm(f, g, f->Wx.r, f->Wx.i, f->Cxi.r, f->Cxi.i, f->Ym.r, f->Ym.i, 1, -1, 0.f, f->st));
    }
    {   // 4b. G_R[md][c][o] = sum_b conj(S[b][c][md]) G_Ym[b][o][md]
        Geo g;
        g.M = w; g.N = w; g.K = B;
        g.am = NM; g.ak = (long long)w * NM; g.a1 = 1;
        g.bk = (long long)w * NM; g.bn = NM; g.b1 = 1;
        g.cm = w; g.cn = 1; g.c1 = (long long)w * w;
        g.nb1 = NM;
        FCHK(cgemm(f, g, S.r, S.i, f->Ym.r, f->Ym.i, dRr, dRi, -1, 1, 0.f, f->st));
    }
    {   // 4b'. G_S[b][c][md] = sum_o G_Ym[b][o][md] conj(R[md][c][o])  (into S: the spectrum is no longer needed)
        Geo g;
        g.M = B; g.N = w; g.K = w;
        g.am = (long long)w * NM; g.ak = NM; g.a1 = 1;
        g.bk = 1; g.bn = w; g.b1 = (long long)w * w;
        g.cm = (long long)w * NM; g.cn = NM; g.c1 = 1;
        g.nb1 = NM;
        FCHK(cgemm(f, g, f->Ym.r, f->Ym.i, Rr, Ri, S.r, S.i, 1, -1, 0.f, f->st));
    }
    {   // 3b. G_Vy = G_S Cx^H along x
        Geo g;
        g.M = (int)(BC * KZ * KY); g.N = X; g.K = KX;
        g.am = KX; g.ak = 1;
        g.bk = 1; g.bn = KX;
        g.cm = X; g.cn = 1;
        FCHK(cgemm(f, g, S.r, S.i, f->Cx.r, f->Cx.i, f->Vy.r, f->Vy.i, 1, -1, 0.f, f->st));
    }
    {   // 2b. G_Vz = ByT^H G_Vy along y
        Geo g;
        g.M = Y; g.N = X; g.K = KY;
        g.am = 1; g.ak = Y;
        g.bk = X; g.bn = 1; g.b1 = (long long)KY * X;
        g.cm = X; g.cn = 1; g.c1 = (long long)Y * X;
        g.nb1 = (int)(BC * KZ);
        FCHK(cgemm(f, g, f->ByT.r, f->ByT.i, f->Vy.r, f->Vy.i, f->Vz.r, f->Vz.i, -1, 1, 0.f, f->st));
    }
    {   // 1b. dv (+)= Re(AzT^H G_Vz) = AzT_r^T G_r + AzT_i^T G_i
        Geo g;
        g.M = Z; g.N = (int)P; g.K = KZ;
        g.am = 1; g.ak = Z;
        g.bk = P; g.bn = 1; g.b1 = KZ * P;
        g.cm = P; g.cn = 1; g.c1 = Z * P;
        g.nb1 = (int)BC;
        FCHK(xgemm(f, g, f->AzT.r, f->AzT.i, f->Vz.r, f->Vz.i, dv, nullptr, -1, 1, 1.f, beta, f->st));
    }
    return EHD_OK;
}

int forward(ehd_fno* f, int B)
{
    const int w = f->c.width, L = f->c.layers;
    int rc;
    // each layer's bias and GELU run in the epilogue of the GEMM that completes its sum
    Epi ep;
    ep.bias = f->p + f->oPb;
    if ((rc = pointwise(f, B, f->c.d_in, w, f->p + f->oPw, f->in, f->v[0], 0.f, ep))) return rc;
    for (int l = 0; l < L; ++l) {
        if ((rc = spectral_fwd(f, l, B, f->v[l], f->v[l + 1], 0.f))) return rc;
        Epi el;
        el.bias = f->p + f->ob[l];
        if (l < L - 1) {
            el.act = kActGelu;
            el.pre = f->u[l];
        }
        if ((rc = pointwise(f, B, w, w, f->p + f->oW[l], f->v[l], f->v[l + 1], 1.f, el))) return rc;
    }
    Epi eq1;
    eq1.bias = f->p + f->oQ1b;
    eq1.act = kActGelu;
    eq1.pre = f->z1;
    if ((rc = pointwise(f, B, w, f->c.d_hidden, f->p + f->oQ1w, f->v[L], f->h, 0.f, eq1))) return rc;
    Epi eq2;
    eq2.bias = f->p + f->oQ2b;
    if ((rc = pointwise(f, B, f->c.d_hidden, f->c.d_out, f->p + f->oQ2w, f->h, f->out, 0.f, eq2))) return rc;
    return EHD_OK;
}

// gradients of the loss whose output gradient is in f->dout
int backward(ehd_fno* f, int B)
{
    const int w = f->c.width, L = f->c.layers, H = f->c.d_hidden, O = f->c.d_out;
    const long long ZP = f->ZP;
    int rc;
    if ((rc = pointwise_dw(f, B, H, O, f->h, f->dout, f->g + f->oQ2w))) return rc;
    FCHK(channel_sum(f->dout, f->g + f->oQ2b, B, O, ZP, false, f->red, f->st));
    if ((rc = pointwise_dx(f, B, H, O, f->p + f->oQ2w, f->dout, f->dh, 0.f, f->z1))) return rc;
    if ((rc = pointwise_dw(f, B, w, H, f->v[L], f->dh, f->g + f->oQ1w))) return rc;
    FCHK(channel_sum(f->dh, f->g + f->oQ1b, B, H, ZP, false, f->red, f->st));
    float* dv = f->dv0;
    float* dn = f->dv1;
    if ((rc = pointwise_dx(f, B, w, H, f->p + f->oQ1w, f->dh, dv, 0.f))) return rc;
    // dv enters iteration l as the gradient of layer l's pre-activation: the
    // input gradient of layer l + 1 leaves its GEMM times gelu'(u[l])
    for (int l = L - 1; l >= 0; --l) {
        if ((rc = pointwise_dw(f, B, w, w, f->v[l], dv, f->g + f->oW[l]))) return rc;
        FCHK(channel_sum(dv, f->g + f->ob[l], B, w, ZP, false, f->red, f->st));
        if ((rc = spectral_bwd(f, l, B, dv, dn, 0.f))) return rc;
        if ((rc = pointwise_dx(f, B, w, w, f->p + f->oW[l], dv, dn, 1.f, l >= 1 ? f->u[l - 1] : nullptr))) return rc;
        float* t = dv;
        dv = dn;
        dn = t;
    }
    if ((rc = pointwise_dw(f, B, f->c.d_in, w, f->in, dv, f->g + f->oPw))) return rc;
    FCHK(channel_sum(dv, f->g + f->oPb, B, w, ZP, false, f->red, f->st));
    return EHD_OK;
}

int loss_and_grad(ehd_fno* f, int B, const float* in, const float* target, double* loss)
{
    const long long nin = (long long)B * f->c.d_in * f->ZP, nout = (long long)B * f->c.d_out * f->ZP;
    if (in || target) {
        if (!in || !target) return set_error(EHD_EINVAL, "inputs and targets go together (or both NULL)");
        HIPCHK(hipMemcpyAsync(f->in, in, nin * sizeof(float), hipMemcpyHostToDevice, f->st));
        HIPCHK(hipMemcpyAsync(f->tgt, target, nout * sizeof(float), hipMemcpyHostToDevice, f->st));
        f->resident = B;
    } else if (B > f->resident) {
        return set_error(EHD_EINVAL, "no resident batch of %d samples (last loaded: %d)", B, f->resident);
    }
    int rc = forward(f, B);
    if (rc) return rc;
    FCHK(mse(f->out, f->tgt, f->dout, f->part, nout, f->st));
    if ((rc = backward(f, B))) return rc;
    std::vector<double> part(mse_blocks());
    HIPCHK(hipMemcpyAsync(part.data(), f->part, part.size() * sizeof(double), hipMemcpyDeviceToHost, f->st));
    HIPCHK(hipStreamSynchronize(f->st));
    double s = 0.0;
    for (double x : part) s += x;
    if (loss) *loss = s / (double)nout;
    return EHD_OK;
}

int check_batch(const ehd_fno* f, int B)
{
    if (B < 1 || B > f->c.max_batch) return set_error(EHD_EINVAL, "batch %d outside 1..%d", B, f->c.max_batch);
    return EHD_OK;
}

}  // namespace

extern "C" {

int ehd_fno_create(const ehd_fno_config* cfg, ehd_fno** out)
{
    if (!cfg || !out) return set_error(EHD_EINVAL, "null argument");
    *out = nullptr;
    const ehd_fno_config& c = *cfg;
    if (c.X < 2 || c.Y < 2 || c.Z < 2 || (c.X & 1) || (c.Y & 1) || c.width < 1 || c.layers < 1 || c.d_in < 1 ||
        c.d_hidden < 1 || c.d_out < 1 || c.max_batch < 1)
        return set_error(EHD_EINVAL, "bad FNO configuration");
    if (c.mx < 1 || c.my < 1 || c.mz < 1 || c.mx > c.X / 2 || c.my > c.Y / 2 || c.mz > c.Z / 2)
        return set_error(EHD_EINVAL, "modes must lie in 1..X/2, 1..Y/2, 1..Z/2");
    ehd_fno* f = new (std::nothrow) ehd_fno();
    if (!f) return set_error(EHD_ENOMEM, "host allocation failed");
    auto bail = [&](int code) {
        free_fno(f);
        delete f;
        return code;
    };
    f->c = c;
    f->KX = 2 * c.mx;
    f->KY = 2 * c.my;
    f->KZ = c.mz;
    f->NM = f->KX * f->KY * f->KZ;
    f->P = (long long)c.Y * c.X;
    f->ZP = (long long)c.Z * f->P;
    const long long w = c.width, B = c.max_batch, NMw2 = (long long)f->NM * w * w;
    // flat parameter layout
    long long o = 0;
    f->oPw = o; o += (long long)c.d_in * w;
    f->oPb = o; o += w;
    for (int l = 0; l < c.layers; ++l) {
        f->oW.push_back(o); o += w * w;
        f->ob.push_back(o); o += w;
        f->oRr.push_back(o); o += NMw2;
        f->oRi.push_back(o); o += NMw2;
    }
    f->oQ1w = o; o += w * c.d_hidden;
    f->oQ1b = o; o += c.d_hidden;
    f->oQ2w = o; o += (long long)c.d_hidden * c.d_out;
    f->oQ2b = o; o += c.d_out;
    f->nparam = o;
    if (hipStreamCreateWithFlags(&f->st, hipStreamNonBlocking) != hipSuccess)
        return bail(set_error(EHD_EHIP, "hipStreamCreate failed"));
    int rc;
    const long long act = B * w * f->ZP;
    const long long sp_z = B * w * f->KZ * f->P, sp_y = B * w * f->KZ * f->KY * c.X, sp_m = B * w * f->NM;
    if ((rc = falloc(f, &f->p, o)) || (rc = falloc(f, &f->g, o)) || (rc = falloc(f, &f->m1, o)) ||
        (rc = falloc(f, &f->m2, o)) || (rc = falloc(f, &f->in, B * c.d_in * f->ZP)) ||
        (rc = falloc(f, &f->z1, B * c.d_hidden * f->ZP)) || (rc = falloc(f, &f->h, B * c.d_hidden * f->ZP)) ||
        (rc = falloc(f, &f->dh, B * c.d_hidden * f->ZP)) || (rc = falloc(f, &f->out, B * c.d_out * f->ZP)) ||
        (rc = falloc(f, &f->tgt, B * c.d_out * f->ZP)) || (rc = falloc(f, &f->dout, B * c.d_out * f->ZP)) ||
        (rc = falloc(f, &f->dv0, act)) || (rc = falloc(f, &f->dv1, act)))
        return bail(rc);
    f->v.assign(c.layers + 1, nullptr);
    for (auto& p : f->v)
        if ((rc = falloc(f, &p, act))) return bail(rc);
    f->u.assign(c.layers > 1 ? c.layers - 1 : 0, nullptr);
    for (auto& p : f->u)
        if ((rc = falloc(f, &p, act))) return bail(rc);
    f->S.assign(c.layers, Cbuf{});
    for (auto& s : f->S)
        if ((rc = calloc_c(f, &s, sp_m))) return bail(rc);
    if ((rc = calloc_c(f, &f->Vz, sp_z)) || (rc = calloc_c(f, &f->Wy, sp_z)) || (rc = calloc_c(f, &f->Vy, sp_y)) ||
        (rc = calloc_c(f, &f->Wx, sp_y)) || (rc = calloc_c(f, &f->Ym, sp_m)) ||
        (rc = calloc_c(f, &f->AzT, (long long)f->KZ * c.Z)) || (rc = calloc_c(f, &f->Azi, (long long)f->KZ * c.Z)) ||
        (rc = calloc_c(f, &f->ByT, (long long)f->KY * c.Y)) || (rc = calloc_c(f, &f->Byi, (long long)f->KY * c.Y)) ||
        (rc = calloc_c(f, &f->Cx, (long long)c.X * f->KX)) || (rc = calloc_c(f, &f->Cxi, (long long)f->KX * c.X)))
        return bail(rc);
    f->ksplit = 2048;
    f->nsplit = (int)((f->ZP + f->ksplit - 1) / f->ksplit);
    {
        const long long wmax = std::max<long long>({(long long)c.d_in * w, w * w, w * (long long)c.d_hidden,
                                                    (long long)c.d_hidden * c.d_out});
        f->wscr_n = B * f->nsplit * wmax;
        if ((rc = falloc(f, &f->wscr, f->wscr_n))) return bail(rc);
        const int cmax = std::max({c.width, c.d_hidden, c.d_out});
        const long long nred = std::max(sum_slices_part_floats(wmax), channel_sum_part_floats((int)B, cmax, f->ZP));
        if ((rc = falloc(f, &f->red, nred))) return bail(rc);
    }
    {
        void* q = nullptr;
        if (hipMalloc(&q, sizeof(double) * mse_blocks()) != hipSuccess)
            return bail(set_error(EHD_ENOMEM, "FNO allocation failed"));
        f->allocs.push_back(q);
        f->alloc_bytes.push_back(sizeof(double) * mse_blocks());
        f->part = (double*)q;
    }
    if ((rc = make_dft(f))) return bail(rc);
    if (hipMemset(f->p, 0, o * sizeof(float)) != hipSuccess || hipMemset(f->m1, 0, o * sizeof(float)) != hipSuccess ||
        hipMemset(f->m2, 0, o * sizeof(float)) != hipSuccess)
        return bail(set_error(EHD_EHIP, "FNO clear failed"));
    *out = f;
    return EHD_OK;
}

void ehd_fno_destroy(ehd_fno* f)
{
    if (!f) return;
    free_fno(f);
    delete f;
}

long long ehd_fno_num_params(const ehd_fno* f) { return f ? f->nparam : 0; }

int ehd_fno_set_params(ehd_fno* f, const float* params)
{
    if (!f || !params) return set_error(EHD_EINVAL, "null argument");
    HIPCHK(hipMemcpy(f->p, params, f->nparam * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(f->m1, 0, f->nparam * sizeof(float)));
    HIPCHK(hipMemset(f->m2, 0, f->nparam * sizeof(float)));
    f->adam_t = 0;
    return EHD_OK;
}

int ehd_fno_get_params(ehd_fno* f, float* params)
{
    if (!f || !params) return set_error(EHD_EINVAL, "null argument");
    HIPCHK(hipMemcpy(params, f->p, f->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return EHD_OK;
}

int ehd_fno_forward(ehd_fno* f, int batch, const float* in, float* out)
{
    if (!f || !in || !out) return set_error(EHD_EINVAL, "null argument");
    int rc = check_batch(f, batch);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(f->in, in, (size_t)batch * f->c.d_in * f->ZP * sizeof(float), hipMemcpyHostToDevice, f->st));
    f->resident = 0;   // the inputs no longer match the resident targets
    if ((rc = forward(f, batch))) return rc;
    HIPCHK(hipMemcpyAsync(out, f->out, (size_t)batch * f->c.d_out * f->ZP * sizeof(float), hipMemcpyDeviceToHost, f->st));
    HIPCHK(hipStreamSynchronize(f->st));
    return EHD_OK;
}

int ehd_fno_gradients(ehd_fno* f, int batch, const float* in, const float* target, double* loss, float* grad)
{
    if (!f) return set_error(EHD_EINVAL, "null argument");
    int rc = check_batch(f, batch);
    if (rc) return rc;
    if ((rc = loss_and_grad(f, batch, in, target, loss))) return rc;
    if (grad) HIPCHK(hipMemcpy(grad, f->g, f->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return EHD_OK;
}

int ehd_fno_train_step(ehd_fno* f, int batch, const float* in, const float* target, float lr, double* loss)
{
    if (!f) return set_error(EHD_EINVAL, "null argument");
    int rc = check_batch(f, batch);
    if (rc) return rc;
    if ((rc = loss_and_grad(f, batch, in, target, loss))) return rc;
    f->adam_t += 1;
    FCHK(adam(f->p, f->g, f->m1, f->m2, f->nparam, lr, 0.9f, 0.999f, 1e-8f, f->adam_t, f->st));
    HIPCHK(hipStreamSynchronize(f->st));
    return EHD_OK;
}

int ehd_fno_set_dataset(ehd_fno* f, int n, const float* in, const float* target)
{
    if (!f || n < 1 || !in || !target) return set_error(EHD_EINVAL, "bad dataset arguments");
    free_dataset(f);
    const size_t nin = (size_t)n * f->c.d_in * f->ZP, nout = (size_t)n * f->c.d_out * f->ZP;
    if (hipMalloc((void**)&f->ds_in, nin * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&f->ds_tgt, nout * sizeof(float)) != hipSuccess) {
        free_dataset(f);
        return set_error(EHD_ENOMEM, "dataset of %d samples (%.1f GB) does not fit", n,
                         (double)(nin + nout) * 4e-9);
    }
    HIPCHK(hipMemcpyAsync(f->ds_in, in, nin * sizeof(float), hipMemcpyHostToDevice, f->st));
    HIPCHK(hipMemcpyAsync(f->ds_tgt, target, nout * sizeof(float), hipMemcpyHostToDevice, f->st));
    HIPCHK(hipStreamSynchronize(f->st));
    f->ds_n = n;
    return EHD_OK;
}

int ehd_fno_train_step_indices(ehd_fno* f, int batch, const int* idx, float lr, double* loss)
{
    if (!f || !idx) return set_error(EHD_EINVAL, "null argument");
    int rc = check_batch(f, batch);
    if (rc) return rc;
    if (f->ds_n == 0) return set_error(EHD_ESTATE, "no resident dataset (ehd_fno_set_dataset)");
    const size_t sin = (size_t)f->c.d_in * f->ZP, sout = (size_t)f->c.d_out * f->ZP;
    for (int i = 0; i < batch; ++i) {
        if (idx[i] < 0 || idx[i] >= f->ds_n)
            return set_error(EHD_EINVAL, "sample index %d outside 0..%d", idx[i], f->ds_n - 1);
    }
    // gather the batch on the device; the step then runs on it as on a resident batch
    for (int i = 0; i < batch; ++i) {
        HIPCHK(hipMemcpyAsync(f->in + i * sin, f->ds_in + (size_t)idx[i] * sin, sin * sizeof(float),
                              hipMemcpyDeviceToDevice, f->st));
        HIPCHK(hipMemcpyAsync(f->tgt + i * sout, f->ds_tgt + (size_t)idx[i] * sout, sout * sizeof(float),
                              hipMemcpyDeviceToDevice, f->st));
    }
    f->resident = batch;
    return ehd_fno_train_step(f, batch, nullptr, nullptr, lr, loss);
}

int ehd_fno_gemm_work(ehd_fno* f, int reset, ehd_fno_work* out)
{
    if (!f || !out) return set_error(EHD_EINVAL, "null argument");
    *out = f->work;
    if (reset) f->work = ehd_fno_work{};
    return EHD_OK;
}

int ehd_fno_gemm_strided(int M, int N, int K, int nb, int ksplit, const float* A, const long long as[3],
                         long long a_len, const float* B, const long long bs[3], long long b_len, float* C,
                         const long long cs[4], long long c_len, float alpha, float beta)
{
    if (M < 1 || N < 1 || K < 1 || nb < 1 || ksplit < 0 || !A || !B || !C || !as || !bs || !cs)
        return set_error(EHD_EINVAL, "bad GEMM arguments");
    for (int i = 0; i < 3; ++i)
        if (as[i] < 0 || bs[i] < 0 || cs[i] < 0 || cs[3] < 0) return set_error(EHD_EINVAL, "negative stride");
    const int ns = ksplit > 0 ? (K + ksplit - 1) / ksplit : 1;
    const long long la = (nb - 1) * as[2] + (M - 1) * as[0] + (long long)(K - 1) * as[1];
    const long long lb = (nb - 1) * bs[2] + (long long)(K - 1) * bs[0] + (N - 1) * bs[1];
    const long long lc = (nb - 1) * cs[2] + (ns - 1) * cs[3] + (M - 1) * cs[0] + (N - 1) * cs[1];
    if (la >= a_len || lb >= b_len || lc >= c_len) return set_error(EHD_EINVAL, "strides reach outside the arrays");
    float *dA = nullptr, *dB = nullptr, *dC = nullptr;
    auto cleanup = [&]() {
        if (dA) (void)hipFree(dA);
        if (dB) (void)hipFree(dB);
        if (dC) (void)hipFree(dC);
    };
    if (hipMalloc(&dA, (size_t)a_len * 4) != hipSuccess || hipMalloc(&dB, (size_t)b_len * 4) != hipSuccess ||
        hipMalloc(&dC, (size_t)c_len * 4) != hipSuccess) {
        cleanup();
        return set_error(EHD_ENOMEM, "GEMM allocation failed");
    }
    hipError_t e = hipMemcpy(dA, A, (size_t)a_len * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, B, (size_t)b_len * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dC, C, (size_t)c_len * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        GemmArgs g{M, N, K, dA, as[0], as[1], dB, bs[0], bs[1], dC, cs[0], cs[1], nb, ns,
                   as[2], 0, bs[2], 0, cs[2], cs[3], alpha, beta, ksplit};
        e = gemm_bf16(g, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(C, dC, (size_t)c_len * 4, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return set_error(EHD_EHIP, "GEMM failed: %s", hipGetErrorString(e));
    return EHD_OK;
}

int ehd_fno_cgemm_strided(int M, int N, int K, int nb, const float* Ar, const float* Ai, const long long as[3],
                          long long a_len, const float* Br, const float* Bi, const long long bs[3], long long b_len,
                          float* Cr, float* Ci, const long long cs[3], long long c_len, int sa, int sb, float alpha,
                          float beta)
{
    if (M < 1 || N < 1 || K < 1 || nb < 1 || !Ar || !Br || !Cr || !as || !bs || !cs || (sa != 1 && sa != -1) ||
        (sb != 1 && sb != -1))
        return set_error(EHD_EINVAL, "bad complex GEMM arguments");
    for (int i = 0; i < 3; ++i)
        if (as[i] < 0 || bs[i] < 0 || cs[i] < 0) return set_error(EHD_EINVAL, "negative stride");
    const long long la = (nb - 1) * as[2] + (M - 1) * as[0] + (long long)(K - 1) * as[1];
    const long long lb = (nb - 1) * bs[2] + (long long)(K - 1) * bs[0] + (N - 1) * bs[1];
    const long long lc = (nb - 1) * cs[2] + (M - 1) * cs[0] + (N - 1) * cs[1];
    if (la >= a_len || lb >= b_len || lc >= c_len) return set_error(EHD_EINVAL, "strides reach outside the arrays");
    // device copies of the planes given: [Ar, Ai, Br, Bi, Cr, Ci]
    const float* src[6] = {Ar, Ai, Br, Bi, Cr, Ci};
    const long long len[6] = {a_len, a_len, b_len, b_len, c_len, c_len};
    float* d[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    auto cleanup = [&]() {
        for (float* q : d)
            if (q) (void)hipFree(q);
    };
    hipError_t e = hipSuccess;
    for (int i = 0; i < 6 && e == hipSuccess; ++i) {
        if (!src[i]) continue;
        if (hipMalloc(&d[i], (size_t)len[i] * 4) != hipSuccess) {
            cleanup();
            return set_error(EHD_ENOMEM, "complex GEMM allocation failed");
        }
        e = hipMemcpy(d[i], src[i], (size_t)len[i] * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
        CGemmArgs x{GemmArgs{M, N, K, d[0], as[0], as[1], d[2], bs[0], bs[1], d[4], cs[0], cs[1], nb, 1, as[2], 0,
                             bs[2], 0, cs[2], 0, alpha, beta, 0},
                    d[1], d[3], d[5], (float)sa, (float)sb};
        e = cgemm_bf16(x, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(Cr, d[4], (size_t)c_len * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && Ci) e = hipMemcpy(Ci, d[5], (size_t)c_len * 4, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return set_error(EHD_EHIP, "complex GEMM failed: %s", hipGetErrorString(e));
    return EHD_OK;
}

int ehd_fno_gemm(int M, int N, int K, const float* A, const float* B, float* C)
{
    if (M < 1 || N < 1 || K < 1 || !A || !B || !C) return set_error(EHD_EINVAL, "bad GEMM arguments");
    float *dA = nullptr, *dB = nullptr, *dC = nullptr;
    auto cleanup = [&]() {
        if (dA) (void)hipFree(dA);
        if (dB) (void)hipFree(dB);
        if (dC) (void)hipFree(dC);
    };
    if (hipMalloc(&dA, (size_t)M * K * 4) != hipSuccess || hipMalloc(&dB, (size_t)K * N * 4) != hipSuccess ||
        hipMalloc(&dC, (size_t)M * N * 4) != hipSuccess) {
        cleanup();
        return set_error(EHD_ENOMEM, "GEMM allocation failed");
    }
    hipError_t e = hipMemcpy(dA, A, (size_t)M * K * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dB, B, (size_t)K * N * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        GemmArgs g{M, N, K, dA, K, 1, dB, N, 1, dC, N, 1, 1, 1, 0, 0, 0, 0, 0, 0, 1.f, 0.f, 0};
        e = gemm_bf16(g, nullptr);
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) return set_error(EHD_EHIP, "GEMM failed: %s", hipGetErrorString(e));
    return EHD_OK;
}

}  // extern "C"
