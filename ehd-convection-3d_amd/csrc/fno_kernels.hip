// fno_kernels.hip -- the surrogate's GPU kernels: a strided, batched GEMM on
// gfx950's bf16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulation)
// that carries every contraction of the Fourier neural operator (the
// truncated DFTs written as matrices, the per-mode channel mixing, the
// pointwise layers and all their adjoints), plus the elementwise kernels
// (loss, column sums, Adam); bias and GELU run in the GEMM epilogue.
//
// Operands are FP32 in memory and rounded to bf16 (nearest even) when staged
// into LDS; accumulation and the epilogue are FP32.  Any element strides are
// allowed, so transposes, real/imaginary planes and sub-blocks cost no copy;
// an operand whose k or m/n axis is unit-stride (and 16-byte aligned) is
// staged with 16-byte loads, anything else element by element.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fno_internal.h"

namespace ehd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return u >> 16;
}

__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) { return to_bf16(lo) | (to_bf16(hi) << 16); }

// GELU with the exact (erf) form, and its derivative
__device__ __forceinline__ float gelu(float v) { return 0.5f * v * (1.0f + erff(v * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x)
{
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.3989422804014327f;
}

// Tile 64 x 64 x 32; 256 threads = 4 waves, each owning a 32 x 32 block of C
// (2 x 2 MFMA blocks of 16 x 16).  Both operands sit in LDS as bf16 rows of
// k (A as [m][k], B as [n][k]), so every lane fetches its 8 k-values of an
// MFMA operand with one 16-byte LDS read; the row pitch of 40 elements (80
// bytes) keeps those reads 16-byte aligned and spreads the 16 rows of an
// operand over the banks.  Two LDS buffers: the global loads of k-step s+1
// are in flight while the matrix cores work on step s.
constexpr int TM = 64, TN = 64, TK = 32, LP = TK + 8;
constexpr int kSliceGroups = 32;     // first-pass groups of slices in sum_slices
constexpr int kRowChunk = 8192;      // row elements per workgroup in channel_sum

// operand staging layouts
constexpr int L_K = 0;       // k unit-stride: 16-byte loads along k
constexpr int L_MN = 1;      // m (A) or n (B) unit-stride: 16-byte loads along m / n
constexpr int L_ANY = 2;     // element loads

// result layouts: with a unit-stride axis the C tile goes through LDS and
// leaves as 16-byte stores along that axis (a wave writes four 256-byte
// rows); otherwise each lane stores its MFMA elements (64-byte pieces)
constexpr int C_ANY = 0;
constexpr int C_N = 1;       // n unit-stride
constexpr int C_M = 2;       // m unit-stride
constexpr int CP = TN + 4;   // C tile pitch in LDS (floats)
constexpr int kOperandLds = 2 * TM * LP * 2 + 2 * TN * LP * 2;
constexpr int kGemmLds = kOperandLds > TM * CP * 4 ? kOperandLds : TM * CP * 4;
constexpr int kCgemmLds = 2 * kOperandLds > TM * CP * 4 ? 2 * kOperandLds : TM * CP * 4;

// One operand tile (64 rows of m or n, 32 of k) in flight: 8 floats per thread.
struct Stage {
    float v[8];
};

// P[r][k] at P + r*sr + k*sk for r < R, k in [k0, ke): rows r0.. of the tile
template <int L>
__device__ __forceinline__ void load_tile(Stage& st, const float* __restrict__ P, long long sr, long long sk, int r0,
                                          int R, int k0, int ke, int tid)
{
    if constexpr (L == L_K) {
        // 64 rows x 8 quads of k; thread: quads tid and tid + 256
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = tid + 256 * h, r = q >> 3, k = k0 + 4 * (q & 7);
            float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r0 + r < R && k < ke) x = *(const float4*)(P + (long long)(r0 + r) * sr + k);
            st.v[4 * h] = x.x;
            st.v[4 * h + 1] = x.y;
            st.v[4 * h + 2] = x.z;
            st.v[4 * h + 3] = x.w;
        }
    } else if constexpr (L == L_MN) {
        // 16 quads of rows x 16 pairs of k; thread: one quad, k = 2p and 2p+1
        const int rq = 4 * (tid & 15), k = k0 + 2 * (tid >> 4);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
            if (r0 + rq < R && k + h < ke) x = *(const float4*)(P + (r0 + rq) + (long long)(k + h) * sk);
            st.v[4 * h] = x.x;
            st.v[4 * h + 1] = x.y;
            st.v[4 * h + 2] = x.z;
            st.v[4 * h + 3] = x.w;
        }
    } else {
        // element loads, the smaller stride fastest across the threads
        const bool kfast = sk <= sr;
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const int e = tid + 256 * h;
            const int r = kfast ? (e >> 5) : (e & 63), k = k0 + (kfast ? (e & 31) : (e >> 6));
            st.v[h] = (r0 + r < R && k < ke) ? P[(long long)(r0 + r) * sr + (long long)k * sk] : 0.f;
        }
    }
}

template <int L>
__device__ __forceinline__ void store_tile(const Stage& st, unsigned short (*T)[LP], int tid, long long sr,
                                           long long sk)
{
    if constexpr (L == L_K) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = tid + 256 * h, r = q >> 3, k = 4 * (q & 7);
            uint2 w;
            w.x = pack_bf16(st.v[4 * h], st.v[4 * h + 1]);
            w.y = pack_bf16(st.v[4 * h + 2], st.v[4 * h + 3]);
            *(uint2*)&T[r][k] = w;
        }
    } else if constexpr (L == L_MN) {
        const int rq = 4 * (tid & 15), k = 2 * (tid >> 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) *(uint32_t*)&T[rq + e][k] = pack_bf16(st.v[e], st.v[4 + e]);
    } else {
        const bool kfast = sk <= sr;
#pragma unroll
        for (int h = 0; h < 8; ++h) {
            const int e = tid + 256 * h;
            const int r = kfast ? (e >> 5) : (e & 63), k = kfast ? (e & 31) : (e >> 6);
            T[r][k] = (unsigned short)to_bf16(st.v[h]);
        }
    }
}

// The epilogue's elementwise tail on a result v at offset o of C (row m):
// + bias[m], then GELU (the value before it kept in pre) or the product with
// gelu'(dact).
__device__ __forceinline__ float finish(const GemmArgs& g, float v, int m, long long o)
{
    if (g.bias) v = v + g.bias[m];
    if (g.act == kActGelu) {
        g.pre[o] = v;
        v = gelu(v);
    } else if (g.act == kActGeluGrad) {
        v = v * gelu_grad(g.dact[o]);
    }
    return v;
}

// The same on a quad of C at offset o (16-byte aligned, as pre and dact
// there): bias bm, bm + dm, bm + 2 dm, bm + 3 dm.
__device__ __forceinline__ float4 finish4(const GemmArgs& g, float4 v, int m, int dm, long long o)
{
    if (g.bias) {
        v.x = v.x + g.bias[m];
        v.y = v.y + g.bias[m + dm];
        v.z = v.z + g.bias[m + 2 * dm];
        v.w = v.w + g.bias[m + 3 * dm];
    }
    if (g.act == kActGelu) {
        *(float4*)(g.pre + o) = v;
        v = make_float4(gelu(v.x), gelu(v.y), gelu(v.z), gelu(v.w));
    } else if (g.act == kActGeluGrad) {
        const float4 d = *(const float4*)(g.dact + o);
        v.x = v.x * gelu_grad(d.x);
        v.y = v.y * gelu_grad(d.y);
        v.z = v.z * gelu_grad(d.z);
        v.w = v.w * gelu_grad(d.w);
    }
    return v;
}

// Writes one 64 x 64 tile of C = alpha * acc (+ beta * C) at batch offset oc;
// acc in the MFMA result layout (D[i][j] of a 16x16 block: lane j +
// 16*(i/4), element i%4).  The vector layouts stage the tile in lds (TM * CP
// floats, which the caller has retired from every other use) and leave with
// 16-byte stores.  TAIL: then the bias and activation of finish().
template <int CL, bool TAIL>
__device__ __forceinline__ void store_result(const GemmArgs& g, float* __restrict__ C, long long oc,
                                             const f32x4 (&acc)[2][2], unsigned char* lds, int m0, int n0, int wm,
                                             int wn, int tid)
{
    const int lane = tid & 63;
    const int lj = lane & 15, li4 = (lane >> 4) * 4;
    if constexpr (CL == C_ANY) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int gm = m0 + wm + a * 16 + li4 + e, gn = n0 + wn + b * 16 + lj;
                    if (gm < g.M && gn < g.N) {
                        const long long o = oc + (long long)gm * g.cm + (long long)gn * g.cn;
                        float v = g.alpha * acc[a][b][e];
                        if (g.beta != 0.f) v = v + g.beta * C[o];
                        if constexpr (TAIL) v = finish(g, v, gm, o);
                        C[o] = v;
                    }
                }
    } else {
        float(*Cs)[CP] = reinterpret_cast<float(*)[CP]>(lds);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int i = wm + a * 16 + li4 + e, j = wn + b * 16 + lj;
                    if constexpr (CL == C_N) Cs[i][j] = acc[a][b][e];
                    else Cs[j][i] = acc[a][b][e];
                }
        __syncthreads();
        // 64 rows (m for C_N, n for C_M) of 16 quads along the unit-stride axis
        const int r0 = CL == C_N ? m0 : n0, c0 = CL == C_N ? n0 : m0;
        const int R = CL == C_N ? g.M : g.N, Q = CL == C_N ? g.N : g.M;
        const long long sr = CL == C_N ? g.cm : g.cn;
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int q = tid + 256 * h, r = q >> 4, c4 = 4 * (q & 15);
            if (r0 + r < R && c0 + c4 < Q) {   // Q is a multiple of 4: the quad lies inside
                const long long o = oc + (long long)(r0 + r) * sr + (c0 + c4);
                float4 v = *(const float4*)&Cs[r][c4];
                v.x = g.alpha * v.x;
                v.y = g.alpha * v.y;
                v.z = g.alpha * v.z;
                v.w = g.alpha * v.w;
                if (g.beta != 0.f) {
                    const float4 w = *(const float4*)(C + o);
                    v.x = v.x + g.beta * w.x;
                    v.y = v.y + g.beta * w.y;
                    v.z = v.z + g.beta * w.z;
                    v.w = v.w + g.beta * w.w;
                }
                // row m (C_N) or four consecutive m (C_M)
                if constexpr (TAIL) v = finish4(g, v, CL == C_N ? r0 + r : c0 + c4, CL == C_N ? 0 : 1, o);
                *(float4*)(C + o) = v;
            }
        }
    }
}

template <int AL, int BL, int CL>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g)
{
    // operand buffers during the k-steps, then the C tile
    __shared__ __align__(16) unsigned char lds[kGemmLds];
    unsigned short(*As)[TM][LP] = reinterpret_cast<unsigned short(*)[TM][LP]>(lds);
    unsigned short(*Bs)[TN][LP] = reinterpret_cast<unsigned short(*)[TN][LP]>(lds + 2 * TM * LP * 2);
    const int bz = blockIdx.z;
    const int i1 = bz / g.nb2, i2 = bz - i1 * g.nb2;
    const float* __restrict__ A = g.A + i1 * g.a1 + i2 * g.a2;
    const float* __restrict__ B = g.B + i1 * g.b1 + i2 * g.b2;
    const long long oc = i1 * g.c1 + i2 * g.c2;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int kb = g.ksplit > 0 ? i2 * g.ksplit : 0;
    const int ke = g.ksplit > 0 ? min(g.K, kb + g.ksplit) : g.K;
    Stage sa, sb;
    load_tile<AL>(sa, A, g.am, g.ak, m0, g.M, kb, ke, tid);
    load_tile<BL>(sb, B, g.bn, g.bk, n0, g.N, kb, ke, tid);
    store_tile<AL>(sa, As[0], tid, g.am, g.ak);
    store_tile<BL>(sb, Bs[0], tid, g.bn, g.bk);
    __syncthreads();
    // MFMA 16x16x32 bf16 operands: A[i][k] in lane i + 16*(k/8), element k%8;
    // B[k][j] in lane j + 16*(k/8), element k%8
    const int li = lane & 15, lk = (lane >> 4) * 8;
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += TK) {
        const bool more = k0 + TK < ke;
        if (more) {
            load_tile<AL>(sa, A, g.am, g.ak, m0, g.M, k0 + TK, ke, tid);
            load_tile<BL>(sb, B, g.bn, g.bk, n0, g.N, k0 + TK, ke, tid);
        }
        bf16x8 af[2], bfr[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) __builtin_memcpy(&af[a], &As[buf][wm + a * 16 + li][lk], 16);
#pragma unroll
        for (int b = 0; b < 2; ++b) __builtin_memcpy(&bfr[b], &Bs[buf][wn + b * 16 + li][lk], 16);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        if (more) {
            store_tile<AL>(sa, As[buf ^ 1], tid, g.am, g.ak);
            store_tile<BL>(sb, Bs[buf ^ 1], tid, g.bn, g.bk);
        }
        __syncthreads();
        buf ^= 1;
    }
    // the last k-step's barrier has retired every LDS read of the operands
    if (g.bias || g.act != kActNone) store_result<CL, true>(g, g.C, oc, acc, lds, m0, n0, wm, wn, tid);
    else store_result<CL, false>(g, g.C, oc, acc, lds, m0, n0, wm, wn, tid);
}

__device__ __forceinline__ void negate(Stage& s)
{
#pragma unroll
    for (int i = 0; i < 8; ++i) s.v[i] = -s.v[i];
}

// The complex product (Ar + i sa Ai)(Br + i sb Bi) on the same tiles, real
// and imaginary parts in separate planes with shared strides: an absent
// imaginary operand is zero, an absent Ci is not written.  The signs are
// applied while staging (exact in bf16); three accumulator sets hold
// Ar Br, (sa Ai)(sb Bi) and Ar (sb Bi) + (sa Ai) Br, so each output plane is
// written once and each operand plane read once.
template <int AL, int BL, int CL>
__global__ void __launch_bounds__(256) cgemm_bf16_kernel(CGemmArgs x)
{
    const GemmArgs& g = x.g;
    __shared__ __align__(16) unsigned char lds[kCgemmLds];
    // [buffer][part: real, imaginary][row][k]
    unsigned short(*As)[2][TM][LP] = reinterpret_cast<unsigned short(*)[2][TM][LP]>(lds);
    unsigned short(*Bs)[2][TN][LP] = reinterpret_cast<unsigned short(*)[2][TN][LP]>(lds + 2 * 2 * TM * LP * 2);
    const bool hai = x.Ai != nullptr, hbi = x.Bi != nullptr, hci = x.Ci != nullptr;
    const bool nega = x.sa < 0.f, negb = x.sb < 0.f;
    const int bz = blockIdx.z;
    const int i1 = bz / g.nb2, i2 = bz - i1 * g.nb2;
    const long long oa = i1 * g.a1 + i2 * g.a2, ob = i1 * g.b1 + i2 * g.b2, oc = i1 * g.c1 + i2 * g.c2;
    const float* __restrict__ Ar = g.A + oa;
    const float* __restrict__ Ai = hai ? x.Ai + oa : nullptr;
    const float* __restrict__ Br = g.B + ob;
    const float* __restrict__ Bi = hbi ? x.Bi + ob : nullptr;
    const int m0 = blockIdx.x * TM, n0 = blockIdx.y * TN;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x4 rr[2][2], ii[2][2], ri[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            rr[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            ii[a][b] = rr[a][b];
            ri[a][b] = rr[a][b];
        }
    Stage sar, sai, sbr, sbi;
    auto load = [&](int k0) {
        load_tile<AL>(sar, Ar, g.am, g.ak, m0, g.M, k0, g.K, tid);
        if (hai) load_tile<AL>(sai, Ai, g.am, g.ak, m0, g.M, k0, g.K, tid);
        load_tile<BL>(sbr, Br, g.bn, g.bk, n0, g.N, k0, g.K, tid);
        if (hbi) load_tile<BL>(sbi, Bi, g.bn, g.bk, n0, g.N, k0, g.K, tid);
    };
    auto stage = [&](int buf) {
        store_tile<AL>(sar, As[buf][0], tid, g.am, g.ak);
        if (hai) {
            if (nega) negate(sai);
            store_tile<AL>(sai, As[buf][1], tid, g.am, g.ak);
        }
        store_tile<BL>(sbr, Bs[buf][0], tid, g.bn, g.bk);
        if (hbi) {
            if (negb) negate(sbi);
            store_tile<BL>(sbi, Bs[buf][1], tid, g.bn, g.bk);
        }
    };
    load(0);
    stage(0);
    __syncthreads();
    const int li = lane & 15, lk = (lane >> 4) * 8;
    int buf = 0;
    for (int k0 = 0; k0 < g.K; k0 += TK) {
        const bool more = k0 + TK < g.K;
        if (more) load(k0 + TK);
        bf16x8 ar[2], br[2], ai[2] = {}, bi[2] = {};
#pragma unroll
        for (int a = 0; a < 2; ++a) __builtin_memcpy(&ar[a], &As[buf][0][wm + a * 16 + li][lk], 16);
#pragma unroll
        for (int b = 0; b < 2; ++b) __builtin_memcpy(&br[b], &Bs[buf][0][wn + b * 16 + li][lk], 16);
        if (hai)
#pragma unroll
            for (int a = 0; a < 2; ++a) __builtin_memcpy(&ai[a], &As[buf][1][wm + a * 16 + li][lk], 16);
        if (hbi)
#pragma unroll
            for (int b = 0; b < 2; ++b) __builtin_memcpy(&bi[b], &Bs[buf][1][wn + b * 16 + li][lk], 16);
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                rr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[a], br[b], rr[a][b], 0, 0, 0);
                if (hai && hbi) ii[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[a], bi[b], ii[a][b], 0, 0, 0);
                if (hci && hbi) ri[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ar[a], bi[b], ri[a][b], 0, 0, 0);
                if (hci && hai) ri[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ai[a], br[b], ri[a][b], 0, 0, 0);
            }
        if (more) stage(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) rr[a][b] -= ii[a][b];
    store_result<CL, false>(g, g.C, oc, rr, lds, m0, n0, wm, wn, tid);
    if (hci) {
        __syncthreads();   // every lane is done with the staged real tile
        store_result<CL, false>(g, x.Ci, oc, ri, lds, m0, n0, wm, wn, tid);
    }
}

// Sums of many slices and of channel rows, deterministic (fixed order, no
// atomics) and spread over enough workgroups to fill the chip: a first pass
// writes partial sums per group of slices / chunk of a row, a second adds
// the partials in a fixed order.

// 256 threads = 64 consecutive elements x 4 slice lanes; one workgroup per
// (64 elements, group of slices): part[g][i] = sum of src[j*n + i] over the
// slices j of group g
__global__ void __launch_bounds__(256) slice_group_sum_kernel(const float* __restrict__ src, float* __restrict__ part,
                                                              long long n, int nslices)
{
    __shared__ float red[4][64];
    const int e = threadIdx.x & 63, ls = threadIdx.x >> 6, g = blockIdx.y;
    const long long i = (long long)blockIdx.x * 64 + e;
    const int per = (nslices + kSliceGroups - 1) / kSliceGroups;
    const int j0 = g * per, j1 = min(nslices, j0 + per);
    float s = 0.f;
    if (i < n)
        for (int j = j0 + ls; j < j1; j += 4) s += src[(long long)j * n + i];
    red[ls][e] = s;
    __syncthreads();
    if (ls == 0 && i < n) part[(long long)g * n + i] = (red[0][e] + red[1][e]) + (red[2][e] + red[3][e]);
}

__global__ void __launch_bounds__(256) group_total_kernel(const float* __restrict__ part, float* __restrict__ dst,
                                                          long long n, float beta)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int g = 0; g < kSliceGroups; ++g) s += part[(long long)g * n + i];
    dst[i] = (beta != 0.f) ? beta * dst[i] + s : s;
}

__device__ __forceinline__ float block_sum_256(float s, float* red)
{
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// part[row][chunk] = sum of g[row][chunk*kRowChunk ...) for the rows (b, c)
template <bool VEC>
__global__ void __launch_bounds__(256) row_chunk_sum_kernel(const float* __restrict__ g, float* __restrict__ part,
                                                            long long S, int nchunk)
{
    __shared__ float red[256];
    const long long row = blockIdx.y;
    const float* __restrict__ p = g + row * S;
    const long long s0 = (long long)blockIdx.x * kRowChunk, s1 = min(S, s0 + kRowChunk);
    float s = 0.f;
    if constexpr (VEC) {
        for (long long i = s0 + 4 * threadIdx.x; i < s1; i += 1024) {
            const float4 v = *(const float4*)(p + i);
            s += (v.x + v.y) + (v.z + v.w);
        }
    } else {
        for (long long i = s0 + threadIdx.x; i < s1; i += 256) s += p[i];
    }
    const float t = block_sum_256(s, red);
    if (threadIdx.x == 0) part[row * nchunk + blockIdx.x] = t;
}

// db[c] (+)= sum over b and chunks of part[(b C + c) nchunk + chunk]: one workgroup per channel
__global__ void __launch_bounds__(256) channel_total_kernel(const float* __restrict__ part, float* __restrict__ db,
                                                            int B, int C, int nchunk, int accumulate)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int t = threadIdx.x; t < B * nchunk; t += 256) {
        const int b = t / nchunk, k = t - b * nchunk;
        s += part[((long long)b * C + c) * nchunk + k];
    }
    const float tot = block_sum_256(s, red);
    if (threadIdx.x == 0) db[c] = accumulate ? db[c] + tot : tot;
}

// d = out - target; grad = 2 d / n; partial sums of d^2 per workgroup
__global__ void __launch_bounds__(256) mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt,
                                                  float* __restrict__ grad, double* __restrict__ part, long long n)
{
    __shared__ double red[256];
    double s = 0.0;
    const float sc = (float)(2.0 / (double)n);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = out[i] - tgt[i];
        grad[i] = sc * d;
        s += (double)d * (double)d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, long long n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

unsigned grid_of(long long n)
{
    long long g = (n + 255) / 256;
    return (unsigned)(g < 65536 ? (g < 1 ? 1 : g) : 65536);
}

bool aligned16(const float* p) { return ((uintptr_t)p & 15u) == 0; }
bool mul4(long long v) { return (v & 3) == 0; }

}  // namespace

// Staging layout of A: 16-byte loads need the unit-stride axis in quads that
// lie wholly inside or outside the matrix (extent a multiple of 4) and every
// other offset, batch strides included, a multiple of 4 floats from a 16-byte
// aligned base.
int gemm_layout_a(const GemmArgs& g)
{
    const bool base = aligned16(g.A) && mul4(g.a1) && mul4(g.a2);
    if (base && g.ak == 1 && mul4(g.am) && mul4(g.K) && mul4(g.ksplit)) return L_K;
    if (base && g.am == 1 && mul4(g.ak) && mul4(g.M)) return L_MN;
    return L_ANY;
}

int gemm_layout_b(const GemmArgs& g)
{
    const bool base = aligned16(g.B) && mul4(g.b1) && mul4(g.b2);
    if (base && g.bk == 1 && mul4(g.bn) && mul4(g.K) && mul4(g.ksplit)) return L_K;
    if (base && g.bn == 1 && mul4(g.bk) && mul4(g.N)) return L_MN;
    return L_ANY;
}

// Layout of C: 16-byte stores need the unit-stride axis in whole quads
// (extent a multiple of 4) and every other offset a multiple of 4 floats
// from a 16-byte aligned base (and so do the epilogue's pre and dact, which
// share C's offsets).
int gemm_layout_c(const GemmArgs& g)
{
    const bool base = aligned16(g.C) && mul4(g.c1) && mul4(g.c2) && (g.act != kActGelu || aligned16(g.pre)) &&
                      (g.act != kActGeluGrad || aligned16(g.dact));
    if (base && g.cn == 1 && mul4(g.cm) && mul4(g.N)) return C_N;
    if (base && g.cm == 1 && mul4(g.cn) && mul4(g.M)) return C_M;
    return C_ANY;
}

namespace {
template <int AL, int BL, int CL>
void launch(dim3 grid, const GemmArgs& g, hipStream_t st)
{
    gemm_bf16_kernel<AL, BL, CL><<<grid, 256, 0, st>>>(g);
}

template <int AL, int BL, int CL>
void launch(dim3 grid, const CGemmArgs& x, hipStream_t st)
{
    cgemm_bf16_kernel<AL, BL, CL><<<grid, 256, 0, st>>>(x);
}

template <int AL, int BL, class Args>
void launch_c(int cl, dim3 grid, const Args& a, hipStream_t st)
{
    if (cl == C_N) launch<AL, BL, C_N>(grid, a, st);
    else if (cl == C_M) launch<AL, BL, C_M>(grid, a, st);
    else launch<AL, BL, C_ANY>(grid, a, st);
}

template <class Args>
void dispatch(int al, int bl, int cl, dim3 grid, const Args& a, hipStream_t st)
{
    switch (al * 3 + bl) {
    case 0: launch_c<L_K, L_K>(cl, grid, a, st); break;
    case 1: launch_c<L_K, L_MN>(cl, grid, a, st); break;
    case 2: launch_c<L_K, L_ANY>(cl, grid, a, st); break;
    case 3: launch_c<L_MN, L_K>(cl, grid, a, st); break;
    case 4: launch_c<L_MN, L_MN>(cl, grid, a, st); break;
    case 5: launch_c<L_MN, L_ANY>(cl, grid, a, st); break;
    case 6: launch_c<L_ANY, L_K>(cl, grid, a, st); break;
    case 7: launch_c<L_ANY, L_MN>(cl, grid, a, st); break;
    default: launch_c<L_ANY, L_ANY>(cl, grid, a, st); break;
    }
}
}  // namespace

hipError_t gemm_bf16(const GemmArgs& g, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.nb1 <= 0 || g.nb2 <= 0) return hipErrorInvalidValue;
    dim3 grid((g.M + TM - 1) / TM, (g.N + TN - 1) / TN, g.nb1 * g.nb2);
    if (grid.y > 65535 || grid.z > 65535) return hipErrorInvalidValue;
    dispatch(gemm_layout_a(g), gemm_layout_b(g), gemm_layout_c(g), grid, g, st);
    return hipGetLastError();
}

hipError_t cgemm_bf16(const CGemmArgs& x, hipStream_t st)
{
    const GemmArgs& g = x.g;
    if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.nb1 <= 0 || g.nb2 <= 0 || g.ksplit != 0) return hipErrorInvalidValue;
    dim3 grid((g.M + TM - 1) / TM, (g.N + TN - 1) / TN, g.nb1 * g.nb2);
    if (grid.y > 65535 || grid.z > 65535) return hipErrorInvalidValue;
    // the two planes of an operand share its strides; a 16-byte form needs both bases aligned
    int al = gemm_layout_a(g), bl = gemm_layout_b(g), cl = gemm_layout_c(g);
    GemmArgs im = g;
    im.A = x.Ai ? x.Ai : g.A;
    im.B = x.Bi ? x.Bi : g.B;
    im.C = x.Ci ? x.Ci : g.C;
    if (gemm_layout_a(im) != al) al = L_ANY;
    if (gemm_layout_b(im) != bl) bl = L_ANY;
    if (gemm_layout_c(im) != cl) cl = C_ANY;
    dispatch(al, bl, cl, grid, x, st);
    return hipGetLastError();
}

long long sum_slices_part_floats(long long n) { return (long long)kSliceGroups * n; }

hipError_t sum_slices(const float* src, float* dst, long long n, int nslices, float beta, float* part, hipStream_t st)
{
    if (n <= 0 || nslices <= 0) return hipErrorInvalidValue;
    const long long nb = (n + 63) / 64;
    if (nb > 2147483647LL) return hipErrorInvalidValue;
    slice_group_sum_kernel<<<dim3((unsigned)nb, kSliceGroups), 256, 0, st>>>(src, part, n, nslices);
    group_total_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(part, dst, n, beta);
    return hipGetLastError();
}

long long channel_sum_part_floats(int B, int C, long long S)
{
    return (long long)B * C * ((S + kRowChunk - 1) / kRowChunk);
}

hipError_t channel_sum(const float* g, float* db, int B, int C, long long S, bool accumulate, float* part,
                       hipStream_t st)
{
    const long long rows = (long long)B * C;
    if (rows <= 0 || rows > 65535 || S <= 0) return hipErrorInvalidValue;
    const int nchunk = (int)((S + kRowChunk - 1) / kRowChunk);
    const dim3 grid((unsigned)nchunk, (unsigned)rows);
    if (S % 4 == 0 && (uintptr_t)g % 16 == 0)
        row_chunk_sum_kernel<true><<<grid, 256, 0, st>>>(g, part, S, nchunk);
    else
        row_chunk_sum_kernel<false><<<grid, 256, 0, st>>>(g, part, S, nchunk);
    channel_total_kernel<<<C, 256, 0, st>>>(part, db, B, C, nchunk, accumulate ? 1 : 0);
    return hipGetLastError();
}

int mse_blocks() { return 256; }

hipError_t mse(const float* out, const float* tgt, float* grad, double* part, long long n, hipStream_t st)
{
    mse_kernel<<<mse_blocks(), 256, 0, st>>>(out, tgt, grad, part, n);
    return hipGetLastError();
}

hipError_t adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
                int t, hipStream_t st)
{
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    adam_kernel<<<grid_of(n), 256, 0, st>>>(p, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace ehd
