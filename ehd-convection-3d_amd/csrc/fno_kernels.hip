// fno_kernels.hip -- the surrogate's GPU kernels: a strided, batched GEMM on
// the bf16 matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation) that
// carries every contraction of the Fourier neural operator (the truncated
// DFTs written as matrices, the per-mode channel mixing, the pointwise
// layers and all their adjoints), plus the elementwise kernels (bias, GELU,
// loss, column sums, Adam).
//
// Operands are FP32 in memory and rounded to bf16 (nearest even) when staged
// into LDS; accumulation and the epilogue are FP32.  Any element strides are
// allowed, so transposes, real/imaginary planes and sub-blocks cost no copy.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fno_internal.h"

namespace ehd {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even (finite inputs)
    return (short)(u >> 16);
}

constexpr int TM = 64, TN = 64, TK = 16;

// Workgroup: 256 threads = 4 waves; the 64x64 C tile is split 2x2 over the
// waves (32x32 each = 2x2 MFMA blocks of 16x16).  Per K step of 16 the A
// tile (64x16) and B tile (16x64) are staged in LDS as bf16.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g)
{
    __shared__ short As[TM][TK + 2];
    __shared__ short Bs[TK][TN + 2];
    const int bz = blockIdx.z;
    const int i1 = bz / g.nb2, i2 = bz - i1 * g.nb2;
    const float* __restrict__ A = g.A + i1 * g.a1 + i2 * g.a2;
    const float* __restrict__ B = g.B + i1 * g.b1 + i2 * g.b2;
    float* __restrict__ C = g.C + i1 * g.c1 + i2 * g.c2;
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
    for (int k0 = kb; k0 < ke; k0 += TK) {
        // stage: 64*16 = 1024 elements of A and of B, 4 per thread
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = tid + 256 * r;
            const int am = e / TK, ak = e - am * TK;          // A tile row-major [64][16]
            const int gm = m0 + am, gk = k0 + ak;
            As[am][ak] = (gm < g.M && gk < ke) ? to_bf16(A[(long long)gm * g.am + (long long)gk * g.ak]) : (short)0;
            const int bk = e / TN, bn = e - bk * TN;          // B tile row-major [16][64]
            const int hk = k0 + bk, hn = n0 + bn;
            Bs[bk][bn] = (hk < ke && hn < g.N) ? to_bf16(B[(long long)hk * g.bk + (long long)hn * g.bn]) : (short)0;
        }
        __syncthreads();
        // MFMA 16x16x16 bf16 operand layout: A[i][k] in lane i + 16*(k/4),
        // element k%4; B[k][j] in lane j + 16*(k/4), element k%4.
        const int li = lane & 15, lk = (lane >> 4) * 4;
        bf16x4 af[2], bfr[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int row = wm + a * 16 + li;
            af[a] = bf16x4{As[row][lk], As[row][lk + 1], As[row][lk + 2], As[row][lk + 3]};
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int col = wn + b * 16 + li;
            bfr[b] = bf16x4{Bs[lk][col], Bs[lk + 1][col], Bs[lk + 2][col], Bs[lk + 3][col]};
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    // D[i][j] of a 16x16 block: lane j + 16*(i/4), element i%4
    const int lj = lane & 15, li4 = (lane >> 4) * 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int gm = m0 + wm + a * 16 + li4 + e, gn = n0 + wn + b * 16 + lj;
                if (gm < g.M && gn < g.N) {
                    float* c = C + (long long)gm * g.cm + (long long)gn * g.cn;
                    const float v = g.alpha * acc[a][b][e];
                    *c = (g.beta != 0.f) ? v + g.beta * *c : v;
                }
            }
}

__global__ void __launch_bounds__(256) sum_slices_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                         long long n, int nslices, float beta)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int j = 0; j < nslices; ++j) s += src[(long long)j * n + i];
        dst[i] = (beta != 0.f) ? beta * dst[i] + s : s;
    }
}

// y[b][c][s] = x[b][c][s] + bias[c]; act: u = that, y = gelu(u)
__global__ void __launch_bounds__(256) bias_act_kernel(float* __restrict__ y, float* __restrict__ u,
                                                       const float* __restrict__ bias, int C, long long S,
                                                       long long total, int act)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int c = (int)((i / S) % C);
        const float v = y[i] + bias[c];
        if (act) {
            u[i] = v;
            y[i] = 0.5f * v * (1.0f + erff(v * 0.70710678118654752f));
        } else {
            y[i] = v;
        }
    }
}

// g[i] *= gelu'(u[i])
__global__ void __launch_bounds__(256) gelu_bwd_kernel(float* __restrict__ g, const float* __restrict__ u,
                                                       long long total)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const float x = u[i];
        const float d = 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * expf(-0.5f * x * x) * 0.3989422804014327f;
        g[i] *= d;
    }
}

// db[c] (+)= sum over b, s of g[b][c][s]: one workgroup per channel
__global__ void __launch_bounds__(256) channel_sum_kernel(const float* __restrict__ g, float* __restrict__ db, int B,
                                                          int C, long long S, int accumulate)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.f;
    for (int b = 0; b < B; ++b) {
        const float* p = g + ((long long)b * C + c) * S;
        for (long long i = threadIdx.x; i < S; i += 256) s += p[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = accumulate ? db[c] + red[0] : red[0];
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

}  // namespace

hipError_t gemm_bf16(const GemmArgs& g, hipStream_t st)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.nb1 <= 0 || g.nb2 <= 0) return hipErrorInvalidValue;
    dim3 grid((g.M + TM - 1) / TM, (g.N + TN - 1) / TN, g.nb1 * g.nb2);
    if (grid.y > 65535 || grid.z > 65535) return hipErrorInvalidValue;
    gemm_bf16_kernel<<<grid, 256, 0, st>>>(g);
    return hipGetLastError();
}

hipError_t sum_slices(const float* src, float* dst, long long n, int nslices, float beta, hipStream_t st)
{
    sum_slices_kernel<<<grid_of(n), 256, 0, st>>>(src, dst, n, nslices, beta);
    return hipGetLastError();
}

hipError_t bias_act(float* y, float* u, const float* bias, int B, int C, long long S, bool gelu, hipStream_t st)
{
    const long long total = (long long)B * C * S;
    bias_act_kernel<<<grid_of(total), 256, 0, st>>>(y, u, bias, C, S, total, gelu ? 1 : 0);
    return hipGetLastError();
}

hipError_t gelu_bwd(float* g, const float* u, long long n, hipStream_t st)
{
    gelu_bwd_kernel<<<grid_of(n), 256, 0, st>>>(g, u, n);
    return hipGetLastError();
}

hipError_t channel_sum(const float* g, float* db, int B, int C, long long S, bool accumulate, hipStream_t st)
{
    channel_sum_kernel<<<C, 256, 0, st>>>(g, db, B, C, S, accumulate ? 1 : 0);
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
