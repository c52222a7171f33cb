// fno_internal.h -- kernels of the Fourier-neural-operator surrogate.
#pragma once

#include <hip/hip_runtime.h>

namespace ehd {

// C(i1,i2)[m][n] = alpha * sum_k A(i1,i2)[m][k] * B(i1,i2)[k][n] + beta * C(i1,i2)[m][n]
// with A[m][k] at A + i1*a1 + i2*a2 + m*am + k*ak (likewise B, C): FP32 in
// memory, bf16 matrix-core products, FP32 accumulation.
struct GemmArgs {
    int M, N, K;
    const float* A;
    long long am, ak;
    const float* B;
    long long bk, bn;
    float* C;
    long long cm, cn;
    int nb1, nb2;
    long long a1, a2, b1, b2, c1, c2;
    float alpha, beta;
    int ksplit;        // > 0: batch index i2 takes k in [i2*ksplit, (i2+1)*ksplit) (split-K partial sums)
    // epilogue after alpha/beta: + bias[m] (null: none), then act:
    //   kActGelu: pre (C's offsets) = the value, C = gelu(value);
    //   kActGeluGrad: C = value * gelu'(dact) (dact at C's offsets)
    const float* bias = nullptr;
    int act = 0;
    float* pre = nullptr;
    const float* dact = nullptr;
};
constexpr int kActNone = 0, kActGelu = 1, kActGeluGrad = 2;

hipError_t gemm_bf16(const GemmArgs& g, hipStream_t st);
// Complex product on the matrix cores: with Ar = g.A, Br = g.B, Cr = g.C,
//   Cr = alpha * Re[(Ar + i sa Ai)(Br + i sb Bi)] + beta * Cr,
//   Ci = alpha * Im[(Ar + i sa Ai)(Br + i sb Bi)] + beta * Ci,
// each imaginary plane at the same strides as its real plane; Ai or Bi null:
// that operand is real; Ci null: only Cr is written.  sa, sb = +1 or -1
// (-1 conjugates).  No split-K.
struct CGemmArgs {
    GemmArgs g;
    const float* Ai;
    const float* Bi;
    float* Ci;
    float sa, sb;
};
hipError_t cgemm_bf16(const CGemmArgs& x, hipStream_t st);
// staging layout gemm_bf16 picks for A and B: 0 k unit-stride, 1 m (n) unit-stride
// (16-byte loads), 2 element loads
int gemm_layout_a(const GemmArgs& g);
int gemm_layout_b(const GemmArgs& g);
// result layout: 0 element stores, 1 n unit-stride, 2 m unit-stride (16-byte
// stores through LDS)
int gemm_layout_c(const GemmArgs& g);
// dst[i] = beta*dst[i] + sum_j src[j*n + i], j < nslices; part holds
// sum_slices_part_floats(n) floats of workspace
long long sum_slices_part_floats(long long n);
hipError_t sum_slices(const float* src, float* dst, long long n, int nslices, float beta, float* part, hipStream_t st);
// db[c] (+)= sum_{b,s} g[b][c][s]; part holds channel_sum_part_floats(B, C, S) floats of workspace
long long channel_sum_part_floats(int B, int C, long long S);
hipError_t channel_sum(const float* g, float* db, int B, int C, long long S, bool accumulate, float* part,
                       hipStream_t st);
// grad = 2 (out - tgt) / n, part[mse_blocks()] = partial sums of (out - tgt)^2
int mse_blocks();
hipError_t mse(const float* out, const float* tgt, float* grad, double* part, long long n, hipStream_t st);
// Adam (Kingma & Ba 2015), t counts from 1
hipError_t adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
                int t, hipStream_t st);

}  // namespace ehd
