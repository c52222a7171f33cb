/*
 * ehd_fno.h -- C-ABI of the Fourier-neural-operator surrogate (warm starts).
 *
 * No counterpart in the reference (which has no learned component).  A 3-D
 * FNO maps the governing groups of a box and its grid coordinates to the
 * box's steady fields; the lattice Boltzmann run then starts from that
 * prediction (ehd_init with EHD_INIT_FIELDS) instead of the hydrostatic
 * state.  Network (activations [batch][channel][z][y][x], FP32 in memory,
 * every contraction on the bf16 matrix cores with FP32 accumulation):
 *   lift     v0 = a Pw + Pb                      d_in -> width
 *   layer l  v' = gelu(K_l v + v W_l + b_l)      (no gelu after the last)
 *   K_l      truncated real DFT in z (modes 0..mz-1), complex DFTs in y and x
 *            (modes 0..m-1 and N-m..N-1), per-mode complex channel mixing
 *            R_l[kz][ky][kx][in][out], the inverse transforms
 *   project  out = gelu(v Q1w + Q1b) Q2w + Q2b  width -> d_hidden -> d_out
 * Loss: mean squared error; optimiser: Adam.
 *
 * Flat parameter order (float32): Pw[d_in][width], Pb[width], then per layer
 * W[width][width], b[width], Rr[mz][2my][2mx][width][width], Ri[same], then
 * Q1w[width][d_hidden], Q1b[d_hidden], Q2w[d_hidden][d_out], Q2b[d_out].
 * Host arrays: inputs [batch][d_in][Z][Y][X], outputs [batch][d_out][Z][Y][X].
 * Return codes as in ehd.h; ehd_last_error() holds the message.
 */
#ifndef EHD_FNO_H
#define EHD_FNO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ehd_fno_config {
    int X, Y, Z;            /* grid; X, Y even, Z >= 2 */
    int width;              /* channels of the Fourier layers (32) */
    int mx, my, mz;         /* retained modes per axis (12, capped at X/2, Y/2, Z/2) */
    int layers;             /* Fourier layers (4) */
    int d_in, d_hidden, d_out;   /* 6, 128, 5 */
    int max_batch;          /* largest batch of one call */
} ehd_fno_config;

typedef struct ehd_fno ehd_fno;

int ehd_fno_create(const ehd_fno_config* cfg, ehd_fno** out);
void ehd_fno_destroy(ehd_fno* f);
long long ehd_fno_num_params(const ehd_fno* f);
int ehd_fno_set_params(ehd_fno* f, const float* params);
int ehd_fno_get_params(ehd_fno* f, float* params);

/* out = network(in) for batch <= max_batch samples */
int ehd_fno_forward(ehd_fno* f, int batch, const float* in, float* out);

/* mean-squared-error loss of network(in) against target and its gradient
 * with respect to every parameter (flat order; grad may be NULL).  in and
 * target both NULL: the batch the last gradients / train_step call copied to
 * the device is used again (no host transfer; batch must not exceed it, and
 * an ehd_fno_forward call in between invalidates it). */
int ehd_fno_gradients(ehd_fno* f, int batch, const float* in, const float* target, double* loss, float* grad);

/* one Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the loss of the batch;
 * loss (may be NULL) receives the loss before the update; in and target NULL
 * as in ehd_fno_gradients */
int ehd_fno_train_step(ehd_fno* f, int batch, const float* in, const float* target, float lr, double* loss);

/* A training set resident in HBM: n samples of inputs [n][d_in][Z][Y][X] and
 * targets [n][d_out][Z][Y][X], copied once (replacing any earlier set; the
 * library allocates it, EHD_ENOMEM if it does not fit). */
int ehd_fno_set_dataset(ehd_fno* f, int n, const float* in, const float* target);

/* One Adam step (as ehd_fno_train_step) on the samples idx[0..batch) of the
 * resident set, gathered on the device: no host transfer per step. */
int ehd_fno_train_step_indices(ehd_fno* f, int batch, const int* idx, float lr, double* loss);

/* C = A B for row-major host matrices (M x K, K x N) through the bf16
 * matrix-core GEMM the network uses (checks of its operand layout) */
int ehd_fno_gemm(int M, int N, int K, const float* A, const float* B, float* C);

/* The same GEMM with the element strides and batching the network's stages
 * use (checks of every staging layout):
 *   C(b,s)[m][n] = alpha sum_k A(b)[m][k] B(b)[k][n] + beta C(b,s)[m][n]
 * with A(b)[m][k] = A[b*as[2] + m*as[0] + k*as[1]], B(b)[k][n] =
 * B[b*bs[2] + k*bs[0] + n*bs[1]], C(b,s)[m][n] = C[b*cs[2] + s*cs[3] +
 * m*cs[0] + n*cs[1]] for b < nb; with ksplit > 0, slice s < ceil(K/ksplit)
 * sums only k in [s*ksplit, (s+1)*ksplit) (split-K partial products).
 * Host arrays of a_len, b_len, c_len floats; every index the strides reach
 * must lie inside them (checked; strides >= 0). */
int ehd_fno_gemm_strided(int M, int N, int K, int nb, int ksplit, const float* A, const long long as[3],
                         long long a_len, const float* B, const long long bs[3], long long b_len, float* C,
                         const long long cs[4], long long c_len, float alpha, float beta);

/* The complex product the spectral layers use, in one launch:
 *   Cr = alpha Re[(Ar + i sa Ai)(Br + i sb Bi)] + beta Cr,
 *   Ci = alpha Im[(Ar + i sa Ai)(Br + i sb Bi)] + beta Ci,
 * each plane indexed as A, B and C above (the imaginary plane of an operand
 * at the same strides as its real plane; cs[3] unused).  sa, sb = +1 or -1
 * (-1: conjugate).  Ai or Bi NULL: that operand is real; Ci NULL: only Cr is
 * written.  Host arrays of a_len, b_len, c_len floats per plane (checked as
 * above). */
int ehd_fno_cgemm_strided(int M, int N, int K, int nb, const float* Ar, const float* Ai, const long long as[3],
                          long long a_len, const float* Br, const float* Bi, const long long bs[3], long long b_len,
                          float* Cr, float* Ci, const long long cs[3], long long c_len, int sa, int sb, float alpha,
                          float beta);

/* Work of the GEMMs issued since the last reset (forward, gradients and
 * training steps): algorithmic FLOPs (2 M N K per real product; a complex
 * product counts as its four real ones), algorithmic HBM bytes (A and B read
 * once, C written once and read too when accumulated, all FP32), launches.
 * reset != 0 clears the counters after reading. */
typedef struct ehd_fno_work {
    double flops;
    double bytes;
    long long launches;
} ehd_fno_work;
int ehd_fno_gemm_work(ehd_fno* f, int reset, ehd_fno_work* out);

#ifdef __cplusplus
}
#endif
#endif
