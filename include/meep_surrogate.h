/*
 * meep_surrogate.h -- C ABI of the spectral neural-operator kernels in
 * libmeepfdtd.so (BF16 matrix-core GEMM and the elementwise pieces of
 * training and inference).
 *
 * The reference has no learned component: these entry points have no
 * counterpart in it.  They serve the surrogate-assisted inverse-design loop
 * (meep_torch_amd.surrogate), which predicts transmission spectra and their
 * density gradients and falls back to the FDTD engine when a verification
 * solve disagrees.
 *
 * Conventions as in meep_fdtd.h: device pointers, 0 on success, negative
 * code plus mt_last_error() on failure, all work enqueued on `stream`.
 * BF16 buffers hold raw bfloat16 bit patterns (uint16, round-to-nearest-even
 * from FP32).
 */
#ifndef MEEP_SURROGATE_H
#define MEEP_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

/* C[z] = alpha * sum_{r<R} A[z,r] (M x K) * B[z,r] (K x N) + beta * C[z],
 * z < batch.  A, B: BF16; C: FP32.  Element (i, k) of A[z,r] sits at
 * A + z*a_b + r*a_r + i*a_m + k*a_k, and likewise for B and C (no r axis).
 * The reduction space R x ceil(K/32) is cut into `splits` contiguous parts
 * over grid.z; for splits > 1, `work` holds splits*batch*M*N floats of
 * partial tiles, summed in split order (deterministic). */
typedef struct {
  int M, N, K, R, batch, splits;
  const void* A;
  long long a_m, a_k, a_b, a_r;
  const void* B;
  long long b_k, b_n, b_b, b_r;
  float* C;
  long long c_m, c_n, c_b;
  float* work;
  float alpha, beta;
} mt_gemm;

int mt_gemm_bf16(const mt_gemm* g, void* stream);

/* y[i] = bf16(x[i]), i < n. */
int mt_cast_bf16(const float* x, void* y, long long n, void* stream);

/* x viewed as (rows, C, P): x += bias[c] (pre-activation, stored in place);
 * act = gelu(x) if `act` else x, written to act16 (BF16) and/or act32
 * (FP32) when non-NULL.  gelu is the exact erf form. */
int mt_bias_act(float* x, const float* bias, int C, int P, long long n, int act, void* act16, float* act32,
                void* stream);

/* d = dy * gelu'(pre) (or dy if !act); with row_bcast, dy is one value per
 * row of length P, times `scale` (backward of a mean).  Written to d32
 * and/or d16 when non-NULL. */
int mt_act_grad(const float* dy, const float* pre, long long n, int act, int P, int row_bcast, float scale,
                float* d32, void* d16, void* stream);

/* out[c] = scale * sum_{b<B, p<P} x[b, c, p] (fixed-order, deterministic).
 * work: device scratch of B*C floats (the per-(b, c) row sums). */
int mt_channel_sum(const float* x, float* out, float* work, int B, int C, int P, float scale, void* stream);

/* out[r] = scale * sum_{p<P} x[r, p]. */
int mt_row_sum(const float* x, float* out, long long rows, int P, float scale, void* stream);

/* Spectral weights W[i, o, m] = wre + i*wim (Ci x Co x Mo, FP32) -> per-mode
 * BF16 real block matrices wp[m] (2Ci x 2Co) = [[Wre, Wim], [-Wim, Wre]]. */
int mt_pack_spectral(const float* wre, const float* wim, void* wp, int Ci, int Co, int Mo, void* stream);

/* Gradient of mt_pack_spectral: dwre = d11 + d22, dwim = d12 - d21. */
int mt_unpack_spectral_grad(const float* dwp, float* dwre, float* dwim, int Ci, int Co, int Mo, void* stream);

/* Adam (with decoupled-from-nothing L2 term wd*w added to the gradient):
 * m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 * w -= lr * (m / (1-b1^step)) / (sqrt(v / (1-b2^step)) + eps). */
int mt_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
            int step, float wd, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MEEP_SURROGATE_H */
