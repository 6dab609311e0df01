// Kernels of the spectral neural-operator surrogate: one batched, strided
// BF16 x BF16 -> FP32 GEMM on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_bf16) that carries every layer's forward and
// backward (truncated 2D DFT, its inverse, per-mode complex channel mixing,
// pointwise channel mixing, and their transposes), plus the elementwise
// pieces of training and inference: FP32 -> BF16 rounding, bias + GELU and
// its derivative, spectral-weight packing, channel and row reductions, Adam.
//
// GEMM:  C[z] = alpha * sum_{r<R} A[z,r] (MxK) * B[z,r] (KxN) + beta * C[z]
// with independent element strides for every operand axis, so transposed
// operands and interleaved layouts need no copies.  A 256-thread workgroup
// (four waves of 64) computes a 64x64 tile of C; each wave owns a 32x32
// quarter in one MFMA accumulator (16 FP32 values per lane).  The reduction
// space (R x ceil(K/32) tiles of 32) can be split over grid.z; split partial
// tiles go to a workspace and are summed in split order by a second kernel,
// so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.h"
#include "meep_surrogate.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

namespace mt {
namespace {

constexpr int GT = 64;    // C tile edge
constexpr int GK = 32;    // reduction tile
constexpr int LDA = 40;   // LDS row pitch in bf16 (80 B: 16-byte aligned rows)

struct GemmK {
  int M, N, K, R, batch, splits, ktiles, f_per_split;
  const uint16_t* A;
  long long a_m, a_k, a_b, a_r;
  const uint16_t* B;
  long long b_k, b_n, b_b, b_r;
  float* C;
  long long c_m, c_n, c_b;
  float* work;
  float alpha, beta;
  int vec_a, vec_b;
};

// Tile loads.  MC: A contiguous along m (else along k).  KC: B contiguous
// along k (else along n).  Every thread moves 8 elements; a chunk that is
// fully in range and 16-byte aligned moves as one 128-bit load.
template <bool MC>
__device__ __forceinline__ void load_a(const GemmK& g, const uint16_t* A, int bm, int k0,
                                       uint16_t (*As)[LDA]) {
  const int t = threadIdx.x;
  if (!MC) {  // k contiguous: thread -> (row, 8 k)
    const int i = t >> 2, kc = (t & 3) * 8;
    const int gi = bm + i, gk = k0 + kc;
    if (g.vec_a && gi < g.M && gk + 8 <= g.K) {
      uint4 v = *reinterpret_cast<const uint4*>(A + (long long)gi * g.a_m + gk);
      *reinterpret_cast<uint4*>(&As[i][kc]) = v;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        As[i][kc + q] = (gi < g.M && gk + q < g.K) ? A[(long long)gi * g.a_m + (long long)(gk + q) * g.a_k]
                                                  : (uint16_t)0;
    }
  } else {  // m contiguous: thread -> (k, 8 rows)
    const int kk = t >> 3, ic = (t & 7) * 8;
    const int gk = k0 + kk, gi = bm + ic;
    uint16_t v[8];
    if (g.vec_a && gk < g.K && gi + 8 <= g.M) {
      uint4 u = *reinterpret_cast<const uint4*>(A + gi + (long long)gk * g.a_k);
      *reinterpret_cast<uint4*>(v) = u;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        v[q] = (gk < g.K && gi + q < g.M) ? A[(long long)(gi + q) * g.a_m + (long long)gk * g.a_k] : (uint16_t)0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) As[ic + q][kk] = v[q];
  }
}

template <bool KC>
__device__ __forceinline__ void load_b(const GemmK& g, const uint16_t* B, int bn, int k0,
                                       uint16_t (*Bs)[LDA]) {
  const int t = threadIdx.x;
  if (KC) {  // k contiguous: thread -> (col, 8 k)
    const int j = t >> 2, kc = (t & 3) * 8;
    const int gj = bn + j, gk = k0 + kc;
    if (g.vec_b && gj < g.N && gk + 8 <= g.K) {
      uint4 v = *reinterpret_cast<const uint4*>(B + (long long)gj * g.b_n + gk);
      *reinterpret_cast<uint4*>(&Bs[j][kc]) = v;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        Bs[j][kc + q] = (gj < g.N && gk + q < g.K) ? B[(long long)(gk + q) * g.b_k + (long long)gj * g.b_n]
                                                  : (uint16_t)0;
    }
  } else {  // n contiguous: thread -> (k, 8 cols)
    const int kk = t >> 3, jc = (t & 7) * 8;
    const int gk = k0 + kk, gj = bn + jc;
    uint16_t v[8];
    if (g.vec_b && gk < g.K && gj + 8 <= g.N) {
      uint4 u = *reinterpret_cast<const uint4*>(B + (long long)gk * g.b_k + gj);
      *reinterpret_cast<uint4*>(v) = u;
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q)
        v[q] = (gk < g.K && gj + q < g.N) ? B[(long long)gk * g.b_k + (long long)(gj + q) * g.b_n] : (uint16_t)0;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) Bs[jc + q][kk] = v[q];
  }
}

template <bool MC, bool KC>
__global__ __launch_bounds__(256) void gemm_kernel(GemmK g) {
  __shared__ __align__(16) uint16_t As[GT][LDA];
  __shared__ __align__(16) uint16_t Bs[GT][LDA];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int bm = blockIdx.y * GT, bn = blockIdx.x * GT;
  const int bz = blockIdx.z % g.batch, split = blockIdx.z / g.batch;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const long long f_total = (long long)g.R * g.ktiles;
  const long long f0 = (long long)split * g.f_per_split;
  const long long f1 = f0 + g.f_per_split < f_total ? f0 + g.f_per_split : f_total;

  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;

  for (long long f = f0; f < f1; ++f) {
    const int r = (int)(f / g.ktiles), k0 = (int)(f % g.ktiles) * GK;
    const uint16_t* A = g.A + (long long)bz * g.a_b + (long long)r * g.a_r;
    const uint16_t* B = g.B + (long long)bz * g.b_b + (long long)r * g.b_r;
    load_a<MC>(g, A, bm, k0, As);
    load_b<KC>(g, B, bn, k0, Bs);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 16) {
      const int kq = kk + (lane >> 5) * 8;
      bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + (lane & 31)][kq]);
      bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }

  const int j = bn + wn + (lane & 31);
  if (g.splits == 1) {
    float* C = g.C + (long long)bz * g.c_b;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = bm + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (i < g.M && j < g.N) {
        float* c = C + (long long)i * g.c_m + (long long)j * g.c_n;
        float v = g.alpha * acc[r];
        if (g.beta != 0.f) v += g.beta * *c;
        *c = v;
      }
    }
  } else {
    float* W = g.work + ((long long)split * g.batch + bz) * g.M * g.N;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = bm + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (i < g.M && j < g.N) W[(long long)i * g.N + j] = acc[r];
    }
  }
}

// C = alpha * sum_s work[s] + beta * C, splits summed in order
__global__ void split_reduce_kernel(GemmK g) {
  const long long mn = (long long)g.M * g.N;
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= mn * g.batch) return;
  const int bz = (int)(e / mn);
  const long long ij = e % mn;
  const int i = (int)(ij / g.N), j = (int)(ij % g.N);
  float s = 0.f;
  for (int q = 0; q < g.splits; ++q) s += g.work[((long long)q * g.batch + bz) * mn + ij];
  float* c = g.C + (long long)bz * g.c_b + (long long)i * g.c_m + (long long)j * g.c_n;
  float v = g.alpha * s;
  if (g.beta != 0.f) v += g.beta * *c;
  *c = v;
}

__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__global__ void cast_kernel(const float* x, uint16_t* y, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}

// x (rows, C, P): pre = x + bias[c] (stored back into x); act = gelu(pre)
// or pre; written as BF16 and, if act32 is given, as FP32.
__global__ void bias_act_kernel(float* x, const float* bias, int C, int P, long long n, int act,
                                uint16_t* act16, float* act32) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = x[i];
  if (bias) v += bias[(i / P) % C];
  x[i] = v;
  float y = act ? gelu(v) : v;
  if (act16) act16[i] = to_bf16(y);
  if (act32) act32[i] = y;
}

// dpre = dy * gelu'(pre) (or dy); dy may be a per-row value (rows of length
// P) times scale, for the backward of a mean over P.
__global__ void act_grad_kernel(const float* dy, const float* pre, long long n, int act, int P, int row_bcast,
                                float scale, float* d32, uint16_t* d16) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float d = row_bcast ? dy[i / P] * scale : dy[i];
  if (act) d *= gelu_d(pre[i]);
  if (d32) d32[i] = d;
  if (d16) d16[i] = to_bf16(d);
}

// out[c] = scale * sum_{b, p} x[b, c, p]; one block per channel, fixed
// per-thread strides and a fixed tree, so the sum is deterministic.
// out[c] = scale * sum_b rows[b, c]: the per-channel fold of the (b, c) row
// sums, in order of b.
__global__ __launch_bounds__(256) void channel_fold_kernel(const float* rows, float* out, int B, int C,
                                                           float scale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int b = 0; b < B; ++b) s += rows[(long long)b * C + c];
  out[c] = s * scale;
}

// out[row] = scale * sum_p x[row, p]
__global__ __launch_bounds__(256) void row_sum_kernel(const float* x, float* out, int P, float scale) {
  __shared__ float red[256];
  const float* row = x + (long long)blockIdx.x * P;
  float s = 0.f;
  for (int p = threadIdx.x; p < P; p += 256) s += row[p];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0] * scale;
}

// Spectral weights W (Ci, Co, Mo) complex -> per-mode real block matrix
// Wp[m] (2Ci x 2Co) = [[Wre, Wim], [-Wim, Wre]] in BF16, so that
// [Vre | Vim] * Wp[m] = [Re(V W) | Im(V W)].
__global__ void pack_spectral_kernel(const float* wre, const float* wim, uint16_t* wp, int Ci, int Co, int Mo) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = 4LL * Ci * Co;
  if (e >= per * Mo) return;
  const int m = (int)(e / per);
  const int rem = (int)(e % per);
  const int k = rem / (2 * Co), n = rem % (2 * Co);
  const int pi = k / Ci, i = k % Ci, po = n / Co, o = n % Co;
  const long long w = ((long long)i * Co + o) * Mo + m;
  float v = (pi == po) ? wre[w] : (pi == 0 ? wim[w] : -wim[w]);
  wp[e] = to_bf16(v);
}

// Gradient of the packing: dWre = d11 + d22, dWim = d12 - d21 (+= into the
// gradient buffers when accumulate is set).
__global__ void unpack_spectral_grad_kernel(const float* dwp, float* dwre, float* dwim, int Ci, int Co, int Mo) {
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)Ci * Co * Mo) return;
  const int m = (int)(e % Mo);
  const int o = (int)((e / Mo) % Co);
  const int i = (int)(e / ((long long)Mo * Co));
  const float* d = dwp + (long long)m * 4 * Ci * Co;
  const int W2 = 2 * Co;
  float d11 = d[(long long)i * W2 + o], d22 = d[(long long)(Ci + i) * W2 + Co + o];
  float d12 = d[(long long)i * W2 + Co + o], d21 = d[(long long)(Ci + i) * W2 + o];
  dwre[e] = d11 + d22;
  dwim[e] = d12 - d21;
}

__global__ void adam_kernel(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float wd) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i] + wd * w[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline unsigned blocks(long long n, int t) { return (unsigned)((n + t - 1) / t); }

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace
}  // namespace mt

using namespace mt;

extern "C" int mt_gemm_bf16(const mt_gemm* q, void* stream) {
  if (!q) return fail(MT_EINVAL, "gemm: null descriptor");
  const mt_gemm& d = *q;
  if (d.M < 0 || d.N < 0 || d.K < 0 || d.R < 1 || d.batch < 1 || d.splits < 1)
    return fail(MT_EINVAL, "gemm: bad sizes");
  if (d.M == 0 || d.N == 0) return MT_OK;
  if (!d.A || !d.B || !d.C) return fail(MT_EINVAL, "gemm: null operand");
  if ((long long)d.batch * d.splits > 65535) return fail(MT_EINVAL, "gemm: batch x splits exceeds 65535");
  const bool a_mc = d.a_m == 1 && d.a_k != 1;
  const bool b_kc = d.b_k == 1 && d.b_n != 1;
  GemmK g{};
  g.M = d.M; g.N = d.N; g.K = d.K; g.R = d.R; g.batch = d.batch;
  g.ktiles = d.K > 0 ? (d.K + GK - 1) / GK : 0;
  const long long f_total = (long long)d.R * g.ktiles;
  int splits = d.splits;
  if (f_total == 0) splits = 1;
  if (splits > f_total && f_total > 0) splits = (int)f_total;
  g.f_per_split = f_total > 0 ? (int)((f_total + splits - 1) / splits) : 0;
  if (f_total > 0) splits = (int)((f_total + g.f_per_split - 1) / g.f_per_split);
  g.splits = splits;
  if (splits > 1 && !d.work) return fail(MT_EINVAL, "gemm: splits > 1 needs a workspace");
  g.A = static_cast<const uint16_t*>(d.A); g.a_m = d.a_m; g.a_k = d.a_k; g.a_b = d.a_b; g.a_r = d.a_r;
  g.B = static_cast<const uint16_t*>(d.B); g.b_k = d.b_k; g.b_n = d.b_n; g.b_b = d.b_b; g.b_r = d.b_r;
  g.C = d.C; g.c_m = d.c_m; g.c_n = d.c_n; g.c_b = d.c_b;
  g.work = d.work; g.alpha = d.alpha; g.beta = d.beta;
  // 128-bit loads need the contiguous axis at unit stride and every other
  // stride (and the base) on 16-byte boundaries
  const long long a_other = a_mc ? d.a_k : d.a_m;
  const long long b_other = b_kc ? d.b_n : d.b_k;
  g.vec_a = ((a_mc ? d.a_m : d.a_k) == 1) && aligned16(d.A) && a_other % 8 == 0 && d.a_b % 8 == 0 &&
            d.a_r % 8 == 0;
  g.vec_b = ((b_kc ? d.b_k : d.b_n) == 1) && aligned16(d.B) && b_other % 8 == 0 && d.b_b % 8 == 0 &&
            d.b_r % 8 == 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  dim3 grid((d.N + GT - 1) / GT, (d.M + GT - 1) / GT, d.batch * splits);
  if (grid.y > 65535) return fail(MT_EINVAL, "gemm: M too large");
  if (a_mc && b_kc) hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, s, g);
  else if (a_mc) hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, s, g);
  else if (b_kc) hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, s, g);
  if (splits > 1) {
    long long n = (long long)d.M * d.N * d.batch;
    hipLaunchKernelGGL(split_reduce_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, g);
  }
  return check_launch("gemm_bf16");
}

extern "C" int mt_cast_bf16(const float* x, void* y, long long n, void* stream) {
  if (n <= 0) return MT_OK;
  if (!x || !y) return fail(MT_EINVAL, "cast: null pointer");
  hipLaunchKernelGGL(cast_kernel, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), x,
                     static_cast<uint16_t*>(y), n);
  return check_launch("cast_bf16");
}

extern "C" int mt_bias_act(float* x, const float* bias, int C, int P, long long n, int act, void* act16,
                           float* act32, void* stream) {
  if (n <= 0) return MT_OK;
  if (!x || C < 1 || P < 1) return fail(MT_EINVAL, "bias_act: bad arguments");
  hipLaunchKernelGGL(bias_act_kernel, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), x,
                     bias, C, P, n, act, static_cast<uint16_t*>(act16), act32);
  return check_launch("bias_act");
}

extern "C" int mt_act_grad(const float* dy, const float* pre, long long n, int act, int P, int row_bcast,
                           float scale, float* d32, void* d16, void* stream) {
  if (n <= 0) return MT_OK;
  if (!dy || (act && !pre) || P < 1) return fail(MT_EINVAL, "act_grad: bad arguments");
  hipLaunchKernelGGL(act_grad_kernel, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), dy,
                     pre, n, act, P, row_bcast, scale, d32, static_cast<uint16_t*>(d16));
  return check_launch("act_grad");
}

extern "C" int mt_channel_sum(const float* x, float* out, float* work, int B, int C, int P, float scale,
                              void* stream) {
  if (B <= 0 || C <= 0) return MT_OK;
  if (!x || !out || !work || P < 1 || (long long)B * C > 0x7fffffff) return fail(MT_EINVAL, "channel_sum: bad arguments");
  // one block per (b, c) row fills the whole GPU (a block per channel left
  // most CUs idle), then a deterministic fold over b
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(row_sum_kernel, dim3((unsigned)(B * C)), dim3(256), 0, s, x, work, P, 1.0f);
  hipLaunchKernelGGL(channel_fold_kernel, dim3((C + 255) / 256), dim3(256), 0, s, work, out, B, C, scale);
  return check_launch("channel_sum");
}

extern "C" int mt_row_sum(const float* x, float* out, long long rows, int P, float scale, void* stream) {
  if (rows <= 0) return MT_OK;
  if (!x || !out || P < 1 || rows > 0x7fffffff) return fail(MT_EINVAL, "row_sum: bad arguments");
  hipLaunchKernelGGL(row_sum_kernel, dim3((unsigned)rows), dim3(256), 0, static_cast<hipStream_t>(stream), x, out,
                     P, scale);
  return check_launch("row_sum");
}

extern "C" int mt_pack_spectral(const float* wre, const float* wim, void* wp, int Ci, int Co, int Mo,
                                void* stream) {
  if (!wre || !wim || !wp || Ci < 1 || Co < 1 || Mo < 1) return fail(MT_EINVAL, "pack_spectral: bad arguments");
  long long n = 4LL * Ci * Co * Mo;
  hipLaunchKernelGGL(pack_spectral_kernel, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     wre, wim, static_cast<uint16_t*>(wp), Ci, Co, Mo);
  return check_launch("pack_spectral");
}

extern "C" int mt_unpack_spectral_grad(const float* dwp, float* dwre, float* dwim, int Ci, int Co, int Mo,
                                       void* stream) {
  if (!dwp || !dwre || !dwim || Ci < 1 || Co < 1 || Mo < 1)
    return fail(MT_EINVAL, "unpack_spectral_grad: bad arguments");
  long long n = (long long)Ci * Co * Mo;
  hipLaunchKernelGGL(unpack_spectral_grad_kernel, dim3(blocks(n, 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), dwp, dwre, dwim, Ci, Co, Mo);
  return check_launch("unpack_spectral_grad");
}

extern "C" int mt_adam(float* w, const float* g, float* m, float* v, long long n, float lr, float b1, float b2,
                       float eps, int step, float wd, void* stream) {
  if (n <= 0) return MT_OK;
  if (!w || !g || !m || !v || step < 1) return fail(MT_EINVAL, "adam: bad arguments");
  float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m, v,
                     n, lr, b1, b2, eps, bc1, bc2, wd);
  return check_launch("adam");
}
