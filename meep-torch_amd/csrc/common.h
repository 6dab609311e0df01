// Shared helpers for the FDTD engine kernels (gfx950 / CDNA4, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "meep_fdtd.h"

namespace mt {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);

// CPML profile on the device side: psi recursion only inside the shells.
template <typename T>
struct Pml {
  const T* b;
  const T* a;
  int lo_end, hi_begin;
  __device__ __forceinline__ bool in(int i) const { return i < lo_end || i >= hi_begin; }
};

template <typename T>
inline Pml<T> pml_of(const mt_pml_axis& ax) {
  return Pml<T>{static_cast<const T*>(ax.b), static_cast<const T*>(ax.a), ax.lo_end, ax.hi_begin};
}

// Four-point averages of the TE fixed point (stepper_te.py:52-61): the value
// of an Ey-shaped field at Ex (i, j) (i < nx-1), and of an Ex-shaped field at
// Ey (i, j) (j < ny-1), summed in the reference's order, zero past the
// walls.  Each is also the transpose of the other (the adjoint of the
// fixed point scatters cotangents with them).
template <typename T>
__device__ __forceinline__ T avg4_to_ex(const T* ey, int nx, int ny, int i, int j) {
  const T a = j >= 1 ? ey[(long long)i * (ny - 1) + j - 1] : T(0);
  const T b = j < ny - 1 ? ey[(long long)i * (ny - 1) + j] : T(0);
  const T c = j >= 1 ? ey[(long long)(i + 1) * (ny - 1) + j - 1] : T(0);
  const T d = j < ny - 1 ? ey[(long long)(i + 1) * (ny - 1) + j] : T(0);
  return T(0.25) * (a + b + c + d);
}

template <typename T>
__device__ __forceinline__ T avg4_to_ey(const T* ex, int nx, int ny, int i, int j) {
  const T a = i >= 1 ? ex[(long long)(i - 1) * ny + j] : T(0);
  const T b = i >= 1 ? ex[(long long)(i - 1) * ny + j + 1] : T(0);
  const T c = i < nx - 1 ? ex[(long long)i * ny + j] : T(0);
  const T d = i < nx - 1 ? ex[(long long)i * ny + j + 1] : T(0);
  return T(0.25) * (a + b + c + d);
}

int check_axis(const mt_pml_axis& ax, int n, const char* name);
// A 2D source box inside an n0 x n1 field (or absent).
int check_src(const mt_source2d* s, int n0, int n1);

// RN(1/b) when div_c's fast sequence equals IEEE division for every finite
// FP32 dividend with divisor b (checked exhaustively on the host once per b
// and cached), else 0; div_c then divides.
float div_recip_f32(float b);
template <typename T>
inline T const_div_recip(T b);
template <>
inline float const_div_recip<float>(float b) { return div_recip_f32(b); }
template <>
inline double const_div_recip<double>(double) { return 0.0; }

// a / b for the constant grid step b > 0 with y = RN(1/b): q0 = a y, then
// one FMA correction with the exact remainder a - q0 b (Markstein's
// sequence), the sign of a restored with copysign (a zero quotient takes a's
// sign, as IEEE division gives it).  That this equals the IEEE quotient is
// not assumed from a theorem: the host passes y only for FP32 divisors whose
// sequence it has checked against IEEE division for every dividend of
// magnitude in [2^-100, 2^100] and for zero (const_div_recip), and y = 0
// otherwise (FP64 included), which selects true division.  Other dividends
// (tiny values in the tail of a wavefront, huge ones) take true division.
// Four instructions plus the exponent test, against about ten for IEEE
// division (a reciprocal among them).
__device__ __forceinline__ bool dx_div_covered(float a) {
  return (unsigned)(__builtin_amdgcn_frexp_expf(a) + 99) <= 200u;  // 0 or 2^-100 <= |a| <= 2^100
}

template <typename T>
__device__ __forceinline__ T div_c(T a, T b, T y) {
  if (y == T(0)) return a / b;
  const T q0 = a * y;
  const T r = fma(-q0, b, a);
  T q = copysign(fma(r, y, q0), a);
  if (!dx_div_covered(static_cast<float>(a))) {
    __asm__ volatile("");  // a real branch: keeps the division off the common path
    q = a / b;
  }
  return q;
}

// The same for a group of quotients that the caller checks once: each call
// takes the fast sequence and records its dividend's exponent; when exact()
// is false for a lane, the caller recomputes that lane's whole group with
// IeeeDiv (one branch per group instead of one per quotient).  Fast = false
// (no verified y) divides every quotient.
template <typename T, bool Fast>
struct DxDiv {
  T b, y;
  unsigned worst;
  __device__ __forceinline__ T operator()(T a) {
    if constexpr (!Fast) return a / b;
    const T q0 = a * y;
    const T r = fma(-q0, b, a);
    worst = max(worst, (unsigned)(__builtin_amdgcn_frexp_expf(static_cast<float>(a)) + 99));
    return copysign(fma(r, y, q0), a);
  }
  __device__ __forceinline__ bool exact() const { return !Fast || worst <= 200u; }
};

template <typename T>
struct IeeeDiv {
  T b;
  __device__ __forceinline__ T operator()(T a) const { return a / b; }
};

// Monitor support shared by the 2D and 3D steppers (dft.hip).
struct MonitorPlan {
  const mt_monitor* mons;
  int n_mons;
  void* stage;
  long long stage_bytes;
  int stage_steps;
  int dtype;
};
int monitors_validate(const MonitorPlan& mp);
// Kernel timing hook (mt_kernel_timing_enable): brackets launches of the
// given kind (1: forward stencil, 2: batched adjoint step) with HIP events
// on their stream when that kind is enabled; no-ops otherwise.
void timing_begin(hipStream_t s, int kind = 1);
void timing_end(hipStream_t s, long long cell_steps, int kind = 1);
// Gather the values of every term for step slot `slot` (after the step).
int monitors_sample(const MonitorPlan& mp, int slot, int t_rel, hipStream_t s);
// Fold `n_filled` staged steps whose first global step index is t0 into the accumulators.
int monitors_flush(const MonitorPlan& mp, int n_filled, int t0, double dt, hipStream_t s);

inline dim3 grid2(int n_fast, int n_slow, dim3 blk) {
  return dim3((n_fast + blk.x - 1) / blk.x, (n_slow + blk.y - 1) / blk.y, 1);
}

}  // namespace mt
