// Reverse-time discrete adjoint of the 2D TM Yee step (eps + CPML + soft
// source + DFT monitors), for dL/deps.
//
// Forward step t maps x_t -> x_{t+1}:
//   B half  gy = Dy Ez, pzy' = b pzy + a gy, Hx' = Hx - dt (gy + pzy')
//           gx = Dx Ez, pzx' = b pzx + a gx, Hy' = Hy + dt (gx + pzx')
//   D half  dHy = Dx^T-stencil Hy', dHx = Dy-stencil Hx', psi recursions,
//           Dz' = Dz + dt curl - c_t mask,   Ez' = Dz' / eps
// and the monitors sample (Ez', Hy', Hx') after the step.  eps enters only
// through Ez' = Dz'/eps, so dL/deps = sum_t -(dL/dEz'_t) Ez'_t / eps.  One
// reverse step is the exact transpose of these linear maps, in four passes:
//   inject  monitor cotangents: sum_f dt (cos(w t) dacc_re - sin(w t) dacc_im)
//   A       Ez' cotangent -> eps gradient and Dz cotangent; curl cotangent ->
//           D-half psi cotangents and the H-difference cotangents gHy, gHx
//   B       Hy/Hx cotangents (transposed differences of gHy/gHx), B-half psi
//           cotangents, and the E-difference cotangents gEx, gEy
//   C       Ez cotangent of x_t = transposed differences of gEx, gEy
// psi cotangents are tracked inside the CPML shells only: outside them a = 0,
// so they never feed back into the fields or eps.
#include <hip/hip_runtime.h>

#include "common.h"
#include "adjoint2d.h"

namespace mt {
namespace {

template <typename T>
__global__ __launch_bounds__(ABX* ABY) void adj_a_kernel(AdjK<T> p, const T* ez_after) {
  int j = blockIdx.x * ABX + threadIdx.x;
  int i = blockIdx.y * ABY + threadIdx.y;
  if (i < p.nx && j < p.ny) adj_a_cell(p, ez_after, i, j);
}

template <typename T>
__global__ __launch_bounds__(ABX* ABY) void adj_b_kernel(AdjK<T> p) {
  int j = blockIdx.x * ABX + threadIdx.x;
  int i = blockIdx.y * ABY + threadIdx.y;
  if (i < p.nx && j < p.ny) adj_b_cell(p, i, j);
}

template <typename T>
__global__ __launch_bounds__(ABX* ABY) void adj_c_kernel(AdjK<T> p) {
  int j = blockIdx.x * ABX + threadIdx.x;
  int i = blockIdx.y * ABY + threadIdx.y;
  if (i < p.nx && j < p.ny) adj_c_cell(p, i, j);
}

template <typename T>
__global__ __launch_bounds__(kInjE* kInjS) void adj_inject_kernel(mt_monitor m, int t, double dt_d, int t_rel) {
  adj_inject_block<T>(m, blockIdx.x * kInjE, t, dt_d, t_rel);
}

template <typename T>
int adj_t(const mt_tm2d* d, const mt_tm2d_adj* a, const void* tape, const mt_monitor* mons, int n_mons,
          int step0, int k, hipStream_t s) {
  AdjK<T> p = make_adjk<T>(d, a);
  const T* tp = static_cast<const T*>(tape);
  dim3 blk(ABX, ABY), grd = grid2(p.ny, p.nx, blk);
  long long cells = (long long)p.nx * p.ny;
  for (int s_ = k - 1; s_ >= 0; --s_) {
    int t = step0 + s_;
    for (int m = 0; m < n_mons; ++m) {
      const mt_monitor& mm = mons[m];
      int n = mm.n0 * mm.n1;
      if (n == 0) continue;
      hipLaunchKernelGGL(adj_inject_kernel<T>, dim3((n + kInjE - 1) / kInjE), dim3(kInjE * kInjS), 0, s, mm, t,
                         d->dt, s_);
    }
    hipLaunchKernelGGL(adj_a_kernel<T>, grd, blk, 0, s, p, tp + (long long)s_ * cells);
    hipLaunchKernelGGL(adj_b_kernel<T>, grd, blk, 0, s, p);
    hipLaunchKernelGGL(adj_c_kernel<T>, grd, blk, 0, s, p);
    int rc = check_launch("tm2d adjoint step");
    if (rc) return rc;
  }
  return MT_OK;
}

}  // namespace
}  // namespace mt

using namespace mt;

extern "C" int mt_tm2d_adjoint(const mt_tm2d* p, const mt_tm2d_adj* a, const void* tape,
                               const mt_monitor* mons, int n_mons, int step0, int k, void* stream) {
  if (!p || !a) return fail(MT_EINVAL, "null problem or adjoint state");
  if (p->nx < 3 || p->ny < 3) return fail(MT_EINVAL, "grid must be at least 3x3");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  if (k > 0 && !tape) return fail(MT_EINVAL, "adjoint needs the forward Ez tape");
  if (p->dtype != MT_F32 && p->dtype != MT_F64) return fail(MT_EINVAL, "dtype must be f32 or f64");
  if (!p->eps || !a->aEz || !a->aDz || !a->aHx || !a->aHy || !a->apsi_Ezy || !a->apsi_Ezx ||
      !a->apsi_Hyx || !a->apsi_Hxy || !a->gHy || !a->gHx || !a->gEx || !a->gEy || !a->eps_grad)
    return fail(MT_EINVAL, "null adjoint pointer");
  int rc;
  if ((rc = check_axis(p->hy, p->ny - 1, "hy")) || (rc = check_axis(p->hx, p->nx - 1, "hx")) ||
      (rc = check_axis(p->ex, p->nx - 2, "ex")) ||
      (rc = check_axis(p->ey, p->y_pmc ? p->ny : p->ny - 2, "ey")))
    return rc;
  for (int i = 0; i < n_mons; ++i) {
    const mt_monitor& m = mons[i];
    if (m.n_terms < 1 || m.n_terms > MT_MAX_TERMS || !m.acc) return fail(MT_EINVAL, "bad monitor");
    if (m.kind == 0 && (m.nfreq < 1 || !m.omega)) return fail(MT_EINVAL, "monitor: no frequencies");
    for (int q = 0; q < m.n_terms; ++q)
      if (!m.term[q].a) return fail(MT_EINVAL, "monitor: null adjoint view");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == MT_F64 ? adj_t<double>(p, a, tape, mons, n_mons, step0, k, s)
                            : adj_t<float>(p, a, tape, mons, n_mons, step0, k, s);
}
