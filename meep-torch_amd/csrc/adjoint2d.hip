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

namespace mt {
namespace {

constexpr int BX = 64, BY = 4;

template <typename T>
struct AdjK {
  int nx, ny, y_pmc, ny_e;
  T dt, dx;
  const T* eps;
  Pml<T> hy, hx, ex, ey;
  T *aEz, *aDz, *aHx, *aHy, *apzy, *apzx, *apyx, *apxy;
  T *gHy, *gHx, *gEx, *gEy;
  T* eps_grad;
};

template <typename T>
__global__ __launch_bounds__(BX* BY) void adj_a_kernel(AdjK<T> p, const T* ez_after) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  long long c = (long long)i * ny + j;
  T eps = p.eps[c];
  T ezb = p.aEz[c];
  p.eps_grad[c] = p.eps_grad[c] + (-ezb * ez_after[c]) / eps;
  T dzb = p.aDz[c] + ezb / eps;
  p.aDz[c] = dzb;
  p.aEz[c] = T(0);
  bool live = i >= 1 && i < nx - 1 && (p.y_pmc || (j >= 1 && j < ny - 1));
  T ghy = T(0), ghx = T(0);
  if (live) {
    T cb = p.dt * dzb;
    int ii = i - 1, jj = p.y_pmc ? j : j - 1;
    long long q = (long long)ii * p.ny_e + jj;
    ghy = cb;
    if (p.ex.in(ii)) {
      T pt = p.apyx[q] + cb;
      ghy = cb + p.ex.a[ii] * pt;
      p.apyx[q] = p.ex.b[ii] * pt;
    }
    ghx = -cb;
    if (p.ey.in(jj)) {
      T pt = p.apxy[q] - cb;
      ghx = -cb + p.ey.a[jj] * pt;
      p.apxy[q] = p.ey.b[jj] * pt;
    }
  }
  p.gHy[c] = ghy;
  p.gHx[c] = ghx;
}

template <typename T>
__global__ __launch_bounds__(BX* BY) void adj_b_kernel(AdjK<T> p) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {  // Hy (nx-1, ny)
    long long h = (long long)i * ny + j;
    T hb = p.aHy[h] + (p.gHy[h] - p.gHy[h + ny]) / p.dx;
    p.aHy[h] = hb;
    T gb = p.dt * hb;
    if (p.hx.in(i)) {
      T pt = p.apzx[h] + p.dt * hb;
      gb = gb + p.hx.a[i] * pt;
      p.apzx[h] = p.hx.b[i] * pt;
    }
    p.gEx[h] = gb;
  }
  if (j < ny - 1) {  // Hx (nx, ny-1)
    long long h = (long long)i * (ny - 1) + j;
    long long c = (long long)i * ny + j;
    T hb = p.aHx[h] + (p.gHx[c] - p.gHx[c + 1]) / p.dx;
    p.aHx[h] = hb;
    T gb = -p.dt * hb;
    if (p.hy.in(j)) {
      T pt = p.apzy[h] - p.dt * hb;
      gb = gb + p.hy.a[j] * pt;
      p.apzy[h] = p.hy.b[j] * pt;
    }
    p.gEy[h] = gb;
  }
}

template <typename T>
__global__ __launch_bounds__(BX* BY) void adj_c_kernel(AdjK<T> p) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  T gxm = i >= 1 ? p.gEx[(long long)(i - 1) * ny + j] : T(0);
  T gx0 = i < nx - 1 ? p.gEx[(long long)i * ny + j] : T(0);
  T gym = j >= 1 ? p.gEy[(long long)i * (ny - 1) + j - 1] : T(0);
  T gy0 = j < ny - 1 ? p.gEy[(long long)i * (ny - 1) + j] : T(0);
  p.aEz[(long long)i * ny + j] = (gxm - gx0) / p.dx + (gym - gy0) / p.dx;
}

// Monitor cotangent injection for one monitor at global step t.
template <typename T>
__global__ void adj_inject_kernel(mt_monitor m, int t, double dt_d, int t_rel) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  int q = blockIdx.y;
  int n = m.n0 * m.n1;
  if (e >= n || q >= m.n_terms) return;
  const mt_dft_term& tm = m.term[q];
  int i0 = e / m.n1, i1 = e - i0 * m.n1;
  long long off = (long long)i0 * tm.s0 + (long long)i1 * tm.s1;
  T vb;
  if (m.kind == 1) {
    vb = static_cast<const T*>(m.acc)[t_rel];
  } else {
    T dt = static_cast<T>(dt_d);
    double tt = ((double)t + (tm.t_half ? 0.5 : 1.0)) * dt_d;
    T tq = static_cast<T>(tt);
    const T* om = static_cast<const T*>(m.omega);
    const T* ac = static_cast<const T*>(m.acc);
    vb = T(0);
    for (int f = 0; f < m.nfreq; ++f) {
      T ph = om[f] * tq;
      T re = ac[((long long)f * 2 * m.n_terms + 2 * q) * n + e];
      T im = ac[((long long)f * 2 * m.n_terms + 2 * q + 1) * n + e];
      vb = vb + (dt * cos(ph)) * re + (-(dt * sin(ph))) * im;
    }
  }
  T* a = const_cast<T*>(static_cast<const T*>(tm.a));
  if (tm.b) {
    T* b = const_cast<T*>(static_cast<const T*>(tm.b));
    T h = T(0.5) * vb;
    a[off] = a[off] + h;
    b[off] = b[off] + h;
  } else {
    a[off] = a[off] + vb;
  }
}

template <typename T>
int adj_t(const mt_tm2d* d, const mt_tm2d_adj* a, const void* tape, const mt_monitor* mons, int n_mons,
          int step0, int k, hipStream_t s) {
  AdjK<T> p;
  p.nx = d->nx;
  p.ny = d->ny;
  p.y_pmc = d->y_pmc;
  p.ny_e = d->y_pmc ? d->ny : d->ny - 2;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.eps = static_cast<const T*>(d->eps);
  p.hy = pml_of<T>(d->hy);
  p.hx = pml_of<T>(d->hx);
  p.ex = pml_of<T>(d->ex);
  p.ey = pml_of<T>(d->ey);
  p.aEz = static_cast<T*>(a->aEz);
  p.aDz = static_cast<T*>(a->aDz);
  p.aHx = static_cast<T*>(a->aHx);
  p.aHy = static_cast<T*>(a->aHy);
  p.apzy = static_cast<T*>(a->apsi_Ezy);
  p.apzx = static_cast<T*>(a->apsi_Ezx);
  p.apyx = static_cast<T*>(a->apsi_Hyx);
  p.apxy = static_cast<T*>(a->apsi_Hxy);
  p.gHy = static_cast<T*>(a->gHy);
  p.gHx = static_cast<T*>(a->gHx);
  p.gEx = static_cast<T*>(a->gEx);
  p.gEy = static_cast<T*>(a->gEy);
  p.eps_grad = static_cast<T*>(a->eps_grad);
  const T* tp = static_cast<const T*>(tape);
  dim3 blk(BX, BY), grd = grid2(p.ny, p.nx, blk);
  long long cells = (long long)p.nx * p.ny;
  for (int s_ = k - 1; s_ >= 0; --s_) {
    int t = step0 + s_;
    for (int m = 0; m < n_mons; ++m) {
      const mt_monitor& mm = mons[m];
      int n = mm.n0 * mm.n1;
      if (n == 0) continue;
      dim3 g((n + 255) / 256, mm.n_terms);
      hipLaunchKernelGGL(adj_inject_kernel<T>, g, dim3(256), 0, s, mm, t, d->dt, s_);
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
