// 3D Yee leapfrog with CPML shells (sigma = 0), point / masked soft sources.
//
// Two passes per step: H from curl E, then E from curl H.  One thread per
// cell (i, j, k) of the nx*ny*nz lattice updates every component whose grid
// covers that cell; z is the contiguous axis and runs along the wave.
// Arithmetic follows meep_torch/core/stepper_3d.py:363-445 op by op (lean:
// stepper_3d_lean.py:234-296), so FP32 rounds exactly like the reference.
// psi arrays keep the reference's full-grid shapes; only cells inside a CPML
// shell touch them.
#include <hip/hip_runtime.h>

#include "common.h"

namespace mt {
namespace {

constexpr int BZ = 64;
constexpr int BYY = 4;

template <typename T>
struct K3 {
  int nx, ny, nz, lean;
  T dt, dx;
  const T* eps;
  Pml<T> h[3], e[3];
  T *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
  T* psi[12];
};

enum { P_EZ_Y, P_EY_Z, P_EX_Z, P_EZ_X, P_EY_X, P_EX_Y, P_HZ_Y, P_HY_Z, P_HX_Z, P_HZ_X, P_HY_X, P_HX_Y };

// d + psi with psi' = b psi + a d inside a shell (stepper_3d.py:368-370)
template <typename T>
__device__ __forceinline__ T ce(const Pml<T>& pr, int idx, T* psi, long long q, T d) {
  if (!pr.in(idx)) return d;
  T ps = pr.b[idx] * psi[q] + pr.a[idx] * d;
  psi[q] = ps;
  return d + ps;
}

template <typename T>
__global__ __launch_bounds__(BZ* BYY) void h3_kernel(K3<T> p) {
  int k = blockIdx.x * BZ + threadIdx.x;
  int j = blockIdx.y * BYY + threadIdx.y;
  int i = blockIdx.z;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (j >= ny || k >= nz) return;
  const T dx = p.dx, dt = p.dt;
  if (j < ny - 1 && k < nz - 1) {  // Hx (nx, ny-1, nz-1)
    long long h = ((long long)i * (ny - 1) + j) * (nz - 1) + k;
    long long ez = ((long long)i * ny + j) * (nz - 1) + k;
    long long ey = ((long long)i * (ny - 1) + j) * nz + k;
    T cy = ce(p.h[1], j, p.psi[P_EZ_Y], h, (p.Ez[ez + (nz - 1)] - p.Ez[ez]) / dx);
    T cz = ce(p.h[2], k, p.psi[P_EY_Z], h, (p.Ey[ey + 1] - p.Ey[ey]) / dx);
    p.Hx[h] = p.Hx[h] - dt * (cy - cz);
  }
  if (i < nx - 1 && k < nz - 1) {  // Hy (nx-1, ny, nz-1)
    long long h = ((long long)i * ny + j) * (nz - 1) + k;
    long long ex = ((long long)i * ny + j) * nz + k;
    long long ez = h;  // Ez (nx, ny, nz-1) shares the index for i < nx-1
    T cz = ce(p.h[2], k, p.psi[P_EX_Z], h, (p.Ex[ex + 1] - p.Ex[ex]) / dx);
    T cx = ce(p.h[0], i, p.psi[P_EZ_X], h, (p.Ez[ez + (long long)ny * (nz - 1)] - p.Ez[ez]) / dx);
    p.Hy[h] = p.Hy[h] - dt * (cz - cx);
  }
  if (i < nx - 1 && j < ny - 1) {  // Hz (nx-1, ny-1, nz)
    long long h = ((long long)i * (ny - 1) + j) * nz + k;
    long long ey = h;  // Ey (nx, ny-1, nz) shares the index for i < nx-1
    long long ex = ((long long)i * ny + j) * nz + k;
    T cx = ce(p.h[0], i, p.psi[P_EY_X], h, (p.Ey[ey + (long long)(ny - 1) * nz] - p.Ey[ey]) / dx);
    T cy = ce(p.h[1], j, p.psi[P_EX_Y], h, (p.Ex[ex + nz] - p.Ex[ex]) / dx);
    p.Hz[h] = p.Hz[h] - dt * (cx - cy);
  }
}

template <typename T>
__device__ __forceinline__ T e_update(const K3<T>& p, T E, T eps_c, T curl) {
  if (p.lean) {
    T inv = T(1) / eps_c;
    return E + (curl * inv) * p.dt;
  }
  return (E * eps_c + p.dt * curl) / eps_c;
}

template <typename T>
__device__ __forceinline__ T inject(const K3<T>& p, const mt_source3d& s, int comp, int i, int j, int k,
                                    T E, T eps_c, T c) {
  if (s.kind == 0 || s.comp != comp || c == T(0)) return E;
  if (i < s.i0 || i >= s.i1 || j < s.j0 || j >= s.j1 || k < s.k0 || k >= s.k1) return E;
  if (s.kind == 1) {
    T inv = T(1) / eps_c;
    return p.lean ? E - c * inv : E + inv * c;
  }
  const T* w = static_cast<const T*>(s.w);
  T m = w[((long long)(i - s.i0) * (s.j1 - s.j0) + (j - s.j0)) * (s.k1 - s.k0) + (k - s.k0)];
  if (p.lean) return E - (m * c) * (T(1) / eps_c);
  return E - (c * m) / eps_c;
}

template <typename T>
__global__ __launch_bounds__(BZ* BYY) void e3_kernel(K3<T> p, mt_source3d src, T c, T* tape) {
  int k = blockIdx.x * BZ + threadIdx.x;
  int j = blockIdx.y * BYY + threadIdx.y;
  int i = blockIdx.z;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (j >= ny || k >= nz) return;
  const T dx = p.dx;
  long long ce_ = ((long long)i * ny + j) * nz + k;  // cell index in eps
  T eps0 = p.eps[ce_];
  if (i < nx - 1) {  // Ex (nx-1, ny, nz)
    long long x = ce_;
    T chy = T(0), chz = T(0);
    bool jy = j >= 1 && j < ny - 1, kz = k >= 1 && k < nz - 1;
    if (jy) {
      long long hz = ((long long)i * (ny - 1) + j) * nz + k;
      long long q = ((long long)i * (ny - 2) + (j - 1)) * nz + k;
      chy = ce(p.e[1], j - 1, p.psi[P_HZ_Y], q, (p.Hz[hz] - p.Hz[hz - nz]) / dx);
    }
    if (kz) {
      long long hy = ((long long)i * ny + j) * (nz - 1) + k;
      long long q = ((long long)i * ny + j) * (nz - 2) + (k - 1);
      chz = ce(p.e[2], k - 1, p.psi[P_HY_Z], q, (p.Hy[hy] - p.Hy[hy - 1]) / dx);
    }
    T curl = (jy && kz) ? chy - chz : T(0);
    T epsc = T(0.5) * (eps0 + p.eps[ce_ + (long long)ny * nz]);
    T E = e_update(p, p.Ex[x], epsc, curl);
    E = inject(p, src, 0, i, j, k, E, epsc, c);
    p.Ex[x] = E;
    if (tape) tape[x] = E;
  }
  if (j < ny - 1) {  // Ey (nx, ny-1, nz)
    long long y = ((long long)i * (ny - 1) + j) * nz + k;
    T chz = T(0), chx = T(0);
    bool kz = k >= 1 && k < nz - 1, ix = i >= 1 && i < nx - 1;
    if (kz) {
      long long hx = ((long long)i * (ny - 1) + j) * (nz - 1) + k;
      long long q = ((long long)i * (ny - 1) + j) * (nz - 2) + (k - 1);
      chz = ce(p.e[2], k - 1, p.psi[P_HX_Z], q, (p.Hx[hx] - p.Hx[hx - 1]) / dx);
    }
    if (ix) {
      long long hz = ((long long)i * (ny - 1) + j) * nz + k;
      long long q = ((long long)(i - 1) * (ny - 1) + j) * nz + k;
      chx = ce(p.e[0], i - 1, p.psi[P_HZ_X], q, (p.Hz[hz] - p.Hz[hz - (long long)(ny - 1) * nz]) / dx);
    }
    T curl = (kz && ix) ? chz - chx : T(0);
    T epsc = T(0.5) * (eps0 + p.eps[ce_ + nz]);
    T E = e_update(p, p.Ey[y], epsc, curl);
    E = inject(p, src, 1, i, j, k, E, epsc, c);
    p.Ey[y] = E;
    if (tape) tape[(long long)(nx - 1) * ny * nz + y] = E;
  }
  if (k < nz - 1) {  // Ez (nx, ny, nz-1)
    long long z = ((long long)i * ny + j) * (nz - 1) + k;
    T chx = T(0), chy = T(0);
    bool ix = i >= 1 && i < nx - 1, jy = j >= 1 && j < ny - 1;
    if (ix) {
      long long hy = z;  // Hy (nx-1, ny, nz-1) at (i, j, k) for i < nx-1
      long long q = ((long long)(i - 1) * ny + j) * (nz - 1) + k;
      chx = ce(p.e[0], i - 1, p.psi[P_HY_X], q, (p.Hy[hy] - p.Hy[hy - (long long)ny * (nz - 1)]) / dx);
    }
    if (jy) {
      long long hx = ((long long)i * (ny - 1) + j) * (nz - 1) + k;
      long long q = ((long long)i * (ny - 2) + (j - 1)) * (nz - 1) + k;
      chy = ce(p.e[1], j - 1, p.psi[P_HX_Y], q, (p.Hx[hx] - p.Hx[hx - (nz - 1)]) / dx);
    }
    T curl = (ix && jy) ? chx - chy : T(0);
    T epsc = T(0.5) * (eps0 + p.eps[ce_ + 1]);
    T E = e_update(p, p.Ez[z], epsc, curl);
    E = inject(p, src, 2, i, j, k, E, epsc, c);
    p.Ez[z] = E;
    if (tape) tape[(long long)(nx - 1) * ny * nz + (long long)nx * (ny - 1) * nz + z] = E;
  }
}

template <typename T>
int run3_t(const mt_yee3d* d, const mt_source3d* src, const double* coef, const MonitorPlan& mp,
           int step0, int k, hipStream_t s) {
  K3<T> p;
  p.nx = d->nx;
  p.ny = d->ny;
  p.nz = d->nz;
  p.lean = d->lean;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.eps = static_cast<const T*>(d->eps);
  for (int a = 0; a < 3; ++a) {
    p.h[a] = pml_of<T>(d->h[a]);
    p.e[a] = pml_of<T>(d->e[a]);
  }
  p.Ex = static_cast<T*>(d->Ex);
  p.Ey = static_cast<T*>(d->Ey);
  p.Ez = static_cast<T*>(d->Ez);
  p.Hx = static_cast<T*>(d->Hx);
  p.Hy = static_cast<T*>(d->Hy);
  p.Hz = static_cast<T*>(d->Hz);
  for (int q = 0; q < 12; ++q) p.psi[q] = static_cast<T*>(d->psi[q]);
  mt_source3d sv{};
  if (src && src->kind) sv = *src;
  dim3 blk(BZ, BYY, 1);
  dim3 grd((p.nz + BZ - 1) / BZ, (p.ny + BYY - 1) / BYY, p.nx);
  int S = mp.stage_steps > 0 ? mp.stage_steps : 1;
  int filled = 0, first = step0;
  const long long e_cells = (long long)(p.nx - 1) * p.ny * p.nz + (long long)p.nx * (p.ny - 1) * p.nz +
                            (long long)p.nx * p.ny * (p.nz - 1);
  for (int s_ = 0; s_ < k; ++s_) {
    hipLaunchKernelGGL(h3_kernel<T>, grd, blk, 0, s, p);
    T c = sv.kind ? static_cast<T>(coef[s_]) : T(0);
    T* slot = d->tape ? static_cast<T*>(d->tape) + (long long)s_ * e_cells : nullptr;
    hipLaunchKernelGGL(e3_kernel<T>, grd, blk, 0, s, p, sv, c, slot);
    int rc = check_launch("yee3d step");
    if (rc) return rc;
    if (mp.n_mons) {
      if ((rc = monitors_sample(mp, filled, s_, s))) return rc;
      if (++filled == S || s_ == k - 1) {
        if ((rc = monitors_flush(mp, filled, first, d->dt, s))) return rc;
        first += filled;
        filled = 0;
      }
    }
  }
  return MT_OK;
}

}  // namespace
}  // namespace mt

using namespace mt;

extern "C" int mt_yee3d_run(const mt_yee3d* p, const mt_source3d* src, const double* coef,
                            const mt_monitor* mons, int n_mons, void* stage, long long stage_bytes,
                            int stage_steps, int step0, int k, void* stream) {
  if (!p) return fail(MT_EINVAL, "null problem");
  if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(MT_EINVAL, "grid must be at least 3x3x3");
  if ((long long)p->ny * p->nz > (1LL << 31) || p->nx > 65535)
    return fail(MT_EINVAL, "grid too large for the launch shape");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  if (p->dtype != MT_F32 && p->dtype != MT_F64) return fail(MT_EINVAL, "dtype must be f32 or f64");
  if (!p->eps || !p->Ex || !p->Ey || !p->Ez || !p->Hx || !p->Hy || !p->Hz)
    return fail(MT_EINVAL, "null field pointer");
  for (int q = 0; q < 12; ++q)
    if (!p->psi[q]) return fail(MT_EINVAL, "null psi pointer");
  int n[3] = {p->nx, p->ny, p->nz};
  int rc;
  for (int a = 0; a < 3; ++a) {
    if ((rc = check_axis(p->h[a], n[a] - 1, "h"))) return rc;
    if ((rc = check_axis(p->e[a], n[a] - 2, "e"))) return rc;
  }
  if (src && src->kind) {
    if (src->kind != 1 && src->kind != 2) return fail(MT_EINVAL, "source kind must be 1 or 2");
    if (src->comp < 0 || src->comp > 2) return fail(MT_EINVAL, "source component must be Ex/Ey/Ez");
    int ext[3] = {p->nx, p->ny, p->nz};
    ext[src->comp] -= 1;
    if (src->i0 < 0 || src->j0 < 0 || src->k0 < 0 || src->i1 > ext[0] || src->j1 > ext[1] ||
        src->k1 > ext[2] || src->i0 >= src->i1 || src->j0 >= src->j1 || src->k0 >= src->k1)
      return fail(MT_EINVAL, "source box outside the component grid");
    if (src->kind == 2 && !src->w) return fail(MT_EINVAL, "mask source without weights");
    if (!coef && k > 0) return fail(MT_EINVAL, "source without coefficients");
  }
  MonitorPlan mp{mons, n_mons, stage, stage_bytes, stage_steps, p->dtype};
  if (n_mons && (rc = monitors_validate(mp))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == MT_F64 ? run3_t<double>(p, src, coef, mp, step0, k, s)
                            : run3_t<float>(p, src, coef, mp, step0, k, s);
}
