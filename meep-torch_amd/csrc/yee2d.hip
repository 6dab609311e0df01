// 2D TM / TE Yee leapfrog with CPML shells and soft-source injection.
//
// Each step is two passes over the grid (H half, then E half for TM; D half
// then B half for TE).  Arithmetic follows the reference op by op so that FP32
// rounds exactly like its PyTorch stepper: derivatives are true divisions by
// dx, the psi recursion is b*psi + a*d, curls are summed left to right, E is a
// true division D/eps, and the build disables FMA contraction.  Cells outside
// the CPML shells skip the psi recursion (there b = 1, a = 0, psi stays 0, so
// the result is unchanged).  Fields keep the reference shapes (last axis y
// contiguous); threads run along y for coalesced access.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "common.h"
#include "tiled2d.h"

namespace mt {
namespace {

constexpr int BX = 64;  // threads along y (contiguous)
constexpr int BY = 4;   // threads along x

template <typename T>
struct TmK {
  int nx, ny, y_pmc;
  T dt, dx;
  const T* eps;
  Pml<T> hy, hx, ex, ey;
  T *Ez, *Dz, *Hx, *Hy, *pEzy, *pEzx, *pHyx, *pHxy;
  int ny_e;
  T* tape;  // optional (k, nx, ny) record of Ez after each step
  // optional media (null / 0 when absent)
  int n_poles, kerr;
  const T *c1, *c2, *c3, *chi2, *chi3, *d1, *d2, *h1x, *h2x, *h1y, *h2y;
  T *Pc, *Pp;
  // magnetic Lorentz poles (n_mpoles > 0): B, the M stacks and their scalars
  int n_mpoles;
  const T *mc1x, *mc2x, *mc3x, *mc1y, *mc2y, *mc3y;
  T *Bx, *By, *Mxc, *Mxp, *Myc, *Myp;
  // symmetry cuts at the low edges: ycut / xcut 0 none, 1 zero ghost H (even
  // mirror), 2 on-plane row Hx0 / Hy0; parities of the ghost Ez
  int ycut, xcut, ymir, y_rot2, rot4;
  T *Hx0, *Hy0;
};

// H = B - sum M' with M' = (mc1 M + mc2 M_prev) + mc3 H_old per pole, in pole
// order (stepper_tm.py:209-224); q is the cell's index in a grid of nq cells.
template <typename T>
__device__ __forceinline__ T h_from_b(int nm, const T* mc1, const T* mc2, const T* mc3, T* Mc, T* Mp, long long q,
                                      long long nq, T b, T h_old) {
  T sum = T(0);
  for (int k = 0; k < nm; ++k) {
    const long long o = (long long)k * nq + q;
    const T mc = Mc[o];
    const T mn = (mc1[o] * mc + mc2[o] * Mp[o]) + mc3[o] * h_old;
    Mp[o] = mc;
    Mc[o] = mn;
    sum = k ? sum + mn : mn;
  }
  return b - sum;
}

// meep_torch/core/stepper_tm.py:85-92,126-128 (B half, H from curl E)
// Weight of a magnetic source box at (i, j), or 0 outside it.
template <typename T>
__device__ __forceinline__ bool in_box(const mt_source2d& s, int i, int j) {
  return s.present && i >= s.i0 && i < s.i1 && j >= s.j0 && j < s.j1;
}

template <typename T>
__device__ __forceinline__ T box_w(const mt_source2d& s, int i, int j) {
  return static_cast<const T*>(s.w)[(long long)(i - s.i0) * (s.j1 - s.j0) + (j - s.j0)];
}

// The source of step s of a call: per-step weight boxes (wstep > 0) advance
// the weight pointer; a static box is used as is.
template <typename T>
mt_source2d step_source(const mt_source2d& s, int step) {
  mt_source2d v = s;
  if (v.present && v.wstep) v.w = static_cast<const T*>(v.w) + (long long)step * v.wstep;
  return v;
}

// B half plus the optional TFSF magnetic currents (stepper_tm.py:169-193)
template <typename T>
__global__ __launch_bounds__(BX* BY) void tm_b_kernel(TmK<T> p, mt_source2d hxs, mt_source2d hys, T hc) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  if (i >= p.nx || j >= p.ny) return;
  const int ny = p.ny;
  T ez = p.Ez[(long long)i * ny + j];
  // on-plane H rows of the odd cuts (stepper_tm.py:94-115): the ghost Ez
  // beyond the cut is -Ez_cut (odd mirror), the C2 partner phase *
  // Ez[nx-1-i, 0] (Rotate2) or the other cut's line phase * Ez[0, i]
  // (Rotate4); H0 -/+= (dt (Ez_cut - ghost)) / dx from Ez at time n
  if (j == 0 && p.ycut == 2) {
    const T g = p.rot4 ? T(p.rot4) * p.Ez[i] : p.y_rot2 ? T(p.y_rot2) * p.Ez[(long long)(p.nx - 1 - i) * ny] : -ez;
    p.Hx0[i] = p.Hx0[i] - (p.dt * (ez - g)) / p.dx;
  }
  if (i == 0 && p.xcut == 2) {
    const T g = p.rot4 ? T(p.rot4) * p.Ez[(long long)j * ny] : -ez;
    p.Hy0[j] = p.Hy0[j] + (p.dt * (ez - g)) / p.dx;
  }
  if (j < ny - 1) {
    long long h = (long long)i * (ny - 1) + j;
    T d = (p.Ez[(long long)i * ny + j + 1] - ez) / p.dx;
    T curl = d;
    if (p.hy.in(j)) {
      T ps = p.hy.b[j] * p.pEzy[h] + p.hy.a[j] * d;
      p.pEzy[h] = ps;
      curl = d + ps;
    }
    if (p.n_mpoles) {  // stepper_tm.py:103-106,187-193,214-219: B integrates the curl and takes the current
      T b = p.Bx[h] - p.dt * curl;
      if (in_box<T>(hxs, i, j)) b = b - hc * box_w<T>(hxs, i, j);
      p.Bx[h] = b;
      p.Hx[h] = h_from_b(p.n_mpoles, p.mc1x, p.mc2x, p.mc3x, p.Mxc, p.Mxp, h, (long long)p.nx * (ny - 1), b, p.Hx[h]);
    } else {
      // stepper_tm.py:116-121: sigma_b / mu fold into h1, h2
      T v = p.h1x ? p.h1x[h] * p.Hx[h] - p.h2x[h] * curl : p.Hx[h] - p.dt * curl;
      if (in_box<T>(hxs, i, j)) v = v - hc * box_w<T>(hxs, i, j);
      p.Hx[h] = v;
    }
  }
  if (i < p.nx - 1) {
    long long h = (long long)i * ny + j;
    T d = (p.Ez[(long long)(i + 1) * ny + j] - ez) / p.dx;
    T curl = d;
    if (p.hx.in(i)) {
      T ps = p.hx.b[i] * p.pEzx[h] + p.hx.a[i] * d;
      p.pEzx[h] = ps;
      curl = d + ps;
    }
    if (p.n_mpoles) {
      T b = p.By[h] + p.dt * curl;
      if (in_box<T>(hys, i, j)) b = b - hc * box_w<T>(hys, i, j);
      p.By[h] = b;
      p.Hy[h] = h_from_b(p.n_mpoles, p.mc1y, p.mc2y, p.mc3y, p.Myc, p.Myp, h, (long long)(p.nx - 1) * ny, b, p.Hy[h]);
    } else {
      T v = p.h1y ? p.h1y[h] * p.Hy[h] + p.h2y[h] * curl : p.Hy[h] + p.dt * curl;
      if (in_box<T>(hys, i, j)) v = v - hc * box_w<T>(hys, i, j);
      p.Hy[h] = v;
    }
  }
}

// meep_torch/core/stepper_tm.py:130-166 (D half), :197-200 (source), :226-227 (Ez)
template <typename T>
__global__ __launch_bounds__(BX* BY) void tm_d_kernel(TmK<T> p, mt_source2d src, T coef, int s_rel) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  if (i >= p.nx || j >= p.ny) return;
  const int nx = p.nx, ny = p.ny;
  long long c = (long long)i * ny + j;
  T dz = p.Dz[c];
  // live Dz: the interior, all rows under y_pmc, and from the cut row /
  // column on under a symmetry cut (stepper_tm.py:134-155)
  bool live = (p.xcut ? i < nx - 1 : (i >= 1 && i < nx - 1)) &&
              (p.y_pmc || (p.ycut ? j < ny - 1 : (j >= 1 && j < ny - 1)));
  T curl = T(0);  // zero on the PEC rows, as the reference's padding
  if (live) {
    // H beyond the wall: 0 (PMC, even mirror) or the on-plane row (odd / Rotate cuts)
    T hx1 = (j < ny - 1) ? p.Hx[(long long)i * (ny - 1) + j] : T(0);
    T hx0 = (j >= 1) ? p.Hx[(long long)i * (ny - 1) + j - 1] : (p.ycut == 2 ? p.Hx0[i] : T(0));
    T hy0 = (i >= 1) ? p.Hy[(long long)(i - 1) * ny + j] : (p.xcut == 2 ? p.Hy0[j] : T(0));
    T dHx = (hx1 - hx0) / p.dx;
    T dHy = (p.Hy[(long long)i * ny + j] - hy0) / p.dx;
    int ii = p.xcut ? i : i - 1, jj = (p.y_pmc || p.ycut) ? j : j - 1;
    long long q = (long long)ii * p.ny_e + jj;
    T a1 = dHy;
    if (p.ex.in(ii)) {
      T ps = p.ex.b[ii] * p.pHyx[q] + p.ex.a[ii] * dHy;
      p.pHyx[q] = ps;
      a1 = dHy + ps;
    }
    curl = a1 - dHx;
    if (p.ey.in(jj)) {
      T ps = p.ey.b[jj] * p.pHxy[q] + p.ey.a[jj] * dHx;
      p.pHxy[q] = ps;
      curl = curl - ps;
    }
    if (!p.d1) dz = dz + p.dt * curl;
  }
  if (p.d1) dz = p.d1[c] * dz + p.d2[c] * curl;  // stepper_tm.py:161-163 (sigma_d)
  if (src.present && i >= src.i0 && i < src.i1 && j >= src.j0 && j < src.j1) {
    const T* w = static_cast<const T*>(src.w);
    const T wv = w[(long long)(i - src.i0) * (src.j1 - src.j0) + (j - src.j0)];
    // stepper_tm.py:198-200: under sigma_d the scale is d2 and coef the raw amplitude
    dz = p.d2 ? dz - (p.d2[c] * coef) * wv : dz - coef * wv;
  }
  p.Dz[c] = dz;
  T rhs = dz;
  if (p.n_poles) {  // stepper_tm.py:203-208, 224: P(n+1) from Ez(n), then Dz - sum P
    const T ezo = p.Ez[c];
    const long long plane = (long long)nx * ny;
    T sum = T(0);
    for (int k = 0; k < p.n_poles; ++k) {
      const long long q = k * plane + c;
      const T pc = p.Pc[q], pp = p.Pp[q];
      const T pn = p.c1[q] * pc + p.c2[q] * pp + p.c3[q] * ezo;
      p.Pp[q] = pc;
      p.Pc[q] = pn;
      sum = k == 0 ? pn : sum + pn;
    }
    rhs = dz - sum;
  }
  const T e0 = p.eps[c];
  T ez = rhs / e0;
  if (p.kerr) {  // stepper_tm.py:228-231: D = eps E + chi2 E^2 + chi3 E^3 by fixed point
    const T x2 = p.chi2[c], x3 = p.chi3[c];
    for (int it = 0; it < 3; ++it) ez = rhs / (e0 + x2 * ez + x3 * ez * ez);
  }
  p.Ez[c] = ez;
  // under chi2/chi3 the tape keeps rhs: the adjoint recomputes the iterates from it
  if (p.tape) p.tape[(long long)s_rel * nx * ny + c] = p.kerr ? rhs : ez;
}

template <typename T>
struct TeK {
  int nx, ny;
  T dt, dx;
  const T *eps_x, *eps_y;
  Pml<T> hx, hy, ey, ex;
  T *Hz, *Ex, *Ey, *Dx, *Dy, *pEyx, *pExy, *pHzy, *pHzx;
  T* tape;  // optional: [Ex | Ey] after each step
  // optional media (null / 0 when absent)
  int npx, npy;
  const T *c1x, *c2x, *c3x, *c1y, *c2y, *c3y, *d1x, *d2x, *d1y, *d2y, *hz1, *hz2;
  T *Pxc, *Pxp, *Pyc, *Pyp;
  int kerr;
  const T *chi2x, *chi3x, *chi2y, *chi3y;
  T *kx, *ky;  // scratch of the Ex / Ey shapes for the Kerr passes, tensor eps and oriented poles
  // optional tensor eps (inverse components on the E grids) and oriented poles
  const T *ixx, *ixyx, *iyy, *ixyy, *c3xy, *c3yx;
  int scratch;  // the D half writes E (or, with tensor eps, rhs) to kx / ky instead of Ex / Ey
  int ymir, xmir;  // mirror parities of the low-edge cuts
};

// D - sum P of one E cell with the pole stacks already advanced (the value
// te_rhs_from_d returned, recomputed in the same order).
template <typename T>
__device__ __forceinline__ T te_rhs(T d, int np, long long plane, long long c, const T* pc) {
  if (!np) return d;
  T sum = pc[c];
  for (int k = 1; k < np; ++k) sum = sum + pc[k * plane + c];
  return d - sum;
}

// One Jacobi pass of the TE chi2 / chi3 fixed point (stepper_te.py:161-175):
// |E|^2 at Ex = Ex^2 + avg4(Ey)^2 and at Ey = Ey^2 + avg4(Ex)^2, the
// four-point averages with zero padding beyond the walls
// (stepper_te.py:52-61), all from the input pair (exi, eyi).
template <typename T>
__global__ __launch_bounds__(BX* BY) void te_kerr_kernel(TeK<T> p, const T* exi, const T* eyi, T* exo, T* eyo) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {  // Ex (nx-1, ny); Ey neighbours (i, j-1) (i, j) (i+1, j-1) (i+1, j)
    const long long c = (long long)i * ny + j;
    const T av = avg4_to_ex(eyi, nx, ny, i, j);
    const T e = exi[c];
    const T e2 = e * e + av * av;
    const T rhs = te_rhs(p.Dx[c], p.npx, (long long)(nx - 1) * ny, c, p.Pxc);
    exo[c] = rhs / (p.eps_x[c] + p.chi2x[c] * e + p.chi3x[c] * e2);
  }
  if (j < ny - 1) {  // Ey (nx, ny-1); Ex neighbours (i-1, j) (i-1, j+1) (i, j) (i, j+1)
    const long long c = (long long)i * (ny - 1) + j;
    const T av = avg4_to_ey(exi, nx, ny, i, j);
    const T e = eyi[c];
    const T e2 = e * e + av * av;
    const T rhs = te_rhs(p.Dy[c], p.npy, (long long)nx * (ny - 1), c, p.Pyc);
    eyo[c] = rhs / (p.eps_y[c] + p.chi2y[c] * e + p.chi3y[c] * e2);
  }
}

// rhs = D - sum P after advancing the pole stacks from the E before the step
// (stepper_te.py:137-147, 156-157); with oriented poles each P' also gains
// c3o * av, the other component's four-point average (c3o = 0 for the
// scalar poles); the stacks have `np` planes of `plane` cells.
template <typename T>
__device__ __forceinline__ T te_rhs_from_d(T d, T e_old, int np, long long plane, long long c, const T* c1,
                                           const T* c2, const T* c3, T* pc, T* pp, const T* c3o, T av) {
  T rhs = d;
  if (np) {
    T sum = T(0);
    for (int k = 0; k < np; ++k) {
      const long long q = k * plane + c;
      const T a = pc[q], b = pp[q];
      T pn = c1[q] * a + c2[q] * b + c3[q] * e_old;
      if (c3o) pn = pn + c3o[q] * av;
      pp[q] = a;
      pc[q] = pn;
      sum = k == 0 ? pn : sum + pn;
    }
    rhs = d - sum;
  }
  return rhs;
}

// Tensor eps (stepper_te.py:158-160): E = inv_xx rhs_x + inv_xy avg4(rhs_y)
// and E_y = inv_yy rhs_y + inv_xy avg4(rhs_x), from the rhs pair (rx, ry).
template <typename T>
__global__ __launch_bounds__(BX* BY) void te_tensor_kernel(TeK<T> p, const T* rx, const T* ry, T* exo, T* eyo) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {
    const long long c = (long long)i * ny + j;
    exo[c] = p.ixx[c] * rx[c] + p.ixyx[c] * avg4_to_ex(ry, nx, ny, i, j);
  }
  if (j < ny - 1) {
    const long long c = (long long)i * (ny - 1) + j;
    eyo[c] = p.iyy[c] * ry[c] + p.ixyy[c] * avg4_to_ey(rx, nx, ny, i, j);
  }
}

// meep_torch/core/stepper_te.py:80-103 (D half), :131-134 (source on Dy),
// :160-168 (E = D / eps)
template <typename T>
__global__ __launch_bounds__(BX* BY) void te_d_kernel(TeK<T> p, mt_source2d src, T coef, int s_rel) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {  // Dx (nx-1, ny)
    long long c = (long long)i * ny + j;
    T dxv = p.Dx[c];
    T curl = T(0);  // zero on the PEC rows, as the reference's padding
    if (j >= 1 && j < ny - 1) {
      T d = (p.Hz[(long long)i * (ny - 1) + j] - p.Hz[(long long)i * (ny - 1) + j - 1]) / p.dx;
      curl = d;
      int jj = j - 1;
      if (p.ey.in(jj)) {
        long long q = (long long)i * (ny - 2) + jj;
        T ps = p.ey.b[jj] * p.pHzy[q] + p.ey.a[jj] * d;
        p.pHzy[q] = ps;
        curl = d + ps;
      }
      if (!p.d1x) dxv = dxv + p.dt * curl;
    } else if (j == 0 && p.ymir == 1) {  // even y mirror: on-plane Ex live, odd Hz ghost (stepper_te.py:88-93)
      curl = (T(2) * p.Hz[(long long)i * (ny - 1)]) / p.dx;
      if (!p.d1x) dxv = dxv + p.dt * curl;
    }
    if (p.d1x) dxv = p.d1x[c] * dxv + p.d2x[c] * curl;  // stepper_te.py:96-99 (sigma_d)
    p.Dx[c] = dxv;
    const T av = p.c3xy ? avg4_to_ex(p.Ey, nx, ny, i, j) : T(0);  // E(n-1/2): the D half writes E to scratch
    const T rhs = te_rhs_from_d(dxv, p.npx ? p.Ex[c] : T(0), p.npx, (long long)(nx - 1) * ny, c, p.c1x, p.c2x,
                                p.c3x, p.Pxc, p.Pxp, p.c3xy, av);
    const T ex = p.ixx ? rhs : rhs / p.eps_x[c];  // tensor eps: E is formed from the rhs pair afterwards
    (p.scratch ? p.kx : p.Ex)[c] = ex;
    // under chi2/chi3 the tape keeps rhs: the adjoint recomputes the iterates from it
    if (p.tape) p.tape[(long long)s_rel * ((long long)(nx - 1) * ny + (long long)nx * (ny - 1)) + c] = p.kerr ? rhs : ex;
  }
  if (j < ny - 1) {  // Dy (nx, ny-1)
    long long c = (long long)i * (ny - 1) + j;
    T dyv = p.Dy[c];
    T curl = T(0);
    if (i >= 1 && i < nx - 1) {
      T d = (p.Hz[(long long)i * (ny - 1) + j] - p.Hz[(long long)(i - 1) * (ny - 1) + j]) / p.dx;
      curl = d;
      int ii = i - 1;
      if (p.ex.in(ii)) {
        long long q = (long long)ii * (ny - 1) + j;
        T ps = p.ex.b[ii] * p.pHzx[q] + p.ex.a[ii] * d;
        p.pHzx[q] = ps;
        curl = d + ps;
      }
      if (!p.d1y) dyv = dyv - p.dt * curl;
    } else if (i == 0 && p.xmir == 1) {  // even x mirror: on-plane Ey live (stepper_te.py:94-96)
      curl = (T(2) * p.Hz[j]) / p.dx;
      if (!p.d1y) dyv = dyv - p.dt * curl;
    }
    if (p.d1y) dyv = p.d1y[c] * dyv - p.d2y[c] * curl;
    if (src.present && i >= src.i0 && i < src.i1 && j >= src.j0 && j < src.j1) {
      const T* w = static_cast<const T*>(src.w);
      const T wv = w[(long long)(i - src.i0) * (src.j1 - src.j0) + (j - src.j0)];
      // stepper_te.py:131-134: under sigma_d the scale is d2y and coef the raw amplitude
      dyv = p.d2y ? dyv - (p.d2y[c] * coef) * wv : dyv - coef * wv;
    }
    p.Dy[c] = dyv;
    const T av = p.c3yx ? avg4_to_ey(p.Ex, nx, ny, i, j) : T(0);
    const T rhs = te_rhs_from_d(dyv, p.npy ? p.Ey[c] : T(0), p.npy, (long long)nx * (ny - 1), c, p.c1y, p.c2y,
                                p.c3y, p.Pyc, p.Pyp, p.c3yx, av);
    const T ey = p.iyy ? rhs : rhs / p.eps_y[c];
    (p.scratch ? p.ky : p.Ey)[c] = ey;
    if (p.tape)
      p.tape[(long long)s_rel * ((long long)(nx - 1) * ny + (long long)nx * (ny - 1)) + (long long)(nx - 1) * ny + c] =
          p.kerr ? rhs : ey;
  }
}

// meep_torch/core/stepper_te.py:106-117 (B half, Hz from curl E)
template <typename T>
__global__ __launch_bounds__(BX* BY) void te_b_kernel(TeK<T> p, mt_source2d hzs, T hc) {
  int j = blockIdx.x * BX + threadIdx.x;
  int i = blockIdx.y * BY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx - 1 || j >= ny - 1) return;
  long long h = (long long)i * (ny - 1) + j;
  T dEy = (p.Ey[(long long)(i + 1) * (ny - 1) + j] - p.Ey[h]) / p.dx;
  T dEx = (p.Ex[(long long)i * ny + j + 1] - p.Ex[(long long)i * ny + j]) / p.dx;
  T a1 = dEy;
  if (p.hx.in(i)) {
    T ps = p.hx.b[i] * p.pEyx[h] + p.hx.a[i] * dEy;
    p.pEyx[h] = ps;
    a1 = dEy + ps;
  }
  T curl = a1 - dEx;
  if (p.hy.in(j)) {
    T ps = p.hy.b[j] * p.pExy[h] + p.hy.a[j] * dEx;
    p.pExy[h] = ps;
    curl = curl - ps;
  }
  // stepper_te.py:110-115: sigma_b / mu fold into hz1, hz2
  T v = p.hz1 ? p.hz1[h] * p.Hz[h] - p.hz2[h] * curl : p.Hz[h] - p.dt * curl;
  if (in_box<T>(hzs, i, j)) v = v - hc * box_w<T>(hzs, i, j);  // stepper_te.py:120-128
  p.Hz[h] = v;
}

template <typename T>
int tm_run_t(const mt_tm2d* d, const mt_source2d* src, const double* coef, const MonitorPlan& mp,
             int step0, int k, hipStream_t s) {
  TmK<T> p;
  p.nx = d->nx;
  p.ny = d->ny;
  p.y_pmc = d->y_pmc;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.eps = static_cast<const T*>(d->eps);
  p.hy = pml_of<T>(d->hy);
  p.hx = pml_of<T>(d->hx);
  p.ex = pml_of<T>(d->ex);
  p.ey = pml_of<T>(d->ey);
  p.Ez = static_cast<T*>(d->Ez);
  p.Dz = static_cast<T*>(d->Dz);
  p.Hx = static_cast<T*>(d->Hx);
  p.Hy = static_cast<T*>(d->Hy);
  p.pEzy = static_cast<T*>(d->psi_Ezy);
  p.pEzx = static_cast<T*>(d->psi_Ezx);
  p.pHyx = static_cast<T*>(d->psi_Hyx);
  p.pHxy = static_cast<T*>(d->psi_Hxy);
  p.ny_e = d->y_pmc ? d->ny : d->ny - 2;
  p.tape = static_cast<T*>(d->tape);
  p.n_poles = d->n_poles;
  p.kerr = d->kerr;
  p.c1 = static_cast<const T*>(d->c1);
  p.c2 = static_cast<const T*>(d->c2);
  p.c3 = static_cast<const T*>(d->c3);
  p.Pc = static_cast<T*>(d->P_cur);
  p.Pp = static_cast<T*>(d->P_prev);
  p.chi2 = static_cast<const T*>(d->chi2);
  p.chi3 = static_cast<const T*>(d->chi3);
  p.d1 = static_cast<const T*>(d->d1);
  p.d2 = static_cast<const T*>(d->d2);
  p.h1x = static_cast<const T*>(d->h1x);
  p.h2x = static_cast<const T*>(d->h2x);
  p.h1y = static_cast<const T*>(d->h1y);
  p.h2y = static_cast<const T*>(d->h2y);
  p.n_mpoles = d->n_mpoles;
  p.mc1x = static_cast<const T*>(d->mc1x);
  p.mc2x = static_cast<const T*>(d->mc2x);
  p.mc3x = static_cast<const T*>(d->mc3x);
  p.mc1y = static_cast<const T*>(d->mc1y);
  p.mc2y = static_cast<const T*>(d->mc2y);
  p.mc3y = static_cast<const T*>(d->mc3y);
  p.Bx = static_cast<T*>(d->Bx);
  p.By = static_cast<T*>(d->By);
  p.Mxc = static_cast<T*>(d->Mx_cur);
  p.Mxp = static_cast<T*>(d->Mx_prev);
  p.Myc = static_cast<T*>(d->My_cur);
  p.Myp = static_cast<T*>(d->My_prev);
  p.ycut = d->y_mirror == 1 ? 1 : (d->y_mirror == -1 || d->y_rot2 || d->rot4) ? 2 : 0;
  p.xcut = d->x_mirror == 1 ? 1 : (d->x_mirror == -1 || d->rot4) ? 2 : 0;
  p.ymir = d->y_mirror;
  p.y_rot2 = d->y_rot2;
  p.rot4 = d->rot4;
  p.Hx0 = static_cast<T*>(d->Hx0);
  p.Hy0 = static_cast<T*>(d->Hy0);
  p.ny_e = d->y_pmc ? d->ny : (p.ycut ? d->ny - 1 : d->ny - 2);
  mt_source2d sv{};
  if (src && src->present) sv = *src;
  dim3 blk(BX, BY), grd = grid2(p.ny, p.nx, blk);
  int S = mp.stage_steps > 0 ? mp.stage_steps : 1;
  int filled = 0, first = step0;
  for (int s_ = 0; s_ < k; ++s_) {
    const T hc = d->hcoef ? static_cast<T>(d->hcoef[s_]) : T(0);
    hipLaunchKernelGGL(tm_b_kernel<T>, grd, blk, 0, s, p, d->hsrc[0], d->hsrc[1], hc);
    T c = sv.present ? static_cast<T>(coef[s_]) : T(0);
    hipLaunchKernelGGL(tm_d_kernel<T>, grd, blk, 0, s, p, step_source<T>(sv, s_), c, s_);
    int rc = check_launch("tm2d step");
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

template <typename T>
int te_run_t(const mt_te2d* d, const mt_source2d* src, const double* coef, const MonitorPlan& mp,
             int step0, int k, hipStream_t s) {
  TeK<T> p;
  p.nx = d->nx;
  p.ny = d->ny;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.eps_x = static_cast<const T*>(d->eps_x);
  p.eps_y = static_cast<const T*>(d->eps_y);
  p.hx = pml_of<T>(d->hx);
  p.hy = pml_of<T>(d->hy);
  p.ey = pml_of<T>(d->ey);
  p.ex = pml_of<T>(d->ex);
  p.Hz = static_cast<T*>(d->Hz);
  p.Ex = static_cast<T*>(d->Ex);
  p.Ey = static_cast<T*>(d->Ey);
  p.Dx = static_cast<T*>(d->Dx);
  p.Dy = static_cast<T*>(d->Dy);
  p.pEyx = static_cast<T*>(d->psi_Eyx);
  p.pExy = static_cast<T*>(d->psi_Exy);
  p.pHzy = static_cast<T*>(d->psi_Hzy);
  p.pHzx = static_cast<T*>(d->psi_Hzx);
  p.tape = static_cast<T*>(d->tape);
  p.npx = d->npx;
  p.npy = d->npy;
  p.c1x = static_cast<const T*>(d->c1x);
  p.c2x = static_cast<const T*>(d->c2x);
  p.c3x = static_cast<const T*>(d->c3x);
  p.c1y = static_cast<const T*>(d->c1y);
  p.c2y = static_cast<const T*>(d->c2y);
  p.c3y = static_cast<const T*>(d->c3y);
  p.Pxc = static_cast<T*>(d->Px_cur);
  p.Pxp = static_cast<T*>(d->Px_prev);
  p.Pyc = static_cast<T*>(d->Py_cur);
  p.Pyp = static_cast<T*>(d->Py_prev);
  p.d1x = static_cast<const T*>(d->d1x);
  p.d2x = static_cast<const T*>(d->d2x);
  p.d1y = static_cast<const T*>(d->d1y);
  p.d2y = static_cast<const T*>(d->d2y);
  p.hz1 = static_cast<const T*>(d->hz1);
  p.hz2 = static_cast<const T*>(d->hz2);
  p.kerr = d->kerr;
  p.chi2x = static_cast<const T*>(d->chi2x);
  p.chi3x = static_cast<const T*>(d->chi3x);
  p.chi2y = static_cast<const T*>(d->chi2y);
  p.chi3y = static_cast<const T*>(d->chi3y);
  p.kx = static_cast<T*>(d->kx);
  p.ky = static_cast<T*>(d->ky);
  p.ixx = static_cast<const T*>(d->inv_xx_x);
  p.ixyx = static_cast<const T*>(d->inv_xy_x);
  p.iyy = static_cast<const T*>(d->inv_yy_y);
  p.ixyy = static_cast<const T*>(d->inv_xy_y);
  p.c3xy = static_cast<const T*>(d->c3xy);
  p.c3yx = static_cast<const T*>(d->c3yx);
  p.scratch = p.kerr || p.ixx || p.c3xy;
  p.ymir = d->y_mirror;
  p.xmir = d->x_mirror;
  const size_t bx = (size_t)(p.nx - 1) * p.ny * sizeof(T), by = (size_t)p.nx * (p.ny - 1) * sizeof(T);
  mt_source2d sv{};
  if (src && src->present) sv = *src;
  dim3 blk(BX, BY), grd = grid2(p.ny, p.nx, blk);
  int S = mp.stage_steps > 0 ? mp.stage_steps : 1;
  int filled = 0, first = step0;
  for (int s_ = 0; s_ < k; ++s_) {
    T c = sv.present ? static_cast<T>(coef[s_]) : T(0);
    hipLaunchKernelGGL(te_d_kernel<T>, grd, blk, 0, s, p, step_source<T>(sv, s_), c, s_);
    if (p.ixx) {  // tensor eps: E from the rhs pair in scratch; the chi2/chi3 passes start from it
      hipLaunchKernelGGL(te_tensor_kernel<T>, grd, blk, 0, s, p, p.kx, p.ky, p.Ex, p.Ey);
      if (p.kerr && (hipMemcpyAsync(p.kx, p.Ex, bx, hipMemcpyDeviceToDevice, s) != hipSuccess ||
                     hipMemcpyAsync(p.ky, p.Ey, by, hipMemcpyDeviceToDevice, s) != hipSuccess))
        return fail(MT_EHIP, "te2d: scratch copy failed");
    } else if (p.scratch && !p.kerr) {  // oriented poles only: E was staged in scratch
      if (hipMemcpyAsync(p.Ex, p.kx, bx, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(p.Ey, p.ky, by, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(MT_EHIP, "te2d: scratch copy failed");
    }
    if (p.kerr) {  // three passes: scratch -> E -> scratch -> E
      hipLaunchKernelGGL(te_kerr_kernel<T>, grd, blk, 0, s, p, p.kx, p.ky, p.Ex, p.Ey);
      hipLaunchKernelGGL(te_kerr_kernel<T>, grd, blk, 0, s, p, p.Ex, p.Ey, p.kx, p.ky);
      hipLaunchKernelGGL(te_kerr_kernel<T>, grd, blk, 0, s, p, p.kx, p.ky, p.Ex, p.Ey);
    }
    const T hc = d->hcoef ? static_cast<T>(d->hcoef[s_]) : T(0);
    hipLaunchKernelGGL(te_b_kernel<T>, grd, blk, 0, s, p, d->hsrc, hc);
    int rc = check_launch("te2d step");
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

// Cells whose tiles carry no CPML psi: rows [i0, i1) x columns [j0, j1).
void tm_interior(const mt_tm2d* d, int* i0, int* i1, int* j0, int* j1) {
  const int joff = d->y_pmc ? 0 : 1;
  *i0 = std::max(d->hx.lo_end, d->ex.lo_end + 1) + 1;
  *i1 = std::min(d->hx.hi_begin, d->ex.hi_begin + 1) - 1;
  *j0 = std::max(d->hy.lo_end, d->ey.lo_end + joff) + 1;
  *j1 = std::min(d->hy.hi_begin, d->ey.hi_begin + joff) - 1;
}

void te_interior(const mt_te2d* d, int* i0, int* i1, int* j0, int* j1) {
  *i0 = std::max(d->hx.lo_end, d->ex.lo_end + 1) + 1;
  *i1 = std::min(d->hx.hi_begin, d->ex.hi_begin + 1) - 1;
  *j0 = std::max(d->hy.lo_end, d->ey.lo_end + 1) + 1;
  *j1 = std::min(d->hy.hi_begin, d->ey.hi_begin + 1) - 1;
}

int tile_counts(int variant, int dtype, int nx, int ny, int i0, int i1, int j0, int j1, int* n_in, int* n_ring) {
  if (variant < 0 || variant >= tiled_variant_count()) return fail(MT_EINVAL, "unknown tiled variant");
  const TileMap m = tiled_map(variant, dtype, nx, ny, i0, i1, j0, j1);
  const int in = (m.r1 - m.r0) * (m.c1 - m.c0);
  if (n_in) *n_in = in;
  if (n_ring) *n_ring = m.R * m.C - in;
  return MT_OK;
}

template <typename T, typename P>
int tm_run_tiled(const mt_tm2d* d, const mt_source2d* src, const double* coef, const MonitorPlan& mp,
                 int step0, int k, hipStream_t s) {
  TmTiled<T> p{};
  p.nx = d->nx;
  p.ny = d->ny;
  p.y_pmc = d->y_pmc;
  p.ny_e = d->y_pmc ? d->ny : d->ny - 2;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.rdx = const_div_recip<T>(p.dx);
  p.eps = static_cast<const T*>(d->eps);
  p.hy = pml_of<T>(d->hy);
  p.hx = pml_of<T>(d->hx);
  p.ex = pml_of<T>(d->ex);
  p.ey = pml_of<T>(d->ey);
  tm_interior(d, &p.in_i0, &p.in_i1, &p.in_j0, &p.in_j1);
  T* A[8] = {static_cast<T*>(d->Ez), static_cast<T*>(d->Dz), static_cast<T*>(d->Hx), static_cast<T*>(d->Hy),
             static_cast<T*>(d->psi_Ezy), static_cast<T*>(d->psi_Ezx), static_cast<T*>(d->psi_Hyx),
             static_cast<T*>(d->psi_Hxy)};
  T* B[8];
  for (int q = 0; q < 8; ++q) B[q] = static_cast<T*>(d->alt[q]);
  size_t nbytes[8] = {(size_t)d->nx * d->ny, (size_t)d->nx * d->ny, (size_t)d->nx * (d->ny - 1),
                      (size_t)(d->nx - 1) * d->ny, (size_t)d->nx * (d->ny - 1), (size_t)(d->nx - 1) * d->ny,
                      (size_t)(d->nx - 2) * p.ny_e, (size_t)(d->nx - 2) * p.ny_e};
  p.tape = static_cast<T*>(d->tape);
  if (src && src->present) p.src = *src;
  p.hsrc[0] = d->hsrc[0];
  p.hsrc[1] = d->hsrc[1];
  const bool hsrc = d->hsrc[0].present || d->hsrc[1].present;
  const int NS = tiled_variant_steps(d->variant, d->dtype);
  const int S = mp.n_mons ? mp.stage_steps : k;
  int done = 0, slot = 0, first = step0, cur = 0, rc;
  while (done < k) {
    int nsub = std::min(NS, k - done);
    if (mp.n_mons) nsub = std::min(nsub, S - slot);
    T** in = cur ? B : A;
    T** out = cur ? A : B;
    p.Ez = in[0], p.Dz = in[1], p.Hx = in[2], p.Hy = in[3];
    p.pEzy = in[4], p.pEzx = in[5], p.pHyx = in[6], p.pHxy = in[7];
    p.oEz = out[0], p.oDz = out[1], p.oHx = out[2], p.oHy = out[3];
    p.oEzy = out[4], p.oEzx = out[5], p.oHyx = out[6], p.oHxy = out[7];
    for (int u = 0; u < kMaxSub; ++u) p.coef[u] = (p.src.present && u < nsub) ? static_cast<T>(coef[done + u]) : T(0);
    for (int u = 0; u < kMaxSub; ++u) p.hcoef[u] = (hsrc && u < nsub) ? static_cast<T>(d->hcoef[done + u]) : T(0);
    if (src && src->present) p.src = step_source<T>(*src, done);
    TapSet taps;
    taps.n = 0;
    if (mp.n_mons && !build_taps(mp, slot, done, sizeof(T), &taps)) return fail(MT_EINVAL, "tap build failed");
    rc = P::tm(d->variant, p, taps, nsub, done, s);
    if (rc) return rc;
    cur ^= 1;
    done += nsub;
    slot += nsub;
    if (mp.n_mons && (slot == S || done == k)) {
      if ((rc = monitors_flush(mp, slot, first, d->dt, s))) return rc;
      first += slot;
      slot = 0;
    }
  }
  if (cur)
    for (int q = 0; q < 8; ++q)
      if (hipMemcpyAsync(A[q], B[q], nbytes[q] * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(MT_EHIP, "tiled TM: copy back failed");
  return MT_OK;
}

template <typename T, typename P>
int te_run_tiled(const mt_te2d* d, const mt_source2d* src, const double* coef, const MonitorPlan& mp,
                 int step0, int k, hipStream_t s) {
  TeTiled<T> p{};
  p.nx = d->nx;
  p.ny = d->ny;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.rdx = const_div_recip<T>(p.dx);
  p.eps_x = static_cast<const T*>(d->eps_x);
  p.eps_y = static_cast<const T*>(d->eps_y);
  p.hx = pml_of<T>(d->hx);
  p.hy = pml_of<T>(d->hy);
  p.ey = pml_of<T>(d->ey);
  p.ex = pml_of<T>(d->ex);
  te_interior(d, &p.in_i0, &p.in_i1, &p.in_j0, &p.in_j1);
  T* A[7] = {static_cast<T*>(d->Hz), static_cast<T*>(d->Dx), static_cast<T*>(d->Dy),
             static_cast<T*>(d->psi_Eyx), static_cast<T*>(d->psi_Exy), static_cast<T*>(d->psi_Hzy),
             static_cast<T*>(d->psi_Hzx)};
  T* B[7];
  for (int q = 0; q < 7; ++q) B[q] = static_cast<T*>(d->alt[q]);
  const size_t nx = d->nx, ny = d->ny;
  size_t nel[7] = {(nx - 1) * (ny - 1), (nx - 1) * ny, nx * (ny - 1), (nx - 1) * (ny - 1), (nx - 1) * (ny - 1),
                   (nx - 1) * (ny - 2), (nx - 2) * (ny - 1)};
  p.Ex = static_cast<T*>(d->Ex);
  p.Ey = static_cast<T*>(d->Ey);
  if (src && src->present) p.src = *src;
  const int NS = tiled_variant_steps(d->variant, d->dtype);
  const int S = mp.n_mons ? mp.stage_steps : k;
  int done = 0, slot = 0, first = step0, cur = 0, rc;
  while (done < k) {
    int nsub = std::min(NS, k - done);
    if (mp.n_mons) nsub = std::min(nsub, S - slot);
    T** in = cur ? B : A;
    T** out = cur ? A : B;
    p.Hz = in[0], p.Dx = in[1], p.Dy = in[2], p.pEyx = in[3], p.pExy = in[4], p.pHzy = in[5], p.pHzx = in[6];
    p.oHz = out[0], p.oDx = out[1], p.oDy = out[2], p.oEyx = out[3], p.oExy = out[4], p.oHzy = out[5],
    p.oHzx = out[6];
    for (int u = 0; u < kMaxSub; ++u) p.coef[u] = (p.src.present && u < nsub) ? static_cast<T>(coef[done + u]) : T(0);
    if (src && src->present) p.src = step_source<T>(*src, done);
    TapSet taps;
    taps.n = 0;
    if (mp.n_mons && !build_taps(mp, slot, done, sizeof(T), &taps)) return fail(MT_EINVAL, "tap build failed");
    rc = P::te(d->variant, p, taps, nsub, done + nsub == k ? 1 : 0, s);
    if (rc) return rc;
    cur ^= 1;
    done += nsub;
    slot += nsub;
    if (mp.n_mons && (slot == S || done == k)) {
      if ((rc = monitors_flush(mp, slot, first, d->dt, s))) return rc;
      first += slot;
      slot = 0;
    }
  }
  if (cur)
    for (int q = 0; q < 7; ++q)
      if (hipMemcpyAsync(A[q], B[q], nel[q] * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(MT_EHIP, "tiled TE: copy back failed");
  return MT_OK;
}

struct DispF32 {
  static int tm(int v, const TmTiled<float>& p, const TapSet& t, int n, int s0, hipStream_t s) {
    return tm_tiled_dispatch_f32(v, p, t, n, s0, s);
  }
  static int te(int v, const TeTiled<float>& p, const TapSet& t, int n, int we, hipStream_t s) {
    return te_tiled_dispatch_f32(v, p, t, n, we, s);
  }
};
struct DispF64 {
  static int tm(int v, const TmTiled<double>& p, const TapSet& t, int n, int s0, hipStream_t s) {
    return tm_tiled_dispatch_f64(v, p, t, n, s0, s);
  }
  static int te(int v, const TeTiled<double>& p, const TapSet& t, int n, int we, hipStream_t s) {
    return te_tiled_dispatch_f64(v, p, t, n, we, s);
  }
};

bool taps_ok(const MonitorPlan& mp) {
  for (int i = 0; i < mp.n_mons; ++i)
    for (int q = 0; q < mp.mons[i].n_terms; ++q)
      if (mp.mons[i].term[q].line == 0) return false;
  return true;
}

bool alt_ok(void* const* alt, int n) {
  for (int q = 0; q < n; ++q)
    if (!alt[q]) return false;
  return true;
}

}  // namespace
}  // namespace mt

using namespace mt;

extern "C" int mt_tm2d_run(const mt_tm2d* p, const mt_source2d* src, const double* coef,
                           const mt_monitor* mons, int n_mons, void* stage, long long stage_bytes,
                           int stage_steps, int step0, int k, void* stream) {
  if (!p) return fail(MT_EINVAL, "null problem");
  if (p->nx < 3 || p->ny < 3) return fail(MT_EINVAL, "grid must be at least 3x3");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  if (p->dtype != MT_F32 && p->dtype != MT_F64) return fail(MT_EINVAL, "dtype must be f32 or f64");
  if (!p->eps || !p->Ez || !p->Dz || !p->Hx || !p->Hy || !p->psi_Ezy || !p->psi_Ezx || !p->psi_Hyx ||
      !p->psi_Hxy)
    return fail(MT_EINVAL, "null field pointer");
  int rc;
  const int cuts[4] = {p->y_mirror, p->x_mirror, p->y_rot2, p->rot4};
  for (int q = 0; q < 4; ++q)
    if (cuts[q] < -1 || cuts[q] > 1) return fail(MT_EINVAL, "symmetry parity must be 0/+1/-1");
  if (p->y_mirror && p->y_rot2) return fail(MT_EINVAL, "y_mirror and y_rot2 both cut the low-y edge");
  if (p->rot4 && (p->y_mirror || p->x_mirror || p->y_rot2)) return fail(MT_EINVAL, "rot4 cuts both low edges on its own");
  if (p->rot4 && p->nx != p->ny) return fail(MT_EINVAL, "rot4 needs a square grid");
  const bool ycut = p->y_mirror || p->y_rot2 || p->rot4, xcut = p->x_mirror || p->rot4;
  if (p->y_pmc && (ycut || xcut)) return fail(MT_EINVAL, "symmetry cuts and y_pmc are mutually exclusive");
  if ((p->y_mirror == -1 || p->y_rot2 || p->rot4) && !p->Hx0) return fail(MT_EINVAL, "odd / Rotate y cut needs Hx0");
  if ((p->x_mirror == -1 || p->rot4) && !p->Hy0) return fail(MT_EINVAL, "odd / Rotate4 x cut needs Hy0");
  if ((p->y_mirror == -1 || p->x_mirror == -1 || p->y_rot2 || p->rot4) && (p->h1x || p->n_mpoles))
    return fail(MT_EINVAL, "odd mirror / Rotate2 / Rotate4 with sigma_b/mu/magnetic poles");
  if ((rc = check_axis(p->hy, p->ny - 1, "hy")) || (rc = check_axis(p->hx, p->nx - 1, "hx")) ||
      (rc = check_axis(p->ex, xcut ? p->nx - 1 : p->nx - 2, "ex")) ||
      (rc = check_axis(p->ey, p->y_pmc ? p->ny : (ycut ? p->ny - 1 : p->ny - 2), "ey")))
    return rc;
  if ((rc = check_src(src, p->nx, p->ny))) return rc;
  if (src && src->present && !coef && k > 0) return fail(MT_EINVAL, "source without coefficients");
  if ((rc = check_src(&p->hsrc[0], p->nx, p->ny - 1)) || (rc = check_src(&p->hsrc[1], p->nx - 1, p->ny))) return rc;
  const bool hsrc = p->hsrc[0].present || p->hsrc[1].present;
  if (hsrc && !p->hcoef && k > 0) return fail(MT_EINVAL, "magnetic source without coefficients");
  if (p->n_poles < 0) return fail(MT_EINVAL, "negative pole count");
  if (p->n_poles && (!p->c1 || !p->c2 || !p->c3 || !p->P_cur || !p->P_prev))
    return fail(MT_EINVAL, "poles without coefficients or pole stacks");
  if (p->kerr && (!p->chi2 || !p->chi3)) return fail(MT_EINVAL, "chi2/chi3 fixed point without chi2 and chi3");
  if (!p->d1 != !p->d2) return fail(MT_EINVAL, "sigma_d needs both d1 and d2");
  if (!p->h1x != !p->h2x || !p->h1x != !p->h1y || !p->h1x != !p->h2y)
    return fail(MT_EINVAL, "sigma_b/mu needs h1x, h2x, h1y and h2y");
  if (p->n_mpoles < 0) return fail(MT_EINVAL, "negative magnetic pole count");
  if (p->n_mpoles) {
    if (p->h1x) return fail(MT_EINVAL, "magnetic poles are not combined with sigma_b/mu (h1x..h2y)");
    if (!p->mc1x || !p->mc2x || !p->mc3x || !p->mc1y || !p->mc2y || !p->mc3y || !p->Bx || !p->By ||
        !p->Mx_cur || !p->Mx_prev || !p->My_cur || !p->My_prev)
      return fail(MT_EINVAL, "magnetic poles need mc1..mc3 x/y, Bx, By and the M stacks");
  }
  const bool media = p->n_poles || p->kerr || p->d1 || p->h1x || p->n_mpoles || ycut || xcut;  // two-pass kernels
  MonitorPlan mp{mons, n_mons, stage, stage_bytes, stage_steps, p->dtype};
  if (n_mons && (rc = monitors_validate(mp))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p->variant >= 0 && !media) {   // media: two-pass kernels
    if (p->variant >= tiled_variant_count()) return fail(MT_EINVAL, "unknown tiled variant");
    if (!alt_ok(p->alt, 8)) return fail(MT_EINVAL, "tiled path needs the alt buffers");
    if (taps_ok(mp))
      return p->dtype == MT_F64 ? tm_run_tiled<double, DispF64>(p, src, coef, mp, step0, k, s)
                                : tm_run_tiled<float, DispF32>(p, src, coef, mp, step0, k, s);
  }
  return p->dtype == MT_F64 ? tm_run_t<double>(p, src, coef, mp, step0, k, s)
                            : tm_run_t<float>(p, src, coef, mp, step0, k, s);
}

extern "C" int mt_te2d_run(const mt_te2d* p, const mt_source2d* src, const double* coef,
                           const mt_monitor* mons, int n_mons, void* stage, long long stage_bytes,
                           int stage_steps, int step0, int k, void* stream) {
  if (!p) return fail(MT_EINVAL, "null problem");
  if (p->nx < 3 || p->ny < 3) return fail(MT_EINVAL, "grid must be at least 3x3");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  if (p->dtype != MT_F32 && p->dtype != MT_F64) return fail(MT_EINVAL, "dtype must be f32 or f64");
  if (!p->eps_x || !p->eps_y || !p->Hz || !p->Ex || !p->Ey || !p->Dx || !p->Dy || !p->psi_Eyx ||
      !p->psi_Exy || !p->psi_Hzy || !p->psi_Hzx)
    return fail(MT_EINVAL, "null field pointer");
  int rc;
  if ((rc = check_axis(p->hx, p->nx - 1, "hx")) || (rc = check_axis(p->hy, p->ny - 1, "hy")) ||
      (rc = check_axis(p->ey, p->ny - 2, "ey")) || (rc = check_axis(p->ex, p->nx - 2, "ex")))
    return rc;
  if ((rc = check_src(src, p->nx, p->ny - 1))) return rc;
  if (src && src->present && !coef && k > 0) return fail(MT_EINVAL, "source without coefficients");
  if ((rc = check_src(&p->hsrc, p->nx - 1, p->ny - 1))) return rc;
  if (p->hsrc.present && !p->hcoef && k > 0) return fail(MT_EINVAL, "magnetic source without coefficients");
  if (p->npx < 0 || p->npy < 0) return fail(MT_EINVAL, "negative pole count");
  if (p->npx && (!p->c1x || !p->c2x || !p->c3x || !p->Px_cur || !p->Px_prev))
    return fail(MT_EINVAL, "Ex poles without coefficients or pole stacks");
  if (p->npy && (!p->c1y || !p->c2y || !p->c3y || !p->Py_cur || !p->Py_prev))
    return fail(MT_EINVAL, "Ey poles without coefficients or pole stacks");
  if (!p->d1x != !p->d2x || !p->d1x != !p->d1y || !p->d1x != !p->d2y)
    return fail(MT_EINVAL, "sigma_d needs d1x, d2x, d1y and d2y");
  if (!p->hz1 != !p->hz2) return fail(MT_EINVAL, "sigma_b/mu needs hz1 and hz2");
  if (p->kerr && (!p->chi2x || !p->chi3x || !p->chi2y || !p->chi3y || !p->kx || !p->ky))
    return fail(MT_EINVAL, "TE chi2/chi3 fixed point needs chi2x, chi3x, chi2y, chi3y and the kx, ky scratch");
  const bool tensor = p->inv_xx_x || p->inv_xy_x || p->inv_yy_y || p->inv_xy_y;
  if (tensor && !(p->inv_xx_x && p->inv_xy_x && p->inv_yy_y && p->inv_xy_y))
    return fail(MT_EINVAL, "tensor eps needs inv_xx_x, inv_xy_x, inv_yy_y and inv_xy_y");
  if (!p->c3xy != !p->c3yx) return fail(MT_EINVAL, "oriented poles need both c3xy and c3yx");
  if (p->c3xy && (!p->npx || !p->npy)) return fail(MT_EINVAL, "oriented poles need Ex and Ey pole stacks");
  if ((tensor || p->c3xy) && (!p->kx || !p->ky))
    return fail(MT_EINVAL, "tensor eps and oriented poles need the kx, ky scratch");
  if (p->y_mirror < -1 || p->y_mirror > 1 || p->x_mirror < -1 || p->x_mirror > 1)
    return fail(MT_EINVAL, "mirror parity must be 0/+1/-1");
  const bool media = p->npx || p->npy || p->d1x || p->hz1 || p->kerr || tensor || p->c3xy || p->y_mirror ||
                     p->x_mirror;  // two-pass kernels
  MonitorPlan mp{mons, n_mons, stage, stage_bytes, stage_steps, p->dtype};
  if (n_mons && (rc = monitors_validate(mp))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the tape and magnetic currents are handled by the two-pass kernels
  if (p->variant >= 0 && !p->tape && !p->hsrc.present && !media) {
    if (p->variant >= tiled_variant_count()) return fail(MT_EINVAL, "unknown tiled variant");
    if (!alt_ok(p->alt, 7)) return fail(MT_EINVAL, "tiled path needs the alt buffers");
    if (taps_ok(mp))
      return p->dtype == MT_F64 ? te_run_tiled<double, DispF64>(p, src, coef, mp, step0, k, s)
                                : te_run_tiled<float, DispF32>(p, src, coef, mp, step0, k, s);
  }
  return p->dtype == MT_F64 ? te_run_t<double>(p, src, coef, mp, step0, k, s)
                            : te_run_t<float>(p, src, coef, mp, step0, k, s);
}

extern "C" int mt_tm2d_tiles(const mt_tm2d* p, int* n_interior, int* n_ring) {
  if (!p) return fail(MT_EINVAL, "null problem");
  int i0, i1, j0, j1;
  tm_interior(p, &i0, &i1, &j0, &j1);
  return tile_counts(p->variant, p->dtype, p->nx, p->ny, i0, i1, j0, j1, n_interior, n_ring);
}

extern "C" int mt_te2d_tiles(const mt_te2d* p, int* n_interior, int* n_ring) {
  if (!p) return fail(MT_EINVAL, "null problem");
  int i0, i1, j0, j1;
  te_interior(p, &i0, &i1, &j0, &j1);
  return tile_counts(p->variant, p->dtype, p->nx, p->ny, i0, i1, j0, j1, n_interior, n_ring);
}

extern "C" float mt_div_recip_f32(float b) { return div_recip_f32(b); }

// ---------------------------------------------------------------------------
// Allocation-free variant of mt_tm2d_run: the DFT staging ring lives in a
// file-scope device array instead of a caller workspace.  Every GPU the
// library runs on holds one copy of the array until the process ends; its
// device address is looked up once per device (hipGetSymbolAddress) and
// used for every launch -- the array's name is a host-side stand-in and is
// never handed to a kernel.  The ring needs no clearing: each slot is
// written by the gather kernel before the flush reads it, and the phase
// table is rebuilt at every flush.  Calls that overlap on one GPU share the
// array and race.
// ---------------------------------------------------------------------------
namespace mt {
namespace {
constexpr long long kStaticStageBytes = 16LL << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[kStaticStageBytes];
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(MT_EHIP, "static staging: no current device");
  if (dev < 0 || dev >= kMaxDevices) return fail(MT_EINVAL, "static staging: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(MT_EHIP, std::string("static staging: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return MT_OK;
}
}  // namespace
}  // namespace mt

extern "C" long long mt_static_stage_bytes(void) { return mt::kStaticStageBytes; }

extern "C" int mt_tm2d_run_static(const mt_tm2d* p, const mt_source2d* src, const double* coef,
                                  const mt_monitor* mons, int n_mons, int step0, int k, void* stream) {
  if (!p) return fail(MT_EINVAL, "null problem");
  if (p->dtype != MT_F32 && p->dtype != MT_F64) return fail(MT_EINVAL, "dtype must be f32 or f64");
  void* stage = nullptr;
  long long bytes = 0;
  int steps = 1;
  if (n_mons > 0) {
    if (!mons) return fail(MT_EINVAL, "null monitor array");
    const long long per_step = mt_stage_bytes(mons, n_mons, 1, p->dtype);
    if (per_step > kStaticStageBytes)
      return fail(MT_EINVAL, "monitors need " + std::to_string(per_step) + " bytes of staging per step, above the " +
                                 std::to_string(kStaticStageBytes) + "-byte static staging array");
    steps = per_step > 0 ? (int)std::min<long long>(64, kStaticStageBytes / per_step) : 64;
    bytes = mt_stage_bytes(mons, n_mons, steps, p->dtype);
    int rc = scratch_address(&stage);
    if (rc) return rc;
  }
  return mt_tm2d_run(p, src, coef, mons, n_mons, stage, bytes, steps, step0, k, stream);
}
