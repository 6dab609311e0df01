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
__global__ __launch_bounds__(ABX* ABY) void adj_a_kernel(AdjK<T> p, const T* ez_after, int s_rel) {
  int j = blockIdx.x * ABX + threadIdx.x;
  int i = blockIdx.y * ABY + threadIdx.y;
  if (i < p.nx && j < p.ny) adj_a_cell(p, ez_after, i, j, s_rel);
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
    hipLaunchKernelGGL(adj_a_kernel<T>, grd, blk, 0, s, p, tp + (long long)s_ * cells, s_);
    hipLaunchKernelGGL(adj_b_kernel<T>, grd, blk, 0, s, p);
    hipLaunchKernelGGL(adj_c_kernel<T>, grd, blk, 0, s, p);
    int rc = check_launch("tm2d adjoint step");
    if (rc) return rc;
  }
  return MT_OK;
}

// ------------------------------------------------------------------- TE
// Reverse step of the TE Yee step (stepper_te.py:73-190): forward step t
//   D half  gy = Dy Hz, psi_Hzy' = b psi + a gy, Dx' = Dx + dt (gy + psi')
//           gx = Dx Hz, psi_Hzx' = b psi + a gx, Dy' = Dy - dt (gx + psi') - c_t mask
//   E       Ex' = Dx' / eps_x, Ey' = Dy' / eps_y
//   B half  Hz' = Hz - dt ((dEy + psi_Eyx') - dEx - psi_Exy')
// and the monitors sample (Ex', Ey', Hz') after the step.  After the
// monitor injection, three passes give the exact transpose:
//   P1 (Hz cells)  curl cotangent -> B-half psi cotangents and the
//                  E-difference cotangents gEy, gEx
//   P2 (E/D cells) Ex', Ey' cotangents (transposed differences of gEx, gEy)
//                  -> eps_x / eps_y gradients and Dx', Dy' cotangents; D-half
//                  psi cotangents and the H-difference cotangents gDx, gDy
//   P3 (Hz cells)  Hz cotangent += transposed differences of gDx, gDy
template <typename T>
struct AdjTe {
  int nx, ny;
  T dt, dx;
  const T *eps_x, *eps_y;
  Pml<T> hx, hy, ey, ex;
  T *aHz, *aEx, *aEy, *aDx, *aDy, *apEyx, *apExy, *apHzy, *apHzx;
  T *gEy, *gEx, *gDx, *gDy, *gepsx, *gepsy;
  // conductivity (sigma_d, sigma_b, mu): Dx' = d1x Dx + d2x cx,
  // Dy' = d1y Dy - d2y cy, Hz' = hz1 Hz - hz2 curl; NULL: 1 and dt
  const T *d1x, *d2x, *d1y, *d2y, *hz1, *hz2;
  // ADE poles per axis: per-cell c1, c2, c3 stacks and the cotangents of the
  // P / P_prev stacks (npx on the Ex grid, npy on the Ey grid)
  int npx, npy;
  const T *c1x, *c2x, *c3x, *c1y, *c2y, *c3y;
  T *aPxc, *aPxp, *aPyc, *aPyp;
  // chi2/chi3 fixed point (kerr; the tape then holds rhs = D - sum P): the
  // recomputed iterates kx[k] (Ex grid) and ky[k] (Ey grid), k = 0..2, the
  // cotangents kgx / kgy of the iterate being transposed, those of its
  // four-point averages (kga on the Ex grid, kgb on the Ey grid) and the
  // rhs cotangents krx / kry
  int kerr;
  const T *chi2x, *chi3x, *chi2y, *chi3y;
  T *kx[3], *ky[3], *kgx, *kgy, *kga, *kgb, *krx, *kry;
  // tensor eps (stepper_te.py:160-166): E = ixx rhs_x + ixyx avg(rhs_y),
  // E_y = iyy rhs_y + ixyy avg(rhs_x), and the gradients of the four maps
  const T *ixx, *ixyx, *iyy, *ixyy;
  T *gixx, *gixyx, *giyy, *gixyy;
  // oriented poles (stepper_te.py:140-145): each Px' also gains c3xy avg(Ey),
  // each Py' c3yx avg(Ex), with the old E; their cross cotangents are summed
  // into kga (Ex grid) / kgb (Ey grid) and scattered by te_oriented_scatter
  const T *c3xy, *c3yx;
  // even mirror cuts at the low edges (stepper_te.py:87-97): the on-plane Dx
  // row (ymir == 1) / Dy column (xmir == 1) took c2 * 2 Hz[0] / dx
  int ymir, xmir;
};

// Four-point averages of the product a * b (same neighbours and order as
// avg4_to_ex / avg4_to_ey), for the transposes of the tensor-eps cross terms.
template <typename T>
__device__ __forceinline__ T avg4_prod_to_ex(const T* a, const T* b, int nx, int ny, int i, int j) {
  auto at = [&](int ii, int jj, bool ok) { const long long q = (long long)ii * (ny - 1) + jj; return ok ? a[q] * b[q] : T(0); };
  return T(0.25) * (at(i, j - 1, j >= 1) + at(i, j, j < ny - 1) + at(i + 1, j - 1, j >= 1) + at(i + 1, j, j < ny - 1));
}

template <typename T>
__device__ __forceinline__ T avg4_prod_to_ey(const T* a, const T* b, int nx, int ny, int i, int j) {
  auto at = [&](int ii, int jj, bool ok) { const long long q = (long long)ii * ny + jj; return ok ? a[q] * b[q] : T(0); };
  return T(0.25) * (at(i - 1, j, i >= 1) + at(i - 1, j + 1, i >= 1) + at(i, j, i < nx - 1) + at(i, j + 1, i < nx - 1));
}

// Transpose of one axis's poles at a cell: P' = c1 P + c2 P_prev + c3 E(old)
// and E' = (D' - sum P') / eps, with r = E'b / eps the cotangent that E'
// passes to every P'.  Updates the stacks' cotangents in place and returns
// the old E's cotangent (the old E feeds nothing else in a TE step).
// With oriented poles (c3o) *cross gets sum c3o P'b, the cotangent of the
// other component's averaged old E at this cell.
template <typename T>
__device__ __forceinline__ T te_pole_adj(int np, const T* c1, const T* c2, const T* c3, T* apc, T* app,
                                         long long c, long long ncell, T r, const T* c3o = nullptr,
                                         T* cross = nullptr) {
  T e_old = T(0), x = T(0);
  for (int n = 0; n < np; ++n) {
    const long long o = (long long)n * ncell + c;
    const T apn = apc[o] - r, prev = app[o];
    apc[o] = c1[o] * apn + prev;
    app[o] = c2[o] * apn;
    e_old = e_old + c3[o] * apn;
    if (c3o) x = x + c3o[o] * apn;
  }
  if (cross) *cross = x;
  return e_old;
}

// The oriented poles' cross terms: the old Ey was averaged onto every Ex
// cell's Px' and the old Ex onto every Ey cell's Py', so Ey gains the
// Ey-grid average of kga and Ex the Ex-grid average of kgb (each average is
// the other's transpose).
template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_oriented_scatter(AdjTe<T> p) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {
    const long long c = (long long)i * ny + j;
    p.aEx[c] = p.aEx[c] + avg4_to_ex(p.kgb, nx, ny, i, j);
  }
  if (j < ny - 1) {
    const long long c = (long long)i * (ny - 1) + j;
    p.aEy[c] = p.aEy[c] + avg4_to_ey(p.kga, nx, ny, i, j);
  }
}

template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_adj_p1(AdjTe<T> p) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx - 1 || j >= ny - 1) return;
  const long long h = (long long)i * (ny - 1) + j;
  const bool sx = p.hx.in(i), sy = p.hy.in(j);
  const T ahz = p.aHz[h];
  T pyx = T(0), pxy = T(0), bx = T(0), ax = T(0), by = T(0), ay = T(0);
  if (sx) {
    pyx = p.apEyx[h];
    bx = p.hx.b[i];
    ax = p.hx.a[i];
  }
  if (sy) {
    pxy = p.apExy[h];
    by = p.hy.b[j];
    ay = p.hy.a[j];
  }
  const T h1 = p.hz1 ? p.hz1[h] : T(1), h2 = p.hz1 ? p.hz2[h] : p.dt;
  if (p.hz1) p.aHz[h] = h1 * ahz;  // Hz' = hz1 Hz - hz2 curl
  const T gc = -h2 * ahz;          // cotangent of the curl
  T gy = gc, gx = -gc;
  if (sx) {
    T t = pyx + gc;
    gy = gc + ax * t;
    p.apEyx[h] = bx * t;
  }
  if (sy) {
    T t = pxy - gc;
    gx = -gc + ay * t;
    p.apExy[h] = by * t;
  }
  p.gEy[h] = gy;
  p.gEx[h] = gx;
}

// P2 under chi2/chi3, first half: the cotangents of the final iterates
// (E after the step) go to kgx / kgy for the fixed point's transpose.
template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_adj_p2a(AdjTe<T> p) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {
    const long long c = (long long)i * ny + j, h = (long long)i * (ny - 1) + j;
    const T gxm = j >= 1 ? p.gEx[h - 1] : T(0), gx0 = j < ny - 1 ? p.gEx[h] : T(0);
    p.kgx[c] = p.aEx[c] + (gxm - gx0) / p.dx;
  }
  if (j < ny - 1) {
    const long long c = (long long)i * (ny - 1) + j;
    const T gym = i >= 1 ? p.gEy[c - (ny - 1)] : T(0), gy0 = i < nx - 1 ? p.gEy[c] : T(0);
    p.kgy[c] = p.aEy[c] + (gym - gy0) / p.dx;
  }
}

// The fixed point's iterates from the taped rhs (stepper_te.py:164-175):
// k = 0 the first iterate rhs / eps, k = 1, 2 one Jacobi pass from the
// previous pair, in the forward's arithmetic.
template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_kerr_iter(AdjTe<T> p, const T* tex, const T* tey, int k) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {
    const long long c = (long long)i * ny + j;
    if (k == 0) {   // the first iterate: rhs / eps_x, or the tensor eps product
      p.kx[0][c] = p.ixx ? p.ixx[c] * tex[c] + p.ixyx[c] * avg4_to_ex(tey, nx, ny, i, j) : tex[c] / p.eps_x[c];
    } else {
      const T av = avg4_to_ex(p.ky[k - 1], nx, ny, i, j), e = p.kx[k - 1][c];
      const T e2 = e * e + av * av;
      p.kx[k][c] = tex[c] / (p.eps_x[c] + p.chi2x[c] * e + p.chi3x[c] * e2);
    }
  }
  if (j < ny - 1) {
    const long long c = (long long)i * (ny - 1) + j;
    if (k == 0) {
      p.ky[0][c] = p.iyy ? p.iyy[c] * tey[c] + p.ixyy[c] * avg4_to_ey(tex, nx, ny, i, j) : tey[c] / p.eps_y[c];
    } else {
      const T av = avg4_to_ey(p.kx[k - 1], nx, ny, i, j), e = p.ky[k - 1][c];
      const T e2 = e * e + av * av;
      p.ky[k][c] = tey[c] / (p.eps_y[c] + p.chi2y[c] * e + p.chi3y[c] * e2);
    }
  }
}

// Transpose of Jacobi pass k (iterate k+1 from iterate k), pointwise part:
// E' = rhs / den, den = eps + chi2 E + chi3 (E^2 + av^2), so with g the
// cotangent of E' and t = -g E' / den: rhs gains g / den, eps gains t, E
// gains t (chi2 + 2 chi3 E) (into kgx / kgy) and the average av gains
// t chi3 2 av (kga / kgb, scattered by te_kerr_scatter).
template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_kerr_back(AdjTe<T> p, const T* tex, const T* tey, int k) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {
    const long long c = (long long)i * ny + j;
    const T av = avg4_to_ex(p.ky[k], nx, ny, i, j), e = p.kx[k][c];
    const T x2 = p.chi2x[c], x3 = p.chi3x[c];
    const T den = p.eps_x[c] + x2 * e + x3 * (e * e + av * av);
    const T en = tex[c] / den, g = p.kgx[c];
    const T t = (-g * en) / den;
    p.krx[c] = (k == 2 ? T(0) : p.krx[c]) + g / den;
    p.gepsx[c] = p.gepsx[c] + t;
    p.kgx[c] = t * (x2 + T(2) * x3 * e);
    p.kga[c] = t * (x3 * (T(2) * av));
  }
  if (j < ny - 1) {
    const long long c = (long long)i * (ny - 1) + j;
    const T av = avg4_to_ey(p.kx[k], nx, ny, i, j), e = p.ky[k][c];
    const T x2 = p.chi2y[c], x3 = p.chi3y[c];
    const T den = p.eps_y[c] + x2 * e + x3 * (e * e + av * av);
    const T en = tey[c] / den, g = p.kgy[c];
    const T t = (-g * en) / den;
    p.kry[c] = (k == 2 ? T(0) : p.kry[c]) + g / den;
    p.gepsy[c] = p.gepsy[c] + t;
    p.kgy[c] = t * (x2 + T(2) * x3 * e);
    p.kgb[c] = t * (x3 * (T(2) * av));
  }
}

// ... and the scatter of the averages' cotangents: the Ey-grid average of an
// Ex-grid field and the Ex-grid average of an Ey-grid field are each other's
// transposes, so Ex gains avg4_to_ex(kgb) and Ey gains avg4_to_ey(kga).
template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_kerr_scatter(AdjTe<T> p) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {
    const long long c = (long long)i * ny + j;
    p.kgx[c] = p.kgx[c] + avg4_to_ex(p.kgb, nx, ny, i, j);
  }
  if (j < ny - 1) {
    const long long c = (long long)i * (ny - 1) + j;
    p.kgy[c] = p.kgy[c] + avg4_to_ey(p.kga, nx, ny, i, j);
  }
}

// Transpose of the tensor-eps product with kgx / kgy the cotangents of the E
// it forms (after the fixed point's transpose under chi2/chi3): the rhs
// cotangents gain ixx g + the averaged ixy_y g of the Ey neighbours (and the
// mirror for Ey), the maps' gradients gain g times their factor.
template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_tensor_back(AdjTe<T> p, const T* tex, const T* tey) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {
    const long long c = (long long)i * ny + j;
    const T g = p.kgx[c];
    p.krx[c] = (p.kerr ? p.krx[c] : T(0)) + (p.ixx[c] * g + avg4_prod_to_ex(p.ixyy, p.kgy, nx, ny, i, j));
    p.gixx[c] = p.gixx[c] + g * tex[c];
    p.gixyx[c] = p.gixyx[c] + g * avg4_to_ex(tey, nx, ny, i, j);
  }
  if (j < ny - 1) {
    const long long c = (long long)i * (ny - 1) + j;
    const T g = p.kgy[c];
    p.kry[c] = (p.kerr ? p.kry[c] : T(0)) + (p.iyy[c] * g + avg4_prod_to_ey(p.ixyx, p.kgx, nx, ny, i, j));
    p.giyy[c] = p.giyy[c] + g * tey[c];
    p.gixyy[c] = p.gixyy[c] + g * avg4_to_ey(tex, nx, ny, i, j);
  }
}

// MODE 0: E = rhs / eps from the E' cotangent; 1: after the chi2/chi3
// transpose (krx holds the passes' rhs cotangent, kgx that of rhs / eps);
// 2: tensor eps (krx complete).
template <typename T, int MODE>
__global__ __launch_bounds__(ABX* ABY) void te_adj_p2(AdjTe<T> p, const T* tex, const T* tey) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx || j >= ny) return;
  if (i < nx - 1) {  // Ex / Dx cell (nx-1, ny)
    const long long c = (long long)i * ny + j;
    const long long h = (long long)i * (ny - 1) + j;
    const bool live = j >= 1 && j < ny - 1;
    const int jj = j - 1;
    const bool sh = live && p.ey.in(jj);
    const long long q = (long long)i * (ny - 2) + jj;
    const T gxm = j >= 1 ? p.gEx[h - 1] : T(0);
    const T gx0 = j < ny - 1 ? p.gEx[h] : T(0);
    const T aex = p.aEx[c], adx = p.aDx[c], epsx = p.eps_x[c], ex = tex[c], ge = p.gepsx[c];
    const T c1 = p.d1x ? p.d1x[c] : T(1), c2 = p.d1x ? p.d2x[c] : p.dt;
    T ps = T(0), b = T(0), a = T(0);
    if (sh) {
      ps = p.apHzy[q];
      b = p.ey.b[jj];
      a = p.ey.a[jj];
    }
    T rb;   // cotangent of rhs = Dx' - sum P'
    if constexpr (MODE == 2) {
      rb = p.krx[c];
    } else if constexpr (MODE == 1) {   // the first iterate rhs / eps_x; tex holds rhs
      const T g0 = p.kgx[c];
      rb = p.krx[c] + g0 / epsx;
      p.gepsx[c] = ge + (-g0 * p.kx[0][c]) / epsx;
    } else {
      const T exb = aex + (gxm - gx0) / p.dx;
      p.gepsx[c] = ge + (-exb * ex) / epsx;
      rb = exb / epsx;
    }
    const T dxb = adx + rb;
    p.aDx[c] = p.d1x ? c1 * dxb : dxb;
    T cr = T(0);
    p.aEx[c] = p.npx ? te_pole_adj(p.npx, p.c1x, p.c2x, p.c3x, p.aPxc, p.aPxp, c, (long long)(nx - 1) * ny, rb,
                                   p.c3xy, &cr)
                     : T(0);
    if (p.c3xy) p.kga[c] = cr;
    T gd = T(0);
    if (live) {
      const T g = c2 * dxb;
      gd = g;
      if (sh) {
        T t = ps + g;
        gd = g + a * t;
        p.apHzy[q] = b * t;
      }
    } else if (j == 0 && p.ymir == 1) {   // the odd Hz ghost: curl = 2 Hz[i, 0] / dx, no psi at the cut
      gd = T(2) * (c2 * dxb);
    }
    p.gDx[c] = gd;
  }
  if (j < ny - 1) {  // Ey / Dy cell (nx, ny-1)
    const long long c = (long long)i * (ny - 1) + j;
    const bool live = i >= 1 && i < nx - 1;
    const int ii = i - 1;
    const bool sh = live && p.ex.in(ii);
    const long long q = (long long)ii * (ny - 1) + j;
    const T gym = i >= 1 ? p.gEy[c - (ny - 1)] : T(0);
    const T gy0 = i < nx - 1 ? p.gEy[c] : T(0);
    const T aey = p.aEy[c], ady = p.aDy[c], epsy = p.eps_y[c], ey = tey[c], ge = p.gepsy[c];
    const T c1 = p.d1x ? p.d1y[c] : T(1), c2 = p.d1x ? p.d2y[c] : p.dt;
    T ps = T(0), b = T(0), a = T(0);
    if (sh) {
      ps = p.apHzx[q];
      b = p.ex.b[ii];
      a = p.ex.a[ii];
    }
    T rb;
    if constexpr (MODE == 2) {
      rb = p.kry[c];
    } else if constexpr (MODE == 1) {
      const T g0 = p.kgy[c];
      rb = p.kry[c] + g0 / epsy;
      p.gepsy[c] = ge + (-g0 * p.ky[0][c]) / epsy;
    } else {
      const T eyb = aey + (gym - gy0) / p.dx;
      p.gepsy[c] = ge + (-eyb * ey) / epsy;
      rb = eyb / epsy;
    }
    const T dyb = ady + rb;
    p.aDy[c] = p.d1x ? c1 * dyb : dyb;
    T cr = T(0);
    p.aEy[c] = p.npy ? te_pole_adj(p.npy, p.c1y, p.c2y, p.c3y, p.aPyc, p.aPyp, c, (long long)nx * (ny - 1), rb,
                                   p.c3yx, &cr)
                     : T(0);
    if (p.c3yx) p.kgb[c] = cr;
    T gd = T(0);
    if (live) {
      const T g = -c2 * dyb;
      gd = g;
      if (sh) {
        T t = ps + g;
        gd = g + a * t;
        p.apHzx[q] = b * t;
      }
    } else if (i == 0 && p.xmir == 1) {
      gd = T(2) * (-c2 * dyb);
    }
    p.gDy[c] = gd;
  }
}

template <typename T>
__global__ __launch_bounds__(ABX* ABY) void te_adj_p3(AdjTe<T> p) {
  const int j = blockIdx.x * ABX + threadIdx.x;
  const int i = blockIdx.y * ABY + threadIdx.y;
  const int nx = p.nx, ny = p.ny;
  if (i >= nx - 1 || j >= ny - 1) return;
  const long long h = (long long)i * (ny - 1) + j;
  const T d0 = p.gDx[(long long)i * ny + j], d1 = p.gDx[(long long)i * ny + j + 1];
  const T e0 = p.gDy[(long long)i * (ny - 1) + j], e1 = p.gDy[(long long)(i + 1) * (ny - 1) + j];
  p.aHz[h] = p.aHz[h] + (d0 - d1) / p.dx + (e0 - e1) / p.dx;
}

template <typename T>
int te_adj_t(const mt_te2d* d, const mt_te2d_adj* a, const void* tape, const mt_monitor* mons, int n_mons,
             int step0, int k, hipStream_t s) {
  AdjTe<T> p;
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
  p.aHz = static_cast<T*>(a->aHz);
  p.aEx = static_cast<T*>(a->aEx);
  p.aEy = static_cast<T*>(a->aEy);
  p.aDx = static_cast<T*>(a->aDx);
  p.aDy = static_cast<T*>(a->aDy);
  p.apEyx = static_cast<T*>(a->apsi_Eyx);
  p.apExy = static_cast<T*>(a->apsi_Exy);
  p.apHzy = static_cast<T*>(a->apsi_Hzy);
  p.apHzx = static_cast<T*>(a->apsi_Hzx);
  p.gEy = static_cast<T*>(a->gEy);
  p.gEx = static_cast<T*>(a->gEx);
  p.gDx = static_cast<T*>(a->gDx);
  p.gDy = static_cast<T*>(a->gDy);
  p.gepsx = static_cast<T*>(a->epsx_grad);
  p.gepsy = static_cast<T*>(a->epsy_grad);
  p.d1x = static_cast<const T*>(d->d1x);
  p.d2x = static_cast<const T*>(d->d2x);
  p.d1y = static_cast<const T*>(d->d1y);
  p.d2y = static_cast<const T*>(d->d2y);
  p.hz1 = static_cast<const T*>(d->hz1);
  p.hz2 = static_cast<const T*>(d->hz2);
  p.npx = (d->npx > 0 && a->aPx_cur) ? d->npx : 0;
  p.npy = (d->npy > 0 && a->aPy_cur) ? d->npy : 0;
  p.c1x = static_cast<const T*>(d->c1x);
  p.c2x = static_cast<const T*>(d->c2x);
  p.c3x = static_cast<const T*>(d->c3x);
  p.c1y = static_cast<const T*>(d->c1y);
  p.c2y = static_cast<const T*>(d->c2y);
  p.c3y = static_cast<const T*>(d->c3y);
  p.aPxc = static_cast<T*>(a->aPx_cur);
  p.aPxp = static_cast<T*>(a->aPx_prev);
  p.aPyc = static_cast<T*>(a->aPy_cur);
  p.aPyp = static_cast<T*>(a->aPy_prev);
  p.kerr = d->kerr;
  p.chi2x = static_cast<const T*>(d->chi2x);
  p.chi3x = static_cast<const T*>(d->chi3x);
  p.chi2y = static_cast<const T*>(d->chi2y);
  p.chi3y = static_cast<const T*>(d->chi3y);
  p.ixx = static_cast<const T*>(d->inv_xx_x);
  p.ixyx = static_cast<const T*>(d->inv_xy_x);
  p.iyy = static_cast<const T*>(d->inv_yy_y);
  p.ixyy = static_cast<const T*>(d->inv_xy_y);
  p.gixx = static_cast<T*>(a->ixx_grad);
  p.gixyx = static_cast<T*>(a->ixy_x_grad);
  p.giyy = static_cast<T*>(a->iyy_grad);
  p.gixyy = static_cast<T*>(a->ixy_y_grad);
  p.c3xy = static_cast<const T*>(d->c3xy);
  p.c3yx = static_cast<const T*>(d->c3yx);
  p.ymir = d->y_mirror;
  p.xmir = d->x_mirror;
  if (p.kerr || p.ixx || p.c3xy) {   // workspace: six Ex-grid arrays, then six Ey-grid arrays
    const long long nex = (long long)(p.nx - 1) * p.ny, ney = (long long)p.nx * (p.ny - 1);
    T* w = static_cast<T*>(a->kerr_ws);
    T* wy = w + 6 * nex;
    for (int q = 0; q < 3; ++q) p.kx[q] = w + q * nex, p.ky[q] = wy + q * ney;
    p.kgx = w + 3 * nex, p.kga = w + 4 * nex, p.krx = w + 5 * nex;
    p.kgy = wy + 3 * ney, p.kgb = wy + 4 * ney, p.kry = wy + 5 * ney;
  }
  const T* tp = static_cast<const T*>(tape);
  const long long slot = (long long)(p.nx - 1) * p.ny + (long long)p.nx * (p.ny - 1);
  dim3 blk(ABX, ABY), grd = grid2(p.ny, p.nx, blk);
  for (int s_ = k - 1; s_ >= 0; --s_) {
    const int t = step0 + s_;
    for (int m = 0; m < n_mons; ++m) {
      const mt_monitor& mm = mons[m];
      const int n = mm.n0 * mm.n1;
      if (n == 0) continue;
      hipLaunchKernelGGL(adj_inject_kernel<T>, dim3((n + kInjE - 1) / kInjE), dim3(kInjE * kInjS), 0, s, mm, t,
                         d->dt, s_);
    }
    const T* ts = tp + (long long)s_ * slot;
    const T* tsy = ts + (long long)(p.nx - 1) * p.ny;
    hipLaunchKernelGGL(te_adj_p1<T>, grd, blk, 0, s, p);
    if (p.kerr || p.ixx) {   // E' from rhs through the tensor product and / or the three Jacobi passes
      hipLaunchKernelGGL(te_adj_p2a<T>, grd, blk, 0, s, p);
      if (p.kerr) {   // recompute the iterates, then transpose the passes last to first
        for (int q = 0; q < 3; ++q) hipLaunchKernelGGL(te_kerr_iter<T>, grd, blk, 0, s, p, ts, tsy, q);
        for (int q = 2; q >= 0; --q) {
          hipLaunchKernelGGL(te_kerr_back<T>, grd, blk, 0, s, p, ts, tsy, q);
          hipLaunchKernelGGL(te_kerr_scatter<T>, grd, blk, 0, s, p);
        }
      }
      if (p.ixx) {
        hipLaunchKernelGGL(te_tensor_back<T>, grd, blk, 0, s, p, ts, tsy);
        hipLaunchKernelGGL((te_adj_p2<T, 2>), grd, blk, 0, s, p, ts, tsy);
      } else {
        hipLaunchKernelGGL((te_adj_p2<T, 1>), grd, blk, 0, s, p, ts, tsy);
      }
    } else {
      hipLaunchKernelGGL((te_adj_p2<T, 0>), grd, blk, 0, s, p, ts, tsy);
    }
    if (p.c3xy) hipLaunchKernelGGL(te_oriented_scatter<T>, grd, blk, 0, s, p);
    hipLaunchKernelGGL(te_adj_p3<T>, grd, blk, 0, s, p);
    int rc = check_launch("te2d adjoint step");
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
  if (p->kerr && (!p->chi2 || !p->chi3))
    return fail(MT_EINVAL, "adjoint through chi2/chi3 needs chi2 and chi3 (and a tape of Dz - sum P)");
  if ((a->chi2_grad || a->chi3_grad) && !p->kerr)
    return fail(MT_EINVAL, "chi2/chi3 gradients need the chi2/chi3 fixed point (kerr = 1)");
  if (a->c3_grad && p->n_poles > 0 && k > 0 && !a->ez_init)
    return fail(MT_EINVAL, "the c3 gradient needs ez_init (Ez before the first step)");
  if (p->n_poles < 0 || (p->n_poles && (!a->aP_cur || !a->aP_prev || !p->c1 || !p->c2 || !p->c3)))
    return fail(MT_EINVAL, "adjoint with poles needs c1, c2, c3 and the aP_cur / aP_prev cotangents");
  if (p->n_mpoles < 0 ||
      (p->n_mpoles && (!a->aBx || !a->aBy || !a->aMx_cur || !a->aMx_prev || !a->aMy_cur || !a->aMy_prev ||
                       !p->mc1x || !p->mc2x || !p->mc3x || !p->mc1y || !p->mc2y || !p->mc3y)))
    return fail(MT_EINVAL, "adjoint with magnetic poles needs mc1x..mc3y and the B / M cotangents");
  if (p->n_mpoles && p->h1x) return fail(MT_EINVAL, "magnetic poles do not combine with sigma_b / mu");
  if (p->y_mirror == -1 || p->x_mirror == -1 || p->y_rot2 || p->rot4)
    return fail(MT_EINVAL, "the TM adjoint transposes even mirror cuts only (no odd or Rotate cuts)");
  if (p->y_mirror < -1 || p->y_mirror > 1 || p->x_mirror < -1 || p->x_mirror > 1 || (p->y_mirror && p->y_pmc))
    return fail(MT_EINVAL, "mirror parity must be 0/+1/-1, without y_pmc");
  if (!p->d1 != !p->d2 || !p->h1x != !p->h2x || !p->h1y != !p->h2y || !p->d1 != !p->h1x || !p->d1 != !p->h1y)
    return fail(MT_EINVAL, "conductivity needs all of d1, d2, h1x, h2x, h1y, h2y");
  if (!p->eps || !a->aEz || !a->aDz || !a->aHx || !a->aHy || !a->apsi_Ezy || !a->apsi_Ezx ||
      !a->apsi_Hyx || !a->apsi_Hxy || !a->gHy || !a->gHx || !a->gEx || !a->gEy || !a->eps_grad)
    return fail(MT_EINVAL, "null adjoint pointer");
  int rc;
  if ((rc = check_axis(p->hy, p->ny - 1, "hy")) || (rc = check_axis(p->hx, p->nx - 1, "hx")) ||
      (rc = check_axis(p->ex, p->x_mirror == 1 ? p->nx - 1 : p->nx - 2, "ex")) ||
      (rc = check_axis(p->ey, p->y_pmc ? p->ny : (p->y_mirror == 1 ? p->ny - 1 : p->ny - 2), "ey")))
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

extern "C" int mt_te2d_adjoint(const mt_te2d* p, const mt_te2d_adj* a, const void* tape, const mt_monitor* mons,
                               int n_mons, int step0, int k, void* stream) {
  if (!p || !a) return fail(MT_EINVAL, "null problem or adjoint state");
  if (p->nx < 3 || p->ny < 3) return fail(MT_EINVAL, "grid must be at least 3x3");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  if (k > 0 && !tape) return fail(MT_EINVAL, "adjoint needs the forward E tape");
  if (p->kerr && (!a->kerr_ws || !p->chi2x || !p->chi3x || !p->chi2y || !p->chi3y))
    return fail(MT_EINVAL, "adjoint through chi2/chi3 needs the four chi maps and kerr_ws (and a tape of D - sum P)");
  if ((p->c3xy || p->c3yx) && (!p->c3xy || !p->c3yx || !a->kerr_ws))
    return fail(MT_EINVAL, "adjoint with oriented poles needs c3xy, c3yx and kerr_ws");
  if (p->inv_xx_x && (!p->inv_xy_x || !p->inv_yy_y || !p->inv_xy_y || !a->kerr_ws || !a->ixx_grad ||
                      !a->ixy_x_grad || !a->iyy_grad || !a->ixy_y_grad))
    return fail(MT_EINVAL, "adjoint with tensor eps needs the four inverse maps, their gradients and kerr_ws");
  if (p->y_mirror < -1 || p->y_mirror > 1 || p->x_mirror < -1 || p->x_mirror > 1)
    return fail(MT_EINVAL, "mirror parity must be 0/+1/-1");
  if (p->npx < 0 || p->npy < 0 ||
      (p->npx && (!a->aPx_cur || !a->aPx_prev || !p->c1x || !p->c2x || !p->c3x)) ||
      (p->npy && (!a->aPy_cur || !a->aPy_prev || !p->c1y || !p->c2y || !p->c3y)))
    return fail(MT_EINVAL, "adjoint with poles needs c1, c2, c3 and the pole-stack cotangents");
  if (!p->d1x != !p->d2x || !p->d1x != !p->d1y || !p->d1x != !p->d2y || !p->d1x != !p->hz1 || !p->d1x != !p->hz2)
    return fail(MT_EINVAL, "conductivity needs all of d1x, d2x, d1y, d2y, hz1, hz2");
  if (p->dtype != MT_F32 && p->dtype != MT_F64) return fail(MT_EINVAL, "dtype must be f32 or f64");
  if (!p->eps_x || !p->eps_y || !a->aHz || !a->aEx || !a->aEy || !a->aDx || !a->aDy || !a->apsi_Eyx ||
      !a->apsi_Exy || !a->apsi_Hzy || !a->apsi_Hzx || !a->gEy || !a->gEx || !a->gDx || !a->gDy ||
      !a->epsx_grad || !a->epsy_grad)
    return fail(MT_EINVAL, "null adjoint pointer");
  int rc;
  if ((rc = check_axis(p->hx, p->nx - 1, "hx")) || (rc = check_axis(p->hy, p->ny - 1, "hy")) ||
      (rc = check_axis(p->ey, p->ny - 2, "ey")) || (rc = check_axis(p->ex, p->nx - 2, "ex")))
    return rc;
  for (int i = 0; i < n_mons; ++i) {
    const mt_monitor& m = mons[i];
    if (m.n_terms < 1 || m.n_terms > MT_MAX_TERMS || !m.acc) return fail(MT_EINVAL, "bad monitor");
    if (m.kind == 0 && (m.nfreq < 1 || !m.omega)) return fail(MT_EINVAL, "monitor: no frequencies");
    for (int q = 0; q < m.n_terms; ++q)
      if (!m.term[q].a) return fail(MT_EINVAL, "monitor: null adjoint view");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == MT_F64 ? te_adj_t<double>(p, a, tape, mons, n_mons, step0, k, s)
                            : te_adj_t<float>(p, a, tape, mons, n_mons, step0, k, s);
}
