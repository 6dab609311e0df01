// Reverse-time discrete adjoint of the 3D functional Yee step (CPML shells,
// material conductivity, soft sources, DFT monitors), for dL/deps.
//
// Forward step t (yee3d.hip) maps x_t -> x_{t+1}:
//   H half  for each H component, two differences of the old E (÷ dx), each
//           through its CPML recursion psi' = b psi + a g inside a shell:
//           Hx' = Hx - dt (cy - cz), Hy' = Hy - dt (cz - cx), Hz' = Hz - dt (cx - cy)
//   E half  for each E component, two differences of the new H through the
//           E-side psi; curl = 0 on the PEC faces; eps_c = the two-point
//           average of the cell eps along the component's axis;
//           E' = E + (dt curl + source) / eps_c  (in exact arithmetic)
//   then the monitors sample (E', H').
// One reverse step transposes this in four passes:
//   inject  monitor cotangents into the adjoint arrays
//   A       per E cell: eps_c cotangent -(dL/dE') (E' - E) / eps_c (E, E'
//           from the tape), curl cotangent dt/eps_c dL/dE', E-side psi
//           cotangents, and the cotangents of the H differences (g[0..5])
//   B       per H cell: dL/dH' += transposed H differences of g[0..5];
//           H-side psi cotangents and the E-difference cotangents (g[6..11])
//   C       per cell: dL/dE += transposed E differences of g[6..11];
//           dL/deps += 1/2 of the eps_c cotangents of the components that
//           average this cell
// Identity parts (dE'/dE = 1, dH'/dH = 1) leave the adjoint arrays in place.
// With poles, A also transposes the pole recurrence (pole3_adj).
// With conductivity the step is E' = d1 E + (d2 curl + source) / eps_c and
// H' = h1 H - h2 curl: A scales dL/dE' by d1 in place, uses d2 for dt and
// -(dL/dE') (E' - d1 E) / eps_c for the eps_c cotangent; B scales dL/dH' by
// h1 and uses h2 for dt.
#include <hip/hip_runtime.h>

#include "adjoint2d.h"
#include "common.h"

namespace mt {
namespace {

constexpr int ABZ = 64;
constexpr int ABY = 4;

enum { P_EZ_Y, P_EY_Z, P_EX_Z, P_EZ_X, P_EY_X, P_EX_Y, P_HZ_Y, P_HY_Z, P_HX_Z, P_HZ_X, P_HY_X, P_HX_Y };
// difference-cotangent scratch slots
enum { G_HZ_Y, G_HY_Z, G_HX_Z, G_HZ_X, G_HY_X, G_HX_Y, G_EZ_Y, G_EY_Z, G_EX_Z, G_EZ_X, G_EY_X, G_EX_Y };

template <typename T>
struct A3 {
  int nx, ny, nz;
  T dt, dx;
  const T* eps;
  Pml<T> h[3], e[3];
  T *aEx, *aEy, *aEz, *aHx, *aHy, *aHz;
  T* apsi[12];
  T* g[12];
  T* geps[3];
  T* eps_grad;
  // conductivity pairs (d1, d2) of ex ey ez hx hy hz, or all NULL
  const T* dmp[12];
  // np ADE poles: scalars c1[n], c2[n]; per E component the c3 stack and the
  // cotangents of the P / P_prev stacks (np component grids each)
  int np;
  const T *c1, *c2;
  const T* c3[3];
  T *aPc[3], *aPp[3];
};

// Transpose of one E cell's poles: P' = (c1 P + c2 P_prev) + c3 E and
// E' = (d1 (E eps_c + sum P) + d2 curl - sum P') / eps_c, with r = E'b / eps_c.
// P' gains -r, P gains d1 r + c1 P'b + P_prev'b (P_prev' = P), P_prev gains
// c2 P'b; returns the old E's cotangent from the poles, sum c3 P'b.
template <typename T>
__device__ __forceinline__ T pole3_adj(const A3<T>& p, int a, long long q, long long nq, T d1, T r) {
  T e_old = T(0);
  for (int n = 0; n < p.np; ++n) {
    const long long o = (long long)n * nq + q;
    const T apn = p.aPc[a][o] - r, prev = p.aPp[a][o];
    p.aPc[a][o] = d1 * r + p.c1[n] * apn + prev;
    p.aPp[a][o] = p.c2[n] * apn;
    e_old = e_old + p.c3[a][o] * apn;
  }
  return e_old;
}

// Transpose of c = d + psi', psi' = b psi + a d (psi' only inside a shell):
// given the cotangent cb of c, returns the cotangent of d and updates the
// psi cotangent in place (psi' cotangent from later steps + cb -> b * that).
template <typename T>
__device__ __forceinline__ T ce_t(const Pml<T>& pr, int idx, T* apsi, long long q, T cb) {
  if (!pr.in(idx)) return cb;
  T tot = apsi[q] + cb;
  apsi[q] = pr.b[idx] * tot;
  return cb + pr.a[idx] * tot;
}

template <typename T>
__global__ __launch_bounds__(ABZ* ABY) void adj3_a(A3<T> p, const T* e0, const T* e1) {
  const int k = blockIdx.x * ABZ + threadIdx.x, j = blockIdx.y * ABY + threadIdx.y, i = blockIdx.z;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (j >= ny || k >= nz) return;
  const T dt = p.dt, rdx = T(1) / p.dx;
  const long long nEx = (long long)(nx - 1) * ny * nz, nEy = (long long)nx * (ny - 1) * nz;
  const long long c_ = ((long long)i * ny + j) * nz + k;
  const T eps0 = p.eps[c_];
  if (i < nx - 1) {  // Ex: curl = chy (Hz y-diff) - chz (Hy z-diff)
    const long long x = c_;
    const T ab = p.aEx[x];
    const T epsc = T(0.5) * (eps0 + p.eps[c_ + (long long)ny * nz]);
    const T d1 = p.dmp[0] ? p.dmp[0][x] : T(1), d2 = p.dmp[0] ? p.dmp[1][x] : dt;
    p.geps[0][x] = -ab * (e1[x] - (p.dmp[0] ? d1 * e0[x] : e0[x])) / epsc;
    if (p.np) p.aEx[x] = d1 * ab + pole3_adj(p, 0, x, nEx, d1, ab / epsc);
    else if (p.dmp[0]) p.aEx[x] = d1 * ab;
    const bool jy = j >= 1 && j < ny - 1, kz = k >= 1 && k < nz - 1;
    const T cb = (jy && kz) ? ab * d2 / epsc : T(0);
    T gy = T(0), gz = T(0);
    if (jy) gy = ce_t(p.e[1], j - 1, p.apsi[P_HZ_Y], ((long long)i * (ny - 2) + (j - 1)) * nz + k, cb);
    if (kz) gz = ce_t(p.e[2], k - 1, p.apsi[P_HY_Z], ((long long)i * ny + j) * (nz - 2) + (k - 1), -cb);
    p.g[G_HZ_Y][x] = gy * rdx;
    p.g[G_HY_Z][x] = gz * rdx;
  }
  if (j < ny - 1) {  // Ey: curl = chz (Hx z-diff) - chx (Hz x-diff)
    const long long y = ((long long)i * (ny - 1) + j) * nz + k;
    const T ab = p.aEy[y];
    const T epsc = T(0.5) * (eps0 + p.eps[c_ + nz]);
    const T d1 = p.dmp[0] ? p.dmp[2][y] : T(1), d2 = p.dmp[0] ? p.dmp[3][y] : dt;
    p.geps[1][y] = -ab * (e1[nEx + y] - (p.dmp[0] ? d1 * e0[nEx + y] : e0[nEx + y])) / epsc;
    if (p.np) p.aEy[y] = d1 * ab + pole3_adj(p, 1, y, nEy, d1, ab / epsc);
    else if (p.dmp[0]) p.aEy[y] = d1 * ab;
    const bool kz = k >= 1 && k < nz - 1, ix = i >= 1 && i < nx - 1;
    const T cb = (kz && ix) ? ab * d2 / epsc : T(0);
    T gz = T(0), gx = T(0);
    if (kz) gz = ce_t(p.e[2], k - 1, p.apsi[P_HX_Z], ((long long)i * (ny - 1) + j) * (nz - 2) + (k - 1), cb);
    if (ix) gx = ce_t(p.e[0], i - 1, p.apsi[P_HZ_X], ((long long)(i - 1) * (ny - 1) + j) * nz + k, -cb);
    p.g[G_HX_Z][y] = gz * rdx;
    p.g[G_HZ_X][y] = gx * rdx;
  }
  if (k < nz - 1) {  // Ez: curl = chx (Hy x-diff) - chy (Hx y-diff)
    const long long z = ((long long)i * ny + j) * (nz - 1) + k;
    const T ab = p.aEz[z];
    const T epsc = T(0.5) * (eps0 + p.eps[c_ + 1]);
    const T d1 = p.dmp[0] ? p.dmp[4][z] : T(1), d2 = p.dmp[0] ? p.dmp[5][z] : dt;
    p.geps[2][z] = -ab * (e1[nEx + nEy + z] - (p.dmp[0] ? d1 * e0[nEx + nEy + z] : e0[nEx + nEy + z])) / epsc;
    if (p.np) p.aEz[z] = d1 * ab + pole3_adj(p, 2, z, (long long)nx * ny * (nz - 1), d1, ab / epsc);
    else if (p.dmp[0]) p.aEz[z] = d1 * ab;
    const bool ix = i >= 1 && i < nx - 1, jy = j >= 1 && j < ny - 1;
    const T cb = (ix && jy) ? ab * d2 / epsc : T(0);
    T gx = T(0), gy = T(0);
    if (ix) gx = ce_t(p.e[0], i - 1, p.apsi[P_HY_X], ((long long)(i - 1) * ny + j) * (nz - 1) + k, cb);
    if (jy) gy = ce_t(p.e[1], j - 1, p.apsi[P_HX_Y], ((long long)i * (ny - 2) + (j - 1)) * (nz - 1) + k, -cb);
    p.g[G_HY_X][z] = gx * rdx;
    p.g[G_HX_Y][z] = gy * rdx;
  }
}

template <typename T>
__global__ __launch_bounds__(ABZ* ABY) void adj3_b(A3<T> p) {
  const int k = blockIdx.x * ABZ + threadIdx.x, j = blockIdx.y * ABY + threadIdx.y, i = blockIdx.z;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (j >= ny || k >= nz) return;
  const T dt = p.dt, rdx = T(1) / p.dx;
  // E-component linear indices of neighbours
  auto ex = [&](int a, int b, int c) { return ((long long)a * ny + b) * nz + c; };
  auto ey = [&](int a, int b, int c) { return ((long long)a * (ny - 1) + b) * nz + c; };
  auto ez = [&](int a, int b, int c) { return ((long long)a * ny + b) * (nz - 1) + c; };
  if (j < ny - 1 && k < nz - 1) {  // Hx (nx, ny-1, nz-1)
    const long long h = ((long long)i * (ny - 1) + j) * (nz - 1) + k;
    T a = p.aHx[h] + (p.g[G_HX_Z][ey(i, j, k)] - p.g[G_HX_Z][ey(i, j, k + 1)]) +
          (p.g[G_HX_Y][ez(i, j, k)] - p.g[G_HX_Y][ez(i, j + 1, k)]);
    const T h2 = p.dmp[0] ? p.dmp[7][h] : dt;
    p.aHx[h] = p.dmp[0] ? p.dmp[6][h] * a : a;
    // Hx' = h1 Hx - h2 (cy - cz): cy from Ez y-diff, cz from Ey z-diff
    p.g[G_EZ_Y][h] = ce_t(p.h[1], j, p.apsi[P_EZ_Y], h, -h2 * a) * rdx;
    p.g[G_EY_Z][h] = ce_t(p.h[2], k, p.apsi[P_EY_Z], h, h2 * a) * rdx;
  }
  if (i < nx - 1 && k < nz - 1) {  // Hy (nx-1, ny, nz-1)
    const long long h = ((long long)i * ny + j) * (nz - 1) + k;
    T a = p.aHy[h] + (p.g[G_HY_Z][ex(i, j, k)] - p.g[G_HY_Z][ex(i, j, k + 1)]) +
          (p.g[G_HY_X][ez(i, j, k)] - p.g[G_HY_X][ez(i + 1, j, k)]);
    const T h2 = p.dmp[0] ? p.dmp[9][h] : dt;
    p.aHy[h] = p.dmp[0] ? p.dmp[8][h] * a : a;
    // Hy' = h1 Hy - h2 (cz - cx): cz from Ex z-diff, cx from Ez x-diff
    p.g[G_EX_Z][h] = ce_t(p.h[2], k, p.apsi[P_EX_Z], h, -h2 * a) * rdx;
    p.g[G_EZ_X][h] = ce_t(p.h[0], i, p.apsi[P_EZ_X], h, h2 * a) * rdx;
  }
  if (i < nx - 1 && j < ny - 1) {  // Hz (nx-1, ny-1, nz)
    const long long h = ((long long)i * (ny - 1) + j) * nz + k;
    T a = p.aHz[h] + (p.g[G_HZ_Y][ex(i, j, k)] - p.g[G_HZ_Y][ex(i, j + 1, k)]) +
          (p.g[G_HZ_X][ey(i, j, k)] - p.g[G_HZ_X][ey(i + 1, j, k)]);
    const T h2 = p.dmp[0] ? p.dmp[11][h] : dt;
    p.aHz[h] = p.dmp[0] ? p.dmp[10][h] * a : a;
    // Hz' = h1 Hz - h2 (cx - cy): cx from Ey x-diff, cy from Ex y-diff
    p.g[G_EY_X][h] = ce_t(p.h[0], i, p.apsi[P_EY_X], h, -h2 * a) * rdx;
    p.g[G_EX_Y][h] = ce_t(p.h[1], j, p.apsi[P_EX_Y], h, h2 * a) * rdx;
  }
}

template <typename T>
__global__ __launch_bounds__(ABZ* ABY) void adj3_c(A3<T> p) {
  const int k = blockIdx.x * ABZ + threadIdx.x, j = blockIdx.y * ABY + threadIdx.y, i = blockIdx.z;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (j >= ny || k >= nz) return;
  auto hx = [&](int a, int b, int c) { return ((long long)a * (ny - 1) + b) * (nz - 1) + c; };
  auto hy = [&](int a, int b, int c) { return ((long long)a * ny + b) * (nz - 1) + c; };
  auto hz = [&](int a, int b, int c) { return ((long long)a * (ny - 1) + b) * nz + c; };
  const long long c_ = ((long long)i * ny + j) * nz + k;
  T ge = T(0);
  if (i < nx - 1) {  // Ex (nx-1, ny, nz): in Hy (z-diff) and Hz (y-diff)
    T a = p.aEx[c_];
    if (k >= 1) a += p.g[G_EX_Z][hy(i, j, k - 1)];
    if (k < nz - 1) a -= p.g[G_EX_Z][hy(i, j, k)];
    if (j >= 1) a += p.g[G_EX_Y][hz(i, j - 1, k)];
    if (j < ny - 1) a -= p.g[G_EX_Y][hz(i, j, k)];
    p.aEx[c_] = a;
    ge += p.geps[0][c_];
  }
  if (i >= 1) ge += p.geps[0][c_ - (long long)ny * nz];
  if (j < ny - 1) {  // Ey (nx, ny-1, nz): in Hx (z-diff) and Hz (x-diff)
    const long long y = ((long long)i * (ny - 1) + j) * nz + k;
    T a = p.aEy[y];
    if (k >= 1) a += p.g[G_EY_Z][hx(i, j, k - 1)];
    if (k < nz - 1) a -= p.g[G_EY_Z][hx(i, j, k)];
    if (i >= 1) a += p.g[G_EY_X][hz(i - 1, j, k)];
    if (i < nx - 1) a -= p.g[G_EY_X][hz(i, j, k)];
    p.aEy[y] = a;
    ge += p.geps[1][y];
  }
  if (j >= 1) ge += p.geps[1][((long long)i * (ny - 1) + (j - 1)) * nz + k];
  if (k < nz - 1) {  // Ez (nx, ny, nz-1): in Hx (y-diff) and Hy (x-diff)
    const long long z = ((long long)i * ny + j) * (nz - 1) + k;
    T a = p.aEz[z];
    if (j >= 1) a += p.g[G_EZ_Y][hx(i, j - 1, k)];
    if (j < ny - 1) a -= p.g[G_EZ_Y][hx(i, j, k)];
    if (i >= 1) a += p.g[G_EZ_X][hy(i - 1, j, k)];
    if (i < nx - 1) a -= p.g[G_EZ_X][hy(i, j, k)];
    p.aEz[z] = a;
    ge += p.geps[2][z];
  }
  if (k >= 1) ge += p.geps[2][((long long)i * ny + j) * (nz - 1) + (k - 1)];
  p.eps_grad[c_] += T(0.5) * ge;
}

template <typename T>
__global__ __launch_bounds__(kInjE* kInjS) void adj3_inject(mt_monitor m, int t, double dt_d) {
  adj_inject_block<T>(m, blockIdx.x * kInjE, t, dt_d, 0);
}

template <typename T>
int adj3_t(const mt_yee3d* d, const mt_yee3d_adj* a, const void* tape, const mt_monitor* mons, int n_mons,
           int step0, int k, hipStream_t s) {
  A3<T> p;
  p.nx = d->nx;
  p.ny = d->ny;
  p.nz = d->nz;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.eps = static_cast<const T*>(d->eps);
  for (int q = 0; q < 3; ++q) {
    p.h[q] = pml_of<T>(d->h[q]);
    p.e[q] = pml_of<T>(d->e[q]);
  }
  p.aEx = static_cast<T*>(a->aEx);
  p.aEy = static_cast<T*>(a->aEy);
  p.aEz = static_cast<T*>(a->aEz);
  p.aHx = static_cast<T*>(a->aHx);
  p.aHy = static_cast<T*>(a->aHy);
  p.aHz = static_cast<T*>(a->aHz);
  for (int q = 0; q < 12; ++q) {
    p.apsi[q] = static_cast<T*>(a->apsi[q]);
    p.g[q] = static_cast<T*>(a->g[q]);
  }
  for (int q = 0; q < 3; ++q) p.geps[q] = static_cast<T*>(a->geps[q]);
  p.eps_grad = static_cast<T*>(a->eps_grad);
  for (int q = 0; q < 12; ++q) p.dmp[q] = static_cast<const T*>(d->damp[q]);
  p.np = (d->n_poles > 0 && a->aP_cur[0]) ? d->n_poles : 0;
  p.c1 = static_cast<const T*>(d->c1);
  p.c2 = static_cast<const T*>(d->c2);
  for (int q = 0; q < 3; ++q) {
    p.c3[q] = static_cast<const T*>(d->c3[q]);
    p.aPc[q] = static_cast<T*>(a->aP_cur[q]);
    p.aPp[q] = static_cast<T*>(a->aP_prev[q]);
  }
  const long long slot = (long long)(p.nx - 1) * p.ny * p.nz + (long long)p.nx * (p.ny - 1) * p.nz +
                         (long long)p.nx * p.ny * (p.nz - 1);
  const T* tp = static_cast<const T*>(tape);
  dim3 blk(ABZ, ABY, 1), grd((p.nz + ABZ - 1) / ABZ, (p.ny + ABY - 1) / ABY, p.nx);
  for (int s_ = k - 1; s_ >= 0; --s_) {
    const int t = step0 + s_;
    for (int m = 0; m < n_mons; ++m) {
      const mt_monitor& mm = mons[m];
      const int n = mm.n0 * mm.n1;
      if (n == 0) continue;
      hipLaunchKernelGGL(adj3_inject<T>, dim3((n + kInjE - 1) / kInjE), dim3(kInjE * kInjS), 0, s, mm, t, d->dt);
    }
    hipLaunchKernelGGL(adj3_a<T>, grd, blk, 0, s, p, tp + (long long)s_ * slot, tp + (long long)(s_ + 1) * slot);
    hipLaunchKernelGGL(adj3_b<T>, grd, blk, 0, s, p);
    hipLaunchKernelGGL(adj3_c<T>, grd, blk, 0, s, p);
    int rc = check_launch("yee3d adjoint step");
    if (rc) return rc;
  }
  return MT_OK;
}

}  // namespace
}  // namespace mt

using namespace mt;

extern "C" int mt_yee3d_adjoint(const mt_yee3d* p, const mt_yee3d_adj* a, const void* tape,
                                const mt_monitor* mons, int n_mons, int step0, int k, void* stream) {
  if (!p || !a) return fail(MT_EINVAL, "null problem or adjoint state");
  if (p->nx < 3 || p->ny < 3 || p->nz < 3) return fail(MT_EINVAL, "grid must be at least 3x3x3");
  if ((long long)p->ny * p->nz > (1LL << 31) || p->nx > 65535)
    return fail(MT_EINVAL, "grid too large for the launch shape");
  if (p->lean) return fail(MT_EINVAL, "the 3D adjoint transposes the functional stepper (lean = 0)");
  if (p->absorber) return fail(MT_EINVAL, "the 3D adjoint transposes the CPML path (absorber = 0)");
  if (p->n_poles < 0) return fail(MT_EINVAL, "negative pole count");
  if (p->n_poles) {
    if (!p->damp[0] || !p->c1 || !p->c2) return fail(MT_EINVAL, "3D adjoint with poles needs damp, c1 and c2");
    for (int q = 0; q < 3; ++q)
      if (!p->c3[q] || !a->aP_cur[q] || !a->aP_prev[q])
        return fail(MT_EINVAL, "3D adjoint with poles needs c3 and the pole-stack cotangents");
  }
  for (int q = 1; q < 12; ++q)
    if (!p->damp[q] != !p->damp[0]) return fail(MT_EINVAL, "3D conductivity needs all 12 damping arrays");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  if (k > 0 && !tape) return fail(MT_EINVAL, "adjoint needs the forward E tape");
  if (p->dtype != MT_F32 && p->dtype != MT_F64) return fail(MT_EINVAL, "dtype must be f32 or f64");
  if (!p->eps || !a->aEx || !a->aEy || !a->aEz || !a->aHx || !a->aHy || !a->aHz || !a->eps_grad)
    return fail(MT_EINVAL, "null adjoint pointer");
  for (int q = 0; q < 12; ++q)
    if (!a->apsi[q] || !a->g[q]) return fail(MT_EINVAL, "null adjoint psi or scratch pointer");
  for (int q = 0; q < 3; ++q)
    if (!a->geps[q]) return fail(MT_EINVAL, "null eps_c cotangent pointer");
  int n[3] = {p->nx, p->ny, p->nz};
  int rc;
  for (int q = 0; q < 3; ++q) {
    if ((rc = check_axis(p->h[q], n[q] - 1, "h"))) return rc;
    if ((rc = check_axis(p->e[q], n[q] - 2, "e"))) return rc;
  }
  for (int i = 0; i < n_mons; ++i) {
    const mt_monitor& m = mons[i];
    if (m.n_terms < 1 || m.n_terms > MT_MAX_TERMS || !m.acc) return fail(MT_EINVAL, "bad monitor");
    if (m.kind == 0 && (m.nfreq < 1 || !m.omega)) return fail(MT_EINVAL, "monitor: no frequencies");
    if (m.kind != 0) return fail(MT_EINVAL, "3D adjoint: DFT monitors only");
    for (int q = 0; q < m.n_terms; ++q)
      if (!m.term[q].a) return fail(MT_EINVAL, "monitor: null adjoint view");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == MT_F64 ? adj3_t<double>(p, a, tape, mons, n_mons, step0, k, s)
                            : adj3_t<float>(p, a, tape, mons, n_mons, step0, k, s);
}
