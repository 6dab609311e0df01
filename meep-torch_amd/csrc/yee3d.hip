// 3D Yee leapfrog with CPML shells, point / masked soft sources, and (two-pass
// path) material conductivity and ADE poles.
//
// Two passes per step: H from curl E, then E from curl H.  One thread per
// cell (i, j, k) of the nx*ny*nz lattice updates every component whose grid
// covers that cell; z is the contiguous axis and runs along the wave.
// Arithmetic follows meep_torch/core/stepper_3d.py:363-445 op by op (lean:
// stepper_3d_lean.py:234-296), so FP32 rounds exactly like the reference.
// psi arrays keep the reference's full-grid shapes; only cells inside a CPML
// shell touch them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

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
  // media (two-pass kernels with M = true): (d1, d2) of ex ey ez hx hy hz,
  // np poles with scalars c1[n], c2[n] and per E component the stacks
  // c3[a], Pc[a] (P), Pp[a] (P_prev) of np component grids each
  const T* dmp[12];
  int np;
  const T *c1, *c2;
  const T* c3[3];
  T *Pc[3], *Pp[3];
};

enum { P_EZ_Y, P_EY_Z, P_EX_Z, P_EZ_X, P_EY_X, P_EX_Y, P_HZ_Y, P_HY_Z, P_HX_Z, P_HZ_X, P_HY_X, P_HX_Y };

// The CPML curl term d + psi' with psi' = b psi + a d inside a shell, d
// outside (stepper_3d.py:368-370), split into its loads (ce_load) and its
// arithmetic + psi store (ce_apply), so that a kernel can issue every load
// of a cell before its first store: the field, psi and profile arrays never
// alias, but the compiler cannot know that and would otherwise keep each
// component's loads behind the previous component's stores.
template <typename T>
struct CeIn {
  bool in;
  T b, a, ps;
};

template <typename T>
__device__ __forceinline__ CeIn<T> ce_load(const Pml<T>& pr, int idx, bool use, const T* psi, long long q) {
  CeIn<T> c{use && pr.in(idx), T(0), T(0), T(0)};
  if (c.in) {
    c.b = pr.b[idx];
    c.a = pr.a[idx];
    c.ps = psi[q];
  }
  return c;
}

template <typename T>
__device__ __forceinline__ T ce_apply(const CeIn<T>& c, T* psi, long long q, T d) {
  if (!c.in) return d;
  T ps = c.b * c.ps + c.a * d;
  psi[q] = ps;
  return d + ps;
}

template <typename T, bool M>
__global__ __launch_bounds__(BZ* BYY) void h3_kernel(K3<T> p) {
  const int k = blockIdx.x * BZ + threadIdx.x;
  const int j = blockIdx.y * BYY + threadIdx.y;
  const int i = blockIdx.z;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (j >= ny || k >= nz) return;
  const T dx = p.dx, dt = p.dt;
  const bool fx = j < ny - 1 && k < nz - 1;                    // Hx (nx, ny-1, nz-1)
  const bool fy = i < nx - 1 && k < nz - 1;                    // Hy (nx-1, ny, nz-1)
  const bool fz = i < nx - 1 && j < ny - 1;                    // Hz (nx-1, ny-1, nz)
  const long long ix = ((long long)i * (ny - 1) + j) * (nz - 1) + k;  // Hx
  const long long iy = ((long long)i * ny + j) * (nz - 1) + k;        // Hy, Ez (i, j, k)
  const long long iz = ((long long)i * (ny - 1) + j) * nz + k;        // Hz, Ey (i, j, k)
  const long long ie = ((long long)i * ny + j) * nz + k;              // Ex (i, j, k)
  // ---- loads
  T ez0 = T(0), ezj = T(0), ezx = T(0), ey0 = T(0), eyk = T(0), eyx = T(0), ex0 = T(0), exk = T(0), exj = T(0);
  T hx = T(0), hy = T(0), hz = T(0);
  if (fx || fy) ez0 = p.Ez[iy];
  if (fx || fz) ey0 = p.Ey[iz];
  if (fy || fz) ex0 = p.Ex[ie];
  if (fx) {
    ezj = p.Ez[iy + (nz - 1)];
    eyk = p.Ey[iz + 1];
    hx = p.Hx[ix];
  }
  if (fy) {
    exk = p.Ex[ie + 1];
    ezx = p.Ez[iy + (long long)ny * (nz - 1)];
    hy = p.Hy[iy];
  }
  if (fz) {
    eyx = p.Ey[iz + (long long)(ny - 1) * nz];
    exj = p.Ex[ie + nz];
    hz = p.Hz[iz];
  }
  const CeIn<T> cxy = ce_load(p.h[1], j, fx, p.psi[P_EZ_Y], ix), cxz = ce_load(p.h[2], k, fx, p.psi[P_EY_Z], ix);
  const CeIn<T> cyz = ce_load(p.h[2], k, fy, p.psi[P_EX_Z], iy), cyx = ce_load(p.h[0], i, fy, p.psi[P_EZ_X], iy);
  const CeIn<T> czx = ce_load(p.h[0], i, fz, p.psi[P_EY_X], iz), czy = ce_load(p.h[1], j, fz, p.psi[P_EX_Y], iz);
  // conductivity pairs: H = d1 H - d2 curl (stepper_3d.py:375,382,389); at
  // sigma = 0 the reference's pairs are exactly (1, dt)
  T d1x = T(1), d2x = dt, d1y = T(1), d2y = dt, d1z = T(1), d2z = dt;
  if constexpr (M) {
    if (fx) d1x = p.dmp[6][ix], d2x = p.dmp[7][ix];
    if (fy) d1y = p.dmp[8][iy], d2y = p.dmp[9][iy];
    if (fz) d1z = p.dmp[10][iz], d2z = p.dmp[11][iz];
  }
  // ---- arithmetic and stores (stepper_3d.py:363-390)
  if (fx) {
    T cy = ce_apply(cxy, p.psi[P_EZ_Y], ix, (ezj - ez0) / dx);
    T cz = ce_apply(cxz, p.psi[P_EY_Z], ix, (eyk - ey0) / dx);
    p.Hx[ix] = M ? d1x * hx - d2x * (cy - cz) : hx - dt * (cy - cz);
  }
  if (fy) {
    T cz = ce_apply(cyz, p.psi[P_EX_Z], iy, (exk - ex0) / dx);
    T cx = ce_apply(cyx, p.psi[P_EZ_X], iy, (ezx - ez0) / dx);
    p.Hy[iy] = M ? d1y * hy - d2y * (cz - cx) : hy - dt * (cz - cx);
  }
  if (fz) {
    T cx = ce_apply(czx, p.psi[P_EY_X], iz, (eyx - ey0) / dx);
    T cy = ce_apply(czy, p.psi[P_EX_Y], iz, (exj - ex0) / dx);
    p.Hz[iz] = M ? d1z * hz - d2z * (cx - cy) : hz - dt * (cx - cy);
  }
}

template <typename T, typename P>
__device__ __forceinline__ T e_update(const P& p, T E, T eps_c, T curl) {
  if (p.lean) {
    T inv = T(1) / eps_c;
    return E + (curl * inv) * p.dt;
  }
  return (E * eps_c + p.dt * curl) / eps_c;
}

// E update with media (stepper_3d.py:295-301,318-329,402-437): the poles
// advance from the incoming E(n), P' = (c1 P + c2 P_prev) + c3 E(n), then
//   E = (d1 (E eps_c + sum P) + d2 curl - sum P') / eps_c
// (without poles E = (d1 E eps_c + d2 curl) / eps_c); q is the cell's index
// in component a's grid of nq cells.
template <typename T>
__device__ __forceinline__ T e_media(const K3<T>& p, int a, long long q, long long nq, T E, T eps_c, T curl) {
  const T d1 = p.dmp[2 * a][q], d2 = p.dmp[2 * a + 1][q];
  if (p.np == 0) return (d1 * E * eps_c + d2 * curl) / eps_c;
  T sc = T(0), sn = T(0);
  for (int n = 0; n < p.np; ++n) {
    const long long o = (long long)n * nq + q;
    const T pc = p.Pc[a][o], pp = p.Pp[a][o];
    const T pn = (p.c1[n] * pc + p.c2[n] * pp) + p.c3[a][o] * E;
    p.Pp[a][o] = pc;
    p.Pc[a][o] = pn;
    sc = n ? sc + pc : pc;
    sn = n ? sn + pn : pn;
  }
  return (d1 * (E * eps_c + sc) + d2 * curl - sn) / eps_c;
}

template <typename T, typename P>
__device__ __forceinline__ T inject(const P& p, const mt_source3d& s, int comp, int i, int j, int k,
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

template <typename T, bool M>
__global__ __launch_bounds__(BZ* BYY) void e3_kernel(K3<T> p, mt_source3d src, T c, T* tape) {
  const int k = blockIdx.x * BZ + threadIdx.x;
  const int j = blockIdx.y * BYY + threadIdx.y;
  const int i = blockIdx.z;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (j >= ny || k >= nz) return;
  const T dx = p.dx;
  const bool ax = i < nx - 1, ay = j < ny - 1, az = k < nz - 1;  // Ex, Ey, Ez exist at this cell
  const bool ix = i >= 1 && ax, jy = j >= 1 && ay, kz = k >= 1 && az;
  const long long ce_ = ((long long)i * ny + j) * nz + k;                 // eps, Ex
  const long long iy = ((long long)i * (ny - 1) + j) * nz + k;            // Ey, Hz (i, j, k)
  const long long iz = ((long long)i * ny + j) * (nz - 1) + k;            // Ez, Hy (i, j, k)
  const long long ihx = ((long long)i * (ny - 1) + j) * (nz - 1) + k;     // Hx (i, j, k)
  // ---- loads (every one before the first store; see ce_load)
  const T eps0 = p.eps[ce_];
  T epx = T(0), epy = T(0), epz = T(0), ex = T(0), ey = T(0), ez = T(0);
  T hz0 = T(0), hzj = T(0), hzx = T(0), hy0 = T(0), hyk = T(0), hyx = T(0), hx0 = T(0), hxk = T(0), hxj = T(0);
  if (ax) {
    epx = p.eps[ce_ + (long long)ny * nz];
    ex = p.Ex[ce_];
  }
  if (ay) {
    epy = p.eps[ce_ + nz];
    ey = p.Ey[iy];
  }
  if (az) {
    epz = p.eps[ce_ + 1];
    ez = p.Ez[iz];
  }
  if (ax && ay) hz0 = p.Hz[iy];
  if (ax && jy) hzj = p.Hz[iy - nz];
  if (ay && ix) hzx = p.Hz[iy - (long long)(ny - 1) * nz];
  if (ax && az) hy0 = p.Hy[iz];
  if (ax && kz) hyk = p.Hy[iz - 1];
  if (az && ix) hyx = p.Hy[iz - (long long)ny * (nz - 1)];
  if (ay && az) hx0 = p.Hx[ihx];
  if (ay && kz) hxk = p.Hx[ihx - 1];
  if (az && jy) hxj = p.Hx[ihx - (nz - 1)];
  const long long q_xy = ((long long)i * (ny - 2) + (j - 1)) * nz + k;
  const long long q_xz = ((long long)i * ny + j) * (nz - 2) + (k - 1);
  const long long q_yz = ((long long)i * (ny - 1) + j) * (nz - 2) + (k - 1);
  const long long q_yx = ((long long)(i - 1) * (ny - 1) + j) * nz + k;
  const long long q_zx = ((long long)(i - 1) * ny + j) * (nz - 1) + k;
  const long long q_zy = ((long long)i * (ny - 2) + (j - 1)) * (nz - 1) + k;
  const CeIn<T> cxy = ce_load(p.e[1], j - 1, ax && jy, p.psi[P_HZ_Y], q_xy);
  const CeIn<T> cxz = ce_load(p.e[2], k - 1, ax && kz, p.psi[P_HY_Z], q_xz);
  const CeIn<T> cyz = ce_load(p.e[2], k - 1, ay && kz, p.psi[P_HX_Z], q_yz);
  const CeIn<T> cyx = ce_load(p.e[0], i - 1, ay && ix, p.psi[P_HZ_X], q_yx);
  const CeIn<T> czx = ce_load(p.e[0], i - 1, az && ix, p.psi[P_HY_X], q_zx);
  const CeIn<T> czy = ce_load(p.e[1], j - 1, az && jy, p.psi[P_HX_Y], q_zy);
  // ---- arithmetic and stores (stepper_3d.py:392-445)
  if (ax) {  // Ex (nx-1, ny, nz)
    T chy = T(0), chz = T(0);
    if (jy) chy = ce_apply(cxy, p.psi[P_HZ_Y], q_xy, (hz0 - hzj) / dx);
    if (kz) chz = ce_apply(cxz, p.psi[P_HY_Z], q_xz, (hy0 - hyk) / dx);
    T curl = (jy && kz) ? chy - chz : T(0);
    T epsc = T(0.5) * (eps0 + epx);
    T E = M ? e_media(p, 0, ce_, (long long)(nx - 1) * ny * nz, ex, epsc, curl) : e_update(p, ex, epsc, curl);
    E = inject(p, src, 0, i, j, k, E, epsc, c);
    p.Ex[ce_] = E;
    if (tape) tape[ce_] = E;
  }
  if (ay) {  // Ey (nx, ny-1, nz)
    T chz = T(0), chx = T(0);
    if (kz) chz = ce_apply(cyz, p.psi[P_HX_Z], q_yz, (hx0 - hxk) / dx);
    if (ix) chx = ce_apply(cyx, p.psi[P_HZ_X], q_yx, (hz0 - hzx) / dx);
    T curl = (kz && ix) ? chz - chx : T(0);
    T epsc = T(0.5) * (eps0 + epy);
    T E = M ? e_media(p, 1, iy, (long long)nx * (ny - 1) * nz, ey, epsc, curl) : e_update(p, ey, epsc, curl);
    E = inject(p, src, 1, i, j, k, E, epsc, c);
    p.Ey[iy] = E;
    if (tape) tape[(long long)(nx - 1) * ny * nz + iy] = E;
  }
  if (az) {  // Ez (nx, ny, nz-1)
    T chx = T(0), chy = T(0);
    if (ix) chx = ce_apply(czx, p.psi[P_HY_X], q_zx, (hy0 - hyx) / dx);
    if (jy) chy = ce_apply(czy, p.psi[P_HX_Y], q_zy, (hx0 - hxj) / dx);
    T curl = (ix && jy) ? chx - chy : T(0);
    T epsc = T(0.5) * (eps0 + epz);
    T E = M ? e_media(p, 2, iz, (long long)nx * ny * (nz - 1), ez, epsc, curl) : e_update(p, ez, epsc, curl);
    E = inject(p, src, 2, i, j, k, E, epsc, c);
    p.Ez[iz] = E;
    if (tape) tape[(long long)(nx - 1) * ny * nz + (long long)nx * (ny - 1) * nz + iz] = E;
  }
}

// ---------------------------------------------------------------------------
// Fused single-pass step, streaming along x.
//
// A workgroup of 64 x TJ threads owns the (j, k) column block
// [j0, j0+TJ) x [k0, k0+64) of one x segment and marches over its planes.
// Per plane i it computes H'(i) for its cells plus the low halo row
// (j0-1: Hx', Hz') and column (k0-1: Hx', Hy') that its E' update reads,
// puts them in LDS, then computes E'(i).  The x differences of E' use the
// thread's own H'(i-1), kept in registers.  Old E and eps planes i, i+1 are
// staged in LDS (three rotating buffers, halo included), so each value
// comes from memory once per workgroup.  Fields and psi are read from the
// old buffers and written to the new ones (ping-pong): halo values are
// recomputed, never read from a neighbour's output, so workgroups need no
// ordering.  A segment starting at i_begin > 0 first recomputes H'(i_begin-1)
// without writing it.  The arithmetic of every update is the two-pass
// kernels' expression for expression, so results are bit-identical.
// ---------------------------------------------------------------------------
template <typename T>
struct F3 {
  int nx, ny, nz, lean, i_seg;
  T dt, dx;
  const T* eps;
  Pml<T> h[3], e[3];
  const T *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;
  T *Exn, *Eyn, *Ezn, *Hxn, *Hyn, *Hzn;
  const T* psi[12];
  T* psin[12];
};

// The CPML curl term with the psi recursion reading the old buffer and (if W) writing the new one
template <typename T, bool W>
__device__ __forceinline__ T ce2(const Pml<T>& pr, int idx, const T* ps, T* pn, long long q, T d) {
  if (!pr.in(idx)) return d;
  T v = pr.b[idx] * ps[q] + pr.a[idx] * d;
  if (W) pn[q] = v;
  return d + v;
}

constexpr int FK = 64;

template <typename T, int TJ>
struct FusedSmem {
  T e[3][3][TJ + 2][FK + 2];  // [buffer][Ex,Ey,Ez][j - (j0-1)][k - (k0-1)]
  T eps[3][TJ + 1][FK + 1];   // [buffer][j - j0][k - k0]
  T h[3][TJ + 1][FK + 1];     // H' of the plane: [Hx,Hy,Hz][j - (j0-1)][k - (k0-1)]
};

// Plane staging: each thread owns NQ fixed slots of the (TJ+2) x (FK+2) E
// region and of the (TJ+1) x (FK+1) eps region; loads go to registers first
// (issued a plane ahead) and are stored to LDS after the current plane's
// compute, so memory latency overlaps the arithmetic.
constexpr int kPlaneQ = 2;

template <typename T, int TJ>
struct PlaneRegs {
  T e[kPlaneQ][3];
  T eps[kPlaneQ];
};

template <typename T, int TJ>
__device__ __forceinline__ void fused_fetch_plane(const F3<T>& p, PlaneRegs<T, TJ>& pr, int ip, int j0, int k0) {
  const int nt = FK * TJ, tid = threadIdx.y * FK + threadIdx.x;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  static_assert((TJ + 2) * (FK + 2) <= kPlaneQ * FK * TJ, "plane slots");
  static_assert((TJ + 1) * (FK + 1) <= kPlaneQ * FK * TJ, "eps slots");
#pragma unroll
  for (int q = 0; q < kPlaneQ; ++q) {
    const int e = tid + q * nt;
    T a = T(0), b = T(0), c = T(0), d = T(0);
    if (e < (TJ + 2) * (FK + 2)) {
      const int r = e / (FK + 2), cc = e - r * (FK + 2);
      const int j = j0 - 1 + r, k = k0 - 1 + cc;
      if (j >= 0 && j < ny && k >= 0 && k < nz) {
        if (ip < nx - 1) a = p.Ex[((long long)ip * ny + j) * nz + k];
        if (j < ny - 1) b = p.Ey[((long long)ip * (ny - 1) + j) * nz + k];
        if (k < nz - 1) c = p.Ez[((long long)ip * ny + j) * (nz - 1) + k];
      }
    }
    if (e < (TJ + 1) * (FK + 1)) {
      const int r = e / (FK + 1), cc = e - r * (FK + 1);
      const int j = j0 + r, k = k0 + cc;
      if (j < ny && k < nz) d = p.eps[((long long)ip * ny + j) * nz + k];
    }
    pr.e[q][0] = a;
    pr.e[q][1] = b;
    pr.e[q][2] = c;
    pr.eps[q] = d;
  }
}

template <typename T, int TJ>
__device__ __forceinline__ void fused_store_plane(FusedSmem<T, TJ>& sm, const PlaneRegs<T, TJ>& pr, int buf) {
  const int nt = FK * TJ, tid = threadIdx.y * FK + threadIdx.x;
#pragma unroll
  for (int q = 0; q < kPlaneQ; ++q) {
    const int e = tid + q * nt;
    if (e < (TJ + 2) * (FK + 2)) {
      const int r = e / (FK + 2), cc = e - r * (FK + 2);
      sm.e[buf][0][r][cc] = pr.e[q][0];
      sm.e[buf][1][r][cc] = pr.e[q][1];
      sm.e[buf][2][r][cc] = pr.e[q][2];
    }
    if (e < (TJ + 1) * (FK + 1)) {
      const int r = e / (FK + 1), cc = e - r * (FK + 1);
      sm.eps[buf][r][cc] = pr.eps[q];
    }
  }
}

// Per-plane views: base pointers of plane ip of every array (uniform across
// the workgroup) so that per-thread offsets stay 32-bit.  Valid for arrays
// below 2^31 elements per plane set (checked on the host).
template <typename T>
struct PlaneP {
  const T *hx, *hy, *hz;        // old H, plane ip
  T *hxn, *hyn, *hzn;           // new H, plane ip
  const T *ps[12];              // old psi at the plane the update indexes
  T* pn[12];
};

template <typename T>
__device__ __forceinline__ PlaneP<T> plane_ptrs(const F3<T>& p, int ip) {
  const long long nx = p.nx, ny = p.ny, nz = p.nz;
  PlaneP<T> q;
  const long long sHx = (ny - 1) * (nz - 1), sHy = ny * (nz - 1), sHz = (ny - 1) * nz;
  q.hx = p.Hx + ip * sHx;
  q.hy = p.Hy + ip * sHy;
  q.hz = p.Hz + ip * sHz;
  q.hxn = p.Hxn + ip * sHx;
  q.hyn = p.Hyn + ip * sHy;
  q.hzn = p.Hzn + ip * sHz;
  // H-side psi share the H component's plane stride
  const long long sh[6] = {sHx, sHx, sHy, sHy, sHz, sHz};
  for (int a = 0; a < 6; ++a) {
    q.ps[a] = p.psi[a] + ip * sh[a];
    q.pn[a] = p.psin[a] + ip * sh[a];
  }
  // E-side psi: plane ip (y/z shells) or ip-1 (x shells)
  const long long se[6] = {(ny - 2) * nz, ny * (nz - 2), (ny - 1) * (nz - 2), (ny - 1) * nz, ny * (nz - 1),
                           (ny - 2) * (nz - 1)};
  const long long pl[6] = {ip, ip, ip, ip - 1, ip - 1, ip};
  for (int a = 0; a < 6; ++a) {
    q.ps[6 + a] = p.psi[6 + a] + pl[a] * se[a];
    q.pn[6 + a] = p.psin[6 + a] + pl[a] * se[a];
  }
  return q;
}

// H' components of cell (ip, j, k) from its old value hold; sm.e planes ip
// (cur) and ip+1 (nxt) in LDS coordinates (r, c) = (j - (j0-1), k - (k0-1)).
template <typename T, int TJ, bool W>
__device__ __forceinline__ T fused_hx(const F3<T>& p, const PlaneP<T>& q, const FusedSmem<T, TJ>& sm, int cur,
                                      int j, int k, int r, int c, T hold) {
  const int h = j * (p.nz - 1) + k;
  T cy = ce2<T, W>(p.h[1], j, q.ps[0], q.pn[0], h, (sm.e[cur][2][r + 1][c] - sm.e[cur][2][r][c]) / p.dx);
  T cz = ce2<T, W>(p.h[2], k, q.ps[1], q.pn[1], h, (sm.e[cur][1][r][c + 1] - sm.e[cur][1][r][c]) / p.dx);
  T v = hold - p.dt * (cy - cz);
  if (W) q.hxn[h] = v;
  return v;
}

template <typename T, int TJ, bool W>
__device__ __forceinline__ T fused_hy(const F3<T>& p, const PlaneP<T>& q, const FusedSmem<T, TJ>& sm, int cur,
                                      int nxt, int ip, int j, int k, int r, int c, T hold) {
  const int h = j * (p.nz - 1) + k;
  T cz = ce2<T, W>(p.h[2], k, q.ps[2], q.pn[2], h, (sm.e[cur][0][r][c + 1] - sm.e[cur][0][r][c]) / p.dx);
  T cx = ce2<T, W>(p.h[0], ip, q.ps[3], q.pn[3], h, (sm.e[nxt][2][r][c] - sm.e[cur][2][r][c]) / p.dx);
  T v = hold - p.dt * (cz - cx);
  if (W) q.hyn[h] = v;
  return v;
}

template <typename T, int TJ, bool W>
__device__ __forceinline__ T fused_hz(const F3<T>& p, const PlaneP<T>& q, const FusedSmem<T, TJ>& sm, int cur,
                                      int nxt, int ip, int j, int k, int r, int c, T hold) {
  const int h = j * p.nz + k;
  T cx = ce2<T, W>(p.h[0], ip, q.ps[4], q.pn[4], h, (sm.e[nxt][1][r][c] - sm.e[cur][1][r][c]) / p.dx);
  T cy = ce2<T, W>(p.h[1], j, q.ps[5], q.pn[5], h, (sm.e[cur][0][r + 1][c] - sm.e[cur][0][r][c]) / p.dx);
  T v = hold - p.dt * (cx - cy);
  if (W) q.hzn[h] = v;
  return v;
}

// Old H values one thread needs for plane ip: its own cell, and (halo
// lanes) the low-row / low-column cells it computes.
template <typename T>
struct HoldRegs {
  T x, y, z, a, b;  // own Hx, Hy, Hz; halo: (row) Hx, Hz or (column) Hx, Hy
};

template <typename T, int TJ>
__device__ __forceinline__ void fused_fetch_h(const F3<T>& p, HoldRegs<T>& h, int ip, int j0, int k0, int wcol) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int tx = threadIdx.x, ty = threadIdx.y, j = j0 + ty, k = k0 + tx;
  h = HoldRegs<T>{T(0), T(0), T(0), T(0), T(0)};
  if (ip >= nx) return;
  const T* Hx = p.Hx + (long long)ip * (ny - 1) * (nz - 1);
  const T* Hy = p.Hy + (long long)ip * ny * (nz - 1);
  const T* Hz = p.Hz + (long long)ip * (ny - 1) * nz;
  if (j < ny && k < nz) {
    if (j < ny - 1 && k < nz - 1) h.x = Hx[j * (nz - 1) + k];
    if (ip < nx - 1 && k < nz - 1) h.y = Hy[j * (nz - 1) + k];
    if (ip < nx - 1 && j < ny - 1) h.z = Hz[j * nz + k];
  }
  if (ty == 0 && j0 >= 1 && k < nz) {
    const int jh = j0 - 1;
    if (k < nz - 1) h.a = Hx[jh * (nz - 1) + k];
    if (ip < nx - 1) h.b = Hz[jh * nz + k];
  } else if (ty == wcol && tx < TJ && k0 >= 1 && j0 + tx < ny) {
    const int jc = j0 + tx, kh = k0 - 1;
    if (jc < ny - 1) h.a = Hx[jc * (nz - 1) + kh];
    if (ip < nx - 1) h.b = Hy[jc * (nz - 1) + kh];
  }
}

template <typename T, int TJ>
__global__ __launch_bounds__(FK* TJ, 4) void fused3_kernel(F3<T> p, mt_source3d src, T c_src, T* tape) {
  __shared__ FusedSmem<T, TJ> sm;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int k0 = blockIdx.x * FK, j0 = blockIdx.y * TJ;
  const int k = k0 + tx, j = j0 + ty;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const int i_begin = blockIdx.z * p.i_seg;
  const int i_end = min(nx, i_begin + p.i_seg);
  const int i_first = i_begin > 0 ? i_begin - 1 : 0;
  const bool own = j < ny && k < nz;
  const int r = ty + 1, c = tx + 1;
  const int wcol = TJ >= 2 ? 1 : 0;  // wave computing the halo column
  const T dx = p.dx;
  T hyp = T(0), hzp = T(0);  // own H'(i-1)

  PlaneRegs<T, TJ> pre;
  fused_fetch_plane<T, TJ>(p, pre, i_first, j0, k0);
  fused_store_plane<T, TJ>(sm, pre, i_first % 3);
  if (i_first + 1 < nx) {
    fused_fetch_plane<T, TJ>(p, pre, i_first + 1, j0, k0);
    fused_store_plane<T, TJ>(sm, pre, (i_first + 1) % 3);
  }
  HoldRegs<T> hold;
  fused_fetch_h<T, TJ>(p, hold, i_first, j0, k0, wcol);
  __syncthreads();
  for (int ip = i_first; ip < i_end; ++ip) {
    const int cur = ip % 3, nxt = (ip + 1) % 3;
    const bool wr = ip >= i_begin;
    const bool more = ip + 2 < nx;
    const PlaneP<T> q = plane_ptrs(p, ip);
    // ---- H'(ip): own cell, low halo row (j0-1), low halo column (k0-1)
    T hx = T(0), hy = T(0), hz = T(0);
    if (own) {
      if (j < ny - 1 && k < nz - 1)
        hx = wr ? fused_hx<T, TJ, true>(p, q, sm, cur, j, k, r, c, hold.x)
                : fused_hx<T, TJ, false>(p, q, sm, cur, j, k, r, c, hold.x);
      if (ip < nx - 1 && k < nz - 1)
        hy = wr ? fused_hy<T, TJ, true>(p, q, sm, cur, nxt, ip, j, k, r, c, hold.y)
                : fused_hy<T, TJ, false>(p, q, sm, cur, nxt, ip, j, k, r, c, hold.y);
      if (ip < nx - 1 && j < ny - 1)
        hz = wr ? fused_hz<T, TJ, true>(p, q, sm, cur, nxt, ip, j, k, r, c, hold.z)
                : fused_hz<T, TJ, false>(p, q, sm, cur, nxt, ip, j, k, r, c, hold.z);
      sm.h[0][r][c] = hx;
      sm.h[1][r][c] = hy;
      sm.h[2][r][c] = hz;
    }
    if (ty == 0 && j0 >= 1 && k < nz) {  // halo row j0-1: Hx', Hz'
      const int jh = j0 - 1;
      if (k < nz - 1) sm.h[0][0][c] = fused_hx<T, TJ, false>(p, q, sm, cur, jh, k, 0, c, hold.a);
      if (ip < nx - 1) sm.h[2][0][c] = fused_hz<T, TJ, false>(p, q, sm, cur, nxt, ip, jh, k, 0, c, hold.b);
    } else if (ty == wcol && tx < TJ && k0 >= 1 && j0 + tx < ny) {  // halo column k0-1: Hx', Hy'
      const int jc = j0 + tx, kh = k0 - 1, rc = tx + 1;
      if (jc < ny - 1) sm.h[0][rc][0] = fused_hx<T, TJ, false>(p, q, sm, cur, jc, kh, rc, 0, hold.a);
      if (ip < nx - 1) sm.h[1][rc][0] = fused_hy<T, TJ, false>(p, q, sm, cur, nxt, ip, jc, kh, rc, 0, hold.b);
    }
    __syncthreads();
    // prefetch the old H of the next plane and the E / eps plane after it
    fused_fetch_h<T, TJ>(p, hold, ip + 1, j0, k0, wcol);
    if (more) fused_fetch_plane<T, TJ>(p, pre, ip + 2, j0, k0);
    // ---- E'(ip) of the own cell
    if (own && wr) {
      const T eps0 = sm.eps[cur][ty][tx];
      if (ip < nx - 1) {  // Ex
        const int x = j * nz + k;
        T chy = T(0), chz = T(0);
        const bool jy = j >= 1 && j < ny - 1, kz = k >= 1 && k < nz - 1;
        if (jy) chy = ce2<T, true>(p.e[1], j - 1, q.ps[6], q.pn[6], (j - 1) * nz + k, (sm.h[2][r][c] - sm.h[2][r - 1][c]) / dx);
        if (kz) chz = ce2<T, true>(p.e[2], k - 1, q.ps[7], q.pn[7], j * (nz - 2) + (k - 1), (sm.h[1][r][c] - sm.h[1][r][c - 1]) / dx);
        T curl = (jy && kz) ? chy - chz : T(0);
        T epsc = T(0.5) * (eps0 + sm.eps[nxt][ty][tx]);
        T E = e_update(p, sm.e[cur][0][r][c], epsc, curl);
        E = inject(p, src, 0, ip, j, k, E, epsc, c_src);
        const long long o = (long long)ip * ny * nz + x;
        p.Exn[o] = E;
        if (tape) tape[o] = E;
      }
      if (j < ny - 1) {  // Ey
        const int y = j * nz + k;
        T chz = T(0), chx = T(0);
        const bool kz = k >= 1 && k < nz - 1, ix = ip >= 1 && ip < nx - 1;
        if (kz) chz = ce2<T, true>(p.e[2], k - 1, q.ps[8], q.pn[8], j * (nz - 2) + (k - 1), (sm.h[0][r][c] - sm.h[0][r][c - 1]) / dx);
        if (ix) chx = ce2<T, true>(p.e[0], ip - 1, q.ps[9], q.pn[9], y, (hz - hzp) / dx);
        T curl = (kz && ix) ? chz - chx : T(0);
        T epsc = T(0.5) * (eps0 + sm.eps[cur][ty + 1][tx]);
        T E = e_update(p, sm.e[cur][1][r][c], epsc, curl);
        E = inject(p, src, 1, ip, j, k, E, epsc, c_src);
        const long long o = (long long)ip * (ny - 1) * nz + y;
        p.Eyn[o] = E;
        if (tape) tape[(long long)(nx - 1) * ny * nz + o] = E;
      }
      if (k < nz - 1) {  // Ez
        const int z = j * (nz - 1) + k;
        T chx = T(0), chy = T(0);
        const bool ix = ip >= 1 && ip < nx - 1, jy = j >= 1 && j < ny - 1;
        if (ix) chx = ce2<T, true>(p.e[0], ip - 1, q.ps[10], q.pn[10], z, (hy - hyp) / dx);
        if (jy) chy = ce2<T, true>(p.e[1], j - 1, q.ps[11], q.pn[11], (j - 1) * (nz - 1) + k, (sm.h[0][r][c] - sm.h[0][r - 1][c]) / dx);
        T curl = (ix && jy) ? chx - chy : T(0);
        T epsc = T(0.5) * (eps0 + sm.eps[cur][ty][tx + 1]);
        T E = e_update(p, sm.e[cur][2][r][c], epsc, curl);
        E = inject(p, src, 2, ip, j, k, E, epsc, c_src);
        const long long o = (long long)ip * ny * (nz - 1) + z;
        p.Ezn[o] = E;
        if (tape) tape[(long long)(nx - 1) * ny * nz + (long long)nx * (ny - 1) * nz + o] = E;
      }
    }
    hyp = hy;
    hzp = hz;
    if (more) fused_store_plane<T, TJ>(sm, pre, (ip + 2) % 3);
    __syncthreads();
  }
}

// Monitor descriptors with every field term moved from the primary buffers
// to the matching alt buffers (same offsets).
void retarget(const mt_monitor* in, int n, mt_monitor* out, void* const* prim, const long long* bytes,
              void* const* alt) {
  for (int m = 0; m < n; ++m) {
    out[m] = in[m];
    for (int q = 0; q < in[m].n_terms; ++q) {
      const void* ptrs[2] = {in[m].term[q].a, in[m].term[q].b};
      for (int w = 0; w < 2; ++w) {
        if (!ptrs[w]) continue;
        const char* a = static_cast<const char*>(ptrs[w]);
        for (int f = 0; f < 18; ++f) {
          const char* b0 = static_cast<const char*>(prim[f]);
          if (a >= b0 && a < b0 + bytes[f]) {
            const void* moved = static_cast<const char*>(alt[f]) + (a - b0);
            if (w == 0) out[m].term[q].a = moved;
            else out[m].term[q].b = moved;
            break;
          }
        }
      }
    }
  }
}

template <typename T, int TJ>
int run3_fused(const mt_yee3d* d, const mt_source3d& sv, const double* coef, const MonitorPlan& mp, int step0,
               int k, hipStream_t s) {
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  void* prim[18] = {d->Ex, d->Ey, d->Ez, d->Hx, d->Hy, d->Hz};
  for (int q = 0; q < 12; ++q) prim[6 + q] = d->psi[q];
  long long n[18];
  const long long nE[3] = {(long long)(nx - 1) * ny * nz, (long long)nx * (ny - 1) * nz, (long long)nx * ny * (nz - 1)};
  const long long nH[3] = {(long long)nx * (ny - 1) * (nz - 1), (long long)(nx - 1) * ny * (nz - 1),
                           (long long)(nx - 1) * (ny - 1) * nz};
  const long long nP[12] = {nH[0], nH[0], nH[1], nH[1], nH[2], nH[2],
                            (long long)(nx - 1) * (ny - 2) * nz, (long long)(nx - 1) * ny * (nz - 2),
                            (long long)nx * (ny - 1) * (nz - 2), (long long)(nx - 2) * (ny - 1) * nz,
                            (long long)(nx - 2) * ny * (nz - 1), (long long)nx * (ny - 2) * (nz - 1)};
  for (int q = 0; q < 3; ++q) {
    n[q] = nE[q];
    n[3 + q] = nH[q];
  }
  for (int q = 0; q < 12; ++q) n[6 + q] = nP[q];
  long long bytes[18];
  for (int q = 0; q < 18; ++q) bytes[q] = n[q] * (long long)sizeof(T);
  std::vector<mt_monitor> alt_mons(mp.n_mons > 0 ? mp.n_mons : 1);
  if (mp.n_mons) retarget(mp.mons, mp.n_mons, alt_mons.data(), prim, bytes, d->alt);
  MonitorPlan mpa = mp;
  mpa.mons = alt_mons.data();

  F3<T> p;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.lean = d->lean;
  p.dt = static_cast<T>(d->dt);
  p.dx = static_cast<T>(d->dx);
  p.eps = static_cast<const T*>(d->eps);
  for (int a = 0; a < 3; ++a) {
    p.h[a] = pml_of<T>(d->h[a]);
    p.e[a] = pml_of<T>(d->e[a]);
  }
  const int gk = (nz + FK - 1) / FK, gj = (ny + TJ - 1) / TJ;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  int nseg = (4 * cus + gk * gj - 1) / (gk * gj);
  nseg = std::max(1, std::min(nseg, nx / 16));
  p.i_seg = (nx + nseg - 1) / nseg;
  nseg = (nx + p.i_seg - 1) / p.i_seg;
  dim3 blk(FK, TJ, 1), grd(gk, gj, nseg);
  const long long e_cells = nE[0] + nE[1] + nE[2];
  int S = mp.stage_steps > 0 ? mp.stage_steps : 1;
  int filled = 0, first = step0;
  for (int s_ = 0; s_ < k; ++s_) {
    void* const* src_b = (s_ % 2 == 0) ? prim : d->alt;
    void* const* dst_b = (s_ % 2 == 0) ? d->alt : prim;
    p.Ex = static_cast<const T*>(src_b[0]);
    p.Ey = static_cast<const T*>(src_b[1]);
    p.Ez = static_cast<const T*>(src_b[2]);
    p.Hx = static_cast<const T*>(src_b[3]);
    p.Hy = static_cast<const T*>(src_b[4]);
    p.Hz = static_cast<const T*>(src_b[5]);
    p.Exn = static_cast<T*>(dst_b[0]);
    p.Eyn = static_cast<T*>(dst_b[1]);
    p.Ezn = static_cast<T*>(dst_b[2]);
    p.Hxn = static_cast<T*>(dst_b[3]);
    p.Hyn = static_cast<T*>(dst_b[4]);
    p.Hzn = static_cast<T*>(dst_b[5]);
    for (int q = 0; q < 12; ++q) {
      p.psi[q] = static_cast<const T*>(src_b[6 + q]);
      p.psin[q] = static_cast<T*>(dst_b[6 + q]);
    }
    T c = sv.kind ? static_cast<T>(coef[s_]) : T(0);
    T* slot = d->tape ? static_cast<T*>(d->tape) + (long long)s_ * e_cells : nullptr;
    timing_begin(s);
    hipLaunchKernelGGL((fused3_kernel<T, TJ>), grd, blk, 0, s, p, sv, c, slot);
    timing_end(s, (long long)nx * ny * nz);
    int rc = check_launch("yee3d fused step");
    if (rc) return rc;
    if (mp.n_mons) {
      if ((rc = monitors_sample(s_ % 2 == 0 ? mpa : mp, filled, s_, s))) return rc;
      if (++filled == S || s_ == k - 1) {
        if ((rc = monitors_flush(mp, filled, first, d->dt, s))) return rc;
        first += filled;
        filled = 0;
      }
    }
  }
  if (k % 2 == 1) {
    for (int q = 0; q < 18; ++q)
      if (hipMemcpyAsync(prim[q], d->alt[q], bytes[q], hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(MT_EHIP, "yee3d fused: copy-back failed");
  }
  return MT_OK;
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
  const bool media = d->damp[0] != nullptr;
  for (int q = 0; q < 12; ++q) p.dmp[q] = static_cast<const T*>(d->damp[q]);
  p.np = media ? d->n_poles : 0;
  p.c1 = static_cast<const T*>(d->c1);
  p.c2 = static_cast<const T*>(d->c2);
  for (int a = 0; a < 3; ++a) {
    p.c3[a] = static_cast<const T*>(d->c3[a]);
    p.Pc[a] = static_cast<T*>(d->P_cur[a]);
    p.Pp[a] = static_cast<T*>(d->P_prev[a]);
  }
  mt_source3d sv{};
  if (src && src->kind) sv = *src;
  if (d->variant == 1) return run3_fused<T, 8>(d, sv, coef, mp, step0, k, s);
  if (d->variant == 2) return run3_fused<T, 4>(d, sv, coef, mp, step0, k, s);
  if (d->variant == 3) return run3_fused<T, 16>(d, sv, coef, mp, step0, k, s);
  dim3 blk(BZ, BYY, 1);
  dim3 grd((p.nz + BZ - 1) / BZ, (p.ny + BYY - 1) / BYY, p.nx);
  int S = mp.stage_steps > 0 ? mp.stage_steps : 1;
  int filled = 0, first = step0;
  const long long e_cells = (long long)(p.nx - 1) * p.ny * p.nz + (long long)p.nx * (p.ny - 1) * p.nz +
                            (long long)p.nx * p.ny * (p.nz - 1);
  for (int s_ = 0; s_ < k; ++s_) {
    timing_begin(s);
    T c = sv.kind ? static_cast<T>(coef[s_]) : T(0);
    T* slot = d->tape ? static_cast<T*>(d->tape) + (long long)s_ * e_cells : nullptr;
    if (media) {
      hipLaunchKernelGGL((h3_kernel<T, true>), grd, blk, 0, s, p);
      hipLaunchKernelGGL((e3_kernel<T, true>), grd, blk, 0, s, p, sv, c, slot);
    } else {
      hipLaunchKernelGGL((h3_kernel<T, false>), grd, blk, 0, s, p);
      hipLaunchKernelGGL((e3_kernel<T, false>), grd, blk, 0, s, p, sv, c, slot);
    }
    timing_end(s, (long long)p.nx * p.ny * p.nz);
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
  if (p->variant < 0 || p->variant > 3) return fail(MT_EINVAL, "3D variant must be 0..3");
  if (p->variant >= 1 && p->variant <= 3) {
    for (int q = 0; q < 18; ++q)
      if (!p->alt[q]) return fail(MT_EINVAL, "fused 3D variant needs all 18 alt buffers");
    if (p->dtype == MT_F64 && p->variant == 3) return fail(MT_EINVAL, "3D variant 3 (16-row tiles) is FP32 only");
  }
  bool any_damp = false, all_damp = true;
  for (int q = 0; q < 12; ++q) {
    any_damp = any_damp || p->damp[q];
    all_damp = all_damp && p->damp[q];
  }
  if (any_damp && !all_damp) return fail(MT_EINVAL, "3D media need all 12 damping arrays");
  if (p->n_poles < 0 || p->n_poles > 64) return fail(MT_EINVAL, "3D n_poles must be 0..64");
  if (p->n_poles > 0 && !any_damp) return fail(MT_EINVAL, "3D poles need the damping arrays (d1 = 1, d2 = dt)");
  if (any_damp) {
    if (p->variant != 0) return fail(MT_EINVAL, "3D media run on the two-pass variant 0 only");
    if (p->lean) return fail(MT_EINVAL, "3D media are not supported by the lean stepper");
    if (p->n_poles > 0) {
      if (!p->c1 || !p->c2) return fail(MT_EINVAL, "3D poles without c1 / c2");
      for (int a = 0; a < 3; ++a)
        if (!p->c3[a] || !p->P_cur[a] || !p->P_prev[a]) return fail(MT_EINVAL, "3D poles: null c3 or P stack");
    }
  }
  MonitorPlan mp{mons, n_mons, stage, stage_bytes, stage_steps, p->dtype};
  if (n_mons && (rc = monitors_validate(mp))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == MT_F64 ? run3_t<double>(p, src, coef, mp, step0, k, s)
                            : run3_t<float>(p, src, coef, mp, step0, k, s);
}
