// 3D Yee leapfrog with CPML shells, point / masked soft sources, and (two-pass
// path) material conductivity and ADE poles.
//
// Variant 0, two passes per step: H from curl E, then E from curl H, in
// place.  One thread per cell (i, j, k) of the nx*ny*nz lattice updates every
// component whose grid covers that cell; z is the contiguous axis and runs
// along the wave.  Variants 1-3 split each step into a one-pass interior
// kernel (no psi) and the same per-cell updates over the CPML shell, on
// ping-pong field buffers.
// Arithmetic follows meep_torch/core/stepper_3d.py:363-445 op by op (lean:
// stepper_3d_lean.py:234-296), so FP32 rounds exactly like the reference.
// psi arrays keep the reference's full-grid shapes; only cells inside a CPML
// shell touch them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
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
  T rdx;  // RN(1/dx) when division by dx may take div_c's fast sequence, else 0
  const T* eps;
  Pml<T> h[3], e[3];
  T *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;      // read: E, and old H
  T *oEx, *oEy, *oEz, *oHx, *oHy, *oHz;  // written: new E, new H (== the read set in place)
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

// H half-step of one cell (i, j, k): every H component whose grid covers the
// cell.  Reads E and the old H from Ex..Hz and writes the new H to oHx..oHz
// (the same buffers for the in-place two-pass step; the ping-pong buffers of
// the split step).
// A (with M): the matched-absorber path of a run without CPML
// (stepper_3d.py:329-335): H = d1 H - (d2 (dE_a - dE_b)) / dx, one division
// per component, and no psi.
template <typename T, bool M, bool A = false>
__device__ __forceinline__ void h3_cell(const K3<T>& p, int i, int j, int k) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const T dx = p.dx, dt = p.dt, rdx = p.rdx;
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
  // conductivity pairs: H = d1 H - d2 curl (stepper_3d.py:375,382,389); at
  // sigma = 0 the reference's pairs are exactly (1, dt)
  T d1x = T(1), d2x = dt, d1y = T(1), d2y = dt, d1z = T(1), d2z = dt;
  if constexpr (M) {
    if (fx) d1x = p.dmp[6][ix], d2x = p.dmp[7][ix];
    if (fy) d1y = p.dmp[8][iy], d2y = p.dmp[9][iy];
    if (fz) d1z = p.dmp[10][iz], d2z = p.dmp[11][iz];
  }
  if constexpr (A) {
    static_assert(M, "the absorber path carries the damping pairs");
    if (fx) p.oHx[ix] = d1x * hx - div_c(d2x * ((ezj - ez0) - (eyk - ey0)), dx, rdx);
    if (fy) p.oHy[iy] = d1y * hy - div_c(d2y * ((exk - ex0) - (ezx - ez0)), dx, rdx);
    if (fz) p.oHz[iz] = d1z * hz - div_c(d2z * ((eyx - ey0) - (exj - ex0)), dx, rdx);
    return;
  }
  const CeIn<T> cxy = ce_load(p.h[1], j, fx, p.psi[P_EZ_Y], ix), cxz = ce_load(p.h[2], k, fx, p.psi[P_EY_Z], ix);
  const CeIn<T> cyz = ce_load(p.h[2], k, fy, p.psi[P_EX_Z], iy), cyx = ce_load(p.h[0], i, fy, p.psi[P_EZ_X], iy);
  const CeIn<T> czx = ce_load(p.h[0], i, fz, p.psi[P_EY_X], iz), czy = ce_load(p.h[1], j, fz, p.psi[P_EX_Y], iz);
  // ---- arithmetic and stores (stepper_3d.py:363-390)
  if (fx) {
    T cy = ce_apply(cxy, p.psi[P_EZ_Y], ix, div_c(ezj - ez0, dx, rdx));
    T cz = ce_apply(cxz, p.psi[P_EY_Z], ix, div_c(eyk - ey0, dx, rdx));
    p.oHx[ix] = M ? d1x * hx - d2x * (cy - cz) : hx - dt * (cy - cz);
  }
  if (fy) {
    T cz = ce_apply(cyz, p.psi[P_EX_Z], iy, div_c(exk - ex0, dx, rdx));
    T cx = ce_apply(cyx, p.psi[P_EZ_X], iy, div_c(ezx - ez0, dx, rdx));
    p.oHy[iy] = M ? d1y * hy - d2y * (cz - cx) : hy - dt * (cz - cx);
  }
  if (fz) {
    T cx = ce_apply(czx, p.psi[P_EY_X], iz, div_c(eyx - ey0, dx, rdx));
    T cy = ce_apply(czy, p.psi[P_EX_Y], iz, div_c(exj - ex0, dx, rdx));
    p.oHz[iz] = M ? d1z * hz - d2z * (cx - cy) : hz - dt * (cx - cy);
  }
}

template <typename T, bool M, bool A = false>
__global__ __launch_bounds__(BZ* BYY) void h3_kernel(K3<T> p) {
  const int k = blockIdx.x * BZ + threadIdx.x;
  const int j = blockIdx.y * BYY + threadIdx.y;
  if (j >= p.ny || k >= p.nz) return;
  h3_cell<T, M, A>(p, blockIdx.z, j, k);
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
// Lean form (stepper_3d_lean.py:207-231, no conductivity): the in-place
// recurrence P' = fma(c3, E, fma(c1, P, c2 P_prev)) (torch's addcmul_ rounds
// each multiply-add once), then E + ((curl dt + sum P) - sum P') / eps_c as
// a product with the reciprocal.
template <typename T>
__device__ __forceinline__ T e_media(const K3<T>& p, int a, long long q, long long nq, T E, T eps_c, T curl) {
  if (p.lean) {
    T sc = T(0), sn = T(0);
    for (int n = 0; n < p.np; ++n) {
      const long long o = (long long)n * nq + q;
      const T pc = p.Pc[a][o], pp = p.Pp[a][o];
      const T pn = fma(p.c3[a][o], E, fma(p.c1[n], pc, p.c2[n] * pp));
      p.Pp[a][o] = pc;
      p.Pc[a][o] = pn;
      sc = n ? sc + pc : pc;
      sn = n ? sn + pn : pn;
    }
    return E + (((curl * p.dt) + sc) - sn) * (T(1) / eps_c);
  }
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

// E half-step of one cell: reads the new H from oHx..oHz and the old E from
// Ex..Ez, writes the new E to oEx..oEz (same buffers when in place).  A: the
// absorber path's curl, (dH_a - dH_b) / dx with one division and no psi
// (stepper_3d.py:336-360).
template <typename T, bool M, bool A = false>
__device__ __forceinline__ void e3_cell(const K3<T>& p, const mt_source3d& src, T c, T* tape, int i, int j,
                                        int k) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const T dx = p.dx, rdx = p.rdx;
  const bool ax = i < nx - 1, ay = j < ny - 1, az = k < nz - 1;  // Ex, Ey, Ez exist at this cell
  const bool ix = i >= 1 && ax, jy = j >= 1 && ay, kz = k >= 1 && az;
  const long long ce_ = ((long long)i * ny + j) * nz + k;                 // eps, Ex
  const long long iy = ((long long)i * (ny - 1) + j) * nz + k;            // Ey, Hz (i, j, k)
  const long long iz = ((long long)i * ny + j) * (nz - 1) + k;            // Ez, Hy (i, j, k)
  const long long ihx = ((long long)i * (ny - 1) + j) * (nz - 1) + k;     // Hx (i, j, k)
  const T *Hx = p.oHx, *Hy = p.oHy, *Hz = p.oHz;
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
  if (ax && ay) hz0 = Hz[iy];
  if (ax && jy) hzj = Hz[iy - nz];
  if (ay && ix) hzx = Hz[iy - (long long)(ny - 1) * nz];
  if (ax && az) hy0 = Hy[iz];
  if (ax && kz) hyk = Hy[iz - 1];
  if (az && ix) hyx = Hy[iz - (long long)ny * (nz - 1)];
  if (ay && az) hx0 = Hx[ihx];
  if (ay && kz) hxk = Hx[ihx - 1];
  if (az && jy) hxj = Hx[ihx - (nz - 1)];
  if constexpr (A) {
    static_assert(M, "the absorber path carries the damping pairs");
    if (ax) {
      T curl = (jy && kz) ? div_c((hz0 - hzj) - (hy0 - hyk), dx, rdx) : T(0);
      T epsc = T(0.5) * (eps0 + epx);
      T E = inject(p, src, 0, i, j, k, e_media(p, 0, ce_, (long long)(nx - 1) * ny * nz, ex, epsc, curl), epsc, c);
      p.oEx[ce_] = E;
      if (tape) tape[ce_] = E;
    }
    if (ay) {
      T curl = (kz && ix) ? div_c((hx0 - hxk) - (hz0 - hzx), dx, rdx) : T(0);
      T epsc = T(0.5) * (eps0 + epy);
      T E = inject(p, src, 1, i, j, k, e_media(p, 1, iy, (long long)nx * (ny - 1) * nz, ey, epsc, curl), epsc, c);
      p.oEy[iy] = E;
      if (tape) tape[(long long)(nx - 1) * ny * nz + iy] = E;
    }
    if (az) {
      T curl = (ix && jy) ? div_c((hy0 - hyx) - (hx0 - hxj), dx, rdx) : T(0);
      T epsc = T(0.5) * (eps0 + epz);
      T E = inject(p, src, 2, i, j, k, e_media(p, 2, iz, (long long)nx * ny * (nz - 1), ez, epsc, curl), epsc, c);
      p.oEz[iz] = E;
      if (tape) tape[(long long)(nx - 1) * ny * nz + (long long)nx * (ny - 1) * nz + iz] = E;
    }
    return;
  }
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
    if (jy) chy = ce_apply(cxy, p.psi[P_HZ_Y], q_xy, div_c(hz0 - hzj, dx, rdx));
    if (kz) chz = ce_apply(cxz, p.psi[P_HY_Z], q_xz, div_c(hy0 - hyk, dx, rdx));
    T curl = (jy && kz) ? chy - chz : T(0);
    T epsc = T(0.5) * (eps0 + epx);
    T E = M ? e_media(p, 0, ce_, (long long)(nx - 1) * ny * nz, ex, epsc, curl) : e_update(p, ex, epsc, curl);
    E = inject(p, src, 0, i, j, k, E, epsc, c);
    p.oEx[ce_] = E;
    if (tape) tape[ce_] = E;
  }
  if (ay) {  // Ey (nx, ny-1, nz)
    T chz = T(0), chx = T(0);
    if (kz) chz = ce_apply(cyz, p.psi[P_HX_Z], q_yz, div_c(hx0 - hxk, dx, rdx));
    if (ix) chx = ce_apply(cyx, p.psi[P_HZ_X], q_yx, div_c(hz0 - hzx, dx, rdx));
    T curl = (kz && ix) ? chz - chx : T(0);
    T epsc = T(0.5) * (eps0 + epy);
    T E = M ? e_media(p, 1, iy, (long long)nx * (ny - 1) * nz, ey, epsc, curl) : e_update(p, ey, epsc, curl);
    E = inject(p, src, 1, i, j, k, E, epsc, c);
    p.oEy[iy] = E;
    if (tape) tape[(long long)(nx - 1) * ny * nz + iy] = E;
  }
  if (az) {  // Ez (nx, ny, nz-1)
    T chx = T(0), chy = T(0);
    if (ix) chx = ce_apply(czx, p.psi[P_HY_X], q_zx, div_c(hy0 - hyx, dx, rdx));
    if (jy) chy = ce_apply(czy, p.psi[P_HX_Y], q_zy, div_c(hx0 - hxj, dx, rdx));
    T curl = (ix && jy) ? chx - chy : T(0);
    T epsc = T(0.5) * (eps0 + epz);
    T E = M ? e_media(p, 2, iz, (long long)nx * ny * (nz - 1), ez, epsc, curl) : e_update(p, ez, epsc, curl);
    E = inject(p, src, 2, i, j, k, E, epsc, c);
    p.oEz[iz] = E;
    if (tape) tape[(long long)(nx - 1) * ny * nz + (long long)nx * (ny - 1) * nz + iz] = E;
  }
}

template <typename T, bool M, bool A = false>
__global__ __launch_bounds__(BZ* BYY) void e3_kernel(K3<T> p, mt_source3d src, T c, T* tape) {
  const int k = blockIdx.x * BZ + threadIdx.x;
  const int j = blockIdx.y * BYY + threadIdx.y;
  if (j >= p.ny || k >= p.nz) return;
  e3_cell<T, M, A>(p, src, c, tape, blockIdx.z, j, k);
}

// ---------------------------------------------------------------------------
// Split step (variants 1-3): a one-pass interior kernel over the box of cells
// whose H and E updates carry no CPML psi, and the two-pass cell updates over
// the six shell pieces around it.  Fields are ping-ponged: every kernel of a
// step reads E and H from buffer set A and writes set B (psi arrays are
// updated in place: each psi value is read and written by its own cell
// only).  Order per step: interior kernel, shell H pass (both read only A),
// then shell E pass (reads the new H in B, written by both before it).
// ---------------------------------------------------------------------------

// One shell piece: the cell box [i0, i0+ni) x [j0, j0+nj) x [k0, k0+nk),
// covered by blocks of 256 threads laid out 64 (k) x 4 (j) ("wide") or
// 16 x 16 ("narrow", for pieces thin along k); b0 = its first block.
struct Piece {
  int i0, j0, k0, ni, nj, nk, narrow, gk, gj, b0;
};

struct Pieces {
  Piece p[6];
  int n, total;
};

__device__ __forceinline__ Piece piece_of(const Pieces& P, int b) {
  // uniform selection without dynamic indexing of the argument block
  Piece q = P.p[0];
#pragma unroll
  for (int t = 1; t < 6; ++t)
    if (t < P.n && b >= P.p[t].b0) q = P.p[t];
  return q;
}

__device__ __forceinline__ bool piece_cell(const Pieces& P, int* i, int* j, int* k) {
  const Piece q = piece_of(P, blockIdx.x);
  const int b = blockIdx.x - q.b0;
  const int bk = b % q.gk, r = b / q.gk, bj = r % q.gj, bi = r / q.gj;
  const int tk = q.narrow ? (threadIdx.x & 15) : (threadIdx.x & 63);
  const int tj = q.narrow ? (threadIdx.x >> 4) : (threadIdx.x >> 6);
  const int kk = bk * (q.narrow ? 16 : 64) + tk, jj = bj * (q.narrow ? 16 : 4) + tj;
  if (kk >= q.nk || jj >= q.nj) return false;
  *i = q.i0 + bi;
  *j = q.j0 + jj;
  *k = q.k0 + kk;
  return true;
}

template <typename T>
__global__ __launch_bounds__(256) void h3_pieces(K3<T> p, Pieces P) {
  int i, j, k;
  if (piece_cell(P, &i, &j, &k)) h3_cell<T, false>(p, i, j, k);
}

template <typename T>
__global__ __launch_bounds__(256) void e3_pieces(K3<T> p, Pieces P, mt_source3d src, T c, T* tape) {
  int i, j, k;
  if (piece_cell(P, &i, &j, &k)) e3_cell<T, false>(p, src, c, tape, i, j, k);
}

// Interior kernel.  A workgroup owns a tile of TJ rows (j) x 64 columns (k)
// and one x segment [ib, ie) of the interior box, and streams along x.  It
// has TJ + 2 waves, one per tile row from j0 - 1 to j0 + TJ, lane = k - k0:
// the first row computes the halo Hx', Hz' that the tile's E' reads, the last
// one only supplies E and eps; lane 0 also holds column k0 - 1 (halo Hx', Hy')
// and lane 63 column k0 + 64 (Ex, Ey, eps).  Per plane i, after one barrier:
// H'(i) from E(i) (LDS) and E(i+1) (registers), then E'(i-1) from H'(i-1)
// (LDS) and H'(i-2) (registers).  E planes are double-buffered in LDS, eps
// triple-buffered, H' double-buffered, so one barrier per plane orders every
// LDS write after the reads of the same buffer; the barrier orders LDS only
// (lds_barrier), so global loads, issued two planes ahead, stay in flight.  Halo values are recomputed from the read-only set A,
// never read from another workgroup's output, so workgroups are independent.
// Arithmetic is the two-pass cell updates' expression for expression (the box
// is psi-free and away from the PEC walls), so results are bit-identical.
// Workgroup barrier that orders LDS only: this wave's LDS reads and writes
// complete (lgkmcnt 0), then s_barrier.  __syncthreads() would also wait for
// every outstanding global load and store (its workgroup fences wait on
// vmcnt), draining the prefetch of the streaming kernel at every plane; here
// global loads land only in the issuing lane's registers and the stores go
// to buffers no workgroup reads, so they stay in flight across the barrier.
// The empty asm statements keep the compiler from moving memory accesses
// across it.
__device__ __forceinline__ void lds_barrier() {
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);   // vmcnt 63 (any), expcnt 7 (any), lgkmcnt 0
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
}

template <typename T>
struct I3 {
  int nx, ny, nz, lean;
  int I0, I1, J0, J1, K0, K1;   // interior box
  int kbase, njt, nkt, iseg;    // k tiles from kbase (64-aligned); j tiles from J0; x segments of iseg planes
  int ntiles, per_xcd;
  T dt, dx, rdx;
  const T* eps;
  const T *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;   // set A (read)
  T *oEx, *oEy, *oEz, *oHx, *oHy, *oHz;   // set B (written)
};

template <typename T>
struct Col {          // one field column value set
  T x, y, z;
};

template <typename T, int TJ, bool FD>
__global__ __launch_bounds__(64 * (TJ + 2)) void y3_interior(I3<T> p, mt_source3d src, T c_src, T* tape) {
  constexpr int NR = TJ + 2;
  __shared__ T sE[4][3][NR][66];      // E planes i-1 .. i+2 (ring of 4)
  __shared__ T sEps[3][TJ + 1][65];   // eps planes i-1 .. i+1 (ring of 3)
  __shared__ T sH[2][3][TJ + 1][65];  // H' planes i-1, i
  // XCD-aware tile order: consecutive workgroups go to different XCDs
  // (round robin), so XCD x takes the contiguous tile range
  // [x * per_xcd, (x + 1) * per_xcd), j fastest: j-neighbour tiles, which
  // share halo rows, run on the same XCD (same L2) at about the same time.
  const int t = (blockIdx.x & 7) * p.per_xcd + (blockIdx.x >> 3);
  if (t >= p.ntiles) return;
  const int jt = t % p.njt, r2 = t / p.njt, kt = r2 % p.nkt, sg = r2 / p.nkt;
  const int j0 = p.J0 + jt * TJ, k0 = p.kbase + kt * 64;
  const int ib = p.I0 + sg * p.iseg, ie = min(p.I1, ib + p.iseg);
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  // One wave per row, so the row and everything derived from it is uniform
  // across the wave (scalar registers and branches); only k varies by lane.
  const int lane = threadIdx.x;
  const int wr = __builtin_amdgcn_readfirstlane((int)threadIdx.y) - 1;   // -1 .. TJ
  const int j = j0 + wr, k = k0 + lane;
  const bool own = wr >= 0 && wr < TJ;                        // rows j0 .. j0+TJ-1
  const bool wantH = wr < TJ;                                 // own rows and the halo row j0-1
  const bool active = own && j < p.J1 && k >= p.K0 && k < p.K1;
  const bool haloL = k0 >= p.K0;                              // column k0-1 feeds active cells
  const int r = wr + 1, c = lane + 1;                         // LDS coordinates of the own cell
  const T dx = p.dx, dt = p.dt, rdx = p.rdx;

  // Every load is in bounds without a mask: row, column and plane indices are
  // clamped into the grid.  The interior box keeps every neighbour an active
  // cell reads inside the grid, so a clamped (duplicate) value only ever
  // reaches cells that are not stored.  Lanes 0 and 63 read the halo columns
  // k0-1 and k0+64 with the same instruction that reads the other lanes' own
  // column.  Offsets are byte offsets within one x plane.
  const int jE = min(max(j, 0), ny - 1), jY = min(max(j, 0), ny - 2);
  const int kE = min(k, nz - 1), kZ = min(k, nz - 2);
  const int kh = lane == 0 ? max(k0 - 1, 0) : (lane == 63 ? min(k0 + 64, nz - 1) : kE);
  const int khZ = min(kh, nz - 2);
  constexpr unsigned S = sizeof(T);
  const unsigned oEx = (unsigned)(jE * nz + kE) * S, oEy = (unsigned)(jY * nz + kE) * S;  // also eps; Hz
  const unsigned oEz = (unsigned)(jE * (nz - 1) + kZ) * S;                                // also Hy
  const unsigned oHx = (unsigned)(jY * (nz - 1) + kZ) * S;
  const unsigned hEx = (unsigned)(jE * nz + kh) * S, hEy = (unsigned)(jY * nz + kh) * S;  // also eps
  const unsigned hEz = (unsigned)(jE * (nz - 1) + khZ) * S;                               // also Hy
  const unsigned hHx = (unsigned)(jY * (nz - 1) + khZ) * S;
  const long long pE = (long long)ny * nz, pY = (long long)(ny - 1) * nz, pZ = (long long)ny * (nz - 1),
                  pX = (long long)(ny - 1) * (nz - 1);
  auto at = [](const T* base, long long plane, int ip, unsigned off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base + ip * plane) + off);
  };
  auto put = [](T* base, long long plane, int ip, unsigned off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base + ip * plane) + off) = v;
  };

  struct Stage {
    Col<T> e, eh, h;   // E own column, E halo columns, old H own column
    T eps, epsh;       // eps own column, eps halo column (lane 63)
    T hlx, hly;        // old Hx, Hy of column k0-1 (lane 0)
  };
  // the stage iteration i consumes: E(i+1), eps(i), H(i); planes past the
  // segment are clamped into the grid and never used
  auto fetch = [&](Stage& st, int i) {
    const int pe = min(i + 1, nx - 1), pex = min(i + 1, nx - 2), pp = min(i, nx - 1), ph = min(i, nx - 2);
    st.e.x = at(p.Ex, pE, pex, oEx);
    st.e.y = at(p.Ey, pY, pe, oEy);
    st.e.z = at(p.Ez, pZ, pe, oEz);
    st.eh.x = at(p.Ex, pE, pex, hEx);
    st.eh.y = at(p.Ey, pY, pe, hEy);
    st.eh.z = at(p.Ez, pZ, pe, hEz);
    st.eps = at(p.eps, pE, pp, oEx);
    st.epsh = at(p.eps, pE, pp, hEx);
    st.h.x = at(p.Hx, pX, ph, oHx);
    st.h.y = at(p.Hy, pZ, ph, oEz);
    st.h.z = at(p.Hz, pY, ph, oEy);
    st.hlx = at(p.Hx, pX, ph, hHx);
    st.hly = at(p.Hy, pZ, ph, hEz);
  };
  auto putE = [&](int ip, const Col<T>& e, const Col<T>& eh) {
    const int b = ip & 3;
    sE[b][0][r][c] = e.x;
    sE[b][1][r][c] = e.y;
    sE[b][2][r][c] = e.z;
    if (lane == 0) {
      sE[b][0][r][0] = eh.x;
      sE[b][1][r][0] = eh.y;
      sE[b][2][r][0] = eh.z;
    } else if (lane == 63) {
      sE[b][0][r][65] = eh.x;
      sE[b][1][r][65] = eh.y;
    }
  };

  // Pipeline (two planes of load latency for every stream): at iteration i
  // plane i+1 of E and plane i of eps go into LDS from registers loaded two
  // iterations earlier, and E(i+3), eps(i+2) and the old H(i+2) are loaded.
  // Three register stages rotate by name (the loop is unrolled by three), so
  // no value is copied out of a load still in flight: a copy would wait for it.
  const int ip0 = ib - 1;
  {
    const int pex = min(ip0, nx - 2);
    putE(ip0, Col<T>{at(p.Ex, pE, pex, oEx), at(p.Ey, pY, ip0, oEy), at(p.Ez, pZ, ip0, oEz)},
         Col<T>{at(p.Ex, pE, pex, hEx), at(p.Ey, pY, ip0, hEy), at(p.Ez, pZ, ip0, hEz)});
  }
  Stage S0, S1, S2;
  fetch(S0, ip0);
  fetch(S1, ip0 + 1);
  Col<T> hp{T(0), T(0), T(0)}, hpp{T(0), T(0), T(0)};

  auto body = [&](int i, const Stage& cur, Stage& ahead) {
    // ---- into LDS: E(i+1), eps(i)
    putE(i + 1, cur.e, cur.eh);
    if (wr >= 0) {
      sEps[i % 3][wr][lane] = cur.eps;
      if (lane == 63 && wr < TJ) sEps[i % 3][wr][64] = cur.epsh;
    }
    // ---- loads two planes ahead
    fetch(ahead, i + 2);
    lds_barrier();
    // ---- H'(i): own cells, halo row j0-1 (Hx', Hz'), halo column k0-1 (Hx', Hy')
    Col<T> hn{T(0), T(0), T(0)};
    if (i < ie && wantH) {
      const int b0 = i & 3, b1 = (i + 1) & 3;
      const T ex0 = sE[b0][0][r][c], ey0 = sE[b0][1][r][c], ez0 = sE[b0][2][r][c];
      const T ezj = sE[b0][2][r + 1][c], eyk = sE[b0][1][r][c + 1];
      const T exk = sE[b0][0][r][c + 1], exj = sE[b0][0][r + 1][c];
      const T eyx = sE[b1][1][r][c], ezx = sE[b1][2][r][c];
      // stepper_3d.py:371-375 (Hx'), and the same for Hy', Hz'
      auto hcalc = [&](auto& D) {
        Col<T> h;
        h.x = cur.h.x - dt * (D(ezj - ez0) - D(eyk - ey0));
        h.y = own ? cur.h.y - dt * (D(exk - ex0) - D(ezx - ez0)) : T(0);
        h.z = cur.h.z - dt * (D(eyx - ey0) - D(exj - ex0));
        return h;
      };
      DxDiv<T, FD> D{dx, rdx, 0u};
      hn = hcalc(D);
      if (!D.exact()) {
        __asm__ volatile("");
        IeeeDiv<T> I{dx};
        hn = hcalc(I);
      }
      const int bH = i & 1;
      sH[bH][0][r][c] = hn.x;
      sH[bH][1][r][c] = hn.y;
      sH[bH][2][r][c] = hn.z;
      if (own && haloL && lane == 0) {  // column k0-1
        const T hez0 = sE[b0][2][r][0], hey0 = sE[b0][1][r][0], hex0 = sE[b0][0][r][0];
        const T a0 = sE[b0][2][r + 1][0] - hez0, a1 = sE[b0][1][r][1] - hey0;
        const T a2 = sE[b0][0][r][1] - hex0, a3 = sE[b1][2][r][0] - hez0;
        sH[bH][0][r][0] = cur.hlx - dt * (div_c(a0, dx, rdx) - div_c(a1, dx, rdx));
        sH[bH][1][r][0] = cur.hly - dt * (div_c(a2, dx, rdx) - div_c(a3, dx, rdx));
      }
      if (active && i >= ib) {
        put(p.oHx, pX, i, oHx, hn.x);
        put(p.oHy, pZ, i, oEz, hn.y);
        put(p.oHz, pY, i, oEy, hn.z);
      }
    }
    // ---- E'(i-1) of the own cell (stepper_3d.py:392-437)
    if (i > ib && active) {
      const int im = i - 1, bH = im & 1, bP = im % 3, bE = im & 3;
      const T epsm = sEps[bP][wr][lane], eps1 = sEps[i % 3][wr][lane];
      const T hzj = sH[bH][2][r - 1][c], hyk = sH[bH][1][r][c - 1];
      const T hxk = sH[bH][0][r][c - 1], hxj = sH[bH][0][r - 1][c];
      auto curls = [&](auto& D) {
        return Col<T>{D(hp.z - hzj) - D(hp.y - hyk), D(hp.x - hxk) - D(hp.z - hpp.z),
                      D(hp.y - hpp.y) - D(hp.x - hxj)};
      };
      DxDiv<T, FD> D{dx, rdx, 0u};
      Col<T> cv = curls(D);
      if (!D.exact()) {
        __asm__ volatile("");
        IeeeDiv<T> I{dx};
        cv = curls(I);
      }
      T epsc = T(0.5) * (epsm + eps1);
      T ex = e_update(p, sE[bE][0][r][c], epsc, cv.x);
      ex = inject(p, src, 0, im, j, k, ex, epsc, c_src);
      epsc = T(0.5) * (epsm + sEps[bP][wr + 1][lane]);
      T ey = e_update(p, sE[bE][1][r][c], epsc, cv.y);
      ey = inject(p, src, 1, im, j, k, ey, epsc, c_src);
      epsc = T(0.5) * (epsm + sEps[bP][wr][lane + 1]);
      T ez = e_update(p, sE[bE][2][r][c], epsc, cv.z);
      ez = inject(p, src, 2, im, j, k, ez, epsc, c_src);
      put(p.oEx, pE, im, oEx, ex);
      put(p.oEy, pY, im, oEy, ey);
      put(p.oEz, pZ, im, oEz, ez);
      if (tape) {  // [Ex | Ey | Ez] of this step
        T* ty = tape + (long long)(nx - 1) * pE;
        put(tape, pE, im, oEx, ex);
        put(ty, pY, im, oEy, ey);
        put(ty + (long long)nx * pY, pZ, im, oEz, ez);
      }
    }
    hpp = hp;
    hp = hn;
  };
  for (int i = ip0;;) {
    body(i, S0, S2);
    if (++i > ie) break;
    body(i, S1, S0);
    if (++i > ie) break;
    body(i, S2, S1);
    if (++i > ie) break;
  }
}

// The interior box: cells whose H (and the H of their low neighbours, which
// the interior kernel recomputes) and E carry no psi, away from the walls.
struct Box3 {
  int I0, I1, J0, J1, K0, K1;
};

inline Box3 interior_box(const mt_yee3d* d) {
  const int n[3] = {d->nx, d->ny, d->nz};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    // H at index i is psi-free on [h.lo_end, h.hi_begin); E at i uses the E
    // profile at i-1: psi-free on [e.lo_end + 1, e.hi_begin + 1); the halo
    // H'(i-1) must be psi-free too
    lo[a] = std::max({1, d->h[a].lo_end + 1, d->e[a].lo_end + 1});
    hi[a] = std::min({n[a] - 1, d->h[a].hi_begin, d->e[a].hi_begin + 1});
  }
  return Box3{lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]};
}

inline Piece make_piece(int i0, int i1, int j0, int j1, int k0, int k1, int narrow) {
  Piece q{};
  q.i0 = i0, q.j0 = j0, q.k0 = k0;
  q.ni = std::max(0, i1 - i0), q.nj = std::max(0, j1 - j0), q.nk = std::max(0, k1 - k0);
  q.narrow = narrow;
  q.gk = (q.nk + (narrow ? 15 : 63)) / (narrow ? 16 : 64);
  q.gj = (q.nj + (narrow ? 15 : 3)) / (narrow ? 16 : 4);
  return q;
}

inline Pieces shell_pieces(const Box3& b, int nx, int ny, int nz) {
  Piece all[6] = {make_piece(0, b.I0, 0, ny, 0, nz, 0),          make_piece(b.I1, nx, 0, ny, 0, nz, 0),
                  make_piece(b.I0, b.I1, 0, b.J0, 0, nz, 0),     make_piece(b.I0, b.I1, b.J1, ny, 0, nz, 0),
                  make_piece(b.I0, b.I1, b.J0, b.J1, 0, b.K0, 1), make_piece(b.I0, b.I1, b.J0, b.J1, b.K1, nz, 1)};
  Pieces P{};
  int blocks = 0;
  for (int t = 0; t < 6; ++t) {
    const Piece& q = all[t];
    if (q.ni == 0 || q.nj == 0 || q.nk == 0) continue;
    P.p[P.n] = q;
    P.p[P.n].b0 = blocks;
    blocks += q.gk * q.gj * q.ni;
    ++P.n;
  }
  P.total = blocks;
  return P;
}

// Monitor descriptors with every field term moved from the primary buffers
// to the matching alternate buffers (same offsets).
void retarget(const mt_monitor* in, int n, mt_monitor* out, void* const* prim, const long long* bytes,
              void* const* alt, int n_arrays) {
  for (int m = 0; m < n; ++m) {
    out[m] = in[m];
    for (int q = 0; q < in[m].n_terms; ++q) {
      const void* ptrs[2] = {in[m].term[q].a, in[m].term[q].b};
      for (int w = 0; w < 2; ++w) {
        if (!ptrs[w]) continue;
        const char* a = static_cast<const char*>(ptrs[w]);
        for (int f = 0; f < n_arrays; ++f) {
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
int run3_split(const mt_yee3d* d, K3<T> p, const mt_source3d& sv, const double* coef, const MonitorPlan& mp,
               int step0, int k, hipStream_t s) {
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const Box3 b = interior_box(d);
  void* prim[6] = {d->Ex, d->Ey, d->Ez, d->Hx, d->Hy, d->Hz};
  const long long nE[3] = {(long long)(nx - 1) * ny * nz, (long long)nx * (ny - 1) * nz, (long long)nx * ny * (nz - 1)};
  const long long nH[3] = {(long long)nx * (ny - 1) * (nz - 1), (long long)(nx - 1) * ny * (nz - 1),
                           (long long)(nx - 1) * (ny - 1) * nz};
  long long bytes[6];
  for (int q = 0; q < 3; ++q) {
    bytes[q] = nE[q] * (long long)sizeof(T);
    bytes[3 + q] = nH[q] * (long long)sizeof(T);
  }
  std::vector<mt_monitor> alt_mons(mp.n_mons > 0 ? mp.n_mons : 1);
  if (mp.n_mons) retarget(mp.mons, mp.n_mons, alt_mons.data(), prim, bytes, d->alt, 6);
  MonitorPlan mpa = mp;
  mpa.mons = alt_mons.data();

  I3<T> q{};
  q.nx = nx, q.ny = ny, q.nz = nz, q.lean = d->lean;
  q.I0 = b.I0, q.I1 = b.I1, q.J0 = b.J0, q.J1 = b.J1, q.K0 = b.K0, q.K1 = b.K1;
  q.kbase = (b.K0 / 64) * 64;
  q.njt = (b.J1 - b.J0 + TJ - 1) / TJ;
  q.nkt = (b.K1 - q.kbase + 63) / 64;
  // x segments of about 32 planes: many workgroups (small tail), and the
  // one recomputed halo plane per segment costs about 3%
  const int tiles2d = q.njt * q.nkt;
  int nseg = std::max(1, (b.I1 - b.I0 + 31) / 32);
  q.iseg = (b.I1 - b.I0 + nseg - 1) / nseg;
  nseg = (b.I1 - b.I0 + q.iseg - 1) / q.iseg;
  q.ntiles = tiles2d * nseg;
  q.per_xcd = (q.ntiles + 7) / 8;
  q.dt = static_cast<T>(d->dt);
  q.dx = static_cast<T>(d->dx);
  q.rdx = const_div_recip<T>(q.dx);
  q.eps = static_cast<const T*>(d->eps);
  const Pieces P = shell_pieces(b, nx, ny, nz);
  const long long box_cells = (long long)(b.I1 - b.I0) * (b.J1 - b.J0) * (b.K1 - b.K0);
  const long long e_cells = nE[0] + nE[1] + nE[2];
  int S = mp.stage_steps > 0 ? mp.stage_steps : 1;
  int filled = 0, first = step0;
  for (int s_ = 0; s_ < k; ++s_) {
    void* const* A = (s_ % 2 == 0) ? prim : d->alt;
    void* const* B = (s_ % 2 == 0) ? d->alt : prim;
    q.Ex = static_cast<const T*>(A[0]), q.Ey = static_cast<const T*>(A[1]), q.Ez = static_cast<const T*>(A[2]);
    q.Hx = static_cast<const T*>(A[3]), q.Hy = static_cast<const T*>(A[4]), q.Hz = static_cast<const T*>(A[5]);
    q.oEx = static_cast<T*>(B[0]), q.oEy = static_cast<T*>(B[1]), q.oEz = static_cast<T*>(B[2]);
    q.oHx = static_cast<T*>(B[3]), q.oHy = static_cast<T*>(B[4]), q.oHz = static_cast<T*>(B[5]);
    p.Ex = static_cast<T*>(A[0]), p.Ey = static_cast<T*>(A[1]), p.Ez = static_cast<T*>(A[2]);
    p.Hx = static_cast<T*>(A[3]), p.Hy = static_cast<T*>(A[4]), p.Hz = static_cast<T*>(A[5]);
    p.oEx = q.oEx, p.oEy = q.oEy, p.oEz = q.oEz, p.oHx = q.oHx, p.oHy = q.oHy, p.oHz = q.oHz;
    const T c = sv.kind ? static_cast<T>(coef[s_]) : T(0);
    T* slot = d->tape ? static_cast<T*>(d->tape) + (long long)s_ * e_cells : nullptr;
    timing_begin(s);
    if (q.rdx != T(0))
      hipLaunchKernelGGL((y3_interior<T, TJ, true>), dim3(8 * q.per_xcd), dim3(64, TJ + 2), 0, s, q, sv, c, slot);
    else
      hipLaunchKernelGGL((y3_interior<T, TJ, false>), dim3(8 * q.per_xcd), dim3(64, TJ + 2), 0, s, q, sv, c, slot);
    timing_end(s, box_cells);
    if (P.total) {
      hipLaunchKernelGGL(h3_pieces<T>, dim3(P.total), dim3(256), 0, s, p, P);
      hipLaunchKernelGGL(e3_pieces<T>, dim3(P.total), dim3(256), 0, s, p, P, sv, c, slot);
    }
    int rc = check_launch("yee3d split step");
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
    for (int q6 = 0; q6 < 6; ++q6)
      if (hipMemcpyAsync(prim[q6], d->alt[q6], bytes[q6], hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(MT_EHIP, "yee3d split: copy-back failed");
  }
  return MT_OK;
}

// ---------------------------------------------------------------------------
// Variants 4-6: two Yee steps per pass (temporal blocking) over the slab I
// of cells whose two-step dependency cone is psi-free along x and y and away
// from the x and y walls (the interior box shrunk by two cells along x and
// y; every z index, z walls and z CPML shells included), and the two steps
// of the rest of the grid (the "ring": the x and y CPML slabs) with the
// per-cell updates.  Per pair of steps, with A the current field set, B the
// other one and C a third (alt[6..11]):
//   y3_tb2:  I cells, steps 1 and 2, A -> B; the z-CPML memory of I cells
//            read from the psi arrays, written to scratch copies
//            (alt[12..15]) that a copy returns after the pass, so no wave
//            reads a value another wave has already advanced; step-1 fields
//            of the listed x planes (the monitors' planes) and of the cells
//            on I's x and y faces -> C;
//   ring 1:  step 1 of every cell outside I, A -> C (psi in place), reading
//            the H' of I's high faces from C;
//   ring 2:  step 2 of every cell outside I, C -> B (psi in place), reading
//            the E' of I's low faces from C and the H'' of the I cells next
//            to the ring from B.
// Every value of B, of the psi arrays and of the monitored planes of C is
// the two-pass step's value bit for bit: each cell runs the two-pass cell
// update's expressions.
// ---------------------------------------------------------------------------
constexpr int TB_PLANES = 8;

template <typename T>
struct TB3 {
  int nx, ny, nz, lean;
  int I0, I1, J0, J1;                 // the slab I (all of z)
  int njt, nkl, kt_lo, kt_hi, iseg, ntiles, per_xcd, nsp;
  int sp[TB_PLANES];                  // x planes whose step-1 fields go to C
  T dt, dx, rdx, c1, c2;              // source coefficients of steps 1 and 2
  const T* eps;
  Pml<T> hz, ez;                      // z profiles of the H and E updates
  const T *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;   // A (read)
  T *oEx, *oEy, *oEz, *oHx, *oHy, *oHz;   // B (step 2)
  T *cEx, *cEy, *cEz, *cHx, *cHy, *cHz;   // C (step 1, listed planes and faces)
  const T *psEYZ, *psEXZ, *psHYZ, *psHXZ; // z-CPML memory (read)
  T *osEYZ, *osEXZ, *osHYZ, *osHXZ;       // its scratch copies (written)
};

// Cross-lane moves for the k neighbours of the two-step kernel: DPP
// wavefront shifts (every lane active), the value held by lane l + 1 (next)
// or l - 1 (prev); lane 63 / lane 0 receive 0.
__device__ __forceinline__ float from_next_lane(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_prev_lane(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

// One wave per tile, no LDS and no barrier: lane l holds column
// k = k0 - 2 + l and every lane holds the R rows j0 - 2 .. j0 + R - 3 in
// registers, so j neighbours are registers of the same lane and k neighbours
// come from the next / previous lane.  Outputs: rows j0 .. j0 + R - 5, lanes
// 2 .. 61 (columns k0 .. k0 + 59), planes [ib, ie).  The wave streams along
// x; at plane i it advances step 1 at plane i (H'(i) from E(i), E(i+1);
// E'(i) from H'(i), H'(i-1)) and step 2 at plane i-1 (H''(i-1) from E'(i-1),
// E'(i); E''(i-1) from H''(i-1), H''(i-2)), so each plane of A is read once
// and each plane of B written once.  Halo rows, lanes and the two planes
// before ib are recomputed, never read from another wave's output.
// ZP: the tile's columns reach the z walls or the z CPML shells (the curl
// terms along z vanish at the walls; the z-CPML memory of step 1 stays in
// registers for step 2); else every column is psi-free and off the walls.
// Issuing the next plane's loads before this plane's arithmetic (one more
// register set) and a three-waves-per-SIMD register budget (spills) were
// both measured slower on MI355X than this form at two waves per SIMD.
// Division by dx: FD, the verified constant-divisor sequence (div_c's fast
// path) with the dividends' exponents recorded; a row of a stage in which
// any lane of the wave met a dividend outside the verified range is
// recomputed with IEEE division by the whole wave (a uniform branch), which
// is div_c's result for every quotient.
template <int R, bool FD, bool ZP>
__global__ __launch_bounds__(64) void y3_tb2(TB3<float> p, mt_source3d src) {
  using T = float;
  constexpr int TR = R - 4;
  const int t = (blockIdx.x & 7) * p.per_xcd + (blockIdx.x >> 3);   // XCD-aware, j tiles fastest
  if (t >= p.ntiles) return;
  const int jt = t % p.njt, r2 = t / p.njt, e = r2 % p.nkl, sg = r2 / p.nkl;
  // k tiles kt_lo .. kt_hi-1 are the psi-free ones; the others take ZP
  const int kt = ZP ? (e < p.kt_lo ? e : p.kt_hi + (e - p.kt_lo)) : p.kt_lo + e;
  const int j0 = p.J0 + jt * TR, k0 = kt * 60;
  const int ib = p.I0 + sg * p.iseg, ie = min(p.I1, ib + p.iseg);
  const int ny = p.ny, nz = p.nz;
  const int lane = threadIdx.x;
  const int kk = k0 - 2 + lane;
  const bool lane_own = lane >= 2 && lane < 62 && kk < nz;
  const bool hasZ = kk < nz - 1;    // Ez, Hx, Hy exist in this column
  // Indices past the grid are clamped: they feed only cells that no output
  // reads (halo cells beyond a wall, or past I's last row).
  const int kA = min(max(kk, 0), nz - 1), kB = min(max(kk, 0), nz - 2), kC = min(max(kk - 1, 0), nz - 3);
  // z walls and z CPML (ZP only): Ex and Ey carry no curl on the wall
  // columns (stepper_3d.py:392-445 masks), the z curl terms of Hx, Hy (H
  // profile at k) and Ex, Ey (E profile at k-1) carry psi inside the shells
  const bool kz = !ZP || (kk >= 1 && kk < nz - 1);
  const bool hin = ZP && kk >= 0 && kk < nz - 1 && p.hz.in(kk);
  const bool ein = ZP && kz && p.ez.in(kk - 1);
  T hb = T(0), ha = T(0), eb = T(0), ea = T(0);
  if (hin) hb = p.hz.b[kk], ha = p.hz.a[kk];
  if (ein) eb = p.ez.b[kk - 1], ea = p.ez.a[kk - 1];
  const T dx = p.dx, dt = p.dt;
  // byte offsets within a plane of the nz- / (nz-1)- / (nz-2)-wide grids:
  // a wave-uniform row part (scalar registers) plus this lane's column part
  unsigned rA[R], rB[R], rC[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int jr = min(j0 - 2 + r, ny - 2);
    rA[r] = (unsigned)(jr * nz) * 4u;
    rB[r] = (unsigned)(jr * (nz - 1)) * 4u;
    rC[r] = (unsigned)(jr * (nz - 2)) * 4u;
  }
  const unsigned lA = (unsigned)kA * 4u, lB = (unsigned)kB * 4u, lC = (unsigned)kC * 4u;
  // plane strides in bytes: Ex, eps (ny, nz); Ey, Hz (ny-1, nz); Ez, Hy (ny, nz-1); Hx (ny-1, nz-1);
  // psi Hy_z (ny, nz-2), Hx_z (ny-1, nz-2)
  const long long sA = 4LL * ny * nz, sY = 4LL * (ny - 1) * nz, sZ = 4LL * ny * (nz - 1),
                  sX = 4LL * (ny - 1) * (nz - 1), sQ = 4LL * ny * (nz - 2), sW = 4LL * (ny - 1) * (nz - 2);
  auto at = [](const void* base, long long plane_bytes, int ip) {
    return static_cast<const char*>(base) + ip * plane_bytes;
  };
  auto ld = [](const char* plane, unsigned off) { return *reinterpret_cast<const T*>(plane + off); };
  auto st = [](T* base, long long plane_bytes, int ip, unsigned off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + ip * plane_bytes + off) = v;
  };
  // The source term only where the wave's rows, columns and planes meet
  // the source box (a uniform test per plane).
  const bool src_rows = src.kind != 0 && j0 - 2 < src.j1 && j0 + R - 2 > src.j0 && k0 - 2 < src.k1 &&
                        k0 + 62 > src.k0;
  auto inj = [&](int comp, int ip, int j, T E, T epsc, T c) {
    return inject(p, src, comp, ip, j, kk, E, epsc, c);
  };

  T e0x[R], e0y[R], e0z[R], epc[R], epp[R];   // E(i), eps(i), eps(i-1)
  T h1px[R], h1py[R], h1pz[R];                // H'(i-1)
  T e1px[R], e1py[R], e1pz[R];                // E'(i-1)
  T h2py[R], h2pz[R];                         // H''(i-2)
  T q1EYZ[R], q1EXZ[R], q1HYZ[R], q1HXZ[R];   // step-1 z-CPML memory of plane i-1 (ZP)
  int i = ib - 2;
  {
    const char *pEx = at(p.Ex, sA, i), *pEy = at(p.Ey, sY, i), *pEz = at(p.Ez, sZ, i), *pe = at(p.eps, sA, i);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      e0x[r] = ld(pEx + rA[r], lA);
      e0y[r] = ld(pEy + rA[r], lA);
      e0z[r] = ld(pEz + rB[r], lB);
      epc[r] = ld(pe + rA[r], lA);
      epp[r] = h1px[r] = h1py[r] = h1pz[r] = e1px[r] = e1py[r] = e1pz[r] = h2py[r] = h2pz[r] = T(0);
      q1EYZ[r] = q1EXZ[r] = q1HYZ[r] = q1HXZ[r] = T(0);
    }
  }
  // the loads iteration i consumes: E(i+1), eps(i+1), H(i) (planes past the
  // grid are clamped and never used)
  T nEx[R], nEy[R], nEz[R], epn[R], hx0[R], hy0[R], hz0[R];
  auto fetch = [&](int ip, T* fEx, T* fEy, T* fEz, T* fep, T* fHx, T* fHy, T* fHz) {
    const int pe = min(ip + 1, p.nx - 1), pex = min(ip + 1, p.nx - 2), ph = min(ip, p.nx - 2);
    const char *pEx = at(p.Ex, sA, pex), *pEy = at(p.Ey, sY, pe), *pEz = at(p.Ez, sZ, pe), *pp = at(p.eps, sA, pe),
               *pHx = at(p.Hx, sX, ph), *pHy = at(p.Hy, sZ, ph), *pHz = at(p.Hz, sY, ph);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      fEx[r] = ld(pEx + rA[r], lA);
      fEy[r] = ld(pEy + rA[r], lA);
      fEz[r] = ld(pEz + rB[r], lB);
      fep[r] = ld(pp + rA[r], lA);
      if (r < R - 1) {
        fHx[r] = ld(pHx + rB[r], lB);
        fHy[r] = ld(pHy + rB[r], lB);
        fHz[r] = ld(pHz + rA[r], lA);
      }
    }
  };
  for (; i <= ie; ++i) {
    fetch(i, nEx, nEy, nEz, epn, hx0, hy0, hz0);
    // z-CPML memory of plane i (step 0), shell columns only
    T q0EYZ[R], q0EXZ[R], q0HYZ[R], q0HXZ[R];
    if constexpr (ZP) {
      const char *a1 = at(p.psEYZ, sX, i), *a2 = at(p.psEXZ, sZ, min(i, p.nx - 2)),
                 *a3 = at(p.psHYZ, sQ, min(i, p.nx - 2)), *a4 = at(p.psHXZ, sW, i);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        q0EYZ[r] = hin ? ld(a1 + rB[r], lB) : T(0);
        q0EXZ[r] = hin ? ld(a2 + rB[r], lB) : T(0);
        q0HYZ[r] = ein ? ld(a3 + rC[r], lC) : T(0);
        q0HXZ[r] = ein ? ld(a4 + rC[r], lC) : T(0);
      }
    }
    const bool src_here = src_rows && i >= src.i0 - 1 && i < src.i1 + 1;
    // ---- step 1 at plane i: H'(i) on rows 0 .. R-2 (stepper_3d.py:363-390),
    // E'(i) on rows 1 .. R-2 (stepper_3d.py:392-445); each row's quotients
    // are one group of the division check
    T h1x[R], h1y[R], h1z[R], e1x[R], e1y[R], e1z[R];
    T q1nEYZ[R], q1nEXZ[R], q1nHYZ[R], q1nHXZ[R];
#pragma unroll
    for (int r = 0; r < R - 1; ++r) {
      const T eyk = from_next_lane(e0y[r]), exk = from_next_lane(e0x[r]);
      auto row = [&](auto& dv) {
        T cz = dv(eyk - e0y[r]);
        if constexpr (ZP) {   // ce_apply: psi' = b psi + a d, d + psi'
          const T ps = hb * q0EYZ[r] + ha * cz;
          q1nEYZ[r] = ps;
          cz = hin ? cz + ps : cz;
        }
        h1x[r] = hx0[r] - dt * (dv(e0z[r + 1] - e0z[r]) - cz);
        T cz2 = dv(exk - e0x[r]);
        if constexpr (ZP) {
          const T ps = hb * q0EXZ[r] + ha * cz2;
          q1nEXZ[r] = ps;
          cz2 = hin ? cz2 + ps : cz2;
        }
        h1y[r] = hy0[r] - dt * (cz2 - dv(nEz[r] - e0z[r]));
        h1z[r] = hz0[r] - dt * (dv(nEy[r] - e0y[r]) - dv(e0x[r + 1] - e0x[r]));
      };
      DxDiv<T, FD> dv{dx, p.rdx, 0u};
      row(dv);
      if (FD && __any(!dv.exact())) {
        DxDiv<T, false> dq{dx, p.rdx, 0u};
        row(dq);
      }
    }
#pragma unroll
    for (int r = 1; r < R - 1; ++r) {
      const int j = j0 - 2 + r;
      const T hyk = from_prev_lane(h1y[r]), hxk = from_prev_lane(h1x[r]), epk = from_next_lane(epc[r]);
      auto row = [&](auto& dv) {
        T chz = dv(h1y[r] - hyk);
        if constexpr (ZP) {
          const T ps = eb * q0HYZ[r] + ea * chz;
          q1nHYZ[r] = ps;
          chz = ein ? chz + ps : chz;
        }
        T curl = dv(h1z[r] - h1z[r - 1]) - chz;
        if constexpr (ZP) curl = kz ? curl : T(0);
        T epsc = T(0.5) * (epc[r] + epn[r]);
        e1x[r] = e_update(p, e0x[r], epsc, curl);
        if (src_here) e1x[r] = inj(0, i, j, e1x[r], epsc, p.c1);
        T chz2 = dv(h1x[r] - hxk);
        if constexpr (ZP) {
          const T ps = eb * q0HXZ[r] + ea * chz2;
          q1nHXZ[r] = ps;
          chz2 = ein ? chz2 + ps : chz2;
        }
        curl = chz2 - dv(h1z[r] - h1pz[r]);
        if constexpr (ZP) curl = kz ? curl : T(0);
        epsc = T(0.5) * (epc[r] + epc[r + 1]);
        e1y[r] = e_update(p, e0y[r], epsc, curl);
        if (src_here) e1y[r] = inj(1, i, j, e1y[r], epsc, p.c1);
        curl = dv(h1y[r] - h1py[r]) - dv(h1x[r] - h1x[r - 1]);
        epsc = T(0.5) * (epc[r] + epk);
        e1z[r] = e_update(p, e0z[r], epsc, curl);
        if (src_here) e1z[r] = inj(2, i, j, e1z[r], epsc, p.c1);
      };
      DxDiv<T, FD> dv{dx, p.rdx, 0u};
      row(dv);
      if (FD && __any(!dv.exact())) {
        DxDiv<T, false> dq{dx, p.rdx, 0u};
        row(dq);
      }
    }
    // step-1 fields into C: whole monitored planes, and the cells on I's x
    // and y faces (the ring's step 1 reads H' of the high faces, its step 2
    // reads E' of the low faces)
    if (i >= ib && i < ie && lane_own) {
      bool plane = i == p.I0 || i == p.I1 - 1;
      for (int q = 0; q < p.nsp; ++q) plane = plane || p.sp[q] == i;
#pragma unroll
      for (int r = 2; r < R - 2; ++r) {
        const int j = j0 - 2 + r;
        if (j >= p.J1) break;
        if (plane || j == p.J0 || j == p.J1 - 1) {
          st(p.cEx, sA, i, rA[r] + lA, e1x[r]);
          st(p.cEy, sY, i, rA[r] + lA, e1y[r]);
          st(p.cHz, sY, i, rA[r] + lA, h1z[r]);
          if (hasZ) {
            st(p.cEz, sZ, i, rB[r] + lB, e1z[r]);
            st(p.cHx, sX, i, rB[r] + lB, h1x[r]);
            st(p.cHy, sZ, i, rB[r] + lB, h1y[r]);
          }
        }
      }
    }
    // ---- step 2 at plane i-1: H''(i-1) on rows 1 .. R-3, E''(i-1) on the
    // output rows 2 .. R-3
    if (i >= ib) {
      const int ip = i - 1;
      const bool src2 = src_rows && ip >= src.i0 - 1 && ip < src.i1 + 1;
      T h2x[R], h2y[R], h2z[R], ex2[R], ey2[R], ez2[R];
      T q2EYZ[R], q2EXZ[R], q2HYZ[R], q2HXZ[R];
#pragma unroll
      for (int r = 1; r < R - 2; ++r) {
        const T eyk = from_next_lane(e1py[r]), exk = from_next_lane(e1px[r]);
        auto row = [&](auto& dv) {
          T cz = dv(eyk - e1py[r]);
          if constexpr (ZP) {
            const T ps = hb * q1EYZ[r] + ha * cz;
            q2EYZ[r] = ps;
            cz = hin ? cz + ps : cz;
          }
          h2x[r] = h1px[r] - dt * (dv(e1pz[r + 1] - e1pz[r]) - cz);
          T cz2 = dv(exk - e1px[r]);
          if constexpr (ZP) {
            const T ps = hb * q1EXZ[r] + ha * cz2;
            q2EXZ[r] = ps;
            cz2 = hin ? cz2 + ps : cz2;
          }
          h2y[r] = h1py[r] - dt * (cz2 - dv(e1z[r] - e1pz[r]));
          h2z[r] = h1pz[r] - dt * (dv(e1y[r] - e1py[r]) - dv(e1px[r + 1] - e1px[r]));
        };
        DxDiv<T, FD> dv{dx, p.rdx, 0u};
        row(dv);
        if (FD && __any(!dv.exact())) {
          DxDiv<T, false> dq{dx, p.rdx, 0u};
          row(dq);
        }
      }
#pragma unroll
      for (int r = 2; r < R - 2; ++r) {
        const int j = j0 - 2 + r;
        const T hyk = from_prev_lane(h2y[r]), hxk = from_prev_lane(h2x[r]), epk = from_next_lane(epp[r]);
        auto row = [&](auto& dv) {
          T chz = dv(h2y[r] - hyk);
          if constexpr (ZP) {
            const T ps = eb * q1HYZ[r] + ea * chz;
            q2HYZ[r] = ps;
            chz = ein ? chz + ps : chz;
          }
          T curl = dv(h2z[r] - h2z[r - 1]) - chz;
          if constexpr (ZP) curl = kz ? curl : T(0);
          T epsc = T(0.5) * (epp[r] + epc[r]);
          ex2[r] = e_update(p, e1px[r], epsc, curl);
          if (src2) ex2[r] = inj(0, ip, j, ex2[r], epsc, p.c2);
          T chz2 = dv(h2x[r] - hxk);
          if constexpr (ZP) {
            const T ps = eb * q1HXZ[r] + ea * chz2;
            q2HXZ[r] = ps;
            chz2 = ein ? chz2 + ps : chz2;
          }
          curl = chz2 - dv(h2z[r] - h2pz[r]);
          if constexpr (ZP) curl = kz ? curl : T(0);
          epsc = T(0.5) * (epp[r] + epp[r + 1]);
          ey2[r] = e_update(p, e1py[r], epsc, curl);
          if (src2) ey2[r] = inj(1, ip, j, ey2[r], epsc, p.c2);
          curl = dv(h2y[r] - h2py[r]) - dv(h2x[r] - h2x[r - 1]);
          epsc = T(0.5) * (epp[r] + epk);
          ez2[r] = e_update(p, e1pz[r], epsc, curl);
          if (src2) ez2[r] = inj(2, ip, j, ez2[r], epsc, p.c2);
        };
        DxDiv<T, FD> dv{dx, p.rdx, 0u};
        row(dv);
        if (FD && __any(!dv.exact())) {
          DxDiv<T, false> dq{dx, p.rdx, 0u};
          row(dq);
        }
      }
      if (ip >= ib && lane_own) {
#pragma unroll
        for (int r = 2; r < R - 2; ++r) {
          if (j0 - 2 + r >= p.J1) break;
          st(p.oEx, sA, ip, rA[r] + lA, ex2[r]);
          st(p.oEy, sY, ip, rA[r] + lA, ey2[r]);
          st(p.oHz, sY, ip, rA[r] + lA, h2z[r]);
          if (hasZ) {
            st(p.oEz, sZ, ip, rB[r] + lB, ez2[r]);
            st(p.oHx, sX, ip, rB[r] + lB, h2x[r]);
            st(p.oHy, sZ, ip, rB[r] + lB, h2y[r]);
          }
          if constexpr (ZP) {
            if (hin) {
              st(p.osEYZ, sX, ip, rB[r] + lB, q2EYZ[r]);
              st(p.osEXZ, sZ, ip, rB[r] + lB, q2EXZ[r]);
            }
            if (ein) {
              st(p.osHYZ, sQ, ip, rC[r] + lC, q2HYZ[r]);
              st(p.osHXZ, sW, ip, rC[r] + lC, q2HXZ[r]);
            }
          }
        }
      }
#pragma unroll
      for (int r = 1; r < R - 2; ++r) {
        h2py[r] = h2y[r];
        h2pz[r] = h2z[r];
      }
    }
    // rotate the planes (each array over the rows it holds)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (r < R - 1) h1px[r] = h1x[r], h1py[r] = h1y[r], h1pz[r] = h1z[r];
      if (r >= 1 && r < R - 1) e1px[r] = e1x[r], e1py[r] = e1y[r], e1pz[r] = e1z[r];
      if constexpr (ZP) {
        if (r < R - 1) q1EYZ[r] = q1nEYZ[r], q1EXZ[r] = q1nEXZ[r];
        if (r >= 1 && r < R - 1) q1HYZ[r] = q1nHYZ[r], q1HXZ[r] = q1nHXZ[r];
      }
      e0x[r] = nEx[r], e0y[r] = nEy[r], e0z[r] = nEz[r];
      epp[r] = epc[r], epc[r] = epn[r];
    }
  }
}

// Returns the z-CPML memory the two-step kernel advanced (scratch copies) to
// the psi arrays: for each of the four z-psi arrays, the rows [j0, j1) of
// the planes [i0, i1) over its shell index ranges [0, lo) and [hi, nk).
struct ZCopy {
  const float* src[4];
  float* dst[4];
  long long plane[4];
  int row[4], lo[4], hi[4], nk[4];
  int i0, i1, j0, j1;
};

__global__ __launch_bounds__(256) void zpsi_copy(ZCopy c) {
  const int a = blockIdx.y;
  const int nlo = c.lo[a], nhi = c.nk[a] - c.hi[a], w = nlo + nhi;
  const long long n = (long long)(c.i1 - c.i0) * (c.j1 - c.j0) * w;
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const int q = (int)(e % w);
    const long long rj = e / w;
    const int j = c.j0 + (int)(rj % (c.j1 - c.j0)), i = c.i0 + (int)(rj / (c.j1 - c.j0));
    const int k = q < nlo ? q : c.hi[a] + (q - nlo);
    const long long o = i * c.plane[a] + (long long)j * c.row[a] + k;
    c.dst[a][o] = c.src[a][o];
  }
}

inline Box3 shrink(const Box3& b, int m) {
  return Box3{b.I0 + m, b.I1 - m, b.J0 + m, b.J1 - m, b.K0 + m, b.K1 - m};
}

inline bool box_ok(const Box3& b) { return b.I1 > b.I0 && b.J1 > b.J0 && b.K1 > b.K0; }

// The x planes the monitors read (each term must lie in one x plane), into
// planes[0 .. *n); false when a term spans several planes or more than
// TB_PLANES planes are read.
inline bool monitor_planes(const MonitorPlan& mp, void* const* prim, const long long* bytes,
                           const long long* plane_elems, int elem, int* planes, int* n) {
  *n = 0;
  for (int m = 0; m < mp.n_mons; ++m) {
    const mt_monitor& M = mp.mons[m];
    for (int q = 0; q < M.n_terms; ++q) {
      const void* ptrs[2] = {M.term[q].a, M.term[q].b};
      for (int w = 0; w < 2; ++w) {
        if (!ptrs[w]) continue;
        const char* a = static_cast<const char*>(ptrs[w]);
        int f = -1;
        for (int g = 0; g < 6; ++g) {
          const char* b0 = static_cast<const char*>(prim[g]);
          if (a >= b0 && a < b0 + bytes[g]) f = g;
        }
        if (f < 0) return false;
        const long long o = (a - static_cast<const char*>(prim[f])) / elem;
        const long long n0 = M.kind == 1 ? 1 : M.n0, n1 = M.kind == 1 ? 1 : M.n1;
        const long long last = o + (n0 - 1) * M.term[q].s0 + (n1 - 1) * M.term[q].s1;
        const long long p0 = o / plane_elems[f];
        if (last / plane_elems[f] != p0 || M.term[q].s0 < 0 || M.term[q].s1 < 0) return false;
        bool have = false;
        for (int u = 0; u < *n; ++u) have = have || planes[u] == (int)p0;
        if (!have) {
          if (*n == TB_PLANES) return false;
          planes[(*n)++] = (int)p0;
        }
      }
    }
  }
  return true;
}

// Variant 4-6 host loop (RI: tile rows of the psi-free columns, RZ: of the
// z-edge columns); returns 1 (nothing launched) when the run does not fit
// it, and the caller takes the two-pass path.
template <typename T, int RI, int RZ>
int run3_tb2(const mt_yee3d* d, K3<T> p, const mt_source3d& sv, const double* coef, const MonitorPlan& mp,
             int step0, int k, hipStream_t s) {
  static_assert(sizeof(T) == 4, "the two-step pass is FP32");
  const int nx = d->nx, ny = d->ny, nz = d->nz;
  const Box3 P = interior_box(d);
  const Box3 I{P.I0 + 2, P.I1 - 2, P.J0 + 2, P.J1 - 2, 0, nz};
  if (!box_ok(I) || nz < 4 || d->tape || d->lean) return 1;
  void* prim[6] = {d->Ex, d->Ey, d->Ez, d->Hx, d->Hy, d->Hz};
  const long long nE[3] = {(long long)(nx - 1) * ny * nz, (long long)nx * (ny - 1) * nz, (long long)nx * ny * (nz - 1)};
  const long long nH[3] = {(long long)nx * (ny - 1) * (nz - 1), (long long)(nx - 1) * ny * (nz - 1),
                           (long long)(nx - 1) * (ny - 1) * nz};
  const long long plane[6] = {(long long)ny * nz, (long long)(ny - 1) * nz, (long long)ny * (nz - 1),
                              (long long)(ny - 1) * (nz - 1), (long long)ny * (nz - 1), (long long)(ny - 1) * nz};
  long long bytes[6];
  for (int q = 0; q < 3; ++q) {
    bytes[q] = nE[q] * (long long)sizeof(T);
    bytes[3 + q] = nH[q] * (long long)sizeof(T);
  }
  TB3<T> q{};
  if (mp.n_mons && !monitor_planes(mp, prim, bytes, plane, (int)sizeof(T), q.sp, &q.nsp)) return 1;
  void* const* A0 = prim;
  void* const* B0 = d->alt;
  void* const* C = d->alt + 6;
  std::vector<mt_monitor> alt_mons(mp.n_mons > 0 ? mp.n_mons : 1), c_mons(mp.n_mons > 0 ? mp.n_mons : 1);
  if (mp.n_mons) {
    retarget(mp.mons, mp.n_mons, alt_mons.data(), prim, bytes, d->alt, 6);
    retarget(mp.mons, mp.n_mons, c_mons.data(), prim, bytes, d->alt + 6, 6);
  }
  MonitorPlan mpa = mp, mpc = mp;
  mpa.mons = alt_mons.data();
  mpc.mons = c_mons.data();

  q.nx = nx, q.ny = ny, q.nz = nz, q.lean = 0;
  q.I0 = I.I0, q.I1 = I.I1, q.J0 = I.J0, q.J1 = I.J1;
  q.hz = p.h[2], q.ez = p.e[2];
  q.psEYZ = static_cast<const T*>(d->psi[P_EY_Z]), q.psEXZ = static_cast<const T*>(d->psi[P_EX_Z]);
  q.psHYZ = static_cast<const T*>(d->psi[P_HY_Z]), q.psHXZ = static_cast<const T*>(d->psi[P_HX_Z]);
  q.osEYZ = static_cast<T*>(d->alt[12]), q.osEXZ = static_cast<T*>(d->alt[13]);
  q.osHYZ = static_cast<T*>(d->alt[14]), q.osHXZ = static_cast<T*>(d->alt[15]);
  // k tiles of 60 columns from 0; the ones whose 64 lanes stay inside the
  // psi-free, wall-free z range take the plain kernel
  const int nkt = (nz + 59) / 60;
  int kt_lo = nkt, kt_hi = nkt;
  for (int t = 0; t < nkt; ++t)
    if (60 * t - 2 >= P.K0 && 60 * t + 62 <= P.K1) {
      if (kt_lo == nkt) kt_lo = t;
      kt_hi = t + 1;
    }
  if (kt_lo == nkt) kt_lo = kt_hi = 0;
  q.kt_lo = kt_lo, q.kt_hi = kt_hi;
  q.dt = p.dt, q.dx = p.dx, q.rdx = p.rdx;
  q.eps = p.eps;
  // launch shapes of the two kernels: x segments, each three planes longer
  // than its output; the count minimises rounds x (planes + 3), a round
  // being the waves the chip holds at once, so the last round is no sliver
  const int nxI = I.I1 - I.I0;
  auto shape = [&](TB3<T>& x, int R, int nkl, const void* kern) -> int {
    x.nkl = nkl;
    x.njt = (I.J1 - I.J0 + (R - 4) - 1) / (R - 4);
    const int tiles2d = x.njt * nkl;
    int per_cu = 0, cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0) != hipSuccess)
      return fail(MT_EHIP, "yee3d two-step: occupancy query failed");
    const long long cap = std::max(1, per_cu * cus);
    int nseg = 1;
    long long best = -1;
    for (int ns = 1; ns <= std::min(nxI, 64); ++ns) {
      const int len = (nxI + ns - 1) / ns;
      const int used = (nxI + len - 1) / len;
      const long long cost = (((long long)tiles2d * used + cap - 1) / cap) * (len + 3);
      if (best < 0 || cost < best) best = cost, nseg = used;
    }
    x.iseg = (nxI + nseg - 1) / nseg;
    nseg = (nxI + x.iseg - 1) / x.iseg;
    x.ntiles = tiles2d * nseg;
    x.per_xcd = (x.ntiles + 7) / 8;
    return MT_OK;
  };
  const bool fd = q.rdx != T(0);
  TB3<T> qi = q, qz = q;
  int rc;
  const void* ki_fast = (const void*)y3_tb2<RI, true, false>;
  const void* ki_ieee = (const void*)y3_tb2<RI, false, false>;
  if (kt_hi > kt_lo && (rc = shape(qi, RI, kt_hi - kt_lo, fd ? ki_fast : ki_ieee))) return rc;
  if (nkt - (kt_hi - kt_lo) > 0 &&
      (rc = shape(qz, RZ, nkt - (kt_hi - kt_lo),
                  fd ? (const void*)y3_tb2<RZ, true, true> : (const void*)y3_tb2<RZ, false, true>)))
    return rc;
  const bool run_i = kt_hi > kt_lo, run_z = nkt - (kt_hi - kt_lo) > 0;
  // the z-psi cells the edge kernel advances
  ZCopy zc{};
  {
    const int src_q[4] = {P_EY_Z, P_EX_Z, P_HY_Z, P_HX_Z};
    const long long pl[4] = {(long long)(ny - 1) * (nz - 1), (long long)ny * (nz - 1), (long long)ny * (nz - 2),
                             (long long)(ny - 1) * (nz - 2)};
    const int rw[4] = {nz - 1, nz - 1, nz - 2, nz - 2};
    for (int a = 0; a < 4; ++a) {
      zc.src[a] = static_cast<const float*>(d->alt[12 + a]);
      zc.dst[a] = static_cast<float*>(d->psi[src_q[a]]);
      zc.plane[a] = pl[a];
      zc.row[a] = rw[a];
      zc.nk[a] = rw[a];
      const Pml<T>& pr = a < 2 ? q.hz : q.ez;
      zc.lo[a] = std::min(std::max(pr.lo_end, 0), rw[a]);
      zc.hi[a] = std::max(std::min(pr.hi_begin, rw[a]), zc.lo[a]);
    }
    zc.i0 = I.I0, zc.i1 = I.I1, zc.j0 = I.J0, zc.j1 = I.J1;
  }
  const bool zcopy = run_z && (zc.lo[0] + zc.nk[0] - zc.hi[0] + zc.lo[2] + zc.nk[2] - zc.hi[2]) > 0;
  const Pieces P2 = shell_pieces(I, nx, ny, nz);   // the ring: every cell outside I
  const long long box_cells = (long long)(I.I1 - I.I0) * (I.J1 - I.J0) * nz;
  auto set_io = [](auto& x, void* const* rd, void* const* wr) {
    using U = std::remove_const_t<std::remove_pointer_t<decltype(x.oEx)>>;
    x.Ex = static_cast<U*>(rd[0]), x.Ey = static_cast<U*>(rd[1]), x.Ez = static_cast<U*>(rd[2]);
    x.Hx = static_cast<U*>(rd[3]), x.Hy = static_cast<U*>(rd[4]), x.Hz = static_cast<U*>(rd[5]);
    x.oEx = static_cast<U*>(wr[0]), x.oEy = static_cast<U*>(wr[1]), x.oEz = static_cast<U*>(wr[2]);
    x.oHx = static_cast<U*>(wr[3]), x.oHy = static_cast<U*>(wr[4]), x.oHz = static_cast<U*>(wr[5]);
  };
  for (TB3<T>* x : {&qi, &qz}) {
    x->cEx = static_cast<T*>(C[0]), x->cEy = static_cast<T*>(C[1]), x->cEz = static_cast<T*>(C[2]);
    x->cHx = static_cast<T*>(C[3]), x->cHy = static_cast<T*>(C[4]), x->cHz = static_cast<T*>(C[5]);
  }
  int S = mp.stage_steps > 0 ? mp.stage_steps : 1;
  int filled = 0, first = step0;
  auto sample = [&](const MonitorPlan& X, int step) -> int {
    if (!mp.n_mons) return 0;
    int e;
    if ((e = monitors_sample(X, filled, step, s))) return e;
    if (++filled == S || step == k - 1) {
      if ((e = monitors_flush(mp, filled, first, d->dt, s))) return e;
      first += filled;
      filled = 0;
    }
    return 0;
  };
  bool in_alt = false;   // the current set: primary or alt[0..5]
  int s_ = 0;
  for (; s_ + 1 < k; s_ += 2) {
    void* const* A = in_alt ? B0 : A0;
    void* const* B = in_alt ? A0 : B0;
    const T c1 = sv.kind ? static_cast<T>(coef[s_]) : T(0), c2 = sv.kind ? static_cast<T>(coef[s_ + 1]) : T(0);
    qi.c1 = qz.c1 = c1;
    qi.c2 = qz.c2 = c2;
    set_io(qi, A, B);
    set_io(qz, A, B);
    timing_begin(s);
    if (run_i) {
      void* args[] = {&qi, const_cast<mt_source3d*>(&sv)};
      if (hipLaunchKernel(fd ? ki_fast : ki_ieee, dim3(8 * qi.per_xcd), dim3(64), args, 0, s) != hipSuccess)
        return fail(MT_EHIP, "yee3d two-step: launch failed");
    }
    if (run_z) {
      if (fd)
        hipLaunchKernelGGL((y3_tb2<RZ, true, true>), dim3(8 * qz.per_xcd), dim3(64), 0, s, qz, sv);
      else
        hipLaunchKernelGGL((y3_tb2<RZ, false, true>), dim3(8 * qz.per_xcd), dim3(64), 0, s, qz, sv);
    }
    timing_end(s, 2 * box_cells);
    if (zcopy) hipLaunchKernelGGL(zpsi_copy, dim3(512, 4), dim3(256), 0, s, zc);
    // ring, step 1: A -> C
    set_io(p, A, C);
    hipLaunchKernelGGL(h3_pieces<T>, dim3(P2.total), dim3(256), 0, s, p, P2);
    hipLaunchKernelGGL(e3_pieces<T>, dim3(P2.total), dim3(256), 0, s, p, P2, sv, c1, (T*)nullptr);
    if ((rc = check_launch("yee3d two-step pass"))) return rc;
    if ((rc = sample(mpc, s_))) return rc;
    // ring, step 2: C -> B
    set_io(p, C, B);
    hipLaunchKernelGGL(h3_pieces<T>, dim3(P2.total), dim3(256), 0, s, p, P2);
    hipLaunchKernelGGL(e3_pieces<T>, dim3(P2.total), dim3(256), 0, s, p, P2, sv, c2, (T*)nullptr);
    if ((rc = check_launch("yee3d two-step ring"))) return rc;
    in_alt = !in_alt;
    if ((rc = sample(in_alt ? mpa : mp, s_ + 1))) return rc;
  }
  if (s_ < k) {   // odd count: the last step in place, two-pass
    void* const* A = in_alt ? B0 : A0;
    set_io(p, A, A);
    dim3 blk(BZ, BYY, 1);
    dim3 grd((nz + BZ - 1) / BZ, (ny + BYY - 1) / BYY, nx);
    const T c = sv.kind ? static_cast<T>(coef[s_]) : T(0);
    hipLaunchKernelGGL((h3_kernel<T, false>), grd, blk, 0, s, p);
    hipLaunchKernelGGL((e3_kernel<T, false>), grd, blk, 0, s, p, sv, c, (T*)nullptr);
    if ((rc = check_launch("yee3d step"))) return rc;
    if ((rc = sample(in_alt ? mpa : mp, s_))) return rc;
  }
  if (in_alt) {
    for (int q6 = 0; q6 < 6; ++q6)
      if (hipMemcpyAsync(prim[q6], d->alt[q6], bytes[q6], hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(MT_EHIP, "yee3d two-step: copy-back failed");
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
  p.rdx = const_div_recip<T>(p.dx);
  p.eps = static_cast<const T*>(d->eps);
  for (int a = 0; a < 3; ++a) {
    p.h[a] = pml_of<T>(d->h[a]);
    p.e[a] = pml_of<T>(d->e[a]);
  }
  p.Ex = p.oEx = static_cast<T*>(d->Ex);
  p.Ey = p.oEy = static_cast<T*>(d->Ey);
  p.Ez = p.oEz = static_cast<T*>(d->Ez);
  p.Hx = p.oHx = static_cast<T*>(d->Hx);
  p.Hy = p.oHy = static_cast<T*>(d->Hy);
  p.Hz = p.oHz = static_cast<T*>(d->Hz);
  for (int q = 0; q < 12; ++q) p.psi[q] = static_cast<T*>(d->psi[q]);
  const bool media = d->damp[0] != nullptr;
  const bool lean_poles = d->lean && d->n_poles > 0;   // poles of the lean form (no damping arrays)
  for (int q = 0; q < 12; ++q) p.dmp[q] = static_cast<const T*>(d->damp[q]);
  p.np = (media || lean_poles) ? d->n_poles : 0;
  p.c1 = static_cast<const T*>(d->c1);
  p.c2 = static_cast<const T*>(d->c2);
  for (int a = 0; a < 3; ++a) {
    p.c3[a] = static_cast<const T*>(d->c3[a]);
    p.Pc[a] = static_cast<T*>(d->P_cur[a]);
    p.Pp[a] = static_cast<T*>(d->P_prev[a]);
  }
  mt_source3d sv{};
  if (src && src->kind) sv = *src;
  const Box3 box = interior_box(d);
  const bool small = (long long)d->ny * d->nz * (long long)sizeof(T) < (1LL << 31);  // 32-bit in-plane offsets
  const bool split = d->variant >= 1 && d->variant <= 3 && !media && !lean_poles && !d->absorber && small &&
                     box.I1 > box.I0 && box.J1 > box.J0 &&
                     box.K1 > box.K0;
  // two steps per pass (variants 4, 5: tile rows 10 / 8 on the psi-free
  // columns, 8 on the z-edge columns), FP32 CPML runs
  // without media; runs it does not fit take the two-pass path
  if constexpr (sizeof(T) == 4) {
    if (d->variant >= 4 && !media && !d->absorber && small) {
      const int rc = d->variant == 4 ? run3_tb2<T, 10, 8>(d, p, sv, coef, mp, step0, k, s)
                                     : run3_tb2<T, 8, 8>(d, p, sv, coef, mp, step0, k, s);
      if (rc != 1) return rc;
    }
  }
  // tile rows: FP32 6 / 10 / 4, FP64 4 (the LDS rings must fit 64 KB)
  if (split) {
    if constexpr (sizeof(T) == 4) {
      if (d->variant == 1) return run3_split<T, 6>(d, p, sv, coef, mp, step0, k, s);
      if (d->variant == 2) return run3_split<T, 10>(d, p, sv, coef, mp, step0, k, s);
    }
    return run3_split<T, 4>(d, p, sv, coef, mp, step0, k, s);
  }
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
    if (d->absorber) {
      hipLaunchKernelGGL((h3_kernel<T, true, true>), grd, blk, 0, s, p);
      hipLaunchKernelGGL((e3_kernel<T, true, true>), grd, blk, 0, s, p, sv, c, slot);
    } else if (media) {
      hipLaunchKernelGGL((h3_kernel<T, true>), grd, blk, 0, s, p);
      hipLaunchKernelGGL((e3_kernel<T, true>), grd, blk, 0, s, p, sv, c, slot);
    } else if (lean_poles) {   // H without damping, E through the lean pole form
      hipLaunchKernelGGL((h3_kernel<T, false>), grd, blk, 0, s, p);
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
  if (p->absorber != 0 && p->absorber != 1) return fail(MT_EINVAL, "absorber must be 0 or 1");
  int n[3] = {p->nx, p->ny, p->nz};
  int rc;
  if (p->absorber) {  // no CPML: psi unused, every axis without shells
    for (int a = 0; a < 3; ++a)
      if (p->h[a].n != n[a] - 1 || p->e[a].n != n[a] - 2 || p->h[a].lo_end != 0 || p->e[a].lo_end != 0 ||
          p->h[a].hi_begin != p->h[a].n || p->e[a].hi_begin != p->e[a].n)
        return fail(MT_EINVAL, "absorber run: every profile must have its length and no shell");
  } else {
    for (int q = 0; q < 12; ++q)
      if (!p->psi[q]) return fail(MT_EINVAL, "null psi pointer");
    for (int a = 0; a < 3; ++a) {
      if ((rc = check_axis(p->h[a], n[a] - 1, "h"))) return rc;
      if ((rc = check_axis(p->e[a], n[a] - 2, "e"))) return rc;
    }
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
  if (p->variant < 0 || p->variant > 5) return fail(MT_EINVAL, "3D variant must be 0..5");
  if (p->variant >= 1 && p->variant <= 3) {
    for (int q = 0; q < 6; ++q)
      if (!p->alt[q]) return fail(MT_EINVAL, "split 3D variant needs the 6 alt field buffers");
  }
  if (p->variant >= 4) {
    for (int q = 0; q < 16; ++q)
      if (!p->alt[q]) return fail(MT_EINVAL, "two-step 3D variant needs alt[0..15] (two field sets, four z-psi)");
  }
  bool any_damp = false, all_damp = true;
  for (int q = 0; q < 12; ++q) {
    any_damp = any_damp || p->damp[q];
    all_damp = all_damp && p->damp[q];
  }
  if (any_damp && !all_damp) return fail(MT_EINVAL, "3D media need all 12 damping arrays");
  if (p->n_poles < 0 || p->n_poles > 64) return fail(MT_EINVAL, "3D n_poles must be 0..64");
  if (p->n_poles > 0 && !any_damp && !p->lean)
    return fail(MT_EINVAL, "3D poles need the damping arrays (d1 = 1, d2 = dt)");
  if (p->absorber && (!any_damp || p->lean))
    return fail(MT_EINVAL, "absorber run needs the 12 damping arrays and the functional stepper");
  if (any_damp && p->lean) return fail(MT_EINVAL, "3D conductivity is not supported by the lean stepper");
  if (p->n_poles > 0) {
    if (!p->c1 || !p->c2) return fail(MT_EINVAL, "3D poles without c1 / c2");
    for (int a = 0; a < 3; ++a)
      if (!p->c3[a] || !p->P_cur[a] || !p->P_prev[a]) return fail(MT_EINVAL, "3D poles: null c3 or P stack");
  }
  MonitorPlan mp{mons, n_mons, stage, stage_bytes, stage_steps, p->dtype};
  if (n_mons && (rc = monitors_validate(mp))) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  return p->dtype == MT_F64 ? run3_t<double>(p, src, coef, mp, step0, k, s)
                            : run3_t<float>(p, src, coef, mp, step0, k, s);
}
