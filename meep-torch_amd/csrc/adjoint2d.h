// Device bodies of the 2D TM adjoint passes (shared by the single-problem
// entry point in adjoint2d.hip and the batched one in batch2d.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace mt {

constexpr int ABX = 64, ABY = 4;

template <typename T>
struct AdjK {
  int nx, ny, y_pmc, ny_e;
  int ycut, xcut;  // even mirror cuts at the low edges: the cut row / column of Dz is live
  T dt, dx;
  const T* eps;
  Pml<T> hy, hx, ex, ey;
  T *aEz, *aDz, *aHx, *aHy, *apzy, *apzx, *apyx, *apxy;
  T *gHy, *gHx, *gEx, *gEy;
  T* eps_grad;
  // conductivity (sigma_d, sigma_b, mu): Dz' = d1 Dz + d2 curl,
  // Hx' = h1x Hx - h2x curl, Hy' = h1y Hy + h2y curl; NULL: d1 = h1 = 1,
  // d2 = h2 = dt
  const T *d1, *d2, *h1x, *h2x, *h1y, *h2y;
  // np ADE poles: per-cell scalars c1, c2, c3 and the cotangents of the
  // stacks P (aPc) and P_prev (aPp), all (np, nx, ny)
  int np;
  const T *c1, *c2, *c3;
  T *aPc, *aPp;
  // chi2/chi3 fixed point (kerr = 1): the tape holds rhs = Dz - sum P
  int kerr;
  const T *chi2, *chi3;
  // optional gradients of chi2, chi3 and the poles' c3; ez_init = Ez before
  // the first step of the call (c3's gradient needs the Ez each step starts from)
  T *chi2_grad, *chi3_grad, *c3_grad;
  const T* ez_init;
  // magnetic Lorentz poles (nm > 0; no conductivity): per-cell stacks mc1..mc3
  // of the Hx and Hy grids and the cotangents of B and of the M stacks
  int nm;
  const T *mc1x, *mc2x, *mc3x, *mc1y, *mc2y, *mc3y;
  T *aBx, *aBy, *aMxc, *aMxp, *aMyc, *aMyp;
};

// Transpose of one H component's magnetic poles at a cell (stepper_tm.py:
// 103-106,187-193,209-224): H' = B' - sum M', M' = (mc1 M + mc2 M_prev) +
// mc3 H, B' = B +- dt curl.  hb is the cotangent of H'; the B cotangent
// (updated in place) becomes that of B' (= of the old B), every M' gains
// -hb, the stacks take the recurrence's transpose, and the old H's
// cotangent (it feeds only the M') is returned; *bb gets B''s cotangent,
// which the curl receives times +-dt.
template <typename T>
__device__ __forceinline__ T mpole_adj(int nm, const T* mc1, const T* mc2, const T* mc3, T* aB, T* aMc, T* aMp,
                                       long long q, long long nq, T hb, T* bb) {
  const T b = aB[q] + hb;
  aB[q] = b;
  *bb = b;
  T h_old = T(0);
  for (int k = 0; k < nm; ++k) {
    const long long o = (long long)k * nq + q;
    const T mb = aMc[o] - hb, prev = aMp[o];
    aMc[o] = mc1[o] * mb + prev;
    aMp[o] = mc2[o] * mb;
    h_old = h_old + mc3[o] * mb;
  }
  return h_old;
}

// Ez before step s_rel of a call: ez_init for the first step, else the tape
// slot of the step before (under chi2/chi3 that slot is Dz - sum P, and Ez
// is the forward's three fixed-point passes from it).
template <typename T>
__device__ __forceinline__ T adj_ez_before(const AdjK<T>& p, const T* ez_after, int s_rel, long long c) {
  if (s_rel == 0) return p.ez_init[c];
  const T v = ez_after[c - (long long)p.nx * p.ny];
  if (!p.kerr) return v;
  const T e0 = p.eps[c], x2 = p.chi2[c], x3 = p.chi3[c];
  T ez = v / e0;
  for (int it = 0; it < 3; ++it) ez = v / (e0 + x2 * ez + x3 * ez * ez);
  return ez;
}

// Each pass below issues all of a cell's loads before its first store: the
// adjoint, scratch and profile arrays never alias, but the compiler cannot
// know that and would otherwise serialise every load behind the stores that
// precede it in program order.

// A: Ez' cotangent -> eps gradient and Dz cotangent; curl cotangent -> D-half
// psi cotangents and the H-difference cotangents gHy, gHx.
template <typename T>
__device__ __forceinline__ void adj_a_cell(const AdjK<T>& p, const T* ez_after, int i, int j, int s_rel = 0) {
  const int nx = p.nx, ny = p.ny;
  const long long c = (long long)i * ny + j;
  // live Dz as in the forward (stepper_tm.py:134-155); beyond an even cut H is 0
  const bool live = (p.xcut ? i < nx - 1 : (i >= 1 && i < nx - 1)) &&
                    (p.y_pmc || (p.ycut ? j < ny - 1 : (j >= 1 && j < ny - 1)));
  const int ii = p.xcut ? i : i - 1, jj = (p.y_pmc || p.ycut) ? j : j - 1;
  const long long q = (long long)ii * p.ny_e + jj;
  const T eps = p.eps[c], ezb = p.aEz[c], eg = p.eps_grad[c], ea = ez_after[c], dz0 = p.aDz[c];
  const T d1 = p.d1 ? p.d1[c] : T(1), d2 = p.d1 ? p.d2[c] : p.dt;
  const bool sx = live && p.ex.in(ii), sy = live && p.ey.in(jj);
  T pyx = T(0), ax = T(0), bx = T(0), pxy = T(0), ay = T(0), by = T(0);
  if (sx) {
    pyx = p.apyx[q];
    ax = p.ex.a[ii];
    bx = p.ex.b[ii];
  }
  if (sy) {
    pxy = p.apxy[q];
    ay = p.ey.a[jj];
    by = p.ey.b[jj];
  }
  // cotangents of rhs = Dz - sum P and of eps through E = rhs / eps, or
  // through the fixed point E(k+1) = rhs / den(k), den(k) = eps + chi2 E(k)
  // + chi3 E(k)^2 from E(0) = rhs / eps (stepper_tm.py:224-231): the
  // iterates are recomputed from the taped rhs, then the three passes are
  // transposed last to first
  T rhs_b, eps_b;
  if (p.kerr) {
    const T x2 = p.chi2[c], x3 = p.chi3[c], rhs = ea;
    T E[4], den[3];
    E[0] = rhs / eps;
    for (int k = 0; k < 3; ++k) {
      den[k] = eps + x2 * E[k] + x3 * E[k] * E[k];
      E[k + 1] = rhs / den[k];
    }
    T eb = ezb;
    rhs_b = T(0);
    eps_b = T(0);
    T x2_b = T(0), x3_b = T(0);
    for (int k = 2; k >= 0; --k) {
      const T g = eb / den[k];      // dE(k+1)/drhs
      const T t = -g * E[k + 1];    // dE(k+1)/dden(k)
      rhs_b = rhs_b + g;
      eps_b = eps_b + t;
      x2_b = x2_b + t * E[k];
      x3_b = x3_b + t * (E[k] * E[k]);
      eb = t * (x2 + T(2) * x3 * E[k]);
    }
    rhs_b = rhs_b + eb / eps;
    eps_b = eps_b + (-eb * E[0]) / eps;
    if (p.chi2_grad) p.chi2_grad[c] = p.chi2_grad[c] + x2_b;
    if (p.chi3_grad) p.chi3_grad[c] = p.chi3_grad[c] + x3_b;
  } else {
    rhs_b = ezb / eps;
    eps_b = (-ezb * ea) / eps;
  }
  p.eps_grad[c] = eg + eps_b;
  T dzb = dz0 + rhs_b;
  p.aDz[c] = p.d1 ? d1 * dzb : dzb;          // Dz' = d1 Dz + d2 curl
  // poles: Ez' = (Dz' - sum P') / eps and P' = c1 P + c2 P_prev + c3 Ez, so
  // P' gains -Ez'b / eps; P then carries c1 P'b + P_prev'b (P_prev' = P),
  // P_prev carries c2 P'b, and the old Ez gains c3 P'b (phase C adds the
  // curl part to it)
  T ez_old = T(0);
  if (p.np) {
    const long long ncell = (long long)nx * ny;
    const T r = rhs_b;
    // P' = c1 P + c2 P_prev + c3 Ez(n): c3 gains P'b Ez(n)
    const T ezn = p.c3_grad ? adj_ez_before(p, ez_after, s_rel, c) : T(0);
    for (int n = 0; n < p.np; ++n) {
      const long long o = (long long)n * ncell + c;
      const T apn = p.aPc[o] - r, app = p.aPp[o];
      if (p.c3_grad) p.c3_grad[o] = p.c3_grad[o] + apn * ezn;
      p.aPc[o] = p.c1[o] * apn + app;
      p.aPp[o] = p.c2[o] * apn;
      ez_old = ez_old + p.c3[o] * apn;
    }
  }
  p.aEz[c] = ez_old;
  T ghy = T(0), ghx = T(0);
  if (live) {
    T cb = d2 * dzb;
    ghy = cb;
    if (sx) {
      T pt = pyx + cb;
      ghy = cb + ax * pt;
      p.apyx[q] = bx * pt;
    }
    ghx = -cb;
    if (sy) {
      T pt = pxy - cb;
      ghx = -cb + ay * pt;
      p.apxy[q] = by * pt;
    }
  }
  p.gHy[c] = ghy;
  p.gHx[c] = ghx;
}

// B: Hy/Hx cotangents (transposed differences of gHy/gHx), B-half psi
// cotangents, E-difference cotangents gEx, gEy.
template <typename T>
__device__ __forceinline__ void adj_b_cell(const AdjK<T>& p, int i, int j) {
  const int nx = p.nx, ny = p.ny;
  const bool fy = i < nx - 1, fx = j < ny - 1;
  const long long hy = (long long)i * ny + j;            // Hy (nx-1, ny), gHy, gEx
  const long long hx = (long long)i * (ny - 1) + j;      // Hx (nx, ny-1), gEy
  T ahy = T(0), gy0 = T(0), gy1 = T(0), ahx = T(0), gx0 = T(0), gx1 = T(0);
  T pzx = T(0), a_x = T(0), b_x = T(0), pzy = T(0), a_y = T(0), b_y = T(0);
  const bool sx = fy && p.hx.in(i), sy = fx && p.hy.in(j);
  if (fy) {
    ahy = p.aHy[hy];
    gy0 = p.gHy[hy];
    gy1 = p.gHy[hy + ny];
  }
  if (fx) {
    ahx = p.aHx[hx];
    gx0 = p.gHx[hy];
    gx1 = p.gHx[hy + 1];
  }
  if (sx) {
    pzx = p.apzx[hy];
    a_x = p.hx.a[i];
    b_x = p.hx.b[i];
  }
  if (sy) {
    pzy = p.apzy[hx];
    a_y = p.hy.a[j];
    b_y = p.hy.b[j];
  }
  // Hy' = h1y Hy + h2y curl, Hx' = h1x Hx - h2x curl (h1 = 1, h2 = dt without
  // conductivity)
  T h1y = T(1), h2y = p.dt, h1x = T(1), h2x = p.dt;
  if (p.h1y) {
    if (fy) h1y = p.h1y[hy], h2y = p.h2y[hy];
    if (fx) h1x = p.h1x[hx], h2x = p.h2x[hx];
  }
  if (fy) {
    T hb = ahy + (gy0 - gy1) / p.dx;
    if (p.nm) {   // the curl reaches H' through B'
      T bb;
      p.aHy[hy] = mpole_adj(p.nm, p.mc1y, p.mc2y, p.mc3y, p.aBy, p.aMyc, p.aMyp, hy, (long long)(nx - 1) * ny, hb,
                            &bb);
      hb = bb;
    } else {
      p.aHy[hy] = p.h1y ? h1y * hb : hb;
    }
    T gb = h2y * hb;
    if (sx) {
      T pt = pzx + h2y * hb;
      gb = gb + a_x * pt;
      p.apzx[hy] = b_x * pt;
    }
    p.gEx[hy] = gb;
  }
  if (fx) {
    T hb = ahx + (gx0 - gx1) / p.dx;
    if (p.nm) {
      T bb;
      p.aHx[hx] = mpole_adj(p.nm, p.mc1x, p.mc2x, p.mc3x, p.aBx, p.aMxc, p.aMxp, hx, (long long)nx * (ny - 1), hb,
                            &bb);
      hb = bb;
    } else {
      p.aHx[hx] = p.h1y ? h1x * hb : hb;
    }
    T gb = -h2x * hb;
    if (sy) {
      T pt = pzy - h2x * hb;
      gb = gb + a_y * pt;
      p.apzy[hx] = b_y * pt;
    }
    p.gEy[hx] = gb;
  }
}

// C: Ez cotangent of x_t = transposed differences of gEx, gEy.
template <typename T>
__device__ __forceinline__ void adj_c_cell(const AdjK<T>& p, int i, int j) {
  const int nx = p.nx, ny = p.ny;
  T gxm = i >= 1 ? p.gEx[(long long)(i - 1) * ny + j] : T(0);
  T gx0 = i < nx - 1 ? p.gEx[(long long)i * ny + j] : T(0);
  T gym = j >= 1 ? p.gEy[(long long)i * (ny - 1) + j - 1] : T(0);
  T gy0 = j < ny - 1 ? p.gEy[(long long)i * (ny - 1) + j] : T(0);
  const long long c = (long long)i * ny + j;
  const T v = (gxm - gx0) / p.dx + (gym - gy0) / p.dx;
  p.aEz[c] = p.np ? p.aEz[c] + v : v;
}

// ---------------------------------------------------------------------------
// One-pass form of A, B and C for problems without conductivity or poles
// (the batched fused adjoint step): the same arithmetic, split so that a
// tile can evaluate A and B on its cells and on a one-cell halo, keep gHy,
// gHx, gEx, gEy in LDS, and write the new cotangents of its own cells to a
// second set (r reads the cotangents of step t+1, w receives those of step
// t; eps, profiles and eps_grad are shared).

// A at (i, j): returns gHy, gHx; with own, also writes eps_grad, w.aDz and
// the D-half psi cotangents of the cell.  (aEz is not written: without
// poles C overwrites it.)
template <typename T>
__device__ __forceinline__ void adj_a_fused(const AdjK<T>& r, const AdjK<T>& w, const T* ez_after, int i, int j,
                                            bool own, T& ghy, T& ghx) {
  const int nx = r.nx, ny = r.ny;
  const long long c = (long long)i * ny + j;
  const bool live = i >= 1 && i < nx - 1 && (r.y_pmc || (j >= 1 && j < ny - 1));
  const int ii = i - 1, jj = r.y_pmc ? j : j - 1;
  const long long q = (long long)ii * r.ny_e + jj;
  const T eps = r.eps[c], ezb = r.aEz[c], dz0 = r.aDz[c];
  const bool sx = live && r.ex.in(ii), sy = live && r.ey.in(jj);
  T pyx = T(0), ax = T(0), bx = T(0), pxy = T(0), ay = T(0), by = T(0);
  if (sx) {
    pyx = r.apyx[q];
    ax = r.ex.a[ii];
    bx = r.ex.b[ii];
  }
  if (sy) {
    pxy = r.apxy[q];
    ay = r.ey.a[jj];
    by = r.ey.b[jj];
  }
  if (own) {
    const T eg = r.eps_grad[c], ea = ez_after[c];
    w.eps_grad[c] = eg + (-ezb * ea) / eps;
  }
  const T dzb = dz0 + ezb / eps;
  if (own) w.aDz[c] = dzb;
  ghy = T(0);
  ghx = T(0);
  if (live) {
    const T cb = r.dt * dzb;
    ghy = cb;
    if (sx) {
      const T pt = pyx + cb;
      ghy = cb + ax * pt;
      if (own) w.apyx[q] = bx * pt;
    }
    ghx = -cb;
    if (sy) {
      const T pt = pxy - cb;
      ghx = -cb + ay * pt;
      if (own) w.apxy[q] = by * pt;
    }
  }
}

// Hy part of B at (i, j) (needs i < nx-1): gy0 = gHy(i, j), gy1 = gHy(i+1, j);
// returns gEx, with own writes the Hy cotangent and its psi cotangent.
template <typename T>
__device__ __forceinline__ T adj_b_hy_fused(const AdjK<T>& r, const AdjK<T>& w, int i, int j, T gy0, T gy1,
                                            bool own) {
  const long long hy = (long long)i * r.ny + j;
  const bool sx = r.hx.in(i);
  const T ahy = r.aHy[hy];
  T pzx = T(0), a_x = T(0), b_x = T(0);
  if (sx) {
    pzx = r.apzx[hy];
    a_x = r.hx.a[i];
    b_x = r.hx.b[i];
  }
  const T h2y = r.dt;
  const T hb = ahy + (gy0 - gy1) / r.dx;
  if (own) w.aHy[hy] = hb;
  T gb = h2y * hb;
  if (sx) {
    const T pt = pzx + h2y * hb;
    gb = gb + a_x * pt;
    if (own) w.apzx[hy] = b_x * pt;
  }
  return gb;
}

// The B-phase inputs of one cell (the H cotangents and their psi
// cotangents with the profile coefficients), loaded ahead: they do not
// depend on the A phase, so a kernel can issue them before its first barrier.
template <typename T>
struct BIn {
  T ahy, pzx, a_x, b_x, ahx, pzy, a_y, b_y;
  bool sx, sy;
};

template <typename T>
__device__ __forceinline__ BIn<T> adj_b_load(const AdjK<T>& r, int i, int j, bool hy_ok, bool hx_ok) {
  BIn<T> v{T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), false, false};
  if (hy_ok) {
    const long long hy = (long long)i * r.ny + j;
    v.sx = r.hx.in(i);
    v.ahy = r.aHy[hy];
    if (v.sx) {
      v.pzx = r.apzx[hy];
      v.a_x = r.hx.a[i];
      v.b_x = r.hx.b[i];
    }
  }
  if (hx_ok) {
    const long long hx = (long long)i * (r.ny - 1) + j;
    v.sy = r.hy.in(j);
    v.ahx = r.aHx[hx];
    if (v.sy) {
      v.pzy = r.apzy[hx];
      v.a_y = r.hy.a[j];
      v.b_y = r.hy.b[j];
    }
  }
  return v;
}

// adj_b_hy_fused / adj_b_hx_fused from preloaded inputs (same arithmetic);
// Store = false for halo cells, whose cotangents another tile owns
template <typename T, bool Store = true>
__device__ __forceinline__ T adj_b_hy_pre(const AdjK<T>& r, const AdjK<T>& w, int i, int j, const BIn<T>& v,
                                          T gy0, T gy1) {
  const long long hy = (long long)i * r.ny + j;
  const T h2y = r.dt;
  const T hb = v.ahy + (gy0 - gy1) / r.dx;
  if (Store) w.aHy[hy] = hb;
  T gb = h2y * hb;
  if (v.sx) {
    const T pt = v.pzx + h2y * hb;
    gb = gb + v.a_x * pt;
    if (Store) w.apzx[hy] = v.b_x * pt;
  }
  return gb;
}

template <typename T, bool Store = true>
__device__ __forceinline__ T adj_b_hx_pre(const AdjK<T>& r, const AdjK<T>& w, int i, int j, const BIn<T>& v,
                                          T gx0, T gx1) {
  const long long hx = (long long)i * (r.ny - 1) + j;
  const T h2x = r.dt;
  const T hb = v.ahx + (gx0 - gx1) / r.dx;
  if (Store) w.aHx[hx] = hb;
  T gb = -h2x * hb;
  if (v.sy) {
    const T pt = v.pzy - h2x * hb;
    gb = gb + v.a_y * pt;
    if (Store) w.apzy[hx] = v.b_y * pt;
  }
  return gb;
}

// Hx part of B at (i, j) (needs j < ny-1): gx0 = gHx(i, j), gx1 = gHx(i, j+1);
// returns gEy, with own writes the Hx cotangent and its psi cotangent.
template <typename T>
__device__ __forceinline__ T adj_b_hx_fused(const AdjK<T>& r, const AdjK<T>& w, int i, int j, T gx0, T gx1,
                                            bool own) {
  const long long hx = (long long)i * (r.ny - 1) + j;
  const bool sy = r.hy.in(j);
  const T ahx = r.aHx[hx];
  T pzy = T(0), a_y = T(0), b_y = T(0);
  if (sy) {
    pzy = r.apzy[hx];
    a_y = r.hy.a[j];
    b_y = r.hy.b[j];
  }
  const T h2x = r.dt;
  const T hb = ahx + (gx0 - gx1) / r.dx;
  if (own) w.aHx[hx] = hb;
  T gb = -h2x * hb;
  if (sy) {
    const T pt = pzy - h2x * hb;
    gb = gb + a_y * pt;
    if (own) w.apzy[hx] = b_y * pt;
  }
  return gb;
}

// Adds the cotangent vb of element e of monitor term tm into the adjoint
// array(s) it samples (half to each of the two averaged arrays).
template <typename T>
__device__ __forceinline__ void adj_inject_add(const mt_monitor& m, const mt_dft_term& tm, int e, T vb) {
  int i0 = e / m.n1, i1 = e - i0 * m.n1;
  long long off = (long long)i0 * tm.s0 + (long long)i1 * tm.s1;
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

// Monitor cotangent of element e of term q at global step t (one thread
// sums every frequency).
template <typename T>
__device__ __forceinline__ void adj_inject_elem(const mt_monitor& m, int q, int e, int t, double dt_d, int t_rel) {
  int n = m.n0 * m.n1;
  const mt_dft_term& tm = m.term[q];
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
  adj_inject_add<T>(m, tm, e, vb);
}

// Block-cooperative injection: a block of kInjE * kInjS threads owns the
// monitor elements [e0, e0 + kInjE).  The phase factors dt cos(w t),
// -dt sin(w t) of every (term, frequency) are evaluated once per block into
// LDS (not once per element), each of the kInjS lane groups sums every
// kInjS-th frequency, and lane group 0 adds the partial sums in group order.
// Monitors with more than kInjTab (term, frequency) pairs, and probe stacks,
// take the one-thread-per-element form.
constexpr int kInjE = 32, kInjS = 8, kInjTab = 1024;

template <typename T>
__device__ void adj_inject_block(const mt_monitor& m, int e0, int t, double dt_d, int t_rel) {
  __shared__ T cs[2 * kInjTab];
  __shared__ T part[kInjS][kInjE];
  const int n = m.n0 * m.n1;
  const int el = threadIdx.x % kInjE, sl = threadIdx.x / kInjE;
  const int e = e0 + el;
  const int nq = m.n_terms, nf = m.nfreq;
  if (m.kind == 1 || nq * nf > kInjTab) {   // uniform over the block
    if (sl == 0 && e < n)
      for (int q = 0; q < nq; ++q) adj_inject_elem<T>(m, q, e, t, dt_d, t_rel);
    return;
  }
  const T dt = static_cast<T>(dt_d);
  const T* om = static_cast<const T*>(m.omega);
  const T* ac = static_cast<const T*>(m.acc);
  for (int x = threadIdx.x; x < nq * nf; x += kInjE * kInjS) {
    int q = x / nf, f = x - q * nf;
    double tt = ((double)t + (m.term[q].t_half ? 0.5 : 1.0)) * dt_d;
    T ph = om[f] * static_cast<T>(tt);
    cs[2 * x] = dt * cos(ph);
    cs[2 * x + 1] = -(dt * sin(ph));
  }
  __syncthreads();
  for (int q = 0; q < nq; ++q) {
    T v = T(0);
    if (e < n)
      for (int f = sl; f < nf; f += kInjS) {
        const T* c = cs + 2 * (q * nf + f);
        T re = ac[((long long)f * 2 * nq + 2 * q) * n + e];
        T im = ac[((long long)f * 2 * nq + 2 * q + 1) * n + e];
        v = v + c[0] * re + c[1] * im;
      }
    part[sl][el] = v;
    __syncthreads();
    if (sl == 0 && e < n) {
      T vb = part[0][el];
      for (int g = 1; g < kInjS; ++g) vb = vb + part[g][el];
      adj_inject_add<T>(m, m.term[q], e, vb);
    }
    __syncthreads();
  }
}

template <typename T>
AdjK<T> make_adjk(const mt_tm2d* d, const mt_tm2d_adj* a) {
  AdjK<T> p;
  p.nx = d->nx;
  p.ny = d->ny;
  p.y_pmc = d->y_pmc;
  p.ycut = d->y_mirror == 1;
  p.xcut = d->x_mirror == 1;
  p.ny_e = d->y_pmc ? d->ny : (p.ycut ? d->ny - 1 : d->ny - 2);
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
  p.d1 = static_cast<const T*>(d->d1);
  p.d2 = static_cast<const T*>(d->d2);
  p.h1x = static_cast<const T*>(d->h1x);
  p.h2x = static_cast<const T*>(d->h2x);
  p.h1y = static_cast<const T*>(d->h1y);
  p.h2y = static_cast<const T*>(d->h2y);
  p.np = (d->n_poles > 0 && a->aP_cur) ? d->n_poles : 0;
  p.c1 = static_cast<const T*>(d->c1);
  p.c2 = static_cast<const T*>(d->c2);
  p.c3 = static_cast<const T*>(d->c3);
  p.aPc = static_cast<T*>(a->aP_cur);
  p.aPp = static_cast<T*>(a->aP_prev);
  p.kerr = d->kerr;
  p.chi2 = static_cast<const T*>(d->chi2);
  p.chi3 = static_cast<const T*>(d->chi3);
  p.chi2_grad = static_cast<T*>(a->chi2_grad);
  p.chi3_grad = static_cast<T*>(a->chi3_grad);
  p.c3_grad = p.np ? static_cast<T*>(a->c3_grad) : nullptr;
  p.ez_init = static_cast<const T*>(a->ez_init);
  p.nm = (d->n_mpoles > 0 && a->aBx) ? d->n_mpoles : 0;
  p.mc1x = static_cast<const T*>(d->mc1x);
  p.mc2x = static_cast<const T*>(d->mc2x);
  p.mc3x = static_cast<const T*>(d->mc3x);
  p.mc1y = static_cast<const T*>(d->mc1y);
  p.mc2y = static_cast<const T*>(d->mc2y);
  p.mc3y = static_cast<const T*>(d->mc3y);
  p.aBx = static_cast<T*>(a->aBx);
  p.aBy = static_cast<T*>(a->aBy);
  p.aMxc = static_cast<T*>(a->aMx_cur);
  p.aMxp = static_cast<T*>(a->aMx_prev);
  p.aMyc = static_cast<T*>(a->aMy_cur);
  p.aMyp = static_cast<T*>(a->aMy_prev);
  return p;
}

}  // namespace mt
