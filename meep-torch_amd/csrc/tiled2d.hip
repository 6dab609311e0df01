// Temporally blocked 2D TM / TE steppers: NS Yee steps per launch on
// overlapped LDS tiles.
//
// A workgroup owns a TX x TY block of the (i, j) cell lattice.  It loads the
// block plus a halo (NS+1 cells on the low sides, NS on the high sides) of
// the evolving fields into LDS, advances NS steps there, and writes back only
// the cells it owns.  Per half-step the stencil reaches one cell (D/E halves
// towards lower indices, H/B halves towards higher ones), so after NS steps
// every owned cell, and the one row/column below it that averaged H monitor
// samples need, is exact.  Halo cells are recomputed redundantly; cells whose
// neighbours fall outside the tile compute throw-away values that never reach
// an owned cell.  Reads come from one set of global buffers and writes go to
// the other (ping-pong), so tiles never see each other's partial results.
//
// Arithmetic per cell is the per-half-step kernels' (yee2d.hip), op for op,
// so results are bit-identical to them and to the reference's FP32 stepper.
// Tiles whose halo touches a CPML shell also run the four psi recursions
// (template PML = true), with each cell's psi values in the registers of
// the thread that updates it (the recursion is pointwise, so psi needs no
// LDS); interior tiles carry none and run in a separate launch.  Monitor samples are taken inside the tile after each step by the
// tile that owns the sampled cell and written to the staging ring consumed by
// the DFT flush (dft.hip).
#include <hip/hip_runtime.h>

#include "common.h"
#include "tiled2d.h"

#include <algorithm>

namespace mt {
namespace {

// Threads per tile workgroup.  TM tiles run 1024 threads (3-6 cells per
// thread at NS = 4): the TM step is light and the design loops' grids are
// small, so more waves per tile shorten each launch.  TE interior tiles run
// 512: on 2048^2 the interior TE launch is arithmetic-bound and 1024-thread
// interior tiles measured slower (98 -> 123 us per launch).
constexpr int BLK_TM = 1024;
// TE: 512 threads for interior tiles; CPML ring tiles (few, each carrying
// the psi recursions, so the ring launch is one latency chain per CU) run
// 1024 threads, which shortens each thread's share of a tile.
constexpr int te_block(bool pml) { return pml ? 1024 : 512; }

template <typename T>
__device__ __forceinline__ T dget(const T* p, long long idx, bool ok) {
  return ok ? p[idx] : T(0);
}

// Tile (tr, tc) for linear index b of the interior rectangle or of the ring.
__device__ __forceinline__ void tile_linear(bool pml, const TileMap& m, int b, int* tr, int* tc) {
  if (!pml) {
    int w = m.c1 - m.c0;
    *tr = m.r0 + b / w;
    *tc = m.c0 + b % w;
    return;
  }
  if (b < m.r0 * m.C) {
    *tr = b / m.C;
    *tc = b % m.C;
    return;
  }
  b -= m.r0 * m.C;
  if (b < (m.R - m.r1) * m.C) {
    *tr = m.r1 + b / m.C;
    *tc = b % m.C;
    return;
  }
  b -= (m.R - m.r1) * m.C;
  int per = m.c0 + (m.C - m.c1);
  *tr = m.r0 + b / per;
  int k = b % per;
  *tc = k < m.c0 ? k : m.c1 + (k - m.c0);
}

// Tile (tr, tc) of this workgroup: interior launches cover the rectangle
// [r0, r1) x [c0, c1); PML launches enumerate the ring of tiles around it.
__device__ __forceinline__ void tile_of(bool pml, const TileMap& m, int* tr, int* tc) {
  if (!pml) {
    *tr = m.r0 + blockIdx.y;
    *tc = m.c0 + blockIdx.x;
    return;
  }
  int b = blockIdx.x;
  if (b < m.r0 * m.C) {
    *tr = b / m.C;
    *tc = b % m.C;
    return;
  }
  b -= m.r0 * m.C;
  if (b < (m.R - m.r1) * m.C) {
    *tr = m.r1 + b / m.C;
    *tc = b % m.C;
    return;
  }
  b -= (m.R - m.r1) * m.C;
  int per = m.c0 + (m.C - m.c1);
  *tr = m.r0 + b / per;
  int k = b % per;
  *tc = k < m.c0 ? k : m.c1 + (k - m.c0);
}

// ---------------------------------------------------------------- taps
// Whether any monitor term has cells in the owned block [I0, I0+TXo) x
// [J0, J0+TYo) (the cells write_taps writes).  Uniform over the workgroup:
// a tile without taps skips write_taps and the barrier after it.
__device__ __forceinline__ bool taps_touch(const TapSet& taps, int I0, int J0, int TXo, int TYo) {
  for (int q = 0; q < taps.n; ++q) {
    const Tap& t = taps.t[q];
    if (t.line == 1) {
      if (t.fix >= I0 && t.fix < I0 + TXo && t.lo < J0 + TYo && t.lo + t.n > J0) return true;
    } else if (t.line == 2) {
      if (t.fix >= J0 && t.fix < J0 + TYo && t.lo < I0 + TXo && t.lo + t.n > I0) return true;
    } else if (t.line == 3) {
      if (I0 < t.n0 && J0 < t.n1) return true;
    } else if (t.line == 4) {
      if (t.fix >= I0 && t.fix < I0 + TXo && t.lo >= J0 && t.lo < J0 + TYo) return true;
    } else {
      return true;
    }
  }
  return false;
}

// Write this tile's share of every monitor term after sub-step `s`.
// FIELD(id, li, lj) returns the LDS value of field `id` at local cell.
template <typename T, int RX, int RY, typename F>
__device__ void write_taps(const TapSet& taps, int s, int I0, int J0, int TXo, int TYo, int a0, int b0,
                           F field, int slot_base = 0) {
  const int tid = threadIdx.x;
  for (int q = 0; q < taps.n; ++q) {
    const Tap& t = taps.t[q];
    T* out = static_cast<T*>(t.out) + (long long)(t.slot0 + slot_base + s) * t.slot_stride;
    if (t.line == 1) {  // row: fixed i = fix, j = lo + e
      if (t.fix < I0 || t.fix >= I0 + TXo) continue;
      int jlo = max(t.lo, J0), jhi = min(t.lo + t.n, J0 + TYo);
      for (int j = jlo + tid; j < jhi; j += blockDim.x) {
        int li = t.fix - a0, lj = j - b0;
        T v = field(t.field, li, lj);
        if (t.avg) v = T(0.5) * (field(t.field, li - 1, lj) + v);
        out[j - t.lo] = v;
      }
    } else if (t.line == 2) {  // column: fixed j = fix, i = lo + e
      if (t.fix < J0 || t.fix >= J0 + TYo) continue;
      int ilo = max(t.lo, I0), ihi = min(t.lo + t.n, I0 + TXo);
      for (int i = ilo + tid; i < ihi; i += blockDim.x) {
        int li = i - a0, lj = t.fix - b0;
        T v = field(t.field, li, lj);
        if (t.avg) v = T(0.5) * (field(t.field, li, lj - 1) + v);
        out[i - t.lo] = v;
      }
    } else if (t.line == 3) {  // whole field, row-major with row length t.n1
      int ihi = min(I0 + TXo, t.n0), jhi = min(J0 + TYo, t.n1);
      int w = jhi - J0;
      if (I0 >= ihi || w <= 0) continue;
      int cnt = (ihi - I0) * w;
      for (int c = tid; c < cnt; c += blockDim.x) {
        int i = I0 + c / w, j = J0 + c % w;
        out[(long long)i * t.n1 + j] = field(t.field, i - a0, j - b0);
      }
    } else if (t.line == 4) {  // single cell (fix, lo); stacked per step
      if (tid == 0 && t.fix >= I0 && t.fix < I0 + TXo && t.lo >= J0 && t.lo < J0 + TYo)
        out[0] = field(t.field, t.fix - a0, t.lo - b0);
    }
  }
}

// ------------------------------------------------------------------- TE
template <typename T, int TX, int TY, int NS, bool PML>
__global__ __launch_bounds__(te_block(PML)) void te_tiled(TeTiled<T> p, TapSet taps, int nsub, int write_e) {
  constexpr int BLK_TE = te_block(PML);
  constexpr int RX = TX + 2 * NS + 1, RY = TY + 2 * NS + 1, RC = RX * RY;
  constexpr int PER = (RC + BLK_TE - 1) / BLK_TE;
  __shared__ T sHz[RC], sDx[RC], sDy[RC], sEx[RC], sEy[RC];
  // The psi recursions are pointwise (a cell's psi feeds only that cell's
  // update), so the thread that updates a cell keeps its four psi values in
  // registers and writes back those of its owned cells at the end.
  constexpr int PP = PML ? PER : 1;
  T vPzy[PP], vPzx[PP], vPyx[PP], vPxy[PP];  // psi_Hzy (Dx, y shell, (i, j-1)), psi_Hzx (Dy, x shell,
                                             // (i-1, j)), psi_Eyx / psi_Exy (Hz, x / y shell, (i, j))
  const int nx = p.nx, ny = p.ny;
  int tr, tc;
  tile_of(PML, p.map, &tr, &tc);
  const int I0 = tr * TX, J0 = tc * TY;
  const int a0 = I0 - NS - 1, b0 = J0 - NS - 1;
  const int tid = threadIdx.x;
  const T dt = p.dt, dx = p.dx, rdx = p.rdx;
  // A thread updates the same PER cells of the LDS region in every
  // sub-step, so eps and the D fields of those cells live in its registers:
  // the update reads them without LDS traffic, and the LDS holds only what
  // neighbours, taps and the write-back read (fewer bytes per workgroup,
  // more workgroups per CU).
  T vDx[PER], vDy[PER], vKx[PER], vKy[PER];
  {
    T vHz[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TE;
      int i = a0 + c / RY, j = b0 + c % RY;
      bool in = c < RC;
      bool hz = in && i >= 0 && i < nx - 1 && j >= 0 && j < ny - 1;
      bool ex = in && i >= 0 && i < nx - 1 && j >= 0 && j < ny;
      bool ey = in && i >= 0 && i < nx && j >= 0 && j < ny - 1;
      vHz[u] = dget(p.Hz, (long long)i * (ny - 1) + j, hz);
      vDx[u] = dget(p.Dx, (long long)i * ny + j, ex);
      vDy[u] = dget(p.Dy, (long long)i * (ny - 1) + j, ey);
      vKx[u] = ex ? p.eps_x[(long long)i * ny + j] : T(1);
      vKy[u] = ey ? p.eps_y[(long long)i * (ny - 1) + j] : T(1);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TE;
      if (c < RC) {
        sHz[c] = vHz[u];
        sDx[c] = vDx[u];
        sDy[c] = vDy[u];
      }
    }
  }
  if constexpr (PML) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TE;
      int i = a0 + c / RY, j = b0 + c % RY;
      bool in = c < RC;
      bool hz = in && i >= 0 && i < nx - 1 && j >= 0 && j < ny - 1;
      bool ex = in && i >= 0 && i < nx - 1 && j >= 0 && j < ny;
      bool ey = in && i >= 0 && i < nx && j >= 0 && j < ny - 1;
      vPzy[u] = dget(p.pHzy, (long long)i * (ny - 2) + (j - 1), ex && j >= 1 && j < ny - 1 && p.ey.in(j - 1));
      vPzx[u] = dget(p.pHzx, (long long)(i - 1) * (ny - 1) + j, ey && i >= 1 && i < nx - 1 && p.ex.in(i - 1));
      vPyx[u] = dget(p.pEyx, (long long)i * (ny - 1) + j, hz && p.hx.in(i));
      vPxy[u] = dget(p.pExy, (long long)i * (ny - 1) + j, hz && p.hy.in(j));
    }
  }
  __syncthreads();
  // uniform per tile: does the source reach the LDS region, do taps reach the owned block
  const bool src_tile = p.src.present && p.src.i0 < a0 + RX && p.src.i1 > a0 && p.src.j0 < b0 + RY &&
                        p.src.j1 > b0;
  const bool tap_tile = taps.n && taps_touch(taps, I0, J0, TX, TY);
  for (int s = 0; s < nsub; ++s) {
    const T coef = p.coef[s];
    // D half (stepper_te.py:80-103,131-134,160-168)
    int li = tid / RY, lj = tid % RY;
#pragma unroll
    for (int u = 0; u < PER; ++u, li += BLK_TE / RY, lj += BLK_TE % RY) {
      if (lj >= RY) {
        lj -= RY;
        ++li;
      }
      int c = li * RY + lj;
      if (c >= RC) break;
      int i = a0 + li, j = b0 + lj;
      if (!PML || (i >= 0 && i < nx - 1 && j >= 0 && j < ny)) {
        T dxv = vDx[u];
        if ((!PML || (j >= 1 && j < ny - 1)) && lj >= 1) {
          T d = div_c(sHz[c] - sHz[c - 1], dx, rdx);
          T curl = d;
          if (PML && p.ey.in(j - 1)) {
            T ps = p.ey.b[j - 1] * vPzy[PML ? u : 0] + p.ey.a[j - 1] * d;
            vPzy[PML ? u : 0] = ps;
            curl = d + ps;
          }
          dxv = dxv + dt * curl;
        }
        vDx[u] = dxv;
        sDx[c] = dxv;
        sEx[c] = dxv / vKx[u];
      }
      if (!PML || (i >= 0 && i < nx && j >= 0 && j < ny - 1)) {
        T dyv = vDy[u];
        if ((!PML || (i >= 1 && i < nx - 1)) && li >= 1) {
          T d = div_c(sHz[c] - sHz[c - RY], dx, rdx);
          T curl = d;
          if (PML && p.ex.in(i - 1)) {
            T ps = p.ex.b[i - 1] * vPzx[PML ? u : 0] + p.ex.a[i - 1] * d;
            vPzx[PML ? u : 0] = ps;
            curl = d + ps;
          }
          dyv = dyv - dt * curl;
        }
        if (src_tile && i >= p.src.i0 && i < p.src.i1 && j >= p.src.j0 && j < p.src.j1)
          dyv = dyv - coef * static_cast<const T*>(p.src.w)[(long long)s * p.src.wstep + (long long)(i - p.src.i0) * (p.src.j1 - p.src.j0) +
                                                             (j - p.src.j0)];
        vDy[u] = dyv;
        sDy[c] = dyv;
        sEy[c] = dyv / vKy[u];
      }
    }
    __syncthreads();
    // B half (stepper_te.py:106-117)
    li = tid / RY;
    lj = tid % RY;
#pragma unroll
    for (int u = 0; u < PER; ++u, li += BLK_TE / RY, lj += BLK_TE % RY) {
      if (lj >= RY) {
        lj -= RY;
        ++li;
      }
      int c = li * RY + lj;
      if (c >= RC) break;
      int i = a0 + li, j = b0 + lj;
      if ((!PML || (i >= 0 && i < nx - 1 && j >= 0 && j < ny - 1)) && li < RX - 1 && lj < RY - 1) {
        T dEy = div_c(sEy[c + RY] - sEy[c], dx, rdx);
        T dEx = div_c(sEx[c + 1] - sEx[c], dx, rdx);
        T a1 = dEy;
        if (PML && p.hx.in(i)) {
          T ps = p.hx.b[i] * vPyx[PML ? u : 0] + p.hx.a[i] * dEy;
          vPyx[PML ? u : 0] = ps;
          a1 = dEy + ps;
        }
        T curl = a1 - dEx;
        if (PML && p.hy.in(j)) {
          T ps = p.hy.b[j] * vPxy[PML ? u : 0] + p.hy.a[j] * dEx;
          vPxy[PML ? u : 0] = ps;
          curl = curl - ps;
        }
        sHz[c] = sHz[c] - dt * curl;
      }
    }
    __syncthreads();
    if (tap_tile) {
      write_taps<T, RX, RY>(taps, s, I0, J0, TX, TY, a0, b0, [&](int f, int li, int lj) -> T {
        int c = li * RY + lj;
        return f == 0 ? sHz[c] : f == 1 ? sEx[c] : f == 2 ? sEy[c] : f == 3 ? sDx[c] : sDy[c];
      });
      __syncthreads();
    }
  }
  // write back the owned block
  for (int c = tid; c < TX * TY; c += BLK_TE) {
    int i = I0 + c / TY, j = J0 + c % TY;
    int lc = (i - a0) * RY + (j - b0);
    if (i < nx - 1 && j < ny - 1) {
      long long h = (long long)i * (ny - 1) + j;
      p.oHz[h] = sHz[lc];
    }
    if (i < nx - 1 && j < ny) {
      long long h = (long long)i * ny + j;
      p.oDx[h] = sDx[lc];
      if (write_e) p.Ex[h] = sEx[lc];
    }
    if (i < nx && j < ny - 1) {
      long long h = (long long)i * (ny - 1) + j;
      p.oDy[h] = sDy[lc];
      if (write_e) p.Ey[h] = sEy[lc];
    }
  }
  if constexpr (PML) {  // psi of the owned cells, from the registers of the threads that updated them
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TE;
      int i = a0 + c / RY, j = b0 + c % RY;
      if (c >= RC || i < I0 || i >= I0 + TX || j < J0 || j >= J0 + TY) continue;
      if (i < nx - 1 && j < ny - 1) {
        long long h = (long long)i * (ny - 1) + j;
        if (p.hx.in(i)) p.oEyx[h] = vPyx[u];
        if (p.hy.in(j)) p.oExy[h] = vPxy[u];
      }
      if (i < nx - 1 && j < ny && j >= 1 && j < ny - 1 && p.ey.in(j - 1))
        p.oHzy[(long long)i * (ny - 2) + (j - 1)] = vPzy[u];
      if (i < nx && j < ny - 1 && i >= 1 && i < nx - 1 && p.ex.in(i - 1))
        p.oHzx[(long long)(i - 1) * (ny - 1) + j] = vPzx[u];
    }
  }
}

// ------------------------------------------------------------------- TM
template <typename T, int TX, int TY, int NS, bool PML>
__device__ __forceinline__ void tm_tile(const TmTiled<T>& p, const TapSet& taps, int nsub, int tape_s0, int tr,
                                        int tc, int slot_base, int coef_t0) {
  constexpr int RX = TX + 2 * NS + 1, RY = TY + 2 * NS + 1, RC = RX * RY;
  constexpr int PER = (RC + BLK_TM - 1) / BLK_TM;
  __shared__ T sEz[RC], sDz[RC], sHx[RC], sHy[RC];
  // psi in the registers of the thread that updates the cell (as in te_tiled)
  constexpr int PP = PML ? PER : 1;
  T vPzy[PP], vPzx[PP], vPyx[PP], vPxy[PP];  // psi_Ezy (Hx, y shell, (i, j)), psi_Ezx (Hy, x shell, (i, j)),
                                             // psi_Hyx / psi_Hxy (Dz, x / y shell, (i-1, jj))
  const int nx = p.nx, ny = p.ny, ny_e = p.ny_e;
  const int I0 = tr * TX, J0 = tc * TY;
  const int a0 = I0 - NS - 1, b0 = J0 - NS - 1;
  const int tid = threadIdx.x;
  const T dt = p.dt, dx = p.dx, rdx = p.rdx;
  const int joff = p.y_pmc ? 0 : 1;
  // eps and Dz of the PER cells this thread updates stay in its registers
  // across sub-steps (as in te_tiled)
  T vDz[PER], vK[PER];
  {
    T vEz[PER], vHx[PER], vHy[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TM;
      int i = a0 + c / RY, j = b0 + c % RY;
      bool ez = c < RC && i >= 0 && i < nx && j >= 0 && j < ny;
      vEz[u] = dget(p.Ez, (long long)i * ny + j, ez);
      vDz[u] = dget(p.Dz, (long long)i * ny + j, ez);
      vHx[u] = dget(p.Hx, (long long)i * (ny - 1) + j, ez && j < ny - 1);
      vHy[u] = dget(p.Hy, (long long)i * ny + j, ez && i < nx - 1);
      vK[u] = ez ? p.eps[(long long)i * ny + j] : T(1);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TM;
      if (c < RC) {
        sEz[c] = vEz[u];
        sDz[c] = vDz[u];
        sHx[c] = vHx[u];
        sHy[c] = vHy[u];
      }
    }
  }
  if constexpr (PML) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TM;
      int i = a0 + c / RY, j = b0 + c % RY;
      bool ez = c < RC && i >= 0 && i < nx && j >= 0 && j < ny;
      bool hx = ez && j < ny - 1;
      bool hy = ez && i < nx - 1;
      vPzy[u] = dget(p.pEzy, (long long)i * (ny - 1) + j, hx && p.hy.in(j));
      vPzx[u] = dget(p.pEzx, (long long)i * ny + j, hy && p.hx.in(i));
      int jj = j - joff;
      bool live = ez && i >= 1 && i < nx - 1 && jj >= 0 && jj < ny_e;
      long long q = (long long)(i - 1) * ny_e + jj;
      vPyx[u] = dget(p.pHyx, q, live && p.ex.in(i - 1));
      vPxy[u] = dget(p.pHxy, q, live && p.ey.in(jj));
    }
  }
  __syncthreads();
  // uniform per tile: does the source reach the LDS region, do taps reach the owned block
  const bool src_tile = p.src.present && p.src.i0 < a0 + RX && p.src.i1 > a0 && p.src.j0 < b0 + RY &&
                        p.src.j1 > b0;
  const bool hx_tile = p.hsrc[0].present && p.hsrc[0].i0 < a0 + RX && p.hsrc[0].i1 > a0 &&
                       p.hsrc[0].j0 < b0 + RY && p.hsrc[0].j1 > b0;
  const bool hy_tile = p.hsrc[1].present && p.hsrc[1].i0 < a0 + RX && p.hsrc[1].i1 > a0 &&
                       p.hsrc[1].j0 < b0 + RY && p.hsrc[1].j1 > b0;
  const bool tap_tile = taps.n && taps_touch(taps, I0, J0, TX, TY);
  for (int s = 0; s < nsub; ++s) {
    const T coef = p.coef_tab ? p.coef_tab[coef_t0 + s] : p.coef[s];
    const T hc = p.hcoef_tab ? p.hcoef_tab[coef_t0 + s] : p.hcoef[s];
    // B half (stepper_tm.py:85-92,126-128)
    int li = tid / RY, lj = tid % RY;
#pragma unroll
    for (int u = 0; u < PER; ++u, li += BLK_TM / RY, lj += BLK_TM % RY) {
      if (lj >= RY) {
        lj -= RY;
        ++li;
      }
      int c = li * RY + lj;
      if (c >= RC) break;
      int i = a0 + li, j = b0 + lj;
      if (PML && (i < 0 || i >= nx || j < 0 || j >= ny)) continue;
      T ez = sEz[c];
      if ((!PML || j < ny - 1) && lj < RY - 1) {
        T d = div_c(sEz[c + 1] - ez, dx, rdx);
        T curl = d;
        if (PML && p.hy.in(j)) {
          T ps = p.hy.b[j] * vPzy[PML ? u : 0] + p.hy.a[j] * d;
          vPzy[PML ? u : 0] = ps;
          curl = d + ps;
        }
        T hxv = sHx[c] - dt * curl;
        if (hx_tile && i >= p.hsrc[0].i0 && i < p.hsrc[0].i1 && j >= p.hsrc[0].j0 && j < p.hsrc[0].j1)
          hxv = hxv - hc * static_cast<const T*>(p.hsrc[0].w)[(long long)(i - p.hsrc[0].i0) *
                                                                   (p.hsrc[0].j1 - p.hsrc[0].j0) +
                                                               (j - p.hsrc[0].j0)];  // stepper_tm.py:190-191
        sHx[c] = hxv;
      }
      if ((!PML || i < nx - 1) && li < RX - 1) {
        T d = div_c(sEz[c + RY] - ez, dx, rdx);
        T curl = d;
        if (PML && p.hx.in(i)) {
          T ps = p.hx.b[i] * vPzx[PML ? u : 0] + p.hx.a[i] * d;
          vPzx[PML ? u : 0] = ps;
          curl = d + ps;
        }
        T hyv = sHy[c] + dt * curl;
        if (hy_tile && i >= p.hsrc[1].i0 && i < p.hsrc[1].i1 && j >= p.hsrc[1].j0 && j < p.hsrc[1].j1)
          hyv = hyv - hc * static_cast<const T*>(p.hsrc[1].w)[(long long)(i - p.hsrc[1].i0) *
                                                                   (p.hsrc[1].j1 - p.hsrc[1].j0) +
                                                               (j - p.hsrc[1].j0)];  // stepper_tm.py:188-189
        sHy[c] = hyv;
      }
    }
    __syncthreads();
    // D half (stepper_tm.py:130-166,197-200,226-227)
    li = tid / RY;
    lj = tid % RY;
#pragma unroll
    for (int u = 0; u < PER; ++u, li += BLK_TM / RY, lj += BLK_TM % RY) {
      if (lj >= RY) {
        lj -= RY;
        ++li;
      }
      int c = li * RY + lj;
      if (c >= RC) break;
      int i = a0 + li, j = b0 + lj;
      if (PML && (i < 0 || i >= nx || j < 0 || j >= ny)) continue;
      T dz = vDz[u];
      bool live = !PML || (i >= 1 && i < nx - 1 && (p.y_pmc || (j >= 1 && j < ny - 1)));
      if (live && li >= 1 && lj >= 1) {
        T hx1 = (!PML || j < ny - 1) ? sHx[c] : T(0);
        T hx0 = (!PML || j >= 1) ? sHx[c - 1] : T(0);
        T dHx = div_c(hx1 - hx0, dx, rdx);
        T dHy = div_c(sHy[c] - sHy[c - RY], dx, rdx);
        int ii = i - 1, jj = j - joff;
        T a1 = dHy;
        if (PML && p.ex.in(ii)) {
          T ps = p.ex.b[ii] * vPyx[PML ? u : 0] + p.ex.a[ii] * dHy;
          vPyx[PML ? u : 0] = ps;
          a1 = dHy + ps;
        }
        T curl = a1 - dHx;
        if (PML && p.ey.in(jj)) {
          T ps = p.ey.b[jj] * vPxy[PML ? u : 0] + p.ey.a[jj] * dHx;
          vPxy[PML ? u : 0] = ps;
          curl = curl - ps;
        }
        dz = dz + dt * curl;
      }
      if (src_tile && i >= p.src.i0 && i < p.src.i1 && j >= p.src.j0 && j < p.src.j1)
        dz = dz - coef * static_cast<const T*>(p.src.w)[(long long)s * p.src.wstep + (long long)(i - p.src.i0) * (p.src.j1 - p.src.j0) +
                                                        (j - p.src.j0)];
      vDz[u] = dz;
      sDz[c] = dz;
      sEz[c] = dz / vK[u];
    }
    __syncthreads();
    if (p.tape) {
      for (int c = tid; c < TX * TY; c += BLK_TM) {
        int i = I0 + c / TY, j = J0 + c % TY;
        if (i < nx && j < ny)
          p.tape[(long long)(tape_s0 + s) * nx * ny + (long long)i * ny + j] = sEz[(i - a0) * RY + (j - b0)];
      }
    }
    if (tap_tile) {
      write_taps<T, RX, RY>(taps, s, I0, J0, TX, TY, a0, b0, [&](int f, int li, int lj) -> T {
        int c = li * RY + lj;
        return f == 0 ? sEz[c] : f == 1 ? sHx[c] : f == 2 ? sHy[c] : sDz[c];
      }, slot_base);
      __syncthreads();
    }
  }
  for (int c = tid; c < TX * TY; c += BLK_TM) {
    int i = I0 + c / TY, j = J0 + c % TY;
    if (i >= nx || j >= ny) continue;
    int lc = (i - a0) * RY + (j - b0);
    long long e = (long long)i * ny + j;
    p.oEz[e] = sEz[lc];
    p.oDz[e] = sDz[lc];
    if (j < ny - 1) p.oHx[(long long)i * (ny - 1) + j] = sHx[lc];
    if (i < nx - 1) p.oHy[e] = sHy[lc];
  }
  if constexpr (PML) {  // psi of the owned cells, from the registers of the threads that updated them
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      int c = tid + u * BLK_TM;
      int i = a0 + c / RY, j = b0 + c % RY;
      if (c >= RC || i < I0 || i >= I0 + TX || j < J0 || j >= J0 + TY || i >= nx || j >= ny) continue;
      if (j < ny - 1 && p.hy.in(j)) p.oEzy[(long long)i * (ny - 1) + j] = vPzy[u];
      if (i < nx - 1 && p.hx.in(i)) p.oEzx[(long long)i * ny + j] = vPzx[u];
      int jj = j - joff;
      if (i >= 1 && i < nx - 1 && jj >= 0 && jj < ny_e) {
        long long q = (long long)(i - 1) * ny_e + jj;
        if (p.ex.in(i - 1)) p.oHyx[q] = vPyx[u];
        if (p.ey.in(jj)) p.oHxy[q] = vPxy[u];
      }
    }
  }
}

template <typename T, int TX, int TY, int NS, bool PML>
__global__ __launch_bounds__(BLK_TM) void tm_tiled(TmTiled<T> p, TapSet taps, int nsub, int tape_s0) {
  int tr, tc;
  tile_of(PML, p.map, &tr, &tc);
  tm_tile<T, TX, TY, NS, PML>(p, taps, nsub, tape_s0, tr, tc, 0, 0);
}

// FP32 interior tiles held to 64 VGPRs (eight waves per SIMD): two
// 1024-thread workgroups fit a CU (their LDS allows it), so one hides the
// other's barrier and load latency.  Ring tiles (CPML psi) and FP64 keep the
// plain bound.
template <typename T, int TX, int TY, int NS>
__global__ __launch_bounds__(BLK_TM) __attribute__((amdgpu_waves_per_eu(8, 8))) void tm_tiled_in2(
    TmTiled<T> p, TapSet taps, int nsub, int tape_s0) {
  int tr, tc;
  tile_of(false, p.map, &tr, &tc);
  tm_tile<T, TX, TY, NS, false>(p, taps, nsub, tape_s0, tr, tc, 0, 0);
}

// Persistent batched form: every workgroup walks the flattened (variant,
// tile) work list of all variants.  Per-variant problems (both ping-pong
// orientations), taps and tile maps live in device memory, so the launch
// depends only on scalars and can be captured in a graph and replayed.
template <typename T, int TX, int TY, int NS, bool PML>
__device__ __forceinline__ void tm_batch_body(const TmBatchVar<T>* vars, const int* prefix, int nvar, int nsub,
                                              int parity, int t0, int slot_base, int tape_s0) {
  const int total = prefix[nvar];
  // XCD-aware order: workgroups are dealt to the eight XCDs round robin, so
  // the blocks of XCD x (blockIdx % 8 == x) share the contiguous item range
  // [x * chunk, (x + 1) * chunk): neighbouring tiles, whose halos overlap,
  // run on the same XCD (same L2) at about the same time
  const int x = blockIdx.x & 7, nx8 = ((int)gridDim.x - x + 7) >> 3;
  const int chunk = (total + 7) >> 3, end = min(total, (x + 1) * chunk);
  for (int item = x * chunk + (int)(blockIdx.x >> 3); item < end; item += nx8) {
    int lo = 0, hi = nvar;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (prefix[mid] <= item) lo = mid;
      else hi = mid;
    }
    const TmBatchVar<T>& V = vars[lo];
    const TmTiled<T>& p = parity ? V.pB : V.pA;
    int tr, tc;
    tile_linear(PML, p.map, item - prefix[lo], &tr, &tc);
    tm_tile<T, TX, TY, NS, PML>(p, V.taps, nsub, tape_s0, tr, tc, slot_base, t0);
    __syncthreads();
  }
}

template <typename T, int TX, int TY, int NS, bool PML>
__global__ __launch_bounds__(BLK_TM) void tm_batch(const TmBatchVar<T>* vars, const int* prefix, int nvar,
                                                int nsub, int parity, int t0, int slot_base, int tape_s0) {
  tm_batch_body<T, TX, TY, NS, PML>(vars, prefix, nvar, nsub, parity, t0, slot_base, tape_s0);
}

// the interior form held to 64 VGPRs, as tm_tiled_in2
template <typename T, int TX, int TY, int NS>
__global__ __launch_bounds__(BLK_TM) __attribute__((amdgpu_waves_per_eu(8, 8))) void tm_batch_in2(
    const TmBatchVar<T>* vars, const int* prefix, int nvar, int nsub, int parity, int t0, int slot_base,
    int tape_s0) {
  tm_batch_body<T, TX, TY, NS, false>(vars, prefix, nvar, nsub, parity, t0, slot_base, tape_s0);
}

}  // namespace

// ------------------------------------------------------------- launchers
// Interior tiles: the rectangle of tiles whose whole LDS region avoids every
// CPML shell; PML tiles: the ring around it.
template <int TX, int TY, int NS>
TileMap make_map(int nx, int ny, int in_i0, int in_i1, int in_j0, int in_j1) {
  TileMap m;
  m.R = (nx + TX - 1) / TX;
  m.C = (ny + TY - 1) / TY;
  auto lo = [](int edge, int t, int ns) { return (edge + ns + 1 + t - 1) / t; };
  auto hi = [](int edge, int t, int ns) { int v = edge - t - ns; return v < 0 ? 0 : v / t + 1; };
  m.r0 = std::min(m.R, lo(in_i0, TX, NS));
  m.r1 = std::max(m.r0, std::min(m.R, hi(in_i1, TX, NS)));
  m.c0 = std::min(m.C, lo(in_j0, TY, NS));
  m.c1 = std::max(m.c0, std::min(m.C, hi(in_j1, TY, NS)));
  if (m.r1 == m.r0 || m.c1 == m.c0) m.r0 = m.r1 = m.c0 = m.c1 = 0;
  return m;
}

// The interior launch is the one the kernel timing hook measures; its
// cell-steps are the interior tiles' cells times the substeps.
template <typename P, typename K1, typename K2>
int launch_pair(const TileMap& m, K1 kin, K2 kpml, hipStream_t s, long long tile_cells, int nsub) {
  int n_in = (m.r1 - m.r0) * (m.c1 - m.c0);
  int n_ring = m.R * m.C - n_in;
  if (n_in) {
    timing_begin(s);
    kin(dim3(m.c1 - m.c0, m.r1 - m.r0));
    timing_end(s, (long long)n_in * tile_cells * nsub);
  }
  if (n_ring) kpml(dim3(n_ring));
  return MT_OK;
}

template <typename T, int TX, int TY, int NS>
int launch_te_tiled(TeTiled<T> p, const TapSet& taps, int nsub, int write_e, hipStream_t s) {
  p.map = make_map<TX, TY, NS>(p.nx, p.ny, p.in_i0, p.in_i1, p.in_j0, p.in_j1);
  launch_pair<T>(
      p.map,
      [&](dim3 g) { hipLaunchKernelGGL((te_tiled<T, TX, TY, NS, false>), g, dim3(te_block(false)), 0, s, p, taps, nsub, write_e); },
      [&](dim3 g) { hipLaunchKernelGGL((te_tiled<T, TX, TY, NS, true>), g, dim3(te_block(true)), 0, s, p, taps, nsub, write_e); },
      s, (long long)TX * TY, nsub);
  return check_launch("te2d tiled step");
}

template <typename T, int TX, int TY, int NS>
int launch_tm_tiled(TmTiled<T> p, const TapSet& taps, int nsub, int tape_s0, hipStream_t s) {
  p.map = make_map<TX, TY, NS>(p.nx, p.ny, p.in_i0, p.in_i1, p.in_j0, p.in_j1);
  launch_pair<T>(
      p.map,
      [&](dim3 g) {
        if constexpr (sizeof(T) == 4)
          hipLaunchKernelGGL((tm_tiled_in2<T, TX, TY, NS>), g, dim3(BLK_TM), 0, s, p, taps, nsub, tape_s0);
        else
          hipLaunchKernelGGL((tm_tiled<T, TX, TY, NS, false>), g, dim3(BLK_TM), 0, s, p, taps, nsub, tape_s0);
      },
      [&](dim3 g) { hipLaunchKernelGGL((tm_tiled<T, TX, TY, NS, true>), g, dim3(BLK_TM), 0, s, p, taps, nsub, tape_s0); },
      s, (long long)TX * TY, nsub);
  return check_launch("tm2d tiled step");
}

template <typename T, int TX, int TY, int NS>
int launch_tm_batch(const TmBatchVar<T>* vars, const int* pre_in, const int* pre_ring, int nvar, int n_in,
                    int n_ring, int grid_in, int grid_ring, int nsub, int parity, int t0, int slot_base,
                    int tape_s0, hipStream_t s) {
  if (n_in) {
    timing_begin(s);
    if constexpr (sizeof(T) == 4)
      hipLaunchKernelGGL((tm_batch_in2<T, TX, TY, NS>), dim3(std::min(n_in, grid_in)), dim3(BLK_TM), 0, s, vars,
                         pre_in, nvar, nsub, parity, t0, slot_base, tape_s0);
    else
      hipLaunchKernelGGL((tm_batch<T, TX, TY, NS, false>), dim3(std::min(n_in, grid_in)), dim3(BLK_TM), 0, s, vars,
                       pre_in, nvar, nsub, parity, t0, slot_base, tape_s0);
    timing_end(s, (long long)n_in * TX * TY * nsub);
  }
  if (n_ring)
    hipLaunchKernelGGL((tm_batch<T, TX, TY, NS, true>), dim3(std::min(n_ring, grid_ring)), dim3(BLK_TM), 0, s, vars,
                       pre_ring, nvar, nsub, parity, t0, slot_base, tape_s0);
  return check_launch("tm2d batched step");
}

template <int TX, int TY, int NS>
TileMap map_for(int nx, int ny, int in_i0, int in_i1, int in_j0, int in_j1) {
  return make_map<TX, TY, NS>(nx, ny, in_i0, in_i1, in_j0, in_j1);
}

// Tile-shape variants the autotuner chooses from (index = variant id).
// (TX, TY, NS): owned rows, owned columns, Yee steps per launch.  FP64 tiles
// are smaller so LDS (fields + psi) fits one workgroup.
#define MT_VARIANTS_F32(X, T) \
  X(T, 16, 64, 1) X(T, 16, 64, 2) X(T, 16, 64, 4) X(T, 32, 64, 1) X(T, 32, 64, 2) X(T, 32, 64, 4) X(T, 8, 128, 2)
#define MT_VARIANTS_F64(X, T) \
  X(T, 8, 64, 1) X(T, 8, 64, 2) X(T, 8, 64, 4) X(T, 16, 64, 1) X(T, 16, 64, 2) X(T, 16, 32, 2) X(T, 8, 128, 1)
static const int kVarF32[7][3] = {{16, 64, 1}, {16, 64, 2}, {16, 64, 4}, {32, 64, 1},
                                  {32, 64, 2}, {32, 64, 4}, {8, 128, 2}};
static const int kVarF64[7][3] = {{8, 64, 1}, {8, 64, 2}, {8, 64, 4}, {16, 64, 1},
                                  {16, 64, 2}, {16, 32, 2}, {8, 128, 1}};

int te_tiled_dispatch_f32(int variant, const TeTiled<float>& p, const TapSet& taps, int nsub, int write_e,
                          hipStream_t s) {
  int v = 0;
#define X(T, A, B, C) \
  if (variant == v++) return launch_te_tiled<T, A, B, C>(p, taps, nsub, write_e, s);
  MT_VARIANTS_F32(X, float)
#undef X
  return fail(MT_EINVAL, "unknown tiled variant");
}

int te_tiled_dispatch_f64(int variant, const TeTiled<double>& p, const TapSet& taps, int nsub, int write_e,
                          hipStream_t s) {
  int v = 0;
#define X(T, A, B, C) \
  if (variant == v++) return launch_te_tiled<T, A, B, C>(p, taps, nsub, write_e, s);
  MT_VARIANTS_F64(X, double)
#undef X
  return fail(MT_EINVAL, "unknown tiled variant");
}

int tm_tiled_dispatch_f32(int variant, const TmTiled<float>& p, const TapSet& taps, int nsub, int tape_s0,
                          hipStream_t s) {
  int v = 0;
#define X(T, A, B, C) \
  if (variant == v++) return launch_tm_tiled<T, A, B, C>(p, taps, nsub, tape_s0, s);
  MT_VARIANTS_F32(X, float)
#undef X
  return fail(MT_EINVAL, "unknown tiled variant");
}

int tm_tiled_dispatch_f64(int variant, const TmTiled<double>& p, const TapSet& taps, int nsub, int tape_s0,
                          hipStream_t s) {
  int v = 0;
#define X(T, A, B, C) \
  if (variant == v++) return launch_tm_tiled<T, A, B, C>(p, taps, nsub, tape_s0, s);
  MT_VARIANTS_F64(X, double)
#undef X
  return fail(MT_EINVAL, "unknown tiled variant");
}

int tm_batch_dispatch_f32(int variant, const TmBatchVar<float>* vars, const int* pre_in, const int* pre_ring,
                          int nvar, int n_in, int n_ring, int grid_in, int grid_ring, int nsub, int parity,
                          int t0, int slot_base, int tape_s0, hipStream_t s) {
  int v = 0;
#define X(T, A, B, C)                                                                                  \
  if (variant == v++)                                                                                  \
    return launch_tm_batch<T, A, B, C>(vars, pre_in, pre_ring, nvar, n_in, n_ring, grid_in, grid_ring, \
                                       nsub, parity, t0, slot_base, tape_s0, s);
  MT_VARIANTS_F32(X, float)
#undef X
  return fail(MT_EINVAL, "unknown tiled variant");
}

int tm_batch_dispatch_f64(int variant, const TmBatchVar<double>* vars, const int* pre_in, const int* pre_ring,
                          int nvar, int n_in, int n_ring, int grid_in, int grid_ring, int nsub, int parity,
                          int t0, int slot_base, int tape_s0, hipStream_t s) {
  int v = 0;
#define X(T, A, B, C)                                                                                  \
  if (variant == v++)                                                                                  \
    return launch_tm_batch<T, A, B, C>(vars, pre_in, pre_ring, nvar, n_in, n_ring, grid_in, grid_ring, \
                                       nsub, parity, t0, slot_base, tape_s0, s);
  MT_VARIANTS_F64(X, double)
#undef X
  return fail(MT_EINVAL, "unknown tiled variant");
}

TileMap tiled_map(int variant, int dtype, int nx, int ny, int in_i0, int in_i1, int in_j0, int in_j1) {
  int v = 0;
#define X(T, A, B, C) \
  if (variant == v++) return map_for<A, B, C>(nx, ny, in_i0, in_i1, in_j0, in_j1);
  if (dtype == MT_F64) {
    MT_VARIANTS_F64(X, double)
  } else {
    MT_VARIANTS_F32(X, float)
  }
#undef X
  return TileMap{};
}

int tiled_variant_count() { return 7; }

int tiled_variant_steps(int variant, int dtype) {
  if (variant < 0 || variant >= 7) return 0;
  return dtype == MT_F64 ? kVarF64[variant][2] : kVarF32[variant][2];
}

}  // namespace mt

extern "C" int mt_tiled_variants(int dtype, int* steps, int* tx, int* ty) {
  const int(*tab)[3] = dtype == MT_F64 ? mt::kVarF64 : mt::kVarF32;
  for (int v = 0; v < 7; ++v) {
    if (steps) steps[v] = tab[v][2];
    if (tx) tx[v] = tab[v][0];
    if (ty) ty[v] = tab[v][1];
  }
  return 7;
}
