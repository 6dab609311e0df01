// Batched 2D TM design variants of uneven size: one persistent launch pair
// (interior tiles, CPML-ring tiles) advances every variant a block of steps;
// the batched adjoint runs each reverse pass over the cells of all variants
// in one launch.  Per-variant descriptors live in caller-provided device
// memory, so the forward launches depend only on scalars (graph-capturable).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "adjoint2d.h"
#include "common.h"
#include "tiled2d.h"

namespace {
constexpr int kMaxMonPerVar = 8;
constexpr int kCellBlk = 256;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct Layout {
  size_t vars, pre_in, pre_ring, adjk, adjk2, blk_pre, tile_pre, inj, inj2, inj_pre, tapes, dts, total;
};

template <typename T>
Layout layout(int nvar) {
  Layout L{};
  size_t o = 0;
  L.vars = o;
  o = align256(o + nvar * sizeof(mt::TmBatchVar<T>));
  L.pre_in = o;
  o = align256(o + (nvar + 1) * sizeof(int));
  L.pre_ring = o;
  o = align256(o + (nvar + 1) * sizeof(int));
  L.adjk = o;
  o = align256(o + nvar * sizeof(mt::AdjK<T>));
  L.adjk2 = o;
  o = align256(o + nvar * sizeof(mt::AdjK<T>));
  L.blk_pre = o;
  o = align256(o + (nvar + 1) * sizeof(int));
  L.tile_pre = o;
  o = align256(o + (nvar + 1) * sizeof(int));
  L.inj = o;
  o = align256(o + (size_t)kMaxMonPerVar * nvar * sizeof(mt_monitor));
  L.inj2 = o;
  o = align256(o + (size_t)kMaxMonPerVar * nvar * sizeof(mt_monitor));
  L.inj_pre = o;
  o = align256(o + (size_t)kMaxMonPerVar * (nvar + 1) * sizeof(int));
  L.tapes = o;
  o = align256(o + nvar * sizeof(void*));
  L.dts = o;
  o = align256(o + nvar * sizeof(double));
  L.total = o;
  return L;
}
}  // namespace

struct mt_batch {
  int nvar, dtype, variant, S, NS;
  int n_in, n_ring, grid_in, grid_ring, cell_blocks, max_mons;
  char* ws;
  Layout L;
  std::vector<mt_tm2d> probs;
  std::vector<std::vector<mt_monitor>> mons;
  std::vector<void*> stage;
  std::vector<long long> stage_bytes;
  // adjoint descriptors uploaded by mt_tm2d_batch_adjoint_prepare
  bool adj_ready = false;
  int adj_max_m = 0;
  std::vector<int> adj_inj_blocks;
  // second cotangent set of the fused adjoint step (mt_tm2d_batch_adjoint_pingpong)
  bool adj_fused = false;
  int fused_tiles = 0;
  long long total_cells = 0;
  std::vector<mt_tm2d_adj> adj0, adj1;
};

namespace mt {
namespace {

template <typename T, int PH>
__global__ __launch_bounds__(kCellBlk) void adj_batch_cells(const AdjK<T>* ks, void* const* tapes,
                                                            const int* blk_pre, int nvar, int s_rel) {
  int b = blockIdx.x;
  int lo = 0, hi = nvar;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (blk_pre[mid] <= b) lo = mid;
    else hi = mid;
  }
  const AdjK<T>& p = ks[lo];
  long long cells = (long long)p.nx * p.ny;
  long long cell = (long long)(b - blk_pre[lo]) * kCellBlk + threadIdx.x;
  if (cell >= cells) return;
  int i = (int)(cell / p.ny), j = (int)(cell % p.ny);
  if (PH == 0) adj_a_cell(p, static_cast<const T*>(tapes[lo]) + (long long)s_rel * cells, i, j);
  else if (PH == 1) adj_b_cell(p, i, j);
  else adj_c_cell(p, i, j);
}

template <typename T>
__global__ __launch_bounds__(kCellBlk) void adj_batch_inject(const mt_monitor* inj, const int* pre, int nvar,
                                                             int t, const double* dts, int t_rel) {
  int b = blockIdx.x;
  int lo = 0, hi = nvar;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid;
  }
  const mt_monitor& m = inj[lo];
  if (m.n_terms == 0) return;
  adj_inject_block<T>(m, (b - pre[lo]) * kInjE, t, dts[lo], t_rel);
}

// Fused reverse step (no conductivity, no poles): one pass per tile of
// FTI x FTJ cells of one variant.  A runs on the tile and a one-cell halo
// (gHy on the rows above and below, gHx on the columns left and right),
// B's Hy part on the tile and the row above, its Hx part on the tile and
// the column to the left, C on the tile; gHy, gHx, gEx, gEy stay in LDS.
// Every cotangent is read from set r (step t+1) and the tile's own cells
// are written to set w (step t), so the halo reads never race with a
// neighbour tile's writes; eps_grad (own cells only) is updated in place.
// Same arithmetic as adj_a_cell / adj_b_cell / adj_c_cell.
constexpr int FTI = 8, FTJ = 64;

template <typename T>
__global__ __launch_bounds__(FTI * FTJ) void adj_batch_fused(const AdjK<T>* rs, const AdjK<T>* wsd,
                                                             void* const* tapes, const int* tile_pre, int nvar,
                                                             int s_rel, int per_xcd) {
  __shared__ T sGHy[FTI + 2][FTJ];   // rows i0-1 .. i0+FTI
  __shared__ T sGHx[FTI][FTJ + 2];   // columns j0-1 .. j0+FTJ
  __shared__ T sGEx[FTI + 1][FTJ];   // rows i0-1 .. i0+FTI-1
  __shared__ T sGEy[FTI][FTJ + 1];   // columns j0-1 .. j0+FTJ-1
  // XCD-aware order (as in tm_batch): XCD x takes the contiguous tile range
  // [x * per_xcd, (x + 1) * per_xcd), so the tiles whose halos overlap share an L2
  const int b = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
  if (b >= tile_pre[nvar]) return;
  int lo = 0, hi = nvar;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (tile_pre[mid] <= b) lo = mid;
    else hi = mid;
  }
  const AdjK<T>& r = rs[lo];
  const AdjK<T>& w = wsd[lo];
  const int nx = r.nx, ny = r.ny;
  const int ntj = (ny + FTJ - 1) / FTJ, t = b - tile_pre[lo];
  const int i0 = (t / ntj) * FTI, j0 = (t % ntj) * FTJ;
  const int tx = threadIdx.x, ty = threadIdx.y, q = ty * FTJ + tx;
  const int i = i0 + ty, j = j0 + tx;
  const bool own = i < nx && j < ny;
  const T* ez_after = static_cast<const T*>(tapes[lo]) + (long long)s_rel * nx * ny;
  // Every load of the A and B phases that does not depend on the A results
  // is issued before the first store and barrier: the B inputs of the own
  // cell and of this thread's B halo cell, then the A halo cell (which
  // stores nothing), then the own A cell.
  const bool hrow = q < FTJ, hcol = !hrow && q < FTJ + FTI;
  const int rr = q - FTJ;
  const BIn<T> pre = adj_b_load(r, i, j, own && i < nx - 1, own && j < ny - 1);
  const BIn<T> preh = adj_b_load(r, hrow ? i0 - 1 : i0 + rr, hrow ? j0 + q : j0 - 1,
                                 hrow && i0 >= 1 && j0 + q < ny, hcol && j0 >= 1 && i0 + rr < nx);
  // ---- A
  if (q < 2 * FTJ) {                      // rows i0-1 and i0+FTI: gHy
    const int hi_row = q >= FTJ, jj = j0 + (q % FTJ), ii = hi_row ? i0 + FTI : i0 - 1;
    T gy = T(0), gx;
    if (ii >= 0 && ii < nx && jj < ny) adj_a_fused(r, w, ez_after, ii, jj, false, gy, gx);
    sGHy[hi_row ? FTI + 1 : 0][q % FTJ] = gy;
  } else if (q < 2 * FTJ + 2 * FTI) {     // columns j0-1 and j0+FTJ: gHx
    const int q2 = q - 2 * FTJ, hi_col = q2 >= FTI, ii = i0 + (q2 % FTI), jj = hi_col ? j0 + FTJ : j0 - 1;
    T gy, gx = T(0);
    if (jj >= 0 && jj < ny && ii < nx) adj_a_fused(r, w, ez_after, ii, jj, false, gy, gx);
    sGHx[q2 % FTI][hi_col ? FTJ + 1 : 0] = gx;
  }
  T ghy = T(0), ghx = T(0);
  if (own) adj_a_fused(r, w, ez_after, i, j, true, ghy, ghx);
  sGHy[ty + 1][tx] = ghy;
  sGHx[ty][tx + 1] = ghx;
  __syncthreads();
  // ---- B
  T gex = T(0), gey = T(0);
  if (own && i < nx - 1) gex = adj_b_hy_pre(r, w, i, j, pre, sGHy[ty + 1][tx], sGHy[ty + 2][tx]);
  if (own && j < ny - 1) gey = adj_b_hx_pre(r, w, i, j, pre, sGHx[ty][tx + 1], sGHx[ty][tx + 2]);
  sGEx[ty + 1][tx] = gex;
  sGEy[ty][tx + 1] = gey;
  if (hrow) {                             // row i0-1: gEx
    T g = T(0);
    if (i0 >= 1 && j0 + q < ny) g = adj_b_hy_pre<T, false>(r, w, i0 - 1, j0 + q, preh, sGHy[0][q], sGHy[1][q]);
    sGEx[0][q] = g;
  } else if (hcol) {                      // column j0-1: gEy
    T g = T(0);
    if (j0 >= 1 && i0 + rr < nx) g = adj_b_hx_pre<T, false>(r, w, i0 + rr, j0 - 1, preh, sGHx[rr][0], sGHx[rr][1]);
    sGEy[rr][0] = g;
  }
  __syncthreads();
  // ---- C
  if (own) {
    const T gxm = i >= 1 ? sGEx[ty][tx] : T(0);
    const T gx0 = i < nx - 1 ? sGEx[ty + 1][tx] : T(0);
    const T gym = j >= 1 ? sGEy[ty][tx] : T(0);
    const T gy0 = j < ny - 1 ? sGEy[ty][tx + 1] : T(0);
    w.aEz[(long long)i * ny + j] = (gxm - gx0) / r.dx + (gym - gy0) / r.dx;
  }
}

template <typename T>
TmTiled<T> tiled_of(const mt_tm2d* d, const mt_source2d* src, const void* coef_tab, const void* hcoef_tab,
                    int variant, bool swap) {
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
  const int joff = d->y_pmc ? 0 : 1;
  p.in_i0 = std::max(d->hx.lo_end, d->ex.lo_end + 1) + 1;
  p.in_i1 = std::min(d->hx.hi_begin, d->ex.hi_begin + 1) - 1;
  p.in_j0 = std::max(d->hy.lo_end, d->ey.lo_end + joff) + 1;
  p.in_j1 = std::min(d->hy.hi_begin, d->ey.hi_begin + joff) - 1;
  p.map = tiled_map(variant, d->dtype, d->nx, d->ny, p.in_i0, p.in_i1, p.in_j0, p.in_j1);
  T* A[8] = {static_cast<T*>(d->Ez), static_cast<T*>(d->Dz), static_cast<T*>(d->Hx), static_cast<T*>(d->Hy),
             static_cast<T*>(d->psi_Ezy), static_cast<T*>(d->psi_Ezx), static_cast<T*>(d->psi_Hyx),
             static_cast<T*>(d->psi_Hxy)};
  T* B[8];
  for (int q = 0; q < 8; ++q) B[q] = static_cast<T*>(d->alt[q]);
  T** in = swap ? B : A;
  T** out = swap ? A : B;
  p.Ez = in[0], p.Dz = in[1], p.Hx = in[2], p.Hy = in[3];
  p.pEzy = in[4], p.pEzx = in[5], p.pHyx = in[6], p.pHxy = in[7];
  p.oEz = out[0], p.oDz = out[1], p.oHx = out[2], p.oHy = out[3];
  p.oEzy = out[4], p.oEzx = out[5], p.oHyx = out[6], p.oHxy = out[7];
  p.tape = static_cast<T*>(d->tape);
  if (src && src->present) p.src = *src;
  p.coef_tab = static_cast<const T*>(coef_tab);
  p.hsrc[0] = d->hsrc[0];
  p.hsrc[1] = d->hsrc[1];
  p.hcoef_tab = static_cast<const T*>(hcoef_tab);
  return p;
}

template <typename T>
int create_t(mt_batch* b, const mt_tm2d_batch_item* items, hipStream_t s) {
  const int nvar = b->nvar;
  std::vector<TmBatchVar<T>> vars(nvar);
  std::vector<int> pre_in(nvar + 1, 0), pre_ring(nvar + 1, 0), blk(nvar + 1, 0);
  std::vector<double> dts(nvar);
  for (int v = 0; v < nvar; ++v) {
    const mt_tm2d_batch_item& it = items[v];
    vars[v].pA = tiled_of<T>(&it.prob, &it.src, it.coef_tab, it.hcoef_tab, b->variant, false);
    vars[v].pB = tiled_of<T>(&it.prob, &it.src, it.coef_tab, it.hcoef_tab, b->variant, true);
    MonitorPlan mp{it.mons, it.n_mons, it.stage, it.stage_bytes, b->S, b->dtype};
    vars[v].taps.n = 0;
    if (it.n_mons && !build_taps(mp, 0, 0, sizeof(T), &vars[v].taps))
      return fail(MT_EINVAL, "batched run: every monitor term needs its semantic (tile) form");
    const TileMap& m = vars[v].pA.map;
    int n_in = (m.r1 - m.r0) * (m.c1 - m.c0);
    pre_in[v + 1] = pre_in[v] + n_in;
    pre_ring[v + 1] = pre_ring[v] + (m.R * m.C - n_in);
    blk[v + 1] = blk[v] + (int)(((long long)it.prob.nx * it.prob.ny + kCellBlk - 1) / kCellBlk);
    dts[v] = it.prob.dt;
  }
  b->n_in = pre_in[nvar];
  b->n_ring = pre_ring[nvar];
  b->cell_blocks = blk[nvar];
  char* ws = b->ws;
  if (hipMemcpyAsync(ws + b->L.vars, vars.data(), nvar * sizeof(TmBatchVar<T>), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.pre_in, pre_in.data(), (nvar + 1) * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.pre_ring, pre_ring.data(), (nvar + 1) * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.blk_pre, blk.data(), (nvar + 1) * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.dts, dts.data(), nvar * sizeof(double), hipMemcpyHostToDevice, s) ||
      hipStreamSynchronize(s))
    return fail(MT_EHIP, "batched run: descriptor upload failed");
  return MT_OK;
}

template <typename T>
int run_t(mt_batch* b, int step0, int k, hipStream_t s) {
  const int S = b->max_mons ? b->S : k;
  int done = 0, slot = 0, first = step0, cur = 0, rc;
  const TmBatchVar<T>* vars = reinterpret_cast<const TmBatchVar<T>*>(b->ws + b->L.vars);
  const int* pin = reinterpret_cast<const int*>(b->ws + b->L.pre_in);
  const int* pring = reinterpret_cast<const int*>(b->ws + b->L.pre_ring);
  while (done < k) {
    int nsub = std::min(b->NS, k - done);
    if (b->max_mons) nsub = std::min(nsub, S - slot);
    rc = (b->dtype == MT_F64)
             ? tm_batch_dispatch_f64(b->variant, reinterpret_cast<const TmBatchVar<double>*>(vars), pin, pring,
                                     b->nvar, b->n_in, b->n_ring, b->grid_in, b->grid_ring, nsub, cur,
                                     step0 + done, slot, done, s)
             : tm_batch_dispatch_f32(b->variant, reinterpret_cast<const TmBatchVar<float>*>(vars), pin, pring,
                                     b->nvar, b->n_in, b->n_ring, b->grid_in, b->grid_ring, nsub, cur,
                                     step0 + done, slot, done, s);
    if (rc) return rc;
    cur ^= 1;
    done += nsub;
    slot += nsub;
    if (b->max_mons && (slot == S || done == k)) {
      for (int v = 0; v < b->nvar; ++v) {
        if (b->mons[v].empty()) continue;
        MonitorPlan mp{b->mons[v].data(), (int)b->mons[v].size(), b->stage[v], b->stage_bytes[v], b->S, b->dtype};
        if ((rc = monitors_flush(mp, slot, first, b->probs[v].dt, s))) return rc;
      }
      first += slot;
      slot = 0;
    }
  }
  if (cur) {
    for (int v = 0; v < b->nvar; ++v) {
      const mt_tm2d& d = b->probs[v];
      size_t nx = d.nx, ny = d.ny, nye = d.y_pmc ? d.ny : d.ny - 2;
      size_t n[8] = {nx * ny, nx * ny, nx * (ny - 1), (nx - 1) * ny, nx * (ny - 1), (nx - 1) * ny,
                     (nx - 2) * nye, (nx - 2) * nye};
      void* A[8] = {d.Ez, d.Dz, d.Hx, d.Hy, d.psi_Ezy, d.psi_Ezx, d.psi_Hyx, d.psi_Hxy};
      for (int q = 0; q < 8; ++q)
        if (hipMemcpyAsync(A[q], d.alt[q], n[q] * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
          return fail(MT_EHIP, "batched run: copy back failed");
    }
  }
  return MT_OK;
}

// Upload the adjoint descriptors of every variant (cotangent fields, tapes,
// monitor injection terms) into the plan's device workspace; the sweeps that
// follow only pass scalars, so they can be captured in a graph.
template <typename T>
int adjoint_prepare_t(mt_batch* b, const mt_tm2d_adj* adj, void* const* tapes, const mt_monitor* mons,
                      const int* mon_off, hipStream_t s) {
  const int nvar = b->nvar;
  std::vector<AdjK<T>> ks(nvar);
  for (int v = 0; v < nvar; ++v) ks[v] = make_adjk<T>(&b->probs[v], &adj[v]);
  int max_m = 0;
  for (int v = 0; v < nvar; ++v) max_m = std::max(max_m, mon_off[v + 1] - mon_off[v]);
  if (max_m > kMaxMonPerVar) return fail(MT_EINVAL, "batched adjoint: too many monitors per variant");
  std::vector<mt_monitor> inj((size_t)kMaxMonPerVar * nvar);
  std::vector<int> ipre((size_t)kMaxMonPerVar * (nvar + 1), 0);
  for (int m = 0; m < max_m; ++m) {
    for (int v = 0; v < nvar; ++v) {
      mt_monitor& dst = inj[(size_t)m * nvar + v];
      dst = mt_monitor{};
      int blocks = 0;
      if (m < mon_off[v + 1] - mon_off[v]) {
        dst = mons[mon_off[v] + m];
        blocks = (dst.n0 * dst.n1 + kInjE - 1) / kInjE;
      }
      ipre[(size_t)m * (nvar + 1) + v + 1] = ipre[(size_t)m * (nvar + 1) + v] + blocks;
    }
  }
  char* ws = b->ws;
  if (hipMemcpyAsync(ws + b->L.adjk, ks.data(), nvar * sizeof(AdjK<T>), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.tapes, tapes, nvar * sizeof(void*), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.inj, inj.data(), inj.size() * sizeof(mt_monitor), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.inj_pre, ipre.data(), ipre.size() * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipStreamSynchronize(s))
    return fail(MT_EHIP, "batched adjoint: descriptor upload failed");
  b->adj_max_m = max_m;
  b->adj0.assign(adj, adj + nvar);
  b->adj_fused = false;
  b->adj_inj_blocks.assign(max_m, 0);
  for (int m = 0; m < max_m; ++m) b->adj_inj_blocks[m] = ipre[(size_t)m * (nvar + 1) + nvar];
  b->adj_ready = true;
  return MT_OK;
}

// Second cotangent set for the fused reverse step: descriptors of set 1
// (same scratch and eps_grad as set 0) and monitor injections into it, plus
// the tile prefix.  Applies when no variant has conductivity or poles;
// *fused says whether the sweeps will use it.
template <typename T>
int adjoint_pingpong_t(mt_batch* b, const mt_tm2d_adj* alt, const mt_monitor* mons, const int* mon_off, int* fused,
                       hipStream_t s) {
  if (!b->adj_ready) return fail(MT_EINVAL, "batched adjoint: descriptors not prepared");
  const int nvar = b->nvar;
  *fused = 0;
  b->adj_fused = false;
  for (int v = 0; v < nvar; ++v) {
    const mt_tm2d& d = b->probs[v];
    if (d.d1 || d.h1x || d.h1y || d.n_poles > 0) return MT_OK;
  }
  std::vector<AdjK<T>> ks(nvar);
  std::vector<int> tpre(nvar + 1, 0);
  for (int v = 0; v < nvar; ++v) {
    ks[v] = make_adjk<T>(&b->probs[v], &alt[v]);
    const mt_tm2d& d = b->probs[v];
    tpre[v + 1] = tpre[v] + ((d.nx + FTI - 1) / FTI) * ((d.ny + FTJ - 1) / FTJ);
  }
  std::vector<mt_monitor> inj((size_t)kMaxMonPerVar * nvar);
  for (int m = 0; m < b->adj_max_m; ++m)
    for (int v = 0; v < nvar; ++v) {
      mt_monitor& dst = inj[(size_t)m * nvar + v];
      dst = mt_monitor{};
      if (m < mon_off[v + 1] - mon_off[v]) dst = mons[mon_off[v] + m];
    }
  char* ws = b->ws;
  if (hipMemcpyAsync(ws + b->L.adjk2, ks.data(), nvar * sizeof(AdjK<T>), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.tile_pre, tpre.data(), (nvar + 1) * sizeof(int), hipMemcpyHostToDevice, s) ||
      hipMemcpyAsync(ws + b->L.inj2, inj.data(), inj.size() * sizeof(mt_monitor), hipMemcpyHostToDevice, s) ||
      hipStreamSynchronize(s))
    return fail(MT_EHIP, "batched adjoint: second-set upload failed");
  b->adj1.assign(alt, alt + nvar);
  b->fused_tiles = tpre[nvar];
  b->total_cells = 0;
  for (int v = 0; v < nvar; ++v) b->total_cells += (long long)b->probs[v].nx * b->probs[v].ny;
  b->adj_fused = true;
  *fused = 1;
  return MT_OK;
}

// Reverse-time sweep over steps [step0, step0 + k) with the prepared
// descriptors: per step, the monitor injection then either the fused pass
// (alternating the two cotangent sets, set 0 again at the end) or the three
// transpose passes over the cells of every variant.  Launches only.
template <typename T>
int adjoint_sweep_t(mt_batch* b, int step0, int k, hipStream_t s) {
  if (!b->adj_ready) return fail(MT_EINVAL, "batched adjoint: descriptors not prepared");
  if (b->adj_fused) {
    const int nvar = b->nvar;
    char* ws = b->ws;
    const AdjK<T>* k0 = reinterpret_cast<const AdjK<T>*>(ws + b->L.adjk);
    const AdjK<T>* k1 = reinterpret_cast<const AdjK<T>*>(ws + b->L.adjk2);
    void* const* dtp = reinterpret_cast<void* const*>(ws + b->L.tapes);
    const int* tpre = reinterpret_cast<const int*>(ws + b->L.tile_pre);
    const mt_monitor* inj0 = reinterpret_cast<const mt_monitor*>(ws + b->L.inj);
    const mt_monitor* inj1 = reinterpret_cast<const mt_monitor*>(ws + b->L.inj2);
    const int* dipre = reinterpret_cast<const int*>(ws + b->L.inj_pre);
    const double* ddts = reinterpret_cast<const double*>(ws + b->L.dts);
    int par = 0;
    for (int s_ = k - 1; s_ >= 0; --s_) {
      const int t = step0 + s_;
      for (int m = 0; m < b->adj_max_m; ++m) {
        const int nb = b->adj_inj_blocks[m];
        if (nb)
          hipLaunchKernelGGL(adj_batch_inject<T>, dim3(nb), dim3(kInjE * kInjS), 0, s,
                             (par ? inj1 : inj0) + (size_t)m * nvar, dipre + (size_t)m * (nvar + 1), nvar, t, ddts,
                             s_);
      }
      timing_begin(s, 2);
      const int per_xcd = (b->fused_tiles + 7) / 8;
      hipLaunchKernelGGL(adj_batch_fused<T>, dim3(8 * per_xcd), dim3(FTJ, FTI), 0, s, par ? k1 : k0,
                         par ? k0 : k1, dtp, tpre, nvar, s_, per_xcd);
      timing_end(s, b->total_cells, 2);
      par ^= 1;
      int rc = check_launch("batched tm2d fused adjoint step");
      if (rc) return rc;
    }
    if (par) {  // odd step count: the cotangents are in set 1
      for (int v = 0; v < nvar; ++v) {
        const mt_tm2d& d = b->probs[v];
        const mt_tm2d_adj &a0 = b->adj0[v], &a1 = b->adj1[v];
        size_t nx = d.nx, ny = d.ny, nye = d.y_pmc ? d.ny : d.ny - 2;
        size_t n[8] = {nx * ny, nx * ny, nx * (ny - 1), (nx - 1) * ny, nx * (ny - 1), (nx - 1) * ny,
                       (nx - 2) * nye, (nx - 2) * nye};
        void* dst[8] = {a0.aEz, a0.aDz, a0.aHx, a0.aHy, a0.apsi_Ezy, a0.apsi_Ezx, a0.apsi_Hyx, a0.apsi_Hxy};
        void* src[8] = {a1.aEz, a1.aDz, a1.aHx, a1.aHy, a1.apsi_Ezy, a1.apsi_Ezx, a1.apsi_Hyx, a1.apsi_Hxy};
        for (int q = 0; q < 8; ++q)
          if (hipMemcpyAsync(dst[q], src[q], n[q] * sizeof(T), hipMemcpyDeviceToDevice, s) != hipSuccess)
            return fail(MT_EHIP, "batched adjoint: copy back failed");
      }
    }
    return MT_OK;
  }
  const int nvar = b->nvar;
  char* ws = b->ws;
  const AdjK<T>* dks = reinterpret_cast<const AdjK<T>*>(ws + b->L.adjk);
  void* const* dtp = reinterpret_cast<void* const*>(ws + b->L.tapes);
  const int* dblk = reinterpret_cast<const int*>(ws + b->L.blk_pre);
  const mt_monitor* dinj = reinterpret_cast<const mt_monitor*>(ws + b->L.inj);
  const int* dipre = reinterpret_cast<const int*>(ws + b->L.inj_pre);
  const double* ddts = reinterpret_cast<const double*>(ws + b->L.dts);
  for (int s_ = k - 1; s_ >= 0; --s_) {
    int t = step0 + s_;
    for (int m = 0; m < b->adj_max_m; ++m) {
      int nb = b->adj_inj_blocks[m];
      if (nb)
        hipLaunchKernelGGL(adj_batch_inject<T>, dim3(nb), dim3(kInjE * kInjS), 0, s, dinj + (size_t)m * nvar,
                           dipre + (size_t)m * (nvar + 1), nvar, t, ddts, s_);
    }
    hipLaunchKernelGGL((adj_batch_cells<T, 0>), dim3(b->cell_blocks), dim3(kCellBlk), 0, s, dks, dtp, dblk, nvar, s_);
    hipLaunchKernelGGL((adj_batch_cells<T, 1>), dim3(b->cell_blocks), dim3(kCellBlk), 0, s, dks, dtp, dblk, nvar, s_);
    hipLaunchKernelGGL((adj_batch_cells<T, 2>), dim3(b->cell_blocks), dim3(kCellBlk), 0, s, dks, dtp, dblk, nvar, s_);
    int rc = check_launch("batched tm2d adjoint step");
    if (rc) return rc;
  }
  return MT_OK;
}

}  // namespace
}  // namespace mt

using namespace mt;

extern "C" long long mt_tm2d_batch_workspace(int nvar, int dtype) {
  if (nvar < 1) return 0;
  return dtype == MT_F64 ? (long long)layout<double>(nvar).total : (long long)layout<float>(nvar).total;
}

extern "C" mt_batch* mt_tm2d_batch_create(const mt_tm2d_batch_item* items, int nvar, int variant, int stage_steps,
                                          void* dev_ws, long long ws_bytes, void* stream) {
  if (!items || nvar < 1) {
    fail(MT_EINVAL, "batched run: no variants");
    return nullptr;
  }
  int dtype = items[0].prob.dtype;
  if (variant < 0 || variant >= tiled_variant_count()) {
    fail(MT_EINVAL, "batched run: invalid tile variant");
    return nullptr;
  }
  if (ws_bytes < mt_tm2d_batch_workspace(nvar, dtype) || !dev_ws) {
    fail(MT_EINVAL, "batched run: workspace too small");
    return nullptr;
  }
  int max_mons = 0;
  for (int v = 0; v < nvar; ++v) {
    const mt_tm2d& d = items[v].prob;
    if (d.dtype != dtype) {
      fail(MT_EINVAL, "batched run: all variants must share the dtype");
      return nullptr;
    }
    if (d.nx < 3 || d.ny < 3 || !d.eps || !d.Ez || !d.Dz || !d.Hx || !d.Hy || !d.psi_Ezy || !d.psi_Ezx ||
        !d.psi_Hyx || !d.psi_Hxy) {
      fail(MT_EINVAL, "batched run: incomplete problem");
      return nullptr;
    }
    for (int q = 0; q < 8; ++q)
      if (!d.alt[q]) {
        fail(MT_EINVAL, "batched run: missing ping-pong buffers");
        return nullptr;
      }
    if ((d.hsrc[0].present || d.hsrc[1].present) && !items[v].hcoef_tab) {
      fail(MT_EINVAL, "batched run: magnetic (TFSF) source without its coefficient table");
      return nullptr;
    }
    if (check_src(&d.hsrc[0], d.nx, d.ny - 1) || check_src(&d.hsrc[1], d.nx - 1, d.ny)) return nullptr;
    if (d.n_poles || d.kerr || d.d1 || d.h1x || d.n_mpoles || d.y_mirror || d.x_mirror || d.y_rot2 || d.rot4) {
      fail(MT_EINVAL, "batched run: media and symmetry cuts are not supported");
      return nullptr;
    }
    if (check_axis(d.hy, d.ny - 1, "hy") || check_axis(d.hx, d.nx - 1, "hx") || check_axis(d.ex, d.nx - 2, "ex") ||
        check_axis(d.ey, d.y_pmc ? d.ny : d.ny - 2, "ey"))
      return nullptr;
    const mt_source2d& sr = items[v].src;
    if (sr.present && (!items[v].coef_tab || sr.i0 < 0 || sr.j0 < 0 || sr.i1 > d.nx || sr.j1 > d.ny || !sr.w ||
                       sr.wstep != 0)) {
      fail(MT_EINVAL, "batched run: bad source");
      return nullptr;
    }
    for (int m = 0; m < items[v].n_mons; ++m)
      if (items[v].mons[m].kind != 0) {
        fail(MT_EINVAL, "batched run: probes are not supported (DFT monitors only)");
        return nullptr;
      }
    if (items[v].n_mons) {
      MonitorPlan mp{items[v].mons, items[v].n_mons, items[v].stage, items[v].stage_bytes, stage_steps, dtype};
      if (monitors_validate(mp)) return nullptr;
    }
    max_mons = std::max(max_mons, items[v].n_mons);
  }
  mt_batch* b = new mt_batch();
  b->nvar = nvar;
  b->dtype = dtype;
  b->variant = variant;
  b->S = stage_steps > 0 ? stage_steps : 1;
  b->NS = tiled_variant_steps(variant, dtype);
  b->max_mons = max_mons;
  b->ws = static_cast<char*>(dev_ws);
  b->L = dtype == MT_F64 ? layout<double>(nvar) : layout<float>(nvar);
  for (int v = 0; v < nvar; ++v) {
    b->probs.push_back(items[v].prob);
    b->mons.emplace_back(items[v].mons, items[v].mons + items[v].n_mons);
    b->stage.push_back(items[v].stage);
    b->stage_bytes.push_back(items[v].stage_bytes);
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  b->grid_in = 4 * cus;
  b->grid_ring = 4 * cus;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = dtype == MT_F64 ? create_t<double>(b, items, s) : create_t<float>(b, items, s);
  if (rc) {
    delete b;
    return nullptr;
  }
  return b;
}

extern "C" int mt_tm2d_batch_run(mt_batch* b, int step0, int k, void* stream) {
  if (!b) return fail(MT_EINVAL, "null batch");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->dtype == MT_F64 ? run_t<double>(b, step0, k, s) : run_t<float>(b, step0, k, s);
}

extern "C" int mt_tm2d_batch_adjoint(mt_batch* b, const mt_tm2d_adj* adj, void* const* tapes, const mt_monitor* mons,
                                     const int* mon_off, int step0, int k, void* stream) {
  if (!b || !adj || !tapes || !mon_off) return fail(MT_EINVAL, "batched adjoint: null argument");
  for (int v = 0; v < b->nvar; ++v) {
    const mt_tm2d_adj& a = adj[v];
    if (!a.aEz || !a.aDz || !a.aHx || !a.aHy || !a.apsi_Ezy || !a.apsi_Ezx || !a.apsi_Hyx || !a.apsi_Hxy ||
        !a.gHy || !a.gHx || !a.gEx || !a.gEy || !a.eps_grad || (k > 0 && !tapes[v]))
      return fail(MT_EINVAL, "batched adjoint: null adjoint pointer");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc = b->dtype == MT_F64 ? adjoint_prepare_t<double>(b, adj, tapes, mons, mon_off, s)
                              : adjoint_prepare_t<float>(b, adj, tapes, mons, mon_off, s);
  if (rc) return rc;
  return b->dtype == MT_F64 ? adjoint_sweep_t<double>(b, step0, k, s) : adjoint_sweep_t<float>(b, step0, k, s);
}

extern "C" int mt_tm2d_batch_adjoint_prepare(mt_batch* b, const mt_tm2d_adj* adj, void* const* tapes,
                                             const mt_monitor* mons, const int* mon_off, void* stream) {
  if (!b || !adj || !tapes || !mon_off) return fail(MT_EINVAL, "batched adjoint: null argument");
  for (int v = 0; v < b->nvar; ++v) {
    const mt_tm2d_adj& a = adj[v];
    if (!a.aEz || !a.aDz || !a.aHx || !a.aHy || !a.apsi_Ezy || !a.apsi_Ezx || !a.apsi_Hyx || !a.apsi_Hxy ||
        !a.gHy || !a.gHx || !a.gEx || !a.gEy || !a.eps_grad || !tapes[v])
      return fail(MT_EINVAL, "batched adjoint: null adjoint pointer");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->dtype == MT_F64 ? adjoint_prepare_t<double>(b, adj, tapes, mons, mon_off, s)
                            : adjoint_prepare_t<float>(b, adj, tapes, mons, mon_off, s);
}

extern "C" int mt_tm2d_batch_adjoint_pingpong(mt_batch* b, const mt_tm2d_adj* alt, const mt_monitor* mons,
                                              const int* mon_off, int* fused, void* stream) {
  if (!b || !alt || !mon_off || !fused) return fail(MT_EINVAL, "batched adjoint: null argument");
  if (!b->adj_ready || (int)b->adj0.size() != b->nvar)
    return fail(MT_EINVAL, "batched adjoint: descriptors not prepared");
  for (int v = 0; v < b->nvar; ++v) {
    const mt_tm2d_adj& a = alt[v];
    if (!a.aEz || !a.aDz || !a.aHx || !a.aHy || !a.apsi_Ezy || !a.apsi_Ezx || !a.apsi_Hyx || !a.apsi_Hxy ||
        !a.eps_grad)
      return fail(MT_EINVAL, "batched adjoint: null pointer in the second cotangent set");
    if (a.eps_grad != b->adj0[v].eps_grad)
      return fail(MT_EINVAL, "batched adjoint: both cotangent sets must share eps_grad");
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->dtype == MT_F64 ? adjoint_pingpong_t<double>(b, alt, mons, mon_off, fused, s)
                            : adjoint_pingpong_t<float>(b, alt, mons, mon_off, fused, s);
}

extern "C" int mt_tm2d_batch_adjoint_sweep(mt_batch* b, int step0, int k, void* stream) {
  if (!b) return fail(MT_EINVAL, "batched adjoint: null plan");
  if (k < 0) return fail(MT_EINVAL, "negative step count");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return b->dtype == MT_F64 ? adjoint_sweep_t<double>(b, step0, k, s) : adjoint_sweep_t<float>(b, step0, k, s);
}

extern "C" void mt_batch_destroy(mt_batch* b) { delete b; }
