// Argument blocks of the temporally blocked 2D steppers (tiled2d.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace mt {

constexpr int kMaxSub = 8;    // max steps per tiled launch
constexpr int kMaxTaps = 32;

// One monitor term in tile coordinates (see mt_dft_term's semantic fields).
struct Tap {
  void* out;                 // staging base of this term (or probe stack)
  long long slot_stride;     // elements between consecutive step slots
  int slot0;                 // slot of the launch's first step
  int field, line, fix, lo, n, avg;
  int n0, n1;                // whole-field extents (line 3)
};

struct TapSet {
  int n;
  Tap t[kMaxTaps];
};

// Tile rectangle [r0, r1) x [c0, c1) of interior tiles in an R x C tiling.
struct TileMap {
  int R, C, r0, r1, c0, c1;
};

template <typename T>
struct TeTiled {
  int nx, ny;
  int in_i0, in_i1, in_j0, in_j1;  // cells whose tiles carry no psi
  TileMap map;
  T dt, dx, rdx;                   // rdx = RN(1/dx)
  const T *eps_x, *eps_y;
  Pml<T> hx, hy, ey, ex;
  const T *Hz, *Dx, *Dy, *pHzy, *pHzx, *pEyx, *pExy;   // inputs
  T *oHz, *oDx, *oDy, *oHzy, *oHzx, *oEyx, *oExy;     // outputs
  T *Ex, *Ey;                                          // final E (write_e)
  mt_source2d src;
  T coef[kMaxSub];
};

template <typename T>
struct TmTiled {
  int nx, ny, y_pmc, ny_e;
  int in_i0, in_i1, in_j0, in_j1;
  TileMap map;
  T dt, dx, rdx;
  const T* eps;
  Pml<T> hy, hx, ex, ey;
  const T *Ez, *Dz, *Hx, *Hy, *pEzy, *pEzx, *pHyx, *pHxy;
  T *oEz, *oDz, *oHx, *oHy, *oEzy, *oEzx, *oHyx, *oHxy;
  T* tape;
  mt_source2d src;
  T coef[kMaxSub];
  const T* coef_tab;  // batched runs: dt*amp for every global step (else NULL)
  // magnetic (TFSF) currents of the B half on Hx (hsrc[0]) and Hy (hsrc[1]):
  // H -= hc * w after the curl update, hc = hcoef[substep] or hcoef_tab[step]
  mt_source2d hsrc[2];
  T hcoef[kMaxSub];
  const T* hcoef_tab;
};

// One variant of a batched TM run: its problem in both ping-pong
// orientations (pA reads buffer set A and writes B, pB the reverse).
template <typename T>
struct TmBatchVar {
  TmTiled<T> pA, pB;
  TapSet taps;
};

int te_tiled_dispatch_f32(int variant, const TeTiled<float>& p, const TapSet& taps, int nsub, int write_e,
                          hipStream_t s);
int te_tiled_dispatch_f64(int variant, const TeTiled<double>& p, const TapSet& taps, int nsub, int write_e,
                          hipStream_t s);
int tm_tiled_dispatch_f32(int variant, const TmTiled<float>& p, const TapSet& taps, int nsub, int tape_s0,
                          hipStream_t s);
int tm_tiled_dispatch_f64(int variant, const TmTiled<double>& p, const TapSet& taps, int nsub, int tape_s0,
                          hipStream_t s);
int tm_batch_dispatch_f32(int variant, const TmBatchVar<float>* vars, const int* pre_in, const int* pre_ring,
                          int nvar, int n_in, int n_ring, int grid_in, int grid_ring, int nsub, int parity,
                          int t0, int slot_base, int tape_s0, hipStream_t s);
int tm_batch_dispatch_f64(int variant, const TmBatchVar<double>* vars, const int* pre_in, const int* pre_ring,
                          int nvar, int n_in, int n_ring, int grid_in, int grid_ring, int nsub, int parity,
                          int t0, int slot_base, int tape_s0, hipStream_t s);
TileMap tiled_map(int variant, int dtype, int nx, int ny, int in_i0, int in_i1, int in_j0, int in_j1);
int tiled_variant_count();
int tiled_variant_steps(int variant, int dtype);

// Build the tap set of a tiled launch from the monitor plan; false if some
// term has no semantic form (then the per-half-step path must be used).
bool build_taps(const MonitorPlan& mp, int slot0, int t_rel0, int elem_size, TapSet* out);

}  // namespace mt
