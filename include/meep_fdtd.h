/*
 * meep_fdtd.h -- C ABI of the MI355X FDTD time-stepping engine (libmeepfdtd.so).
 *
 * Plain C: device pointers, sizes and doubles; no torch types.  All pointers
 * marked "device" address HIP device memory of the run's dtype (MT_F32 or
 * MT_F64).  Arrays use the reference's tensor shapes, row-major, last axis
 * contiguous, so a caller hands over the storage of the reference state
 * tensors unchanged.  Every entry point returns 0 on success and a negative
 * code on error; mt_last_error() then returns a message (thread-local).
 * All work is enqueued on `stream` (a hipStream_t, NULL = default stream);
 * no entry point synchronises the device.
 */
#ifndef MEEP_FDTD_H
#define MEEP_FDTD_H

#ifdef __cplusplus
extern "C" {
#endif

#define MT_F32 0
#define MT_F64 1

#define MT_OK 0
#define MT_EINVAL (-1)
#define MT_EHIP (-2)
#define MT_ENOTIMPL (-3)

#define MT_MAX_TERMS 4

/* One 1D CPML profile (b, a) of length n.  Indices outside the shells
 * [0, lo_end) and [hi_begin, n) must hold exactly b = 1, a = 0 (the
 * reference's interior identity, meep_torch/core/pml.py:53-58), so the
 * kernels skip the psi recursion there. */
typedef struct {
  const void* b; /* device, length n */
  const void* a; /* device, length n */
  int n;
  int lo_end;
  int hi_begin;
} mt_pml_axis;

/* Soft source on a 2D field: dense weights on the bounding box
 * [i0,i1) x [j0,j1) of the reference's mask (sources.py:49-78); at step s the
 * field loses (coef[s] cast to dtype) * w (stepper_tm.py:197-200,
 * stepper_te.py:131-134).  coef is a HOST array of k doubles = dt * amp(t).
 * wstep = 0: one weight box for every step; wstep > 0 (electric sources of
 * mt_tm2d_run / mt_te2d_run only): step s of the call uses the box at
 * w + s * wstep elements -- the per-step sum of several sources'
 * amp_i(t) * mask_i that the reference injects as one mask
 * (mp/cpp/core.py:584-589), with coef[s] = dt. */
typedef struct {
  int present;
  int i0, i1, j0, j1;
  const void* w;                          /* device (i1-i0, j1-j0) [per step] */
  long long wstep;
} mt_source2d;

/* 2D TM problem + state (meep_torch/core/coeffs.py:21-74 TMCoeffs and
 * meep_torch/core/stepper_tm.py:12-34 TMState, eps/CPML/y_pmc subset).
 * Profiles follow the reference naming: hy (ny-1, half-y), hx (nx-1, half-x),
 * ex (nx-2, integer-x interior), ey (ny-2 interior, or ny under y_pmc). */
typedef struct {
  int nx, ny, dtype, y_pmc;
  double dt, dx;
  const void* eps;                       /* device (nx, ny) eps_inf */
  mt_pml_axis hy, hx, ex, ey;
  void *Ez, *Dz, *Hx, *Hy;               /* device (nx,ny) (nx,ny) (nx,ny-1) (nx-1,ny) */
  void *psi_Ezy, *psi_Ezx;               /* device (nx,ny-1) (nx-1,ny) */
  void *psi_Hyx, *psi_Hxy;               /* device (nx-2, ny_e), ny_e = ny-2 or ny (y_pmc) */
  void* tape;                            /* NULL, or device (k, nx, ny): Ez after each step of the call
                                          * (with kerr = 1: rhs = Dz - sum P after each step, from which
                                          * the adjoint recomputes the fixed-point iterates) */
  /* Temporally blocked path: variant >= 0 selects a tile shape / steps per
   * launch (mt_tiled_variants); -1 runs two passes per step.  alt[] are
   * second buffers of the same shapes as Ez Dz Hx Hy psi_Ezy psi_Ezx psi_Hyx
   * psi_Hxy holding an identical copy on entry (ping-pong targets); results
   * always end in the primary buffers. */
  int variant, pad_;
  void* alt[8];
  /* optional magnetic (TFSF) currents of the B phase (stepper_tm.py:169-193,
   * run.py:17-23): hsrc[0] on Hx (nx, ny-1), hsrc[1] on Hy (nx-1, ny); at
   * step s the field loses (hcoef[s] cast to dtype) * w right after its curl
   * update.  hcoef is a HOST array of k doubles = dt * h_amplitude(t).  Runs
   * with a magnetic current take the two-pass path. */
  mt_source2d hsrc[2];
  const double* hcoef;
  /* optional media (coeffs.py:87-106,154-200; stepper_tm.py:116-121,161-163,
   * 198-231), two-pass path only:
   *   n_poles electric ADE poles: c1, c2, c3 and the pole stacks P_cur,
   *     P_prev, all (n_poles, nx, ny); P(n+1) = c1 P(n) + c2 P(n-1) + c3 Ez(n),
   *     Ez = (Dz - sum P) / eps;
   *   kerr = 1: three fixed-point passes Ez = rhs / (eps + chi2 Ez + chi3 Ez^2)
   *     with chi2, chi3 (nx, ny);
   *   d1, d2 (nx, ny): Dz = d1 Dz + d2 curl (sigma_d); the electric source
   *     is then (d2 * coef[s]) * w with coef[s] the raw amplitude, not dt*amp;
   *   h1x, h2x (nx, ny-1), h1y, h2y (nx-1, ny): Hx = h1x Hx - h2x curl,
   *     Hy = h1y Hy + h2y curl (sigma_b and mu). */
  int n_poles, kerr;
  const void *c1, *c2, *c3;
  void *P_cur, *P_prev;
  const void *chi2, *chi3;
  const void *d1, *d2, *h1x, *h2x, *h1y, *h2y;
  /* optional magnetic Lorentz poles (coeffs.py:177-196, stepper_tm.py:103-106,
   * 180-193,209-224), two-pass path only, not with h1x..h2y: B integrates the
   * curl, Bx = Bx - dt curl, By = By + dt curl, the TFSF currents go into B,
   * then per pole M' = (mc1 M + mc2 M_prev) + mc3 H(n-1/2), M_prev' = M, and
   * H = B - sum M' (pole order).  mc1x, mc2x, mc3x, Mx_cur, Mx_prev are
   * (n_mpoles, nx, ny-1), the y ones (n_mpoles, nx-1, ny); Bx (nx, ny-1),
   * By (nx-1, ny). */
  int n_mpoles, pad4_;
  const void *mc1x, *mc2x, *mc3x, *mc1y, *mc2y, *mc3y;
  void *Bx, *By, *Mx_cur, *Mx_prev, *My_cur, *My_prev;
  /* symmetry cuts at the low edges (coeffs.py:211-281, stepper_tm.py:
   * 94-155), two-pass path only, not with y_pmc: y_mirror / x_mirror parity
   * +-1, y_rot2 (Rotate2 phase at the low-y cut) and rot4 (Rotate4 phase at
   * both cuts, square grids) +-1, 0 = off.  A cut makes the Dz row / column
   * on the cut live: profile ey (ny-1) / ex (nx-1), psi_Hxy / psi_Hyx one
   * node longer on that axis, and the H just beyond the cut is 0 (even
   * mirror) or the on-plane row Hx0 (nx) under an odd y mirror, Rotate2 or
   * Rotate4, Hy0 (ny) under an odd x mirror or Rotate4, each advanced in
   * the B phase from a ghost Ez: H0 -/+= (dt (Ez_cut - ghost)) / dx. */
  int y_mirror, x_mirror, y_rot2, rot4;
  void *Hx0, *Hy0;
} mt_tm2d;

/* 2D TE problem + state (coeffs.py:294-345 TECoeffs, stepper_te.py:11-24). */
typedef struct {
  int nx, ny, dtype, pad_;
  double dt, dx;
  const void *eps_x, *eps_y;             /* device (nx-1,ny) (nx,ny-1) */
  mt_pml_axis hx, hy, ey, ex;            /* nx-1, ny-1, ny-2, nx-2 */
  void *Hz, *Ex, *Ey, *Dx, *Dy;          /* (nx-1,ny-1) (nx-1,ny) (nx,ny-1) (nx-1,ny) (nx,ny-1) */
  void *psi_Eyx, *psi_Exy;               /* (nx-1,ny-1) each */
  void *psi_Hzy, *psi_Hzx;               /* (nx-1,ny-2) (nx-2,ny-1) */
  /* Tiled path as for mt_tm2d; alt[] shadows Hz Dx Dy psi_Eyx psi_Exy
   * psi_Hzy psi_Hzx (alt[7] unused).  Ex/Ey are written at the end of the call. */
  int variant, pad2_;
  void* alt[8];
  /* optional device tape (two-pass path only): E after step s of a call is
   * written to slot s = tape + s * ((nx-1)*ny + nx*(ny-1)) as [Ex | Ey], for
   * mt_te2d_adjoint's eps gradient; under chi2/chi3 (kerr) or tensor eps the slot holds
   * [Dx - sum Px | Dy - sum Py], from which the adjoint recomputes the
   * fixed-point iterates */
  void* tape;
  /* optional magnetic (TFSF) current on Hz (nx-1, ny-1) of the B half
   * (stepper_te.py:120-128), as hsrc/hcoef of mt_tm2d */
  mt_source2d hsrc;
  const double* hcoef;
  /* optional media (coeffs.py:357-440; stepper_te.py:96-103,110-115,
   * 131-134,137-168), two-pass path only:
   *   npx / npy electric ADE poles on Ex / Ey: c1x, c2x, c3x and the stacks
   *     Px_cur, Px_prev (npx, nx-1, ny); c1y, c2y, c3y, Py_cur, Py_prev
   *     (npy, nx, ny-1); P(n+1/2) = c1 P(n-1/2) + c2 P(n-3/2) + c3 E(n-1/2),
   *     E = (D - sum P) / eps;
   *   d1x, d2x (nx-1, ny), d1y, d2y (nx, ny-1): Dx = d1x Dx + d2x curl_x,
   *     Dy = d1y Dy - d2y curl_y (sigma_d); the electric source is then
   *     (d2y * coef[s]) * w with coef[s] the raw amplitude;
   *   hz1, hz2 (nx-1, ny-1): Hz = hz1 Hz - hz2 curl (sigma_b and mu). */
  int npx, npy;
  const void *c1x, *c2x, *c3x, *c1y, *c2y, *c3y;
  void *Px_cur, *Px_prev, *Py_cur, *Py_prev;
  const void *d1x, *d2x, *d1y, *d2y, *hz1, *hz2;
  /* kerr = 1: three Jacobi passes of E = rhs / (eps + chi2 E + chi3 |E|^2)
   * with |E|^2 = E^2 + (four-point average of the other component)^2
   * (stepper_te.py:52-61,161-175); chi2x, chi3x (nx-1, ny), chi2y, chi3y
   * (nx, ny-1); kx, ky: device scratch of the Ex and Ey shapes. */
  int kerr, pad3_;
  const void *chi2x, *chi3x, *chi2y, *chi3y;
  void *kx, *ky;
  /* optional in-plane tensor eps (coeffs.py:390-400, stepper_te.py:158-160):
   * inv_xx_x, inv_xy_x (nx-1, ny), inv_yy_y, inv_xy_y (nx, ny-1), all four or
   * none; E = inv_xx_x rhs_x + inv_xy_x avg4(rhs_y) (and the y analogue)
   * with rhs = D - sum P, before any chi2/chi3 passes.  Optional oriented
   * poles (coeffs.py:425-455, stepper_te.py:137-147): c3xy (npx, nx-1, ny)
   * and c3yx (npy, nx, ny-1), both or none: P' += c3xy avg4(Ey(n-1/2)) on Ex
   * (c3yx avg4(Ex) on Ey).  avg4 is the four-point average with zeros
   * beyond the walls.  Either needs kx, ky. */
  const void *inv_xx_x, *inv_xy_x, *inv_yy_y, *inv_xy_y;
  const void *c3xy, *c3yx;
  /* mirror cuts at the low edges (coeffs.py:457-469, stepper_te.py:80-97),
   * two-pass path only: parity +-1, 0 = off; an even cut makes the on-plane
   * tangential E row live with curl (2 Hz[0]) / dx (Dx row 0 under y_mirror
   * = +1, Dy column 0 under x_mirror = +1); the profiles carry no low-side
   * shell. */
  int y_mirror, x_mirror;
} mt_te2d;


/* 3D problem + state (stepper_3d.py:28-67 Coeffs3D / State3D, CPML path;
 * material conductivity and ADE poles through the media fields at the end).
 * lean = 1 selects the in-place lean stepper's E update and source sign
 * (stepper_3d_lean.py:234-296).  absorber = 1 selects the path of a run
 * without CPML (stepper_3d.py:328-361: Coeffs3D.cpml is None, the boundary
 * is the matched-sigma absorber folded into the damping pairs, or plain PEC):
 *   H = d1 H - (d2 (dE_a - dE_b)) / dx,  E = (d1 E eps_c + d2 (dH_a - dH_b) / dx) / eps_c
 * (poles as below); it needs the 12 damping arrays, lean = 0, every profile
 * with lo_end = 0 and hi_begin = n (b, a and psi may be NULL), and runs the
 * two-pass kernels.  A CPML run whose absorber axes carry the reference's
 * identity profile (b = a = 0, k = 1) is an ordinary absorber = 0 run whose
 * shells cover those axes. */
typedef struct {
  int nx, ny, nz, dtype, lean, absorber;
  double dt, dx;
  const void* eps;                        /* device (nx,ny,nz) cell-centred */
  mt_pml_axis h[3];                       /* per axis, length n-1 (offset 1/2) */
  mt_pml_axis e[3];                       /* per axis, length n-2 (interior) */
  void *Ex, *Ey, *Ez, *Hx, *Hy, *Hz;      /* reference State3D shapes */
  /* psi, reference shapes (stepper_3d.py:47-58), in this order:
   * Ez_y Ey_z Ex_z Ez_x Ey_x Ex_y Hz_y Hy_z Hx_z Hz_x Hy_x Hx_y */
  void* psi[12];
  /* optional device tape: when set, E after step s of a call is written to
   * slot s = tape + s * (|Ex| + |Ey| + |Ez|) as [Ex | Ey | Ez]
   * (reference shapes), for the adjoint's eps gradient */
  void* tape;
  /* variant 0: two passes per step (H kernel, E kernel), in place; 1-3:
   * split step -- one pass streaming along x over the psi-free interior box
   * (tile rows 6, 14, 4; FP64 uses 6 for variant 2) plus the per-cell H and
   * E updates over the CPML shell pieces -- which needs alt[0..5]: second
   * buffers of Ex Ey Ez Hx Hy Hz (same shapes), used ping-pong (psi stay in
   * place); 4, 5: two steps per pass over the slab inside the x and y CPML
   * (every z; tile rows 10, 8) plus two per-cell steps of the x and y CPML
   * slabs, which needs alt[0..5] as above, alt[6..11] (a third field set)
   * and alt[12..15] (scratch copies of psi Ey_z Ex_z Hy_z Hx_z); FP32 runs
   * without tape whose monitors each read one x plane only, other runs take
   * variant 0.  The state ends in the primary buffers either way.  Media
   * run the two-pass kernels whatever the variant. */
  int variant, pad2_;
  void* alt[18];
  /* Media of the functional stepper (stepper_3d.py:217-266,295-301); two-pass
   * variant 0 (other variants take it).  damp: all NULL (sigma = 0, no poles) or 12
   * device arrays, (d1, d2) of ex, ey, ez on the E component grids then of hx,
   * hy, hz on the H component grids, from the Yee-averaged sigma_d / sigma_b:
   * d1 = (1 - s dt/2) / (1 + s dt/2), d2 = dt / (1 + s dt/2), and
   *   H = d1 H - d2 curlE,   E = (d1 E eps_c + d2 curlH) / eps_c.
   * n_poles > 0 (needs damp): c1, c2 device arrays of n_poles scalars, and per
   * E component a = x, y, z the stacks c3[a], P_cur[a], P_prev[a] of shape
   * (n_poles, component grid):
   *   P' = c1 P + c2 P_prev + c3 E(n),  P_prev' = P,
   *   E = (d1 (E eps_c + sum P) + d2 curlH - sum P') / eps_c,
   * the sums taken in pole order.  With lean = 1 (stepper_3d_lean.py:143-165,
   * 207-231) damp stays NULL (the lean form has no conductivity) and poles
   * take the lean form:
   *   P' = fma(c3, E(n), fma(c1, P, c2 P_prev)),
   *   E = E + ((curlH dt + sum P) - sum P') * (1 / eps_c). */
  const void* damp[12];
  int n_poles, pad3_;
  const void *c1, *c2;
  const void* c3[3];
  void* P_cur[3];
  void* P_prev[3];
} mt_yee3d;

/* 3D soft source: kind 1 = point at (i0, j0, k0), kind 2 = mask given as dense
 * weights w on the box [i0,i1) x [j0,j1) x [k0,k1) of the component grid.
 * comp: 0 Ex, 1 Ey, 2 Ez.  Functional form adds (dt*amp) * (1/eps_c) at a
 * point (stepper_3d.py:475-481) and subtracts ((dt*amp) * w) / eps_c for a
 * mask (stepper_3d.py:452-461); lean form subtracts (dt*amp) * inv_eps and
 * (w * (dt*amp)) * inv_eps (stepper_3d_lean.py:285-296).  A step whose
 * coefficient is exactly 0 injects nothing (stepper_3d.py:449). */
typedef struct {
  int kind, comp;
  int i0, i1, j0, j1, k0, k1;
  const void* w;                          /* device (i1-i0, j1-j0, k1-k0), kind 2 */
} mt_source3d;

/* A DFT monitor: up to MT_MAX_TERMS strided 2D field views sampled after
 * every step.  Term value = a[e] or 0.5*(a[e]+b[e]) (H averaged onto the E
 * nodes).  Per step t the accumulator gains
 *   dt*(cos(w*t_q) v),  dt*(-sin(w*t_q) v)        (monitors.py:84-95)
 * in slots 2q, 2q+1 of acc (nfreq, 2*n_terms, n0, n1), summed in step order.
 * t_q = (t + 1/2) dt if t_half else (t + 1) dt, rounded to dtype.
 * kind 1 = probe: single element stacked per step into out[t - step0]. */
typedef struct {
  const void* a;
  const void* b;                          /* NULL: no averaging */
  long long s0, s1;                       /* element strides */
  int t_half;
  /* Semantic form of the same view, used by the tiled 2D steppers (which
   * sample from on-chip tiles, not from global arrays): field id (TM: 0 Ez,
   * 1 Hx, 2 Hy, 3 Dz; TE: 0 Hz, 1 Ex, 2 Ey, 3 Dx, 4 Dy), line = 1 row
   * (fixed first index `fix`, second index lo + e), 2 column (fixed second
   * index, first index lo + e), 3 whole field (row-major), 4 single cell
   * (fix, lo); averaged terms average positions fix-1 and fix.
   * line = 0: no semantic form (the per-half-step steppers are used). */
  int field, line, fix, lo, pad_;
} mt_dft_term;

typedef struct {
  int kind;                               /* 0 DFT (reduce "sum"), 1 probe (reduce "stack") */
  int n_terms, n0, n1, nfreq, pad_;
  mt_dft_term term[MT_MAX_TERMS];
  const void* omega;                      /* device, nfreq values of 2*pi*f in dtype */
  void* acc;                              /* device accumulators (DFT) or (k,) stack (probe) */
} mt_monitor;

/* Tile-shape variants of the temporally blocked 2D steppers for `dtype`:
 * returns their number; for 0 <= v < count fills steps[v] (Yee steps per
 * launch) and tx[v], ty[v] (owned tile rows/columns) when the arrays are
 * non-NULL.  Used by the host-side autotuner. */
int mt_tiled_variants(int dtype, int* steps, int* tx, int* ty);

/* Tile counts of the temporally blocked launch that mt_tm2d_run /
 * mt_te2d_run (and the batched runner) make for p->variant on this problem:
 * interior tiles (no CPML psi in their LDS region) and ring tiles.  Host
 * only, no device work.  No counterpart in the reference: an introspection
 * aid so tests can show that a grid exercises the interior kernels. */
int mt_tm2d_tiles(const mt_tm2d* p, int* n_interior, int* n_ring);
int mt_te2d_tiles(const mt_te2d* p, int* n_interior, int* n_ring);

/* RN(1/b) if the tiled kernels' constant-divisor sequence (q0 = a*y, one
 * FMA correction) equals IEEE a / b for every FP32 dividend, checked
 * exhaustively on the host (cached per b); 0 otherwise, in which case the
 * kernels divide.  Host only.  No counterpart in the reference, whose steps
 * divide by dx with torch's true division (stepper_tm.py:86-89). */
float mt_div_recip_f32(float b);

/* Bytes of staging workspace one call needs for `stage_steps` buffered steps. */
long long mt_stage_bytes(const mt_monitor* mons, int n_mons, int stage_steps, int dtype);

/* Advance a 2D TM state k steps starting at global step index step0.
 * Replaces the per-step loop of meep_torch/core/run.py:10-40 (_run_chunk)
 * driving step_tm (meep_torch/core/stepper_tm.py:234-254) and the monitor
 * samples of meep_torch/core/monitors.py:11-100.  Monitor accumulators are
 * summed in place (the caller zeroes them at a chunk start). */
int mt_tm2d_run(const mt_tm2d* p, const mt_source2d* src, const double* coef,
                const mt_monitor* mons, int n_mons, void* stage, long long stage_bytes,
                int stage_steps, int step0, int k, void* stream);

/* Allocation-free variant of mt_tm2d_run.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as mt_tm2d_run without the staging workspace: the
 * DFT staging ring lives in the library's static device array of
 * mt_static_stage_bytes() bytes (16 MiB), one per GPU, held until the
 * process ends; a call whose monitors need more than that for one step fails
 * with MT_EINVAL.  Nothing is allocated (the tiled variants still step
 * through the caller's ping-pong field copies in p->alt, as mt_tm2d_run
 * does).  Calls that overlap on one GPU share the array and race. */
int mt_tm2d_run_static(const mt_tm2d* p, const mt_source2d* src, const double* coef,
                       const mt_monitor* mons, int n_mons, int step0, int k, void* stream);

/* Bytes of the static staging array behind mt_tm2d_run_static. */
long long mt_static_stage_bytes(void);

/* Same for 2D TE: _run_chunk driving step_te (meep_torch/core/stepper_te.py:178-190)
 * with FluxMonitorTE/FluxMonitorTEY/DFTFieldsMonitor (monitors.py:24-155). */
int mt_te2d_run(const mt_te2d* p, const mt_source2d* src, const double* coef,
                const mt_monitor* mons, int n_mons, void* stage, long long stage_bytes,
                int stage_steps, int step0, int k, void* stream);

/* Advance a 3D state k steps: replaces run_3d (meep_torch/core/monitors_3d.py:192-210)
 * over step_3d (stepper_3d.py:303-445) or, with lean = 1, run_3d_lean's loop
 * (stepper_3d_lean.py:307-349), with FluxMonitor3D / DFTFieldsMonitor3D. */
int mt_yee3d_run(const mt_yee3d* p, const mt_source3d* src, const double* coef,
                 const mt_monitor* mons, int n_mons, void* stage, long long stage_bytes,
                 int stage_steps, int step0, int k, void* stream);

/* Adjoint state of a 2D TM run: cotangents of every TMState field (same
 * shapes), scratch for the transposed differences, and the eps gradient. */
typedef struct {
  void *aEz, *aDz, *aHx, *aHy;
  void *apsi_Ezy, *apsi_Ezx, *apsi_Hyx, *apsi_Hxy;
  void *gHy, *gHx;                        /* scratch (nx, ny) each */
  void *gEx, *gEy;                        /* scratch (nx-1, ny) and (nx, ny-1) */
  void* eps_grad;                         /* device (nx, ny), accumulated */
  /* cotangents of the pole stacks P_cur, P_prev (n_poles, nx, ny) when the
   * problem has ADE poles; NULL otherwise */
  void *aP_cur, *aP_prev;
  /* optional gradients of the other differentiable coefficients, accumulated
   * like eps_grad (NULL: not wanted): chi2_grad, chi3_grad (nx, ny) of the
   * chi2 / chi3 maps, c3_grad (n_poles, nx, ny) of the poles' c3 (the pole
   * masks enter the step through it).  c3_grad needs ez_init, the Ez before
   * the first step of the call (the tape holds the Ez, or under chi2/chi3
   * the Dz - sum P, after each step). */
  void *chi2_grad, *chi3_grad, *c3_grad;
  const void* ez_init;
  /* magnetic Lorentz poles (p->n_mpoles > 0): cotangents of Bx (nx, ny-1),
   * By (nx-1, ny) and of the M stacks Mx_cur, Mx_prev (n_mpoles, nx, ny-1),
   * My_cur, My_prev (n_mpoles, nx-1, ny); NULL otherwise */
  void *aBx, *aBy, *aMx_cur, *aMx_prev, *aMy_cur, *aMy_prev;
} mt_tm2d_adj;

/* Reverse-time adjoint of k TM steps [step0, step0+k): on entry the mt_tm2d_adj
 * state holds the cotangent of the state after step step0+k-1, on return the
 * cotangent of the state before step step0; eps_grad gains dL/deps of the k
 * steps.  `tape` (k, nx, ny) holds the forward Ez after each step (recorded by
 * mt_tm2d_run with p->tape set).  Monitor descriptors are those of the forward
 * run with term views pointing into the ADJOINT fields and `acc` holding the
 * accumulator cotangent dL/dacc; each step injects
 *   dt*(cos(w t) dacc_re - sin(w t) dacc_im)  (halved onto both rows of an
 * averaged term).  Conductivity (d1, d2, h1x, h2x, h1y, h2y of p, all or none)
 * and ADE poles (c1, c2, c3 of p with a->aP_cur, a->aP_prev) are transposed
 * with the recurrences; so are magnetic poles (mc1x..mc3y of p with the aB* /
 * aM* cotangents of a: B' = B +- dt curl, M' = (mc1 M + mc2 M_prev) + mc3 H,
 * H' = B' - sum M').  Replaces
 * PyTorch autograd through run_tm's checkpointed chunks
 * (meep_torch/core/run.py:99-120, eps reached via stepper_tm.py:226-227).
 * Even mirror cuts (y_mirror / x_mirror = +1) are transposed with the
 * forward's live Dz set and psi indexing; odd and Rotate cuts are refused. */
int mt_tm2d_adjoint(const mt_tm2d* p, const mt_tm2d_adj* a, const void* tape,
                    const mt_monitor* mons, int n_mons, int step0, int k, void* stream);

/* Adjoint state and scratch of one TE problem (same shapes as the forward
 * state; psi cotangents are tracked inside the CPML shells only). */
typedef struct {
  void *aHz, *aEx, *aEy, *aDx, *aDy;
  void *apsi_Eyx, *apsi_Exy, *apsi_Hzy, *apsi_Hzx;
  void *gEy, *gEx;            /* scratch (nx-1,ny-1) each */
  void *gDx, *gDy;            /* scratch (nx-1,ny) (nx,ny-1) */
  void *epsx_grad, *epsy_grad;  /* (nx-1,ny) (nx,ny-1), accumulated */
  /* cotangents of the pole stacks Px_cur, Px_prev (npx, nx-1, ny) and
   * Py_cur, Py_prev (npy, nx, ny-1) when the problem has them; else NULL */
  void *aPx_cur, *aPx_prev, *aPy_cur, *aPy_prev;
  /* chi2/chi3 problems (p->kerr), tensor eps and oriented poles (p->c3xy):
   * workspace of 6 * ((nx-1)*ny + nx*(ny-1)) elements for the recomputed
   * fixed-point iterates and the cotangents of the four-point averages (under
   * chi2/chi3 the forward tape holds rhs = D - sum P, not E); else NULL */
  void* kerr_ws;
  /* tensor eps (p->inv_xx_x set; needs kerr_ws too): gradients of the four
   * inverse maps inv_xx_x, inv_xy_x (nx-1, ny), inv_yy_y, inv_xy_y (nx, ny-1),
   * accumulated (the tape then holds D - sum P); else NULL */
  void *ixx_grad, *ixy_x_grad, *iyy_grad, *ixy_y_grad;
} mt_te2d_adj;

/* Reverse sweep of k TE steps starting at global step step0 (the gradient of
 * torch.utils.checkpoint'ed run_te chunks, meep_torch/core/run.py:81-118,
 * through step_te, meep_torch/core/stepper_te.py:73-190): for s = k-1..0 the
 * monitor cotangents of step step0+s are injected, then the exact transpose
 * of the step runs, with E after the step read from tape slot s (as written
 * by mt_te2d_run).  Accumulates dL/deps_x and dL/deps_y.  Conductivity
 * (d1x, d2x, d1y, d2y, hz1, hz2 of p, all or none) and ADE poles (c1x..c3y
 * of p with the aP* cotangents of a) are transposed with the recurrences.
 * chi2/chi3 (kerr = 1, with a->kerr_ws and the rhs tape): the three Jacobi
 * passes of stepper_te.py:169-175 are recomputed from the taped rhs and
 * transposed last to first, their four-point averages scattered back with
 * the other average (each is the other's transpose).  Oriented poles (c3xy,
 * c3yx, with a->kerr_ws): each pole's cotangent times c3xy is summed per Ex
 * cell and scattered onto the old Ey by the same transposed average (and
 * c3yx onto the old Ex).  Even mirror cuts (y_mirror / x_mirror = +1): the
 * on-plane Dx row / Dy column took 2 Hz[0] / dx, whose cotangent returns to
 * Hz[0]; odd cuts keep the zero row. */
int mt_te2d_adjoint(const mt_te2d* p, const mt_te2d_adj* a, const void* tape, const mt_monitor* mons,
                    int n_mons, int step0, int k, void* stream);

/* Adjoint state of the 3D functional stepper (sigma = 0, CPML): cotangents of
 * every state array (reference shapes, updated in place), 12 scratch arrays
 * of difference cotangents -- the first six shaped like the E components
 * that read them (Hz_y: Ex, Hy_z: Ex, Hx_z: Ey, Hz_x: Ey, Hy_x: Ez, Hx_y: Ez),
 * the last six like the H components (Ez_y: Hx, Ey_z: Hx, Ex_z: Hy, Ez_x: Hy,
 * Ey_x: Hz, Ex_y: Hz) --, three eps_c cotangents shaped like Ex, Ey, Ez, and
 * eps_grad (nx,ny,nz) into which dL/deps is accumulated. */
typedef struct {
  void *aEx, *aEy, *aEz, *aHx, *aHy, *aHz;
  void* apsi[12];        /* order of mt_yee3d.psi */
  void* g[12];
  void* geps[3];
  void* eps_grad;
  /* cotangents of the pole stacks P_cur[a], P_prev[a] of mt_yee3d when it
   * has ADE poles (a = x, y, z); NULL otherwise */
  void* aP_cur[3];
  void* aP_prev[3];
} mt_yee3d_adj;

/* Reverse-time adjoint of k steps of mt_yee3d_run (lean = 0) from global step
 * step0: inject the monitor cotangents (monitor term views pointing into the
 * adjoint arrays, acc = dL/dacc), then transpose the E and H half-steps.
 * tape holds k + 1 slots: slot 0 = E at step0, slot s + 1 = E after step
 * step0 + s.  Replaces the autograd reverse sweep of run_3d_adjoint
 * (meep_torch/core/adjoint3d.py:56-115). */
int mt_yee3d_adjoint(const mt_yee3d* p, const mt_yee3d_adj* a, const void* tape, const mt_monitor* mons,
                     int n_mons, int step0, int k, void* stream);

/* ---- Batched TM design variants (uneven grid sizes, one persistent launch) ----
 * Each item is a full 2D TM problem with its own state (primary buffers plus
 * alt[] ping-pong buffers), source, per-step coefficient table on the device
 * (coef_tab[t] = dt*amp(t) for every global step t of the run, dtype of the
 * problem) and monitors with their staging workspace.  All items share the
 * dtype.  The plan copies per-variant descriptors into `dev_ws` (device
 * memory of mt_tm2d_batch_workspace bytes) once; afterwards every call only
 * passes scalars, so the launches can be captured in a graph and replayed. */
typedef struct {
  mt_tm2d prob;
  mt_source2d src;
  const void* coef_tab;
  const mt_monitor* mons;
  int n_mons, pad_;
  void* stage;
  long long stage_bytes;
  /* magnetic (TFSF) currents prob.hsrc[0..1] (ports/mode_source.py:60-81):
   * device table hcoef_tab[t] = dt*h_amplitude(t) for every global step t
   * (dtype of the problem); NULL without them */
  const void* hcoef_tab;
} mt_tm2d_batch_item;

typedef struct mt_batch mt_batch;       /* opaque host-side plan */

long long mt_tm2d_batch_workspace(int nvar, int dtype);

/* Build the plan (tile `variant` of mt_tiled_variants; stage_steps = DFT
 * staging ring length shared by all variants).  Returns NULL on error. */
mt_batch* mt_tm2d_batch_create(const mt_tm2d_batch_item* items, int nvar, int variant, int stage_steps,
                               void* dev_ws, long long ws_bytes, void* stream);

/* Advance every variant k steps from global step step0 in one persistent
 * launch pair per block of steps (interior tiles, CPML-ring tiles); monitor
 * spectra are folded per variant; states end in the primary buffers.  With
 * tapes set in the items, Ez after each step s of the call goes to tape[s].
 * Replaces a Python loop of run_tm calls over variants
 * (examples_2/y_splitter/c_y_splitter.py:154-200 design pattern). */
int mt_tm2d_batch_run(mt_batch* b, int step0, int k, void* stream);

/* Reverse-time adjoint of k steps for every variant at once (the batched form
 * of mt_tm2d_adjoint): adj[v], tape[v] and the monitor cotangent descriptors
 * of variant v (mons + mon_off[v] .. mons + mon_off[v+1]). */
int mt_tm2d_batch_adjoint(mt_batch* b, const mt_tm2d_adj* adj, void* const* tapes, const mt_monitor* mons,
                          const int* mon_off, int step0, int k, void* stream);

/* The same adjoint in two parts, for callers that replay it from a graph:
 * _prepare uploads the descriptors (adj, tapes, monitor cotangents) into the
 * plan's workspace and waits for the upload; _sweep then runs the reverse
 * steps [step0, step0+k) with launches only (capturable).  The buffers the
 * descriptors point to must stay in place between prepare and every sweep.
 * Same math as mt_tm2d_batch_adjoint (core/run.py:99-120 checkpointed
 * backward, batched). */
int mt_tm2d_batch_adjoint_prepare(mt_batch* b, const mt_tm2d_adj* adj, void* const* tapes, const mt_monitor* mons,
                                  const int* mon_off, void* stream);
int mt_tm2d_batch_adjoint_sweep(mt_batch* b, int step0, int k, void* stream);

/* Optional second cotangent set for the batched sweep (call after
 * mt_tm2d_batch_adjoint_prepare): alt[v] has the same shapes as adj[v] and
 * the same eps_grad; mons are the monitor injections into that set (same
 * order and offsets).  When no variant has conductivity or ADE poles
 * (*fused = 1) every later sweep runs each reverse step as one fused pass
 * over tiles that reads one set and writes the other (the transposed
 * differences stay in LDS), and leaves the result in the first set; the
 * gradient is bit-identical to the three-pass sweep.  Same math as
 * mt_tm2d_batch_adjoint (the checkpointed backward of core/run.py:99-120). */
int mt_tm2d_batch_adjoint_pingpong(mt_batch* b, const mt_tm2d_adj* alt, const mt_monitor* mons, const int* mon_off,
                                   int* fused, void* stream);

void mt_batch_destroy(mt_batch* b);

/* Kernel timing hook.  No counterpart in the reference: a measurement aid
 * for benchmarks.  on = 1: every launch of the dominant stencil kernel (2D:
 * the temporally blocked interior-tile launch; 3D: one full step) is
 * bracketed by HIP events on its own stream and its cell-steps are counted;
 * on = 2: the same for every fused batched adjoint step; 3: both; 0: off.  mt_kernel_timing_read waits for those events, returns the
 * summed duration, launch count and cell-steps, and clears the record.
 * Host-thread-unsafe: use from one thread. */
int mt_kernel_timing_enable(int on);
int mt_kernel_timing_read(double* total_ms, long long* launches, long long* cell_steps);

/* Message of the last failing call on this thread ("" if none). */
const char* mt_last_error(void);

/* ABI version (bumped on any struct layout change). */
int mt_abi_version(void);

/* Layout of the ABI structs as this library was compiled, for bindings to
 * check their own declarations against (no counterpart in the reference: a
 * binding aid).  Writes up to n values and returns how many there are:
 * sizeof of mt_pml_axis, mt_source2d, mt_tm2d, mt_te2d, mt_yee3d,
 * mt_source3d, mt_dft_term, mt_monitor, mt_tm2d_adj, mt_te2d_adj,
 * mt_yee3d_adj, mt_tm2d_batch_item, then offsetof of mt_tm2d.hsrc,
 * mt_tm2d.hcoef, mt_te2d.tape, mt_te2d.hsrc, mt_te2d.hcoef, mt_yee3d.alt,
 * mt_monitor.acc, mt_tm2d.n_poles, mt_tm2d.h2y, mt_te2d.npx, mt_te2d.hz2,
 * mt_te2d.ky, mt_yee3d.damp, mt_yee3d.P_prev. */
int mt_abi_layout(long long* out, int n);

#ifdef __cplusplus
}
#endif
#endif /* MEEP_FDTD_H */
