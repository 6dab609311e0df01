// On-the-fly DFT accumulation for flux lines / planes and field monitors.
//
// After every step a gather kernel writes each monitor term's sample value
// (E, or the two-point H average 0.5*(a+b)) into a staging ring of
// `stage_steps` slots.  Every `stage_steps` steps (and at the end of a call) a
// phase-table kernel evaluates cos/sin(w*t) once per (frequency, step) and a
// flush kernel folds the staged samples into the accumulators, one thread per
// (term, element), looping over frequencies and then over the staged steps in
// step order -- the same left-to-right summation as the reference's
// per-step `partials + smp` (meep_torch/core/run.py:29-33), so FP32 spectra
// round like the reference's.
#include <hip/hip_runtime.h>

#include <vector>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <thread>

#include <xmmintrin.h>

#include "common.h"
#include "tiled2d.h"

namespace mt {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(MT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return MT_OK;
}

// Exhaustive host check of div_c's quotient sequence for one FP32 divisor b:
// with y = RN(1/b), q0 = RN(a y), r = fma(-q0, b, a), q = copysign(fma(r, y,
// q0), a) must equal the IEEE quotient a / b for every dividend of magnitude
// in [2^-100, 2^100] (all 1.68e9 positive bit patterns there) and for zero;
// the sequence and the quotient are odd in a (copysign gives a zero quotient
// a's sign), so the negative patterns follow.  The kernels send every other
// dividend to true division.  Run once per divisor (cached), split over host
// threads, vectorised with AVX2/FMA where the host has them; the threads
// clear the FTZ/DAZ flags so that the check sees IEEE arithmetic.
namespace {

inline void ieee_denormals() { _mm_setcsr(_mm_getcsr() & ~0x8040u); }

bool markstein_range_generic(uint32_t lo, uint32_t hi, float b, float y) {
  ieee_denormals();
  for (uint32_t w = lo; w < hi; ++w) {
    float a;
    std::memcpy(&a, &w, 4);
    const float q = a / b;
    const float q0 = a * y;
    const float r = std::fma(-q0, b, a);
    const float q1 = std::copysign(std::fma(r, y, q0), a);
    if (std::memcmp(&q, &q1, 4) != 0) return false;
  }
  return true;
}

__attribute__((target("avx2,fma"))) bool markstein_range_fma(uint32_t lo, uint32_t hi, float b, float y) {
  ieee_denormals();
  constexpr uint32_t kBlock = 1u << 16;
  for (uint32_t w0 = lo; w0 < hi; w0 += kBlock) {
    const uint32_t w1 = std::min(hi, w0 + kBlock);
    uint32_t bad = 0;
    for (uint32_t w = w0; w < w1; ++w) {
      float a;
      std::memcpy(&a, &w, 4);
      const float q = a / b;
      const float q0 = a * y;
      const float r = __builtin_fmaf(-q0, b, a);
      const float q1 = __builtin_copysignf(__builtin_fmaf(r, y, q0), a);
      uint32_t u, v;
      std::memcpy(&u, &q, 4);
      std::memcpy(&v, &q1, 4);
      bad |= u ^ v;
    }
    if (bad) return false;
  }
  return true;
}

}  // namespace

static bool div_const_exact_f32(float b) {
  if (!(b >= 0x1p-16f && b <= 0x1p16f)) return false;
  volatile float vb = b;
  const float y = 1.0f / vb;
  const bool hw = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  constexpr uint32_t kLo = 27u << 23, kHi = (227u << 23) + 1;  // bit patterns of 2^-100 .. 2^100 inclusive
  const bool zero_ok = hw ? markstein_range_fma(0, 1, b, y) : markstein_range_generic(0, 1, b, y);
  if (!zero_ok) return false;
  unsigned nt = std::thread::hardware_concurrency();
  nt = std::max(1u, std::min(16u, nt));
  std::atomic<bool> ok{true};
  std::atomic<uint32_t> next{0};
  constexpr uint32_t kChunk = 1u << 24;
  auto worker = [&]() {
    for (;;) {
      const uint32_t c = next.fetch_add(1);
      const uint64_t lo = kLo + (uint64_t)c * kChunk;
      if (lo >= kHi || !ok.load()) return;
      const uint32_t hi = (uint32_t)std::min<uint64_t>(kHi, lo + kChunk);
      const bool pass = hw ? markstein_range_fma((uint32_t)lo, hi, b, y) : markstein_range_generic((uint32_t)lo, hi, b, y);
      if (!pass) ok.store(false);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nt; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  return ok.load();
}

float div_recip_f32(float b) {
  static std::mutex mu;
  static std::map<uint32_t, bool> seen;
  uint32_t key;
  std::memcpy(&key, &b, 4);
  std::lock_guard<std::mutex> lock(mu);
  auto it = seen.find(key);
  if (it == seen.end()) it = seen.emplace(key, div_const_exact_f32(b)).first;
  return it->second ? 1.0f / b : 0.0f;
}

int check_src(const mt_source2d* s, int n0, int n1) {
  if (!s || !s->present) return MT_OK;
  if (s->wstep < 0) return fail(MT_EINVAL, "source: negative weight step");
  if (s->i0 < 0 || s->j0 < 0 || s->i1 > n0 || s->j1 > n1 || s->i0 >= s->i1 || s->j0 >= s->j1 || !s->w)
    return fail(MT_EINVAL, "source box outside the field");
  return MT_OK;
}

int check_axis(const mt_pml_axis& ax, int n, const char* name) {
  if (ax.n != n) return fail(MT_EINVAL, std::string("profile ") + name + ": length " +
                                 std::to_string(ax.n) + " != " + std::to_string(n));
  if (n > 0 && (!ax.b || !ax.a)) return fail(MT_EINVAL, std::string("profile ") + name + ": null");
  if (ax.lo_end < 0 || ax.hi_begin > n || ax.lo_end > ax.hi_begin)
    return fail(MT_EINVAL, std::string("profile ") + name + ": bad shell bounds");
  return MT_OK;
}

namespace {

constexpr int kMaxTermsAll = 32;  // terms over all monitors of one call

struct GatherTerm {
  const void* a;
  const void* b;
  void* out;        // staging base for this term (slot stride = n) or probe stack
  long long s0, s1;
  int n0, n1;
  long long slot_stride;  // elements between consecutive slots
};

struct GatherArgs {
  int n_terms;
  int prefix[kMaxTermsAll + 1];
  GatherTerm t[kMaxTermsAll];
};

template <typename T>
__global__ void gather_kernel(GatherArgs g, int slot) {
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.prefix[g.n_terms]) return;
  int q = 0;
  while (tid >= g.prefix[q + 1]) ++q;
  const GatherTerm& t = g.t[q];
  int e = tid - g.prefix[q];
  int i0 = e / t.n1, i1 = e - i0 * t.n1;
  long long off = (long long)i0 * t.s0 + (long long)i1 * t.s1;
  T v = static_cast<const T*>(t.a)[off];
  if (t.b) v = T(0.5) * (v + static_cast<const T*>(t.b)[off]);
  static_cast<T*>(t.out)[(long long)slot * t.slot_stride + e] = v;
}

// table[(f * S + s) * 4 + {0,1,2,3}] = cos(w tE), sin(w tE), cos(w tH), sin(w tH)
template <typename T>
__global__ void phase_kernel(const T* omega, int nfreq, int n_filled, int t0, double dt, T* table) {
  int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nfreq * n_filled) return;
  int f = tid / n_filled, s = tid - f * n_filled;
  double tE = ((double)(t0 + s) + 1.0) * dt;
  double tH = ((double)(t0 + s) + 0.5) * dt;
  T w = omega[f];
  T pe = w * static_cast<T>(tE);
  T ph = w * static_cast<T>(tH);
  T* o = table + (long long)(f * n_filled + s) * 4;
  o[0] = cos(pe);
  o[1] = sin(pe);
  o[2] = cos(ph);
  o[3] = sin(ph);
}

// One thread per (element, term) and group of G frequencies: the
// accumulator pairs are read and written once per flush, each staged sample
// is read once per group (coalesced along the element axis) and feeds G
// independent accumulations, and the phase rows are uniform across the
// block.  Every accumulator still adds its steps one by one in step order,
// so the result does not depend on the grouping.  G = 4 for large monitors
// (3D planes: the staged samples, re-read once per group, are far larger
// than L2); G = 1 where that would leave too few threads.  Accumulator
// layout (nfreq, 2*n_terms, n).
template <typename T, int G>
__global__ void flush_kernel(const T* stage, const T* table, T* acc, int n_terms, int n, int nfreq,
                             int n_filled, int half_mask, T dt) {
  constexpr int kFlushF = G;
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int f0 = blockIdx.y * kFlushF;
  int q = blockIdx.z;
  if (e >= n) return;
  const int nf = nfreq - f0 < kFlushF ? nfreq - f0 : kFlushF;   // uniform over the block
  const T* st = stage + (long long)q * n + e;  // slot stride = n_terms * n
  long long slot_stride = (long long)n_terms * n;
  int toff = ((half_mask >> q) & 1) ? 2 : 0;
  T re[kFlushF], im[kFlushF];
#pragma unroll
  for (int g = 0; g < kFlushF; ++g) {
    if (g < nf) {
      const T* a = acc + ((long long)(f0 + g) * 2 * n_terms + 2 * q) * n + e;
      re[g] = a[0];
      im[g] = a[n];
    }
  }
  for (int s = 0; s < n_filled; ++s) {
    T v = st[(long long)s * slot_stride];
#pragma unroll
    for (int g = 0; g < kFlushF; ++g) {
      if (g < nf) {
        const T* tb = table + ((long long)(f0 + g) * n_filled + s) * 4 + toff;
        T c = tb[0];
        T sn = tb[1];
        re[g] = re[g] + dt * (c * v);
        im[g] = im[g] + dt * ((-sn) * v);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < kFlushF; ++g) {
    if (g < nf) {
      T* a = acc + ((long long)(f0 + g) * 2 * n_terms + 2 * q) * n + e;
      a[0] = re[g];
      a[n] = im[g];
    }
  }
}

template <typename T>
size_t esz() { return sizeof(T); }

long long mon_elems(const mt_monitor& m) { return (long long)m.n0 * m.n1; }

}  // namespace

int monitors_validate(const MonitorPlan& mp) {
  int total = 0;
  for (int i = 0; i < mp.n_mons; ++i) {
    const mt_monitor& m = mp.mons[i];
    if (m.kind != 0 && m.kind != 1) return fail(MT_EINVAL, "monitor: unknown kind");
    if (m.n_terms < 1 || m.n_terms > MT_MAX_TERMS) return fail(MT_EINVAL, "monitor: bad n_terms");
    if (m.n0 < 0 || m.n1 < 0) return fail(MT_EINVAL, "monitor: bad extents");
    if (!m.acc) return fail(MT_EINVAL, "monitor: null accumulator");
    if (m.kind == 0 && (m.nfreq < 1 || !m.omega)) return fail(MT_EINVAL, "monitor: no frequencies");
    if (m.kind == 1 && (m.n_terms != 1 || m.n0 * m.n1 != 1))
      return fail(MT_EINVAL, "probe monitor must have one single-element term");
    for (int q = 0; q < m.n_terms; ++q)
      if (!m.term[q].a) return fail(MT_EINVAL, "monitor: null field view");
    total += m.n_terms;
  }
  if (total > kMaxTermsAll) return fail(MT_EINVAL, "too many monitor terms in one call (max 32)");
  long long need = mt_stage_bytes(mp.mons, mp.n_mons, mp.stage_steps, mp.dtype);
  if (need > mp.stage_bytes) return fail(MT_EINVAL, "staging workspace too small");
  if (need > 0 && !mp.stage) return fail(MT_EINVAL, "staging workspace missing");
  return MT_OK;
}

template <typename T>
static int sample_t(const MonitorPlan& mp, int slot, int t_rel, hipStream_t s) {
  GatherArgs g{};
  int nt = 0;
  char* base = static_cast<char*>(mp.stage);
  long long off = 0;
  for (int i = 0; i < mp.n_mons; ++i) {
    const mt_monitor& m = mp.mons[i];
    long long n = mon_elems(m);
    for (int q = 0; q < m.n_terms; ++q) {
      GatherTerm& t = g.t[nt];
      t.a = m.term[q].a;
      t.b = m.term[q].b;
      t.s0 = m.term[q].s0;
      t.s1 = m.term[q].s1;
      t.n0 = m.n0;
      t.n1 = m.n1;
      if (m.kind == 1) {
        t.out = static_cast<T*>(m.acc);
        t.slot_stride = 1;
      } else {
        // staging layout per monitor: [slot][term][elem]
        t.out = reinterpret_cast<T*>(base + off) + (long long)q * n;
        t.slot_stride = (long long)m.n_terms * n;
      }
      g.prefix[nt + 1] = g.prefix[nt] + (int)n;
      ++nt;
    }
    // staging layout per DFT monitor: [slot][term][elem] then its phase table
    if (m.kind == 0)
      off += ((long long)mp.stage_steps * m.n_terms * n + (long long)m.nfreq * mp.stage_steps * 4) * sizeof(T);
  }
  g.n_terms = nt;
  int total = g.prefix[nt];
  if (total == 0) return MT_OK;
  // probes write at the call-relative step index, DFT terms at the ring slot
  bool any_probe = false;
  for (int i = 0; i < mp.n_mons; ++i) any_probe |= mp.mons[i].kind == 1;
  if (!any_probe) {
    hipLaunchKernelGGL(gather_kernel<T>, dim3((total + 255) / 256), dim3(256), 0, s, g, slot);
    return check_launch("monitor gather");
  }
  // probes and DFT terms use different slot indices: launch the two sets apart
  GatherArgs gp{}, gd{};
  int np = 0, nd = 0;
  int ti = 0;
  for (int i = 0; i < mp.n_mons; ++i) {
    const mt_monitor& m = mp.mons[i];
    for (int q = 0; q < m.n_terms; ++q, ++ti) {
      GatherArgs& dst = m.kind == 1 ? gp : gd;
      int& c = m.kind == 1 ? np : nd;
      dst.t[c] = g.t[ti];
      dst.prefix[c + 1] = dst.prefix[c] + (g.prefix[ti + 1] - g.prefix[ti]);
      ++c;
    }
  }
  gp.n_terms = np;
  gd.n_terms = nd;
  if (nd) {
    int tot = gd.prefix[nd];
    if (tot) hipLaunchKernelGGL(gather_kernel<T>, dim3((tot + 255) / 256), dim3(256), 0, s, gd, slot);
  }
  if (np) hipLaunchKernelGGL(gather_kernel<T>, dim3((gp.prefix[np] + 255) / 256), dim3(256), 0, s, gp, t_rel);
  return check_launch("monitor gather");
}

bool build_taps(const MonitorPlan& mp, int slot0, int t_rel0, int esz, TapSet* out) {
  out->n = 0;
  char* base = static_cast<char*>(mp.stage);
  long long off = 0;
  for (int i = 0; i < mp.n_mons; ++i) {
    const mt_monitor& m = mp.mons[i];
    long long n = mon_elems(m);
    for (int q = 0; q < m.n_terms; ++q) {
      const mt_dft_term& tm = m.term[q];
      if (tm.line < 1 || tm.line > 4 || out->n >= kMaxTaps) return false;
      Tap& t = out->t[out->n++];
      t.field = tm.field;
      t.line = tm.line;
      t.fix = tm.fix;
      t.lo = tm.lo;
      t.n = (int)n;
      t.avg = tm.b != nullptr;
      t.n0 = tm.line == 3 ? tm.fix : 0;  // whole field: fix = rows, lo = row length
      t.n1 = tm.line == 3 ? tm.lo : 0;
      if (m.kind == 1) {
        t.out = m.acc;
        t.slot_stride = 1;
        t.slot0 = t_rel0;
      } else {
        t.out = base + off + (long long)q * n * esz;
        t.slot_stride = (long long)m.n_terms * n;
        t.slot0 = slot0;
      }
    }
    if (m.kind == 0)
      off += ((long long)mp.stage_steps * m.n_terms * n + (long long)m.nfreq * mp.stage_steps * 4) * esz;
  }
  return true;
}

int monitors_sample(const MonitorPlan& mp, int slot, int t_rel, hipStream_t s) {
  if (mp.n_mons == 0) return MT_OK;
  return mp.dtype == MT_F64 ? sample_t<double>(mp, slot, t_rel, s) : sample_t<float>(mp, slot, t_rel, s);
}

template <typename T>
static int flush_t(const MonitorPlan& mp, int n_filled, int t0, double dt, hipStream_t s) {
  char* base = static_cast<char*>(mp.stage);
  long long off = 0;
  for (int i = 0; i < mp.n_mons; ++i) {
    const mt_monitor& m = mp.mons[i];
    if (m.kind != 0) continue;
    long long n = mon_elems(m);
    T* stage = reinterpret_cast<T*>(base + off);
    off += (long long)mp.stage_steps * m.n_terms * n * sizeof(T);
    T* table = reinterpret_cast<T*>(base + off);
    off += (long long)m.nfreq * mp.stage_steps * 4 * sizeof(T);
    if (n == 0) continue;
    int nt = m.nfreq * n_filled;
    hipLaunchKernelGGL(phase_kernel<T>, dim3((nt + 255) / 256), dim3(256), 0, s,
                       static_cast<const T*>(m.omega), m.nfreq, n_filled, t0, dt, table);
    int half_mask = 0;
    for (int q = 0; q < m.n_terms; ++q) half_mask |= (m.term[q].t_half ? 1 : 0) << q;
    int blk = n >= 256 ? 256 : 64;
    const bool grouped = (long long)n * m.n_terms * ((m.nfreq + 3) / 4) >= (1LL << 18);
    const int G = grouped ? 4 : 1;
    dim3 grid((unsigned)((n + blk - 1) / blk), (unsigned)((m.nfreq + G - 1) / G), (unsigned)m.n_terms);
    if (grouped)
      hipLaunchKernelGGL((flush_kernel<T, 4>), grid, dim3(blk), 0, s, (const T*)stage, (const T*)table,
                         static_cast<T*>(m.acc), m.n_terms, (int)n, m.nfreq, n_filled, half_mask,
                         static_cast<T>(dt));
    else
      hipLaunchKernelGGL((flush_kernel<T, 1>), grid, dim3(blk), 0, s, (const T*)stage, (const T*)table,
                         static_cast<T*>(m.acc), m.n_terms, (int)n, m.nfreq, n_filled, half_mask,
                         static_cast<T>(dt));
    int rc = check_launch("monitor flush");
    if (rc) return rc;
  }
  return MT_OK;
}

int monitors_flush(const MonitorPlan& mp, int n_filled, int t0, double dt, hipStream_t s) {
  if (mp.n_mons == 0 || n_filled <= 0) return MT_OK;
  return mp.dtype == MT_F64 ? flush_t<double>(mp, n_filled, t0, dt, s)
                            : flush_t<float>(mp, n_filled, t0, dt, s);
}

namespace {
struct TimingRec {
  hipEvent_t a, b;
  long long cells;  // -1 while open
};
std::vector<TimingRec> g_rec;
std::vector<hipEvent_t> g_pool;
int g_timing = 0;  // bit 0: forward stencil launches, bit 1: batched adjoint steps
constexpr size_t kMaxTimingRecs = 1 << 16;

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

void timing_begin(hipStream_t s, int kind) {
  if (!(g_timing & kind) || g_rec.size() >= kMaxTimingRecs) return;
  TimingRec r{take_event(), take_event(), -1};
  if (!r.a || !r.b) return;
  if (hipEventRecord(r.a, s) != hipSuccess) return;
  g_rec.push_back(r);
}

void timing_end(hipStream_t s, long long cell_steps, int kind) {
  if (!(g_timing & kind) || g_rec.empty() || g_rec.back().cells != -1) return;
  TimingRec& r = g_rec.back();
  if (hipEventRecord(r.b, s) != hipSuccess) {
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
    g_rec.pop_back();
    return;
  }
  r.cells = cell_steps;
}

}  // namespace mt

extern "C" int mt_kernel_timing_enable(int on) {
  mt::g_timing = on < 0 ? 0 : on;
  return MT_OK;
}

extern "C" int mt_kernel_timing_read(double* total_ms, long long* launches, long long* cell_steps) {
  double ms = 0.0;
  long long n = 0, cells = 0;
  for (auto& r : mt::g_rec) {
    if (r.cells >= 0) {
      float t = 0.f;
      if (hipEventSynchronize(r.b) != hipSuccess || hipEventElapsedTime(&t, r.a, r.b) != hipSuccess)
        return mt::fail(MT_EHIP, "kernel timing: event query failed");
      ms += t;
      ++n;
      cells += r.cells;
    }
    mt::g_pool.push_back(r.a);
    mt::g_pool.push_back(r.b);
  }
  mt::g_rec.clear();
  if (total_ms) *total_ms = ms;
  if (launches) *launches = n;
  if (cell_steps) *cell_steps = cells;
  return MT_OK;
}

extern "C" long long mt_stage_bytes(const mt_monitor* mons, int n_mons, int stage_steps, int dtype) {
  long long es = dtype == MT_F64 ? 8 : 4;
  long long total = 0;
  for (int i = 0; i < n_mons; ++i) {
    const mt_monitor& m = mons[i];
    if (m.kind != 0) continue;
    long long n = (long long)m.n0 * m.n1;
    total += (long long)stage_steps * m.n_terms * n * es;
    total += (long long)m.nfreq * stage_steps * 4 * es;
  }
  return total;
}

extern "C" const char* mt_last_error(void) { return mt::g_err.c_str(); }

extern "C" int mt_abi_version(void) { return 27; }

extern "C" int mt_abi_layout(long long* out, int n) {
  const long long v[] = {
      (long long)sizeof(mt_pml_axis),        (long long)sizeof(mt_source2d),   (long long)sizeof(mt_tm2d),
      (long long)sizeof(mt_te2d),            (long long)sizeof(mt_yee3d),      (long long)sizeof(mt_source3d),
      (long long)sizeof(mt_dft_term),        (long long)sizeof(mt_monitor),    (long long)sizeof(mt_tm2d_adj),
      (long long)sizeof(mt_te2d_adj),        (long long)sizeof(mt_yee3d_adj),  (long long)sizeof(mt_tm2d_batch_item),
      (long long)offsetof(mt_tm2d, hsrc),    (long long)offsetof(mt_tm2d, hcoef), (long long)offsetof(mt_te2d, tape),
      (long long)offsetof(mt_te2d, hsrc),    (long long)offsetof(mt_te2d, hcoef), (long long)offsetof(mt_yee3d, alt),
      (long long)offsetof(mt_monitor, acc),  (long long)offsetof(mt_tm2d, n_poles), (long long)offsetof(mt_tm2d, h2y),
      (long long)offsetof(mt_te2d, npx),     (long long)offsetof(mt_te2d, hz2),
      (long long)offsetof(mt_te2d, ky),      (long long)offsetof(mt_yee3d, damp), (long long)offsetof(mt_yee3d, P_prev)};
  const int total = (int)(sizeof(v) / sizeof(v[0]));
  for (int q = 0; q < n && q < total; ++q) out[q] = v[q];
  return total;
}
