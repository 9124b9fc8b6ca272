// pmx: C-ABI implementation (include/pmx.h) -- host orchestration of the
// gfx950 kernels in pmx_kernels.h.  One HIP stream per context; all device
// state of a system stays resident between calls.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/pmx.h"
#include "pmx_kernels.h"
#include "pmx_batch.h"

using namespace pmx;

struct pmx_ctx {
  int device = 0;
  int cus = 256;  // compute units of the device
  void *scratch = nullptr;  // device address of this device's copy of `scratch` (no-allocation force pass)
  hipStream_t stream = nullptr;   // main stream: sort, list, pair kernel, boundary pairs, finalize
  hipStream_t stream2 = nullptr;  // PME, exclusion correction and bonded terms, concurrent with the pair kernel
  // batched launches: pinned host staging (grown as needed) and the fork/join
  // events, kept from one batch to the next
  void *batch_host = nullptr;
  size_t batch_host_bytes = 0;
  hipEvent_t batch_fork = nullptr, batch_join = nullptr;
  std::string err;
};

// Scratch of pmx_sys_compute_noalloc: the per-pass partial results, counters
// and records of a force pass (per-tile energies and pair statistics, the
// boundary-pair records and their counters) live here instead of in the
// system's own buffers.  Each device holds one copy for the life of the
// process, shared by every call on that device.  In host code the name is a
// stand-in, not a device address: host code uses only the address
// hipGetSymbolAddress returns for the current device (pmx_ctx::scratch).
constexpr size_t kScratchBytes = 32ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Device address of `scratch` per device, looked up once per device (with
// that device current) by the first pmx_ctx_create on it.
constexpr int kMaxDevices = 64;
void *g_scratch_addr[kMaxDevices] = {};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  bool listed = false;  // registered in Sys::all (freed by pmx_sys_destroy)
};

// Timed stages: sort (+ integrate), list (+ accumulator zeroing), pair,
// spread, fft (+ convolution), gather, fixups (exclusion correction, bonded
// terms), finalize (boundary pairs, force unsort).  Events: 0 start, 1 sorted,
// 2 list built and accumulators zeroed, 4 spread done = fork (main stream), 3
// pair done (main stream), 5 fft, 6 gather, 7 correction + bonded done
// (second stream), 8 end.  The second stream's stages run beside the pair
// kernel.
constexpr int kStages = 8;
constexpr int kStageEv[kStages][2] = {{0, 1}, {1, 2}, {4, 3}, {2, 4}, {4, 5}, {5, 6}, {6, 7}, {7, 8}};
static_assert(kStages == PMX_STAGES, "stage count of the C ABI");

struct Sys {
  pmx_ctx *ctx = nullptr;
  pmx_sys_desc d{};
  bool mixed = false;
  int n = 0, npad = 0, ncl = 0, nsuper = 0, ncols = 0;  // nsuper: upper bound (slots = 64 nsuper)
  // host copies
  std::vector<double> q, lj;
  std::vector<float> lj32;         // FP32 copy of lj (source of ljt32)
  std::vector<int> type;
  std::vector<int4> etab;          // de-duplicated exclusion table a1 a2 cls
  std::vector<int> eptr;           // CSR of partners per atom (orig index)
  std::vector<int2> elist;
  double alpha = 0, self_e = 0, charged_e = 0;
  double qpair = 1.0;              // charge factor of the pair pass (0: PME real term off)
  Box box{};
  BinParams bin{};
  PairParams pp{};
  ListParams lp{};
  PmeParams pme{};
  FftPlan plx{}, ply{}, plz{};
  bool pme_on = false;
  int cap = 0, maxmaskwork = 0, ntiles_cap = 0;
  // launch shapes the autotuner may change (pmx_sys_set_tuning)
  int tune_fft_lines = 16;     // most lines per FFT block
  int tune_pair_blocks = 0;    // pair-kernel blocks per CU (0: as many as fit)
  int tune_schedule = 0;       // 0: runs round robin over the waves; 1: later runs from a counter
  int tune_spread_stream2 = 0; // 1: the PME spread runs on the second stream, beside the pair kernel
  int tune_fixup_main = 0;     // 1: exclusion correction and bonded terms on the main stream after the pair kernel
                               // (default 1 below kFixupMainBelow atoms)
  int tune_fused_spread = 1;   // 1: inside an MD chunk the next pass's spread joins the kick-drift
  int tune_fft_planes = 1;     // 1: y, x and convolution passes in one launch per plane on small planes
  bool fft_plane = false;      // the plane launch is in use
  int border_blocks = 4;       // blocks of the boundary-pair pass
  int tune_prune = 1;          // 1: the first pair pass after a list build prunes the list's cluster pairs
  bool prune_pending = false;  // a list was built and its first pair pass has not run
  int pair_grid = 0;           // blocks of the pair-kernel launch
  long long nbins = 0;
  int nscanblocks = 0;
  int recip_blocks = 0, corr_blocks = 0;
  int zl = 1, tky = 1, tkx = 1;
  // device buffers
  std::vector<DevBuf *> all;
  DevBuf xorig, xwrap, q_d, type_d, key, hist, start, cursor, scansum, perm, permp, slot, refcell, xs, xs64, ts,
      perm2, sc_info, sc_col, colsc, sctr, shalf, sorg, ctr, half, entries, count, ecost, lcost, tep, tcp, tpos, tposS, wstart, maskwork, nmaskwork, masks, overflow, tiles, ntiles, tilectr, ljtab, ljt32, facc, fout,
      pair_e, npairs, hash, eptr_d, elist_d, etab_d, corr_e, recip_e, eout, np_out, hash_out, grid64, gridfx, spec, spec2,
      greal, theta, twx, twy, twz, border, nborder, border_acc;
  int maxborder = 65536;
  // results
  long long last_np = 0;
  unsigned long long last_hash = 0;
  float stage_ms[kStages] = {0};
  hipEvent_t *ev = nullptr;                    // event set of the current timed call
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;  // main stream <-> second stream per force pass
  std::vector<std::vector<hipEvent_t>> evsets;  // one set of kStages + 1 per timed call since timing_begin
  size_t evused = 0;
  bool timing_active = false;                   // MD steps record events (timing_begin .. timing_end)
  bool have_positions = false;
  bool cached = false;
  bool list_valid = false;             // entries, counts and work of a built list are in place
  bool acc_clean = false;              // force accumulators, PME grid and pass counters are zero
  bool spread_ready = false;           // the grid holds the next pass's spread (fused into the last kick-drift)
  bool dry = false;                    // size only: allocations are counted, nothing runs
  unsigned long long dry_bytes = 0;
  double eout_h[4] = {0, 0, 0, 0};
  // device-resident MD state
  struct Md {
    bool on = false;
    int nb = 0, na = 0;
    double h = 0, skin = 0;
    DevBuf bonds, bpar, angles, apar, minv, v, xbuild, xbak, vbak, fbak, dmax, bond_e, ke_e;
    int chunk = 16;
    long long steps = 0, rebuilds = 0, rollbacks = 0;
    bool forces_valid = false;
    int integrator = PMX_MD_LEAPFROG;
    double vdt = 1, fdt = 0, ndt = 0, fc = 0;  // Langevin half-kick constants
    unsigned long long seed = 0;
    // two chunks in flight: backups, displacement records and their pinned
    // host copies (per slot: kMaxChunk displacements + the overflow flag)
    unsigned *host = nullptr;
    hipEvent_t chunk_ev[2] = {nullptr, nullptr};
    // chunks captured as HIP graphs, keyed by (steps, first, slot); the
    // chunk's first step index is written to step_dev before each launch
    bool graphs = false;
    DevBuf step_dev;
    std::map<long long, hipGraphExec_t> gexec;
  } md;
};
constexpr int kMaxChunk = 128;
constexpr int kFixupMainBelow = 100000;  // atoms; default of PMX_TUNE_FIXUP_STREAM
constexpr double kInvTimeFactor = 0.02045482828087295;  // protomol/base/PMConstants.cpp:54
constexpr double kBoltzmann = 0.001987191;              // protomol/base/PMConstants.cpp:70

int fail(pmx_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

#define HIPCHK(expr)                                                                        \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) return fail(s->ctx, PMX_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

// Runtime calls that may synchronize the device (allocation, free, pinned
// allocation, event and stream creation, synchronous copies) can invalidate a
// stream capture that another thread has in progress, so they and the
// captures themselves are serialized on this lock.
std::recursive_mutex g_rt_mu;
using RtLock = std::lock_guard<std::recursive_mutex>;

// Device block cache.  Systems come and go (a request loop builds one per
// request), so released device blocks are kept per device in size classes
// (powers of 2 split in quarters above 4 KiB) and handed to the next request
// of the same class instead of going back to hipFree / hipMalloc.  A block is
// released only after the streams that used it are synchronized.  When
// hipMalloc fails, the cached blocks of that device are freed and the
// allocation is tried once more.  Guarded by g_rt_mu.
size_t size_class(size_t bytes) {
  if (bytes <= 4096) return bytes <= 256 ? 256 : (size_t)1 << (64 - __builtin_clzll(bytes - 1));
  const int e = 63 - __builtin_clzll(bytes);  // 2^e <= bytes < 2^(e+1)
  const size_t q = (size_t)1 << (e - 2);
  return (bytes + q - 1) / q * q;
}
std::map<int, std::multimap<size_t, void *>> g_block_cache;
size_t g_cached_bytes = 0;
constexpr size_t kBlockCacheMax = 16ull << 30;  // cached bytes kept per process, at most

void *block_get(int device, size_t cls) {
  auto &m = g_block_cache[device];
  auto it = m.find(cls);
  if (it != m.end()) {
    void *p = it->second;
    m.erase(it);
    g_cached_bytes -= cls;
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, cls) == hipSuccess) return p;
  for (auto &kv : m) hipFree(kv.second), g_cached_bytes -= kv.first;
  m.clear();
  p = nullptr;
  return hipMalloc(&p, cls) == hipSuccess ? p : nullptr;
}

void block_put(int device, void *p, size_t cls) {
  if (!p) return;
  if (g_cached_bytes + cls > kBlockCacheMax) {
    hipFree(p);
    return;
  }
  g_block_cache[device].emplace(cls, p);
  g_cached_bytes += cls;
}

int dalloc(Sys *s, DevBuf &b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (s->dry) {
    s->dry_bytes += size_class(bytes);
    return PMX_OK;
  }
  RtLock lk(g_rt_mu);
  if (b.p && b.bytes >= bytes) return PMX_OK;
  if (b.p) {
    // a growing buffer: its last users must be done before the block is reused
    hipStreamSynchronize(s->ctx->stream);
    hipStreamSynchronize(s->ctx->stream2);
    block_put(s->ctx->device, b.p, b.bytes);
  }
  const size_t cls = size_class(bytes);
  b.p = block_get(s->ctx->device, cls);
  if (!b.p) return fail(s->ctx, PMX_ERR_OOM, "hipMalloc failed (" + std::to_string(cls) + " bytes)");
  b.bytes = cls;
  if (!b.listed) s->all.push_back(&b);
  b.listed = true;
  return PMX_OK;
}

template <class T>
T *P(DevBuf &b) {
  return reinterpret_cast<T *>(b.p);
}

unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Abramowitz-free alpha choice (NonbondedPMEwaldSystemForce.h:520-537).
double choose_alpha(double rc, double acc) {
  double a = 1.0;
  while (std::erfc(a * rc) / rc >= acc) a *= 2.0;
  double lo = 0., hi = a;
  for (int i = 0; i < 100; ++i) {
    a = 0.5 * (lo + hi);
    if (std::erfc(a * rc) / rc >= acc)
      lo = a;
    else
      hi = a;
  }
  return a;
}

bool make_plan(int n, FftPlan &pl) {
  pl.n = n;
  pl.nrad = 0;
  int m = n;
  while (m % 4 == 0 && pl.nrad < 16) {
    pl.rad[pl.nrad++] = 4;
    m /= 4;
  }
  for (int r : {2, 3, 5})
    while (m % r == 0 && pl.nrad < 16) {
      pl.rad[pl.nrad++] = r;
      m /= r;
    }
  return m == 1;
}

void bspline_host(int order, double w, double *th, double *dth) {
  th[order - 1] = 0.0;
  th[1] = w;
  th[0] = 1.0 - w;
  for (int k = 2; k < order - 1; k++) {
    double div = 1.0 / k;
    th[k] = div * w * th[k - 1];
    for (int j = 1; j < k; j++) th[k - j] = div * ((w + j) * th[k - j - 1] + (k + 1 - j - w) * th[k - j]);
    th[0] = div * (1 - w) * th[0];
  }
  dth[0] = -th[0];
  for (int j = 1; j <= order - 1; j++) dth[j] = th[j - 1] - th[j];
  double div = 1.0 / (order - 1.0);
  th[order - 1] = div * w * th[order - 2];
  for (int j = 1; j < order - 1; j++)
    th[order - 1 - j] = div * ((w + j) * th[order - j - 2] + (order - j - w) * th[order - 1 - j]);
  th[0] = div * (1 - w) * th[0];
}

SwitchC make_switch(int kind, double on, double rc) {
  SwitchC s{};
  s.kind = kind;
  s.rc2 = rc * rc;
  s.on2 = on * on;
  double t = rc * rc - on * on;
  s.s1 = kind == SW_C2 ? 1.0 / (t * t * t) : 0.0;
  s.s2 = rc * rc - 3.0 * on * on;
  s.s3 = kind == SW_C2 ? 4.0 / (t * t * t) : 0.0;
  s.c15_1 = 1.5 / rc;
  s.c05_3 = 0.5 / (rc * rc * rc);
  s.c15_3 = 1.5 / (rc * rc * rc);
  return s;
}

int validate(pmx_ctx *c, const pmx_sys_desc *d) {
  if (!d) return fail(c, PMX_ERR_ARG, "null descriptor");
  if (d->n <= 0 || d->ntypes <= 0 || d->nexcl < 0) return fail(c, PMX_ERR_ARG, "bad sizes");
  if (!d->charge || !d->type || !d->lj || (d->nexcl > 0 && !d->excl)) return fail(c, PMX_ERR_ARG, "null array");
  if (d->ntypes > kMaxTypesTotal) return fail(c, PMX_ERR_UNSUPPORTED, "more than 4096 LJ types");
  for (int k = 0; k < 3; ++k)
    if (!(d->box[k] > 0)) return fail(c, PMX_ERR_ARG, "box edge must be > 0");
  if (d->lj_switch < 0 || d->lj_switch > 3 || d->coul_kind < 0 || d->coul_kind > 2)
    return fail(c, PMX_ERR_ARG, "bad force kind");
  if (d->lj_switch == SW_NONE && d->coul_kind == COUL_NONE) return fail(c, PMX_ERR_ARG, "no force selected");
  if (d->lj_switch != SW_NONE && !(d->lj_rc > 0)) return fail(c, PMX_ERR_ARG, "LJ cutoff must be > 0");
  if (d->lj_switch == SW_C2 && !(d->lj_on >= 0 && d->lj_on < d->lj_rc))
    return fail(c, PMX_ERR_ARG, "C2 switching function: 0 <= switchon < cutoff");
  if (d->coul_kind != COUL_NONE && !(d->coul_rc > 0)) return fail(c, PMX_ERR_ARG, "Coulomb cutoff must be > 0");
  if (d->coul_kind == COUL_CUTOFF && d->coul_switch == SW_C2 && !(d->coul_on >= 0 && d->coul_on < d->coul_rc))
    return fail(c, PMX_ERR_ARG, "C2 switching function: 0 <= switchon < cutoff");
  if (d->coul_kind == COUL_CUTOFF && (d->coul_switch < 1 || d->coul_switch > 3))
    return fail(c, PMX_ERR_ARG, "bad Coulomb switching function");
  if (d->coul_kind == COUL_PME) {
    if (d->coul_switch != SW_CUTOFF && d->coul_switch != SW_C1)
      return fail(c, PMX_ERR_ARG, "PME real term takes the Cutoff or C1 switching function");
    const int p = d->pme_order;
    if (p < 2 || p % 2 != 0 || p > 8) return fail(c, PMX_ERR_ARG, "PMEwald order must be even, 2..8");
    for (int k = 0; k < 3; ++k)
      if (d->pme_grid[k] < p) return fail(c, PMX_ERR_ARG, "PMEwald grid must be >= order");
    if (!(d->pme_accuracy > 0) && !(d->pme_alpha > 0)) return fail(c, PMX_ERR_ARG, "PMEwald accuracy > 0");
    if ((d->pme_terms & 7) == 0) return fail(c, PMX_ERR_ARG, "no PME term selected");
  }
  if (!(d->cellsize > 0)) return fail(c, PMX_ERR_ARG, "cellsize must be > 0");
  if (d->precision != PMX_FP64 && d->precision != PMX_MIXED) return fail(c, PMX_ERR_ARG, "bad precision");
  double rcmax = std::max(d->lj_switch ? d->lj_rc : 0.0, d->coul_kind ? d->coul_rc : 0.0);
  for (int k = 0; k < 3; ++k)
    if (rcmax > 0.5 * d->box[k])
      return fail(c, PMX_ERR_UNSUPPORTED, "cutoff larger than half the box (multiple images per pair)");
  for (int i = 0; i < d->n; ++i)
    if (d->type[i] < 0 || d->type[i] >= d->ntypes) return fail(c, PMX_ERR_ARG, "atom type out of range");
  for (int e = 0; e < d->nexcl; ++e) {
    int a = d->excl[3 * e], b = d->excl[3 * e + 1], cl = d->excl[3 * e + 2];
    if (a < 0 || b < 0 || a >= d->n || b >= d->n || cl < 0 || cl > 2)
      return fail(c, PMX_ERR_ARG, "exclusion entry out of range");
  }
  return PMX_OK;
}

// FFT blocking: lines per block, two LDS buffers of lines * n complex
// within 24 KiB (so that FFT blocks fit beside the pair kernel's blocks,
// which leave about that much LDS free on each CU), at most
// tune_fft_lines lines (the transposed segments); one energy partial per
// x-pass block.
int fft_blocking(Sys *s) {
  const PmeParams &pm = s->pme;
  const int cpx = s->mixed ? 8 : 16;
  const int budget = 24576 / (2 * cpx);
  const int lines = s->tune_fft_lines;
  s->zl = std::max(1, std::min(lines, budget / pm.nz));
  s->tky = std::max(1, std::min(lines, budget / pm.ny));
  s->tkx = std::max(1, std::min(lines, budget / pm.nx));
  s->fft_plane = s->tune_fft_planes && pm.nx * pm.ny <= kFftPlanePoints;
  s->recip_blocks = s->fft_plane ? pm.nzh : pm.nzh * ((pm.ny + s->tkx - 1) / s->tkx);
  return dalloc(s, s->recip_e, 8 * (size_t)s->recip_blocks);
}

template <bool MIXED>
void run_pair_phase(Sys *s, bool stats, int phase);

int setup(Sys *s) {
  const pmx_sys_desc &d = s->d;
  s->acc_clean = false;  // buffers are (re)allocated below
  s->list_valid = false;
  s->mixed = d.precision == PMX_MIXED;
  s->n = d.n;
  s->q.assign(d.charge, d.charge + d.n);
  s->type.assign(d.type, d.type + d.n);
  s->lj.assign(d.lj, d.lj + (size_t)d.ntypes * d.ntypes * 4);

  // Exclusions (ExclusionTable.cpp:103-120): the correction table holds each
  // pair once, in insertion order, with the class it was first added with
  // (its filter set orders pairs on (a1, a2) only, ExclusionTable.cpp:20-27);
  // the lookup class of a pair is the last one added.
  std::set<std::pair<int, int>> seen;
  std::map<std::pair<int, int>, int> cls;
  s->etab.clear();
  for (int e = 0; e < d.nexcl; ++e) {
    int a = d.excl[3 * e], b = d.excl[3 * e + 1], c = d.excl[3 * e + 2];
    if (a > b) std::swap(a, b);
    if (c != 0 && a != b && !seen.count({a, b})) {
      s->etab.push_back(make_int4(a, b, c, 0));
      seen.insert({a, b});
    }
    if (a != b) cls[{a, b}] = c;
  }
  std::vector<std::vector<int2>> part(d.n);
  for (auto &kv : cls) {
    if (kv.second == 0) continue;
    part[kv.first.first].push_back(make_int2(kv.first.second, kv.second));
    part[kv.first.second].push_back(make_int2(kv.first.first, kv.second));
  }
  s->eptr.assign(d.n + 1, 0);
  s->elist.clear();
  for (int a = 0; a < d.n; ++a) {
    std::sort(part[a].begin(), part[a].end(), [](int2 x, int2 y) { return x.x < y.x; });
    s->eptr[a] = (int)s->elist.size();
    for (auto &v : part[a]) s->elist.push_back(v);
  }
  s->eptr[d.n] = (int)s->elist.size();

  // Box (PeriodicBoundaryConditions.cpp:32-77)
  Box &bx = s->box;
  for (int k = 0; k < 3; ++k) {
    bx.L[k] = d.box[k];
    bx.Linv[k] = 1.0 / d.box[k];
    bx.o[k] = d.origin[k];
    double a = d.origin[k] - d.box[k] * 0.5, b = d.origin[k] + d.box[k] * 0.5;
    bx.mn[k] = std::min(a, b);
    double mx = std::max(a, b);
    bx.H[k] = mx - bx.mn[k];
    bx.H2[k] = bx.H[k] * 0.5;
  }
  bx.Lr[0] = (bx.L[1] * bx.L[2]) / (bx.L[0] * (bx.L[1] * bx.L[2]));
  bx.Lr[1] = (bx.L[2] * bx.L[0]) / (bx.L[1] * (bx.L[2] * bx.L[0]));
  bx.Lr[2] = (bx.L[0] * bx.L[1]) / (bx.L[2] * (bx.L[0] * bx.L[1]));
  {
    double dx = (bx.L[0] * 0.5) * (bx.L[0] * 0.5), dy = (bx.L[1] * 0.5) * (bx.L[1] * 0.5),
           dz = (bx.L[2] * 0.5) * (bx.L[2] * 0.5);
    bx.D = std::min(dx, std::min(dy, dz));
  }

  // Binning parameters: x-y columns about one supercluster (64 atoms) wide
  // and deep, z bins of about two atoms per column; reference cell grid
  // (CubicCellManager.cpp:46-65, ArrayCellListStructure.cpp:40-42).
  BinParams &bp = s->bin;
  bp.box = bx;
  const double vol = bx.L[0] * bx.L[1] * bx.L[2];
  const double rho = std::max(1, d.n) / vol;
  const double a_sc = std::cbrt(kSuperSize / rho);
  bp.ncx = (int)std::max(1L, std::min(1024L, std::lround(bx.L[0] / a_sc)));
  bp.ncy = (int)std::max(1L, std::min(1024L, std::lround(bx.L[1] / a_sc)));
  {
    const double area = (bx.L[0] / bp.ncx) * (bx.L[1] / bp.ncy);
    const long long nzb = std::lround(bx.L[2] * rho * area / 2.0);
    bp.nzb = (int)std::max(1LL, std::min(1LL << 16, nzb));
  }
  bp.cxinv = bp.ncx / bx.L[0];
  bp.cyinv = bp.ncy / bx.L[1];
  bp.zinv = bp.nzb / bx.L[2];
  s->ncols = bp.ncx * bp.ncy;
  // each column's last supercluster may be partly padding
  s->nsuper = (d.n + kSuperSize - 1) / kSuperSize + s->ncols;
  s->npad = s->nsuper * kSuperSize;
  s->ncl = s->npad / kClusterSize;
  for (int k = 0; k < 3; ++k) {
    double H = bx.H[k];
    double e = H / std::max(1.0, std::floor(H / d.cellsize));
    bp.rc_inv[k] = 1.0 / e;
    bp.rc_n[k] = std::max(1, (int)std::floor(H / e + 1.0e-14));
    bp.delta[k] = bx.o[k] - bx.mn[k];
  }
  s->nbins = (long long)s->ncols * bp.nzb;
  s->nscanblocks = (int)((s->nbins + kScanBlock - 1) / kScanBlock);

  // Pair parameters
  PairParams &pp = s->pp;
  pp.box = bx;
  pp.has_lj = d.lj_switch != SW_NONE;
  pp.lj = make_switch(pp.has_lj ? d.lj_switch : SW_CUTOFF, d.lj_on, pp.has_lj ? d.lj_rc : 0.0);
  pp.coul = d.coul_kind;
  if (d.coul_kind == COUL_CUTOFF)
    pp.co = make_switch(d.coul_switch, d.coul_on, d.coul_rc);
  else if (d.coul_kind == COUL_PME)
    pp.co = make_switch(d.coul_switch, 0.0, d.coul_rc);
  else
    pp.co = make_switch(SW_CUTOFF, 0.0, 0.0);
  const double rcmax = std::max(pp.has_lj ? d.lj_rc : 0.0, d.coul_kind ? d.coul_rc : 0.0);
  pp.rc2max = rcmax * rcmax;
  const double Lmax = std::max(bx.L[0], std::max(bx.L[1], bx.L[2]));
  const double eps = s->mixed ? std::ldexp(1.0, -24) : std::ldexp(1.0, -53);
  // |delta r2| bound: coordinates rounded once (<= L eps), differences and
  // image shift (<= 3 L eps per component), squares and sums (3 rc^2 eps).
  pp.guard = 4.0 * (2.0 * std::sqrt(3.0) * rcmax * 3.0 * Lmax * eps + 3.0 * pp.rc2max * eps);
  pp.csf = d.csf;
  pp.ntypes = d.ntypes;
  for (int k = 0; k < 3; ++k) {
    pp.Lhi[k] = origin_round(bx.L[k]);
    pp.Llo[k] = bx.L[k] - pp.Lhi[k];
  }

  s->pme_on = d.coul_kind == COUL_PME;
  // PME without -real: the pair pass still decides membership (pair
  // statistics) but its Coulomb part sees zero charges
  s->qpair = s->pme_on && !(d.pme_terms & PMX_PME_REAL) ? 0.0 : 1.0;
  s->alpha = 0;
  s->self_e = s->charged_e = 0;
  if (s->pme_on) {
    s->alpha = d.pme_alpha > 0 ? d.pme_alpha : choose_alpha(d.coul_rc, d.pme_accuracy);
    pp.alpha = s->alpha;
    pp.a2pi = 2.0 * s->alpha / std::sqrt(M_PI);
    if (d.pme_terms & PMX_PME_CORR) {
      double q2 = 0, qs = 0;
      for (int i = 0; i < d.n; ++i) q2 += d.charge[i] * d.charge[i];
      for (int i = 0; i < d.n; ++i) qs += d.charge[i];
      s->self_e = -q2 * s->alpha / std::sqrt(M_PI);
      if (std::fabs(qs * 0.00268283) > 1.0e-5) s->charged_e = -M_PI / (2.0 * vol * s->alpha * s->alpha) * qs * qs;
    }
  }

  // List parameters
  ListParams &lp = s->lp;
  const double margin = 4.0 * Lmax * std::ldexp(1.0, -23) + 1e-4;
  for (int k = 0; k < 3; ++k) {
    lp.L[k] = (float)bx.L[k];
    lp.Linv[k] = (float)(1.0 / bx.L[k]);
  }
  lp.skin = (float)s->md.skin;
  lp.rlist2 = (float)((rcmax + margin + s->md.skin) * (rcmax + margin + s->md.skin));
  s->pp.prune_r2 = (double)lp.rlist2 * (1.0 + 1e-5);
  lp.rc = (float)rcmax;
  lp.ncl = s->ncl;
  lp.nsuper = s->nsuper;
  {
    // supercluster pruning radius: the list radius plus a slack that covers
    // the FP32 rounding of the boxes; column reach from the column widths
    // (atoms of columns kx + 1 apart are more than kx widths apart in x)
    const double rl = rcmax + margin + s->md.skin + 0.02;
    lp.rlist2_sc = (float)(rl * rl);
    lp.ncx = bp.ncx;
    lp.ncy = bp.ncy;
    lp.kx = (int)std::floor(rl / (bx.L[0] / bp.ncx)) + 1;
    lp.ky = (int)std::floor(rl / (bx.L[1] / bp.ncy)) + 1;
  }
  {
    // generous per-supercluster capacity from the cluster density
    const double rho_cl = s->ncl / vol;
    const double reach = 2.0 * rcmax + 24.0;
    long long c = (long long)std::ceil(rho_cl * reach * reach * reach) + 64;
    s->cap = (int)std::min<long long>(c, s->ncl);
    s->cap = (s->cap + kTileEntries - 1) / kTileEntries * kTileEntries;
  }
  lp.cap = s->cap;
  s->ntiles_cap = kMaxRuns + 2 * s->nsuper;  // a run touches one supercluster or more
  s->maxmaskwork = std::max(1024, s->nsuper * 128);

  int rc;
  auto A = [&](DevBuf &b, size_t bytes) { return dalloc(s, b, bytes); };
  const size_t N = d.n, NP = s->npad;
  if ((rc = A(s->xorig, 24 * N)) || (rc = A(s->xwrap, 24 * N)) || (rc = A(s->q_d, 8 * N)) ||
      (rc = A(s->type_d, 4 * N)) || (rc = A(s->key, 4 * N)) || (rc = A(s->hist, 4 * s->nbins)) ||
      (rc = A(s->start, 4 * s->nbins)) || (rc = A(s->cursor, 4 * s->nbins)) ||
      (rc = A(s->scansum, 4 * (size_t)s->nscanblocks)) || (rc = A(s->perm, 4 * N)) || (rc = A(s->perm2, 4 * N)) ||
      (rc = A(s->permp, 4 * NP)) ||
      (rc = A(s->slot, 4 * N)) || (rc = A(s->refcell, 4 * N)) || (rc = A(s->xs, 32 * NP)) ||
      (rc = A(s->xs64, 32 * NP)) || (rc = A(s->ts, 4 * NP)) || (rc = A(s->sc_info, 8 * (size_t)s->nsuper)) ||
      (rc = A(s->sc_col, 4 * (size_t)s->nsuper)) || (rc = A(s->colsc, 4 * ((size_t)s->ncols + 1))) ||
      (rc = A(s->sctr, 16 * (size_t)s->nsuper)) || (rc = A(s->shalf, 16 * (size_t)s->nsuper)) ||
      (rc = A(s->sorg, 16 * (size_t)s->nsuper)) ||
      (rc = A(s->ctr, 16 * (size_t)s->ncl)) ||
      (rc = A(s->half, 16 * (size_t)s->ncl)) || (rc = A(s->entries, 16 * (size_t)s->nsuper * s->cap)) ||
      (rc = A(s->count, 4 * (size_t)s->nsuper)) || (rc = A(s->lcost, 4 * (size_t)s->nsuper)) ||
      (rc = A(s->ecost, 4 * (size_t)s->nsuper * s->cap)) ||
      (rc = A(s->tep, 4 * ((size_t)s->nsuper + 1))) || (rc = A(s->tcp, 8 * ((size_t)s->nsuper + 1))) ||
      (rc = A(s->tpos, 4 * ((size_t)kMaxRuns + 1))) ||
      (rc = A(s->tposS, 4 * ((size_t)kMaxRuns + 1))) ||
      (rc = A(s->wstart, 4 * ((size_t)kMaxRuns + 1))) ||
      (rc = A(s->maskwork, 4 * (size_t)s->maxmaskwork)) ||
      (rc = A(s->nmaskwork, 8)) || (rc = A(s->masks, 16 * 8 * (size_t)s->maxmaskwork)) ||
      (rc = A(s->overflow, 8)) || (rc = A(s->tiles, 16 * (size_t)s->ntiles_cap)) || (rc = A(s->ntiles, 16)) ||
      (rc = A(s->tilectr, 8)) || (rc = A(s->ljtab, 8 * s->lj.size())) ||
      (rc = A(s->ljt32, 4 * s->lj.size())) || (rc = A(s->facc, 24 * NP)) ||
      (rc = A(s->fout, 24 * N)) || (rc = A(s->pair_e, 16 * (size_t)s->ntiles_cap)) ||
      (rc = A(s->npairs, 8 * (size_t)s->ntiles_cap)) || (rc = A(s->hash, 8 * (size_t)s->ntiles_cap)) ||
      (rc = A(s->eptr_d, 4 * (N + 1))) || (rc = A(s->elist_d, 8 * s->elist.size())) ||
      (rc = A(s->etab_d, 16 * s->etab.size())) || (rc = A(s->eout, 64)) || (rc = A(s->np_out, 8)) ||
      (rc = A(s->hash_out, 8)) || (rc = A(s->border, 8 * (size_t)s->maxborder)) ||
      (rc = A(s->nborder, 8)) || (rc = A(s->border_acc, 32)))
    return rc;
  s->corr_blocks = std::max(1, (int)nblk((long long)s->etab.size(), 256));
  // boundary pairs grow with the pair count: one block per 16 superclusters
  s->border_blocks = std::min(128, std::max(4, (int)(s->nsuper / 16)));
  if ((rc = A(s->corr_e, 8 * (size_t)s->corr_blocks))) return rc;
  if (s->dry) {
    // PME buffers as allocated below (grid, half spectrum, real potential,
    // influence table, twiddles, energy partials bounded by ny * nzh)
    if (s->pme_on) {
      const size_t nx = d.pme_grid[0], ny = d.pme_grid[1], nz = d.pme_grid[2], nzh = nz / 2 + 1;
      const size_t G = nx * ny * nz, GH = nx * ny * nzh, cxb = s->mixed ? 8 : 16, rb = s->mixed ? 4 : 8;
      for (size_t b : {8 * G, cxb * GH, cxb * GH, rb * G, rb * GH, cxb * nx, cxb * ny, cxb * nz, 8 * ny * nzh, (size_t)8,
                       (size_t)8, (size_t)8, (size_t)8, (size_t)8})
        s->dry_bytes += size_class(b);
    }
    return PMX_OK;
  }

  hipStream_t st = s->ctx->stream;
  // the cell histogram starts cleared; each rebuild's scan clears it again
  hipLaunchKernelGGL(k_fill_u32, dim3(nblk(s->nbins, 256)), dim3(256), 0, st, P<uint32_t>(s->hist), 0u, s->nbins);
  {
    RtLock lk(g_rt_mu);
    // the fork and join order the two streams of one device only (the host
    // never waits on them), so they need no system-scope fence
    if (!s->fork_ev) HIPCHK(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming | hipEventDisableSystemFence));
    if (!s->join_ev) HIPCHK(hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming | hipEventDisableSystemFence));
  }
  HIPCHK(hipMemcpyAsync(s->q_d.p, s->q.data(), 8 * N, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->type_d.p, s->type.data(), 4 * N, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->ljtab.p, s->lj.data(), 8 * s->lj.size(), hipMemcpyHostToDevice, st));
  // FP32 copy for the mixed pair kernel when the table does not fit its LDS
  s->lj32.assign(s->lj.begin(), s->lj.end());
  HIPCHK(hipMemcpyAsync(s->ljt32.p, s->lj32.data(), 4 * s->lj32.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(s->eptr_d.p, s->eptr.data(), 4 * (N + 1), hipMemcpyHostToDevice, st));
  if (!s->elist.empty())
    HIPCHK(hipMemcpyAsync(s->elist_d.p, s->elist.data(), 8 * s->elist.size(), hipMemcpyHostToDevice, st));
  if (!s->etab.empty())
    HIPCHK(hipMemcpyAsync(s->etab_d.p, s->etab.data(), 16 * s->etab.size(), hipMemcpyHostToDevice, st));

  if (s->pme_on) {
    PmeParams &pm = s->pme;
    pm.nx = d.pme_grid[0];
    pm.ny = d.pme_grid[1];
    pm.nz = d.pme_grid[2];
    pm.nzh = pm.nz / 2 + 1;
    pm.order = d.pme_order;
    pm.hxr = pm.nx / bx.L[0];
    pm.hyr = pm.ny / bx.L[1];
    pm.hzr = pm.nz / bx.L[2];
    pm.offx = -(pm.order - 1) / 2 + pm.nx;
    pm.offy = -(pm.order - 1) / 2 + pm.ny;
    pm.offz = -(pm.order - 1) / 2 + pm.nz;
    if (!make_plan(pm.nx, s->plx) || !make_plan(pm.ny, s->ply) || !make_plan(pm.nz, s->plz))
      return fail(s->ctx, PMX_ERR_UNSUPPORTED, "PME grid sizes must be products of 2, 3 and 5");
    const size_t G = (size_t)pm.nx * pm.ny * pm.nz, GH = (size_t)pm.nx * pm.ny * pm.nzh;
    const size_t cxb = s->mixed ? 8 : 16;
    if ((rc = A(s->grid64, s->mixed ? 8 : 8 * G)) || (rc = A(s->gridfx, s->mixed ? 8 * G : 8)) ||
        (rc = A(s->spec, cxb * GH)) || (rc = A(s->spec2, cxb * GH)) || (rc = A(s->greal, (s->mixed ? 4 : 8) * G)) ||
        (rc = A(s->theta, (s->mixed ? 4 : 8) * GH)) || (rc = A(s->twx, cxb * pm.nx)) ||
        (rc = A(s->twy, cxb * pm.ny)) || (rc = A(s->twz, cxb * pm.nz)))
      return rc;
    // Influence function on the half spectrum (Grid.h:237-341 with the
    // tables of Grid.h:526-575 and dftmod :589-600).
    const int p = pm.order;
    std::vector<double> th0(p), dth0(p);
    bspline_host(p, 0.0, th0.data(), dth0.data());
    auto dftmod = [&](int nn, std::vector<double> &out) {
      out.resize(nn);
      for (int i = 0; i < nn; i++) {
        double sc = 0.0, ss = 0.0;
        for (int j = 0; j < p; j++) {
          double x = M_PI * 2.0 * i * j / (double)nn;
          sc += th0[j] * std::cos(x);
          ss += th0[j] * std::sin(x);
        }
        out[i] = 1.0 / (sc * sc + ss * ss);
      }
    };
    std::vector<double> bxm, bym, bzm;
    dftmod(pm.nx, bxm);
    dftmod(pm.ny, bym);
    dftmod(pm.nz, bzm);
    const double W = bx.L[0], Lh = bx.L[1], Hh = bx.L[2];
    const double Wr = 1.0 / W, Lr = 1.0 / Lh, Hr = 1.0 / Hh, V = W * Lh * Hh;
    const double fac = M_PI * M_PI / (s->alpha * s->alpha);
    std::vector<double> ex(pm.nx), ey(pm.ny), ez(pm.nzh);
    for (int i = 0; i < pm.nx; i++) {
      int i0 = i <= pm.nx / 2 ? i : i - pm.nx;
      ex[i] = std::exp(-fac * (i0 * Wr) * (i0 * Wr));
    }
    for (int j = 0; j < pm.ny; j++) {
      int j0 = j <= pm.ny / 2 ? j : j - pm.ny;
      ey[j] = std::exp(-fac * (j0 * Lr) * (j0 * Lr));
    }
    for (int k = 0; k < pm.nzh; k++) ez[k] = std::exp(-fac * (k * Hr) * (k * Hr));
    std::vector<double> th(GH);
    const double piVr = 1.0 / (M_PI * V);
    for (int i = 0; i < pm.nx; i++) {
      int i0 = i <= pm.nx / 2 ? i : i - pm.nx;
      double mi = i0 * Wr, exi = ex[i] * piVr;
      for (int j = 0; j < pm.ny; j++) {
        int j0 = j <= pm.ny / 2 ? j : j - pm.ny;
        double bxy = bxm[i] * bym[j], mj = j0 * Lr, mij = mi * mi + mj * mj, exy = exi * ey[j];
        for (int k = 0; k < pm.nzh; k++) {
          double v = 0.0;
          if (i != 0 || j != 0 || k != 0) {
            double mk = k * Hr, m2 = mij + mk * mk;
            v = bxy * bzm[k] * exy * ez[k] / m2;
          }
          th[((size_t)k * pm.ny + j) * pm.nx + i] = v;  // the x pass's layout [k][j][i]
        }
      }
    }
    // uploads on the system's stream; the host arrays live until the
    // synchronisation at the end of the set-up
    std::vector<double> tw64[3];
    std::vector<float> tw32[3], thf;
    auto tw = [&](int axis, int nn, DevBuf &b) -> hipError_t {
      std::vector<double> &t = tw64[axis];
      t.resize(2 * nn);
      for (int m = 0; m < nn; ++m) {
        t[2 * m] = std::cos(-2.0 * M_PI * m / nn);
        t[2 * m + 1] = std::sin(-2.0 * M_PI * m / nn);
      }
      if (s->mixed) {
        tw32[axis].assign(t.begin(), t.end());
        return hipMemcpyAsync(b.p, tw32[axis].data(), 4 * tw32[axis].size(), hipMemcpyHostToDevice, st);
      }
      return hipMemcpyAsync(b.p, t.data(), 8 * t.size(), hipMemcpyHostToDevice, st);
    };
    HIPCHK(tw(0, pm.nx, s->twx));
    HIPCHK(tw(1, pm.ny, s->twy));
    HIPCHK(tw(2, pm.nz, s->twz));
    if (s->mixed) {
      thf.assign(th.begin(), th.end());
      HIPCHK(hipMemcpyAsync(s->theta.p, thf.data(), 4 * GH, hipMemcpyHostToDevice, st));
    } else {
      HIPCHK(hipMemcpyAsync(s->theta.p, th.data(), 8 * GH, hipMemcpyHostToDevice, st));
    }
    if ((rc = fft_blocking(s))) return rc;
    HIPCHK(hipStreamSynchronize(st));  // the host tables above are still alive
  }
  HIPCHK(hipStreamSynchronize(st));
  if (s->mixed)  // the pair-kernel grid (phase -1 sizes it)
    run_pair_phase<true>(s, false, -1);
  else
    run_pair_phase<false>(s, false, -1);
  s->cached = true;
  return PMX_OK;
}

// Launches of the kernels templated on the force kinds: the pair kernel
// (phase 0, main stream), the boundary pairs (phase 1, main stream), the
// exclusion correction with the bonded terms (phase 2, second stream), or all
// three fixups in one launch on the main stream (phase 3); the fixups are one
// kernel with different block ranges.
template <bool MIXED>
FixupArgs<MIXED> fixup_args(Sys *s, bool stats, bool border_here, bool rest_here, int *nblocks);

// Resident blocks per CU of a pair-kernel launch (at least 1, at most the
// kernel's wave budget per CU); each caller keeps its kernel's value in a
// function-local static, initialized once even under concurrent first calls.
int pair_blocks_per_cu(const void *kernel) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kPairBlock, 0) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return std::min(per_cu, kMaxPairWavesPerCu / (kPairBlock / kWave));
}

template <bool MIXED, int C, int LJ, int CS, bool BIGLJ = false>
void launch_pair_phase(Sys *s, bool stats, int phase) {
  hipStream_t st = phase == 2 ? s->ctx->stream2 : s->ctx->stream;
  const bool border_here = phase == 1 || phase == 3, rest_here = phase == 2 || phase == 3;
  using R4 = typename Real4<typename std::conditional<MIXED, float, double>::type>::type;
  auto *facc = P<typename AccT<MIXED>::type>(s->facc);
  if (phase < 0) {
    // as many blocks as are resident at once (the schedules assign each
    // wave its first tile, or its whole run, statically)
    static const int per_cu =
        pair_blocks_per_cu(reinterpret_cast<const void *>(k_pair<MIXED, C, LJ, CS, BIGLJ, kPairPlain>));
    // knob > 0: blocks per CU; 0: as many as fit; < 0: that many fewer
    const int bpc = s->tune_pair_blocks > 0 ? std::min(s->tune_pair_blocks, per_cu)
                                            : std::max(1, per_cu + s->tune_pair_blocks);
    s->pair_grid = std::max(1, s->ctx->cus * bpc);
    return;
  }
  if (phase == 0) {
    const unsigned grid = (unsigned)s->pair_grid;
    PairOut out{P<double>(s->pair_e), P<long long>(s->npairs), P<unsigned long long>(s->hash),
                P<int2>(s->border), P<int>(s->nborder), s->maxborder, P<int>(s->overflow),
                P<unsigned long long>(s->border_acc)};
    // the list prune runs in the first pass after a build, at the build
    // positions (a pruned cluster pair cannot gain a member before the next
    // build only when judged there); not in a pass that counts pairs, nor
    // for tables beyond the LDS size, which are rare
    const bool prune = !BIGLJ && !stats && s->tune_prune && s->prune_pending;
    s->prune_pending = false;
    const R4 *ljt = MIXED ? reinterpret_cast<const R4 *>(s->ljt32.p) : reinterpret_cast<const R4 *>(s->ljtab.p);
#define PMX_PAIR_ARGS                                                                                              \
  s->pp, s->tune_schedule, P<int4>(s->tiles), P<int>(s->ntiles), P<int>(s->wstart), P<int>(s->tilectr),            \
      P<int4>(s->entries), P<int4>(s->entries), s->cap, P<unsigned short>(s->masks), P<R4>(s->xs),                 \
      P<float4>(s->sorg), P<int>(s->ts), P<int>(s->permp), P<double>(s->ljtab), ljt, facc, out
    if (stats)
      hipLaunchKernelGGL((k_pair<MIXED, C, LJ, CS, BIGLJ, kPairStats>), dim3(grid), dim3(kPairBlock), 0, st,
                         PMX_PAIR_ARGS);
    else if (prune)
      hipLaunchKernelGGL((k_pair<MIXED, C, LJ, CS, false, kPairPrune>), dim3(grid), dim3(kPairBlock), 0, st,
                         PMX_PAIR_ARGS);
    else
      hipLaunchKernelGGL((k_pair<MIXED, C, LJ, CS, BIGLJ, kPairPlain>), dim3(grid), dim3(kPairBlock), 0, st,
                         PMX_PAIR_ARGS);
#undef PMX_PAIR_ARGS
    return;
  }
  int nfix = 0;
  const FixupArgs<MIXED> a = fixup_args<MIXED>(s, stats, border_here, rest_here, &nfix);
  if (nfix > 0)
    hipLaunchKernelGGL((k_fixups<MIXED, C, LJ, CS>), dim3(nfix), dim3(256), 0, st, a);
}

// Arguments of k_fixups for one force pass: the boundary pairs
// (border_here), the exclusion correction and the bonded terms (rest_here);
// *nblocks gets the launch's block count.
template <bool MIXED>
FixupArgs<MIXED> fixup_args(Sys *s, bool stats, bool border_here, bool rest_here, int *nblocks) {
  using R4 = typename Real4<typename std::conditional<MIXED, float, double>::type>::type;
  FixupArgs<MIXED> a{};
  a.border_blocks = border_here ? s->border_blocks : 0;
  a.pp = s->pp;
  a.stats = stats ? 1 : 0;
  a.border = P<int2>(s->border);
  a.nborder = P<int>(s->nborder);
  a.maxborder = s->maxborder;
  a.xs = P<R4>(s->xs);
  a.sorg = P<float4>(s->sorg);
  a.ts = P<int>(s->ts);
  a.perm = P<int>(s->permp);
  a.xorig = P<double>(s->xorig);
  a.ljtab = P<double>(s->ljtab);
  a.eptr = P<int>(s->eptr_d);
  a.elist = P<int2>(s->elist_d);
  a.border_e = P<unsigned long long>(s->border_acc);
  a.border_stat = P<unsigned long long>(s->border_acc) + 2;
  const bool corr = s->pme_on && (s->d.pme_terms & PMX_PME_CORR) && !s->etab.empty();
  a.corr_blocks = corr && rest_here ? s->corr_blocks : 0;
  a.ne = (int)s->etab.size();
  a.etab = P<int4>(s->etab_d);
  a.q = P<double>(s->q_d);
  a.slot = P<int>(s->slot);
  a.alpha = s->alpha;
  a.csf = s->d.csf;
  a.corr_e = P<double>(s->corr_e);
  int bonded_blocks = 0;
  if (rest_here && s->md.on && s->md.nb + s->md.na > 0) {
    bonded_blocks = (int)nblk(s->md.nb + s->md.na, 256);
    a.nbond = s->md.nb;
    a.nangle = s->md.na;
    a.bonds = P<int2>(s->md.bonds);
    a.bpar = P<double2>(s->md.bpar);
    a.angles = P<int4>(s->md.angles);
    a.apar = P<double2>(s->md.apar);
    a.bond_e = P<double>(s->md.bond_e);
  }
  a.facc = P<typename AccT<MIXED>::type>(s->facc);
  *nblocks = a.border_blocks + a.corr_blocks + bonded_blocks;
  return a;
}

template <bool MIXED>
void run_pair_phase(Sys *s, bool stats, int phase) {
  const int lj = s->d.lj_switch, cs = s->d.coul_switch;
  if (s->d.ntypes > kMaxTypes && lj != SW_NONE) {
    // LJ table too large for the kernel's LDS: read from the caches
    if (s->pp.coul == COUL_PME)
      launch_pair_phase<MIXED, COUL_PME, SW_ANY, SW_ANY, true>(s, stats, phase);
    else if (s->pp.coul == COUL_CUTOFF)
      launch_pair_phase<MIXED, COUL_CUTOFF, SW_ANY, SW_ANY, true>(s, stats, phase);
    else
      launch_pair_phase<MIXED, COUL_NONE, SW_ANY, SW_ANY, true>(s, stats, phase);
    return;
  }
  if (s->pp.coul == COUL_PME) {
    if (lj == SW_C2 && cs == SW_CUTOFF)
      launch_pair_phase<MIXED, COUL_PME, SW_C2, SW_CUTOFF>(s, stats, phase);
    else if (lj == SW_NONE)
      launch_pair_phase<MIXED, COUL_PME, SW_NONE, SW_ANY>(s, stats, phase);
    else
      launch_pair_phase<MIXED, COUL_PME, SW_ANY, SW_ANY>(s, stats, phase);
  } else if (s->pp.coul == COUL_CUTOFF) {
    if (lj == SW_C2 && cs == SW_C1)
      launch_pair_phase<MIXED, COUL_CUTOFF, SW_C2, SW_C1>(s, stats, phase);
    else if (lj == SW_NONE)
      launch_pair_phase<MIXED, COUL_CUTOFF, SW_NONE, SW_ANY>(s, stats, phase);
    else
      launch_pair_phase<MIXED, COUL_CUTOFF, SW_ANY, SW_ANY>(s, stats, phase);
  } else {
    launch_pair_phase<MIXED, COUL_NONE, SW_ANY, SW_ANY>(s, stats, phase);
  }
}

// Spreading and gathering for the stencil order ORDER (validated: even, 2..8).
template <bool MIXED, int ORDER>
void launch_spread(Sys *s, hipStream_t st) {
  if (ORDER >= 4) {
    hipLaunchKernelGGL((k_spread_patch<MIXED, ORDER>), dim3(nblk(s->npad, kSpreadSlots)), dim3(256), 0, st, s->npad,
                       s->pme, P<double>(s->xorig), P<int>(s->permp), P<double>(s->q_d), P<double>(s->grid64),
                       P<unsigned long long>(s->gridfx));
    return;
  }
  constexpr int P3 = ORDER * ORDER * ORDER;
  constexpr int apw = P3 >= 64 ? 1 : 64 / P3;
  const int waves = (int)nblk(s->npad, apw);
  hipLaunchKernelGGL((k_spread<MIXED, ORDER>), dim3(nblk((long long)waves * 64, 256)), dim3(256), 0, st, s->npad,
                     s->pme, P<double>(s->xorig), P<int>(s->permp), P<double>(s->q_d), P<double>(s->grid64),
                     P<unsigned long long>(s->gridfx));
}

template <bool MIXED, int ORDER>
void launch_gather(Sys *s, hipStream_t st) {
  using T = typename std::conditional<MIXED, float, double>::type;
  if (ORDER >= 4) {
    hipLaunchKernelGGL((k_gather_patch<MIXED, ORDER>), dim3(nblk(s->npad, kSpreadSlots)), dim3(256), 0, st, s->npad,
                       s->pme, P<double>(s->xorig), P<int>(s->permp), P<double>(s->q_d), P<T>(s->greal),
                       P<typename AccT<MIXED>::type>(s->facc));
    return;
  }
  const int waves = (int)nblk(s->npad, 64 / pme_tpa(ORDER));
  hipLaunchKernelGGL((k_gather<MIXED, ORDER>), dim3(nblk((long long)waves * 64, 256)), dim3(256), 0, st, s->npad,
                     waves,
                     s->pme, P<double>(s->xorig), P<int>(s->permp), P<double>(s->q_d), P<T>(s->greal),
                     P<typename AccT<MIXED>::type>(s->facc));
}

// The fused unsort + kick + drift + next spread (orders 4 to 8).
template <bool MIXED>
void launch_unsort_kick_drift_spread(Sys *s, const Kick &kk, unsigned *dmax) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using R4 = typename Real4<R>::type;
  hipStream_t st = s->ctx->stream;
#define PMX_UKDS(ORD)                                                                                              \
  hipLaunchKernelGGL((k_unsort_kick_drift_spread<MIXED, R, ORD>), dim3(nblk(s->npad, kSpreadSlots)), dim3(256), 0, \
                     st, s->npad, kk, s->md.h, P<int>(s->permp), P<typename AccT<MIXED>::type>(s->facc),           \
                     P<double>(s->fout), P<int>(s->tilectr), P<int>(s->nborder), P<double>(s->md.minv),             \
                     P<double>(s->md.v), P<double>(s->xorig), P<double>(s->md.xbuild), P<R4>(s->xs),                \
                     P<double4>(s->xs64), P<float4>(s->sorg), dmax, s->pme, P<double>(s->q_d),                      \
                     P<double>(s->grid64), P<unsigned long long>(s->gridfx))
  switch (s->pme.order) {
    case 4: PMX_UKDS(4); break;
    case 6: PMX_UKDS(6); break;
    default: PMX_UKDS(8); break;
  }
#undef PMX_UKDS
}

// Charge spreading on the main stream (before the pair kernel: alone it
// runs at full rate, next to the pair kernel its many small waves would
// not); the charge grid was zeroed by the k_zero launch of the pass.
template <bool MIXED>
void run_spread(Sys *s, hipStream_t st) {
  switch (s->pme.order) {
    case 2: launch_spread<MIXED, 2>(s, st); break;
    case 4: launch_spread<MIXED, 4>(s, st); break;
    case 6: launch_spread<MIXED, 6>(s, st); break;
    default: launch_spread<MIXED, 8>(s, st); break;
  }
}

// FFT convolution and force gather on the second stream.
template <bool MIXED>
int run_pme(Sys *s, bool timing) {
  using T = typename std::conditional<MIXED, float, double>::type;
  hipStream_t st = s->ctx->stream2;
  const PmeParams &pm = s->pme;
  if (s->d.pme_terms & PMX_PME_RECIP) {
    const size_t cx = sizeof(Cx<T>);
    // z pass per (i, TY lines of j), y passes per (k, TI lines of i), x pass
    // per (k, TJ lines of j)
    const int zb = pm.nx * ((pm.ny + s->zl - 1) / s->zl);
    const int yb = pm.nzh * ((pm.nx + s->tky - 1) / s->tky);
    hipLaunchKernelGGL((k_fft_z_r2c<T, MIXED>), dim3(zb), dim3(256), 2 * s->zl * pm.nz * cx, st, pm, s->plz, s->zl,
                       P<double>(s->grid64), P<unsigned long long>(s->gridfx), P<Cx<T>>(s->twz), P<Cx<T>>(s->spec));
    if (s->fft_plane) {
      hipLaunchKernelGGL((k_fft_plane_conv<T>), dim3(pm.nzh), dim3(256), 2 * (size_t)pm.nx * pm.ny * cx, st, pm,
                         s->ply, s->plx, P<Cx<T>>(s->twy), P<Cx<T>>(s->twx), P<T>(s->theta), P<Cx<T>>(s->spec),
                         P<double>(s->recip_e));
    } else {
      hipLaunchKernelGGL((k_fft_y<T>), dim3(yb), dim3(256), 2 * s->tky * pm.ny * cx, st, pm, s->ply, s->tky, false,
                         P<Cx<T>>(s->twy), P<Cx<T>>(s->spec), P<Cx<T>>(s->spec2));
      hipLaunchKernelGGL((k_fft_x_conv<T>), dim3(s->recip_blocks), dim3(256), 2 * s->tkx * pm.nx * cx, st, pm,
                         s->plx, s->tkx, P<Cx<T>>(s->twx), P<T>(s->theta), P<Cx<T>>(s->spec2),
                         P<double>(s->recip_e));
      hipLaunchKernelGGL((k_fft_y<T>), dim3(yb), dim3(256), 2 * s->tky * pm.ny * cx, st, pm, s->ply, s->tky, true,
                         P<Cx<T>>(s->twy), P<Cx<T>>(s->spec2), P<Cx<T>>(s->spec));
    }
    hipLaunchKernelGGL((k_fft_z_c2r<T>), dim3(zb), dim3(256), 2 * s->zl * pm.nz * cx, st, pm, s->plz, s->zl,
                       P<Cx<T>>(s->twz), P<Cx<T>>(s->spec), P<T>(s->greal));
    if (timing) hipEventRecord(s->ev[5], st);
    switch (pm.order) {
      case 2: launch_gather<MIXED, 2>(s, st); break;
      case 4: launch_gather<MIXED, 4>(s, st); break;
      case 6: launch_gather<MIXED, 6>(s, st); break;
      default: launch_gather<MIXED, 8>(s, st); break;
    }
  } else if (timing) {
    hipEventRecord(s->ev[5], st);
  }
  if (timing) hipEventRecord(s->ev[6], st);
  return PMX_OK;
}

void reduce_energies(Sys *s, bool stats) {
  const bool recip = s->pme_on && (s->d.pme_terms & PMX_PME_RECIP);
  const bool corr = s->pme_on && (s->d.pme_terms & PMX_PME_CORR) && !s->etab.empty();
  hipLaunchKernelGGL(k_reduce_energy, dim3(1), dim3(1024), 0, s->ctx->stream, s->ntiles_cap, P<int>(s->ntiles),
                     P<double>(s->pair_e), s->mixed ? 1 : 0, recip ? s->recip_blocks : 0, P<double>(s->recip_e),
                     corr ? s->corr_blocks : 0, P<double>(s->corr_e), P<long long>(s->npairs),
                     P<unsigned long long>(s->hash), stats ? 1 : 0, P<unsigned long long>(s->border_acc),
                     P<unsigned long long>(s->border_acc) + 2, P<double>(s->eout), P<long long>(s->np_out),
                     P<unsigned long long>(s->hash_out));
}

void timing_begin_call(Sys *s, bool timing) {
  if (!timing) return;
  RtLock lk(g_rt_mu);
  if (s->evused == s->evsets.size()) {
    std::vector<hipEvent_t> set(kStages + 1);
    for (auto &e : set) hipEventCreate(&e);
    s->evsets.push_back(set);
  }
  s->ev = s->evsets[s->evused++].data();
  hipEventRecord(s->ev[0], s->ctx->stream);
}

// Cell sort of the current positions and the cluster-pair list built from it
// (the reference rebuilds its cell lists at every force evaluation,
// ForceGroup.cpp:57; the MD path rebuilds when the skin requires it).
// Work tiles of the current list for the current schedule and pair grid.
void make_tiles(Sys *s, hipStream_t st) {
  int *const nruns = P<int>(s->ntiles) + 2;
  hipLaunchKernelGGL(k_runs_prefix, dim3(1), dim3(1024), 0, st, s->nsuper, P<int>(s->count), P<int>(s->lcost), nruns,
                     P<int>(s->tep), P<long long>(s->tcp), P<int>(s->tpos), P<int>(s->tposS));
  hipLaunchKernelGGL(k_runs_cut, dim3(s->nsuper), dim3(256), 0, st, s->nsuper, P<int>(s->count), P<int>(s->ecost),
                     s->cap, nruns, P<int>(s->tep), P<long long>(s->tcp), P<int>(s->tpos), P<int>(s->tposS));
  hipLaunchKernelGGL(k_runs_tiles, dim3(1), dim3(1024), 0, st, P<int>(s->count), P<int>(s->tep), P<int>(s->tpos),
                     P<int>(s->tposS), P<int>(s->wstart), P<int4>(s->tiles), P<int>(s->ntiles), s->ntiles_cap);
}

int rebuild(Sys *s, bool timing) {
  hipStream_t st = s->ctx->stream;
  const int n = s->n;
  // the histogram starts cleared (set-up, then each scan clears it)
  hipLaunchKernelGGL(k_bin, dim3(nblk(n, 256)), dim3(256), 0, st, n, P<double>(s->xorig), s->bin, P<uint32_t>(s->key),
                     P<double>(s->xwrap), P<int>(s->refcell), P<uint32_t>(s->hist));
  if (s->nbins <= kScanSingleMax) {
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kScanBlock), 0, st, P<uint32_t>(s->hist), s->nbins,
                       P<uint32_t>(s->start), P<uint32_t>(s->cursor));
  } else {
    hipLaunchKernelGGL(k_scan_blocksums, dim3(s->nscanblocks), dim3(kScanBlock), 0, st, P<uint32_t>(s->hist),
                       s->nbins, P<uint32_t>(s->scansum));
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanBlock), 0, st, P<uint32_t>(s->scansum), s->nscanblocks);
    hipLaunchKernelGGL(k_scan_final, dim3(s->nscanblocks), dim3(kScanBlock), 0, st, P<uint32_t>(s->hist), s->nbins,
                       P<uint32_t>(s->scansum), P<uint32_t>(s->start), P<uint32_t>(s->cursor));
  }
  hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, st, n, P<uint32_t>(s->key), P<uint32_t>(s->cursor),
                     P<int>(s->perm));
  hipLaunchKernelGGL(k_binrank, dim3(nblk(n, 256)), dim3(256), 0, st, n, s->nbins, P<uint32_t>(s->key),
                     P<uint32_t>(s->start), P<int>(s->perm), P<int>(s->perm2));
  hipLaunchKernelGGL(k_columns, dim3(1), dim3(1024), 0, st, s->ncols, s->bin.nzb, n, P<uint32_t>(s->start),
                     s->nsuper, P<int2>(s->sc_info), P<int>(s->sc_col), P<int>(s->colsc));
  double *xbuild = s->md.on ? P<double>(s->md.xbuild) : nullptr;
  if (s->mixed)
    hipLaunchKernelGGL(k_split_fill<float>, dim3(nblk((long long)s->nsuper * 64, 256)), dim3(256), 0, st, s->nsuper,
                       P<int2>(s->sc_info), P<int>(s->perm2), P<double>(s->xwrap), P<int>(s->permp), P<int>(s->slot),
                       P<double>(s->q_d), P<int>(s->type_d), P<float4>(s->xs), P<double4>(s->xs64), P<int>(s->ts),
                       P<double>(s->xorig), xbuild, P<float4>(s->sorg), P<float4>(s->ctr), P<float4>(s->half),
                       P<float4>(s->sctr), P<float4>(s->shalf), P<int>(s->nmaskwork), s->qpair);
  else
    hipLaunchKernelGGL(k_split_fill<double>, dim3(nblk((long long)s->nsuper * 64, 256)), dim3(256), 0, st, s->nsuper,
                       P<int2>(s->sc_info), P<int>(s->perm2), P<double>(s->xwrap), P<int>(s->permp), P<int>(s->slot),
                       P<double>(s->q_d), P<int>(s->type_d), P<double4>(s->xs), P<double4>(s->xs64), P<int>(s->ts),
                       P<double>(s->xorig), xbuild, P<float4>(s->sorg), P<float4>(s->ctr), P<float4>(s->half),
                       P<float4>(s->sctr), P<float4>(s->shalf), P<int>(s->nmaskwork), s->qpair);
  if (timing) hipEventRecord(s->ev[1], st);
  hipLaunchKernelGGL(k_build_list, dim3(s->nsuper), dim3(kListBlock), 0, st, s->lp, P<int>(s->sc_col),
                     P<int>(s->colsc), P<float4>(s->sctr), P<float4>(s->shalf), P<float4>(s->ctr),
                     P<float4>(s->half), P<int4>(s->entries), P<int>(s->count), P<int>(s->ecost), P<int>(s->lcost),
                     P<int>(s->maskwork), P<int>(s->nmaskwork), s->maxmaskwork, P<int>(s->overflow));
  if (!s->etab.empty())
    hipLaunchKernelGGL(k_mark_exclusions, dim3(nblk((long long)s->etab.size(), 256)), dim3(256), 0, st,
                       (int)s->etab.size(), P<int4>(s->etab_d), P<int>(s->slot), P<int4>(s->entries),
                       P<int>(s->count), s->cap, P<int>(s->maskwork), P<int>(s->nmaskwork), s->maxmaskwork,
                       P<int>(s->overflow));
  // a wave per mask slot, up to 16 blocks per CU (the slots in use are known
  // on the device only; waves beyond them return at once)
  hipLaunchKernelGGL(k_build_masks,
                     dim3((unsigned)std::min<long long>(nblk((long long)s->maxmaskwork, 4), 16ll * s->ctx->cus)),
                     dim3(256), 0, st, P<int>(s->maskwork),
                     P<int>(s->nmaskwork), s->maxmaskwork, P<int4>(s->entries), s->cap, P<int>(s->permp),
                     P<int>(s->eptr_d), P<int2>(s->elist_d), P<unsigned short>(s->masks));
  make_tiles(s, st);
  s->list_valid = true;
  s->prune_pending = true;
  if (timing) hipEventRecord(s->ev[2], st);
  return PMX_OK;
}

// Forces (and optionally energies) at the current positions with the
// current list: pair + boundary pairs, PME, correction, bonded terms, then
// the FP64 force array in the caller's atom order.
// next: the following MD step's kick (with next_dmax, its displacement
// record): the pass ends with that step's kick and drift fused into the
// force unsort.
// Clears the force accumulators, the pass counters and the PME charge grid.
void launch_zero(Sys *s) {
  const int npad = s->npad;
  ZeroRanges z{};
  int nr = 0;
  z.p[nr] = P<unsigned long long>(s->facc);
  z.n[nr++] = 3ll * npad;
  z.p[nr] = P<unsigned long long>(s->tilectr);
  z.n[nr++] = 1;
  z.p[nr] = P<unsigned long long>(s->nborder);
  z.n[nr++] = 1;
  z.p[nr] = P<unsigned long long>(s->border_acc);
  z.n[nr++] = 4;
  if (s->pme_on && (s->d.pme_terms & PMX_PME_RECIP)) {
    const long long G = (long long)s->pme.nx * s->pme.ny * s->pme.nz;
    z.p[nr] = s->mixed ? P<unsigned long long>(s->gridfx) : P<unsigned long long>(s->grid64);
    z.n[nr++] = G;
  }
  hipLaunchKernelGGL(k_zero, dim3(std::max(1u, nblk(3ll * npad, 256)), nr), dim3(256), 0, s->ctx->stream, z);
}

int forces(Sys *s, unsigned flags, bool energies, const Kick *next = nullptr, unsigned *next_dmax = nullptr,
           bool spread_done = false) {
  hipStream_t st = s->ctx->stream;
  const bool timing = flags & PMX_EVAL_TIMING;
  const bool stats = flags & PMX_EVAL_PAIRSTATS;
  const int n = s->n, npad = s->npad;
  // A completed pass leaves the accumulators, the PME grid and the counters
  // cleared (k_unsort_forces, k_fft_z_r2c, k_pair); only the first pass after
  // (re)allocation or after an interrupted pass clears them here.  A spread
  // fused into the previous kick-drift is this pass's only when the caller
  // says so; otherwise it belongs to other positions and is cleared.
  if (s->spread_ready && !spread_done) s->acc_clean = false;
  s->spread_ready = false;
  if (!s->acc_clean) launch_zero(s);
  s->acc_clean = false;
  if (timing) hipEventRecord(s->ev[2], st);
  // fork: FFT, gather, exclusion correction and bonded terms on the second
  // stream beside the pair kernel; all add into the zeroed accumulators.
  // The spread runs before the fork on the main stream, or first on the
  // second stream (tuning knob).  The pair stage's start event is recorded
  // after the fork, right before the pair kernel.
  hipStream_t st2 = s->ctx->stream2;
  const bool spread = s->pme_on && (s->d.pme_terms & PMX_PME_RECIP);
  const bool spread2 = spread && s->tune_spread_stream2;
  if (spread && !spread2 && !spread_done) {
    if (s->mixed)
      run_spread<true>(s, st);
    else
      run_spread<false>(s, st);
  }
  HIPCHK(hipEventRecord(s->fork_ev, st));
  HIPCHK(hipStreamWaitEvent(st2, s->fork_ev, 0));
  if (spread2) {
    if (s->mixed)
      run_spread<true>(s, st2);
    else
      run_spread<false>(s, st2);
  }
  if (timing) hipEventRecord(s->ev[4], st);
  if (s->mixed)
    run_pair_phase<true>(s, stats, 0);
  else
    run_pair_phase<false>(s, stats, 0);
  if (timing) hipEventRecord(s->ev[3], st);
  if (s->pme_on) {
    int rc = s->mixed ? run_pme<true>(s, timing) : run_pme<false>(s, timing);
    if (rc) return rc;
  } else if (timing) {
    hipEventRecord(s->ev[5], st2);
    hipEventRecord(s->ev[6], st2);
  }
  // the exclusion correction and bonded terms on the second stream (after
  // the PME chain), or with the boundary pairs on the main stream after the
  // pair kernel (tuning knob)
  const int fix_phase = s->tune_fixup_main ? 3 : 1;
  if (!s->tune_fixup_main) {
    if (s->mixed)
      run_pair_phase<true>(s, stats, 2);
    else
      run_pair_phase<false>(s, stats, 2);
  }
  if (timing) hipEventRecord(s->ev[7], st2);
  HIPCHK(hipEventRecord(s->join_ev, st2));
  // boundary pairs after the pair kernel, then join
  if (s->mixed)
    run_pair_phase<true>(s, stats, fix_phase);
  else
    run_pair_phase<false>(s, stats, fix_phase);
  HIPCHK(hipStreamWaitEvent(st, s->join_ev, 0));
  // inside an MD chunk the next pass's spread joins the kick-drift
  const bool fuse_spread = next && spread && !spread2 && s->pme.order >= 4 && s->tune_fused_spread;
  if (fuse_spread) {
    if (s->mixed)
      launch_unsort_kick_drift_spread<true>(s, *next, next_dmax);
    else
      launch_unsort_kick_drift_spread<false>(s, *next, next_dmax);
    s->spread_ready = true;
  } else if (next && s->mixed)
    hipLaunchKernelGGL((k_unsort_kick_drift<true, float>), dim3(nblk(npad, 256)), dim3(256), 0, st, npad, *next,
                       s->md.h, P<int>(s->permp), P<unsigned long long>(s->facc), P<double>(s->fout),
                       P<int>(s->tilectr), P<int>(s->nborder), P<double>(s->md.minv), P<double>(s->md.v),
                       P<double>(s->xorig), P<double>(s->md.xbuild), P<float4>(s->xs), P<double4>(s->xs64),
                       P<float4>(s->sorg), next_dmax);
  else if (next)
    hipLaunchKernelGGL((k_unsort_kick_drift<false, double>), dim3(nblk(npad, 256)), dim3(256), 0, st, npad, *next,
                       s->md.h, P<int>(s->permp), P<double>(s->facc), P<double>(s->fout), P<int>(s->tilectr),
                       P<int>(s->nborder), P<double>(s->md.minv), P<double>(s->md.v), P<double>(s->xorig),
                       P<double>(s->md.xbuild), P<double4>(s->xs), P<double4>(s->xs64), P<float4>(s->sorg),
                       next_dmax);
  else if (s->mixed)
    hipLaunchKernelGGL(k_unsort_forces<true>, dim3(nblk(n, 256)), dim3(256), 0, st, n, P<int>(s->slot),
                       P<unsigned long long>(s->facc), P<double>(s->fout), P<int>(s->tilectr),
                       P<int>(s->nborder));
  else
    hipLaunchKernelGGL(k_unsort_forces<false>, dim3(nblk(n, 256)), dim3(256), 0, st, n, P<int>(s->slot),
                       P<double>(s->facc), P<double>(s->fout), P<int>(s->tilectr), P<int>(s->nborder));
  s->acc_clean = true;
  if (energies) reduce_energies(s, stats);
  if (timing) hipEventRecord(s->ev[8], st);
  HIPCHK(hipGetLastError());
  return PMX_OK;
}

int compute(Sys *s, unsigned flags) {
  const bool timing = flags & PMX_EVAL_TIMING;
  timing_begin_call(s, timing);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s->ctx->stream, P<int>(s->overflow), 0, 2ll);
  int rc = rebuild(s, timing);
  if (rc) return rc;
  return forces(s, flags, true);
}

int finish(Sys *s, unsigned flags, double *f_aos, double *energies) {
  hipStream_t st = s->ctx->stream;
  int ovf = 0;
  HIPCHK(hipMemcpyAsync(&ovf, s->overflow.p, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(s->eout_h, s->eout.p, 32, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&s->last_np, s->np_out.p, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&s->last_hash, s->hash_out.p, 8, hipMemcpyDeviceToHost, st));
  std::vector<double> f;
  if (f_aos) {
    f.resize(3 * (size_t)s->n);
    HIPCHK(hipMemcpyAsync(f.data(), s->fout.p, 24 * (size_t)s->n, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  if (ovf) return fail(s->ctx, PMX_ERR_OVERFLOW, "cluster pair list capacity exceeded");
  if ((flags & PMX_EVAL_TIMING) && s->ev)
    for (int k = 0; k < kStages; ++k)
      hipEventElapsedTime(&s->stage_ms[k], s->ev[kStageEv[k][0]], s->ev[kStageEv[k][1]]);
  if (f_aos)
    for (size_t k = 0; k < f.size(); ++k) f_aos[k] += f[k];
  if (energies) {
    double ec = s->eout_h[0];
    if (s->pme_on) ec += s->eout_h[2] + s->eout_h[3] + s->self_e + s->charged_e;
    energies[PMX_E_COULOMB] += ec;
    energies[PMX_E_LENNARDJONES] += s->eout_h[1];
  }
  return PMX_OK;
}

void drop_graphs(Sys *s) {
  for (auto &kv : s->md.gexec) hipGraphExecDestroy(kv.second);
  s->md.gexec.clear();
}

int grow_and_retry(Sys *s) {
  drop_graphs(s);  // captured chunks hold the old buffer addresses
  s->list_valid = false;
  // double the list capacity and the mask pool, then rebuild buffers
  s->cap = std::min(s->ncl + kTileEntries, 2 * s->cap);
  s->cap = (s->cap + kTileEntries - 1) / kTileEntries * kTileEntries;
  s->lp.cap = s->cap;
  s->ntiles_cap = kMaxRuns + 2 * s->nsuper;
  s->maxmaskwork *= 2;
  s->maxborder *= 2;
  if (int rc0 = dalloc(s, s->border, 8 * (size_t)s->maxborder)) return rc0;
  int rc;
  if ((rc = dalloc(s, s->entries, 16 * (size_t)s->nsuper * s->cap)) || (rc = dalloc(s, s->ecost, 4 * (size_t)s->nsuper * s->cap)) ||
      (rc = dalloc(s, s->tiles, 16 * (size_t)s->ntiles_cap)) ||
      (rc = dalloc(s, s->pair_e, 16 * (size_t)s->ntiles_cap)) || (rc = dalloc(s, s->npairs, 8 * (size_t)s->ntiles_cap)) ||
      (rc = dalloc(s, s->hash, 8 * (size_t)s->ntiles_cap)) || (rc = dalloc(s, s->maskwork, 4 * (size_t)s->maxmaskwork)) ||
      (rc = dalloc(s, s->masks, 16 * 8 * (size_t)s->maxmaskwork)))
    return rc;
  return PMX_OK;
}

}  // namespace

struct pmx_sys {
  Sys s;
};

extern "C" {

unsigned long long pmx_pair_hash(int i, int j) {
  if (i > j) std::swap(i, j);
  return pair_hash((uint32_t)i, (uint32_t)j);
}

int pmx_ctx_create(int device, pmx_ctx **out) {
  RtLock lk(g_rt_mu);
  if (!out) return PMX_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PMX_ERR_NODEV;
  if (device < 0 || device >= ndev) return PMX_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return PMX_ERR_HIP;
  // The code objects in this library are gfx950 only.
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return PMX_ERR_HIP;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return PMX_ERR_NODEV;
  if (device >= kMaxDevices) return PMX_ERR_UNSUPPORTED;
  if (!g_scratch_addr[device]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PMX_ERR_HIP;
    g_scratch_addr[device] = ptr;
  }
  pmx_ctx *c = new pmx_ctx();
  c->device = device;
  c->scratch = g_scratch_addr[device];
  c->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return PMX_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking) != hipSuccess) {
    hipStreamDestroy(c->stream);
    delete c;
    return PMX_ERR_HIP;
  }
  *out = c;
  return PMX_OK;
}

int pmx_ctx_destroy(pmx_ctx *ctx) {
  RtLock lk(g_rt_mu);
  if (!ctx) return PMX_ERR_ARG;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  if (ctx->stream2) hipStreamDestroy(ctx->stream2);
  if (ctx->batch_host) hipHostFree(ctx->batch_host);
  if (ctx->batch_fork) hipEventDestroy(ctx->batch_fork);
  if (ctx->batch_join) hipEventDestroy(ctx->batch_join);
  delete ctx;
  return PMX_OK;
}

const char *pmx_last_error(const pmx_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int pmx_sys_create(pmx_ctx *ctx, const pmx_sys_desc *desc, pmx_sys **out) {
  if (!ctx || !out) return PMX_ERR_ARG;
  *out = nullptr;
  int rc = validate(ctx, desc);
  if (rc) return rc;
  hipSetDevice(ctx->device);
  pmx_sys *ps = new pmx_sys();
  ps->s.ctx = ctx;
  ps->s.d = *desc;
  // below this size the PME chain (a fixed number of short dependent
  // launches) is the longer branch of a force pass
  ps->s.tune_fixup_main = desc->n < kFixupMainBelow ? 1 : 0;
  rc = setup(&ps->s);
  if (rc) {
    pmx_sys_destroy(ps);
    return rc;
  }
  *out = ps;
  return PMX_OK;
}

int pmx_sys_destroy(pmx_sys *sys) {
  if (!sys) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  hipSetDevice(s->ctx->device);
  hipStreamSynchronize(s->ctx->stream);
  hipStreamSynchronize(s->ctx->stream2);
  RtLock lk(g_rt_mu);
  drop_graphs(s);
  for (DevBuf *b : s->all) block_put(s->ctx->device, b->p, b->bytes);
  if (s->fork_ev) hipEventDestroy(s->fork_ev);
  if (s->join_ev) hipEventDestroy(s->join_ev);
  for (auto &e : s->md.chunk_ev)
    if (e) hipEventDestroy(e);
  if (s->md.host) hipHostFree(s->md.host);
  for (auto &set : s->evsets)
    for (auto &e : set) hipEventDestroy(e);
  delete sys;
  return PMX_OK;
}

int pmx_sys_uncache(pmx_sys *sys) {
  if (!sys) return PMX_ERR_ARG;
  sys->s.cached = false;
  return PMX_OK;
}

int pmx_sys_set_positions(pmx_sys *sys, const double *x_aos) {
  if (!sys || !x_aos) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  hipSetDevice(s->ctx->device);
  for (size_t k = 0; k < 3 * (size_t)s->n; ++k)
    if (!std::isfinite(x_aos[k])) return fail(s->ctx, PMX_ERR_ARG, "non-finite position");
  HIPCHK(hipMemcpyAsync(s->xorig.p, x_aos, 24 * (size_t)s->n, hipMemcpyHostToDevice, s->ctx->stream));
  s->have_positions = true;
  return PMX_OK;
}

int pmx_sys_compute(pmx_sys *sys, unsigned flags) {
  if (!sys) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  if (!s->have_positions) return fail(s->ctx, PMX_ERR_ARG, "positions not set");
  hipSetDevice(s->ctx->device);
  if (!s->cached) {
    int rc = setup(s);
    if (rc) return rc;
  }
  return compute(s, flags);
}

int pmx_sys_compute_noalloc(pmx_sys *sys, unsigned flags) {
  if (!sys) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  if (flags & PMX_EVAL_TIMING)
    return fail(s->ctx, PMX_ERR_ARG, "pmx_sys_compute_noalloc: PMX_EVAL_TIMING needs event creation");
  if (!s->have_positions) return fail(s->ctx, PMX_ERR_ARG, "positions not set");
  if (!s->cached)
    return fail(s->ctx, PMX_ERR_ARG,
                "pmx_sys_compute_noalloc: tables dropped by pmx_sys_uncache (pmx_sys_compute rebuilds them)");
  if (!s->ctx->scratch) return fail(s->ctx, PMX_ERR_NODEV, "no scratch address for this device");
  // layout: counters (tile counter, boundary count, 4 boundary sums), then
  // boundary records, per-tile energies, pair counts and hashes
  auto up = [](size_t v) { return (v + 255) / 256 * 256; };
  const size_t nb = (size_t)s->maxborder, nt = (size_t)s->ntiles_cap;
  const size_t o_border = 256, o_pe = up(o_border + 8 * nb), o_np = up(o_pe + 16 * nt), o_h = up(o_np + 8 * nt),
               end = o_h + 8 * nt;
  if (end > kScratchBytes)
    return fail(s->ctx, PMX_ERR_ARG,
                "pmx_sys_compute_noalloc: the pass needs " + std::to_string(end) + " bytes of scratch, the maximum is " +
                    std::to_string(kScratchBytes) + " (about 1,000,000 atoms at a 12 A list radius)");
  hipSetDevice(s->ctx->device);
  unsigned char *base = static_cast<unsigned char *>(s->ctx->scratch);
  DevBuf *bufs[7] = {&s->tilectr, &s->nborder, &s->border_acc, &s->border, &s->pair_e, &s->npairs, &s->hash};
  void *const addr[7] = {base, base + 8, base + 16, base + o_border, base + o_pe, base + o_np, base + o_h};
  void *saved[7];
  for (int k = 0; k < 7; ++k) {
    saved[k] = bufs[k]->p;
    bufs[k]->p = addr[k];
  }
  // the counters keep the last call's values (another system may have used
  // them): clear them with a fill kernel, which graph capture also records
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->ctx->stream, reinterpret_cast<unsigned long long *>(base),
                     0ull, 6ll);
  const int rc = compute(s, flags);
  for (int k = 0; k < 7; ++k) bufs[k]->p = saved[k];
  return rc;
}

int pmx_sys_get_forces(pmx_sys *sys, double *f_aos, double *energies) {
  if (!sys) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  hipSetDevice(s->ctx->device);
  return finish(s, 0, f_aos, energies);
}

int pmx_sys_sync(pmx_sys *sys) {
  if (!sys) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return PMX_OK;
}

int pmx_sys_evaluate(pmx_sys *sys, const double *x_aos, double *f_aos, double *energies, unsigned flags) {
  if (!sys || !x_aos || !f_aos || !energies) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  int rc = pmx_sys_set_positions(sys, x_aos);
  if (rc) return rc;
  for (int attempt = 0; attempt < 6; ++attempt) {
    s->evused = 0;
    rc = pmx_sys_compute(sys, flags);
    if (rc) return rc;
    rc = finish(s, flags, f_aos, energies);
    if (rc != PMX_ERR_OVERFLOW) return rc;
    if ((rc = grow_and_retry(s))) return rc;
  }
  return fail(s->ctx, PMX_ERR_OVERFLOW, "cluster pair list capacity exceeded after retries");
}

int pmx_sys_alpha(const pmx_sys *sys, double *alpha) {
  if (!sys || !alpha) return PMX_ERR_ARG;
  *alpha = sys->s.alpha;
  return PMX_OK;
}

int pmx_sys_pair_stats(const pmx_sys *sys, long long *npairs, unsigned long long *hash) {
  if (!sys || !npairs || !hash) return PMX_ERR_ARG;
  *npairs = sys->s.last_np;
  *hash = sys->s.last_hash;
  return PMX_OK;
}

int pmx_sys_cells(pmx_sys *sys, int *cell_of_atom, int *dims) {
  if (!sys || !cell_of_atom || !dims) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  hipSetDevice(s->ctx->device);
  HIPCHK(hipMemcpyAsync(cell_of_atom, s->refcell.p, 4 * (size_t)s->n, hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  for (int k = 0; k < 3; ++k) dims[k] = s->bin.rc_n[k];
  return PMX_OK;
}

int pmx_sys_timing_begin(pmx_sys *sys) {
  if (!sys) return PMX_ERR_ARG;
  sys->s.evused = 0;
  sys->s.timing_active = true;
  return PMX_OK;
}

int pmx_sys_timing_end(pmx_sys *sys, float *avg_ms, int *ncalls) {
  if (!sys || !avg_ms || !ncalls) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  for (int k = 0; k < kStages; ++k) avg_ms[k] = 0.f;
  const size_t m = s->evused;
  for (size_t c = 0; c < m; ++c)
    for (int k = 0; k < kStages; ++k) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, s->evsets[c][kStageEv[k][0]], s->evsets[c][kStageEv[k][1]]));
      avg_ms[k] += t;
    }
  for (int k = 0; k < kStages; ++k) avg_ms[k] = m ? avg_ms[k] / (float)m : 0.f;
  *ncalls = (int)m;
  s->evused = 0;
  s->timing_active = false;
  return PMX_OK;
}

int pmx_sys_time_pair(pmx_sys *sys, int reps, float *ms_per_launch) {
  if (!sys || reps < 1 || reps > 4096 || !ms_per_launch) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  if (!s->have_positions) return fail(s->ctx, PMX_ERR_ARG, "no positions set");
  hipSetDevice(s->ctx->device);
  hipStream_t st = s->ctx->stream;
  if (int rc = compute(s, 0)) return rc;
  // each launch gets its own zeroed tile counter and boundary-pair count
  int *ctrs = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  {
    RtLock lk(g_rt_mu);
    if (hipMalloc(&ctrs, 8 * (size_t)reps) != hipSuccess) return fail(s->ctx, PMX_ERR_OOM, "hipMalloc failed");
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      if (e0) hipEventDestroy(e0);
      hipFree(ctrs);
      return fail(s->ctx, PMX_ERR_HIP, "hipEventCreate failed");
    }
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(nblk(2ll * reps, 256)), dim3(256), 0, st, ctrs, 0, 2ll * reps);
  void *const tc = s->tilectr.p, *const nb = s->nborder.p;
  hipEventRecord(e0, st);
  for (int r = 0; r < reps; ++r) {
    s->tilectr.p = ctrs + r;
    s->nborder.p = ctrs + reps + r;
    if (s->mixed)
      run_pair_phase<true>(s, false, 0);
    else
      run_pair_phase<false>(s, false, 0);
  }
  s->tilectr.p = tc;
  s->nborder.p = nb;
  hipEventRecord(e1, st);
  const hipError_t e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) hipEventElapsedTime(&t, e0, e1);
  {
    RtLock lk(g_rt_mu);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(ctrs);
  }
  s->acc_clean = false;  // the repeated launches added into the accumulators
  if (e != hipSuccess) return fail(s->ctx, PMX_ERR_HIP, hipGetErrorString(e));
  *ms_per_launch = t / (float)reps;
  return PMX_OK;
}

int pmx_sys_set_tuning(pmx_sys *sys, int knob, int value) {
  if (!sys) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  switch (knob) {
    case PMX_TUNE_FFT_LINES:
      if (value < 1 || value > 64) return fail(s->ctx, PMX_ERR_ARG, "FFT lines per block must be 1..64");
      s->tune_fft_lines = value;
      break;
    case PMX_TUNE_PAIR_BLOCKS_PER_CU:
      if (value < -7 || value > 64) return fail(s->ctx, PMX_ERR_ARG, "pair blocks per CU must be -7..64");
      s->tune_pair_blocks = value;
      break;
    case PMX_TUNE_PAIR_SCHEDULE:
      if (value < 0 || value > 1) return fail(s->ctx, PMX_ERR_ARG, "pair schedule must be 0 or 1");
      s->tune_schedule = value;
      break;
    case PMX_TUNE_SPREAD_STREAM:
      if (value < 0 || value > 1) return fail(s->ctx, PMX_ERR_ARG, "spread stream must be 0 or 1");
      s->tune_spread_stream2 = value;
      break;
    case PMX_TUNE_LIST_PRUNE:
      if (value < 0 || value > 1) return fail(s->ctx, PMX_ERR_ARG, "list prune must be 0 or 1");
      s->tune_prune = value;
      break;
    case PMX_TUNE_FIXUP_STREAM:
      if (value < 0 || value > 1) return fail(s->ctx, PMX_ERR_ARG, "fixup stream must be 0 or 1");
      s->tune_fixup_main = value;
      break;
    case PMX_TUNE_FUSED_SPREAD:
      if (value < 0 || value > 1) return fail(s->ctx, PMX_ERR_ARG, "fused spread must be 0 or 1");
      s->tune_fused_spread = value;
      break;
    case PMX_TUNE_FFT_PLANES:
      if (value < 0 || value > 1) return fail(s->ctx, PMX_ERR_ARG, "FFT planes must be 0 or 1");
      s->tune_fft_planes = value;
      break;
    default:
      return fail(s->ctx, PMX_ERR_ARG, "unknown tuning knob");
  }
  // captured chunks hold the old launch shapes; the energy partials of the
  // x pass may be reallocated
  hipSetDevice(s->ctx->device);
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream2));
  {
    RtLock lk(g_rt_mu);
    drop_graphs(s);
  }
  if (s->mixed)
    run_pair_phase<true>(s, false, -1);
  else
    run_pair_phase<false>(s, false, -1);
  // the tiles follow the schedule and the grid: cut the current list again
  // (the next MD step may use it without a rebuild)
  if (s->list_valid) {
    make_tiles(s, s->ctx->stream);
    HIPCHK(hipStreamSynchronize(s->ctx->stream));
  }
  return s->pme_on ? fft_blocking(s) : PMX_OK;
}

int pmx_sys_stage_times(const pmx_sys *sys, float *ms) {
  if (!sys || !ms) return PMX_ERR_ARG;
  for (int k = 0; k < kStages; ++k) ms[k] = sys->s.stage_ms[k];
  return PMX_OK;
}

// ---------------------------------------------------------------------------
// Device-resident MD

int pmx_md_init(pmx_sys *sys, const pmx_md_desc *d) {
  if (!sys || !d) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  hipSetDevice(s->ctx->device);
  if (d->nbonds < 0 || d->nangles < 0 || !d->mass || !(d->timestep_fs > 0) || d->skin < 0 || d->skin > 4.0 ||
      (d->nbonds > 0 && (!d->bonds || !d->bond_par)) || (d->nangles > 0 && (!d->angles || !d->angle_par)))
    return fail(s->ctx, PMX_ERR_ARG, "bad MD descriptor");
  if ((d->integrator != PMX_MD_LEAPFROG && d->integrator != PMX_MD_LANGEVIN) ||
      (d->integrator == PMX_MD_LANGEVIN && (!(d->temperature >= 0) || !(d->gamma >= 0))))
    return fail(s->ctx, PMX_ERR_ARG, "bad integrator parameters");
  for (int i = 0; i < s->n; ++i)
    if (!(d->mass[i] > 0)) return fail(s->ctx, PMX_ERR_ARG, "masses must be > 0");
  for (int b = 0; b < 2 * d->nbonds; ++b)
    if (d->bonds[b] < 0 || d->bonds[b] >= s->n) return fail(s->ctx, PMX_ERR_ARG, "bond atom out of range");
  for (int a = 0; a < 3 * d->nangles; ++a)
    if (d->angles[a] < 0 || d->angles[a] >= s->n) return fail(s->ctx, PMX_ERR_ARG, "angle atom out of range");
  double rcmax = std::max(s->pp.has_lj ? s->d.lj_rc : 0.0, s->d.coul_kind ? s->d.coul_rc : 0.0);
  for (int k = 0; k < 3; ++k)
    if (rcmax + d->skin > 0.5 * s->box.L[k])
      return fail(s->ctx, PMX_ERR_UNSUPPORTED, "cutoff + skin larger than half the box");
  Sys::Md &m = s->md;
  {
    RtLock lk(g_rt_mu);
    drop_graphs(s);
  }
  m.on = true;
  m.graphs = d->graphs != 0;
  m.nb = d->nbonds;
  m.na = d->nangles;
  m.h = d->timestep_fs * kInvTimeFactor;
  m.skin = d->skin;
  m.chunk = 16;
  m.steps = m.rebuilds = m.rollbacks = 0;
  m.forces_valid = false;
  m.integrator = d->integrator;
  m.seed = d->seed;
  if (d->integrator == PMX_MD_LANGEVIN) {
    // LangevinLeapfrogIntegrator.cpp:28 (gamma per internal time unit) and
    // :46-51; gamma = 0 takes the limits (plain leapfrog half kick, no noise)
    const double g = d->gamma / (1000.0 * kInvTimeFactor), dt = m.h;
    m.vdt = std::exp(-0.5 * g * dt);
    m.fdt = g > 0 ? (1.0 - std::exp(-0.5 * g * dt)) / g : 0.5 * dt;
    m.ndt = g > 0 ? std::sqrt((1.0 - std::exp(-g * dt)) / (2.0 * g)) : 0.0;
    m.fc = 2.0 * kBoltzmann * d->temperature * g;
  }
  const size_t N = s->n;
  std::vector<int2> b(m.nb);
  std::vector<double2> bp(m.nb);
  for (int i = 0; i < m.nb; ++i) {
    b[i] = make_int2(d->bonds[2 * i], d->bonds[2 * i + 1]);
    bp[i] = make_double2(d->bond_par[2 * i], d->bond_par[2 * i + 1]);
  }
  std::vector<int4> a(m.na);
  std::vector<double2> ap(m.na);
  for (int i = 0; i < m.na; ++i) {
    a[i] = make_int4(d->angles[3 * i], d->angles[3 * i + 1], d->angles[3 * i + 2], 0);
    ap[i] = make_double2(d->angle_par[2 * i], d->angle_par[2 * i + 1]);
  }
  std::vector<double> minv(N);
  for (size_t i = 0; i < N; ++i) minv[i] = 1.0 / d->mass[i];
  const int nbonded_blocks = std::max(1, (int)nblk(m.nb + m.na, 256));
  int rc;
  if ((rc = dalloc(s, m.bonds, 8 * b.size())) || (rc = dalloc(s, m.bpar, 16 * bp.size())) ||
      (rc = dalloc(s, m.angles, 16 * a.size())) || (rc = dalloc(s, m.apar, 16 * ap.size())) ||
      (rc = dalloc(s, m.minv, 8 * N)) || (rc = dalloc(s, m.v, 24 * N)) || (rc = dalloc(s, m.xbuild, 24 * N)) ||
      (rc = dalloc(s, m.xbak, 2 * 24 * N)) || (rc = dalloc(s, m.vbak, 2 * 24 * N)) ||
      (rc = dalloc(s, m.fbak, 2 * 24 * N)) || (rc = dalloc(s, m.dmax, 4 * 2 * kMaxChunk)) ||
      (rc = dalloc(s, m.step_dev, 8)) || (rc = dalloc(s, m.bond_e, 16 * (size_t)nbonded_blocks)) ||
      (rc = dalloc(s, m.ke_e, 8 * (size_t)nblk(N, 256))))
    return rc;
  if (!m.host) {
    RtLock lk(g_rt_mu);
    HIPCHK(hipHostMalloc((void **)&m.host, 4 * 2 * (kMaxChunk + 1), hipHostMallocDefault));
    for (auto &e : m.chunk_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  hipStream_t st = s->ctx->stream;
  if (m.nb) HIPCHK(hipMemcpyAsync(m.bonds.p, b.data(), 8 * b.size(), hipMemcpyHostToDevice, st));
  if (m.nb) HIPCHK(hipMemcpyAsync(m.bpar.p, bp.data(), 16 * bp.size(), hipMemcpyHostToDevice, st));
  if (m.na) HIPCHK(hipMemcpyAsync(m.angles.p, a.data(), 16 * a.size(), hipMemcpyHostToDevice, st));
  if (m.na) HIPCHK(hipMemcpyAsync(m.apar.p, ap.data(), 16 * ap.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(m.minv.p, minv.data(), 8 * N, hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  // the pair-list radius includes the skin
  s->cached = false;
  return setup(s);
}

int pmx_md_set_state(pmx_sys *sys, const double *x, const double *v) {
  if (!sys || !x || !v) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  if (!s->md.on) return fail(s->ctx, PMX_ERR_ARG, "pmx_md_init first");
  int rc = pmx_sys_set_positions(sys, x);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(s->md.v.p, v, 24 * (size_t)s->n, hipMemcpyHostToDevice, s->ctx->stream));
  s->md.forces_valid = false;
  return PMX_OK;
}

namespace {

// StandardIntegrator::initializeForces: forces at the current positions with
// a freshly built list.  A list capacity overflow grows the buffers and
// recomputes, so the forces the first step kicks with are complete.
int md_ensure_forces(Sys *s) {
  Sys::Md &m = s->md;
  if (m.forces_valid) return PMX_OK;
  hipStream_t st = s->ctx->stream;
  for (int tries = 0;; ++tries) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, P<int>(s->overflow), 0, 2ll);
    int rc = rebuild(s, false);
    if (rc) return rc;
    ++m.rebuilds;
    if ((rc = forces(s, 0, false))) return rc;
    int ovf = 0;
    HIPCHK(hipMemcpyAsync(&ovf, s->overflow.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!ovf) break;
    if (tries >= 6) return fail(s->ctx, PMX_ERR_OVERFLOW, "pair list capacity exceeded");
    if ((rc = grow_and_retry(s))) return rc;
  }
  m.forces_valid = true;
  return PMX_OK;
}

int md_check_overflow(Sys *s) {
  int ovf = 0;
  HIPCHK(hipMemcpyAsync(&ovf, s->overflow.p, 4, hipMemcpyDeviceToHost, s->ctx->stream));
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  if (ovf) return fail(s->ctx, PMX_ERR_OVERFLOW, "pair list capacity exceeded during MD");
  return PMX_OK;
}

// One leapfrog step: kick + drift (+ optional list rebuild at the new
// positions) + forces.  dslot: per-step displacement slot.
// The kicks before the drift of global step `step`: the first step of a run
// gets one half kick (kick = 0.5), later steps the previous step's closing
// half kick and their own opening one (kick = 1).  Langevin half kicks are
// numbered 2 step (opening) and 2 step + 1 (closing).
// With step_dev, `step` is relative to the step index stored there.
Kick make_kick(const Sys::Md &m, double kick, long long step, const unsigned long long *step_dev = nullptr) {
  Kick kk{};
  kk.h_kick = kick * m.h;
  kk.langevin = m.integrator == PMX_MD_LANGEVIN;
  kk.nhalf = kick == 0.5 ? 1 : 2;
  kk.half0 = (unsigned long long)(kick == 0.5 ? 2 * step : 2 * step - 1);
  kk.seed = m.seed;
  kk.vdt = m.vdt;
  kk.fdt = m.fdt;
  kk.ndt = m.ndt;
  kk.fc = m.fc;
  kk.step_dev = step_dev;
  return kk;
}

// kicked: the previous step's force pass already did this step's kick and
// drift; next (with next_dslot): fuse the following step's into this one's.
int md_step(Sys *s, double kick, long long step, bool rebuild_here, int dslot,
            const unsigned long long *step_dev = nullptr, bool kicked = false, const Kick *next = nullptr,
            int next_dslot = 0) {
  hipStream_t st = s->ctx->stream;
  Sys::Md &m = s->md;
  const bool timing = s->timing_active;
  const Kick kk = make_kick(m, kick, step, step_dev);
  // stage 0 of a timed MD step is integrate (+ sort when the list is rebuilt)
  timing_begin_call(s, timing);
  if (kicked)
    ;
  else if (s->mixed)
    hipLaunchKernelGGL(k_kick_drift<float>, dim3(nblk(s->npad, 256)), dim3(256), 0, st, s->npad, kk, m.h,
                       P<int>(s->permp), P<double>(s->fout), P<double>(m.minv), P<double>(m.v), P<double>(s->xorig),
                       P<double>(m.xbuild), P<float4>(s->xs), P<double4>(s->xs64),
                       P<float4>(s->sorg), P<unsigned>(m.dmax) + dslot);
  else
    hipLaunchKernelGGL(k_kick_drift<double>, dim3(nblk(s->npad, 256)), dim3(256), 0, st, s->npad, kk, m.h,
                       P<int>(s->permp), P<double>(s->fout), P<double>(m.minv), P<double>(m.v), P<double>(s->xorig),
                       P<double>(m.xbuild), P<double4>(s->xs), P<double4>(s->xs64),
                       P<float4>(s->sorg), P<unsigned>(m.dmax) + dslot);
  if (rebuild_here) {
    int rc = rebuild(s, timing);
    if (rc) return rc;
  } else if (timing) {
    hipEventRecord(s->ev[1], st);
    hipEventRecord(s->ev[2], st);
  }
  // the previous pass's kick-drift spread these positions when it was fused
  const bool spread_done = kicked && s->spread_ready;
  return forces(s, timing ? PMX_EVAL_TIMING : 0u, false, next, next ? P<unsigned>(m.dmax) + next_dslot : nullptr,
                spread_done);
}

// A chunk of K steps from the committed state: its state is backed up in
// slot `slot`, step 0 rebuilds the list, and the per-step displacements and
// the overflow flag are copied to the slot's pinned host record, after which
// the slot's event fires.
struct Chunk {
  int K;
  long long step0;
  bool first;
  int slot;
};

// The work of a chunk.  rel: step indices relative to *step_dev (capture).
int chunk_body(Sys *s, const Chunk &c, bool rel) {
  Sys::Md &m = s->md;
  hipStream_t st = s->ctx->stream;
  const size_t n = s->n, off = c.slot * 3 * n;
  // backups of the committed state, cleared overflow flags and displacement
  // record, in one launch
  const long long m3 = 3ll * (long long)n;
  hipLaunchKernelGGL(k_chunk_begin, dim3(nblk(std::max<long long>(m3, kMaxChunk), 256)), dim3(256), 0, st, m3,
                     P<double>(s->xorig), P<double>(m.v), P<double>(s->fout), P<double>(m.xbak) + off,
                     P<double>(m.vbak) + off, P<double>(m.fbak) + off, P<int>(s->overflow),
                     P<unsigned>(m.dmax) + c.slot * kMaxChunk, kMaxChunk);
  const unsigned long long *sd = rel ? P<unsigned long long>(m.step_dev) : nullptr;
  // inside the chunk each force pass ends with the next step's kick and
  // drift (one launch instead of two); timed steps keep them apart
  const bool fuse = !s->timing_active;
  for (int k = 0; k < c.K; ++k) {
    Kick nk{};
    const bool nx = fuse && k + 1 < c.K;
    if (nx) nk = make_kick(m, 1.0, rel ? k + 1 : c.step0 + k + 1, sd);
    int rc = md_step(s, (c.first && k == 0) ? 0.5 : 1.0, rel ? k : c.step0 + k, k == 0 || m.skin == 0.0,
                     c.slot * kMaxChunk + k, sd, fuse && k > 0, nx ? &nk : nullptr, c.slot * kMaxChunk + k + 1);
    if (rc) return rc;
  }
  unsigned *h = m.host + c.slot * (kMaxChunk + 1);
  HIPCHK(hipMemcpyAsync(h, P<uint32_t>(m.dmax) + c.slot * kMaxChunk, 4 * (size_t)c.K, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(h + kMaxChunk, s->overflow.p, 4, hipMemcpyDeviceToHost, st));
  return PMX_OK;
}

constexpr size_t kMaxGraphs = 64;

// Queues a chunk: replays its captured graph (capturing it on first use), or
// launches its kernels one by one when graphs are off or stage timing is on.
int enqueue_chunk(Sys *s, const Chunk &c) {
  Sys::Md &m = s->md;
  hipStream_t st = s->ctx->stream;
  hipGraphExec_t x = nullptr;
  // only chunks whose length is a power of two are captured (the others are
  // launched kernel by kernel), so a system holds at most 2 x 2 x 8 graphs
  if (m.graphs && !s->timing_active && (c.K & (c.K - 1)) == 0) {
    const long long key = ((long long)c.K << 2) | ((long long)c.first << 1) | c.slot;
    auto it = m.gexec.find(key);
    if (it != m.gexec.end()) {
      x = it->second;
    } else {
      RtLock lk(g_rt_mu);
      if (m.gexec.size() >= kMaxGraphs) {
        // a launched graph may still run: wait before destroying
        HIPCHK(hipStreamSynchronize(st));
        drop_graphs(s);
      }
      hipGraph_t g = nullptr;
      bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess;
      if (ok) {
        const int rc = chunk_body(s, c, true);
        ok = hipStreamEndCapture(st, &g) == hipSuccess && rc == PMX_OK && g;
      }
      if (ok) ok = hipGraphInstantiate(&x, g, nullptr, nullptr, 0) == hipSuccess;
      if (g) hipGraphDestroy(g);
      if (ok) {
        m.gexec.emplace(key, x);
      } else {
        // nothing captured runs: take this system off graphs and launch the
        // chunk's kernels one by one instead
        x = nullptr;
        m.graphs = false;
        (void)hipGetLastError();
      }
    }
  }
  if (x) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, P<unsigned long long>(m.step_dev),
                       (unsigned long long)c.step0);
    HIPCHK(hipGraphLaunch(x, st));
  } else if (int rc = chunk_body(s, c, false)) {
    return rc;
  }
  HIPCHK(hipEventRecord(m.chunk_ev[c.slot], st));
  return PMX_OK;
}

// Waits for everything queued (a speculative chunk included), then restores
// the state backed up in slot `slot`.
int restore_chunk(Sys *s, int slot) {
  Sys::Md &m = s->md;
  hipStream_t st = s->ctx->stream;
  HIPCHK(hipStreamSynchronize(st));
  const size_t n = s->n, off = slot * 3 * n;
  HIPCHK(hipMemcpyAsync(s->xorig.p, P<double>(m.xbak) + off, 24 * n, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(m.v.p, P<double>(m.vbak) + off, 24 * n, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(s->fout.p, P<double>(m.fbak) + off, 24 * n, hipMemcpyDeviceToDevice, st));
  return PMX_OK;
}

// Length of the next chunk from the largest squared displacement d2 that
// the last chunk reached at step k after its list build.  Displacement is
// taken to grow linearly with time (the fastest atoms move ballistically
// over a few steps; diffusive motion grows slower, so the estimate errs
// short): the list stays valid up to step L = k sqrt(lim2 / d2).  The chunk
// ends at 80 % of L, since a chunk that outlives its list is rolled back and
// the chunk queued behind it discarded, and grows at most 2x per chunk.
int next_chunk(int k, float d2, float lim2, int cur) {
  const int grow = std::min(kMaxChunk, 2 * std::max(1, cur));
  if (!(d2 > 0.f)) return grow;
  const double L = k * std::sqrt((double)lim2 / (double)d2);
  return std::max(1, std::min(grow, (int)std::floor(0.8 * L) + 1));
}

// Progress of one pmx_md_run on one system (see md_advance).
struct MdRun {
  int nsteps = 0;
  int done = 0;     // committed steps of this run
  int planned = 0;  // committed + queued
  int next_slot = 0;
  int overflow_retries = 0;
  std::deque<Chunk> inflight;
  bool finished = false;
};

int md_begin(Sys *s, MdRun &r, int nsteps) {
  Sys::Md &m = s->md;
  if (!m.on || !s->have_positions) return fail(s->ctx, PMX_ERR_ARG, "pmx_md_init and pmx_md_set_state first");
  r = MdRun{};
  r.nsteps = nsteps;
  hipSetDevice(s->ctx->device);
  if (nsteps == 0) {
    r.finished = true;
    return PMX_OK;
  }
  return md_ensure_forces(s);
}

// The steps run in chunks that each start with a list rebuild.  Two chunks
// are in flight: while the device runs chunk c + 1 (queued on the assumption
// that chunk c keeps every atom within half the skin of its build position),
// the host reads chunk c's displacements.  If chunk c crossed the limit at
// step b, the speculative chunk is discarded, the state of chunk c's start is
// restored and its first b steps (which used a valid list) are replayed; the
// trajectory is the same as with a rebuild at every step.  The device never
// waits for the host between chunks.
//
// One call queues chunks up to two in flight, then settles the oldest one if
// it has completed (waiting for it when `block`).  *progressed tells whether
// anything was queued or settled.  After the last step the closing half kick
// is queued and r.finished is set.
int md_advance(Sys *s, MdRun &r, bool block, bool *progressed) {
  Sys::Md &m = s->md;
  hipStream_t st = s->ctx->stream;
  const int n = s->n;
  const float lim2 = (float)(0.25 * m.skin * m.skin);
  *progressed = false;
  if (r.finished) return PMX_OK;
  while (r.inflight.size() < 2 && r.planned < r.nsteps) {
    Chunk c{std::min(m.chunk, r.nsteps - r.planned), m.steps + (r.planned - r.done), r.planned == 0, r.next_slot};
    if (int rc = enqueue_chunk(s, c)) return rc;
    r.inflight.push_back(c);
    r.planned += c.K;
    r.next_slot ^= 1;
    *progressed = true;
  }
  if (!r.inflight.empty()) {
    const Chunk c = r.inflight.front();
    if (block) {
      HIPCHK(hipEventSynchronize(m.chunk_ev[c.slot]));
    } else {
      const hipError_t q = hipEventQuery(m.chunk_ev[c.slot]);
      if (q == hipErrorNotReady) return PMX_OK;
      if (q != hipSuccess) return fail(s->ctx, PMX_ERR_HIP, hipGetErrorString(q));
    }
    r.inflight.pop_front();
    *progressed = true;
    const unsigned *h = m.host + c.slot * (kMaxChunk + 1);
    if (h[kMaxChunk]) {
      // a list or mask capacity was exceeded: grow, roll back, replay
      if (++r.overflow_retries > 6) return fail(s->ctx, PMX_ERR_OVERFLOW, "pair list capacity exceeded during MD");
      if (int rc = restore_chunk(s, c.slot)) return rc;
      if (int rc = grow_and_retry(s)) return rc;
      r.inflight.clear();
      r.planned = r.done;
      return PMX_OK;
    }
    int bad = -1;
    float worst = 0.f;
    for (int k = 1; k < c.K; ++k) {
      const float d = __builtin_bit_cast(float, h[k]);
      worst = std::max(worst, d);
      if (m.skin > 0 && d > lim2) {
        bad = k;
        break;
      }
    }
    if (bad < 0) {
      r.done += c.K;
      m.steps += c.K;
      m.rebuilds += m.skin > 0 ? 1 : c.K;
      if (m.skin > 0 && c.K > 1) m.chunk = next_chunk(c.K - 1, worst, lim2, m.chunk);
    } else {
      // roll back to the chunk's start and replay the steps that used a
      // valid list (deterministic: the replay reproduces them exactly)
      ++m.rollbacks;
      if (int rc = restore_chunk(s, c.slot)) return rc;
      r.inflight.clear();
      for (int k = 0; k < bad; ++k) {
        int rc = md_step(s, (c.first && k == 0) ? 0.5 : 1.0, c.step0 + k, k == 0, k);
        if (rc) return rc;
      }
      r.done += bad;
      m.steps += bad;
      m.rebuilds += 1;
      r.planned = r.done;
      m.chunk = next_chunk(bad, worst, lim2, m.chunk);
    }
  }
  if (r.done == r.nsteps && r.inflight.empty()) {
    // closing half kick of the last step (global index m.steps - 1)
    Kick kk = make_kick(m, 0.5, m.steps);
    kk.half0 = (unsigned long long)(2 * m.steps - 1);
    hipLaunchKernelGGL(k_half_kick, dim3(nblk(n, 256)), dim3(256), 0, st, n, kk, P<double>(s->fout),
                       P<double>(m.minv), P<double>(m.v));
    r.finished = true;
  }
  return PMX_OK;
}

int md_end(Sys *s) {
  HIPCHK(hipStreamSynchronize(s->ctx->stream));
  return md_check_overflow(s);
}

// ---- batched MD: every kernel of a step launched once for many systems ----
// (C++ linkage for the templates inside the C ABI's extern "C" block)
extern "C++" {

#define BHIPCHK(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t _e = (expr);                                                                             \
    if (_e != hipSuccess) return fail(b.ctx, PMX_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

// A device block of the batch (from the block cache, given back at the end).
struct BatchBuf {
  int device = 0;
  void *p = nullptr;
  size_t cls = 0;
  BatchBuf() = default;
  BatchBuf(const BatchBuf &) = delete;
  BatchBuf &operator=(const BatchBuf &) = delete;
  ~BatchBuf() { release(); }
  void release() {
    if (!p) return;
    RtLock lk(g_rt_mu);
    block_put(device, p, cls);
    p = nullptr;
  }
  bool get(int dev, size_t bytes) {
    release();
    device = dev;
    cls = size_class(std::max<size_t>(bytes, 8));
    RtLock lk(g_rt_mu);
    p = block_get(dev, cls);
    return p != nullptr;
  }
};

// Smallest run of a list built in batched launches, in list-cost units (about
// 35 ns of wave time each): a system's runs share the batch's waves with
// every other system's, so its list is cut into runs of real work rather than
// into as many runs as the device has resident waves (the single-system cut,
// which keeps a small system's pass short alone).  The runs, and so a mixed
// result, are those of the same list in any batch.
constexpr long long kBatchRunCost = 2048;

// One batched launch: the argument table, each system's first block and the
// system of each block.
struct BatchLaunch {
  BatchBuf tab, first, b2s;
  unsigned blocks = 0;
  size_t lds = 0;
};

struct Batch {
  pmx_ctx *ctx = nullptr;
  std::vector<Sys *> sys;
  int nsys = 0;
  BatchLaunch bin, scan, scatter, binrank, columns, split, list, excl, masks, rprefix, rcut, rtiles;
  BatchLaunch kick, spread, pair, zr2c, yfwd, xconv, yinv, plane, zc2r, gather, fixups, ukds, unsort, begin, half;
  BatchLaunch energy;
  BatchBuf bsys, need, base, flags, rebuilds;
  // slices of the context's pinned staging, and its events
  BatchSys *bsys_h = nullptr;
  unsigned long long *base_h = nullptr;
  int *flags_h = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// Uploads the table of one batched launch (args[s], nblocks[s] blocks of
// system s; dynamic LDS `lds`).
template <class Op>
int batch_table(Batch &b, BatchLaunch &L, const std::vector<typename Op::Args> &args,
                const std::vector<int> &nblocks, size_t lds = 0) {
  std::vector<int> first(b.nsys + 1, 0);
  for (int i = 0; i < b.nsys; ++i) first[i + 1] = first[i] + nblocks[i];
  std::vector<int> b2s(std::max(1, first[b.nsys]));
  for (int i = 0; i < b.nsys; ++i)
    for (int k = first[i]; k < first[i + 1]; ++k) b2s[k] = i;
  L.blocks = (unsigned)first[b.nsys];
  L.lds = lds;
  const int dev = b.ctx->device;
  if (!L.tab.get(dev, sizeof(typename Op::Args) * args.size()) || !L.first.get(dev, 4 * first.size()) ||
      !L.b2s.get(dev, 4 * b2s.size()))
    return fail(b.ctx, PMX_ERR_OOM, "batch tables");
  RtLock lk(g_rt_mu);
  BHIPCHK(hipMemcpy(L.tab.p, args.data(), sizeof(typename Op::Args) * args.size(), hipMemcpyHostToDevice));
  BHIPCHK(hipMemcpy(L.first.p, first.data(), 4 * first.size(), hipMemcpyHostToDevice));
  BHIPCHK(hipMemcpy(L.b2s.p, b2s.data(), 4 * b2s.size(), hipMemcpyHostToDevice));
  return PMX_OK;
}

template <class Op>
void batch_go(const BatchLaunch &L, hipStream_t st, const BatchStep &bs) {
  if (!L.blocks) return;
  hipLaunchKernelGGL(k_batch<Op>, dim3(L.blocks), dim3(Op::kBlock), L.lds, st,
                     reinterpret_cast<const typename Op::Args *>(L.tab.p), reinterpret_cast<const int *>(L.b2s.p),
                     reinterpret_cast<const int *>(L.first.p), bs);
}

// The batch's kernel case: the pair kernel's cases of run_pair_phase for PME
// systems (0: LJ C2 + Coulomb cutoff switch, 1: no LJ, 2: any switches, 3:
// LJ table beyond the kernel's LDS).
int batch_case(const Sys *s) {
  const int lj = s->d.lj_switch, cs = s->d.coul_switch;
  if (s->d.ntypes > kMaxTypes && lj != SW_NONE) return 3;
  if (lj == SW_C2 && cs == SW_CUTOFF) return 0;
  return lj == SW_NONE ? 1 : 2;
}

template <bool MIXED, int C, int LJ, int CS, bool BIGLJ, int ORDER>
using PairOp = OpPair<MIXED, C, LJ, CS, BIGLJ>;

// Fills every table of the batch from its systems (again after a system's
// buffers grow).
template <bool MIXED, int C, int LJ, int CS, bool BIGLJ, int ORDER>
int batch_tables(Batch &b) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using R4 = typename Real4<R>::type;
  using T = R;
  using Acc = typename AccT<MIXED>::type;
  const int ns = b.nsys;
  const size_t cx = sizeof(Cx<T>);
  std::vector<int> nb(ns);
  int rc;
  {
    std::vector<OpBin::Args> args_bin(ns);
    std::vector<OpScanSingle::Args> args_scan(ns);
    std::vector<OpScatter::Args> args_scatter(ns);
    std::vector<OpBinrank::Args> args_binrank(ns);
    std::vector<OpColumns::Args> args_columns(ns);
    std::vector<typename OpSplitFill<R>::Args> args_split(ns);
    std::vector<OpBuildList::Args> args_list(ns);
    std::vector<OpMarkExclusions::Args> args_excl(ns);
    std::vector<OpBuildMasks::Args> args_masks(ns);
    std::vector<OpRunsPrefix::Args> args_rprefix(ns);
    std::vector<OpRunsCut::Args> args_rcut(ns);
    std::vector<OpRunsTiles::Args> args_rtiles(ns);
    for (int i = 0; i < ns; ++i) {
      Sys *s = b.sys[i];
      args_bin[i] = {s->n, P<double>(s->xorig), s->bin, P<uint32_t>(s->key), P<uint32_t>(s->hist),
                     P<double>(s->xwrap), P<int>(s->refcell)};
      args_scan[i] = {P<uint32_t>(s->hist), P<uint32_t>(s->start), P<uint32_t>(s->cursor), s->nbins};
      args_scatter[i] = {s->n, P<uint32_t>(s->key), P<uint32_t>(s->cursor), P<int>(s->perm)};
      args_binrank[i] = {s->n, s->nbins, P<uint32_t>(s->key), P<uint32_t>(s->start), P<int>(s->perm),
                         P<int>(s->perm2)};
      args_columns[i] = {s->ncols, s->bin.nzb, s->n, s->nsuper, P<uint32_t>(s->start), P<int2>(s->sc_info),
                         P<int>(s->sc_col), P<int>(s->colsc)};
      auto &sp = args_split[i];
      sp.nsuper = s->nsuper;
      sp.sc_info = P<int2>(s->sc_info);
      sp.order = P<int>(s->perm2);
      sp.xwrap = P<double>(s->xwrap);
      sp.q = P<double>(s->q_d);
      sp.permp = P<int>(s->permp);
      sp.slot = P<int>(s->slot);
      sp.type = P<int>(s->type_d);
      sp.xs = P<R4>(s->xs);
      sp.xs64 = P<double4>(s->xs64);
      sp.ts = P<int>(s->ts);
      sp.xorig = P<double>(s->xorig);
      sp.xbuild = P<double>(s->md.xbuild);
      sp.sorg = P<float4>(s->sorg);
      sp.ctr = P<float4>(s->ctr);
      sp.half = P<float4>(s->half);
      sp.sctr = P<float4>(s->sctr);
      sp.shalf = P<float4>(s->shalf);
      sp.nmaskwork = P<int>(s->nmaskwork);
      sp.qpair = s->qpair;
      args_list[i] = {s->lp, P<int>(s->sc_col), P<int>(s->colsc), P<float4>(s->sctr), P<float4>(s->shalf),
                      P<float4>(s->ctr), P<float4>(s->half), P<int4>(s->entries), P<int>(s->count),
                      P<int>(s->ecost), P<int>(s->lcost), P<int>(s->maskwork), P<int>(s->nmaskwork),
                      s->maxmaskwork, P<int>(s->overflow)};
      args_excl[i] = {(int)s->etab.size(), s->cap, s->maxmaskwork, P<int4>(s->etab_d), P<int>(s->slot),
                      P<int>(s->count), P<int4>(s->entries), P<int>(s->maskwork), P<int>(s->nmaskwork),
                      P<int>(s->overflow)};
      args_masks[i] = {P<int>(s->maskwork), P<int>(s->nmaskwork), s->maxmaskwork, s->cap, P<int4>(s->entries),
                       P<int>(s->permp), P<int>(s->eptr_d), P<int2>(s->elist_d), P<unsigned short>(s->masks)};
      int *const nruns = P<int>(s->ntiles) + 2;
      args_rprefix[i] = {s->nsuper, P<int>(s->count), P<int>(s->lcost), nruns, P<int>(s->tep), P<long long>(s->tcp),
                         P<int>(s->tpos), P<int>(s->tposS), kBatchRunCost};
      args_rcut[i] = {s->nsuper, s->cap, P<int>(s->count), P<int>(s->ecost), nruns, P<int>(s->tep),
                      P<long long>(s->tcp), P<int>(s->tpos), P<int>(s->tposS)};
      args_rtiles[i] = {P<int>(s->count), P<int>(s->tep), P<int>(s->tpos), P<int>(s->tposS), P<int>(s->wstart),
                        P<int4>(s->tiles), P<int>(s->ntiles), s->ntiles_cap};
    }
    for (int i = 0; i < ns; ++i) nb[i] = (int)nblk(b.sys[i]->n, 256);
    if ((rc = batch_table<OpBin>(b, b.bin, args_bin, nb, 0))) return rc;
    if ((rc = batch_table<OpScatter>(b, b.scatter, args_scatter, nb, 0))) return rc;
    if ((rc = batch_table<OpBinrank>(b, b.binrank, args_binrank, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = 1;
    if ((rc = batch_table<OpScanSingle>(b, b.scan, args_scan, nb, 0))) return rc;
    if ((rc = batch_table<OpColumns>(b, b.columns, args_columns, nb, 0))) return rc;
    if ((rc = batch_table<OpRunsPrefix>(b, b.rprefix, args_rprefix, nb, 0))) return rc;
    if ((rc = batch_table<OpRunsTiles>(b, b.rtiles, args_rtiles, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = (int)nblk((long long)b.sys[i]->nsuper * 64, 256);
    if ((rc = batch_table<OpSplitFill<R>>(b, b.split, args_split, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = b.sys[i]->nsuper;
    if ((rc = batch_table<OpBuildList>(b, b.list, args_list, nb, 0))) return rc;
    if ((rc = batch_table<OpRunsCut>(b, b.rcut, args_rcut, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = b.sys[i]->etab.empty() ? 0 : (int)nblk((long long)b.sys[i]->etab.size(), 256);
    if ((rc = batch_table<OpMarkExclusions>(b, b.excl, args_excl, nb, 0))) return rc;
    // a wave per mask slot in use (known on the device only), up to 64 blocks
    for (int i = 0; i < ns; ++i) nb[i] = (int)std::min<long long>(nblk((long long)b.sys[i]->maxmaskwork, 4), 64);
    if ((rc = batch_table<OpBuildMasks>(b, b.masks, args_masks, nb, 0))) return rc;
  }
  {
    std::vector<typename OpKickDrift<R>::Args> args_kick(ns);
    std::vector<typename OpSpreadPatch<MIXED, ORDER>::Args> args_spread(ns);
    std::vector<typename PairOp<MIXED, C, LJ, CS, BIGLJ, ORDER>::Args> args_pair(ns);
    std::vector<typename OpFftZr2c<T, MIXED>::Args> args_zr2c(ns);
    std::vector<typename OpFftY<T>::Args> args_yfwd(ns), args_yinv(ns);
    std::vector<typename OpFftXConv<T>::Args> args_xconv(ns);
    std::vector<typename OpFftPlane<T>::Args> args_plane(ns);
    std::vector<typename OpFftZc2r<T>::Args> args_zc2r(ns);
    std::vector<typename OpGatherPatch<MIXED, ORDER>::Args> args_gather(ns);
    std::vector<FixupArgs<MIXED>> args_fixups(ns);
    std::vector<typename OpUnsortKickDriftSpread<MIXED, R, ORDER>::Args> args_ukds(ns);
    std::vector<typename OpUnsortForces<MIXED>::Args> args_unsort(ns);
    std::vector<OpChunkBegin::Args> args_begin(ns);
    std::vector<OpHalfKick::Args> args_half(ns);
    std::vector<OpReduceEnergy::Args> args_energy(ns);
    std::vector<int> nfix(ns);
    long long natoms = 0;
    size_t lds_z = 0, lds_y = 0, lds_x = 0, lds_p = 0;
    for (int i = 0; i < ns; ++i) {
      Sys *s = b.sys[i];
      Sys::Md &m = s->md;
      natoms += s->n;
      const Kick kk = make_kick(m, 1.0, 0, P<unsigned long long>(m.step_dev));
      Acc *facc = P<Acc>(s->facc);
      args_kick[i] = {s->npad, kk, m.h, P<int>(s->permp), P<double>(s->fout), P<double>(m.minv), P<double>(m.xbuild),
                      P<double>(m.v), P<double>(s->xorig), P<R4>(s->xs), P<double4>(s->xs64), P<float4>(s->sorg),
                      P<unsigned>(m.dmax)};
      args_spread[i] = {s->npad, s->pme, P<double>(s->xorig), P<double>(s->q_d), P<int>(s->permp),
                        P<double>(s->grid64), P<unsigned long long>(s->gridfx)};
      const R4 *ljt = MIXED ? reinterpret_cast<const R4 *>(s->ljt32.p) : reinterpret_cast<const R4 *>(s->ljtab.p);
      args_pair[i] = {s->pp, P<int4>(s->tiles), P<int>(s->ntiles), P<int>(s->wstart), P<int>(s->tilectr),
                      P<int4>(s->entries), s->cap, P<unsigned short>(s->masks), P<R4>(s->xs), P<float4>(s->sorg),
                      P<int>(s->ts), P<int>(s->permp), P<double>(s->ljtab), ljt, facc,
                      PairOut{P<double>(s->pair_e), P<long long>(s->npairs), P<unsigned long long>(s->hash),
                              P<int2>(s->border), P<int>(s->nborder), s->maxborder, P<int>(s->overflow),
                              P<unsigned long long>(s->border_acc)}};
      const PmeParams &pm = s->pme;
      args_zr2c[i] = {pm, s->plz, s->zl, P<double>(s->grid64), P<unsigned long long>(s->gridfx), P<Cx<T>>(s->twz),
                      P<Cx<T>>(s->spec)};
      args_yfwd[i] = {pm, s->ply, s->tky, false, P<Cx<T>>(s->twy), P<Cx<T>>(s->spec), P<Cx<T>>(s->spec2)};
      args_xconv[i] = {pm, s->plx, s->tkx, P<Cx<T>>(s->twx), P<T>(s->theta), P<Cx<T>>(s->spec2),
                       P<double>(s->recip_e)};
      args_yinv[i] = {pm, s->ply, s->tky, true, P<Cx<T>>(s->twy), P<Cx<T>>(s->spec2), P<Cx<T>>(s->spec)};
      args_plane[i] = {pm, s->ply, s->plx, P<Cx<T>>(s->twy), P<Cx<T>>(s->twx), P<T>(s->theta), P<Cx<T>>(s->spec),
                       P<double>(s->recip_e)};
      args_zc2r[i] = {pm, s->plz, s->zl, P<Cx<T>>(s->twz), P<Cx<T>>(s->spec), P<T>(s->greal)};
      args_gather[i] = {s->npad, pm, P<double>(s->xorig), P<double>(s->q_d), P<int>(s->permp), P<T>(s->greal), facc};
      args_fixups[i] = fixup_args<MIXED>(s, false, true, true, &nfix[i]);
      args_ukds[i] = {s->npad, kk, m.h, P<int>(s->permp), facc, P<double>(s->fout), P<int>(s->tilectr),
                      P<int>(s->nborder), P<double>(m.minv), P<double>(m.xbuild), P<double>(s->q_d), P<double>(m.v),
                      P<double>(s->xorig), P<R4>(s->xs), P<double4>(s->xs64), P<float4>(s->sorg),
                      P<unsigned>(m.dmax), pm, P<double>(s->grid64), P<unsigned long long>(s->gridfx)};
      args_unsort[i] = {s->n, P<int>(s->slot), facc, P<double>(s->fout), P<int>(s->tilectr), P<int>(s->nborder)};
      args_begin[i] = {3ll * s->n, P<double>(s->xorig), P<double>(m.v), P<double>(s->fout), P<double>(m.xbak),
                       P<double>(m.vbak), P<double>(m.fbak), P<int>(s->overflow), P<unsigned>(m.dmax), kMaxChunk};
      args_half[i] = {s->n, kk, m.h, P<double>(s->fout), P<double>(m.minv), P<double>(m.v)};
      const bool recip = s->pme_on && (s->d.pme_terms & PMX_PME_RECIP);
      const bool corr = s->pme_on && (s->d.pme_terms & PMX_PME_CORR) && !s->etab.empty();
      args_energy[i] = {s->ntiles_cap, P<int>(s->ntiles), P<double>(s->pair_e), s->mixed ? 1 : 0,
                        recip ? s->recip_blocks : 0, corr ? s->corr_blocks : 0, P<double>(s->recip_e),
                        P<double>(s->corr_e), P<long long>(s->npairs), P<unsigned long long>(s->hash),
                        P<unsigned long long>(s->border_acc), P<unsigned long long>(s->border_acc) + 2,
                        P<double>(s->eout), P<long long>(s->np_out), P<unsigned long long>(s->hash_out)};
      // dynamic LDS of each launch: the largest of the systems it serves
      lds_z = std::max(lds_z, 2 * (size_t)s->zl * pm.nz * cx);
      if (s->fft_plane) {
        lds_p = std::max(lds_p, 2 * (size_t)pm.nx * pm.ny * cx);
      } else {
        lds_y = std::max(lds_y, 2 * (size_t)s->tky * pm.ny * cx);
        lds_x = std::max(lds_x, 2 * (size_t)s->tkx * pm.nx * cx);
      }
    }
    for (int i = 0; i < ns; ++i) nb[i] = (int)nblk(b.sys[i]->npad, 256);
    if ((rc = batch_table<OpKickDrift<R>>(b, b.kick, args_kick, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = (int)nblk(b.sys[i]->npad, kSpreadSlots);
    if ((rc = batch_table<OpSpreadPatch<MIXED, ORDER>>(b, b.spread, args_spread, nb, 0))) return rc;
    if ((rc = batch_table<OpGatherPatch<MIXED, ORDER>>(b, b.gather, args_gather, nb, 0))) return rc;
    if ((rc = batch_table<OpUnsortKickDriftSpread<MIXED, R, ORDER>>(b, b.ukds, args_ukds, nb, 0))) return rc;
    // pair-kernel blocks in proportion to the atoms (a run's share of a
    // system's waves changes no result), about the device's resident blocks
    // in all
    static const int per_cu =
        pair_blocks_per_cu(reinterpret_cast<const void *>(k_batch<PairOp<MIXED, C, LJ, CS, BIGLJ, ORDER>>));
    const double pair_blocks = (double)b.ctx->cus * per_cu;
    for (int i = 0; i < ns; ++i)
      nb[i] = std::max(1, (int)std::lround(pair_blocks * b.sys[i]->n / (double)natoms));
    if ((rc = batch_table<PairOp<MIXED, C, LJ, CS, BIGLJ, ORDER>>(b, b.pair, args_pair, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) {
      const Sys *s = b.sys[i];
      nb[i] = s->pme.nx * ((s->pme.ny + s->zl - 1) / s->zl);
    }
    if ((rc = batch_table<OpFftZr2c<T, MIXED>>(b, b.zr2c, args_zr2c, nb, lds_z))) return rc;
    if ((rc = batch_table<OpFftZc2r<T>>(b, b.zc2r, args_zc2r, nb, lds_z))) return rc;
    // systems with small planes take the one-launch plane pass, the others
    // the three line passes
    for (int i = 0; i < ns; ++i) {
      const Sys *s = b.sys[i];
      nb[i] = s->fft_plane ? 0 : s->pme.nzh * ((s->pme.nx + s->tky - 1) / s->tky);
    }
    if ((rc = batch_table<OpFftY<T>>(b, b.yfwd, args_yfwd, nb, lds_y))) return rc;
    if ((rc = batch_table<OpFftY<T>>(b, b.yinv, args_yinv, nb, lds_y))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = b.sys[i]->fft_plane ? 0 : b.sys[i]->recip_blocks;
    if ((rc = batch_table<OpFftXConv<T>>(b, b.xconv, args_xconv, nb, lds_x))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = b.sys[i]->fft_plane ? b.sys[i]->pme.nzh : 0;
    if ((rc = batch_table<OpFftPlane<T>>(b, b.plane, args_plane, nb, lds_p))) return rc;
    if ((rc = batch_table<OpFixups<MIXED, C, LJ, CS>>(b, b.fixups, args_fixups, nfix, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = (int)nblk(b.sys[i]->n, 256);
    if ((rc = batch_table<OpUnsortForces<MIXED>>(b, b.unsort, args_unsort, nb, 0))) return rc;
    if ((rc = batch_table<OpHalfKick>(b, b.half, args_half, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = 1;
    if ((rc = batch_table<OpReduceEnergy>(b, b.energy, args_energy, nb, 0))) return rc;
    for (int i = 0; i < ns; ++i) nb[i] = (int)nblk(std::max<long long>(3ll * b.sys[i]->n, kMaxChunk), 256);
    if ((rc = batch_table<OpChunkBegin>(b, b.begin, args_begin, nb, 0))) return rc;
  }
  // per-system bookkeeping
  for (int i = 0; i < ns; ++i) {
    Sys *s = b.sys[i];
    b.bsys_h[i] = BatchSys{P<unsigned>(s->md.dmax), s->md.skin > 0 ? (float)(0.25 * s->md.skin * s->md.skin) : -1.f, 0,
                           P<unsigned long long>(s->md.step_dev), P<int>(s->overflow),
                           reinterpret_cast<long long *>(b.rebuilds.p) + i};
  }
  return PMX_OK;
}

// Queues one chunk of K steps of every system of the batch (the first step
// of each system's run when `opening`), then the gathering of the overflow
// flags into b.flags_h.  Step 0 kicks and drifts and spreads on its own;
// every later step's kick, drift and spread run in the previous step's last
// launch.
template <bool MIXED, int C, int LJ, int CS, bool BIGLJ, int ORDER>
int batch_chunk(Batch &b, int K, bool opening) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using T = R;
  hipStream_t st = b.ctx->stream, st2 = b.ctx->stream2;
  const int ns = b.nsys;
  BHIPCHK(hipMemcpyAsync(b.bsys.p, b.bsys_h, sizeof(BatchSys) * ns, hipMemcpyHostToDevice, st));
  BHIPCHK(hipMemcpyAsync(b.base.p, b.base_h, 8 * (size_t)ns, hipMemcpyHostToDevice, st));
  const BatchSys *bs = reinterpret_cast<const BatchSys *>(b.bsys.p);
  unsigned char *need = reinterpret_cast<unsigned char *>(b.need.p);
  hipLaunchKernelGGL(k_batch_steps, dim3(nblk(ns, 256)), dim3(256), 0, st, ns, bs,
                     reinterpret_cast<const unsigned long long *>(b.base.p));
  const BatchStep b0{0, opening ? 1 : 0, need};
  batch_go<OpChunkBegin>(b.begin, st, b0);
  for (int k = 0; k < K; ++k) {
    const BatchStep bk{k, opening ? 1 : 0, need};
    if (k == 0) batch_go<OpKickDrift<R>>(b.kick, st, bk);
    hipLaunchKernelGGL(k_batch_need, dim3(nblk(ns, 256)), dim3(256), 0, st, ns, bs, k, need);
    batch_go<OpBin>(b.bin, st, bk);
    batch_go<OpScanSingle>(b.scan, st, bk);
    batch_go<OpScatter>(b.scatter, st, bk);
    batch_go<OpBinrank>(b.binrank, st, bk);
    batch_go<OpColumns>(b.columns, st, bk);
    batch_go<OpSplitFill<R>>(b.split, st, bk);
    batch_go<OpBuildList>(b.list, st, bk);
    batch_go<OpMarkExclusions>(b.excl, st, bk);
    batch_go<OpBuildMasks>(b.masks, st, bk);
    batch_go<OpRunsPrefix>(b.rprefix, st, bk);
    batch_go<OpRunsCut>(b.rcut, st, bk);
    batch_go<OpRunsTiles>(b.rtiles, st, bk);
    if (k == 0) batch_go<OpSpreadPatch<MIXED, ORDER>>(b.spread, st, bk);
    // fork: the PME chain beside the pair kernel
    BHIPCHK(hipEventRecord(b.fork, st));
    BHIPCHK(hipStreamWaitEvent(st2, b.fork, 0));
    batch_go<PairOp<MIXED, C, LJ, CS, BIGLJ, ORDER>>(b.pair, st, bk);
    batch_go<OpFftZr2c<T, MIXED>>(b.zr2c, st2, bk);
    batch_go<OpFftPlane<T>>(b.plane, st2, bk);
    batch_go<OpFftY<T>>(b.yfwd, st2, bk);
    batch_go<OpFftXConv<T>>(b.xconv, st2, bk);
    batch_go<OpFftY<T>>(b.yinv, st2, bk);
    batch_go<OpFftZc2r<T>>(b.zc2r, st2, bk);
    batch_go<OpGatherPatch<MIXED, ORDER>>(b.gather, st2, bk);
    BHIPCHK(hipEventRecord(b.join, st2));
    // boundary pairs, exclusion correction and bonded terms after the pair
    // kernel, then the join
    batch_go<OpFixups<MIXED, C, LJ, CS>>(b.fixups, st, bk);
    BHIPCHK(hipStreamWaitEvent(st, b.join, 0));
    if (k + 1 < K)
      batch_go<OpUnsortKickDriftSpread<MIXED, R, ORDER>>(b.ukds, st, bk);
    else
      batch_go<OpUnsortForces<MIXED>>(b.unsort, st, bk);
  }
  hipLaunchKernelGGL(k_batch_flags, dim3(nblk(ns, 256)), dim3(256), 0, st, ns, bs,
                     reinterpret_cast<int *>(b.flags.p));
  BHIPCHK(hipMemcpyAsync(b.flags_h, b.flags.p, 4 * (size_t)ns, hipMemcpyDeviceToHost, st));
  BHIPCHK(hipGetLastError());
  return PMX_OK;
}

// One force pass of every system of the batch at its set positions (list
// rebuilt for all), with the energies reduced: the work of compute() for
// each.  Systems whose list outgrew its buffers are grown and computed again
// alone.
template <bool MIXED, int C, int LJ, int CS, bool BIGLJ, int ORDER>
int batch_forces(Batch &b) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using T = R;
  hipStream_t st = b.ctx->stream, st2 = b.ctx->stream2;
  const int ns = b.nsys;
  int rc = batch_tables<MIXED, C, LJ, CS, BIGLJ, ORDER>(b);
  if (rc) return rc;
  for (Sys *s : b.sys) {
    if (!s->acc_clean || s->spread_ready) launch_zero(s);
    s->spread_ready = false;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, st, P<int>(s->overflow), 0, 2ll);
  }
  std::vector<unsigned char> all(ns, 1);
  BHIPCHK(hipMemcpyAsync(b.need.p, all.data(), ns, hipMemcpyHostToDevice, st));
  BHIPCHK(hipMemcpyAsync(b.bsys.p, b.bsys_h, sizeof(BatchSys) * ns, hipMemcpyHostToDevice, st));
  const BatchStep bk{0, 0, reinterpret_cast<unsigned char *>(b.need.p)};
  batch_go<OpBin>(b.bin, st, bk);
  batch_go<OpScanSingle>(b.scan, st, bk);
  batch_go<OpScatter>(b.scatter, st, bk);
  batch_go<OpBinrank>(b.binrank, st, bk);
  batch_go<OpColumns>(b.columns, st, bk);
  batch_go<OpSplitFill<R>>(b.split, st, bk);
  batch_go<OpBuildList>(b.list, st, bk);
  batch_go<OpMarkExclusions>(b.excl, st, bk);
  batch_go<OpBuildMasks>(b.masks, st, bk);
  batch_go<OpRunsPrefix>(b.rprefix, st, bk);
  batch_go<OpRunsCut>(b.rcut, st, bk);
  batch_go<OpRunsTiles>(b.rtiles, st, bk);
  batch_go<OpSpreadPatch<MIXED, ORDER>>(b.spread, st, bk);
  BHIPCHK(hipEventRecord(b.fork, st));
  BHIPCHK(hipStreamWaitEvent(st2, b.fork, 0));
  batch_go<PairOp<MIXED, C, LJ, CS, BIGLJ, ORDER>>(b.pair, st, bk);
  batch_go<OpFftZr2c<T, MIXED>>(b.zr2c, st2, bk);
  batch_go<OpFftPlane<T>>(b.plane, st2, bk);
  batch_go<OpFftY<T>>(b.yfwd, st2, bk);
  batch_go<OpFftXConv<T>>(b.xconv, st2, bk);
  batch_go<OpFftY<T>>(b.yinv, st2, bk);
  batch_go<OpFftZc2r<T>>(b.zc2r, st2, bk);
  batch_go<OpGatherPatch<MIXED, ORDER>>(b.gather, st2, bk);
  BHIPCHK(hipEventRecord(b.join, st2));
  batch_go<OpFixups<MIXED, C, LJ, CS>>(b.fixups, st, bk);
  BHIPCHK(hipStreamWaitEvent(st, b.join, 0));
  batch_go<OpUnsortForces<MIXED>>(b.unsort, st, bk);
  batch_go<OpReduceEnergy>(b.energy, st, bk);
  hipLaunchKernelGGL(k_batch_flags, dim3(nblk(ns, 256)), dim3(256), 0, st, ns,
                     reinterpret_cast<const BatchSys *>(b.bsys.p), reinterpret_cast<int *>(b.flags.p));
  BHIPCHK(hipMemcpyAsync(b.flags_h, b.flags.p, 4 * (size_t)ns, hipMemcpyDeviceToHost, st));
  BHIPCHK(hipGetLastError());
  BHIPCHK(hipStreamSynchronize(st));
  for (int i = 0; i < ns; ++i) {
    Sys *s = b.sys[i];
    s->list_valid = true;
    s->prune_pending = false;
    s->acc_clean = true;
    for (int attempt = 0; b.flags_h[i] && attempt < 6; ++attempt) {
      if ((rc = grow_and_retry(s)) || (rc = compute(s, 0))) return rc;
      int ovf = 0;
      if (hipMemcpyAsync(&ovf, s->overflow.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return fail(b.ctx, PMX_ERR_HIP, "overflow flag");
      b.flags_h[i] = ovf;
    }
  }
  return PMX_OK;
}

constexpr int kBatchChunk = 64;  // steps between the host's checks of a batch

// nsteps MD steps of every system of the batch.  Chunks of up to
// kBatchChunk steps; after each the host reads the overflow flags, and a
// chunk in which a list outgrew its buffers is run again from its start
// (backups of every system) after those systems' buffers grow.
template <bool MIXED, int C, int LJ, int CS, bool BIGLJ, int ORDER>
int batch_run(Batch &b, int nsteps) {
  hipStream_t st = b.ctx->stream;
  int rc = batch_tables<MIXED, C, LJ, CS, BIGLJ, ORDER>(b);
  if (rc) return rc;
  BHIPCHK(hipMemsetAsync(b.rebuilds.p, 0, 8 * (size_t)b.nsys, st));
  int done = 0, retries = 0, last = 0;
  while (done < nsteps) {
    const int K = std::min(kBatchChunk, nsteps - done);
    last = K;
    for (int i = 0; i < b.nsys; ++i) b.base_h[i] = (unsigned long long)(b.sys[i]->md.steps + done);
    if ((rc = batch_chunk<MIXED, C, LJ, CS, BIGLJ, ORDER>(b, K, done == 0))) return rc;
    BHIPCHK(hipStreamSynchronize(st));
    bool grown = false;
    for (int i = 0; i < b.nsys; ++i) {
      if (!b.flags_h[i]) continue;
      if (!grown && ++retries > 6) return fail(b.ctx, PMX_ERR_OVERFLOW, "pair list capacity exceeded during MD");
      if (!grown) {
        // every system back to the chunk's start
        for (Sys *s : b.sys)
          if ((rc = restore_chunk(s, 0))) return rc;
      }
      grown = true;
      if ((rc = grow_and_retry(b.sys[i]))) return rc;
    }
    if (grown) {
      if ((rc = batch_tables<MIXED, C, LJ, CS, BIGLJ, ORDER>(b))) return rc;
      for (int i = 0; i < b.nsys; ++i) b.bsys_h[i].force = 1;
      BHIPCHK(hipStreamSynchronize(st));
      continue;
    }
    for (int i = 0; i < b.nsys; ++i) b.bsys_h[i].force = 0;
    done += K;
  }
  // closing half kick of every system's last step
  const BatchStep bk{last, 0, nullptr};
  batch_go<OpHalfKick>(b.half, st, bk);
  std::vector<long long> nreb(b.nsys);
  BHIPCHK(hipMemcpyAsync(nreb.data(), b.rebuilds.p, 8 * (size_t)b.nsys, hipMemcpyDeviceToHost, st));
  BHIPCHK(hipStreamSynchronize(st));
  for (int i = 0; i < b.nsys; ++i) {
    Sys *s = b.sys[i];
    s->md.steps += nsteps;
    s->md.rebuilds += nreb[i];
    s->list_valid = true;
    s->prune_pending = false;
    s->spread_ready = false;
    s->acc_clean = true;
  }
  return PMX_OK;
}

}  // extern "C++"


}  // namespace

int pmx_md_run(pmx_sys *sys, int nsteps) {
  if (!sys || nsteps < 0) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  MdRun r;
  if (int rc = md_begin(s, r, nsteps)) return rc;
  if (r.finished) return PMX_OK;
  bool progressed;
  while (!r.finished)
    if (int rc = md_advance(s, r, true, &progressed)) return rc;
  return md_end(s);
}

int pmx_md_run_many(pmx_sys *const *systems, int nsys, int nsteps) {
  if (!systems || nsys < 0 || nsteps < 0) return PMX_ERR_ARG;
  for (int i = 0; i < nsys; ++i) {
    if (!systems[i]) return PMX_ERR_ARG;
    for (int j = 0; j < i; ++j)
      if (systems[j] == systems[i]) return fail(systems[i]->s.ctx, PMX_ERR_ARG, "a system is listed twice");
  }
  // systems grouped by context; one host thread drives each group
  std::vector<pmx_ctx *> ctxs;
  std::vector<std::vector<Sys *>> groups;
  for (int i = 0; i < nsys; ++i) {
    Sys *s = &systems[i]->s;
    size_t g = std::find(ctxs.begin(), ctxs.end(), s->ctx) - ctxs.begin();
    if (g == ctxs.size()) {
      ctxs.push_back(s->ctx);
      groups.emplace_back();
    }
    groups[g].push_back(s);
  }
  std::vector<int> status(groups.size(), PMX_OK);
  auto drive = [&](size_t g) {
    std::vector<Sys *> &sys = groups[g];
    std::vector<MdRun> runs(sys.size());
    for (size_t i = 0; i < sys.size(); ++i)
      if (int rc = md_begin(sys[i], runs[i], nsteps)) {
        status[g] = rc;
        return;
      }
    for (;;) {
      bool any = false, progressed_any = false;
      for (size_t i = 0; i < sys.size(); ++i) {
        if (runs[i].finished) continue;
        any = true;
        bool p = false;
        if (int rc = md_advance(sys[i], runs[i], false, &p)) {
          status[g] = rc;
          return;
        }
        progressed_any |= p;
      }
      if (!any) break;
      if (!progressed_any) {
        // nothing completed: wait for the oldest chunk of the first
        // unfinished system
        for (size_t i = 0; i < sys.size(); ++i)
          if (!runs[i].finished) {
            bool p = false;
            if (int rc = md_advance(sys[i], runs[i], true, &p)) {
              status[g] = rc;
              return;
            }
            break;
          }
      }
    }
    for (Sys *s : sys)
      if (int rc = md_end(s)) {
        status[g] = rc;
        return;
      }
  };
  if (groups.size() == 1) {
    drive(0);
  } else {
    std::vector<std::thread> threads;
    for (size_t g = 0; g < groups.size(); ++g) threads.emplace_back(drive, g);
    for (auto &t : threads) t.join();
  }
  for (int rc : status)
    if (rc) return rc;
  return PMX_OK;
}

extern "C++" {
namespace {

// nsteps < 0: one force pass (batch_forces), else nsteps MD steps.
template <bool MIXED, int C, int LJ, int CS, bool BIG, int ORDER>
int batch_exec(Batch &b, int nsteps) {
  return nsteps < 0 ? batch_forces<MIXED, C, LJ, CS, BIG, ORDER>(b) : batch_run<MIXED, C, LJ, CS, BIG, ORDER>(b, nsteps);
}

template <bool MIXED, int CASE>
int batch_run_case(Batch &b, int order, int nsteps) {
  constexpr int LJ = CASE == 0 ? SW_C2 : (CASE == 1 ? SW_NONE : SW_ANY);
  constexpr int CS = CASE == 0 ? SW_CUTOFF : SW_ANY;
  constexpr bool BIG = CASE == 3;
  switch (order) {
    case 4: return batch_exec<MIXED, COUL_PME, LJ, CS, BIG, 4>(b, nsteps);
    case 6: return batch_exec<MIXED, COUL_PME, LJ, CS, BIG, 6>(b, nsteps);
    default: return batch_exec<MIXED, COUL_PME, LJ, CS, BIG, 8>(b, nsteps);
  }
}

template <bool MIXED>
int batch_run_form(Batch &b, int cs, int order, int nsteps) {
  switch (cs) {
    case 0: return batch_run_case<MIXED, 0>(b, order, nsteps);
    case 1: return batch_run_case<MIXED, 1>(b, order, nsteps);
    case 2: return batch_run_case<MIXED, 2>(b, order, nsteps);
    default: return batch_run_case<MIXED, 3>(b, order, nsteps);
  }
}

// Checks a batch's systems, makes the batch's buffers and runs it
// (nsteps < 0: one force pass).
int batch_entry(pmx_sys *const *systems, int nsys, int nsteps) {
  Sys *s0 = &systems[0]->s;
  pmx_ctx *ctx = s0->ctx;
  const bool md = nsteps >= 0;
  const int cs = batch_case(s0), order = s0->pme.order;
  std::set<const pmx_sys *> seen;
  for (int i = 0; i < nsys; ++i) {
    Sys *s = &systems[i]->s;
    if (!seen.insert(systems[i]).second) return fail(ctx, PMX_ERR_ARG, "a system is listed twice");
    if (s->ctx != ctx) return fail(ctx, PMX_ERR_ARG, "the systems of a batch share one context");
    if (md && !s->md.on) return fail(ctx, PMX_ERR_ARG, "pmx_md_init first (every system of the batch)");
    if (!s->have_positions) return fail(ctx, PMX_ERR_ARG, "positions not set (every system of the batch)");
    if (s->mixed != s0->mixed || batch_case(s) != cs)
      return fail(ctx, PMX_ERR_ARG, "the systems of a batch share the precision and the kinds of force");
    if (!s->pme_on || !(s->d.pme_terms & PMX_PME_RECIP) || s->pme.order < 4 || s->pme.order != order)
      return fail(ctx, PMX_ERR_UNSUPPORTED, "batched launches take PME systems of one stencil order (4, 6 or 8)");
    if (s->nbins > kScanSingleMax)
      return fail(ctx, PMX_ERR_UNSUPPORTED, "batched launches take systems of at most 65,536 cell bins");
  }
  if (nsteps == 0) return PMX_OK;
  hipSetDevice(ctx->device);
  Batch b;
  b.ctx = ctx;
  b.nsys = nsys;
  for (int i = 0; i < nsys; ++i) {
    Sys *s = &systems[i]->s;
    b.sys.push_back(s);
    if (!s->cached)
      if (int rc = setup(s)) return rc;
    if (md) {
      if (int rc = md_ensure_forces(s)) return rc;
      if (!s->acc_clean || s->spread_ready) {
        s->spread_ready = false;
        launch_zero(s);
        s->acc_clean = true;
      }
    }
  }
  const int dev = ctx->device;
  if (!b.bsys.get(dev, sizeof(BatchSys) * nsys) || !b.need.get(dev, nsys) || !b.base.get(dev, 8 * (size_t)nsys) ||
      !b.flags.get(dev, 4 * (size_t)nsys) || !b.rebuilds.get(dev, 8 * (size_t)nsys))
    return fail(ctx, PMX_ERR_OOM, "batch buffers");
  {
    RtLock lk(g_rt_mu);
    const size_t need = (sizeof(BatchSys) + 16) * (size_t)nsys;
    if (ctx->batch_host_bytes < need) {
      if (ctx->batch_host) hipHostFree(ctx->batch_host);
      ctx->batch_host = nullptr;
      ctx->batch_host_bytes = 0;
      if (hipHostMalloc(&ctx->batch_host, need, hipHostMallocDefault) != hipSuccess)
        return fail(ctx, PMX_ERR_HIP, "batch host buffers");
      ctx->batch_host_bytes = need;
    }
    if ((!ctx->batch_fork &&
         hipEventCreateWithFlags(&ctx->batch_fork, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess) ||
        (!ctx->batch_join &&
         hipEventCreateWithFlags(&ctx->batch_join, hipEventDisableTiming | hipEventDisableSystemFence) != hipSuccess))
      return fail(ctx, PMX_ERR_HIP, "batch events");
  }
  b.bsys_h = static_cast<BatchSys *>(ctx->batch_host);
  b.base_h = reinterpret_cast<unsigned long long *>(b.bsys_h + nsys);
  b.flags_h = reinterpret_cast<int *>(b.base_h + nsys);
  b.fork = ctx->batch_fork;
  b.join = ctx->batch_join;
  // the batch runs on the context's two streams, idle from here on
  if (hipStreamSynchronize(ctx->stream) != hipSuccess || hipStreamSynchronize(ctx->stream2) != hipSuccess)
    return fail(ctx, PMX_ERR_HIP, "stream synchronize");
  return s0->mixed ? batch_run_form<true>(b, cs, order, nsteps) : batch_run_form<false>(b, cs, order, nsteps);
}

}  // namespace
}  // extern "C++"

int pmx_md_run_batch(pmx_sys *const *systems, int nsys, int nsteps) {
  if (!systems || nsys <= 0 || nsteps < 0) return PMX_ERR_ARG;
  for (int i = 0; i < nsys; ++i)
    if (!systems[i]) return PMX_ERR_ARG;
  return batch_entry(systems, nsys, nsteps);
}

int pmx_sys_compute_batch(pmx_sys *const *systems, int nsys, unsigned flags) {
  if (!systems || nsys <= 0) return PMX_ERR_ARG;
  for (int i = 0; i < nsys; ++i)
    if (!systems[i]) return PMX_ERR_ARG;
  if (flags) return fail(systems[0]->s.ctx, PMX_ERR_UNSUPPORTED, "batched force passes take no flags");
  return batch_entry(systems, nsys, -1);
}

int pmx_md_get_state(pmx_sys *sys, double *x, double *v, double *energies) {
  if (!sys) return PMX_ERR_ARG;
  Sys *s = &sys->s;
  Sys::Md &m = s->md;
  if (!m.on || !s->have_positions) return fail(s->ctx, PMX_ERR_ARG, "pmx_md_init and pmx_md_set_state first");
  hipSetDevice(s->ctx->device);
  hipStream_t st = s->ctx->stream;
  const int n = s->n;
  if (x) HIPCHK(hipMemcpyAsync(x, s->xorig.p, 24 * (size_t)n, hipMemcpyDeviceToHost, st));
  if (v) HIPCHK(hipMemcpyAsync(v, m.v.p, 24 * (size_t)n, hipMemcpyDeviceToHost, st));
  if (energies) {
    // potential energies at the current positions (computed first if the
    // state was just set) and kinetic energy of v
    if (int rc = md_ensure_forces(s)) return rc;
    reduce_energies(s, false);
    hipLaunchKernelGGL(k_kinetic, dim3(nblk(n, 256)), dim3(256), 0, st, n, P<double>(m.v), P<double>(m.minv),
                       P<double>(m.ke_e));
    const int nbb = std::max(1, (int)nblk(m.nb + m.na, 256));
    std::vector<double> be(2 * nbb, 0.0), ke(nblk(n, 256));
    if (m.nb + m.na > 0) HIPCHK(hipMemcpyAsync(be.data(), m.bond_e.p, 16 * (size_t)nbb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ke.data(), m.ke_e.p, 8 * ke.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(s->eout_h, s->eout.p, 32, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double eb = 0, ea = 0, ek = 0;
    for (int i = 0; i < nbb; ++i) {
      eb += be[2 * i];
      ea += be[2 * i + 1];
    }
    for (double e : ke) ek += e;
    double ec = s->eout_h[0];
    if (s->pme_on) ec += s->eout_h[2] + s->eout_h[3] + s->self_e + s->charged_e;
    energies[PMX_MD_E_COULOMB] = ec;
    energies[PMX_MD_E_LJ] = s->eout_h[1];
    energies[PMX_MD_E_BOND] = eb;
    energies[PMX_MD_E_ANGLE] = ea;
    energies[PMX_MD_E_KINETIC] = ek;
  }
  HIPCHK(hipStreamSynchronize(st));
  return PMX_OK;
}

unsigned long long pmx_estimate_bytes(const pmx_sys_desc *desc, int kind) {
  if (!desc || validate(nullptr, desc) != PMX_OK) return 0;
  pmx_sys tmp;
  Sys *s = &tmp.s;
  s->d = *desc;
  s->dry = true;
  if (setup(s) != PMX_OK) return 0;
  unsigned long long b = s->dry_bytes;
  // device MD state: velocities, masses, build positions, two chunk backups
  // of x, v and f, bonded tables (bonds and angles about one per atom in the
  // systems served)
  if (kind == 1) b += 200ull * (unsigned long long)desc->n + 48ull * (unsigned long long)desc->n + 8192;
  return b;
}

int pmx_sys_device_bytes(const pmx_sys *sys, unsigned long long *bytes) {
  if (!sys || !bytes) return PMX_ERR_ARG;
  unsigned long long b = 0;
  for (const DevBuf *d : sys->s.all)
    if (d->p) b += d->bytes;
  *bytes = b;
  return PMX_OK;
}

int pmx_langevin_noise(unsigned long long seed, unsigned long long half, int n, double *g) {
  if (n < 0 || (n > 0 && !g)) return PMX_ERR_ARG;
  for (int i = 0; i < n; ++i) langevin_gauss3(seed, half, (uint32_t)i, g + 3 * i);
  return PMX_OK;
}

int pmx_md_stats(const pmx_sys *sys, long long *steps, long long *rebuilds, long long *rollbacks) {
  if (!sys || !steps || !rebuilds || !rollbacks) return PMX_ERR_ARG;
  *steps = sys->s.md.steps;
  *rebuilds = sys->s.md.rebuilds;
  *rollbacks = sys->s.md.rollbacks;
  return PMX_OK;
}

}  // extern "C"
