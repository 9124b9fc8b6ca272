// Batched launches: one launch runs a kernel over the blocks of many systems.
//
// A batch is a set of systems of one context that share the kernel case (form,
// force kinds, PME order).  Each kernel of the MD step (and of a force pass)
// gets, per batch, a
// table with one argument block per system in device memory, the first block
// of every system (first[s], first[nsys] = all blocks) and the system of every
// block (b2s).  Block b then does the work of block b - first[s] of system
// b2s[b] in a launch of first[s + 1] - first[s] blocks, through the same
// kernel body as the single-system launch, so a system's results do not depend
// on the batch it runs in.
//
// The list rebuild runs on the device: per step a small kernel marks the
// systems whose atoms moved more than half the skin since their last build
// (the displacement record of the step's kick-drift), and the blocks of the
// rebuild kernels of the other systems return at once.
#pragma once

#include "pmx_kernels.h"

namespace pmx {

// Per-launch values shared by the systems of a batched step.
struct BatchStep {
  int k;                          // step within the chunk
  int opening;                    // 1: step 0 of the chunk opens the run (half kick)
  const unsigned char *need;      // per system: the list is rebuilt at this step
};

template <class Op>
__global__ void __launch_bounds__(Op::kBlock, Op::kMinBlocks) k_batch(const typename Op::Args *__restrict__ tab,
                                                                      const int *__restrict__ b2s,
                                                                      const int *__restrict__ first, BatchStep st) {
  // the blocks of one XCD take a contiguous range of the table (whole
  // systems share its L2)
  const int b = xcd_block(blockIdx.x, gridDim.x);
  const int s = __builtin_amdgcn_readfirstlane(b2s[b]);
  if (Op::kGated && !st.need[s]) return;
  const int f = first[s];
  Op::run(tab[s], b - f, first[s + 1] - f, st);
}

// The kick of step k of a chunk for a system whose kick parameters are `base`
// (half0 relative to the chunk's first step in *base.step_dev): step 0 of an
// opening chunk gets one half kick, every other step the previous step's
// closing and its own opening half kick.
__device__ inline Kick batch_kick(const Kick &base, double h, int k, bool opening) {
  Kick kk = base;
  const bool half = opening && k == 0;
  kk.h_kick = (half ? 0.5 : 1.0) * h;
  kk.nhalf = half ? 1 : 2;
  kk.half0 = half ? 2ull * (unsigned long long)k : 2ull * (unsigned long long)k - 1ull;
  return kk;
}

// ---- list rebuild (gated) ----

struct OpBin {
  struct Args {
    int n;
    const double *xorig;
    BinParams p;
    uint32_t *key, *hist;
    double *xwrap;
    int *refcell;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    bin_body(bid, nbk, a.n, a.xorig, a.p, a.key, a.xwrap, a.refcell, a.hist);
  }
};

struct OpScanSingle {
  struct Args {
    uint32_t *hist, *start, *cursor;
    long long nbins;
  };
  static constexpr int kBlock = kScanBlock, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    scan_single_body(bid, nbk, a.hist, a.nbins, a.start, a.cursor);
  }
};

struct OpScatter {
  struct Args {
    int n;
    const uint32_t *key;
    uint32_t *cursor;
    int *perm;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    scatter_body(bid, nbk, a.n, a.key, a.cursor, a.perm);
  }
};

struct OpBinrank {
  struct Args {
    int n;
    long long nbins;
    const uint32_t *key, *start;
    const int *in;
    int *out;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    binrank_body(bid, nbk, a.n, a.nbins, a.key, a.start, a.in, a.out);
  }
};

struct OpColumns {
  struct Args {
    int ncols, nzb, n, nsuper;
    const uint32_t *start;
    int2 *sc_info;
    int *sc_col, *colsc;
  };
  static constexpr int kBlock = 1024, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    columns_body(bid, nbk, a.ncols, a.nzb, a.n, a.start, a.nsuper, a.sc_info, a.sc_col, a.colsc);
  }
};

template <class R>
struct OpSplitFill {
  struct Args {
    int nsuper;
    const int2 *sc_info;
    const int *order;
    const double *xwrap, *q;
    int *permp, *slot;
    const int *type;
    typename Real4<R>::type *xs;
    double4 *xs64;
    int *ts;
    const double *xorig;
    double *xbuild;
    float4 *sorg, *ctr, *half, *sctr, *shalf;
    int *nmaskwork;
    double qpair;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    split_fill_body<R>(bid, nbk, a.nsuper, a.sc_info, a.order, a.xwrap, a.permp, a.slot, a.q, a.type, a.xs, a.xs64,
                       a.ts, a.xorig, a.xbuild, a.sorg, a.ctr, a.half, a.sctr, a.shalf, a.nmaskwork, a.qpair);
  }
};

struct OpBuildList {
  struct Args {
    ListParams q;
    const int *sc_col, *colsc;
    const float4 *sctr, *shalf, *ctr, *half;
    int4 *entries;
    int *count, *ecost, *cost_out, *maskwork, *nmaskwork;
    int maxmaskwork;
    int *overflow;
  };
  static constexpr int kBlock = kListBlock, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    build_list_body(bid, nbk, a.q, a.sc_col, a.colsc, a.sctr, a.shalf, a.ctr, a.half, a.entries, a.count, a.ecost,
                    a.cost_out, a.maskwork, a.nmaskwork, a.maxmaskwork, a.overflow);
  }
};

struct OpMarkExclusions {
  struct Args {
    int ne, cap, maxmaskwork;
    const int4 *etab;
    const int *slot, *count;
    int4 *entries;
    int *maskwork, *nmaskwork, *overflow;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    mark_exclusions_body(bid, nbk, a.ne, a.etab, a.slot, a.entries, a.count, a.cap, a.maskwork, a.nmaskwork,
                         a.maxmaskwork, a.overflow);
  }
};

struct OpBuildMasks {
  struct Args {
    const int *maskwork, *nmaskwork;
    int maxmaskwork, cap;
    int4 *entries;
    const int *perm, *eptr;
    const int2 *elist;
    unsigned short *masks;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    build_masks_body(bid, nbk, a.maskwork, a.nmaskwork, a.maxmaskwork, a.entries, a.cap, a.perm, a.eptr, a.elist,
                     a.masks);
  }
};

struct OpRunsPrefix {
  struct Args {
    int nsuper;
    const int *count, *cost;
    int *nruns, *ep;
    long long *cp;
    int *pos, *posS;
    long long min_run_cost;
  };
  static constexpr int kBlock = 1024, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    runs_prefix_body(bid, nbk, a.nsuper, a.count, a.cost, a.nruns, a.ep, a.cp, a.pos, a.posS, a.min_run_cost);
  }
};

struct OpRunsCut {
  struct Args {
    int nsuper, cap;
    const int *count, *ecost, *nruns, *ep;
    const long long *cp;
    int *pos, *posS;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    runs_cut_body(bid, nbk, a.nsuper, a.count, a.ecost, a.cap, a.nruns, a.ep, a.cp, a.pos, a.posS);
  }
};

struct OpRunsTiles {
  struct Args {
    const int *count, *ep, *pos, *posS;
    int *wstart;
    int4 *tiles;
    int *ntiles;
    int ntiles_cap;
  };
  static constexpr int kBlock = 1024, kMinBlocks = 1;
  static constexpr bool kGated = true;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    runs_tiles_body(bid, nbk, a.count, a.ep, a.pos, a.posS, a.wstart, a.tiles, a.ntiles, a.ntiles_cap);
  }
};

// ---- force pass and integration ----

template <class R>
struct OpKickDrift {
  struct Args {
    int n;
    Kick kick;
    double h;
    const int *perm;
    const double *fout, *minv, *xbuild;
    double *v, *x;
    typename Real4<R>::type *xs;
    const double4 *xs64;
    const float4 *sorg;
    unsigned *dmax;  // displacement record of the chunk, one word per step
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &st) {
    kick_drift_body<R>(bid, nbk, a.n, batch_kick(a.kick, a.h, st.k, st.opening), a.h, a.perm, a.fout, a.minv, a.v,
                       a.x, a.xbuild, a.xs, a.xs64, a.sorg, a.dmax + st.k);
  }
};

template <bool MIXED, int P>
struct OpSpreadPatch {
  struct Args {
    int n;
    PmeParams p;
    const double *xorig, *q;
    const int *perm;
    double *grid64;
    unsigned long long *gridfx;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    spread_patch_body<MIXED, P>(bid, nbk, a.n, a.p, a.xorig, a.perm, a.q, a.grid64, a.gridfx);
  }
};

template <bool MIXED, int COUL, int LJSW, int CSW, bool BIGLJ>
struct OpPair {
  using R4 = typename Real4<typename std::conditional<MIXED, float, double>::type>::type;
  struct Args {
    PairParams pp;
    const int4 *tiles;
    const int *ntiles, *wstart;
    int *tilectr;
    int4 *entries;
    int cap;
    const unsigned short *masks;
    const R4 *xs;
    const float4 *sorg;
    const int *ts, *perm;
    const double *ljtab;
    const R4 *ljt;
    typename AccT<MIXED>::type *facc;
    PairOut out;
  };
  static constexpr int kBlock = kPairBlock, kMinBlocks = 4;
  static constexpr bool kGated = false;
  // the runs in contiguous shares per wave (schedule 0)
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    pair_body<MIXED, COUL, LJSW, CSW, BIGLJ, kPairPlain>(bid, nbk, a.pp, 0, a.tiles, a.ntiles, a.wstart, a.tilectr,
                                                         a.entries, a.entries, a.cap, a.masks, a.xs, a.sorg, a.ts,
                                                         a.perm, a.ljtab, a.ljt, a.facc, a.out);
  }
};

template <class T, bool FIXED>
struct OpFftZr2c {
  struct Args {
    PmeParams p;
    FftPlan pl;
    int lines;
    double *g64;
    unsigned long long *gfx;
    const Cx<T> *tw;
    Cx<T> *sa;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    fft_z_r2c_body<T, FIXED>(bid, nbk, a.p, a.pl, a.lines, a.g64, a.gfx, a.tw, a.sa);
  }
};

template <class T>
struct OpFftY {
  struct Args {
    PmeParams p;
    FftPlan pl;
    int lines;
    bool inverse;
    const Cx<T> *tw, *src;
    Cx<T> *dst;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    fft_y_body<T>(bid, nbk, a.p, a.pl, a.lines, a.inverse, a.tw, a.src, a.dst);
  }
};

template <class T>
struct OpFftXConv {
  struct Args {
    PmeParams p;
    FftPlan pl;
    int lines;
    const Cx<T> *tw;
    const T *theta;
    Cx<T> *sb;
    double *epart;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    fft_x_conv_body<T>(bid, nbk, a.p, a.pl, a.lines, a.tw, a.theta, a.sb, a.epart);
  }
};

template <class T>
struct OpFftPlane {
  struct Args {
    PmeParams p;
    FftPlan ply, plx;
    const Cx<T> *twy, *twx;
    const T *theta;
    Cx<T> *sa;
    double *epart;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    fft_plane_conv_body<T>(bid, nbk, a.p, a.ply, a.plx, a.twy, a.twx, a.theta, a.sa, a.epart);
  }
};

template <class T>
struct OpFftZc2r {
  struct Args {
    PmeParams p;
    FftPlan pl;
    int lines;
    const Cx<T> *tw, *sa;
    T *greal;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    fft_z_c2r_body<T>(bid, nbk, a.p, a.pl, a.lines, a.tw, a.sa, a.greal);
  }
};

template <bool MIXED, int P>
struct OpGatherPatch {
  struct Args {
    int n;
    PmeParams p;
    const double *xorig, *q;
    const int *perm;
    const typename std::conditional<MIXED, float, double>::type *greal;
    typename AccT<MIXED>::type *facc;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    gather_patch_body<MIXED, P>(bid, nbk, a.n, a.p, a.xorig, a.perm, a.q, a.greal, a.facc);
  }
};

template <bool MIXED, int COUL, int LJSW, int CSW>
struct OpFixups {
  using Args = FixupArgs<MIXED>;
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    fixups_body<MIXED, COUL, LJSW, CSW>(bid, nbk, a);
  }
};

// unsort of step k's forces, kick and drift of step k + 1, its charge spread
template <bool MIXED, class R, int P>
struct OpUnsortKickDriftSpread {
  struct Args {
    int n;
    Kick kick;
    double h;
    const int *perm;
    typename AccT<MIXED>::type *facc;
    double *fout;
    int *tilectr, *nborder;
    const double *minv, *xbuild, *q;
    double *v, *x;
    typename Real4<R>::type *xs;
    const double4 *xs64;
    const float4 *sorg;
    unsigned *dmax;
    PmeParams pme;
    double *grid64;
    unsigned long long *gridfx;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &st) {
    unsort_kick_drift_spread_body<MIXED, R, P>(bid, nbk, a.n, batch_kick(a.kick, a.h, st.k + 1, false), a.h, a.perm,
                                               a.facc, a.fout, a.tilectr, a.nborder, a.minv, a.v, a.x, a.xbuild,
                                               a.xs, a.xs64, a.sorg, a.dmax + st.k + 1, a.pme, a.q, a.grid64,
                                               a.gridfx);
  }
};

template <bool MIXED>
struct OpUnsortForces {
  struct Args {
    int n;
    const int *slot;
    typename AccT<MIXED>::type *facc;
    double *fout;
    int *tilectr, *nborder;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    unsort_forces_body<MIXED>(bid, nbk, a.n, a.slot, a.facc, a.fout, a.tilectr, a.nborder);
  }
};

// a force pass's energy partials into the system's energy slots (one block)
struct OpReduceEnergy {
  struct Args {
    int ntiles_cap;
    const int *ntiles;
    const double *pair_e;
    int pair_fixed, nrecip, ncorr;
    const double *recip_e, *corr_e;
    const long long *npairs;
    const unsigned long long *hash, *border_e, *border_stat;
    double *out;
    long long *out_np;
    unsigned long long *out_hash;
  };
  static constexpr int kBlock = 1024, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    reduce_energy_body(bid, nbk, a.ntiles_cap, a.ntiles, a.pair_e, a.pair_fixed, a.nrecip, a.recip_e, a.ncorr,
                       a.corr_e, a.npairs, a.hash, 0, a.border_e, a.border_stat, a.out, a.out_np, a.out_hash);
  }
};

// backups for a restore, cleared overflow flags and displacement record
struct OpChunkBegin {
  struct Args {
    long long m3;
    const double *x, *v, *f;
    double *xbak, *vbak, *fbak;
    int *overflow;
    unsigned *dmax;
    int nrec;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &) {
    chunk_begin_body(bid, nbk, a.m3, a.x, a.v, a.f, a.xbak, a.vbak, a.fbak, a.overflow, a.dmax, a.nrec);
  }
};

// the closing half kick after the chunk's last step (st.k = its length)
struct OpHalfKick {
  struct Args {
    int n;
    Kick kick;
    double h;
    const double *fout, *minv;
    double *v;
  };
  static constexpr int kBlock = 256, kMinBlocks = 1;
  static constexpr bool kGated = false;
  __device__ static void run(const Args &a, int bid, int nbk, const BatchStep &st) {
    Kick kk = a.kick;
    kk.h_kick = 0.5 * a.h;
    kk.nhalf = 1;
    kk.half0 = 2ull * (unsigned long long)st.k - 1ull;
    half_kick_body(bid, nbk, a.n, kk, a.fout, a.minv, a.v);
  }
};

// ---- per-system bookkeeping of a batch ----

struct BatchSys {
  const unsigned *dmax;          // the chunk's displacement record
  float lim2;                    // (skin / 2)^2
  int force;                     // rebuild at step 0 of the chunk whatever the displacement
  unsigned long long *step_dev;  // first step of the chunk
  const int *overflow;
  long long *rebuilds;           // device count of the builds in the batch
};

// need[s]: system s rebuilds its list at step k (first thread of each
// system); the count of builds goes up.
__global__ void k_batch_need(int nsys, const BatchSys *__restrict__ bs, int k, unsigned char *__restrict__ need) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const BatchSys b = bs[s];
  const float d = __uint_as_float(b.dmax[k]);
  const bool r = (k == 0 && b.force) || d > b.lim2;
  need[s] = r ? 1 : 0;
  if (r) ++*b.rebuilds;
}

// The chunk's first step index of every system (base[s]) and the overflow
// flags of the last chunk gathered into one array for the host.
__global__ void k_batch_steps(int nsys, const BatchSys *__restrict__ bs, const unsigned long long *__restrict__ base) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) *bs[s].step_dev = base[s];
}

__global__ void k_batch_flags(int nsys, const BatchSys *__restrict__ bs, int *__restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsys) out[s] = *bs[s].overflow;
}

}  // namespace pmx
