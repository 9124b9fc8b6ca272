// HIP kernels for the ProtoMol nonbonded hot path on gfx950 (CDNA4).
//
// Data layout in HBM (per system):
//   xorig  double[3N]  positions as the caller hands them (AoS, original order)
//   permp  int[Npad]   sorted slot -> original atom (-1 for padding slots)
//   slot   int[N]      original atom -> sorted slot
//   xs     R4[Npad]    coordinates relative to the slot's supercluster origin
//                      + scaled charge, sorted (R = double for the FP64 form,
//                      float for the mixed form)
//   sorg   float4[Nsuper]  supercluster origins (multiples of 1/16 A)
//   ts     int[Npad]   LJ type per sorted slot (-1 padding)
//   facc   Acc[3Npad]  force accumulators in sorted order (double, or int64
//                      fixed point 2^-32 for the mixed form)
// Superclusters are 64 slots (one wavefront): up to 64 atoms of one x-y
// column of the box, consecutive in z, split by x, then y, then z into 8
// clusters of 8 slots.  Padding slots fill the last supercluster of a column.
#pragma once
#include <type_traits>

#include "pmx_common.h"

namespace pmx {

// ---------------------------------------------------------------------------
// helpers

template <class T>
struct Real4;
template <>
struct Real4<double> {
  using type = double4;
};
template <>
struct Real4<float> {
  using type = float4;
};

// Reference minimal image in FP64 with the reference's operation order and
// no FMA contraction (PeriodicBoundaryConditions.h:72-110).  Returns d^2.
__device__ inline double exact_d2(const double *__restrict__ x, int a, int b, const Box &bx) {
#pragma clang fp contract(off)
  double d[3] = {x[3 * b] - x[3 * a], x[3 * b + 1] - x[3 * a + 1], x[3 * b + 2] - x[3 * a + 2]};
  double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (d2 > bx.D) {
    for (int k = 0; k < 3; ++k) {
      if (d[k] < -bx.H2[k])
        d[k] += bx.H[k];
      else if (d[k] > bx.H2[k])
        d[k] -= bx.H[k];
    }
    d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (d2 > bx.D) {
      for (int k = 0; k < 3; ++k) d[k] -= bx.L[k] * rint(bx.Lr[k] * d[k]);
      d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    }
  }
  return d2;
}

// Reference minimalPosition (PeriodicBoundaryConditions.h:113-147), FP64, no contraction.
__device__ inline void exact_minpos(const double *p, const Box &bx, double *out) {
#pragma clang fp contract(off)
  double d[3] = {p[0] - bx.o[0], p[1] - bx.o[1], p[2] - bx.o[2]};
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= bx.D) {
    out[0] = d[0];
    out[1] = d[1];
    out[2] = d[2];
    return;
  }
  for (int k = 0; k < 3; ++k) {
    if (d[k] < -bx.H2[k])
      d[k] += bx.H[k];
    else if (d[k] > bx.H2[k])
      d[k] -= bx.H[k];
  }
  if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] <= bx.D) {
    out[0] = d[0];
    out[1] = d[1];
    out[2] = d[2];
    return;
  }
  for (int k = 0; k < 3; ++k) out[k] = d[k] - bx.L[k] * rint(bx.Lr[k] * d[k]);
}

template <class T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  return v;
}
__device__ inline int wave_max(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = max(v, __shfl_xor(v, m));
  return v;
}

// ---------------------------------------------------------------------------
// 1. Binning: wrapped coordinates, column x z-bin sort key, reference cell.
struct BinParams {
  Box box;
  int ncx, ncy, nzb;  // column grid in x-y, z bins per column (sort key)
  double cxinv, cyinv, zinv;  // ncx / Lx, ncy / Ly, nzb / Lz
  int rc_n[3];        // reference cell grid
  double rc_inv[3];   // 1 / reference cell edge
  double delta[3];    // origin - min
};

// Block number for work laid out in order: the hardware deals consecutive
// blocks of a launch to the 8 XCDs in turn, so this renumbering gives the
// blocks of one XCD one contiguous range of the work (block bx runs on XCD
// bx mod 8; XCD x's ranks 0, 1, ... map to the x-th of 8 consecutive ranges,
// the first nb mod 8 of them one block longer), and neighbouring work shares
// that XCD's L2.  A bijection on [0, nb).
__device__ inline int xcd_block(int bx, int nb) {
  const int x = bx & 7, q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + (bx >> 3);
}

__device__ __forceinline__ void bin_body(const int bid_, const int nbk_, int n, const double *__restrict__ xorig,
                                         BinParams p, uint32_t *__restrict__ key, double *__restrict__ xwrap,
                                         int *__restrict__ refcell, uint32_t *__restrict__ hist) {
  int a = bid_ * blockDim.x + threadIdx.x;
  if (a >= n) return;
  double w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double x = xorig[3 * a + k];
    double v = x - p.box.L[k] * floor(x * p.box.Linv[k]);
    if (v >= p.box.L[k]) v -= p.box.L[k];
    if (v < 0.0) v = 0.0;
    w[k] = v;
  }
  // key: x-y column, then z bin
  const int cx = min((int)(w[0] * p.cxinv), p.ncx - 1), cy = min((int)(w[1] * p.cyinv), p.ncy - 1);
  const int zb = min((int)(w[2] * p.zinv), p.nzb - 1);
  const uint32_t kb = (uint32_t)((cx * p.ncy + cy) * p.nzb + zb);
  key[a] = kb;
  atomicAdd(&hist[kb], 1u);  // the scan clears hist after reading it
  xwrap[3 * a] = w[0];
  xwrap[3 * a + 1] = w[1];
  xwrap[3 * a + 2] = w[2];
  // Reference cell: findCell(delta + minimalPosition(x)) (Topology.h:64-66,
  // CubicCellManager.cpp:40-44), an upper-face index wrapped to its image.
  double mp[3];
  exact_minpos(&xorig[3 * a], p.box, mp);
  int c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double t;
    {
#pragma clang fp contract(off)
      t = (p.delta[k] + mp[k]) * p.rc_inv[k];
    }
    int ci = (int)floor(t);
    ci %= p.rc_n[k];
    if (ci < 0) ci += p.rc_n[k];
    c[k] = ci;
  }
  refcell[a] = (c[0] * p.rc_n[1] + c[1]) * p.rc_n[2] + c[2];
}

__global__ void k_bin(int n, const double *__restrict__ xorig, BinParams p, uint32_t *__restrict__ key,
                      double *__restrict__ xwrap, int *__restrict__ refcell, uint32_t *__restrict__ hist) {
  bin_body(blockIdx.x, gridDim.x, n, xorig, p, key, xwrap, refcell, hist);
}

// Counting sort by key (one radix digit spanning the key range): histogram,
// exclusive scan, scatter, then order each bin by atom index so the result is
// deterministic and stable.
__global__ void k_fill_u32(uint32_t *p, uint32_t v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_i32(int *p, int v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, double v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Exclusive scan of m uint32 in 3 launches: per-block sums, scan of sums in
// one block, then per-block rescan with the block offset.
constexpr int kScanBlock = 1024;
__global__ void k_scan_blocksums(const uint32_t *__restrict__ in, long long m, uint32_t *__restrict__ sums) {
  __shared__ uint32_t s[kScanBlock / kWave];
  long long i = (long long)blockIdx.x * kScanBlock + threadIdx.x;
  uint32_t v = i < m ? in[i] : 0u;
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kScanBlock / kWave; ++w) t += s[w];
    sums[blockIdx.x] = t;
  }
}
__global__ void k_scan_sums(uint32_t *sums, int nb) {
  // single block: serial over chunks of 1024 with a block-wide scan
  __shared__ uint32_t s[kScanBlock];
  uint32_t carry = 0;
  for (int base = 0; base < nb; base += kScanBlock) {
    int i = base + threadIdx.x;
    uint32_t v = i < nb ? sums[i] : 0u;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
      __syncthreads();
      s[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nb) sums[i] = carry + s[threadIdx.x] - v;  // exclusive
    __syncthreads();
    carry += s[kScanBlock - 1];
    __syncthreads();
  }
}
// Exclusive scan result to out and to out2 (the scatter cursors).
__global__ void k_scan_final(uint32_t *__restrict__ in, long long m, const uint32_t *__restrict__ sums,
                             uint32_t *__restrict__ out, uint32_t *__restrict__ out2) {
  __shared__ uint32_t s[kScanBlock];
  long long i = (long long)blockIdx.x * kScanBlock + threadIdx.x;
  uint32_t v = i < m ? in[i] : 0u;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    uint32_t t = threadIdx.x >= off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < m) {
    const uint32_t r = sums[blockIdx.x] + s[threadIdx.x] - v;
    out[i] = r;
    out2[i] = r;
    in[i] = 0u;  // cleared for the next histogram
  }
}

// Exclusive scan of m <= kScanSingleMax counts in one block: each wave owns
// a contiguous segment and walks it in coalesced 64-count chunks (a wave
// prefix per chunk, a running offset across chunks); segment totals are
// combined across the block first.  The counts are cleared for the next
// histogram.
constexpr long long kScanSingleMax = 1 << 16;
__device__ __forceinline__ void scan_single_body(const int bid_, const int nbk_, uint32_t *__restrict__ in, long long m,
                                                 uint32_t *__restrict__ out, uint32_t *__restrict__ out2) {
  constexpr int NW = kScanBlock / kWave;
  __shared__ uint32_t s_tot[NW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long seg = ((m + NW - 1) / NW + kWave - 1) / kWave * kWave;
  const long long b = min(m, (long long)w * seg), e = min(m, b + seg);
  uint32_t t = 0;
  for (long long i = b + lane; i < e; i += kWave) t += in[i];
  t = wave_sum(t);
  if (lane == 0) s_tot[w] = t;
  __syncthreads();
  uint32_t base = 0;
  for (int k = 0; k < w; ++k) base += s_tot[k];
  for (long long c = b; c < e; c += kWave) {
    const long long i = c + lane;
    const uint32_t v = i < e ? in[i] : 0u;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t y = __shfl_up(x, off);
      if (lane >= off) x += y;
    }
    if (i < e) {
      out[i] = base + x - v;
      out2[i] = base + x - v;
      in[i] = 0u;
    }
    base += __shfl(x, kWave - 1);
  }
}

__global__ void k_scan_single(uint32_t *__restrict__ in, long long m, uint32_t *__restrict__ out,
                              uint32_t *__restrict__ out2) {
  scan_single_body(blockIdx.x, gridDim.x, in, m, out, out2);
}

__device__ __forceinline__ void scatter_body(const int bid_, const int nbk_, int n, const uint32_t *__restrict__ key,
                                             uint32_t *__restrict__ cursor, int *__restrict__ perm) {
  int a = bid_ * blockDim.x + threadIdx.x;
  if (a < n) perm[atomicAdd(&cursor[key[a]], 1u)] = a;
}

__global__ void k_scatter(int n, const uint32_t *__restrict__ key, uint32_t *__restrict__ cursor,
                          int *__restrict__ perm) {
  scatter_body(blockIdx.x, gridDim.x, n, key, cursor, perm);
}

// Orders each bin by atom index: position p of the scatter output moves to
// its bin's start plus the number of smaller atom indices in the bin (bins
// hold a handful of atoms; the result does not depend on the scatter order).
__device__ __forceinline__ void binrank_body(const int bid_, const int nbk_, int n, long long nbins,
                                             const uint32_t *__restrict__ key, const uint32_t *__restrict__ start,
                                             const int *__restrict__ in, int *__restrict__ out) {
  const int p = bid_ * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int a = in[p];
  const uint32_t b = key[a];
  const int s0 = (int)start[b], s1 = b + 1 < (uint64_t)nbins ? (int)start[b + 1] : n;
  int rank = 0;
  for (int q = s0; q < s1; ++q) rank += in[q] < a;
  out[s0 + rank] = a;
}

__global__ void k_binrank(int n, long long nbins, const uint32_t *__restrict__ key,
                          const uint32_t *__restrict__ start, const int *__restrict__ in, int *__restrict__ out) {
  binrank_body(blockIdx.x, gridDim.x, n, nbins, key, start, in, out);
}

// Superclusters of each column: sc_info[sc] = (first sorted position, atom
// count) and sc_col[sc] = its column for sc < total, (0, 0) and -1 beyond;
// colsc[col] = first supercluster of column col, colsc[ncols] = total.  One
// block; columns in chunks of the block size with a running offset.
__device__ __forceinline__ void columns_body(const int bid_, const int nbk_, int ncols, int nzb, int n,
                                             const uint32_t *__restrict__ start, int nsuper_max,
                                             int2 *__restrict__ sc_info, int *__restrict__ sc_col,
                                             int *__restrict__ colsc) {
  __shared__ int s[1024];
  int carry = 0;
  for (int base = 0; base < ncols; base += blockDim.x) {
    const int col = base + threadIdx.x;
    int cs = 0, cnt = 0;
    if (col < ncols) {
      cs = (int)start[(size_t)col * nzb];
      const int ce = col + 1 < ncols ? (int)start[(size_t)(col + 1) * nzb] : n;
      cnt = ce - cs;
    }
    const int nsc = (cnt + kSuperSize - 1) / kSuperSize;
    s[threadIdx.x] = nsc;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
      int v = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
      __syncthreads();
      s[threadIdx.x] += v;
      __syncthreads();
    }
    const int first = carry + s[threadIdx.x] - nsc;
    if (col < ncols) colsc[col] = first;
    for (int q = 0; q < nsc && first + q < nsuper_max; ++q) {
      sc_info[first + q] = make_int2(cs + q * kSuperSize, min(kSuperSize, cnt - q * kSuperSize));
      sc_col[first + q] = col;
    }
    __syncthreads();
    carry += s[blockDim.x - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) colsc[ncols] = carry;
  for (int sc = carry + threadIdx.x; sc < nsuper_max; sc += blockDim.x) {
    sc_info[sc] = make_int2(0, 0);
    sc_col[sc] = -1;
  }
}

__global__ void k_columns(int ncols, int nzb, int n, const uint32_t *__restrict__ start, int nsuper_max,
                          int2 *__restrict__ sc_info, int *__restrict__ sc_col, int *__restrict__ colsc) {
  columns_body(blockIdx.x, gridDim.x, ncols, nzb, n, start, nsuper_max, sc_info, sc_col, colsc);
}

// Ascending bitonic sort of (key, idx) within aligned segments of SEG lanes.
template <int SEG>
__device__ inline void wave_sort(double &key, int &idx) {
  const int lane = threadIdx.x & 63, loc = lane & (SEG - 1);
#pragma unroll
  for (int kk = 2; kk <= SEG; kk <<= 1) {
#pragma unroll
    for (int j = kk >> 1; j > 0; j >>= 1) {
      const double ok = __shfl_xor(key, j);
      const int oi = __shfl_xor(idx, j);
      const bool up = (loc & kk) == 0;
      const bool lower = (lane & j) == 0;
      const bool less = ok < key || (ok == key && oi < idx);
      if ((lower == up) ? less : !less) {
        key = ok;
        idx = oi;
      }
    }
  }
}


// Supercluster boxes are widened by this slack (A) beyond the union of their
// cluster boxes, so that FP32 rounding keeps every cluster box inside.
constexpr float kScSlack = 1e-3f;

// One wave per supercluster (the atoms of sc_info[sc], in z order within
// their column):
//  - split: its atoms (lanes, padding last) are sorted by x, halves by y,
//    quarters by z, so slots 8c..8c+7 form compact clusters; permp (slot ->
//    atom, -1 padding) and slot (atom -> slot) are written;
//  - sorted SoA: coordinates relative to the supercluster origin (the
//    position of the first slot's atom rounded to 1/16 A, in sorg) + charge
//    times qpair (0 when the pair pass carries no Coulomb term),
//    FP64 copy in xs64, LJ type (padding: type -1); xbuild (MD only, else
//    null) receives the positions the list is built from;
//  - bounding boxes (FP32, list building only): per cluster (8-lane groups;
//    ctr.w = real atoms, empty clusters get negative half extents), and per
//    supercluster (the whole wave, widened by kScSlack).
// Thread 0 also clears the mask work counter of the list build that follows.
template <class R>
__device__ __forceinline__ void split_fill_body(const int bid_, const int nbk_, int nsuper_max,
                                                const int2 *__restrict__ sc_info, const int *__restrict__ order,
                                                const double *__restrict__ xwrap, int *__restrict__ permp,
                                                int *__restrict__ slot, const double *__restrict__ q,
                                                const int *__restrict__ type, typename Real4<R>::type *__restrict__ xs,
                                                double4 *__restrict__ xs64, int *__restrict__ ts,
                                                const double *__restrict__ xorig, double *__restrict__ xbuild,
                                                float4 *__restrict__ sorg, float4 *__restrict__ ctr,
                                                float4 *__restrict__ half, float4 *__restrict__ sctr,
                                                float4 *__restrict__ shalf, int *__restrict__ nmaskwork, double qpair) {
  const int sc = (bid_ * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (bid_ == 0 && threadIdx.x == 0) *nmaskwork = 0;
  if (sc >= nsuper_max) return;  // whole waves
  const int2 info = sc_info[sc];
  int a = lane < info.y ? order[info.x + lane] : -1;
  int idx = a < 0 ? 0x7fffffff : a;
  double key = a < 0 ? 1.0e300 : xwrap[3 * a];
  wave_sort<64>(key, idx);
  key = idx == 0x7fffffff ? 1.0e300 : xwrap[3 * idx + 1];
  wave_sort<32>(key, idx);
  key = idx == 0x7fffffff ? 1.0e300 : xwrap[3 * idx + 2];
  wave_sort<16>(key, idx);
  const int s = sc * kSuperSize + lane;
  a = idx == 0x7fffffff ? -1 : idx;
  permp[s] = a;
  if (a >= 0) slot[a] = s;
  // sorted SoA
  double p[3] = {0.0, 0.0, 0.0};
  if (a >= 0)
    for (int k = 0; k < 3; ++k) p[k] = xwrap[3 * a + k];
  const int a0 = __shfl(a, 0);
  double o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double p0 = __shfl(p[k], 0);
    o[k] = a0 >= 0 ? origin_round(p0) : 0.0;
  }
  if (lane == 0) sorg[sc] = make_float4((float)o[0], (float)o[1], (float)o[2], 0.f);
  typename Real4<R>::type v;
  if (a >= 0) {
    if (xbuild) {
      xbuild[3 * a] = xorig[3 * a];
      xbuild[3 * a + 1] = xorig[3 * a + 1];
      xbuild[3 * a + 2] = xorig[3 * a + 2];
    }
    v.x = (R)(p[0] - o[0]);
    v.y = (R)(p[1] - o[1]);
    v.z = (R)(p[2] - o[2]);
    v.w = (R)(q[a] * qpair);
    ts[s] = type[a];
    xs64[s] = make_double4(p[0], p[1], p[2], q[a]);
  } else {
    v.x = v.y = v.z = v.w = (R)0;
    ts[s] = -1;
    xs64[s] = make_double4(0, 0, 0, 0);
  }
  xs[s] = v;
  // boxes: clusters (8-lane groups), then the supercluster (the wave)
  double lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = a >= 0 ? p[k] : 1e300;
    hi[k] = a >= 0 ? p[k] : -1e300;
  }
  int cnt = a >= 0 ? 1 : 0;
#pragma unroll
  for (int m = 1; m < kClusterSize; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], m));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], m));
    }
    cnt += __shfl_xor(cnt, m);
  }
  if ((lane & (kClusterSize - 1)) == 0) {
    const int c = s / kClusterSize;
    if (cnt == 0) {
      ctr[c] = make_float4(0, 0, 0, 0);
      half[c] = make_float4(-1, -1, -1, 0);  // empty
    } else {
      ctr[c] = make_float4((float)(0.5 * (lo[0] + hi[0])), (float)(0.5 * (lo[1] + hi[1])),
                           (float)(0.5 * (lo[2] + hi[2])), (float)cnt);
      half[c] = make_float4((float)(0.5 * (hi[0] - lo[0])) * 1.0000005f + 1e-6f,
                            (float)(0.5 * (hi[1] - lo[1])) * 1.0000005f + 1e-6f,
                            (float)(0.5 * (hi[2] - lo[2])) * 1.0000005f + 1e-6f, 0.f);
    }
  }
#pragma unroll
  for (int m = kClusterSize; m < kWave; m <<= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], m));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], m));
    }
    cnt += __shfl_xor(cnt, m);
  }
  if (lane == 0) {
    if (cnt == 0) {
      sctr[sc] = make_float4(0, 0, 0, 0);
      shalf[sc] = make_float4(-1, -1, -1, 0);
    } else {
      sctr[sc] = make_float4((float)(0.5 * (lo[0] + hi[0])), (float)(0.5 * (lo[1] + hi[1])),
                             (float)(0.5 * (lo[2] + hi[2])), (float)cnt);
      shalf[sc] = make_float4((float)(0.5 * (hi[0] - lo[0])) + kScSlack, (float)(0.5 * (hi[1] - lo[1])) + kScSlack,
                              (float)(0.5 * (hi[2] - lo[2])) + kScSlack, 0.f);
    }
  }
}

template <class R>
__global__ void __launch_bounds__(256) k_split_fill(int nsuper_max, const int2 *__restrict__ sc_info,
                                                    const int *__restrict__ order, const double *__restrict__ xwrap,
                                                    int *__restrict__ permp, int *__restrict__ slot,
                                                    const double *__restrict__ q, const int *__restrict__ type,
                                                    typename Real4<R>::type *__restrict__ xs,
                                                    double4 *__restrict__ xs64, int *__restrict__ ts,
                                                    const double *__restrict__ xorig, double *__restrict__ xbuild,
                                                    float4 *__restrict__ sorg, float4 *__restrict__ ctr,
                                                    float4 *__restrict__ half, float4 *__restrict__ sctr,
                                                    float4 *__restrict__ shalf, int *__restrict__ nmaskwork,
                                                    double qpair) {
  split_fill_body<R>(blockIdx.x, gridDim.x, nsuper_max, sc_info, order, xwrap, permp, slot, q, type, xs, xs64, ts,
                     xorig, xbuild, sorg, ctr, half, sctr, shalf, nmaskwork, qpair);
}


// ---------------------------------------------------------------------------
// 2. The cluster list, from the cluster and supercluster boxes.

__device__ inline float bbox_dist2(float4 ca, float4 ha, float4 cb, float4 hb, const float *L) {
  float d2 = 0.f;
  float dc[3] = {fabsf(ca.x - cb.x), fabsf(ca.y - cb.y), fabsf(ca.z - cb.z)};
  float hs[3] = {ha.x + hb.x, ha.y + hb.y, ha.z + hb.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float d = fminf(dc[k], L[k] - dc[k]);
    float g = fmaxf(0.f, d - hs[k]);
    d2 += g * g;
  }
  return d2;
}

// Largest squared distance between points of two boxes (minimum image of
// the centers).
__device__ inline float bbox_far2(float4 ca, float4 ha, float4 cb, float4 hb, const float *L) {
  float d2 = 0.f;
  float dc[3] = {fabsf(ca.x - cb.x), fabsf(ca.y - cb.y), fabsf(ca.z - cb.z)};
  float hs[3] = {ha.x + hb.x, ha.y + hb.y, ha.z + hb.z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float g = fminf(dc[k], L[k] - dc[k]) + hs[k];
    d2 += g * g;
  }
  return d2;
}

// Image shift of a cluster pair: per axis s in {-1, 0, 1} with the j-center
// + s L nearest the i-center, coded (sx+1) | (sy+1) << 2 | (sz+1) << 4.
// bad is set when j-atoms shifted by s could be more than L/2 from an
// i-atom on some axis, allowing each atom to move by up to skin/2 before the
// next build (then the shift is not the minimum image of every pair and the
// pair kernel takes the per-pair image instead).
constexpr int kShiftPerPair = 64;
__device__ inline int entry_shift(float4 ci, float4 hi, float4 cj, float4 hj, const float *L, const float *Linv,
                                  float skin, bool &bad) {
  const float dc[3] = {cj.x - ci.x, cj.y - ci.y, cj.z - ci.z};
  const float hs[3] = {hi.x + hj.x, hi.y + hj.y, hi.z + hj.z};
  int code = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float s = -rintf(dc[k] * Linv[k]);
    const float d = dc[k] + s * L[k];
    bad |= fabsf(d) + hs[k] + skin >= 0.5f * L[k] * 0.999f;
    code |= ((int)s + 1) << (2 * k);
  }
  return code;
}

struct ListParams {
  float L[3];
  float Linv[3];
  float skin;        // MD pair-list buffer (0 outside MD)
  float rlist2;      // (rc_max + margin)^2
  float rlist2_sc;   // supercluster pruning radius^2 (rlist plus slack)
  float rc;          // largest cutoff (work estimate of an entry)
  int ncl;           // clusters
  int nsuper;
  int cap;           // entries per supercluster
  int ncx, ncy;      // x-y column grid of the sort
  int kx, ky;        // column reach: columns within +-kx, +-ky can hold a j-cluster in range
};

// Contiguous runs of the column window c-k..c+k on a ring of n, ascending:
// returns the number of runs (1 or 2) in r.
__device__ inline int col_runs(int c, int k, int n, int r[2][2]) {
  if (2 * k + 1 >= n) {
    r[0][0] = 0;
    r[0][1] = n - 1;
    return 1;
  }
  const int lo = c - k, hi = c + k;
  if (lo < 0) {
    r[0][0] = 0;
    r[0][1] = hi;
    r[1][0] = lo + n;
    r[1][1] = n - 1;
    return 2;
  }
  if (hi >= n) {
    r[0][0] = 0;
    r[0][1] = hi - n;
    r[1][0] = lo;
    r[1][1] = n - 1;
    return 2;
  }
  r[0][0] = lo;
  r[0][1] = hi;
  return 1;
}

// Tests j-cluster cj (< 0: none) against the eight i-clusters of S and
// appends a list entry when at least one is within the list radius.  Called
// by every thread of the block with cj ascending over threads; a block-wide
// prefix count keeps the entries sorted by cj.  Diagonal entries and entries
// with a partial cluster get a mask work slot (explicit pair masks).
constexpr int kListBlock = 256;
// Work estimate of a list entry for cutting the list into runs, in units of
// about 35 ns of one pair-kernel wave (fitted to per-wave times): kCpCost per
// cluster pair (its 64 candidates: membership and force passes), kMemCost
// times the estimated member fraction of the cluster pair (evaluation; the
// fraction is linear in the cutoff between the boxes' nearest and farthest
// distances), kEntryCost per entry (its dependent loads, the last partial
// batch, j-force sums and atomics), kTileCost on the first entry of a
// supercluster (a wave whose run crosses into it loads its i-atoms and adds
// their forces once more).
constexpr int kCpCost = 8, kMemCost = 9, kEntryCost = 14, kTileCost = 80;

__device__ inline void list_append(const ListParams &q, int S, const float4 (&ic)[8], const float4 (&ih)[8], int cj,
                                   float4 jc, float4 jh, int4 *__restrict__ entries, int *__restrict__ maskwork,
                                   int *__restrict__ nmaskwork, int maxmaskwork, int *__restrict__ overflow,
                                   int *s_cnt, int *s_cost, int *__restrict__ ecost, int &total, int &ctotal) {
  constexpr int NW = kListBlock / kWave;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t imask = 0;
  int mc = kEntryCost;  // work estimate of this entry
  bool need = false;
  int shift = -1;         // image shift code of the entry (see entry_shift)
  bool per_pair = false;  // no single shift is the minimum image for every pair
  if (cj >= 0 && jh.x >= 0.f) {
    const bool jpartial = jc.w < (float)kClusterSize;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int ci = S * 8 + c;
      if (ih[c].x < 0.f || ci > cj) continue;
      const float d2 = bbox_dist2(ic[c], ih[c], jc, jh, q.L);
      if (d2 <= q.rlist2) {
        imask |= 1u << c;
        {
          const float dn = sqrtf(d2), df = sqrtf(bbox_far2(ic[c], ih[c], jc, jh, q.L));
          const float frac = fminf(1.f, fmaxf(0.f, (q.rc - dn) / fmaxf(df - dn, 1e-3f)));
          mc += kCpCost + (int)(kMemCost * frac + 0.5f);
        }
        const bool ipartial = ic[c].w < (float)kClusterSize;
        if (ci == cj || ipartial || jpartial) need = true;
        const int code = entry_shift(ic[c], ih[c], jc, jh, q.L, q.Linv, q.skin, per_pair);
        if (shift < 0) shift = code;
        per_pair |= code != shift;
      }
    }
  }
  const unsigned long long act = __ballot(imask != 0);
  // inclusive prefix of the work over this call's entries (lanes in order)
  int wc = imask != 0 ? mc : 0;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int v = __shfl_up(wc, off);
    if (lane >= off) wc += v;
  }
  if (lane == 0) s_cnt[wv] = __popcll(act);
  if (lane == kWave - 1) s_cost[wv] = wc;
  __syncthreads();
  int before = 0, all = 0, cbefore = 0, call = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const int c = s_cnt[w], k = s_cost[w];
    before += w < wv ? c : 0;
    all += c;
    cbefore += w < wv ? k : 0;
    call += k;
  }
  __syncthreads();
  const int e = total + before + __popcll(act & ((1ull << lane) - 1ull));
  // mask work slots for this wave's entries that need explicit masks, one
  // atomic per wave
  const unsigned long long nw = __ballot(imask != 0 && need && e < q.cap);
  int wbase = 0;
  if (nw) {
    if (lane == 0) wbase = atomicAdd(nmaskwork, (int)__popcll(nw));
    wbase = __builtin_amdgcn_readfirstlane(__shfl(wbase, 0));
  }
  if (imask != 0) {
    if (e < q.cap) {
      int moff = -1;
      if (need) {
        const int w = wbase + __popcll(nw & ((1ull << lane) - 1ull));
        if (w < maxmaskwork) {
          maskwork[w] = S * q.cap + e;
          moff = -2;
        } else {
          atomicExch(overflow, 1);
        }
      }
      entries[(size_t)S * q.cap + e] = make_int4(cj, (int)imask, moff, per_pair ? kShiftPerPair : shift);
      ecost[(size_t)S * q.cap + e] = kTileCost + ctotal + cbefore + wc;  // the tile cost counts from entry 0
    } else {
      atomicExch(overflow, 1);
    }
  }
  total += all;
  ctotal += call;
}

// Cluster-pair list of supercluster S (one block each).  Superclusters are
// numbered column by column (x-major, then y, then z inside a column), so the
// j-superclusters of the columns within the column reach of S's column form
// at most four ascending index ranges per column row x.  Each range is
// pruned by supercluster boxes (a conservative test: a cluster box lies in
// its supercluster box), the survivors' clusters are tested against the
// eight i-clusters of S, and entries are appended in ascending cj.  Only
// cj >= 8S is listed, so each cluster pair appears once.  The work per S is
// bounded by the neighbourhood, independent of the system size.
__device__ __forceinline__ void build_list_body(const int bid_, const int nbk_, const ListParams &q,
                                                const int *__restrict__ sc_col, const int *__restrict__ colsc,
                                                const float4 *__restrict__ sctr, const float4 *__restrict__ shalf,
                                                const float4 *__restrict__ ctr, const float4 *__restrict__ half,
                                                int4 *__restrict__ entries, int *__restrict__ count,
                                                int *__restrict__ ecost, int *__restrict__ cost_out,
                                                int *__restrict__ maskwork, int *__restrict__ nmaskwork,
                                                int maxmaskwork, int *__restrict__ overflow) {
  constexpr int NW = kListBlock / kWave;
  constexpr int kScPerBatch = kListBlock / kClusterSize;
  __shared__ int s_cnt[NW];
  __shared__ int s_cost[NW];
  __shared__ int s_sc[kListBlock];
  __shared__ int s_ra[kWave], s_rpre[kWave], s_rtot;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int S = bid_;
  const int col = sc_col[S];
  if (col < 0) {  // past the last occupied supercluster: no atoms
    if (threadIdx.x == 0) {
      count[S] = 0;
      cost_out[S] = 0;
    }
    return;
  }
  float4 ic[8], ih[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    ic[c] = ctr[S * 8 + c];
    ih[c] = half[S * 8 + c];
  }
  const float4 Sc = sctr[S], Sh = shalf[S];
  int xr[2][2], yr[2][2];
  const int cx = col / q.ncy, cy = col - cx * q.ncy;
  const int nxr = col_runs(cx, q.kx, q.ncx, xr), nyr = col_runs(cy, q.ky, q.ncy, yr);
  // the candidate j-superclusters form index ranges, one per (x, y run) in
  // ascending order; they are enumerated as one flat sequence (64 ranges at
  // a time) so that each pass of 256 threads tests 256 candidates
  const int nx0 = xr[0][1] - xr[0][0] + 1, nxt = nx0 + (nxr > 1 ? xr[1][1] - xr[1][0] + 1 : 0);
  const int R = nxt * nyr;
  int total = 0;   // block-uniform
  int ctotal = 0;  // block-uniform work of the entries so far
  for (int g0 = 0; g0 < R; g0 += kWave) {
    const int nrg = min(kWave, R - g0);
    __syncthreads();  // every thread has read the previous group's ranges
    if (threadIdx.x < kWave) {
      const int r = g0 + lane;
      int a = 0, len = 0;
      if (r < R) {
        const int xi = r / nyr, ry = r - xi * nyr;
        const int x = xi < nx0 ? xr[0][0] + xi : xr[1][0] + (xi - nx0);
        a = max(S, colsc[x * q.ncy + yr[ry][0]]);
        len = max(0, colsc[x * q.ncy + yr[ry][1] + 1] - a);
      }
      int incl = len;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const int y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
      }
      s_ra[lane] = a;
      s_rpre[lane] = incl - len;
      if (lane == kWave - 1) s_rtot = incl;
    }
    __syncthreads();
    const int T = s_rtot;
    for (int base = 0; base < T; base += kListBlock) {
      const int k = base + threadIdx.x;
      bool pass = false;
      int sc = 0;
      if (k < T) {
        // range of candidate k: the last range starting at or before k
        int lo = 0, hi = nrg - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_rpre[mid] <= k)
            lo = mid;
          else
            hi = mid - 1;
        }
        sc = s_ra[lo] + (k - s_rpre[lo]);
        const float4 jh = shalf[sc];
        pass = jh.x >= 0.f && bbox_dist2(Sc, Sh, sctr[sc], jh, q.L) <= q.rlist2_sc;
      }
      const unsigned long long act = __ballot(pass);
      if (lane == 0) s_cnt[wv] = __popcll(act);
      __syncthreads();
      int before = 0, npass = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const int c = s_cnt[w];
        before += w < wv ? c : 0;
        npass += c;
      }
      if (pass) s_sc[before + __popcll(act & ((1ull << lane) - 1ull))] = sc;
      __syncthreads();
      for (int k0 = 0; k0 < npass; k0 += kScPerBatch) {
        const int kk = k0 + (int)(threadIdx.x / kClusterSize);
        int cj = -1;
        float4 jc = make_float4(0, 0, 0, 0), jh = make_float4(-1, -1, -1, 0);
        if (kk < npass) {
          cj = s_sc[kk] * kClusterSize + (int)(threadIdx.x & (kClusterSize - 1));
          jc = ctr[cj];
          jh = half[cj];
        }
        list_append(q, S, ic, ih, cj, jc, jh, entries, maskwork, nmaskwork, maxmaskwork, overflow, s_cnt, s_cost,
                    ecost, total, ctotal);
      }
      __syncthreads();  // s_sc and s_cnt are rewritten by the next pass
    }
  }
  if (threadIdx.x == 0) {
    const int n = total < q.cap ? total : q.cap;
    count[S] = n;
    cost_out[S] = n > 0 ? ecost[(size_t)S * q.cap + n - 1] : 0;
  }
}

__global__ void __launch_bounds__(kListBlock) k_build_list(ListParams q, const int *__restrict__ sc_col,
                                                           const int *__restrict__ colsc,
                                                           const float4 *__restrict__ sctr,
                                                           const float4 *__restrict__ shalf,
                                                           const float4 *__restrict__ ctr,
                                                           const float4 *__restrict__ half,
                                                           int4 *__restrict__ entries, int *__restrict__ count,
                                                           int *__restrict__ ecost, int *__restrict__ cost_out,
                                                           int *__restrict__ maskwork, int *__restrict__ nmaskwork,
                                                           int maxmaskwork, int *__restrict__ overflow) {
  build_list_body(blockIdx.x, gridDim.x, q, sc_col, colsc, sctr, shalf, ctr, half, entries, count, ecost, cost_out,
                  maskwork, nmaskwork, maxmaskwork, overflow);
}

// Flags the list entry that holds each excluded pair (FULL or MODIFIED) so
// that k_build_masks gives it explicit pair masks.  Entries of supercluster S
// are sorted by j-cluster, so the entry is found by binary search; pairs whose
// clusters are not in the list are beyond the list radius and never evaluated.
__device__ __forceinline__ void mark_exclusions_body(const int bid_, const int nbk_, int ne,
                                                     const int4 *__restrict__ etab, const int *__restrict__ slot,
                                                     int4 *__restrict__ entries, const int *__restrict__ count, int cap,
                                                     int *__restrict__ maskwork, int *__restrict__ nmaskwork,
                                                     int maxmaskwork, int *__restrict__ overflow) {
  const int e = bid_ * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int hit = -1;  // entry index (S * cap + k) this thread claimed
  if (e < ne) {
    const int4 t = etab[e];
    const int c1 = slot[t.x] / kClusterSize, c2 = slot[t.y] / kClusterSize;
    if (c1 != c2) {  // diagonal entries always carry masks
      const int ci = c1 < c2 ? c1 : c2, cj = c1 < c2 ? c2 : c1;
      const int S = ci / 8;
      int lo = 0, hi = count[S] - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = entries[(size_t)S * cap + mid].x;
        if (v == cj) {
          if (atomicCAS(&entries[(size_t)S * cap + mid].z, -1, -3) == -1) hit = S * cap + mid;
          break;
        }
        if (v < cj)
          lo = mid + 1;
        else
          hi = mid - 1;
      }
    }
  }
  // one mask work slot per claimed entry, one counter atomic per wave
  const unsigned long long got = __ballot(hit >= 0);
  int wbase = 0;
  if (got) {
    if (lane == 0) wbase = atomicAdd(nmaskwork, (int)__popcll(got));
    wbase = __builtin_amdgcn_readfirstlane(__shfl(wbase, 0));
  }
  if (hit >= 0) {
    const int w = wbase + __popcll(got & ((1ull << lane) - 1ull));
    if (w < maxmaskwork) {
      maskwork[w] = hit;
      atomicExch(&entries[hit].z, -2);
    } else {
      atomicExch(overflow, 1);
    }
  }
}

__global__ void k_mark_exclusions(int ne, const int4 *__restrict__ etab, const int *__restrict__ slot,
                                  int4 *__restrict__ entries, const int *__restrict__ count, int cap,
                                  int *__restrict__ maskwork, int *__restrict__ nmaskwork, int maxmaskwork,
                                  int *__restrict__ overflow) {
  mark_exclusions_body(blockIdx.x, gridDim.x, ne, etab, slot, entries, count, cap, maskwork, nmaskwork, maxmaskwork,
                       overflow);
}

// Binary search of b in the sorted exclusion partners of a (original indices).
__device__ inline int excl_class(const int *__restrict__ eptr, const int2 *__restrict__ elist, int a, int b) {
  int lo = eptr[a], hi = eptr[a + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    int v = elist[mid].x;
    if (v == b) return elist[mid].y;
    if (v < b)
      lo = mid + 1;
    else
      hi = mid - 1;
  }
  return 0;
}

// One wave per entry needing explicit masks.  Mask slot w holds one 16-bit
// word per lane l = tj*8 + ti of the pair kernel: bit c set when i = slot
// 8ci+ti (ci = 8S + c, c in the entry's mask) and j = slot 8cj+tj are real
// atoms (perm >= 0), i precedes j within a shared cluster, and the pair is
// not FULL-excluded; bit 8 + c marks MODIFIED.  Each lane of the pair kernel
// loads its own word once per entry.  The entry's mask field becomes w.
__device__ __forceinline__ void build_masks_body(const int bid_, const int nbk_, const int *__restrict__ maskwork,
                                                 const int *__restrict__ nmaskwork, int maxmaskwork,
                                                 int4 *__restrict__ entries, int cap, const int *__restrict__ perm,
                                                 const int *__restrict__ eptr, const int2 *__restrict__ elist,
                                                 unsigned short *__restrict__ masks) {
  const int lane = threadIdx.x & 63;
  int nw = *nmaskwork;
  if (nw > maxmaskwork) nw = maxmaskwork;
  const int nwaves = nbk_ * (blockDim.x / kWave);
  for (int w = bid_ * (blockDim.x / kWave) + (threadIdx.x >> 6); w < nw; w += nwaves) {
    const int eidx = maskwork[w];
    const int S = eidx / cap;
    const int4 ent = entries[eidx];
    const int cj = ent.x;
    const int ti = lane & 7, tj = lane >> 3;
    const int aj = perm[cj * kClusterSize + tj];
    // the eight i-clusters' exclusion lookups advance together (one
    // dependent load chain instead of eight)
    int ai[8], lo[8], hi[8], cls[8];
    bool ok[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int ci = S * 8 + c;
      ai[c] = (ent.y & (1 << c)) ? perm[ci * kClusterSize + ti] : -1;
      ok[c] = ai[c] >= 0 && aj >= 0 && (ci != cj || ti < tj);
      lo[c] = ok[c] ? eptr[ai[c]] : 0;
      hi[c] = ok[c] ? eptr[ai[c] + 1] - 1 : -1;
      cls[c] = 0;
    }
    bool any = true;
    while (any) {
      any = false;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if (lo[c] <= hi[c]) {
          const int mid = (lo[c] + hi[c]) >> 1;
          const int2 v = elist[mid];
          if (v.x == aj) {
            cls[c] = v.y;
            hi[c] = lo[c] - 1;
          } else if (v.x < aj) {
            lo[c] = mid + 1;
          } else {
            hi[c] = mid - 1;
          }
          any = true;
        }
      }
    }
    unsigned bits = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      // ok[c] implies c is in the entry's mask
      if (ok[c] && cls[c] != 1) bits |= 1u << c;
      if (ok[c] && cls[c] == 2) bits |= 1u << (8 + c);
    }
    masks[(size_t)w * kWave + lane] = (unsigned short)bits;
    if (lane == 0) entries[eidx].z = w;
  }
}

__global__ void __launch_bounds__(256) k_build_masks(const int *__restrict__ maskwork, const int *__restrict__ nmaskwork,
                                                     int maxmaskwork, int4 *__restrict__ entries, int cap,
                                                     const int *__restrict__ perm, const int *__restrict__ eptr,
                                                     const int2 *__restrict__ elist,
                                                     unsigned short *__restrict__ masks) {
  build_masks_body(blockIdx.x, gridDim.x, maskwork, nmaskwork, maxmaskwork, entries, cap, perm, eptr, elist, masks);
}

// Compact tile list: tiles of up to kTileEntries entries per supercluster.
template <class T>
__device__ inline int last_le(const T *__restrict__ a, int n, T v) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Run schedule: the list entries of all superclusters, in supercluster
// order, are cut into nruns contiguous runs of equal estimated work (the
// work estimates of k_build_list: per supercluster, and as an inclusive
// prefix over its entries); nruns, a multiple of kMinRuns set by C, the
// total, depends only on the list.  Run w starts at entry
// position pos[w] (ep[S] + offset, offsets even) in supercluster posS[w],
// where the global work prefix first exceeds b_w = floor(w C / nruns).  A run
// is split into one tile per supercluster it touches (at most nruns + nsuper
// tiles); run r is tiles [wstart[r], wstart[r + 1]).  Three passes, none
// searching:
//   k_runs_prefix (one block): prefixes ep (entries) and cp (work) over the
//     superclusters, the run count, run boundaries preset to the end;
//   k_runs_cut (a block per supercluster): each entry places the run
//     boundaries b_w that fall in its work interval;
//   k_runs_tiles (one block): tiles per run, their prefix, the tiles.
__device__ __forceinline__ void runs_prefix_body(const int bid_, const int nbk_, int nsuper,
                                                 const int *__restrict__ count, const int *__restrict__ cost,
                                                 int *__restrict__ nruns_p, int *__restrict__ ep,
                                                 long long *__restrict__ cp, int *__restrict__ pos,
                                                 int *__restrict__ posS, long long min_run_cost = 0) {
  __shared__ int s_e[1024];
  __shared__ long long s_c[1024];
  __shared__ int s_last[1024 / kWave];
  const int tid = threadIdx.x;
  int ecarry = 0, last = -1;
  long long ccarry = 0;
  for (int base = 0; base < nsuper; base += 1024) {
    const int S = base + tid;
    const int c = S < nsuper ? count[S] : 0;
    const long long w = S < nsuper ? (long long)cost[S] : 0;
    if (c > 0) last = S;
    s_e[tid] = c;
    s_c[tid] = w;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int ve = tid >= off ? s_e[tid - off] : 0;
      const long long vc = tid >= off ? s_c[tid - off] : 0;
      __syncthreads();
      s_e[tid] += ve;
      s_c[tid] += vc;
      __syncthreads();
    }
    if (S < nsuper) {
      ep[S] = ecarry + s_e[tid] - c;
      cp[S] = ccarry + s_c[tid] - w;
    }
    ecarry += s_e[1023];
    ccarry += s_c[1023];
    __syncthreads();
  }
  last = wave_max(last);
  if ((tid & 63) == 0) s_last[tid >> 6] = last;
  __syncthreads();
  int lastS = -1;
  for (int k = 0; k < 1024 / kWave; ++k) lastS = max(lastS, s_last[k]);
  const long long per = (long long)kMinRuns * kRunCost;
  const long long k = (ccarry + per - 1) / per;
  int nruns = kMinRuns * (int)(k < 1 ? 1 : (k > kMaxRuns / kMinRuns ? kMaxRuns / kMinRuns : k));
  if (min_run_cost > 0) {
    // runs of at least min_run_cost units, in whole waves' worth of 64 runs:
    // a small list's runs would otherwise hold an entry or two each, and the
    // wave would spend its time on the i-force flush at every tile end
    const long long r = (ccarry + min_run_cost - 1) / min_run_cost;
    nruns = (int)min((long long)nruns, max(64ll, (r + 63) / 64 * 64));
  }
  if (tid == 0) {
    ep[nsuper] = ecarry;
    cp[nsuper] = ccarry;
    *nruns_p = nruns;
  }
  // every boundary starts at the end (k_runs_cut moves those that fall in
  // the list); run 0 starts at the beginning
  for (int w = tid; w <= nruns; w += 1024) {
    pos[w] = w == 0 ? 0 : ecarry;
    posS[w] = w == 0 ? 0 : lastS + 1;
  }
}

__global__ void k_runs_prefix(int nsuper, const int *__restrict__ count, const int *__restrict__ cost,
                              int *__restrict__ nruns_p, int *__restrict__ ep, long long *__restrict__ cp,
                              int *__restrict__ pos, int *__restrict__ posS) {
  runs_prefix_body(blockIdx.x, gridDim.x, nsuper, count, cost, nruns_p, ep, cp, pos, posS);
}

__device__ __forceinline__ void runs_cut_body(const int bid_, const int nbk_, int nsuper, const int *__restrict__ count,
                                              const int *__restrict__ ecost, int cap, const int *__restrict__ nruns_p,
                                              const int *__restrict__ ep, const long long *__restrict__ cp,
                                              int *__restrict__ pos, int *__restrict__ posS) {
  const int S = bid_;
  const int nruns = *nruns_p;
  const int n = count[S];
  const long long C = cp[nsuper], base = cp[S];
  if (C <= 0) return;
  const int *ec = ecost + (size_t)S * cap;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const long long lo = base + (e > 0 ? ec[e - 1] : 0), hi = base + ec[e];
    // runs w in [1, nruns) with lo <= floor(w C / nruns) < hi
    long long w0 = (lo * nruns + C - 1) / C, w1 = (hi * nruns + C - 1) / C;
    w0 = max(w0, 1ll);
    w1 = min(w1, (long long)nruns);
    for (long long w = w0; w < w1; ++w) {
      pos[w] = ep[S] + (e & ~1);  // runs hold whole pair-kernel groups
      posS[w] = S;
    }
  }
}

__global__ void k_runs_cut(int nsuper, const int *__restrict__ count, const int *__restrict__ ecost, int cap,
                           const int *__restrict__ nruns_p, const int *__restrict__ ep,
                           const long long *__restrict__ cp, int *__restrict__ pos, int *__restrict__ posS) {
  runs_cut_body(blockIdx.x, gridDim.x, nsuper, count, ecost, cap, nruns_p, ep, cp, pos, posS);
}

__device__ __forceinline__ void runs_tiles_body(const int bid_, const int nbk_, const int *__restrict__ count,
                                                const int *__restrict__ ep, const int *__restrict__ pos,
                                                const int *__restrict__ posS, int *__restrict__ wstart,
                                                int4 *__restrict__ tiles, int *__restrict__ ntiles, int ntiles_cap) {
  __shared__ int s_n[1024];
  const int tid = threadIdx.x;
  const int nruns = ntiles[2];  // from k_runs_prefix
  // run w: entries [pos[w], pos[w+1]), superclusters S0 .. S1
  auto span = [&](int w, int &S0, int &S1) {
    const int p0 = pos[w], p1 = pos[w + 1];
    S0 = posS[w];
    S1 = posS[w + 1] - (p1 == ep[posS[w + 1]] ? 1 : 0);
    return p1 > p0 ? S1 - S0 + 1 : 0;
  };
  int carry = 0;
  for (int base = 0; base < nruns; base += 1024) {
    const int w = base + tid;
    int S0, S1;
    const int n = w < nruns ? span(w, S0, S1) : 0;
    s_n[tid] = n;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int v = tid >= off ? s_n[tid - off] : 0;
      __syncthreads();
      s_n[tid] += v;
      __syncthreads();
    }
    if (w < nruns) {
      int t = carry + s_n[tid] - n;
      wstart[w] = min(t, ntiles_cap);
      for (int S = S0; S < S0 + n && t < ntiles_cap; ++S, ++t) {
        const int e0 = S == S0 ? pos[w] - ep[S] : 0;
        const int e1 = S == S1 ? pos[w + 1] - ep[S] : count[S];
        tiles[t] = make_int4(S, e0, e1, 0);
      }
    }
    carry += s_n[1023];
    __syncthreads();
  }
  if (tid == 0) {
    wstart[nruns] = min(carry, ntiles_cap);
    ntiles[0] = min(carry, ntiles_cap);
  }
}

__global__ void k_runs_tiles(const int *__restrict__ count, const int *__restrict__ ep, const int *__restrict__ pos,
                             const int *__restrict__ posS, int *__restrict__ wstart, int4 *__restrict__ tiles,
                             int *__restrict__ ntiles, int ntiles_cap) {
  runs_tiles_body(blockIdx.x, gridDim.x, count, ep, pos, posS, wstart, tiles, ntiles, ntiles_cap);
}

// ---------------------------------------------------------------------------
// 3. Pair kernel: LJ + Coulomb (cutoff or PME real space) with switching.
//
// Wavefront layout: lane = 8*tj + ti.  A work tile is (i-supercluster S,
// up to kTileEntries list entries).  Each entry names one j-cluster (8 atoms,
// lane tj holds atom tj) and an 8-bit mask of the i-clusters of S it is in
// range of; for every set i-cluster c the 64 lanes evaluate the 8x8 atom
// pairs (8c+ti, tj).  i-forces stay in registers for the whole tile, j-forces
// are summed over the 8 i-lanes with DPP and added once per entry.
//
// Membership is decided exactly: a pair whose FP32/FP64 r^2 lies within
// `guard` of a cutoff^2 is not evaluated here but recorded, and the
// boundary-pair pass of k_fixups re-tests it in FP64 with the reference's
// operation order on the original coordinates before evaluating it.

__device__ inline float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double fast_rsqrt(double x) { return 1.0 / sqrt(x); }
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double fast_rcp(double x) { return 1.0 / x; }
__device__ inline float fast_exp(float x) { return __expf(x); }
__device__ inline double fast_exp(double x) { return exp(x); }
__device__ inline float fast_rint(float x) { return rintf(x); }
__device__ inline double fast_rint(double x) { return rint(x); }

template <class R>
__device__ inline R poly5(R ar) {
  // Abramowitz & Stegun 7.1.26 (reference: protomol/base/MathUtilities.cpp:39-49)
  const R E1 = (R)0.254829592, E2 = (R)-0.284496736, E3 = (R)1.421413741, E4 = (R)-1.453152027,
          E5 = (R)1.061405429, PP = (R)0.3275911;
  R t = fast_rcp((R)1 + PP * ar);
  return t * (E1 + t * (E2 + t * (E3 + t * (E4 + t * E5))));
}

template <bool MIXED>
struct AccT;
template <>
struct AccT<false> {
  using type = double;
  __device__ static void add(double *p, double v) { unsafeAtomicAdd(p, v); }
};
template <>
struct AccT<true> {
  using type = unsigned long long;
  __device__ static void add(unsigned long long *p, float v) {
    long long q = __float2ll_rn(v * 4294967296.0f);
    atomicAdd(p, (unsigned long long)q);
  }
  __device__ static void add(unsigned long long *p, double v) {
    long long q = __double2ll_rn(v * kForceScale);
    atomicAdd(p, (unsigned long long)q);
  }
};

// Switch constants in the arithmetic type of the kernel.
template <class R>
struct SwR {
  R rc2, on2, s1, s2, s3, c15_1, c05_3, c15_3;
  __device__ void load(const SwitchC &s) {
    rc2 = (R)s.rc2;
    on2 = (R)s.on2;
    s1 = (R)s.s1;
    s2 = (R)s.s2;
    s3 = (R)s.s3;
    c15_1 = (R)s.c15_1;
    c05_3 = (R)s.c05_3;
    c15_3 = (R)s.c15_3;
  }
};

// Switching function value and (1/r) dS/dr at r2 (in range), branch-free.
// KIND = SW_ANY selects at run time from `kind`.
constexpr int SW_ANY = 9;
template <int KIND, class R>
__device__ inline void switch_val(int kind, const SwR<R> &s, R r2, R rinv, R &val, R &der) {
  const int k = KIND == SW_ANY ? kind : KIND;
  if (k == SW_C2) {
    // C2SwitchingFunction.h:42-53: 1 below r_on, polynomial on [r_on, rc]
    R c2 = s.rc2 - r2;
    c2 = c2 > (R)0 ? c2 : (R)0;
    const R c4 = c2 * (s.s2 + (R)2 * r2);
    const bool on = r2 >= s.on2;
    val = on ? s.s1 * (c2 * c4) : (R)1;
    der = on ? s.s3 * (c2 * c2 - c4) : (R)0;
  } else if (k == SW_C1) {
    // C1SwitchingFunction.h:41-50
    const R dist = r2 * rinv;
    val = (R)1 - dist * (s.c15_1 - r2 * s.c05_3);
    der = -s.c15_1 * rinv + dist * s.c15_3;
  } else {
    val = (R)1;
    der = (R)0;
  }
}

// One pair's energies and scalar force (F = f * d), given membership flags.
template <int COUL, int LJSW, int CSW, class R>
__device__ inline void pair_eval(const PairParams &pp, const SwR<R> &sl, const SwR<R> &sc, R alpha, R a2pi, R csf,
                                 R r2, R qq, R A, R B, bool in_lj, bool in_c, R &elj, R &eco, R &fr) {
  const R r2s = (in_lj || in_c) ? r2 : (R)1;
  const R rinv = fast_rsqrt(r2s);
  const R rinv2 = rinv * rinv;
  fr = (R)0;
  elj = (R)0;
  eco = (R)0;
  if (LJSW != SW_NONE) {
    // LennardJonesForce.h:28-60
    const R r6 = rinv2 * rinv2 * rinv2;
    const R r6B = B * r6, r12A = A * r6 * r6;
    R e = r12A - r6B;
    R f = ((R)12 * r12A - (R)6 * r6B) * rinv2;
    R sv, sd;
    switch_val<LJSW, R>(pp.lj.kind, sl, r2s, rinv, sv, sd);
    f = f * sv - e * sd;
    e = e * sv;
    elj = in_lj ? e : (R)0;
    fr += in_lj ? f : (R)0;
  }
  if (COUL == COUL_CUTOFF) {
    // CoulombForce.h:29-40
    R e = qq * rinv;
    R f = e * rinv2;
    if constexpr (CSW != SW_CUTOFF) {  // a plain cutoff switches nothing (1 and 0)
      R sv, sd;
      switch_val<CSW, R>(pp.co.kind, sc, r2s, rinv, sv, sd);
      f = f * sv - e * sd;
      e = e * sv;
    }
    eco = in_c ? e : (R)0;
    fr += in_c ? f : (R)0;
  } else if (COUL == COUL_PME) {
    // NonbondedPMEwaldSystemForce.h:657-683 (poly5 form)
    const R r = r2s * rinv;
    const R ar = alpha * r;
    const R ex = qq * fast_exp(-ar * ar);
    R e = poly5<R>(ar) * ex * rinv;
    R f = (e + a2pi * ex) * rinv2;
    if constexpr (CSW != SW_CUTOFF) {
      R sv, sd;
      switch_val<CSW, R>(pp.co.kind, sc, r2s, rinv, sv, sd);
      f = f * sv - e * sd;
      e = e * sv;
    }
    eco = in_c ? e : (R)0;
    fr += in_c ? f : (R)0;
  }
}

// Sum over the 8 lanes of a cluster group (lanes differing in bits 0..2).
__device__ inline float group8_sum(float v) {
  // quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, true));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, true));
  return v;
}
__device__ inline double group8_sum(double v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  return v;
}
struct PairOut {
  double *energy;            // [2 * ntiles_max]: per tile (coulomb, lj); int64 fixed point 2^-32 in the mixed form
  long long *npairs;         // [ntiles_max] when stats
  unsigned long long *hash;  // [ntiles_max]
  int2 *border;              // recorded boundary pairs (sorted slots i, j)
  int *nborder;
  int maxborder;
  int *overflow;
  unsigned long long *border_acc;  // [4] boundary-pair energies and stats, cleared here
};

// Queue entry of the compacted pair kernel: one in-cutoff pair awaiting its
// force.  info packs its force slot (bits 0..9: entry of the group, i-cluster
// and lane, kGroup * 8 * 64 slots), MODIFIED (bit 10), the i-atom of the
// supercluster (bits 11..16) and the group's j-atom (bits 17..20); the
// evaluating lane reads the atoms' charges and types from LDS.
template <class R>
struct PairQ {
  R r2;
  int info;
};
constexpr int kSlotBits = 10;
// r^2 given to an excluded candidate pair: beyond every cutoff, band and
// list radius
template <class R>
constexpr R kFarR2 = (R)1e30;
constexpr int kPairPlain = 0, kPairPrune = 1, kPairStats = 2;  // k_pair MODE

// Charge and LJ type (clamped to 0 for padding) of a j-atom of the group.
template <class R>
struct JAtom {
  R q;
  int ty;
};

constexpr int kGroup = 2;    // list entries whose pairs share one queue flush
constexpr int kQueue = 128;  // queue ring per wave (holds < 64 + 64 pairs)
static_assert(kGroup == 2, "runs are cut at even entries (k_runs_cut)");

// Wave-private LDS area of the pair kernel: the force slots of a group and
// the queue; at a segment end, when both are empty, it holds the transpose of
// one component of the i-force partial sums (8 values per lane).  (Kept
// small: an FFT block of the PME chain must fit on a CU beside the pair
// kernel's blocks.)
template <class R>
struct PairArea {
  static constexpr int kSlotBytes = (int)sizeof(R) * kGroup * 8 * kWave;
  static constexpr int kQueueBytes = (int)sizeof(PairQ<R>) * kQueue;
  static constexpr int kTransBytes = (int)sizeof(R) * 8 * kWave;
  static constexpr int kBytes =
      kSlotBytes + kQueueBytes > kTransBytes ? kSlotBytes + kQueueBytes : kTransBytes;
};

// Fixed-point (mixed form) or FP64 sums of a wave's own i-atom forces and of
// its pair energies.
template <bool MIXED>
struct WaveAcc;
template <>
struct WaveAcc<true> {
  using T = long long;
  __device__ static T of(float v) { return __float2ll_rn(v * 4294967296.0f); }
  __device__ static void add(unsigned long long *p, T v) { atomicAdd(p, (unsigned long long)v); }
  __device__ static double energy_bits(T v) { return __longlong_as_double(v); }
};
template <>
struct WaveAcc<false> {
  using T = double;
  __device__ static T of(double v) { return v; }
  __device__ static void add(double *p, T v) { unsafeAtomicAdd(p, v); }
  __device__ static double energy_bits(T v) { return v; }
};

// Box edges in the kernel's type, split as hi + lo (see PairParams::Lhi).
template <class R>
struct EdgesR {
  R hi[3], lo[3], inv[3];
  __device__ void load(const PairParams &pp) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      hi[k] = (R)pp.Lhi[k];
      lo[k] = (R)pp.Llo[k];
      inv[k] = (R)pp.box.Linv[k];
    }
  }
  // image k of a difference: d - k (hi + lo), hi exact
  __device__ R image(R d, int ax) const {
    const R k = fast_rint(d * inv[ax]);
    return (d - k * hi[ax]) - k * lo[ax];
  }
};

// Minimum-image difference j - i in the kernel's arithmetic type.  Sorted
// coordinates are relative to supercluster origins; xj has already been
// moved by the entry offset (origin of j's supercluster - origin of i's,
// plus the entry's image shift unless the entry needs the per-pair image,
// per_pair, wave-uniform).  Phase A and phase C of k_pair call it with the
// same operands.
template <class R, class R4>
__device__ inline void pair_delta(const R4 &xi, const R4 &xj, const EdgesR<R> &E, int per_pair, R &dx, R &dy,
                                  R &dz) {
  dx = xj.x - xi.x;
  dy = xj.y - xi.y;
  dz = xj.z - xi.z;
  if (per_pair) {
    dx = E.image(dx, 0);
    dy = E.image(dy, 1);
    dz = E.image(dz, 2);
  }
}

// Cluster-pair kernel with compaction.  For a group of kGroup list entries:
//   A. every candidate lane (i = 8c+ti of the supercluster, j = lane tj of
//      the entry's j-cluster) computes r^2 and its membership; members are
//      appended to a per-wave LDS queue (ballot + mbcnt), and every full
//      64-pair batch of the queue is evaluated at once, all lanes busy;
//   B. the last partial batch is evaluated; each pair's scalar force lands in
//      its slot of an LDS array (zero for non-members);
//   C. the candidate lanes read their slot and accumulate F = f d into the
//      i-forces (registers) and j-forces (summed over the 8 i-lanes with DPP,
//      one fixed-point atomic add per j-atom component and entry).
// Only phase B does the expensive math (rsqrt, exp, erfc polynomial,
// switches), on members only, which are about a quarter of the candidates.
// Deterministic sums: a tile (a run's part of one supercluster's list) is the
// same whichever wave takes it, and at its end the i-force partial sums are
// transposed through LDS, lane 8 c + ti sums its i-atom's eight partials in
// a fixed order and adds the sum to the force array in fixed point; the
// tile's FP32 pair-energy sums go to fixed point too.  So the mixed form's forces and energies do not depend
// on the schedule, the grid or the device.
//
// MODE kPairPrune (the first pass after a list build): cluster pairs without
// a candidate pair inside the list radius are dropped from the entries'
// masks (through entries_w, which aliases entries; each entry is read and
// written by the one wave that takes it).  MODE kPairStats: the pass also
// counts the in-cutoff pairs and sums their hashes.
template <bool MIXED, int COUL, int LJSW, int CSW, bool BIGLJ, int MODE>
__device__ __forceinline__ void pair_body(const int bid_, const int nbk_, PairParams pp, int dynamic,
    const int4 *__restrict__ tiles, const int *__restrict__ ntiles_p, const int *__restrict__ wstart,
    int *__restrict__ tile_counter, const int4 *__restrict__ entries, int4 *entries_w, int cap,
    const unsigned short *__restrict__ masks,
    const typename Real4<typename std::conditional<MIXED, float, double>::type>::type *__restrict__ xs,
    const float4 *__restrict__ sorg, const int *__restrict__ ts, const int *__restrict__ perm,
    const double *__restrict__ ljtab,
    const typename Real4<typename std::conditional<MIXED, float, double>::type>::type *__restrict__ ljt,
    typename AccT<MIXED>::type *__restrict__ facc, PairOut out) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using R4 = typename Real4<R>::type;
  using Acc = AccT<MIXED>;
  using WA = WaveAcc<MIXED>;
  using AT = typename WA::T;
  constexpr int W = kPairBlock / kWave;
  constexpr bool PRUNE = MODE == kPairPrune, STATS = MODE == kPairStats;
  __shared__ R4 s_i[W][kSuperSize];
  __shared__ int s_ty[W][kSuperSize];
  __shared__ R4 s_lj[kMaxTypes * kMaxTypes];  // {A, B, A14, B14}
  __shared__ __align__(16) unsigned char s_area[W][PairArea<R>::kBytes];
  __shared__ JAtom<R> s_j[W][kGroup * kClusterSize];

  const int nt = pp.ntypes;
  // up to kMaxTypes LJ types the table sits in LDS; beyond that (BIGLJ) the
  // typed copy ljt is read through the caches
  for (int k = threadIdx.x; !BIGLJ && k < nt * nt; k += blockDim.x) {
    R4 v;
    v.x = (R)ljtab[4 * k];
    v.y = (R)ljtab[4 * k + 1];
    v.z = (R)ljtab[4 * k + 2];
    v.w = (R)ljtab[4 * k + 3];
    s_lj[k] = v;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int ti = lane & 7, tj = lane >> 3;
  // queue info of this lane's pairs: force slot (bits 0..9, with the group
  // entry and i-cluster added), MODIFIED (bit 10), i-atom c * 8 + ti (bits
  // 11..16, c added), j-atom g * 8 + tj (bits 17..20, g added)
  const int info_lane = lane | (ti << 11) | (tj << 17);
  EdgesR<R> E;
  E.load(pp);
  SwR<R> sl, sc;
  sl.load(pp.lj);
  sc.load(pp.co);
  const R guard = (R)pp.guard;
  const R alpha = (R)pp.alpha, a2pi = (R)pp.a2pi, csf = (R)pp.csf;
  const bool one_cut = sl.rc2 == sc.rc2;  // one guard band serves both cutoffs
  const bool same_cut = LJSW == SW_NONE || COUL == COUL_NONE || one_cut;
  // membership: r^2 within the larger cutoff of the kinds present
  const R rmax2 = LJSW == SW_NONE ? sc.rc2 : (COUL == COUL_NONE ? sl.rc2 : (sl.rc2 > sc.rc2 ? sl.rc2 : sc.rc2));
  const R prune_r2 = (R)pp.prune_r2;
  // the boundary-pair sums are filled after this kernel (same stream)
  if (bid_ == 0 && threadIdx.x < 4) out.border_acc[threadIdx.x] = 0ull;
  R4 *const si = s_i[wv];
  R *const fr_slot = reinterpret_cast<R *>(s_area[wv]);
  PairQ<R> *const q = reinterpret_cast<PairQ<R> *>(s_area[wv] + PairArea<R>::kSlotBytes);
  R *const tr = reinterpret_cast<R *>(s_area[wv]);

  // Runs (the grid holds the waves that fit on the device at once): wave w
  // takes the contiguous runs [w nruns / nwaves, (w + 1) nruns / nwaves)
  // (dynamic == 0: consecutive tiles often share a supercluster, whose
  // i-atoms then stay loaded), or
  // run w and then runs in order from a device-wide counter (dynamic == 1).
  // Run r is tiles [wstart[r], wstart[r + 1]).  Which wave takes a run
  // changes no result.
  const int nwaves = nbk_ * W;
  const int gw = bid_ * W + wv;
  const int nruns = ntiles_p[2];
  const int r0 = dynamic ? gw : (int)((long long)gw * nruns / nwaves);
  const int r1 = dynamic ? min(gw + 1, nruns) : (int)((long long)(gw + 1) * nruns / nwaves);
  int t = r0 < r1 ? wstart[r0] : 0, tstop = r0 < r1 ? wstart[r1] : 0;
  int curS = -1;      // supercluster whose i-atoms are loaded
  bool ireal = false;  // slot `lane` of curS holds an atom
  while (true) {
    if (t >= tstop) {
      if (!dynamic) break;
      int rn = 0;
      if (lane == 0) rn = nwaves + atomicAdd(tile_counter, 1);
      const int r = __builtin_amdgcn_readfirstlane(__shfl(rn, 0));
      if (r >= nruns) break;
      t = wstart[r];
      tstop = wstart[r + 1];
      continue;
    }
    const int4 tile = tiles[t];
    const int S = __builtin_amdgcn_readfirstlane(tile.x);
    const int e0 = __builtin_amdgcn_readfirstlane(tile.y);
    const float4 oS = sorg[S];
    if (S != curS) {
      curS = S;
      si[lane] = xs[S * kSuperSize + lane];
      const int v = ts[S * kSuperSize + lane];
      ireal = v >= 0;
      s_ty[wv][lane] = (v < 0 ? 0 : v) * nt;  // row of the LJ table
      __builtin_amdgcn_wave_barrier();
    }

    R fi[8][3];
#pragma unroll
    for (int c = 0; c < 8; ++c) fi[c][0] = fi[c][1] = fi[c][2] = (R)0;
    R elT = 0, ecT = 0;  // this lane's pair energies over the tile
    int npair = 0;
    unsigned long long hsum = 0;
    int qhead = 0, qtail = 0;  // wave-uniform queue positions

    // Evaluates queue pairs [qhead, qhead + cnt) with lane l on pair qhead + l.
    // A queued pair passed the membership test: it is in range of at least
    // one of the two cutoffs, and not within the guard band of either.
    // `inside` is std::true_type when every queued pair is inside the cutoff
    // of each kind present (one cutoff for both kinds), so no per-pair test
    // or select is needed.
    auto eval_batch = [&](auto inside, int cnt, R &qel, R &qec) {
      __builtin_amdgcn_wave_barrier();
      if (lane < cnt) {
        const PairQ<R> e = q[(qhead + lane) & (kQueue - 1)];
        const int slot = e.info & ((1 << kSlotBits) - 1);
        const bool mod = (e.info >> kSlotBits) & 1;
        const int ia = (e.info >> 11) & 63;  // c * 8 + ti
        const int ja = (e.info >> 17) & 15;  // g * 8 + tj
        const JAtom<R> jd = s_j[wv][ja];
        R qq = si[ia].w * jd.q;
        if (mod) qq *= csf;
        const bool il = LJSW != SW_NONE && (decltype(inside)::value || e.r2 <= sl.rc2);
        const bool ic = COUL != COUL_NONE && (decltype(inside)::value || e.r2 <= sc.rc2);
        R A = 0, B = 0;
        if (LJSW != SW_NONE) {
          const int li = s_ty[wv][ia] + jd.ty;
          const R4 lj = BIGLJ ? ljt[li] : s_lj[li];
          A = mod ? lj.z : lj.x;
          B = mod ? lj.w : lj.y;
        }
        R e1, e2, f;
        pair_eval<COUL, LJSW, CSW, R>(pp, sl, sc, alpha, a2pi, csf, e.r2, qq, A, B, il, ic, e1, e2, f);
        qel += e1;
        qec += e2;
        fr_slot[slot] = f;
      }
      qhead += cnt;
    };

    // Tile end (queue empty, slots free): the tile's FP32 i-force partial
    // sums go through LDS one component at a time so that lane 8 c + ti sums
    // the eight partials of its i-atom (from lanes (ti, tj'), tj' = 0..7 in
    // order) and adds the sum, in fixed point, to the atom's force.
    auto flush = [&]() {
      const int sidx = S * kSuperSize + lane;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int c = 0; c < 8; ++c) tr[lane * 8 + c] = fi[c][k];
        __builtin_amdgcn_wave_barrier();
        R sk = 0;
#pragma unroll
        for (int tq = 0; tq < 8; ++tq) sk += tr[(8 * tq + ti) * 8 + tj];
        if (ireal) WA::add(&facc[3 * sidx + k], WA::of(sk));
      }
      __builtin_amdgcn_wave_barrier();
    };

    const int eend = __builtin_amdgcn_readfirstlane(tile.z);
    for (int eg = e0; eg < eend; eg += kGroup) {
      const int ng = eend - eg < kGroup ? eend - eg : kGroup;
      int cjv[kGroup], emv[kGroup], mov[kGroup], shv[kGroup], tyj0[kGroup];
      unsigned mbv[kGroup];  // this lane's exclusion bits (candidate per i-cluster, MODIFIED << 8)
      R4 xj[kGroup];
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        int4 ent = make_int4(0, 0, -1, 0);
        if (g < ng) ent = entries[(size_t)S * cap + eg + g];
        cjv[g] = __builtin_amdgcn_readfirstlane(ent.x);
        emv[g] = __builtin_amdgcn_readfirstlane(ent.y) & 0xff;
        mov[g] = __builtin_amdgcn_readfirstlane(ent.z);
        shv[g] = __builtin_amdgcn_readfirstlane(ent.w);
        const int sj = cjv[g] * kClusterSize + tj;
        xj[g] = xs[sj];
        {
          // j-cluster into the frame of S: origin difference (exact) plus
          // the entry's image shift (hi part exact, lo part rounded once)
          const float4 oJ = sorg[cjv[g] / (kSuperSize / kClusterSize)];
          R ox = (R)(oJ.x - oS.x), oy = (R)(oJ.y - oS.y), oz = (R)(oJ.z - oS.z);
          if (!(shv[g] & kShiftPerPair)) {
            const R sx = (R)((shv[g] & 3) - 1), sy = (R)(((shv[g] >> 2) & 3) - 1), sz = (R)(((shv[g] >> 4) & 3) - 1);
            ox = (ox + sx * E.hi[0]) + sx * E.lo[0];
            oy = (oy + sy * E.hi[1]) + sy * E.lo[1];
            oz = (oz + sz * E.hi[2]) + sz * E.lo[2];
          }
          xj[g].x += ox;
          xj[g].y += oy;
          xj[g].z += oz;
        }
        tyj0[g] = ts[sj];
        mbv[g] = mov[g] >= 0 ? (unsigned)masks[(size_t)mov[g] * kWave + lane] : 0xffu;
        if (ti == 0) s_j[wv][g * kClusterSize + tj] = JAtom<R>{xj[g].w, tyj0[g] < 0 ? 0 : tyj0[g]};
      }
      // every force slot of the group starts at zero (phase B fills the
      // members' slots, phase C reads the slots of the entries' masks)
#pragma unroll
      for (int k = 0; k < kGroup * 8 / 4; ++k) {
        R4 z;
        z.x = z.y = z.z = z.w = (R)0;
        reinterpret_cast<R4 *>(fr_slot)[k * kWave + lane] = z;
      }
      __builtin_amdgcn_wave_barrier();

      // ---- phase A: membership, queue, full batches
      R qel = 0, qec = 0;
      unsigned used = 0;  // bit g * 8 + c: the cluster pair has a member
      unsigned kept = 0;  // bit g * 8 + c: (prune pass) a candidate pair within the list radius
      // Candidate scan of entry g, specialised on one cutoff for both kinds
      // (SAME) and on the entry carrying an exclusion mask (MASKED), so that
      // each membership flag is a single compare whose ballot is its lane
      // mask.  An excluded candidate's r^2 becomes kFarR2, which fails every
      // range and band test.  The flags are those of the plain form: with one
      // cutoff, "outside the band and within range" is d < -guard.
      auto scan = [&](auto same, auto masked, const int g) {
        constexpr bool SAME = decltype(same)::value, MASKED = decltype(masked)::value;
        const int sj = cjv[g] * kClusterSize + tj;
        unsigned mrem = emv[g];
        while (mrem) {
          const int c = __builtin_ctz(mrem);
          mrem &= mrem - 1;
          const R4 xi = si[c * 8 + ti];
          R dx, dy, dz;
          pair_delta<R, R4>(xi, xj[g], E, shv[g] & kShiftPerPair, dx, dy, dz);
          R r2 = dx * dx + dy * dy + dz * dz;
          if (MASKED && !(mbv[g] & (1u << c))) r2 = kFarR2<R>;
          // a pair within the guard band of a cutoff is left to the FP64
          // boundary pass; a member is in range of a cutoff and outside
          // every band
          bool border, in;
          if (SAME) {
            const R d = r2 - rmax2;
            border = fabs(d) <= guard;
            in = d < -guard;
          } else {
            border = fmin(fabs(r2 - sl.rc2), fabs(r2 - sc.rc2)) <= guard;
            in = (border ? kFarR2<R> : r2) <= rmax2;
          }
          const unsigned long long bm = __builtin_amdgcn_ballot_w64(in);
          if (PRUNE && __builtin_amdgcn_ballot_w64(r2 <= prune_r2)) kept |= 1u << (g * 8 + c);
          if (__builtin_amdgcn_ballot_w64(border) && border) {
            const int b = atomicAdd(out.nborder, 1);
            if (b < out.maxborder)
              out.border[b] = make_int2(S * kSuperSize + c * 8 + ti, sj);
            else
              atomicExch(out.overflow, 1);
          }
          if (STATS && in) {
            const int ai = perm[S * kSuperSize + c * 8 + ti], aj = perm[sj];
            ++npair;
            hsum += pair_hash(ai < aj ? ai : aj, ai < aj ? aj : ai);
          }
          if (bm) {
            used |= 1u << (g * 8 + c);
            if (in) {
              // queue position: qtail + members in lower lanes
              const unsigned pos = __builtin_amdgcn_mbcnt_hi((unsigned)(bm >> 32),
                                                             __builtin_amdgcn_mbcnt_lo((unsigned)bm, (unsigned)qtail));
              const int modified = MASKED ? (int)((mbv[g] >> (8 + c)) & 1u) : 0;
              PairQ<R> e;
              e.r2 = r2;
              e.info = info_lane | (((g * 8 + c) << 6) | (c << 14) | (g << 20)) | (modified << kSlotBits);
              q[pos & (kQueue - 1)] = e;
            }
            qtail += __popcll(bm);
            while (qtail - qhead >= kWave) eval_batch(same, kWave, qel, qec);
          }
        }
      };
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        if (same_cut) {
          if (mov[g] >= 0)
            scan(std::true_type{}, std::true_type{}, g);
          else
            scan(std::true_type{}, std::false_type{}, g);
        } else {
          if (mov[g] >= 0)
            scan(std::false_type{}, std::true_type{}, g);
          else
            scan(std::false_type{}, std::false_type{}, g);
        }
      }
      // prune pass: cluster pairs without a candidate pair inside the list
      // radius cannot have a member before the next list build (every atom
      // stays within half the skin of its build position); later passes skip
      // them.  They have no member now either, so no result changes.
      if (PRUNE) {
#pragma unroll
        for (int g = 0; g < kGroup; ++g) {
          const int km = (kept >> (g * 8)) & 0xff;
          if (g < ng && lane == 0 && km != emv[g]) entries_w[(size_t)S * cap + eg + g].y = km;
        }
      }
      // ---- phase B: the rest of the queue
      if (qtail > qhead) {
        if (same_cut)
          eval_batch(std::true_type{}, qtail - qhead, qel, qec);
        else
          eval_batch(std::false_type{}, qtail - qhead, qel, qec);
      }
      elT += qel;
      ecT += qec;
      __builtin_amdgcn_wave_barrier();

      // ---- phase C: forces from the slots
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        if (g >= ng) break;
        // cluster pairs without members add nothing (their slots are zero)
        const int em = (used >> (g * 8)) & 0xff;
        if (em == 0) continue;
        const int sj = cjv[g] * kClusterSize + tj;
        R fj0 = 0, fj1 = 0, fj2 = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          if (!(em & (1 << c))) continue;
          const R4 xi = si[c * 8 + ti];
          R dx, dy, dz;
          pair_delta<R, R4>(xi, xj[g], E, shv[g] & kShiftPerPair, dx, dy, dz);
          const R f = fr_slot[(g * 8 + c) * kWave + lane];
          const R fx = dx * f, fy = dy * f, fz = dz * f;
          fi[c][0] -= fx;
          fi[c][1] -= fy;
          fi[c][2] -= fz;
          fj0 += fx;
          fj1 += fy;
          fj2 += fz;
        }
        fj0 = group8_sum(fj0);
        fj1 = group8_sum(fj1);
        fj2 = group8_sum(fj2);
        if (ti < 3 && tyj0[g] >= 0) {
          const R v = ti == 0 ? fj0 : (ti == 1 ? fj1 : fj2);
          Acc::add(&facc[3 * sj + ti], v);
        }
      }
      __builtin_amdgcn_wave_barrier();
    }
    flush();
    const AT eljA = wave_sum(WA::of(elT));
    const AT ecoA = wave_sum(WA::of(ecT));
    long long np = 0;
    if (STATS) {
      np = wave_sum((long long)npair);
      hsum = wave_sum(hsum);
    }
    if (lane == 0) {
      out.energy[2 * t] = WA::energy_bits(ecoA);
      out.energy[2 * t + 1] = WA::energy_bits(eljA);
      if (STATS) {
        out.npairs[t] = np;
        out.hash[t] = hsum;
      }
    }
    __builtin_amdgcn_wave_barrier();
    ++t;
  }
}

template <bool MIXED, int COUL, int LJSW, int CSW, bool BIGLJ, int MODE>
__global__ void __launch_bounds__(kPairBlock, 4) k_pair(
    PairParams pp, int dynamic, const int4 *__restrict__ tiles, const int *__restrict__ ntiles_p,
    const int *__restrict__ wstart, int *__restrict__ tile_counter, const int4 *__restrict__ entries, int4 *entries_w,
    int cap,
    const unsigned short *__restrict__ masks,
    const typename Real4<typename std::conditional<MIXED, float, double>::type>::type *__restrict__ xs,
    const float4 *__restrict__ sorg, const int *__restrict__ ts, const int *__restrict__ perm,
    const double *__restrict__ ljtab,
    const typename Real4<typename std::conditional<MIXED, float, double>::type>::type *__restrict__ ljt,
    typename AccT<MIXED>::type *__restrict__ facc, PairOut out) {
  // the blocks of one XCD take a contiguous range of runs
  pair_body<MIXED, COUL, LJSW, CSW, BIGLJ, MODE>(xcd_block(blockIdx.x, gridDim.x), gridDim.x, pp, dynamic, tiles,
                                                 ntiles_p, wstart, tile_counter, entries, entries_w, cap, masks, xs,
                                                 sorg, ts, perm, ljtab, ljt, facc, out);
}

// Boundary pairs: exact FP64 membership on the original coordinates
// (reference operation order), then the same pair evaluation.  Energies go
// to int64 fixed point (2^-32) so the total is independent of record order.
// Blocks [0, nblocks) of the fixup launch walk the records with a grid stride.
template <bool MIXED, int COUL, int LJSW, int CSW>
__device__ void border_pairs(int blk, int nblocks, const PairParams &pp, int stats, const int2 *__restrict__ border,
                             const int *__restrict__ nborder, int maxborder,
                             const typename Real4<typename std::conditional<MIXED, float, double>::type>::type *__restrict__ xs,
                             const float4 *__restrict__ sorg, const int *__restrict__ ts, const int *__restrict__ perm,
                             const double *__restrict__ xorig,
                             const double *__restrict__ ljtab, const int *__restrict__ eptr,
                             const int2 *__restrict__ elist, typename AccT<MIXED>::type *__restrict__ facc,
                             unsigned long long *__restrict__ eacc, unsigned long long *__restrict__ stat) {
  using R = typename std::conditional<MIXED, float, double>::type;
  const int nb = min(*nborder, maxborder);
  for (int b = blk * blockDim.x + threadIdx.x; b < nb; b += nblocks * blockDim.x) {
    const int2 rec = border[b];
    const int si = rec.x, sj = rec.y;
    const int ai = perm[si], aj = perm[sj];
    const int lo = ai < aj ? ai : aj, hi = ai < aj ? aj : ai;
    const double d2 = exact_d2(xorig, lo, hi, pp.box);
    const bool in_lj = LJSW != SW_NONE && d2 <= pp.lj.rc2;
    const bool in_c = COUL != COUL_NONE && d2 <= pp.co.rc2;
    if (!(in_lj || in_c)) continue;
    // exclusion class (recorded pairs already passed the FULL mask)
    const int cls = excl_class(eptr, elist, ai, aj);
    const bool modified = cls == 2;
    SwR<R> sl, sc;
    sl.load(pp.lj);
    sc.load(pp.co);
    const auto xi = xs[si], xj = xs[sj];
    const float4 oi = sorg[si / kSuperSize], oj = sorg[sj / kSuperSize];
    EdgesR<R> E;
    E.load(pp);
    R dx = (xj.x - xi.x) + (R)(oj.x - oi.x), dy = (xj.y - xi.y) + (R)(oj.y - oi.y),
      dz = (xj.z - xi.z) + (R)(oj.z - oi.z);
    dx = E.image(dx, 0);
    dy = E.image(dy, 1);
    dz = E.image(dz, 2);
    R r2 = dx * dx + dy * dy + dz * dz;
    // evaluate at the cutoff side the exact test chose
    R A = 0, B = 0;
    if (LJSW != SW_NONE) {
      const int k = 4 * (ts[si] * pp.ntypes + ts[sj]);
      A = (R)ljtab[k + (modified ? 2 : 0)];
      B = (R)ljtab[k + (modified ? 3 : 1)];
    }
    R qq = xi.w * xj.w;
    if (modified) qq *= (R)pp.csf;
    R e1, e2, fr;
    pair_eval<COUL, LJSW, CSW, R>(pp, sl, sc, (R)pp.alpha, (R)pp.a2pi, (R)pp.csf, r2, qq, A, B, in_lj, in_c, e1, e2, fr);
    AccT<MIXED>::add(&facc[3 * si], -dx * fr);
    AccT<MIXED>::add(&facc[3 * si + 1], -dy * fr);
    AccT<MIXED>::add(&facc[3 * si + 2], -dz * fr);
    AccT<MIXED>::add(&facc[3 * sj], dx * fr);
    AccT<MIXED>::add(&facc[3 * sj + 1], dy * fr);
    AccT<MIXED>::add(&facc[3 * sj + 2], dz * fr);
    atomicAdd(&eacc[0], (unsigned long long)__double2ll_rn((double)e2 * kForceScale));
    atomicAdd(&eacc[1], (unsigned long long)__double2ll_rn((double)e1 * kForceScale));
    if (stats) {
      atomicAdd(&stat[0], 1ull);
      atomicAdd(&stat[1], pair_hash(lo, hi));
    }
  }
}

// ---------------------------------------------------------------------------
// 4. PME: B-spline spreading, LDS FFT passes with the influence function,
//    force gather.  Grid [nx][ny][nz], z fastest (Grid.h:70).

struct PmeParams {
  int nx, ny, nz, nzh, order;
  double hxr, hyr, hzr;   // K/L per axis (Grid.h:520-522)
  int offx, offy, offz;   // -(p-1)/2 + K (Grid.h:509-511)
};

// BSpline::set (mdl/interface/force/nonbonded/BSpline.cpp:55-105) for a
// compile-time order P (2..8): weights th[0..P-1] and derivatives dth.
template <int P, class R>
__device__ inline void bspline_w(R w, R *th, R *dth) {
  th[P - 1] = 0;
  th[1] = w;
  th[0] = (R)1 - w;
#pragma unroll
  for (int k = 2; k < P - 1; k++) {
    const R div = (R)1 / (R)k;
    th[k] = div * w * th[k - 1];
#pragma unroll
    for (int j = 1; j < k; j++) th[k - j] = div * ((w + (R)j) * th[k - j - 1] + ((R)(k + 1 - j) - w) * th[k - j]);
    th[0] = div * ((R)1 - w) * th[0];
  }
  dth[0] = -th[0];
#pragma unroll
  for (int j = 1; j <= P - 1; j++) dth[j] = th[j - 1] - th[j];
  const R div = (R)1 / (R)(P - 1);
  th[P - 1] = div * w * th[P - 2];
#pragma unroll
  for (int j = 1; j < P - 1; j++)
    th[P - 1 - j] = div * ((w + (R)j) * th[P - j - 2] + ((R)(P - j) - w) * th[P - 1 - j]);
  th[0] = div * ((R)1 - w) * th[0];
}

// Element m of a register array as a sum of 0/1-weighted terms.  A select
// chain would be folded into a dynamically indexed load, which moves the
// array out of registers into LDS or scratch.
template <int P, class R>
__device__ inline R pick(const R *v, int m) {
  R r = (R)0;
#pragma unroll
  for (int t = 0; t < P; ++t) r += (R)(m == t) * v[t];
  return r;
}

// Threads per atom in spreading and gathering: one per (i, j) stencil row,
// rounded up to a power of two (order 2: 4, order 4: 16, orders 6 and 8: 64).
__host__ __device__ constexpr int pme_tpa(int order) {
  return order * order <= 4 ? 4 : order * order <= 16 ? 16 : 64;
}

// Grid index (base + offset) mod K for base + offset in [0, 3K).
__device__ inline int wrap_grid(int v, int K) {
  v -= v >= K ? K : 0;
  v -= v >= K ? K : 0;
  return v;
}

// Grid coordinate of one axis: u = x K/L wrapped to [0, K) (Grid.h:111-120).
__device__ inline void grid_coord(double x, double hr, int K, int off, int &base, double &w) {
  double u = x * hr;
  while (u < 0.0) u += K;
  while (u >= K) u -= K;
  int iu = (int)u;
  base = iu + off;
  w = u - iu;
}

// Charge spreading through an LDS patch (Grid::anterpolateCharge,
// Grid.h:107-154).  One block takes kSpreadSlots consecutive sorted slots
// (half a supercluster: a compact group of atoms).  Wave 0 computes each
// atom's grid base and fractional offsets and the box of grid points the
// group's stencils cover (bases taken relative to the first atom's, on the
// periodic grid); the block accumulates every stencil contribution into an
// LDS copy of that box with LDS atomics, then adds the box's nonzero points
// to the grid with one global atomic each.  The accumulator type is the
// grid's (int64 fixed point 2^-40 in the mixed form, so the result does not
// depend on the order of the adds; double in the FP64 form).  A group whose
// box exceeds the patch capacity, or wraps onto itself, adds to the grid
// directly.
constexpr int kSpreadSlots = 32;
template <int P>
__host__ __device__ constexpr int spread_patch_max() {
  return P <= 4 ? 2048 : 4096;
}

// The spread of one group (the block's kSpreadSlots slots from s0): lane t
// of wave 0 brings atom a (-1: none) of slot s0 + t at position xa; the
// other threads' a and xa are not read.
template <bool MIXED, int P>
__device__ void spread_patch_group(int s0, const PmeParams &p, int a, const double xa[3],
                                   const double *__restrict__ q, double *__restrict__ grid64,
                                   unsigned long long *__restrict__ gridfx) {
  using R = typename std::conditional<MIXED, float, double>::type;
  using G = typename std::conditional<MIXED, unsigned long long, double>::type;
  constexpr int P3 = P * P * P;
  constexpr int PMAX = spread_patch_max<P>();
  __shared__ G s_patch[PMAX];
  __shared__ int s_rb[kSpreadSlots][3];  // base relative to the patch origin
  __shared__ R s_w[kSpreadSlots][3];
  __shared__ R s_q[kSpreadSlots];
  __shared__ int s_org[3], s_ext[3], s_direct, s_any;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int K[3] = {p.nx, p.ny, p.nz};
  if (wv == 0) {
    const int t = lane;
    int b[3] = {0, 0, 0};
    double w[3] = {0, 0, 0};
    if (a >= 0) {
      grid_coord(xa[0], p.hxr, p.nx, p.offx, b[0], w[0]);
      grid_coord(xa[1], p.hyr, p.ny, p.offy, b[1], w[1]);
      grid_coord(xa[2], p.hzr, p.nz, p.offz, b[2], w[2]);
    }
    const unsigned long long ok = __ballot(a >= 0);
    if (ok) {
      const int first = __builtin_ctzll(ok);
      int lo[3], hi[3], db[3], ref[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        ref[k] = __shfl(b[k], first);
        int d = b[k] - ref[k];  // on the periodic grid, nearest to the reference
        d -= K[k] * (int)floorf(((float)d + 0.5f * (float)K[k]) / (float)K[k]);
        db[k] = d;
        lo[k] = a >= 0 ? d : 0x3fffffff;
        hi[k] = a >= 0 ? d : -0x3fffffff;
      }
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          lo[k] = min(lo[k], __shfl_xor(lo[k], m));
          hi[k] = max(hi[k], __shfl_xor(hi[k], m));
        }
      const int ex = hi[0] - lo[0] + P, ey = hi[1] - lo[1] + P, ez = hi[2] - lo[2] + P;
      const bool direct = ex > K[0] || ey > K[1] || ez > K[2] || (long long)ex * ey * ez > PMAX;
      if (t < kSpreadSlots) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          // direct adds use the plain base; patch adds the base within the box
          s_rb[t][k] = direct ? b[k] : db[k] - lo[k];
          s_w[t][k] = (R)w[k];
        }
        s_q[t] = a >= 0 ? (R)q[a] : (R)0;
      }
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) s_org[k] = ref[k] + lo[k];
        s_ext[0] = ex;
        s_ext[1] = ey;
        s_ext[2] = ez;
        s_direct = direct ? 1 : 0;
        s_any = 1;
      }
    } else if (lane == 0) {
      s_any = 0;
    }
  }
  __syncthreads();
  if (!s_any) return;  // block-uniform: no atoms in this group
  const bool direct = s_direct;
  const int ex = s_ext[0], ey = s_ext[1], ez = s_ext[2];
  const int vol = ex * ey * ez;
  if (!direct)
    for (int i = tid; i < vol; i += blockDim.x) s_patch[i] = (G)0;
  __syncthreads();
  for (int t = wv; t < kSpreadSlots; t += blockDim.x / kWave) {
    const R qa = s_q[t];
    if (qa == (R)0) continue;  // padding slot or uncharged atom: no contribution
    R tx[P], ty[P], tz[P], dd[P];
    bspline_w<P, R>(s_w[t][0], tx, dd);
    bspline_w<P, R>(s_w[t][1], ty, dd);
    bspline_w<P, R>(s_w[t][2], tz, dd);
    const int rb0 = s_rb[t][0], rb1 = s_rb[t][1], rb2 = s_rb[t][2];
#pragma unroll
    for (int pt0 = 0; pt0 < P3; pt0 += kWave) {
      const int pt = pt0 + lane;
      if (P3 % kWave != 0 && pt >= P3) break;
      const int i = pt / (P * P), j = (pt / P) % P, kk = pt % P;
      const R v = ((qa * pick<P>(tx, i)) * pick<P>(ty, j)) * pick<P>(tz, kk);
      G add;
      if (MIXED)
        add = (G)(unsigned long long)__double2ll_rn((double)v * kGridScale);
      else
        add = (G)v;
      if (!direct) {
        atomicAdd(&s_patch[((rb0 + i) * ey + (rb1 + j)) * ez + (rb2 + kk)], add);
      } else {
        const size_t g = ((size_t)wrap_grid(i + rb0, p.nx) * p.ny + wrap_grid(j + rb1, p.ny)) * p.nz +
                         wrap_grid(kk + rb2, p.nz);
        if (MIXED)
          atomicAdd(&gridfx[g], (unsigned long long)add);
        else
          unsafeAtomicAdd(&grid64[g], (double)add);
      }
    }
  }
  if (direct) return;
  __syncthreads();
  const int ox = s_org[0], oy = s_org[1], oz = s_org[2];
  for (int i = tid; i < vol; i += blockDim.x) {
    const G v = s_patch[i];
    if (v == (G)0) continue;
    const int iz = i % ez, iy = (i / ez) % ey, ix = i / (ez * ey);
    const int gx = (ox + ix) % p.nx, gy = (oy + iy) % p.ny, gz = (oz + iz) % p.nz;
    const size_t g = ((size_t)(gx < 0 ? gx + p.nx : gx) * p.ny + (gy < 0 ? gy + p.ny : gy)) * p.nz +
                     (gz < 0 ? gz + p.nz : gz);
    if (MIXED)
      atomicAdd(&gridfx[g], (unsigned long long)v);
    else
      unsafeAtomicAdd(&grid64[g], (double)v);
  }
}

template <bool MIXED, int P>
__device__ __forceinline__ void spread_patch_body(const int bid_, const int nbk_, int n, PmeParams p,
                                                  const double *__restrict__ xorig, const int *__restrict__ perm,
                                                  const double *__restrict__ q, double *__restrict__ grid64,
                                                  unsigned long long *__restrict__ gridfx) {
  const int s0 = bid_ * kSpreadSlots;
  const int t = threadIdx.x, s = s0 + t;
  int a = -1;
  double xa[3] = {0, 0, 0};
  if (t < kSpreadSlots && s < n) {
    a = perm[s];
    if (a >= 0) {
      xa[0] = xorig[3 * a];
      xa[1] = xorig[3 * a + 1];
      xa[2] = xorig[3 * a + 2];
    }
  }
  spread_patch_group<MIXED, P>(s0, p, a, xa, q, grid64, gridfx);
}

template <bool MIXED, int P>
__global__ void __launch_bounds__(256) k_spread_patch(int n, PmeParams p, const double *__restrict__ xorig,
                                                      const int *__restrict__ perm, const double *__restrict__ q,
                                                      double *__restrict__ grid64,
                                                      unsigned long long *__restrict__ gridfx) {
  // the blocks of one XCD take a contiguous range of slots (one region of
  // the grid)
  spread_patch_body<MIXED, P>(xcd_block(blockIdx.x, gridDim.x), gridDim.x, n, p, xorig, perm, q, grid64, gridfx);
}

// Charge spreading (Grid::anterpolateCharge, Grid.h:107-154).  The P^3
// stencil points of one atom go to consecutive lanes (z fastest), so each
// atomic instruction of a wave updates P^2 short contiguous rows of one
// atom's stencil; orders 2 (8 points) pack 8 atoms into a wave, orders 6 and
// 8 loop over their points in steps of 64.  Slots run in sorted order; n
// counts slots (padding slots return at once).
template <bool MIXED, int P>
__global__ void __launch_bounds__(256) k_spread(int n, PmeParams p, const double *__restrict__ xorig,
                                                const int *__restrict__ perm, const double *__restrict__ q,
                                                double *__restrict__ grid64, unsigned long long *__restrict__ gridfx) {
  using R = typename std::conditional<MIXED, float, double>::type;
  constexpr int P3 = P * P * P;
  constexpr int LPA = P3 >= kWave ? kWave : P3;  // lanes per atom
  constexpr int APW = kWave / LPA;               // atoms per wave
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = gt >> 6, lane = gt & 63;
  const int s = wave * APW + lane / LPA;
  if (s >= n) return;
  const int a = perm[s];
  if (a < 0) return;  // padding slot
  const int l0 = lane % LPA;
  int b[3];
  double w[3];
  grid_coord(xorig[3 * a], p.hxr, p.nx, p.offx, b[0], w[0]);
  grid_coord(xorig[3 * a + 1], p.hyr, p.ny, p.offy, b[1], w[1]);
  grid_coord(xorig[3 * a + 2], p.hzr, p.nz, p.offz, b[2], w[2]);
  R tx[P], ty[P], tz[P], d[P];
  bspline_w<P, R>((R)w[0], tx, d);
  bspline_w<P, R>((R)w[1], ty, d);
  bspline_w<P, R>((R)w[2], tz, d);
  const R qa = (R)q[a];
#pragma unroll
  for (int pt0 = 0; pt0 < P3; pt0 += LPA) {
    const int pt = pt0 + l0;
    const int i = pt / (P * P), j = (pt / P) % P, kk = pt % P;
    const R v = ((qa * pick<P>(tx, i)) * pick<P>(ty, j)) * pick<P>(tz, kk);
    const size_t g = ((size_t)wrap_grid(i + b[0], p.nx) * p.ny + wrap_grid(j + b[1], p.ny)) * p.nz +
                     wrap_grid(kk + b[2], p.nz);
    if (MIXED) {
      long long fx = __double2ll_rn((double)v * kGridScale);
      atomicAdd(&gridfx[g], (unsigned long long)fx);
    } else {
      unsafeAtomicAdd(&grid64[g], (double)v);
    }
  }
}

template <class T>
struct Cx {
  T x, y;
};

template <class T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return Cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// Forward DFT of small radix R in registers, e^{-2 pi i rk/R}.
template <int R, class T>
__device__ inline void dft_small(Cx<T> *v) {
  if constexpr (R == 2) {
    Cx<T> a = v[0], b = v[1];
    v[0] = Cx<T>{a.x + b.x, a.y + b.y};
    v[1] = Cx<T>{a.x - b.x, a.y - b.y};
  } else if constexpr (R == 4) {
    Cx<T> a = v[0], b = v[1], c = v[2], d = v[3];
    Cx<T> s0{a.x + c.x, a.y + c.y}, s1{a.x - c.x, a.y - c.y};
    Cx<T> s2{b.x + d.x, b.y + d.y}, s3{b.x - d.x, b.y - d.y};
    v[0] = Cx<T>{s0.x + s2.x, s0.y + s2.y};
    v[2] = Cx<T>{s0.x - s2.x, s0.y - s2.y};
    // -i * s3 = (s3.y, -s3.x)
    v[1] = Cx<T>{s1.x + s3.y, s1.y - s3.x};
    v[3] = Cx<T>{s1.x - s3.y, s1.y + s3.x};
  } else if constexpr (R == 3) {
    const T c = (T)-0.5, s = (T)0.86602540378443864676;
    Cx<T> a = v[0], b = v[1], d = v[2];
    Cx<T> t{b.x + d.x, b.y + d.y}, u{b.x - d.x, b.y - d.y};
    v[0] = Cx<T>{a.x + t.x, a.y + t.y};
    Cx<T> m{a.x + c * t.x, a.y + c * t.y};
    // -i s u = (s u.y, -s u.x)
    v[1] = Cx<T>{m.x + s * u.y, m.y - s * u.x};
    v[2] = Cx<T>{m.x - s * u.y, m.y + s * u.x};
  } else {
    static_assert(R == 5, "radix 2, 3, 4 or 5");
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    Cx<T> a = v[0];
    Cx<T> t1{v[1].x + v[4].x, v[1].y + v[4].y}, u1{v[1].x - v[4].x, v[1].y - v[4].y};
    Cx<T> t2{v[2].x + v[3].x, v[2].y + v[3].y}, u2{v[2].x - v[3].x, v[2].y - v[3].y};
    v[0] = Cx<T>{a.x + t1.x + t2.x, a.y + t1.y + t2.y};
    Cx<T> m1{a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
    Cx<T> m2{a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
    // y1 = m1 - i(s1 u1 + s2 u2); y4 = m1 + i(...); y2 = m2 - i(s2 u1 - s1 u2); y3 = m2 + i(...)
    Cx<T> n1{s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y};
    Cx<T> n2{s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y};
    v[1] = Cx<T>{m1.x + n1.y, m1.y - n1.x};
    v[4] = Cx<T>{m1.x - n1.y, m1.y + n1.x};
    v[2] = Cx<T>{m2.x + n2.y, m2.y - n2.x};
    v[3] = Cx<T>{m2.x - n2.y, m2.y + n2.x};
  }
}

struct FftPlan {
  int n;
  int nrad;
  int rad[16];
};

// One Stockham pass of radix R over `nb` lines of length n (src -> dst):
// butterfly j reads src[j + r n/R], twiddles by tw[k r n/(Ns R)] (k = j mod
// Ns; the index stays below n), writes dst[(j - k) R + k + r Ns].
template <int R, class T>
__device__ inline void fft_pass(const Cx<T> *__restrict__ src, Cx<T> *__restrict__ dst, int nb, int n, int Ns,
                                const Cx<T> *__restrict__ tw) {
  const int nbut = n / R;
  const int tstep = n / (Ns * R);
  for (int t = threadIdx.x; t < nbut * nb; t += blockDim.x) {
    const int line = t / nbut, j = t - line * nbut;
    const int k = j % Ns;
    const Cx<T> *s = src + line * n;
    Cx<T> *d = dst + line * n;
    Cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[r] = s[j + r * nbut];
      if (r > 0 && k > 0) v[r] = cmul(v[r], tw[k * r * tstep]);
    }
    dft_small<R, T>(v);
    const int idxD = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[idxD + r * Ns] = v[r];
  }
}

// Batched forward Stockham FFT of `nb` lines of length plan.n in LDS
// (a -> b ping-pong); twiddle table tw[m] = e^{-2 pi i m/n}.  Returns the
// buffer holding the result.  All threads of the block participate.
template <class T>
__device__ Cx<T> *lds_fft(Cx<T> *a, Cx<T> *b, int nb, const FftPlan &pl, const Cx<T> *__restrict__ tw) {
  const int n = pl.n;
  int Ns = 1;
  for (int s = 0; s < pl.nrad; ++s) {
    const int R = pl.rad[s];
    if (R == 4)
      fft_pass<4, T>(a, b, nb, n, Ns, tw);
    else if (R == 2)
      fft_pass<2, T>(a, b, nb, n, Ns, tw);
    else if (R == 3)
      fft_pass<3, T>(a, b, nb, n, Ns, tw);
    else
      fft_pass<5, T>(a, b, nb, n, Ns, tw);
    __syncthreads();
    Cx<T> *tmp = a;
    a = b;
    b = tmp;
    Ns *= R;
  }
  return a;
}

// The 3-D FFT passes move the data between three layouts so that every pass
// reads and writes whole lines: the charge grid and the real potential are
// [i][j][k] (k fastest); the half spectrum is SA = [i][k][j] after the z pass
// and SB = [k][j][i] after the y pass.  A block takes a few consecutive lines
// (TY, TI or TJ), so the transposing reads and writes move segments of that
// many consecutive elements.

// z pass, real -> half complex, for the TY lines (i, j0 .. j0 + TY - 1):
// the grid points are read once and cleared for the next spread; out: SA.
template <class T, bool FIXED>
__device__ __forceinline__ void fft_z_r2c_body(const int bid_, const int nbk_, const PmeParams &p, const FftPlan &pl,
                                               int TY, double *__restrict__ g64, unsigned long long *__restrict__ gfx,
                                               const Cx<T> *__restrict__ tw, Cx<T> *__restrict__ sa) {
  extern __shared__ unsigned char smem_raw[];
  Cx<T> *a = reinterpret_cast<Cx<T> *>(smem_raw);
  Cx<T> *b = a + TY * p.nz;
  const int jt = (p.ny + TY - 1) / TY;
  const int i = bid_ / jt, j0 = (bid_ - i * jt) * TY;
  const int nb = min(TY, p.ny - j0);
  const size_t base = ((size_t)i * p.ny + j0) * p.nz;
  for (int t = threadIdx.x; t < nb * p.nz; t += blockDim.x) {
    T v;
    if (FIXED) {
      v = (T)((double)(long long)gfx[base + t] * (1.0 / kGridScale));
      gfx[base + t] = 0ull;
    } else {
      v = (T)g64[base + t];
      g64[base + t] = 0.0;
    }
    a[t] = Cx<T>{v, (T)0};
  }
  __syncthreads();
  const Cx<T> *r = lds_fft<T>(a, b, nb, pl, tw);
  for (int t = threadIdx.x; t < nb * p.nzh; t += blockDim.x) {
    const int k = t / nb, jj = t - k * nb;
    sa[((size_t)i * p.nzh + k) * p.ny + j0 + jj] = r[jj * p.nz + k];
  }
}

template <class T, bool FIXED>
__global__ void k_fft_z_r2c(PmeParams p, FftPlan pl, int TY, double *__restrict__ g64,
                            unsigned long long *__restrict__ gfx, const Cx<T> *__restrict__ tw, Cx<T> *__restrict__ sa) {
  fft_z_r2c_body<T, FIXED>(blockIdx.x, gridDim.x, p, pl, TY, g64, gfx, tw, sa);
}

// y pass for the TI lines (i0 .. i0 + TI - 1, k): forward SA -> SB; inverse
// (conj-in / conj-out around the forward transform) SB -> SA.
template <class T>
__device__ __forceinline__ void fft_y_body(const int bid_, const int nbk_, const PmeParams &p, const FftPlan &pl,
                                           int TI, bool inverse, const Cx<T> *__restrict__ tw,
                                           const Cx<T> *__restrict__ src, Cx<T> *__restrict__ dst) {
  extern __shared__ unsigned char smem_raw[];
  Cx<T> *a = reinterpret_cast<Cx<T> *>(smem_raw);
  Cx<T> *b = a + TI * p.ny;
  const int it = (p.nx + TI - 1) / TI;
  const int k = bid_ / it, i0 = (bid_ - k * it) * TI;
  const int nb = min(TI, p.nx - i0);
  if (!inverse) {
    for (int t = threadIdx.x; t < nb * p.ny; t += blockDim.x) {
      const int ii = t / p.ny, j = t - ii * p.ny;
      a[t] = src[((size_t)(i0 + ii) * p.nzh + k) * p.ny + j];
    }
  } else {
    for (int t = threadIdx.x; t < nb * p.ny; t += blockDim.x) {
      const int j = t / nb, ii = t - j * nb;
      Cx<T> v = src[((size_t)k * p.ny + j) * p.nx + i0 + ii];
      v.y = -v.y;
      a[ii * p.ny + j] = v;
    }
  }
  __syncthreads();
  const Cx<T> *r = lds_fft<T>(a, b, nb, pl, tw);
  if (!inverse) {
    for (int t = threadIdx.x; t < nb * p.ny; t += blockDim.x) {
      const int j = t / nb, ii = t - j * nb;
      dst[((size_t)k * p.ny + j) * p.nx + i0 + ii] = r[ii * p.ny + j];
    }
  } else {
    for (int t = threadIdx.x; t < nb * p.ny; t += blockDim.x) {
      const int ii = t / p.ny, j = t - ii * p.ny;
      Cx<T> v = r[t];
      v.y = -v.y;
      dst[((size_t)(i0 + ii) * p.nzh + k) * p.ny + j] = v;
    }
  }
}

template <class T>
__global__ void k_fft_y(PmeParams p, FftPlan pl, int TI, bool inverse, const Cx<T> *__restrict__ tw,
                        const Cx<T> *__restrict__ src, Cx<T> *__restrict__ dst) {
  fft_y_body<T>(blockIdx.x, gridDim.x, p, pl, TI, inverse, tw, src, dst);
}

// x pass for the TJ lines (k, j0 .. j0 + TJ - 1) of SB, in place: forward
// FFT, influence function multiply with the energy (Grid.h:237-341; theta in
// the SB layout), inverse FFT.
template <class T>
__device__ __forceinline__ void fft_x_conv_body(const int bid_, const int nbk_, const PmeParams &p, const FftPlan &pl,
                                                int TJ, const Cx<T> *__restrict__ tw, const T *__restrict__ theta,
                                                Cx<T> *__restrict__ sb, double *__restrict__ epart) {
  extern __shared__ unsigned char smem_raw[];
  Cx<T> *a = reinterpret_cast<Cx<T> *>(smem_raw);
  Cx<T> *b = a + TJ * p.nx;
  __shared__ double s_e[1024 / kWave];
  const int jt = (p.ny + TJ - 1) / TJ;
  const int k = bid_ / jt, j0 = (bid_ - k * jt) * TJ;
  const int nb = min(TJ, p.ny - j0);
  const size_t base = ((size_t)k * p.ny + j0) * p.nx;
  for (int t = threadIdx.x; t < nb * p.nx; t += blockDim.x) a[t] = sb[base + t];
  __syncthreads();
  Cx<T> *r = lds_fft<T>(a, b, nb, pl, tw);
  const double wgt = (k > 0 && ((k != p.nz / 2) || (p.nz & 1))) ? 2.0 : 1.0;
  double e = 0.0;
  for (int t = threadIdx.x; t < nb * p.nx; t += blockDim.x) {
    const T th = theta[base + t];
    const Cx<T> v = r[t];
    e += (double)th * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
    // conj for the inverse transform (conj-in / conj-out)
    r[t] = Cx<T>{v.x * th, -(v.y * th)};
  }
  __syncthreads();
  Cx<T> *o = (r == a) ? b : a;
  const Cx<T> *r2 = lds_fft<T>(r, o, nb, pl, tw);
  for (int t = threadIdx.x; t < nb * p.nx; t += blockDim.x) {
    Cx<T> v = r2[t];
    v.y = -v.y;
    sb[base + t] = v;
  }
  e = wave_sum(wgt * e);
  if ((threadIdx.x & 63) == 0) s_e[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) tot += s_e[w];
    epart[bid_] = tot;
  }
}

template <class T>
__global__ void k_fft_x_conv(PmeParams p, FftPlan pl, int TJ, const Cx<T> *__restrict__ tw,
                             const T *__restrict__ theta, Cx<T> *__restrict__ sb, double *__restrict__ epart) {
  fft_x_conv_body<T>(blockIdx.x, gridDim.x, p, pl, TJ, tw, theta, sb, epart);
}

// Largest plane (nx ny points) that k_fft_plane_conv takes: two LDS copies
// of 1,024 points take 16 KiB (mixed) or 32 KiB (FP64).  On larger planes
// the three line launches, with many more blocks, are faster (measured on
// 45 x 48 and 60 x 60 planes).
constexpr int kFftPlanePoints = 1024;

// The y pass, the x pass with the convolution and the inverse y pass of one
// plane k of the half spectrum in one block (the work of k_fft_y, k_fft_x_conv
// and k_fft_y in one launch, for small planes): SA plane k
// is read as nx lines along y, transformed, transposed in LDS to ny lines
// along x, transformed, multiplied by the influence function (energy summed
// as in k_fft_x_conv), transformed back along x and, transposed, along y, and
// written back to SA.  The conj-out of the inverse x transform and the
// conj-in of the inverse y transform cancel, so only the last conj remains.
template <class T>
__device__ __forceinline__ void fft_plane_conv_body(const int bid_, const int nbk_, const PmeParams &p,
                                                    const FftPlan &ply, const FftPlan &plx,
                                                    const Cx<T> *__restrict__ twy, const Cx<T> *__restrict__ twx,
                                                    const T *__restrict__ theta, Cx<T> *__restrict__ sa,
                                                    double *__restrict__ epart) {
  extern __shared__ unsigned char smem_raw[];
  const int nx = p.nx, ny = p.ny, nxy = nx * ny;
  Cx<T> *const a = reinterpret_cast<Cx<T> *>(smem_raw);
  Cx<T> *const b = a + nxy;
  __shared__ double s_e[256 / kWave];
  const int k = bid_;
  for (int t = threadIdx.x; t < nxy; t += blockDim.x) {
    const int i = t / ny, j = t - i * ny;
    a[t] = sa[((size_t)i * p.nzh + k) * ny + j];
  }
  __syncthreads();
  // forward y: lines i, layout [i][j]
  Cx<T> *r = lds_fft<T>(a, b, nx, ply, twy);
  Cx<T> *o = (r == a) ? b : a;
  for (int t = threadIdx.x; t < nxy; t += blockDim.x) {
    const int j = t / nx, i = t - j * nx;
    o[t] = r[i * ny + j];
  }
  __syncthreads();
  // forward x: lines j, layout [j][i] (the SB order of theta)
  r = lds_fft<T>(o, r, ny, plx, twx);
  const double wgt = (k > 0 && ((k != p.nz / 2) || (p.nz & 1))) ? 2.0 : 1.0;
  const T *th_k = theta + (size_t)k * nxy;
  double e = 0.0;
  for (int t = threadIdx.x; t < nxy; t += blockDim.x) {
    const T th = th_k[t];
    const Cx<T> v = r[t];
    e += (double)th * ((double)v.x * (double)v.x + (double)v.y * (double)v.y);
    r[t] = Cx<T>{v.x * th, -(v.y * th)};  // conj-in of the inverse x transform
  }
  __syncthreads();
  o = (r == a) ? b : a;
  r = lds_fft<T>(r, o, ny, plx, twx);
  o = (r == a) ? b : a;
  for (int t = threadIdx.x; t < nxy; t += blockDim.x) {
    const int i = t / ny, j = t - i * ny;
    o[t] = r[j * nx + i];
  }
  __syncthreads();
  // inverse y (forward transform of the conjugate), conj-out into SA
  r = lds_fft<T>(o, r, nx, ply, twy);
  for (int t = threadIdx.x; t < nxy; t += blockDim.x) {
    const int i = t / ny, j = t - i * ny;
    Cx<T> v = r[t];
    v.y = -v.y;
    sa[((size_t)i * p.nzh + k) * ny + j] = v;
  }
  e = wave_sum(wgt * e);
  if ((threadIdx.x & 63) == 0) s_e[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) tot += s_e[w];
    epart[k] = tot;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_fft_plane_conv(PmeParams p, FftPlan ply, FftPlan plx,
                                                        const Cx<T> *__restrict__ twy, const Cx<T> *__restrict__ twx,
                                                        const T *__restrict__ theta, Cx<T> *__restrict__ sa,
                                                        double *__restrict__ epart) {
  fft_plane_conv_body<T>(blockIdx.x, gridDim.x, p, ply, plx, twy, twx, theta, sa, epart);
}

// z pass, half complex -> real for the TY lines (i, j0 .. j0 + TY - 1) of SA
// (Hermitian completion along z, inverse sign); out: the real potential.
template <class T>
__device__ __forceinline__ void fft_z_c2r_body(const int bid_, const int nbk_, const PmeParams &p, const FftPlan &pl,
                                               int TY, const Cx<T> *__restrict__ tw, const Cx<T> *__restrict__ sa,
                                               T *__restrict__ greal) {
  extern __shared__ unsigned char smem_raw[];
  Cx<T> *a = reinterpret_cast<Cx<T> *>(smem_raw);
  Cx<T> *b = a + TY * p.nz;
  const int jt = (p.ny + TY - 1) / TY;
  const int i = bid_ / jt, j0 = (bid_ - i * jt) * TY;
  const int nb = min(TY, p.ny - j0);
  for (int t = threadIdx.x; t < nb * p.nz; t += blockDim.x) {
    const int k = t / nb, jj = t - k * nb;
    Cx<T> v;
    if (k < p.nzh) {
      v = sa[((size_t)i * p.nzh + k) * p.ny + j0 + jj];
      v.y = -v.y;  // conj-in
    } else {
      v = sa[((size_t)i * p.nzh + (p.nz - k)) * p.ny + j0 + jj];  // conj(conj(X[nz-k]))
    }
    a[jj * p.nz + k] = v;
  }
  __syncthreads();
  const Cx<T> *r = lds_fft<T>(a, b, nb, pl, tw);
  const size_t base = ((size_t)i * p.ny + j0) * p.nz;
  for (int t = threadIdx.x; t < nb * p.nz; t += blockDim.x) greal[base + t] = r[t].x;
}

template <class T>
__global__ void k_fft_z_c2r(PmeParams p, FftPlan pl, int TY, const Cx<T> *__restrict__ tw,
                            const Cx<T> *__restrict__ sa, T *__restrict__ greal) {
  fft_z_c2r_body<T>(blockIdx.x, gridDim.x, p, pl, TY, tw, sa, greal);
}

// Force interpolation (Grid.h:159-189) into the sorted accumulators.  Lane
// groups as in k_spread: each lane sums one (i, j) stencil row, the group
// reduces with cross-lane adds and its first lane adds the atom's force.
template <bool MIXED, int P>
__global__ void __launch_bounds__(256) k_gather(int n, int waves, PmeParams p, const double *__restrict__ xorig,
                                                const int *__restrict__ perm, const double *__restrict__ q,
                                                const typename std::conditional<MIXED, float, double>::type *__restrict__ greal,
                                                typename AccT<MIXED>::type *__restrict__ facc) {
  using R = typename std::conditional<MIXED, float, double>::type;
  constexpr int TPA = pme_tpa(P);
  const int gt = blockIdx.x * blockDim.x + threadIdx.x;
  const int wave = gt >> 6, lane = gt & 63;
  if (wave >= waves) return;  // whole waves only: the group reduction below needs every lane
  const int r = lane % TPA;
  const int s = (lane / TPA) * waves + wave;
  R fx = 0, fy = 0, fz = 0;
  int a = -1;
  if (s < n) a = perm[s];  // n counts slots; padding slots have a < 0
  if (a >= 0 && r < P * P) {
    const int i = r / P, j = r - (r / P) * P;
    int b[3];
    double w[3];
    grid_coord(xorig[3 * a], p.hxr, p.nx, p.offx, b[0], w[0]);
    grid_coord(xorig[3 * a + 1], p.hyr, p.ny, p.offy, b[1], w[1]);
    grid_coord(xorig[3 * a + 2], p.hzr, p.nz, p.offz, b[2], w[2]);
    R tx[P], ty[P], tz[P], dx[P], dy[P], dz[P];
    bspline_w<P, R>((R)w[0], tx, dx);
    bspline_w<P, R>((R)w[1], ty, dy);
    bspline_w<P, R>((R)w[2], tz, dz);
    const R txi = pick<P>(tx, i), tyj = pick<P>(ty, j);
    const R xij = pick<P>(dx, i) * tyj, yij = txi * pick<P>(dy, j), zij = txi * tyj;
    const int i1 = wrap_grid(i + b[0], p.nx), j1 = wrap_grid(j + b[1], p.ny);
    const size_t row = ((size_t)i1 * p.ny + j1) * p.nz;
#pragma unroll
    for (int kk = 0; kk < P; kk++) {
      const int k1 = wrap_grid(kk + b[2], p.nz);
      const R term = greal[row + k1];
      fx -= term * xij * tz[kk];
      fy -= term * yij * tz[kk];
      fz -= term * zij * dz[kk];
    }
  }
#pragma unroll
  for (int m = 1; m < TPA; m <<= 1) {
    fx += __shfl_xor(fx, m);
    fy += __shfl_xor(fy, m);
    fz += __shfl_xor(fz, m);
  }
  if (r == 0 && a >= 0) {
    const R qa = (R)q[a];
    AccT<MIXED>::add(&facc[3 * s], fx * (R)p.hxr * qa);
    AccT<MIXED>::add(&facc[3 * s + 1], fy * (R)p.hyr * qa);
    AccT<MIXED>::add(&facc[3 * s + 2], fz * (R)p.hzr * qa);
  }
}

// Force interpolation through an LDS patch of the real potential: one block
// takes kSpreadSlots consecutive sorted slots (a compact group of atoms, as
// k_spread_patch); wave 0 finds their grid bases and the box of grid points
// their stencils cover; the block loads that box once, then 8 lanes per atom
// sum its P^2 stencil rows from LDS (row r on lane r mod 8), reduce across
// the 8 lanes and add the force.  A group whose box exceeds the patch reads
// the potential from the grid directly.  The grid is read about once per
// block instead of once per atom's stencil row.
template <bool MIXED, int P>
__device__ __forceinline__ void gather_patch_body(const int bid_, const int nbk_, int n, PmeParams p,
    const double *__restrict__ xorig, const int *__restrict__ perm, const double *__restrict__ q,
    const typename std::conditional<MIXED, float, double>::type *__restrict__ greal,
    typename AccT<MIXED>::type *__restrict__ facc) {
  using R = typename std::conditional<MIXED, float, double>::type;
  constexpr int PMAX = spread_patch_max<P>();
  constexpr int LPA = 8;  // lanes per atom
  static_assert(kSpreadSlots * LPA == 256, "one block of 256 threads per group");
  __shared__ R s_patch[PMAX];
  __shared__ int s_rb[kSpreadSlots][3];  // base relative to the patch origin (direct: the plain base)
  __shared__ R s_w[kSpreadSlots][3];
  __shared__ int s_org[3], s_ext[3], s_direct, s_any;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int s0 = bid_ * kSpreadSlots;
  const int K[3] = {p.nx, p.ny, p.nz};
  if (wv == 0) {
    const int t = lane;
    const int sl = s0 + t;
    const int a = (t < kSpreadSlots && sl < n) ? perm[sl] : -1;
    int b[3] = {0, 0, 0};
    double w[3] = {0, 0, 0};
    if (a >= 0) {
      grid_coord(xorig[3 * a], p.hxr, p.nx, p.offx, b[0], w[0]);
      grid_coord(xorig[3 * a + 1], p.hyr, p.ny, p.offy, b[1], w[1]);
      grid_coord(xorig[3 * a + 2], p.hzr, p.nz, p.offz, b[2], w[2]);
    }
    const unsigned long long ok = __ballot(a >= 0);
    if (ok) {
      const int first = __builtin_ctzll(ok);
      int lo[3], hi[3], db[3], ref[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        ref[k] = __shfl(b[k], first);
        int d = b[k] - ref[k];  // on the periodic grid, nearest to the reference
        d -= K[k] * (int)floorf(((float)d + 0.5f * (float)K[k]) / (float)K[k]);
        db[k] = d;
        lo[k] = a >= 0 ? d : 0x3fffffff;
        hi[k] = a >= 0 ? d : -0x3fffffff;
      }
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          lo[k] = min(lo[k], __shfl_xor(lo[k], m));
          hi[k] = max(hi[k], __shfl_xor(hi[k], m));
        }
      const int ex = hi[0] - lo[0] + P, ey = hi[1] - lo[1] + P, ez = hi[2] - lo[2] + P;
      const bool direct = ex > K[0] || ey > K[1] || ez > K[2] || (long long)ex * ey * ez > PMAX;
      if (t < kSpreadSlots) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          s_rb[t][k] = a < 0 ? -1 : (direct ? b[k] : db[k] - lo[k]);
          s_w[t][k] = (R)w[k];
        }
      }
      if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) s_org[k] = ref[k] + lo[k];
        s_ext[0] = ex;
        s_ext[1] = ey;
        s_ext[2] = ez;
        s_direct = direct ? 1 : 0;
        s_any = 1;
      }
    } else if (lane == 0) {
      s_any = 0;
    }
  }
  __syncthreads();
  if (!s_any) return;  // block-uniform: no atoms in this group
  const bool direct = s_direct;
  const int ex = s_ext[0], ey = s_ext[1], ez = s_ext[2];
  if (!direct) {
    const int vol = ex * ey * ez;
    const int o0 = s_org[0], o1 = s_org[1], o2 = s_org[2];
    for (int i = tid; i < vol; i += blockDim.x) {
      const int x = i / (ey * ez), rem = i - x * (ey * ez), y = rem / ez, z = rem - y * ez;
      s_patch[i] = greal[((size_t)wrap_grid(o0 + x, p.nx) * p.ny + wrap_grid(o1 + y, p.ny)) * p.nz +
                         wrap_grid(o2 + z, p.nz)];
    }
  }
  __syncthreads();
  const int t = tid / LPA, r0 = tid % LPA;  // atom t of the group, lane r0 of its 8
  R fx = 0, fy = 0, fz = 0;
  const int rb0 = s_rb[t][0], rb1 = s_rb[t][1], rb2 = s_rb[t][2];
  if (rb0 >= 0) {
    R tx[P], ty[P], tz[P], dx[P], dy[P], dz[P];
    bspline_w<P, R>(s_w[t][0], tx, dx);
    bspline_w<P, R>(s_w[t][1], ty, dy);
    bspline_w<P, R>(s_w[t][2], tz, dz);
    for (int rr = r0; rr < P * P; rr += LPA) {
      const int i = rr / P, j = rr - (rr / P) * P;
      const R txi = pick<P>(tx, i), tyj = pick<P>(ty, j);
      const R xij = pick<P>(dx, i) * tyj, yij = txi * pick<P>(dy, j), zij = txi * tyj;
      if (!direct) {
        const R *row = s_patch + ((rb0 + i) * ey + (rb1 + j)) * ez + rb2;
#pragma unroll
        for (int kk = 0; kk < P; kk++) {
          const R term = row[kk];
          fx -= term * xij * tz[kk];
          fy -= term * yij * tz[kk];
          fz -= term * zij * dz[kk];
        }
      } else {
        const int i1 = wrap_grid(i + rb0, p.nx), j1 = wrap_grid(j + rb1, p.ny);
        const size_t row = ((size_t)i1 * p.ny + j1) * p.nz;
#pragma unroll
        for (int kk = 0; kk < P; kk++) {
          const R term = greal[row + wrap_grid(kk + rb2, p.nz)];
          fx -= term * xij * tz[kk];
          fy -= term * yij * tz[kk];
          fz -= term * zij * dz[kk];
        }
      }
    }
  }
#pragma unroll
  for (int m = 1; m < LPA; m <<= 1) {
    fx += __shfl_xor(fx, m);
    fy += __shfl_xor(fy, m);
    fz += __shfl_xor(fz, m);
  }
  const int sl = s0 + t;
  if (r0 == 0 && rb0 >= 0) {
    const R qa = (R)q[perm[sl]];
    AccT<MIXED>::add(&facc[3 * sl], fx * (R)p.hxr * qa);
    AccT<MIXED>::add(&facc[3 * sl + 1], fy * (R)p.hyr * qa);
    AccT<MIXED>::add(&facc[3 * sl + 2], fz * (R)p.hzr * qa);
  }
}

template <bool MIXED, int P>
__global__ void __launch_bounds__(256) k_gather_patch(int n, PmeParams p, const double *__restrict__ xorig,
                                                      const int *__restrict__ perm, const double *__restrict__ q,
                                                      const typename std::conditional<MIXED, float, double>::type
                                                          *__restrict__ greal,
                                                      typename AccT<MIXED>::type *__restrict__ facc) {
  gather_patch_body<MIXED, P>(xcd_block(blockIdx.x, gridDim.x), gridDim.x, n, p, xorig, perm, q, greal, facc);
}

// ---------------------------------------------------------------------------
// 5. Exclusion correction (NonbondedPMEwaldSystemForce.h:874-911), FP64, exact erf.
// One block of 256 entries (block blk of the correction range).
template <bool MIXED>
__device__ void correction_block(int blk, int ne, const int4 *__restrict__ etab, const double *__restrict__ xorig,
                                 const double *__restrict__ q, const int *__restrict__ slot, Box box, double alpha,
                                 double csf, typename AccT<MIXED>::type *__restrict__ facc, double *__restrict__ epart) {
  __shared__ double s_e[kPairBlock / kWave];
  int e = blk * blockDim.x + threadIdx.x;
  double en = 0.0;
  if (e < ne) {
    const int4 t = etab[e];
    const int a1 = t.x, a2 = t.y;
    double d[3] = {xorig[3 * a2] - xorig[3 * a1], xorig[3 * a2 + 1] - xorig[3 * a1 + 1],
                   xorig[3 * a2 + 2] - xorig[3 * a1 + 2]};
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 > box.D) {
      for (int k = 0; k < 3; ++k) {
        if (d[k] < -box.H2[k])
          d[k] += box.H[k];
        else if (d[k] > box.H2[k])
          d[k] -= box.H[k];
      }
      r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 > box.D) {
        for (int k = 0; k < 3; ++k) d[k] -= box.L[k] * rint(box.Lr[k] * d[k]);
        r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      }
    }
    double qq = q[a1] * q[a2];
    if (t.z == 2) qq *= 1 - csf;
    const double r = sqrt(r2), rr = 1.0 / r;
    const double ee = erf(alpha * r) * rr;
    en = -qq * ee;
    const double fr = qq * (2.0 * alpha / sqrt(M_PI) * exp(-alpha * alpha * r2) - ee) * rr * rr;
    const int s1 = slot[a1], s2 = slot[a2];
    for (int k = 0; k < 3; ++k) {
      AccT<MIXED>::add(&facc[3 * s1 + k], -d[k] * fr);
      AccT<MIXED>::add(&facc[3 * s2 + k], d[k] * fr);
    }
  }
  en = wave_sum(en);
  if ((threadIdx.x & 63) == 0) s_e[threadIdx.x >> 6] = en;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) tot += s_e[w];
    epart[blk] = tot;
  }
}

// ---------------------------------------------------------------------------
// 6. Finalize: sorted accumulators -> caller's AoS order, FP64.
// The accumulators are cleared as they are read, and so are the pair
// kernel's tile counter and the boundary-pair count, so the next force pass
// starts from zero without a separate clearing launch (padding slots are
// never accumulated into).
template <bool MIXED>
__device__ __forceinline__ void unsort_forces_body(const int bid_, const int nbk_, int n, const int *__restrict__ slot,
                                                   typename AccT<MIXED>::type *__restrict__ facc,
                                                   double *__restrict__ fout, int *__restrict__ tilectr,
                                                   int *__restrict__ nborder) {
  int a = bid_ * blockDim.x + threadIdx.x;
  if (a == 0) {
    *tilectr = 0;
    *nborder = 0;
  }
  if (a >= n) return;
  const int s = slot[a];
  for (int k = 0; k < 3; ++k) {
    if (MIXED)
      fout[3 * a + k] = (double)(long long)facc[3 * s + k] * (1.0 / kForceScale);
    else
      fout[3 * a + k] = ((const double *)facc)[3 * s + k];
    facc[3 * s + k] = 0;
  }
}

template <bool MIXED>
__global__ void k_unsort_forces(int n, const int *__restrict__ slot, typename AccT<MIXED>::type *__restrict__ facc,
                                double *__restrict__ fout, int *__restrict__ tilectr, int *__restrict__ nborder) {
  unsort_forces_body<MIXED>(blockIdx.x, gridDim.x, n, slot, facc, fout, tilectr, nborder);
}

// Deterministic energy/pair-stat reduction in one block, in index order.
// Pair energies per tile are int64 fixed point (2^-32) when pair_fixed (the
// mixed form: their sum is exact, whatever tiles the schedule made), FP64
// otherwise.
__device__ __forceinline__ void reduce_energy_body(const int bid_, const int nbk_, int ntiles_cap,
                                                   const int *__restrict__ ntiles_p, const double *__restrict__ pair_e,
                                                   int pair_fixed, int nrecip, const double *__restrict__ recip_e,
                                                   int ncorr, const double *__restrict__ corr_e,
                                                   const long long *__restrict__ npairs,
                                                   const unsigned long long *__restrict__ hash, int stats,
                                                   const unsigned long long *__restrict__ border_e,
                                                   const unsigned long long *__restrict__ border_stat,
                                                   double *__restrict__ out, long long *__restrict__ out_np,
                                                   unsigned long long *__restrict__ out_hash) {
  __shared__ double sd[4][1024 / kWave];
  __shared__ long long sn[1024 / kWave];
  __shared__ unsigned long long sh[1024 / kWave];
  __shared__ long long sf[2][1024 / kWave];
  const int nt = min(*ntiles_p, ntiles_cap);
  double c = 0, l = 0, r = 0, k = 0;
  long long np = 0, cf = 0, lf = 0;
  unsigned long long h = 0;
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    if (pair_fixed) {
      cf += __double_as_longlong(pair_e[2 * i]);
      lf += __double_as_longlong(pair_e[2 * i + 1]);
    } else {
      c += pair_e[2 * i];
      l += pair_e[2 * i + 1];
    }
    if (stats) {
      np += npairs[i];
      h += hash[i];
    }
  }
  for (int i = threadIdx.x; i < nrecip; i += blockDim.x) r += recip_e[i];
  for (int i = threadIdx.x; i < ncorr; i += blockDim.x) k += corr_e[i];
  c = wave_sum(c);
  l = wave_sum(l);
  cf = wave_sum(cf);
  lf = wave_sum(lf);
  r = wave_sum(r);
  k = wave_sum(k);
  np = wave_sum(np);
  h = wave_sum(h);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sd[0][w] = c;
    sd[1][w] = l;
    sd[2][w] = r;
    sd[3][w] = k;
    sn[w] = np;
    sh[w] = h;
    sf[0][w] = cf;
    sf[1][w] = lf;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double C = 0, Lj = 0, Rr = 0, K = 0;
    long long N = 0, CF = 0, LF = 0;
    unsigned long long H = 0;
    for (int i = 0; i < (int)(blockDim.x / kWave); ++i) {
      CF += sf[0][i];
      LF += sf[1][i];
      C += sd[0][i];
      Lj += sd[1][i];
      Rr += sd[2][i];
      K += sd[3][i];
      N += sn[i];
      H += sh[i];
    }
    out[0] = C + (double)(CF + (long long)border_e[0]) * (1.0 / kForceScale);
    out[1] = Lj + (double)(LF + (long long)border_e[1]) * (1.0 / kForceScale);
    out[2] = 0.5 * Rr;
    out[3] = K;
    *out_np = N + (long long)border_stat[0];
    *out_hash = H + border_stat[1];
  }
}

__global__ void k_reduce_energy(int ntiles_cap, const int *__restrict__ ntiles_p, const double *__restrict__ pair_e,
                                int pair_fixed, int nrecip, const double *__restrict__ recip_e, int ncorr,
                                const double *__restrict__ corr_e, const long long *__restrict__ npairs,
                                const unsigned long long *__restrict__ hash, int stats,
                                const unsigned long long *__restrict__ border_e,
                                const unsigned long long *__restrict__ border_stat, double *__restrict__ out,
                                long long *__restrict__ out_np, unsigned long long *__restrict__ out_hash) {
  reduce_energy_body(blockIdx.x, gridDim.x, ntiles_cap, ntiles_p, pair_e, pair_fixed, nrecip, recip_e, ncorr, corr_e,
                     npairs, hash, stats, border_e, border_stat, out, out_np, out_hash);
}

// ---------------------------------------------------------------------------
// 7. Device-resident MD: bonded terms, leapfrog kick/drift, kinetic energy.

// Zeroes up to 8 ranges of 64-bit words in one launch (blockIdx.y = range).
struct ZeroRanges {
  unsigned long long *p[8];
  long long n[8];
};
__global__ void k_zero(ZeroRanges z) {
  const int r = blockIdx.y;
  unsigned long long *p = z.p[r];
  const long long n = z.n[r];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// Reference minimalDifference(c1, c2) = c2 - c1 folded to the minimal image.
__device__ inline void min_diff(const double *x, int c1, int c2, const Box &bx, double *d) {
  double d2;
  for (int k = 0; k < 3; ++k) d[k] = x[3 * c2 + k] - x[3 * c1 + k];
  d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  if (d2 > bx.D) {
    for (int k = 0; k < 3; ++k) {
      if (d[k] < -bx.H2[k])
        d[k] += bx.H[k];
      else if (d[k] > bx.H2[k])
        d[k] -= bx.H[k];
    }
    d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (d2 > bx.D)
      for (int k = 0; k < 3; ++k) d[k] -= bx.L[k] * rint(bx.Lr[k] * d[k]);
  }
}

// Harmonic bonds (BondSystemForce.h:97-118) and angles without Urey-Bradley
// (AngleSystemForce.h:95-147), FP64.  Threads [0, nb) bonds, [nb, nb+na) angles.
template <bool MIXED>
__device__ void bonded_block(int blk, int nb, int na, const int2 *__restrict__ bonds, const double2 *__restrict__ bpar,
                             const int4 *__restrict__ angles, const double2 *__restrict__ apar,
                             const double *__restrict__ x, const int *__restrict__ slot, Box box,
                             typename AccT<MIXED>::type *__restrict__ facc, double *__restrict__ epart) {
  __shared__ double s_e[2][kPairBlock / kWave];
  const int t = blk * blockDim.x + threadIdx.x;
  double eb = 0.0, ea = 0.0;
  if (t < nb) {
    const int2 b = bonds[t];
    const double2 p = bpar[t];
    double r12[3];
    min_diff(x, b.y, b.x, box, r12);  // atom1 - atom2
    const double r = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
    const double dpot = 2.0 * p.x * (r - p.y);
    const double s1 = slot[b.x], s2 = slot[b.y];
    for (int k = 0; k < 3; ++k) {
      const double f1 = r12[k] * (-dpot / r);
      AccT<MIXED>::add(&facc[3 * (int)s1 + k], f1);
      AccT<MIXED>::add(&facc[3 * (int)s2 + k], -f1);
    }
    eb = p.x * (r - p.y) * (r - p.y);
  } else if (t < nb + na) {
    const int4 a = angles[t - nb];
    const double2 p = apar[t - nb];
    double r12[3], r32[3];
    min_diff(x, a.y, a.x, box, r12);  // atom1 - atom2
    min_diff(x, a.y, a.z, box, r32);  // atom3 - atom2
    const double d12 = sqrt(r12[0] * r12[0] + r12[1] * r12[1] + r12[2] * r12[2]);
    const double d32 = sqrt(r32[0] * r32[0] + r32[1] * r32[1] + r32[2] * r32[2]);
    const double cx = r12[1] * r32[2] - r12[2] * r32[1], cy = r12[2] * r32[0] - r12[0] * r32[2],
                 cz = r12[0] * r32[1] - r12[1] * r32[0];
    const double theta = atan2(sqrt(cx * cx + cy * cy + cz * cz), r12[0] * r32[0] + r12[1] * r32[1] + r12[2] * r32[2]);
    const double st = sin(theta), ct = cos(theta);
    const double dpot = 2.0 * p.x * (theta - p.y);
    double f1[3], f3[3];
    for (int k = 0; k < 3; ++k) {
      const double dr12 = r12[k] / d12, dr32 = r32[k] / d32;
      const double dt1 = (dr12 * ct - dr32) / (st * d12);
      const double dt3 = (dr32 * ct - dr12) / (st * d32);
      f1[k] = -dt1 * dpot;
      f3[k] = -dt3 * dpot;
    }
    const int s1 = slot[a.x], s2 = slot[a.y], s3 = slot[a.z];
    for (int k = 0; k < 3; ++k) {
      AccT<MIXED>::add(&facc[3 * s1 + k], f1[k]);
      AccT<MIXED>::add(&facc[3 * s2 + k], -f1[k] - f3[k]);
      AccT<MIXED>::add(&facc[3 * s3 + k], f3[k]);
    }
    ea = p.x * (theta - p.y) * (theta - p.y);
  }
  eb = wave_sum(eb);
  ea = wave_sum(ea);
  if ((threadIdx.x & 63) == 0) {
    s_e[0][threadIdx.x >> 6] = eb;
    s_e[1][threadIdx.x >> 6] = ea;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = 0, a = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) {
      b += s_e[0][w];
      a += s_e[1][w];
    }
    epart[2 * blk] = b;
    epart[2 * blk + 1] = a;
  }
}

// The small FP64 passes that follow the pair and PME kernels, in one launch
// of 256-thread blocks: [0, border) boundary pairs, then exclusion
// correction blocks, then bonded-term blocks.
template <bool MIXED>
struct FixupArgs {
  using R4 = typename Real4<typename std::conditional<MIXED, float, double>::type>::type;
  int border_blocks, corr_blocks;
  // boundary pairs
  PairParams pp;
  int stats;
  const int2 *border;
  const int *nborder;
  int maxborder;
  const R4 *xs;
  const float4 *sorg;
  const int *ts, *perm;
  const double *xorig, *ljtab;
  const int *eptr;
  const int2 *elist;
  unsigned long long *border_e, *border_stat;
  // correction
  int ne;
  const int4 *etab;
  const double *q;
  const int *slot;
  double alpha, csf;
  double *corr_e;
  // bonded
  int nbond, nangle;
  const int2 *bonds;
  const double2 *bpar;
  const int4 *angles;
  const double2 *apar;
  double *bond_e;
  typename AccT<MIXED>::type *facc;
};

template <bool MIXED, int COUL, int LJSW, int CSW>
__device__ __forceinline__ void fixups_body(const int bid_, const int nbk_, FixupArgs<MIXED> a) {
  int b = bid_;
  if (b < a.border_blocks) {
    border_pairs<MIXED, COUL, LJSW, CSW>(b, a.border_blocks, a.pp, a.stats, a.border, a.nborder, a.maxborder, a.xs,
                                         a.sorg, a.ts, a.perm, a.xorig, a.ljtab, a.eptr, a.elist, a.facc, a.border_e,
                                         a.border_stat);
    return;
  }
  b -= a.border_blocks;
  if (b < a.corr_blocks) {
    correction_block<MIXED>(b, a.ne, a.etab, a.xorig, a.q, a.slot, a.pp.box, a.alpha, a.csf, a.facc, a.corr_e);
    return;
  }
  b -= a.corr_blocks;
  bonded_block<MIXED>(b, a.nbond, a.nangle, a.bonds, a.bpar, a.angles, a.apar, a.xorig, a.slot, a.pp.box, a.facc,
                      a.bond_e);
}

template <bool MIXED, int COUL, int LJSW, int CSW>
__global__ void __launch_bounds__(256) k_fixups(FixupArgs<MIXED> a) {
  fixups_body<MIXED, COUL, LJSW, CSW>(blockIdx.x, gridDim.x, a);
}

// Velocity kicks of one leapfrog step.  Leapfrog (LeapfrogIntegrator.cpp):
// v += F h_kick / m.  Langevin leapfrog (LangevinLeapfrogIntegrator.cpp:
// 44-71): nhalf half kicks v = v vdt + F fdt / m + g sqrt(fc / m) ndt, g a
// standard normal triple for (atom, global half-kick index).
// One kick: h_kick = fraction of the step times h; Langevin half kicks
// numbered half0, half0 + 1 (plus 2 x *step_dev when step_dev is set: a
// captured chunk reads its first step index from device memory, half0 then
// being relative to it, modulo 2^64).
struct Kick {
  double h_kick;
  int langevin, nhalf;
  unsigned long long half0, seed;
  double vdt, fdt, ndt, fc;
  const unsigned long long *step_dev;
};

__device__ inline void apply_kick(const Kick &kk, int a, const double f[3], double minv, double v[3]) {
  if (!kk.langevin) {
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] += f[k] * kk.h_kick * minv;
    return;
  }
  const double sfm = sqrt(kk.fc * minv) * kk.ndt;
  const unsigned long long h0 = kk.half0 + (kk.step_dev ? 2ull * *kk.step_dev : 0ull);
  for (int hh = 0; hh < kk.nhalf; ++hh) {
    double g[3];
    langevin_gauss3(kk.seed, h0 + hh, (uint32_t)a, g);
#pragma unroll
    for (int k = 0; k < 3; ++k) v[k] = v[k] * kk.vdt + f[k] * kk.fdt * minv + g[k] * sfm;
  }
}

// Kick (v += F kick h / m) and drift (x += v h) of every atom, FP64
// (LeapfrogIntegrator.cpp:33-87).  The pair kernel's sorted coordinates then
// follow each atom continuously from its wrapped position at the last list
// build (xs64), since list entries carry a fixed image shift, and the largest
// squared displacement since that build is folded into dmax (float bits).
template <class R>
__device__ __forceinline__ void kick_drift_body(const int bid_, const int nbk_, int n, Kick kk, double h,
                                                const int *__restrict__ perm, const double *__restrict__ fout,
                                                const double *__restrict__ minv, double *__restrict__ v,
                                                double *__restrict__ x, const double *__restrict__ xbuild,
                                                typename Real4<R>::type *__restrict__ xs,
                                                const double4 *__restrict__ xs64, const float4 *__restrict__ sorg,
                                                unsigned int *__restrict__ dmax) {
  const int s = bid_ * blockDim.x + threadIdx.x;
  float d2f = 0.f;
  const int a = s < n ? perm[s] : -1;  // n counts slots; padding slots have a < 0
  if (a >= 0) {
    double dd[3], vv[3] = {v[3 * a], v[3 * a + 1], v[3 * a + 2]};
    const double ff[3] = {fout[3 * a], fout[3 * a + 1], fout[3 * a + 2]};
    apply_kick(kk, a, ff, minv[a], vv);
    double d2 = 0;
    for (int k = 0; k < 3; ++k) {
      v[3 * a + k] = vv[k];
      const double xk = x[3 * a + k] + vv[k] * h;
      x[3 * a + k] = xk;
      dd[k] = xk - xbuild[3 * a + k];
      d2 += dd[k] * dd[k];
    }
    d2f = (float)d2;
    const double4 b = xs64[s];
    const float4 o = sorg[s / kSuperSize];
    typename Real4<R>::type r = xs[s];
    r.x = (R)((b.x + dd[0]) - (double)o.x);
    r.y = (R)((b.y + dd[1]) - (double)o.y);
    r.z = (R)((b.z + dd[2]) - (double)o.z);
    xs[s] = r;
  }
  // wave max, one atomic per wave
  for (int m = 32; m >= 1; m >>= 1) d2f = fmaxf(d2f, __shfl_xor(d2f, m));
  if ((threadIdx.x & 63) == 0) atomicMax(dmax, __float_as_uint(d2f));
}

template <class R>
__global__ void k_kick_drift(int n, Kick kk, double h, const int *__restrict__ perm,
                             const double *__restrict__ fout, const double *__restrict__ minv, double *__restrict__ v,
                             double *__restrict__ x, const double *__restrict__ xbuild,
                             typename Real4<R>::type *__restrict__ xs, const double4 *__restrict__ xs64,
                             const float4 *__restrict__ sorg, unsigned int *__restrict__ dmax) {
  kick_drift_body<R>(blockIdx.x, gridDim.x, n, kk, h, perm, fout, minv, v, x, xbuild, xs, xs64, sorg, dmax);
}

// The end of one MD step and the start of the next in one pass over the
// sorted slots: the step's forces leave the accumulators (cleared, with the
// pass counters) for the caller's atom order, and the next step's kick and
// drift use them at once (k_unsort_forces followed by k_kick_drift).
template <bool MIXED, class R>
__global__ void k_unsort_kick_drift(int n, Kick kk, double h, const int *__restrict__ perm,
                                    typename AccT<MIXED>::type *__restrict__ facc, double *__restrict__ fout,
                                    int *__restrict__ tilectr, int *__restrict__ nborder,
                                    const double *__restrict__ minv, double *__restrict__ v,
                                    double *__restrict__ x, const double *__restrict__ xbuild,
                                    typename Real4<R>::type *__restrict__ xs, const double4 *__restrict__ xs64,
                                    const float4 *__restrict__ sorg, unsigned int *__restrict__ dmax) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) {
    *tilectr = 0;
    *nborder = 0;
  }
  float d2f = 0.f;
  const int a = s < n ? perm[s] : -1;  // n counts slots; padding slots have a < 0
  if (a >= 0) {
    double ff[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (MIXED)
        ff[k] = (double)(long long)facc[3 * s + k] * (1.0 / kForceScale);
      else
        ff[k] = ((const double *)facc)[3 * s + k];
      facc[3 * s + k] = 0;
      fout[3 * a + k] = ff[k];
    }
    double dd[3], vv[3] = {v[3 * a], v[3 * a + 1], v[3 * a + 2]};
    apply_kick(kk, a, ff, minv[a], vv);
    double d2 = 0;
    for (int k = 0; k < 3; ++k) {
      v[3 * a + k] = vv[k];
      const double xk = x[3 * a + k] + vv[k] * h;
      x[3 * a + k] = xk;
      dd[k] = xk - xbuild[3 * a + k];
      d2 += dd[k] * dd[k];
    }
    d2f = (float)d2;
    const double4 b = xs64[s];
    const float4 o = sorg[s / kSuperSize];
    typename Real4<R>::type r = xs[s];
    r.x = (R)((b.x + dd[0]) - (double)o.x);
    r.y = (R)((b.y + dd[1]) - (double)o.y);
    r.z = (R)((b.z + dd[2]) - (double)o.z);
    xs[s] = r;
  }
  for (int m = 32; m >= 1; m >>= 1) d2f = fmaxf(d2f, __shfl_xor(d2f, m));
  if ((threadIdx.x & 63) == 0) atomicMax(dmax, __float_as_uint(d2f));
}

// k_unsort_kick_drift and the charge spread of the next force pass in one
// launch: block b takes the kSpreadSlots slots from 32 b; lanes 0..31 of wave
// 0 unsort their slot's force, kick and drift its atom as k_unsort_kick_drift
// does, and bring the new position to the group's spread (the grid was left
// cleared by the previous pass).  Used inside an MD chunk, where the next
// pass spreads these positions.
template <bool MIXED, class R, int P>
__device__ __forceinline__ void unsort_kick_drift_spread_body(const int bid_, const int nbk_, int n, Kick kk, double h,
                                                              const int *__restrict__ perm,
                                                              typename AccT<MIXED>::type *__restrict__ facc,
                                                              double *__restrict__ fout, int *__restrict__ tilectr,
                                                              int *__restrict__ nborder,
                                                              const double *__restrict__ minv, double *__restrict__ v,
                                                              double *__restrict__ x, const double *__restrict__ xbuild,
                                                              typename Real4<R>::type *__restrict__ xs,
                                                              const double4 *__restrict__ xs64,
                                                              const float4 *__restrict__ sorg,
                                                              unsigned int *__restrict__ dmax, PmeParams pme,
                                                              const double *__restrict__ q, double *__restrict__ grid64,
                                                              unsigned long long *__restrict__ gridfx) {
  const int s0 = bid_ * kSpreadSlots;
  const int t = threadIdx.x, s = s0 + t;
  if (bid_ == 0 && t == 0) {
    *tilectr = 0;
    *nborder = 0;
  }
  int a = -1;
  double xa[3] = {0, 0, 0};
  if (t < kWave) {
    float d2f = 0.f;
    if (t < kSpreadSlots && s < n) a = perm[s];
    if (a >= 0) {
      double ff[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        if (MIXED)
          ff[k] = (double)(long long)facc[3 * s + k] * (1.0 / kForceScale);
        else
          ff[k] = ((const double *)facc)[3 * s + k];
        facc[3 * s + k] = 0;
        fout[3 * a + k] = ff[k];
      }
      double dd[3], vv[3] = {v[3 * a], v[3 * a + 1], v[3 * a + 2]};
      apply_kick(kk, a, ff, minv[a], vv);
      double d2 = 0;
      for (int k = 0; k < 3; ++k) {
        v[3 * a + k] = vv[k];
        const double xk = x[3 * a + k] + vv[k] * h;
        x[3 * a + k] = xk;
        xa[k] = xk;
        dd[k] = xk - xbuild[3 * a + k];
        d2 += dd[k] * dd[k];
      }
      d2f = (float)d2;
      const double4 b = xs64[s];
      const float4 o = sorg[s / kSuperSize];
      typename Real4<R>::type r = xs[s];
      r.x = (R)((b.x + dd[0]) - (double)o.x);
      r.y = (R)((b.y + dd[1]) - (double)o.y);
      r.z = (R)((b.z + dd[2]) - (double)o.z);
      xs[s] = r;
    }
    for (int m = 32; m >= 1; m >>= 1) d2f = fmaxf(d2f, __shfl_xor(d2f, m));
    if (t == 0) atomicMax(dmax, __float_as_uint(d2f));
  }
  spread_patch_group<MIXED, P>(s0, pme, a, xa, q, grid64, gridfx);
}

template <bool MIXED, class R, int P>
__global__ void __launch_bounds__(256) k_unsort_kick_drift_spread(
    int n, Kick kk, double h, const int *__restrict__ perm, typename AccT<MIXED>::type *__restrict__ facc,
    double *__restrict__ fout, int *__restrict__ tilectr, int *__restrict__ nborder, const double *__restrict__ minv,
    double *__restrict__ v, double *__restrict__ x, const double *__restrict__ xbuild,
    typename Real4<R>::type *__restrict__ xs, const double4 *__restrict__ xs64, const float4 *__restrict__ sorg,
    unsigned int *__restrict__ dmax, PmeParams pme, const double *__restrict__ q, double *__restrict__ grid64,
    unsigned long long *__restrict__ gridfx) {
  // the blocks of one XCD take a contiguous range of slots
  unsort_kick_drift_spread_body<MIXED, R, P>(xcd_block(blockIdx.x, gridDim.x), gridDim.x, n, kk, h, perm, facc, fout,
                                             tilectr, nborder, minv, v, x, xbuild, xs, xs64, sorg, dmax, pme, q,
                                             grid64, gridfx);
}

// Start of an MD chunk in one launch: the committed state (x, v, f: m3 =
// 3N doubles each) is backed up for a rollback, the overflow flags (2) and
// the chunk's displacement record (nrec) are cleared.
__device__ __forceinline__ void chunk_begin_body(const int bid_, const int nbk_, long long m3,
                                                 const double *__restrict__ x, const double *__restrict__ v,
                                                 const double *__restrict__ f, double *__restrict__ xbak,
                                                 double *__restrict__ vbak, double *__restrict__ fbak,
                                                 int *__restrict__ overflow, unsigned *__restrict__ dmax, int nrec) {
  const long long i = (long long)bid_ * blockDim.x + threadIdx.x;
  if (i < m3) {
    xbak[i] = x[i];
    vbak[i] = v[i];
    fbak[i] = f[i];
  }
  if (i < 2) overflow[i] = 0;
  if (i < nrec) dmax[i] = 0u;
}

__global__ void k_chunk_begin(long long m3, const double *__restrict__ x, const double *__restrict__ v,
                              const double *__restrict__ f, double *__restrict__ xbak, double *__restrict__ vbak,
                              double *__restrict__ fbak, int *__restrict__ overflow, unsigned *__restrict__ dmax,
                              int nrec) {
  chunk_begin_body(blockIdx.x, gridDim.x, m3, x, v, f, xbak, vbak, fbak, overflow, dmax, nrec);
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

__device__ __forceinline__ void half_kick_body(const int bid_, const int nbk_, int n, Kick kk,
                                               const double *__restrict__ fout, const double *__restrict__ minv,
                                               double *__restrict__ v) {
  const int a = bid_ * blockDim.x + threadIdx.x;
  if (a >= n) return;
  double vv[3] = {v[3 * a], v[3 * a + 1], v[3 * a + 2]};
  const double ff[3] = {fout[3 * a], fout[3 * a + 1], fout[3 * a + 2]};
  apply_kick(kk, a, ff, minv[a], vv);
  for (int k = 0; k < 3; ++k) v[3 * a + k] = vv[k];
}

__global__ void k_half_kick(int n, Kick kk, const double *__restrict__ fout, const double *__restrict__ minv,
                            double *__restrict__ v) {
  half_kick_body(blockIdx.x, gridDim.x, n, kk, fout, minv, v);
}

// Kinetic energy 0.5 sum m v^2 into per-block partials.
__global__ void k_kinetic(int n, const double *__restrict__ v, const double *__restrict__ minv,
                          double *__restrict__ epart) {
  __shared__ double s_e[kPairBlock / kWave];
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (a < n) e = 0.5 * (v[3 * a] * v[3 * a] + v[3 * a + 1] * v[3 * a + 1] + v[3 * a + 2] * v[3 * a + 2]) / minv[a];
  e = wave_sum(e);
  if ((threadIdx.x & 63) == 0) s_e[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_e[w];
    epart[blockIdx.x] = t;
  }
}

}  // namespace pmx
