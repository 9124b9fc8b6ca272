// Shared definitions for the pmx HIP kernels and host code (gfx950).
#pragma once
#include <cmath>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmx {

constexpr int kClusterSize = 8;    // atoms per cluster (one 8-lane group of a wave)
constexpr int kSuperSize = 64;     // atoms per supercluster (8 clusters, one wave)
constexpr int kTileEntries = 16;   // granularity of the per-supercluster list capacity
constexpr int kMaxPairWavesPerCu = 32;  // resident pair-kernel waves per CU, at most
// Work runs of the pair kernel: the list is cut into runs of equal estimated
// work whose number depends only on the list (never on the device or the
// launch shape): kMinRuns k, k = ceil(total work / (kMinRuns kRunCost)),
// 1 <= k <= kMaxRuns / kMinRuns.  A wave sums its i-forces in FP32 over a
// run's part of one supercluster (a tile) and adds the sums to the force
// array in fixed point, so forces and energies are the same whichever wave
// takes which run.  kMinRuns is the pair kernel's resident wave count on
// MI355X (256 CUs x 4 blocks of 4 waves): the static schedule gives every
// wave k equal runs there (one for systems up to about 10^5 atoms: a run
// boundary costs a flush of the i-forces, so runs are long).
constexpr int kMinRuns = 4096;
constexpr int kMaxRuns = 5 * kMinRuns;
constexpr int kRunCost = 8192;     // work units (list cost model) per run, at least
constexpr int kWave = 64;
constexpr int kPairBlock = 256;    // 4 waves
constexpr double kForceScale = 4294967296.0;      // 2^32: int64 fixed-point forces
constexpr double kGridScale = 1099511627776.0;    // 2^40: int64 fixed-point PME charge grid
constexpr int kMaxTypes = 16;        // LJ types whose table the pair kernel keeps in LDS
constexpr int kMaxTypesTotal = 4096; // LJ types accepted (tables beyond kMaxTypes are read from HBM)

enum { SW_NONE = 0, SW_CUTOFF = 1, SW_C1 = 2, SW_C2 = 3 };
enum { COUL_NONE = 0, COUL_CUTOFF = 1, COUL_PME = 2 };

// Orthogonal periodic box, with the reference's minimal-image constants
// (protomol/src/protomol/topology/PeriodicBoundaryConditions.cpp:32-77).
struct Box {
  double L[3], Linv[3], Lr[3], o[3], mn[3], H[3], H2[3];
  double D;
};

// Per-switch constants (C2SwitchingFunction.cpp:13-19, C1SwitchingFunction.cpp:12-16).
struct SwitchC {
  int kind;
  double rc2, on2, s1, s2, s3, c15_1, c05_3, c15_3;
};

struct PairParams {
  Box box;
  SwitchC lj, co;
  double rc2max;     // membership radius^2 (max of the two cutoffs)
  double guard;      // |r2 - rc^2| band re-tested in FP64 with the reference's operation order
  double alpha, a2pi, csf;
  int ntypes;
  int coul;          // COUL_*
  int has_lj;
  // box edges split as Lhi + Llo with Lhi a multiple of 1/16 A, so that image
  // shifts of supercluster origins (also multiples of 1/16) add exactly
  double Lhi[3], Llo[3];
  // pair-list radius^2 (cutoff + margin + skin, with slack for the FP32 r^2):
  // the pass after a list build drops cluster pairs without a candidate
  // pair inside it, which no later pass before the next build can need
  double prune_r2;
};

// Sorted coordinates are stored relative to the origin of their supercluster
// (the position of its first atom rounded to a multiple of kOriginQuantum),
// so FP32 keeps about 1e-6 A of precision in any box size; origin
// differences and multiples of Lhi are exact in FP32.
constexpr double kOriginQuantum = 1.0 / 16.0;
__host__ __device__ inline double origin_round(double x) { return rint(x * 16.0) * kOriginQuantum; }

// Order-independent pair hash: splitmix64 of (i << 32 | j), i < j.
__host__ __device__ inline uint64_t pair_hash(uint32_t i, uint32_t j) {
  uint64_t z = ((uint64_t)i << 32) | (uint64_t)j;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC 2011): ten rounds of
// two 32x32 -> 64-bit products with the published multipliers and Weyl key
// increments.
__host__ __device__ inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Three standard normal numbers for (atom, half kick) of a Langevin run:
// Philox block at counter (atom, half lo, half hi, 'LANG'), uniforms
// (x + 1/2) 2^-32, Box-Muller on (u0, u1) and (u2, u3).
__host__ __device__ inline void langevin_gauss3(uint64_t seed, uint64_t half, uint32_t atom, double g[3]) {
  const uint32_t ctr[4] = {atom, (uint32_t)half, (uint32_t)(half >> 32), 0x4C414E47u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t x[4];
  philox4x32_10(ctr, key, x);
  const double s = 2.3283064365386963e-10;  // 2^-32
  const double u0 = (x[0] + 0.5) * s, u1 = (x[1] + 0.5) * s, u2 = (x[2] + 0.5) * s, u3 = (x[3] + 0.5) * s;
  const double twopi = 6.283185307179586;
  const double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
  g[0] = r0 * cos(twopi * u1);
  g[1] = r0 * sin(twopi * u1);
  g[2] = r1 * cos(twopi * u3);
}

}  // namespace pmx
