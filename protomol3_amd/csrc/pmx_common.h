// Shared definitions for the pmx HIP kernels and host code (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmx {

constexpr int kClusterSize = 8;    // atoms per cluster (one 8-lane group of a wave)
constexpr int kSuperSize = 64;     // atoms per supercluster (8 clusters, one wave)
constexpr int kTileEntries = 16;   // list entries per work tile
constexpr int kWave = 64;
constexpr int kPairBlock = 256;    // 4 waves
constexpr double kForceScale = 4294967296.0;      // 2^32: int64 fixed-point forces
constexpr double kGridScale = 1099511627776.0;    // 2^40: int64 fixed-point PME charge grid
constexpr int kMaxTypes = 16;

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
};

// Order-independent pair hash: splitmix64 of (i << 32 | j), i < j.
__host__ __device__ inline uint64_t pair_hash(uint32_t i, uint32_t j) {
  uint64_t z = ((uint64_t)i << 32) | (uint64_t)j;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace pmx
