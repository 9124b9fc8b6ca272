// CPU oracle for the ProtoMol nonbonded hot path -- TEST INFRASTRUCTURE ONLY.
// FP64 restatement of the reference algorithms; see pm_oracle.h.  Each block
// cites the reference file:line it follows (paths relative to the reference
// tree; PM = protomol/src/protomol, MDLNB = mdl/interface/force/nonbonded).
//
// Build: g++ -O2 -std=c++17 -fPIC -shared (oracle/Makefile).  Compiled
// without FMA contraction so that distance tests evaluate in the same
// operation order as the reference.
#include "pm_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <thread>
#include <tuple>
#include <vector>

namespace {

struct V3 {
  double x, y, z;
};
inline double nsq(const V3 &v) { return v.x * v.x + v.y * v.y + v.z * v.z; }

// ---------------------------------------------------------------------------
// Periodic boundary conditions, orthogonal box.
// PM/topology/PeriodicBoundaryConditions.cpp:32-77 (set),
// PeriodicBoundaryConditions.h:72-110 (minimalDifference),
// PeriodicBoundaryConditions.h:113-147 (minimalPosition).
struct Pbc {
  double L[3], Lr[3], o[3], mn[3], mx[3], H[3], H2[3], D, V;
  void set(const double *box, const double *origin) {
    for (int k = 0; k < 3; ++k) {
      L[k] = box[k];
      o[k] = origin[k];
    }
    // e1r = (e2 x e3) / e1.(e2 x e3): only the diagonal survives.
    Lr[0] = (L[1] * L[2]) / (L[0] * (L[1] * L[2]));
    Lr[1] = (L[2] * L[0]) / (L[1] * (L[2] * L[0]));
    Lr[2] = (L[0] * L[1]) / (L[2] * (L[0] * L[1]));
    V = std::fabs(L[0] * L[1] * L[2]);
    for (int k = 0; k < 3; ++k) {
      double a = o[k] - L[k] * 0.5, b = o[k] + L[k] * 0.5;
      mn[k] = std::min(a, b);
      mx[k] = std::max(a, b);
      H[k] = mx[k] - mn[k];
      H2[k] = H[k] * 0.5;
    }
    double dx = (L[0] * 0.5) * (L[0] * 0.5), dy = (L[1] * 0.5) * (L[1] * 0.5),
           dz = (L[2] * 0.5) * (L[2] * 0.5);
    D = std::min(dx, std::min(dy, dz));
  }
  V3 minimalDifference(const V3 &c1, const V3 &c2, double &d2) const {
    V3 d{c2.x - c1.x, c2.y - c1.y, c2.z - c1.z};
    d2 = nsq(d);
    if (d2 > D) {
      double *c = &d.x;
      for (int k = 0; k < 3; ++k) {
        if (c[k] < -H2[k])
          c[k] += H[k];
        else if (c[k] > H2[k])
          c[k] -= H[k];
      }
      d2 = nsq(d);
      if (d2 > D) {
        for (int k = 0; k < 3; ++k) c[k] -= L[k] * std::rint(Lr[k] * c[k]);
        d2 = nsq(d);
      }
    }
    return d;
  }
  V3 minimalPosition(const V3 &p) const {
    V3 d{p.x - o[0], p.y - o[1], p.z - o[2]};
    if (nsq(d) <= D) return d;
    double *c = &d.x;
    for (int k = 0; k < 3; ++k) {
      if (c[k] < -H2[k])
        c[k] += H[k];
      else if (c[k] > H2[k])
        c[k] -= H[k];
    }
    if (nsq(d) <= D) return d;
    return V3{d.x - L[0] * std::rint(Lr[0] * d.x), d.y - L[1] * std::rint(Lr[1] * d.y),
              d.z - L[2] * std::rint(Lr[2] * d.z)};
  }
  // PM/topology/PeriodicBoundaryConditions.cpp:79-104 plus the (0,0,0) that
  // NonbondedPMEwaldSystemForce.h:577-578 prepends.
  std::vector<V3> latticeVectors(double cutoff) const {
    std::vector<V3> lat;
    int bk = (int)std::floor(cutoff / std::fabs(L[0]) + 1.0);
    int bl = (int)std::floor(cutoff / std::fabs(L[1]) + 1.0);
    int bm = (int)std::floor(cutoff / std::fabs(L[2]) + 1.0);
    for (int k = -bk; k <= bk; k++)
      for (int l = -bl; l <= bl; l++)
        for (int m = -bm; m <= bm; m++)
          if (k != 0 || l != 0 || m != 0) {
            V3 c{L[0] * k, L[1] * l, L[2] * m};
            if (std::sqrt(nsq(c)) <= 2 * cutoff) lat.push_back(c);
          }
    std::stable_sort(lat.begin(), lat.end(), [](const V3 &a, const V3 &b) { return nsq(a) < nsq(b); });
    lat.insert(lat.begin(), V3{0, 0, 0});
    return lat;
  }
};

// ---------------------------------------------------------------------------
// Exclusion lookup: PM/topology/ExclusionTable.cpp:103-120 (add keeps the
// last class per pair for lookup, and a table in insertion order for the PME
// correction that holds each pair once, with the class it was first added
// with: the set it is filtered by orders ExclusionPair on (a1, a2) only,
// ExclusionTable.cpp:20-27), ExclusionTable.h:116-131 (check).
struct Excl {
  std::map<std::pair<int, int>, int> cls;
  std::vector<std::array<int, 3>> table;
  void build(const pmo_system *s) {
    std::set<std::pair<int, int>> seen;
    for (int e = 0; e < s->nexcl; ++e) {
      int a = s->excl[3 * e], b = s->excl[3 * e + 1], c = s->excl[3 * e + 2];
      if (a > b) std::swap(a, b);
      if (c != 0 && a != b && !seen.count({a, b})) {
        table.push_back({a, b, c});
        seen.insert({a, b});
      }
      cls[{a, b}] = c;
    }
  }
  int check(int a, int b) const {
    if (a > b) std::swap(a, b);
    auto it = cls.find({a, b});
    return it == cls.end() ? 0 : it->second;
  }
};

// ---------------------------------------------------------------------------
// Cell list.  PM/topology/CubicCellManager.cpp:40-65 (edge and findCell),
// ArrayCellListStructure.cpp:22-128 (grid dims, MAX_INT_2 offsets),
// ArrayCellListStructure.h:94-121 (findPeriodic, basisCell),
// Topology.h:44-81 (binning: i = N-1..0 prepends, so each list ascends).
constexpr double kEps = 1.0e-14;  // PM/base/PMConstants.cpp:50
constexpr int kMaxInt2 = 2147483647 / 2;

struct CellList {
  int nx, ny, nz;
  double edge[3], inv[3];
  std::vector<int> head, next;
  int maxN[3], n1[3], maxN1[3];

  void build(const Pbc &bc, const V3 *pos, int n, double cellsize) {
    const double diff[3] = {bc.mx[0] - bc.mn[0], bc.mx[1] - bc.mn[1], bc.mx[2] - bc.mn[2]};
    for (int k = 0; k < 3; ++k) {
      edge[k] = diff[k] / std::max(1.0, std::floor(diff[k] / cellsize));
      inv[k] = 1.0 / edge[k];
    }
    int d[3];
    for (int k = 0; k < 3; ++k) d[k] = std::max(1, (int)std::floor(diff[k] / edge[k] + kEps));
    nx = d[0];
    ny = d[1];
    nz = d[2];
    for (int k = 0; k < 3; ++k) {
      maxN[k] = kMaxInt2 - kMaxInt2 % d[k];
      n1[k] = -d[k] / 2;
      maxN1[k] = maxN[k] - n1[k];
    }
    head.assign((size_t)nx * ny * nz, -1);
    next.assign(n, -1);
    const V3 delta{bc.o[0] - bc.mn[0], bc.o[1] - bc.mn[1], bc.o[2] - bc.mn[2]};
    for (int i = n - 1; i >= 0; --i) {
      int c = cellOf(bc, delta, pos[i]);
      next[i] = head[c];
      head[c] = i;
    }
  }
  // findCell(delta + minimalPosition(x)).  A coordinate exactly on the upper
  // face (+L/2) gives index n, which the reference writes out of bounds; the
  // periodic image of that point is index 0, which is used here.
  int cellOf(const Pbc &bc, const V3 &delta, const V3 &p) const {
    V3 m = bc.minimalPosition(p);
    int c[3] = {(int)std::floor((delta.x + m.x) * inv[0]), (int)std::floor((delta.y + m.y) * inv[1]),
                (int)std::floor((delta.z + m.z) * inv[2])};
    const int dim[3] = {nx, ny, nz};
    for (int k = 0; k < 3; ++k) c[k] = ((c[k] % dim[k]) + dim[k]) % dim[k];
    return (c[0] * ny + c[1]) * nz + c[2];
  }
  std::array<int, 3> basisCell(int x, int y, int z) const {
    return {(x + maxN1[0]) % nx + n1[0], (y + maxN1[1]) % ny + n1[1], (z + maxN1[2]) % nz + n1[2]};
  }
  int periodicIndex(int x, int y, int z) const {
    return (((x + maxN[0]) % nx) * ny + (y + maxN[1]) % ny) * nz + (z + maxN[2]) % nz;
  }
};

// PM/topology/CellListEnumeratorPeriodicBoundaries.h:25-86 (delta list) and
// :114-134 (visit order): cell pairs (A,A) then (A,B) with B > A in array
// order, deltas sorted lexicographically.
std::vector<std::pair<int, int>> cellPairs(const CellList &cl, double cutoff) {
  int nxd = (int)(cutoff / cl.edge[0] + 1.0 + kEps);
  int nyd = (int)(cutoff / cl.edge[1] + 1.0 + kEps);
  int nzd = (int)(cutoff / cl.edge[2] + 1.0 + kEps);
  int n0 = std::min(nxd, cl.nx), n1 = std::min(nyd, cl.ny), n2 = std::min(nzd, cl.nz);
  double xx = cl.edge[0] * cl.edge[0], yy = cl.edge[1] * cl.edge[1], zz = cl.edge[2] * cl.edge[2];
  double c2 = cutoff * cutoff;
  std::set<std::array<int, 3>> seen;
  seen.insert({0, 0, 0});
  std::vector<std::array<int, 3>> deltas;
  for (int k = -n0; k <= n0; k++) {
    int x = std::max(std::abs(k) - 1, 0);
    double d0 = x * x * xx;
    for (int l = -n1; l <= n1; l++) {
      int y = std::max(std::abs(l) - 1, 0);
      double d1 = d0 + y * y * yy;
      for (int m = -n2; m <= n2; m++) {
        int z = std::max(std::abs(m) - 1, 0);
        if (d1 + z * z * zz < c2) {
          auto b = cl.basisCell(k, l, m);
          if (!seen.count(b)) {
            deltas.push_back(b);
            seen.insert(b);
          }
        }
      }
    }
  }
  std::sort(deltas.begin(), deltas.end());
  std::vector<std::pair<int, int>> out;
  const int ncell = cl.nx * cl.ny * cl.nz;
  for (int a = 0; a < ncell; ++a) {
    if (cl.head[a] < 0) continue;
    out.push_back({a, a});
    int ax = a / (cl.ny * cl.nz), ay = (a / cl.nz) % cl.ny, az = a % cl.nz;
    for (auto &d : deltas) {
      int b = cl.periodicIndex(ax + d[0], ay + d[1], az + d[2]);
      if (cl.head[b] < 0 || b < a) continue;
      out.push_back({a, b});
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// Switching functions.  PM/switch/C2SwitchingFunction.{h:42-53,cpp:13-19},
// C1SwitchingFunction.{h:41-50,cpp:12-16}, CutoffSwitchingFunction.h:41-44.
struct Switch {
  int kind = PMO_SW_NONE;
  double on, rc, on2, rc2, s1, s2, s3, c15_1, c05_3, c15_3;
  void set(int k, double son, double cut) {
    kind = k;
    on = son;
    rc = cut;
    rc2 = cut * cut;
    on2 = son * son;
    double t = cut * cut - son * son;
    s1 = 1.0 / (t * t * t);
    s2 = cut * cut - 3.0 * son * son;
    s3 = 4.0 / (t * t * t);
    c15_1 = 1.5 / cut;
    c05_3 = 0.5 / (cut * cut * cut);
    c15_3 = 1.5 / (cut * cut * cut);
  }
  void operator()(double &value, double &deriv, double d2) const {
    switch (kind) {
      case PMO_SW_C2:
        deriv = 0.0;
        if (d2 > rc2)
          value = 0.0;
        else if (d2 >= on2) {
          double c2 = rc2 - d2;
          double c4 = c2 * (s2 + 2.0 * d2);
          value = s1 * (c2 * c4);
          deriv = s3 * (c2 * c2 - c4);
        } else
          value = 1.0;
        return;
      case PMO_SW_C1: {
        if (d2 > rc2) {
          value = 0.0;
          deriv = 0.0;
          return;
        }
        double dist = std::sqrt(d2);
        value = 1.0 - dist * (c15_1 - d2 * c05_3);
        deriv = -c15_1 / dist + dist * c15_3;
        return;
      }
      default:
        deriv = 0;
        value = (d2 > rc2 ? 0.0 : 1.0);
        return;
    }
  }
};

// PM/base/MathUtilities.cpp:39-49 (Abramowitz & Stegun 7.1.26 polynomial).
inline double poly5(double ar) {
  const double E1 = 0.254829592, E2 = -0.284496736, E3 = 1.421413741, E4 = -1.453152027, E5 = 1.061405429,
               PP = 0.3275911;
  double t = 1.0 / (1.0 + PP * ar);
  return t * (E1 + t * (E2 + t * (E3 + t * (E4 + t * E5))));
}

struct Ctx {
  const pmo_system *s;
  Pbc bc;
  Excl ex;
  std::vector<V3> pos;
  CellList cl;
};

void initCtx(Ctx &c, const pmo_system *s, double cellsize) {
  c.s = s;
  c.bc.set(s->box, s->origin);
  c.ex.build(s);
  c.pos.resize(s->n);
  for (int i = 0; i < s->n; ++i) c.pos[i] = V3{s->pos[3 * i], s->pos[3 * i + 1], s->pos[3 * i + 2]};
  c.cl.build(c.bc, c.pos.data(), s->n, cellsize);
}

inline int exclusionOf(const Ctx &c, int i, int j) {
  // OneAtomPair.h:63-69: the table is consulted only within a molecule.
  if (c.s->molecule[i] != c.s->molecule[j]) return 0;
  return c.ex.check(i, j);
}

// One NonbondedCutoff pass (NonbondedCutoffForce.h:37-54 over
// OneAtomPair.h:47-106 with LennardJonesForce.h:28-60 or CoulombForce.h:29-40).
// kind 0 = LJ, 1 = Coulomb.  Cell pairs [cp0, cp1) of the visit list.
void cutoffPass(const Ctx &c, int kind, const Switch &sw, const std::vector<std::pair<int, int>> &cps, size_t cp0,
                size_t cp1, double *f, double &energy) {
  const pmo_system *s = c.s;
  const double rc2 = sw.rc2;
  for (size_t p = cp0; p < cp1; ++p) {
    const int A = cps[p].first, B = cps[p].second;
    const bool notSame = A != B;
    for (int i = c.cl.head[A]; i != -1; i = c.cl.next[i])
      for (int j = notSame ? c.cl.head[B] : c.cl.next[i]; j != -1; j = c.cl.next[j]) {
        double d2;
        V3 d = c.bc.minimalDifference(c.pos[i], c.pos[j], d2);
        if (d2 > rc2) continue;
        int ex = exclusionOf(c, i, j);
        if (ex == 1) continue;
        double r2 = 1.0 / d2;
        double e, fr;
        if (kind == 0) {
          const double *p4 = s->lj + 4 * (s->type[i] * s->ntypes + s->type[j]);
          double A_ = ex == 2 ? p4[2] : p4[0], B_ = ex == 2 ? p4[3] : p4[1];
          double r6 = r2 * r2 * r2, r12 = r6 * r6;
          double r6B = B_ * r6, r12A = A_ * r12;
          e = r12A - r6B;
          fr = 12.0 * r12A * r2 - 6.0 * r6B * r2;
        } else {
          e = s->charge[i] * s->charge[j] * std::sqrt(r2);
          if (ex == 2) e *= s->csf;
          fr = e * r2;
        }
        double sv, sd;
        sw(sv, sd, d2);
        fr = fr * sv - e * sd;
        e = e * sv;
        energy += e;
        double fx = d.x * fr, fy = d.y * fr, fz = d.z * fr;
        f[3 * i] -= fx;
        f[3 * i + 1] -= fy;
        f[3 * i + 2] -= fz;
        f[3 * j] += fx;
        f[3 * j + 1] += fy;
        f[3 * j + 2] += fz;
      }
  }
}

// PME real-space pass: MDLNB/NonbondedPMEwaldSystemForce.h:605-707.
// Includes j == i (self images) and the lattice-vector loop.
void pmeRealPass(const Ctx &c, const Switch &sw, double alpha, const std::vector<V3> &lat,
                 const std::vector<std::pair<int, int>> &cps, size_t cp0, size_t cp1, double *f, double &energy) {
  const pmo_system *s = c.s;
  const double rc2 = sw.rc2;
  const double a2pi = 2.0 * alpha / std::sqrt(M_PI);
  for (size_t p = cp0; p < cp1; ++p) {
    const int A = cps[p].first, B = cps[p].second;
    const bool notSame = A != B;
    for (int i = c.cl.head[A]; i != -1; i = c.cl.next[i]) {
      const double qi = s->charge[i];
      double fi[3] = {0, 0, 0};
      for (int j = notSame ? c.cl.head[B] : i; j != -1; j = c.cl.next[j]) {
        double dummy;
        V3 rm = c.bc.minimalDifference(c.pos[i], c.pos[j], dummy);
        const double qj = s->charge[j];
        int ex = exclusionOf(c, i, j);
        if (i == j) ex = 1;
        for (size_t k = 0; k < lat.size(); ++k) {
          if (ex == 1) {
            ex = 0;
            continue;
          }
          V3 rij{rm.x + lat[k].x, rm.y + lat[k].y, rm.z + lat[k].z};
          double rsq = nsq(rij);
          if (rsq > rc2) continue;
          double r = std::sqrt(rsq);
          double qq = qi * qj;
          if (ex == 2) qq *= s->csf;
          double rr = 1.0 / r;
          double ar = alpha * r;
          double e = qq * std::exp(-ar * ar);
          double en = poly5(ar) * e * rr;
          double fr = (en + a2pi * e) * rr * rr;
          double sv, sd;
          sw(sv, sd, rsq);
          fr = fr * sv - en * sd;
          en = en * sv;
          energy += en;
          double fx = rij.x * fr, fy = rij.y * fr, fz = rij.z * fr;
          fi[0] -= fx;
          fi[1] -= fy;
          fi[2] -= fz;
          f[3 * j] += fx;
          f[3 * j + 1] += fy;
          f[3 * j + 2] += fz;
          ex = 0;
        }
      }
      f[3 * i] += fi[0];
      f[3 * i + 1] += fi[1];
      f[3 * i + 2] += fi[2];
    }
  }
}

// ---------------------------------------------------------------------------
// FFT: mixed-radix recursive Cooley-Tukey for any length (radix 2, 3, 5 and
// a direct DFT for other prime factors).  Only the reference's semantics
// (unnormalised, opposite directions for the two transforms,
// PM/parallel/FFTComplex.cpp:56-96) matter here, not its backend.
using cd = std::complex<double>;
void fft1d(cd *a, int n, int stride, int sign, std::vector<cd> &scratch) {
  if (n == 1) return;
  int p = 2;
  while (p * p <= n && n % p) ++p;
  if (n % p) p = n;  // prime length
  const int m = n / p;
  std::vector<cd> tmp(n);
  for (int r = 0; r < p; ++r)
    for (int k = 0; k < m; ++k) tmp[r * m + k] = a[(k * p + r) * (size_t)stride];
  for (int r = 0; r < p; ++r) fft1d(tmp.data() + r * m, m, 1, sign, scratch);
  for (int k = 0; k < n; ++k) {
    cd acc = 0;
    const int kk = k % m;
    for (int r = 0; r < p; ++r) {
      double ang = sign * 2.0 * M_PI * (double)((long)r * k % n) / n;
      acc += tmp[r * m + kk] * cd(std::cos(ang), std::sin(ang));
    }
    a[(size_t)k * stride] = acc;
  }
}

void fft3d(cd *q, int nx, int ny, int nz, int sign) {
  std::vector<cd> scratch, line;
  line.resize(std::max(nx, std::max(ny, nz)));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) fft1d(q + ((size_t)i * ny + j) * nz, nz, 1, sign, scratch);
  for (int i = 0; i < nx; ++i)
    for (int k = 0; k < nz; ++k) fft1d(q + (size_t)i * ny * nz + k, ny, nz, sign, scratch);
  for (int j = 0; j < ny; ++j)
    for (int k = 0; k < nz; ++k) fft1d(q + (size_t)j * nz + k, nx, ny * nz, sign, scratch);
}

// MDLNB/BSpline.cpp:55-105.
void bspline(int order, double w, double *theta, double *dtheta) {
  if (order == 1) {
    theta[0] = 1;
    dtheta[0] = 0;
    return;
  }
  if (order == 2) {
    theta[1] = w;
    theta[0] = 1.0 - w;
    dtheta[1] = 1;
    dtheta[0] = -1;
    return;
  }
  theta[order - 1] = 0.0;
  theta[1] = w;
  theta[0] = 1.0 - w;
  for (int k = 2; k < order - 1; k++) {
    double div = 1.0 / k;
    theta[k] = div * w * theta[k - 1];
    for (int j = 1; j < k; j++) theta[k - j] = div * ((w + j) * theta[k - j - 1] + (k + 1 - j - w) * theta[k - j]);
    theta[0] = div * (1 - w) * theta[0];
  }
  dtheta[0] = -theta[0];
  for (int j = 1; j <= order - 1; j++) dtheta[j] = theta[j - 1] - theta[j];
  double div = 1.0 / (order - 1.0);
  theta[order - 1] = div * w * theta[order - 2];
  for (int j = 1; j < order - 1; j++)
    theta[order - 1 - j] = div * ((w + j) * theta[order - j - 2] + (order - j - w) * theta[order - 1 - j]);
  theta[0] = div * (1 - w) * theta[0];
}

// MDLNB/NonbondedPMEwaldSystemForce.h:520-537.
double alphaFor(double rc, double acc) {
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

// Reciprocal term: MDLNB/NonbondedPMEwaldSystemForce.h:812-864 with
// Grid.h:107-154 (anterpolateCharge), :237-341 (scalarSum), :159-189
// (interpolateForce), :499-600 (initialize, dftmod).
double pmeReciprocal(const Ctx &c, const pmo_spec *sp, double alpha, double *f) {
  const pmo_system *s = c.s;
  const int nx = sp->pme_grid[0], ny = sp->pme_grid[1], nz = sp->pme_grid[2], p = sp->pme_order;
  const double W = c.bc.L[0], Lh = c.bc.L[1], Hh = c.bc.L[2];
  const double Wr = 1.0 / W, Lr = 1.0 / Lh, Hr = 1.0 / Hh, V = W * Lh * Hh;
  const double hxr = nx / W, hyr = ny / Lh, hzr = nz / Hh;
  const int offx = -(p - 1) / 2 + nx, offy = -(p - 1) / 2 + ny, offz = -(p - 1) / 2 + nz;
  const double fac = M_PI * M_PI / (alpha * alpha);
  std::vector<double> ex(nx), ey(ny), ez(nz / 2 + 1), bx(nx), by(ny), bz(nz);
  for (int i = 0; i < nx; i++) {
    int i0 = i <= nx / 2 ? i : i - nx;
    ex[i] = std::exp(-fac * (i0 * Wr) * (i0 * Wr));
  }
  for (int j = 0; j < ny; j++) {
    int j0 = j <= ny / 2 ? j : j - ny;
    ey[j] = std::exp(-fac * (j0 * Lr) * (j0 * Lr));
  }
  for (int k = 0; k <= nz / 2; k++) ez[k] = std::exp(-fac * (k * Hr) * (k * Hr));
  std::vector<double> th0(p), dth0(p);
  bspline(p, 0.0, th0.data(), dth0.data());
  auto dftmod = [&](int n, std::vector<double> &out) {
    for (int i = 0; i < n; i++) {
      double sc = 0.0, ss = 0.0;
      for (int j = 0; j < p; j++) {
        double x = M_PI * 2.0 * i * j / (double)n;
        sc += th0[j] * std::cos(x);
        ss += th0[j] * std::sin(x);
      }
      out[i] = 1.0 / (sc * sc + ss * ss);
    }
  };
  dftmod(nx, bx);
  dftmod(ny, by);
  dftmod(nz, bz);

  const size_t G = (size_t)nx * ny * nz;
  std::vector<cd> Q(G, cd(0, 0));
  const int n = s->n;
  std::vector<int> base(3 * n);
  std::vector<double> th(3 * n * p), dth(3 * n * p);
  for (int a = 0; a < n; ++a) {
    double u[3] = {c.pos[a].x * hxr, c.pos[a].y * hyr, c.pos[a].z * hzr};
    const int dims[3] = {nx, ny, nz}, offs[3] = {offx, offy, offz};
    for (int d = 0; d < 3; ++d) {
      while (u[d] < 0.0) u[d] += dims[d];
      while (u[d] >= dims[d]) u[d] -= dims[d];
      int iu = (int)u[d];
      base[3 * a + d] = iu + offs[d];
      bspline(p, u[d] - iu, &th[(3 * a + d) * p], &dth[(3 * a + d) * p]);
    }
    const double *tx = &th[(3 * a) * p], *ty = &th[(3 * a + 1) * p], *tz = &th[(3 * a + 2) * p];
    for (int i = 0; i < p; i++) {
      double aq = s->charge[a] * tx[i];
      int i1 = (i + base[3 * a]) % nx;
      for (int j = 0; j < p; j++) {
        double ab = aq * ty[j];
        int j1 = (j + base[3 * a + 1]) % ny;
        for (int k = 0; k < p; k++) {
          int k1 = (k + base[3 * a + 2]) % nz;
          Q[((size_t)i1 * ny + j1) * nz + k1] += cd(ab * tz[k], 0.0);
        }
      }
    }
  }
  fft3d(Q.data(), nx, ny, nz, -1);  // fftBack
  double energy = 0.0;
  const double piVr = 1.0 / (M_PI * V);
  for (int i = 0; i < nx; i++) {
    int i0 = i <= nx / 2 ? i : i - nx;
    double mi = i0 * Wr;
    double exi = ex[i] * piVr;
    for (int j = 0; j < ny; j++) {
      int j0 = j <= ny / 2 ? j : j - ny;
      double bxy = bx[i] * by[j];
      double mj = j0 * Lr;
      double mij = mi * mi + mj * mj;
      double exy = exi * ey[j];
      for (int k = (i != 0 || j != 0 ? 0 : 1); k <= nz / 2; k++) {
        double mk = k * Hr;
        double m2 = mij + mk * mk;
        double theta = bxy * bz[k] * exy * ez[k] / m2;
        cd &q = Q[((size_t)i * ny + j) * nz + k];
        double e = (q.real() * q.real() + q.imag() * q.imag()) * theta;
        if (k > 0 && ((k != nz / 2) || (nz & 1))) {
          e *= 2.0;
          Q[((size_t)((nx - i) % nx) * ny + (ny - j) % ny) * nz + (nz - k) % nz] *= theta;
        }
        energy += e;
        q *= theta;
      }
    }
  }
  Q[0] = cd(0, 0);
  fft3d(Q.data(), nx, ny, nz, +1);  // fftForward
  for (int a = 0; a < n; ++a) {
    const double *tx = &th[(3 * a) * p], *ty = &th[(3 * a + 1) * p], *tz = &th[(3 * a + 2) * p];
    const double *dx = &dth[(3 * a) * p], *dy = &dth[(3 * a + 1) * p], *dz = &dth[(3 * a + 2) * p];
    double fx = 0, fy = 0, fz = 0;
    for (int i = 0; i < p; i++) {
      int i1 = (i + base[3 * a]) % nx;
      for (int j = 0; j < p; j++) {
        int j1 = (j + base[3 * a + 1]) % ny;
        double xij = dx[i] * ty[j], yij = tx[i] * dy[j], zij = tx[i] * ty[j];
        for (int k = 0; k < p; k++) {
          int k1 = (k + base[3 * a + 2]) % nz;
          double term = Q[((size_t)i1 * ny + j1) * nz + k1].real();
          fx -= term * xij * tz[k];
          fy -= term * yij * tz[k];
          fz -= term * zij * dz[k];
        }
      }
    }
    const double q = s->charge[a];
    f[3 * a] += fx * hxr * q;
    f[3 * a + 1] += fy * hyr * q;
    f[3 * a + 2] += fz * hzr * q;
  }
  return energy * 0.5;
}

// Correction term: MDLNB/NonbondedPMEwaldSystemForce.h:874-911 (exact erf).
double pmeCorrection(const Ctx &c, double alpha, double *f) {
  const pmo_system *s = c.s;
  const double a2pi = 2.0 * alpha / std::sqrt(M_PI), a2 = alpha * alpha;
  double intra = 0.0;
  for (auto &e : c.ex.table) {
    double r2;
    V3 rij = c.bc.minimalDifference(c.pos[e[0]], c.pos[e[1]], r2);
    double qq = s->charge[e[0]] * s->charge[e[1]];
    if (e[2] == 2) qq *= 1 - s->csf;
    double r = std::sqrt(r2), rr = 1 / r;
    double ee = std::erf(alpha * r) * rr;
    intra -= qq * ee;
    double fr = qq * (a2pi * std::exp(-a2 * r2) - ee) * rr * rr;
    f[3 * e[0]] -= rij.x * fr;
    f[3 * e[0] + 1] -= rij.y * fr;
    f[3 * e[0] + 2] -= rij.z * fr;
    f[3 * e[1]] += rij.x * fr;
    f[3 * e[1] + 1] += rij.y * fr;
    f[3 * e[1] + 2] += rij.z * fr;
  }
  return intra;
}

bool validSystem(const pmo_system *s) {
  if (!s || s->n < 0 || s->ntypes <= 0 || s->nexcl < 0) return false;
  if (s->n > 0 && (!s->pos || !s->charge || !s->type || !s->molecule)) return false;
  for (int k = 0; k < 3; ++k)
    if (!(s->box[k] > 0)) return false;
  return true;
}

// Runs the pass over the visit list, optionally split over threads, each
// thread with its own force buffer (summed in thread order).
template <class Fn>
double runPass(const Ctx &c, const std::vector<std::pair<int, int>> &cps, int nthreads, double *f, Fn fn) {
  if (nthreads <= 1) {
    double e = 0;
    fn(size_t(0), cps.size(), f, e);
    return e;
  }
  const int n = c.s->n;
  std::vector<std::vector<double>> bufs(nthreads, std::vector<double>(3 * (size_t)n, 0.0));
  std::vector<double> es(nthreads, 0.0);
  std::vector<std::thread> th;
  const size_t m = cps.size();
  for (int t = 0; t < nthreads; ++t) {
    size_t a = m * t / nthreads, b = m * (t + 1) / nthreads;
    th.emplace_back([&, a, b, t] { fn(a, b, bufs[t].data(), es[t]); });
  }
  for (auto &x : th) x.join();
  double e = 0;
  for (int t = 0; t < nthreads; ++t) {
    for (size_t k = 0; k < 3 * (size_t)n; ++k) f[k] += bufs[t][k];
    e += es[t];
  }
  return e;
}

}  // namespace

extern "C" {

double pmo_poly5(double x) { return poly5(x); }

double pmo_pme_alpha(double rc, double accuracy) { return alphaFor(rc, accuracy); }

int pmo_bspline(int order, double w, double *theta, double *dtheta) {
  if (order < 1 || !theta || !dtheta) return -1;
  bspline(order, w, theta, dtheta);
  return 0;
}

int pmo_fft3d(double *data, int nx, int ny, int nz, int sign) {
  if (!data || nx < 1 || ny < 1 || nz < 1) return -1;
  fft3d(reinterpret_cast<cd *>(data), nx, ny, nz, sign);
  return 0;
}

int pmo_cells(const pmo_system *s, double cellsize, int *cell_of_atom, int *dims) {
  if (!validSystem(s) || !(cellsize > 0)) return -1;
  Ctx c;
  initCtx(c, s, cellsize);
  const V3 delta{c.bc.o[0] - c.bc.mn[0], c.bc.o[1] - c.bc.mn[1], c.bc.o[2] - c.bc.mn[2]};
  for (int i = 0; i < s->n; ++i) cell_of_atom[i] = c.cl.cellOf(c.bc, delta, c.pos[i]);
  dims[0] = c.cl.nx;
  dims[1] = c.cl.ny;
  dims[2] = c.cl.nz;
  return 0;
}

int pmo_pairs(const pmo_system *s, double cellsize, double rc, int *pairs, long cap, long *npairs) {
  if (!validSystem(s) || !(cellsize > 0) || !(rc > 0) || !npairs) return -1;
  Ctx c;
  initCtx(c, s, cellsize);
  auto cps = cellPairs(c.cl, rc);
  long k = 0;
  const double rc2 = rc * rc;
  for (auto &cp : cps) {
    const bool notSame = cp.first != cp.second;
    for (int i = c.cl.head[cp.first]; i != -1; i = c.cl.next[i])
      for (int j = notSame ? c.cl.head[cp.second] : c.cl.next[i]; j != -1; j = c.cl.next[j]) {
        double d2;
        c.bc.minimalDifference(c.pos[i], c.pos[j], d2);
        if (d2 > rc2 || exclusionOf(c, i, j) == 1) continue;
        if (k < cap) {
          pairs[2 * k] = i;
          pairs[2 * k + 1] = j;
        }
        ++k;
      }
  }
  *npairs = k;
  return 0;
}

int pmo_pairs_bruteforce(const pmo_system *s, double rc, int *pairs, long cap, long *npairs) {
  if (!validSystem(s) || !(rc > 0) || !npairs) return -1;
  Ctx c;
  c.s = s;
  c.bc.set(s->box, s->origin);
  c.ex.build(s);
  c.pos.resize(s->n);
  for (int i = 0; i < s->n; ++i) c.pos[i] = V3{s->pos[3 * i], s->pos[3 * i + 1], s->pos[3 * i + 2]};
  long k = 0;
  const double rc2 = rc * rc;
  for (int i = 0; i < s->n; ++i)
    for (int j = i + 1; j < s->n; ++j) {
      double d2;
      c.bc.minimalDifference(c.pos[i], c.pos[j], d2);
      if (d2 > rc2 || exclusionOf(c, i, j) == 1) continue;
      if (k < cap) {
        pairs[2 * k] = i;
        pairs[2 * k + 1] = j;
      }
      ++k;
    }
  *npairs = k;
  return 0;
}

int pmo_evaluate(const pmo_system *s, const pmo_spec *sp, double *forces, double *energies, int nthreads) {
  if (!validSystem(s) || !sp || !forces || !energies || !(sp->cellsize > 0)) return -1;
  Ctx c;
  initCtx(c, s, sp->cellsize);
  if (sp->lj_switch != PMO_SW_NONE) {
    Switch sw;
    sw.set(sp->lj_switch, sp->lj_on, sp->lj_rc);
    auto cps = cellPairs(c.cl, sp->lj_rc);
    energies[1] += runPass(c, cps, nthreads, forces, [&](size_t a, size_t b, double *f, double &e) {
      cutoffPass(c, 0, sw, cps, a, b, f, e);
    });
  }
  if (sp->coul_kind == PMO_COUL_CUTOFF) {
    Switch sw;
    sw.set(sp->coul_switch, sp->coul_on, sp->coul_rc);
    auto cps = cellPairs(c.cl, sp->coul_rc);
    energies[0] += runPass(c, cps, nthreads, forces, [&](size_t a, size_t b, double *f, double &e) {
      cutoffPass(c, 1, sw, cps, a, b, f, e);
    });
  } else if (sp->coul_kind == PMO_COUL_PME) {
    const int p = sp->pme_order;
    if (p < 2 || p % 2 || sp->pme_grid[0] < p || sp->pme_grid[1] < p || sp->pme_grid[2] < p) return -2;
    const double alpha = sp->pme_alpha > 0 ? sp->pme_alpha : alphaFor(sp->coul_rc, sp->pme_accuracy);
    double intra = 0, self = 0, charged = 0, real = 0, recip = 0;
    if (sp->pme_corr) {
      intra = pmeCorrection(c, alpha, forces);
      double q2 = 0, q = 0;
      for (int i = 0; i < s->n; ++i) q2 += s->charge[i] * s->charge[i];
      for (int i = 0; i < s->n; ++i) q += s->charge[i];
      self = -q2 * alpha / std::sqrt(M_PI);
      if (std::fabs(q * 0.00268283) > 1.0e-5) charged = -M_PI / (2.0 * c.bc.V * alpha * alpha) * q * q;
    }
    if (sp->pme_real) {
      Switch sw;
      sw.set(sp->coul_switch == PMO_SW_C1 ? PMO_SW_C1 : PMO_SW_CUTOFF, 0.0, sp->coul_rc);
      auto lat = c.bc.latticeVectors(sp->coul_rc);
      auto cps = cellPairs(c.cl, sp->coul_rc);
      real = runPass(c, cps, nthreads, forces, [&](size_t a, size_t b, double *f, double &e) {
        pmeRealPass(c, sw, alpha, lat, cps, a, b, f, e);
      });
    }
    if (sp->pme_recip) recip = pmeReciprocal(c, sp, alpha, forces);
    energies[0] += real + recip + intra + 0.0 + self + charged;
  }
  return 0;
}

int pmo_pme_terms(const pmo_system *s, const pmo_spec *sp, double *out) {
  if (!validSystem(s) || !sp || !out || sp->coul_kind != PMO_COUL_PME) return -1;
  Ctx c;
  initCtx(c, s, sp->cellsize);
  std::vector<double> f(3 * (size_t)s->n, 0.0);
  const double alpha = sp->pme_alpha > 0 ? sp->pme_alpha : alphaFor(sp->coul_rc, sp->pme_accuracy);
  Switch sw;
  sw.set(sp->coul_switch == PMO_SW_C1 ? PMO_SW_C1 : PMO_SW_CUTOFF, 0.0, sp->coul_rc);
  auto lat = c.bc.latticeVectors(sp->coul_rc);
  auto cps = cellPairs(c.cl, sp->coul_rc);
  double real = 0;
  pmeRealPass(c, sw, alpha, lat, cps, 0, cps.size(), f.data(), real);
  double recip = pmeReciprocal(c, sp, alpha, f.data());
  double intra = pmeCorrection(c, alpha, f.data());
  double q2 = 0, q = 0;
  for (int i = 0; i < s->n; ++i) q2 += s->charge[i] * s->charge[i];
  for (int i = 0; i < s->n; ++i) q += s->charge[i];
  out[0] = real;
  out[1] = recip;
  out[2] = intra;
  out[3] = -q2 * alpha / std::sqrt(M_PI);
  out[4] = std::fabs(q * 0.00268283) > 1.0e-5 ? -M_PI / (2.0 * c.bc.V * alpha * alpha) * q * q : 0.0;
  out[5] = alpha;
  return 0;
}

// Classic Ewald sum (independent of the PME grid): exact erfc real space over
// every image within rc_real, reciprocal space over all m with
// |m_x|,|m_y|,|m_z| <= kmax, self term, and removal of FULL-excluded pairs
// (erf term), which is what PME's correction term approximates exactly.
int pmo_ewald_direct(const pmo_system *s, double alpha, double rc_real, int kmax, double *forces, double *energy) {
  if (!validSystem(s) || !(alpha > 0) || kmax < 1 || !forces || !energy) return -1;
  Pbc bc;
  bc.set(s->box, s->origin);
  Excl ex;
  ex.build(s);
  const int n = s->n;
  const double L[3] = {s->box[0], s->box[1], s->box[2]};
  const double V = L[0] * L[1] * L[2];
  const double a2pi = 2.0 * alpha / std::sqrt(M_PI);
  double E = 0;
  const int nb[3] = {(int)std::ceil(rc_real / L[0]), (int)std::ceil(rc_real / L[1]), (int)std::ceil(rc_real / L[2])};
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j) {
      int exc = (s->molecule[i] == s->molecule[j]) ? ex.check(i, j) : 0;
      double qq = s->charge[i] * s->charge[j];
      double d0[3] = {s->pos[3 * j] - s->pos[3 * i], s->pos[3 * j + 1] - s->pos[3 * i + 1],
                      s->pos[3 * j + 2] - s->pos[3 * i + 2]};
      for (int k = 0; k < 3; ++k) d0[k] -= L[k] * std::rint(d0[k] / L[k]);
      for (int a = -nb[0]; a <= nb[0]; ++a)
        for (int b = -nb[1]; b <= nb[1]; ++b)
          for (int cc = -nb[2]; cc <= nb[2]; ++cc) {
            const bool zero = a == 0 && b == 0 && cc == 0;
            if (i == j && zero) continue;
            double d[3] = {d0[0] + a * L[0], d0[1] + b * L[1], d0[2] + cc * L[2]};
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 > rc_real * rc_real) continue;
            double r = std::sqrt(r2);
            double w = (i == j) ? 0.5 : 1.0;
            double e, fr;
            if (zero && exc == 1) {
              // excluded pair: subtract the reciprocal-space part erf/r
              e = -qq * std::erf(alpha * r) / r;
              fr = qq * (a2pi * std::exp(-alpha * alpha * r2) - std::erf(alpha * r) / r) / r2;
            } else {
              e = qq * std::erfc(alpha * r) / r;
              fr = qq * (std::erfc(alpha * r) / r + a2pi * std::exp(-alpha * alpha * r2)) / r2;
            }
            E += w * e;
            if (i != j)
              for (int k = 0; k < 3; ++k) {
                forces[3 * i + k] -= d[k] * fr;
                forces[3 * j + k] += d[k] * fr;
              }
          }
    }
  // reciprocal space
  std::vector<double> Sre, Sim;
  for (int mx = -kmax; mx <= kmax; ++mx)
    for (int my = -kmax; my <= kmax; ++my)
      for (int mz = -kmax; mz <= kmax; ++mz) {
        if (!mx && !my && !mz) continue;
        double m[3] = {mx / L[0], my / L[1], mz / L[2]};
        double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
        double g = std::exp(-M_PI * M_PI * m2 / (alpha * alpha)) / (M_PI * V * m2);
        double sr = 0, si = 0;
        for (int i = 0; i < n; ++i) {
          double ph = 2 * M_PI * (m[0] * s->pos[3 * i] + m[1] * s->pos[3 * i + 1] + m[2] * s->pos[3 * i + 2]);
          sr += s->charge[i] * std::cos(ph);
          si += s->charge[i] * std::sin(ph);
        }
        E += 0.5 * g * (sr * sr + si * si);
        for (int i = 0; i < n; ++i) {
          double ph = 2 * M_PI * (m[0] * s->pos[3 * i] + m[1] * s->pos[3 * i + 1] + m[2] * s->pos[3 * i + 2]);
          // dE/dr_i = g q_i 2 pi m (-sin(ph) sr + cos(ph) si)
          double c = g * s->charge[i] * 2 * M_PI * (std::sin(ph) * sr - std::cos(ph) * si);
          for (int k = 0; k < 3; ++k) forces[3 * i + k] += c * m[k];
        }
      }
  double q2 = 0;
  for (int i = 0; i < n; ++i) q2 += s->charge[i] * s->charge[i];
  E += -q2 * alpha / std::sqrt(M_PI);
  *energy = E;
  return 0;
}

}  // extern "C"
