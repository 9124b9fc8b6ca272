// Test driver for the reference ProtoMol nonbonded path, compiled against the
// reference's own sources (oracle/ref/Makefile).  It is test infrastructure:
// it builds a Topology<PeriodicBoundaryConditions, CubicCellManager> straight
// from a plain-text system file, instantiates the reference's force templates
// and writes the forces and energies they produce.  Nothing here changes the
// reference's arithmetic; it only replaces the .conf/PSF front end with a
// direct fill of the same topology fields.
//
// System file (whitespace separated, '#' comment lines allowed):
//   PMXSYS 1
//   N T NE
//   box  ex ey ez  ox oy oz        (orthogonal box vectors, origin)
//   csf  c                         (coulombScalingFactor)
//   T*T lines: A B A14 B14         (row-major LJ table)
//   N lines:   type scaledCharge molecule
//   NE lines:  a1 a2 cls           (cls 1 = FULL, 2 = MODIFIED)
//   N lines:   x y z
//
// Usage: pm_ref <sysfile> <outfile> cellsize <cs> <force spec> ...
// Force specs (each may repeat; energies/forces of all are summed):
//   lj_c2 <switchon> <cutoff>        LennardJones NonbondedCutoff C2
//   lj_cutoff <cutoff>               LennardJones NonbondedCutoff Cutoff
//   coul_c1 <cutoff>                 Coulomb NonbondedCutoff C1
//   coul_c2 <switchon> <cutoff>      Coulomb NonbondedCutoff C2
//   coul_cutoff <cutoff>             Coulomb NonbondedCutoff Cutoff
//   ljcoul_c2c2 <on> <rc> <on> <rc>  LennardJones Coulomb (OneAtomPairTwo)
//   pme <nx> <ny> <nz> <order> <rc> <accuracy> <alpha> <RRC>
//        Coulomb PMEwald, Cutoff switch; RRC is three 0/1 flags for
//        -real -reciprocal -correction
//   pme_c1 ...                       same with the C1 switch
// Output: line 1 "E_coulomb E_lj alpha", then N lines "fx fy fz" (%.17g).
// With env PM_REF_PAIRS=<file>, also writes the cell of every atom and the
// ordered (i, j) visit list of the first cutoff force (membership oracle).
// With env PM_REF_MASSES=<file> (N numbers) the atoms get those masses
// (scaledMass; 1 otherwise).  With env PM_REF_COMPARE=1 the first two force
// specs are wrapped as the reference's compare pair (the `compare` force
// keyword: SystemCompareForce of the first, then of the second compared with
// the first) and evaluated twice; the reference prints its "Comparing"
// lines, and its averaged summary when the pair is destroyed.

#include <protomol/topology/Topology.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/CubicCellManager.h>
#include <protomol/topology/CellListEnumeratorPeriodicBoundaries.h>
#include <protomol/topology/ExclusionType.h>
#include <protomol/force/OneAtomPair.h>
#include <protomol/force/OneAtomPairTwo.h>
#include <protomol/force/LennardJonesForce.h>
#include <protomol/force/CoulombForce.h>
#include <protomol/force/nonbonded/NonbondedCutoffSystemForce.h>
#include <protomol/switch/C1SwitchingFunction.h>
#include <protomol/switch/C2SwitchingFunction.h>
#include <protomol/switch/CutoffSwitchingFunction.h>
#include <protomol/type/ScalarStructure.h>
#include <protomol/type/Vector3DBlock.h>
#include <protomol/base/Report.h>
#include <protomol/force/CompareForce.h>
using namespace ProtoMol::Report;
#include "BSpline.h"
#include "NonbondedPMEwaldSystemForce.h"

#include <sys/time.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

using namespace ProtoMol;
typedef PeriodicBoundaryConditions PBC;
typedef Topology<PBC, CubicCellManager> Topo;

static std::string next_tok(std::istream &in) {
  std::string t;
  while (in >> t) {
    if (t[0] == '#') {
      std::string rest;
      std::getline(in, rest);
      continue;
    }
    return t;
  }
  std::cerr << "pm_ref: unexpected end of system file\n";
  std::exit(2);
}
static double rd(std::istream &in) { return std::strtod(next_tok(in).c_str(), 0); }
static long ri(std::istream &in) { return std::strtol(next_tok(in).c_str(), 0, 10); }

int main(int argc, char **argv) {
  if (argc < 5) {
    std::cerr << "usage: pm_ref <sysfile> <outfile> cellsize <cs> <force spec>...\n";
    return 2;
  }
  std::ifstream in(argv[1]);
  if (!in) { std::cerr << "pm_ref: cannot open " << argv[1] << "\n"; return 2; }
  if (next_tok(in) != "PMXSYS") { std::cerr << "pm_ref: bad magic\n"; return 2; }
  ri(in);
  const int N = (int)ri(in), T = (int)ri(in), NE = (int)ri(in);
  next_tok(in);  // "box"
  double ex = rd(in), ey = rd(in), ez = rd(in), ox = rd(in), oy = rd(in), oz = rd(in);
  next_tok(in);  // "csf"
  double csf = rd(in);

  double cellsize = 0;
  int argi = 3;
  if (std::string(argv[argi]) != "cellsize") { std::cerr << "pm_ref: need cellsize\n"; return 2; }
  cellsize = std::atof(argv[argi + 1]);
  argi += 2;

  PBC bc(Vector3D(ex, 0, 0), Vector3D(0, ey, 0), Vector3D(0, 0, ez), Vector3D(ox, oy, oz));
  Topo topo(csf, ExclusionType(ExclusionType::ONE4MODIFIED), bc, CubicCellManager(cellsize), 16.0f);

  topo.lennardJonesParameters.resize(T);
  for (int a = 0; a < T; ++a)
    for (int b = 0; b < T; ++b) {
      double A = rd(in), B = rd(in), A14 = rd(in), B14 = rd(in);
      if (b >= a) topo.lennardJonesParameters.set(a, b, LennardJonesParameters(A, B, A14, B14));
    }
  topo.atoms.resize(N);
  int nmol = 0;
  for (int i = 0; i < N; ++i) {
    topo.atoms[i].type = (int)ri(in);
    topo.atoms[i].scaledCharge = rd(in);
    topo.atoms[i].molecule = (int)ri(in);
    topo.atoms[i].scaledMass = 1.0;
    if (topo.atoms[i].molecule + 1 > nmol) nmol = topo.atoms[i].molecule + 1;
  }
  if (const char *mf = std::getenv("PM_REF_MASSES")) {
    std::ifstream min(mf);
    for (int i = 0; i < N; ++i) topo.atoms[i].scaledMass = rd(min);
  }
  topo.exclusions.resize(N);
  topo.exclusions.clear();
  for (int e = 0; e < NE; ++e) {
    int a1 = (int)ri(in), a2 = (int)ri(in), c = (int)ri(in);
    topo.exclusions.add(a1, a2, c == 1 ? EXCLUSION_FULL : EXCLUSION_MODIFIED);
  }
  topo.exclusions.optimize();
  topo.molecules.resize(nmol);
  Vector3DBlock pos;
  pos.resize(N);
  for (int i = 0; i < N; ++i) {
    double x = rd(in), y = rd(in), z = rd(in);
    pos[i] = Vector3D(x, y, z);
  }

  Vector3DBlock forces;
  forces.resize(N);
  forces.zero();
  ScalarStructure energies;
  energies.clear();
  double alpha_used = 0.0;

  std::vector<SystemForce *> owned;
  while (argi < argc) {
    std::string k = argv[argi++];
    SystemForce *f = 0;
    if (k == "lj_c2") {
      double on = std::atof(argv[argi]), rc = std::atof(argv[argi + 1]);
      argi += 2;
      typedef OneAtomPair<PBC, C2SwitchingFunction, LennardJonesForce> P;
      f = new NonbondedCutoffSystemForce<CubicCellManager, P>(rc, P(LennardJonesForce(), C2SwitchingFunction(on, rc)));
    } else if (k == "lj_cutoff") {
      double rc = std::atof(argv[argi++]);
      typedef OneAtomPair<PBC, CutoffSwitchingFunction, LennardJonesForce> P;
      f = new NonbondedCutoffSystemForce<CubicCellManager, P>(rc, P(LennardJonesForce(), CutoffSwitchingFunction(rc)));
    } else if (k == "coul_c1") {
      double rc = std::atof(argv[argi++]);
      typedef OneAtomPair<PBC, C1SwitchingFunction, CoulombForce> P;
      f = new NonbondedCutoffSystemForce<CubicCellManager, P>(rc, P(CoulombForce(), C1SwitchingFunction(rc)));
    } else if (k == "coul_c2") {
      double on = std::atof(argv[argi]), rc = std::atof(argv[argi + 1]);
      argi += 2;
      typedef OneAtomPair<PBC, C2SwitchingFunction, CoulombForce> P;
      f = new NonbondedCutoffSystemForce<CubicCellManager, P>(rc, P(CoulombForce(), C2SwitchingFunction(on, rc)));
    } else if (k == "coul_cutoff") {
      double rc = std::atof(argv[argi++]);
      typedef OneAtomPair<PBC, CutoffSwitchingFunction, CoulombForce> P;
      f = new NonbondedCutoffSystemForce<CubicCellManager, P>(rc, P(CoulombForce(), CutoffSwitchingFunction(rc)));
    } else if (k == "ljcoul_c2c2") {
      double on1 = std::atof(argv[argi]), rc1 = std::atof(argv[argi + 1]);
      double on2 = std::atof(argv[argi + 2]), rc2 = std::atof(argv[argi + 3]);
      argi += 4;
      typedef OneAtomPairTwo<PBC, C2SwitchingFunction, LennardJonesForce, C2SwitchingFunction, CoulombForce> P;
      double rc = rc1 > rc2 ? rc1 : rc2;
      f = new NonbondedCutoffSystemForce<CubicCellManager, P>(
          rc, P(LennardJonesForce(), C2SwitchingFunction(on1, rc1), CoulombForce(), C2SwitchingFunction(on2, rc2)));
    } else if (k == "pme" || k == "pme_c1") {
      unsigned nx = std::atoi(argv[argi]), ny = std::atoi(argv[argi + 1]), nz = std::atoi(argv[argi + 2]);
      unsigned order = std::atoi(argv[argi + 3]);
      double rc = std::atof(argv[argi + 4]), acc = std::atof(argv[argi + 5]), alpha = std::atof(argv[argi + 6]);
      std::string m = argv[argi + 7];
      argi += 8;
      int code = (m[0] == '1') * 4 + (m[1] == '1') * 2 + (m[2] == '1');
#define PMX_PME(R, Q, C, SW)                                                                              \
  f = new NonbondedPMEwaldSystemForce<PBC, CubicCellManager, R, Q, C, BSpline, SW>(nx, ny, nz, order, rc, \
                                                                                  acc, alpha, 3.0)
      if (k == "pme") {
        switch (code) {
          case 7: PMX_PME(true, true, true, CutoffSwitchingFunction); break;
          case 4: PMX_PME(true, false, false, CutoffSwitchingFunction); break;
          case 2: PMX_PME(false, true, false, CutoffSwitchingFunction); break;
          case 1: PMX_PME(false, false, true, CutoffSwitchingFunction); break;
          case 6: PMX_PME(true, true, false, CutoffSwitchingFunction); break;
          default: std::cerr << "pm_ref: unsupported PME mask\n"; return 2;
        }
      } else {
        switch (code) {
          case 7: PMX_PME(true, true, true, C1SwitchingFunction); break;
          case 4: PMX_PME(true, false, false, C1SwitchingFunction); break;
          default: std::cerr << "pm_ref: unsupported PME mask\n"; return 2;
        }
      }
#undef PMX_PME
    } else {
      std::cerr << "pm_ref: unknown force spec " << k << "\n";
      return 2;
    }
    owned.push_back(f);
  }

  if (std::getenv("PM_REF_COMPARE") && owned.size() >= 2) {
    SystemForce *a = dynamic_cast<SystemForce *>(owned[0]->makeCompareForce(owned[0], 0));
    SystemForce *b =
        dynamic_cast<SystemForce *>(owned[1]->makeCompareForce(owned[1], dynamic_cast<CompareForce *>(a)));
    for (int rep = 0; rep < 2; ++rep) {
      forces.zero();
      energies.clear();
      topo.uncacheCellList();
      a->evaluate(&topo, &pos, &forces, &energies);
      b->evaluate(&topo, &pos, &forces, &energies);
    }
    delete b;  // prints the averaged summary; deletes the wrapped forces
    delete a;
    std::fflush(stdout);
    return 0;
  }

  // Same sequence as ForceGroup::evaluateSystemForces: cell lists are marked
  // stale once per evaluation, then each force is evaluated in turn.
  topo.uncacheCellList();
  for (size_t n = 0; n < owned.size(); ++n) {
    owned[n]->preProcess(&topo, &pos);
    owned[n]->evaluate(&topo, &pos, &forces, &energies);
    owned[n]->postProcess(&topo, &energies, &forces);
  }
  // With env PM_REF_TIME_REPS=k: k more evaluations of the same forces into
  // scratch arrays (the first one above set up the PME tables), timed
  // together; prints "eval_s <seconds per evaluation>".
  if (const char *reps_s = std::getenv("PM_REF_TIME_REPS")) {
    const int reps = std::atoi(reps_s);
    Vector3DBlock fs;
    fs.resize(N);
    ScalarStructure es;
    struct timeval tv0, tv1;
    gettimeofday(&tv0, 0);
    for (int r = 0; r < reps; ++r) {
      fs.zero();
      es.clear();
      topo.uncacheCellList();
      for (size_t n = 0; n < owned.size(); ++n) {
        owned[n]->preProcess(&topo, &pos);
        owned[n]->evaluate(&topo, &pos, &fs, &es);
        owned[n]->postProcess(&topo, &es, &fs);
      }
    }
    gettimeofday(&tv1, 0);
    const double dt = (tv1.tv_sec - tv0.tv_sec) + 1e-6 * (tv1.tv_usec - tv0.tv_usec);
    std::printf("eval_s %.6f\n", reps > 0 ? dt / reps : 0.0);
  }

  // Membership dump: every atom's cell and the cell-pair x atom-pair visit order
  // of the reference enumerator at the largest algorithm cutoff requested.
  const char *pairs_path = std::getenv("PM_REF_PAIRS");
  if (pairs_path) {
    double rcut = std::atof(std::getenv("PM_REF_PAIRS_RC") ? std::getenv("PM_REF_PAIRS_RC") : "0");
    FILE *pf = std::fopen(pairs_path, "w");
    topo.uncacheCellList();
    topo.updateCellLists(&pos);
    std::fprintf(pf, "cells %d %d %d\n", topo.cellLists.getDimX(), topo.cellLists.getDimY(),
                 topo.cellLists.getDimZ());
    for (int i = 0; i < N; ++i) {
      CubicCellManager::Cell c = topo.cellManager.findCell(
          Vector3D(topo.boundaryConditions.origin() - topo.min) + topo.boundaryConditions.minimalPosition(pos[i]));
      std::fprintf(pf, "%d %d %d %d\n", i, c.x, c.y, c.z);
    }
    Topo::Enumerator en;
    en.initialize(&topo, rcut);
    CellPair cp;
    long npairs = 0;
    std::vector<std::pair<int, int> > out;
    for (; !en.done(); en.next()) {
      en.get(cp);
      bool notSame = en.notSameCell();
      for (int i = cp.first; i != -1; i = topo.atoms[i].cellListNext)
        for (int j = (notSame ? cp.second : topo.atoms[i].cellListNext); j != -1; j = topo.atoms[j].cellListNext) {
          Real d2;
          topo.boundaryConditions.minimalDifference(pos[i], pos[j], d2);
          if (d2 > rcut * rcut) continue;
          if (topo.exclusions.check(i, j) == EXCLUSION_FULL) continue;
          out.push_back(std::make_pair(i, j));
          ++npairs;
        }
    }
    // order-independent set hash: sum of splitmix64(min << 32 | max), mod 2^64
    unsigned long long hsum = 0;
    for (size_t n = 0; n < out.size(); ++n) {
      unsigned long long a = (unsigned)std::min(out[n].first, out[n].second);
      unsigned long long b = (unsigned)std::max(out[n].first, out[n].second);
      unsigned long long z = (a << 32) | b;
      z += 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      hsum += z ^ (z >> 31);
    }
    std::fprintf(pf, "pairs %ld %llu\n", npairs, hsum);
    if (!std::getenv("PM_REF_PAIRS_NOLIST"))
      for (size_t n = 0; n < out.size(); ++n) std::fprintf(pf, "%d %d\n", out[n].first, out[n].second);
    std::fclose(pf);
  }

  FILE *fo = std::fopen(argv[2], "w");
  if (!fo) { std::cerr << "pm_ref: cannot write " << argv[2] << "\n"; return 2; }
  std::fprintf(fo, "%.17g %.17g %.17g\n", (double)energies[ScalarStructure::COULOMB],
               (double)energies[ScalarStructure::LENNARDJONES], alpha_used);
  for (int i = 0; i < N; ++i) std::fprintf(fo, "%.17g %.17g %.17g\n", forces[i].c[0], forces[i].c[1], forces[i].c[2]);
  std::fclose(fo);
  for (size_t n = 0; n < owned.size(); ++n) delete owned[n];
  return 0;
}
