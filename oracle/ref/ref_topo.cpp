// Test driver for the reference ProtoMol topology builder, compiled against
// the reference's own sources (oracle/ref/Makefile).  It is test
// infrastructure: it reads a CHARMM PSF and PAR file with the reference's
// PSFReader / PARReader, runs buildTopology, buildMoleculeTable and
// buildExclusionTable (protomol/topology/BuildTopology.cpp:40-1158) on a
// periodic topology, and writes what the nonbonded path reads from it.
//
// Usage: pm_ref_topo <psf> <par> <exclude> <outfile>
//   exclude: none | 1-2 | 1-3 | 1-4 | scaled1-4   (the .conf "exclude" keyword)
// Output (whitespace separated):
//   N T NE
//   N lines:      type scaledCharge molecule
//   T*T lines:    A B A14 B14            (lennardJonesParameters, row-major)
//   NE lines:     a1 a2 class check      (exclusion table entries in insertion
//                                         order; check = exclusions.check(a1,a2);
//                                         classes as in the system files:
//                                         0 none, 1 FULL, 2 MODIFIED)
//   NB NA ND NI   bonded term counts

#include <protomol/topology/Topology.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/CubicCellManager.h>
#include <protomol/topology/BuildTopology.h>
#include <protomol/topology/ExclusionType.h>
#include <protomol/io/PSFReader.h>
#include <protomol/io/PARReader.h>
#include <protomol/type/PSF.h>
#include <protomol/type/PAR.h>

#include <cstdio>
#include <iostream>
#include <string>

using namespace ProtoMol;
typedef Topology<PeriodicBoundaryConditions, CubicCellManager> Topo;

static int cls(ExclusionClass c) { return c == EXCLUSION_FULL ? 1 : (c == EXCLUSION_MODIFIED ? 2 : 0); }

int main(int argc, char **argv) {
  if (argc != 5) {
    std::cerr << "usage: pm_ref_topo <psf> <par> <exclude> <outfile>\n";
    return 2;
  }
  PSF psf;
  PAR par;
  PSFReader psfr(argv[1]);
  if (!psfr.read(psf)) { std::cerr << "pm_ref_topo: cannot read PSF\n"; return 2; }
  PARReader parr(argv[2]);
  if (!parr.read(par)) { std::cerr << "pm_ref_topo: cannot read PAR\n"; return 2; }
  ExclusionType ex(argv[3]);
  if (!ex.valid()) { std::cerr << "pm_ref_topo: bad exclusion type\n"; return 2; }
  PeriodicBoundaryConditions bc(Vector3D(100, 0, 0), Vector3D(0, 100, 0), Vector3D(0, 0, 100), Vector3D(0, 0, 0));
  Topo topo(1.0, ex, bc, CubicCellManager(5.0), 16.0f);
  buildTopology(&topo, psf, par, false, 0);
  buildMoleculeTable(&topo);
  buildExclusionTable(&topo, ex);
  const int N = (int)topo.atoms.size(), T = (int)topo.atomTypes.size();
  const std::vector<ExclusionPair> &tab = topo.exclusions.getTable();
  FILE *fo = std::fopen(argv[4], "w");
  if (!fo) return 2;
  std::fprintf(fo, "%d %d %d\n", N, T, (int)tab.size());
  for (int i = 0; i < N; ++i)
    std::fprintf(fo, "%d %.17g %d\n", topo.atoms[i].type, topo.atoms[i].scaledCharge, topo.atoms[i].molecule);
  for (int a = 0; a < T; ++a)
    for (int b = 0; b < T; ++b) {
      const LennardJonesParameters &p = topo.lennardJonesParameters(a, b);
      std::fprintf(fo, "%.17g %.17g %.17g %.17g\n", p.A, p.B, p.A14, p.B14);
    }
  for (size_t k = 0; k < tab.size(); ++k)
    std::fprintf(fo, "%d %d %d %d\n", tab[k].a1, tab[k].a2, cls(tab[k].excl),
                 cls(topo.exclusions.check(tab[k].a1, tab[k].a2)));
  std::fprintf(fo, "%d %d %d %d\n", (int)topo.bonds.size(), (int)topo.angles.size(), (int)topo.dihedrals.size(),
               (int)topo.impropers.size());
  std::fclose(fo);
  return 0;
}
