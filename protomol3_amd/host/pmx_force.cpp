// pmx_force: evaluates ProtoMol force lines on the GPU through the host
// boundary (ForceFactory -> GpuNonbondedSystemForce::evaluate -> pmx C ABI)
// for a system in the plain-text PMXSYS format, and writes the result in the
// same layout as the reference driver: "E_coulomb E_lj alpha", then N force
// rows.
//
//   pmx_force <sysfile> <outfile> cellsize <cs> [--mixed] [--device d]
//             [--ranks R] --force "<ProtoMol force line>" [--force "<line>"]
//
// With --ranks R the force group is evaluated the way R MPI ranks evaluate it
// under ProtoMol's static work distribution, one rank after the other in this
// process: each rank calls parallelEvaluate into its own force and energy
// arrays, and the ranks' arrays are summed (Parallel's reduction).  stderr
// reports which ranks evaluated.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "pmx_protomol.h"

using namespace pmx_host;

static std::string tok(std::istream &in) {
  std::string t;
  while (in >> t) {
    if (t[0] == '#') {
      std::string rest;
      std::getline(in, rest);
      continue;
    }
    return t;
  }
  std::cerr << "pmx_force: unexpected end of system file\n";
  std::exit(2);
}

int main(int argc, char **argv) {
  if (argc < 6) {
    std::cerr << "usage: pmx_force <sysfile> <outfile> cellsize <cs> [--mixed] --force \"<line>\" ...\n";
    return 2;
  }
  std::ifstream in(argv[1]);
  if (!in || tok(in) != "PMXSYS") {
    std::cerr << "pmx_force: cannot read " << argv[1] << "\n";
    return 2;
  }
  tok(in);
  Topology topo;
  const int n = std::atoi(tok(in).c_str());
  topo.ntypes = std::atoi(tok(in).c_str());
  const int ne = std::atoi(tok(in).c_str());
  tok(in);
  for (int k = 0; k < 3; ++k) topo.cellBasis[k] = std::atof(tok(in).c_str());
  for (int k = 0; k < 3; ++k) topo.cellOrigin[k] = std::atof(tok(in).c_str());
  tok(in);
  topo.coulombScalingFactor = std::atof(tok(in).c_str());
  topo.lj.resize((size_t)topo.ntypes * topo.ntypes * 4);
  for (double &v : topo.lj) v = std::strtod(tok(in).c_str(), nullptr);
  topo.scaledCharge.resize(n);
  topo.type.resize(n);
  topo.molecule.resize(n);
  for (int i = 0; i < n; ++i) {
    topo.type[i] = std::atoi(tok(in).c_str());
    topo.scaledCharge[i] = std::strtod(tok(in).c_str(), nullptr);
    topo.molecule[i] = std::atoi(tok(in).c_str());
  }
  topo.exclusions.resize(3 * (size_t)ne);
  for (int &v : topo.exclusions) v = std::atoi(tok(in).c_str());
  Vector3DBlock pos(n), forces(n);
  for (double &v : pos.c) v = std::strtod(tok(in).c_str(), nullptr);

  int precision = PMX_FP64, device = 0, ranks = 0;
  std::vector<std::string> lines;
  for (int a = 3; a < argc; ++a) {
    std::string k = argv[a];
    if (k == "cellsize" && a + 1 < argc)
      topo.cellsize = std::atof(argv[++a]);
    else if (k == "--mixed")
      precision = PMX_MIXED;
    else if (k == "--device" && a + 1 < argc)
      device = std::atoi(argv[++a]);
    else if (k == "--ranks" && a + 1 < argc)
      ranks = std::atoi(argv[++a]);
    else if (k == "--force" && a + 1 < argc)
      lines.push_back(argv[++a]);
    else {
      std::cerr << "pmx_force: unknown argument " << k << "\n";
      return 2;
    }
  }
  ForceFactory factory;
  std::unique_ptr<SystemForce> f = factory.make(lines, precision, device);
  if (!f) {
    std::cerr << "pmx_force: " << factory.lastError() << "\n";
    return 3;
  }
  ScalarStructure energies;
  energies.clear();
  forces.zero();
  if (ranks <= 0) {
    f->preProcess(&topo, &pos);
    f->evaluate(&topo, &pos, &forces, &energies);
    f->postProcess(&topo, &energies, &forces);
  } else {
    for (int r = 0; r < ranks; ++r) {
      Parallel::init(Parallel::STATIC, r, ranks);
      Parallel::resetNext();
      Vector3DBlock fr(n);
      ScalarStructure er;
      er.clear();
      f->parallelEvaluate(&topo, &pos, &fr, &er);
      double mag = 0;
      for (double v : fr.c) mag += v * v;
      std::fprintf(stderr, "rank %d of %d: blocks %u, %s\n", r, ranks, f->numberOfBlocks(&topo, &pos),
                   mag > 0 ? "evaluated" : "nothing");
      for (size_t k = 0; k < fr.c.size(); ++k) forces.c[k] += fr.c[k];
      for (int k = 0; k < ScalarStructure::LAST; ++k) energies.v[k] += er.v[k];
    }
  }
  const double alpha = static_cast<GpuNonbondedSystemForce *>(f.get())->alpha();
  FILE *fo = std::fopen(argv[2], "w");
  if (!fo) return 2;
  std::fprintf(fo, "%.17g %.17g %.17g\n", energies[ScalarStructure::COULOMB], energies[ScalarStructure::LENNARDJONES],
               alpha);
  for (int i = 0; i < n; ++i)
    std::fprintf(fo, "%.17g %.17g %.17g\n", forces.c[3 * i], forces.c[3 * i + 1], forces.c[3 * i + 2]);
  std::fclose(fo);
  return 0;
}
