// Test driver for protomol3_amd/host/protomol/GpuNonbondedForces.h, compiled
// against the reference ProtoMol sources (oracle/ref/Makefile, target
// module_check).  It is test infrastructure: it fills one ForceFactory with
// the reference's own CPU prototypes of the periodic nonbonded forces and a
// second one with the GPU prototypes, makes every .conf force line of the
// list below through both factories (ForceFactory::make, the path
// IntegratorFactory.cpp:176 takes for each `force` line), and prints for each
// line the id, parameter values and the class that each factory produced, and
// the GPU force's work-package count.  No force is evaluated (the reference
// never runs on a GPU host).
//
// Usage: pm_module_check           (all lines; exit status 0)
//        pm_module_check bad cpu|gpu N
//                                  (makes invalid line N through one factory:
//                                   the reference's report-and-abort)

// the module header brings in the reference force, factory and topology headers
#include "../../protomol3_amd/host/protomol/GpuNonbondedForces.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace ProtoMol;
typedef PeriodicBoundaryConditions PBC;
typedef CubicCellManager CCM;

// The reference's CPU prototypes for the same ids
// (module/NonbondedCutoffForceModule.cpp:55-100; the PME template of
// mdl/interface/force/nonbonded/NonbondedPMEwaldSystemForce.h).
static void registerCpu(ForceFactory &f) {
  typedef C1SwitchingFunction C1;
  typedef C2SwitchingFunction C2;
  typedef CutoffSwitchingFunction Cut;
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, C2, LennardJonesForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, C1, LennardJonesForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, Cut, LennardJonesForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, C2, CoulombForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, C1, CoulombForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPair<PBC, Cut, CoulombForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPairTwo<PBC, C2, LennardJonesForce, C1, CoulombForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPairTwo<PBC, C2, LennardJonesForce, C2, CoulombForce> >());
  f.reg(new NonbondedCutoffSystemForce<CCM, OneAtomPairTwo<PBC, C2, LennardJonesForce, Cut, CoulombForce> >());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, true, true, BSpline, Cut>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, true, false, BSpline, Cut>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, false, true, BSpline, Cut>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, false, true, true, BSpline, Cut>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, false, false, BSpline, Cut>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, false, true, false, BSpline, Cut>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, false, false, true, BSpline, Cut>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, true, true, BSpline, C1>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, true, false, BSpline, C1>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, false, true, BSpline, C1>());
  f.reg(new NonbondedPMEwaldSystemForce<PBC, CCM, true, false, false, BSpline, C1>());
}

static const char *kLines[] = {
    "LennardJones -algorithm NonbondedCutoff -switchingFunction C2 -switchon 10 -cutoff 12",
    "LennardJones -algorithm NonbondedCutoff -switchingFunction Cutoff -cutoff 9.5",
    "LennardJones -algorithm NonbondedCutoff -switchingFunction C1 -cutoff 9",
    "Coulomb -algorithm NonbondedCutoff -switchingFunction C1 -cutoff 6.5",
    "Coulomb -algorithm NonbondedCutoff -switchingFunction C2 -switchon 7 -cutoff 9.5",
    "Coulomb -algorithm NonbondedCutoff -switchingFunction Cutoff -cutoff 10",
    "LennardJones Coulomb -algorithm NonbondedCutoff -switchingFunction C2 -switchingFunction C1 "
    "-switchon 0.1 -cutoff 6.5 -cutoff 6.5 -cutoff 6.5",
    "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline -cutoff 12 "
    "-gridsize 64 64 64 -order 4 -accuracy 1e-6",
    "Coulomb -algorithm PMEwald -real -interpolation BSpline -cutoff 10 -gridsize 32 32 32 -order 4 -alpha 0.32",
    "Coulomb -algorithm PMEwald -reciprocal -correction -interpolation BSpline -cutoff 10 -gridsize 30 30 30 "
    "-order 6 -alpha 0.34",
    "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline -switchingFunction C1 "
    "-cutoff 10 -gridsize 32 32 32 -order 4 -alpha 0.32",
};

static const char *kBad[] = {
    "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline -cutoff 10 "
    "-gridsize 4 4 4 -order 6",
    "Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline -cutoff 10 "
    "-gridsize 32 32 32 -order 5",
};

static std::string params(const Force *f) {
  std::vector<Parameter> p;
  f->getParameters(p);
  std::string s;
  for (unsigned int i = 0; i < p.size(); ++i) s += (i ? " " : "") + p[i].keyword + "=" + p[i].value.getString();
  return s;
}

static bool isGpu(Force *f) {
  pmxgpu::Terms t;
  return f && pmxgpu::gpuTerms(f, &t);
}

int main(int argc, char **argv) {
  ForceFactory cpu, gpu;
  registerCpu(cpu);
  pmxgpu::registerGpuNonbondedForces(gpu);
  if (argc == 4 && std::string(argv[1]) == "bad") {
    const int k = std::atoi(argv[3]);
    Force *f = (std::string(argv[2]) == "gpu" ? gpu : cpu).make(kBad[k]);
    std::printf("made %s\n", f ? "a force" : "NULL");
    return 0;
  }
  int bad = 0;
  for (unsigned int k = 0; k < sizeof(kLines) / sizeof(kLines[0]); ++k) {
    Force *c = cpu.make(kLines[k]);
    Force *g = gpu.make(kLines[k]);
    const bool same = c && g && c->getId() == g->getId() && params(c) == params(g);
    std::printf("line %u: cpu=%s gpu=%s same=%d gpu_class=%d blocks=%u\n", k, c ? c->getId().c_str() : "NULL",
                g ? g->getId().c_str() : "NULL", (int)same, (int)isGpu(g),
                g ? g->numberOfBlocks(0, 0) : 0u);
    std::printf("  params %s\n", g ? params(g).c_str() : "");
    pmxgpu::Terms t;
    if (g && pmxgpu::gpuTerms(g, &t))
      std::printf("  terms lj %d %.17g %.17g coul %d %d %.17g %.17g pme %d %d %d %d %d %.17g %.17g\n", t.lj_switch,
                  t.lj_on, t.lj_rc, t.coul_kind, t.coul_switch, t.coul_on, t.coul_rc, t.grid[0], t.grid[1], t.grid[2],
                  t.order, t.pme_terms, t.accuracy, t.alpha);
    if (!same || !isGpu(g) || g->numberOfBlocks(0, 0) != 1) ++bad;
    delete c;
    delete g;
  }
  std::printf("%s\n", bad ? "FAILED" : "OK");
  return bad ? 1 : 0;
}
