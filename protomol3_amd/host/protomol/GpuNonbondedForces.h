/* -*- c++ -*- */
// ProtoMol-side module of the pmx engine: GPU versions of ProtoMol's periodic
// nonbonded system forces, written against ProtoMol's own headers and
// registered under ProtoMol's own force ids, so that a .conf force line
//
//   force LennardJones -algorithm NonbondedCutoff -switchingFunction C2 -switchon 10 -cutoff 12
//   force Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline
//         -cutoff 12 -gridsize 64 64 64 -order 4 -accuracy 1e-6
//
// selects the GPU force when the application registers these prototypes in
// place of the CPU ones (module/NonbondedCutoffForceModule.cpp:40-201 for the
// cutoff forces; the PME force of mdl/interface/force/nonbonded/
// NonbondedPMEwaldSystemForce.h is registered the same way).
//
// Each GPU force derives from the reference force type it replaces
// (NonbondedCutoffSystemForce<CubicCellManager, OneAtomPair<...>> and
// NonbondedPMEwaldSystemForce<PBC, CubicCellManager, real, reciprocal,
// correction, BSpline, Switch>), so its keyword, id, parameter list and
// .conf grammar are the reference's; it overrides only
//   evaluate          SystemForce.h:32-35: forces accumulated into `forces`,
//                     energies into the COULOMB / LENNARDJONES slots;
//   parallelEvaluate  SystemForce.h:107-112 + NonbondedCutoffSystemForce.h:57-77:
//                     the force is one work package (numberOfBlocks = 1), so
//                     under MPI exactly the rank that Parallel::next() hands
//                     it to evaluates it on its GPU and the others add nothing;
//                     Parallel's reduction then gives every rank the sum;
//   numberOfBlocks    Force.h:36-37;
//   uncache           Force.h:55 (the engine rebuilds its tables);
//   doMake            the reference's constructor arguments and checks.
// The system (charges, LJ table, exclusions, box) is taken from the
// topology at the first evaluation and kept resident on the device; a
// different topology object rebuilds it.  Inside a multiple-time-step
// integrator each GPU force is an ordinary member of its force group
// (integrator/MTSIntegrator.cpp:33-56): every evaluate is independent.
// Errors follow the reference: configuration errors make doMake report and
// return NULL (NonbondedPMEwaldSystemForce.h:1054-1058), run-time errors
// report and abort (base/Report.cpp:248-253).
//
// Build: compile the translation unit that includes this header with the
// ProtoMol include directories (protomol/src and
// mdl/interface/force/nonbonded) and link libpmx (include/pmx.h).
#ifndef PMX_GPU_NONBONDED_FORCES_H
#define PMX_GPU_NONBONDED_FORCES_H

#include <protomol/base/Report.h>
#include <protomol/config/Parameter.h>
#include <protomol/factory/ForceFactory.h>
#include <protomol/force/CoulombForce.h>
#include <protomol/force/LennardJonesForce.h>
#include <protomol/force/OneAtomPair.h>
#include <protomol/force/OneAtomPairTwo.h>
#include <protomol/force/nonbonded/NonbondedCutoffSystemForce.h>
#include <protomol/parallel/Parallel.h>
#include <protomol/switch/C1SwitchingFunction.h>
#include <protomol/switch/C2SwitchingFunction.h>
#include <protomol/switch/CutoffSwitchingFunction.h>
#include <protomol/topology/CellListEnumeratorPeriodicBoundaries.h>
#include <protomol/topology/CubicCellManager.h>
#include <protomol/topology/ExclusionType.h>
#include <protomol/topology/PeriodicBoundaryConditions.h>
#include <protomol/topology/Topology.h>
#include <protomol/type/ScalarStructure.h>
#include <protomol/type/Vector3DBlock.h>

// The PME force header of the MDL interface reports through an unqualified
// `report` and expects the periodic topology types declared before it.
using namespace ProtoMol::Report;
#include "BSpline.h"
#include "NonbondedPMEwaldSystemForce.h"
#include "pmx.h"

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace ProtoMol {
namespace pmxgpu {

typedef PeriodicBoundaryConditions PBC;
typedef CubicCellManager CCM;
typedef Topology<PBC, CCM> PeriodicTopology;

// Engine code of a switching function type.
template <class S>
struct SwitchCode;
template <>
struct SwitchCode<C2SwitchingFunction> {
  enum { value = PMX_SW_C2, nparams = 2 };  // -switchon, -cutoff
};
template <>
struct SwitchCode<C1SwitchingFunction> {
  enum { value = PMX_SW_C1, nparams = 1 };  // -cutoff
};
template <>
struct SwitchCode<CutoffSwitchingFunction> {
  enum { value = PMX_SW_CUTOFF, nparams = 1 };  // -cutoff
};

// Engine code of a pair functor type.
template <class F>
struct ForceCode;
template <>
struct ForceCode<LennardJonesForce> {
  enum { lj = 1, coulomb = 0 };
};
template <>
struct ForceCode<CoulombForce> {
  enum { lj = 0, coulomb = 1 };
};

// What one GPU force computes, as the engine's descriptor fields.
struct Terms {
  int lj_switch, coul_kind, coul_switch;
  double lj_on, lj_rc, coul_on, coul_rc;
  int grid[3], order, pme_terms;
  double accuracy, alpha;
  Terms()
      : lj_switch(PMX_SW_NONE), coul_kind(PMX_COUL_NONE), coul_switch(PMX_SW_NONE), lj_on(0), lj_rc(0), coul_on(0),
        coul_rc(0), order(4), pme_terms(0), accuracy(1e-6), alpha(-1.0) {
    grid[0] = grid[1] = grid[2] = 0;
  }
};

// Where the GPU forces run: engine precision (PMX_FP64, the plain form, or
// PMX_MIXED) and device.  Set before the prototypes are registered.
struct Placement {
  int precision, device;
  Placement() : precision(PMX_FP64), device(0) {}
};

// Values of a force's parameter list, in order.
inline std::vector<Real> parameterValues(const Force &m) {
  std::vector<Parameter> p;
  m.getParameters(p);
  std::vector<Real> v;
  for (unsigned int i = 0; i < p.size(); ++i) {
    Real x = 0;
    p[i].value.get(x);
    v.push_back(x);
  }
  return v;
}

// The resident device system of one GPU force.
class Device {
 public:
  Device() : ctx_(0), sys_(0), built_for_(0) {}
  ~Device() { release(); }
  Placement where;

  void uncache() {
    if (sys_) pmx_sys_uncache(sys_);
  }

  void evaluate(const GenericTopology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                ScalarStructure *energies, const Terms &t) {
    if (!sys_ || built_for_ != topo) build(topo, t);
    double e[PMX_E_COUNT] = {0, 0};
    if (pmx_sys_evaluate(sys_, &((*positions)[0].c[0]), &((*forces)[0].c[0]), e, 0) != PMX_OK)
      Report::report << Report::error << "pmx: " << pmx_last_error(ctx_) << Report::endr;
    (*energies)[ScalarStructure::COULOMB] += e[PMX_E_COULOMB];
    (*energies)[ScalarStructure::LENNARDJONES] += e[PMX_E_LENNARDJONES];
  }

 private:
  Device(const Device &);
  Device &operator=(const Device &);

  void release() {
    if (sys_) pmx_sys_destroy(sys_);
    if (ctx_) pmx_ctx_destroy(ctx_);
    sys_ = 0;
    ctx_ = 0;
  }

  // pmx_sys_desc from the topology (GenericTopology.h:155-161,
  // SemiGenericTopology.h, Topology.h:236).
  void build(const GenericTopology *topo, const Terms &t) {
    const PeriodicTopology *pt = dynamic_cast<const PeriodicTopology *>(topo);
    if (!pt)
      Report::report << Report::error << "pmx: the GPU nonbonded forces need a periodic topology with a cubic cell manager"
                     << Report::endr;
    const PBC &bc = pt->boundaryConditions;
    const Vector3D e1 = bc.e1(), e2 = bc.e2(), e3 = bc.e3();
    if (e1.c[1] != 0 || e1.c[2] != 0 || e2.c[0] != 0 || e2.c[2] != 0 || e3.c[0] != 0 || e3.c[1] != 0)
      Report::report << Report::error << "pmx: the GPU nonbonded forces need an orthogonal box" << Report::endr;
    const int n = (int)pt->atoms.size();
    charge_.resize(n);
    type_.resize(n);
    molecule_.resize(n);
    for (int i = 0; i < n; ++i) {
      charge_[i] = pt->atoms[i].scaledCharge;
      type_[i] = pt->atoms[i].type;
      molecule_[i] = pt->atoms[i].molecule;
    }
    const int nt = pt->lennardJonesParameters.size();
    lj_.resize(4 * (size_t)nt * nt);
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        const LennardJonesParameters &p = pt->lennardJonesParameters(a, b);
        double *o = &lj_[4 * ((size_t)a * nt + b)];
        o[0] = p.A;
        o[1] = p.B;
        o[2] = p.A14;
        o[3] = p.B14;
      }
    // exclusion table, in its order (ExclusionTable.h:87); the engine codes
    // FULL as 1 and MODIFIED as 2
    const std::vector<ExclusionPair> &tab = pt->exclusions.getTable();
    excl_.clear();
    for (unsigned int k = 0; k < tab.size(); ++k) {
      if (tab[k].excl == EXCLUSION_NONE) continue;
      excl_.push_back(tab[k].a1);
      excl_.push_back(tab[k].a2);
      excl_.push_back(tab[k].excl == EXCLUSION_FULL ? 1 : 2);
    }
    pmx_sys_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n = n;
    d.ntypes = nt;
    d.nexcl = (int)(excl_.size() / 3);
    d.charge = &charge_[0];
    d.type = &type_[0];
    d.molecule = &molecule_[0];
    d.lj = &lj_[0];
    d.excl = excl_.empty() ? 0 : &excl_[0];
    d.box[0] = e1.c[0];
    d.box[1] = e2.c[1];
    d.box[2] = e3.c[2];
    for (int k = 0; k < 3; ++k) d.origin[k] = bc.origin().c[k];
    d.csf = pt->coulombScalingFactor;
    d.cellsize = pt->cellManager.getCellSize();
    d.precision = where.precision;
    d.lj_switch = t.lj_switch;
    d.lj_on = t.lj_on;
    d.lj_rc = t.lj_rc;
    d.coul_kind = t.coul_kind;
    d.coul_switch = t.coul_switch;
    d.coul_on = t.coul_on;
    d.coul_rc = t.coul_rc;
    for (int k = 0; k < 3; ++k) d.pme_grid[k] = t.grid[k];
    d.pme_order = t.order;
    d.pme_accuracy = t.accuracy;
    d.pme_alpha = t.alpha;
    d.pme_terms = t.pme_terms;
    release();
    if (pmx_ctx_create(where.device, &ctx_) != PMX_OK)
      Report::report << Report::error << "pmx: no usable GPU " << where.device << Report::endr;
    if (pmx_sys_create(ctx_, &d, &sys_) != PMX_OK)
      Report::report << Report::error << "pmx: " << pmx_last_error(ctx_) << Report::endr;
    built_for_ = topo;
  }

  pmx_ctx *ctx_;
  pmx_sys *sys_;
  const GenericTopology *built_for_;
  std::vector<double> charge_, lj_;
  std::vector<int> type_, molecule_, excl_;
};

// The switching-function parameters of a OneAtomPair / OneAtomPairTwo, in
// the order of its getParameters (functor parameters are empty for
// LennardJonesForce and CoulombForce).
template <class Sw, class F>
void addLine(const std::vector<Real> &v, size_t &k, Terms &t) {
  const int code = SwitchCode<Sw>::value;
  double on = 0, rc;
  if (SwitchCode<Sw>::nparams == 2) on = v[k++];
  rc = v[k++];
  if (ForceCode<F>::lj) {
    t.lj_switch = code;
    t.lj_on = on;
    t.lj_rc = rc;
  } else {
    t.coul_kind = PMX_COUL_CUTOFF;
    t.coul_switch = code;
    t.coul_on = on;
    t.coul_rc = rc;
  }
}

template <class TOneAtomPair>
struct PairTerms;
template <class Sw, class F>
struct PairTerms<OneAtomPair<PBC, Sw, F> > {
  static Terms of(const std::vector<Real> &v) {
    Terms t;
    size_t k = 0;
    addLine<Sw, F>(v, k, t);
    return t;
  }
};
template <class SwA, class FA, class SwB, class FB>
struct PairTerms<OneAtomPairTwo<PBC, SwA, FA, SwB, FB> > {
  static Terms of(const std::vector<Real> &v) {
    Terms t;
    size_t k = 0;
    addLine<SwA, FA>(v, k, t);
    addLine<SwB, FB>(v, k, t);
    return t;
  }
};

// NonbondedCutoffSystemForce<CubicCellManager, TOneAtomPair> on the GPU
// (NonbondedCutoffSystemForce.h:45-77, NonbondedCutoffForce.h:60-91).
template <class TOneAtomPair>
class GpuNonbondedCutoffSystemForce : public NonbondedCutoffSystemForce<CCM, TOneAtomPair> {
  typedef NonbondedCutoffSystemForce<CCM, TOneAtomPair> Base;

 public:
  explicit GpuNonbondedCutoffSystemForce(const Placement &w = Placement()) : Base() { dev_.where = w; }
  GpuNonbondedCutoffSystemForce(Real cutoff, TOneAtomPair oneAtomPair, const Placement &w) : Base(cutoff, oneAtomPair) {
    dev_.where = w;
  }

  // the engine's descriptor fields of this force's lines
  Terms terms() const { return PairTerms<TOneAtomPair>::of(parameterValues(*this)); }
  virtual void evaluate(const GenericTopology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                        ScalarStructure *energies) {
    dev_.evaluate(topo, positions, forces, energies, terms());
  }
  virtual void parallelEvaluate(const GenericTopology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                                ScalarStructure *energies) {
    if (Parallel::next()) evaluate(topo, positions, forces, energies);
  }
  virtual unsigned int numberOfBlocks(const GenericTopology *, const Vector3DBlock *) { return 1; }
  virtual void uncache() { dev_.uncache(); }

 private:
  virtual Force *doMake(const std::vector<Value> &values) const {
    std::vector<Value> pairValues(values.begin(), values.end() - 1);
    return new GpuNonbondedCutoffSystemForce(values[values.size() - 1], TOneAtomPair::make(pairValues), dev_.where);
  }
  Device dev_;
};

// NonbondedPMEwaldSystemForce<PBC, CCM, real, reciprocal, correction,
// BSpline, TSwitch> on the GPU (NonbondedPMEwaldSystemForce.h:245-319,
// 985-1061).
template <bool real, bool reciprocal, bool correction, class TSwitch>
class GpuPMEwaldSystemForce : public NonbondedPMEwaldSystemForce<PBC, CCM, real, reciprocal, correction, BSpline, TSwitch> {
  typedef NonbondedPMEwaldSystemForce<PBC, CCM, real, reciprocal, correction, BSpline, TSwitch> Base;

 public:
  explicit GpuPMEwaldSystemForce(const Placement &w = Placement()) : Base() { dev_.where = w; }
  GpuPMEwaldSystemForce(unsigned int nx, unsigned int ny, unsigned int nz, unsigned int order, Real cutoff,
                        Real accuracy, Real alpha, const Placement &w)
      : Base(nx, ny, nz, order, cutoff, accuracy, alpha, 3.0) {
    dev_.where = w;
  }

  virtual void evaluate(const GenericTopology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                        ScalarStructure *energies) {
    dev_.evaluate(topo, positions, forces, energies, terms());
  }
  // the engine's descriptor fields: -gridsize nx ny nz, -cutoff, -order,
  // -accuracy, -alpha, the terms of the template and its switch
  Terms terms() const {
    const std::vector<Real> v = parameterValues(*this);
    Terms t;
    t.coul_kind = PMX_COUL_PME;
    t.coul_switch = SwitchCode<TSwitch>::value;
    for (int k = 0; k < 3; ++k) t.grid[k] = (int)v[k];
    t.coul_rc = v[3];
    t.order = (int)v[4];
    t.accuracy = v[5];
    t.alpha = v[6];
    t.pme_terms = (real ? PMX_PME_REAL : 0) | (reciprocal ? PMX_PME_RECIP : 0) | (correction ? PMX_PME_CORR : 0);
    return t;
  }
  virtual void parallelEvaluate(const GenericTopology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                                ScalarStructure *energies) {
    if (Parallel::next()) evaluate(topo, positions, forces, energies);
  }
  virtual unsigned int numberOfBlocks(const GenericTopology *, const Vector3DBlock *) { return 1; }
  virtual void uncache() { dev_.uncache(); }

 private:
  // the reference's checks and messages (NonbondedPMEwaldSystemForce.h:1024-1061)
  virtual Force *doMake(const std::vector<Value> &values) const {
    const int nx = values[0], ny = values[1], nz = values[2], order = values[4];
    const Real cutoff = values[3], accuracy = values[5], alpha = values[6];
    const std::string kw = this->getKeyword();
    std::string err;
    if (order < 2 || (order % 2) != 0 || !values[4].valid())
      err += kw + " order (=" + values[4].getString() + ") > 1 and even.";
    if (!values[0].valid() || !values[1].valid() || !values[2].valid() || nx < order || ny < order || nz < order)
      err += kw + " force: " + values[4].getString() + " <= nx (=" + values[0].getString() + "), " +
             values[4].getString() + " <= ny (=" + values[1].getString() + "), " + values[4].getString() +
             " <= nz (=" + values[2].getString() + ").";
    if (cutoff <= 0.0 || !values[3].valid()) err += kw + " cutoff (=" + values[3].getString() + ") > 0.";
    if (accuracy <= 0.0 || !values[5].valid()) err += kw + " accuracy (=" + values[5].getString() + ") > 0.";
    if (!values[6].valid()) err += " alpha '" + values[6].getString() + "' not valid.";
    if (!err.empty()) {
      // the reference's message: the checks, then " force PMEwald :" and the
      // checks again
      err += " force " + kw + " :" + err;
      Report::report << Report::error << err << Report::endr;
      return 0;
    }
    return new GpuPMEwaldSystemForce((unsigned int)nx, (unsigned int)ny, (unsigned int)nz, (unsigned int)order,
                                     cutoff, accuracy, alpha, dev_.where);
  }
  Device dev_;
};

// The descriptor fields of a force made by a factory holding these
// prototypes; false if it is not one of them.
template <class G>
bool termsIf(const Force *f, Terms *t) {
  const G *g = dynamic_cast<const G *>(f);
  if (g) *t = g->terms();
  return g != 0;
}
inline bool gpuTerms(const Force *f, Terms *t) {
  typedef C1SwitchingFunction C1;
  typedef C2SwitchingFunction C2;
  typedef CutoffSwitchingFunction Cut;
  return termsIf<GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C2, LennardJonesForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C1, LennardJonesForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, Cut, LennardJonesForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C2, CoulombForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C1, CoulombForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, Cut, CoulombForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPairTwo<PBC, C2, LennardJonesForce, C1, CoulombForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPairTwo<PBC, C2, LennardJonesForce, C2, CoulombForce> > >(f, t) ||
         termsIf<GpuNonbondedCutoffSystemForce<OneAtomPairTwo<PBC, C2, LennardJonesForce, Cut, CoulombForce> > >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, true, true, Cut> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, true, false, Cut> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, false, true, Cut> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<false, true, true, Cut> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, false, false, Cut> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<false, true, false, Cut> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<false, false, true, Cut> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, true, true, C1> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, true, false, C1> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, false, true, C1> >(f, t) ||
         termsIf<GpuPMEwaldSystemForce<true, false, false, C1> >(f, t);
}

// Registers the GPU prototypes under the reference's ids: the periodic
// NonbondedCutoff LennardJones and Coulomb forces with the C1, C2 and Cutoff
// switches this engine serves (NonbondedCutoffForceModule.cpp:55-100), the
// LennardJones + Coulomb pair force of OneAtomPairTwo (C2 LJ with C1, C2 or
// Cutoff Coulomb), and PMEwald with every combination of -real, -reciprocal
// and -correction, BSpline interpolation, Cutoff or C1 switch.  An
// application registers these instead of the CPU prototypes with the same
// ids (ForceFactory keeps one prototype per id).
inline void registerGpuNonbondedForces(ForceFactory &f, const Placement &w = Placement()) {
  typedef C1SwitchingFunction C1;
  typedef C2SwitchingFunction C2;
  typedef CutoffSwitchingFunction Cut;
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C2, LennardJonesForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C1, LennardJonesForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, Cut, LennardJonesForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C2, CoulombForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, C1, CoulombForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPair<PBC, Cut, CoulombForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPairTwo<PBC, C2, LennardJonesForce, C1, CoulombForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPairTwo<PBC, C2, LennardJonesForce, C2, CoulombForce> >(w));
  f.reg(new GpuNonbondedCutoffSystemForce<OneAtomPairTwo<PBC, C2, LennardJonesForce, Cut, CoulombForce> >(w));
  f.reg(new GpuPMEwaldSystemForce<true, true, true, Cut>(w));
  f.reg(new GpuPMEwaldSystemForce<true, true, false, Cut>(w));
  f.reg(new GpuPMEwaldSystemForce<true, false, true, Cut>(w));
  f.reg(new GpuPMEwaldSystemForce<false, true, true, Cut>(w));
  f.reg(new GpuPMEwaldSystemForce<true, false, false, Cut>(w));
  f.reg(new GpuPMEwaldSystemForce<false, true, false, Cut>(w));
  f.reg(new GpuPMEwaldSystemForce<false, false, true, Cut>(w));
  f.reg(new GpuPMEwaldSystemForce<true, true, true, C1>(w));
  f.reg(new GpuPMEwaldSystemForce<true, true, false, C1>(w));
  f.reg(new GpuPMEwaldSystemForce<true, false, true, C1>(w));
  f.reg(new GpuPMEwaldSystemForce<true, false, false, C1>(w));
}

}  // namespace pmxgpu
}  // namespace ProtoMol

#endif
