// Host-side C++ mirror of ProtoMol's force boundary over the pmx C ABI.
//
// ProtoMol selects forces from .conf lines through ForceFactory::make
// (protomol/src/protomol/factory/ForceFactory.cpp:21-200) and evaluates them
// through SystemForce::evaluate (protomol/src/protomol/force/system/
// SystemForce.h:32-35), accumulating into a Vector3DBlock of forces and the
// COULOMB / LENNARDJONES slots of a ScalarStructure.  The types here keep
// those names, argument meanings and error behaviour (config errors: make()
// returns nullptr after reporting; runtime errors: report and abort) so that
// a ProtoMol build can register GpuNonbondedSystemForce under the same
// keywords as the CPU forces it replaces.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pmx.h"

namespace pmx_host {

// Vector3DBlock: contiguous AoS FP64 coordinates (protomol/type/Vector3DBlock.h:32).
struct Vector3DBlock {
  std::vector<double> c;
  explicit Vector3DBlock(size_t n = 0) : c(3 * n, 0.0) {}
  size_t size() const { return c.size() / 3; }
  void zero() { std::fill(c.begin(), c.end(), 0.0); }
  double *data() { return c.data(); }
  const double *data() const { return c.data(); }
};

// ScalarStructure energy slots (protomol/type/ScalarStructure.h:26-34).
struct ScalarStructure {
  enum Index { COULOMB = 0, LENNARDJONES, BOND, ANGLE, DIHEDRAL, IMPROPER, OTHER, LAST };
  double v[LAST] = {0};
  void clear() {
    for (double &x : v) x = 0.0;
  }
  double &operator[](Index i) { return v[i]; }
  double operator[](Index i) const { return v[i]; }
};

// The parts of ProtoMol's periodic topology the nonbonded path reads
// (protomol/topology/GenericTopology.h, Topology.h, ExclusionTable.h).
struct Topology {
  std::vector<double> scaledCharge;   // q * SQRTCOULOMBCONSTANT
  std::vector<int> type;              // LJ type
  std::vector<int> molecule;
  int ntypes = 0;
  std::vector<double> lj;             // [ntypes*ntypes*4] {A, B, A14, B14}
  std::vector<int> exclusions;        // [3*m] a1 a2 class (1 FULL, 2 MODIFIED)
  double cellBasis[3] = {0, 0, 0};    // orthogonal cellBasisVector1..3
  double cellOrigin[3] = {0, 0, 0};
  double coulombScalingFactor = 1.0;
  double cellsize = 0.0;              // cellManager Cubic -cellsize
  size_t size() const { return scaledCharge.size(); }
};

// Report-and-abort convention of the reference (protomol/base/Report.cpp:248-253).
[[noreturn]] inline void reportErrorAndAbort(const std::string &msg) {
  std::fprintf(stderr, "Fatal Error: %s\n", msg.c_str());
  std::abort();
}

// Work-package distribution of ProtoMol's Parallel as a force sees it
// (protomol/src/protomol/parallel/Parallel.cpp:911-1040): SEQUENTIAL, where
// next() is always true, and STATIC, where package k of a force group goes to
// rank k mod size.  (The reference's dynamic MASTER/SLAVE mode hands packages
// out on request over MPI; a one-package force is evaluated by exactly one
// rank there too.)
class Parallel {
 public:
  enum Mode { SEQUENTIAL, STATIC };
  static void init(Mode mode, int rank, int size) {
    mode_ = mode;
    rank_ = rank;
    size_ = size;
    next_ = 0;
  }
  static bool isParallel() { return mode_ != SEQUENTIAL && size_ > 1; }
  // Parallel::distribute at the start of a force group's evaluation
  static void resetNext() { next_ = 0; }
  static bool next() {
    if (mode_ == SEQUENTIAL) return true;
    const bool mine = next_ == rank_;
    next_ = (next_ + 1) % size_;
    return mine;
  }

 private:
  static inline Mode mode_ = SEQUENTIAL;
  static inline int rank_ = 0, size_ = 1, next_ = 0;
};

class SystemForce {
 public:
  virtual ~SystemForce() = default;
  virtual void evaluate(const Topology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                        ScalarStructure *energies) = 0;
  // Parallel evaluation (SystemForce.h:107-112, NonbondedCutoffSystemForce.h:
  // 57-77): a force split into numberOfBlocks packages evaluates the packages
  // Parallel::next() gives this rank; the ranks' forces and energies are then
  // summed.  The GPU force is one package, evaluated whole by one rank.
  virtual unsigned numberOfBlocks(const Topology *, const Vector3DBlock *) { return 1; }
  virtual void parallelEvaluate(const Topology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                                ScalarStructure *energies) {
    if (Parallel::next()) evaluate(topo, positions, forces, energies);
  }
  // ForceGroup.cpp:65-69 calls these around evaluate; the nonbonded forces
  // have nothing to do there
  virtual void preProcess(const Topology *, const Vector3DBlock *) {}
  virtual void postProcess(const Topology *, ScalarStructure *, Vector3DBlock *) {}
  virtual void uncache() {}
  virtual std::string getKeyword() const = 0;
  virtual std::string getIdNoAlias() const = 0;
};

// Parsed form of the force lines this engine serves:
//   force LennardJones [Coulomb] -algorithm NonbondedCutoff -switchingFunction X ... -cutoff rc
//   force Coulomb -algorithm PMEwald [-real] [-reciprocal] [-correction]
//         -interpolation BSpline [-switchingFunction C1] -gridsize nx ny nz -cutoff rc
//         [-order p] [-accuracy e] [-alpha a]
struct ForceLine {
  bool lj = false, coulomb = false, pme = false;
  int lj_switch = PMX_SW_NONE, coul_switch = PMX_SW_NONE;
  double lj_on = 0, lj_rc = 0, coul_on = 0, coul_rc = 0;
  int grid[3] = {0, 0, 0};
  int order = 4;
  double accuracy = 1e-6, alpha = -1.0;
  int terms = 0;
};

// Parses one force line (without the leading "force"); returns false and
// fills err on a grammar error, as ForceFactory::make reports them.
bool parseForceLine(const std::string &line, ForceLine *out, std::string *err);

// GPU implementation of the NonbondedCutoff LJ/Coulomb and PMEwald system
// forces.  Several lines of one force group fuse into one device system.
class GpuNonbondedSystemForce : public SystemForce {
 public:
  GpuNonbondedSystemForce(const std::vector<ForceLine> &lines, int precision, int device = 0);
  ~GpuNonbondedSystemForce() override;
  void evaluate(const Topology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                ScalarStructure *energies) override;
  void uncache() override;
  std::string getKeyword() const override { return keyword_; }
  std::string getIdNoAlias() const override { return id_; }
  double alpha() const;

 private:
  void build(const Topology *topo);
  std::vector<ForceLine> lines_;
  int precision_, device_;
  pmx_ctx *ctx_ = nullptr;
  pmx_sys *sys_ = nullptr;
  const Topology *built_for_ = nullptr;
  std::string keyword_, id_;
};

// ForceFactory-shaped registry for the keywords this engine serves.
class ForceFactory {
 public:
  // make("LennardJones Coulomb -algorithm NonbondedCutoff ...", precision)
  // nullptr on a configuration error (message in lastError()).
  std::unique_ptr<SystemForce> make(const std::vector<std::string> &lines, int precision, int device = 0);
  const std::string &lastError() const { return err_; }

 private:
  std::string err_;
};

}  // namespace pmx_host
