// Host-side C++ mirror of ProtoMol's force boundary (see pmx_protomol.h).
#include "pmx_protomol.h"

#include <sstream>

namespace pmx_host {

namespace {

int switchKind(const std::string &name) {
  if (name == "C2") return PMX_SW_C2;
  if (name == "C1") return PMX_SW_C1;
  if (name == "Cutoff") return PMX_SW_CUTOFF;
  return -1;
}

bool num(const std::vector<std::string> &t, size_t i, double *v) {
  if (i >= t.size()) return false;
  char *end = nullptr;
  *v = std::strtod(t[i].c_str(), &end);
  return end && *end == 0;
}

}  // namespace

bool parseForceLine(const std::string &line, ForceLine *out, std::string *err) {
  std::istringstream in(line);
  std::vector<std::string> t;
  for (std::string w; in >> w;) t.push_back(w);
  ForceLine f;
  size_t i = 0;
  for (; i < t.size() && t[i][0] != '-'; ++i) {
    if (t[i] == "LennardJones")
      f.lj = true;
    else if (t[i] == "Coulomb")
      f.coulomb = true;
    else {
      *err = "unknown force '" + t[i] + "'";
      return false;
    }
  }
  std::string algorithm;
  std::vector<int> sw;                 // switching functions in order of appearance
  std::vector<double> cutoffs, switchons;
  for (; i < t.size(); ++i) {
    const std::string &k = t[i];
    double v;
    if (k == "-algorithm" && i + 1 < t.size()) {
      algorithm = t[++i];
    } else if (k == "-switchingFunction" && i + 1 < t.size()) {
      int s = switchKind(t[++i]);
      if (s < 0) {
        *err = "unknown switching function '" + t[i] + "'";
        return false;
      }
      sw.push_back(s);
    } else if (k == "-cutoff" && num(t, i + 1, &v)) {
      cutoffs.push_back(v);
      ++i;
    } else if (k == "-switchon" && num(t, i + 1, &v)) {
      switchons.push_back(v);
      ++i;
    } else if (k == "-real") {
      f.terms |= PMX_PME_REAL;
    } else if (k == "-reciprocal") {
      f.terms |= PMX_PME_RECIP;
    } else if (k == "-correction") {
      f.terms |= PMX_PME_CORR;
    } else if (k == "-interpolation" && i + 1 < t.size()) {
      if (t[++i] != "BSpline") {
        *err = "only -interpolation BSpline is served";
        return false;
      }
    } else if (k == "-gridsize" && i + 3 < t.size()) {
      for (int d = 0; d < 3; ++d) f.grid[d] = std::atoi(t[i + 1 + d].c_str());
      i += 3;
    } else if (k == "-order" && num(t, i + 1, &v)) {
      f.order = (int)v;
      ++i;
    } else if (k == "-accuracy" && num(t, i + 1, &v)) {
      f.accuracy = v;
      ++i;
    } else if (k == "-alpha" && num(t, i + 1, &v)) {
      f.alpha = v;
      ++i;
    } else {
      *err = "unexpected parameter '" + k + "'";
      return false;
    }
  }
  if (algorithm == "NonbondedCutoff") {
    if (cutoffs.empty()) {
      *err = "NonbondedCutoff needs -cutoff";
      return false;
    }
    // one switching function per force, in the order LennardJones, Coulomb
    size_t si = 0, on = 0;
    auto take = [&](int *kind, double *swon, double *rc) {
      *kind = si < sw.size() ? sw[si] : PMX_SW_CUTOFF;
      *rc = cutoffs[std::min(si, cutoffs.size() - 1)];
      *swon = (*kind == PMX_SW_C2 && on < switchons.size()) ? switchons[on++] : 0.0;
      ++si;
    };
    if (f.lj) take(&f.lj_switch, &f.lj_on, &f.lj_rc);
    if (f.coulomb) take(&f.coul_switch, &f.coul_on, &f.coul_rc);
  } else if (algorithm == "PMEwald") {
    if (f.lj || !f.coulomb) {
      *err = "PMEwald is a Coulomb algorithm";
      return false;
    }
    f.pme = true;
    f.coul_switch = sw.empty() ? PMX_SW_CUTOFF : sw[0];
    f.coul_rc = cutoffs.empty() ? 0.0 : cutoffs.back();
    if (f.terms == 0) f.terms = PMX_PME_REAL | PMX_PME_RECIP | PMX_PME_CORR;
  } else {
    *err = "algorithm '" + algorithm + "' is not served by this engine";
    return false;
  }
  *out = f;
  return true;
}

GpuNonbondedSystemForce::GpuNonbondedSystemForce(const std::vector<ForceLine> &lines, int precision, int device)
    : lines_(lines), precision_(precision), device_(device) {
  for (const ForceLine &l : lines_) {
    if (!keyword_.empty()) keyword_ += "+";
    keyword_ += l.pme ? "PMEwald" : "NonbondedCutoff";
  }
  id_ = "GPU " + keyword_;
}

GpuNonbondedSystemForce::~GpuNonbondedSystemForce() {
  if (sys_) pmx_sys_destroy(sys_);
  if (ctx_) pmx_ctx_destroy(ctx_);
}

void GpuNonbondedSystemForce::uncache() {
  if (sys_) pmx_sys_uncache(sys_);
}

double GpuNonbondedSystemForce::alpha() const {
  double a = 0;
  if (sys_) pmx_sys_alpha(sys_, &a);
  return a;
}

void GpuNonbondedSystemForce::build(const Topology *topo) {
  if (!ctx_) {
    int rc = pmx_ctx_create(device_, &ctx_);
    if (rc) reportErrorAndAbort("pmx_ctx_create failed (" + std::to_string(rc) + ")");
  }
  if (sys_) pmx_sys_destroy(sys_);
  sys_ = nullptr;
  pmx_sys_desc d{};
  d.n = (int)topo->size();
  d.ntypes = topo->ntypes;
  d.nexcl = (int)(topo->exclusions.size() / 3);
  d.charge = topo->scaledCharge.data();
  d.type = topo->type.data();
  d.molecule = topo->molecule.data();
  d.lj = topo->lj.data();
  d.excl = topo->exclusions.data();
  for (int k = 0; k < 3; ++k) {
    d.box[k] = topo->cellBasis[k];
    d.origin[k] = topo->cellOrigin[k];
  }
  d.csf = topo->coulombScalingFactor;
  d.cellsize = topo->cellsize;
  d.precision = precision_;
  for (const ForceLine &l : lines_) {
    if (l.lj) {
      d.lj_switch = l.lj_switch;
      d.lj_on = l.lj_on;
      d.lj_rc = l.lj_rc;
    }
    if (l.coulomb) {
      d.coul_kind = l.pme ? PMX_COUL_PME : PMX_COUL_CUTOFF;
      d.coul_switch = l.coul_switch;
      d.coul_on = l.coul_on;
      d.coul_rc = l.coul_rc;
      if (l.pme) {
        for (int k = 0; k < 3; ++k) d.pme_grid[k] = l.grid[k];
        d.pme_order = l.order;
        d.pme_accuracy = l.accuracy;
        d.pme_alpha = l.alpha;
        d.pme_terms = l.terms;
      }
    }
  }
  int rc = pmx_sys_create(ctx_, &d, &sys_);
  if (rc) reportErrorAndAbort(std::string("pmx_sys_create: ") + pmx_last_error(ctx_));
  built_for_ = topo;
}

void GpuNonbondedSystemForce::evaluate(const Topology *topo, const Vector3DBlock *positions, Vector3DBlock *forces,
                                       ScalarStructure *energies) {
  if (!sys_ || built_for_ != topo) build(topo);
  double e[PMX_E_COUNT] = {0, 0};
  int rc = pmx_sys_evaluate(sys_, positions->data(), forces->data(), e, 0);
  if (rc) reportErrorAndAbort(std::string("pmx_sys_evaluate: ") + pmx_last_error(ctx_));
  (*energies)[ScalarStructure::COULOMB] += e[PMX_E_COULOMB];
  (*energies)[ScalarStructure::LENNARDJONES] += e[PMX_E_LENNARDJONES];
}

std::unique_ptr<SystemForce> ForceFactory::make(const std::vector<std::string> &lines, int precision, int device) {
  std::vector<ForceLine> parsed;
  bool lj = false, coul = false;
  for (const std::string &ln : lines) {
    ForceLine f;
    if (!parseForceLine(ln, &f, &err_)) return nullptr;
    if ((f.lj && lj) || (f.coulomb && coul)) {
      err_ = "more than one LennardJones or Coulomb line for one device system";
      return nullptr;
    }
    lj |= f.lj;
    coul |= f.coulomb;
    parsed.push_back(f);
  }
  if (parsed.empty()) {
    err_ = "no force line";
    return nullptr;
  }
  return std::unique_ptr<SystemForce>(new GpuNonbondedSystemForce(parsed, precision, device));
}

}  // namespace pmx_host
