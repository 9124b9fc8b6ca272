/* CPU oracle for the ProtoMol nonbonded hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch FP64 restatement of the reference algorithms (cell binning,
 * periodic cell-pair enumeration, LJ/Coulomb pair functors with switching,
 * and the PME real / reciprocal / correction terms).  It is the checker that
 * the HIP path is compared against; it is never linked into the product
 * library (libpmx.so), which fails loudly when its HIP kernels are missing.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every entry point returns 0 on success, a negative code on bad input.
 */
#ifndef PM_ORACLE_H
#define PM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n;                  /* atoms */
  int ntypes;             /* LJ types */
  int nexcl;              /* exclusion entries */
  const double *pos;      /* [3n] AoS, as ProtoMol's Vector3DBlock */
  const double *charge;   /* [n] scaled charges */
  const int *type;        /* [n] */
  const int *molecule;    /* [n] */
  const double *lj;       /* [ntypes*ntypes*4] {A, B, A14, B14} */
  const int *excl;        /* [3*nexcl] a1 a2 cls (1 FULL, 2 MODIFIED) */
  double box[3];          /* orthogonal edge lengths */
  double origin[3];
  double csf;             /* coulombScalingFactor */
} pmo_system;

enum { PMO_SW_NONE = 0, PMO_SW_CUTOFF = 1, PMO_SW_C1 = 2, PMO_SW_C2 = 3 };
enum { PMO_COUL_NONE = 0, PMO_COUL_CUTOFF = 1, PMO_COUL_PME = 2 };

typedef struct {
  double cellsize;
  int lj_switch;          /* PMO_SW_* ; NONE disables LJ */
  double lj_on, lj_rc;
  int coul_kind;          /* PMO_COUL_* */
  int coul_switch;        /* PMO_SW_CUTOFF / C1 / C2 (PME: CUTOFF or C1) */
  double coul_on, coul_rc;
  int pme_grid[3];
  int pme_order;
  double pme_accuracy;
  double pme_alpha;       /* <= 0: bisection from accuracy and rc */
  int pme_real, pme_recip, pme_corr;
} pmo_spec;

/* energies[0] = Coulomb slot, energies[1] = Lennard-Jones slot.
 * forces [3n] are ACCUMULATED (+=), as SystemForce::evaluate does.
 * nthreads <= 1: serial in the reference's visit order. */
int pmo_evaluate(const pmo_system *s, const pmo_spec *spec, double *forces, double *energies,
                 int nthreads);

/* Detailed PME terms (energies only): out[0..5] = real, reciprocal,
 * intra-molecular correction, point self, charged system, alpha. */
int pmo_pme_terms(const pmo_system *s, const pmo_spec *spec, double *out);

/* Ewald splitting parameter chosen like the reference when alpha <= 0. */
double pmo_pme_alpha(double rc, double accuracy);

/* Cell binning: cell index (x*ny + y)*nz + z per atom and the grid dims. */
int pmo_cells(const pmo_system *s, double cellsize, int *cell_of_atom, int *dims);

/* In-cutoff non-FULL-excluded pairs in the reference's visit order
 * (pairs[2k], pairs[2k+1] = i, j).  *npairs receives the count even when it
 * exceeds cap (then only cap pairs are written). */
int pmo_pairs(const pmo_system *s, double cellsize, double rc, int *pairs, long cap, long *npairs);

/* Brute-force O(N^2) minimal-image pair set (i < j), same predicate. */
int pmo_pairs_bruteforce(const pmo_system *s, double rc, int *pairs, long cap, long *npairs);

/* Classic Ewald sum with exact erfc: real space over all images within
 * rc_real, reciprocal over |m| <= kmax per axis; exclusion (FULL pairs) and
 * self corrections applied.  Returns Coulomb energy and accumulates forces. */
int pmo_ewald_direct(const pmo_system *s, double alpha, double rc_real, int kmax, double *forces,
                     double *energy);

/* B-spline weights (order p) at fractional offset w: theta[p], dtheta[p]. */
int pmo_bspline(int order, double w, double *theta, double *dtheta);

/* poly5 approximation used by the PME real term. */
double pmo_poly5(double x);

/* In-place unnormalised 3-D complex DFT on [nx][ny][nz] (z fastest),
 * interleaved re/im; sign = +1 or -1 in the exponent. */
int pmo_fft3d(double *data, int nx, int ny, int nz, int sign);

#ifdef __cplusplus
}
#endif
#endif
