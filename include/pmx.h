/* pmx -- MI355X (gfx950) nonbonded force engine for ProtoMol.
 *
 * C ABI of the drop-in boundary.  Plain pointers and sizes only; nothing
 * throws across it; every call returns PMX_OK (0) or a negative pmx_status.
 *
 * It replaces, for one ProtoMol topology on one GPU, the work done by
 *   SystemForce::evaluate(const GenericTopology*, const Vector3DBlock*,
 *                         Vector3DBlock*, ScalarStructure*)
 *     (protomol/src/protomol/force/system/SystemForce.h:32-35)
 * for the force lines
 *   force LennardJones [Coulomb] -algorithm NonbondedCutoff -switchingFunction ...
 *     (protomol/src/protomol/force/nonbonded/NonbondedCutoffSystemForce.h:45-55,
 *      registered in protomol/src/protomol/module/NonbondedCutoffForceModule.cpp:40-201)
 *   force Coulomb -algorithm PMEwald -real -reciprocal -correction -interpolation BSpline ...
 *     (mdl/interface/force/nonbonded/NonbondedPMEwaldSystemForce.h:245-319)
 * When both lines are present they share one device system and one pair pass.
 *
 * Ownership: the caller keeps the host arrays; the library owns device
 * buffers (positions, cell/cluster lists, PME grids) which stay resident in
 * HBM between calls.  A pmx_ctx (and every pmx_sys made from it) is used by
 * one host thread at a time.
 */
#ifndef PMX_H
#define PMX_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmx_ctx pmx_ctx;
typedef struct pmx_sys pmx_sys;

typedef enum {
  PMX_OK = 0,
  PMX_ERR_ARG = -1,         /* invalid argument or descriptor */
  PMX_ERR_OOM = -2,         /* device allocation failed */
  PMX_ERR_HIP = -3,         /* HIP runtime error */
  PMX_ERR_UNSUPPORTED = -4, /* valid ProtoMol input outside this engine (e.g. non-orthogonal box) */
  PMX_ERR_NODEV = -5,       /* no GPU or kernels missing for this device */
  PMX_ERR_OVERFLOW = -6     /* internal list capacity exceeded (retry is automatic; reported if persistent) */
} pmx_status;

typedef enum { PMX_FP64 = 0, PMX_MIXED = 1 } pmx_precision;
/* Switching functions: CutoffSwitchingFunction, C1SwitchingFunction,
 * C2SwitchingFunction (protomol/src/protomol/switch/). */
typedef enum { PMX_SW_NONE = 0, PMX_SW_CUTOFF = 1, PMX_SW_C1 = 2, PMX_SW_C2 = 3 } pmx_switch;
typedef enum { PMX_COUL_NONE = 0, PMX_COUL_CUTOFF = 1, PMX_COUL_PME = 2 } pmx_coulomb;
/* PME term selection (-real -reciprocal -correction). */
enum { PMX_PME_REAL = 1, PMX_PME_RECIP = 2, PMX_PME_CORR = 4 };
/* Energy slots written by pmx_sys_evaluate: the first two entries of
 * ScalarStructure (protomol/src/protomol/type/ScalarStructure.h:26-29). */
enum { PMX_E_COULOMB = 0, PMX_E_LENNARDJONES = 1, PMX_E_COUNT = 2 };
/* evaluate flags */
enum {
  PMX_EVAL_PAIRSTATS = 1u,  /* also count in-cutoff pairs and hash the pair set */
  PMX_EVAL_TIMING = 2u      /* record per-stage device times (pmx_sys_stage_times) */
};

typedef struct {
  int n;                  /* atoms */
  int ntypes;             /* LJ types */
  int nexcl;              /* exclusion entries */
  const double *charge;   /* [n] scaled charges, q * SQRTCOULOMBCONSTANT */
  const int *type;        /* [n] LJ type */
  const int *molecule;    /* [n] molecule id (unused for pairs: exclusions are explicit) */
  const double *lj;       /* [ntypes*ntypes*4] {A, B, A14, B14} (LennardJonesParameterTable) */
  const int *excl;        /* [3*nexcl] a1 a2 class, class 1 = FULL, 2 = MODIFIED (ExclusionTable) */
  double box[3];          /* orthogonal cell basis lengths (cellBasisVector1..3) */
  double origin[3];       /* cellorigin */
  double csf;             /* coulombScalingFactor */
  double cellsize;        /* cellsize of the reference cell manager (membership queries) */
  int lj_switch;          /* pmx_switch; PMX_SW_NONE = no LennardJones line */
  double lj_on, lj_rc;    /* -switchon, -cutoff */
  int coul_kind;          /* pmx_coulomb */
  int coul_switch;        /* cutoff Coulomb: CUTOFF/C1/C2; PME real: CUTOFF or C1 */
  double coul_on, coul_rc;
  int pme_grid[3];        /* -gridsize nx ny nz */
  int pme_order;          /* -order (even, >= 2) */
  double pme_accuracy;    /* -accuracy */
  double pme_alpha;       /* -alpha (<= 0: chosen from accuracy and cutoff) */
  int pme_terms;          /* PMX_PME_* mask */
  int precision;          /* pmx_precision */
} pmx_sys_desc;

/* Context: one GPU, one HIP stream. */
int pmx_ctx_create(int device, pmx_ctx **out);
int pmx_ctx_destroy(pmx_ctx *ctx);
/* Last error message for this context ("" if none). Never NULL. */
const char *pmx_last_error(const pmx_ctx *ctx);

/* Builds the resident device system (doMake + initialize, including the
 * PME alpha choice and influence tables;
 * NonbondedPMEwaldSystemForce.h:460-595). */
int pmx_sys_create(pmx_ctx *ctx, const pmx_sys_desc *desc, pmx_sys **out);
int pmx_sys_destroy(pmx_sys *sys);
/* Drops cached tables so the next evaluate rebuilds them (Force::uncache). */
int pmx_sys_uncache(pmx_sys *sys);

/* SystemForce::evaluate: x_aos [3n] positions; f_aos [3n] is ACCUMULATED
 * (+=); energies[PMX_E_COUNT] are accumulated into the two slots.
 * Synchronous. */
int pmx_sys_evaluate(pmx_sys *sys, const double *x_aos, double *f_aos, double *energies, unsigned flags);

/* Device-resident form (positions already in HBM): positions are uploaded
 * once with pmx_sys_set_positions; pmx_sys_compute runs the force pass on
 * the device without host transfers; pmx_sys_get_forces fetches them. */
int pmx_sys_set_positions(pmx_sys *sys, const double *x_aos);
int pmx_sys_compute(pmx_sys *sys, unsigned flags);
int pmx_sys_get_forces(pmx_sys *sys, double *f_aos, double *energies);
int pmx_sys_sync(pmx_sys *sys);
/* No counterpart in the reference.  pmx_sys_compute of every system, each at
 * the positions set with pmx_sys_set_positions, with every kernel of the
 * force pass launched once for all of them (the batched launches of
 * pmx_md_run_batch, with the same conditions on the systems; flags must be 0).
 * Each system's forces and energies are those it gets in a batch of its own
 * (the mixed form to the last bit) and pmx_sys_compute's to FP32 rounding
 * (a list built in batched launches is cut into coarser runs); fetch them
 * with pmx_sys_get_forces.  For many
 * small systems at once (a request loop's force requests, ensemble force
 * calls), whose passes alone are bound by launch latency. */
int pmx_sys_compute_batch(pmx_sys *const *systems, int nsys, unsigned flags);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): pmx_sys_compute with no allocation, ever.  The force
 * pass's per-pass scratch (per-tile energy and pair-statistic partials, the
 * boundary-pair records and their counters) lives in one static device
 * array of 32 MiB per device, held until the process ends, instead of in the
 * system's buffers; a pass that would need more (about 1,000,000 atoms at
 * a 12 A list radius) is rejected with PMX_ERR_ARG.  A list that overflows
 * its capacity is reported by pmx_sys_get_forces as PMX_ERR_OVERFLOW and is
 * not regrown (pmx_sys_evaluate grows it and retries).  Also rejected:
 * PMX_EVAL_TIMING, and a system whose tables pmx_sys_uncache dropped.
 * Every call on one device shares that device's array: calls that overlap
 * on one GPU (systems of different contexts, or threads) race.  Results are
 * read with pmx_sys_get_forces as after pmx_sys_compute. */
int pmx_sys_compute_noalloc(pmx_sys *sys, unsigned flags);

/* Device bytes pmx_sys_create would allocate for desc, computed on the host
 * without allocating (kind 1 adds the device MD state of pmx_md_init, with
 * about one bond and one angle per atom); 0 for an invalid descriptor. */
unsigned long long pmx_estimate_bytes(const pmx_sys_desc *desc, int kind);
/* Device memory held by the system (all its buffers, bytes). */
int pmx_sys_device_bytes(const pmx_sys *sys, unsigned long long *bytes);

/* PME splitting parameter in use (after pmx_sys_create). */
int pmx_sys_alpha(const pmx_sys *sys, double *alpha);

/* Pair statistics of the last evaluate run with PMX_EVAL_PAIRSTATS:
 * number of non-excluded pairs with minimal-image d^2 <= max cutoff^2
 * (each unordered pair once) and a 64-bit order-independent hash of the
 * set (sum over pairs of pmx_pair_hash(min(i,j), max(i,j)), wrapping). */
int pmx_sys_pair_stats(const pmx_sys *sys, long long *npairs, unsigned long long *hash);
/* The hash used by pmx_sys_pair_stats (so callers can hash a reference set). */
unsigned long long pmx_pair_hash(int i, int j);

/* Reference cell of every atom for the CubicCellManager with the descriptor's
 * cellsize (Topology::updateCellLists binning, Topology.h:44-81), computed
 * on the device from the last positions; index (x*ny + y)*nz + z. */
int pmx_sys_cells(pmx_sys *sys, int *cell_of_atom, int *dims);

/* Device milliseconds of the last evaluate with PMX_EVAL_TIMING, in the
 * order: sort (+ MD integrate), list (+ accumulator zeroing), pair (the
 * cluster-pair kernel alone), spread, fft+convolve, gather, fixups (FP64
 * re-test of pairs at the cutoff, exclusion correction, bonded terms),
 * finalize.  ms must hold PMX_STAGES floats. */
enum { PMX_STAGES = 8 };
int pmx_sys_stage_times(const pmx_sys *sys, float *ms);
/* Per-stage device times averaged over every pmx_sys_compute call made with
 * PMX_EVAL_TIMING since pmx_sys_timing_begin (HIP events on the system's
 * stream, read after one synchronisation). avg_ms holds PMX_STAGES floats. */
int pmx_sys_timing_begin(pmx_sys *sys);
int pmx_sys_timing_end(pmx_sys *sys, float *avg_ms, int *ncalls);
/* Average device milliseconds of one launch of the cluster-pair kernel alone:
 * one full evaluation at the last positions builds the list, then the kernel
 * is launched `reps` times back to back between two HIP events on the
 * system's main stream.  Measurement only, no counterpart in the reference;
 * the forces of the last evaluation are kept. */
int pmx_sys_time_pair(pmx_sys *sys, int reps, float *ms_per_launch);
/* Launch shapes for the autotuner (no counterpart in the reference):
 * PMX_TUNE_FFT_LINES     most lines per FFT block (1..64, default 16; the
 *                        LDS limit of 24 KiB per block also applies);
 * PMX_TUNE_PAIR_BLOCKS_PER_CU  pair-kernel blocks per CU (0: as many as fit,
 *                        -k: k fewer than fit, at least 1; fewer leave room
 *                        for the PME chain beside it);
 * PMX_TUNE_PAIR_SCHEDULE 0 (default): the pair list's runs (a number fixed by
 *                        the list, a multiple of 4,096, of equal estimated
 *                        work each) are dealt to the resident waves in
 *                        contiguous shares; 1: each wave's first run is
 *                        static, later ones come one at a time from a
 *                        device-wide counter (better when fewer waves are
 *                        resident than the default launch's 4,096);
 * PMX_TUNE_SPREAD_STREAM 0 (default): the PME charge spread runs on the main
 *                        stream before the pair kernel; 1: it runs first on
 *                        the second stream, beside the pair kernel;
 * PMX_TUNE_LIST_PRUNE    1 (default): the first pair pass after each list
 *                        build drops, for the later passes, the cluster pairs
 *                        with no candidate pair inside the list radius;
 *                        0: keep them;
 * PMX_TUNE_FIXUP_STREAM  0 (default from 100,000 atoms): the exclusion
 *                        correction and the bonded terms run on the second
 *                        stream after the PME chain; 1 (default below 100,000
 *                        atoms): on the main stream, in one launch with the
 *                        boundary pairs after the pair kernel (better where
 *                        the PME chain is the longer branch);
 * PMX_TUNE_FUSED_SPREAD  1 (default): inside an MD chunk the PME charge spread
 *                        of the next force pass runs in the launch that kicks
 *                        and drifts the atoms; 0: in a launch of its own.
 * PMX_TUNE_FFT_PLANES    1 (default): on PME grids whose x-y planes have at
 *                        most 1,024 points, the y transform, the x transform
 *                        with the convolution and the inverse y transform run
 *                        as one launch with a block per plane; 0: three
 *                        launches over lines (also the only form on larger
 *                        planes).
 * The resident state is kept.  No knob changes a result: the mixed form's
 * forces and energies are the same to the last bit under every setting (the
 * runs, and the FP32 partial sums within them, depend on the list only; a
 * pruned cluster pair has no member), the FP64 form's up to the order of its
 * atomic additions.  (Knobs 1 and 5, tiles per CU and runs per wave, are
 * gone: the run count no longer depends on the launch.) */
enum {
  PMX_TUNE_FFT_LINES = 2,
  PMX_TUNE_PAIR_BLOCKS_PER_CU = 3,
  PMX_TUNE_PAIR_SCHEDULE = 4,
  PMX_TUNE_SPREAD_STREAM = 6,
  PMX_TUNE_LIST_PRUNE = 7,
  PMX_TUNE_FIXUP_STREAM = 8,
  PMX_TUNE_FUSED_SPREAD = 9,
  PMX_TUNE_FFT_PLANES = 10
};
int pmx_sys_set_tuning(pmx_sys *sys, int knob, int value);

/* ---------------------------------------------------------------------
 * Device-resident MD (the caller of the hot path: StandardIntegrator /
 * LeapfrogIntegrator, protomol/src/protomol/integrator/leapfrog/
 * LeapfrogIntegrator.cpp:89-104).  Positions, velocities and forces stay in
 * HBM for whole runs; the bonded terms of the systems in this engine
 * (harmonic bonds and angles, force/bonded/BondSystemForce.h:97-118,
 * AngleSystemForce.h:95-147) are evaluated on the device too.
 *
 * The cluster-pair list is reused between steps while no atom has moved
 * more than skin/2 since it was built; membership is still decided exactly
 * against the cutoff every step.  Steps run in chunks; if an atom crossed
 * the skin/2 bound inside a chunk, the chunk is rolled back and replayed
 * with an earlier rebuild, so every committed step used a valid list.
 */
typedef struct {
  int nbonds, nangles;
  const int *bonds;          /* [2*nbonds] atom1 atom2 */
  const double *bond_par;    /* [2*nbonds] k, r0; E = k (r - r0)^2 */
  const int *angles;         /* [3*nangles] atom1 atom2(vertex) atom3 */
  const double *angle_par;   /* [2*nangles] k, theta0 (rad); E = k (theta - theta0)^2 */
  const double *mass;        /* [n] amu */
  double timestep_fs;
  double skin;               /* pair-list buffer in A (0: rebuild every step) */
  int integrator;            /* PMX_MD_LEAPFROG or PMX_MD_LANGEVIN */
  double temperature;        /* Langevin: K */
  double gamma;              /* Langevin: friction in 1/ps (the reference's -gamma) */
  unsigned long long seed;   /* Langevin: noise stream (Philox4x32-10 key) */
  int graphs;                /* 1: each chunk of steps is captured once as a HIP graph and replayed
                                (one launch per chunk); 0: kernels are launched one by one */
} pmx_md_desc;
/* Integrators: LeapfrogIntegrator (integrator/leapfrog/LeapfrogIntegrator.cpp:
 * 33-104) and LangevinLeapfrogIntegrator (integrator/base/
 * LangevinLeapfrogIntegrator.cpp:44-71), the latter with its noise drawn from
 * a counter-based generator keyed by (seed, atom, half-kick index) instead of
 * the reference's serial drand48 stream. */
enum { PMX_MD_LEAPFROG = 0, PMX_MD_LANGEVIN = 1 };

/* Energy slots of pmx_md_get_state. */
enum { PMX_MD_E_COULOMB = 0, PMX_MD_E_LJ = 1, PMX_MD_E_BOND = 2, PMX_MD_E_ANGLE = 3, PMX_MD_E_KINETIC = 4,
       PMX_MD_E_COUNT = 5 };

int pmx_md_init(pmx_sys *sys, const pmx_md_desc *desc);
/* x, v: [3n] AoS, v in ProtoMol internal units (A per internal time unit). */
int pmx_md_set_state(pmx_sys *sys, const double *x, const double *v);
/* LeapfrogIntegrator::run(nsteps): half kick + drift, forces, (kick + drift,
 * forces) x (nsteps-1), half kick.  Returns after queueing and checking the
 * chunks (synchronous).  Two chunks are in flight at a time, so the device
 * does not wait for the host's check of a chunk. */
int pmx_md_run(pmx_sys *sys, int nsteps);
/* Batched replicas: pmx_md_run(systems[i], nsteps) for every i, with the
 * systems' chunks interleaved.  Systems on the same context share its
 * streams and one host thread; systems on different contexts run
 * concurrently on the device, one host thread per context.  Each system's
 * trajectory is the one pmx_md_run gives it alone.  No counterpart in the
 * reference, which runs one system per process (ProtoMolApp.cpp); it serves
 * many independent replicas of uneven size (Hybrid Monte Carlo proposals,
 * replica ensembles).  Returns the first error of any system. */
int pmx_md_run_many(pmx_sys *const *systems, int nsys, int nsteps);
/* Batched replicas in batched launches: pmx_md_run(systems[i], nsteps) for
 * every i, with every kernel of a step launched once for all the systems (a
 * table in device memory maps each block to its system and that system's
 * arguments).  The systems share one context, the precision, the kinds of
 * force and a PME stencil order of 4, 6 or 8, and each has at most 65,536
 * cell bins.  The list of a system is rebuilt on the device at the first step
 * at which an atom has moved more than half the skin since the last build
 * (no chunks rolled back); the host checks the lists' capacities every 64
 * steps and runs such a stretch again after growing a list that overflowed.
 * A system's trajectory does not depend on the batch it runs in (the mixed
 * form to the last bit); it may differ from pmx_md_run's in the mixed form's
 * last bits where pmx_md_run rebuilds the list at other steps (its chunk
 * starts).  No counterpart in the reference, which runs one system per
 * process (ProtoMolApp.cpp); it serves many independent small replicas, whose
 * steps alone are bound by kernel launch latency. */
int pmx_md_run_batch(pmx_sys *const *systems, int nsys, int nsteps);
/* Current positions, velocities and energies (energies[PMX_MD_E_COUNT]:
 * potential terms at x, kinetic energy of v; right after pmx_md_set_state
 * the forces are computed first).  Any pointer may be NULL. */
int pmx_md_get_state(pmx_sys *sys, double *x, double *v, double *energies);
/* The Langevin noise of one half kick, computed on the host: g[3*i..3*i+2]
 * for atoms i < n, the numbers the device draws for (seed, half kick `half`).
 * No GPU needed. */
int pmx_langevin_noise(unsigned long long seed, unsigned long long half, int n, double *g);
/* Counters: committed steps, list rebuilds of committed chunks and rolled-back
 * chunks since pmx_md_init. */
int pmx_md_stats(const pmx_sys *sys, long long *steps, long long *rebuilds, long long *rollbacks);

#ifdef __cplusplus
}
#endif
#endif
