/* pmx_serve -- request loop over the pmx engine on one GPU.
 *
 * No counterpart in the reference: ProtoMol runs one simulation per process
 * (ProtoMolApp, protomol/src/protomol/ProtoMolApp.cpp).  The loop accepts
 * force-evaluation requests (SystemForce::evaluate of one configuration) and
 * N-step trajectory requests (LeapfrogIntegrator / LangevinLeapfrogIntegrator
 * runs) for independent systems of any size.
 *
 * `slots` worker threads each own one engine context (two HIP streams), so up
 * to `slots` requests run on the device at once and their kernels overlap.
 *
 * Two ways to submit work:
 *  - one-shot requests (pmx_request.session == 0) carry their whole system;
 *    the system is built on the device, used and released;
 *  - sessions (pmx_server_open ... pmx_server_close) keep one system resident
 *    in HBM between requests: positions, velocities, the cell sort and pair
 *    list, PME tables and grids stay on the device, and each trajectory
 *    request continues from where the previous one stopped (the resume
 *    counterpart of ProtoMol's checkpoint module,
 *    protomol/src/protomol/module/CheckpointModule.cpp:13-51).
 *
 * Admission by HBM footprint: every one-shot request and every session has
 * an estimated device footprint (pmx_estimate_bytes).  A one-shot request or
 * a session open, taken in submission order, starts when a slot is free and
 * its footprint fits in the budget beside the running requests and the open
 * sessions (a single item larger than the budget runs when nothing else is
 * resident).  A session holds its footprint from open to close; its requests
 * need no further admission and run in order on the slot that opened it.
 * Device blocks of finished systems are kept by the engine for reuse, so a
 * steady request stream does not allocate device memory.
 *
 * Continuous batching: a slot that starts a one-shot trajectory request also
 * starts the queued one-shot trajectory requests that can share its kernel
 * launches (same precision, kinds of force, LJ type count, PME settings and
 * step count) and fit in the budget, up to 64 requests and 200,000 atoms
 * beyond its own, and runs them with
 * pmx_md_run_batch (every kernel of a step launched once for all of them).
 * Every one-shot trajectory runs that way, alone or batched, so its result
 * does not depend on the requests it was batched with (a system that no batch
 * takes runs with pmx_md_run).
 *
 * All inputs are copied at submit; output pointers must stay valid until
 * pmx_server_wait returns for the ticket.
 */
#ifndef PMX_SERVE_H
#define PMX_SERVE_H

#include "pmx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmx_server pmx_server;

enum { PMX_REQ_FORCES = 0, PMX_REQ_TRAJECTORY = 1 };

typedef struct {
  int kind;                 /* PMX_REQ_FORCES or PMX_REQ_TRAJECTORY */
  const pmx_sys_desc *sys;  /* one-shot: the system (its arrays are copied); session: ignored */
  const pmx_md_desc *md;    /* one-shot trajectory: bonded terms, integrator (copied); session: ignored */
  const double *x;          /* [3n] positions (session trajectory: NULL continues from the resident state) */
  const double *v;          /* trajectory: [3n] velocities (session: with x, or NULL) */
  int nsteps;               /* trajectory: steps */
  double *out_a;            /* forces: [3n] forces; trajectory: [3n] final positions */
  double *out_v;            /* trajectory: [3n] final velocities (may be NULL) */
  double *out_energies;     /* forces: [PMX_E_COUNT]; trajectory: [PMX_MD_E_COUNT] (may be NULL) */
  long long session;        /* 0: one-shot; else a session from pmx_server_open */
} pmx_request;

int pmx_server_create(int device, int slots, unsigned long long hbm_budget, pmx_server **out);
/* Queues a request; *ticket identifies it. */
int pmx_server_submit(pmx_server *srv, const pmx_request *req, long long *ticket);
/* Blocks until the request is done; returns its pmx_status.  Times in ms
 * since its submit: when it started on a slot and when it finished.  One
 * wait per ticket: a ticket already collected, or one another thread is
 * waiting on, gives PMX_ERR_ARG. */
int pmx_server_wait(pmx_server *srv, long long ticket, double *start_ms, double *done_ms);
/* pmx_server_wait with the three times (submit, start, done) in ms since
 * pmx_server_create, so that requests can be placed on one time line. */
int pmx_server_wait_times(pmx_server *srv, long long ticket, double times[3]);
/* Opens a session: a resident system made from `sys` (and, for trajectory
 * requests, `md`; NULL for a force-only session) at positions x (and
 * velocities v with md).  Queued like a request; *session identifies it at
 * once.  If the open fails, the session's requests fail with its status. */
int pmx_server_open(pmx_server *srv, const pmx_sys_desc *sys, const pmx_md_desc *md, const double *x,
                    const double *v, long long *session);
/* Queues the release of a session after its queued requests. */
int pmx_server_close(pmx_server *srv, long long session);
/* Requests queued + running. */
int pmx_server_pending(pmx_server *srv, int *n);
/* Admission record since create: the largest sum of footprints resident at
 * once (running one-shot requests + open sessions) and the most items
 * running at once. */
int pmx_server_stats(pmx_server *srv, unsigned long long *peak_bytes, int *peak_running);
/* Waits for all requests, closes open sessions, stops the workers, frees
 * everything. */
int pmx_server_destroy(pmx_server *srv);
/* Message of the last failed request ("" if none). */
const char *pmx_server_last_error(const pmx_server *srv);

#ifdef __cplusplus
}
#endif
#endif
