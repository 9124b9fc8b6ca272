/* pmx_serve -- request loop over the pmx engine on one GPU.
 *
 * No counterpart in the reference: ProtoMol runs one simulation per process
 * (ProtoMolApp, protomol/src/protomol/ProtoMolApp.cpp).  The loop accepts
 * force-evaluation requests (SystemForce::evaluate of one configuration) and
 * N-step trajectory requests (LeapfrogIntegrator / LangevinLeapfrogIntegrator
 * runs) for independent systems of any size.
 *
 * `slots` worker threads each own one engine context (one HIP stream), so up
 * to `slots` requests run on the device at once and their kernels overlap.
 * A queued request starts when a slot is free and its estimated device
 * footprint fits in the HBM budget beside the running ones (a request larger
 * than the budget runs alone).  A system stays resident in HBM (coordinates,
 * lists, PME grids) for the whole request.
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
  const pmx_sys_desc *sys;  /* the system (its arrays are copied) */
  const pmx_md_desc *md;    /* trajectory: bonded terms, integrator (copied) */
  const double *x;          /* [3n] positions */
  const double *v;          /* trajectory: [3n] velocities */
  int nsteps;               /* trajectory: steps */
  double *out_a;            /* forces: [3n] forces; trajectory: [3n] final positions */
  double *out_v;            /* trajectory: [3n] final velocities (may be NULL) */
  double *out_energies;     /* forces: [PMX_E_COUNT]; trajectory: [PMX_MD_E_COUNT] (may be NULL) */
} pmx_request;

int pmx_server_create(int device, int slots, unsigned long long hbm_budget, pmx_server **out);
/* Queues a request; *ticket identifies it. */
int pmx_server_submit(pmx_server *srv, const pmx_request *req, long long *ticket);
/* Blocks until the request is done; returns its pmx_status.  Times in ms
 * since its submit: when it started on a slot and when it finished.  One
 * wait per ticket: a ticket already collected, or one another thread is
 * waiting on, gives PMX_ERR_ARG. */
int pmx_server_wait(pmx_server *srv, long long ticket, double *start_ms, double *done_ms);
/* Requests queued + running. */
int pmx_server_pending(pmx_server *srv, int *n);
/* Waits for all requests, stops the workers, frees everything. */
int pmx_server_destroy(pmx_server *srv);
/* Message of the last failed request ("" if none). */
const char *pmx_server_last_error(const pmx_server *srv);

#ifdef __cplusplus
}
#endif
#endif
