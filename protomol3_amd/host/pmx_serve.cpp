// Request loop over the pmx engine (include/pmx_serve.h).
//
// Worker threads each own one pmx_ctx.  One-shot requests and session opens
// wait in one submission-ordered queue; a worker takes the head once the sum
// of the footprints already resident (running one-shot requests and open
// sessions) plus the head's fits in the budget, or nothing is resident.  A
// session belongs to the worker that opened it: its system lives on that
// worker's context, and its requests and its close go to that worker's own
// queue, which the worker serves before the shared one.  Requests that
// arrive while others run are admitted as soon as a slot and memory free up,
// so work of different requests overlaps on the device.
//
// Continuous batching: a worker that takes a one-shot trajectory request also
// takes the queued one-shot trajectory requests that can share its launches
// (same precision, kinds of force, LJ types, PME settings and step count)
// and fit in the budget, up to kMaxBatch requests and kMaxBatchAtoms atoms
// beyond its own, and runs them with
// pmx_md_run_batch: every kernel of a step is launched once for all of them.
// Every one-shot trajectory runs through pmx_md_run_batch, alone or not, so
// its result does not depend on what it was batched with.
#include "../../include/pmx_serve.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

// Deep copy of a system descriptor (and its MD descriptor, if any).
struct SysCopy {
  pmx_sys_desc sys{};
  std::vector<double> charge, lj, mass, bond_par, angle_par;
  std::vector<int> type, molecule, excl, bonds, angles;
  bool has_md = false;
  pmx_md_desc md{};
};

template <class T>
void copy_in(std::vector<T> &dst, const T *src, size_t n) {
  dst.assign(src, src + (src ? n : 0));
}

int copy_desc(SysCopy &c, const pmx_sys_desc &d, const pmx_md_desc *m) {
  if (d.n <= 0 || d.ntypes <= 0 || d.nexcl < 0 || !d.charge || !d.type || !d.lj) return PMX_ERR_ARG;
  c.sys = d;
  const size_t n = (size_t)d.n;
  copy_in(c.charge, d.charge, n);
  copy_in(c.type, d.type, n);
  copy_in(c.molecule, d.molecule, n);
  copy_in(c.lj, d.lj, (size_t)d.ntypes * d.ntypes * 4);
  copy_in(c.excl, d.excl, 3 * (size_t)d.nexcl);
  c.sys.charge = c.charge.data();
  c.sys.type = c.type.data();
  c.sys.molecule = d.molecule ? c.molecule.data() : nullptr;
  c.sys.lj = c.lj.data();
  c.sys.excl = c.excl.empty() ? nullptr : c.excl.data();
  if (m) {
    if (m->nbonds < 0 || m->nangles < 0 || !m->mass) return PMX_ERR_ARG;
    c.has_md = true;
    c.md = *m;
    copy_in(c.mass, m->mass, n);
    copy_in(c.bonds, m->bonds, 2 * (size_t)m->nbonds);
    copy_in(c.bond_par, m->bond_par, 2 * (size_t)m->nbonds);
    copy_in(c.angles, m->angles, 3 * (size_t)m->nangles);
    copy_in(c.angle_par, m->angle_par, 2 * (size_t)m->nangles);
    c.md.mass = c.mass.data();
    c.md.bonds = c.bonds.empty() ? nullptr : c.bonds.data();
    c.md.bond_par = c.bond_par.empty() ? nullptr : c.bond_par.data();
    c.md.angles = c.angles.empty() ? nullptr : c.angles.data();
    c.md.angle_par = c.angle_par.empty() ? nullptr : c.angle_par.data();
  }
  return PMX_OK;
}

struct Job;

struct Session {
  long long id = 0;
  unsigned long long bytes = 0;
  int owner = -1;              // worker that opened it (-1 until its open is taken)
  bool closing = false;
  int open_status = PMX_OK;
  bool has_md = false;
  int n = 0;
  std::deque<std::shared_ptr<Job>> pending;  // requests submitted before the owner was known
  pmx_sys *sys = nullptr;      // used by the owner worker only
};

enum Op { OP_ONESHOT, OP_OPEN, OP_SESSION, OP_CLOSE };

struct Job {
  long long ticket = 0;
  Op op = OP_ONESHOT;
  int kind = PMX_REQ_FORCES;
  std::shared_ptr<SysCopy> desc;  // one-shot and open
  std::shared_ptr<Session> sess;  // open, session request, close
  std::vector<double> x, v;
  int nsteps = 0;
  double *out_a = nullptr, *out_v = nullptr, *out_e = nullptr;
  unsigned long long bytes = 0;  // footprint taken at admission (one-shot, open)
  Clock::time_point t_submit, t_start, t_done;
  bool done = false;
  bool waited = false;  // a pmx_server_wait holds the ticket
  bool tracked = false; // has a ticket the caller may wait on
  int status = PMX_OK;
};

}  // namespace

struct pmx_server {
  int device = 0;
  unsigned long long budget = 0, in_use = 0, peak_bytes = 0;
  int running = 0, peak_running = 0;
  std::vector<std::thread> workers;
  std::deque<std::shared_ptr<Job>> queue;               // one-shot requests and opens
  std::vector<std::deque<std::shared_ptr<Job>>> priv;   // per worker: session requests and closes
  std::map<long long, std::shared_ptr<Job>> jobs;       // tickets not yet collected
  std::map<long long, std::shared_ptr<Session>> sessions;
  mutable std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  long long next = 1;
  Clock::time_point t0 = Clock::now();
  std::string err;
};

namespace {

// A one-shot force request (one-shot trajectories run in batches, below).
int run_oneshot(pmx_ctx *ctx, Job &j) {
  pmx_sys *sys = nullptr;
  int rc = pmx_sys_create(ctx, &j.desc->sys, &sys);
  if (rc == PMX_OK) {
    const size_t n = (size_t)j.desc->sys.n;
    std::vector<double> f(3 * n, 0.0);
    double e[PMX_E_COUNT] = {0, 0};
    rc = pmx_sys_evaluate(sys, j.x.data(), f.data(), e, 0);
    if (rc == PMX_OK) {
      if (j.out_a) std::memcpy(j.out_a, f.data(), 24 * n);
      if (j.out_e) std::memcpy(j.out_e, e, sizeof(e));
    }
  }
  if (sys) pmx_sys_destroy(sys);
  return rc;
}

int run_open(pmx_ctx *ctx, Job &j) {
  Session &ss = *j.sess;
  int rc = pmx_sys_create(ctx, &j.desc->sys, &ss.sys);
  if (rc == PMX_OK && j.desc->has_md) {
    if ((rc = pmx_md_init(ss.sys, &j.desc->md)) == PMX_OK) rc = pmx_md_set_state(ss.sys, j.x.data(), j.v.data());
  }
  if (rc != PMX_OK && ss.sys) {
    pmx_sys_destroy(ss.sys);
    ss.sys = nullptr;
  }
  ss.open_status = rc;
  return rc;
}

int run_session(Job &j) {
  Session &ss = *j.sess;
  if (ss.open_status != PMX_OK) return ss.open_status;
  const size_t n = (size_t)ss.n;
  int rc;
  if (j.kind == PMX_REQ_FORCES) {
    std::vector<double> f(3 * n, 0.0);
    double e[PMX_E_COUNT] = {0, 0};
    rc = pmx_sys_evaluate(ss.sys, j.x.data(), f.data(), e, 0);
    if (rc == PMX_OK) {
      if (j.out_a) std::memcpy(j.out_a, f.data(), 24 * n);
      if (j.out_e) std::memcpy(j.out_e, e, sizeof(e));
    }
    return rc;
  }
  double e[PMX_MD_E_COUNT] = {0, 0, 0, 0, 0};
  std::vector<double> x(3 * n), v(3 * n);
  if (!j.x.empty() && (rc = pmx_md_set_state(ss.sys, j.x.data(), j.v.data())) != PMX_OK) return rc;
  if ((rc = pmx_md_run(ss.sys, j.nsteps)) == PMX_OK && (rc = pmx_md_get_state(ss.sys, x.data(), v.data(), e)) == PMX_OK) {
    if (j.out_a) std::memcpy(j.out_a, x.data(), 24 * n);
    if (j.out_v) std::memcpy(j.out_v, v.data(), 24 * n);
    if (j.out_e) std::memcpy(j.out_e, e, sizeof(e));
  }
  return rc;
}

bool admissible(const pmx_server *srv, const Job &j) {
  return srv->in_use == 0 || srv->in_use + j.bytes <= srv->budget;
}

constexpr size_t kMaxBatch = 64;            // one-shot trajectories run in one batch, at most
constexpr long long kMaxBatchAtoms = 200000;  // atoms of the requests a batch adds to its first

bool batch_head(const Job &j) { return j.op == OP_ONESHOT && j.kind == PMX_REQ_TRAJECTORY; }

// j can join a batch headed by h (the engine takes the batch as one kernel
// case: same precision, kinds of force, LJ type count and PME settings).
bool joins(const Job &h, const Job &j) {
  if (!batch_head(j) || j.nsteps != h.nsteps) return false;
  const pmx_sys_desc &a = h.desc->sys, &b = j.desc->sys;
  return a.precision == b.precision && a.coul_kind == b.coul_kind && a.lj_switch == b.lj_switch &&
         a.coul_switch == b.coul_switch && a.ntypes == b.ntypes && a.pme_order == b.pme_order &&
         a.pme_terms == b.pme_terms;
}

// One-shot trajectories in one pmx_md_run_batch call.  A batch the engine
// does not take runs system by system: each in a batch of its own, or, for a
// system no batch takes (no PME, say), with pmx_md_run.  Sets each job's
// status.
void run_trajectories(pmx_ctx *ctx, std::vector<std::shared_ptr<Job>> &jobs, std::string &err) {
  const size_t nj = jobs.size();
  std::vector<pmx_sys *> sys(nj, nullptr), live;
  std::vector<size_t> at;
  for (size_t i = 0; i < nj; ++i) {
    Job &j = *jobs[i];
    int rc = pmx_sys_create(ctx, &j.desc->sys, &sys[i]);
    if (rc == PMX_OK && (rc = pmx_md_init(sys[i], &j.desc->md)) == PMX_OK)
      rc = pmx_md_set_state(sys[i], j.x.data(), j.v.data());
    j.status = rc;
    if (rc == PMX_OK) {
      live.push_back(sys[i]);
      at.push_back(i);
    } else {
      err = pmx_last_error(ctx);
    }
  }
  if (!live.empty()) {
    const int rc = pmx_md_run_batch(live.data(), (int)live.size(), jobs[at[0]]->nsteps);
    if (rc == PMX_ERR_UNSUPPORTED || rc == PMX_ERR_ARG) {
      for (size_t k = 0; k < live.size(); ++k) {
        int r1 = pmx_md_run_batch(&live[k], 1, jobs[at[k]]->nsteps);
        if (r1 == PMX_ERR_UNSUPPORTED || r1 == PMX_ERR_ARG) r1 = pmx_md_run(live[k], jobs[at[k]]->nsteps);
        jobs[at[k]]->status = r1;
      }
    } else {
      for (size_t k : at) jobs[k]->status = rc;
    }
  }
  for (size_t k = 0; k < live.size(); ++k) {
    Job &j = *jobs[at[k]];
    const size_t n = (size_t)j.desc->sys.n;
    double e[PMX_MD_E_COUNT] = {0, 0, 0, 0, 0};
    std::vector<double> x(3 * n), v(3 * n);
    if (j.status == PMX_OK) j.status = pmx_md_get_state(live[k], x.data(), v.data(), e);
    if (j.status == PMX_OK) {
      if (j.out_a) std::memcpy(j.out_a, x.data(), 24 * n);
      if (j.out_v) std::memcpy(j.out_v, v.data(), 24 * n);
      if (j.out_e) std::memcpy(j.out_e, e, sizeof(e));
    } else {
      err = pmx_last_error(ctx);
    }
  }
  for (pmx_sys *q : sys)
    if (q) pmx_sys_destroy(q);
}

void worker(pmx_server *srv, int w) {
  pmx_ctx *ctx = nullptr;
  const int crc = pmx_ctx_create(srv->device, &ctx);
  for (;;) {
    std::shared_ptr<Job> job;
    std::vector<std::shared_ptr<Job>> batch;  // one-shot trajectories run together
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      srv->cv_work.wait(lk, [&] {
        if (!srv->priv[w].empty()) return true;
        if (!srv->queue.empty()) return admissible(srv, *srv->queue.front());
        return srv->stop;
      });
      if (!srv->priv[w].empty()) {
        job = srv->priv[w].front();
        srv->priv[w].pop_front();
      } else if (!srv->queue.empty()) {
        job = srv->queue.front();
        srv->queue.pop_front();
        srv->in_use += job->bytes;
        srv->peak_bytes = std::max(srv->peak_bytes, srv->in_use);
        if (job->op == OP_OPEN) {
          // this worker owns the session from now on; requests that came
          // before follow the open on this worker's own queue
          job->sess->owner = w;
          for (auto &p : job->sess->pending) srv->priv[w].push_back(p);
          job->sess->pending.clear();
        }
        if (batch_head(*job)) {
          batch.push_back(job);
          // a batch's requests finish together: the atoms added are capped
          // so that small requests do not wait long behind large ones
          long long added = 0;
          for (auto it = srv->queue.begin(); it != srv->queue.end() && batch.size() < kMaxBatch;) {
            if (joins(*job, **it) && admissible(srv, **it) && added + (*it)->desc->sys.n <= kMaxBatchAtoms) {
              added += (*it)->desc->sys.n;
              srv->in_use += (*it)->bytes;
              srv->peak_bytes = std::max(srv->peak_bytes, srv->in_use);
              batch.push_back(*it);
              it = srv->queue.erase(it);
            } else {
              ++it;
            }
          }
        }
      } else {
        break;  // stop requested and nothing left for this worker
      }
      const auto now = Clock::now();
      if (batch.empty()) {
        ++srv->running;
        job->t_start = now;
      } else {
        srv->running += (int)batch.size();
        for (auto &b : batch) b->t_start = now;
      }
      srv->peak_running = std::max(srv->peak_running, srv->running);
    }
    if (!batch.empty()) {
      std::string err = crc != PMX_OK ? std::string("no device context") : std::string();
      if (crc == PMX_OK)
        run_trajectories(ctx, batch, err);
      else
        for (auto &b : batch) b->status = crc;
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        const auto now = Clock::now();
        for (auto &b : batch) {
          srv->in_use -= b->bytes;
          --srv->running;
          b->t_done = now;
          b->done = true;
          if (b->status != PMX_OK) srv->err = err;
        }
        batch.clear();
        job.reset();
      }
      srv->cv_done.notify_all();
      srv->cv_work.notify_all();
      continue;
    }
    int rc = crc;
    if (crc == PMX_OK) {
      switch (job->op) {
        case OP_ONESHOT: rc = run_oneshot(ctx, *job); break;
        case OP_OPEN: rc = run_open(ctx, *job); break;
        case OP_SESSION: rc = run_session(*job); break;
        case OP_CLOSE:
          if (job->sess->sys) pmx_sys_destroy(job->sess->sys);
          job->sess->sys = nullptr;
          rc = PMX_OK;
          break;
      }
    }
    std::string err = rc == PMX_OK ? "" : (crc != PMX_OK ? std::string("no device context") : pmx_last_error(ctx));
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      if (job->op == OP_ONESHOT) srv->in_use -= job->bytes;
      if (job->op == OP_CLOSE) {
        srv->in_use -= job->sess->bytes;
        srv->sessions.erase(job->sess->id);
      }
      --srv->running;
      job->status = rc;
      job->t_done = Clock::now();
      job->done = true;
      if (!job->tracked) job.reset();
      if (rc != PMX_OK) srv->err = err;
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
  }
  // sessions this worker still owns (never closed) go with it
  std::vector<std::shared_ptr<Session>> mine;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    for (auto &kv : srv->sessions)
      if (kv.second->owner == w) mine.push_back(kv.second);
  }
  for (auto &ss : mine)
    if (ss->sys) pmx_sys_destroy(ss->sys), ss->sys = nullptr;
  if (ctx) pmx_ctx_destroy(ctx);
}

// Queues a job under the lock: one-shot and open jobs in submission order,
// session jobs on the owner's queue (or the session's pending list).
void enqueue(pmx_server *srv, const std::shared_ptr<Job> &j) {
  if (j->op == OP_ONESHOT || j->op == OP_OPEN) {
    srv->queue.push_back(j);
  } else if (j->sess->owner >= 0) {
    srv->priv[j->sess->owner].push_back(j);
  } else {
    j->sess->pending.push_back(j);
  }
}

}  // namespace

extern "C" {

int pmx_server_create(int device, int slots, unsigned long long hbm_budget, pmx_server **out) {
  if (!out || slots < 1 || slots > 64) return PMX_ERR_ARG;
  auto *srv = new pmx_server;
  srv->device = device;
  srv->budget = hbm_budget;
  srv->priv.resize(slots);
  for (int i = 0; i < slots; ++i) srv->workers.emplace_back(worker, srv, i);
  *out = srv;
  return PMX_OK;
}

int pmx_server_submit(pmx_server *srv, const pmx_request *r, long long *ticket) {
  if (!srv || !r || !ticket) return PMX_ERR_ARG;
  if (r->kind != PMX_REQ_FORCES && r->kind != PMX_REQ_TRAJECTORY) return PMX_ERR_ARG;
  if (r->kind == PMX_REQ_TRAJECTORY && r->nsteps < 0) return PMX_ERR_ARG;
  auto j = std::make_shared<Job>();
  j->kind = r->kind;
  j->nsteps = r->nsteps;
  j->out_a = r->out_a;
  j->out_v = r->out_v;
  j->out_e = r->out_energies;
  j->tracked = true;
  std::shared_ptr<Session> ss;
  size_t n;
  if (r->session == 0) {
    if (!r->sys || !r->x || (r->kind == PMX_REQ_TRAJECTORY && (!r->md || !r->v))) return PMX_ERR_ARG;
    j->op = OP_ONESHOT;
    j->desc = std::make_shared<SysCopy>();
    if (int rc = copy_desc(*j->desc, *r->sys, r->kind == PMX_REQ_TRAJECTORY ? r->md : nullptr)) return rc;
    n = (size_t)r->sys->n;
    j->bytes = pmx_estimate_bytes(&j->desc->sys, r->kind);
    if (j->bytes == 0) return PMX_ERR_ARG;  // descriptor the engine rejects
  } else {
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->sessions.find(r->session);
    if (it == srv->sessions.end() || it->second->closing) return PMX_ERR_ARG;
    ss = it->second;
    n = (size_t)ss->n;
    // force requests need positions and a session without an integrator;
    // trajectory requests need one
    if (r->kind == PMX_REQ_FORCES && (!r->x || ss->has_md)) return PMX_ERR_ARG;
    if (r->kind == PMX_REQ_TRAJECTORY && (!ss->has_md || (r->x && !r->v))) return PMX_ERR_ARG;
    j->op = OP_SESSION;
    j->sess = ss;
  }
  if (r->x) copy_in(j->x, r->x, 3 * n);
  if (r->kind == PMX_REQ_TRAJECTORY && r->x) copy_in(j->v, r->v, 3 * n);
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (ss && ss->closing) return PMX_ERR_ARG;
    j->ticket = srv->next++;
    srv->jobs[j->ticket] = j;
    enqueue(srv, j);
  }
  srv->cv_work.notify_all();
  *ticket = j->ticket;
  return PMX_OK;
}

int pmx_server_open(pmx_server *srv, const pmx_sys_desc *sys, const pmx_md_desc *md, const double *x,
                    const double *v, long long *session) {
  if (!srv || !sys || !session || (md && (!x || !v))) return PMX_ERR_ARG;
  auto j = std::make_shared<Job>();
  j->op = OP_OPEN;
  j->desc = std::make_shared<SysCopy>();
  if (int rc = copy_desc(*j->desc, *sys, md)) return rc;
  const size_t n = (size_t)sys->n;
  if (md) {
    copy_in(j->x, x, 3 * n);
    copy_in(j->v, v, 3 * n);
  }
  j->bytes = pmx_estimate_bytes(&j->desc->sys, md ? 1 : 0);
  if (j->bytes == 0) return PMX_ERR_ARG;
  auto ss = std::make_shared<Session>();
  ss->bytes = j->bytes;
  ss->has_md = md != nullptr;
  ss->n = sys->n;
  j->sess = ss;
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    ss->id = srv->next++;
    j->ticket = ss->id;
    srv->sessions[ss->id] = ss;
    enqueue(srv, j);
  }
  srv->cv_work.notify_all();
  *session = ss->id;
  return PMX_OK;
}

int pmx_server_close(pmx_server *srv, long long session) {
  if (!srv) return PMX_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->sessions.find(session);
    if (it == srv->sessions.end() || it->second->closing) return PMX_ERR_ARG;
    auto j = std::make_shared<Job>();
    j->op = OP_CLOSE;
    j->sess = it->second;
    j->sess->closing = true;
    j->t_submit = Clock::now();
    enqueue(srv, j);
  }
  srv->cv_work.notify_all();
  return PMX_OK;
}

static int wait_job(pmx_server *srv, long long ticket, std::shared_ptr<Job> &j) {
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(ticket);
  // unknown, already collected, or another thread is waiting on it
  if (it == srv->jobs.end() || it->second->waited) return PMX_ERR_ARG;
  j = it->second;
  j->waited = true;
  srv->cv_done.wait(lk, [&] { return j->done; });
  srv->jobs.erase(ticket);
  return PMX_OK;
}

int pmx_server_wait(pmx_server *srv, long long ticket, double *start_ms, double *done_ms) {
  if (!srv) return PMX_ERR_ARG;
  std::shared_ptr<Job> j;
  if (int rc = wait_job(srv, ticket, j)) return rc;
  auto ms = [&](Clock::time_point t) { return std::chrono::duration<double, std::milli>(t - j->t_submit).count(); };
  if (start_ms) *start_ms = ms(j->t_start);
  if (done_ms) *done_ms = ms(j->t_done);
  return j->status;
}

int pmx_server_wait_times(pmx_server *srv, long long ticket, double times[3]) {
  if (!srv || !times) return PMX_ERR_ARG;
  std::shared_ptr<Job> j;
  if (int rc = wait_job(srv, ticket, j)) return rc;
  auto ms = [&](Clock::time_point t) { return std::chrono::duration<double, std::milli>(t - srv->t0).count(); };
  times[0] = ms(j->t_submit);
  times[1] = ms(j->t_start);
  times[2] = ms(j->t_done);
  return j->status;
}

int pmx_server_pending(pmx_server *srv, int *n) {
  if (!srv || !n) return PMX_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  size_t q = srv->queue.size();
  for (auto &p : srv->priv) q += p.size();
  for (auto &kv : srv->sessions) q += kv.second->pending.size();
  *n = (int)q + srv->running;
  return PMX_OK;
}

int pmx_server_stats(pmx_server *srv, unsigned long long *peak_bytes, int *peak_running) {
  if (!srv) return PMX_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (peak_bytes) *peak_bytes = srv->peak_bytes;
  if (peak_running) *peak_running = srv->peak_running;
  return PMX_OK;
}

int pmx_server_destroy(pmx_server *srv) {
  if (!srv) return PMX_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  for (auto &t : srv->workers) t.join();
  delete srv;
  return PMX_OK;
}

const char *pmx_server_last_error(const pmx_server *srv) {
  thread_local std::string copy;
  if (!srv) return "";
  std::lock_guard<std::mutex> lk(srv->mu);
  copy = srv->err;
  return copy.c_str();
}

}  // extern "C"
