// Request loop over the pmx engine (include/pmx_serve.h).
//
// Worker threads each own one pmx_ctx (one HIP stream on the device).  A
// worker takes the oldest queued request once the sum of the device
// footprints of the running requests plus this one fits in the budget (or
// nothing else runs), builds the resident system, runs it and releases it.
// Requests that arrive while others run are admitted as soon as a slot and
// memory free up, so work of different requests overlaps on the device.
#include "../../include/pmx_serve.h"

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

struct Job {
  long long ticket = 0;
  int kind = PMX_REQ_FORCES;
  // copies of the request's inputs
  pmx_sys_desc sys{};
  std::vector<double> charge, lj, x, v, mass, bond_par, angle_par;
  std::vector<int> type, molecule, excl, bonds, angles;
  pmx_md_desc md{};
  int nsteps = 0;
  double *out_a = nullptr, *out_v = nullptr, *out_e = nullptr;
  unsigned long long bytes = 0;
  Clock::time_point t_submit, t_start, t_done;
  bool done = false;
  bool waited = false;  // a pmx_server_wait holds the ticket
  int status = PMX_OK;
};

}  // namespace

struct pmx_server {
  int device = 0;
  unsigned long long budget = 0, in_use = 0;
  std::vector<std::thread> workers;
  std::deque<std::shared_ptr<Job>> queue;
  std::map<long long, std::shared_ptr<Job>> jobs;
  mutable std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int running = 0;
  long long next = 1;
  std::string err;
};

namespace {

int run_job(pmx_ctx *ctx, Job &j, std::string &err) {
  pmx_sys *sys = nullptr;
  int rc = pmx_sys_create(ctx, &j.sys, &sys);
  if (rc == PMX_OK) {
    const size_t n = (size_t)j.sys.n;
    if (j.kind == PMX_REQ_FORCES) {
      std::vector<double> f(3 * n, 0.0);
      double e[PMX_E_COUNT] = {0, 0};
      rc = pmx_sys_evaluate(sys, j.x.data(), f.data(), e, 0);
      if (rc == PMX_OK) {
        if (j.out_a) std::memcpy(j.out_a, f.data(), 24 * n);
        if (j.out_e) std::memcpy(j.out_e, e, sizeof(e));
      }
    } else {
      double e[PMX_MD_E_COUNT] = {0, 0, 0, 0, 0};
      std::vector<double> x(3 * n), v(3 * n);
      if ((rc = pmx_md_init(sys, &j.md)) == PMX_OK && (rc = pmx_md_set_state(sys, j.x.data(), j.v.data())) == PMX_OK &&
          (rc = pmx_md_run(sys, j.nsteps)) == PMX_OK && (rc = pmx_md_get_state(sys, x.data(), v.data(), e)) == PMX_OK) {
        if (j.out_a) std::memcpy(j.out_a, x.data(), 24 * n);
        if (j.out_v) std::memcpy(j.out_v, v.data(), 24 * n);
        if (j.out_e) std::memcpy(j.out_e, e, sizeof(e));
      }
    }
  }
  if (rc != PMX_OK) err = pmx_last_error(ctx);
  if (sys) pmx_sys_destroy(sys);
  return rc;
}

void worker(pmx_server *srv) {
  pmx_ctx *ctx = nullptr;
  const int crc = pmx_ctx_create(srv->device, &ctx);
  for (;;) {
    std::shared_ptr<Job> job;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      srv->cv_work.wait(lk, [&] {
        if (srv->queue.empty()) return srv->stop;
        const auto &h = srv->queue.front();
        return srv->running == 0 || srv->in_use + h->bytes <= srv->budget;
      });
      if (srv->queue.empty()) break;  // stop requested and nothing left
      job = srv->queue.front();
      srv->queue.pop_front();
      srv->in_use += job->bytes;
      ++srv->running;
      job->t_start = Clock::now();
    }
    std::string err;
    const int rc = crc == PMX_OK ? run_job(ctx, *job, err) : crc;
    if (crc != PMX_OK) err = "no device context";
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->in_use -= job->bytes;
      --srv->running;
      job->status = rc;
      job->t_done = Clock::now();
      job->done = true;
      if (rc != PMX_OK) srv->err = err;
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
  }
  if (ctx) pmx_ctx_destroy(ctx);
}

template <class T>
void copy_in(std::vector<T> &dst, const T *src, size_t n) {
  dst.assign(src, src + (src ? n : 0));
}

}  // namespace

extern "C" {

int pmx_server_create(int device, int slots, unsigned long long hbm_budget, pmx_server **out) {
  if (!out || slots < 1 || slots > 64) return PMX_ERR_ARG;
  auto *srv = new pmx_server;
  srv->device = device;
  srv->budget = hbm_budget;
  for (int i = 0; i < slots; ++i) srv->workers.emplace_back(worker, srv);
  *out = srv;
  return PMX_OK;
}

int pmx_server_submit(pmx_server *srv, const pmx_request *r, long long *ticket) {
  if (!srv || !r || !r->sys || !r->x || !ticket) return PMX_ERR_ARG;
  if (r->kind != PMX_REQ_FORCES && r->kind != PMX_REQ_TRAJECTORY) return PMX_ERR_ARG;
  if (r->kind == PMX_REQ_TRAJECTORY && (!r->md || !r->v || r->nsteps < 0)) return PMX_ERR_ARG;
  const pmx_sys_desc &d = *r->sys;
  if (d.n <= 0 || d.ntypes <= 0 || d.nexcl < 0) return PMX_ERR_ARG;
  auto j = std::make_shared<Job>();
  j->kind = r->kind;
  j->sys = d;
  const size_t n = (size_t)d.n;
  copy_in(j->charge, d.charge, n);
  copy_in(j->type, d.type, n);
  copy_in(j->molecule, d.molecule, n);
  copy_in(j->lj, d.lj, (size_t)d.ntypes * d.ntypes * 4);
  copy_in(j->excl, d.excl, 3 * (size_t)d.nexcl);
  j->sys.charge = j->charge.data();
  j->sys.type = j->type.data();
  j->sys.molecule = d.molecule ? j->molecule.data() : nullptr;
  j->sys.lj = j->lj.data();
  j->sys.excl = j->excl.empty() ? nullptr : j->excl.data();
  copy_in(j->x, r->x, 3 * n);
  if (r->kind == PMX_REQ_TRAJECTORY) {
    const pmx_md_desc &m = *r->md;
    if (m.nbonds < 0 || m.nangles < 0 || !m.mass) return PMX_ERR_ARG;
    j->md = m;
    copy_in(j->v, r->v, 3 * n);
    copy_in(j->mass, m.mass, n);
    copy_in(j->bonds, m.bonds, 2 * (size_t)m.nbonds);
    copy_in(j->bond_par, m.bond_par, 2 * (size_t)m.nbonds);
    copy_in(j->angles, m.angles, 3 * (size_t)m.nangles);
    copy_in(j->angle_par, m.angle_par, 2 * (size_t)m.nangles);
    j->md.mass = j->mass.data();
    j->md.bonds = j->bonds.empty() ? nullptr : j->bonds.data();
    j->md.bond_par = j->bond_par.empty() ? nullptr : j->bond_par.data();
    j->md.angles = j->angles.empty() ? nullptr : j->angles.data();
    j->md.angle_par = j->angle_par.empty() ? nullptr : j->angle_par.data();
    j->nsteps = r->nsteps;
  }
  j->out_a = r->out_a;
  j->out_v = r->out_v;
  j->out_e = r->out_energies;
  j->bytes = pmx_estimate_bytes(&j->sys, r->kind);
  if (j->bytes == 0) return PMX_ERR_ARG;  // descriptor the engine rejects
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    j->ticket = srv->next++;
    srv->jobs[j->ticket] = j;
    srv->queue.push_back(j);
  }
  srv->cv_work.notify_one();
  *ticket = j->ticket;
  return PMX_OK;
}

int pmx_server_wait(pmx_server *srv, long long ticket, double *start_ms, double *done_ms) {
  if (!srv) return PMX_ERR_ARG;
  std::shared_ptr<Job> j;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->jobs.find(ticket);
    // unknown, already collected, or another thread is waiting on it
    if (it == srv->jobs.end() || it->second->waited) return PMX_ERR_ARG;
    j = it->second;
    j->waited = true;
    srv->cv_done.wait(lk, [&] { return j->done; });
    srv->jobs.erase(ticket);
  }
  auto ms = [&](Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(t - j->t_submit).count();
  };
  if (start_ms) *start_ms = ms(j->t_start);
  if (done_ms) *done_ms = ms(j->t_done);
  return j->status;
}

int pmx_server_pending(pmx_server *srv, int *n) {
  if (!srv || !n) return PMX_ERR_ARG;
  std::lock_guard<std::mutex> lk(srv->mu);
  *n = (int)srv->queue.size() + srv->running;
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
