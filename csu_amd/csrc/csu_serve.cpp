// csu_serve.cpp -- continuous-batching solve server (host code).
//
// Each admitted solve owns a csu_dev context and its stream.  A step admits
// queued solves while slots and the HBM budget allow (first come, first
// served), keeps a chunk of iterations enqueued on every active context,
// and retires solves whose density residual has dropped by the requested
// factor.  Contexts are independent, so the GPU overlaps their kernels and
// each solve's numbers do not depend on what else is running.  Contexts are
// created on worker threads (host geometry, uploads, wall distance), so
// admission overlaps with the iterations of the solves already running.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>
#include <string.h>
#include "csu.h"
#include "csu_serve.h"

namespace {

struct Req {
  int id;
  const csu_mesh_desc* mesh;
  const csu_params* prm;
  const double* U0;
  double drop;
  int max_iter;
  int keep;
  double t_submit;
  double cfl_max = 0.0, cfl_growth = 1.0;  // CFL ramp: x growth per chunk up to cfl_max
  int peak_ref = 0;          // drops measured from the largest residual so far (1) or the first (0)
  double switch_drop = 0.0;  // mixed precision: FP64 from this drop on (0: never)
};

struct Job {
  Req r;
  csu_dev* dev = nullptr;
  csu_job_result res{};
  int chunk_n = 0;  // iterations of the chunk in flight
  bool res0_set = false;
  double peak = 0.0;  // largest density residual so far
  double best = 0.0;  // smallest density residual so far (mixed precision: stall test)
  int stalled = 0;    // chunks since the smallest residual last fell by 10 %
  double cfl = 0.0;  // current CFL of a ramped job
  int64_t est = 0;   // footprint estimate charged while the context is created
  int64_t charged = 0;  // device bytes of the context charged against the budget
  int create_rc = 0;
};

}  // namespace

struct csu_server {
  int max_slots, chunk;
  int64_t budget, used = 0;
  int next_id = 0;
  std::chrono::steady_clock::time_point t0;
  std::deque<Req> queue;
  std::vector<Job*> active;
  std::deque<csu_job_result> done;
  std::map<int, std::vector<double>> states;
  std::vector<double> rbuf, cbuf;
  // context creation on the worker thread: jobs handed over, jobs back
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job*> to_create, created;
  int creating = 0;  // handed over and not yet taken back (main thread's count)
  bool stop = false;
  std::vector<std::thread> workers;  // creators: up to 4, never more than the slots
};

static void creator(csu_server* s) {
  for (;;) {
    Job* j;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [s] { return s->stop || !s->to_create.empty(); });
      if (s->stop) return;
      j = s->to_create.front();
      s->to_create.pop_front();
    }
    int rc = csu_dev_create(j->r.mesh, j->r.prm, &j->dev);
    if (rc) j->dev = nullptr;
    if (!rc && j->r.U0) rc = csu_dev_set_state(j->dev, j->r.U0);
    if (rc && j->dev) {  // a context that cannot run holds no memory
      csu_dev_destroy(j->dev);
      j->dev = nullptr;
    }
    j->create_rc = rc;
    std::lock_guard<std::mutex> lk(s->mu);
    s->created.push_back(j);
  }
}

static double now(const csu_server* s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

extern "C" double csu_server_clock(const csu_server* s) { return s ? now(s) : -1.0; }

extern "C" int csu_server_create(int32_t max_slots, int64_t hbm_budget, int32_t chunk, csu_server** out) {
  if (!out || max_slots <= 0 || hbm_budget <= 0 || chunk <= 0) return CSU_ERR_ARG;
  csu_server* s = new csu_server();
  s->max_slots = max_slots;
  s->budget = hbm_budget;
  s->chunk = chunk;
  s->t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < std::min(4, max_slots); k++) s->workers.emplace_back(creator, s);
  *out = s;
  return CSU_OK;
}

extern "C" void csu_server_destroy(csu_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (std::thread& t : s->workers) t.join();
  for (Job* j : s->to_create) delete j;
  for (Job* j : s->created) {
    if (j->dev) csu_dev_destroy(j->dev);
    delete j;
  }
  for (Job* j : s->active) {
    csu_dev_destroy(j->dev);
    delete j;
  }
  delete s;
}

extern "C" int csu_server_submit(csu_server* s, const csu_mesh_desc* mesh, const csu_params* prm, const double* U0,
                                 double drop, int32_t max_iter, int32_t keep_state, int32_t* id) {
  if (!s || !mesh || !prm || !(drop > 0) || max_iter <= 0) return CSU_ERR_ARG;
  Req r{s->next_id++, mesh, prm, U0, drop, max_iter, keep_state ? 1 : 0, now(s)};
  s->queue.push_back(r);
  if (id) *id = r.id;
  return CSU_OK;
}

extern "C" int csu_server_set_ramp(csu_server* s, int32_t id, double cfl_max, double growth) {
  if (!s || !(growth >= 1.0) || !(cfl_max > 0.0)) return CSU_ERR_ARG;
  for (Req& r : s->queue)
    if (r.id == id) {
      r.cfl_max = cfl_max;
      r.cfl_growth = growth;
      return CSU_OK;
    }
  return CSU_ERR_ARG;  // unknown, or already admitted
}

extern "C" int csu_server_set_options(csu_server* s, int32_t id, int32_t peak_ref, double switch_drop) {
  if (!s || !(switch_drop >= 0.0)) return CSU_ERR_ARG;
  for (Req& r : s->queue)
    if (r.id == id) {
      r.peak_ref = peak_ref ? 1 : 0;
      r.switch_drop = switch_drop;
      return CSU_OK;
    }
  return CSU_ERR_ARG;  // unknown, or already admitted
}

extern "C" int csu_server_pending(const csu_server* s, int32_t* queued, int32_t* active) {
  if (!s) return CSU_ERR_ARG;
  if (queued) *queued = (int32_t)s->queue.size();
  if (active) *active = (int32_t)s->active.size() + s->creating;
  return CSU_OK;
}

static void retire(csu_server* s, size_t k, int status) {
  Job* j = s->active[k];
  j->res.status = status;
  j->res.t_done = now(s);
  if (j->dev) {
    j->res.bytes = csu_dev_bytes(j->dev);
    if (j->r.keep && (status == CSU_OK || status == CSU_ERR_DIVERGED)) {
      std::vector<double> U((size_t)j->r.prm->nvar * j->r.mesh->ncell);
      if (csu_dev_get_state(j->dev, U.data()) == CSU_OK) s->states[j->r.id] = std::move(U);
    }
    csu_dev_destroy(j->dev);
  }
  s->used -= j->charged;  // exactly what take_created added
  s->done.push_back(j->res);
  delete j;
  s->active.erase(s->active.begin() + (long)k);
}

static int enqueue_chunk(csu_server* s, Job* j) {
  int n = j->r.max_iter - j->res.iters;
  if (n > s->chunk) n = s->chunk;
  j->chunk_n = n;
  return csu_dev_iterate_async(j->dev, n);
}

// Hand queued solves to the creator thread while slots and the budget allow
// (contexts being created count against both, at their footprint estimate).
static void admit(csu_server* s) {
  bool handed = false;
  while (!s->queue.empty() && (int)s->active.size() + s->creating < s->max_slots) {
    const int64_t est = csu_dev_footprint(s->queue.front().mesh, s->queue.front().prm, s->chunk);
    // a solve larger than the whole budget still runs, alone
    if ((!s->active.empty() || s->creating) && s->used + est > s->budget) break;
    Job* j = new Job();
    j->r = s->queue.front();
    s->queue.pop_front();
    j->res.id = j->r.id;
    j->res.ncell = j->r.mesh->ncell;
    j->res.t_submit = j->r.t_submit;
    j->res.t_admit = now(s);
    j->res.iter_conv = -1;
    j->res.iter_switch = -1;
    j->cfl = j->r.prm->cfl;
    j->est = est;
    s->used += est;
    s->creating++;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->to_create.push_back(j);
    }
    handed = true;
  }
  if (handed) s->cv.notify_all();
}

// Take back the contexts the creator has finished and start their first chunk.
static int take_created(csu_server* s) {
  std::deque<Job*> got;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    got.swap(s->created);
  }
  int retired = 0;
  for (Job* j : got) {
    s->creating--;
    s->used -= j->est;
    s->active.push_back(j);
    int rc = j->create_rc;
    if (!rc) {
      // charged after the first chunk, which allocates the history ring
      rc = enqueue_chunk(s, j);
      j->charged = csu_dev_bytes(j->dev);
      s->used += j->charged;
    }
    if (rc) { retire(s, s->active.size() - 1, rc); retired++; }
  }
  return retired;
}

// A finished chunk: read its history, decide, and feed the next chunk.
static bool harvest(csu_server* s, size_t k) {
  Job* j = s->active[k];
  const int nv = j->r.prm->nvar;
  s->rbuf.resize((size_t)j->chunk_n * nv);
  s->cbuf.resize((size_t)j->chunk_n * 4);
  int rc = csu_dev_async_result(j->dev, s->rbuf.data(), s->cbuf.data());
  if (rc) { retire(s, k, rc); return true; }
  bool diverged = false;
  for (int it = 0; it < j->chunk_n; it++) {
    const double r0 = s->rbuf[(size_t)it * nv];
    j->res.iters++;
    if (!std::isfinite(r0)) {  // the state is gone: stop the solve
      j->res.res_last = r0;
      diverged = true;
      break;
    }
    if (!j->res0_set) { j->res.res0 = r0; j->res0_set = true; }
    j->peak = std::max(j->peak, r0);
    j->res.res_peak = j->peak;
    j->res.res_last = r0;
    memcpy(j->res.coef_last, &s->cbuf[(size_t)it * 4], sizeof(double) * 4);
    const double ref = j->r.peak_ref ? j->peak : j->res.res0;
    if (j->res.iter_conv < 0 && r0 <= j->r.drop * ref) {
      j->res.iter_conv = j->res.iters;
      j->res.converged = 1;
      break;
    }
  }
  if (diverged) { retire(s, k, CSU_ERR_DIVERGED); return true; }
  if (j->res.converged || j->res.iters >= j->r.max_iter) { retire(s, k, CSU_OK); return true; }
  // mixed precision levels off where its FP32 fluxes' rounding dominates
  // the residual: FP64 from the switch drop, or once the residual has not
  // fallen by 10 % in four chunks
  if (j->res.iters <= j->chunk_n || j->res.res_last < 0.9 * j->best) {
    j->best = j->res.res_last;
    j->stalled = 0;
  } else {
    j->stalled++;
  }
  if (j->r.switch_drop > 0.0 && j->res.iter_switch < 0 && j->r.prm->precision == CSU_PREC_MIXED &&
      (j->res.res_last <= j->r.switch_drop * (j->r.peak_ref ? j->peak : j->res.res0) || j->stalled >= 4)) {
    // mixed precision has done its part: FP64 to the converged state
    if ((rc = csu_dev_set_precision(j->dev, CSU_PREC_FP64))) { retire(s, k, rc); return true; }
    j->res.iter_switch = j->res.iters;
  }
  if (j->r.cfl_growth > 1.0 && j->cfl < j->r.cfl_max) {
    j->cfl = std::min(j->cfl * j->r.cfl_growth, j->r.cfl_max);
    if ((rc = csu_dev_set_cfl(j->dev, j->cfl))) { retire(s, k, rc); return true; }
  }
  if ((rc = enqueue_chunk(s, j))) { retire(s, k, rc); return true; }
  return false;
}

extern "C" int csu_server_step(csu_server* s, double max_wait) {
  if (!s) return -CSU_ERR_ARG;
  admit(s);
  const double t_end = now(s) + (max_wait > 0 ? max_wait : 0);
  int retired = take_created(s);
  bool any = retired > 0;
  while (!s->active.empty() || s->creating) {
    const int r0 = take_created(s);
    retired += r0;
    if (r0) any = true;
    for (size_t k = 0; k < s->active.size();) {
      int st = csu_dev_ready(s->active[k]->dev);
      if (st < 0) { retire(s, k, -st); retired++; any = true; continue; }
      if (st == 1) {
        any = true;
        if (harvest(s, k)) { retired++; continue; }
      }
      k++;
    }
    if (retired) admit(s);
    if (any || now(s) >= t_end) break;
    std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
  return retired;
}

extern "C" int csu_server_results(csu_server* s, csu_job_result* out, int32_t max) {
  if (!s || (!out && max > 0)) return -CSU_ERR_ARG;
  int n = 0;
  while (n < max && !s->done.empty()) {
    out[n++] = s->done.front();
    s->done.pop_front();
  }
  return n;
}

extern "C" int csu_server_state(const csu_server* s, int32_t id, double* U) {
  if (!s || !U) return CSU_ERR_ARG;
  auto it = s->states.find(id);
  if (it == s->states.end()) return CSU_ERR_ARG;
  memcpy(U, it->second.data(), it->second.size() * sizeof(double));
  return CSU_OK;
}
