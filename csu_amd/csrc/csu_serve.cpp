// csu_serve.cpp -- continuous-batching solve server (host code).
//
// Admitted solves run in units: a unit is one device context with its own
// stream, holding one solve (csu_dev_create) or, with batching on, several
// solves whose kernel-selecting options match (csu_batch_create: their
// meshes concatenated, one set of launches per iteration).  A step admits
// queued solves while unit slots and the HBM budget allow (first come,
// first served; a batch takes the oldest queued solve and then every later
// queued solve that can share its kernels, up to the batch limits), keeps a
// chunk of iterations enqueued on every unit, and retires solves whose
// density residual has dropped by the requested factor: a solve of a batch
// is held from then on and its result reported at once; a unit's memory is
// freed when its last solve retires.  Units are independent, so the GPU
// overlaps their kernels, and a solve's numbers do not depend on what else
// runs (batched solves are bit-equal to their runs alone; only the
// mixed-precision switch to FP64, taken by a batch as a whole when its
// first solve needs it, depends on the batch).  Contexts are created on
// worker threads (host geometry, uploads, wall distance), so admission
// overlaps with the iterations of the solves already running.
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
  int q = 0;  // index of the solve in its unit's context
  csu_job_result res{};
  bool running = true;
  bool res0_set = false;
  double peak = 0.0;  // largest density residual so far
  double best = 0.0;  // smallest density residual so far (mixed precision: stall test)
  int stalled = 0;    // chunks since the smallest residual last fell by 10 %
  double cfl = 0.0;   // current CFL of a ramped job
};

struct Unit {
  std::vector<Job*> jobs;
  csu_dev* dev = nullptr;
  bool batch = false;   // a csu_batch_create context
  int chunk_n = 0;      // iterations of the chunk in flight
  int running = 0;      // solves not yet retired
  int64_t est = 0;      // footprint estimate charged while the context is created
  int64_t charged = 0;  // device bytes of the context charged against the budget
  int64_t cells = 0;
  int create_rc = 0;
};

}  // namespace

struct csu_server {
  int max_slots, chunk;
  int batch_solves = 1;          // solves per unit (1: one context per solve)
  int64_t batch_cells = 0;       // cells per batch unit (0: no limit)
  int64_t budget, used = 0;
  int next_id = 0;
  std::chrono::steady_clock::time_point t0;
  std::deque<Req> queue;
  std::vector<Unit*> active;
  std::deque<csu_job_result> done;
  std::map<int, std::vector<double>> states;
  std::vector<double> rbuf, cbuf;
  // context creation on the worker threads: units handed over, units back
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Unit*> to_create, created;
  int creating = 0;  // handed over and not yet taken back (main thread's count)
  bool stop = false;
  std::vector<std::thread> workers;  // creators: up to 4, never more than the slots
};

static void destroy_unit(Unit* u) {
  if (u->dev) csu_dev_destroy(u->dev);
  for (Job* j : u->jobs) delete j;
  delete u;
}

static void creator(csu_server* s) {
  for (;;) {
    Unit* u;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [s] { return s->stop || !s->to_create.empty(); });
      if (s->stop) return;
      u = s->to_create.front();
      s->to_create.pop_front();
    }
    int rc;
    if (u->batch) {
      std::vector<const csu_mesh_desc*> meshes;
      std::vector<const csu_params*> prms;
      for (Job* j : u->jobs) {
        meshes.push_back(j->r.mesh);
        prms.push_back(j->r.prm);
      }
      rc = csu_batch_create((int32_t)u->jobs.size(), meshes.data(), prms.data(), &u->dev);
      for (size_t q = 0; !rc && q < u->jobs.size(); q++)
        if (u->jobs[q]->r.U0) rc = csu_batch_set_state(u->dev, (int32_t)q, u->jobs[q]->r.U0);
    } else {
      rc = csu_dev_create(u->jobs[0]->r.mesh, u->jobs[0]->r.prm, &u->dev);
      if (!rc && u->jobs[0]->r.U0) rc = csu_dev_set_state(u->dev, u->jobs[0]->r.U0);
    }
    if (rc && u->dev) {  // a context that cannot run holds no memory
      csu_dev_destroy(u->dev);
    }
    if (rc) u->dev = nullptr;
    u->create_rc = rc;
    std::lock_guard<std::mutex> lk(s->mu);
    s->created.push_back(u);
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

extern "C" int csu_server_set_batching(csu_server* s, int32_t max_solves, int64_t max_cells) {
  if (!s || max_solves <= 0 || max_cells < 0) return CSU_ERR_ARG;
  s->batch_solves = max_solves;
  s->batch_cells = max_cells;
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
  for (Unit* u : s->to_create) destroy_unit(u);
  for (Unit* u : s->created) destroy_unit(u);
  for (Unit* u : s->active) destroy_unit(u);
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
  int solves = 0;
  for (const Unit* u : s->active) solves += u->running;
  if (queued) *queued = (int32_t)s->queue.size();
  if (active) *active = (int32_t)solves + s->creating;
  return CSU_OK;
}

// Report one solve of unit k and take it out of the unit's work (a batch
// holds it; its state stays that of the end of the chunk).
static int finish_job(csu_server* s, Unit* u, Job* j, int status) {
  j->running = false;
  u->running--;
  j->res.status = status;
  j->res.t_done = now(s);
  int rc = CSU_OK;
  if (u->dev) {
    const int64_t bytes = csu_dev_bytes(u->dev);
    j->res.bytes = u->cells > 0 ? (int64_t)((double)bytes * j->r.mesh->ncell / (double)u->cells) : bytes;
    if (j->r.keep && (status == CSU_OK || status == CSU_ERR_DIVERGED)) {
      std::vector<double> U((size_t)j->r.prm->nvar * j->r.mesh->ncell);
      int g = u->batch ? csu_batch_get_state(u->dev, j->q, U.data()) : csu_dev_get_state(u->dev, U.data());
      if (g == CSU_OK) s->states[j->r.id] = std::move(U);
    }
    if (u->batch && u->running > 0) rc = csu_batch_set_active(u->dev, j->q, 0);
  }
  s->done.push_back(j->res);
  return rc;
}

// Remove unit k once none of its solves runs; frees its context.
static void drop_unit(csu_server* s, size_t k) {
  Unit* u = s->active[k];
  s->used -= u->charged;  // exactly what take_created added
  destroy_unit(u);
  s->active.erase(s->active.begin() + (long)k);
}

// End every running solve of unit k with status (an error of its context).
static void fail_unit(csu_server* s, size_t k, int status) {
  Unit* u = s->active[k];
  for (Job* j : u->jobs)
    if (j->running) finish_job(s, u, j, status);
  drop_unit(s, k);
}

static int enqueue_chunk(csu_server* s, Unit* u) {
  // a unit of one solve stops at its cap exactly; a batch runs whole chunks
  // (a capped solve of a batch keeps the state of its last chunk's end)
  int left = 0;
  for (Job* j : u->jobs)
    if (j->running) left = std::max(left, j->r.max_iter - j->res.iters);
  u->chunk_n = std::min(s->chunk, std::max(left, 1));
  if (u->batch) u->chunk_n = s->chunk;
  return csu_dev_iterate_async(u->dev, u->chunk_n);
}

static bool can_share(const Req& a, const Req& b) {
  return csu_batch_compatible(a.mesh, a.prm, b.mesh, b.prm) == 1;
}

// Hand queued solves to the creator threads while unit slots and the
// budget allow (units being created count against both, at their footprint
// estimate).  With batching, a unit takes the oldest queued solve and every
// later one that can share its kernels, up to the batch limits.
static void admit(csu_server* s) {
  bool handed = false;
  while (!s->queue.empty() && (int)s->active.size() + s->creating < s->max_slots) {
    const bool first = s->active.empty() && !s->creating;
    const Req& head = s->queue.front();
    const int64_t est0 = csu_dev_footprint(head.mesh, head.prm, s->chunk);
    // a solve larger than the whole budget still runs, alone
    if (!first && s->used + est0 > s->budget) break;
    Unit* u = new Unit();
    u->est = est0;
    u->cells = head.mesh->ncell;
    std::vector<size_t> take{0};
    if (s->batch_solves > 1) {
      for (size_t i = 1; i < s->queue.size() && (int)take.size() < s->batch_solves; i++) {
        const Req& r = s->queue[i];
        if (!can_share(head, r)) continue;
        if (s->batch_cells > 0 && u->cells + r.mesh->ncell > s->batch_cells) continue;
        const int64_t e = csu_dev_footprint(r.mesh, r.prm, s->chunk);
        if (s->used + u->est + e > s->budget) continue;
        take.push_back(i);
        u->est += e;
        u->cells += r.mesh->ncell;
      }
      u->batch = true;
    }
    const double t = now(s);
    for (size_t i : take) {
      Job* j = new Job();
      j->r = s->queue[i];
      j->q = (int)u->jobs.size();
      j->res.id = j->r.id;
      j->res.ncell = j->r.mesh->ncell;
      j->res.t_submit = j->r.t_submit;
      j->res.t_admit = t;
      j->res.iter_conv = -1;
      j->res.iter_switch = -1;
      j->cfl = j->r.prm->cfl;
      u->jobs.push_back(j);
    }
    for (size_t k = take.size(); k-- > 0;) s->queue.erase(s->queue.begin() + (long)take[k]);
    u->running = (int)u->jobs.size();
    s->used += u->est;
    s->creating++;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->to_create.push_back(u);
    }
    handed = true;
  }
  if (handed) s->cv.notify_all();
}

// Take back the contexts the creators have finished and start their first
// chunk.
static int take_created(csu_server* s) {
  std::deque<Unit*> got;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    got.swap(s->created);
  }
  int retired = 0;
  for (Unit* u : got) {
    s->creating--;
    s->used -= u->est;
    s->active.push_back(u);
    int rc = u->create_rc;
    if (!rc) {
      // charged after the first chunk, which allocates the history ring
      rc = enqueue_chunk(s, u);
      u->charged = csu_dev_bytes(u->dev);
      s->used += u->charged;
    }
    if (rc) {
      retired += u->running;
      fail_unit(s, s->active.size() - 1, rc);
    }
  }
  return retired;
}

// One solve's rows of a finished chunk: res/coef of iteration it at
// res[it * stride * nv], coef[it * stride * 4].  Returns 1 when the solve
// stops (status in *status), 0 when it runs on.
static int scan_chunk(Job* j, int chunk_n, const double* res, const double* coef, int stride, int* status) {
  const int nv = j->r.prm->nvar;
  for (int it = 0; it < chunk_n; it++) {
    const double r0 = res[(size_t)it * stride * nv];
    j->res.iters++;
    if (!std::isfinite(r0)) {  // the state is gone: stop the solve
      j->res.res_last = r0;
      *status = CSU_ERR_DIVERGED;
      return 1;
    }
    if (!j->res0_set) { j->res.res0 = r0; j->res0_set = true; }
    j->peak = std::max(j->peak, r0);
    j->res.res_peak = j->peak;
    j->res.res_last = r0;
    memcpy(j->res.coef_last, coef + (size_t)it * stride * 4, sizeof(double) * 4);
    const double ref = j->r.peak_ref ? j->peak : j->res.res0;
    if (j->res.iter_conv < 0 && r0 <= j->r.drop * ref) {
      j->res.iter_conv = j->res.iters;
      j->res.converged = 1;
      *status = CSU_OK;
      return 1;
    }
    if (j->res.iters >= j->r.max_iter) {
      *status = CSU_OK;
      return 1;
    }
  }
  // mixed precision levels off where its FP32 fluxes' rounding dominates
  // the residual: the stall count (chunks without a 10 % fall of the
  // smallest residual), counted once the residual is below a tenth of its
  // peak (before that a rising start-up transient is no stall)
  if (j->res.iters <= chunk_n || j->res.res_last < 0.9 * j->best) {
    j->best = j->res.res_last;
    j->stalled = 0;
  } else if (j->res.res_last < 0.1 * j->peak) {
    j->stalled++;
  }
  return 0;
}

// A finished chunk of unit k: read its history, decide per solve, and feed
// the next chunk.  Returns the number of solves retired.
static int harvest(csu_server* s, size_t k) {
  Unit* u = s->active[k];
  const int ns = (int)u->jobs.size();
  const int nv = u->jobs[0]->r.prm->nvar;
  s->rbuf.resize((size_t)u->chunk_n * ns * nv);
  s->cbuf.resize((size_t)u->chunk_n * ns * 4);
  int rc = csu_dev_async_result(u->dev, s->rbuf.data(), s->cbuf.data());
  if (rc) {
    const int n = u->running;
    fail_unit(s, k, rc);
    return n;
  }
  int retired = 0;
  bool to_fp64 = false;
  for (Job* j : u->jobs) {
    if (!j->running) continue;
    int status = CSU_OK;
    if (scan_chunk(j, u->chunk_n, &s->rbuf[(size_t)j->q * nv], &s->cbuf[(size_t)j->q * 4], ns, &status)) {
      retired++;
      if ((rc = finish_job(s, u, j, status))) break;
      continue;
    }
    if (j->r.prm->precision == CSU_PREC_MIXED && j->r.switch_drop > 0.0 && j->res.iter_switch < 0) {
      const double ref = j->r.peak_ref ? j->peak : j->res.res0;
      if (j->res.res_last <= j->r.switch_drop * ref || j->stalled >= 4) to_fp64 = true;
    }
    if (j->r.cfl_growth > 1.0 && j->cfl < j->r.cfl_max) {
      j->cfl = std::min(j->cfl * j->r.cfl_growth, j->r.cfl_max);
      rc = u->batch ? csu_batch_set_cfl(u->dev, j->q, j->cfl) : csu_dev_set_cfl(u->dev, j->cfl);
      if (rc) break;
    }
  }
  if (!rc && u->running > 0 && to_fp64) {
    // mixed precision has done its part: FP64 to the converged state (a
    // batch switches as soon as one running solve has reached its switch
    // drop or stalled: from there on mixed precision would hold that solve
    // back, and FP64 is never wrong for the others)
    if (!(rc = csu_dev_set_precision(u->dev, CSU_PREC_FP64)))
      for (Job* j : u->jobs)
        if (j->running) j->res.iter_switch = j->res.iters;
  }
  if (!rc && u->running > 0) rc = enqueue_chunk(s, u);
  if (rc) {
    retired += u->running;
    fail_unit(s, k, rc);
    return retired;
  }
  if (u->running == 0) drop_unit(s, k);
  return retired;
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
      Unit* u = s->active[k];
      int st = csu_dev_ready(u->dev);
      if (st < 0) {
        retired += u->running;
        any = true;
        fail_unit(s, k, -st);
        continue;
      }
      if (st == 1) {
        any = true;
        const size_t before = s->active.size();
        retired += harvest(s, k);
        if (s->active.size() < before) continue;  // the unit is gone
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
