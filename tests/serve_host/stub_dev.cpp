// Host stand-in for the device library, for testing the solve server's
// admission, batching, retire and accounting logic on the CPU (with the
// sanitizers).  Each solve's density residual follows a closed form of its
// parameters: r(t) = 10 exp(-0.01 cfl t) at iteration t (1-based), with cfl
// the solve's current CFL; a solve whose inlet Mach is above 5 gives NaN
// from iteration 25 on.  Every context records how many iterations each of
// its solves ran, so the driver can check that held solves stop.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "csu.h"
#include "stub_dev.h"

struct csu_dev {
  int nsolve = 1;
  bool batch = false;
  std::vector<int> ncell, nvar, active, iters;
  std::vector<double> cfl, mach;
  int prec = 0;
  int pending = 0;  // iterations of the last async call
  std::vector<double> res, coef;
  int64_t bytes = 0;
};

static int g_live = 0, g_created = 0, g_max_batch = 0;

extern "C" int stub_live(void) { return g_live; }
extern "C" int stub_created(void) { return g_created; }
extern "C" int stub_max_batch(void) { return g_max_batch; }

static double resid(const csu_dev* d, int q, int t) {
  if (d->mach[q] > 5.0 && t >= 25) return NAN;
  if (d->mach[q] == 4.0)  // a start-up transient rising for 100 iterations, then a floor
    return t < 100 ? 10.0 * (1.0 + 0.05 * t) : std::fmax(60.0 * std::exp(-0.05 * (t - 100)), 0.1);
  return 10.0 * std::exp(-0.01 * d->cfl[q] * t);
}

static csu_dev* make(int n, const csu_mesh_desc* const* m, const csu_params* const* p, bool batch) {
  csu_dev* d = new csu_dev();
  d->nsolve = n;
  d->batch = batch;
  for (int q = 0; q < n; q++) {
    d->ncell.push_back(m[q]->ncell);
    d->nvar.push_back(p[q]->nvar);
    d->active.push_back(1);
    d->iters.push_back(0);
    d->cfl.push_back(p[q]->cfl);
    d->mach.push_back(p[q]->mach_inf);
    d->bytes += 1000 * (int64_t)m[q]->ncell;
  }
  d->prec = p[0]->precision;
  g_live++;
  g_created++;
  if (n > g_max_batch) g_max_batch = n;
  return d;
}

extern "C" int csu_dev_create(const csu_mesh_desc* mesh, const csu_params* prm, csu_dev** out) {
  if (!mesh || !prm || !out) return CSU_ERR_ARG;
  if (mesh->ncell <= 0) return CSU_ERR_ARG;
  *out = make(1, &mesh, &prm, false);
  return CSU_OK;
}

extern "C" int csu_batch_create(int32_t n, const csu_mesh_desc* const* meshes, const csu_params* const* prms,
                                csu_dev** out) {
  if (n <= 0 || !meshes || !prms || !out) return CSU_ERR_ARG;
  for (int q = 0; q < n; q++)
    if (meshes[q]->ncell <= 0 || csu_batch_compatible(meshes[0], prms[0], meshes[q], prms[q]) != 1)
      return CSU_ERR_ARG;
  *out = make(n, meshes, prms, true);
  return CSU_OK;
}

extern "C" int csu_batch_compatible(const csu_mesh_desc* ma, const csu_params* pa, const csu_mesh_desc* mb,
                                    const csu_params* pb) {
  if (!ma || !pa || !mb || !pb) return -CSU_ERR_ARG;
  return pa->flux == pb->flux && pa->time_scheme == pb->time_scheme && pa->nvar == pb->nvar &&
         pa->precision == pb->precision;
}

extern "C" void csu_dev_destroy(csu_dev* d) {
  if (!d) return;
  g_live--;
  delete d;
}

extern "C" int csu_dev_set_state(csu_dev* d, const double* U) { return d && U ? CSU_OK : CSU_ERR_ARG; }

extern "C" int csu_batch_set_state(csu_dev* d, int32_t q, const double* U) {
  return d && U && q >= 0 && q < d->nsolve ? CSU_OK : CSU_ERR_ARG;
}

// state of a solve: every entry is its iteration count
static void fill(const csu_dev* d, int q, double* U) {
  const size_t n = (size_t)d->nvar[q] * d->ncell[q];
  for (size_t k = 0; k < n; k++) U[k] = d->iters[q];
}

extern "C" int csu_dev_get_state(csu_dev* d, double* U) {
  if (!d || !U || d->nsolve != 1) return CSU_ERR_ARG;
  fill(d, 0, U);
  return CSU_OK;
}

extern "C" int csu_batch_get_state(csu_dev* d, int32_t q, double* U) {
  if (!d || !U || !d->batch || q < 0 || q >= d->nsolve) return CSU_ERR_ARG;
  fill(d, q, U);
  return CSU_OK;
}

extern "C" int csu_batch_set_active(csu_dev* d, int32_t q, int32_t on) {
  if (!d || !d->batch || q < 0 || q >= d->nsolve) return CSU_ERR_ARG;
  d->active[q] = on ? 1 : 0;
  return CSU_OK;
}

extern "C" int csu_batch_set_cfl(csu_dev* d, int32_t q, double cfl) {
  if (!d || !d->batch || q < 0 || q >= d->nsolve || !(cfl > 0)) return CSU_ERR_ARG;
  d->cfl[q] = cfl;
  return CSU_OK;
}

extern "C" int csu_dev_set_cfl(csu_dev* d, double cfl) {
  if (!d || d->batch || !(cfl > 0)) return CSU_ERR_ARG;
  d->cfl[0] = cfl;
  return CSU_OK;
}

extern "C" int csu_dev_set_precision(csu_dev* d, int32_t prec) {
  if (!d) return CSU_ERR_ARG;
  d->prec = prec;
  return CSU_OK;
}

extern "C" int csu_dev_iterate_async(csu_dev* d, int32_t n) {
  if (!d || n <= 0) return CSU_ERR_ARG;
  const int ns = d->nsolve, nv = d->nvar[0];
  d->pending = n;
  d->res.assign((size_t)n * ns * nv, 0.0);
  d->coef.assign((size_t)n * ns * 4, 0.0);
  for (int it = 0; it < n; it++)
    for (int q = 0; q < ns; q++) {
      if (!d->active[q]) continue;
      const int t = ++d->iters[q];
      d->res[((size_t)it * ns + q) * nv] = resid(d, q, t);
      for (int k = 0; k < 4; k++) d->coef[((size_t)it * ns + q) * 4 + k] = t + 0.25 * k;
    }
  d->bytes += 8 * (int64_t)n * ns * (nv + 4);  // the history ring, as the device's first call allocates
  return CSU_OK;
}

extern "C" int csu_dev_ready(csu_dev* d) { return d ? 1 : -CSU_ERR_ARG; }

extern "C" int csu_dev_async_result(csu_dev* d, double* res, double* coef) {
  if (!d || !d->pending) return CSU_ERR_ARG;
  if (res) memcpy(res, d->res.data(), d->res.size() * sizeof(double));
  if (coef) memcpy(coef, d->coef.data(), d->coef.size() * sizeof(double));
  d->pending = 0;
  return CSU_OK;
}

extern "C" int64_t csu_dev_bytes(const csu_dev* d) { return d ? d->bytes : 0; }

extern "C" int64_t csu_dev_footprint(const csu_mesh_desc* m, const csu_params* p, int32_t hist) {
  return 1000 * (int64_t)m->ncell + 8 * (int64_t)hist * (p->nvar + 4);
}
