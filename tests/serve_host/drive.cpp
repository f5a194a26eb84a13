// Drives the solve server (csu_serve.cpp) over the host stand-in of the
// device library (stub_dev.cpp): the same requests served one context per
// solve and in batched units must give the same results, equal to the
// stand-in's closed-form residual; batches must hold only solves that can
// share kernels and respect their limits; every context must be freed.
// Prints "ok" and exits 0 when every check holds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>
#include "csu.h"
#include "csu_serve.h"
#include "stub_dev.h"

#define CHECK(c)                                                       \
  do {                                                                 \
    if (!(c)) {                                                        \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      std::exit(1);                                                    \
    }                                                                  \
  } while (0)

struct Spec {
  int ncell, flux, prec;
  double cfl, mach;
  int max_iter;
  double ramp_max, ramp_growth, switch_drop;
};

struct Run {
  std::map<int, csu_job_result> res;
  std::map<int, double> state0;  // first entry of the kept state: the solve's iterations run
};

static Run serve(const std::vector<Spec>& specs, int slots, int batch, int64_t batch_cells, int64_t budget,
                 int chunk, double drop) {
  std::vector<csu_mesh_desc> meshes(specs.size());
  std::vector<csu_params> prms(specs.size());
  std::vector<std::vector<double>> U0(specs.size());
  csu_server* s = nullptr;
  CHECK(csu_server_create(slots, budget, chunk, &s) == CSU_OK);
  if (batch > 1) CHECK(csu_server_set_batching(s, batch, batch_cells) == CSU_OK);
  for (size_t k = 0; k < specs.size(); k++) {
    meshes[k] = csu_mesh_desc{};
    meshes[k].ncell = specs[k].ncell;
    prms[k] = csu_params{};
    prms[k].nvar = 4;
    prms[k].flux = specs[k].flux;
    prms[k].precision = specs[k].prec;
    prms[k].cfl = specs[k].cfl;
    prms[k].mach_inf = specs[k].mach;
    U0[k].assign(4 * (size_t)specs[k].ncell, 0.0);
    int32_t id = -1;
    CHECK(csu_server_submit(s, &meshes[k], &prms[k], (k % 3 == 0) ? U0[k].data() : nullptr, drop,
                            specs[k].max_iter, 1, &id) == CSU_OK);
    CHECK(id == (int)k);
    if (specs[k].ramp_growth > 1.0) CHECK(csu_server_set_ramp(s, id, specs[k].ramp_max, specs[k].ramp_growth) == 0);
    if (specs[k].switch_drop > 0) CHECK(csu_server_set_options(s, id, 0, specs[k].switch_drop) == CSU_OK);
  }
  Run out;
  std::vector<csu_job_result> buf(8);
  for (int guard = 0; out.res.size() < specs.size(); guard++) {
    CHECK(guard < 100000);
    CHECK(csu_server_step(s, 0.001) >= 0);
    int n;
    while ((n = csu_server_results(s, buf.data(), (int32_t)buf.size())) > 0)
      for (int i = 0; i < n; i++) {
        CHECK(out.res.count(buf[i].id) == 0);  // reported once
        out.res[buf[i].id] = buf[i];
      }
  }
  int32_t q = -1, a = -1;
  CHECK(csu_server_pending(s, &q, &a) == CSU_OK && q == 0 && a == 0);
  CHECK(stub_live() == 0);  // the last retire freed every context
  for (auto& kv : out.res) {
    const Spec& sp = specs[kv.first];
    std::vector<double> U(4 * (size_t)sp.ncell);
    CHECK(csu_server_state(s, kv.first, U.data()) == CSU_OK);
    out.state0[kv.first] = U[0];
  }
  csu_server_destroy(s);
  return out;
}

// The stand-in's residual run alone, chunk by chunk as the server runs it.
static void expect(const Spec& sp, int chunk, double drop, const csu_job_result& r, double state0) {
  double cfl = sp.cfl, r1 = 0;
  int t = 0, conv = -1, status = CSU_OK;
  bool stop = false;
  while (!stop) {
    for (int it = 0; it < chunk && !stop; it++) {
      t++;
      const double rt = (sp.mach > 5.0 && t >= 25) ? NAN : 10.0 * std::exp(-0.01 * cfl * t);
      if (!std::isfinite(rt)) { status = CSU_ERR_DIVERGED; stop = true; break; }
      if (t == 1) r1 = rt;
      if (rt <= drop * r1) { conv = t; stop = true; break; }
      if (t >= sp.max_iter) stop = true;
    }
    if (!stop && sp.ramp_growth > 1.0 && cfl < sp.ramp_max) cfl = std::fmin(cfl * sp.ramp_growth, sp.ramp_max);
  }
  CHECK(r.status == status);
  CHECK(r.iters == t);
  CHECK(r.iter_conv == conv && r.converged == (conv > 0));
  // the kept state is that of the end of the chunk in which the solve stopped
  CHECK(state0 == chunk * ((t + chunk - 1) / chunk));
}

int main() {
  const int chunk = 20;
  const double drop = 1e-6;
  std::vector<Spec> specs;
  for (int k = 0; k < 14; k++) {
    Spec sp{};
    sp.ncell = 1000 + 700 * ((k * 7) % 11);
    sp.flux = k % 3 == 2 ? 1 : 0;  // two kinds of kernels
    sp.cfl = 8.0 + 3.0 * ((k * 5) % 7);
    sp.mach = k == 4 ? 6.0 : 2.0;  // solve 4 diverges at iteration 25
    sp.max_iter = k == 6 ? 60 : 400;  // solve 6 stops at its cap, a whole number of chunks
    if (k == 9) { sp.ramp_max = 30.0; sp.ramp_growth = 1.2; }
    specs.push_back(sp);
  }
  const int64_t big = (int64_t)1 << 40;
  Run one = serve(specs, 1, 1, 0, big, chunk, drop);
  CHECK(stub_max_batch() == 1);
  Run slots = serve(specs, 4, 1, 0, big, chunk, drop);
  Run bat = serve(specs, 3, 4, 0, big, chunk, drop);
  CHECK(stub_max_batch() == 4);
  Run bcell = serve(specs, 2, 8, 12000, big, chunk, drop);  // cell limit per batch
  Run bbud = serve(specs, 3, 8, 0, 20000000, chunk, drop);  // HBM budget limits the batches
  for (size_t k = 0; k < specs.size(); k++) {
    for (const Run* r : {&one, &slots, &bat, &bcell, &bbud}) expect(specs[k], chunk, drop, r->res.at((int)k), r->state0.at((int)k));
  }
  // the batches held only solves that share kernels: solves admitted together
  // (one unit: the same admission time) have the same flux
  std::map<double, std::vector<int>> units;
  for (auto& kv : bat.res) units[kv.second.t_admit].push_back(kv.first);
  bool packed = false;
  for (auto& u : units) {
    packed |= u.second.size() > 1;
    CHECK(u.second.size() <= 4);
    for (int id : u.second) CHECK(specs[id].flux == specs[u.second[0]].flux);
  }
  CHECK(packed);
  std::map<double, int> cells;
  for (auto& kv : bcell.res) cells[kv.second.t_admit] += specs[kv.first].ncell;
  for (auto& c : cells) CHECK(c.second <= 12000 || c.second == specs[0].ncell);
  // mixed precision: a batch switches to FP64 once every running solve has
  // reached its switch drop
  std::vector<Spec> mixed;
  for (int k = 0; k < 4; k++) {
    Spec sp{};
    sp.ncell = 2000 + 500 * k;
    sp.prec = CSU_PREC_MIXED;
    sp.cfl = 10.0 + 5.0 * k;
    sp.mach = 2.0;
    sp.max_iter = 400;
    sp.switch_drop = 1e-2;
    mixed.push_back(sp);
  }
  Run mb = serve(mixed, 1, 4, 0, big, chunk, drop);
  Run mo = serve(mixed, 1, 1, 0, big, chunk, drop);
  int last = 0;
  for (int k = 0; k < 4; k++) {
    // alone: each switches at the end of the chunk where it reaches 1e-2
    const double tsw = 1 + std::log(100.0) / (0.01 * mixed[k].cfl);
    CHECK(mo.res.at(k).iter_switch == chunk * (int)std::ceil(tsw / chunk));
    last = std::max(last, mo.res.at(k).iter_switch);
  }
  // in one batch: at the first one's switch, for every solve still running
  // then (a solve that stopped before it ran mixed precision to the end)
  int first = 1 << 30;
  for (int k = 0; k < 4; k++) first = std::min(first, mo.res.at(k).iter_switch);
  CHECK(first < last);
  int switched = 0;
  for (int k = 0; k < 4; k++) {
    const csu_job_result& r = mb.res.at(k);
    if (r.iters > first) CHECK(r.iter_switch == first && ++switched);
    else CHECK(r.iter_switch == -1);
  }
  CHECK(switched >= 2);
  // the stall rule: a residual that rises through a start-up transient is no
  // stall; one at its floor, below a tenth of its peak, is after four chunks
  // without a 10 % fall
  std::vector<Spec> st(1);
  st[0].ncell = 3000;
  st[0].prec = CSU_PREC_MIXED;
  st[0].cfl = 10.0;
  st[0].mach = 4.0;
  st[0].max_iter = 600;
  st[0].switch_drop = 1e-6;
  Run sr = serve(st, 1, 1, 0, big, chunk, drop);
  int t_floor = 100 + (int)std::ceil(std::log(600.0) / 0.05);  // first iteration at the floor
  int sw = -1, stall = 0;
  double best = 0, peak = 0;
  for (int c = 1; c * chunk <= 600 && sw < 0; c++) {
    const int t = c * chunk;
    for (int u = t - chunk + 1; u <= t; u++)
      peak = std::fmax(peak, u < 100 ? 10.0 * (1.0 + 0.05 * u) : std::fmax(60.0 * std::exp(-0.05 * (u - 100)), 0.1));
    const double r = t < 100 ? 10.0 * (1.0 + 0.05 * t) : std::fmax(60.0 * std::exp(-0.05 * (t - 100)), 0.1);
    if (c == 1 || r < 0.9 * best) { best = r; stall = 0; }
    else if (r < 0.1 * peak) stall++;
    if (stall >= 4) sw = t;
  }
  CHECK(sw > t_floor && sw > 4 * chunk);
  CHECK(sr.res.at(0).iter_switch == sw);
  std::printf("ok\n");
  return 0;
}
