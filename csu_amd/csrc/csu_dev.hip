// csu_dev.hip -- device context and C ABI of the residual-and-update path
// (include/csu.h).  One context = one mesh on one GPU, all state resident in
// HBM, one HIP stream.  Kernels live in kernels.h.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>
#include "csu.h"
#include "kernels.h"
#include "kern_implicit.h"
#include "kern_setup.h"

using namespace csu;

#define CSU_MAX_DEVICES 64

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local char g_err[512] = "";

static int set_err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return set_err(CSU_ERR_HIP, "%s:%d %s: %s", __FILE__, __LINE__, #x,          \
                     hipGetErrorString(e_));                                       \
  } while (0)

extern "C" const char* csu_last_error(void) { return g_err; }
extern "C" const char* csu_version(void) { return "csu-mi355x 0.2 (gfx950)"; }

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct DevBuf {
  std::vector<void*> owned;
  ~DevBuf() {
    for (void* p : owned) (void)hipFree(p);
  }
};

struct csu_dev {
  int nc, nf, nb, nv, nmark, iter, visc;
  int device;                  // HIP device the context was created on
  int nsolve;                  // solves in the context (a batch: more than one)
  std::vector<SolveP> sol_host;  // batch: host copy of the solve table (device: geo.sol)
  csu_params prm;
  KP kp;
  Geo geo;
  St st;
  hipStream_t stream;
  DevBuf buf;
  int nwall;
  int* wall;                   // [nwall] boundary ids of wall/wallT faces
  double *U, *Uaux, *R, *dtl, *dtg;
  double* red;                 // reduction scratch
  int nred;                    // number of partial blocks of a cell launch
  double *res_hist, *coef_hist;
  int hist_cap;
  int fresh;                   // state set from the host: P and dt not yet derived
  int pu;                      // RK stages recover the primitives from the state (k_grad's PU form)
  int sblk;                    // threads per workgroup of k_grad and k_flux
  int lblk;                    // threads per workgroup of an LU-SGS level launch
  int tile_ok;                 // the tile form's LDS list length (CSU_TILE_FACES_S or CSU_TILE_FACES), 0: no tile form
  int tile;                    // face fluxes evaluated per cell tile inside k_update, no k_flux (1)
  int tuned;                   // launch shapes from the autotuner (1) or the defaults (0)
  // implicit schemes
  double *Dp, *dW0, *dW1, *A;  // diagonal (without the dt term), increments, DPLUR blocks
  double* caux;                // LU-SGS: per cell c, max(4/(3r), gamma/r), mu/Pr of the iteration's state [3][nc]
  int* jstate;                 // DPLUR: [unused, buffer holding the last result]
  double* jred;                // DPLUR: per-sweep block maxima of |dx_rho| [nlinear][cell blocks]
  std::vector<int> lvlL, lvlU; // LU-SGS level offsets (forward / backward)
  int *cellsL, *cellsU;        // cells sorted by level
  // iteration graphs (captured once, replayed): RK keyed by local time
  // step x which buffer holds U; implicit keyed by nothing
  int* hctr;                   // device iteration counter: slot of res/coef history
  hipGraphExec_t gexec[4];
  double* gU[4];               // d->U at capture
  double* Ubase;               // the buffer U starts in (graph parity key)
  size_t bytes;                // device bytes allocated by this context
  // asynchronous iteration: history rows copied to pinned host memory
  double *hres, *hcoef;
  int hcap, async_start, async_n;
  int use_graphs;
  double* tstate;              // time-accurate run: t, tmax, state, steps (k_time_clamp)
  // profiling
  int prof;
  std::vector<std::string> kname;
  std::vector<double> ktime;
  std::vector<int> kcount;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> evpool;
};

static inline int nblk(int n) { return (n + CSU_BLOCK - 1) / CSU_BLOCK; }

template <typename T>
static int dalloc(csu_dev* d, T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(CSU_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  }
  d->buf.owned.push_back((void*)*p);
  d->bytes += n * sizeof(T);
  return CSU_OK;
}

// Free one of the context's own allocations before the context ends.
static void dfree(csu_dev* d, void* p, size_t bytes) {
  if (!p) return;
  auto it = std::find(d->buf.owned.begin(), d->buf.owned.end(), p);
  if (it == d->buf.owned.end()) return;
  d->buf.owned.erase(it);
  (void)hipFree(p);
  d->bytes -= bytes;
}

template <typename T>
static int upload(csu_dev* d, const T** dst, const std::vector<T>& h) {
  T* p;
  int rc = dalloc(d, &p, h.size());
  if (rc) return rc;
  // on the context's own (non-blocking) stream: no legacy-stream work that
  // would have to order itself after graph captures on other threads
  if (!h.empty()) {
    HIPCHK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));  // h may be a temporary
  }
  *dst = p;
  return CSU_OK;
}

static int validate(const csu_mesh_desc* m, const csu_params* p) {
  if (!m || !p) return set_err(CSU_ERR_ARG, "null mesh or params");
  if (m->ncell <= 0 || m->nnode <= 0 || m->nface < 0 || m->nbface < 0 || m->nmark < 0)
    return set_err(CSU_ERR_ARG, "bad mesh sizes");
  if (!m->node_xy || !m->cell_vert || !m->cell_nbr || !m->cell_face || (m->nface && !m->face) ||
      (m->nbface && (!m->bface || !m->bface_mark || !m->mark_type)))
    return set_err(CSU_ERR_ARG, "null mesh array");
  if (p->nvar < 4 || p->nvar > 6) return set_err(CSU_ERR_ARG, "nvar %d", p->nvar);
  for (int i = 0; i < m->ncell; i++) {
    const int32_t* v = m->cell_vert + 4 * i;
    int np = v[3] >= 0 ? 4 : 3;
    for (int k = 0; k < np; k++)
      if (v[k] < 0 || v[k] >= m->nnode) return set_err(CSU_ERR_ARG, "cell %d vertex out of range", i);
    if (v[3] < -1 || v[3] >= m->nnode) return set_err(CSU_ERR_ARG, "cell %d vertex out of range", i);
    for (int k = 0; k < 4; k++) {
      int n = m->cell_nbr[4 * i + k];
      if (k < np) {
        if (n >= m->ncell || n == INT32_MAX)
          return set_err(CSU_ERR_ARG, "cell %d has %d sides but neighbour %d is missing", i, np, k);
        if (n < 0 && (-1 - n) >= m->nbface) return set_err(CSU_ERR_ARG, "cell %d boundary neighbour out of range", i);
      }
      int f = m->cell_face[4 * i + k];
      if (f != 0 && (f > m->nface || -f > m->nface)) return set_err(CSU_ERR_ARG, "cell %d face id out of range", i);
    }
  }
  for (int f = 0; f < m->nface; f++) {
    const int32_t* q = m->face + 4 * f;
    if (q[0] < 0 || q[0] >= m->ncell || q[1] < 0 || q[1] >= m->ncell || q[2] < 0 || q[2] >= m->nnode ||
        q[3] < 0 || q[3] >= m->nnode)
      return set_err(CSU_ERR_ARG, "face %d out of range", f);
  }
  for (int g = 0; g < m->nbface; g++) {
    const int32_t* q = m->bface + 3 * g;
    if (q[0] < 0 || q[0] >= m->ncell || q[1] < 0 || q[1] >= m->nnode || q[2] < 0 || q[2] >= m->nnode)
      return set_err(CSU_ERR_ARG, "boundary face %d out of range", g);
    if (m->bface_mark[g] < 0 || m->bface_mark[g] >= m->nmark) return set_err(CSU_ERR_ARG, "boundary marker out of range");
    int t = m->mark_type[m->bface_mark[g]];
    if (t < 0 || t > 4) return set_err(CSU_ERR_ARG, "boundary type %d", t);
  }
  if (p->time_scheme == CSU_TIME_RK && !(p->stages == 1 || p->stages == 3 || p->stages == 4 || p->stages == 5))
    return set_err(CSU_ERR_UNSUPPORTED, "RK stages %d (the reference's stage tables are 1, 3, 4, 5)", p->stages);
  if (p->order != 1 && p->order != 2) return set_err(CSU_ERR_ARG, "order %d", p->order);
  if (p->flux < 0 || p->flux > 4) return set_err(CSU_ERR_ARG, "flux %d", p->flux);
  if (p->limiter < 0 || p->limiter > 2) return set_err(CSU_ERR_ARG, "limiter %d", p->limiter);
  if (p->gradient != 0 && p->gradient != 1) return set_err(CSU_ERR_ARG, "gradient %d", p->gradient);
  if (p->gradient == 1 && m->axi == 1)
    return set_err(CSU_ERR_UNSUPPORTED, "Green-Gauss gradients of an axisymmetric run");
  if (p->gradient == 1 && p->precision == CSU_PREC_FP64_LD)
    return set_err(CSU_ERR_UNSUPPORTED, "Green-Gauss gradients with the long double least-squares mode");
  if ((p->sa && p->laminar) || (p->sst && (p->laminar || p->sa)))
    return set_err(CSU_ERR_UNSUPPORTED, "more than one of laminar / sa / sst set");
  if (p->nvar != (p->sa ? 5 : p->sst ? 6 : 4)) return set_err(CSU_ERR_ARG, "nvar %d does not match the model", p->nvar);
  if (p->sst && p->order == 2 && p->limiter == 1)
    return set_err(CSU_ERR_UNSUPPORTED, "limiter 1 with SST: the reference's range arrays hold 5 variables (limiter.c:37-38)");
  if (p->time_scheme < 0 || p->time_scheme > 2) return set_err(CSU_ERR_ARG, "time scheme %d", p->time_scheme);
  if (p->time_scheme == CSU_TIME_DPLUR && p->nlinear < 0) return set_err(CSU_ERR_ARG, "Nlinear %d", p->nlinear);
  if (p->precision != CSU_PREC_FP64 && p->precision != CSU_PREC_FP64_LD && p->precision != CSU_PREC_MIXED)
    return set_err(CSU_ERR_ARG, "precision %d", p->precision);
  return CSU_OK;
}

static void graphs_drop(csu_dev* d) {
  for (int k = 0; k < 4; k++) {
    if (d->gexec[k]) (void)hipGraphExecDestroy(d->gexec[k]);
    d->gexec[k] = nullptr;
    d->gU[k] = nullptr;
  }
}

extern "C" void csu_dev_destroy(csu_dev* d) {
  if (!d) return;
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  graphs_drop(d);
  if (d->hres) (void)hipHostFree(d->hres);
  if (d->hcoef) (void)hipHostFree(d->hcoef);
  for (hipEvent_t e : d->evpool) (void)hipEventDestroy(e);
  for (auto& pe : d->pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
}

// geometry in the reference's expressions (mesh.c:371-441, 720-764)
static void elem_center(const csu_mesh_desc* m, int i, double* x, double* y) {
  const int32_t* v = m->cell_vert + 4 * i;
  int np = v[3] >= 0 ? 4 : 3;
  *x = 0.0; *y = 0.0;
  for (int k = 0; k < np; k++) { *x += m->node_xy[2 * v[k]]; *y += m->node_xy[2 * v[k] + 1]; }
  *x /= np; *y /= np;
}
static void calc_ds(const csu_mesh_desc* m, int p0, int p1, double* dSx, double* dSy) {
  double x0 = m->node_xy[2 * p0], y0 = m->node_xy[2 * p0 + 1];
  double x1 = m->node_xy[2 * p1], y1 = m->node_xy[2 * p1 + 1];
  *dSx = y1 - y0;
  *dSy = -(x1 - x0);
  double y = (y0 + y1) / 2;
  if (m->axi == 1) { *dSx = (*dSx) * y; *dSy = (*dSy) * y; }
}
static double tri_omega(const csu_mesh_desc* m, int p0, int p1, int p2) {
  const double* q = m->node_xy;
  double x1 = q[2 * p0], x2 = q[2 * p1], x3 = q[2 * p2], y1 = q[2 * p0 + 1], y2 = q[2 * p1 + 1], y3 = q[2 * p2 + 1];
  return 0.5 * ((x1 - x2) * (y1 + y2) + (x2 - x3) * (y2 + y3) + (x3 - x1) * (y3 + y1));
}

// Wall segments (p0x, p0y, p1x, p1y) SoA [4][nw] of the wall / wallT
// boundary faces, in boundary-face order.
static std::vector<double> wall_segments(const csu_mesh_desc* m, int* nw) {
  std::vector<int> w;
  for (int g = 0; g < m->nbface; g++) {
    const int t = m->mark_type[m->bface_mark[g]];
    if (t == CSU_BC_WALL || t == CSU_BC_WALLT) w.push_back(g);
  }
  const size_t n = w.size();
  std::vector<double> seg(4 * (n ? n : 1));
  for (size_t k = 0; k < n; k++) {
    const int32_t* q = m->bface + 3 * w[k];
    seg[k] = m->node_xy[2 * q[1]];
    seg[n + k] = m->node_xy[2 * q[1] + 1];
    seg[2 * n + k] = m->node_xy[2 * q[2]];
    seg[3 * n + k] = m->node_xy[2 * q[2] + 1];
  }
  *nw = (int)n;
  return seg;
}

// Wall distance on the device into dist[nc] (device memory), from device
// cell centres; synchronous on `st`.
static int wall_distance_launch(const csu_mesh_desc* m, hipStream_t st, const double* xc, const double* yc,
                                double* dist) {
  int nw = 0;
  std::vector<double> seg = wall_segments(m, &nw);
  double* dseg = nullptr;
  if (hipMalloc(&dseg, seg.size() * sizeof(double)) != hipSuccess)
    return set_err(CSU_ERR_NOMEM, "hipMalloc of %zu wall-segment bytes failed", seg.size() * sizeof(double));
  int rc = CSU_OK;
  if (hipMemcpyAsync(dseg, seg.data(), seg.size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
    rc = set_err(CSU_ERR_HIP, "wall-segment upload failed");
  if (!rc && m->ncell > 0) {
    hipLaunchKernelGGL(k_wall_dist, dim3(nblk(m->ncell)), dim3(CSU_BLOCK), 0, st, m->ncell, xc, yc, nw, dseg, dist);
    if (hipGetLastError() != hipSuccess) rc = set_err(CSU_ERR_HIP, "k_wall_dist launch failed");
  }
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = set_err(CSU_ERR_HIP, "k_wall_dist failed");
  (void)hipFree(dseg);
  return rc;
}

extern "C" int csu_wall_distance_dev(const csu_mesh_desc* m, double* wd) {
  if (!m || !wd || m->ncell <= 0 || !m->node_xy || !m->cell_vert) return set_err(CSU_ERR_ARG, "bad mesh or output");
  if (m->nbface > 0 && (!m->bface || !m->bface_mark || !m->mark_type)) return set_err(CSU_ERR_ARG, "bad boundary");
  const int nc = m->ncell;
  std::vector<double> xy(2 * (size_t)nc);
  for (int i = 0; i < nc; i++) elem_center(m, i, &xy[i], &xy[nc + i]);
  double *dxy = nullptr, *dd = nullptr;
  if (hipMalloc(&dxy, xy.size() * sizeof(double)) != hipSuccess) return set_err(CSU_ERR_NOMEM, "hipMalloc failed");
  if (hipMalloc(&dd, (size_t)nc * sizeof(double)) != hipSuccess) {
    (void)hipFree(dxy);
    return set_err(CSU_ERR_NOMEM, "hipMalloc failed");
  }
  int rc = CSU_OK;
  hipStream_t st = nullptr;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) rc = set_err(CSU_ERR_HIP, "stream");
  if (!rc && hipMemcpyAsync(dxy, xy.data(), xy.size() * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
    rc = set_err(CSU_ERR_HIP, "centre upload failed");
  if (!rc) rc = wall_distance_launch(m, st, dxy, dxy + nc, dd);
  if (!rc && hipMemcpyAsync(wd, dd, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    rc = set_err(CSU_ERR_HIP, "wall-distance download failed");
  if (st) {
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = set_err(CSU_ERR_HIP, "wall distance");
    (void)hipStreamDestroy(st);
  }
  (void)hipFree(dxy);
  (void)hipFree(dd);
  return rc;
}

// Host-side arrays of one mesh in the device layout, or of a batch: several
// meshes concatenated, ids offset, each solve's cells, interior faces and
// boundary faces starting on a multiple of CSU_CHUNK (the elements between
// are padding that no kernel computes).
struct HostMesh {
  int nc = 0, nf = 0, nb = 0;
  std::vector<int> cnp, cnbr, cface;                              // [nc], [4][nc], [4][nc]
  std::vector<int> cfn;  // [4][nc] the other cell of face slot j, -1 where cface is 0
  std::vector<double> xc, yc, om, dsxm, dsym, dslat;              // [nc]
  std::vector<double> edx, edy, lwx, lwy;                         // [4][nc]
  std::vector<unsigned char> fsl, bmask;                          // [nf], [nc]
  std::vector<int> fe0, fe1;                                      // [nf]
  std::vector<double> fdsx, fdsy, fds;                            // [nf]
  std::vector<double> fnx, fny, fdd;  // [nf] unit normal, centre distance of the face's cells (LU-SGS)
  std::vector<int> be0, btype, bdonor;                            // [nb]
  std::vector<double> bdsx, bdsy, bds, bxm, bym, bnx, bny, bds2;  // [nb]
  std::vector<int> wall;                                          // boundary ids of wall / wallT faces
  std::vector<int> levL, levU;  // LU-SGS: dependency level of each cell, forward / backward sweep (-1: padding)
  double volmax = 0.0;
};

// Geometry and connectivity of one mesh in the reference's expressions
// (mesh.c:371-441, 720-764), the device face order, least-squares weights,
// boundary data and the LU-SGS levels.
static int host_mesh(const csu_mesh_desc* m, const csu_params* p, HostMesh& h) {
  const int nc = m->ncell, nf = m->nface, nb = m->nbface;
  h.nc = nc; h.nf = nf; h.nb = nb;
  const size_t c4 = 4 * (size_t)nc;
  h.cnp.assign(nc, 0); h.cnbr.assign(c4, 0); h.cface.assign(c4, 0);
  h.xc.assign(nc, 0.0); h.yc.assign(nc, 0.0); h.om.assign(nc, 0.0);
  h.dsxm.assign(nc, 0.0); h.dsym.assign(nc, 0.0); h.dslat.assign(nc, 0.0);
  h.edx.assign(c4, 0.0); h.edy.assign(c4, 0.0); h.lwx.assign(c4, 0.0); h.lwy.assign(c4, 0.0);
  h.volmax = 0.0;
  for (int i = 0; i < nc; i++) {
    const int32_t* v = m->cell_vert + 4 * i;
    int np = v[3] >= 0 ? 4 : 3;
    h.cnp[i] = np;
    elem_center(m, i, &h.xc[i], &h.yc[i]);
    double a = tri_omega(m, v[0], v[1], v[2]);
    if (np == 4) a += tri_omega(m, v[2], v[3], v[0]);
    h.dslat[i] = a;
    if (m->axi == 1) a *= h.yc[i];
    h.om[i] = a;
    h.volmax = fmax(a, h.volmax);
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j < np; j++) {
      int p0 = v[j], p1 = v[(j + 1) % np];
      double dSx, dSy;
      calc_ds(m, p0, p1, &dSx, &dSy);
      sx += fabs(dSx);
      sy += fabs(dSy);
      double xm = (m->node_xy[2 * p0] + m->node_xy[2 * p1]) * 0.5;
      double ym = (m->node_xy[2 * p0 + 1] + m->node_xy[2 * p1 + 1]) * 0.5;
      h.edx[(size_t)j * nc + i] = xm - h.xc[i];
      h.edy[(size_t)j * nc + i] = ym - h.yc[i];
    }
    h.dsxm[i] = sx * 0.5;
    h.dsym[i] = sy * 0.5;
    for (int j = 0; j < 4; j++) {
      h.cnbr[(size_t)j * nc + i] = m->cell_nbr[4 * i + j];
      h.cface[(size_t)j * nc + i] = m->cell_face[4 * i + j];
    }
  }
  // Device face order: sorted by (e0, e1) so that consecutive faces touch
  // neighbouring cells (the reference's hash order scatters them).  Cells
  // keep their face lists in the reference's order; only the ids change, so
  // every per-cell sum keeps its order.
  std::vector<int> ford(nf), finv(nf);
  for (int f = 0; f < nf; f++) ford[f] = f;
  std::stable_sort(ford.begin(), ford.end(), [m](int a, int b) {
    const int32_t *qa = m->face + 4 * a, *qb = m->face + 4 * b;
    return qa[0] != qb[0] ? qa[0] < qb[0] : qa[1] < qb[1];
  });
  for (int f = 0; f < nf; f++) finv[ford[f]] = f;
  for (size_t t = 0; t < h.cface.size(); t++) {
    int f = h.cface[t];
    if (f > 0) h.cface[t] = finv[f - 1] + 1;
    else if (f < 0) h.cface[t] = -(finv[-f - 1] + 1);
  }
  // Green-Gauss (gradient 1): gx = sum_j (u_f - u_i) sx_j / Omega over the
  // cell's faces, u_f the average of the two sides, sx_j the face vector
  // pointing out of the cell (the closed contour's vectors sum to zero, so
  // the difference form equals the face-value form and keeps a uniform
  // field's gradient exactly 0): weights 0.5 sx_j / Omega per slot.
  if (p->gradient == 1) {
    for (int i = 0; i < nc; i++)
      for (int j = 0; j < h.cnp[i]; j++) {
        const int n = m->cell_nbr[4 * i + j];
        double sx, sy;
        if (n >= 0) {
          const int f = m->cell_face[4 * i + j], fi = f > 0 ? f - 1 : -f - 1;
          const int32_t* q = m->face + 4 * fi;
          calc_ds(m, q[2], q[3], &sx, &sy);
          if (f < 0) { sx = -sx; sy = -sy; }
        } else {
          const int32_t* q = m->bface + 3 * (-1 - n);
          calc_ds(m, q[1], q[2], &sx, &sy);
        }
        h.lwx[(size_t)j * nc + i] = 0.5 * sx / h.om[i];
        h.lwy[(size_t)j * nc + i] = 0.5 * sy / h.om[i];
      }
  }
  // Least-squares weights per neighbour slot (solverCalcGrad3, solver.c:1031-1074):
  // gx = sum_j wx_j du_j with wx_j = (c dx_j - b dy_j)/det, wy_j = (a dy_j - b dx_j)/det,
  // a, b, c, det in long double from the products the reference forms in double.
  for (int i = 0; i < nc && p->gradient == 0; i++) {
    double dx[4], dy[4];
    long double a = 0, b = 0, c = 0;
    for (int j = 0; j < h.cnp[i]; j++) {
      const int n = m->cell_nbr[4 * i + j];
      double xn, yn;
      if (n >= 0) { xn = h.xc[n]; yn = h.yc[n]; }
      else {
        const int32_t* q = m->bface + 3 * (-1 - n);
        xn = (m->node_xy[2 * q[1]] + m->node_xy[2 * q[2]]) / 2;
        yn = (m->node_xy[2 * q[1] + 1] + m->node_xy[2 * q[2] + 1]) / 2;
      }
      dx[j] = xn - h.xc[i];
      dy[j] = yn - h.yc[i];
      a += dx[j] * dx[j];
      b += dx[j] * dy[j];
      c += dy[j] * dy[j];
    }
    const long double det = a * c - b * b;
    for (int j = 0; j < h.cnp[i]; j++) {
      h.lwx[(size_t)j * nc + i] = (double)((c * dx[j] - b * dy[j]) / det);
      h.lwy[(size_t)j * nc + i] = (double)((a * dy[j] - b * dx[j]) / det);
    }
  }
  // edge index of a face in a cell's vertex loop (the limiter and the
  // reconstruction offsets are per edge)
  auto edge_of = [m](int cell, int p0, int p1) -> int {
    const int32_t* v = m->cell_vert + 4 * cell;
    const int np = v[3] >= 0 ? 4 : 3;
    for (int j = 0; j < np; j++) {
      const int a = v[j], b = v[(j + 1) % np];
      if ((a == p0 && b == p1) || (a == p1 && b == p0)) return j;
    }
    return -1;
  };
  h.fsl.assign(nf, 0);
  h.bmask.assign(nc, 0);
  for (int i = 0; i < nc; i++)
    for (int j = 0; j < h.cnp[i]; j++)
      if (m->cell_nbr[4 * i + j] < 0) h.bmask[i] |= (unsigned char)(1 << j);
  h.fe0.assign(nf, 0); h.fe1.assign(nf, 0);
  h.fdsx.assign(nf, 0.0); h.fdsy.assign(nf, 0.0); h.fds.assign(nf, 0.0);
  for (int f = 0; f < nf; f++) {
    const int32_t* q = m->face + 4 * ford[f];
    h.fe0[f] = q[0]; h.fe1[f] = q[1];
    const int s0 = edge_of(q[0], q[2], q[3]), s1 = edge_of(q[1], q[2], q[3]);
    if (s0 < 0 || s1 < 0)
      return set_err(CSU_ERR_ARG, "face %d (%d-%d) is not an edge of its cells %d, %d", f, q[2], q[3], q[0], q[1]);
    h.fsl[f] = (unsigned char)(s0 | (s1 << 2));
    double dSx, dSy;
    calc_ds(m, q[2], q[3], &dSx, &dSy);
    h.fdsx[f] = dSx; h.fdsy[f] = dSy; h.fds[f] = sqrt(dSx * dSx + dSy * dSy);
  }
  // the implicit kernels' neighbour of each face slot, so that their
  // neighbour gathers do not wait for a face-array lookup
  h.cfn.assign(c4, -1);
  for (size_t t = 0; t < c4; t++) {
    const int f = h.cface[t];
    if (f != 0) h.cfn[t] = f > 0 ? h.fe1[f - 1] : h.fe0[-f - 1];
  }
  // LU-SGS face terms that depend on the geometry only, in the sweep's own
  // expressions (implicit.c:288-313, 555-595): n = dS/|dS| (seen from e1 it
  // is the exact negation) and the distance of the two cell centres (the
  // same from either side: the differences only change sign)
  h.fnx.assign(nf, 0.0); h.fny.assign(nf, 0.0); h.fdd.assign(nf, 0.0);
  for (int f = 0; f < nf; f++) {
    h.fnx[f] = h.fdsx[f] / h.fds[f];
    h.fny[f] = h.fdsy[f] / h.fds[f];
    const double ddx = h.xc[h.fe1[f]] - h.xc[h.fe0[f]], ddy = h.yc[h.fe1[f]] - h.yc[h.fe0[f]];
    h.fdd[f] = sqrt(ddx * ddx + ddy * ddy);
  }
  h.be0.assign(nb, 0); h.btype.assign(nb, 0); h.bdonor.assign(nb, -1);
  for (auto* v : {&h.bdsx, &h.bdsy, &h.bds, &h.bxm, &h.bym, &h.bnx, &h.bny, &h.bds2}) v->assign(nb, 0.0);
  h.wall.clear();
  int last_writer = -1;  // last boundary face before this one that writes an energy viscous flux
  for (int g = 0; g < nb; g++) {
    const int32_t* q = m->bface + 3 * g;
    h.be0[g] = q[0];
    h.btype[g] = m->mark_type[m->bface_mark[g]];
    double dSx, dSy;
    calc_ds(m, q[1], q[2], &dSx, &dSy);
    h.bdsx[g] = dSx; h.bdsy[g] = dSy; h.bds[g] = sqrt(dSx * dSx + dSy * dSy);
    // boundary element centre (elementCenter of a 2-point element)
    h.bxm[g] = (m->node_xy[2 * q[1]] + m->node_xy[2 * q[2]]) / 2;
    h.bym[g] = (m->node_xy[2 * q[1] + 1] + m->node_xy[2 * q[2] + 1]) / 2;
    // meshCalcDS2 (mesh.c:742-764): unit normal, length scaled for axi
    {
      double x0 = m->node_xy[2 * q[1]], y0 = m->node_xy[2 * q[1] + 1];
      double x1 = m->node_xy[2 * q[2]], y1 = m->node_xy[2 * q[2] + 1];
      double sx = y1 - y0, sy = -(x1 - x0), sl = sqrt(sx * sx + sy * sy);
      h.bnx[g] = sx / sl; h.bny[g] = sy / sl;
      h.bds2[g] = m->axi == 1 ? sl * (y0 + y1) * 0.5 : sl;
    }
    if (h.btype[g] == CSU_BC_WALL || h.btype[g] == CSU_BC_WALLT) h.wall.push_back(g);
    if (h.btype[g] == CSU_BC_WALL) h.bdonor[g] = last_writer;
    else last_writer = g;
  }
  h.levL.clear();
  h.levU.clear();
  if (p->time_scheme == CSU_TIME_LUSGS) {
    // levels of the sweeps' dependency DAGs (lower / upper cell index)
    for (int dir = 0; dir < 2; dir++) {
      std::vector<int>& lev = dir == 0 ? h.levL : h.levU;
      lev.assign(nc, 0);
      for (int t = 0; t < nc; t++) {
        int i = dir == 0 ? t : nc - 1 - t, l = 0;
        for (int j = 0; j < 4; j++) {
          int n = m->cell_nbr[4 * i + j];
          if (m->cell_face[4 * i + j] == 0 || n < 0) continue;
          if (dir == 0 ? n < i : n > i) l = std::max(l, lev[n] + 1);
        }
        lev[i] = l;
      }
    }
  }
  return CSU_OK;
}

static size_t up_chunk(size_t n) { return (n + CSU_CHUNK - 1) / CSU_CHUNK * CSU_CHUNK; }

// Concatenate the meshes of a batch (ids offset by each solve's start) and
// fill the element ranges of the solve table.
static void host_concat(const std::vector<HostMesh>& part, HostMesh& H, std::vector<SolveP>& sol) {
  const size_t n = part.size();
  sol.assign(n, SolveP{});
  size_t NC = 0, NF = 0, NB = 0, NW = 0;
  for (size_t s = 0; s < n; s++) {
    const HostMesh& h = part[s];
    SolveP& q = sol[s];
    q.c0 = (int)NC; q.c1 = (int)(NC + h.nc); NC += up_chunk(h.nc);
    q.f0 = (int)NF; q.f1 = (int)(NF + h.nf); NF += up_chunk(h.nf);
    q.b0 = (int)NB; q.b1 = (int)(NB + h.nb); NB += up_chunk(h.nb);
    q.w0 = (int)NW; q.w1 = (int)(NW + h.wall.size()); NW += h.wall.size();
    q.active = 1;
    q.volmax = h.volmax;
  }
  H.nc = (int)NC; H.nf = (int)NF; H.nb = (int)NB;
  const size_t C4 = 4 * NC;
  // padding: no kernel computes it; finite values keep any stray read harmless
  H.cnp.assign(NC, 0); H.cnbr.assign(C4, INT32_MAX); H.cface.assign(C4, 0); H.cfn.assign(C4, -1);
  H.xc.assign(NC, 0.0); H.yc.assign(NC, 0.0); H.om.assign(NC, 1.0);
  H.dsxm.assign(NC, 1.0); H.dsym.assign(NC, 1.0); H.dslat.assign(NC, 0.0);
  H.edx.assign(C4, 0.0); H.edy.assign(C4, 0.0); H.lwx.assign(C4, 0.0); H.lwy.assign(C4, 0.0);
  H.fsl.assign(NF ? NF : 1, 0); H.bmask.assign(NC, 0);
  H.fe0.assign(NF, 0); H.fe1.assign(NF, 0);
  H.fdsx.assign(NF, 0.0); H.fdsy.assign(NF, 0.0); H.fds.assign(NF, 1.0);
  H.fnx.assign(NF, 0.0); H.fny.assign(NF, 0.0); H.fdd.assign(NF, 1.0);
  H.be0.assign(NB, 0); H.btype.assign(NB, 0); H.bdonor.assign(NB, -1);
  for (auto* v : {&H.bdsx, &H.bdsy, &H.bxm, &H.bym, &H.bnx, &H.bny}) v->assign(NB, 0.0);
  H.bds.assign(NB, 1.0); H.bds2.assign(NB, 1.0);
  H.wall.clear();
  const bool lev = !part.empty() && !part[0].levL.empty();
  H.levL.assign(lev ? NC : 0, -1); H.levU.assign(lev ? NC : 0, -1);
  H.volmax = part.empty() ? 0.0 : part[0].volmax;
  for (size_t s = 0; s < n; s++) {
    const HostMesh& h = part[s];
    const int c0 = sol[s].c0, f0 = sol[s].f0, b0 = sol[s].b0, nc = h.nc;
    for (int i = 0; i < nc; i++) {
      const size_t gi = (size_t)c0 + i;
      H.cnp[gi] = h.cnp[i]; H.xc[gi] = h.xc[i]; H.yc[gi] = h.yc[i]; H.om[gi] = h.om[i];
      H.dsxm[gi] = h.dsxm[i]; H.dsym[gi] = h.dsym[i]; H.dslat[gi] = h.dslat[i]; H.bmask[gi] = h.bmask[i];
      for (int j = 0; j < 4; j++) {
        const size_t lj = (size_t)j * nc + i, gj = (size_t)j * NC + gi;
        const int nbr = h.cnbr[lj], f = h.cface[lj];
        H.cnbr[gj] = nbr == INT32_MAX ? nbr : nbr >= 0 ? nbr + c0 : -1 - ((-1 - nbr) + b0);
        H.cface[gj] = f > 0 ? f + f0 : f < 0 ? f - f0 : 0;
        H.cfn[gj] = h.cfn[lj] >= 0 ? h.cfn[lj] + c0 : -1;
        H.edx[gj] = h.edx[lj]; H.edy[gj] = h.edy[lj]; H.lwx[gj] = h.lwx[lj]; H.lwy[gj] = h.lwy[lj];
      }
      if (lev) { H.levL[gi] = h.levL[i]; H.levU[gi] = h.levU[i]; }
    }
    for (int f = 0; f < h.nf; f++) {
      const size_t gf = (size_t)f0 + f;
      H.fe0[gf] = h.fe0[f] + c0; H.fe1[gf] = h.fe1[f] + c0; H.fsl[gf] = h.fsl[f];
      H.fdsx[gf] = h.fdsx[f]; H.fdsy[gf] = h.fdsy[f]; H.fds[gf] = h.fds[f];
      H.fnx[gf] = h.fnx[f]; H.fny[gf] = h.fny[f]; H.fdd[gf] = h.fdd[f];
    }
    for (int g = 0; g < h.nb; g++) {
      const size_t gg = (size_t)b0 + g;
      H.be0[gg] = h.be0[g] + c0; H.btype[gg] = h.btype[g];
      H.bdonor[gg] = h.bdonor[g] >= 0 ? h.bdonor[g] + b0 : -1;
      H.bdsx[gg] = h.bdsx[g]; H.bdsy[gg] = h.bdsy[g]; H.bds[gg] = h.bds[g]; H.bxm[gg] = h.bxm[g];
      H.bym[gg] = h.bym[g]; H.bnx[gg] = h.bnx[g]; H.bny[gg] = h.bny[g]; H.bds2[gg] = h.bds2[g];
    }
    for (int w : h.wall) H.wall.push_back(w + b0);
  }
}

// Limiter type 0 reference values of a free stream (limiter.c:45-59), [6].
static void limiter_pref(const csu_params* p, double* pref) {
  const int nv = p->nvar;
  for (int k = 0; k < 6; k++) pref[k] = 0.0;
  pref[0] = p->pin[0] * p->pin[0];
  pref[1] = p->pin[1] * p->pin[1] + p->pin[2] * p->pin[2];
  pref[2] = pref[1];
  pref[3] = p->pin[3] * p->pin[3];
  for (int k = 4; k < nv; k++) pref[k] = p->pin[k + 1] * p->pin[k + 1];
}

// ---------------------------------------------------------------------------
// launch-shape autotuner
// ---------------------------------------------------------------------------
// Launch shapes that do not change any result (no reduction depends on
// them): workgroup size of the gradient and face kernels, of the LU-SGS
// level launches, and the XCD-contiguous tile map.  The tuner times each
// candidate on a context and keeps the fastest; the choice is remembered
// per kind of run and mesh-size bucket (log2 of the cell count) and taken
// by every later context of that bucket in the process.
struct TuneKey {
  int nv, visc, scheme, prec, order, bucket;
  bool operator<(const TuneKey& o) const {
    return std::tie(nv, visc, scheme, prec, order, bucket) < std::tie(o.nv, o.visc, o.scheme, o.prec, o.order, o.bucket);
  }
};
struct TuneVal { int sblk, lblk, xcd, tile; };
static std::mutex g_tune_mu;
static std::map<TuneKey, TuneVal> g_tune;

static TuneKey tune_key(const csu_dev* d) {
  int b = 0;
  while ((1 << (b + 1)) <= d->nc && b < 30) b++;
  return TuneKey{d->nv, d->visc, d->prm.time_scheme, d->prm.precision, d->prm.order, b};
}

static void tune_lookup(csu_dev* d) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find(tune_key(d));
  if (it == g_tune.end()) return;
  d->sblk = it->second.sblk;
  d->lblk = it->second.lblk;
  d->tile = it->second.tile;  // dropped again by dev_setup when the context has no tile lists
  d->kp.xcd = it->second.xcd;
  d->tuned = 1;
}

// Face lists of the tile form (k_update TILE): the interior faces of each
// 256-cell tile (of a batch's solves: their real faces only), a face shared
// by two tiles in both lists, and each cell slot's face as its place in its
// own tile's list (signed like cface: + where the cell is the face's e0).
// False when a tile has more faces than the workgroup's LDS list holds.
static bool tile_lists(const HostMesh& H, const std::vector<SolveP>* sol, std::vector<int>& toff,
                       std::vector<int>& tface, std::vector<short>& tsl, int* maxf) {
  const int nc = H.nc, nt = (nc + CSU_BLOCK - 1) / CSU_BLOCK;
  auto each_face = [&](auto fn) {
    if (sol && !sol->empty()) {
      for (const SolveP& q : *sol)
        for (int f = q.f0; f < q.f1; f++) fn(f);
    } else {
      for (int f = 0; f < H.nf; f++) fn(f);
    }
  };
  std::vector<int> cnt(nt, 0);
  each_face([&](int f) {
    const int a = H.fe0[f] / CSU_BLOCK, b = H.fe1[f] / CSU_BLOCK;
    cnt[a]++;
    if (b != a) cnt[b]++;
  });
  toff.assign(nt + 1, 0);
  *maxf = 0;
  for (int t = 0; t < nt; t++) {
    if (cnt[t] > CSU_TILE_FACES) return false;
    *maxf = std::max(*maxf, cnt[t]);
    toff[t + 1] = toff[t] + cnt[t];
  }
  tface.assign(toff[nt] > 0 ? toff[nt] : 1, 0);
  std::vector<int> fill(toff.begin(), toff.end() - 1), loc0(H.nf > 0 ? H.nf : 1, -1), loc1(H.nf > 0 ? H.nf : 1, -1);
  each_face([&](int f) {
    const int a = H.fe0[f] / CSU_BLOCK, b = H.fe1[f] / CSU_BLOCK;
    loc0[f] = fill[a] - toff[a];
    tface[fill[a]++] = f;
    if (b != a) {
      loc1[f] = fill[b] - toff[b];
      tface[fill[b]++] = f;
    } else {
      loc1[f] = loc0[f];
    }
  });
  tsl.assign(4 * (size_t)nc, 0);
  for (int j = 0; j < 4; j++)
    for (int i = 0; i < nc; i++) {
      const int cf = H.cface[(size_t)j * nc + i];
      if (cf == 0) continue;
      const int f = cf > 0 ? cf - 1 : -cf - 1;
      const int l = cf > 0 ? loc0[f] : loc1[f];  // cf > 0: i is the face's e0
      const int t = i / CSU_BLOCK;
      if (l < 0 || l >= cnt[t] || (cf > 0 ? H.fe0[f] : H.fe1[f]) != i) return false;
      tsl[(size_t)j * nc + i] = (short)(cf > 0 ? l + 1 : -(l + 1));
    }
  return true;
}

// Device state of a context from its host arrays: uploads, allocations, wall
// distance, initial free stream.  ms / ps: the ns meshes and options
// (options that change the kernels are those of ps[0]); sol: the batch's
// solve table (nullptr: one solve).
static int dev_setup(csu_dev* d, const HostMesh& H, int ns, const csu_mesh_desc* const* ms,
                     const csu_params* const* ps, std::vector<SolveP>* sol) {
  const csu_params* p = ps[0];
  const int nc = H.nc, nf = H.nf, nb = H.nb, nv = p->nvar;
  int rc;
  d->nc = nc; d->nf = nf; d->nb = nb; d->nv = nv; d->nsolve = ns;
  d->nmark = ms[0]->nmark; d->iter = 0; d->prm = *p; d->prof = 0;
  d->visc = (p->laminar || p->sa || p->sst) ? 1 : 0;
  d->nwall = (int)H.wall.size();

  // ---- kernel parameters ----
  KP& kp = d->kp;
  kp.nc = nc; kp.nf = nf; kp.nb = nb; kp.nv = nv; kp.order = p->order; kp.flux = p->flux;
  kp.cend = nc; kp.fend = nf; kp.bend = nb;
  kp.laminar = p->laminar; kp.sa = p->sa; kp.sst = p->sst; kp.limiter = p->limiter;
  kp.sst_trans = p->sst ? p->sst_trans : 0;
  kp.sa_cc = p->sa ? p->sa_cc : 0;
  kp.turb1order = p->turb1order; kp.axi = ms[0]->axi; kp.visc_blazek = p->visc_blazek;
  kp.lsq_ld = p->precision == CSU_PREC_FP64_LD; kp.store_grad = 0;
  kp.mixed = p->precision == CSU_PREC_MIXED;
  // small meshes: 64-thread workgroups for the gradient and face kernels
  // (no shared memory), so that a few hundred cell tiles reach every CU
  d->sblk = nc <= 262144 ? 64 : CSU_BLOCK;
  d->lblk = 64;
  d->tile = 0;
  d->tile_ok = 0;
  kp.xcd = 1;
  d->tuned = 0;
  tune_lookup(d);
  // inviscid ideal-gas second-order RK with the default limiter (any
  // precision): the stage kernels recover the primitives from the state
  // (k_grad PU); first order, the global-range limiter and the axisymmetric
  // source read stored primitives, so those keep them.
  // CSU_STORED_PRIMITIVES=1 in the environment selects the stored form
  // everywhere (same results; for comparisons of the two forms).
  const char* sp = getenv("CSU_STORED_PRIMITIVES");
  d->pu = nv == 4 && !p->laminar && p->time_scheme == CSU_TIME_RK && p->order == 2 && p->limiter != 1 &&
          kp.axi == 0 && p->tp == 0 && !(sp && sp[0] == '1');
  kp.rlim = p->rlim; kp.pout = p->pout; kp.twall = p->twall; kp.efix = p->efix; kp.minf = p->mach_inf;
  kp.sst_owall = p->sst_owallfactor;
  {
    const double kap = 0.41;  // sstInit (sst.c:31-35)
    kp.sst_g1 = 0.075 / 0.09 - 0.5 * kap * kap / sqrt(0.09);
    kp.sst_g2 = 0.0828 / 0.09 - 0.856 * kap * kap / sqrt(0.09);
  }
  kp.limK = p->lim_k; kp.volmax = H.volmax; kp.pr = p->pr; kp.pr_t = p->pr_t; kp.wimp = p->wimp;
  for (int k = 0; k < 7; k++) kp.pin[k] = p->pin[k];
  kp.gas.gamma = p->gamma; kp.gas.R = p->gas_r;
  kp.gas.Cp = kp.gas.gamma * kp.gas.R / (kp.gas.gamma - 1);
  kp.gas.Cv = kp.gas.Cp - kp.gas.R;
  kp.gas.cc[0] = 1.81716e3; kp.gas.cc[1] = 9.890504e2 - 287.0530; kp.gas.cc[2] = -9.595592e-3;
  kp.gas.cc[3] = 1.041469e-4; kp.gas.cc[4] = -4.433065e-8; kp.gas.cc[5] = 5.879263e-12;
  kp.gas.tp = p->tp;
  std::vector<double> pref(6 * (size_t)ns, 0.0);
  for (int s = 0; s < ns; s++) limiter_pref(ps[s], &pref[6 * (size_t)s]);
  if (sol)
    for (int s = 0; s < ns; s++) {
      SolveP& q = (*sol)[s];
      q.cfl = ps[s]->cfl; q.sref = ps[s]->sref; q.pout = ps[s]->pout; q.twall = ps[s]->twall;
      q.minf = ps[s]->mach_inf;
      for (int k = 0; k < 7; k++) q.pin[k] = ps[s]->pin[k];
    }

  // ---- device allocation ----
  d->nred = nblk(nc);
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    d->stream = nullptr;
    return set_err(CSU_ERR_HIP, "hipStreamCreate failed");
  }
  Geo& g = d->geo;
  St& s = d->st;
  const size_t nvc = (size_t)nv * nc;
  const double* pref_d = nullptr;
  const int* wall_d = nullptr;
  if ((rc = upload(d, &g.cnp, H.cnp)) || (rc = upload(d, &g.cnbr, H.cnbr)) || (rc = upload(d, &g.cface, H.cface)) ||
      (rc = upload(d, &g.xc, H.xc)) || (rc = upload(d, &g.yc, H.yc)) || (rc = upload(d, &g.omega, H.om)) ||
      (rc = upload(d, &g.dsxm, H.dsxm)) || (rc = upload(d, &g.dsym, H.dsym)) || (rc = upload(d, &g.dslat, H.dslat)) ||
      (rc = upload(d, &g.edx, H.edx)) || (rc = upload(d, &g.edy, H.edy)) ||
      (rc = upload(d, &g.lwx, H.lwx)) || (rc = upload(d, &g.lwy, H.lwy)) || (rc = upload(d, &g.fsl, H.fsl)) ||
      (rc = upload(d, &g.bmask, H.bmask)) ||
      (rc = upload(d, &g.fe0, H.fe0)) || (rc = upload(d, &g.fe1, H.fe1)) || (rc = upload(d, &g.fdsx, H.fdsx)) ||
      (rc = upload(d, &g.fdsy, H.fdsy)) || (rc = upload(d, &g.fds, H.fds)) ||
      (rc = upload(d, &g.be0, H.be0)) || (rc = upload(d, &g.btype, H.btype)) || (rc = upload(d, &g.bdonor, H.bdonor)) ||
      (rc = upload(d, &g.bdsx, H.bdsx)) || (rc = upload(d, &g.bdsy, H.bdsy)) || (rc = upload(d, &g.bds, H.bds)) ||
      (rc = upload(d, &g.bxm, H.bxm)) || (rc = upload(d, &g.bym, H.bym)) || (rc = upload(d, &g.bnx, H.bnx)) ||
      (rc = upload(d, &g.bny, H.bny)) || (rc = upload(d, &g.bds2, H.bds2)) || (rc = upload(d, &wall_d, H.wall)) ||
      (rc = upload(d, &pref_d, pref)) ||
      (rc = dalloc(d, &d->U, nvc)) || (rc = dalloc(d, &d->Uaux, nvc)) || (rc = dalloc(d, &d->R, nvc)) ||
      (rc = dalloc(d, &d->dtl, nc + 1)) || (rc = dalloc(d, &d->dtg, ns)) ||
      (rc = dalloc(d, &d->red, 2 * (size_t)nv * d->nred + 16)) ||
      (rc = dalloc(d, &s.P, (size_t)(nv + 1) * nc)) || (rc = dalloc(d, &s.Pb, (size_t)(nv + 1) * nb)) ||
      (rc = dalloc(d, &s.dPx, nvc)) || (rc = dalloc(d, &s.dPy, nvc)) || (rc = dalloc(d, &s.phi, nvc)) ||
      (rc = dalloc(d, &s.fflux, (size_t)nv * nf)) || (rc = dalloc(d, &s.bflux, (size_t)nv * nb)))
    return rc;
  d->wall = const_cast<int*>(wall_d);
  s.pref2 = const_cast<double*>(pref_d);
  if (sol) {
    // the solve of every chunk of cells, faces and boundary faces, and of every cell
    std::vector<int> ck(nc / CSU_CHUNK + 1, 0), fk(nf / CSU_CHUNK + 1, 0), bk(nb / CSU_CHUNK + 1, 0), cs(nc, 0);
    for (int q = 0; q < ns; q++) {
      const SolveP& v = (*sol)[q];
      for (size_t e = v.c0; e < (size_t)v.c0 + up_chunk(v.c1 - v.c0); e++) cs[e] = q;
      for (size_t e = v.c0; e < (size_t)v.c0 + up_chunk(v.c1 - v.c0); e += CSU_CHUNK) ck[e / CSU_CHUNK] = q;
      for (size_t e = v.f0; e < (size_t)v.f0 + up_chunk(v.f1 - v.f0); e += CSU_CHUNK) fk[e / CSU_CHUNK] = q;
      for (size_t e = v.b0; e < (size_t)v.b0 + up_chunk(v.b1 - v.b0); e += CSU_CHUNK) bk[e / CSU_CHUNK] = q;
    }
    const SolveP* sol_d = nullptr;
    if ((rc = upload(d, &sol_d, *sol)) || (rc = upload(d, &g.csolk, ck)) || (rc = upload(d, &g.fsolk, fk)) ||
        (rc = upload(d, &g.bsolk, bk)) || (rc = upload(d, &g.csol, cs)))
      return rc;
    g.sol = sol_d;
    d->sol_host = *sol;
  }
  if (p->order == 2 && !kp.mixed && (rc = dalloc(d, &s.rec, 4 * nvc))) return rc;
  if (d->pu) {
    std::vector<int> toff, tface;
    std::vector<short> tsl;
    int maxf = 0;
    if (tile_lists(H, sol, toff, tface, tsl, &maxf)) {
      if ((rc = upload(d, &g.toff, toff)) || (rc = upload(d, &g.tface, tface)) || (rc = upload(d, &g.tsl, tsl)))
        return rc;
      d->tile_ok = maxf <= CSU_TILE_FACES_S ? CSU_TILE_FACES_S : CSU_TILE_FACES;
    }
  }
  if (!d->tile_ok) d->tile = 0;
  // untuned default: the tile form from 200k cells on, where it measured
  // 4-6 % faster in every shape (below that, launch latency decides)
  else if (!d->tuned && nc >= 200000) d->tile = 1;
  if (kp.mixed) {
    auto f32 = [](const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); };
    if ((rc = upload(d, &g.lwx32, f32(H.lwx))) || (rc = upload(d, &g.lwy32, f32(H.lwy))) ||
        (rc = upload(d, &g.edx32, f32(H.edx))) || (rc = upload(d, &g.edy32, f32(H.edy))) ||
        (rc = upload(d, &g.fdsx32, f32(H.fdsx))) || (rc = upload(d, &g.fdsy32, f32(H.fdsy))) ||
        (rc = upload(d, &g.fds32, f32(H.fds))) || (rc = dalloc(d, &s.fflux32, (size_t)nv * nf)) ||
        (p->order == 2 && (rc = dalloc(d, &s.rec32, 4 * nvc))))
      return rc;
  }
  if (p->sa || p->sst) {
    // wall distance of each solve to its own walls: the mesh's, or saCalcD
    // on the device
    double* wdd = nullptr;
    if ((rc = dalloc(d, &wdd, nc))) return rc;
    HIPCHK(hipMemsetAsync(wdd, 0, sizeof(double) * (size_t)nc, d->stream));
    for (int q = 0; q < ns; q++) {
      const int c0 = sol ? (*sol)[q].c0 : 0;
      if (ms[q]->wall_dist) {
        HIPCHK(hipMemcpyAsync(wdd + c0, ms[q]->wall_dist, sizeof(double) * ms[q]->ncell, hipMemcpyHostToDevice,
                              d->stream));
        HIPCHK(hipStreamSynchronize(d->stream));
      } else if ((rc = wall_distance_launch(ms[q], d->stream, g.xc + c0, g.yc + c0, wdd + c0))) {
        return rc;
      }
    }
    g.wdist = wdd;
  }
  if (d->visc) {
    if ((rc = dalloc(d, &s.dTx, nc)) || (rc = dalloc(d, &s.dTy, nc)) ||
        (rc = dalloc(d, &s.vflux, (size_t)(nv - 1) * nf)) || (rc = dalloc(d, &s.bvflux, (size_t)(nv - 1) * nb)) ||
        (rc = dalloc(d, &s.wfric, 2 * (size_t)nb)) || (rc = dalloc(d, &s.miT, nf)) ||
        (rc = dalloc(d, &s.src, 2 * (size_t)nc)) || (rc = dalloc(d, &s.dq, 6 * (size_t)nc)) ||
        (rc = dalloc(d, &s.wten, 4 * (size_t)nb)))
      return rc;
    HIPCHK(hipMemsetAsync(s.wfric, 0, sizeof(double) * 2 * (size_t)(nb ? nb : 1), d->stream));
  }
  if (p->time_scheme != CSU_TIME_RK) {
    if ((rc = dalloc(d, &d->Dp, nc)) || (rc = dalloc(d, &d->dW0, nvc)) || (rc = dalloc(d, &d->dW1, nvc)) ||
        (rc = upload(d, &g.cfn, H.cfn)))
      return rc;
    HIPCHK(hipMemsetAsync(d->dW0, 0, nvc * sizeof(double), d->stream));
    HIPCHK(hipMemsetAsync(d->dW1, 0, nvc * sizeof(double), d->stream));   // implicitInitDPLUR
  }
  if (p->time_scheme == CSU_TIME_DPLUR) {
    if ((rc = dalloc(d, &d->A, (size_t)4 * nv * nv * nc)) || (rc = dalloc(d, &d->jstate, 2 * (size_t)ns)) ||
        (rc = dalloc(d, &d->jred, (size_t)(p->nlinear > 0 ? p->nlinear : 1) * nblk(nc))))
      return rc;
    HIPCHK(hipMemsetAsync(d->A, 0, (size_t)4 * nv * nv * nc * sizeof(double), d->stream));
    HIPCHK(hipMemsetAsync(d->jstate, 0, 2 * sizeof(int) * (size_t)ns, d->stream));
  }
  if (p->time_scheme == CSU_TIME_LUSGS) {
    if ((rc = upload(d, &g.fnx, H.fnx)) || (rc = upload(d, &g.fny, H.fny)) || (rc = upload(d, &g.fdd, H.fdd)) ||
        (rc = dalloc(d, &d->caux, 3 * (size_t)nc)))
      return rc;
    // cells grouped by sweep level (the levels of a batch's solves merge:
    // cells of different solves never depend on each other)
    auto build = [&](const std::vector<int>& lev, std::vector<int>& off, int** dcells) -> int {
      int depth = 0;
      for (int i = 0; i < nc; i++) depth = std::max(depth, lev[i] + 1);
      off.assign(depth + 1, 0);
      for (int i = 0; i < nc; i++)
        if (lev[i] >= 0) off[lev[i] + 1]++;
      for (int l = 0; l < depth; l++) off[l + 1] += off[l];
      std::vector<int> cells(off[depth]), fill(off.begin(), off.end() - 1);
      for (int i = 0; i < nc; i++)
        if (lev[i] >= 0) cells[fill[lev[i]]++] = i;
      const int* dc = nullptr;
      int r = upload(d, &dc, cells);
      *dcells = const_cast<int*>(dc);
      return r;
    };
    if ((rc = build(H.levL, d->lvlL, &d->cellsL)) || (rc = build(H.levU, d->lvlU, &d->cellsU))) return rc;
  }
  d->Ubase = d->U;
  if ((rc = dalloc(d, &d->hctr, 1)) || (rc = dalloc(d, &d->tstate, 4))) return rc;
  HIPCHK(hipMemsetAsync(d->hctr, 0, sizeof(int), d->stream));
  // free-stream initial state (solverInitU); a batch's padding cells take
  // their solve's, so that every value in the arrays stays finite
  std::vector<double> U0(nvc);
  for (int q = 0; q < ns; q++) {
    const size_t c0 = sol ? (*sol)[q].c0 : 0, c1 = sol ? c0 + up_chunk((*sol)[q].c1 - c0) : nc;
    for (int k = 0; k < nv; k++)
      for (size_t i = c0; i < c1; i++) U0[(size_t)k * nc + i] = ps[q]->uin[k];
  }
  HIPCHK(hipMemcpyAsync(d->U, U0.data(), nvc * sizeof(double), hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipMemcpyAsync(d->Uaux, U0.data(), nvc * sizeof(double), hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipMemsetAsync(s.Pb, 0, sizeof(double) * (size_t)(nv + 1) * (nb ? nb : 1), d->stream));
  // this context's work only: other contexts' streams keep running
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

static csu_dev* dev_new() {
  csu_dev* d = new csu_dev();
  d->stream = nullptr; d->prof = 0; d->nsolve = 1;
  d->hist_cap = 0; d->res_hist = d->coef_hist = nullptr; d->fresh = 1;
  d->Dp = d->dW0 = d->dW1 = d->A = d->caux = nullptr; d->jstate = nullptr; d->jred = nullptr; d->cellsL = d->cellsU = nullptr;
  d->hctr = nullptr; d->use_graphs = 1; d->bytes = 0; d->tstate = nullptr;
  d->hres = d->hcoef = nullptr; d->hcap = 0; d->async_start = 0; d->async_n = 0;
  for (int k = 0; k < 4; k++) { d->gexec[k] = nullptr; d->gU[k] = nullptr; }
  memset((void*)&d->kp, 0, sizeof(KP));
  memset((void*)&d->geo, 0, sizeof(Geo));
  memset((void*)&d->st, 0, sizeof(St));
  return d;
}

extern "C" int csu_dev_create(const csu_mesh_desc* m, const csu_params* p, csu_dev** out) {
  if (!out) return set_err(CSU_ERR_ARG, "null out");
  *out = nullptr;
  int rc = validate(m, p);
  if (rc) return rc;
  csu_dev* d = dev_new();
  if (hipGetDevice(&d->device) != hipSuccess) {
    delete d;
    return set_err(CSU_ERR_HIP, "hipGetDevice failed");
  }
  HostMesh h;
  if ((rc = host_mesh(m, p, h)) || (rc = dev_setup(d, h, 1, &m, &p, nullptr))) {
    csu_dev_destroy(d);
    return rc;
  }
  *out = d;
  return CSU_OK;
}

// Options every solve of a batch must share: they select the kernels and
// the shape of an iteration.
static int batch_compatible(const csu_params* a, const csu_params* b, int axi_a, int axi_b) {
  return a->nvar == b->nvar && a->order == b->order && a->flux == b->flux && a->stages == b->stages &&
         a->time_scheme == b->time_scheme && a->dt_local == b->dt_local && a->dt_local_n == b->dt_local_n &&
         a->laminar == b->laminar && a->sa == b->sa && a->sst == b->sst && a->sst_trans == b->sst_trans &&
         a->tp == b->tp && a->limiter == b->limiter && a->turb1order == b->turb1order &&
         a->visc_blazek == b->visc_blazek && a->nlinear == b->nlinear && a->precision == b->precision &&
         a->sa_cc == b->sa_cc && a->gradient == b->gradient && a->rlim == b->rlim && a->lim_k == b->lim_k &&
         a->efix == b->efix &&
         a->pr == b->pr && a->pr_t == b->pr_t && a->wimp == b->wimp && a->gamma == b->gamma &&
         a->gas_r == b->gas_r && a->sst_owallfactor == b->sst_owallfactor && axi_a == axi_b;
}

extern "C" int csu_batch_compatible(const csu_mesh_desc* ma, const csu_params* pa, const csu_mesh_desc* mb,
                                    const csu_params* pb) {
  if (!ma || !pa || !mb || !pb) return -set_err(CSU_ERR_ARG, "null argument");
  return batch_compatible(pa, pb, ma->axi, mb->axi) ? 1 : 0;
}

extern "C" int csu_batch_create(int32_t n, const csu_mesh_desc* const* meshes, const csu_params* const* prms,
                                csu_dev** out) {
  if (!out) return set_err(CSU_ERR_ARG, "null out");
  *out = nullptr;
  if (n <= 0 || !meshes || !prms) return set_err(CSU_ERR_ARG, "batch of %d solves", n);
  int rc;
  double cells = 0;
  for (int q = 0; q < n; q++) {
    if ((rc = validate(meshes[q], prms[q]))) return set_err(rc, "solve %d: %s", q, std::string(g_err).c_str());
    if (!batch_compatible(prms[0], prms[q], meshes[0]->axi, meshes[q]->axi))
      return set_err(CSU_ERR_ARG, "solve %d: options that select the kernels differ from solve 0's", q);
    cells += up_chunk(meshes[q]->ncell);
  }
  if (cells > 2.0e9) return set_err(CSU_ERR_ARG, "batch of %.0f cells: beyond 32-bit cell ids", cells);
  csu_dev* d = dev_new();
  if (hipGetDevice(&d->device) != hipSuccess) {
    delete d;
    return set_err(CSU_ERR_HIP, "hipGetDevice failed");
  }
  std::vector<HostMesh> parts(n);
  for (int q = 0; q < n && !rc; q++) rc = host_mesh(meshes[q], prms[q], parts[q]);
  if (rc) {
    delete d;
    return rc;
  }
  HostMesh H;
  std::vector<SolveP> sol;
  host_concat(parts, H, sol);
  parts.clear();
  if ((rc = dev_setup(d, H, n, meshes, prms, &sol))) {
    csu_dev_destroy(d);
    return rc;
  }
  *out = d;
  return CSU_OK;
}

extern "C" int csu_dev_set_state(csu_dev* d, const double* U) {
  if (!d || !U) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  HIPCHK(hipMemcpyAsync(d->U, U, sizeof(double) * (size_t)d->nv * d->nc, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  d->fresh = 1;
  return CSU_OK;
}

extern "C" int csu_dev_get_state(csu_dev* d, double* U) {
  if (!d || !U) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  HIPCHK(hipMemcpyAsync(U, d->U, sizeof(double) * (size_t)d->nv * d->nc, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

extern "C" int csu_dev_get_dt(csu_dev* d, double* dtl) {
  if (!d || !dtl) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  HIPCHK(hipMemcpyAsync(dtl, d->dtl, sizeof(double) * (size_t)d->nc, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

extern "C" int csu_dev_iteration(const csu_dev* d) { return d ? d->iter : -1; }

extern "C" int csu_dev_sync(csu_dev* d) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

// ---------------------------------------------------------------------------
// profiling: HIP events around each launch on the library's stream
// ---------------------------------------------------------------------------
static int kid(csu_dev* d, const char* name) {
  for (size_t i = 0; i < d->kname.size(); i++)
    if (d->kname[i] == name) return (int)i;
  d->kname.push_back(name);
  d->ktime.push_back(0.0);
  d->kcount.push_back(0);
  return (int)d->kname.size() - 1;
}
static hipEvent_t ev_get(csu_dev* d) {
  if (!d->evpool.empty()) {
    hipEvent_t e = d->evpool.back();
    d->evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}
struct Prof {
  csu_dev* d; int id; hipEvent_t a, b;
  Prof(csu_dev* dd, const char* name) : d(dd), id(-1) {
    if (d->prof) { id = kid(d, name); a = ev_get(d); b = ev_get(d); (void)hipEventRecord(a, d->stream); }
  }
  ~Prof() {
    if (id >= 0) { (void)hipEventRecord(b, d->stream); d->pending.push_back({id, {a, b}}); }
  }
};
static int prof_collect(csu_dev* d) {
  if (d->pending.empty()) return CSU_OK;
  HIPCHK(hipStreamSynchronize(d->stream));
  for (auto& pe : d->pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, pe.second.first, pe.second.second);
    d->ktime[pe.first] += ms;
    d->kcount[pe.first] += 1;
    d->evpool.push_back(pe.second.first);
    d->evpool.push_back(pe.second.second);
  }
  d->pending.clear();
  return CSU_OK;
}

extern "C" int csu_dev_set_graphs(csu_dev* d, int32_t on) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(d->stream));
  if (!on) graphs_drop(d);
  d->use_graphs = on ? 1 : 0;
  return CSU_OK;
}

extern "C" int csu_dev_set_profiling(csu_dev* d, int32_t on) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  int rc = prof_collect(d);
  if (rc) return rc;
  d->prof = on ? 1 : 0;
  for (size_t i = 0; i < d->ktime.size(); i++) { d->ktime[i] = 0; d->kcount[i] = 0; }
  return CSU_OK;
}

extern "C" int csu_dev_kernel_times(csu_dev* d, const char** names, double* avg_ms, int32_t* launches, int32_t max) {
  if (!d) return -set_err(CSU_ERR_ARG, "null argument");
  int rc = prof_collect(d);
  if (rc) return -rc;
  int n = (int)d->kname.size();
  for (int i = 0; i < n && i < max; i++) {
    if (names) names[i] = d->kname[i].c_str();
    if (avg_ms) avg_ms[i] = d->kcount[i] ? d->ktime[i] / d->kcount[i] : 0.0;
    if (launches) launches[i] = d->kcount[i];
  }
  return n;
}

// ---------------------------------------------------------------------------
// orchestration
// ---------------------------------------------------------------------------
// One residual evaluation (solverCalcR) with P valid, then the fused update.
// Us non-null: the primitives-from-state form (k_grad's PU), where Us is the
// stage's conservative state and P is neither read nor written.
template <int NV, bool VISC>
static int stage(csu_dev* d, int flags, double sigma, double rk_a, double* Un, double* Uout,
                 const double* Us = nullptr) {
  const KP& kp = d->kp;
  const int nc = d->nc;
  if constexpr (NV == 4 && !VISC) {
    if (Us) {
      St st = d->st;
      st.Us = Us;
      if (d->nb > 0) {
        Prof p(d, "k_bcflux");
        hipLaunchKernelGGL((k_bcflux<4, true>), dim3(nblk(d->nb)), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, st);
      }
      {
        Prof p(d, "k_grad");
        if (kp.mixed)
          hipLaunchKernelGGL((k_grad<4, false, false, true, true>), dim3((nc + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp,
                             d->geo, st);
        else if (kp.lsq_ld)
          hipLaunchKernelGGL((k_grad<4, false, true, false, true>), dim3((nc + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp,
                             d->geo, st);
        else
          hipLaunchKernelGGL((k_grad<4, false, false, false, true>), dim3((nc + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp,
                             d->geo, st);
      }
      if (d->nf > 0 && !d->tile) {
        Prof p(d, "k_flux");
        if (kp.mixed)
          hipLaunchKernelGGL((k_flux<4, false, true>), dim3((d->nf + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp, d->geo, st);
        else
          hipLaunchKernelGGL((k_flux<4, false>), dim3((d->nf + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp, d->geo, st);
      }
      {
        Prof p(d, "k_update");
        const dim3 gu(nblk(nc)), bu(CSU_BLOCK);
        double* const rdt = d->red + (size_t)NV * d->nred;
        const bool ts = d->tile_ok == CSU_TILE_FACES_S;
        if (kp.mixed && d->tile && ts)
          hipLaunchKernelGGL((k_update<4, false, true, true, CSU_TILE_FACES_S>), gu, bu, 0, d->stream, kp, d->geo, st,
                             flags, sigma, rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red, rdt);
        else if (kp.mixed && d->tile)
          hipLaunchKernelGGL((k_update<4, false, true, true, CSU_TILE_FACES>), gu, bu, 0, d->stream, kp, d->geo, st,
                             flags, sigma, rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red, rdt);
        else if (kp.mixed)
          hipLaunchKernelGGL((k_update<4, false, true, true>), gu, bu, 0, d->stream, kp, d->geo, st, flags, sigma,
                             rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red, rdt);
        else if (d->tile && ts)
          hipLaunchKernelGGL((k_update<4, false, false, true, CSU_TILE_FACES_S>), gu, bu, 0, d->stream, kp, d->geo, st,
                             flags, sigma, rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red, rdt);
        else if (d->tile)
          hipLaunchKernelGGL((k_update<4, false, false, true, CSU_TILE_FACES>), gu, bu, 0, d->stream, kp, d->geo, st,
                             flags, sigma, rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red, rdt);
        else
          hipLaunchKernelGGL((k_update<4, false, false, true>), gu, bu, 0, d->stream, kp, d->geo, st, flags, sigma,
                             rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red, rdt);
      }
      HIPCHK(hipGetLastError());
      return CSU_OK;
    }
  }
  if (kp.order == 2 && kp.limiter == 1) {
    Prof p(d, "k_lim1");
    hipLaunchKernelGGL(k_lim1_partial<NV>, dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, d->st.P, d->red);
    hipLaunchKernelGGL(k_lim1_final<NV>, dim3(d->nsolve), dim3(CSU_BLOCK), 0, d->stream, kp, nblk(nc), nblk(nc),
                       d->geo.sol, d->red, d->st.pref2);
  }
  if (d->nb > 0) {
    // ghost states first: the gradients read them
    Prof p(d, "k_bcflux");
    hipLaunchKernelGGL(k_bcflux<NV>, dim3(nblk(d->nb)), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, d->st);
  }
  {
    Prof p(d, "k_grad");
    if (kp.mixed)
      hipLaunchKernelGGL((k_grad<NV, VISC, false, true>), dim3((nc + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp, d->geo,
                         d->st);
    else if (kp.lsq_ld)
      hipLaunchKernelGGL((k_grad<NV, VISC, true>), dim3((nc + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp, d->geo, d->st);
    else
      hipLaunchKernelGGL((k_grad<NV, VISC, false>), dim3((nc + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp, d->geo, d->st);
  }
  if (d->nf > 0) {
    Prof p(d, "k_flux");
    if (kp.mixed)
      hipLaunchKernelGGL((k_flux<NV, VISC, true>), dim3((d->nf + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp, d->geo,
                         d->st);
    else
      hipLaunchKernelGGL((k_flux<NV, VISC>), dim3((d->nf + d->sblk - 1) / d->sblk), dim3(d->sblk), 0, d->stream, kp, d->geo, d->st);
  }
  {
    Prof p(d, "k_update");
    if (kp.mixed)
      hipLaunchKernelGGL((k_update<NV, VISC, true>), dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, d->st,
                         flags, sigma, rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red,
                         d->red + (size_t)NV * d->nred);
    else
      hipLaunchKernelGGL((k_update<NV, VISC>), dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, d->st,
                         flags, sigma, rk_a, d->prm.cfl, d->dtl, d->dtg, Un, Uout, d->R, d->red,
                         d->red + (size_t)NV * d->nred);
  }
  HIPCHK(hipGetLastError());
  return CSU_OK;
}

template <int NV>
static int prim_pass(csu_dev* d) {
  Prof p(d, "k_prim");
  hipLaunchKernelGGL(k_prim<NV>, dim3(nblk(d->nc)), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->U, d->st.P, d->geo.sol,
                     d->geo.csolk);
  HIPCHK(hipGetLastError());
  return CSU_OK;
}

template <int NV, bool VISC>
static int residual_t(csu_dev* d) {
  int rc = prim_pass<NV>(d);
  d->kp.store_grad = 1;  // the caller may ask for dPx, dPy, phi
  if (!rc) rc = stage<NV, VISC>(d, F_STORE_R, 0.0, 0.0, nullptr, nullptr);
  d->kp.store_grad = 0;
  d->fresh = 1;
  return rc;
}

extern "C" int csu_dev_residual(csu_dev* d, double* R, double* dPx, double* dPy, double* phi, double* P,
                                double* Pb) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  int rc = d->nv == 6 ? residual_t<6, true>(d)
           : d->nv == 5 ? residual_t<5, true>(d) : d->visc ? residual_t<4, true>(d) : residual_t<4, false>(d);
  if (rc) return rc;
  const size_t nvc = (size_t)d->nv * d->nc;
  const size_t nc = d->nc;
  if (R) HIPCHK(hipMemcpyAsync(R, d->R, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (dPx) HIPCHK(hipMemcpyAsync(dPx, d->st.dPx, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (dPy) HIPCHK(hipMemcpyAsync(dPy, d->st.dPy, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (d->visc) {
    // after a viscous residual the reference's gradient slot 3 holds grad T
    // (solverGrad_T, solver.c:477-490)
    if (dPx) HIPCHK(hipMemcpyAsync(dPx + 3 * nc, d->st.dTx, nc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    if (dPy) HIPCHK(hipMemcpyAsync(dPy + 3 * nc, d->st.dTy, nc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  }
  if (phi) HIPCHK(hipMemcpyAsync(phi, d->st.phi, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  std::vector<double> tmp;
  const int ns = d->nv + 1;
  if (P) {
    // device P is SoA [nv+1][nc]; the ABI returns [nc][nv+1]
    tmp.resize((size_t)ns * d->nc);
    HIPCHK(hipMemcpyAsync(tmp.data(), d->st.P, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (int i = 0; i < d->nc; i++)
      for (int m = 0; m < ns; m++) P[(size_t)i * ns + m] = tmp[(size_t)m * d->nc + i];
  }
  if (Pb && d->nb) {
    tmp.resize((size_t)ns * d->nb);
    HIPCHK(hipMemcpyAsync(tmp.data(), d->st.Pb, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (int g = 0; g < d->nb; g++)
      for (int m = 0; m < ns; m++) Pb[(size_t)g * ns + m] = tmp[(size_t)m * d->nb + g];
  }
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

// solverCalcPrimitive(solver, U) (solver.c:1153-1213) on the device state:
// U clamped in place, P [nc][nv+1] to the host (NULL allowed).
// solverCalcPrimitive: the cells, then the ghost states of the boundary
// faces (its boundaryCalcPrimitive calls, solver.c:1215-1218); k_bcflux's
// boundary fluxes are recomputed by the next stage before they are read.
template <int NV>
static int prim_ghost_pass(csu_dev* d) {
  int rc = prim_pass<NV>(d);
  if (rc || d->nb == 0) return rc;
  Prof p(d, "k_bcflux");
  hipLaunchKernelGGL(k_bcflux<NV>, dim3(nblk(d->nb)), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->geo, d->st);
  HIPCHK(hipGetLastError());
  return CSU_OK;
}

extern "C" int csu_dev_primitives(csu_dev* d, double* P) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  int rc = d->nv == 6 ? prim_ghost_pass<6>(d) : d->nv == 5 ? prim_ghost_pass<5>(d) : prim_ghost_pass<4>(d);
  if (rc) return rc;
  d->fresh = 1;
  if (P) {
    const int ns = d->nv + 1;
    std::vector<double> tmp((size_t)ns * d->nc);
    HIPCHK(hipMemcpyAsync(tmp.data(), d->st.P, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (int i = 0; i < d->nc; i++)
      for (int m = 0; m < ns; m++) P[(size_t)i * ns + m] = tmp[(size_t)m * d->nc + i];
  }
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

// RK coefficients, reference solverStepRK (solver.c:788-840)
static const double RK1[] = {1.0};
static const double RK3[] = {0.1481, 0.4, 1.0};
static const double RK4[] = {0.0833, 0.2069, 0.4265, 1.0};
static const double RK5[] = {0.0533, 0.1263, 0.2375, 0.4414, 1.0};

// Residual norms and force coefficients of the iteration just done, into
// the history slot the device counter names (k_forces advances it).
template <int NV>
static void iteration_tail(csu_dev* d, int nblocks_cells) {
  {
    Prof p(d, "k_res");
    hipLaunchKernelGGL(k_res_final<NV>, dim3(d->nsolve), dim3(RES_THREADS), 0, d->stream, d->nc, nblocks_cells,
                       nblocks_cells, d->geo.sol, d->red,
                       d->res_hist, d->hist_cap, (const int*)d->hctr);
  }
  {
    Prof p(d, "k_forces");
    hipLaunchKernelGGL(k_forces, dim3(d->nsolve), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->nwall, d->wall, d->st.Pb,
                       d->st.wfric, d->geo.bdsx, d->geo.bdsy, d->prm.sref, d->visc, d->coef_hist, d->hist_cap,
                       d->hctr, d->geo.sol);
  }
  if (d->geo.sol) hipLaunchKernelGGL(k_next_slot, dim3(1), dim3(64), 0, d->stream, d->hctr);
}

static int is_local(const csu_dev* d, int it) {
  return d->prm.dt_local == 1 && !(d->prm.dt_local_n >= 0 && it > d->prm.dt_local_n);
}

// Capture `body` (which enqueues work on d->stream) into graph slot k, or
// replay slot k when it was captured with the same U buffer.
template <typename F>
static int run_graph(csu_dev* d, int k, F body) {
  if (!d->gexec[k] || d->gU[k] != d->U) {
    if (d->gexec[k]) (void)hipGraphExecDestroy(d->gexec[k]);
    d->gexec[k] = nullptr;
    double* u0 = d->U;
    HIPCHK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    int rc = body();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(d->stream, &graph);
    if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return set_err(CSU_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&d->gexec[k], graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) { d->gexec[k] = nullptr; return set_err(CSU_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e)); }
    d->gU[k] = u0;
    d->U = u0;  // the body's host-side pointer swaps are replayed by the graph
    // a capture records work without running it: replay once now
  }
  HIPCHK(hipGraphLaunch(d->gexec[k], d->stream));
  return CSU_OK;
}

// One RK iteration: stages, buffer swap, next time step, norms, forces.
template <int NV, bool VISC>
static int rk_iteration(csu_dev* d, int it_index) {
  const int nc = d->nc;
  const double* a = d->prm.stages == 1 ? RK1 : d->prm.stages == 3 ? RK3 : d->prm.stages == 4 ? RK4 : RK5;
  const double sigma = 0.5 * d->prm.stages;
  double* red_dt = d->red + (size_t)NV * d->nred;
  const int cur_local = is_local(d, it_index), next_local = is_local(d, it_index + 1);
  int first_flags = F_CLAMP_UN;
  if (d->fresh) {
    // time step from the unclamped state (solverCalcDt), then primitives
    // (the primitives-from-state form recovers them in the stage kernels
    // and clamps U^n in the first update, as after any other iteration)
    Prof p(d, "k_dt");
    hipLaunchKernelGGL(k_dt<NV>, dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->geo, cur_local, d->prm.cfl,
                       d->U, d->dtl, red_dt);
    if (!cur_local)
      hipLaunchKernelGGL(k_dt_final, dim3(d->nsolve), dim3(CSU_BLOCK), 0, d->stream, nblk(nc), d->prm.cfl, red_dt,
                         d->dtg, d->geo.sol);
    if (!d->pu) {
      int rc = prim_pass<NV>(d);
      if (rc) return rc;
      first_flags = 0;
    }
    d->fresh = 0;
  }
  if (d->st.tube)
    hipLaunchKernelGGL(k_time_clamp, dim3(1), dim3(64), 0, d->stream, d->tstate, d->dtg, d->prm.stages);
  for (int s = 0; s < d->prm.stages; s++) {
    const bool last = s == d->prm.stages - 1;
    int flags = F_RK | (cur_local ? F_CUR_LOCAL : 0) | (s == 0 ? first_flags : 0);
    if (last) flags |= F_LAST | F_STORE_R | (next_local ? F_NEXT_LOCAL : 0);
    int rc = stage<NV, VISC>(d, flags, sigma, a[s], d->U, d->Uaux, d->pu ? (s == 0 ? d->U : d->Uaux) : nullptr);
    if (rc) return rc;
  }
  std::swap(d->U, d->Uaux);
  if (!next_local)
    hipLaunchKernelGGL(k_dt_final, dim3(d->nsolve), dim3(CSU_BLOCK), 0, d->stream, nblk(nc), d->prm.cfl, red_dt,
                       d->dtg, d->geo.sol);
  iteration_tail<NV>(d, nblk(nc));
  HIPCHK(hipGetLastError());
  return CSU_OK;
}

template <int NV, bool VISC>
static int iterate_t(csu_dev* d, int niter) {
  int it = 0;
  while (it < niter) {
    const int i0 = d->iter;
    const bool steady = is_local(d, i0) == is_local(d, i0 + 1) && is_local(d, i0 + 1) == is_local(d, i0 + 2);
    if (d->use_graphs && !d->prof && !d->fresh && niter - it >= 2 && steady) {
      // two iterations per graph, so U ends in the buffer it started in
      const int k = (is_local(d, i0) ? 2 : 0) + (d->U == d->Ubase ? 0 : 1);
      int rc = run_graph(d, k, [&]() -> int {
        int r = rk_iteration<NV, VISC>(d, i0);
        if (!r) r = rk_iteration<NV, VISC>(d, i0 + 1);
        return r;
      });
      if (rc) return rc;
      d->iter += 2;
      it += 2;
      continue;
    }
    int rc = rk_iteration<NV, VISC>(d, d->iter);
    if (rc) return rc;
    d->iter++;
    it++;
  }
  return CSU_OK;
}

// implicit iteration: solverCalcR, then LU-SGS (implicitCalcD + sweeps) or
// DPLUR (implicitCalcDPLUR), then U += dW (solver.c:1871-1884)
template <int NV, bool VISC>
static int implicit_iteration(csu_dev* d) {
  const KP& kp = d->kp;
  const int nc = d->nc, nb = nblk(nc);
  double* red_min = d->red + (size_t)NV * d->nred;
  int rc = prim_pass<NV>(d);
  if (!rc) rc = stage<NV, VISC>(d, F_STORE_R | F_RES, 0.0, 0.0, nullptr, nullptr);
  if (rc) return rc;
  {
    Prof p(d, "k_impD");
    hipLaunchKernelGGL((k_impD<NV, VISC>), dim3(nb), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, d->st, d->prm.wimp,
                       d->Dp, red_min, d->caux);
    hipLaunchKernelGGL(k_dt_final, dim3(d->nsolve), dim3(CSU_BLOCK), 0, d->stream, nb, d->prm.cfl, red_min, d->dtg,
                       d->geo.sol);
  }
  if (d->prm.time_scheme == CSU_TIME_LUSGS) {
    Prof p(d, "k_lusgs");
    // a level holds a fraction of the cells (a colour of the renumbered
    // mesh): small workgroups (64 threads unless tuned) spread it over every CU
    const int LB = d->lblk;
    for (size_t l = 0; l + 1 < d->lvlL.size(); l++) {
      int off = d->lvlL[l], cnt = d->lvlL[l + 1] - off;
      hipLaunchKernelGGL((k_lusgs<NV, VISC, 0>), dim3((cnt + LB - 1) / LB), dim3(LB), 0, d->stream, kp, d->geo, d->st,
                         d->cellsL + off, cnt, d->Dp, d->dtg, d->R, d->dW0, (const double*)nullptr, (double*)nullptr,
                         (const double*)d->caux);
    }
    // the backward levels start each cell from D*dW0 and apply the final
    // increment to the state themselves
    for (size_t l = 0; l + 1 < d->lvlU.size(); l++) {
      int off = d->lvlU[l], cnt = d->lvlU[l + 1] - off;
      hipLaunchKernelGGL((k_lusgs<NV, VISC, 1>), dim3((cnt + LB - 1) / LB), dim3(LB), 0, d->stream, kp, d->geo, d->st,
                         d->cellsU + off, cnt, d->Dp, d->dtg, d->R, d->dW1, (const double*)d->dW0, d->U,
                         (const double*)d->caux);
    }
  } else {
    Prof p(d, "k_dplur");
    hipLaunchKernelGGL((k_blocks<NV, VISC>), dim3(nb), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, d->st, d->U, d->A);
    // one launch per sweep: each tests the previous sweep's maximum itself
    for (int mm = 0; mm < d->prm.nlinear; mm++)
      hipLaunchKernelGGL(k_jacobi<NV>, dim3(nb), dim3(CSU_BLOCK), 0, d->stream, kp, d->geo, d->A, d->Dp, d->dtg,
                         d->R, d->dW0, d->dW1, (const int*)d->jstate, d->jred, mm, d->st.dq);
    hipLaunchKernelGGL(k_upd_imp<NV>, dim3(nb), dim3(CSU_BLOCK), 0, d->stream, nc, d->U, d->dW0, d->dW1,
                       (const int*)d->jstate, (const double*)d->jred, d->prm.nlinear, nb, d->geo.sol, d->geo.csolk);
    hipLaunchKernelGGL(k_jac_base, dim3((d->nsolve + 63) / 64), dim3(64), 0, d->stream, d->jstate,
                       (const double*)d->jred, d->prm.nlinear, nb, d->geo.sol, d->nsolve);
  }
  iteration_tail<NV>(d, nb);
  HIPCHK(hipGetLastError());
  return CSU_OK;
}

template <int NV, bool VISC>
static int iterate_implicit(csu_dev* d, int niter) {
  for (int it = 0; it < niter; it++) {
    int rc = d->use_graphs && !d->prof ? run_graph(d, 0, [&]() { return implicit_iteration<NV, VISC>(d); })
                                       : implicit_iteration<NV, VISC>(d);
    if (rc) return rc;
    d->iter++;
  }
  d->fresh = 1;
  return CSU_OK;
}

// Enqueue niter iterations on the context's stream (no host wait).
static int enqueue_iterations(csu_dev* d, int niter) {
  if (d->hist_cap < niter) {
    // a new history ring (the old one freed): device slots continue from
    // the iteration count
    HIPCHK(hipStreamSynchronize(d->stream));
    graphs_drop(d);
    dfree(d, d->res_hist, sizeof(double) * (size_t)d->hist_cap * d->nsolve * d->nv);
    dfree(d, d->coef_hist, sizeof(double) * (size_t)d->hist_cap * d->nsolve * 4);
    d->res_hist = d->coef_hist = nullptr;
    d->hist_cap = 0;
    int rc;
    if ((rc = dalloc(d, &d->res_hist, (size_t)niter * d->nsolve * d->nv)) ||
        (rc = dalloc(d, &d->coef_hist, (size_t)niter * d->nsolve * 4)))
      return rc;
    d->hist_cap = niter;
  }
  int rc;
  if (d->prm.time_scheme == CSU_TIME_RK)
    rc = d->nv == 6   ? iterate_t<6, true>(d, niter)
         : d->nv == 5 ? iterate_t<5, true>(d, niter)
         : d->visc    ? iterate_t<4, true>(d, niter)
                      : iterate_t<4, false>(d, niter);
  else
    rc = d->nv == 6   ? iterate_implicit<6, true>(d, niter)
         : d->nv == 5 ? iterate_implicit<5, true>(d, niter)
         : d->visc    ? iterate_implicit<4, true>(d, niter)
                      : iterate_implicit<4, false>(d, niter);
  return rc;
}

extern "C" int csu_dev_iterate(csu_dev* d, int32_t niter, double* res, double* coef) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (niter < 0) return set_err(CSU_ERR_ARG, "niter %d", niter);
  if (niter == 0) return CSU_OK;
  const int start = d->iter;
  int rc = enqueue_iterations(d, niter);
  if (rc) return rc;
  if (res || coef) {
    // a slot holds every solve's row: [nsolve][nvar] norms, [nsolve][4] forces
    const size_t rn = (size_t)d->nsolve * d->nv, cn = (size_t)d->nsolve * 4;
    std::vector<double> rh((size_t)d->hist_cap * rn), ch((size_t)d->hist_cap * cn);
    HIPCHK(hipMemcpyAsync(rh.data(), d->res_hist, rh.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipMemcpyAsync(ch.data(), d->coef_hist, ch.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (int it = 0; it < niter; it++) {
      int slot = (start + it) % d->hist_cap;
      if (res) memcpy(res + (size_t)it * rn, rh.data() + (size_t)slot * rn, sizeof(double) * rn);
      if (coef) memcpy(coef + (size_t)it * cn, ch.data() + (size_t)slot * cn, sizeof(double) * cn);
    }
  }
  return prof_collect(d);
}

// The state a context starts from (U as saved), iteration 0, implicit
// increments zero, every time step to be derived again.
static int tune_reset(csu_dev* d, const double* saved) {
  const size_t nvc = (size_t)d->nv * d->nc;
  HIPCHK(hipStreamSynchronize(d->stream));
  graphs_drop(d);
  HIPCHK(hipMemcpyAsync(d->U, saved, nvc * sizeof(double), hipMemcpyDeviceToDevice, d->stream));
  HIPCHK(hipMemcpyAsync(d->Uaux, saved, nvc * sizeof(double), hipMemcpyDeviceToDevice, d->stream));
  HIPCHK(hipMemsetAsync(d->hctr, 0, sizeof(int), d->stream));
  if (d->dW0) HIPCHK(hipMemsetAsync(d->dW0, 0, nvc * sizeof(double), d->stream));
  if (d->dW1) HIPCHK(hipMemsetAsync(d->dW1, 0, nvc * sizeof(double), d->stream));
  if (d->A) HIPCHK(hipMemsetAsync(d->A, 0, (size_t)4 * d->nv * d->nv * d->nc * sizeof(double), d->stream));
  if (d->jstate) HIPCHK(hipMemsetAsync(d->jstate, 0, 2 * sizeof(int) * (size_t)d->nsolve, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  d->iter = 0;
  d->fresh = 1;
  d->Ubase = d->U;
  return CSU_OK;
}

// Time every candidate launch shape over `iters` graph-replayed iterations
// (after two that capture) from the context's state before its first
// iteration, keep the fastest for this context and its bucket, and put the
// context back to that state: results are the same with any shape.
extern "C" int csu_dev_autotune(csu_dev* d, int32_t iters, int32_t* choice) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (d->iter != 0) return set_err(CSU_ERR_ARG, "autotune runs before the first iteration");
  if (d->prof) return set_err(CSU_ERR_ARG, "autotune with profiling on");
  if (iters <= 0) iters = 4;
  const size_t nvc = (size_t)d->nv * d->nc;
  double* saved = nullptr;
  if (hipMalloc(&saved, nvc * sizeof(double)) != hipSuccess) return set_err(CSU_ERR_NOMEM, "autotune state copy");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CSU_OK;
  if (hipMemcpyAsync(saved, d->U, nvc * sizeof(double), hipMemcpyDeviceToDevice, d->stream) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = set_err(CSU_ERR_HIP, "autotune set-up");
  // history ring for the longest run first, so that no candidate's timed
  // iterations wait on an allocation
  if (!rc) rc = tune_reset(d, saved);
  if (!rc) rc = enqueue_iterations(d, 2 * iters + 2);
  const int sbs[3] = {64, 128, 256};
  const bool lus = d->prm.time_scheme == CSU_TIME_LUSGS;
  TuneVal best{d->sblk, d->lblk, d->kp.xcd, d->tile};
  float best_ms = 1e30f;
  for (int u = 0; u < (d->tile_ok ? 2 : 1) && !rc; u++)
  for (int a = 0; a < 3 && !rc; a++)
    for (int b = 0; b < (lus ? 3 : 1) && !rc; b++)
      for (int x = 1; x >= 0 && !rc; x--) {
        d->tile = u;
        d->sblk = sbs[a];
        d->lblk = lus ? sbs[b] : d->lblk;
        d->kp.xcd = x;
        if ((rc = tune_reset(d, saved)) || (rc = enqueue_iterations(d, 2))) break;
        // the better of two timed runs: one run alone is a few per cent noisy
        float ms = 1e30f;
        for (int rep = 0; rep < 2 && !rc; rep++) {
          float t = 0.f;
          if (hipEventRecord(e0, d->stream) != hipSuccess) { rc = set_err(CSU_ERR_HIP, "event"); break; }
          if ((rc = enqueue_iterations(d, iters))) break;
          if (hipEventRecord(e1, d->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
              hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            rc = set_err(CSU_ERR_HIP, "autotune timing");
            break;
          }
          ms = std::min(ms, t);
        }
        if (rc) break;
        if (ms < best_ms) { best_ms = ms; best = TuneVal{d->sblk, d->lblk, d->kp.xcd, d->tile}; }
      }
  d->sblk = best.sblk;
  d->lblk = best.lblk;
  d->kp.xcd = best.xcd;
  d->tile = best.tile;
  const int rc2 = tune_reset(d, saved);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(saved);
  if (rc) return rc;
  if (rc2) return rc2;
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[tune_key(d)] = best;
  }
  d->tuned = 1;
  if (choice) { choice[0] = best.sblk; choice[1] = best.lblk; choice[2] = best.xcd; choice[3] = best.tile; }
  return CSU_OK;
}

// Launch shapes chosen by the caller (a shape the autotuner picked in an
// earlier process, or a fixed one for profiling runs): gradient/face
// workgroup and LU-SGS level workgroup (64, 128 or 256), XCD tile map (0
// or 1), tile form of the face flux + update (0 or 1; primitives-from-state
// runs only).  Results do not depend on them.
extern "C" int csu_dev_set_launch_shape(csu_dev* d, const int32_t* choice) {
  if (!d || !choice) return set_err(CSU_ERR_ARG, "null argument");
  auto ok = [](int b) { return b == 64 || b == 128 || b == 256; };
  if (!ok(choice[0]) || !ok(choice[1]) || (choice[2] != 0 && choice[2] != 1) || (choice[3] != 0 && choice[3] != 1))
    return set_err(CSU_ERR_ARG, "launch shape %d %d %d %d", choice[0], choice[1], choice[2], choice[3]);
  if (choice[3] && !d->tile_ok)
    return set_err(CSU_ERR_ARG, "the tile form needs an inviscid second-order RK run whose tiles fit");
  HIPCHK(hipStreamSynchronize(d->stream));
  graphs_drop(d);  // the shapes are part of the captured launches
  d->sblk = choice[0];
  d->lblk = choice[1];
  d->kp.xcd = choice[2];
  d->tile = choice[3];
  d->tuned = 0;
  return CSU_OK;
}

// Launch shapes in use: gradient/face workgroup, LU-SGS level workgroup,
// XCD tile map, tile form; returns 1 when they came from the autotuner.
extern "C" int csu_dev_launch_shape(const csu_dev* d, int32_t* choice) {
  if (!d || !choice) return -set_err(CSU_ERR_ARG, "null argument");
  choice[0] = d->sblk; choice[1] = d->lblk; choice[2] = d->kp.xcd; choice[3] = d->tile;
  return d->tuned;
}

// ---------------------------------------------------------------------------
// batches: several independent solves in one context
// ---------------------------------------------------------------------------
extern "C" int csu_batch_size(const csu_dev* d) { return d ? d->nsolve : -1; }

static int batch_solve_arg(const csu_dev* d, int q) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (d->sol_host.empty()) return set_err(CSU_ERR_ARG, "not a batch context");
  if (q < 0 || q >= d->nsolve) return set_err(CSU_ERR_ARG, "solve %d of %d", q, d->nsolve);
  return CSU_OK;
}

extern "C" int csu_batch_ncell(const csu_dev* d, int32_t q) {
  if (batch_solve_arg(d, q)) return -1;
  return d->sol_host[q].c1 - d->sol_host[q].c0;
}

extern "C" int csu_batch_set_state(csu_dev* d, int32_t q, const double* U) {
  int rc = batch_solve_arg(d, q);
  if (rc) return rc;
  if (!U) return set_err(CSU_ERR_ARG, "null state");
  const SolveP& v = d->sol_host[q];
  const size_t n = (size_t)(v.c1 - v.c0);
  for (int k = 0; k < d->nv; k++)
    HIPCHK(hipMemcpyAsync(d->U + (size_t)k * d->nc + v.c0, U + k * n, n * sizeof(double), hipMemcpyHostToDevice,
                          d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  d->fresh = 1;
  return CSU_OK;
}

extern "C" int csu_batch_get_state(csu_dev* d, int32_t q, double* U) {
  int rc = batch_solve_arg(d, q);
  if (rc) return rc;
  if (!U) return set_err(CSU_ERR_ARG, "null state");
  const SolveP& v = d->sol_host[q];
  const size_t n = (size_t)(v.c1 - v.c0);
  for (int k = 0; k < d->nv; k++)
    HIPCHK(hipMemcpyAsync(U + k * n, d->U + (size_t)k * d->nc + v.c0, n * sizeof(double), hipMemcpyDeviceToHost,
                          d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

// The solve table on the device follows the host copy (stream-ordered:
// iterations already enqueued see the old values, later ones the new).
static int batch_push(csu_dev* d, int q) {
  HIPCHK(hipMemcpyAsync((void*)(d->geo.sol + q), &d->sol_host[q], sizeof(SolveP), hipMemcpyHostToDevice,
                        d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

// A new CFL number for one solve, with csu_dev_set_cfl's meaning: the next
// iteration takes every time step from the current state (for the other
// solves that gives the values they already hold).
extern "C" int csu_batch_set_cfl(csu_dev* d, int32_t q, double cfl) {
  int rc = batch_solve_arg(d, q);
  if (rc) return rc;
  if (!(cfl > 0.0) || !std::isfinite(cfl)) return set_err(CSU_ERR_ARG, "CFL %g", cfl);
  d->sol_host[q].cfl = cfl;
  d->fresh = 1;
  return batch_push(d, q);
}

// Hold (on = 0) or resume one solve: a held solve's state does not change.
extern "C" int csu_batch_set_active(csu_dev* d, int32_t q, int32_t on) {
  int rc = batch_solve_arg(d, q);
  if (rc) return rc;
  d->sol_host[q].active = on ? 1 : 0;
  d->fresh = 1;
  return batch_push(d, q);
}

// Run every solve of a batch to a drop of its density residual by `drop`
// from its first iteration (solverCalcRes, solver.c:867-892), chunk
// iterations at a time (one graph replay per two iterations), holding each
// solve once it converges, stops being finite or reaches max_iter; with
// growth > 1 a solve's CFL is multiplied by growth after every chunk up to
// cfl_max[q] (NULL: no ramp).  out[q] receives the solve's result.
extern "C" int csu_batch_run(csu_dev* d, double drop, int32_t max_iter, int32_t chunk, const double* cfl_max,
                             double growth, int32_t peak_ref, double switch_drop, csu_batch_result* out) {
  if (!d || !out) return set_err(CSU_ERR_ARG, "null argument");
  if (d->sol_host.empty()) return set_err(CSU_ERR_ARG, "not a batch context");
  if (!(drop > 0) || max_iter <= 0 || chunk <= 0 || !(growth >= 1.0) || !(switch_drop >= 0.0))
    return set_err(CSU_ERR_ARG, "drop %g, max_iter %d, chunk %d, growth %g, switch %g", drop, max_iter, chunk,
                   growth, switch_drop);
  const int ns = d->nsolve, nv = d->nv;
  for (int q = 0; q < ns; q++) {
    memset(&out[q], 0, sizeof(csu_batch_result));
    out[q].iter_conv = -1;
    out[q].iter_switch = -1;
    out[q].cfl = d->sol_host[q].cfl;
    if (!d->sol_host[q].active) out[q].status = -1;  // held by the caller: not run
  }
  std::vector<double> res((size_t)chunk * ns * nv), coef((size_t)chunk * ns * 4);
  std::vector<int> chunks(ns, 0), stall(ns, 0);
  std::vector<double> best(ns, 0.0);
  int running = 0;
  for (int q = 0; q < ns; q++) running += d->sol_host[q].active;
  while (running > 0) {
    int rc = csu_dev_iterate(d, chunk, res.data(), coef.data());
    if (rc) return rc;
    for (int q = 0; q < ns; q++) {
      csu_batch_result& r = out[q];
      if (!d->sol_host[q].active) continue;
      bool stop = false;
      chunks[q]++;
      for (int it = 0; it < chunk && !stop; it++) {
        const double r0 = res[((size_t)it * ns + q) * nv];
        r.iters++;
        if (!std::isfinite(r0)) {
          r.res_last = r0;
          r.status = CSU_ERR_DIVERGED;
          stop = true;
          break;
        }
        if (r.iters == 1) r.res0 = r0;
        r.res_peak = std::max(r.res_peak, r0);
        r.res_last = r0;
        memcpy(r.coef_last, &coef[((size_t)it * ns + q) * 4], sizeof r.coef_last);
        if (r0 <= drop * (peak_ref ? r.res_peak : r.res0)) {
          r.converged = 1;
          r.iter_conv = r.iters;
          stop = true;
        } else if (r.iters >= max_iter) {
          stop = true;
        }
      }
      // the chunk's remaining iterations ran too: the state is that of the
      // chunk's end, the recorded values those of the stopping iteration
      if (stop) {
        d->sol_host[q].active = 0;
        if ((rc = batch_push(d, q))) return rc;
        running--;
        d->fresh = 1;
      } else if (growth > 1.0 && cfl_max && d->sol_host[q].cfl < cfl_max[q]) {
        d->sol_host[q].cfl = std::min(d->sol_host[q].cfl * growth, cfl_max[q]);
        r.cfl = d->sol_host[q].cfl;
        if ((rc = batch_push(d, q))) return rc;
        d->fresh = 1;
      }
    }
    // mixed precision: FP64 as soon as one running solve has reached
    // switch_drop or stalled (its residual, below a tenth of its peak, has
    // not fallen by 10 % in four chunks: FP32 flux rounding dominates it);
    // from there on mixed precision would hold that solve back, and FP64 is
    // never wrong for the others
    for (int q = 0; q < ns; q++) {
      if (!d->sol_host[q].active) continue;
      if (chunks[q] == 1 || out[q].res_last < 0.9 * best[q]) {
        best[q] = out[q].res_last;
        stall[q] = 0;
      } else if (out[q].res_last < 0.1 * out[q].res_peak) {
        stall[q]++;
      }
    }
    if (switch_drop > 0.0 && d->prm.precision == CSU_PREC_MIXED && running > 0) {
      bool any = false;
      for (int q = 0; q < ns; q++)
        if (d->sol_host[q].active && (out[q].res_last <= switch_drop * (peak_ref ? out[q].res_peak : out[q].res0) ||
                                      stall[q] >= 4))
          any = true;
      if (any) {
        if ((rc = csu_dev_set_precision(d, CSU_PREC_FP64))) return rc;
        for (int q = 0; q < ns; q++)
          if (d->sol_host[q].active) out[q].iter_switch = out[q].iters;
      }
    }
  }
  for (int q = 0; q < ns; q++) out[q].iters_run = chunks[q] * chunk;
  return CSU_OK;
}

// New CFL number for the following iterations.  The next iteration takes its
// time step from the current state with the new number, as a run restarted
// from this state with the new CFL would (the chain of restart runs with
// stepped CFL that ramps the reference's CFL, solver.c:1846-1862).
extern "C" int csu_dev_set_cfl(csu_dev* d, double cfl) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  if (!(cfl > 0.0) || !std::isfinite(cfl)) return set_err(CSU_ERR_ARG, "CFL %g", cfl);
  HIPCHK(hipStreamSynchronize(d->stream));
  d->prm.cfl = cfl;
  graphs_drop(d);  // the CFL is a kernel argument of the captured iterations
  d->fresh = 1;
  return CSU_OK;
}

// Precision of the following iterations; the state carries over.  The
// FP32 copies of the geometry (and FP32 edge states and fluxes) are made on
// the device the first time mixed precision is asked for, FP64 edge states
// the first time FP64 is.
extern "C" int csu_dev_set_precision(csu_dev* d, int32_t prec) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (prec != CSU_PREC_FP64 && prec != CSU_PREC_MIXED && prec != CSU_PREC_FP64_LD)
    return set_err(CSU_ERR_ARG, "precision %d", prec);
  if (prec == d->prm.precision) return CSU_OK;
  if (prec == CSU_PREC_FP64_LD && d->prm.gradient == 1)
    return set_err(CSU_ERR_UNSUPPORTED, "Green-Gauss gradients with the long double least-squares mode");
  HIPCHK(hipStreamSynchronize(d->stream));
  const size_t nc = d->nc, nf = d->nf, nvc = (size_t)d->nv * nc;
  Geo& g = d->geo;
  St& s = d->st;
  int rc;
  auto f32 = [&](const double* src, const float** dst, size_t n) -> int {
    float* p = nullptr;
    int r = dalloc(d, &p, n);
    if (r) return r;
    if (n) hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + CSU_BLOCK - 1) / CSU_BLOCK)), dim3(CSU_BLOCK), 0,
                              d->stream, src, p, n);
    *dst = p;
    return hipGetLastError() == hipSuccess ? CSU_OK : set_err(CSU_ERR_HIP, "k_to_f32 launch");
  };
  if (prec == CSU_PREC_MIXED) {
    if (!g.lwx32 && ((rc = f32(g.lwx, &g.lwx32, 4 * nc)) || (rc = f32(g.lwy, &g.lwy32, 4 * nc)) ||
                     (rc = f32(g.edx, &g.edx32, 4 * nc)) || (rc = f32(g.edy, &g.edy32, 4 * nc)) ||
                     (rc = f32(g.fdsx, &g.fdsx32, nf)) || (rc = f32(g.fdsy, &g.fdsy32, nf)) ||
                     (rc = f32(g.fds, &g.fds32, nf))))
      return rc;
    if (!s.fflux32 && (rc = dalloc(d, &s.fflux32, (size_t)d->nv * nf))) return rc;
    if (d->prm.order == 2 && !s.rec32 && (rc = dalloc(d, &s.rec32, 4 * nvc))) return rc;
  } else if (d->prm.order == 2 && !s.rec && (rc = dalloc(d, &s.rec, 4 * nvc))) {
    return rc;
  }
  HIPCHK(hipStreamSynchronize(d->stream));
  d->kp.mixed = prec == CSU_PREC_MIXED;
  d->kp.lsq_ld = prec == CSU_PREC_FP64_LD;
  d->prm.precision = prec;
  graphs_drop(d);  // the captured iterations launch the other kernels
  return CSU_OK;
}

// Time-accurate run to tmax, the tube branch of solverSolve
// (solver.c:1912-1941).  The step control runs on the device
// (k_time_clamp), so iterations are enqueued in chunks; the chunk length
// follows the remaining time over the current step, and iterations a chunk
// holds past the last step leave U unchanged.
extern "C" int csu_dev_advance_time(csu_dev* d, double tmax, int32_t max_iter, double* t, int32_t* niter,
                                    int32_t* finished) {
  if (!d || !t) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  if (d->prm.time_scheme != CSU_TIME_RK || d->prm.dt_local == 1)
    return set_err(CSU_ERR_UNSUPPORTED, "a time-accurate run needs the RK scheme with a global time step");
  if (!(tmax >= *t)) return set_err(CSU_ERR_ARG, "tmax %g before t %g", tmax, *t);
  double ts[4] = {*t, tmax, 0.0, 0.0}, dt = 0.0;
  HIPCHK(hipMemcpyAsync(d->tstate, ts, sizeof ts, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  graphs_drop(d);
  d->st.tube = d->tstate;
  int rc = CSU_OK, chunk = 4;
  for (;;) {
    const int done = (int)ts[3];
    if (max_iter > 0 && chunk > max_iter - done) chunk = max_iter - done;
    if ((rc = enqueue_iterations(d, chunk))) break;
    HIPCHK(hipMemcpyAsync(ts, d->tstate, sizeof ts, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipMemcpyAsync(&dt, d->dtg, sizeof dt, hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    if (ts[2] != 0.0) break;
    if (!(dt > 0.0) || !std::isfinite(dt)) {
      rc = set_err(CSU_ERR_ARG, "time step %g at t = %g after %d steps: the state is no longer physical", dt, ts[0],
                   (int)ts[3]);
      break;
    }
    if (max_iter > 0 && (int)ts[3] >= max_iter) break;
    const double left = (tmax - ts[0]) / (dt * d->prm.stages / 2.0);
    chunk = left > 256 ? 256 : left < 1 ? 1 : (int)left + 1;
  }
  graphs_drop(d);
  d->st.tube = nullptr;
  d->fresh = 1;
  *t = ts[0];
  if (niter) *niter = (int32_t)ts[3];
  if (finished) *finished = ts[2] != 0.0;
  if (rc) return rc;
  return prof_collect(d);
}

extern "C" int csu_dev_iterate_async(csu_dev* d, int32_t niter) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (niter <= 0) return set_err(CSU_ERR_ARG, "niter %d", niter);
  if (d->prof) return set_err(CSU_ERR_ARG, "asynchronous iterations with profiling on");
  const int start = d->iter;
  int rc = enqueue_iterations(d, niter);
  if (rc) return rc;
  if (d->hcap < d->hist_cap) {
    HIPCHK(hipStreamSynchronize(d->stream));
    if (d->hres) (void)hipHostFree(d->hres);
    if (d->hcoef) (void)hipHostFree(d->hcoef);
    d->hres = d->hcoef = nullptr;
    d->hcap = 0;
    if (hipHostMalloc((void**)&d->hres, sizeof(double) * (size_t)d->hist_cap * d->nsolve * d->nv, 0) != hipSuccess ||
        hipHostMalloc((void**)&d->hcoef, sizeof(double) * (size_t)d->hist_cap * d->nsolve * 4, 0) != hipSuccess)
      return set_err(CSU_ERR_NOMEM, "pinned history buffers");
    d->hcap = d->hist_cap;
  }
  HIPCHK(hipMemcpyAsync(d->hres, d->res_hist, sizeof(double) * (size_t)d->hist_cap * d->nsolve * d->nv,
                        hipMemcpyDeviceToHost,
                        d->stream));
  HIPCHK(hipMemcpyAsync(d->hcoef, d->coef_hist, sizeof(double) * (size_t)d->hist_cap * d->nsolve * 4,
                        hipMemcpyDeviceToHost,
                        d->stream));
  d->async_start = start;
  d->async_n = niter;
  return CSU_OK;
}

extern "C" int csu_dev_ready(csu_dev* d) {
  if (!d) return -set_err(CSU_ERR_ARG, "null argument");
  hipError_t e = hipStreamQuery(d->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return -set_err(CSU_ERR_HIP, "stream: %s", hipGetErrorString(e));
}

extern "C" int csu_dev_async_result(csu_dev* d, double* res, double* coef) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(d->stream));
  const size_t rn = (size_t)d->nsolve * d->nv, cn = (size_t)d->nsolve * 4;
  for (int it = 0; it < d->async_n; it++) {
    int slot = (d->async_start + it) % d->hist_cap;
    if (res) memcpy(res + (size_t)it * rn, d->hres + (size_t)slot * rn, sizeof(double) * rn);
    if (coef) memcpy(coef + (size_t)it * cn, d->hcoef + (size_t)slot * cn, sizeof(double) * cn);
  }
  return CSU_OK;
}

extern "C" int csu_dev_wall_tensor(csu_dev* d, double* out) {
  if (!d || !out) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  if (!d->visc) return set_err(CSU_ERR_ARG, "no viscous terms: no wall shear");
  std::vector<double> w(4 * (size_t)d->nb);
  HIPCHK(hipMemcpyAsync(w.data(), d->st.wten, w.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  for (int q = 0; q < d->nb; q++)
    for (int k = 0; k < 4; k++) out[4 * (size_t)q + k] = w[(size_t)k * d->nb + q];
  return CSU_OK;
}

extern "C" int csu_dev_sst_fields(csu_dev* d, double* out) {
  if (!d || !out) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  if (!d->kp.sst || d->nv != 6) return set_err(CSU_ERR_ARG, "not an SST run");
  const size_t n = 5 * (size_t)d->nc;
  double* buf = nullptr;
  if (hipMalloc(&buf, n * sizeof(double)) != hipSuccess)
    return set_err(CSU_ERR_NOMEM, "hipMalloc(%zu bytes) failed", n * sizeof(double));
  hipLaunchKernelGGL(k_sst_fields, dim3(nblk(d->nc)), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->st.P, d->st.dPx,
                     d->st.dPy, d->geo.wdist, buf);
  std::vector<double> w(n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(w.data(), buf, n * sizeof(double), hipMemcpyDeviceToHost, d->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
  const hipError_t ef = hipFree(buf);
  if (e == hipSuccess) e = ef;
  if (e != hipSuccess) return set_err(CSU_ERR_HIP, "sst fields: %s", hipGetErrorString(e));
  for (int i = 0; i < d->nc; i++)
    for (int k = 0; k < 5; k++) out[5 * (size_t)i + k] = w[(size_t)k * d->nc + i];
  return CSU_OK;
}

extern "C" int csu_dev_surf_faces(csu_dev* d, double* out) {
  if (!d || !out) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  if (d->nv != 5 && d->nv != 6) return set_err(CSU_ERR_ARG, "not an SA or SST run");
  const int nb = d->nb;
  if (nb == 0) return CSU_OK;
  const size_t n = 11 * (size_t)nb;
  double* buf = nullptr;
  if (hipMalloc(&buf, n * sizeof(double)) != hipSuccess)
    return set_err(CSU_ERR_NOMEM, "hipMalloc(%zu bytes) failed", n * sizeof(double));
  if (d->nv == 6)
    hipLaunchKernelGGL(k_surf<6>, dim3((nb + CSU_BLOCK - 1) / CSU_BLOCK), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->geo,
                       d->st, buf);
  else
    hipLaunchKernelGGL(k_surf<5>, dim3((nb + CSU_BLOCK - 1) / CSU_BLOCK), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->geo,
                       d->st, buf);
  std::vector<double> w(n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(w.data(), buf, n * sizeof(double), hipMemcpyDeviceToHost, d->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
  const hipError_t ef = hipFree(buf);
  if (e == hipSuccess) e = ef;
  if (e != hipSuccess) return set_err(CSU_ERR_HIP, "surface faces: %s", hipGetErrorString(e));
  for (int q = 0; q < nb; q++)
    for (int k = 0; k < 11; k++) out[11 * (size_t)q + k] = w[(size_t)k * nb + q];
  return CSU_OK;
}

// ---------------------------------------------------------------------------
// allocation-free surface data
// ---------------------------------------------------------------------------
// A fixed staging array in device memory (one copy per GPU, held until the
// process ends) replaces the per-call buffer of csu_dev_surf_faces.  The
// array's name is only a host-side stand-in: host code uses the address
// hipGetSymbolAddress returns on each device.  k_surf writes all 11 entries
// of every face it covers, so nothing in it needs clearing between calls.
static __device__ __align__(256) double scratch[11 * CSU_SURF_NOALLOC_MAX_NB];
static std::mutex g_scratch_mu;
static double* g_scratch_dev[CSU_MAX_DEVICES];

// The context's device's copy of `scratch` (looked up once per device, with
// that device current).
static int scratch_on(const csu_dev* d, double** out) {
  if (d->device < 0 || d->device >= CSU_MAX_DEVICES) return set_err(CSU_ERR_ARG, "device %d", d->device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[d->device]) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != d->device) HIPCHK(hipSetDevice(d->device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (cur != d->device) HIPCHK(hipSetDevice(cur));
    if (e != hipSuccess || !ptr)
      return set_err(CSU_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", d->device, hipGetErrorString(e));
    g_scratch_dev[d->device] = (double*)ptr;
  }
  *out = g_scratch_dev[d->device];
  return CSU_OK;
}

extern "C" int csu_dev_surf_faces_noalloc(csu_dev* d, double* out) {
  if (!d || !out) return set_err(CSU_ERR_ARG, "null argument");
  if (d->nsolve > 1) return set_err(CSU_ERR_UNSUPPORTED, "%s: not on a batch context (use csu_batch_*)", __func__);
  if (d->nv != 5 && d->nv != 6) return set_err(CSU_ERR_ARG, "not an SA or SST run");
  const int nb = d->nb;
  if (nb > CSU_SURF_NOALLOC_MAX_NB)
    return set_err(CSU_ERR_ARG, "%d boundary faces: the allocation-free form stages at most %d", nb,
                   CSU_SURF_NOALLOC_MAX_NB);
  if (nb == 0) return CSU_OK;
  double* buf = nullptr;
  int rc = scratch_on(d, &buf);
  if (rc) return rc;
  if (d->nv == 6)
    hipLaunchKernelGGL((k_surf<6, true>), dim3((nb + CSU_BLOCK - 1) / CSU_BLOCK), dim3(CSU_BLOCK), 0, d->stream,
                       d->kp, d->geo, d->st, buf);
  else
    hipLaunchKernelGGL((k_surf<5, true>), dim3((nb + CSU_BLOCK - 1) / CSU_BLOCK), dim3(CSU_BLOCK), 0, d->stream,
                       d->kp, d->geo, d->st, buf);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, buf, 11 * sizeof(double) * (size_t)nb, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

// HBM read + write rate of a plain FP64 streaming kernel over two arrays
// of `bytes` each (the practical ceiling the stage kernels are compared
// with), best of `reps` timed launches, on the current device.
extern "C" int csu_stream_rate(int64_t bytes, int32_t reps, double* gbps) {
  if (!gbps || bytes < 1024 || reps <= 0) return set_err(CSU_ERR_ARG, "bytes %lld reps %d", (long long)bytes, reps);
  const size_t n2 = (size_t)bytes / 16;
  double2 *x = nullptr, *y = nullptr;
  if (hipMalloc(&x, n2 * 16) != hipSuccess) return set_err(CSU_ERR_NOMEM, "stream buffer");
  if (hipMalloc(&y, n2 * 16) != hipSuccess) {
    (void)hipFree(x);
    return set_err(CSU_ERR_NOMEM, "stream buffer");
  }
  int rc = CSU_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess || hipMemset(x, 0, n2 * 16) != hipSuccess)
    rc = set_err(CSU_ERR_HIP, "stream set-up");
  if (n2 > (size_t)CSU_BLOCK * 0x7fffffffu) rc = set_err(CSU_ERR_ARG, "stream of %lld bytes", (long long)bytes);
  const unsigned grid = (unsigned)((n2 + CSU_BLOCK - 1) / CSU_BLOCK);
  float best = 1e30f;
  for (int r = 0; r <= reps && !rc; r++) {   // launch 0 warms up
    float ms = 0.f;
    if (hipEventRecord(e0, nullptr) != hipSuccess) { rc = set_err(CSU_ERR_HIP, "stream timing"); break; }
    hipLaunchKernelGGL(k_stream, dim3(grid), dim3(CSU_BLOCK), 0, nullptr, x, y, n2, 1.000001);
    if (hipEventRecord(e1, nullptr) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = set_err(CSU_ERR_HIP, "stream timing");
    else if (r > 0 && ms < best)
      best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(x);
  (void)hipFree(y);
  if (rc) return rc;
  *gbps = 2.0 * 16.0 * (double)n2 / (best * 1e-3) / 1e9;
  return CSU_OK;
}

extern "C" int64_t csu_dev_bytes(const csu_dev* d) { return d ? (int64_t)d->bytes : -1; }

// Device bytes csu_dev_create will allocate (plus a history ring of
// hist iterations), from the allocation list of csu_dev_create.
extern "C" int64_t csu_dev_footprint(const csu_mesh_desc* m, const csu_params* p, int32_t hist) {
  if (!m || !p) return -1;
  const double nc = m->ncell, nf = m->nface, nb = m->nbface, nv = p->nvar;
  const bool visc = p->laminar || p->sa || p->sst;
  double b = 0;
  b += nc * (4 + 16 + 16 + 48 + 64 + 64 + 8) + nf * (8 + 1 + 24) + nb * (12 + 64 + 4);  // geometry
  b += nc * (24 * nv + 8) + 2 * nv * 8 * (nc / 256 + 1) + 128;                           // U, Uaux, R, dt, reductions
  b += (nv + 1) * 8 * (nc + nb) + 24 * nv * nc + 8 * nv * nf;                             // P, Pb, grad, phi, fflux
  if (p->order == 2) b += 32 * nv * nc;                                                   // edge states
  if (visc) b += nc * (16 + 16 + 48) + nf * (8 * (nv - 1) + 8) + nb * (8 * (nv - 1) + 16);
  if (p->time_scheme != CSU_TIME_RK) b += nc * (8 + 16 * nv + 8 + 16);
  if (p->time_scheme == CSU_TIME_LUSGS) b += nc * 24 + nf * 24;
  if (p->time_scheme == CSU_TIME_DPLUR) b += 32 * nv * nv * nc + 8.0 * (p->nlinear > 0 ? p->nlinear : 1) * (nc / 256 + 1);
  if (p->precision == 1) b += nc * 64 + nf * (12 + 4 * nv) + (p->order == 2 ? 16 * nv * nc : 0);  // FP32 copies
  if (p->time_scheme == CSU_TIME_RK && !visc && p->order == 2) b += 8 * nc + 8 * nf + 4 * (nc / 256 + 2);  // tile lists
  b += (nv + 4) * 8 * (double)(hist > 0 ? hist : 1);
  // allocation granularity: every array rounds up (at most 2 MiB slack per large array)
  return (int64_t)(b * 1.02) + (int64_t)64 * 4096;
}
