// csu_dev.hip -- device context, kernels and C ABI of the residual-and-update
// path (include/csu.h).  One context = one mesh on one GPU, all state
// resident in HBM, one HIP stream.
//
// Per explicit stage (reference solverCalcR + solverRK, solver.c:685-766):
//   k_prim    cells  U -> primitive P, positivity clamps written back to U
//   k_lim1    cells  (limiter type 1 only) global min/max of P
//   k_grad    cells  ghost states of the cell's boundary faces, least-squares
//                    gradient (double-double sums), Venkatakrishnan phi
//   k_flux    faces  reconstruction + Riemann flux -> per-face flux buffer
//   k_update  cells  gather of face fluxes in the reference's per-cell order,
//                    boundary fluxes, then U_stage = U^n - c*R/Omega
// Per iteration: k_dt (local time step), residual norm, wall forces.
// Every cell gathers its own faces, so no atomics and a fixed summation
// order: results do not depend on scheduling.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <algorithm>
#include <string>
#include "csu.h"
#include "device_math.h"

using namespace csu;

#define CSU_BLOCK 256

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
extern "C" const char* csu_version(void) { return "csu-mi355x 0.1 (gfx950)"; }

// ---------------------------------------------------------------------------
// kernel parameter block (passed by value)
// ---------------------------------------------------------------------------
struct KP {
  int nc, nf, nb, nv, order, flux, laminar, sa, sst, limiter, turb1order;
  double rlim, pout, twall, efix, minf, limK, volmax;
  double pin[7];
  GasD gas;
};

struct csu_dev {
  int nc, nf, nb, nn, nv, nmark, iter;
  csu_params prm;
  KP kp;
  hipStream_t stream;
  // geometry
  int *cnp, *cnbr, *cface;                 // [nc], [4][nc], [4][nc]
  double *xc, *yc, *omega, *dsxm, *dsym;    // [nc]
  double *edx, *edy;                       // [4][nc] own-edge midpoint offsets
  int *fe0, *fe1;                          // [nf]
  double *fdsx, *fdsy, *fds, *fdx0, *fdy0, *fdx1, *fdy1;   // [nf]
  int *be0, *btype;                        // [nb]
  double *bdsx, *bdsy, *bds, *bxm, *bym;   // [nb]
  int nwall;
  int* wall;                               // [nwall] boundary ids of wall/wallT faces
  // state
  double *U, *Uaux, *P, *Pb, *dPx, *dPy, *phi, *R, *fflux, *dtl;
  double *pref2;                           // [nv] limiter reference (type 0 const, type 1 per stage)
  double *red;                             // reduction scratch
  int nred;                                // number of partial blocks
  double* dtg;                             // global time step (device scalar)
  int fresh;                               // state set from the host: P and dt not yet derived
  double *res_hist, *coef_hist;            // [cap][nv], [cap][4]
  int hist_cap;
  // profiling
  int prof;
  std::vector<std::string> kname;
  std::vector<double> ktime;
  std::vector<int> kcount;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> evpool;
};

static inline int nblk(int n) { return (n + CSU_BLOCK - 1) / CSU_BLOCK; }
static inline int slot_of(int kk) { return kk > 3 ? kk + 1 : kk; }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Primitive recovery of one cell with the positivity clamps, reference
// solverCalcPrimitive (solver.c:1153-1213).  u[] (conservative, in
// registers) receives the clamped values the reference writes back into U;
// p[] receives r u v p T [nu~ | k om].
template <int NV>
__device__ __forceinline__ void prim_point(const KP& kp, double* u, double* p) {
  const double r0 = fmax(u[0], kp.rlim);
  u[0] = r0;
  double k = 0;
  if (NV == 6) {
    k = u[4] / r0;
    if (k < 1e-14) { k = 1e-14; u[4] = r0 * k; }
    p[5] = k;
    double om = u[5] / r0;
    if (om <= 0.0) { om = 0.0; u[5] = r0 * om; }
    p[6] = om;
  }
  const double uu = u[1] / r0, vv = u[2] / r0;
  double e = u[3] / r0 - 0.5 * (uu * uu + vv * vv) - k;
  e = fmax(e, 7.176325e+02);
  double T = gas_e2T<double>(kp.gas, e);
  if (T < 1) {
    T = 1;
    e = gas_T2e<double>(kp.gas, T);
    u[3] = (e + 0.5 * (uu * uu + vv * vv) + k) * r0;
  }
  p[0] = r0; p[1] = uu; p[2] = vv;
  p[3] = kp.gas.R * T * r0;
  p[4] = T;
  if (NV == 5) {
    double n = u[4] / r0;
    n = fmax(n, 0.0);
    u[4] = n * r0;
    p[5] = n;
  }
}

// Standalone primitive pass (first stage after the state was set from the
// host).  Clamped values are written back to U, as the reference does.
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_prim(KP kp, double* __restrict__ U, double* __restrict__ P) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc;
  if (i >= nc) return;
  double u[NV], u0[NV], p[NV + 1];
#pragma unroll
  for (int k = 0; k < NV; k++) u0[k] = u[k] = U[k * nc + i];
  prim_point<NV>(kp, u, p);
#pragma unroll
  for (int k = 0; k < NV; k++)
    if (__double_as_longlong(u[k]) != __double_as_longlong(u0[k])) U[k * nc + i] = u[k];
#pragma unroll
  for (int m = 0; m < NV + 1; m++) P[m * nc + i] = p[m];
}

// Ghost state of a boundary face from the interior primitive Pi (nv+1 slots),
// reference boundaryCalcPrimitive (boundary.c:514-693).
template <int NV>
__device__ __forceinline__ void ghost_state(const KP& kp, int type, double dSx, double dSy, double dS,
                                            const double* Pi, double wall_d, double* Pg) {
  double PL[4] = {Pi[0], Pi[1], Pi[2], Pi[3]};
  double Pb[4];
  const double R = kp.gas.R;
  if (type == CSU_BC_SYMMETRY) {
    if (dS > 0) {
      rotate<double>(PL[1], PL[2], dSx, dSy, dS);
      PL[1] = 0.0;
      rotate<double>(PL[1], PL[2], dSx, -dSy, dS);
    } else {
      PL[2] = 0.0;
    }
    for (int k = 0; k < 4; k++) Pg[k] = PL[k];
    Pg[4] = Pg[3] / (Pg[0] * R);
  } else if (type == CSU_BC_INLET) {
    bc_inlet(kp.gas, kp.pin, PL, Pb, dSx / dS, dSy / dS);
    for (int k = 0; k < 4; k++) Pg[k] = Pb[k];
    Pg[4] = Pg[3] / (Pg[0] * R);
  } else if (type == CSU_BC_OUTLET) {
    bc_outlet(kp.gas, kp.pout, PL, Pb, dSx / dS, dSy / dS);
    for (int k = 0; k < 4; k++) Pg[k] = Pb[k];
    Pg[4] = Pg[3] / (Pg[0] * R);
  } else {
    if (kp.laminar == 1 || kp.sa == 1) {
      if (type == CSU_BC_WALLT) {
        Pg[0] = PL[3] / (R * kp.twall);
        Pg[1] = 0.0; Pg[2] = 0.0; Pg[3] = PL[3]; Pg[4] = kp.twall;
      } else {
        Pg[0] = PL[0]; Pg[1] = 0.0; Pg[2] = 0.0; Pg[3] = PL[3];
        Pg[4] = Pg[3] / (Pg[0] * R);
      }
    } else {
      if (dS > 0) {
        bc_wall(kp.gas, PL, Pb, dSx / dS, dSy / dS);
        for (int k = 0; k < 4; k++) Pg[k] = Pb[k];
      } else {
        PL[2] = 0.0;
        for (int k = 0; k < 4; k++) Pg[k] = PL[k];
      }
      Pg[4] = Pg[3] / (Pg[0] * R);
    }
  }
  if (NV == 5) {  // SA nu~
    if (type == CSU_BC_SYMMETRY || type == CSU_BC_OUTLET) Pg[5] = Pi[5];
    else if (type == CSU_BC_INLET) Pg[5] = kp.pin[5];
    else Pg[5] = 0.0;
  }
  (void)wall_d;
}

// Gradient + limiter kernel.  Reference solverUpdateGrad/solverCalcGrad3
// (solver.c:451-474, 1031-1074) and the limiter loop of inter()
// (solver.c:496-548, limiter.c:98-147).
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_grad(
    KP kp, const int* __restrict__ cnp, const int* __restrict__ cnbr,
    const double* __restrict__ xc, const double* __restrict__ yc, const double* __restrict__ omega,
    const double* __restrict__ edx, const double* __restrict__ edy,
    const double* __restrict__ bdsx, const double* __restrict__ bdsy, const double* __restrict__ bds,
    const double* __restrict__ bxm, const double* __restrict__ bym, const int* __restrict__ btype,
    const double* __restrict__ P, double* __restrict__ Pb, const double* __restrict__ pref20,
    double* __restrict__ dPx, double* __restrict__ dPy, double* __restrict__ phi) {
  constexpr int NS = NV + 1;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc, nb = kp.nb;
  if (i >= nc) return;
  const int np = cnp[i];
  const double x0 = xc[i], y0 = yc[i];
  double Pi[NS];
#pragma unroll
  for (int m = 0; m < NS; m++) Pi[m] = P[m * nc + i];
  double Pn[4][NS];
  double dx[4], dy[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (j >= np) break;
    int n = cnbr[j * nc + i];
    if (n >= 0) {
      dx[j] = xc[n] - x0;
      dy[j] = yc[n] - y0;
#pragma unroll
      for (int m = 0; m < NS; m++) Pn[j][m] = P[m * nc + n];
    } else {
      int g = -1 - n;
      dx[j] = bxm[g] - x0;
      dy[j] = bym[g] - y0;
      ghost_state<NV>(kp, btype[g], bdsx[g], bdsy[g], bds[g], Pi, 0.0, Pn[j]);
#pragma unroll
      for (int m = 0; m < NS; m++) Pb[m * nb + g] = Pn[j][m];
    }
  }
  // geometric normal-equation terms, sums of FP64 products in double-double
  dd a = {0, 0}, b = {0, 0}, c = {0, 0};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (j >= np) break;
    a = dd_add_d(a, dx[j] * dx[j]);
    b = dd_add_d(b, dx[j] * dy[j]);
    c = dd_add_d(c, dy[j] * dy[j]);
  }
  const dd det = dd_sub(dd_mul(a, c), dd_mul(b, b));
  const dd nb_ = dd_neg(b);
  double lim_aux = 0.0;
  if (kp.limiter == 0) {
    lim_aux = kp.limK * sqrt(omega[i] / kp.volmax);
    lim_aux = lim_aux * lim_aux * lim_aux;
  }
#pragma unroll
  for (int kk = 0; kk < NV; kk++) {
    const int m = kk > 3 ? kk + 1 : kk;
    const double u0 = Pi[m];
    dd A = {0, 0}, B = {0, 0};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (j >= np) break;
      double du = Pn[j][m] - u0;
      A = dd_add_d(A, dx[j] * du);
      B = dd_add_d(B, dy[j] * du);
    }
    double gx = dd_div(dd_sub(dd_mul(c, A), dd_mul(b, B)), det).hi;
    double gy = dd_div(dd_add(dd_mul(nb_, A), dd_mul(a, B)), det).hi;
    dPx[kk * nc + i] = gx;
    dPy[kk * nc + i] = gy;
    if (kp.order == 2) {
      double pref2 = kp.limiter == 0 ? pref20[kk] * lim_aux : pref20[kk];
      double Pmin = u0, Pmax = u0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j >= np) break;
        Pmax = fmax(Pmax, Pn[j][m]);
        Pmin = fmin(Pmin, Pn[j][m]);
      }
      double ph = 0.0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j >= np) break;
        double d2 = gx * edx[j * nc + i] + gy * edy[j * nc + i];
        double p0 = limiter_v2<double>(u0, Pmin, Pmax, d2, pref2);
        ph = (j == 0) ? p0 : fmin(ph, p0);
      }
      phi[kk * nc + i] = ph;
    }
  }
}

// Limiter type 1 global range (limiter.c:64-94): per-block min/max partials.
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_lim1_partial(KP kp, const double* __restrict__ P,
                                                            double* __restrict__ red) {
  __shared__ double smin[NV][CSU_BLOCK], smax[NV][CSU_BLOCK];
  int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const int nc = kp.nc;
  int ii = i < nc ? i : 0;
#pragma unroll
  for (int jj = 0; jj < NV; jj++) {
    int mm = jj == 4 ? 5 : jj;  // limiter.c:350-357 (slot 5 also for jj==5)
    double v = P[mm * nc + ii];
    smin[jj][t] = v; smax[jj][t] = v;
  }
  __syncthreads();
  for (int s = CSU_BLOCK / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int jj = 0; jj < NV; jj++) {
        smin[jj][t] = fmin(smin[jj][t], smin[jj][t + s]);
        smax[jj][t] = fmax(smax[jj][t], smax[jj][t + s]);
      }
    __syncthreads();
  }
  if (t == 0)
    for (int jj = 0; jj < NV; jj++) {
      red[(2 * jj) * gridDim.x + blockIdx.x] = smin[jj][0];
      red[(2 * jj + 1) * gridDim.x + blockIdx.x] = smax[jj][0];
    }
}

template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_lim1_final(KP kp, int nblocks, const double* __restrict__ red,
                                                          double* __restrict__ pref2) {
  __shared__ double smin[CSU_BLOCK], smax[CSU_BLOCK];
  int t = threadIdx.x;
  for (int jj = 0; jj < NV; jj++) {
    double mn = red[(2 * jj) * nblocks], mx = red[(2 * jj + 1) * nblocks];
    for (int b = t; b < nblocks; b += CSU_BLOCK) {
      mn = fmin(mn, red[(2 * jj) * nblocks + b]);
      mx = fmax(mx, red[(2 * jj + 1) * nblocks + b]);
    }
    smin[t] = mn; smax[t] = mx;
    __syncthreads();
    for (int s = CSU_BLOCK / 2; s > 0; s >>= 1) {
      if (t < s) { smin[t] = fmin(smin[t], smin[t + s]); smax[t] = fmax(smax[t], smax[t + s]); }
      __syncthreads();
    }
    if (t == 0) {
      double aux = (smax[0] - smin[0]) * kp.limK;
      pref2[jj] = aux * aux;
    }
    __syncthreads();
  }
}

// Interior face flux, reference inter() face loop (solver.c:550-644).
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_flux(
    KP kp, const int* __restrict__ fe0, const int* __restrict__ fe1,
    const double* __restrict__ fdsx, const double* __restrict__ fdsy, const double* __restrict__ fds,
    const double* __restrict__ fdx0, const double* __restrict__ fdy0,
    const double* __restrict__ fdx1, const double* __restrict__ fdy1,
    const double* __restrict__ P, const double* __restrict__ dPx, const double* __restrict__ dPy,
    const double* __restrict__ phi, double* __restrict__ fflux) {
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc, nf = kp.nf;
  if (f >= nf) return;
  const int e0 = fe0[f], e1 = fe1[f];
  const double dSx = fdsx[f], dSy = fdsy[f], dS = fds[f];
  double PL[NV], PR[NV];
  if (kp.order == 2) {
    const double dx0 = fdx0[f], dy0 = fdy0[f], dx1 = fdx1[f], dy1 = fdy1[f];
#pragma unroll
    for (int kk = 0; kk < NV; kk++) {
      const int m = kk > 3 ? kk + 1 : kk;
      if (kk > 3 && kp.turb1order) {
        PL[kk] = P[m * nc + e0];
        PR[kk] = P[m * nc + e1];
      } else {
        PL[kk] = P[m * nc + e0] + (dPx[kk * nc + e0] * dx0 + dPy[kk * nc + e0] * dy0) * phi[kk * nc + e0];
        PR[kk] = P[m * nc + e1] + (dPx[kk * nc + e1] * dx1 + dPy[kk * nc + e1] * dy1) * phi[kk * nc + e1];
      }
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) {
      const int m = kk > 3 ? kk + 1 : kk;
      PL[kk] = P[m * nc + e0];
      PR[kk] = P[m * nc + e1];
    }
  }
  rotate<double>(PL[1], PL[2], dSx, dSy, dS);
  rotate<double>(PR[1], PR[2], dSx, dSy, dS);
  double fl[NV];
  constexpr int NSC = NV == 5 ? 1 : 0;
  flux_dispatch<double, NSC>(kp.gas, kp.flux, kp.efix, kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4,
                             PR[0], PR[1], PR[2], PR[3], PR + 4, 0.0, 0.0, fl);
  rotate<double>(fl[1], fl[2], dSx, -dSy, dS);
#pragma unroll
  for (int kk = 0; kk < NV; kk++) fflux[kk * nf + f] = fl[kk] * dS;
}

// Convective boundary flux of one boundary face, reference boundaryCalc /
// boundaryCalc_sa (boundary.c:143-341).  Adds f*dS to R when dS > 0.
template <int NV>
__device__ __forceinline__ void bc_flux(const KP& kp, int type, double dSx, double dSy, double dS,
                                        const double* Pi, double* R) {
  double PL[NV], Pb[NV], fl[NV];
  PL[0] = Pi[0]; PL[1] = Pi[1]; PL[2] = Pi[2]; PL[3] = Pi[3];
  if (NV == 5) PL[4] = Pi[5];
  constexpr int NSC = NV == 5 ? 1 : 0;
  if (type == CSU_BC_SYMMETRY) {
    rotate<double>(PL[1], PL[2], dSx, dSy, dS);
    double Pm[NV];
    Pm[0] = PL[0]; Pm[1] = -PL[1]; Pm[2] = PL[2]; Pm[3] = PL[3];
    if (NV == 5) Pm[4] = PL[4];
    flux_dispatch<double, NSC>(kp.gas, kp.flux, kp.efix, kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4,
                               Pm[0], Pm[1], Pm[2], Pm[3], Pm + 4, 0.0, 0.0, fl);
  } else if (type == CSU_BC_INLET || type == CSU_BC_OUTLET) {
    if (type == CSU_BC_INLET) bc_inlet(kp.gas, kp.pin, PL, Pb, dSx / dS, dSy / dS);
    else bc_outlet(kp.gas, kp.pout, PL, Pb, dSx / dS, dSy / dS);
    if (NV == 5) Pb[4] = (type == CSU_BC_INLET) ? kp.pin[5] : PL[4];
    rotate<double>(PL[1], PL[2], dSx, dSy, dS);
    rotate<double>(Pb[1], Pb[2], dSx, dSy, dS);
    flux_dispatch<double, NSC>(kp.gas, kp.flux, kp.efix, kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4,
                               Pb[0], Pb[1], Pb[2], Pb[3], Pb + 4, 0.0, 0.0, fl);
  } else {
    if (kp.laminar == 1 || NV == 5) {
      rotate<double>(PL[1], PL[2], dSx, dSy, dS);
      double Pm[NV];
      Pm[0] = PL[0]; Pm[1] = -PL[1]; Pm[2] = -PL[2]; Pm[3] = PL[3];
      if (NV == 5) Pm[4] = PL[4];
      flux_dispatch<double, NSC>(kp.gas, kp.flux, kp.efix, kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4,
                                 Pm[0], Pm[1], Pm[2], Pm[3], Pm + 4, 0.0, 0.0, fl);
    } else {
      bc_wall(kp.gas, PL, Pb, dSx / dS, dSy / dS);
      rotate<double>(Pb[1], Pb[2], dSx, dSy, dS);
      flux_free<double>(kp.gas, Pb[0], Pb[1], Pb[2], Pb[3], fl);
    }
  }
  rotate<double>(fl[1], fl[2], dSx, -dSy, dS);
  if (dS > 0) {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) R[kk] += fl[kk] * dS;
  }
}

// Residual gather + boundary fluxes, then (flags) the RK stage update fused
// with the next stage's primitive recovery, and on the last stage the next
// iteration's time step and the residual-norm partial sums.
// Reference: gather solver.c:646-667, boundary() boundary.c:447-464,
// solverRK solver.c:741-766, solverCalcPrimitive solver.c:1153-1213,
// solverLocalTimeStep solver.c:895-923, solverCalcRes solver.c:867-892.
enum { F_RK = 1, F_CLAMP_UN = 2, F_LAST = 4, F_NEXT_LOCAL = 8, F_STORE_R = 16, F_CUR_LOCAL = 32 };

template <int NV>
__device__ __forceinline__ double cell_dt(const KP& kp, const double* u, double om, double sx, double sy) {
  // solverLocalTimeStep + solverCalcVel, from the unclamped conservative state
  double k = 0;
  if (NV == 6) k = u[4] / u[0];
  double E = u[3] / u[0];
  double uu = u[1] / u[0];
  double vv = u[2] / u[0];
  double aux = (E - (uu * uu + vv * vv) / 2 - k);
  double T = gas_e2T<double>(kp.gas, aux);
  double c = gas_T2c<double>(kp.gas, T);
  double Lx = (fabs(uu) + c) * sx;
  double Ly = (fabs(vv) + c) * sy;
  return om / (Lx + Ly);
}

template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_update(
    KP kp, int flags, double sigma, double rk_a, double cfl, const int* __restrict__ cnp,
    const int* __restrict__ cnbr, const int* __restrict__ cface,
    const double* __restrict__ bdsx, const double* __restrict__ bdsy, const double* __restrict__ bds,
    const int* __restrict__ btype, double* P, const double* __restrict__ fflux,
    const double* __restrict__ omega, const double* __restrict__ dsxm, const double* __restrict__ dsym,
    double* dtl, const double* dtg, double* Un, double* __restrict__ Uout, double* __restrict__ Rout,
    double* __restrict__ red_res, double* __restrict__ red_dt) {
  __shared__ double sred[NV + 1][CSU_BLOCK];
  const int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const int nc = kp.nc, nf = kp.nf;
  const bool active = i < nc;
  double R[NV];
#pragma unroll
  for (int kk = 0; kk < NV; kk++) R[kk] = 0.0;
  double dt_next = INFINITY;
  if (active) {
    const int np = cnp[i];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int f = cface[j * nc + i];
      if (f > 0) {
#pragma unroll
        for (int kk = 0; kk < NV; kk++) R[kk] += fflux[kk * nf + f - 1];
      } else if (f < 0) {
#pragma unroll
        for (int kk = 0; kk < NV; kk++) R[kk] -= fflux[kk * nf - f - 1];
      }
    }
    double Pi[NV + 1];
    bool loaded = false;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (j >= np) break;
      int n = cnbr[j * nc + i];
      if (n < 0) {
        int g = -1 - n;
        if (!loaded) {
#pragma unroll
          for (int m = 0; m < NV + 1; m++) Pi[m] = P[m * nc + i];
          loaded = true;
        }
        bc_flux<NV>(kp, btype[g], bdsx[g], bdsy[g], bds[g], Pi, R);
      }
    }
    if (flags & F_STORE_R) {
#pragma unroll
      for (int kk = 0; kk < NV; kk++) Rout[kk * nc + i] = R[kk];
    }
    if (flags & F_RK) {
      double un[NV], un0[NV], p[NV + 1];
#pragma unroll
      for (int kk = 0; kk < NV; kk++) un0[kk] = un[kk] = Un[kk * nc + i];
      if (flags & F_CLAMP_UN) {
        // first stage: the reference clamps U^n in place (solverCalcR(U))
        prim_point<NV>(kp, un, p);
#pragma unroll
        for (int kk = 0; kk < NV; kk++)
          if (__double_as_longlong(un[kk]) != __double_as_longlong(un0[kk])) Un[kk * nc + i] = un[kk];
      }
      const double dt = dtl && (flags & F_CUR_LOCAL) ? dtl[i] : dtg[0];
      const double om = omega[i];
      double unew[NV];
#pragma unroll
      for (int kk = 0; kk < NV; kk++) unew[kk] = un[kk] - dt * sigma * rk_a * R[kk] / om;
      if (flags & F_LAST) {
        // U^{n+1} stays unclamped until the next residual, as in the reference
        dt_next = cell_dt<NV>(kp, unew, om, dsxm[i], dsym[i]);
        if (flags & F_NEXT_LOCAL) dtl[i] = cfl * dt_next;
#pragma unroll
        for (int kk = 0; kk < NV; kk++) Uout[kk * nc + i] = unew[kk];
        prim_point<NV>(kp, unew, p);
      } else {
        prim_point<NV>(kp, unew, p);
#pragma unroll
        for (int kk = 0; kk < NV; kk++) Uout[kk * nc + i] = unew[kk];
      }
#pragma unroll
      for (int m = 0; m < NV + 1; m++) P[m * nc + i] = p[m];
    }
  }
  if (!(flags & F_LAST)) return;
  // residual-norm partials (and global time-step minimum): fixed tree
#pragma unroll
  for (int kk = 0; kk < NV; kk++) sred[kk][t] = R[kk] * R[kk];
  sred[NV][t] = dt_next;
  __syncthreads();
  for (int st = CSU_BLOCK / 2; st > 0; st >>= 1) {
    if (t < st) {
#pragma unroll
      for (int kk = 0; kk < NV; kk++) sred[kk][t] += sred[kk][t + st];
      sred[NV][t] = fmin(sred[NV][t], sred[NV][t + st]);
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) red_res[kk * gridDim.x + blockIdx.x] = sred[kk][0];
    red_dt[blockIdx.x] = sred[NV][0];
  }
}

// Local time step, reference solverLocalTimeStep (solver.c:895-923) with
// solverCalcVel (solver.c:420-439): from the conservative state, unclamped.
// Global mode writes omega/(Lx+Ly) for a min-reduction instead.
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_dt(KP kp, int local, double cfl, const double* __restrict__ U,
                                                  const double* __restrict__ omega, const double* __restrict__ dsxm,
                                                  const double* __restrict__ dsym, double* __restrict__ dtl,
                                                  double* __restrict__ red) {
  __shared__ double smin[CSU_BLOCK];
  int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const int nc = kp.nc;
  double val = INFINITY;
  if (i < nc) {
    double U0 = U[i];
    double k = 0;
    if (NV == 6) k = U[4 * nc + i] / U0;
    double E = U[3 * nc + i] / U0;
    double u = U[nc + i] / U0;
    double v = U[2 * nc + i] / U0;
    double aux = (E - (u * u + v * v) / 2 - k);
    double T = gas_e2T<double>(kp.gas, aux);
    double c = gas_T2c<double>(kp.gas, T);
    double Lx = (fabs(u) + c) * dsxm[i];
    double Ly = (fabs(v) + c) * dsym[i];
    val = omega[i] / (Lx + Ly);
    if (local) dtl[i] = cfl * val;
  }
  if (local) return;
  smin[t] = val;
  __syncthreads();
  for (int s = CSU_BLOCK / 2; s > 0; s >>= 1) {
    if (t < s) smin[t] = fmin(smin[t], smin[t + s]);
    __syncthreads();
  }
  if (t == 0) red[blockIdx.x] = smin[0];
}

__global__ __launch_bounds__(CSU_BLOCK) void k_dt_final(int nblocks, double cfl, const double* __restrict__ red,
                                                        double* __restrict__ dtglob) {
  __shared__ double smin[CSU_BLOCK];
  int t = threadIdx.x;
  double m = INFINITY;
  for (int b = t; b < nblocks; b += CSU_BLOCK) m = fmin(m, red[b]);
  smin[t] = m;
  __syncthreads();
  for (int s = CSU_BLOCK / 2; s > 0; s >>= 1) {
    if (t < s) smin[t] = fmin(smin[t], smin[t + s]);
    __syncthreads();
  }
  if (t == 0) dtglob[0] = smin[0] * cfl;
}

// Residual norm sqrt(sum R^2 / Nc) per variable, reference solverCalcRes
// (solver.c:867-892); fixed two-level tree, independent of scheduling.
// Final level of the residual norm: one block, each thread sums a strided
// slice of the block partials for all variables at once (independent
// chains), then a fixed tree.  Deterministic for a given launch shape.
#define RES_THREADS 1024
template <int NV>
__global__ __launch_bounds__(RES_THREADS) void k_res_final(int nc, int nblocks, const double* __restrict__ red,
                                                           double* __restrict__ res) {
  __shared__ double s[NV][RES_THREADS];
  const int t = threadIdx.x;
  double a[NV];
#pragma unroll
  for (int kk = 0; kk < NV; kk++) a[kk] = 0.0;
  for (int b = t; b < nblocks; b += RES_THREADS) {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) a[kk] += red[kk * nblocks + b];
  }
#pragma unroll
  for (int kk = 0; kk < NV; kk++) s[kk][t] = a[kk];
  __syncthreads();
  for (int st = RES_THREADS / 2; st > 0; st >>= 1) {
    if (t < st) {
#pragma unroll
      for (int kk = 0; kk < NV; kk++) s[kk][t] += s[kk][t + st];
    }
    __syncthreads();
  }
  if (t < NV) res[t] = sqrt(s[t][0] / nc);
}

// Wall pressure forces, reference solverCalcCoeff3 pressure part
// (solver.c:1342-1401): Cx_p, Cx_v, Cy_p, Cy_v.
__global__ __launch_bounds__(CSU_BLOCK) void k_forces(KP kp, int nwall, const int* __restrict__ wall,
                                                      const double* __restrict__ Pb, const double* __restrict__ bdsx,
                                                      const double* __restrict__ bdsy, double sref, int axi,
                                                      double* __restrict__ coef) {
  __shared__ double sx[CSU_BLOCK], sy[CSU_BLOCK];
  int t = threadIdx.x;
  const double r = kp.pin[0], u = kp.pin[1], v = kp.pin[2], Pinf = kp.pin[3];
  const double q = 0.5 * r * (u * u + v * v);
  double ax = 0.0, ay = 0.0;
  for (int w = t; w < nwall; w += CSU_BLOCK) {
    int g = wall[w];
    double cp = (Pb[3 * kp.nb + g] - Pinf) / q;
    ax += cp * bdsx[g];
    ay += cp * bdsy[g];
  }
  sx[t] = ax; sy[t] = ay;
  __syncthreads();
  for (int s = CSU_BLOCK / 2; s > 0; s >>= 1) {
    if (t < s) { sx[t] += sx[t + s]; sy[t] += sy[t + s]; }
    __syncthreads();
  }
  if (t == 0) {
    double cxp = sx[0] / sref, cyp = sy[0] / sref;
    if (axi) cxp *= 2 * M_PI;
    coef[0] = cxp; coef[1] = 0.0; coef[2] = cyp; coef[3] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <typename T>
static int dalloc(T** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return set_err(CSU_ERR_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return CSU_OK;
}

template <typename T>
static int h2d(T* d, const T* h, size_t n) {
  if (n == 0) return CSU_OK;
  HIPCHK(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
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
        if (n >= m->ncell || n == INT32_MAX) return set_err(CSU_ERR_ARG, "cell %d has %d sides but neighbour %d missing", i, np, k);
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
    return set_err(CSU_ERR_UNSUPPORTED, "RK stages %d (reference supports 1,3,4,5)", p->stages);
  if (p->order != 1 && p->order != 2) return set_err(CSU_ERR_ARG, "order %d", p->order);
  if (p->flux < 0 || p->flux > 4) return set_err(CSU_ERR_ARG, "flux %d", p->flux);
  if (p->limiter != 0 && p->limiter != 1) return set_err(CSU_ERR_ARG, "limiter %d", p->limiter);
  if (p->laminar || p->sa || p->sst) return set_err(CSU_ERR_UNSUPPORTED, "viscous/turbulent terms not built yet");
  if (p->time_scheme != CSU_TIME_RK) return set_err(CSU_ERR_UNSUPPORTED, "implicit schemes not built yet");
  if (p->precision != CSU_PREC_FP64) return set_err(CSU_ERR_UNSUPPORTED, "mixed precision not built yet");
  return CSU_OK;
}

static void free_all(csu_dev* d) {
  void* ptrs[] = {d->cnp, d->cnbr, d->cface, d->xc, d->yc, d->omega, d->dsxm, d->dsym, d->edx, d->edy,
                  d->fe0, d->fe1, d->fdsx, d->fdsy, d->fds, d->fdx0, d->fdy0, d->fdx1, d->fdy1,
                  d->be0, d->btype, d->bdsx, d->bdsy, d->bds, d->bxm, d->bym, d->wall,
                  d->U, d->Uaux, d->P, d->Pb, d->dPx, d->dPy, d->phi, d->R, d->fflux, d->dtl,
                  d->pref2, d->red, d->res_hist, d->coef_hist, d->dtg};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : d->evpool) (void)hipEventDestroy(e);
  for (auto& pe : d->pending) {
    (void)hipEventDestroy(pe.second.first);
    (void)hipEventDestroy(pe.second.second);
  }
  if (d->stream) (void)hipStreamDestroy(d->stream);
}

extern "C" void csu_dev_destroy(csu_dev* d) {
  if (!d) return;
  free_all(d);
  delete d;
}

// geometry in the reference's expressions (mesh.c:371-386, 720-740)
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

extern "C" int csu_dev_create(const csu_mesh_desc* m, const csu_params* p, csu_dev** out) {
  if (!out) return set_err(CSU_ERR_ARG, "null out");
  *out = nullptr;
  int rc = validate(m, p);
  if (rc) return rc;
  csu_dev* d = new csu_dev();
  memset((void*)&d->kp, 0, sizeof(KP));
  d->nc = m->ncell; d->nf = m->nface; d->nb = m->nbface; d->nn = m->nnode; d->nv = p->nvar;
  d->nmark = m->nmark; d->iter = 0; d->prm = *p;
  d->stream = nullptr; d->prof = 0;
  d->cnp = d->cnbr = d->cface = d->fe0 = d->fe1 = d->be0 = d->btype = d->wall = nullptr;
  d->xc = d->yc = d->omega = d->dsxm = d->dsym = d->edx = d->edy = nullptr;
  d->fdsx = d->fdsy = d->fds = d->fdx0 = d->fdy0 = d->fdx1 = d->fdy1 = nullptr;
  d->bdsx = d->bdsy = d->bds = d->bxm = d->bym = nullptr;
  d->U = d->Uaux = d->P = d->Pb = d->dPx = d->dPy = d->phi = d->R = d->fflux = d->dtl = nullptr;
  d->pref2 = d->red = d->res_hist = d->coef_hist = nullptr;
  d->hist_cap = 0;
  d->dtg = nullptr;
  d->fresh = 1;
  const int nc = d->nc, nf = d->nf, nb = d->nb, nv = d->nv;

  // ---- host-side geometry ----
  std::vector<int> cnp(nc), cnbr(4 * (size_t)nc), cface(4 * (size_t)nc);
  std::vector<double> xc(nc), yc(nc), om(nc), dsxm(nc), dsym(nc), edx(4 * (size_t)nc, 0.0), edy(4 * (size_t)nc, 0.0);
  double volmax = 0.0;
  for (int i = 0; i < nc; i++) {
    const int32_t* v = m->cell_vert + 4 * i;
    int np = v[3] >= 0 ? 4 : 3;
    cnp[i] = np;
    elem_center(m, i, &xc[i], &yc[i]);
    double a = tri_omega(m, v[0], v[1], v[2]);
    if (np == 4) a += tri_omega(m, v[2], v[3], v[0]);
    if (m->axi == 1) a *= yc[i];
    om[i] = a;
    volmax = fmax(a, volmax);
    double sx = 0.0, sy = 0.0;
    for (int j = 0; j < np; j++) {
      int p0 = v[j], p1 = v[(j + 1) % np];
      double dSx, dSy;
      calc_ds(m, p0, p1, &dSx, &dSy);
      sx += fabs(dSx);
      sy += fabs(dSy);
      double xm = (m->node_xy[2 * p0] + m->node_xy[2 * p1]) * 0.5;
      double ym = (m->node_xy[2 * p0 + 1] + m->node_xy[2 * p1 + 1]) * 0.5;
      edx[(size_t)j * nc + i] = xm - xc[i];
      edy[(size_t)j * nc + i] = ym - yc[i];
    }
    dsxm[i] = sx * 0.5;
    dsym[i] = sy * 0.5;
    for (int j = 0; j < 4; j++) {
      cnbr[(size_t)j * nc + i] = m->cell_nbr[4 * i + j];
      cface[(size_t)j * nc + i] = m->cell_face[4 * i + j];
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
  for (size_t t = 0; t < cface.size(); t++) {
    int f = cface[t];
    if (f > 0) cface[t] = finv[f - 1] + 1;
    else if (f < 0) cface[t] = -(finv[-f - 1] + 1);
  }
  std::vector<int> fe0(nf), fe1(nf);
  std::vector<double> fdsx(nf), fdsy(nf), fds(nf), fdx0(nf), fdy0(nf), fdx1(nf), fdy1(nf);
  for (int f = 0; f < nf; f++) {
    const int32_t* q = m->face + 4 * ford[f];
    fe0[f] = q[0]; fe1[f] = q[1];
    double dSx, dSy;
    calc_ds(m, q[2], q[3], &dSx, &dSy);
    fdsx[f] = dSx; fdsy[f] = dSy; fds[f] = sqrt(dSx * dSx + dSy * dSy);
    double xm = (m->node_xy[2 * q[2]] + m->node_xy[2 * q[3]]) * 0.5;
    double ym = (m->node_xy[2 * q[2] + 1] + m->node_xy[2 * q[3] + 1]) * 0.5;
    fdx0[f] = xm - xc[q[0]]; fdy0[f] = ym - yc[q[0]];
    fdx1[f] = xm - xc[q[1]]; fdy1[f] = ym - yc[q[1]];
  }
  std::vector<int> be0(nb), btype(nb), wall;
  std::vector<double> bdsx(nb), bdsy(nb), bds(nb), bxm(nb), bym(nb);
  for (int g = 0; g < nb; g++) {
    const int32_t* q = m->bface + 3 * g;
    be0[g] = q[0];
    btype[g] = m->mark_type[m->bface_mark[g]];
    double dSx, dSy;
    calc_ds(m, q[1], q[2], &dSx, &dSy);
    bdsx[g] = dSx; bdsy[g] = dSy; bds[g] = sqrt(dSx * dSx + dSy * dSy);
    // boundary element centre (elementCenter of a 2-point element)
    bxm[g] = (m->node_xy[2 * q[1]] + m->node_xy[2 * q[2]]) / 2;
    bym[g] = (m->node_xy[2 * q[1] + 1] + m->node_xy[2 * q[2] + 1]) / 2;
    if (btype[g] == CSU_BC_WALL || btype[g] == CSU_BC_WALLT) wall.push_back(g);
  }
  d->nwall = (int)wall.size();

  // ---- kernel parameters ----
  KP& kp = d->kp;
  kp.nc = nc; kp.nf = nf; kp.nb = nb; kp.nv = nv; kp.order = p->order; kp.flux = p->flux;
  kp.laminar = p->laminar; kp.sa = p->sa; kp.sst = p->sst; kp.limiter = p->limiter;
  kp.turb1order = p->turb1order; kp.rlim = p->rlim; kp.pout = p->pout; kp.twall = p->twall;
  kp.efix = p->efix; kp.minf = p->mach_inf; kp.limK = p->lim_k; kp.volmax = volmax;
  for (int k = 0; k < 7; k++) kp.pin[k] = p->pin[k];
  kp.gas.gamma = p->gamma; kp.gas.R = p->gas_r;
  kp.gas.Cp = kp.gas.gamma * kp.gas.R / (kp.gas.gamma - 1);
  kp.gas.Cv = kp.gas.Cp - kp.gas.R;
  kp.gas.cc[0] = 1.81716e3; kp.gas.cc[1] = 9.890504e2 - 287.0530; kp.gas.cc[2] = -9.595592e-3;
  kp.gas.cc[3] = 1.041469e-4; kp.gas.cc[4] = -4.433065e-8; kp.gas.cc[5] = 5.879263e-12;
  kp.gas.tp = p->tp;
  // limiter type 0 reference values (limiter.c:45-59)
  std::vector<double> pref(nv, 0.0);
  pref[0] = p->pin[0] * p->pin[0];
  pref[1] = p->pin[1] * p->pin[1] + p->pin[2] * p->pin[2];
  pref[2] = pref[1];
  pref[3] = p->pin[3] * p->pin[3];
  for (int k = 4; k < nv; k++) pref[k] = p->pin[k + 1] * p->pin[k + 1];

  // ---- device allocation ----
  const size_t nblocks = (size_t)nblk(nc);
  d->nred = (int)nblocks;
  HIPCHK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
  if ((rc = dalloc(&d->cnp, nc)) || (rc = dalloc(&d->cnbr, 4 * (size_t)nc)) || (rc = dalloc(&d->cface, 4 * (size_t)nc)) ||
      (rc = dalloc(&d->xc, nc)) || (rc = dalloc(&d->yc, nc)) || (rc = dalloc(&d->omega, nc)) ||
      (rc = dalloc(&d->dsxm, nc)) || (rc = dalloc(&d->dsym, nc)) || (rc = dalloc(&d->edx, 4 * (size_t)nc)) ||
      (rc = dalloc(&d->edy, 4 * (size_t)nc)) || (rc = dalloc(&d->fe0, nf)) || (rc = dalloc(&d->fe1, nf)) ||
      (rc = dalloc(&d->fdsx, nf)) || (rc = dalloc(&d->fdsy, nf)) || (rc = dalloc(&d->fds, nf)) ||
      (rc = dalloc(&d->fdx0, nf)) || (rc = dalloc(&d->fdy0, nf)) || (rc = dalloc(&d->fdx1, nf)) ||
      (rc = dalloc(&d->fdy1, nf)) || (rc = dalloc(&d->be0, nb)) || (rc = dalloc(&d->btype, nb)) ||
      (rc = dalloc(&d->bdsx, nb)) || (rc = dalloc(&d->bdsy, nb)) || (rc = dalloc(&d->bds, nb)) ||
      (rc = dalloc(&d->bxm, nb)) || (rc = dalloc(&d->bym, nb)) || (rc = dalloc(&d->wall, wall.size())) ||
      (rc = dalloc(&d->U, (size_t)nv * nc)) || (rc = dalloc(&d->Uaux, (size_t)nv * nc)) ||
      (rc = dalloc(&d->P, (size_t)(nv + 1) * nc)) || (rc = dalloc(&d->Pb, (size_t)(nv + 1) * nb)) ||
      (rc = dalloc(&d->dPx, (size_t)nv * nc)) || (rc = dalloc(&d->dPy, (size_t)nv * nc)) ||
      (rc = dalloc(&d->phi, (size_t)nv * nc)) || (rc = dalloc(&d->R, (size_t)nv * nc)) ||
      (rc = dalloc(&d->fflux, (size_t)nv * nf)) || (rc = dalloc(&d->dtl, nc + 1)) ||
      (rc = dalloc(&d->pref2, nv)) || (rc = dalloc(&d->red, 2 * (size_t)nv * nblocks + 16)) ||
      (rc = dalloc(&d->dtg, 1))) {
    csu_dev_destroy(d);
    return rc;
  }
  if ((rc = h2d(d->cnp, cnp.data(), nc)) || (rc = h2d(d->cnbr, cnbr.data(), cnbr.size())) ||
      (rc = h2d(d->cface, cface.data(), cface.size())) || (rc = h2d(d->xc, xc.data(), nc)) ||
      (rc = h2d(d->yc, yc.data(), nc)) || (rc = h2d(d->omega, om.data(), nc)) || (rc = h2d(d->dsxm, dsxm.data(), nc)) ||
      (rc = h2d(d->dsym, dsym.data(), nc)) || (rc = h2d(d->edx, edx.data(), edx.size())) ||
      (rc = h2d(d->edy, edy.data(), edy.size())) || (rc = h2d(d->fe0, fe0.data(), nf)) ||
      (rc = h2d(d->fe1, fe1.data(), nf)) || (rc = h2d(d->fdsx, fdsx.data(), nf)) || (rc = h2d(d->fdsy, fdsy.data(), nf)) ||
      (rc = h2d(d->fds, fds.data(), nf)) || (rc = h2d(d->fdx0, fdx0.data(), nf)) || (rc = h2d(d->fdy0, fdy0.data(), nf)) ||
      (rc = h2d(d->fdx1, fdx1.data(), nf)) || (rc = h2d(d->fdy1, fdy1.data(), nf)) || (rc = h2d(d->be0, be0.data(), nb)) ||
      (rc = h2d(d->btype, btype.data(), nb)) || (rc = h2d(d->bdsx, bdsx.data(), nb)) ||
      (rc = h2d(d->bdsy, bdsy.data(), nb)) || (rc = h2d(d->bds, bds.data(), nb)) || (rc = h2d(d->bxm, bxm.data(), nb)) ||
      (rc = h2d(d->bym, bym.data(), nb)) || (rc = h2d(d->wall, wall.data(), wall.size())) ||
      (rc = h2d(d->pref2, pref.data(), nv))) {
    csu_dev_destroy(d);
    return rc;
  }
  // free-stream initial state (solverInitU)
  std::vector<double> U0((size_t)nv * nc);
  for (int k = 0; k < nv; k++)
    for (int i = 0; i < nc; i++) U0[(size_t)k * nc + i] = p->uin[k];
  if ((rc = h2d(d->U, U0.data(), U0.size()))) { csu_dev_destroy(d); return rc; }
  // Pb of faces not written before the first forces call stays defined
  HIPCHK(hipMemset(d->Pb, 0, sizeof(double) * (size_t)(nv + 1) * (nb ? nb : 1)));
  HIPCHK(hipDeviceSynchronize());
  *out = d;
  return CSU_OK;
}

extern "C" int csu_dev_set_state(csu_dev* d, const double* U) {
  if (!d || !U) return set_err(CSU_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(d->U, U, sizeof(double) * (size_t)d->nv * d->nc, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  d->fresh = 1;
  return CSU_OK;
}

extern "C" int csu_dev_get_state(csu_dev* d, double* U) {
  if (!d || !U) return set_err(CSU_ERR_ARG, "null argument");
  HIPCHK(hipMemcpyAsync(U, d->U, sizeof(double) * (size_t)d->nv * d->nc, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

extern "C" int csu_dev_get_dt(csu_dev* d, double* dtl) {
  if (!d || !dtl) return set_err(CSU_ERR_ARG, "null argument");
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

// ---- profiling helpers ----
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

extern "C" int csu_dev_set_profiling(csu_dev* d, int32_t on) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  d->prof = on ? 1 : 0;
  for (size_t i = 0; i < d->ktime.size(); i++) { d->ktime[i] = 0; d->kcount[i] = 0; }
  return CSU_OK;
}

extern "C" int csu_dev_kernel_times(csu_dev* d, const char** names, double* avg_ms, int32_t* launches, int32_t max) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
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

// ---- one residual evaluation (solverCalcR) with P already valid ----
// Launches gradient/limiter, face flux and the (fused) cell update.
template <int NV>
static int stage(csu_dev* d, int flags, double sigma, double rk_a, double* Un, double* Uout) {
  const KP& kp = d->kp;
  const int nc = d->nc;
  if (kp.order == 2 && kp.limiter == 1) {
    Prof p(d, "k_lim1");
    hipLaunchKernelGGL(k_lim1_partial<NV>, dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, kp, d->P, d->red);
    hipLaunchKernelGGL(k_lim1_final<NV>, dim3(1), dim3(CSU_BLOCK), 0, d->stream, kp, nblk(nc), d->red, d->pref2);
  }
  {
    Prof p(d, "k_grad");
    hipLaunchKernelGGL(k_grad<NV>, dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, kp, d->cnp, d->cnbr, d->xc, d->yc,
                       d->omega, d->edx, d->edy, d->bdsx, d->bdsy, d->bds, d->bxm, d->bym, d->btype, d->P, d->Pb,
                       d->pref2, d->dPx, d->dPy, d->phi);
  }
  if (d->nf > 0) {
    Prof p(d, "k_flux");
    hipLaunchKernelGGL(k_flux<NV>, dim3(nblk(d->nf)), dim3(CSU_BLOCK), 0, d->stream, kp, d->fe0, d->fe1, d->fdsx,
                       d->fdsy, d->fds, d->fdx0, d->fdy0, d->fdx1, d->fdy1, d->P, d->dPx, d->dPy, d->phi, d->fflux);
  }
  {
    Prof p(d, "k_update");
    hipLaunchKernelGGL(k_update<NV>, dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, kp, flags, sigma, rk_a,
                       d->prm.cfl, d->cnp, d->cnbr, d->cface, d->bdsx, d->bdsy, d->bds, d->btype, d->P, d->fflux,
                       d->omega, d->dsxm, d->dsym, d->dtl, d->dtg, Un, Uout, d->R, d->red, d->red + (size_t)NV * d->nred);
  }
  HIPCHK(hipGetLastError());
  return CSU_OK;
}

template <int NV>
static int prim_pass(csu_dev* d) {
  Prof p(d, "k_prim");
  hipLaunchKernelGGL(k_prim<NV>, dim3(nblk(d->nc)), dim3(CSU_BLOCK), 0, d->stream, d->kp, d->U, d->P);
  HIPCHK(hipGetLastError());
  return CSU_OK;
}

template <int NV>
static int residual_t(csu_dev* d) {
  int rc = prim_pass<NV>(d);
  if (!rc) rc = stage<NV>(d, F_STORE_R, 0.0, 0.0, nullptr, nullptr);
  d->fresh = 1;
  return rc;
}

extern "C" int csu_dev_residual(csu_dev* d, double* R, double* dPx, double* dPy, double* phi, double* P,
                                double* Pb) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  int rc;
  switch (d->nv) {
    case 4: rc = residual_t<4>(d); break;
    case 5: rc = residual_t<5>(d); break;
    default: return set_err(CSU_ERR_UNSUPPORTED, "nvar %d", d->nv);
  }
  if (rc) return rc;
  const size_t nvc = (size_t)d->nv * d->nc;
  if (R) HIPCHK(hipMemcpyAsync(R, d->R, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (dPx) HIPCHK(hipMemcpyAsync(dPx, d->dPx, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (dPy) HIPCHK(hipMemcpyAsync(dPy, d->dPy, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  if (phi) HIPCHK(hipMemcpyAsync(phi, d->phi, nvc * sizeof(double), hipMemcpyDeviceToHost, d->stream));
  std::vector<double> tmp;
  if (P) {
    // device P is SoA [nv+1][nc]; the ABI returns [nc][nv+1]
    const int ns = d->nv + 1;
    tmp.resize((size_t)ns * d->nc);
    HIPCHK(hipMemcpyAsync(tmp.data(), d->P, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (int i = 0; i < d->nc; i++)
      for (int m = 0; m < ns; m++) P[(size_t)i * ns + m] = tmp[(size_t)m * d->nc + i];
  }
  if (Pb && d->nb) {
    const int ns = d->nv + 1;
    tmp.resize((size_t)ns * d->nb);
    HIPCHK(hipMemcpyAsync(tmp.data(), d->Pb, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (int g = 0; g < d->nb; g++)
      for (int m = 0; m < ns; m++) Pb[(size_t)g * ns + m] = tmp[(size_t)m * d->nb + g];
  }
  HIPCHK(hipStreamSynchronize(d->stream));
  return CSU_OK;
}

// RK coefficients, reference solverStepRK (solver.c:788-840)
static const double RK1[] = {1.0};
static const double RK3[] = {0.1481, 0.4, 1.0};
static const double RK4[] = {0.0833, 0.2069, 0.4265, 1.0};
static const double RK5[] = {0.0533, 0.1263, 0.2375, 0.4414, 1.0};

template <int NV>
static int iterate_t(csu_dev* d, int niter) {
  const KP& kp = d->kp;
  const int nc = d->nc;
  const double* a = d->prm.stages == 1 ? RK1 : d->prm.stages == 3 ? RK3 : d->prm.stages == 4 ? RK4 : RK5;
  const double sigma = 0.5 * d->prm.stages;
  double* red_dt = d->red + (size_t)NV * d->nred;
  auto is_local = [d](int it) {
    return d->prm.dt_local == 1 && !(d->prm.dt_local_n >= 0 && it > d->prm.dt_local_n);
  };
  for (int it = 0; it < niter; it++) {
    const int cur_local = is_local(d->iter), next_local = is_local(d->iter + 1);
    int first_flags = F_CLAMP_UN;
    if (d->fresh) {
      // time step from the unclamped state (solverCalcDt), then primitives
      Prof p(d, "k_dt");
      hipLaunchKernelGGL(k_dt<NV>, dim3(nblk(nc)), dim3(CSU_BLOCK), 0, d->stream, kp, cur_local, d->prm.cfl, d->U,
                         d->omega, d->dsxm, d->dsym, d->dtl, red_dt);
      if (!cur_local)
        hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(CSU_BLOCK), 0, d->stream, nblk(nc), d->prm.cfl, red_dt, d->dtg);
      int rc = prim_pass<NV>(d);
      if (rc) return rc;
      first_flags = 0;
      d->fresh = 0;
    }
    for (int s = 0; s < d->prm.stages; s++) {
      const bool last = s == d->prm.stages - 1;
      int flags = F_RK | (cur_local ? F_CUR_LOCAL : 0) | (s == 0 ? first_flags : 0);
      if (last) flags |= F_LAST | F_STORE_R | (next_local ? F_NEXT_LOCAL : 0);
      int rc = stage<NV>(d, flags, sigma, a[s], d->U, d->Uaux);
      if (rc) return rc;
    }
    std::swap(d->U, d->Uaux);
    if (!next_local)
      hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(CSU_BLOCK), 0, d->stream, nblk(nc), d->prm.cfl, red_dt, d->dtg);
    const int slot = d->iter % d->hist_cap;
    {
      Prof p(d, "k_res");
      hipLaunchKernelGGL(k_res_final<NV>, dim3(1), dim3(RES_THREADS), 0, d->stream, nc, nblk(nc), d->red,
                         d->res_hist + (size_t)slot * NV);
    }
    {
      Prof p(d, "k_forces");
      hipLaunchKernelGGL(k_forces, dim3(1), dim3(CSU_BLOCK), 0, d->stream, kp, d->nwall, d->wall, d->Pb, d->bdsx,
                         d->bdsy, d->prm.sref, 0, d->coef_hist + (size_t)slot * 4);
    }
    HIPCHK(hipGetLastError());
    d->iter++;
  }
  return CSU_OK;
}

extern "C" int csu_dev_iterate(csu_dev* d, int32_t niter, double* res, double* coef) {
  if (!d) return set_err(CSU_ERR_ARG, "null argument");
  if (niter < 0) return set_err(CSU_ERR_ARG, "niter %d", niter);
  if (niter == 0) return CSU_OK;
  if (d->hist_cap < niter) {
    if (d->res_hist) (void)hipFree(d->res_hist);
    if (d->coef_hist) (void)hipFree(d->coef_hist);
    d->res_hist = d->coef_hist = nullptr;
    int rc;
    if ((rc = dalloc(&d->res_hist, (size_t)niter * d->nv)) || (rc = dalloc(&d->coef_hist, (size_t)niter * 4)))
      return rc;
    d->hist_cap = niter;
  }
  const int start = d->iter;
  int rc;
  switch (d->nv) {
    case 4: rc = iterate_t<4>(d, niter); break;
    case 5: rc = iterate_t<5>(d, niter); break;
    default: return set_err(CSU_ERR_UNSUPPORTED, "nvar %d", d->nv);
  }
  if (rc) return rc;
  if (res || coef) {
    std::vector<double> rh((size_t)d->hist_cap * d->nv), ch((size_t)d->hist_cap * 4);
    HIPCHK(hipMemcpyAsync(rh.data(), d->res_hist, rh.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipMemcpyAsync(ch.data(), d->coef_hist, ch.size() * sizeof(double), hipMemcpyDeviceToHost, d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
    for (int it = 0; it < niter; it++) {
      int slot = (start + it) % d->hist_cap;
      if (res) memcpy(res + (size_t)it * d->nv, rh.data() + (size_t)slot * d->nv, sizeof(double) * d->nv);
      if (coef) memcpy(coef + (size_t)it * 4, ch.data() + (size_t)slot * 4, sizeof(double) * 4);
    }
  }
  return prof_collect(d);
}
