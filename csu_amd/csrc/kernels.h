// kernels.h -- the residual-and-update kernels (one mesh, SoA FP64 state).
//
// Per explicit stage (reference solverCalcR + solverRK, solver.c:685-766):
//   k_grad    cells  ghost states of the cell's boundary faces, least-squares
//                    gradients (precomputed weights, or double-double sums),
//                    Venkatakrishnan phi, the reconstructed state at every
//                    edge; viscous: T gradient, boundary viscous fluxes, wall
//                    friction (all need only the cell's own fresh data)
//   k_flux    faces  Riemann flux of the two edge states (+ viscous face flux)
//   k_update  cells  per-cell sums in the reference's order, then the RK
//                    update fused with the next stage's primitive recovery;
//                    on the last stage also the next time step and the
//                    residual-norm block partials
// Every cell gathers its own faces: no atomics, fixed summation order.
#pragma once
#include <type_traits>
#include "kern_common.h"
#include "kern_visc.h"

namespace csu {

enum { F_RK = 1, F_CLAMP_UN = 2, F_LAST = 4, F_NEXT_LOCAL = 8, F_STORE_R = 16, F_CUR_LOCAL = 32, F_RES = 64 };

// Standalone primitive pass (first stage after the state was set from the
// host).  Clamped values are written back to U, as the reference does.
// Batch: the cells of a held solve keep their state (a batch's padding
// cells hold their solve's free stream and are converted like the rest).
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_prim(KP kp, double* __restrict__ U, double* __restrict__ P,
                                                    const SolveP* __restrict__ sol, const int* __restrict__ csolk) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc;
  if (i >= nc) return;
  if (sol && !sol[solve_of(csolk, blockIdx.x * blockDim.x)].active) return;
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

// LSQ gradient of one variable from its differences to the neighbours, with
// the normal-equation sums in double-double (the reference uses x87 long
// double accumulators, solver.c:1041-1073).
template <int NP>
__device__ __forceinline__ void lsq_grad(int np, const double* dx, const double* dy, const double* du,
                                         const dd& a, const dd& b, const dd& nb, const dd& c, const dd& det,
                                         double& gx, double& gy) {
  dd A = {0, 0}, B = {0, 0};
#pragma unroll
  for (int j = 0; j < NP; j++) {
    if (j >= np) break;
    A = dd_add_d(A, dx[j] * du[j]);
    B = dd_add_d(B, dy[j] * du[j]);
  }
  gx = dd_div(dd_sub(dd_mul(c, A), dd_mul(b, B)), det).hi;
  gy = dd_div(dd_add(dd_mul(nb, A), dd_mul(a, B)), det).hi;
}

// Gradient + limiter (+ boundary work) kernel.  Reference solverUpdateGrad /
// solverCalcGrad3 (solver.c:451-474, 1031-1074), solverGrad_T
// (solver.c:477-490), the limiter loop of inter() (solver.c:496-548,
// limiter.c:98-147), boundaryCalcPrimitive (boundary.c:514-693),
// boundaryFaceViscFlux (laminar.c:128-346), boundaryCalcFrictionWall
// (boundary.c:695-742).
// MIX (mixed precision): gradients, limiter and reconstruction in FP32 from
// FP64 differences; edge states stored in FP32.
// PU (inviscid second-order RK runs): the primitives of the cell and its
// neighbours are recovered here from the stage's conservative state s.Us,
// which the update kernel then writes unclamped, instead of being stored
// by the update kernel and read back: the same function of the same values
// (solverCalcPrimitive on the unclamped stage state), 40 bytes per cell
// less written and 8 less read per stage.
template <int NV, bool VISC, bool LD, bool MIX = false, bool PU = false>
__global__ __launch_bounds__(CSU_BLOCK) void k_grad(KP kp, Geo g, St s) {
  using TF = typename std::conditional<MIX, float, double>::type;
  constexpr int NS = NV + 1;
  const int e0 = (kp.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * blockDim.x;
  const int i = e0 + threadIdx.x;
  const int nc = kp.nc, nb = kp.nb;
  if (g.sol) {
    const int sid = solve_of(g.csolk, e0);
    solve_kp(kp, g.sol[sid]);
    s.pref2 += 6 * sid;
  }
  if (i >= kp.cend) return;
  const int np = g.cnp[i];
  const double x0 = g.xc[i], y0 = g.yc[i];
  double Pi[NS];
  if (PU) {
    double u[4] = {s.Us[i], s.Us[nc + i], s.Us[2 * nc + i], s.Us[3 * nc + i]};
    prim_point<4>(kp, u, Pi);
  } else {
#pragma unroll
    for (int m = 0; m < NS; m++) Pi[m] = s.P[m * nc + i];
  }
  // neighbours: cells, or the ghost states of boundary faces (k_bcflux)
  double dx[4], dy[4];
  int gb[4], nbr[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    gb[j] = -1;
    nbr[j] = 0;
    if (j >= np) break;
    int n = g.cnbr[j * nc + i];
    nbr[j] = n;
    if (n >= 0) {
      dx[j] = g.xc[n] - x0;
      dy[j] = g.yc[n] - y0;
    } else {
      int q = -1 - n;
      gb[j] = q;
      dx[j] = g.bxm[q] - x0;
      dy[j] = g.bym[q] - y0;
    }
  }
  // Every neighbour value the cell reads, loaded before any store: the
  // compiler cannot tell the stores below from these loads, so loading
  // them inside the variable loop would serialise one memory round trip
  // per variable.  Slots past np read the cell itself and are not used.
  double vnb[4][NV], vT[4];
  if (PU) {
    // all neighbour states first, then their primitives; ghost states as
    // stored; slots past np take the cell's own values
    double un[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool act = j < np;
      const int n = act ? nbr[j] : i;
      const double* base = n >= 0 ? s.Us + n : s.Pb + (-1 - n);
      const int st = n >= 0 ? nc : nb;
#pragma unroll
      for (int kk = 0; kk < 4; kk++) un[j][kk] = base[kk * st];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool act = j < np;
      const int n = act ? nbr[j] : i;
      if (act && n >= 0) {
        double pp[5];
        prim_point<4>(kp, un[j], pp);
#pragma unroll
        for (int kk = 0; kk < NV; kk++) vnb[j][kk] = pp[kk];
      } else {
#pragma unroll
        for (int kk = 0; kk < NV; kk++) vnb[j][kk] = act ? un[j][kk] : Pi[kk];
      }
      vT[j] = 0.0;
    }
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const bool act = j < np;
      const int n = act ? nbr[j] : i;
      const double* base = n >= 0 ? s.P + n : s.Pb + (-1 - n);
      const int st = n >= 0 ? nc : nb;
#pragma unroll
      for (int kk = 0; kk < NV; kk++) {
        const int m = kk > 3 ? kk + 1 : kk;
        vnb[j][kk] = base[m * st];
      }
      vT[j] = VISC ? base[4 * st] : 0.0;
    }
  }
  // every cell has slots 0-2; slot 3 (quads only) is read only by quads,
  // a branch that is uniform on triangle and quad meshes alike
  TF ex[4] = {0, 0, 0, 0}, ey[4] = {0, 0, 0, 0};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if (j == 3 && np < 4) break;
    if (MIX) {
      ex[j] = g.edx32[j * nc + i];
      ey[j] = g.edy32[j * nc + i];
    } else {
      // read once per stage: non-temporal, so they do not displace the
      // neighbour states in the caches
      ex[j] = __builtin_nontemporal_load(&g.edx[j * nc + i]);
      ey[j] = __builtin_nontemporal_load(&g.edy[j * nc + i]);
    }
  }
  // Least squares (solver.c:1031-1074).  Default: gx = sum_j wx_j du_j with
  // the weights (c dx_j - b dy_j)/det, (a dy_j - b dx_j)/det precomputed per
  // cell in long double.  LD: the normal-equation sums of every call in
  // double-double, as the reference's x87 long double accumulators do.
  dd a = {0, 0}, b = {0, 0}, c = {0, 0}, det = {0, 0}, nb_ = {0, 0};
  TF wx[4], wy[4];
  if (LD) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (j >= np) break;
      a = dd_add_d(a, dx[j] * dx[j]);
      b = dd_add_d(b, dx[j] * dy[j]);
      c = dd_add_d(c, dy[j] * dy[j]);
    }
    det = dd_sub(dd_mul(a, c), dd_mul(b, b));
    nb_ = dd_neg(b);
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      // slot 3 only for quads (past np the weight is 0)
      wx[j] = 0;
      wy[j] = 0;
      if (j == 3 && np < 4) break;
      if (MIX) {
        wx[j] = g.lwx32[j * nc + i];
        wy[j] = g.lwy32[j * nc + i];
      } else {
        wx[j] = __builtin_nontemporal_load(&g.lwx[j * nc + i]);
        wy[j] = __builtin_nontemporal_load(&g.lwy[j * nc + i]);
      }
    }
  }
  auto grad = [&](const double* du, double& gx, double& gy) {
    if (LD) {
      lsq_grad<4>(np, dx, dy, du, a, b, nb_, c, det, gx, gy);
    } else {
      TF ax = wx[0] * (TF)du[0], ay = wy[0] * (TF)du[0];
#pragma unroll
      for (int j = 1; j < 4; j++) {
        if (j >= np) break;
        ax += wx[j] * (TF)du[j];
        ay += wy[j] * (TF)du[j];
      }
      gx = ax;
      gy = ay;
    }
  };
  const bool store = VISC || kp.store_grad;
  double lim_aux = 0.0;
  if (kp.limiter == 0) {
    lim_aux = kp.limK * sqrt(g.omega[i] / kp.volmax);
    lim_aux = lim_aux * lim_aux * lim_aux;
  }
  double gux = 0, guy = 0, gvx = 0, gvy = 0, grx = 0, gry = 0, gnx = 0, gny = 0, gox = 0, goy = 0;
#pragma unroll
  for (int kk = 0; kk < NV; kk++) {
    const int m = kk > 3 ? kk + 1 : kk;
    const double u0 = Pi[m];
    double vn[4], du[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      vn[j] = j < np ? vnb[j][kk] : u0;
      du[j] = vn[j] - u0;
    }
    double gx, gy;
    grad(du, gx, gy);
    if (store) {
      s.dPx[kk * nc + i] = gx;
      s.dPy[kk * nc + i] = gy;
    }
    if (kk == 0) { grx = gx; gry = gy; }
    if (kk == 1) { gux = gx; guy = gy; }
    if (kk == 2) { gvx = gx; gvy = gy; }
    if (kk == 4) { gnx = gx; gny = gy; }  // SA nu~ or SST k
    if (kk == 5) { gox = gx; goy = gy; }  // SST omega
    if (kp.order == 2) {
      double pref2 = kp.limiter == 0 ? s.pref2[kk] * lim_aux : s.pref2[kk];
      double Pmin = u0, Pmax = u0;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j >= np) break;
        Pmax = fmax(Pmax, vn[j]);
        Pmin = fmin(Pmin, vn[j]);
      }
      TF ph = 0, d2[4];
      const TF d1max = (TF)(Pmax - u0), d1min = (TF)(Pmin - u0), tgx = (TF)gx, tgy = (TF)gy;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j >= np) break;
        d2[j] = tgx * ex[j] + tgy * ey[j];
      }
      // one limiter per launch (a uniform branch around the whole edge
      // loop: a per-edge select would evaluate both functions' divisions)
      if (kp.limiter == 2) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (j >= np) break;
          const TF p0 = limiter_bj<TF>(d1max, d1min, d2[j]);
          ph = (j == 0) ? p0 : fmin(ph, p0);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          if (j >= np) break;
          const TF p0 = limiter_v2d<TF>(d1max, d1min, d2[j], (TF)pref2);
          ph = (j == 0) ? p0 : fmin(ph, p0);
        }
      }
      if (store) s.phi[kk * nc + i] = ph;
      // the state each face of this cell sees on this side (solver.c:598-603)
      const bool first = kk > 3 && kp.turb1order;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        if (j >= np) break;
        if (MIX) s.rec32[(j * NV + kk) * nc + i] = (float)(first ? u0 : u0 + (double)(d2[j] * ph));
        else __builtin_nontemporal_store(first ? u0 : u0 + d2[j] * ph, &s.rec[(j * NV + kk) * nc + i]);
      }
    }
  }
  if (VISC) {
    double du[4];
#pragma unroll
    for (int j = 0; j < 4; j++) du[j] = j < np ? vT[j] - Pi[4] : 0.0;
    double gTx, gTy;
    grad(du, gTx, gTy);
    s.dTx[i] = gTx;
    s.dTy[i] = gTy;
    const double gx3[5] = {gux, gvx, gTx, gnx, gox}, gy3[5] = {guy, gvy, gTy, gny, goy};
    if (NV == 6) {
      double src[2], dq[6];
      sst_cell_source(kp, Pi, g.wdist[i], g.omega[i], gx3, gy3, src, dq);
      s.src[i] = src[0];
      s.src[nc + i] = src[1];
#pragma unroll
      for (int q = 0; q < 6; q++) s.dq[q * nc + i] = dq[q];
    }
    if (NV == 5) {
      // saInterSource (sa.c:224-260): uses grad rho, du/dy, dv/dx, grad nu~
      const double mi_L = sutherland(Pi[4]);
      const double S = fabs(guy - gvx);
      const double Qt = sa_source(Pi[5], mi_L / Pi[0], S, g.wdist[i], Pi[0], grx, gry, gnx, gny, kp.sa_cc != 0);
      s.src[i] = Qt * g.omega[i];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int q = gb[j];
      if (q < 0) continue;
      const int t = g.btype[q];
      double f[5];
      if (NV == 6) {
        sst_bc_face(kp, t, g.bdsx[q], g.bdsy[q], g.bnx[q], g.bny[q], g.bds2[q], dx[j], dy[j], Pi, g.wdist[i], gx3,
                    gy3, f);
        s.bvflux[3 * nb + q] = f[3];
        s.bvflux[4 * nb + q] = f[4];
      } else if (NV == 5) {
        sa_bc_face(kp, t, g.bdsx[q], g.bdsy[q], g.bnx[q], g.bny[q], g.bds2[q], dx[j], dy[j], Pi, gx3, gy3, grx, gry,
                   f);
        s.bvflux[3 * nb + q] = f[3];
      } else {
        laminar_bc_face(kp, t, g.bdsx[q], g.bdsy[q], g.bnx[q], g.bny[q], g.bds2[q], dx[j], dy[j], Pi, gx3, gy3, f);
      }
      s.bvflux[0 * nb + q] = f[0];
      s.bvflux[1 * nb + q] = f[1];
      s.bvflux[2 * nb + q] = f[2];
      if (t == CSU_BC_WALL || t == CSU_BC_WALLT) {
        double fx, fy;
        wall_friction(g.bdsx[q], g.bdsy[q], dx[j], dy[j], Pi, gx3, gy3, fx, fy);
        s.wfric[q] = fx;
        s.wfric[nb + q] = fy;
        double wt[4];
        wall_tensor(NV > 4, g.bdsx[q], g.bdsy[q], dx[j], dy[j], g.bxm[q], Pi, NV > 4 ? g.wdist[i] : 0.0, gx3, gy3,
                    wt);
#pragma unroll
        for (int k = 0; k < 4; k++) s.wten[k * nb + q] = wt[k];
      }
    }
  }
}

// Limiter type 1 global range (limiter.c:64-94): per-block min/max partials.
// Lanes past the solve's last cell take its first cell's values.
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_lim1_partial(KP kp, Geo g, const double* __restrict__ P,
                                                            double* __restrict__ red) {
  __shared__ double swn[NV][CSU_BLOCK / 64], swx[NV][CSU_BLOCK / 64];
  int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const int nc = kp.nc;
  int c0 = 0;
  if (g.sol) {
    const SolveP& q = g.sol[solve_of(g.csolk, blockIdx.x * blockDim.x)];
    c0 = q.c0;
    kp.cend = q.c1;
  }
  int ii = i < kp.cend ? i : c0;
  double mn[NV], mx[NV];
#pragma unroll
  for (int jj = 0; jj < NV; jj++) {
    int mm = jj == 4 ? 5 : jj;  // limiter.c:350-357 (slot 5 also for jj == 5)
    mn[jj] = mx[jj] = P[mm * nc + ii];
  }
  block_reduce<CSU_BLOCK, NV>(mn, swn, RedMin());
  block_reduce<CSU_BLOCK, NV>(mx, swx, RedMax());
  if (t == 0)
    for (int jj = 0; jj < NV; jj++) {
      red[(2 * jj) * gridDim.x + blockIdx.x] = mn[jj];
      red[(2 * jj + 1) * gridDim.x + blockIdx.x] = mx[jj];
    }
}

// One workgroup per solve: the partials of its blocks (red rows of length
// stride), its reference values into pref2 [solve][6].
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_lim1_final(KP kp, int nblocks, int stride, const SolveP* __restrict__ sol,
                                                          const double* __restrict__ red, double* __restrict__ pref2) {
  __shared__ double swn[NV][CSU_BLOCK / 64], swx[NV][CSU_BLOCK / 64];
  int t = threadIdx.x;
  if (sol) {
    const SolveP& q = sol[blockIdx.x];
    red += q.c0 / CSU_BLOCK;
    nblocks = (q.c1 - q.c0 + CSU_BLOCK - 1) / CSU_BLOCK;
    pref2 += 6 * blockIdx.x;
  }
  double mn[NV], mx[NV];
#pragma unroll
  for (int jj = 0; jj < NV; jj++) {
    mn[jj] = red[(2 * jj) * stride];
    mx[jj] = red[(2 * jj + 1) * stride];
    for (int b = t; b < nblocks; b += CSU_BLOCK) {
      mn[jj] = fmin(mn[jj], red[(2 * jj) * stride + b]);
      mx[jj] = fmax(mx[jj], red[(2 * jj + 1) * stride + b]);
    }
  }
  block_reduce<CSU_BLOCK, NV>(mn, swn, RedMin());
  block_reduce<CSU_BLOCK, NV>(mx, swx, RedMax());
  if (t == 0)
    for (int jj = 0; jj < NV; jj++) {
      double aux = (mx[jj] - mn[jj]) * kp.limK;
      pref2[jj] = aux * aux;
    }
}

// Convective flux of interior face f times dS, reference inter() face loop
// (solver.c:550-644): the two edge states (order 2, reconstructed by k_grad
// at edge s0 of e0 and edge s1 of e1) or the cell states (order 1),
// rotated into the face frame, the Riemann flux, rotated back.  MIX: in
// FP32 from FP32 edge states, as FP32 values.
template <int NV, bool MIX>
__device__ __forceinline__ void conv_flux(const KP& kp, const Geo& g, const St& s, int f, int e0, int e1, int s0,
                                          int s1, typename std::conditional<MIX, float, double>::type* out) {
  using TF = typename std::conditional<MIX, float, double>::type;
  const int nc = kp.nc;
  TF dSx, dSy, dS;
  if (MIX) {
    dSx = g.fdsx32[f]; dSy = g.fdsy32[f]; dS = g.fds32[f];
  } else {
    dSx = g.fdsx[f]; dSy = g.fdsy[f]; dS = g.fds[f];
  }
  TF PL[NV], PR[NV];
  if (kp.order == 2) {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) {
      if (MIX) {
        PL[kk] = s.rec32[(s0 * NV + kk) * nc + e0];
        PR[kk] = s.rec32[(s1 * NV + kk) * nc + e1];
      } else {
        PL[kk] = s.rec[(s0 * NV + kk) * nc + e0];
        PR[kk] = s.rec[(s1 * NV + kk) * nc + e1];
      }
    }
  } else {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) {
      const int m = kk > 3 ? kk + 1 : kk;
      PL[kk] = (TF)s.P[m * nc + e0];
      PR[kk] = (TF)s.P[m * nc + e1];
    }
  }
  rotate<TF>(PL[1], PL[2], dSx, dSy, dS);
  rotate<TF>(PR[1], PR[2], dSx, dSy, dS);
  TF fl[NV];
  constexpr int NSC = NV - 4;
  flux_dispatch<TF, NSC>(kp.gas, kp.flux, (TF)kp.efix, (TF)kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4, PR[0],
                         PR[1], PR[2], PR[3], PR + 4, NV == 6 ? PL[4] : (TF)0, NV == 6 ? PR[4] : (TF)0, fl);
  rotate<TF>(fl[1], fl[2], dSx, -dSy, dS);
#pragma unroll
  for (int kk = 0; kk < NV; kk++) out[kk] = fl[kk] * dS;
}

// Interior face flux, reference inter() face loop (solver.c:550-644), and
// for viscous runs the laminar face flux (laminar.c:17-75).
// MIX: the convective flux in FP32 from FP32 edge states, stored in FP32.
template <int NV, bool VISC, bool MIX = false>
__global__ __launch_bounds__(CSU_BLOCK) void k_flux(KP kp, Geo g, St s) {
  const int e0f = (kp.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x) * blockDim.x;
  const int f = e0f + threadIdx.x;
  const int nc = kp.nc, nf = kp.nf;
  if (g.sol) solve_kp(kp, g.sol[solve_of(g.fsolk, e0f)]);
  if (f >= kp.fend) return;
  const int e0 = g.fe0[f], e1 = g.fe1[f];
  {
    const int sl = g.fsl[f];
    if (MIX) {
      float fl[NV];
      conv_flux<NV, true>(kp, g, s, f, e0, e1, sl & 3, sl >> 2, fl);
#pragma unroll
      for (int kk = 0; kk < NV; kk++) s.fflux32[(size_t)f * NV + kk] = fl[kk];
    } else {
      double fl[NV];
      conv_flux<NV, false>(kp, g, s, f, e0, e1, sl & 3, sl >> 2, fl);
#pragma unroll
      for (int kk = 0; kk < NV; kk++) {
        // a face's 32 bytes (NV = 4) stream out whole; the 40- and 48-byte
        // records of the turbulent forms straddle lines and are stored plainly
        if (NV == 4) __builtin_nontemporal_store(fl[kk], &s.fflux[(size_t)f * NV + kk]);
        else s.fflux[(size_t)f * NV + kk] = fl[kk];
      }
    }
  }
  const double dSx = g.fdsx[f], dSy = g.fdsy[f];
  if (VISC && NV == 4) {
    double P0[5], P1[5];
#pragma unroll
    for (int m = 0; m < 5; m++) { P0[m] = s.P[m * nc + e0]; P1[m] = s.P[m * nc + e1]; }
    const double g0x[3] = {s.dPx[nc + e0], s.dPx[2 * nc + e0], s.dTx[e0]};
    const double g0y[3] = {s.dPy[nc + e0], s.dPy[2 * nc + e0], s.dTy[e0]};
    const double g1x[3] = {s.dPx[nc + e1], s.dPx[2 * nc + e1], s.dTx[e1]};
    const double g1y[3] = {s.dPy[nc + e1], s.dPy[2 * nc + e1], s.dTy[e1]};
    double F[3];
    laminar_face(kp, dSx, dSy, g.xc[e1] - g.xc[e0], g.yc[e1] - g.yc[e0], P0, P1, g0x, g0y, g1x, g1y, F);
    s.vflux[0 * nf + f] = F[0];
    s.vflux[1 * nf + f] = F[1];
    s.vflux[2 * nf + f] = F[2];
  }
  if (VISC && NV == 5) {
    double P0[6], P1[6];
#pragma unroll
    for (int m = 0; m < 6; m++) { P0[m] = s.P[m * nc + e0]; P1[m] = s.P[m * nc + e1]; }
    const bool cc = kp.sa_cc != 0;
    const double g0x[5] = {s.dPx[nc + e0], s.dPx[2 * nc + e0], s.dTx[e0], s.dPx[4 * nc + e0], cc ? s.dPx[e0] : 0.0};
    const double g0y[5] = {s.dPy[nc + e0], s.dPy[2 * nc + e0], s.dTy[e0], s.dPy[4 * nc + e0], cc ? s.dPy[e0] : 0.0};
    const double g1x[5] = {s.dPx[nc + e1], s.dPx[2 * nc + e1], s.dTx[e1], s.dPx[4 * nc + e1], cc ? s.dPx[e1] : 0.0};
    const double g1y[5] = {s.dPy[nc + e1], s.dPy[2 * nc + e1], s.dTy[e1], s.dPy[4 * nc + e1], cc ? s.dPy[e1] : 0.0};
    double F[4], mi_t;
    sa_face(kp, dSx, dSy, g.xc[e1] - g.xc[e0], g.yc[e1] - g.yc[e0], P0, P1, g0x, g0y, g1x, g1y, F, mi_t);
#pragma unroll
    for (int k = 0; k < 4; k++) s.vflux[k * nf + f] = F[k];
    s.miT[f] = mi_t;
  }
  if (VISC && NV == 6) {
    double P0[7], P1[7];
#pragma unroll
    for (int m = 0; m < 7; m++) { P0[m] = s.P[m * nc + e0]; P1[m] = s.P[m * nc + e1]; }
    const double g0x[5] = {s.dPx[nc + e0], s.dPx[2 * nc + e0], s.dTx[e0], s.dPx[4 * nc + e0], s.dPx[5 * nc + e0]};
    const double g0y[5] = {s.dPy[nc + e0], s.dPy[2 * nc + e0], s.dTy[e0], s.dPy[4 * nc + e0], s.dPy[5 * nc + e0]};
    const double g1x[5] = {s.dPx[nc + e1], s.dPx[2 * nc + e1], s.dTx[e1], s.dPx[4 * nc + e1], s.dPx[5 * nc + e1]};
    const double g1y[5] = {s.dPy[nc + e1], s.dPy[2 * nc + e1], s.dTy[e1], s.dPy[4 * nc + e1], s.dPy[5 * nc + e1]};
    double F[5], mi_t;
    sst_face(kp, dSx, dSy, g.xc[e1] - g.xc[e0], g.yc[e1] - g.yc[e0], P0, P1, g.wdist[e0], g.wdist[e1], g0x, g0y,
             g1x, g1y, F, mi_t);
#pragma unroll
    for (int k = 0; k < 5; k++) s.vflux[k * nf + f] = F[k];
    s.miT[f] = mi_t;
  }
}

// Per boundary face, from its cell's primitives: the ghost state
// (boundaryCalcPrimitive, boundary.c:514-693) into Pb, and the convective
// flux from the first-order state, boundaryCalc / _sa / _sst
// (boundary.c:143-446): f*dS, or 0 when dS = 0.
template <int NV, bool PU = false>
__global__ __launch_bounds__(CSU_BLOCK) void k_bcflux(KP kp, Geo g, St s) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc, nb = kp.nb;
  if (g.sol) solve_kp(kp, g.sol[solve_of(g.bsolk, blockIdx.x * blockDim.x)]);
  if (q >= kp.bend) return;
  const int i = g.be0[q];
  double Pi[NV + 1], R[NV], Pg[NV + 1];
  if (PU) {  // primitives of the stage state (see k_grad)
    double u[4] = {s.Us[i], s.Us[nc + i], s.Us[2 * nc + i], s.Us[3 * nc + i]};
    prim_point<4>(kp, u, Pi);
  } else {
#pragma unroll
    for (int m = 0; m < NV + 1; m++) Pi[m] = s.P[m * nc + i];
  }
  ghost_state<NV>(kp, g.btype[q], g.bdsx[q], g.bdsy[q], g.bds[q], Pi, NV == 6 ? g.wdist[i] : 0.0, Pg);
#pragma unroll
  for (int m = 0; m < NV + 1; m++) s.Pb[m * nb + q] = Pg[m];
#pragma unroll
  for (int kk = 0; kk < NV; kk++) R[kk] = 0.0;
  bc_flux<NV>(kp, g.btype[q], g.bdsx[q], g.bdsy[q], g.bds[q], Pi, R);
#pragma unroll
  for (int kk = 0; kk < NV; kk++) s.bflux[kk * nb + q] = R[kk];
}

// Residual sums + RK update.  Order of the per-cell sums follows
// solverCalcR: interior convective faces (solver.c:646-667), boundary
// convective fluxes (boundary.c:447-464), axisymmetric pressure
// (solver.c:670-683), interior viscous faces (laminar.c:77-98), boundary
// viscous fluxes (laminar.c:101-125).  Then solverRK (solver.c:741-766)
// fused with the next stage's solverCalcPrimitive, and on the last stage
// the next time step (solver.c:895-950) and solverCalcRes partials
// (solver.c:867-892).
// TILE > 0 (inviscid runs): no face-flux kernel.  The workgroup of a 256-cell
// tile first evaluates the convective flux of every interior face of its
// cells into LDS (conv_flux, as the face kernel does; a face between two
// tiles is evaluated by both), then each cell sums its faces' fluxes from
// LDS in its slot order: the same values in the same order, without the
// face-flux array's write and gather.  TILE is the LDS list's length (the
// smaller one where every tile of the mesh fits: more workgroups per CU).
template <int NV, bool VISC, bool MIX = false, bool PU = false, int TILE = 0>
__global__ __launch_bounds__(CSU_BLOCK) void k_update(KP kp, Geo g, St s, int flags, double sigma, double rk_a,
                                                      double cfl, double* dtl, const double* dtg, double* Un,
                                                      double* __restrict__ Uout, double* __restrict__ Rout,
                                                      double* __restrict__ red_res, double* __restrict__ red_dt) {
  __shared__ double swr[NV][CSU_BLOCK / 64], swd[1][CSU_BLOCK / 64];
  const int tile = kp.xcd ? xcd_tile(blockIdx.x, gridDim.x) : (int)blockIdx.x;
  const int t = threadIdx.x, i = tile * blockDim.x + t;
  const int nc = kp.nc, nf = kp.nf, nb = kp.nb;
  bool held = false;  // batch: a solve that is not active keeps its state
  if (g.sol) {
    const int sid = solve_of(g.csolk, tile * blockDim.x);
    const SolveP& q = g.sol[sid];
    solve_kp(kp, q);
    cfl = q.cfl;
    dtg += sid;
    held = !q.active;
  }
  const bool active = i < kp.cend;
  using TS = typename std::conditional<MIX, float, double>::type;
  __shared__ TS sfl[TILE > 0 ? TILE * NV : 1];
  if (TILE > 0) {
    static_assert(TILE == 0 || !VISC, "tile form: inviscid runs only");
    // the tile's faces (its cells' interior faces): g.tface[g.toff[tile] ..)
    const int f0 = g.toff[tile], nft = g.toff[tile + 1] - f0;
    for (int k = t; k < nft; k += blockDim.x) {
      const int f = g.tface[f0 + k];
      const int sl = g.fsl[f];
      TS fl[NV];
      conv_flux<NV, MIX>(kp, g, s, f, g.fe0[f], g.fe1[f], sl & 3, sl >> 2, fl);
#pragma unroll
      for (int kk = 0; kk < NV; kk++) sfl[k * NV + kk] = fl[kk];
    }
    __syncthreads();
  }
  double R[NV];
#pragma unroll
  for (int kk = 0; kk < NV; kk++) R[kk] = 0.0;
  double dt_next = INFINITY;
  if (active) {
    // all face indices, then all face fluxes, then the sums in slot order:
    // two memory round trips instead of one pair per slot (slots without
    // a face read face 0 and add nothing).  Tile form: the slot's face as
    // +-(its place in the tile's list + 1), with the face id's sign.
    int cf[4];
#pragma unroll
    for (int j = 0; j < 4; j++) cf[j] = TILE > 0 ? (int)g.tsl[j * nc + i] : g.cface[j * nc + i];
    double fv[4][NV];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int f = cf[j], q = f > 0 ? f - 1 : (f < 0 ? -f - 1 : 0);
#pragma unroll
      for (int kk = 0; kk < NV; kk++) {
        if (TILE > 0) fv[j][kk] = (double)sfl[q * NV + kk];
        else fv[j][kk] = MIX ? (double)s.fflux32[(size_t)q * NV + kk] : s.fflux[(size_t)q * NV + kk];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int f = cf[j];
#pragma unroll
      for (int kk = 0; kk < NV; kk++) R[kk] = f > 0 ? R[kk] + fv[j][kk] : (f < 0 ? R[kk] - fv[j][kk] : R[kk]);
    }
    // boundary fluxes (k_bcflux), in the cell's slot order = marker order
    int gbq[4];
    const int bm = g.bmask[i];  // most cells have no boundary face: skip their neighbour list
#pragma unroll
    for (int j = 0; j < 4; j++) {
      gbq[j] = -1;
      if (!((bm >> j) & 1)) continue;
      const int n = g.cnbr[j * nc + i];
      if (n < 0) {
        const int q = -1 - n;
        gbq[j] = q;
#pragma unroll
        for (int kk = 0; kk < NV; kk++) R[kk] += s.bflux[kk * nb + q];
      }
    }
    if (kp.axi == 1) {
      R[2] -= s.P[3 * nc + i] * g.dslat[i];
      if (NV == 6) R[2] -= (2. / 3.) * s.P[i] * s.P[5 * nc + i] * g.dslat[i];
    }
    if (VISC && NV == 4) {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int f = cf[j];
        if (f > 0) {
#pragma unroll
          for (int kk = 1; kk < 4; kk++) R[kk] -= s.vflux[(kk - 1) * nf + f - 1];
        } else if (f < 0) {
#pragma unroll
          for (int kk = 1; kk < 4; kk++) R[kk] += s.vflux[(kk - 1) * nf - f - 1];
        }
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int q = gbq[j];
        if (q < 0) continue;
        R[1] -= s.bvflux[q];
        R[2] -= s.bvflux[nb + q];
        if (g.btype[q] == CSU_BC_WALL) {
          // adiabatic wall: the reference subtracts the energy flux of the
          // boundary face processed before it (laminar.c:101-117, 230-269)
          const int dn = g.bdonor[q];
          R[3] -= dn >= 0 ? s.bvflux[2 * nb + dn] : 0.0;
        } else {
          R[3] -= s.bvflux[2 * nb + q];
        }
      }
    }
    if (VISC && NV == 5) {
      // saInterFace gather (sa.c:199-220), saInterSource (sa.c:224-260),
      // saBoundary (sa.c:351-368)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int f = cf[j];
        if (f > 0) {
#pragma unroll
          for (int kk = 1; kk < 5; kk++) R[kk] -= s.vflux[(kk - 1) * nf + f - 1];
        } else if (f < 0) {
#pragma unroll
          for (int kk = 1; kk < 5; kk++) R[kk] += s.vflux[(kk - 1) * nf - f - 1];
        }
      }
      R[4] -= s.src[i];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int q = gbq[j];
        if (q < 0) continue;
#pragma unroll
        for (int kk = 1; kk < 5; kk++) R[kk] -= s.bvflux[(kk - 1) * nb + q];
      }
    }
    if (VISC && NV == 6) {
      // sstInterFace gather (sst.c:168-192), sstInterSource (sst.c:638-639),
      // sstBoundaryFace (sst.c:549-567)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int f = cf[j];
        if (f > 0) {
#pragma unroll
          for (int kk = 1; kk < 6; kk++) R[kk] -= s.vflux[(kk - 1) * nf + f - 1];
        } else if (f < 0) {
#pragma unroll
          for (int kk = 1; kk < 6; kk++) R[kk] += s.vflux[(kk - 1) * nf - f - 1];
        }
      }
      R[4] -= s.src[i];
      R[5] -= s.src[nc + i];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int q = gbq[j];
        if (q < 0) continue;
#pragma unroll
        for (int kk = 1; kk < 6; kk++) R[kk] -= s.bvflux[(kk - 1) * nb + q];
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
      // a time-accurate run that has taken its last step holds U bitwise
      const bool hold = (s.tube && s.tube[2] > 1.5) || held;
      if ((flags & F_CLAMP_UN) && !hold) {
        // first stage: the reference clamps U^n in place (solverCalcR(U))
        prim_point<NV>(kp, un, p);
#pragma unroll
        for (int kk = 0; kk < NV; kk++)
          if (__double_as_longlong(un[kk]) != __double_as_longlong(un0[kk])) Un[kk * nc + i] = un[kk];
      }
      const double dt = (flags & F_CUR_LOCAL) ? dtl[i] : dtg[0];
      const double om = g.omega[i];
      double unew[NV];
#pragma unroll
      for (int kk = 0; kk < NV; kk++) unew[kk] = hold ? un0[kk] : un[kk] - dt * sigma * rk_a * R[kk] / om;
      if (flags & F_LAST) {
        // U^{n+1} stays unclamped until the next residual, as in the reference
        dt_next = cell_dt<NV>(kp, unew, om, g.dsxm[i], g.dsym[i]);
        if (flags & F_NEXT_LOCAL) dtl[i] = cfl * dt_next;
#pragma unroll
        for (int kk = 0; kk < NV; kk++) Uout[kk * nc + i] = unew[kk];
        if (!PU) prim_point<NV>(kp, unew, p);
      } else if (PU) {
        // the next stage recovers its primitives from this state; its clamps
        // only ever reach the primitives (the reference's RK overwrites the
        // clamped intermediate state), so it is stored as computed
#pragma unroll
        for (int kk = 0; kk < NV; kk++) Uout[kk * nc + i] = unew[kk];
      } else {
        prim_point<NV>(kp, unew, p);
#pragma unroll
        for (int kk = 0; kk < NV; kk++) Uout[kk * nc + i] = unew[kk];
      }
      if (!PU) {
#pragma unroll
        for (int m = 0; m < NV + 1; m++) s.P[m * nc + i] = p[m];
      }
    }
  }
  if (!(flags & (F_LAST | F_RES))) return;
  // residual-norm partials (and global time-step minimum): fixed-order
  // wavefront and workgroup reductions
  double r2[NV], dmin[1] = {dt_next};
#pragma unroll
  for (int kk = 0; kk < NV; kk++) r2[kk] = R[kk] * R[kk];
  block_reduce<CSU_BLOCK, NV>(r2, swr, RedSum());
  block_reduce<CSU_BLOCK, 1>(dmin, swd, RedMin());
  if (t == 0) {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) red_res[kk * gridDim.x + tile] = r2[kk];
    red_dt[tile] = dmin[0];
  }
}

// Time step from the host-set state (solverCalcDt, solver.c:925-950):
// local values or per-block minima for the global step.
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_dt(KP kp, Geo g, int local, double cfl, const double* __restrict__ U,
                                                  double* __restrict__ dtl, double* __restrict__ red) {
  __shared__ double sw[1][CSU_BLOCK / 64];
  int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const int nc = kp.nc;
  if (g.sol) {
    const SolveP& q = g.sol[solve_of(g.csolk, blockIdx.x * blockDim.x)];
    kp.cend = q.c1;
    cfl = q.cfl;
  }
  double val = INFINITY;
  if (i < kp.cend) {
    double u[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) u[k] = U[k * nc + i];
    val = cell_dt<NV>(kp, u, g.omega[i], g.dsxm[i], g.dsym[i]);
    if (local) dtl[i] = cfl * val;
  }
  if (local) return;
  double m[1] = {val};
  block_reduce<CSU_BLOCK, 1>(m, sw, RedMin());
  if (t == 0) red[blockIdx.x] = m[0];
}

// Global time step: CFL x the minimum of the block partials.  Batch (sol
// non-null): one workgroup per solve, over its own blocks, with its CFL.
__global__ __launch_bounds__(CSU_BLOCK) void k_dt_final(int nblocks, double cfl, const double* __restrict__ red,
                                                        double* __restrict__ dtglob, const SolveP* __restrict__ sol) {
  __shared__ double sw[1][CSU_BLOCK / 64];
  int t = threadIdx.x;
  if (sol) {
    const SolveP& q = sol[blockIdx.x];
    red += q.c0 / CSU_BLOCK;
    nblocks = (q.c1 - q.c0 + CSU_BLOCK - 1) / CSU_BLOCK;
    cfl = q.cfl;
    dtglob += blockIdx.x;
  }
  double m[1] = {INFINITY};
  for (int b = t; b < nblocks; b += CSU_BLOCK) m[0] = fmin(m[0], red[b]);
  block_reduce<CSU_BLOCK, 1>(m, sw, RedMin());
  if (t == 0) dtglob[0] = m[0] * cfl;
}

// Time-accurate step control, the tube branch of solverSolve
// (solver.c:1912-1936): the global step is cut to reach tmax and the run
// stops after that step; t advances by dt*stages/2.  ts = t, tmax, state,
// steps; state 0 running, 1 this is the last step, 2 past the last step
// (dt 0 and the update kernel holds U).
__global__ void k_time_clamp(double* __restrict__ ts, double* __restrict__ dtg, int stages) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ts[2] != 0.0) {
    ts[2] = 2.0;
    dtg[0] = 0.0;
    return;
  }
  const double t = ts[0];
  double dt = dtg[0];
  if (t + dt > ts[1]) {
    dt = ts[1] - t;
    dtg[0] = dt;
    ts[2] = 1.0;
  }
  ts[0] = t + dt * stages / 2.0;
  ts[3] += 1.0;
}

// Final level of the residual norm: one block, each thread sums a strided
// slice of the block partials for all variables at once, then a fixed tree.
#define RES_THREADS 1024
template <int NV>
// res_hist slot = ctr % cap: the iteration counter lives on the device so
// that a captured iteration graph writes each replay's norms to its own slot.
// Partial rows have length stride.  Batch (sol non-null): one workgroup per
// solve, over its own blocks; history slot rows hold gridDim.x solves.
__global__ __launch_bounds__(RES_THREADS) void k_res_final(int nc, int nblocks, int stride, const SolveP* __restrict__ sol,
                                                           const double* __restrict__ red,
                                                           double* __restrict__ res_hist, int cap,
                                                           const int* __restrict__ ctr) {
  double* res = res_hist + ((size_t)(ctr[0] % cap) * gridDim.x + blockIdx.x) * NV;
  if (sol) {
    const SolveP& q = sol[blockIdx.x];
    red += q.c0 / CSU_BLOCK;
    nc = q.c1 - q.c0;
    nblocks = (nc + CSU_BLOCK - 1) / CSU_BLOCK;
  }
  __shared__ double sw[NV][RES_THREADS / 64];
  const int t = threadIdx.x;
  double a[NV];
#pragma unroll
  for (int kk = 0; kk < NV; kk++) a[kk] = 0.0;
  for (int b = t; b < nblocks; b += RES_THREADS) {
#pragma unroll
    for (int kk = 0; kk < NV; kk++) a[kk] += red[kk * stride + b];
  }
  block_reduce<RES_THREADS, NV>(a, sw, RedSum());
  if (t == 0)
#pragma unroll
    for (int kk = 0; kk < NV; kk++) res[kk] = sqrt(a[kk] / nc);
}

// Wall force coefficients Cx_p, Cx_v, Cy_p, Cy_v, reference solverCalcCoeff3
// (solver.c:1342-1401); friction per face comes from k_grad.
// Batch (sol non-null): one workgroup per solve over its wall-list range,
// its free stream and reference length; the slot counter is then advanced
// by k_next_slot (a workgroup cannot advance it while others read it).
__global__ __launch_bounds__(CSU_BLOCK) void k_forces(KP kp, int nwall, const int* __restrict__ wall,
                                                      const double* __restrict__ Pb, const double* __restrict__ wfric,
                                                      const double* __restrict__ bdsx, const double* __restrict__ bdsy,
                                                      double sref, int visc, double* __restrict__ coef_hist, int cap,
                                                      int* __restrict__ ctr, const SolveP* __restrict__ sol) {
  double* coef = coef_hist + ((size_t)(ctr[0] % cap) * gridDim.x + blockIdx.x) * 4;
  if (sol) {
    const SolveP& q = sol[blockIdx.x];
    solve_kp(kp, q);
    wall += q.w0;
    nwall = q.w1 - q.w0;
    sref = q.sref;
  }
  __shared__ double sw[4][CSU_BLOCK / 64];
  int t = threadIdx.x;
  const double r = kp.pin[0], u = kp.pin[1], v = kp.pin[2], Pinf = kp.pin[3];
  const double q = 0.5 * r * (u * u + v * v);
  double ax = 0.0, ay = 0.0, vx = 0.0, vy = 0.0;
  for (int w = t; w < nwall; w += CSU_BLOCK) {
    int gq = wall[w];
    double cp = (Pb[3 * kp.nb + gq] - Pinf) / q;
    ax += cp * bdsx[gq];
    ay += cp * bdsy[gq];
    if (visc) {
      vx -= wfric[gq] / q;
      vy -= wfric[kp.nb + gq] / q;
    }
  }
  double f[4] = {ax, vx, ay, vy};
  block_reduce<CSU_BLOCK, 4>(f, sw, RedSum());
  if (t == 0) {
    double c[4] = {f[0] / sref, f[1] / sref, f[2] / sref, f[3] / sref};
    if (kp.axi) { c[0] *= 2 * M_PI; c[1] *= 2 * M_PI; }
    for (int k = 0; k < 4; k++) coef[k] = c[k];
    if (!sol) ctr[0] = ctr[0] + 1;  // last kernel of an iteration: advance the history slot
  }
}

// Batch: advance the history slot after every solve's forces.
__global__ void k_next_slot(int* __restrict__ ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr[0] = ctr[0] + 1;
}

// Per-cell SST fields of the solution2.csv writer, reference sstInterMiT
// (sst.c:964-1007): sstSources on the primitives and the last residual's
// gradients of u, v, k, omega, with no omega = 0 guard.  out [5][nc] =
// eddy viscosity, F1, F2, wall distance, omega.
__global__ __launch_bounds__(CSU_BLOCK) void k_sst_fields(KP kp, const double* __restrict__ P,
                                                          const double* __restrict__ gx,
                                                          const double* __restrict__ gy,
                                                          const double* __restrict__ wd, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc;
  if (i >= nc) return;
  SstV v;
  v.dux = gx[nc + i]; v.dvx = gx[2 * nc + i]; v.dkx = gx[4 * nc + i]; v.dox = gx[5 * nc + i];
  v.duy = gy[nc + i]; v.dvy = gy[2 * nc + i]; v.dky = gy[4 * nc + i]; v.doy = gy[5 * nc + i];
  v.r = P[i]; v.T = P[4 * nc + i]; v.k = P[5 * nc + i]; v.om = P[6 * nc + i]; v.d = wd[i];
  v.mi_L = sutherland(v.T);
  sst_sources(kp, v);
  out[i] = v.mi_t;
  out[nc + i] = v.F1;
  out[2 * nc + i] = v.F2;
  out[3 * nc + i] = v.d;
  out[4 * nc + i] = v.om;
}

// Per boundary face, what the surfData.csv writers take (saSolverWriteSurf
// sa.c:854-960, saCC_SolverWriteSurf saCC.c:599-705, sstSolverWriteSurf
// sst.c:820-962): the boundary viscous flux and eddy viscosity recomputed
// from the current primitives and the last residual's gradients, the ghost
// state of that residual and the wall distance of the face's cell.
// out [11][nb] = f_x f_y f_e miEddy, ghost r u v p T (nu~ | k), d; with
// AOS, [nb][11] (the host layout, copied out without a transpose).
template <int NV, bool AOS = false>
__global__ __launch_bounds__(CSU_BLOCK) void k_surf(KP kp, Geo g, St s, double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc, nb = kp.nb;
  if (q >= nb) return;
  const int i = g.be0[q];
  double Pi[NV + 1];
#pragma unroll
  for (int m = 0; m <= NV; m++) Pi[m] = s.P[m * nc + i];
  const double gx[5] = {s.dPx[nc + i], s.dPx[2 * nc + i], s.dTx[i], s.dPx[4 * nc + i], NV == 6 ? s.dPx[5 * nc + i] : 0.0};
  const double gy[5] = {s.dPy[nc + i], s.dPy[2 * nc + i], s.dTy[i], s.dPy[4 * nc + i], NV == 6 ? s.dPy[5 * nc + i] : 0.0};
  const double dx = g.bxm[q] - g.xc[i], dy = g.bym[q] - g.yc[i];
  const int t = g.btype[q];
  double f[5], mie = 0.0;
  if (NV == 6)
    sst_bc_face(kp, t, g.bdsx[q], g.bdsy[q], g.bnx[q], g.bny[q], g.bds2[q], dx, dy, Pi, g.wdist[i], gx, gy, f, &mie);
  else
    sa_bc_face(kp, t, g.bdsx[q], g.bdsy[q], g.bnx[q], g.bny[q], g.bds2[q], dx, dy, Pi, gx, gy, s.dPx[i], s.dPy[i], f,
               &mie);
  // every one of the 11 entries of face q is written
  const size_t cs = AOS ? 1 : (size_t)nb, rs = AOS ? 11 : 1;
  double* o = out + (size_t)q * rs;
  o[0] = f[0];
  o[cs] = f[1];
  o[2 * cs] = f[2];
  o[3 * cs] = mie;
#pragma unroll
  for (int m = 0; m < 6; m++) o[(4 + m) * cs] = s.Pb[m * nb + q];
  o[10 * cs] = g.wdist[i];
}

}  // namespace csu
