// kern_implicit.h -- implicit time stepping: LU-SGS by level scheduling and
// DPLUR by damped Jacobi (reference implicit.c).
//
// LU-SGS: the reference sweeps cells in index order, a cell using the final
// increments of its lower-index (forward) or higher-index (backward)
// neighbours.  The host groups cells into levels of that dependency DAG; each
// level is one launch, cells of a level are independent.  Every cell does the
// reference's arithmetic on the same inputs, so the result equals the serial
// sweep for any cell order; a colour-sorted numbering makes the number of
// levels equal to the number of colours.
#pragma once
#include "kern_common.h"

namespace csu {

__device__ __forceinline__ double gamma_of(const GasD& g, double T) { return gas_T2gamma<double>(g, T); }

// implicitCalcD (implicit.c:19-181), everything but the global-time-step
// term: writes Dp = 0.5 w Lc (+ Lv/Omega) and per-block min of Omega/Lc.
template <int NV, bool VISC>
__global__ __launch_bounds__(CSU_BLOCK) void k_impD(KP kp, Geo g, St s, double wimp, double* __restrict__ Dp,
                                                    double* __restrict__ red_min, double* __restrict__ caux) {
  __shared__ double sw[1][CSU_BLOCK / 64];
  const int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const int nc = kp.nc;
  if (g.sol) kp.cend = g.sol[solve_of(g.csolk, blockIdx.x * blockDim.x)].c1;
  double val = INFINITY;
  if (i < kp.cend) {
    const double* P = s.P;
    double Lc = 0, Lv = 0;
    // all slots' face data and neighbours, then all states, then the sums
    // in slot order.  The face's two states are the cell's and the
    // neighbour's (cfn) in either order: the face means add them, and a sum
    // of two doubles does not depend on their order.
    int cfj[4], fij[4], nbj[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      cfj[j] = g.cface[j * nc + i];
      nbj[j] = g.cfn[j * nc + i];
      fij[j] = cfj[j] > 0 ? cfj[j] - 1 : (cfj[j] < 0 ? -cfj[j] - 1 : 0);
    }
    double sxj[4], syj[4], sj[4], Pa[4], Pb[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      sxj[j] = g.fdsx[fij[j]];
      syj[j] = g.fdsy[fij[j]];
      sj[j] = g.fds[fij[j]];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) Pa[q] = P[(q == 3 ? 4 : q) * nc + i];  // r u v T
    if (caux) {
      // LU-SGS: the cell-state factors of its neighbour terms, evaluated
      // once per cell instead of once per face and sweep (same expressions
      // of the same values: implicit.c:307-313, 555-595)
      caux[i] = gas_T2c<double>(kp.gas, Pa[3]);
      if (VISC) {
        caux[nc + i] = fmax(4 / (3 * Pa[0]), gamma_of(kp.gas, Pa[3]) / Pa[0]);
        caux[2 * nc + i] = sutherland(Pa[3]) / kp.pr;
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int b = cfj[j] == 0 ? i : nbj[j];
#pragma unroll
      for (int q = 0; q < 4; q++) Pb[j][q] = P[(q == 3 ? 4 : q) * nc + b];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (cfj[j] == 0) continue;
      const int fi = fij[j];
      const double dSx = sxj[j], dSy = syj[j], dS = sj[j];
      double r = (Pa[0] + Pb[j][0]) * 0.5;
      double u = (Pa[1] + Pb[j][1]) * 0.5;
      double v = (Pa[2] + Pb[j][2]) * 0.5;
      double T = (Pa[3] + Pb[j][3]) * 0.5;
      double c = gas_T2c<double>(kp.gas, T);
      double nx = dSx / dS, ny = dSy / dS;
      Lc += (fabs(nx * u + ny * v) + c) * dS;
      if (VISC) {
        double mi = sutherland(T);
        double mt = NV == 4 ? mi / kp.pr : mi / kp.pr + s.miT[fi] / kp.pr_t;
        Lv += fmax(4 / (3 * r), gamma_of(kp.gas, T) / r) * mt * dS * dS;
      }
    }
    const double om = g.omega[i];
    double dtl = Lc;
    double D = 0.5 * wimp * Lc;
    if (VISC) D += Lv / om;
    const int np = g.cnp[i];
    for (int j = 0; j < np; j++) {
      const int n = g.cnbr[j * nc + i];
      if (n >= 0) continue;
      const int q = -1 - n;
      const double dSx = g.bdsx[q], dSy = g.bdsy[q], dS = g.bds[q];
      double r = P[i], u = P[nc + i], v = P[2 * nc + i], T = P[4 * nc + i];
      double c = gas_T2c<double>(kp.gas, T);
      double nx, ny;
      if (dS > 0) { nx = dSx / dS; ny = dSy / dS; } else { nx = 0.0; ny = 1.0; }
      double Lcb = (fabs(nx * u + ny * v) + c) * dS;
      dtl += Lcb;
      D += 0.5 * wimp * Lcb;
      if (VISC) {
        double mi = sutherland(T);
        double Lvb = fmax(4 / (3 * r), gamma_of(kp.gas, T) / r) * (mi / kp.pr) * dS * dS;
        D += Lvb / om;
      }
    }
    Dp[i] = D;
    val = om / dtl;
  }
  double m[1] = {val};
  block_reduce<CSU_BLOCK, 1>(m, sw, RedMin());
  if (t == 0) red_min[blockIdx.x] = m[0];
}

// diagonal with the global-time-step term (implicit.c:176-180); dtg holds
// one step per solve
__device__ __forceinline__ double diag(const double* Dp, const double* omega, const double* dtg, int i, int sid) {
  return Dp[i] + omega[i] / dtg[sid];
}

// implicitCalcDeltaFlux_sa (implicit.c:495-553): with the positivity clamps
__device__ __forceinline__ void delta_flux_sa(const KP& kp, const double* Pj, const double* d, double nx, double ny,
                                              double* dF) {
  double rho = Pj[0], u = Pj[1], v = Pj[2], p = Pj[3], n = Pj[5];
  double T = p / (kp.gas.R * rho);
  double e0 = gas_T2e<double>(kp.gas, T);
  double E = e0 + (u * u + v * v) / 2;
  double U[5] = {rho, rho * u, rho * v, rho * E, rho * n};
  double V = (nx * U[1] + ny * U[2]) / U[0];
  double F0[5] = {U[0] * V, U[1] * V + nx * p, U[2] * V + ny * p, (U[3] + p) * V, U[4] * V};
#pragma unroll
  for (int k = 0; k < 5; k++) U[k] = U[k] + d[k];
  rho = U[0];
  rho = fmax(rho, kp.rlim);
  U[0] = rho;
  u = U[1] / rho; v = U[2] / rho; E = U[3] / rho;
  double e1 = E - (u * u + v * v) * 0.5;
  e1 = fmax(e1, 7.176325e+02);
  T = T + (e1 - e0) / gas_T2Cv<double>(kp.gas, T);
  if (T < 1) {
    T = 1;
    e1 = gas_T2e<double>(kp.gas, T);
    E = e1 + (u * u + v * v) * 0.5;
    U[3] = rho * E;
  }
  p = kp.gas.R * T * rho;
  n = U[4] / rho;
  n = fmax(n, 0.0);
  U[4] = rho * n;
  V = (nx * U[1] + ny * U[2]) / U[0];
  dF[0] = U[0] * V - F0[0];
  dF[1] = (U[1] * V + nx * p) - F0[1];
  dF[2] = (U[2] * V + ny * p) - F0[2];
  dF[3] = (U[3] + p) * V - F0[3];
  dF[4] = U[4] * V - F0[4];
}

// implicitCalcDeltaFlux_sst (implicit.c:1048-1117): k and omega clamps
__device__ __forceinline__ void delta_flux_sst(const KP& kp, const double* Pj, const double* d, double nx, double ny,
                                               double* dF) {
  double rho = Pj[0], u = Pj[1], v = Pj[2], p = Pj[3], k = Pj[5], om = Pj[6];
  double T = p / (kp.gas.R * rho);
  double e0 = gas_T2e<double>(kp.gas, T);
  double E = e0 + (u * u + v * v) * 0.5 + k;
  double U[6] = {rho, rho * u, rho * v, rho * E, rho * k, rho * om};
  double V = (nx * U[1] + ny * U[2]) / U[0];
  double F0[6] = {U[0] * V, U[1] * V + nx * p, U[2] * V + ny * p, (U[3] + p) * V, U[4] * V, U[5] * V};
#pragma unroll
  for (int q = 0; q < 6; q++) U[q] = U[q] + d[q];
  rho = fmax(U[0], kp.rlim);
  U[0] = rho;
  u = U[1] / rho; v = U[2] / rho;
  k = U[4] / rho;
  if (k < 1e-14) { k = 1e-14; U[4] = rho * k; }
  om = U[5] / rho;
  if (om < 0.0) { om = 0.0; U[5] = rho * om; }
  E = U[3] / rho;
  double e1 = E - (u * u + v * v) * 0.5 - k;
  e1 = fmax(e1, 7.176325e+02);
  T = T + (e1 - e0) / gas_T2Cv<double>(kp.gas, T);
  if (T < 1) {
    T = 1;
    e1 = gas_T2e<double>(kp.gas, T);
    E = e1 + (u * u + v * v) * 0.5 + k;
    U[3] = rho * E;
  }
  p = kp.gas.R * T * rho;
  V = (nx * U[1] + ny * U[2]) / U[0];
  dF[0] = U[0] * V - F0[0];
  dF[1] = (U[1] * V + nx * p) - F0[1];
  dF[2] = (U[2] * V + ny * p) - F0[2];
  dF[3] = (U[3] + p) * V - F0[3];
  dF[4] = U[4] * V - F0[4];
  dF[5] = U[5] * V - F0[5];
}

// implicitCalcDeltaFlux (implicit.c:194-231) at state Pj with increment d
template <int NV>
__device__ __forceinline__ void delta_flux(const KP& kp, const double* Pj, const double* d, double nx, double ny,
                                           double* dF) {
  if (NV == 6) {
    delta_flux_sst(kp, Pj, d, nx, ny, dF);
    return;
  }
  if (NV == 5) {
    delta_flux_sa(kp, Pj, d, nx, ny, dF);
    return;
  }
  double rho = Pj[0], u = Pj[1], v = Pj[2], p = Pj[3];
  double T = p / (kp.gas.R * rho);
  double e0 = gas_T2e<double>(kp.gas, T);
  double E = e0 + (u * u + v * v) / 2;
  double U[4] = {rho, rho * u, rho * v, rho * E};
  double V = (nx * U[1] + ny * U[2]) / U[0];
  double F0[4] = {U[0] * V, U[1] * V + nx * p, U[2] * V + ny * p, (U[3] + p) * V};
#pragma unroll
  for (int k = 0; k < 4; k++) U[k] = U[k] + d[k];
  rho = U[0]; u = U[1] / rho; v = U[2] / rho; E = U[3] / rho;
  double e1 = E - 0.5 * (u * u + v * v);
  T = T + (e1 - e0) / gas_T2Cv<double>(kp.gas, T);
  p = kp.gas.R * T * rho;
  V = (nx * U[1] + ny * U[2]) / U[0];
  dF[0] = U[0] * V - F0[0];
  dF[1] = (U[1] * V + nx * p) - F0[1];
  dF[2] = (U[2] * V + ny * p) - F0[2];
  dF[3] = (U[3] + p) * V - F0[3];
}

// Face slot j of cell i seen from i: neighbour id and outward face vector
// (the reference flips p0/p1 when i is the face's e1, implicit.c:288-305;
// the flipped vector is the exact negation).
__device__ __forceinline__ int face_from(const Geo& g, int nc, int i, int j, double& dSx, double& dSy, double& dS,
                                         int& fi) {
  const int f = g.cface[j * nc + i], nb = g.cfn[j * nc + i];
  if (f == 0) return -1;
  fi = f > 0 ? f - 1 : -f - 1;
  dS = g.fds[fi];
  if (f > 0) { dSx = g.fdsx[fi]; dSy = g.fdsy[fi]; return nb; }
  dSx = -g.fdsx[fi]; dSy = -g.fdsy[fi];
  return nb;
}

// One LU-SGS level.  DIR 0: forward sweep (implicitLUSGS_L, implicit.c:274-363),
// dW = -R then lower neighbours.  DIR 1: backward sweep (implicitLUSGS_U,
// implicit.c:365-482) from the forward result dW0: the cell's start value
// D*dW0 (implicit.c:369-399), then upper neighbours; a cell's dW is final
// once its level has run, so the level also applies it to the state
// (solverUpdateUImplicit, solver.c:1945-1955; a held solve of a batch keeps
// its state).
template <int NV, bool VISC, int DIR>
__global__ __launch_bounds__(CSU_BLOCK) void k_lusgs(KP kp, Geo g, St s, const int* __restrict__ cells, int ncell_lvl,
                                                     const double* __restrict__ Dp, const double* __restrict__ dtg,
                                                     const double* __restrict__ R, double* dW,
                                                     const double* __restrict__ dW0, double* __restrict__ U,
                                                     const double* __restrict__ caux) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncell_lvl) return;
  const int nc = kp.nc;
  const int i = cells[t];
  const int sid = g.sol ? g.csol[i] : 0;
  const double D = diag(Dp, g.omega, dtg, i, sid);
  double acc[NV];
  if (DIR == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = -R[k * nc + i];
  } else {
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = D * dW0[k * nc + i];
    if (NV == 6) {  // implicit.c:372-392: the k-omega block times dW0
      const double a = D - s.dq[nc + i], b = -s.dq[2 * nc + i], c = -s.dq[4 * nc + i], dd = D - s.dq[5 * nc + i];
      const double det = a * dd - b * c;
      if (!(fabs(det) < 1e-14)) {
        acc[NV > 4 ? 4 : 0] = a * dW0[4 * nc + i] + b * dW0[5 * nc + i] - s.dq[i] * dW0[i];
        acc[NV > 5 ? 5 : 0] = c * dW0[4 * nc + i] + dd * dW0[5 * nc + i] - s.dq[3 * nc + i] * dW0[i];
      }
    }
  }
  // Face data of all slots is loaded before any slot is used (one round
  // trip per dependency level, not per slot); neighbour states are loaded
  // per slot, only for the neighbours this sweep uses (loading them all up
  // front needs ~200 VGPRs and halves the occupancy, measured slower).
  int cfj[4], nbj[4], fij[4];
#pragma unroll
  for (int j = 0; j < 4; j++) cfj[j] = g.cface[j * nc + i];
  double nxj[4], nyj[4], sj[4], ddj[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int f = cfj[j];
    const int fi = f > 0 ? f - 1 : (f < 0 ? -f - 1 : 0);
    fij[j] = fi;
    sj[j] = g.fds[fi];
    nxj[j] = g.fnx[fi];
    nyj[j] = g.fny[fi];
    ddj[j] = VISC ? g.fdd[fi] : 1.0;
    const int e = g.cfn[j * nc + i];  // loaded beside cface, not after it
    nbj[j] = f == 0 ? i : e;
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    // the slot's neighbour and outward unit normal (face_from: negated when
    // i is the face's e1, implicit.c:288-305; dS/|dS| with the sign of dS,
    // stored per face)
    if (cfj[j] == 0) continue;
    const int nb = nbj[j], fi = fij[j];
    const double dS = sj[j];
    if (DIR == 0 ? !(nb < i) : !(nb > i)) continue;
    const double nx = cfj[j] > 0 ? nxj[j] : -nxj[j], ny = cfj[j] > 0 ? nyj[j] : -nyj[j];
    double Pj[NV + 1];
#pragma unroll
    for (int m = 0; m < NV + 1; m++) Pj[m] = s.P[m * nc + nb];
    double d[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) d[k] = dW[k * nc + nb];
    double dF[NV];
    delta_flux<NV>(kp, Pj, d, nx, ny, dF);
    // the neighbour's sound speed and viscous factors max(4/(3r), gamma/r)
    // and mu/Pr (k_impD, per cell)
    const double c = caux[nb];
    double ra = kp.wimp * (fabs(nx * Pj[1] + ny * Pj[2]) + c) * dS;
    if (NV >= 5) {
      // implicitFunc_sa / _sst (implicit.c:555-595, 1120-1161)
      ra += caux[nc + nb] * (caux[2 * nc + nb] + s.miT[fi] / kp.pr_t) * dS / ddj[j];
    } else if (VISC && kp.laminar) {
      ra += caux[nc + nb] * caux[2 * nc + nb] * dS / ddj[j];
    }
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] -= 0.5 * (dF[k] * dS - ra * d[k]);
  }
  double out[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) out[k] = acc[k] / D;
  if (NV == 6) {
    // SST: k-omega block with the source Jacobians, Cramer's rule
    // (implicit.c:326-351, 447-472).  The reference divides by a long double
    // determinant that holds a double value; the quotient here is the
    // correctly rounded double one.
    constexpr int K = NV > 4 ? 4 : 0, O = NV > 5 ? 5 : 0;
    const double x0 = acc[0] / D;
    const double A = acc[K] + s.dq[i] * x0;
    const double B = acc[O] + s.dq[3 * nc + i] * x0;
    const double a = D - s.dq[nc + i], b = -s.dq[2 * nc + i], c = -s.dq[4 * nc + i], dd = D - s.dq[5 * nc + i];
    const double det = a * dd - b * c;
    if (!(fabs(det) < 1e-14)) {
      out[K] = (A * dd - B * b) / det;
      out[O] = (-A * c + B * a) / det;
    }
  }
#pragma unroll
  for (int k = 0; k < NV; k++) dW[k * nc + i] = out[k];
  if (DIR == 1 && (!g.sol || g.sol[sid].active)) {
#pragma unroll
    for (int k = 0; k < NV; k++) U[k * nc + i] += out[k];
  }
}

// DPLUR finite-difference blocks of every interior neighbour
// (implicitUpdateA, implicit.c:689-813; implicitUpdateA_sa :815-936),
// h = 1e-6.  Layout [4 slots][NV*NV][nc].  SST (NV 6) has no DPLUR in the
// reference (implicitCalcDPLUR builds 4x4 blocks for it); its 6x6 blocks
// here take implicitUpdateA_sa's form with two convected scalars (the
// fluxes of implicitAuxCalcFlux_sst2, implicit.c:1036-1046) and k taken out
// of the internal energy, as implicitCalcDeltaFlux_sst does (:1048-1117).
template <int NV, bool VISC>
__global__ __launch_bounds__(CSU_BLOCK) void k_blocks(KP kp, Geo g, St s, const double* __restrict__ U,
                                                      double* __restrict__ A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nc = kp.nc;
  if (g.sol) kp.cend = g.sol[solve_of(g.csolk, blockIdx.x * blockDim.x)].c1;
  if (i >= kp.cend) return;
  const double h = 1.0e-6;
  // Every slot's face data and neighbour state first (one memory round trip
  // for all slots instead of one chain per slot); slots without a face
  // read the cell itself and are skipped.
  int nbs[4], fis[4];
  double sxs[4], sys_[4], sds[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    fis[j] = 0;
    sxs[j] = sys_[j] = 0.0;
    sds[j] = 1.0;
    nbs[j] = face_from(g, nc, i, j, sxs[j], sys_[j], sds[j], fis[j]);
  }
  double Ujs[4][NV], P1s[4], P2s[4], P3s[4], P4s[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int nb = nbs[j] >= 0 ? nbs[j] : i;
#pragma unroll
    for (int k = 0; k < NV; k++) Ujs[j][k] = U[k * nc + nb];
    P1s[j] = s.P[nc + nb]; P2s[j] = s.P[2 * nc + nb]; P3s[j] = s.P[3 * nc + nb]; P4s[j] = s.P[4 * nc + nb];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int nb = nbs[j], fi = fis[j];
    if (nb < 0) continue;
    const double dSx = sxs[j], dSy = sys_[j], dS = sds[j];
    const double nx = dSx / dS, ny = dSy / dS;
    const double P3 = P3s[j], P4 = P4s[j];
    double Uj[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) Uj[k] = Ujs[j][k];
    double rho = Uj[0], u = Uj[1] / rho, v = Uj[2] / rho, E = Uj[3] / rho;
    double V = (nx * Uj[1] + ny * Uj[2]) / Uj[0];
    double F0[NV];
    F0[0] = Uj[0] * V; F0[1] = Uj[1] * V + nx * P3; F0[2] = Uj[2] * V + ny * P3; F0[3] = (Uj[3] + P3) * V;
    if (NV >= 5) F0[4] = Uj[4] * V;
    if (NV == 6) F0[5] = Uj[5] * V;
    const double ee0 = NV == 6 ? E - 0.5 * (u * u + v * v) - Uj[4] / rho : E - 0.5 * (u * u + v * v);
    const double cv = gas_T2Cv<double>(kp.gas, P4);
    // the diagonal's spectral-radius term, subtracted from each diagonal
    // entry as it is stored (the reference subtracts it from the stored
    // block, implicit.c:792-809: the same two roundings)
    double c = gas_T2c<double>(kp.gas, P4);
    double ra = kp.wimp * (fabs(nx * P1s[j] + ny * P2s[j]) + c) * dS;
    if (NV >= 5) {
      // implicitUpdateA_sa (implicit.c:815-936); SST the same with its
      // face eddy viscosity
      double r = (s.P[i] + s.P[nb]) * 0.5;
      double T = (s.P[4 * nc + i] + P4) * 0.5;
      double mi = sutherland(T);
      double ddx = g.xc[nb] - g.xc[i], ddy = g.yc[nb] - g.yc[i];
      double dd_ = sqrt(ddx * ddx + ddy * ddy);
      ra += fmax(4 / (3 * r), gamma_of(kp.gas, T) / r) * (mi / kp.pr + s.miT[fi] / kp.pr_t) * dS / dd_;
    } else if (VISC && kp.laminar) {
      double r = (s.P[i] + s.P[nb]) * 0.5;
      double T = (s.P[4 * nc + i] + P4) * 0.5;
      double mi = sutherland(T);
      double ddx = g.xc[nb] - g.xc[i], ddy = g.yc[nb] - g.yc[i];
      double dd_ = sqrt(ddx * ddx + ddy * ddy);
      ra += fmax(4 / (3 * r), gamma_of(kp.gas, T) / r) * (mi / kp.pr) * dS / dd_;
    }
    double* Aj = A + (size_t)j * NV * NV * nc;
#pragma unroll
    for (int mm = 0; mm < NV; mm++) {
      double Up[NV];
#pragma unroll
      for (int n = 0; n < NV; n++) Up[n] = mm == n ? Uj[n] + h : Uj[n];
      rho = Up[0]; u = Up[1] / rho; v = Up[2] / rho; E = Up[3] / rho;
      double ee1 = NV == 6 ? E - 0.5 * (u * u + v * v) - Up[4] / rho : E - 0.5 * (u * u + v * v);
      double T = P4 + (ee1 - ee0) / cv;
      double p = kp.gas.R * rho * T;
      double Vp = (nx * Up[1] + ny * Up[2]) / Up[0];
      double F[NV];
      F[0] = Up[0] * Vp; F[1] = Up[1] * Vp + nx * p; F[2] = Up[2] * Vp + ny * p; F[3] = (Up[3] + p) * Vp;
      if (NV >= 5) F[4] = Up[4] * Vp;
      if (NV == 6) F[5] = Up[5] * Vp;
#pragma unroll
      for (int n = 0; n < NV; n++) {
        const double a = 0.5 * (F[n] - F0[n]) * dS / h;
        Aj[(size_t)(n * NV + mm) * nc + i] = n == mm ? a - 0.5 * ra : a;
      }
    }
  }
}

// Largest of every thread's v in the workgroup (fmax: NaN is the identity,
// the maximum is exact in any order), in every thread.
__device__ __forceinline__ double block_fmax(double v, double (*sw)[CSU_BLOCK / 64]) {
  double r[1] = {v};
  block_reduce<CSU_BLOCK, 1>(r, sw, RedMax());
  return r[0];
}

// One damped Jacobi sweep m (implicitMultA + update, implicit.c:938-1022):
// y = D x + sum_j A_j x_j;  x_new = w (-R - y)/D + x, from buf[(state[1] +
// m) & 1] into the other buffer; state[1] is the buffer holding the last
// implicit step's result.  The block maxima of |dx_rho| go to red_sw[m][.];
// each sweep first reduces the previous sweep's (the reference's test
// after every sweep, implicit.c:1026-1031), and once that is below 1e-14
// (or marked -1: the sweeps had already stopped) it writes -1 and does
// nothing: the reference breaks out of its loop.
// SST: the diagonal is LU-SGS's block (implicit.c:326-351), D with the
// k-omega source Jacobians dq; y takes its product and x_new solves with it
// by Cramer's rule (D alone where its determinant is below 1e-14).
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_jacobi(KP kp, Geo g, const double* __restrict__ A,
                                                      const double* __restrict__ Dp, const double* __restrict__ dtg,
                                                      const double* __restrict__ R, double* buf0, double* buf1,
                                                      const int* __restrict__ state, double* __restrict__ red_sw,
                                                      int m, const double* __restrict__ dq) {
  __shared__ double sw0[1][CSU_BLOCK / 64], sw1[1][CSU_BLOCK / 64];
  const int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const int nc = kp.nc, stride = gridDim.x;
  // batch: the solve's own blocks and buffer state
  int sid = 0, b0 = 0, nblocks = gridDim.x;
  if (g.sol) {
    sid = solve_of(g.csolk, blockIdx.x * blockDim.x);
    const SolveP& q = g.sol[sid];
    kp.cend = q.c1;
    b0 = q.c0 / CSU_BLOCK;
    nblocks = (q.c1 - q.c0 + CSU_BLOCK - 1) / CSU_BLOCK;
  }
  const int par = (state[2 * sid + 1] + m) & 1;
  const double* x = par ? buf1 : buf0;
  double* y = par ? buf0 : buf1;
  const bool act = i < kp.cend;
  // Every load that does not depend on the stop test is issued before it:
  // the cell's x, R, diagonal and neighbour ids (per face slot, cfn), then
  // the neighbours' x.  Slots without a face read the cell itself and are
  // skipped; lanes past the last cell read cell 0 and store nothing.
  const int ii = act ? i : 0;
  double xi[NV], Ri[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) xi[k] = x[k * nc + ii];
  int nbj[4];
  bool has[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int c = act ? g.cfn[j * nc + ii] : -1;
    has[j] = c >= 0;
    nbj[j] = c >= 0 ? c : ii;
  }
#pragma unroll
  for (int k = 0; k < NV; k++) Ri[k] = R[k * nc + ii];
  const double D = diag(Dp, g.omega, dtg, ii, sid);
  double dqi[6];
#pragma unroll
  for (int q = 0; q < 6; q++) dqi[q] = NV == 6 ? dq[q * nc + ii] : 0.0;
  double xj[4][NV];
#pragma unroll
  for (int j = 0; j < 4; j++)
#pragma unroll
    for (int n = 0; n < NV; n++) xj[j][n] = x[n * nc + nbj[j]];
  if (m > 0) {
    // The maximum over the solve's blocks is at least this block's own
    // (a number >= 1e-14 rules the stop out, so the reduction is skipped:
    // the common case while the sweeps run).  Otherwise reduce every
    // block's.  NaN is fmax's identity: the maximum is NaN only when every
    // cell's change is, which the reference's loop (implicit.c:1023-1031)
    // does not take as converged.  The branch is uniform in the block.
    const double* prev = red_sw + (size_t)(m - 1) * stride + b0;
    const double own = red_sw[(size_t)(m - 1) * stride + blockIdx.x];
    if (!(own >= 1e-14)) {
      double pm = __builtin_nan("");
      for (int b = t; b < nblocks; b += CSU_BLOCK) pm = fmax(pm, prev[b]);
      if (block_fmax(pm, sw0) < 1e-14) {
        if (t == 0) red_sw[(size_t)m * stride + blockIdx.x] = -1.0;
        return;
      }
    }
  }
  double diff = __builtin_nan("");  // lanes past the last cell: fmax's identity
  if (act) {
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = D * xi[k];
    double qa = D, qb = 0.0, qc = 0.0, qd = D, q0 = 0.0, q3 = 0.0, det = 0.0;
    bool blk = false;
    if (NV == 6) {
      q0 = dqi[0]; q3 = dqi[3];
      qa = D - dqi[1]; qb = -dqi[2]; qc = -dqi[4]; qd = D - dqi[5];
      det = qa * qd - qb * qc;
      blk = !(fabs(det) < 1e-14);
      if (blk) {
        acc[4] = qa * xi[4] + qb * xi[5] - q0 * xi[0];
        acc[5] = qc * xi[4] + qd * xi[5] - q3 * xi[0];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (!has[j]) continue;
      const double* Aj = A + (size_t)j * NV * NV * nc;
#pragma unroll
      for (int k = 0; k < NV; k++)
#pragma unroll
        for (int n = 0; n < NV; n++) acc[k] += Aj[(size_t)(k * NV + n) * nc + i] * xj[j][n];
    }
    const double w = 2. / 3.;
    if (NV == 6 && blk) {
      double z[NV];
#pragma unroll
      for (int k = 0; k < 4; k++) z[k] = (-Ri[k] - acc[k]) / D;
      const double Ak = -Ri[4] - acc[4] + q0 * z[0];
      const double Bo = -Ri[5] - acc[5] + q3 * z[0];
      z[4] = (Ak * qd - Bo * qb) / det;
      z[5] = (-Ak * qc + Bo * qa) / det;
#pragma unroll
      for (int k = 0; k < NV; k++) {
        const double yk = w * z[k] + xi[k];
        y[k * nc + i] = yk;
        if (k == 0) diff = fabs(yk - xi[0]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < NV; k++) {
        double yk = w * (-Ri[k] - acc[k]) / D + xi[k];
        y[k * nc + i] = yk;
        if (k == 0) diff = fabs(yk - xi[0]);
      }
    }
  }
  const double bm = block_fmax(diff, sw1);
  if (t == 0) red_sw[(size_t)m * stride + blockIdx.x] = bm;
}

// Buffer of the DPLUR result of solve sid, whose first block is b0 (rows of
// red_sw have length stride): sweeps 0..k-1 ran (k = the first sweep marked
// -1, or nlin), so the last one wrote buf[(state[2 sid + 1] + k) & 1].
__device__ __forceinline__ int jac_result(const int* state, const double* red_sw, int nlin, int stride, int sid,
                                          int b0) {
  // eight sweeps' marks per round trip (a loop that tests each mark before
  // loading the next waits for one memory round trip per sweep)
  int k = nlin;
  for (int k0 = 0; k0 < nlin && k == nlin; k0 += 8) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = k0 + q < nlin ? red_sw[(size_t)(k0 + q) * stride + b0] : 0.0;
#pragma unroll
    for (int q = 7; q >= 0; q--)
      if (k0 + q < nlin && v[q] < 0.0) k = k0 + q;
  }
  return (state[2 * sid + 1] + k) & 1;
}

// After the update: the result buffer becomes the next step's start (one
// thread per solve).
__global__ void k_jac_base(int* state, const double* __restrict__ red_sw, int nlin, int stride,
                           const SolveP* __restrict__ sol, int nsolve) {
  const int sid = blockIdx.x * blockDim.x + threadIdx.x;
  if (sid >= nsolve) return;
  const int b0 = sol ? sol[sid].c0 / CSU_BLOCK : 0;
  state[2 * sid + 1] = jac_result(state, red_sw, nlin, stride, sid, b0);
}

// solverUpdateUImplicit (solver.c:1945-1955): U += dW; for DPLUR (red_sw
// non-null) dW is the sweeps' result buffer.  Batch: a solve that is not
// active keeps its state.
template <int NV>
__global__ __launch_bounds__(CSU_BLOCK) void k_upd_imp(int nc, double* __restrict__ U, const double* buf0,
                                                       const double* buf1, const int* __restrict__ state,
                                                       const double* __restrict__ red_sw, int nlin, int stride,
                                                       const SolveP* __restrict__ sol, const int* __restrict__ csolk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int sid = 0, b0 = 0, cend = nc;
  if (sol) {
    sid = solve_of(csolk, blockIdx.x * blockDim.x);
    const SolveP& q = sol[sid];
    if (!q.active) return;
    b0 = q.c0 / CSU_BLOCK;
    cend = q.c1;
  }
  if (i >= cend) return;
  const double* dW = (red_sw && jac_result(state, red_sw, nlin, stride, sid, b0)) ? buf1 : buf0;
#pragma unroll
  for (int k = 0; k < NV; k++) U[k * nc + i] += dW[k * nc + i];
}

}  // namespace csu
