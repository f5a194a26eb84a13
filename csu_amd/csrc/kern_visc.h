// kern_visc.h -- laminar viscous fluxes (reference laminar.c) and wall
// friction (boundary.c:695-742), per face / per boundary face.
#pragma once
#include "kern_common.h"

namespace csu {

// Stress tensor from velocity gradients and viscosity (laminar.c:65-67).
struct Tau { double xx, yy, xy; };
__device__ __forceinline__ Tau tau_of(double mi, double dux, double duy, double dvx, double dvy) {
  Tau t;
  t.xx = 2 * mi * (dux - (dux + dvy) / 3);
  t.yy = 2 * mi * (dvy - (dux + dvy) / 3);
  t.xy = mi * (duy + dvx);
  return t;
}

// Directional correction of an averaged gradient g = (gx, gy) with the
// difference quotient dl along (dx, dy)/L, laminar.c:52-59.
__device__ __forceinline__ void edge_correct(double gxm, double gym, double dl, double dx, double dy, double L,
                                             double& gx, double& gy) {
  gx = gxm + (dl - (gxm * dx + gym * dy) / L) * dx / L;
  gy = gym + (dl - (gxm * dx + gym * dy) / L) * dy / L;
}

// Interior face viscous flux F[1..3], reference laminarInter face loop
// (laminar.c:17-75).  Slots of P: 1 u, 2 v, 4 T; gradients of u, v, T.
__device__ __forceinline__ void laminar_face(const KP& kp, double dSx, double dSy, double dx, double dy,
                                             const double* P0, const double* P1,
                                             const double* g0x, const double* g0y,   // e0: u v T
                                             const double* g1x, const double* g1y,   // e1: u v T
                                             double* F) {
  double duxm = (g0x[0] + g1x[0]) * 0.5, dvxm = (g0x[1] + g1x[1]) * 0.5, dTxm = (g0x[2] + g1x[2]) * 0.5;
  double duym = (g0y[0] + g1y[0]) * 0.5, dvym = (g0y[1] + g1y[1]) * 0.5, dTym = (g0y[2] + g1y[2]) * 0.5;
  double L = sqrt(dx * dx + dy * dy);
  double dul = (P1[1] - P0[1]) / L, dvl = (P1[2] - P0[2]) / L, dTl = (P1[4] - P0[4]) / L;
  double dux, duy, dvx, dvy, dTx, dTy;
  edge_correct(duxm, duym, dul, dx, dy, L, dux, duy);
  edge_correct(dvxm, dvym, dvl, dx, dy, L, dvx, dvy);
  edge_correct(dTxm, dTym, dTl, dx, dy, L, dTx, dTy);
  double T = (P1[4] + P0[4]) * 0.5;
  double mi = sutherland(T);
  double k = gas_T2Cp<double>(kp.gas, T) * mi / kp.pr;
  Tau t = tau_of(mi, dux, duy, dvx, dvy);
  double u = (P1[1] + P0[1]) * 0.5, v = (P1[2] + P0[2]) * 0.5;
  F[0] = t.xx * dSx + t.xy * dSy;
  F[1] = t.xy * dSx + t.yy * dSy;
  F[2] = (t.xx * dSx + t.xy * dSy) * u + (t.xy * dSx + t.yy * dSy) * v + k * (dTx * dSx + dTy * dSy);
}

// Boundary-face viscous flux f[0..2] (= reference f[1..3]),
// boundaryFaceViscFlux (laminar.c:128-346).  Pi: cell primitives; g*: the
// cell's gradients of u, v, T; (dx, dy): face centre minus cell centre.
// The adiabatic-wall energy term is left to the caller (carried value).
__device__ __forceinline__ void laminar_bc_face(const KP& kp, int type, double dSx, double dSy, double nx,
                                                double ny, double dS2, double dx, double dy, const double* Pi,
                                                const double* gx, const double* gy, double* f) {
  double dux, duy, dvx, dvy, dTx, dTy;
  if (type == CSU_BC_SYMMETRY) {
    dux = gx[0] - (gx[0] * nx + gy[0] * ny) * nx;
    duy = gy[0] - (gx[0] * nx + gy[0] * ny) * ny;
    dvx = gx[1] - (gx[1] * nx + gy[1] * ny) * nx;
    dvy = gy[1] - (gx[1] * nx + gy[1] * ny) * ny;
    dTx = gx[2] - (gx[2] * nx + gy[2] * ny) * nx;
    dTy = gy[2] - (gx[2] * nx + gy[2] * ny) * ny;
    double T = Pi[4], mi = sutherland(T);
    double k = gas_T2Cp<double>(kp.gas, T) * mi / kp.pr;
    Tau t = tau_of(mi, dux, duy, dvx, dvy);
    double u = Pi[1], v = Pi[2];
    f[0] = (t.xx * nx + t.xy * ny) * dS2;
    f[1] = (t.xy * nx + t.yy * ny) * dS2;
    f[2] = (u * (t.xx * nx + t.xy * ny) + v * (t.xy * nx + t.yy * ny) + k * (dTx * nx + dTy * ny)) * dS2;
    return;
  }
  if (type == CSU_BC_OUTLET) {
    double T = Pi[4], mi = sutherland(T);
    double k = gas_T2Cp<double>(kp.gas, T) * mi / kp.pr;
    Tau t = tau_of(mi, gx[0], gy[0], gx[1], gy[1]);
    double u = Pi[1], v = Pi[2];
    f[0] = t.xx * dSx + t.xy * dSy;
    f[1] = t.xy * dSx + t.yy * dSy;
    f[2] = u * (t.xx * dSx + t.xy * dSy) + v * (t.xy * dSx + t.yy * dSy) + k * (gx[2] * dSx + gy[2] * dSy);
    return;
  }
  // inlet, wall, wallT: one-sided difference to the boundary value
  double L = sqrt(dx * dx + dy * dy);
  double ub = type == CSU_BC_INLET ? kp.pin[1] : 0.0, vb = type == CSU_BC_INLET ? kp.pin[2] : 0.0;
  double dul = (ub - Pi[1]) / L, dvl = (vb - Pi[2]) / L;
  edge_correct(gx[0], gy[0], dul, dx, dy, L, dux, duy);
  edge_correct(gx[1], gy[1], dvl, dx, dy, L, dvx, dvy);
  double T = type == CSU_BC_WALLT ? kp.twall : Pi[4];
  double mi = sutherland(T);
  Tau t = tau_of(mi, dux, duy, dvx, dvy);
  f[0] = t.xx * dSx + t.xy * dSy;
  f[1] = t.xy * dSx + t.yy * dSy;
  f[2] = 0.0;
  if (type == CSU_BC_WALL) return;
  double Tb = type == CSU_BC_INLET ? kp.pin[4] : kp.twall;
  double dTl = (Tb - Pi[4]) / L;
  edge_correct(gx[2], gy[2], dTl, dx, dy, L, dTx, dTy);
  double k = gas_T2Cp<double>(kp.gas, T) * mi / kp.pr;
  if (type == CSU_BC_WALLT) {
    f[2] = k * (dTx * dSx + dTy * dSy);
    return;
  }
  double u = Pi[1], v = Pi[2];
  f[2] = u * (t.xx * dSx + t.xy * dSy) + v * (t.xy * dSx + t.yy * dSy) + k * (dTx * dSx + dTy * dSy);
}

// Wall friction force (fx, fy), reference boundaryCalcFrictionWall
// (boundary.c:695-742).
__device__ __forceinline__ void wall_friction(double dSx, double dSy, double dx, double dy, const double* Pi,
                                              const double* gx, const double* gy, double& fx, double& fy) {
  double L = sqrt(dx * dx + dy * dy);
  double dul = (0 - Pi[1]) / L, dvl = (0 - Pi[2]) / L;
  double dux, duy, dvx, dvy;
  edge_correct(gx[0], gy[0], dul, dx, dy, L, dux, duy);
  edge_correct(gx[1], gy[1], dvl, dx, dy, L, dvx, dvy);
  double mi = sutherland(Pi[4]);
  Tau t = tau_of(mi, dux, duy, dvx, dvy);
  fx = t.xx * dSx + t.xy * dSy;
  fy = t.xy * dSx + t.yy * dSy;
}

}  // namespace csu
