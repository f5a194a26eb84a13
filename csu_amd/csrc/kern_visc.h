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

namespace csu {

// ---------------------------------------------------------------------------
// Spalart-Allmaras (reference sa.c)
// ---------------------------------------------------------------------------
// saCalcFace (sa.c:273-290): fv1 and the nu~ diffusion flux density.
__device__ __forceinline__ void sa_calc_face(double ni, double ni_L, double r, double dnix, double dniy,
                                             double& fv1, double& tx, double& ty) {
  const double Cv1 = 7.1;
  double X = ni / ni_L;
  double aux = X * X * X;
  fv1 = aux / (aux + Cv1 * Cv1 * Cv1);
  const double sig = 2. / 3.;
  tx = r * (ni_L + ni) * dnix / sig;
  ty = r * (ni_L + ni) * dniy / sig;
}

// saCC_CalcFace (saCC.c:146-163): the nu~ diffusion flux density with the
// density-gradient term 0.5 nu~^2 grad rho.
__device__ __forceinline__ void sacc_calc_face(double ni, double ni_L, double r, double dnix, double dniy,
                                               double drx, double dry, double& fv1, double& tx, double& ty) {
  const double Cv1 = 7.1;
  double X = ni / ni_L;
  double aux = X * X * X;
  fv1 = aux / (aux + Cv1 * Cv1 * Cv1);
  const double sig = 2. / 3.;
  tx = (r * (ni_L + ni) * dnix + 0.5 * ni * ni * drx) / sig;
  ty = (r * (ni_L + ni) * dniy + 0.5 * ni * ni * dry) / sig;
}

// saCalcSource (sa.c:292-349), Allmaras et al. 2012, without trip terms;
// cc: saCC_CalcSource (saCC.c:165-226), its guarded S~ branch and its
// density-gradient terms.
__device__ __forceinline__ double sa_source(double ni, double ni_L, double S, double d, double rho, double drx,
                                            double dry, double dnix, double dniy, bool cc = false) {
  const double Cv1 = 7.1;
  double X = ni / ni_L;
  double aux = X * X * X;
  double fv1 = aux / (aux + Cv1 * Cv1 * Cv1);
  const double k = 0.41;
  double fv2 = 1. - X / (1. + X * fv1);
  double Sbar = ni * fv2 / (k * k * d * d);
  double Stil;
  const double cv2 = 0.7, cv3 = 0.9;
  if (Sbar >= -cv2 * S) {
    Stil = S + Sbar;
  } else if (cc) {
    const double aux2 = ((cv3 - 2 * cv2) * S - Sbar);
    Stil = aux2 == 0 ? S : S + S * (cv2 * cv2 * S + cv3 * Sbar) / aux2;
  } else {
    Stil = S + S * (cv2 * cv2 * S + cv3 * Sbar) / ((cv3 - 2 * cv2) * S - Sbar);
  }
  const double sig = 2. / 3., Cb1 = 0.1355, Cb2 = 0.622;
  const double Cw1 = Cb1 / (k * k) + (1. + Cb2) / sig;
  const double Cw2 = 0.3;
  double r = ni / (Stil * k * k * d * d);
  r = fmin(r, 10.0);
  aux = r * r;
  aux = aux * aux * aux;
  double g = r + Cw2 * (aux - r);
  const double Cw3 = 2.;
  aux = Cw3 * Cw3;
  aux = aux * aux * aux;
  double fw = g * pow((1. + aux) / (g * g * g * g * g * g + aux), 1. / 6.);
  if (cc) {
    double Qt = rho * (Cb1 * Stil * ni + Cb2 * (dnix * dnix + dniy * dniy) / sig - Cw1 * fw * (ni * ni / (d * d)));
    Qt += (Cb2 / sig) * (ni * (dnix * drx + dniy * dry) + 0.25 * ni * ni * (drx * drx + dry * dry) / rho);
    return Qt;
  }
  return rho * (Cb1 * Stil * ni + Cb2 * (dnix * dnix + dniy * dniy) / sig - Cw1 * fw * (ni * ni / (d * d))) -
         (ni_L + ni) * (dnix * drx + dniy * dry) / sig;
}

// Interior SA viscous face flux F[0..3] (= reference faceFlux[1..4]) and
// the face eddy viscosity, saInterFace (sa.c:138-221) or, with
// viscBlazek, saInterFaceB (sa.c:26-134), or with saCC saCC_InterFace
// (saCC.c:18-103, which ignores viscBlazek).  Gradients: u v T nu~ rho.
__device__ __forceinline__ void sa_face(const KP& kp, double dSx, double dSy, double dx, double dy,
                                        const double* P0, const double* P1, const double* g0x, const double* g0y,
                                        const double* g1x, const double* g1y, double* F, double& mi_t) {
  double dux = (g0x[0] + g1x[0]) * 0.5, dvx = (g0x[1] + g1x[1]) * 0.5;
  double dTx = (g0x[2] + g1x[2]) * 0.5, dnx = (g0x[3] + g1x[3]) * 0.5;
  double duy = (g0y[0] + g1y[0]) * 0.5, dvy = (g0y[1] + g1y[1]) * 0.5;
  double dTy = (g0y[2] + g1y[2]) * 0.5, dny = (g0y[3] + g1y[3]) * 0.5;
  if (kp.visc_blazek && !kp.sa_cc) {
    double L = sqrt(dx * dx + dy * dy);
    double dul = (P1[1] - P0[1]) / L, dvl = (P1[2] - P0[2]) / L, dTl = (P1[4] - P0[4]) / L, dnl = (P1[5] - P0[5]) / L;
    double a, b;
    edge_correct(dux, duy, dul, dx, dy, L, a, b); dux = a; duy = b;
    edge_correct(dvx, dvy, dvl, dx, dy, L, a, b); dvx = a; dvy = b;
    edge_correct(dTx, dTy, dTl, dx, dy, L, a, b); dTx = a; dTy = b;
    edge_correct(dnx, dny, dnl, dx, dy, L, a, b); dnx = a; dny = b;
  }
  double r = (P1[0] + P0[0]) * 0.5, u = (P1[1] + P0[1]) * 0.5, v = (P1[2] + P0[2]) * 0.5;
  double T = (P1[4] + P0[4]) * 0.5, n = (P1[5] + P0[5]) * 0.5;
  double mi_L = sutherland(T), n_L = mi_L / r, fv1, tx, ty;
  if (kp.sa_cc)
    sacc_calc_face(n, n_L, r, dnx, dny, (g0x[4] + g1x[4]) * 0.5, (g0y[4] + g1y[4]) * 0.5, fv1, tx, ty);
  else
    sa_calc_face(n, n_L, r, dnx, dny, fv1, tx, ty);
  mi_t = fv1 * r * n;
  double mi = mi_L + mi_t;
  double k = gas_T2Cp<double>(kp.gas, T) * (mi_L / kp.pr + mi_t / kp.pr_t);
  Tau t = tau_of(mi, dux, duy, dvx, dvy);
  F[0] = t.xx * dSx + t.xy * dSy;
  F[1] = t.xy * dSx + t.yy * dSy;
  F[2] = (t.xx * dSx + t.xy * dSy) * u + (t.xy * dSx + t.yy * dSy) * v + k * (dTx * dSx + dTy * dSy);
  F[3] = tx * dSx + ty * dSy;
}

// SA boundary viscous flux f[0..3] (= reference f[1..4]),
// saBoundaryFaceViscFlux (sa.c:370-692), or with saCC
// saCC_BoundaryFaceViscFlux (saCC.c:247-589).  Gradients: u v T nu~; grx,
// gry: the cell's density gradient.
__device__ __forceinline__ void sa_bc_face(const KP& kp, int type, double dSx, double dSy, double nx, double ny,
                                           double dS2, double dx, double dy, const double* Pi, const double* gx,
                                           const double* gy, double grx, double gry, double* f,
                                           double* mie = nullptr) {
  double fv1, tx, ty;
  if (type == CSU_BC_SYMMETRY) {
    double dux = gx[0] - (gx[0] * nx + gy[0] * ny) * nx, duy = gy[0] - (gx[0] * nx + gy[0] * ny) * ny;
    double dvx = gx[1] - (gx[1] * nx + gy[1] * ny) * nx, dvy = gy[1] - (gx[1] * nx + gy[1] * ny) * ny;
    double dTx = gx[2] - (gx[2] * nx + gy[2] * ny) * nx, dTy = gy[2] - (gx[2] * nx + gy[2] * ny) * ny;
    double rho = Pi[0], u = Pi[1], v = Pi[2], T = Pi[4], n = Pi[5];
    double mi_L = sutherland(T), n_L = mi_L / rho;
    sa_calc_face(n, n_L, rho, 0.0, 0.0, fv1, tx, ty);  // tx, ty unused: f[3] = 0
    double mi_t = fv1 * rho * n, mi = mi_L + mi_t;
    if (mie) *mie = mi_t;
    double k = gas_T2Cp<double>(kp.gas, T) * (mi_L / kp.pr + mi_t / kp.pr_t);
    Tau t = tau_of(mi, dux, duy, dvx, dvy);
    f[0] = (t.xx * nx + t.xy * ny) * dS2;
    f[1] = (t.xy * nx + t.yy * ny) * dS2;
    f[2] = (u * (t.xx * nx + t.xy * ny) + v * (t.xy * nx + t.yy * ny) + k * (dTx * nx + dTy * ny)) * dS2;
    f[3] = 0.0;
    return;
  }
  if (type == CSU_BC_OUTLET) {
    double rho = Pi[0], u = Pi[1], v = Pi[2], T = Pi[4], n = Pi[5];
    double mi_L = sutherland(T), n_L = mi_L / rho;
    if (kp.sa_cc) sacc_calc_face(n, n_L, rho, gx[3], gy[3], grx, gry, fv1, tx, ty);
    else sa_calc_face(n, n_L, rho, gx[3], gy[3], fv1, tx, ty);
    double mi_t = fv1 * rho * n, mi = mi_L + mi_t;
    if (mie) *mie = mi_t;
    double k = gas_T2Cp<double>(kp.gas, T) * (mi_L / kp.pr + mi_t / kp.pr_t);
    Tau t = tau_of(mi, gx[0], gy[0], gx[1], gy[1]);
    f[0] = t.xx * dSx + t.xy * dSy;
    f[1] = t.xy * dSx + t.yy * dSy;
    f[2] = u * (t.xx * dSx + t.xy * dSy) + v * (t.xy * dSx + t.yy * dSy) + k * (gx[2] * dSx + gy[2] * dSy);
    f[3] = tx * dSx + ty * dSy;
    return;
  }
  const bool inl = type == CSU_BC_INLET;
  double L = sqrt(dx * dx + dy * dy);
  double dul = ((inl ? kp.pin[1] : 0.0) - Pi[1]) / L, dvl = ((inl ? kp.pin[2] : 0.0) - Pi[2]) / L;
  double dnl = ((inl ? kp.pin[5] : 0.0) - Pi[5]) / L;
  double dux, duy, dvx, dvy, dnx, dny, dTx = 0.0, dTy = 0.0;
  edge_correct(gx[0], gy[0], dul, dx, dy, L, dux, duy);
  edge_correct(gx[1], gy[1], dvl, dx, dy, L, dvx, dvy);
  if (type != CSU_BC_WALL) {
    double dTl = ((inl ? kp.pin[4] : kp.twall) - Pi[4]) / L;
    edge_correct(gx[2], gy[2], dTl, dx, dy, L, dTx, dTy);
  }
  edge_correct(gx[3], gy[3], dnl, dx, dy, L, dnx, dny);
  double rho = inl ? kp.pin[0] : Pi[0];
  double T = inl ? kp.pin[4] : (type == CSU_BC_WALLT ? kp.twall : Pi[4]);
  double n = inl ? kp.pin[5] : 0.0;
  double mi_L = sutherland(T), n_L = mi_L / rho;
  if (kp.sa_cc) {
    // density gradient per branch as saCC.c has it: one-sided to the
    // inlet density (:334-353), the cell's gradient at a wallT, and at a
    // wall the cell's d(rho)/dx in both components (:425-426)
    double drx = grx, dry = gry;
    if (inl) edge_correct(grx, gry, (kp.pin[0] - Pi[0]) / L, dx, dy, L, drx, dry);
    else if (type == CSU_BC_WALL) dry = grx;
    sacc_calc_face(n, n_L, rho, dnx, dny, drx, dry, fv1, tx, ty);
  } else {
    sa_calc_face(n, n_L, rho, dnx, dny, fv1, tx, ty);
  }
  double mi_t = fv1 * rho * n;
  double mi = type == CSU_BC_WALL ? mi_L : mi_L + mi_t;
  if (mie) *mie = type == CSU_BC_WALL ? 0.0 : mi_t;  // the wall reports 0 (sa.c:570)
  Tau t = tau_of(mi, dux, duy, dvx, dvy);
  f[0] = t.xx * dSx + t.xy * dSy;
  f[1] = t.xy * dSx + t.yy * dSy;
  f[3] = tx * dSx + ty * dSy;
  if (type == CSU_BC_WALL) { f[2] = 0.0; return; }
  double k = gas_T2Cp<double>(kp.gas, T) * (mi_L / kp.pr + mi_t / kp.pr_t);
  if (type == CSU_BC_WALLT) { f[2] = k * (dTx * dSx + dTy * dSy); return; }
  double u = kp.pin[1], v = kp.pin[2];
  f[2] = u * (t.xx * dSx + t.xy * dSy) + v * (t.xy * dSx + t.yy * dSy) + k * (dTx * dSx + dTy * dSy);
}

}  // namespace csu

namespace csu {

// ---------------------------------------------------------------------------
// Menter SST (reference sst.c)
// ---------------------------------------------------------------------------
struct SstV {
  double dux, duy, dvx, dvy, dkx, dky, dox, doy;
  double r, T, k, om, d, mi_L, mi_t;
  double tkx, tky, tox, toy, Qtk, Qto;
  double F1, F2;  // blending functions of the last sources evaluation
};
constexpr double SST_SK1 = 0.85, SST_SO1 = 0.5, SST_B1 = 0.075, SST_SK2 = 1.0, SST_SO2 = 0.856,
                 SST_B2 = 0.0828, SST_BS = 0.09, SST_A1 = 0.31;

__device__ __forceinline__ double sst_blend(double x1, double x2, double F1) { return x1 * F1 + x2 * (1 - F1); }

// sstF1 / sstF2 (sst.c:788-818); F1's long double temporaries only hold
// values computed in double.
__device__ __forceinline__ double sst_F1(const SstV& v, double n_L_term, double sqrtk_term) {
  double aux1 = 2 * v.r * SST_SO2 * (v.dkx * v.dox + v.dky * v.doy) / v.om;
  if (v.om == 0) aux1 = 0;
  double CD = fmax(aux1, 1.0e-20);
  aux1 = fmax(sqrtk_term, n_L_term);
  double aux2 = 4 * v.r * SST_SO2 * v.k / (CD * v.d * v.d);
  double arg1 = fmin(aux1, aux2);
  return tanh(arg1 * arg1 * arg1 * arg1);
}
__device__ __forceinline__ double sst_F2(double n_L_term, double sqrtk_term) {
  double arg2 = fmax(2 * sqrtk_term, n_L_term);
  return tanh(arg2 * arg2);
}

// The transition variant (sstTrans.c): its own constants, Ry = Rt = 1e6
// fixed (sstTrans.c:48,58,140,150), F1 without the omega = 0 guard.
constexpr double STR_SK1 = 0.85, STR_SO1 = 0.5, STR_B1 = 0.075, STR_SK2 = 1.0, STR_SO2 = 0.856, STR_B2 = 0.0828,
                 STR_BS2 = 0.09, STR_ALPHAS2 = 1.0, STR_ALPHA2 = 0.4403, STR_A1 = 0.31, STR_RY = 1e6, STR_RT = 1e6;

// sstTransF1 (sstTrans.c:96-115): long double temporaries holding doubles.
__device__ __forceinline__ double str_F1s(const SstV& v, double n_L_term, double sqrtk_term) {
  double aux1 = 2 * v.r * STR_SO2 * (v.dkx * v.dox + v.dky * v.doy) / v.om;
  double CD = fmax(aux1, 1.0e-20);
  aux1 = fmax(sqrtk_term, n_L_term);
  double aux2 = 4 * v.r * STR_SO2 * v.k / (CD * v.d * v.d);
  double arg1 = fmin(aux1, aux2);
  return tanh(arg1 * arg1 * arg1 * arg1);
}

// Blending functions of the transition variant, sstTransFlux /
// sstTransSources preamble (sstTrans.c:34-59, 126-151).
__device__ __forceinline__ void str_blend_fns(const SstV& v, double& omega, double& F1, double& F2) {
  double n_L = v.mi_L / v.r;
  omega = fabs(v.duy - v.dvx);
  double sqrtk = sqrt(v.k);
  double sqrtk_term = sqrtk / (STR_BS2 * v.om * v.d);
  double n_L_term = 500 * n_L / (v.d * v.d * v.om);
  double Fs1 = str_F1s(v, n_L_term, sqrtk_term);
  double aux = STR_RY / 120;
  aux = aux * aux;
  aux = aux * aux;
  aux = aux * aux;
  F1 = fmax(Fs1, exp(-aux));
  double arg2 = fmax(2 * sqrtk_term, n_L_term);
  F2 = tanh(arg2 * arg2);
}

// sstTransFlux (sstTrans.c:34-74).
__device__ __forceinline__ void str_flux(SstV& v) {
  double omega, F1, F2;
  str_blend_fns(v, omega, F1, F2);
  double alphas1 = (.025 + STR_RT / 6) / (1 + STR_RT / 6);
  double alphas = sst_blend(alphas1, STR_ALPHAS2, F1);
  double sk = sst_blend(STR_SK1, STR_SK2, F1), so = sst_blend(STR_SO1, STR_SO2, F1);
  v.mi_t = fmin(alphas * v.r * v.k / v.om, STR_A1 * v.r * v.k / (omega * F2));
  double aux = v.mi_L + sk * v.mi_t;
  v.tkx = aux * v.dkx;
  v.tky = aux * v.dky;
  aux = v.mi_L + so * v.mi_t;
  v.tox = aux * v.dox;
  v.toy = aux * v.doy;
}

// sstTransSources (sstTrans.c:126-174).
__device__ __forceinline__ void str_sources(SstV& v) {
  double omega, F1, F2;
  str_blend_fns(v, omega, F1, F2);
  double alpha1 = (5. / 9.) * (0.1 + STR_RT / 2.7) / (1 + STR_RT / 2.7);
  double alphas1 = (.025 + STR_RT / 6) / (1 + STR_RT / 6);
  double a8 = STR_RT / 8;
  a8 = a8 * a8;
  a8 = a8 * a8;
  double bs1 = 0.09 * (5. / 18. + a8) / (1 + a8);
  double alpha = sst_blend(alpha1, STR_ALPHA2, F1), alphas = sst_blend(alphas1, STR_ALPHAS2, F1);
  double b = sst_blend(STR_B1, STR_B2, F1), bs = sst_blend(bs1, STR_BS2, F1);
  v.mi_t = fmin(alphas * v.r * v.k / v.om, STR_A1 * v.r * v.k / (omega * F2));
  double n_t = v.mi_t / v.r;
  double P = v.mi_t * omega * omega;
  v.F1 = F1;
  v.F2 = F2;
  v.Qtk = P - bs * v.r * v.om * v.k;
  v.Qto = alpha * P / n_t - b * v.r * v.om * v.om + 2 * (1 - F1) * v.r * STR_SO2 * (v.dkx * v.dox + v.dky * v.doy) / v.om;
}

// sstFlux (sst.c:749-786): eddy viscosity, k and omega diffusion fluxes.
__device__ __forceinline__ void sst_flux(const KP& kp, SstV& v) {
  if (kp.sst_trans) {
    str_flux(v);
    return;
  }
  double n_L = v.mi_L / v.r;
  double omega = fabs(v.duy - v.dvx);
  double sqrtk_term = sqrt(v.k) / (SST_BS * v.om * v.d);
  double n_L_term = 500 * n_L / (v.d * v.d * v.om);
  double F1 = sst_F1(v, n_L_term, sqrtk_term);
  double F2 = sst_F2(n_L_term, sqrtk_term);
  v.mi_t = v.r * SST_A1 * v.k / fmax(SST_A1 * v.om, omega * F2);
  double sk = sst_blend(SST_SK1, SST_SK2, F1), so = sst_blend(SST_SO1, SST_SO2, F1);
  double aux = v.mi_L + sk * v.mi_t;
  v.tkx = aux * v.dkx;
  v.tky = aux * v.dky;
  aux = v.mi_L + so * v.mi_t;
  v.tox = aux * v.dox;
  v.toy = aux * v.doy;
}

// sstSources (sst.c:698-747): Qtk, Qto.
__device__ __forceinline__ void sst_sources(const KP& kp, SstV& v) {
  if (kp.sst_trans) {
    str_sources(v);
    return;
  }
  double n_L = v.mi_L / v.r;
  double omega = fabs(v.duy - v.dvx);
  double sqrtk_term = sqrt(v.k) / (SST_BS * v.om * v.d);
  double n_L_term = 500 * n_L / (v.d * v.d * v.om);
  double F1 = sst_F1(v, n_L_term, sqrtk_term);
  double F2 = sst_F2(n_L_term, sqrtk_term);
  v.mi_t = v.r * SST_A1 * v.k / fmax(SST_A1 * v.om, omega * F2);
  double n_t = v.mi_t / v.r;
  double gm = sst_blend(kp.sst_g1, kp.sst_g2, F1), b = sst_blend(SST_B1, SST_B2, F1);
  double txx = 2 * v.mi_t * (v.dux - (v.dux + v.dvy) / 3) - 2. * v.r * v.k / 3;
  double tyy = 2 * v.mi_t * (v.dvy - (v.dux + v.dvy) / 3) - 2. * v.r * v.k / 3;
  double txy = v.mi_t * (v.duy + v.dvx);
  double P = v.dux * txx + v.dvy * tyy + (v.duy + v.dvx) * txy;
  P = fmin(P, 10 * SST_BS * v.r * v.om * v.k);
  v.F1 = F1;
  v.F2 = F2;
  double cd = 2 * (1 - F1) * v.r * SST_SO2 * (v.dkx * v.dox + v.dky * v.doy) / v.om;
  v.Qtk = P - SST_BS * v.r * v.om * v.k;
  v.Qto = gm * P / n_t - b * v.r * v.om * v.om + cd;
}

// Stresses with the -2/3 rho k term (sst.c:163-165).
__device__ __forceinline__ Tau sst_tau(double mi, const SstV& v) {
  Tau t;
  t.xx = 2 * mi * (v.dux - (v.dux + v.dvy) / 3) - 2. * v.r * v.k / 3.;
  t.yy = 2 * mi * (v.dvy - (v.dux + v.dvy) / 3) - 2. * v.r * v.k / 3.;
  t.xy = mi * (v.duy + v.dvx);
  return t;
}

// sstInterSource (sst.c:595-691) for one cell: the two source terms times
// Omega and the six finite-difference source Jacobians of the LU-SGS block.
// g*: cell gradients of u v T k om.
__device__ __forceinline__ void sst_cell_source(const KP& kp, const double* Pi, double d, double om_i,
                                                const double* gx, const double* gy, double* src, double* dq) {
  SstV v;
  v.dux = gx[0]; v.dvx = gx[1]; v.dkx = gx[3]; v.dox = gx[4];
  v.duy = gy[0]; v.dvy = gy[1]; v.dky = gy[3]; v.doy = gy[4];
  v.r = Pi[0]; v.T = Pi[4]; v.k = Pi[5]; v.om = Pi[6]; v.d = d;
  v.mi_L = sutherland(v.T);
  if (v.om == 0) {  // omLim (solver.c:1155): no source, no Jacobians
    src[0] = 0.0; src[1] = 0.0;
#pragma unroll
    for (int q = 0; q < 6; q++) dq[q] = 0.0;
    return;
  }
  sst_sources(kp, v);
  src[0] = v.Qtk * om_i;
  src[1] = v.Qto * om_i;
  const double Qk0 = v.Qtk, Qo0 = v.Qto;
  const double h = 1e-6, dh = 1e-6;
  double hh = fmax(fabs(v.om * h), dh);
  v.om += hh;
  sst_sources(kp, v);
  const double dQkdo = (v.Qtk - Qk0) / hh, dQodo = (v.Qto - Qo0) / hh;
  v.om = Pi[6];
  hh = fmax(fabs(v.k * h), dh);
  v.k += hh;
  sst_sources(kp, v);
  const double dQkdk = (v.Qtk - Qk0) / hh, dQodk = (v.Qto - Qo0) / hh;
  v.k = Pi[5];
  hh = fmax(fabs(v.r * h), dh);
  v.r += hh;
  sst_sources(kp, v);
  const double dQkdr = (v.Qtk - Qk0) / hh, dQodr = (v.Qto - Qo0) / hh;
  v.r = Pi[0];
  dq[0] = (dQkdr - dQkdk * v.k / v.r - dQkdo * v.om / v.r) * om_i;
  dq[1] = dQkdk * om_i / v.r;
  dq[2] = dQkdo * om_i / v.r;
  dq[3] = (dQodr - dQodk * v.k / v.r - dQodo * v.om / v.r) * om_i;
  dq[4] = dQodk * om_i / v.r;
  dq[5] = dQodo * om_i / v.r;
}

// Interior SST face flux F[0..4] (= reference faceFlux[1..5]) and the face
// eddy viscosity, sstInterFace (sst.c:56-166).  Gradients: u v T k om.
// Edge correction along (dx, dy)/L in the form sst.c:121-143 uses.
__device__ __forceinline__ void sst_face(const KP& kp, double dSx, double dSy, double dx, double dy,
                                         const double* P0, const double* P1, double d0, double d1,
                                         const double* g0x, const double* g0y, const double* g1x,
                                         const double* g1y, double* F, double& mi_t) {
  const double L = sqrt(dx * dx + dy * dy);
  const double nx = dx / L, ny = dy / L;
  double dl[5] = {(P1[1] - P0[1]) / L, (P1[2] - P0[2]) / L, (P1[4] - P0[4]) / L, (P1[5] - P0[5]) / L,
                  (P1[6] - P0[6]) / L};
  double gxc[5], gyc[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const double gxm = 0.5 * (g0x[q] + g1x[q]), gym = 0.5 * (g0y[q] + g1y[q]);
    const double aux = gxm * nx + gym * ny;
    gxc[q] = gxm + (dl[q] - aux) * nx;
    gyc[q] = gym + (dl[q] - aux) * ny;
  }
  SstV v;
  v.dux = gxc[0]; v.duy = gyc[0]; v.dvx = gxc[1]; v.dvy = gyc[1];
  v.dkx = gxc[3]; v.dky = gyc[3]; v.dox = gxc[4]; v.doy = gyc[4];
  const double dTx = gxc[2], dTy = gyc[2];
  v.r = (P1[0] + P0[0]) * 0.5;
  const double u = (P1[1] + P0[1]) * 0.5, w = (P1[2] + P0[2]) * 0.5;
  v.T = (P1[4] + P0[4]) * 0.5;
  v.k = (P1[5] + P0[5]) * 0.5;
  v.om = (P1[6] + P0[6]) * 0.5;
  v.d = (d1 + d0) * 0.5;
  v.mi_L = sutherland(v.T);
  sst_flux(kp, v);
  const double mi = v.mi_L + v.mi_t;
  const double kk = gas_T2Cp<double>(kp.gas, v.T) * (v.mi_L / kp.pr + v.mi_t / kp.pr_t);
  mi_t = v.mi_t;
  Tau t = sst_tau(mi, v);
  F[0] = t.xx * dSx + t.xy * dSy;
  F[1] = t.xy * dSx + t.yy * dSy;
  F[2] = (t.xx * dSx + t.xy * dSy) * u + (t.xy * dSx + t.yy * dSy) * w + kk * (dTx * dSx + dTy * dSy);
  F[3] = v.tkx * dSx + v.tky * dSy;
  F[4] = v.tox * dSx + v.toy * dSy;
}

// SST boundary viscous flux f[0..4] (= reference f[1..5]),
// sstBoundaryFaceViscFlux (sst.c:202-547).  Gradients: u v T k om.
__device__ __forceinline__ void sst_bc_face(const KP& kp, int type, double dSx, double dSy, double nx, double ny,
                                            double dS2, double dx, double dy, const double* Pi, double d,
                                            const double* gx, const double* gy, double* f,
                                            double* mie = nullptr) {
  SstV v;
  v.d = d;
  if (type == CSU_BC_SYMMETRY) {
    v.dux = gx[0] - (gx[0] * nx + gy[0] * ny) * nx; v.duy = gy[0] - (gx[0] * nx + gy[0] * ny) * ny;
    v.dvx = gx[1] - (gx[1] * nx + gy[1] * ny) * nx; v.dvy = gy[1] - (gx[1] * nx + gy[1] * ny) * ny;
    v.dkx = gx[3] - (gx[3] * nx + gy[3] * ny) * nx; v.dky = gy[3] - (gx[3] * nx + gy[3] * ny) * ny;
    v.dox = gx[4] - (gx[4] * nx + gy[4] * ny) * nx; v.doy = gy[4] - (gx[4] * nx + gy[4] * ny) * ny;
    v.r = Pi[0]; v.T = Pi[4]; v.k = Pi[5]; v.om = Pi[6];
    v.mi_L = sutherland(v.T);
    sst_flux(kp, v);
    if (mie) *mie = v.mi_t;
    Tau t = sst_tau(v.mi_L + v.mi_t, v);
    const double u = Pi[1], w = Pi[2];
    f[0] = (t.xx * nx + t.xy * ny) * dS2;
    f[1] = (t.xy * nx + t.yy * ny) * dS2;
    f[2] = (u * (t.xx * nx + t.xy * ny) + w * (t.xy * nx + t.yy * ny)) * dS2;  // no heat flux
    f[3] = 0.0;
    f[4] = 0.0;
    return;
  }
  if (type == CSU_BC_INLET || type == CSU_BC_WALL || type == CSU_BC_WALLT) {
    const double L = sqrt(dx * dx + dy * dy);
    double ub, vb, Tb = 0.0, kb, ob;
    if (type == CSU_BC_INLET) {
      v.r = kp.pin[0]; v.T = kp.pin[4]; v.k = kp.pin[5]; v.om = kp.pin[6];
      ub = kp.pin[1]; vb = kp.pin[2]; Tb = kp.pin[4]; kb = kp.pin[5]; ob = kp.pin[6];
      v.mi_L = sutherland(v.T);
    } else {
      v.r = Pi[0]; v.T = Pi[4]; v.k = Pi[5]; v.om = Pi[6];
      v.mi_L = sutherland(v.T);
      const double n_L = v.mi_L / v.r;
      ub = 0.0; vb = 0.0; kb = 0.0;
      ob = kp.sst_owall * 6 * n_L / (SST_B1 * v.d * v.d);
      if (type == CSU_BC_WALLT) Tb = kp.twall;
    }
    edge_correct(gx[0], gy[0], (ub - Pi[1]) / L, dx, dy, L, v.dux, v.duy);
    edge_correct(gx[1], gy[1], (vb - Pi[2]) / L, dx, dy, L, v.dvx, v.dvy);
    double dTx = 0.0, dTy = 0.0;
    if (type != CSU_BC_WALL) edge_correct(gx[2], gy[2], (Tb - Pi[4]) / L, dx, dy, L, dTx, dTy);
    edge_correct(gx[3], gy[3], (kb - Pi[5]) / L, dx, dy, L, v.dkx, v.dky);
    edge_correct(gx[4], gy[4], (ob - Pi[6]) / L, dx, dy, L, v.dox, v.doy);
    sst_flux(kp, v);
    if (mie) *mie = v.mi_t;
    Tau t = sst_tau(v.mi_L + v.mi_t, v);
    f[0] = t.xx * dSx + t.xy * dSy;
    f[1] = t.xy * dSx + t.yy * dSy;
    f[3] = v.tkx * dSx + v.tky * dSy;
    f[4] = v.tox * dSx + v.toy * dSy;
    if (type == CSU_BC_WALL) { f[2] = 0.0; return; }
    const double kk = gas_T2Cp<double>(kp.gas, v.T) * (v.mi_L / kp.pr + v.mi_t / kp.pr_t);
    if (type == CSU_BC_WALLT) { f[2] = kk * (dTx * dSx + dTy * dSy); return; }
    f[2] = ub * (t.xx * dSx + t.xy * dSy) + vb * (t.xy * dSx + t.yy * dSy) + kk * (dTx * dSx + dTy * dSy);
    return;
  }
  // outlet
  v.dux = gx[0]; v.dvx = gx[1]; v.dkx = gx[3]; v.dox = gx[4];
  v.duy = gy[0]; v.dvy = gy[1]; v.dky = gy[3]; v.doy = gy[4];
  v.r = Pi[0]; v.T = Pi[4]; v.k = Pi[5]; v.om = Pi[6];
  v.mi_L = sutherland(v.T);
  sst_flux(kp, v);
  if (mie) *mie = v.mi_t;
  const double kk = gas_T2Cp<double>(kp.gas, v.T) * (v.mi_L / kp.pr + v.mi_t / kp.pr_t);
  Tau t = sst_tau(v.mi_L + v.mi_t, v);
  const double u = Pi[1], w = Pi[2];
  f[0] = t.xx * dSx + t.xy * dSy;
  f[1] = t.xy * dSx + t.yy * dSy;
  f[2] = u * (t.xx * dSx + t.xy * dSy) + w * (t.xy * dSx + t.yy * dSy) + kk * (gx[2] * dSx + gy[2] * dSy);
  f[3] = v.tkx * dSx + v.tky * dSy;
  f[4] = v.tox * dSx + v.toy * dSy;
}

}  // namespace csu

namespace csu {

// Wall shear tensor written to friction.csv (solverCalcCoeff2,
// solver.c:1303-1340): boundaryCalcTensorWall (boundary.c:744-793) for
// laminar runs, saCalcTensorWall (sa.c:795-852) for SA and SST.  Pi: cell
// primitives; g*: cell gradients of u, v; (dx, dy): face centre minus cell
// centre; xm: face centre x; d: the cell's wall distance.  out: txx txy tyy yp.
__device__ __forceinline__ void wall_tensor(bool turb, double dSx, double dSy, double dx, double dy, double xm,
                                            const double* Pi, double d, const double* gx, const double* gy,
                                            double* out) {
  const double L = sqrt(dx * dx + dy * dy);
  const double dul = (0 - Pi[1]) / L, dvl = (0 - Pi[2]) / L;
  double dux, duy, dvx, dvy;
  edge_correct(gx[0], gy[0], dul, dx, dy, L, dux, duy);
  edge_correct(gx[1], gy[1], dvl, dx, dy, L, dvx, dvy);
  const double mi = sutherland(Pi[4]);
  const double txx = 2 * mi * (dux - (dux + dvy) / 3);
  const double tyy = 2 * mi * (dvy - (dux + dvy) / 3);
  const double txy = mi * (duy + dvx);
  if (!turb) {
    const double aux = 5.29 * xm / sqrt(Pi[0] * Pi[1] * xm / mi);
    out[0] = txx; out[1] = txy; out[2] = tyy; out[3] = L / aux;
    return;
  }
  const double dS = sqrt(dSx * dSx + dSy * dSy);
  const double fx = (txx * dSx + txy * dSy) / dS;
  const double fy = (txy * dSx + tyy * dSy) / dS;
  const double ft = (fx * dSy - fy * dSx) / dS;
  const double ut = sqrt(fabs(ft) / Pi[0]);
  out[0] = 0.0; out[1] = ft; out[2] = 0.0; out[3] = Pi[0] * d * ut / mi;
}

}  // namespace csu
