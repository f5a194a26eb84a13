// kern_common.h -- kernel parameter blocks and per-cell / per-boundary-face
// routines shared by the residual kernels.
#pragma once
#include "csu.h"
#include "device_math.h"

namespace csu {

#define CSU_BLOCK 256
// A batch concatenates several meshes; every solve's cells, interior faces
// and boundary faces start on a multiple of CSU_CHUNK elements, so every
// chunk -- and every workgroup of up to CSU_CHUNK threads -- belongs to one
// solve, and a solve's per-block reduction partials are those of its own
// run alone.
#define CSU_CHUNK 256
// tile form of the face flux + update: the LDS face lists of a 256-cell
// tile (its cells' interior faces, those shared with other tiles included):
// the short one where every tile of the mesh fits (a triangle mesh in
// column order has ~513), else the long one
#define CSU_TILE_FACES_S 640
#define CSU_TILE_FACES 1024

// Scalar options and free-stream data, passed by value.  nc, nf, nb are the
// array strides (the whole batch); cend, fend, bend the end of the
// workgroup's solve (nc, nf, nb for a single solve).
struct KP {
  int nc, nf, nb, nv, order, flux, laminar, sa, sst, limiter, turb1order, axi, visc_blazek;
  int cend, fend, bend;
  int xcd;        // 1: cell and face tiles mapped XCD-contiguously (xcd_tile); 0: in workgroup order
  int lsq_ld;     // 1: long double LSQ sums emulated per call; 0: precomputed weights
  int mixed;      // 1: FP32 gradients, limiter, reconstruction and convective flux
  int store_grad; // write dPx/dPy/phi even when no viscous term reads them
  int sst_trans;  // SST: the transition variant's eddy viscosity and sources (sstTrans.c)
  int sa_cc;      // SA: the saCC.c form, with density-gradient terms in the nu~ flux and source
  double rlim, pout, twall, efix, minf, limK, volmax, pr, pr_t, wimp;
  double sst_owall, sst_g1, sst_g2;  // SST: oWallFactor, blended gamma constants (sst.c:18-38)
  double pin[7];
  GasD gas;
};

// Workgroup -> work-tile mapping that gives each of the 8 XCDs (one L2 each)
// a contiguous range of tiles.  Workgroups are dispatched round-robin over
// the XCDs, so physical block b runs on XCD b % 8; with this map XCD x
// processes tiles [x*per, (x+1)*per) and neighbour gathers stay in its L2.
// The tail (nblocks % 8 blocks) maps to itself.  A bijection on [0, nblocks).
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
  const int per = nblocks >> 3;
  if (b >= (per << 3)) return b;
  return (b & 7) * per + (b >> 3);
}

// Reductions: within a 64-lane wavefront by lane exchanges (a butterfly:
// after each step the two partners hold the same value, since a + b and
// b + a round alike, so every lane ends with the same result and the
// combination order is fixed); across a workgroup's wavefronts by one LDS
// round, combined in wavefront order.  Deterministic for sums; minima and
// maxima are exact in any order.
struct RedSum {
  __device__ double operator()(double a, double b) const { return a + b; }
};
struct RedMin {
  __device__ double operator()(double a, double b) const { return fmin(a, b); }
};
struct RedMax {
  __device__ double operator()(double a, double b) const { return fmax(a, b); }
};

template <class Op>
__device__ __forceinline__ double wave_reduce(double v, Op op) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
  return v;
}

// NR values per thread over a workgroup of NT threads (every thread must
// call it); sw: LDS scratch of NR x NT/64 doubles.  Every thread receives
// the NR results.
template <int NT, int NR, class Op>
__device__ __forceinline__ void block_reduce(double (&v)[NR], double (*sw)[NT / 64], Op op) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int r = 0; r < NR; r++) v[r] = wave_reduce(v[r], op);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int r = 0; r < NR; r++) sw[r][threadIdx.x >> 6] = v[r];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NR; r++) {
    double a = sw[r][0];
#pragma unroll
    for (int w = 1; w < NW; w++) a = op(a, sw[r][w]);
    v[r] = a;
  }
}

// Batch: solve of the workgroup whose first element is e0 (uniform across
// the workgroup), from a per-chunk table.
__device__ __forceinline__ int solve_of(const int* solk, int e0) { return solk[e0 / CSU_CHUNK]; }

// One solve of a batch: its element ranges in the concatenated arrays and
// what differs between solves (free stream, outlet pressure, wall
// temperature, AUSM+up reference Mach, limiter volume scale, CFL, reference
// length of the force coefficients).
struct SolveP {
  int c0, c1, f0, f1, b0, b1, w0, w1;  // [begin, end): cells, faces, boundary faces, wall-list entries
  int active;                          // 0: the state is held (a converged or finished solve)
  int pad_;
  double cfl, sref, pout, twall, minf, volmax;
  double pin[7];
};

// Batch: the parameters of solve q in the workgroup's copy of kp.
__device__ __forceinline__ void solve_kp(KP& kp, const SolveP& q) {
  kp.cend = q.c1; kp.fend = q.f1; kp.bend = q.b1;
  kp.pout = q.pout; kp.twall = q.twall; kp.minf = q.minf; kp.volmax = q.volmax;
#pragma unroll
  for (int k = 0; k < 7; k++) kp.pin[k] = q.pin[k];
}

// Geometry (device pointers, read-only), passed by value.
struct Geo {
  // batch (sol non-null): per-solve table, the solve of every CSU_CHUNK
  // cells / faces / boundary faces, and of every cell
  const SolveP* sol;
  const int *csolk, *fsolk, *bsolk, *csol;
  const int *cnp, *cnbr, *cface;                        // [nc], [4][nc], [4][nc]
  const int* cfn;  // implicit schemes: [4][nc] the other cell of face slot j, -1 where cface is 0
  const int *toff, *tface;  // tile form: faces of tile t are tface[toff[t] .. toff[t+1])
  const short* tsl;         // tile form: [4][nc] +-(1 + the slot's face's place in its tile's list), 0: none
  const double *xc, *yc, *omega, *dsxm, *dsym, *dslat;  // [nc]
  const double *edx, *edy;                              // [4][nc] own-edge midpoint offsets
  const double *lwx, *lwy;                              // [4][nc] LSQ weights per neighbour slot
  const unsigned char* fsl;                             // [nf] edge of the face in e0 (bits 0-1), e1 (2-3)
  const unsigned char* bmask;                           // [nc] bit j: neighbour slot j is a boundary face
  const float *lwx32, *lwy32, *edx32, *edy32;           // mixed precision copies of lwx, lwy, edx, edy
  const float *fdsx32, *fdsy32, *fds32;                 // mixed precision face vectors
  const double *wdist;                                  // [nc] wall distance (SA/SST)
  const int *fe0, *fe1;                                 // [nf]
  const double *fdsx, *fdsy, *fds;
  const double *fnx, *fny, *fdd;  // LU-SGS: [nf] dS/|dS| and the distance of the face's cell centres
  const int *be0, *btype, *bdonor;                      // [nb]
  const double *bdsx, *bdsy, *bds, *bxm, *bym, *bnx, *bny, *bds2;
};

// Mutable per-stage arrays, passed by value.
struct St {
  double *P, *Pb, *dPx, *dPy, *phi, *dTx, *dTy;
  double *fflux;        // convective face flux times dS, per face [nf][NV] (one gather per face in k_update)
  double *vflux, *bvflux, *wfric, *miT;
  double *pref2;        // limiter reference values [solves][6]
  double *src;          // SA: Qt*Omega [nc]; SST: Qtk*Omega, Qto*Omega [2][nc]
  double *dq;           // SST source Jacobians dQk/d(r, rk, ro), dQo/d(r, rk, ro) [6][nc]
  double *rec;          // second order: reconstructed state at each cell edge [4][NV][nc]
  float *rec32;         // mixed precision: the same in FP32
  float *fflux32;       // mixed precision: convective face fluxes in FP32 [nf][NV]
  double *wten;         // viscous: wall shear tensor txx txy tyy yp per boundary face [4][nb]
  double *bflux;        // convective boundary-face flux times dS [NV][nb]
  const double *tube;   // time-accurate run: t, tmax, state (0 running, 1 last step, 2 past it), steps
  const double *Us;     // primitives-from-state stages: the stage's conservative state [NV][nc]
};

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

// Ghost state of a boundary face from the interior primitive Pi (nv+1
// slots), reference boundaryCalcPrimitive (boundary.c:514-693).
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
  if (NV == 6) {  // SST k, omega (boundary.c:660-690); walls: omega from the wall distance
    if (type == CSU_BC_SYMMETRY || type == CSU_BC_OUTLET) { Pg[5] = Pi[5]; Pg[6] = Pi[6]; }
    else if (type == CSU_BC_INLET) { Pg[5] = kp.pin[5]; Pg[6] = kp.pin[6]; }
    else {
      const double n = sutherland(Pg[4]) / Pg[0];
      Pg[5] = 1e-14;
      Pg[6] = kp.sst_owall * 6 * n / (0.075 * wall_d * wall_d);
    }
  }
  (void)wall_d;
}

// Convective boundary flux of one boundary face, reference boundaryCalc /
// boundaryCalc_sa (boundary.c:143-341).  Adds f*dS to R when dS > 0.
template <int NV>
__device__ __forceinline__ void bc_flux(const KP& kp, int type, double dSx, double dSy, double dS,
                                        const double* Pi, double* R) {
  double PL[NV], Pb[NV], fl[NV];
  PL[0] = Pi[0]; PL[1] = Pi[1]; PL[2] = Pi[2]; PL[3] = Pi[3];
#pragma unroll
  for (int q = 4; q < NV; q++) PL[q] = Pi[q + 1];
  constexpr int NSC = NV - 4;
  const double kL = NV == 6 ? PL[4] : 0.0;  // SST: the k of both states enters the energy
  if (type == CSU_BC_SYMMETRY) {
    rotate<double>(PL[1], PL[2], dSx, dSy, dS);
    double Pm[NV];
    Pm[0] = PL[0]; Pm[1] = -PL[1]; Pm[2] = PL[2]; Pm[3] = PL[3];
#pragma unroll
    for (int q = 4; q < NV; q++) Pm[q] = PL[q];
    flux_dispatch<double, NSC>(kp.gas, kp.flux, kp.efix, kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4,
                               Pm[0], Pm[1], Pm[2], Pm[3], Pm + 4, kL, kL, fl);
  } else if (type == CSU_BC_INLET || type == CSU_BC_OUTLET) {
    if (type == CSU_BC_INLET) bc_inlet(kp.gas, kp.pin, PL, Pb, dSx / dS, dSy / dS);
    else bc_outlet(kp.gas, kp.pout, PL, Pb, dSx / dS, dSy / dS);
#pragma unroll
    for (int q = 4; q < NV; q++) Pb[q] = (type == CSU_BC_INLET) ? kp.pin[q + 1] : PL[q];
    rotate<double>(PL[1], PL[2], dSx, dSy, dS);
    rotate<double>(Pb[1], Pb[2], dSx, dSy, dS);
    flux_dispatch<double, NSC>(kp.gas, kp.flux, kp.efix, kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4,
                               Pb[0], Pb[1], Pb[2], Pb[3], Pb + 4, kL, NV == 6 ? Pb[4] : 0.0, fl);
  } else {
    if (kp.laminar == 1 || NV >= 5) {
      rotate<double>(PL[1], PL[2], dSx, dSy, dS);
      double Pm[NV];
      Pm[0] = PL[0]; Pm[1] = -PL[1]; Pm[2] = -PL[2]; Pm[3] = PL[3];
#pragma unroll
      for (int q = 4; q < NV; q++) Pm[q] = PL[q];
      flux_dispatch<double, NSC>(kp.gas, kp.flux, kp.efix, kp.minf, PL[0], PL[1], PL[2], PL[3], PL + 4,
                                 Pm[0], Pm[1], Pm[2], Pm[3], Pm + 4, kL, kL, fl);
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

// Local time step of one cell, reference solverLocalTimeStep + solverCalcVel
// (solver.c:895-923, 420-439), from the unclamped conservative state.
template <int NV>
__device__ __forceinline__ double cell_dt(const KP& kp, const double* u, double om, double sx, double sy) {
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

}  // namespace csu
