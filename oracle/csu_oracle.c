/*
 * csu_oracle.c -- CPU restatement of the reference's residual-and-update
 * path, used as the parity oracle for the HIP kernels.
 *
 * TEST INFRASTRUCTURE: loaded only by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py, never by the product.
 *
 * Written from the reference's algorithm, function by function (citations
 * are file:line in the reference), with its operation order and its x87
 * long double least-squares sums, so that on x86-64 it reproduces the
 * reference's FP64 results bit for bit.  It is pinned against the reference
 * built from source (oracle/_ref/csu_ref_trace) by the tests.  Data are SoA
 * over the mesh description of csu.h instead of the reference's per-element
 * structs.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "csu_oracle.h"

typedef struct {
    double gamma, R, Cp, Cv, cc[6];
    int tp;
} ogas;

struct csu_oracle {
    int nc, nf, nb, nv, nn, axi, iter, dt_local;
    csu_params prm;
    ogas gas;
    double *xy;
    int *cv, *face, *bface, *btype, *nbr, *cface;
    double *omega, *U, *Uaux, *R, *dPx, *dPy, *phi, *flux, *dtL, *P, *Pb;
    double dt, volMax, Pref20[6];
    /* implicit schemes */
    double *D, *dW0, *dW1, *A;    /* A: DPLUR blocks [nc][4][nv*nv] */
    double *wd, *miT;             /* SA/SST: wall distance [nc], face eddy viscosity [nf] */
    /* SST: source Jacobians of the LU-SGS 2x2 block (sst.c:664-671) */
    double *dQkdr, *dQkdrk, *dQkdro, *dQodr, *dQodrk, *dQodro;
    double sst_g1, sst_g2;        /* sstInit, sst.c:18-38 */
};

/* ---------------- gas model: gasprop.c:44-274 ---------------- */
static void gas_init(ogas* g, double gamma, double R, int tp)
{
    g->gamma = gamma; g->R = R;
    g->Cp = g->gamma * g->R / (g->gamma - 1);     /* gasprop.c:50 */
    g->Cv = g->Cp - g->R;
    g->tp = tp;
    g->cc[0] = 1.81716e3; g->cc[1] = 9.890504e2 - 287.0530; g->cc[2] = -9.595592e-3;
    g->cc[3] = 1.041469e-4; g->cc[4] = -4.433065e-8; g->cc[5] = 5.879263e-12;
}

static double gas_T2e(const ogas* g, double T)           /* gasprop.c:77-99 */
{
    if (!g->tp) return g->Cv * T;
    double P = 1, a = g->cc[0];
    for (int i = 1; i < 6; i++) { P *= T; a += g->cc[i] * P; }
    return a;
}

static double gas_T2Cv(const ogas* g, double T)          /* gasprop.c:101-124 */
{
    if (!g->tp) return g->Cv;
    double P = 1, a = g->cc[1];
    for (int i = 2; i < 6; i++) { P *= T; a += i * g->cc[i] * P; }
    return a;
}

static void gas_T2eCv(const ogas* g, double T, double* e, double* Cv)   /* gasprop.c:126-142 */
{
    double P = 1;
    *e = g->cc[0]; *Cv = g->cc[1];
    for (int i = 1; i < 6; i++) {
        P *= T;
        *e += g->cc[i] * P;
        if (i + 1 < 6) *Cv += (i + 1) * g->cc[i + 1] * P;
    }
}

static double gas_e2T(const ogas* g, double e)           /* gasprop.c:144-166 */
{
    double T = e / g->Cv;
    if (g->tp) {
        double err = 1, e1, Cv;
        int i = 0;
        while ((fabs(err) > 1e-14) && (i < 8)) {
            gas_T2eCv(g, T, &e1, &Cv);
            err = (e - e1) / Cv;
            T += err;
            i++;
        }
    }
    return T;
}

static double gas_T2c(const ogas* g, double T)           /* gasprop.c:168-184 */
{
    double Cv = gas_T2Cv(g, T);
    return sqrt(((Cv + g->R) / Cv) * g->R * T);
}

static double gas_T2Cp(const ogas* g, double T) { return gas_T2Cv(g, T) + g->R; }   /* gasprop.c:174-177 */

/* SST model constants, sstInit (sst.c:18-38) */
#define SST_SK1 0.85
#define SST_SO1 0.5
#define SST_B1 0.075
#define SST_SK2 1.0
#define SST_SO2 0.856
#define SST_B2 0.0828
#define SST_BS 0.09
#define SST_A1 0.31

static double sutherland(double T) { return 1.458e-6 * T * sqrt(T) / (T + 110.4); }  /* solver.c:1222-1225 */

static double gas_critic_H2c(const ogas* g, double H)    /* gasprop.c:219-274 */
{
    double c2 = 2 * (g->gamma - 1) * H / (g->gamma + 1);
    if (!g->tp) return sqrt(c2);
    double T = c2 / (g->gamma * g->R), e, Cv, H0, dH, err;
    int i = 0;
    for (int first = 1;; first = 0) {
        if (!first && !((fabs(err) > 1.0e-12) && (i < 3))) break;
        gas_T2eCv(g, T, &e, &Cv);
        double RT = g->R * T, gam = (Cv + g->R) / Cv;
        H0 = e + RT + gam * RT * 0.5;
        dH = Cv + g->R + gam * g->R * 0.5;
        err = (H - H0) / dH;
        T += err;
        if (!first) i++;
    }
    return gas_T2c(g, T);
}

/* ---------------- geometry: mesh.c:371-441, 720-740 ---------------- */
static int nvert(const csu_oracle* o, int i) { return o->cv[4 * i + 3] >= 0 ? 4 : 3; }

static void cell_center(const csu_oracle* o, int i, double* x, double* y)
{
    int n = nvert(o, i);
    *x = 0.0; *y = 0.0;
    for (int k = 0; k < n; k++) { *x += o->xy[2 * o->cv[4 * i + k]]; *y += o->xy[2 * o->cv[4 * i + k] + 1]; }
    *x /= n; *y /= n;
}

static void bface_center(const csu_oracle* o, int g, double* x, double* y)
{
    int p0 = o->bface[3 * g + 1], p1 = o->bface[3 * g + 2];
    *x = o->xy[2 * p0] + o->xy[2 * p1]; *y = o->xy[2 * p0 + 1] + o->xy[2 * p1 + 1];
    *x /= 2; *y /= 2;
}

/* centre of neighbour slot j of cell i (interior cell or boundary element) */
static void nbr_center(const csu_oracle* o, int i, int j, double* x, double* y)
{
    int n = o->nbr[4 * i + j];
    if (n >= 0) cell_center(o, n, x, y); else bface_center(o, -1 - n, x, y);
}

static void face_ds(const csu_oracle* o, int p0, int p1, double* dSx, double* dSy)
{
    double x0 = o->xy[2 * p0], y0 = o->xy[2 * p0 + 1], x1 = o->xy[2 * p1], y1 = o->xy[2 * p1 + 1];
    *dSx = y1 - y0;
    *dSy = -(x1 - x0);
    double y = (y0 + y1) / 2;
    if (o->axi == 1) { *dSx = (*dSx) * y; *dSy = (*dSy) * y; }
}

static double tri_omega(const csu_oracle* o, int a, int b, int c)
{
    double x1 = o->xy[2 * a], x2 = o->xy[2 * b], x3 = o->xy[2 * c];
    double y1 = o->xy[2 * a + 1], y2 = o->xy[2 * b + 1], y3 = o->xy[2 * c + 1];
    return 0.5 * ((x1 - x2) * (y1 + y2) + (x2 - x3) * (y2 + y3) + (x3 - x1) * (y3 + y1));
}

static double cell_omega(const csu_oracle* o, int i)
{
    const int* v = o->cv + 4 * i;
    double a = tri_omega(o, v[0], v[1], v[2]);
    if (nvert(o, i) == 4) a += tri_omega(o, v[2], v[3], v[0]);
    if (o->axi == 1) { double x, y; cell_center(o, i, &x, &y); a *= y; }
    return a;
}

static void rotation(double* U, double dSx, double dSy, double dS)   /* solver.c:441-448 */
{
    double aux = (U[1] * dSx + U[2] * dSy) / dS;
    U[2] = (-U[1] * dSy + U[2] * dSx) / dS;
    U[1] = aux;
}

/* ---------------- convective fluxes: flux.c ---------------- */
static void entropy_fix(double efix, double* l)          /* flux.c:48-57 */
{
    if ((*l < efix) && (*l > -efix)) *l = 0.5 * (*l * *l / efix + efix);
}

static void flux_roe(const csu_oracle* o, const double* L, const double* Rt, double* f)   /* flux.c:59-129 */
{
    const double g1 = o->gas.gamma - 1;
    double rL = L[0], uL = L[1], vL = L[2], pL = L[3], rR = Rt[0], uR = Rt[1], vR = Rt[2], pR = Rt[3];
    double U1L = rL * uL, U2L = rL * vL, U3L = pL / g1 + (uL * uL + vL * vL) * rL / 2;
    double HL = (U3L + pL) / rL;
    double U1R = rR * uR, U2R = rR * vR, U3R = pR / g1 + (uR * uR + vR * vR) * rR / 2;
    double HR = (U3R + pR) / rR;
    double sL = sqrt(rL), sR = sqrt(rR);
    double ub = (sL * uL + sR * uR) / (sL + sR);
    double vb = (sL * vL + sR * vR) / (sL + sR);
    double Hb = (sL * HL + sR * HR) / (sL + sR);
    double ab = sqrt(g1 * (Hb - (ub * ub + vb * vb) / 2));
    double l1 = ub - ab, l2 = ub, l4 = ub, l5 = ub + ab;
    double k1[4] = {1.0, ub - ab, vb, Hb - ub * ab};
    double k2[4] = {0.0, 0.0, 1.0, vb};
    double k4[4] = {1.0, ub, vb, 0.5 * (ub * ub + vb * vb)};
    double k5[4] = {1.0, ub + ab, vb, Hb + ub * ab};
    double d1 = rR - rL, d2 = U1R - U1L, d3 = U2R - U2L, d5 = U3R - U3L;
    double a4 = (Hb - (ub * ub + vb * vb)) * d1 + ub * d2 + vb * d3 - d5;
    a4 /= (ab * ab) / g1;
    double a2 = d3 - d1 * vb;
    double a5 = ((d1 - a4) + (d2 - ub * d1) / ab) * 0.5;
    double a1 = ((d1 - a4) - (d2 - ub * d1) / ab) * 0.5;
    double FL[4] = {U1L, U1L * uL + pL, U1L * vL, uL * (U3L + pL)};
    double FR[4] = {U1R, U1R * uR + pR, U1R * vR, uR * (U3R + pR)};
    entropy_fix(o->prm.efix, &l1); entropy_fix(o->prm.efix, &l2);
    entropy_fix(o->prm.efix, &l4); entropy_fix(o->prm.efix, &l5);
    for (int i = 0; i < 4; i++)
        f[i] = 0.5 * (FR[i] + FL[i] - a1 * fabs(l1) * k1[i] - a2 * fabs(l2) * k2[i]
                      - a4 * fabs(l4) * k4[i] - a5 * fabs(l5) * k5[i]);
}

/* enthalpy and temperature from a primitive state (flux.c:138-144 pattern) */
/* total enthalpy; kt is the turbulent-energy term the _sst fluxes add to
   rho*E (k*rho, or k alone in fluxAUSMDV_sst, flux.c:993); 0 otherwise,
   which leaves the sum unchanged */
static void HT(const csu_oracle* o, double r, double u, double v, double p, double kt, double* H, double* T)
{
    *T = p / (o->gas.R * r);
    double U3 = gas_T2e(&o->gas, *T) * r + (u * u + v * v) * r / 2 + kt;
    *H = (U3 + p) / r;
}

/* L/R: r u v p [s0 [s1]]: ns passive scalars (SA nu~; SST k, omega) */
static void flux_ausm(const csu_oracle* o, const double* L, const double* Rt, int ns, double* f)  /* flux.c:132-203, 512-585, 901-972 */
{
    double HL, TL, HR, TR;
    HT(o, L[0], L[1], L[2], L[3], ns == 2 ? L[4] * L[0] : 0.0, &HL, &TL);
    HT(o, Rt[0], Rt[1], Rt[2], Rt[3], ns == 2 ? Rt[4] * Rt[0] : 0.0, &HR, &TR);
    double cL = gas_T2c(&o->gas, TL), cR = gas_T2c(&o->gas, TR);
    double ML = L[1] / cL, MR = Rt[1] / cR, Mp, Mm_, Pp, Pm;
    if (fabs(ML) >= 1) { Mp = 0.5 * (ML + fabs(ML)); Pp = Mp / ML; }
    else { double a = 0.25 * (ML + 1) * (ML + 1); Mp = a; Pp = a * (2 - ML); }
    if (fabs(MR) >= 1) { Mm_ = 0.5 * (MR - fabs(MR)); Pm = Mm_ / MR; }
    else { double a = -0.25 * (MR - 1) * (MR - 1); Mm_ = a; Pm = -a * (2 + MR); }
    double M = Mp + Mm_;
    double pm = Pp * L[3] + Pm * Rt[3];
    const double* S = M > 0 ? L : Rt;
    double c = M > 0 ? cL : cR, H = M > 0 ? HL : HR;
    f[0] = S[0] * c * M;
    f[1] = f[0] * S[1] + pm;
    f[2] = f[0] * S[2];
    f[3] = f[0] * H;
    for (int q = 0; q < ns; q++) f[4 + q] = f[0] * S[4 + q];
}

static void flux_ausmdv(const csu_oracle* o, const double* L, const double* Rt, int ns, double* f)  /* flux.c:206-287, 588-670, 975-1058 */
{
    double rL = L[0], uL = L[1], vL = L[2], pL = L[3], rR = Rt[0], uR = Rt[1], vR = Rt[2], pR = Rt[3];
    double HL, TL, HR, TR, aux;
    HT(o, rL, uL, vL, pL, ns == 2 ? L[4] : 0.0, &HL, &TL);
    HT(o, rR, uR, vR, pR, ns == 2 ? Rt[4] : 0.0, &HR, &TR);
    double cL = gas_T2c(&o->gas, TL), cR = gas_T2c(&o->gas, TR);
    double cm = fmax(cL, cR);
    double aL = (2.0 * pL / rL) / (pL / rL + pR / rR);
    double aR = (2.0 * pR / rR) / (pL / rL + pR / rR);
    double uLp, pLp, uRm, pRm;
    aux = 0.5 * (uL + fabs(uL));
    if (fabs(uL) < cm) {
        uLp = aL * (0.25 * (uL + cm) * (uL + cm) / cm - aux) + aux;
        pLp = 0.25 * pL * (uL + cm) * (uL + cm) * (2.0 - uL / cm) / (cm * cm);
    } else { uLp = aux; pLp = pL * aux / uL; }
    aux = 0.5 * (uR - fabs(uR));
    if (fabs(uR) < cm) {
        uRm = aR * (-0.25 * (uR - cm) * (uR - cm) / cm - aux) + aux;
        pRm = 0.25 * pR * (uR - cm) * (uR - cm) * (2.0 + uR / cm) / (cm * cm);
    } else { uRm = aux; pRm = pR * aux / uR; }
    double rU = uLp * rL + uRm * rR;
    int n = 4 + ns;
    f[0] = rU;
    f[1] = pLp + pRm;
    f[2] = 0.5 * (rU * (vR + vL) - fabs(rU) * (vR - vL));
    f[3] = 0.5 * (rU * (HR + HL) - fabs(rU) * (HR - HL));
    for (int q = 0; q < ns; q++) f[4 + q] = 0.5 * (rU * (Rt[4 + q] + L[4 + q]) - fabs(rU) * (Rt[4 + q] - L[4 + q]));
    double fD = 0.5 * (rU * (uR + uL) - fabs(rU) * (uR - uL));
    double fV = uLp * rL * uL + uRm * rR * uR;
    double s = 0.5 * fmin(1, 10 * fabs(pR - pL) / fmin(pL, pR));
    f[1] += (0.5 + s) * fV + (0.5 - s) * fD;
    int A = (uL - cL < 0.0) && (uR - cR > 0.0);
    int B = (uL + cL < 0.0) && (uR + cR > 0.0);
    double psL[6] = {1.0, uL, vL, HL, ns ? L[4] : 0.0, ns > 1 ? L[5] : 0.0};
    double psR[6] = {1.0, uR, vR, HR, ns ? Rt[4] : 0.0, ns > 1 ? Rt[5] : 0.0};
    if (A && !B) {
        aux = 0.125 * ((uR - cR) - (uL - cL));
        for (int k = 0; k < n; k++) f[k] -= aux * (rR * psR[k] - rL * psL[k]);
    } else if (!A && B) {
        aux = 0.125 * ((uR + cR) - (uL + cL));
        for (int k = 0; k < n; k++) f[k] -= aux * (rR * psR[k] - rL * psL[k]);
    }
}

static void flux_ausmpup(const csu_oracle* o, const double* L, const double* Rt, int ns, int v2, double* f)
{
    /* flux.c:290-398 (v2 == 0) and 401-508 (v2 == 1); _sa 673-899; _sst 1061-1297
       (the _sst forms take the critical speed of sound from H - k) */
    double rL = L[0], uL = L[1], vL = L[2], pL = L[3], rR = Rt[0], uR = Rt[1], vR = Rt[2], pR = Rt[3];
    double HL, TL, HR, TR;
    const double kL = ns == 2 ? L[4] : 0.0, kR = ns == 2 ? Rt[4] : 0.0;
    HT(o, rL, uL, vL, pL, kL * rL, &HL, &TL);
    HT(o, rR, uR, vR, pR, kR * rR, &HR, &TR);
    double as = gas_critic_H2c(&o->gas, HL - kL);
    double ahL = as * as / fmax(as, fabs(uL));
    as = gas_critic_H2c(&o->gas, HR - kR);
    double ahR = as * as / fmax(as, fabs(uR));
    double am = fmin(ahL, ahR);
    double Kp = 0.25, Ku = 0.75, sig = 1.0, beta = 1.0 / 8.0;
    double ML = uL / am, MR = uR / am;
    double Mb = sqrt((uL * uL + uR * uR) / (2 * am * am));
    double Minf = o->prm.mach_inf;
    double M0 = sqrt(fmin(1, fmax(Mb * Mb, Minf * Minf)));
    double fa = M0 * (2 - M0);
    if (fa < 1e-3) fa = 1e-3;
    double alpha = (3.0 / 16.0) * (-4 + 5 * fa * fa);
    double rhom = 0.5 * (rL + rR);
    double Mp, Mn, Pp, Pn, a, b;
    if (fabs(ML) >= 1) { Mp = 0.5 * (ML + fabs(ML)); Pp = Mp / ML; }
    else {
        a = 0.25 * (ML + 1) * (ML + 1); b = -0.25 * (ML - 1) * (ML - 1);
        Mp = a * (1 - 16 * beta * b);
        Pp = a * ((2 - ML) - 16 * alpha * ML * b);
    }
    if (fabs(MR) >= 1) { Mn = 0.5 * (MR - fabs(MR)); Pn = Mn / MR; }
    else {
        a = 0.25 * (MR + 1) * (MR + 1); b = -0.25 * (MR - 1) * (MR - 1);
        Mn = b * (1 + 16 * beta * a);
        Pn = b * ((-2 - MR) + 16 * alpha * MR * a);
    }
    double Mm = Mp + Mn - (Kp / fa) * fmax(1 - sig * Mb * Mb, 0) * (pR - pL) / (rhom * am * am);
    double pm;
    if (!v2) pm = Pp * pL + Pn * pR - Ku * Pp * Pn * (rL + rR) * (fa * am) * (uR - uL);
    else pm = 0.5 * (pL + pR) + 0.5 * (Pp - Pn) * (pL - pR)
              + sqrt(0.5 * (uL * uL + vL * vL + uR * uR + vR * vR)) * (Pp + Pn - 1) * 0.5 * (pL + pR) / am;
    double mm = Mm > 0 ? am * Mm * rL : am * Mm * rR;
    const double* S = mm > 0 ? L : Rt;
    double H = mm > 0 ? HL : HR;
    f[0] = mm; f[1] = mm * S[1] + pm; f[2] = mm * S[2]; f[3] = mm * H;
    for (int q = 0; q < ns; q++) f[4 + q] = mm * S[4 + q];
}

/* Roe with turbulence scalars (no reference counterpart: flux_sa / flux_sst
   have no Roe branch, flux.c:1325-1366): Roe's mean-flow flux, the scalars
   carried by its mass flux from the upwind side (f[0] > 0: left), and for
   SST k added to the total-energy flux the same way. */
static void roe_scalars(const double* L, const double* Rt, int ns, double* f)
{
    const double* S = f[0] > 0 ? L : Rt;
    for (int q = 0; q < ns; q++) f[4 + q] = f[0] * S[4 + q];
    if (ns == 2) f[3] += f[0] * S[4];
}

static void flux(const csu_oracle* o, const double* L, const double* Rt, double* f)  /* flux.c:1300-1366 */
{
    int ns = o->nv - 4;
    switch (o->prm.flux) {
        case 0: flux_roe(o, L, Rt, f); if (ns > 0) roe_scalars(L, Rt, ns, f); break;
        case 1: flux_ausm(o, L, Rt, ns, f); break;
        case 2: flux_ausmdv(o, L, Rt, ns, f); break;
        case 3: flux_ausmpup(o, L, Rt, ns, 0, f); break;
        default: flux_ausmpup(o, L, Rt, ns, 1, f); break;
    }
}

static void flux_free(const csu_oracle* o, const double* P, double* f)   /* flux.c:1369-1380 */
{
    double T = P[3] / (o->gas.R * P[0]);
    double U3 = gas_T2e(&o->gas, T) * P[0] + (P[1] * P[1] + P[2] * P[2]) * P[0] / 2;
    double H = (U3 + P[3]) / P[0];
    f[0] = P[0] * P[1];
    f[1] = P[0] * P[1] * P[1] + P[3];
    f[2] = P[0] * P[1] * P[2];
    f[3] = P[0] * P[1] * H;
}

/* ---------------- boundary states: boundary.c:17-141 ---------------- */
static void bc_inlet(const csu_oracle* o, const double* Pa, const double* Pd, double* Pb, double nx, double ny)
{
    double T = Pd[3] / Pd[0] / o->gas.R;
    double c0 = gas_T2c(&o->gas, T);
    double m = sqrt(Pd[1] * Pd[1] + Pd[2] * Pd[2]) / c0;
    if (m < 1.) {
        double pb = 0.5 * (Pa[3] + Pd[3] - Pd[0] * c0 * (nx * (Pa[1] - Pd[1]) + ny * (Pa[2] - Pd[2])));
        Pb[0] = Pa[0] + (pb - Pa[3]) / (c0 * c0);
        Pb[1] = Pa[1] - nx * (Pa[3] - pb) / (Pd[0] * c0);
        Pb[2] = Pa[2] - ny * (Pa[3] - pb) / (Pd[0] * c0);
        Pb[3] = pb;
    } else {
        for (int k = 0; k < 4; k++) Pb[k] = Pa[k];
    }
}

static void bc_outlet(const csu_oracle* o, const double* Pd, double* Pb, double nx, double ny)
{
    double T = Pd[3] / Pd[0] / o->gas.R;
    double c0 = gas_T2c(&o->gas, T);
    double m = sqrt(Pd[1] * Pd[1] + Pd[2] * Pd[2]) / c0;
    if (m < 1.) {
        double pb = o->prm.pout;
        Pb[0] = Pd[0] + (pb - Pd[3]) / (c0 * c0);
        Pb[1] = Pd[1] + nx * (Pd[3] - pb) / (Pd[0] * c0);
        Pb[2] = Pd[2] + ny * (Pd[3] - pb) / (Pd[0] * c0);
        Pb[3] = pb;
    } else {
        for (int k = 0; k < 4; k++) Pb[k] = Pd[k];
    }
}

static void bc_wall(const csu_oracle* o, const double* Pd, double* Pb, double nx, double ny)
{
    double T = Pd[3] / Pd[0] / o->gas.R;
    double c0 = gas_T2c(&o->gas, T);
    double pb = Pd[3] + Pd[0] * c0 * (nx * Pd[1] + ny * Pd[2]);
    Pb[0] = Pd[0] + (pb - Pd[3]) / (c0 * c0);
    Pb[1] = Pd[1] - nx * (nx * Pd[1] + ny * Pd[2]);
    Pb[2] = Pd[2] - ny * (nx * Pd[1] + ny * Pd[2]);
    Pb[3] = pb;
}

/* ---------------- primitive recovery: solver.c:1153-1220 ----------------
   the cells, then the boundary ghost states (boundaryCalcPrimitive, called
   at the end of solverCalcPrimitive) */
static void calc_primitive(csu_oracle* o, double* U)
{
    const int nc = o->nc, nv = o->nv, ns = nv + 1;
#pragma omp parallel for
    for (int i = 0; i < nc; i++) {
        double* P = o->P + (size_t)i * ns;
        P[0] = fmax(U[i], o->prm.rlim);
        U[i] = P[0];
        double k = 0;
        if (o->prm.sst == 1) {
            k = U[4 * nc + i] / P[0];
            if (k < 1e-14) { k = 1e-14; U[4 * nc + i] = P[0] * k; }
            P[5] = k;
            double om = U[5 * nc + i] / P[0];
            if (om <= 0.0) { om = 0.0; U[5 * nc + i] = P[0] * om; }
            P[6] = om;
        }
        double u = U[nc + i] / P[0], v = U[2 * nc + i] / P[0];
        P[1] = u; P[2] = v;
        double e = U[3 * nc + i] / P[0] - 0.5 * (u * u + v * v) - k;
        e = fmax(e, 7.176325e+02);
        double T = gas_e2T(&o->gas, e);
        if (T < 1) {
            T = 1;
            e = gas_T2e(&o->gas, T);
            U[3 * nc + i] = (e + 0.5 * (u * u + v * v) + k) * P[0];
        }
        P[4] = T;
        P[3] = o->gas.R * P[4] * P[0];
        if (o->prm.sa == 1) {
            P[5] = U[4 * nc + i] / P[0];
            P[5] = fmax(P[5], 0);
            U[4 * nc + i] = P[5] * P[0];
        }
    }
    /* boundary ghost states, boundary.c:514-693 (marker order = g order) */
    for (int g = 0; g < o->nb; g++) {
        const int e0 = o->bface[3 * g];
        const double* E = o->P + (size_t)e0 * ns;
        double* B = o->Pb + (size_t)g * ns;
        double dSx, dSy, dS, PL[4] = {E[0], E[1], E[2], E[3]}, Q[4];
        face_ds(o, o->bface[3 * g + 1], o->bface[3 * g + 2], &dSx, &dSy);
        dS = sqrt(dSx * dSx + dSy * dSy);
        int t = o->btype[g];
        if (t == CSU_BC_SYMMETRY) {
            if (dS > 0) { rotation(PL, dSx, dSy, dS); PL[1] = 0.0; rotation(PL, dSx, -dSy, dS); }
            else PL[2] = 0.0;
            for (int k = 0; k < 4; k++) B[k] = PL[k];
            B[4] = B[3] / (B[0] * o->gas.R);
        } else if (t == CSU_BC_INLET) {
            bc_inlet(o, o->prm.pin, PL, Q, dSx / dS, dSy / dS);
            for (int k = 0; k < 4; k++) B[k] = Q[k];
            B[4] = B[3] / (B[0] * o->gas.R);
        } else if (t == CSU_BC_OUTLET) {
            bc_outlet(o, PL, Q, dSx / dS, dSy / dS);
            for (int k = 0; k < 4; k++) B[k] = Q[k];
            B[4] = B[3] / (B[0] * o->gas.R);
        } else if (o->prm.laminar == 1 || o->prm.sa == 1) {   /* SST walls take the inviscid branch */
            if (t == CSU_BC_WALLT) {
                B[0] = PL[3] / (o->gas.R * o->prm.twall); B[1] = 0.0; B[2] = 0.0; B[3] = PL[3]; B[4] = o->prm.twall;
            } else {
                B[0] = PL[0]; B[1] = 0.0; B[2] = 0.0; B[3] = PL[3]; B[4] = B[3] / (B[0] * o->gas.R);
            }
        } else {
            if (dS > 0) { bc_wall(o, PL, Q, dSx / dS, dSy / dS); for (int k = 0; k < 4; k++) B[k] = Q[k]; }
            else { PL[2] = 0.0; for (int k = 0; k < 4; k++) B[k] = PL[k]; }
            B[4] = B[3] / (B[0] * o->gas.R);
        }
        if (o->prm.sa == 1) {
            if (t == CSU_BC_SYMMETRY || t == CSU_BC_OUTLET) B[5] = E[5];
            else if (t == CSU_BC_INLET) B[5] = o->prm.pin[5];
            else B[5] = 0.0;
        }
        if (o->prm.sst == 1) {        /* boundary.c:660-690 */
            if (t == CSU_BC_SYMMETRY || t == CSU_BC_OUTLET) { B[5] = E[5]; B[6] = E[6]; }
            else if (t == CSU_BC_INLET) { B[5] = o->prm.pin[5]; B[6] = o->prm.pin[6]; }
            else {
                double n = sutherland(B[4]) / B[0], d = o->wd[e0];
                B[5] = 1e-14;
                B[6] = o->prm.sst_owallfactor * 6 * n / (SST_B1 * d * d);
            }
        }
    }
}

/* neighbour j of cell i: primitive slot m */
static double nbr_P(const csu_oracle* o, int i, int j, int m)
{
    int n = o->nbr[4 * i + j], ns = o->nv + 1;
    return n >= 0 ? o->P[(size_t)n * ns + m] : o->Pb[(size_t)(-1 - n) * ns + m];
}

/* ---------------- gradients: solver.c:451-474, 1031-1074 ---------------- */
static void grad_lsq(const csu_oracle* o, int i, int m, double* dUx, double* dUy)
{
    double x[5], y[5], u[5], xx, yy;
    int n = nvert(o, i), ns = o->nv + 1;
    cell_center(o, i, &xx, &yy);
    x[0] = xx; y[0] = yy; u[0] = o->P[(size_t)i * ns + m];
    for (int j = 0; j < n; j++) {
        nbr_center(o, i, j, &xx, &yy);
        x[j + 1] = xx; y[j + 1] = yy; u[j + 1] = nbr_P(o, i, j, m);
    }
    long double a = 0, b = 0, c = 0, A = 0, B = 0;
    for (int j = 1; j < n + 1; j++) {
        a += (x[j] - x[0]) * (x[j] - x[0]);
        b += (x[j] - x[0]) * (y[j] - y[0]);
        c += (y[j] - y[0]) * (y[j] - y[0]);
        A += (x[j] - x[0]) * (u[j] - u[0]);
        B += (y[j] - y[0]) * (u[j] - u[0]);
    }
    *dUx = (c * A - b * B) / (a * c - b * b);
    *dUy = (-b * A + a * B) / (a * c - b * b);
}

/* Green-Gauss gradient (gradient 1; no counterpart in the reference, which
   has least squares only): sum over the cell's faces of (u_f - u_i) S_f /
   Omega, u_f the average of the two sides (ghost state on boundary faces),
   S_f the face vector out of the cell; the face vectors of a closed cell
   sum to zero, so this equals sum u_f S_f / Omega.  Planar runs only. */
static void grad_gg(const csu_oracle* o, int i, int m, double* dUx, double* dUy)
{
    const int n = nvert(o, i), ns = o->nv + 1;
    const double ui = o->P[(size_t)i * ns + m];
    double ax = 0.0, ay = 0.0;
    for (int j = 0; j < n; j++) {
        const int nb = o->nbr[4 * i + j];
        double sx, sy;
        if (nb >= 0) {
            const int f = o->cface[4 * i + j], fi = f > 0 ? f - 1 : -f - 1;
            face_ds(o, o->face[4 * fi + 2], o->face[4 * fi + 3], &sx, &sy);
            if (f < 0) { sx = -sx; sy = -sy; }
        } else {
            face_ds(o, o->bface[3 * (-1 - nb) + 1], o->bface[3 * (-1 - nb) + 2], &sx, &sy);
        }
        const double wx = 0.5 * sx / o->omega[i], wy = 0.5 * sy / o->omega[i];
        const double du = nbr_P(o, i, j, m) - ui;
        if (j == 0) { ax = wx * du; ay = wy * du; }
        else { ax += wx * du; ay += wy * du; }
    }
    *dUx = ax;
    *dUy = ay;
}

static void grad_any(const csu_oracle* o, int i, int m, double* dUx, double* dUy)
{
    if (o->prm.gradient == 1) grad_gg(o, i, m, dUx, dUy);
    else grad_lsq(o, i, m, dUx, dUy);
}

static void update_grad(csu_oracle* o)
{
#pragma omp parallel for
    for (int i = 0; i < o->nc; i++)
        for (int k = 0; k < o->nv; k++) {
            double gx, gy;
            grad_any(o, i, k > 3 ? k + 1 : k, &gx, &gy);
            o->dPx[(size_t)k * o->nc + i] = gx;
            o->dPy[(size_t)k * o->nc + i] = gy;
        }
}

/* ---------------- limiter: limiter.c:34-147 ---------------- */
static double limiter_v2(double Ui, double Umin, double Umax, double d2, double ee)
{
    double ans, dmax = Umax - Ui, dmin = Umin - Ui;
    if (d2 == 0) return 1;
    if (d2 > 0) {
        ans = (dmax * dmax + ee) * d2 + 2 * d2 * d2 * dmax;
        ans /= dmax * dmax + 2 * d2 * d2 + dmax * d2 + ee;
    } else {
        ans = (dmin * dmin + ee) * d2 + 2 * d2 * d2 * dmin;
        ans /= dmin * dmin + 2 * d2 * d2 + dmin * d2 + ee;
    }
    return ans / d2;
}

/* Barth-Jespersen, limiterBJ (solver.c:1132-1151): defined in the
   reference but never called by its limiter loop; limiter 2 here. */
static double limiter_bj(double Ui, double Umin, double Umax, double d2)
{
    if (d2 == 0) return 1;
    if (d2 > 0) return fmin(1, (Umax - Ui) / d2);
    return fmin(1, (Umin - Ui) / d2);
}

static void limiter_update(csu_oracle* o)
{
    const int ns = o->nv + 1;
    if (o->prm.limiter == 0) {
        o->volMax = 0.0;
        for (int i = 0; i < o->nc; i++) o->volMax = fmax(o->omega[i], o->volMax);
        const double* Pin = o->prm.pin;
        o->Pref20[0] = Pin[0] * Pin[0];
        o->Pref20[1] = Pin[1] * Pin[1] + Pin[2] * Pin[2];
        o->Pref20[2] = o->Pref20[1];
        o->Pref20[3] = Pin[3] * Pin[3];
        for (int k = 4; k < o->nv; k++) o->Pref20[k] = Pin[k + 1] * Pin[k + 1];
    } else if (o->prm.limiter == 1) {
        double mn[6], mx[6];
        for (int j = 0; j < o->nv; j++) {
            int m = j == 4 ? 5 : j;
            mn[j] = mx[j] = o->P[m];
        }
        for (int i = 1; i < o->nc; i++)
            for (int j = 0; j < o->nv; j++) {
                int m = j == 4 ? 5 : j;
                mn[j] = fmin(mn[j], o->P[(size_t)i * ns + m]);
                mx[j] = fmax(mx[j], o->P[(size_t)i * ns + m]);
            }
        for (int j = 0; j < o->nv; j++) {
            double a = (mx[j] - mn[j]) * o->prm.lim_k;
            o->Pref20[j] = a * a;
        }
    }
}

/* ---------------- inter(): solver.c:493-668 ---------------- */
static void inter(csu_oracle* o)
{
    const int nc = o->nc, nf = o->nf, nv = o->nv, ns = nv + 1;
    if (o->prm.order == 2) {
        limiter_update(o);
#pragma omp parallel for
        for (int i = 0; i < nc; i++) {
            double x0, y0, pref2[6];
            cell_center(o, i, &x0, &y0);
            if (o->prm.limiter == 0) {
                double a = o->prm.lim_k * sqrt(o->omega[i] / o->volMax);
                a = a * a * a;
                for (int k = 0; k < nv; k++) pref2[k] = o->Pref20[k] * a;
            } else {
                for (int k = 0; k < nv; k++) pref2[k] = o->Pref20[k];
            }
            int n = nvert(o, i);
            for (int k = 0; k < nv; k++) {
                int m = k > 3 ? k + 1 : k;
                double Pi = o->P[(size_t)i * ns + m], mn = Pi, mx = Pi, ph = 0;
                for (int j = 0; j < n; j++) {
                    mx = fmax(mx, nbr_P(o, i, j, m));
                    mn = fmin(mn, nbr_P(o, i, j, m));
                }
                for (int j = 0; j < n; j++) {
                    int p0 = o->cv[4 * i + j], p1 = o->cv[4 * i + (j + 1) % n];
                    double xm = (o->xy[2 * p0] + o->xy[2 * p1]) * 0.5;
                    double ym = (o->xy[2 * p0 + 1] + o->xy[2 * p1 + 1]) * 0.5;
                    double d2 = o->dPx[(size_t)k * nc + i] * (xm - x0) + o->dPy[(size_t)k * nc + i] * (ym - y0);
                    double p = o->prm.limiter == 2 ? limiter_bj(Pi, mn, mx, d2) : limiter_v2(Pi, mn, mx, d2, pref2[k]);
                    ph = j == 0 ? p : fmin(ph, p);
                }
                o->phi[(size_t)k * nc + i] = ph;
            }
        }
    }
#pragma omp parallel for
    for (int f = 0; f < nf; f++) {
        const int* q = o->face + 4 * f;
        int e0 = q[0], e1 = q[1];
        double dSx, dSy, dS, L[6] = {0}, Rr[6] = {0}, fl[6];
        face_ds(o, q[2], q[3], &dSx, &dSy);
        dS = sqrt(dSx * dSx + dSy * dSy);
        const double* P0 = o->P + (size_t)e0 * ns;
        const double* P1 = o->P + (size_t)e1 * ns;
        if (o->prm.order == 2) {
            double xm = (o->xy[2 * q[2]] + o->xy[2 * q[3]]) * 0.5, ym = (o->xy[2 * q[2] + 1] + o->xy[2 * q[3] + 1]) * 0.5;
            double x0, y0, x1, y1;
            cell_center(o, e0, &x0, &y0);
            cell_center(o, e1, &x1, &y1);
            for (int k = 0; k < nv; k++) {
                int m = k > 3 ? k + 1 : k;
                if (k > 3 && o->prm.turb1order) { L[k] = P0[m]; Rr[k] = P1[m]; continue; }
                L[k] = P0[m] + (o->dPx[(size_t)k * nc + e0] * (xm - x0) + o->dPy[(size_t)k * nc + e0] * (ym - y0)) * o->phi[(size_t)k * nc + e0];
                Rr[k] = P1[m] + (o->dPx[(size_t)k * nc + e1] * (xm - x1) + o->dPy[(size_t)k * nc + e1] * (ym - y1)) * o->phi[(size_t)k * nc + e1];
            }
        } else {
            for (int k = 0; k < nv; k++) { int m = k > 3 ? k + 1 : k; L[k] = P0[m]; Rr[k] = P1[m]; }
        }
        rotation(L, dSx, dSy, dS);
        rotation(Rr, dSx, dSy, dS);
        flux(o, L, Rr, fl);
        rotation(fl, dSx, -dSy, dS);
        for (int k = 0; k < nv; k++) o->flux[(size_t)k * nf + f] = fl[k] * dS;
    }
#pragma omp parallel for
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < 4; j++) {
            int f = o->cface[4 * i + j];
            if (f > 0) for (int k = 0; k < nv; k++) o->R[(size_t)k * nc + i] += o->flux[(size_t)k * nf + f - 1];
            else if (f < 0) for (int k = 0; k < nv; k++) o->R[(size_t)k * nc + i] -= o->flux[(size_t)k * nf - f - 1];
        }
}

/* ---------------- boundary fluxes: boundary.c:143-464 ---------------- */
static void boundary(csu_oracle* o)
{
    const int nc = o->nc, nv = o->nv, ns = nv + 1;
    for (int g = 0; g < o->nb; g++) {
        const int e0 = o->bface[3 * g];
        const double* E = o->P + (size_t)e0 * ns;
        double dSx, dSy, dS, L[6] = {E[0], E[1], E[2], E[3], nv > 4 ? E[5] : 0.0, nv > 5 ? E[6] : 0.0};
        double B[6], M[6], fl[6] = {0};
        face_ds(o, o->bface[3 * g + 1], o->bface[3 * g + 2], &dSx, &dSy);
        dS = sqrt(dSx * dSx + dSy * dSy);
        int t = o->btype[g];
        if (t == CSU_BC_SYMMETRY) {
            rotation(L, dSx, dSy, dS);
            M[0] = L[0]; M[1] = -L[1]; M[2] = L[2]; M[3] = L[3]; M[4] = L[4]; M[5] = L[5];
            flux(o, L, M, fl);
        } else if (t == CSU_BC_INLET || t == CSU_BC_OUTLET) {
            if (t == CSU_BC_INLET) bc_inlet(o, o->prm.pin, L, B, dSx / dS, dSy / dS);
            else bc_outlet(o, L, B, dSx / dS, dSy / dS);
            B[4] = t == CSU_BC_INLET ? o->prm.pin[5] : L[4];
            B[5] = t == CSU_BC_INLET ? o->prm.pin[6] : L[5];
            rotation(L, dSx, dSy, dS);
            rotation(B, dSx, dSy, dS);
            flux(o, L, B, fl);
        } else if (o->prm.laminar == 1 || nv > 4) {
            rotation(L, dSx, dSy, dS);
            M[0] = L[0]; M[1] = -L[1]; M[2] = -L[2]; M[3] = L[3]; M[4] = L[4]; M[5] = L[5];
            flux(o, L, M, fl);
        } else {
            bc_wall(o, L, B, dSx / dS, dSy / dS);
            rotation(L, dSx, dSy, dS);
            rotation(B, dSx, dSy, dS);
            flux_free(o, B, fl);
        }
        rotation(fl, dSx, -dSy, dS);
        if (dS > 0)
            for (int k = 0; k < nv; k++) o->R[(size_t)k * nc + e0] += fl[k] * dS;
    }
}

/* interAxisPressure, solver.c:670-683 (lateral area: mesh.c:404-418) */
static void axis_pressure(csu_oracle* o)
{
    const int nc = o->nc, ns = o->nv + 1;
    for (int i = 0; i < nc; i++) {
        const int* v = o->cv + 4 * i;
        double dS = tri_omega(o, v[0], v[1], v[2]);
        if (nvert(o, i) == 4) dS += tri_omega(o, v[2], v[3], v[0]);
        const double* P = o->P + (size_t)i * ns;
        o->R[2 * (size_t)nc + i] -= P[3] * dS;
        if (o->prm.sst) o->R[2 * (size_t)nc + i] -= (2. / 3.) * P[0] * P[5] * dS;
    }
}

/* solverGrad_T, solver.c:477-490: T gradient into slot 3 */
static void grad_T(csu_oracle* o)
{
#pragma omp parallel for
    for (int i = 0; i < o->nc; i++) {
        double gx, gy;
        grad_any(o, i, 4, &gx, &gy);
        o->dPx[3 * (size_t)o->nc + i] = gx;
        o->dPy[3 * (size_t)o->nc + i] = gy;
    }
}

/* face stress/heat flux with the edge-direction correction, laminar.c:33-74 */
static void laminar_face(const csu_oracle* o, int f, double* F)
{
    const int nc = o->nc, ns = o->nv + 1;
    const int* q = o->face + 4 * f;
    int e0 = q[0], e1 = q[1];
    const double *P0 = o->P + (size_t)e0 * ns, *P1 = o->P + (size_t)e1 * ns;
    const double *gx = o->dPx, *gy = o->dPy;
    double dSx, dSy, x0, y0, x1, y1;
    face_ds(o, q[2], q[3], &dSx, &dSy);
    double duxm = (gx[nc + e0] + gx[nc + e1]) * 0.5, dvxm = (gx[2 * nc + e0] + gx[2 * nc + e1]) * 0.5;
    double dTxm = (gx[3 * nc + e0] + gx[3 * nc + e1]) * 0.5;
    double duym = (gy[nc + e0] + gy[nc + e1]) * 0.5, dvym = (gy[2 * nc + e0] + gy[2 * nc + e1]) * 0.5;
    double dTym = (gy[3 * nc + e0] + gy[3 * nc + e1]) * 0.5;
    cell_center(o, e0, &x0, &y0);
    cell_center(o, e1, &x1, &y1);
    double dx = x1 - x0, dy = y1 - y0, L = sqrt(dx * dx + dy * dy);
    double dul = (P1[1] - P0[1]) / L, dvl = (P1[2] - P0[2]) / L, dTl = (P1[4] - P0[4]) / L;
    double dux = duxm + (dul - (duxm * dx + duym * dy) / L) * dx / L;
    double duy = duym + (dul - (duxm * dx + duym * dy) / L) * dy / L;
    double dvx = dvxm + (dvl - (dvxm * dx + dvym * dy) / L) * dx / L;
    double dvy = dvym + (dvl - (dvxm * dx + dvym * dy) / L) * dy / L;
    double dTx = dTxm + (dTl - (dTxm * dx + dTym * dy) / L) * dx / L;
    double dTy = dTym + (dTl - (dTxm * dx + dTym * dy) / L) * dy / L;
    double T = (P1[4] + P0[4]) * 0.5, mi = sutherland(T);
    double k = gas_T2Cp(&o->gas, T) * mi / o->prm.pr;
    double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
    double u = (P1[1] + P0[1]) * 0.5, v = (P1[2] + P0[2]) * 0.5;
    F[1] = txx * dSx + txy * dSy;
    F[2] = txy * dSx + tyy * dSy;
    F[3] = (txx * dSx + txy * dSy) * u + (txy * dSx + tyy * dSy) * v + k * (dTx * dSx + dTy * dSy);
}

/* laminarInter, laminar.c:15-99 */
static void laminar_inter(csu_oracle* o)
{
    const int nc = o->nc, nf = o->nf;
#pragma omp parallel for
    for (int f = 0; f < nf; f++) {
        double F[4];
        laminar_face(o, f, F);
        for (int k = 1; k < 4; k++) o->flux[(size_t)k * nf + f] = F[k];
    }
#pragma omp parallel for
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < 4; j++) {
            int f = o->cface[4 * i + j];
            if (f > 0) for (int k = 1; k < 4; k++) o->R[(size_t)k * nc + i] -= o->flux[(size_t)k * nf + f - 1];
            else if (f < 0) for (int k = 1; k < 4; k++) o->R[(size_t)k * nc + i] += o->flux[(size_t)k * nf - f - 1];
        }
}

/* wall-corrected velocity gradients of boundary face g, laminar.c:230-257 */
static void wall_grads(const csu_oracle* o, int g, double ub, double vb, double* du, double* dv,
                       double* L_out, double* dxy)
{
    const int nc = o->nc, ns = o->nv + 1, e0 = o->bface[3 * g];
    const double* P = o->P + (size_t)e0 * ns;
    double x0, y0, x1, y1;
    cell_center(o, e0, &x0, &y0);
    bface_center(o, g, &x1, &y1);
    double dx = x1 - x0, dy = y1 - y0, L = sqrt(dx * dx + dy * dy);
    double dul = (ub - P[1]) / L, dvl = (vb - P[2]) / L;
    double duxm = o->dPx[nc + e0], dvxm = o->dPx[2 * nc + e0], duym = o->dPy[nc + e0], dvym = o->dPy[2 * nc + e0];
    du[0] = duxm + (dul - (duxm * dx + duym * dy) / L) * dx / L;
    du[1] = duym + (dul - (duxm * dx + duym * dy) / L) * dy / L;
    dv[0] = dvxm + (dvl - (dvxm * dx + dvym * dy) / L) * dx / L;
    dv[1] = dvym + (dvl - (dvxm * dx + dvym * dy) / L) * dy / L;
    *L_out = L; dxy[0] = dx; dxy[1] = dy;
}

/* boundaryFaceViscFlux, laminar.c:128-346 (f[3] of an adiabatic wall is 0:
   u = v = 0 there and no heat flux; the reference leaves it unset) */
static void laminar_bc_flux(const csu_oracle* o, int g, double* f)
{
    const int nc = o->nc, ns = o->nv + 1, e0 = o->bface[3 * g], t = o->btype[g];
    const int p0 = o->bface[3 * g + 1], p1 = o->bface[3 * g + 2];
    const double* P = o->P + (size_t)e0 * ns;
    double dux, duy, dvx, dvy, dTx, dTy, dSx, dSy;
    f[3] = 0.0;
    if (t == CSU_BC_SYMMETRY) {
        double x0 = o->xy[2 * p0], y0 = o->xy[2 * p0 + 1], x1 = o->xy[2 * p1], y1 = o->xy[2 * p1 + 1];
        double sx = y1 - y0, sy = -(x1 - x0), dS = sqrt(sx * sx + sy * sy);    /* meshCalcDS2 */
        double nx = sx / dS, ny = sy / dS;
        if (o->axi == 1) dS = dS * (y0 + y1) * 0.5;
        double duxm = o->dPx[nc + e0], dvxm = o->dPx[2 * nc + e0], dTxm = o->dPx[3 * nc + e0];
        double duym = o->dPy[nc + e0], dvym = o->dPy[2 * nc + e0], dTym = o->dPy[3 * nc + e0];
        dux = duxm - (duxm * nx + duym * ny) * nx; duy = duym - (duxm * nx + duym * ny) * ny;
        dvx = dvxm - (dvxm * nx + dvym * ny) * nx; dvy = dvym - (dvxm * nx + dvym * ny) * ny;
        dTx = dTxm - (dTxm * nx + dTym * ny) * nx; dTy = dTym - (dTxm * nx + dTym * ny) * ny;
        double T = P[4], mi = sutherland(T), k = gas_T2Cp(&o->gas, T) * mi / o->prm.pr;
        double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
        double u = P[1], v = P[2];
        f[1] = (txx * nx + txy * ny) * dS;
        f[2] = (txy * nx + tyy * ny) * dS;
        f[3] = (u * (txx * nx + txy * ny) + v * (txy * nx + tyy * ny) + k * (dTx * nx + dTy * ny)) * dS;
        return;
    }
    face_ds(o, p0, p1, &dSx, &dSy);
    if (t == CSU_BC_INLET || t == CSU_BC_WALL || t == CSU_BC_WALLT) {
        const double* Pin = o->prm.pin;
        double du[2], dv[2], L, dxy[2];
        wall_grads(o, g, t == CSU_BC_INLET ? Pin[1] : 0, t == CSU_BC_INLET ? Pin[2] : 0, du, dv, &L, dxy);
        dux = du[0]; duy = du[1]; dvx = dv[0]; dvy = dv[1];
        double T = t == CSU_BC_WALLT ? o->prm.twall : P[4];
        double mi = sutherland(T);
        double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
        f[1] = txx * dSx + txy * dSy;
        f[2] = txy * dSx + tyy * dSy;
        if (t == CSU_BC_WALL) return;
        double Tb = t == CSU_BC_INLET ? Pin[4] : o->prm.twall;
        double dTl = (Tb - P[4]) / L, dx = dxy[0], dy = dxy[1];
        double dTxm = o->dPx[3 * nc + e0], dTym = o->dPy[3 * nc + e0];
        dTx = dTxm + (dTl - (dTxm * dx + dTym * dy) / L) * dx / L;
        dTy = dTym + (dTl - (dTxm * dx + dTym * dy) / L) * dy / L;
        double k = gas_T2Cp(&o->gas, T) * mi / o->prm.pr;
        if (t == CSU_BC_WALLT) { f[3] = k * (dTx * dSx + dTy * dSy); return; }
        double u = P[1], v = P[2];
        f[3] = u * (txx * dSx + txy * dSy) + v * (txy * dSx + tyy * dSy) + k * (dTx * dSx + dTy * dSy);
        return;
    }
    /* outlet: the cell's own gradients */
    dux = o->dPx[nc + e0]; dvx = o->dPx[2 * nc + e0]; dTx = o->dPx[3 * nc + e0];
    duy = o->dPy[nc + e0]; dvy = o->dPy[2 * nc + e0]; dTy = o->dPy[3 * nc + e0];
    double T = P[4], mi = sutherland(T), k = gas_T2Cp(&o->gas, T) * mi / o->prm.pr;
    double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
    double u = P[1], v = P[2];
    f[1] = txx * dSx + txy * dSy;
    f[2] = txy * dSx + tyy * dSy;
    f[3] = u * (txx * dSx + txy * dSy) + v * (txy * dSx + tyy * dSy) + k * (dTx * dSx + dTy * dSy);
}

/* laminarBoundary, laminar.c:101-125.  The reference's adiabatic-wall
   branch (laminar.c:230-269) never writes f[3], and f[] lives in the frame
   of laminarBoundaryVisc, so a wall face subtracts the energy flux of the
   boundary face processed before it (the last face of the preceding marker
   that writes one).  Restated as such; with no such face before it (wall
   marker first) the reference reads an undefined stack value, here 0. */
static void laminar_boundary(csu_oracle* o)
{
    const int nc = o->nc;
    double carry = 0.0;
    for (int g = 0; g < o->nb; g++) {
        double f[4];
        laminar_bc_flux(o, g, f);
        if (o->btype[g] == CSU_BC_WALL) f[3] = carry;
        else carry = f[3];
        int e0 = o->bface[3 * g];
        for (int k = 1; k < 4; k++) o->R[(size_t)k * nc + e0] -= f[k];
    }
}


/* ---------------- Spalart-Allmaras: sa.c ---------------- */
static void sa_calc_face(double ni, double ni_L, double r, double dnix, double dniy, double* fv1, double* tx,
                         double* ty)       /* saCalcFace, sa.c:273-290 */
{
    double Cv1 = 7.1, X = ni / ni_L, aux = X * X * X;
    *fv1 = aux / (aux + Cv1 * Cv1 * Cv1);
    double sig = 2. / 3.;
    *tx = r * (ni_L + ni) * dnix / sig;
    *ty = r * (ni_L + ni) * dniy / sig;
}

/* saCC_CalcFace, saCC.c:146-163: with the density-gradient term */
static void sacc_calc_face(double ni, double ni_L, double r, double dnix, double dniy, double drx, double dry,
                           double* fv1, double* tx, double* ty)
{
    double Cv1 = 7.1, X = ni / ni_L, aux = X * X * X;
    *fv1 = aux / (aux + Cv1 * Cv1 * Cv1);
    double sig = 2. / 3.;
    *tx = (r * (ni_L + ni) * dnix + 0.5 * ni * ni * drx) / sig;
    *ty = (r * (ni_L + ni) * dniy + 0.5 * ni * ni * dry) / sig;
}

/* saCalcSource, sa.c:292-349 (Allmaras et al. 2012); cc: saCC_CalcSource,
   saCC.c:165-226 (guarded S~ branch, density-gradient terms) */
static double sa_source(double ni, double ni_L, double S, double d, double rho, double drx, double dry,
                        double dnix, double dniy, int cc)
{
    double Cv1 = 7.1, X = ni / ni_L, aux = X * X * X;
    double fv1 = aux / (aux + Cv1 * Cv1 * Cv1);
    double k = 0.41;
    double fv2 = 1. - X / (1. + X * fv1);
    double Sbar = ni * fv2 / (k * k * d * d), Stil;
    double cv2 = 0.7, cv3 = 0.9;
    if (Sbar >= -cv2 * S) Stil = S + Sbar;
    else if (cc) {
        double aux2 = ((cv3 - 2 * cv2) * S - Sbar);
        Stil = aux2 == 0 ? S : S + S * (cv2 * cv2 * S + cv3 * Sbar) / aux2;
    } else Stil = S + S * (cv2 * cv2 * S + cv3 * Sbar) / ((cv3 - 2 * cv2) * S - Sbar);
    double sig = 2. / 3., Cb1 = 0.1355, Cb2 = 0.622;
    double Cw1 = Cb1 / (k * k) + (1. + Cb2) / sig;
    double Cw2 = 0.3;
    double r = ni / (Stil * k * k * d * d);
    r = fmin(r, 10);
    aux = r * r;
    aux = aux * aux * aux;
    double g = r + Cw2 * (aux - r);
    double Cw3 = 2.;
    aux = Cw3 * Cw3;
    aux = aux * aux * aux;
    double fw = g * pow((1. + aux) / (g * g * g * g * g * g + aux), 1. / 6.);
    if (cc) {
        double Qt = rho * (Cb1 * Stil * ni + Cb2 * (dnix * dnix + dniy * dniy) / sig - Cw1 * fw * (ni * ni / (d * d)));
        Qt += (Cb2 / sig) * (ni * (dnix * drx + dniy * dry) + 0.25 * ni * ni * (drx * drx + dry * dry) / rho);
        return Qt;
    }
    return rho * (Cb1 * Stil * ni + Cb2 * (dnix * dnix + dniy * dniy) / sig - Cw1 * fw * (ni * ni / (d * d)))
           - (ni_L + ni) * (dnix * drx + dniy * dry) / sig;
}

/* saInterFace (sa.c:138-221) or saInterFaceB (sa.c:26-134), saInterSource
   (sa.c:224-260); with saCC saCC_InterFace / saCC_InterSource
   (saCC.c:18-137), which ignore viscBlazek */
static void sa_inter(csu_oracle* o)
{
    const int nc = o->nc, nf = o->nf, ns = o->nv + 1;
    const double *gx = o->dPx, *gy = o->dPy;
#pragma omp parallel for
    for (int f = 0; f < nf; f++) {
        const int* q = o->face + 4 * f;
        int e0 = q[0], e1 = q[1];
        const double *P0 = o->P + (size_t)e0 * ns, *P1 = o->P + (size_t)e1 * ns;
        double dSx, dSy;
        face_ds(o, q[2], q[3], &dSx, &dSy);
        double dux = (gx[nc + e0] + gx[nc + e1]) * 0.5, dvx = (gx[2 * nc + e0] + gx[2 * nc + e1]) * 0.5;
        double dTx = (gx[3 * nc + e0] + gx[3 * nc + e1]) * 0.5, dnx = (gx[4 * nc + e0] + gx[4 * nc + e1]) * 0.5;
        double duy = (gy[nc + e0] + gy[nc + e1]) * 0.5, dvy = (gy[2 * nc + e0] + gy[2 * nc + e1]) * 0.5;
        double dTy = (gy[3 * nc + e0] + gy[3 * nc + e1]) * 0.5, dny = (gy[4 * nc + e0] + gy[4 * nc + e1]) * 0.5;
        if (o->prm.visc_blazek && !o->prm.sa_cc) {
            double x0, y0, x1, y1;
            cell_center(o, e0, &x0, &y0);
            cell_center(o, e1, &x1, &y1);
            double dx = x1 - x0, dy = y1 - y0, L = sqrt(dx * dx + dy * dy);
            double dul = (P1[1] - P0[1]) / L, dvl = (P1[2] - P0[2]) / L, dTl = (P1[4] - P0[4]) / L;
            double dnl = (P1[5] - P0[5]) / L;
            double a, b;
            a = dux + (dul - (dux * dx + duy * dy) / L) * dx / L; b = duy + (dul - (dux * dx + duy * dy) / L) * dy / L;
            dux = a; duy = b;
            a = dvx + (dvl - (dvx * dx + dvy * dy) / L) * dx / L; b = dvy + (dvl - (dvx * dx + dvy * dy) / L) * dy / L;
            dvx = a; dvy = b;
            a = dTx + (dTl - (dTx * dx + dTy * dy) / L) * dx / L; b = dTy + (dTl - (dTx * dx + dTy * dy) / L) * dy / L;
            dTx = a; dTy = b;
            a = dnx + (dnl - (dnx * dx + dny * dy) / L) * dx / L; b = dny + (dnl - (dnx * dx + dny * dy) / L) * dy / L;
            dnx = a; dny = b;
        }
        double r = (P1[0] + P0[0]) * 0.5, u = (P1[1] + P0[1]) * 0.5, v = (P1[2] + P0[2]) * 0.5;
        double T = (P1[4] + P0[4]) * 0.5, n = (P1[5] + P0[5]) * 0.5;
        double mi_L = sutherland(T), n_L = mi_L / r, fv1, tx, ty;
        if (o->prm.sa_cc)
            sacc_calc_face(n, n_L, r, dnx, dny, (gx[e0] + gx[e1]) * 0.5, (gy[e0] + gy[e1]) * 0.5, &fv1, &tx, &ty);
        else
            sa_calc_face(n, n_L, r, dnx, dny, &fv1, &tx, &ty);
        double mi_t = fv1 * r * n, mi = mi_L + mi_t;
        double k = gas_T2Cp(&o->gas, T) * (mi_L / o->prm.pr + mi_t / o->prm.pr_t);
        o->miT[f] = mi_t;
        double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
        o->flux[(size_t)1 * nf + f] = txx * dSx + txy * dSy;
        o->flux[(size_t)2 * nf + f] = txy * dSx + tyy * dSy;
        o->flux[(size_t)3 * nf + f] = (txx * dSx + txy * dSy) * u + (txy * dSx + tyy * dSy) * v + k * (dTx * dSx + dTy * dSy);
        o->flux[(size_t)4 * nf + f] = tx * dSx + ty * dSy;
    }
#pragma omp parallel for
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < 4; j++) {
            int f = o->cface[4 * i + j];
            if (f > 0) for (int k = 1; k < 5; k++) o->R[(size_t)k * nc + i] -= o->flux[(size_t)k * nf + f - 1];
            else if (f < 0) for (int k = 1; k < 5; k++) o->R[(size_t)k * nc + i] += o->flux[(size_t)k * nf - f - 1];
        }
#pragma omp parallel for
    for (int i = 0; i < nc; i++) {
        const double* P = o->P + (size_t)i * ns;
        double rho = P[0], T = P[4], n = P[5], mi_L = sutherland(T), n_L = mi_L / rho;
        double S = fabs(gy[nc + i] - gx[2 * nc + i]);
        double Qt = sa_source(n, n_L, S, o->wd[i], rho, gx[i], gy[i], gx[4 * nc + i], gy[4 * nc + i], o->prm.sa_cc);
        o->R[(size_t)4 * nc + i] -= Qt * o->omega[i];
    }
}

/* saBoundaryFaceViscFlux, sa.c:370-692; with saCC
   saCC_BoundaryFaceViscFlux, saCC.c:247-589 */
static void sa_bc_flux(const csu_oracle* o, int g, double* f, double* mie)
{
    const int nc = o->nc, ns = o->nv + 1, e0 = o->bface[3 * g], t = o->btype[g];
    const int p0 = o->bface[3 * g + 1], p1 = o->bface[3 * g + 2];
    const double* P = o->P + (size_t)e0 * ns;
    const double *gx = o->dPx, *gy = o->dPy;
    double dux, duy, dvx, dvy, dTx = 0, dTy = 0, dnx = 0, dny = 0, dSx, dSy, fv1, tx, ty;
    if (t == CSU_BC_SYMMETRY) {
        double x0 = o->xy[2 * p0], y0 = o->xy[2 * p0 + 1], x1 = o->xy[2 * p1], y1 = o->xy[2 * p1 + 1];
        double sx = y1 - y0, sy = -(x1 - x0), dS = sqrt(sx * sx + sy * sy);
        double nx = sx / dS, ny = sy / dS;
        if (o->axi == 1) dS = dS * (y0 + y1) * 0.5;
        double duxm = gx[nc + e0], dvxm = gx[2 * nc + e0], dTxm = gx[3 * nc + e0];
        double duym = gy[nc + e0], dvym = gy[2 * nc + e0], dTym = gy[3 * nc + e0];
        dux = duxm - (duxm * nx + duym * ny) * nx; duy = duym - (duxm * nx + duym * ny) * ny;
        dvx = dvxm - (dvxm * nx + dvym * ny) * nx; dvy = dvym - (dvxm * nx + dvym * ny) * ny;
        dTx = dTxm - (dTxm * nx + dTym * ny) * nx; dTy = dTym - (dTxm * nx + dTym * ny) * ny;
        double rho = P[0], u = P[1], v = P[2], T = P[4], n = P[5], mi_L = sutherland(T), n_L = mi_L / rho;
        sa_calc_face(n, n_L, rho, 0.0, 0.0, &fv1, &tx, &ty);   /* tx, ty unused (f[4] = 0) */
        double mi_t = fv1 * rho * n, mi = mi_L + mi_t;
        if (mie) *mie = mi_t;
        double k = gas_T2Cp(&o->gas, T) * (mi_L / o->prm.pr + mi_t / o->prm.pr_t);
        double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
        f[1] = (txx * nx + txy * ny) * dS;
        f[2] = (txy * nx + tyy * ny) * dS;
        f[3] = (u * (txx * nx + txy * ny) + v * (txy * nx + tyy * ny) + k * (dTx * nx + dTy * ny)) * dS;
        f[4] = 0.0;
        return;
    }
    face_ds(o, p0, p1, &dSx, &dSy);
    if (t == CSU_BC_OUTLET) {
        dux = gx[nc + e0]; dvx = gx[2 * nc + e0]; dTx = gx[3 * nc + e0]; dnx = gx[4 * nc + e0];
        duy = gy[nc + e0]; dvy = gy[2 * nc + e0]; dTy = gy[3 * nc + e0]; dny = gy[4 * nc + e0];
        double rho = P[0], u = P[1], v = P[2], T = P[4], n = P[5], mi_L = sutherland(T), n_L = mi_L / rho;
        if (o->prm.sa_cc) sacc_calc_face(n, n_L, rho, dnx, dny, gx[e0], gy[e0], &fv1, &tx, &ty);
        else sa_calc_face(n, n_L, rho, dnx, dny, &fv1, &tx, &ty);
        double mi_t = fv1 * rho * n, mi = mi_L + mi_t;
        if (mie) *mie = mi_t;
        double k = gas_T2Cp(&o->gas, T) * (mi_L / o->prm.pr + mi_t / o->prm.pr_t);
        double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
        f[1] = txx * dSx + txy * dSy;
        f[2] = txy * dSx + tyy * dSy;
        f[3] = u * (txx * dSx + txy * dSy) + v * (txy * dSx + tyy * dSy) + k * (dTx * dSx + dTy * dSy);
        f[4] = tx * dSx + ty * dSy;
        return;
    }
    /* inlet, wall, wallT: one-sided differences to the boundary values */
    const double* Pin = o->prm.pin;
    double x0, y0, x1, y1;
    cell_center(o, e0, &x0, &y0);
    x1 = (o->xy[2 * p0] + o->xy[2 * p1]) * 0.5; y1 = (o->xy[2 * p0 + 1] + o->xy[2 * p1 + 1]) * 0.5;
    double dx = x1 - x0, dy = y1 - y0, L = sqrt(dx * dx + dy * dy);
    int inl = t == CSU_BC_INLET;
    double dul = ((inl ? Pin[1] : 0) - P[1]) / L, dvl = ((inl ? Pin[2] : 0) - P[2]) / L;
    double dnl = ((inl ? Pin[5] : 0) - P[5]) / L;
    double duxm = gx[nc + e0], dvxm = gx[2 * nc + e0], dnxm = gx[4 * nc + e0];
    double duym = gy[nc + e0], dvym = gy[2 * nc + e0], dnym = gy[4 * nc + e0];
    dux = duxm + (dul - (duxm * dx + duym * dy) / L) * dx / L; duy = duym + (dul - (duxm * dx + duym * dy) / L) * dy / L;
    dvx = dvxm + (dvl - (dvxm * dx + dvym * dy) / L) * dx / L; dvy = dvym + (dvl - (dvxm * dx + dvym * dy) / L) * dy / L;
    if (t != CSU_BC_WALL) {
        double dTl = ((inl ? Pin[4] : o->prm.twall) - P[4]) / L;
        double dTxm = gx[3 * nc + e0], dTym = gy[3 * nc + e0];
        dTx = dTxm + (dTl - (dTxm * dx + dTym * dy) / L) * dx / L;
        dTy = dTym + (dTl - (dTxm * dx + dTym * dy) / L) * dy / L;
    }
    dnx = dnxm + (dnl - (dnxm * dx + dnym * dy) / L) * dx / L; dny = dnym + (dnl - (dnxm * dx + dnym * dy) / L) * dy / L;
    double rho = inl ? Pin[0] : P[0];
    double T = inl ? Pin[4] : (t == CSU_BC_WALLT ? o->prm.twall : P[4]);
    double n = inl ? Pin[5] : 0.0;
    double mi_L = sutherland(T), n_L = mi_L / rho;
    if (o->prm.sa_cc) {
        /* saCC.c:334-353 (inlet), :425-426 (wall: d(rho)/dx in both), :494-495 (wallT) */
        double drx = gx[e0], dry = gy[e0];
        if (inl) {
            double drl = (Pin[0] - P[0]) / L, drxm = gx[e0], drym = gy[e0];
            drx = drxm + (drl - (drxm * dx + drym * dy) / L) * dx / L;
            dry = drym + (drl - (drxm * dx + drym * dy) / L) * dy / L;
        } else if (t == CSU_BC_WALL) {
            dry = gx[e0];
        }
        sacc_calc_face(n, n_L, rho, dnx, dny, drx, dry, &fv1, &tx, &ty);
    } else {
        sa_calc_face(n, n_L, rho, dnx, dny, &fv1, &tx, &ty);
    }
    double mi_t = fv1 * rho * n, mi = t == CSU_BC_WALL ? mi_L : mi_L + mi_t;
    if (mie) *mie = t == CSU_BC_WALL ? 0.0 : mi_t;   /* sa.c:570 */
    double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
    f[1] = txx * dSx + txy * dSy;
    f[2] = txy * dSx + tyy * dSy;
    f[4] = tx * dSx + ty * dSy;
    if (t == CSU_BC_WALL) { f[3] = 0.0; return; }
    double k = gas_T2Cp(&o->gas, T) * (mi_L / o->prm.pr + mi_t / o->prm.pr_t);
    if (t == CSU_BC_WALLT) { f[3] = k * (dTx * dSx + dTy * dSy); return; }
    double u = Pin[1], v = Pin[2];
    f[3] = u * (txx * dSx + txy * dSy) + v * (txy * dSx + tyy * dSy) + k * (dTx * dSx + dTy * dSy);
}

static void sa_boundary(csu_oracle* o)       /* saBoundary / saBoundaryFace, sa.c:351-368 */
{
    const int nc = o->nc;
    for (int g = 0; g < o->nb; g++) {
        double f[5];
        sa_bc_flux(o, g, f, NULL);
        int e0 = o->bface[3 * g];
        for (int k = 1; k < 5; k++) o->R[(size_t)k * nc + e0] -= f[k];
    }
}

/* ---------------- Menter SST: sst.c ---------------- */
typedef struct {
    double dux, duy, dvx, dvy, dkx, dky, dox, doy;
    double r, T, k, om, d, mi_L, mi_t;
    double tkx, tky, tox, toy, F1, F2, Qtk, Qto;
} sstv;

static double sst_blend(double x1, double x2, double F1) { return x1 * F1 + x2 * (1 - F1); }  /* sst.c:693 */

/* sstF1 (sst.c:788-809).  Its long double temporaries only ever hold
   values computed in double, so double restates them exactly. */
static double sst_F1(const sstv* v, double n_L_term, double sqrtk_term)
{
    double aux1 = 2 * v->r * SST_SO2 * (v->dkx * v->dox + v->dky * v->doy) / v->om;
    if (v->om == 0) aux1 = 0;
    double CD = fmax(aux1, 1.0e-20);
    aux1 = fmax(sqrtk_term, n_L_term);
    double aux2 = 4 * v->r * SST_SO2 * v->k / (CD * v->d * v->d);
    double arg1 = fmin(aux1, aux2);
    return tanh(arg1 * arg1 * arg1 * arg1);
}

static double sst_F2(double n_L_term, double sqrtk_term)      /* sst.c:811-818 */
{
    double arg2 = fmax(2 * sqrtk_term, n_L_term);
    return tanh(arg2 * arg2);
}

/* ---- the transition variant, sstTrans.c: its own constants, Ry = Rt = 1e6
   fixed (sstTrans.c:48,58,140,150), F1 without the omega = 0 guard ---- */
#define STR_SK1 0.85
#define STR_SO1 0.5
#define STR_B1 0.075
#define STR_SK2 1.0
#define STR_SO2 0.856
#define STR_B2 0.0828
#define STR_BS2 0.09
#define STR_ALPHAS2 1.0
#define STR_ALPHA2 0.4403
#define STR_A1 0.31
#define STR_RY 1e6
#define STR_RT 1e6

/* sstTransF1 (sstTrans.c:96-115) */
static double str_F1s(const sstv* v, double n_L_term, double sqrtk_term)
{
    double aux1 = 2 * v->r * STR_SO2 * (v->dkx * v->dox + v->dky * v->doy) / v->om;
    double CD = fmax(aux1, 1.0e-20);
    aux1 = fmax(sqrtk_term, n_L_term);
    double aux2 = 4 * v->r * STR_SO2 * v->k / (CD * v->d * v->d);
    double arg1 = fmin(aux1, aux2);
    return tanh(arg1 * arg1 * arg1 * arg1);
}

/* common preamble of sstTransFlux / sstTransSources (sstTrans.c:36-59,128-151) */
static void str_blend_fns(const sstv* v, double* omega, double* F1, double* F2)
{
    double n_L = v->mi_L / v->r;
    *omega = fabs(v->duy - v->dvx);
    double sqrtk = sqrt(v->k);
    double sqrtk_term = sqrtk / (STR_BS2 * v->om * v->d);
    double n_L_term = 500 * n_L / (v->d * v->d * v->om);
    double Fs1 = str_F1s(v, n_L_term, sqrtk_term);
    double aux = STR_RY / 120;
    aux = aux * aux;
    aux = aux * aux;
    aux = aux * aux;
    *F1 = fmax(Fs1, exp(-aux));
    double arg2 = fmax(2 * sqrtk_term, n_L_term);   /* sstTransF2, sstTrans.c:117-124 */
    *F2 = tanh(arg2 * arg2);
}

/* sstTransFlux (sstTrans.c:34-74) */
static void str_flux(sstv* v)
{
    double omega, F1, F2;
    str_blend_fns(v, &omega, &F1, &F2);
    double alphas1 = (.025 + STR_RT / 6) / (1 + STR_RT / 6);          /* sstTransAlphas1 */
    double alphas = sst_blend(alphas1, STR_ALPHAS2, F1);
    double sk = sst_blend(STR_SK1, STR_SK2, F1), so = sst_blend(STR_SO1, STR_SO2, F1);
    v->mi_t = fmin(alphas * v->r * v->k / v->om, STR_A1 * v->r * v->k / (omega * F2));
    double aux = v->mi_L + sk * v->mi_t;
    v->tkx = aux * v->dkx;
    v->tky = aux * v->dky;
    aux = v->mi_L + so * v->mi_t;
    v->tox = aux * v->dox;
    v->toy = aux * v->doy;
}

/* sstTransSources (sstTrans.c:126-174) */
static void str_sources(sstv* v)
{
    double omega, F1, F2;
    str_blend_fns(v, &omega, &F1, &F2);
    double alpha1 = (5. / 9.) * (0.1 + STR_RT / 2.7) / (1 + STR_RT / 2.7);   /* sstTransAlpha1 */
    double alphas1 = (.025 + STR_RT / 6) / (1 + STR_RT / 6);
    double a8 = STR_RT / 8;                                                  /* sstTransBs1 */
    a8 = a8 * a8;
    a8 = a8 * a8;
    double bs1 = 0.09 * (5. / 18. + a8) / (1 + a8);
    double alpha = sst_blend(alpha1, STR_ALPHA2, F1), alphas = sst_blend(alphas1, STR_ALPHAS2, F1);
    double b = sst_blend(STR_B1, STR_B2, F1), bs = sst_blend(bs1, STR_BS2, F1);
    v->mi_t = fmin(alphas * v->r * v->k / v->om, STR_A1 * v->r * v->k / (omega * F2));
    double n_t = v->mi_t / v->r;
    double P = v->mi_t * omega * omega;
    v->F1 = F1;
    v->F2 = F2;
    v->Qtk = P - bs * v->r * v->om * v->k;
    v->Qto = alpha * P / n_t - b * v->r * v->om * v->om
             + 2 * (1 - F1) * v->r * STR_SO2 * (v->dkx * v->dox + v->dky * v->doy) / v->om;
}

/* sstFlux (sst.c:749-786): eddy viscosity and the k, omega diffusion flux densities */
static void sst_flux(const csu_oracle* o, sstv* v)
{
    if (o->prm.sst_trans) { str_flux(v); return; }
    double n_L = v->mi_L / v->r;
    double omega = fabs(v->duy - v->dvx);
    double sqrtk_term = sqrt(v->k) / (SST_BS * v->om * v->d);
    double n_L_term = 500 * n_L / (v->d * v->d * v->om);
    double F1 = sst_F1(v, n_L_term, sqrtk_term);
    double F2 = sst_F2(n_L_term, sqrtk_term);
    v->mi_t = v->r * SST_A1 * v->k / fmax(SST_A1 * v->om, omega * F2);
    double sk = sst_blend(SST_SK1, SST_SK2, F1), so = sst_blend(SST_SO1, SST_SO2, F1);
    double aux = v->mi_L + sk * v->mi_t;
    v->tkx = aux * v->dkx;
    v->tky = aux * v->dky;
    aux = v->mi_L + so * v->mi_t;
    v->tox = aux * v->dox;
    v->toy = aux * v->doy;
}

/* sstSources (sst.c:698-747) */
static void sst_sources(const csu_oracle* o, sstv* v)
{
    if (o->prm.sst_trans) { str_sources(v); return; }
    double n_L = v->mi_L / v->r;
    double omega = fabs(v->duy - v->dvx);
    double sqrtk_term = sqrt(v->k) / (SST_BS * v->om * v->d);
    double n_L_term = 500 * n_L / (v->d * v->d * v->om);
    double F1 = sst_F1(v, n_L_term, sqrtk_term);
    double F2 = sst_F2(n_L_term, sqrtk_term);
    v->mi_t = v->r * SST_A1 * v->k / fmax(SST_A1 * v->om, omega * F2);
    double n_t = v->mi_t / v->r;
    double g = sst_blend(o->sst_g1, o->sst_g2, F1), b = sst_blend(SST_B1, SST_B2, F1);
    double txx = 2 * v->mi_t * (v->dux - (v->dux + v->dvy) / 3) - 2. * v->r * v->k / 3;
    double tyy = 2 * v->mi_t * (v->dvy - (v->dux + v->dvy) / 3) - 2. * v->r * v->k / 3;
    double txy = v->mi_t * (v->duy + v->dvx);
    double P = v->dux * txx + v->dvy * tyy + (v->duy + v->dvx) * txy;
    P = fmin(P, 10 * SST_BS * v->r * v->om * v->k);
    v->F1 = F1;
    v->F2 = F2;
    double cd = 2 * (1 - F1) * v->r * SST_SO2 * (v->dkx * v->dox + v->dky * v->doy) / v->om;
    v->Qtk = P - SST_BS * v->r * v->om * v->k;
    v->Qto = g * P / n_t - b * v->r * v->om * v->om + cd;
}

/* stresses with the -2/3 rho k term (sst.c:163-165) */
static void sst_tau(double mi, const sstv* v, double* txx, double* tyy, double* txy)
{
    *txx = 2 * mi * (v->dux - (v->dux + v->dvy) / 3) - 2. * v->r * v->k / 3.;
    *tyy = 2 * mi * (v->dvy - (v->dux + v->dvy) / 3) - 2. * v->r * v->k / 3.;
    *txy = mi * (v->duy + v->dvx);
}

/* sstInterFace (sst.c:56-194) and sstInterSource (sst.c:595-691) */
static void sst_inter(csu_oracle* o)
{
    const int nc = o->nc, nf = o->nf, ns = o->nv + 1;
    const double *gx = o->dPx, *gy = o->dPy;
#pragma omp parallel for
    for (int f = 0; f < nf; f++) {
        const int* q = o->face + 4 * f;
        int e0 = q[0], e1 = q[1];
        const double *P0 = o->P + (size_t)e0 * ns, *P1 = o->P + (size_t)e1 * ns;
        double dSx, dSy, x0, y0, x1, y1;
        face_ds(o, q[2], q[3], &dSx, &dSy);
        cell_center(o, e0, &x0, &y0);
        cell_center(o, e1, &x1, &y1);
        double dx = x1 - x0, dy = y1 - y0, L = sqrt(dx * dx + dy * dy);
        double nx = dx / L, ny = dy / L;
        double dul = (P1[1] - P0[1]) / L, dvl = (P1[2] - P0[2]) / L, dTl = (P1[4] - P0[4]) / L;
        double dkl = (P1[5] - P0[5]) / L, dol = (P1[6] - P0[6]) / L;
        double duxm = 0.5 * (gx[nc + e0] + gx[nc + e1]), dvxm = 0.5 * (gx[2 * nc + e0] + gx[2 * nc + e1]);
        double dTxm = 0.5 * (gx[3 * nc + e0] + gx[3 * nc + e1]), dkxm = 0.5 * (gx[4 * nc + e0] + gx[4 * nc + e1]);
        double doxm = 0.5 * (gx[5 * nc + e0] + gx[5 * nc + e1]);
        double duym = 0.5 * (gy[nc + e0] + gy[nc + e1]), dvym = 0.5 * (gy[2 * nc + e0] + gy[2 * nc + e1]);
        double dTym = 0.5 * (gy[3 * nc + e0] + gy[3 * nc + e1]), dkym = 0.5 * (gy[4 * nc + e0] + gy[4 * nc + e1]);
        double doym = 0.5 * (gy[5 * nc + e0] + gy[5 * nc + e1]);
        sstv v;
        double aux;
        aux = duxm * nx + duym * ny; v.dux = duxm + (dul - aux) * nx; v.duy = duym + (dul - aux) * ny;
        aux = dvxm * nx + dvym * ny; v.dvx = dvxm + (dvl - aux) * nx; v.dvy = dvym + (dvl - aux) * ny;
        aux = dTxm * nx + dTym * ny;
        double dTx = dTxm + (dTl - aux) * nx, dTy = dTym + (dTl - aux) * ny;
        aux = dkxm * nx + dkym * ny; v.dkx = dkxm + (dkl - aux) * nx; v.dky = dkym + (dkl - aux) * ny;
        aux = doxm * nx + doym * ny; v.dox = doxm + (dol - aux) * nx; v.doy = doym + (dol - aux) * ny;
        v.r = (P1[0] + P0[0]) * 0.5;
        double u = (P1[1] + P0[1]) * 0.5, w = (P1[2] + P0[2]) * 0.5;
        v.T = (P1[4] + P0[4]) * 0.5;
        v.k = (P1[5] + P0[5]) * 0.5;
        v.om = (P1[6] + P0[6]) * 0.5;
        v.d = (o->wd[e1] + o->wd[e0]) * 0.5;
        v.mi_L = sutherland(v.T);
        sst_flux(o, &v);
        double mi = v.mi_L + v.mi_t;
        double kk = gas_T2Cp(&o->gas, v.T) * (v.mi_L / o->prm.pr + v.mi_t / o->prm.pr_t);
        o->miT[f] = v.mi_t;
        double txx, tyy, txy;
        sst_tau(mi, &v, &txx, &tyy, &txy);
        o->flux[(size_t)1 * nf + f] = txx * dSx + txy * dSy;
        o->flux[(size_t)2 * nf + f] = txy * dSx + tyy * dSy;
        o->flux[(size_t)3 * nf + f] = (txx * dSx + txy * dSy) * u + (txy * dSx + tyy * dSy) * w + kk * (dTx * dSx + dTy * dSy);
        o->flux[(size_t)4 * nf + f] = v.tkx * dSx + v.tky * dSy;
        o->flux[(size_t)5 * nf + f] = v.tox * dSx + v.toy * dSy;
    }
#pragma omp parallel for
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < 4; j++) {
            int f = o->cface[4 * i + j];
            if (f > 0) for (int k = 1; k < 6; k++) o->R[(size_t)k * nc + i] -= o->flux[(size_t)k * nf + f - 1];
            else if (f < 0) for (int k = 1; k < 6; k++) o->R[(size_t)k * nc + i] += o->flux[(size_t)k * nf - f - 1];
        }
#pragma omp parallel for
    for (int i = 0; i < nc; i++) {
        const double* P = o->P + (size_t)i * ns;
        const double om_i = o->omega[i];
        sstv v;
        v.dux = gx[nc + i]; v.dvx = gx[2 * nc + i]; v.dkx = gx[4 * nc + i]; v.dox = gx[5 * nc + i];
        v.duy = gy[nc + i]; v.dvy = gy[2 * nc + i]; v.dky = gy[4 * nc + i]; v.doy = gy[5 * nc + i];
        v.r = P[0]; v.T = P[4]; v.k = P[5]; v.om = P[6]; v.d = o->wd[i];
        v.mi_L = sutherland(v.T);
        if (v.om == 0) {             /* omLim, solver.c:1155 */
            o->dQkdr[i] = 0; o->dQkdrk[i] = 0; o->dQkdro[i] = 0;
            o->dQodr[i] = 0; o->dQodrk[i] = 0; o->dQodro[i] = 0;
            continue;
        }
        sst_sources(o, &v);
        o->R[(size_t)4 * nc + i] -= v.Qtk * om_i;
        o->R[(size_t)5 * nc + i] -= v.Qto * om_i;
        double dQkdr = v.Qtk, dQkdo = v.Qtk, dQkdk = v.Qtk;
        double dQodr = v.Qto, dQodo = v.Qto, dQodk = v.Qto;
        const double h = 1e-6, dh = 1e-6;
        double hh;
        hh = fabs(v.om * h); hh = fmax(hh, dh);
        v.om += hh;
        sst_sources(o, &v);
        dQkdo = (v.Qtk - dQkdo) / hh; dQodo = (v.Qto - dQodo) / hh;
        v.om = P[6];
        hh = fabs(v.k * h); hh = fmax(hh, dh);
        v.k += hh;
        sst_sources(o, &v);
        dQkdk = (v.Qtk - dQkdk) / hh; dQodk = (v.Qto - dQodk) / hh;
        v.k = P[5];
        hh = fabs(v.r * h); hh = fmax(hh, dh);
        v.r += hh;
        sst_sources(o, &v);
        dQkdr = (v.Qtk - dQkdr) / hh; dQodr = (v.Qto - dQodr) / hh;
        v.r = P[0];
        o->dQkdr[i] = (dQkdr - dQkdk * v.k / v.r - dQkdo * v.om / v.r) * om_i;
        o->dQkdrk[i] = dQkdk * om_i / v.r;
        o->dQkdro[i] = dQkdo * om_i / v.r;
        o->dQodr[i] = (dQodr - dQodk * v.k / v.r - dQodo * v.om / v.r) * om_i;
        o->dQodrk[i] = dQodk * om_i / v.r;
        o->dQodro[i] = dQodo * om_i / v.r;
    }
}

/* sstBoundaryFaceViscFlux (sst.c:202-547): f[1..5] */
static void sst_bc_flux(const csu_oracle* o, int g, double* f, double* mie)
{
    const int nc = o->nc, ns = o->nv + 1, e0 = o->bface[3 * g], t = o->btype[g];
    const int p0 = o->bface[3 * g + 1], p1 = o->bface[3 * g + 2];
    const double* P = o->P + (size_t)e0 * ns;
    const double *gx = o->dPx, *gy = o->dPy;
    const double* Pin = o->prm.pin;
    sstv v;
    double dTx = 0, dTy = 0, txx, tyy, txy;
    if (t == CSU_BC_SYMMETRY) {
        double x0 = o->xy[2 * p0], y0 = o->xy[2 * p0 + 1], x1 = o->xy[2 * p1], y1 = o->xy[2 * p1 + 1];
        double sx = y1 - y0, sy = -(x1 - x0), dS = sqrt(sx * sx + sy * sy);
        double nx = sx / dS, ny = sy / dS;
        if (o->axi == 1) dS = dS * (y0 + y1) * 0.5;
        double duxm = gx[nc + e0], dvxm = gx[2 * nc + e0], dkxm = gx[4 * nc + e0], doxm = gx[5 * nc + e0];
        double duym = gy[nc + e0], dvym = gy[2 * nc + e0], dkym = gy[4 * nc + e0], doym = gy[5 * nc + e0];
        v.dux = duxm - (duxm * nx + duym * ny) * nx; v.duy = duym - (duxm * nx + duym * ny) * ny;
        v.dvx = dvxm - (dvxm * nx + dvym * ny) * nx; v.dvy = dvym - (dvxm * nx + dvym * ny) * ny;
        v.dkx = dkxm - (dkxm * nx + dkym * ny) * nx; v.dky = dkym - (dkxm * nx + dkym * ny) * ny;
        v.dox = doxm - (doxm * nx + doym * ny) * nx; v.doy = doym - (doxm * nx + doym * ny) * ny;
        v.r = P[0]; v.T = P[4]; v.k = P[5]; v.om = P[6]; v.d = o->wd[e0];
        double u = P[1], w = P[2];
        v.mi_L = sutherland(v.T);
        sst_flux(o, &v);
        if (mie) *mie = v.mi_t;
        sst_tau(v.mi_L + v.mi_t, &v, &txx, &tyy, &txy);
        f[1] = (txx * nx + txy * ny) * dS;
        f[2] = (txy * nx + tyy * ny) * dS;
        f[3] = (u * (txx * nx + txy * ny) + w * (txy * nx + tyy * ny)) * dS;   /* no heat flux */
        f[4] = 0.0;
        f[5] = 0.0;
        return;
    }
    double dSx, dSy;
    face_ds(o, p0, p1, &dSx, &dSy);
    if (t == CSU_BC_INLET || t == CSU_BC_WALL || t == CSU_BC_WALLT) {
        double x0, y0;
        cell_center(o, e0, &x0, &y0);
        double x1 = (o->xy[2 * p0] + o->xy[2 * p1]) * 0.5, y1 = (o->xy[2 * p0 + 1] + o->xy[2 * p1 + 1]) * 0.5;
        double dx = x1 - x0, dy = y1 - y0, L = sqrt(dx * dx + dy * dy);
        double ub, vb, Tb = 0, kb, ob;
        v.d = o->wd[e0];
        if (t == CSU_BC_INLET) {
            v.r = Pin[0]; v.T = Pin[4]; v.k = Pin[5]; v.om = Pin[6];
            ub = Pin[1]; vb = Pin[2]; Tb = Pin[4]; kb = Pin[5]; ob = Pin[6];
            v.mi_L = sutherland(v.T);
        } else {
            v.r = P[0]; v.T = P[4]; v.k = P[5]; v.om = P[6];
            v.mi_L = sutherland(v.T);
            double n_L = v.mi_L / v.r;
            double owall = o->prm.sst_owallfactor * 6 * n_L / (SST_B1 * v.d * v.d);
            ub = 0; vb = 0; kb = 0; ob = owall;
            if (t == CSU_BC_WALLT) Tb = o->prm.twall;
        }
        double dul = (ub - P[1]) / L, dvl = (vb - P[2]) / L, dkl = (kb - P[5]) / L, dol = (ob - P[6]) / L;
        double duxm = gx[nc + e0], dvxm = gx[2 * nc + e0], dkxm = gx[4 * nc + e0], doxm = gx[5 * nc + e0];
        double duym = gy[nc + e0], dvym = gy[2 * nc + e0], dkym = gy[4 * nc + e0], doym = gy[5 * nc + e0];
        v.dux = duxm + (dul - (duxm * dx + duym * dy) / L) * dx / L;
        v.duy = duym + (dul - (duxm * dx + duym * dy) / L) * dy / L;
        v.dvx = dvxm + (dvl - (dvxm * dx + dvym * dy) / L) * dx / L;
        v.dvy = dvym + (dvl - (dvxm * dx + dvym * dy) / L) * dy / L;
        if (t != CSU_BC_WALL) {
            double dTl = (Tb - P[4]) / L, dTxm = gx[3 * nc + e0], dTym = gy[3 * nc + e0];
            dTx = dTxm + (dTl - (dTxm * dx + dTym * dy) / L) * dx / L;
            dTy = dTym + (dTl - (dTxm * dx + dTym * dy) / L) * dy / L;
        }
        v.dkx = dkxm + (dkl - (dkxm * dx + dkym * dy) / L) * dx / L;
        v.dky = dkym + (dkl - (dkxm * dx + dkym * dy) / L) * dy / L;
        v.dox = doxm + (dol - (doxm * dx + doym * dy) / L) * dx / L;
        v.doy = doym + (dol - (doxm * dx + doym * dy) / L) * dy / L;
        sst_flux(o, &v);
        if (mie) *mie = v.mi_t;
        sst_tau(v.mi_L + v.mi_t, &v, &txx, &tyy, &txy);
        f[1] = txx * dSx + txy * dSy;
        f[2] = txy * dSx + tyy * dSy;
        f[4] = v.tkx * dSx + v.tky * dSy;
        f[5] = v.tox * dSx + v.toy * dSy;
        if (t == CSU_BC_WALL) { f[3] = 0.0; return; }
        double kk = gas_T2Cp(&o->gas, v.T) * (v.mi_L / o->prm.pr + v.mi_t / o->prm.pr_t);
        if (t == CSU_BC_WALLT) { f[3] = kk * (dTx * dSx + dTy * dSy); return; }
        f[3] = ub * (txx * dSx + txy * dSy) + vb * (txy * dSx + tyy * dSy) + kk * (dTx * dSx + dTy * dSy);
        return;
    }
    /* outlet */
    v.dux = gx[nc + e0]; v.dvx = gx[2 * nc + e0]; dTx = gx[3 * nc + e0]; v.dkx = gx[4 * nc + e0]; v.dox = gx[5 * nc + e0];
    v.duy = gy[nc + e0]; v.dvy = gy[2 * nc + e0]; dTy = gy[3 * nc + e0]; v.dky = gy[4 * nc + e0]; v.doy = gy[5 * nc + e0];
    v.r = P[0]; v.T = P[4]; v.k = P[5]; v.om = P[6]; v.d = o->wd[e0];
    double u = P[1], w = P[2];
    v.mi_L = sutherland(v.T);
    sst_flux(o, &v);
    if (mie) *mie = v.mi_t;
    double kk = gas_T2Cp(&o->gas, v.T) * (v.mi_L / o->prm.pr + v.mi_t / o->prm.pr_t);
    sst_tau(v.mi_L + v.mi_t, &v, &txx, &tyy, &txy);
    f[1] = txx * dSx + txy * dSy;
    f[2] = txy * dSx + tyy * dSy;
    f[3] = u * (txx * dSx + txy * dSy) + w * (txy * dSx + tyy * dSy) + kk * (dTx * dSx + dTy * dSy);
    f[4] = v.tkx * dSx + v.tky * dSy;
    f[5] = v.tox * dSx + v.toy * dSy;
}

static void sst_boundary(csu_oracle* o)      /* sstBoundary / sstBoundaryFace, sst.c:549-575 */
{
    const int nc = o->nc;
    for (int g = 0; g < o->nb; g++) {
        double f[6];
        sst_bc_flux(o, g, f, NULL);
        int e0 = o->bface[3 * g];
        for (int k = 1; k < 6; k++) o->R[(size_t)k * nc + e0] -= f[k];
    }
}

/* solverCalcR, solver.c:685-739 */
static void calc_R(csu_oracle* o, double* U)
{
    memset(o->R, 0, sizeof(double) * (size_t)o->nv * o->nc);
    calc_primitive(o, U);
    update_grad(o);
    inter(o);
    boundary(o);
    if (o->axi == 1) axis_pressure(o);
    if (o->prm.laminar == 1) {
        grad_T(o);
        laminar_inter(o);
        laminar_boundary(o);
    }
    if (o->prm.sa == 1) {
        grad_T(o);
        sa_inter(o);
        sa_boundary(o);
    }
    if (o->prm.sst == 1) {
        grad_T(o);
        sst_inter(o);
        sst_boundary(o);
    }
}

/* ---------------- time step: solver.c:895-950 ---------------- */
static double local_time_step(const csu_oracle* o, int i)
{
    int n = nvert(o, i), nc = o->nc;
    double sx = 0., sy = 0.;
    for (int j = 0; j < n; j++) {
        double dSx, dSy;
        face_ds(o, o->cv[4 * i + j], o->cv[4 * i + (j + 1) % n], &dSx, &dSy);
        sx += fabs(dSx); sy += fabs(dSy);
    }
    sx *= 0.5; sy *= 0.5;
    const double* U = o->U;
    double k = o->prm.sst ? U[4 * nc + i] / U[i] : 0;
    double E = U[3 * nc + i] / U[i];
    double u = U[nc + i] / U[i], v = U[2 * nc + i] / U[i];
    double aux = (E - (u * u + v * v) / 2 - k);
    double c = gas_T2c(&o->gas, gas_e2T(&o->gas, aux));
    double Lx = (fabs(u) + c) * sx, Ly = (fabs(v) + c) * sy;
    return o->omega[i] / (Lx + Ly);
}

static void calc_dt(csu_oracle* o)
{
    if (o->dt_local == 1) {
#pragma omp parallel for
        for (int i = 0; i < o->nc; i++) o->dtL[i] = o->prm.cfl * local_time_step(o, i);
    } else {
        double dt = local_time_step(o, 0);
        for (int i = 1; i < o->nc; i++) {
            double d = local_time_step(o, i);
            if (d < dt) dt = d;
        }
        o->dt = dt * o->prm.cfl;
    }
}

/* ---------------- RK: solver.c:741-840 ---------------- */
static void rk_stage(csu_oracle* o, double a)
{
    double sigma = 0.5 * o->prm.stages;
    const int nc = o->nc;
#pragma omp parallel for
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < o->nv; k++) {
            double dt = o->dt_local == 1 ? o->dtL[i] : o->dt;
            o->Uaux[(size_t)k * nc + i] = o->U[(size_t)k * nc + i] - dt * sigma * a * o->R[(size_t)k * nc + i] / o->omega[i];
        }
}

static void step_rk(csu_oracle* o)
{
    static const double a1[] = {1.0}, a3[] = {0.1481, 0.4, 1.0}, a4[] = {0.0833, 0.2069, 0.4265, 1.0},
                        a5[] = {0.0533, 0.1263, 0.2375, 0.4414, 1.0};
    int s = o->prm.stages;
    const double* a = s == 1 ? a1 : s == 3 ? a3 : s == 4 ? a4 : a5;
    for (int k = 0; k < s; k++) {
        calc_R(o, k == 0 ? o->U : o->Uaux);
        rk_stage(o, a[k]);
    }
    double* t = o->U; o->U = o->Uaux; o->Uaux = t;
}


/* ---------------- implicit schemes: implicit.c ---------------- */
static double gas_T2gamma(const ogas* g, double T) { double Cv = gas_T2Cv(g, T); return (Cv + g->R) / Cv; }

/* implicitCalcD, implicit.c:19-181 */
static void implicit_D(csu_oracle* o)
{
    const int nc = o->nc, ns = o->nv + 1;
    const int visc = o->prm.laminar || o->prm.sa || o->prm.sst;
#pragma omp parallel for
    for (int i = 0; i < nc; i++) {
        double Lc = 0, Lv = 0;
        for (int j = 0; j < 4; j++) {
            int f = o->cface[4 * i + j];
            if (f == 0) continue;
            int f1 = f > 0 ? f - 1 : -f - 1;
            const int* q = o->face + 4 * f1;
            const double *P0 = o->P + (size_t)q[0] * ns, *P1 = o->P + (size_t)q[1] * ns;
            double dSx, dSy;
            face_ds(o, q[2], q[3], &dSx, &dSy);
            double dS = sqrt(dSx * dSx + dSy * dSy);
            double r = (P0[0] + P1[0]) * 0.5, u = (P0[1] + P1[1]) * 0.5, v = (P0[2] + P1[2]) * 0.5;
            double T = (P0[4] + P1[4]) * 0.5, c = gas_T2c(&o->gas, T);
            double nx = dSx / dS, ny = dSy / dS;
            Lc += (fabs(nx * u + ny * v) + c) * dS;
            if (o->prm.laminar) {
                double mi = sutherland(T);
                Lv += fmax(4 / (3 * r), gas_T2gamma(&o->gas, T) / r) * (mi / o->prm.pr) * dS * dS;
            }
            if (o->prm.sa || o->prm.sst) {
                double mi = sutherland(T);
                Lv += fmax(4 / (3 * r), gas_T2gamma(&o->gas, T) / r) * (mi / o->prm.pr + o->miT[f1] / o->prm.pr_t) * dS * dS;
            }
        }
        o->dtL[i] = Lc;
        o->D[i] = 0.5 * o->prm.wimp * Lc;
        if (visc) o->D[i] += Lv / o->omega[i];
    }
    for (int g = 0; g < o->nb; g++) {
        int e1 = o->bface[3 * g];
        const double* P1 = o->P + (size_t)e1 * ns;
        double dSx, dSy, nx, ny, Lv = 0;
        face_ds(o, o->bface[3 * g + 1], o->bface[3 * g + 2], &dSx, &dSy);
        double dS = sqrt(dSx * dSx + dSy * dSy);
        double r = P1[0], u = P1[1], v = P1[2], T = P1[4], c = gas_T2c(&o->gas, T);
        if (dS > 0) { nx = dSx / dS; ny = dSy / dS; } else { nx = 0.0; ny = 1.0; }
        double Lc = (fabs(nx * u + ny * v) + c) * dS;
        if (o->prm.laminar || o->prm.sa || o->prm.sst) {
            double mi = sutherland(T);
            Lv = fmax(4 / (3 * r), gas_T2gamma(&o->gas, T) / r) * (mi / o->prm.pr) * dS * dS;
        }
        o->dtL[e1] += Lc;
        o->D[e1] += 0.5 * o->prm.wimp * Lc;
        if (visc) o->D[e1] += Lv / o->omega[e1];
    }
    for (int i = 0; i < nc; i++) o->dtL[i] = o->omega[i] / o->dtL[i];
    double dt = o->dtL[0];
    for (int i = 1; i < nc; i++) dt = fmin(dt, o->dtL[i]);
    dt *= o->prm.cfl;
    for (int i = 0; i < nc; i++) o->D[i] += o->omega[i] / dt;
}

/* implicitAuxCalcFlux2 / implicitCalcDeltaFlux, implicit.c:183-231 */
static void aux_flux(const double* U, double p, double nx, double ny, double* F)
{
    double V = (nx * U[1] + ny * U[2]) / U[0];
    F[0] = U[0] * V;
    F[1] = U[1] * V + nx * p;
    F[2] = U[2] * V + ny * p;
    F[3] = (U[3] + p) * V;
}

static void delta_flux(const csu_oracle* o, const double* Pj, const double* d, double nx, double ny, double* dF)
{
    double rho = Pj[0], u = Pj[1], v = Pj[2], p = Pj[3], F0[4], F1[4], U[4];
    double T = p / (o->gas.R * rho);
    double e0 = gas_T2e(&o->gas, T);
    double E = e0 + (u * u + v * v) / 2;
    U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = rho * E;
    aux_flux(U, p, nx, ny, F0);
    for (int k = 0; k < 4; k++) U[k] = U[k] + d[k];
    rho = U[0]; u = U[1] / rho; v = U[2] / rho; E = U[3] / rho;
    double e1 = E - 0.5 * (u * u + v * v);
    T = T + (e1 - e0) / gas_T2Cv(&o->gas, T);       /* gasprop_e2Taprox */
    p = o->gas.R * T * rho;
    aux_flux(U, p, nx, ny, F1);
    for (int k = 0; k < 4; k++) dF[k] = F1[k] - F0[k];
}

/* implicitAuxCalcFlux_sa2 / implicitCalcDeltaFlux_sa, implicit.c:484-553 */
static void aux_flux_sst(const double* U, double p, double nx, double ny, double* F)   /* implicit.c:1036-1046 */
{
    double V = (nx * U[1] + ny * U[2]) / U[0];
    F[0] = U[0] * V;
    F[1] = U[1] * V + nx * p;
    F[2] = U[2] * V + ny * p;
    F[3] = (U[3] + p) * V;
    F[4] = U[4] * V;
    F[5] = U[5] * V;
}

static void aux_flux_sa(const double* U, double p, double nx, double ny, double* F)
{
    double V = (nx * U[1] + ny * U[2]) / U[0];
    F[0] = U[0] * V;
    F[1] = U[1] * V + nx * p;
    F[2] = U[2] * V + ny * p;
    F[3] = (U[3] + p) * V;
    F[4] = U[4] * V;
}

static void delta_flux_sa(const csu_oracle* o, const double* Pj, const double* d, double nx, double ny, double* dF)
{
    double rho = Pj[0], u = Pj[1], v = Pj[2], p = Pj[3], n = Pj[5], F0[5], F1[5], U[5];
    double T = p / (o->gas.R * rho);
    double e0 = gas_T2e(&o->gas, T);
    double E = e0 + (u * u + v * v) / 2;
    U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = rho * E; U[4] = rho * n;
    aux_flux_sa(U, p, nx, ny, F0);
    for (int k = 0; k < 5; k++) U[k] = U[k] + d[k];
    rho = U[0];
    rho = fmax(rho, o->prm.rlim);
    U[0] = rho;
    u = U[1] / rho; v = U[2] / rho; E = U[3] / rho;
    double e1 = E - (u * u + v * v) * 0.5;
    e1 = fmax(e1, 7.176325e+02);
    T = T + (e1 - e0) / gas_T2Cv(&o->gas, T);
    if (T < 1) {
        T = 1;
        e1 = gas_T2e(&o->gas, T);
        E = e1 + (u * u + v * v) * 0.5;
        U[3] = rho * E;
    }
    p = o->gas.R * T * rho;
    n = U[4] / rho;
    n = fmax(n, 0);
    U[4] = rho * n;
    aux_flux_sa(U, p, nx, ny, F1);
    for (int k = 0; k < 5; k++) dF[k] = F1[k] - F0[k];
}

/* implicitCalcDeltaFlux_sst (implicit.c:1048-1117), with its clamps */
static void delta_flux_sst(const csu_oracle* o, const double* Pj, const double* d, double nx, double ny, double* dF)
{
    double rho = Pj[0], u = Pj[1], v = Pj[2], p = Pj[3], k = Pj[5], om = Pj[6], F0[6], F1[6], U[6];
    double T = p / (o->gas.R * rho);
    double e0 = gas_T2e(&o->gas, T);
    double E = e0 + (u * u + v * v) * 0.5 + k;
    U[0] = rho; U[1] = rho * u; U[2] = rho * v; U[3] = rho * E; U[4] = rho * k; U[5] = rho * om;
    aux_flux_sst(U, p, nx, ny, F0);
    for (int q = 0; q < 6; q++) U[q] = U[q] + d[q];
    rho = U[0];
    rho = fmax(rho, o->prm.rlim);
    U[0] = rho;
    u = U[1] / rho; v = U[2] / rho;
    k = U[4] / rho;
    if (k < 1e-14) { k = 1e-14; U[4] = rho * k; }
    om = U[5] / rho;
    if (om < 0) { om = 0; U[5] = rho * om; }
    E = U[3] / rho;
    double e1 = E - (u * u + v * v) * 0.5 - k;
    e1 = fmax(e1, 7.176325e+02);
    T = T + (e1 - e0) / gas_T2Cv(&o->gas, T);        /* gasprop_e2Taprox, gasprop.c:156-159 */
    if (T < 1) {
        T = 1;
        e1 = gas_T2e(&o->gas, T);
        E = e1 + (u * u + v * v) * 0.5 + k;
        U[3] = rho * E;
    }
    p = o->gas.R * T * rho;
    aux_flux_sst(U, p, nx, ny, F1);
    for (int q = 0; q < 6; q++) dF[q] = F1[q] - F0[q];
}

/* face j of cell i seen from i: neighbour, outward normal (implicit.c:288-305) */
static int face_from_id(const csu_oracle* o, int i, int slot, double* dSx, double* dSy, int* fid)
{
    int f = o->cface[4 * i + slot];
    if (f == 0) return -1;
    *fid = f > 0 ? f - 1 : -f - 1;
    const int* q = o->face + 4 * (f > 0 ? f - 1 : -f - 1);
    if (f > 0) { face_ds(o, q[2], q[3], dSx, dSy); return q[1]; }
    face_ds(o, q[3], q[2], dSx, dSy);
    return q[0];
}

static int face_from(const csu_oracle* o, int i, int slot, double* dSx, double* dSy)
{
    int fid;
    return face_from_id(o, i, slot, dSx, dSy, &fid);
}

/* implicitFunc, implicit.c:233-272: dW[.][i] -= 0.5 (dF dS - ra dW[.][j]) */
static void implicit_func(const csu_oracle* o, int i, int j, int fid, double dSx, double dSy, double* dW)
{
    const int nc = o->nc, ns = o->nv + 1;
    double dS = sqrt(dSx * dSx + dSy * dSy), nx = dSx / dS, ny = dSy / dS;
    const double* Pj = o->P + (size_t)j * ns;
    double d[6] = {dW[j], dW[nc + j], dW[2 * nc + j], dW[3 * nc + j], o->nv > 4 ? dW[4 * nc + j] : 0.0,
                   o->nv > 5 ? dW[5 * nc + j] : 0.0}, dF[6];
    if (o->nv == 6) delta_flux_sst(o, Pj, d, nx, ny, dF);
    else if (o->nv == 5) delta_flux_sa(o, Pj, d, nx, ny, dF);
    else delta_flux(o, Pj, d, nx, ny, dF);
    double c = gas_T2c(&o->gas, Pj[4]);
    double ra = o->prm.wimp * (fabs(nx * Pj[1] + ny * Pj[2]) + c) * dS;
    if (o->prm.sa || o->prm.sst) {   /* implicitFunc_sa / _sst, implicit.c:555-595, 1120-1161 */
        double r = Pj[0], T = Pj[4], mi = sutherland(T), x0, y0, x1, y1;
        cell_center(o, i, &x0, &y0);
        cell_center(o, j, &x1, &y1);
        double dd = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
        ra += fmax(4 / (3 * r), gas_T2gamma(&o->gas, T) / r) * (mi / o->prm.pr + o->miT[fid] / o->prm.pr_t) * dS / dd;
    } else if (o->prm.laminar) {
        double r = Pj[0], T = Pj[4], mi = sutherland(T), x0, y0, x1, y1;
        cell_center(o, i, &x0, &y0);
        cell_center(o, j, &x1, &y1);
        double dd = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
        ra += fmax(4 / (3 * r), gas_T2gamma(&o->gas, T) / r) * (mi / o->prm.pr) * dS / dd;
    }
    for (int k = 0; k < o->nv; k++) dW[(size_t)k * nc + i] -= 0.5 * (dF[k] * dS - ra * dW[(size_t)k * nc + j]);
}

/* implicitLUSGS_L / _U, implicit.c:274-482: serial sweeps in cell order */
/* diagonal solve of cell i: SST couples k and omega through the source
   Jacobians, a 2x2 block solved by Cramer's rule with a long double
   determinant (implicit.c:326-351, 447-472) */
static void lusgs_diag(const csu_oracle* o, int i, double* dW)
{
    const int nc = o->nc;
    const double D = o->D[i];
    if (o->nv != 6) {
        for (int k = 0; k < o->nv; k++) dW[(size_t)k * nc + i] /= D;
        return;
    }
    for (int k = 0; k < 4; k++) dW[(size_t)k * nc + i] /= D;
    double A = dW[4 * (size_t)nc + i] + o->dQkdr[i] * dW[i];
    double B = dW[5 * (size_t)nc + i] + o->dQodr[i] * dW[i];
    double a = D - o->dQkdrk[i], b = -o->dQkdro[i], c = -o->dQodrk[i], d = D - o->dQodro[i];
    long double det = a * d - b * c;
    if (fabs(det) < 1e-14) {
        dW[4 * (size_t)nc + i] /= D;
        dW[5 * (size_t)nc + i] /= D;
    } else {
        dW[4 * (size_t)nc + i] = (A * d - B * b) / det;
        dW[5 * (size_t)nc + i] = (-A * c + B * a) / det;
    }
}

static void lusgs(csu_oracle* o)
{
    const int nc = o->nc, nv = o->nv;
    for (int i = 0; i < nc; i++) {
        for (int k = 0; k < nv; k++) o->dW0[(size_t)k * nc + i] = -o->R[(size_t)k * nc + i];
        for (int s = 0; s < 4; s++) {
            double dSx, dSy;
            int fid, j = face_from_id(o, i, s, &dSx, &dSy, &fid);
            if (j >= 0 && j < i) implicit_func(o, i, j, fid, dSx, dSy, o->dW0);
        }
        lusgs_diag(o, i, o->dW0);
    }
    for (int i = 0; i < nc; i++) {
        for (int k = 0; k < nv; k++) o->dW1[(size_t)k * nc + i] = o->D[i] * o->dW0[(size_t)k * nc + i];
        if (nv == 6) {       /* implicit.c:372-392 */
            const double D = o->D[i], *x = o->dW0;
            double a = D - o->dQkdrk[i], b = -o->dQkdro[i], c = -o->dQodrk[i], d = D - o->dQodro[i];
            long double det = a * d - b * c;
            if (!(fabs(det) < 1e-14)) {
                o->dW1[4 * (size_t)nc + i] = a * x[4 * (size_t)nc + i] + b * x[5 * (size_t)nc + i] - o->dQkdr[i] * x[i];
                o->dW1[5 * (size_t)nc + i] = c * x[4 * (size_t)nc + i] + d * x[5 * (size_t)nc + i] - o->dQodr[i] * x[i];
            }
        }
    }
    for (int i = nc - 1; i >= 0; i--) {
        for (int s = 0; s < 4; s++) {
            double dSx, dSy;
            int fid, j = face_from_id(o, i, s, &dSx, &dSy, &fid);
            if (j > i) implicit_func(o, i, j, fid, dSx, dSy, o->dW1);
        }
        lusgs_diag(o, i, o->dW1);
    }
}

/* implicitUpdateA, implicit.c:689-813: finite-difference flux Jacobian
   blocks of every interior neighbour, h = 1e-6 */
/* SST (nv 6) has no DPLUR in the reference: implicitCalcDPLUR builds 4x4
   blocks for it (implicit.c:993-1000) and implicitMultA then reads past
   them.  The device library's SST form, restated here: the blocks of
   implicitUpdateA_sa with two convected scalars (implicitAuxCalcFlux_sst2,
   implicit.c:1036-1046) and k taken out of the internal energy as
   implicitCalcDeltaFlux_sst does (:1048-1117); the diagonal is LU-SGS's
   k-omega block with the source Jacobians (implicit.c:326-351). */
static void dplur_blocks(csu_oracle* o)
{
    const int nc = o->nc, nv = o->nv, ns = nv + 1;
    const double h = 1.0e-6;
#pragma omp parallel for
    for (int i = 0; i < nc; i++) {
        for (int s = 0; s < 4; s++) {
            double dSx, dSy;
            int fid, j = face_from_id(o, i, s, &dSx, &dSy, &fid);
            if (j < 0) continue;
            double* A = o->A + ((size_t)i * 4 + s) * nv * nv;
            double dS = sqrt(dSx * dSx + dSy * dSy), nx = dSx / dS, ny = dSy / dS;
            const double* Pj = o->P + (size_t)j * ns;
            double Uj[6] = {0}, F0[6], F[6], Up[6];
            for (int k = 0; k < nv; k++) Uj[k] = o->U[(size_t)k * nc + j];
            double rho = Uj[0], u = Uj[1] / rho, v = Uj[2] / rho, E = Uj[3] / rho;
            if (nv == 6) aux_flux_sst(Uj, Pj[3], nx, ny, F0);
            else if (nv == 5) aux_flux_sa(Uj, Pj[3], nx, ny, F0);
            else aux_flux(Uj, Pj[3], nx, ny, F0);
            double ee0 = E - 0.5 * (u * u + v * v);
            if (nv == 6) ee0 = ee0 - Uj[4] / rho;
            for (int m = 0; m < nv; m++) {
                for (int n = 0; n < nv; n++) Up[n] = m == n ? Uj[n] + h : Uj[n];
                rho = Up[0]; u = Up[1] / rho; v = Up[2] / rho; E = Up[3] / rho;
                double ee1 = E - 0.5 * (u * u + v * v);
                if (nv == 6) ee1 = ee1 - Up[4] / rho;
                double T = Pj[4] + (ee1 - ee0) / gas_T2Cv(&o->gas, Pj[4]);
                double p = o->gas.R * rho * T;
                if (nv == 6) aux_flux_sst(Up, p, nx, ny, F);
                else if (nv == 5) aux_flux_sa(Up, p, nx, ny, F);
                else aux_flux(Up, p, nx, ny, F);
                for (int n = 0; n < nv; n++) A[n * nv + m] = 0.5 * (F[n] - F0[n]) * dS / h;
            }
            double c = gas_T2c(&o->gas, Pj[4]);
            double ra = o->prm.wimp * (fabs(nx * Pj[1] + ny * Pj[2]) + c) * dS;
            if (o->prm.sa || o->prm.sst) {     /* implicitUpdateA_sa, implicit.c:815-936 */
                const double* Pi = o->P + (size_t)i * ns;
                double r = (Pi[0] + Pj[0]) * 0.5, T = (Pi[4] + Pj[4]) * 0.5, mi = sutherland(T), x0, y0, x1, y1;
                cell_center(o, i, &x0, &y0);
                cell_center(o, j, &x1, &y1);
                double dd = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
                ra += fmax(4 / (3 * r), gas_T2gamma(&o->gas, T) / r) * (mi / o->prm.pr + o->miT[fid] / o->prm.pr_t) * dS / dd;
            } else if (o->prm.laminar) {
                const double* Pi = o->P + (size_t)i * ns;
                double r = (Pi[0] + Pj[0]) * 0.5, T = (Pi[4] + Pj[4]) * 0.5, mi = sutherland(T), x0, y0, x1, y1;
                cell_center(o, i, &x0, &y0);
                cell_center(o, j, &x1, &y1);
                double dd = sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0));
                ra += fmax(4 / (3 * r), gas_T2gamma(&o->gas, T) / r) * (mi / o->prm.pr) * dS / dd;
            }
            for (int n = 0; n < nv; n++) A[n * nv + n] -= 0.5 * ra;
        }
    }
}

/* SST: the k-omega 2x2 of the DPLUR block diagonal; 0 where its
   determinant is below 1e-14 (D alone there, as lusgs_diag) */
static int sst_kw_block(const csu_oracle* o, int i, double* a, double* b, double* c, double* d, double* det)
{
    *a = o->D[i] - o->dQkdrk[i];
    *b = -o->dQkdro[i];
    *c = -o->dQodrk[i];
    *d = o->D[i] - o->dQodro[i];
    *det = *a * *d - *b * *c;
    return !(fabs(*det) < 1e-14);
}

/* implicitCalcDPLUR, implicit.c:990-1034: damped Jacobi, w = 2/3 */
static void dplur(csu_oracle* o)
{
    const int nc = o->nc, nv = o->nv;
    implicit_D(o);
    dplur_blocks(o);
    const double w = 2. / 3.;
    for (int mm = 0; mm < o->prm.nlinear; mm++) {
        double* t = o->dW1; o->dW1 = o->dW0; o->dW0 = t;
        const double* x = o->dW0;
        double* y = o->dW1;
#pragma omp parallel for
        for (int i = 0; i < nc; i++) {
            for (int k = 0; k < nv; k++) y[(size_t)k * nc + i] = o->D[i] * x[(size_t)k * nc + i];
            double a, b, c, d, det;
            if (nv == 6 && sst_kw_block(o, i, &a, &b, &c, &d, &det)) {   /* k-omega rows of the block diagonal */
                y[4 * (size_t)nc + i] = a * x[4 * (size_t)nc + i] + b * x[5 * (size_t)nc + i] - o->dQkdr[i] * x[i];
                y[5 * (size_t)nc + i] = c * x[4 * (size_t)nc + i] + d * x[5 * (size_t)nc + i] - o->dQodr[i] * x[i];
            }
            for (int s = 0; s < 4; s++) {
                double dSx, dSy;
                int j = face_from(o, i, s, &dSx, &dSy);
                if (j < 0) continue;
                const double* A = o->A + ((size_t)i * 4 + s) * nv * nv;
                for (int k = 0; k < nv; k++)
                    for (int n = 0; n < nv; n++) y[(size_t)k * nc + i] += A[k * nv + n] * x[(size_t)n * nc + j];
            }
        }
#pragma omp parallel for
        for (int i = 0; i < nc; i++) {
            double a, b, c, d, det;
            if (nv == 6 && sst_kw_block(o, i, &a, &b, &c, &d, &det)) {
                double z[6];
                for (int k = 0; k < 4; k++) z[k] = (-o->R[(size_t)k * nc + i] - y[(size_t)k * nc + i]) / o->D[i];
                double A = -o->R[4 * (size_t)nc + i] - y[4 * (size_t)nc + i] + o->dQkdr[i] * z[0];
                double B = -o->R[5 * (size_t)nc + i] - y[5 * (size_t)nc + i] + o->dQodr[i] * z[0];
                z[4] = (A * d - B * b) / det;
                z[5] = (-A * c + B * a) / det;
                for (int k = 0; k < nv; k++) y[(size_t)k * nc + i] = w * z[k] + x[(size_t)k * nc + i];
                continue;
            }
            for (int k = 0; k < nv; k++)
                y[(size_t)k * nc + i] = w * (-o->R[(size_t)k * nc + i] - y[(size_t)k * nc + i]) / o->D[i] + x[(size_t)k * nc + i];
        }
        double ans = fabs(y[0] - x[0]);
        for (int i = 1; i < nc; i++) ans = fmax(ans, fabs(y[i] - x[i]));
        if (ans < 1e-14) break;
    }
}

static void update_implicit(csu_oracle* o)   /* solverUpdateUImplicit, solver.c:1945-1955 */
{
    size_t n = (size_t)o->nv * o->nc;
    for (size_t t = 0; t < n; t++) o->U[t] += o->dW1[t];
}

/* ---------------- residual norm and forces ---------------- */
static void calc_res(const csu_oracle* o, double* res)       /* solver.c:867-892 */
{
    for (int k = 0; k < o->nv; k++) res[k] = 0.0;
    for (int i = 0; i < o->nc; i++)
        for (int k = 0; k < o->nv; k++) res[k] += o->R[(size_t)k * o->nc + i] * o->R[(size_t)k * o->nc + i];
    for (int k = 0; k < o->nv; k++) res[k] = sqrt(res[k] / o->nc);
}

static void calc_coeff(const csu_oracle* o, double* c4)      /* solver.c:1342-1401, pressure part */
{
    const double* Pin = o->prm.pin;
    double q = 0.5 * Pin[0] * (Pin[1] * Pin[1] + Pin[2] * Pin[2]);
    double cxp = 0, cxv = 0, cyp = 0, cyv = 0;
    const int ns = o->nv + 1;
    for (int g = 0; g < o->nb; g++) {
        int t = o->btype[g];
        if (t != CSU_BC_WALL && t != CSU_BC_WALLT) continue;
        double cp = (o->Pb[(size_t)g * ns + 3] - Pin[3]) / q, dSx, dSy;
        face_ds(o, o->bface[3 * g + 1], o->bface[3 * g + 2], &dSx, &dSy);
        cxp += cp * dSx;
        cyp += cp * dSy;
        if (o->prm.laminar || o->prm.sa || o->prm.sst) {
            /* boundaryCalcFrictionWall, boundary.c:695-742 */
            double du[2], dv[2], L, dxy[2];
            wall_grads(o, g, 0.0, 0.0, du, dv, &L, dxy);
            double mi = sutherland(o->P[(size_t)o->bface[3 * g] * ns + 4]);
            double txx = 2 * mi * (du[0] - (du[0] + dv[1]) / 3), tyy = 2 * mi * (dv[1] - (du[0] + dv[1]) / 3);
            double txy = mi * (du[1] + dv[0]);
            cxv -= (txx * dSx + txy * dSy) / q;
            cyv -= (txy * dSx + tyy * dSy) / q;
        }
    }
    cxp /= o->prm.sref; cxv /= o->prm.sref; cyp /= o->prm.sref; cyv /= o->prm.sref;
    if (o->axi) { cxp *= 2 * M_PI; cxv *= 2 * M_PI; }
    c4[0] = cxp; c4[1] = cxv; c4[2] = cyp; c4[3] = cyv;
}

/* ---------------- API ---------------- */
int oracle_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

static void* dup(const void* p, size_t n) { void* q = malloc(n ? n : 1); if (n) memcpy(q, p, n); return q; }

int oracle_create(const csu_mesh_desc* m, const csu_params* p, csu_oracle** out)
{
    if ((p->sa || p->sst) && !m->wall_dist) return CSU_ERR_UNSUPPORTED;
    /* SST: no global-range limiter (limiter.c:37-38 sizes its range arrays
       for 5 variables) */
    if (p->sst && p->order == 2 && p->limiter == 1)
        return CSU_ERR_UNSUPPORTED;
    /* Green-Gauss: planar runs only */
    if (p->gradient == 1 && m->axi == 1) return CSU_ERR_UNSUPPORTED;
    csu_oracle* o = calloc(1, sizeof(*o));
    o->nc = m->ncell; o->nf = m->nface; o->nb = m->nbface; o->nv = p->nvar; o->nn = m->nnode; o->axi = m->axi;
    o->prm = *p; o->dt_local = p->dt_local;
    gas_init(&o->gas, p->gamma, p->gas_r, p->tp);
    const int nc = o->nc, nv = o->nv;
    o->xy = dup(m->node_xy, sizeof(double) * 2 * (size_t)m->nnode);
    o->cv = dup(m->cell_vert, sizeof(int) * 4 * (size_t)nc);
    o->face = dup(m->face, sizeof(int) * 4 * (size_t)o->nf);
    o->bface = dup(m->bface, sizeof(int) * 3 * (size_t)o->nb);
    o->nbr = dup(m->cell_nbr, sizeof(int) * 4 * (size_t)nc);
    o->cface = dup(m->cell_face, sizeof(int) * 4 * (size_t)nc);
    o->btype = malloc(sizeof(int) * (size_t)(o->nb ? o->nb : 1));
    for (int g = 0; g < o->nb; g++) o->btype[g] = m->mark_type[m->bface_mark[g]];
    o->omega = malloc(sizeof(double) * nc);
    for (int i = 0; i < nc; i++) o->omega[i] = cell_omega(o, i);
    size_t nvc = (size_t)nv * nc;
    o->U = malloc(sizeof(double) * nvc); o->Uaux = calloc(nvc, sizeof(double));
    o->R = calloc(nvc, sizeof(double)); o->dPx = calloc(nvc, sizeof(double));
    o->dPy = calloc(nvc, sizeof(double)); o->phi = calloc(nvc, sizeof(double));
    o->flux = calloc((size_t)nv * (o->nf ? o->nf : 1), sizeof(double));
    o->dtL = calloc(nc, sizeof(double));
    o->P = calloc((size_t)(nv + 1) * nc, sizeof(double));
    o->Pb = calloc((size_t)(nv + 1) * (o->nb ? o->nb : 1), sizeof(double));
    o->D = calloc(nc, sizeof(double));
    o->miT = calloc((size_t)(o->nf ? o->nf : 1), sizeof(double));
    o->wd = calloc(nc, sizeof(double));
    if (m->wall_dist) memcpy(o->wd, m->wall_dist, sizeof(double) * nc);
    double* sq[6];
    for (int q = 0; q < 6; q++) sq[q] = calloc(nc, sizeof(double));
    o->dQkdr = sq[0]; o->dQkdrk = sq[1]; o->dQkdro = sq[2]; o->dQodr = sq[3]; o->dQodrk = sq[4]; o->dQodro = sq[5];
    {
        const double kap = 0.41;
        o->sst_g1 = SST_B1 / SST_BS - SST_SO1 * kap * kap / sqrt(SST_BS);
        o->sst_g2 = SST_B2 / SST_BS - SST_SO2 * kap * kap / sqrt(SST_BS);
    }
    o->dW0 = calloc(nvc, sizeof(double));
    o->dW1 = calloc(nvc, sizeof(double));      /* implicitInitDPLUR zeroes dW1 */
    if (p->time_scheme == CSU_TIME_DPLUR) o->A = calloc((size_t)nc * 4 * nv * nv, sizeof(double));
    for (int k = 0; k < nv; k++)
        for (int i = 0; i < nc; i++) o->U[(size_t)k * nc + i] = p->uin[k];
    *out = o;
    return CSU_OK;
}

void oracle_destroy(csu_oracle* o)
{
    if (!o) return;
    free(o->xy); free(o->cv); free(o->face); free(o->bface); free(o->nbr); free(o->cface); free(o->btype);
    free(o->omega); free(o->U); free(o->Uaux); free(o->R); free(o->dPx); free(o->dPy); free(o->phi);
    free(o->flux); free(o->dtL); free(o->P); free(o->Pb);
    free(o->D); free(o->dW0); free(o->dW1); free(o->A); free(o->miT); free(o->wd);
    free(o->dQkdr); free(o->dQkdrk); free(o->dQkdro); free(o->dQodr); free(o->dQodrk); free(o->dQodro);
    free(o);
}

int oracle_set_state(csu_oracle* o, const double* U)
{
    memcpy(o->U, U, sizeof(double) * (size_t)o->nv * o->nc);
    return CSU_OK;
}

int oracle_get_state(csu_oracle* o, double* U)
{
    memcpy(U, o->U, sizeof(double) * (size_t)o->nv * o->nc);
    return CSU_OK;
}

int oracle_get_dt(csu_oracle* o, double* dtl)
{
    memcpy(dtl, o->dtL, sizeof(double) * (size_t)o->nc);
    return CSU_OK;
}

int oracle_residual(csu_oracle* o, double* R, double* dPx, double* dPy, double* phi, double* P, double* Pb)
{
    calc_R(o, o->U);
    size_t nvc = sizeof(double) * (size_t)o->nv * o->nc;
    if (R) memcpy(R, o->R, nvc);
    if (dPx) memcpy(dPx, o->dPx, nvc);
    if (dPy) memcpy(dPy, o->dPy, nvc);
    if (phi) memcpy(phi, o->phi, nvc);
    if (P) memcpy(P, o->P, sizeof(double) * (size_t)(o->nv + 1) * o->nc);
    if (Pb) memcpy(Pb, o->Pb, sizeof(double) * (size_t)(o->nv + 1) * o->nb);
    return CSU_OK;
}

/* solverCalcPrimitive (solver.c:1153-1220) on its own, as solverWriteSolution2
   calls it (solver.c:2035): the clamps write back into U and the ghost
   states follow the clamped cells; the gradients of the last residual stay. */
int oracle_primitives(csu_oracle* o, double* P)
{
    calc_primitive(o, o->U);
    if (P) memcpy(P, o->P, sizeof(double) * (size_t)(o->nv + 1) * o->nc);
    return CSU_OK;
}

/* sstInterMiT (sst.c:964-1007): sstSources per cell on the primitives and
   the last residual's gradients of u, v, k, omega, without the omega = 0
   guard of the residual's source term. */
int oracle_sst_fields(const csu_oracle* o, double* out)
{
    if (!o->prm.sst) return CSU_ERR_UNSUPPORTED;
    const size_t nc = o->nc;
    const double *gx = o->dPx, *gy = o->dPy;
    for (size_t i = 0; i < nc; i++) {
        const double* P = o->P + i * (o->nv + 1);
        sstv v;
        memset(&v, 0, sizeof v);
        v.dux = gx[nc + i]; v.dvx = gx[2 * nc + i]; v.dkx = gx[4 * nc + i]; v.dox = gx[5 * nc + i];
        v.duy = gy[nc + i]; v.dvy = gy[2 * nc + i]; v.dky = gy[4 * nc + i]; v.doy = gy[5 * nc + i];
        v.r = P[0]; v.T = P[4]; v.k = P[5]; v.om = P[6]; v.d = o->wd[i];
        v.mi_L = sutherland(v.T);
        sst_sources(o, &v);
        double* q = out + 5 * i;
        q[0] = v.mi_t; q[1] = v.F1; q[2] = v.F2; q[3] = v.d; q[4] = v.om;
    }
    return CSU_OK;
}

/* Per boundary face, what the surfData.csv writers take (saSolverWriteSurf
   sa.c:854-960, saCC_SolverWriteSurf saCC.c:599-705, sstSolverWriteSurf
   sst.c:820-962): the boundary viscous flux (reference f[1..3]) and eddy
   viscosity of saBoundaryFaceViscFlux / saCC_BoundaryFaceViscFlux /
   sstBoundaryFaceViscFlux on the current primitives and the last residual's
   gradients, the ghost state of the last residual (r u v p T and nu~ or
   k) and the wall distance of the face's cell.  out [nbface][11]. */
int oracle_surf_faces(const csu_oracle* o, double* out)
{
    if (!(o->prm.sa || o->prm.sst)) return CSU_ERR_UNSUPPORTED;
    const int ns = o->nv + 1;
    for (int g = 0; g < o->nb; g++) {
        double f[6] = {0, 0, 0, 0, 0, 0}, mie = 0.0;
        if (o->prm.sst) sst_bc_flux(o, g, f, &mie);
        else sa_bc_flux(o, g, f, &mie);
        double* q = out + 11 * (size_t)g;
        q[0] = f[1]; q[1] = f[2]; q[2] = f[3]; q[3] = mie;
        for (int m = 0; m < 6; m++) q[4 + m] = o->Pb[(size_t)g * ns + m];
        q[10] = o->wd[o->bface[3 * g]];
    }
    return CSU_OK;
}

int oracle_iterate(csu_oracle* o, int niter, double* res, double* coef)
{
    for (int it = 0; it < niter; it++) {
        if (o->prm.time_scheme == CSU_TIME_RK) {
            calc_dt(o);
            step_rk(o);
        } else if (o->prm.time_scheme == CSU_TIME_LUSGS) {
            calc_R(o, o->U);
            implicit_D(o);
            lusgs(o);
            update_implicit(o);
        } else {
            calc_R(o, o->U);
            dplur(o);
            update_implicit(o);
        }
        if (o->iter == o->prm.dt_local_n) o->dt_local = 0;   /* solver.c:1886-1889 */
        if (res) calc_res(o, res + (size_t)it * o->nv);
        if (coef) calc_coeff(o, coef + (size_t)it * 4);
        o->iter++;
    }
    return CSU_OK;
}

/* Test hook, no reference counterpart: the DPLUR system of one implicit
   step from the current state (solverCalcR, implicitCalcD, the blocks) and
   the sweeps' result.  D [nc] diagonal with the time-step term; A
   [nc][4][nv*nv] neighbour blocks (row-major, y_i += A x_j); nb [nc][4]
   neighbour of each slot (-1: none); nrm [nc][4][3] its outward unit
   normal and length; dq [6][nc] the SST source Jacobians dQk/d(r,rk,ro),
   dQo/d(r,rk,ro) (zeros otherwise); R [nv][nc]; dW [nv][nc] after the
   sweeps.  The state is not advanced (only clamped, as calc_R does). */
int oracle_face_flux(const csu_oracle* o, const double* L, const double* R, double* f)
{
    if (!o || !L || !R || !f) return CSU_ERR_ARG;
    flux(o, L, R, f);
    return CSU_OK;
}

int oracle_dplur_system(csu_oracle* o, double* D, double* A, int* nb, double* nrm, double* dq, double* R,
                        double* dW)
{
    if (o->prm.time_scheme != CSU_TIME_DPLUR) return CSU_ERR_ARG;
    const int nc = o->nc, nv = o->nv;
    calc_R(o, o->U);
    dplur(o);
    memcpy(D, o->D, sizeof(double) * nc);
    memcpy(A, o->A, sizeof(double) * (size_t)nc * 4 * nv * nv);
    memcpy(R, o->R, sizeof(double) * (size_t)nv * nc);
    memcpy(dW, o->dW1, sizeof(double) * (size_t)nv * nc);
    for (int i = 0; i < nc; i++)
        for (int s = 0; s < 4; s++) {
            double dSx = 0, dSy = 0;
            int j = face_from(o, i, s, &dSx, &dSy);
            nb[4 * i + s] = j;
            double dS = j >= 0 ? sqrt(dSx * dSx + dSy * dSy) : 0.0;
            double* q = nrm + 3 * (4 * (size_t)i + s);
            q[0] = j >= 0 ? dSx / dS : 0.0;
            q[1] = j >= 0 ? dSy / dS : 0.0;
            q[2] = dS;
        }
    for (int k = 0; k < 6 * nc; k++) dq[k] = 0.0;
    if (nv == 6)
        for (int i = 0; i < nc; i++) {
            dq[i] = o->dQkdr[i]; dq[nc + i] = o->dQkdrk[i]; dq[2 * nc + i] = o->dQkdro[i];
            dq[3 * nc + i] = o->dQodr[i]; dq[4 * nc + i] = o->dQodrk[i]; dq[5 * nc + i] = o->dQodro[i];
        }
    return CSU_OK;
}

/* CFL change between iterations: the next solverCalcDt / implicitCalcD uses it */
int oracle_set_cfl(csu_oracle* o, double cfl)
{
    if (!(cfl > 0.0)) return CSU_ERR_ARG;
    o->prm.cfl = cfl;
    return CSU_OK;
}

/* Time-accurate run, the tube branch of solverSolve (solver.c:1912-1941):
   global step from solverCalcDt, cut to tmax - t when t + dt passes tmax
   (that step is the last), solverStepRK, t += dt*stages/2.  The reference
   loops without a cap and with no check of dt; here a non-positive or
   non-finite step is an error and max_iter > 0 caps the steps. */
int oracle_advance_time(csu_oracle* o, double tmax, int max_iter, double* t, int* niter, int* finished)
{
    if (o->prm.time_scheme != CSU_TIME_RK || o->dt_local == 1) return CSU_ERR_UNSUPPORTED;
    double tt = *t;
    int n = 0, stop = 0, rc = CSU_OK;
    while (!stop && !(max_iter > 0 && n >= max_iter)) {
        calc_dt(o);
        if (!(o->dt > 0.0) || !isfinite(o->dt)) { rc = CSU_ERR_ARG; break; }
        if (tt + o->dt > tmax) {
            o->dt = tmax - tt;
            stop = 1;
        }
        step_rk(o);
        tt += o->dt * o->prm.stages / 2.0;
        n++;
        o->iter++;
    }
    *t = tt;
    if (niter) *niter = n;
    if (finished) *finished = stop;
    return rc;
}

/* Wall shear of the last residual evaluation, solverCalcCoeff2
   (solver.c:1303-1340): boundaryCalcTensorWall (boundary.c:744-793) for
   laminar runs, saCalcTensorWall (sa.c:795-852) for SA and SST.
   out: [nb][4] = txx txy tyy yp, rows of wall / wallT faces only. */
int oracle_wall_tensor(const csu_oracle* o, double* out)
{
    const int nc = o->nc, ns = o->nv + 1;
    if (!(o->prm.laminar || o->prm.sa || o->prm.sst)) return CSU_ERR_ARG;
    for (int g = 0; g < o->nb; g++) {
        int t = o->btype[g];
        if (t != CSU_BC_WALL && t != CSU_BC_WALLT) continue;
        const int e0 = o->bface[3 * g], p0 = o->bface[3 * g + 1], p1 = o->bface[3 * g + 2];
        const double* P = o->P + (size_t)e0 * ns;
        double dSx, dSy, x0, y0;
        face_ds(o, p0, p1, &dSx, &dSy);
        cell_center(o, e0, &x0, &y0);
        double x1 = (o->xy[2 * p0] + o->xy[2 * p1]) / 2, y1 = (o->xy[2 * p0 + 1] + o->xy[2 * p1 + 1]) / 2;
        double dx = x1 - x0, dy = y1 - y0, L = sqrt(dx * dx + dy * dy);
        double dul = (0 - P[1]) / L, dvl = (0 - P[2]) / L;
        double duxm = o->dPx[nc + e0], dvxm = o->dPx[2 * nc + e0];
        double duym = o->dPy[nc + e0], dvym = o->dPy[2 * nc + e0];
        double dux = duxm + (dul - (duxm * dx + duym * dy) / L) * dx / L;
        double duy = duym + (dul - (duxm * dx + duym * dy) / L) * dy / L;
        double dvx = dvxm + (dvl - (dvxm * dx + dvym * dy) / L) * dx / L;
        double dvy = dvym + (dvl - (dvxm * dx + dvym * dy) / L) * dy / L;
        double mi = sutherland(P[4]);
        double txx = 2 * mi * (dux - (dux + dvy) / 3), tyy = 2 * mi * (dvy - (dux + dvy) / 3), txy = mi * (duy + dvx);
        double* w = out + 4 * (size_t)g;
        if (o->prm.laminar) {
            double aux = 5.29 * x1 / sqrt(P[0] * P[1] * x1 / mi);
            w[0] = txx; w[1] = txy; w[2] = tyy; w[3] = L / aux;
        } else {
            double dS = sqrt(dSx * dSx + dSy * dSy);
            double fx = (txx * dSx + txy * dSy) / dS, fy = (txy * dSx + tyy * dSy) / dS;
            double ft = (fx * dSy - fy * dSx) / dS;
            double ut = sqrt(fabs(ft) / P[0]);
            w[0] = 0.0; w[1] = ft; w[2] = 0.0; w[3] = P[0] * o->wd[e0] * ut / mi;
        }
    }
    return CSU_OK;
}
