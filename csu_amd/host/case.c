/*
 * case.c -- case loading: input.dat options, free-stream state, initial
 * state or restart, mesh.
 *
 * Follows the reference's solverInit / solverSetData / solverInitDomain
 * (solver.c:1403-1834) for which keys are read, their defaults and the
 * free-stream state (conditionState, solver.c:40-89), so that one case
 * directory means the same run to both programs.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "csu_host.h"

#include "hgas.h"

static double sutherland(double T) { return 1.458e-6 * T * sqrt(T) / (T + 110.4); }

/* ---- options ---- */
static char g_missing[128];

static const char* need(const csu_input* in, const char* key)
{
    const char* v = csu_input_get(in, key);
    if (!v && !g_missing[0]) snprintf(g_missing, sizeof g_missing, "%s", key);
    return v ? v : "0";
}

static int opt_i(const csu_input* in, const char* key, int def)
{
    const char* v = csu_input_get(in, key);
    return v ? atoi(v) : def;
}

static double opt_d(const csu_input* in, const char* key, double def)
{
    const char* v = csu_input_get(in, key);
    return v ? strtod(v, NULL) : def;
}

static int choice(const char* s, const char* const* names, int n)
{
    for (int i = 0; i < n; i++) if (strcmp(s, names[i]) == 0) return i;
    return -1;
}

const char* csu_host_missing_key(void) { return g_missing; }

/* conditionState (solver.c:40-75) plus the SA/SST inlet values
   (saInitU, sstInitU): primitive pin[7] and conservative uin[6] */
static void condition_state(const hgas* g, const csu_params* p, double pinf, double Tinf, double mach, double nx,
                            double ny, double* pin, double* uin)
{
    double r = pinf / (g->R * Tinf);
    double cs = hgas_T2c(g, Tinf);
    double u = nx * mach * cs;
    double v = ny * mach * cs;
    double k = 0.0;
    for (int i = 0; i < 7; i++) pin[i] = 0.0;
    for (int i = 0; i < 6; i++) uin[i] = 0.0;
    if (p->sst == 1) {
        double L = p->sst_l, Uv = cs * mach;
        double ReL = (r * Uv * L) / sutherland(Tinf);
        k = p->sst_kfactor * Uv * Uv / ReL;
        double om = p->sst_ofactor * Uv / L;
        uin[4] = r * k; uin[5] = r * om;
        pin[5] = k; pin[6] = om;
    }
    double E = hgas_T2e(g, Tinf) + (u * u + v * v) / 2 + k;
    uin[0] = r; uin[1] = r * u; uin[2] = r * v; uin[3] = r * E;
    pin[0] = r; pin[1] = u; pin[2] = v; pin[3] = g->R * Tinf * r; pin[4] = Tinf;
    if (p->sa == 1) {
        double n = p->turb_ratio * sutherland(Tinf) / r;
        uin[4] = r * n;
        pin[5] = n;
    }
}

int csu_case_load(const char* wd, csu_case** out)
{
    csu_case* c = calloc(1, sizeof(csu_case));
    if (!c) return CSU_ERR_NOMEM;
    snprintf(c->wd, sizeof c->wd, "%s", wd);
    char path[1200];
    g_missing[0] = '\0';

    snprintf(path, sizeof path, "%sinput.dat", wd);
    int rc = csu_input_read(path, &c->input);
    if (rc) { free(c); return rc; }
    const csu_input* in = c->input;
    csu_params* p = &c->prm;

    c->threads = atoi(need(in, "threads"));
    p->sa = opt_i(in, "sa", 0);
    int sacc = opt_i(in, "saCC", 0);
    p->sst = opt_i(in, "sst", 0);
    p->sst_trans = p->sst ? opt_i(in, "sstTrans", 0) : 0;
    p->nvar = 4;
    if (p->sa == 1) p->nvar = 5;
    if (p->sst == 1) p->nvar = 6;
    int axi = atoi(need(in, "axisymmetric"));
    if (g_missing[0]) { csu_case_free(c); return CSU_ERR_INPUT; }
    p->sa_cc = p->sa == 1 && sacc;   /* read only with SA (solver.c:719-730) */

    snprintf(path, sizeof path, "%smesh.su2", wd);
    rc = csu_mesh_read_su2(path, axi, c);
    if (rc) { csu_case_free(c); return rc; }

    /* solverSetData, solver.c:1403-1658 */
    p->order = atoi(need(in, "order"));
    for (int m = 0; m < c->mesh.nmark; m++) {
        char key[96];
        snprintf(key, sizeof key, "BC:%s", c->mark_name[m]);
        const char* v = need(in, key);
        static const char* const bcn[] = {"symmetry", "inlet", "outlet", "wall", "wallT"};
        int t = choice(v, bcn, 5);
        if (g_missing[0]) { csu_case_free(c); return CSU_ERR_INPUT; }
        if (t < 0) {
            snprintf(g_missing, sizeof g_missing, "BC:%s=%s", c->mark_name[m], v);
            csu_case_free(c); return CSU_ERR_INPUT;
        }
        c->mark_type[m] = t;
    }
    p->tp = opt_i(in, "TP", 0);
    p->gamma = 1.4; p->gas_r = 287.0530;
    p->pr = 0.72; p->pr_t = 0.9; p->efix = 0.1;
    p->laminar = opt_i(in, "laminar", 0);
    c->restart = opt_i(in, "restart", 0);
    p->sref = opt_d(in, "Sref", 1.0);
    p->turb_ratio = opt_d(in, "turbRatio", 10.0);
    p->dt_local = opt_i(in, "dtLocal", 0);
    p->dt_local_n = opt_i(in, "dtLocalN", -1);
    p->wimp = opt_d(in, "wImp", 1.0);
    p->turb1order = opt_i(in, "turb1order", 0);
    const char* ts = csu_input_get(in, "timeScheme");
    static const char* const tsn[] = {"RK", "LUSGS", "DPLUR"};
    p->time_scheme = ts ? choice(ts, tsn, 3) : 0;
    if (p->time_scheme < 0) {
        snprintf(g_missing, sizeof g_missing, "timeScheme=%s", ts);
        csu_case_free(c); return CSU_ERR_INPUT;
    }
    p->nlinear = opt_i(in, "Nlinear", 20);
    p->rlim = opt_d(in, "rLim", 1e-6);
    static const char* const fln[] = {"ROE", "AUSM", "AUSMDV", "AUSMpup", "AUSMpup2"};
    const char* fl = need(in, "flux");
    p->flux = choice(fl, fln, 5);
    p->stages = atoi(need(in, "stages"));
    p->cfl = strtod(need(in, "CFL"), NULL);
    if (g_missing[0]) { csu_case_free(c); return CSU_ERR_INPUT; }
    if (p->flux < 0) {
        snprintf(g_missing, sizeof g_missing, "flux=%s", fl);
        csu_case_free(c); return CSU_ERR_INPUT;
    }
    p->pout = opt_d(in, "pout", 1.0e5);
    p->visc_blazek = opt_i(in, "viscBlazek", 0);
    p->twall = opt_d(in, "Twall", 300);
    const char* ws = csu_input_get(in, "writeSurf");
    snprintf(c->write_surf, sizeof c->write_surf, "%s", ws ? ws : "wall");
    p->limiter = opt_i(in, "limiter", 0);
    p->lim_k = opt_d(in, "limK", 1.0);
    if (p->sst) {
        p->sst_kfactor = opt_d(in, "kFactor", 1.125);
        p->sst_ofactor = opt_d(in, "oFactor", 125.0);
        p->sst_owallfactor = opt_d(in, "oWallFactor", 10);
        p->sst_l = opt_d(in, "Lsst", 1.0);
    }
    p->precision = opt_i(in, "precision", CSU_PREC_FP64);
    /* not a key of the reference: "gradient, GG" selects Green-Gauss
       gradients instead of its least squares */
    {
        const char* gs = csu_input_get(in, "gradient");
        if (!gs || strcmp(gs, "LSQ") == 0) p->gradient = 0;
        else if (strcmp(gs, "GG") == 0) p->gradient = 1;
        else {
            snprintf(g_missing, sizeof g_missing, "gradient=%s", gs);
            csu_case_free(c); return CSU_ERR_INPUT;
        }
    }

    /* saCalcD (sa.c:703) is not run here: csu_dev_create computes the wall
       distance on the device when mesh.wall_dist is NULL; csu_wall_distance
       is the host restatement for callers that want it on the host. */

    /* solverInitDomain, solver.c:1660-1725 */
    c->tube = atoi(need(in, "tube"));
    hgas g;
    hgas_init(&g, p->gamma, p->gas_r, p->tp);
    double pin2[7], uin2[6], xm = 0.0;
    if (c->tube == 0) {
        c->nmax = atoi(need(in, "Nmax"));
        double pinf = strtod(need(in, "pressure"), NULL);
        double Tinf = strtod(need(in, "temperature"), NULL);
        double mach = strtod(need(in, "mach"), NULL);
        double nx = strtod(need(in, "nx"), NULL);
        double ny = strtod(need(in, "ny"), NULL);
        if (g_missing[0]) { csu_case_free(c); return CSU_ERR_INPUT; }
        condition_state(&g, p, pinf, Tinf, mach, nx, ny, p->pin, p->uin);
        p->mach_inf = mach;
    } else {
        /* two states split at x = xm (solverInitUTube, solver.c:952-975);
           the first one is the "inlet" of the boundary conditions */
        double p1 = strtod(need(in, "pressure1"), NULL), T1 = strtod(need(in, "temperature1"), NULL);
        double m1 = strtod(need(in, "mach1"), NULL), nx1 = strtod(need(in, "nx1"), NULL);
        double ny1 = strtod(need(in, "ny1"), NULL);
        double p2 = strtod(need(in, "pressure2"), NULL), T2 = strtod(need(in, "temperature2"), NULL);
        double m2 = strtod(need(in, "mach2"), NULL), nx2 = strtod(need(in, "nx2"), NULL);
        double ny2 = strtod(need(in, "ny2"), NULL);
        xm = strtod(need(in, "xm"), NULL);
        c->tmax = strtod(need(in, "tmax"), NULL);
        if (g_missing[0]) { csu_case_free(c); return CSU_ERR_INPUT; }
        /* solverInitUTube sets 4 variables only: no turbulence model here;
           the tube loop calls solverStepRK whatever timeScheme says, and
           with dtLocal 1 it would advance t by a step it never computes */
        if (p->sa || p->sst || p->dt_local == 1) { csu_case_free(c); return CSU_ERR_UNSUPPORTED; }
        p->time_scheme = CSU_TIME_RK;
        condition_state(&g, p, p1, T1, m1, nx1, ny1, p->pin, p->uin);
        condition_state(&g, p, p2, T2, m2, nx2, ny2, pin2, uin2);
        p->mach_inf = m1;
        c->nmax = 0;
    }

    int nc = c->mesh.ncell, nv = p->nvar;
    c->U = malloc(sizeof(double) * (size_t)nv * (size_t)nc);
    if (!c->U) { csu_case_free(c); return CSU_ERR_NOMEM; }
    if (c->restart) {
        snprintf(path, sizeof path, "%srestart.csv", wd);
        rc = csu_read_restart(c, c->U, path);
        if (rc) { csu_case_free(c); return rc; }
    } else if (c->tube == 0) {
        /* solverInitU (4 variables) + saInitU / sstInitU */
        for (int kk = 0; kk < nv; kk++)
            for (int i = 0; i < nc; i++) c->U[(size_t)kk * nc + i] = p->uin[kk];
    } else {
        for (int i = 0; i < nc; i++) {
            const int32_t* cv = c->mesh.cell_vert + 4 * i;
            int np = cv[3] >= 0 ? 4 : 3;
            double x = 0.0;
            for (int k = 0; k < np; k++) x += c->mesh.node_xy[2 * cv[k]];
            x /= np;
            for (int kk = 0; kk < 4; kk++) c->U[(size_t)kk * nc + i] = x < xm ? p->uin[kk] : uin2[kk];
        }
    }
    *out = c;
    return CSU_OK;
}

void csu_case_free(csu_case* c)
{
    if (!c) return;
    free(c->U); free(c->node_xy); free(c->cell_vert); free(c->face); free(c->bface);
    free(c->bface_mark); free(c->mark_type); free(c->cell_nbr); free(c->cell_face);
    free(c->wall_dist); free(c->mark_name);
    csu_input_free(c->input);
    free(c);
}
