/*
 * output.c -- restart.csv and solution.csv, in the reference's formats.
 *
 *   restart.csv  "1,\n" "0, Nelem, Nvar,\n" then one row per cell of
 *                "%.10e, " values (reference solver.c:327-355); read back with
 *                the comma/newline tokenizer of readTables.c:5-142.
 *   solution.csv "0, Np, Nvar,\n" then per node the inverse-distance average
 *                of the surrounding cells' U (reference solver.c:222-325),
 *                velocity zeroed on wall nodes for viscous cases.
 *   solution2.csv a header line, then per node the same average of the
 *                primitive state and Mach, total enthalpy and entropy
 *                (reference solver.c:2009-2220).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "csu_host.h"
#include "hgas.h"

int csu_write_restart(const csu_case* c, const double* U, const char* path)
{
    FILE* f = fopen(path, "w");
    if (!f) return CSU_ERR_IO;
    int nc = c->mesh.ncell, nv = c->prm.nvar;
    fprintf(f, "1,\n");
    fprintf(f, "0, %i, %i,\n", nc, nv);
    for (int i = 0; i < nc; i++) {
        for (int k = 0; k < nv; k++) fprintf(f, "%.10e, ", U[(size_t)k * nc + i]);
        fprintf(f, "\n");
    }
    fclose(f);
    return CSU_OK;
}

/* readTables.c getWord: a token ends at ',' or '\n' */
static int rt_word(FILE* f, char* s, int cap)
{
    int n = 0, ch;
    for (;;) {
        ch = fgetc(f);
        if (ch == EOF || ch == ',' || ch == '\n') break;
        if (n < cap - 1) s[n++] = (char)ch;
    }
    s[n] = '\0';
    return ch;
}

int csu_read_restart(const csu_case* c, double* U, const char* path)
{
    FILE* f = fopen(path, "r");
    if (!f) return CSU_ERR_IO;
    char s[128];
    int nc = c->mesh.ncell, nv = c->prm.nvar;
    rt_word(f, s, sizeof s);                 /* number of tables */
    if (atoi(s) < 1) { fclose(f); return CSU_ERR_IO; }
    rt_word(f, s, sizeof s);
    rt_word(f, s, sizeof s);                 /* table number */
    rt_word(f, s, sizeof s); int nrow = atoi(s);
    rt_word(f, s, sizeof s); int ncol = atoi(s);
    rt_word(f, s, sizeof s);
    if (nrow < nc || ncol < nv) { fclose(f); return CSU_ERR_IO; }
    for (int i = 0; i < nc; i++) {
        for (int k = 0; k < ncol; k++) {
            if (rt_word(f, s, sizeof s) == EOF && k < ncol - 1) { fclose(f); return CSU_ERR_IO; }
            if (k < nv) U[(size_t)k * nc + i] = strtod(s, NULL);
        }
        rt_word(f, s, sizeof s);
    }
    fclose(f);
    return CSU_OK;
}

int csu_write_solution(const csu_case* c, const double* U, const char* path)
{
    const csu_mesh_desc* m = &c->mesh;
    int nv = c->prm.nvar, np = m->nnode, nc = m->ncell;
    double* Up = calloc((size_t)nv * (size_t)np, sizeof(double));
    double* den = calloc((size_t)np, sizeof(double));
    if (!Up || !den) { free(Up); free(den); return CSU_ERR_NOMEM; }
    for (int i = 0; i < nc; i++) {
        const int32_t* v = m->cell_vert + 4 * i;
        int nvert = v[3] >= 0 ? 4 : 3;
        double xc = 0.0, yc = 0.0;
        for (int k = 0; k < nvert; k++) { xc += m->node_xy[2 * v[k]]; yc += m->node_xy[2 * v[k] + 1]; }
        xc /= nvert; yc /= nvert;
        for (int j = 0; j < nvert; j++) {
            int p = v[j];
            double xp = m->node_xy[2 * p], yp = m->node_xy[2 * p + 1];
            double L = sqrt((xp - xc) * (xp - xc) + (yp - yc) * (yp - yc));
            for (int k = 0; k < nv; k++) Up[(size_t)k * np + p] += U[(size_t)k * nc + i] / L;
            den[p] += 1 / L;
        }
    }
    for (int p = 0; p < np; p++)
        if (den[p] != 0)
            for (int k = 0; k < nv; k++) Up[(size_t)k * np + p] /= den[p];
    if (c->prm.laminar || c->prm.sa || c->prm.sst) {
        for (int g = 0; g < m->nbface; g++) {
            int t = m->mark_type[m->bface_mark[g]];
            if (t != CSU_BC_WALL && t != CSU_BC_WALLT) continue;
            int p0 = m->bface[3 * g + 1], p1 = m->bface[3 * g + 2];
            Up[(size_t)1 * np + p0] = 0.0; Up[(size_t)2 * np + p0] = 0.0;
            Up[(size_t)1 * np + p1] = 0.0; Up[(size_t)2 * np + p1] = 0.0;
        }
    }
    FILE* f = fopen(path, "w");
    if (!f) { free(Up); free(den); return CSU_ERR_IO; }
    fprintf(f, "0, %i, %i,\n", np, nv);
    for (int p = 0; p < np; p++) {
        for (int k = 0; k < nv; k++) fprintf(f, "%.10e, ", Up[(size_t)k * np + p]);
        fprintf(f, "\n");
    }
    fclose(f);
    free(Up); free(den);
    return CSU_OK;
}

/* solverWriteSolution2 (solver.c:2009-2220): P [ncell][nvar+1] = r u v p T
   [nu~ | k om] of the clamped state (solverCalcPrimitive), averaged to the
   nodes with inverse-distance weights; on wall (not wallT) nodes of viscous
   runs u = v = 0, and for SA nu~ = 0 (SST: k = 0) and T = Twall when
   Twall >= 0; then Mach, total enthalpy and the entropy relative to the
   inlet's.  SST runs add the node averages of the per-cell fields sst
   [ncell][5] (sstInterMiT: eddy viscosity, F1, F2, wall distance, omega;
   csu_dev_sst_fields); sst is ignored for the other models. */
int csu_write_solution2(const csu_case* c, const double* P, const double* sst, const char* path)
{
    const csu_mesh_desc* m = &c->mesh;
    const csu_params* prm = &c->prm;
    if (prm->sst && !sst) return CSU_ERR_ARG;
    const int nv1 = prm->nvar + 1, np = m->nnode, nc = m->ncell;
    const int nx = prm->sst ? 5 : 0;   /* SST: eddy viscosity, F1, F2, d, omega */
    double* Pn = calloc((size_t)(nv1 + nx) * (size_t)np, sizeof(double));
    double* den = calloc((size_t)np, sizeof(double));
    if (!Pn || !den) { free(Pn); free(den); return CSU_ERR_NOMEM; }
    double* Xn = Pn + (size_t)nv1 * np;
    /* the SST fields take the same node weights in a pass of their own
       (solver.c:2133-2172); summed in the same cell order, so one pass
       gives the same sums */
    for (int i = 0; i < nc; i++) {
        const int32_t* v = m->cell_vert + 4 * i;
        int nvert = v[3] >= 0 ? 4 : 3;
        double xc = 0.0, yc = 0.0;
        for (int k = 0; k < nvert; k++) { xc += m->node_xy[2 * v[k]]; yc += m->node_xy[2 * v[k] + 1]; }
        xc /= nvert; yc /= nvert;
        for (int j = 0; j < nvert; j++) {
            int p = v[j];
            double xp = m->node_xy[2 * p], yp = m->node_xy[2 * p + 1];
            double L = sqrt((xp - xc) * (xp - xc) + (yp - yc) * (yp - yc));
            for (int k = 0; k < nv1; k++) Pn[(size_t)k * np + p] += P[(size_t)i * nv1 + k] / L;
            for (int k = 0; k < nx; k++) Xn[(size_t)k * np + p] += sst[(size_t)i * nx + k] / L;
            den[p] += 1 / L;
        }
    }
    for (int p = 0; p < np; p++)
        if (den[p] != 0)
            for (int k = 0; k < nv1 + nx; k++) Pn[(size_t)k * np + p] /= den[p];
    if (prm->laminar || prm->sa || prm->sst) {
        for (int g = 0; g < m->nbface; g++) {
            if (m->mark_type[m->bface_mark[g]] != CSU_BC_WALL) continue;
            int q[2] = {m->bface[3 * g + 1], m->bface[3 * g + 2]};
            for (int e = 0; e < 2; e++) {
                Pn[(size_t)1 * np + q[e]] = 0.0;
                Pn[(size_t)2 * np + q[e]] = 0.0;
                if (prm->sa || prm->sst) {   /* nu~ or k; omega keeps its average */
                    Pn[(size_t)5 * np + q[e]] = 0.0;
                    if (prm->twall >= 0) Pn[(size_t)4 * np + q[e]] = prm->twall;
                }
            }
        }
    }
    hgas gas;
    hgas_init(&gas, prm->gamma, prm->gas_r, prm->tp);
    /* the inlet condition's own p and T (conditionInit, solver.c:24-38) */
    const char* ks_p = c->tube ? "pressure1" : "pressure";
    const char* ks_T = c->tube ? "temperature1" : "temperature";
    const char *sp = csu_input_get(c->input, ks_p), *sT = csu_input_get(c->input, ks_T);
    double s0 = hgas_Tp2entropy(&gas, sT ? strtod(sT, NULL) : prm->pin[4], sp ? strtod(sp, NULL) : prm->pin[3]);
    FILE* f = fopen(path, "w");
    if (!f) { free(Pn); free(den); return CSU_ERR_IO; }
    fprintf(f, prm->sst ? "r,u,v,p,T,k,o,mach,H,s,miEddy,F1,F2,d,om2,\n"
               : prm->sa ? "r,u,v,p,T,n,mach,H,s,\n" : "r,u,v,p,T,mach,H,s,\n");
    for (int p = 0; p < np; p++) {
        const double* q = Pn + p;
        double V2 = q[np] * q[np] + q[2 * (size_t)np] * q[2 * (size_t)np];
        double T = q[4 * (size_t)np], pr = q[3 * (size_t)np];
        double Q0 = sqrt(V2) / hgas_T2c(&gas, T);
        double Q1 = hgas_T2e(&gas, T) + pr / q[0] + 0.5 * V2;
        double Q2 = hgas_Tp2entropy(&gas, T, pr) - s0;
        for (int k = 0; k < nv1; k++) fprintf(f, "%.10e, ", q[(size_t)k * np]);
        fprintf(f, "%.10e, %.10e, %.10e, ", Q0, Q1, Q2);
        for (int k = 0; k < nx; k++) fprintf(f, "%.10e, ", Xn[(size_t)k * np + p]);
        fputc('\n', f);
    }
    fclose(f);
    free(Pn); free(den);
    return CSU_OK;
}

/* surfData.csv for SA and SST runs: saSolverWriteSurf (sa.c:854-960),
   saCC_SolverWriteSurf (saCC.c:599-705), sstSolverWriteSurf (sst.c:820-962).
   For each marker named by the writeSurf key (default "wall"): per face,
   from faces [nbface][11] (csu_dev_surf_faces): the ghost r u v p T and
   nu~ (SST: k), Cp, Cfx, Cfy, the wall heat flux, y+, Mach and the eddy
   viscosity; then averaged to the marker's nodes.  The nodes are numbered
   as the reference numbers them, p0 of each face in file order and p1 of
   the last face, so a closed marker's first node is left without faces
   (0/0, as in the reference).  A face node outside that list is skipped:
   the reference reads an unset index there. */
int csu_write_surf(const csu_case* c, const double* faces, const char* path)
{
    const csu_mesh_desc* m = &c->mesh;
    const csu_params* prm = &c->prm;
    if (!(prm->sa || prm->sst)) return CSU_ERR_UNSUPPORTED;
    if (!faces && m->nbface) return CSU_ERR_ARG;
    const char* name = csu_input_get(c->input, "writeSurf");
    if (!name) name = "wall";
    hgas gas;
    hgas_init(&gas, prm->gamma, prm->gas_r, prm->tp);
    const double rin = prm->pin[0], uin = prm->pin[1], vin = prm->pin[2], pin = prm->pin[3];
    const double qdin = 0.5 * rin * (uin * uin + vin * vin);
    const int nb = m->nbface, np = m->nnode, NV = 13;
    int* list = malloc(sizeof(int) * (size_t)(nb + 1));
    int* pn = malloc(sizeof(int) * (size_t)(nb + 1));
    int* cnt = malloc(sizeof(int) * (size_t)(nb + 1));
    int* pn2 = malloc(sizeof(int) * (size_t)(np ? np : 1));
    double* D = malloc(sizeof(double) * (size_t)NV * (nb ? nb : 1));
    double* Dp = malloc(sizeof(double) * (size_t)NV * (nb + 1));
    FILE* f = fopen(path, "w");
    if (!list || !pn || !cnt || !pn2 || !D || !Dp || !f) {
        free(list); free(pn); free(cnt); free(pn2); free(D); free(Dp);
        if (f) fclose(f);
        return f ? CSU_ERR_NOMEM : CSU_ERR_IO;
    }
    for (int mk = 0; mk < m->nmark; mk++) {
        if (strcmp(c->mark_name[mk], name) != 0) continue;
        fprintf(f, "x,y,rho,u,v,p,T,n,Cp,Cfx,Cfy,q,yplus,mach,miEddy,\n");
        int ne = 0;
        for (int g = 0; g < nb; g++)
            if (m->bface_mark[g] == mk) list[ne++] = g;
        if (ne == 0) continue;
        for (int p = 0; p < np; p++) pn2[p] = -1;
        for (int ii = 0; ii < ne; ii++) pn[ii] = m->bface[3 * list[ii] + 1];
        pn[ne] = m->bface[3 * list[ne - 1] + 2];
        for (int ii = 0; ii <= ne; ii++) pn2[pn[ii]] = ii;
        for (int ii = 0; ii <= ne; ii++) {
            cnt[ii] = 0;
            for (int k = 0; k < NV; k++) Dp[(size_t)k * (ne + 1) + ii] = 0.0;
        }
        for (int ii = 0; ii < ne; ii++) {
            const int g = list[ii];
            const double* q = faces + 11 * (size_t)g;
            const int p0 = m->bface[3 * g + 1], p1 = m->bface[3 * g + 2];
            const double x0 = m->node_xy[2 * p0], y0 = m->node_xy[2 * p0 + 1];
            const double x1 = m->node_xy[2 * p1], y1 = m->node_xy[2 * p1 + 1];
            double dSx = y1 - y0, dSy = -(x1 - x0);
            if (m->axi == 1) { double y = (y0 + y1) / 2; dSx = dSx * y; dSy = dSy * y; }   /* meshCalcDS */
            const double dS = sqrt(dSx * dSx + dSy * dSy);
            const double fx = q[0], fy = q[1], fe = q[2];
            const double r = q[4], u = q[5], v = q[6], p = q[7], T = q[8];
            const double tau = sqrt(fx * fx + fy * fy) / dS;
            const double uplus = sqrt(tau / r);
            const double mi_L = 1.458e-6 * T * sqrt(T) / (T + 110.4);   /* sutherland, solver.c:1222 */
            double* Dc = D + ii;
            Dc[0 * (size_t)ne] = r;
            Dc[1 * (size_t)ne] = u;
            Dc[2 * (size_t)ne] = v;
            Dc[3 * (size_t)ne] = p;
            Dc[4 * (size_t)ne] = T;
            Dc[5 * (size_t)ne] = q[9];
            Dc[6 * (size_t)ne] = (p - pin) / qdin;
            Dc[7 * (size_t)ne] = fx / (dS * qdin);
            Dc[8 * (size_t)ne] = fy / (dS * qdin);
            Dc[9 * (size_t)ne] = (fe - u * fx - v * fy) / dS;
            Dc[10 * (size_t)ne] = fabs(r * uplus * q[10]) / mi_L;
            Dc[11 * (size_t)ne] = sqrt(u * u + v * v) / hgas_T2c(&gas, T);
            Dc[12 * (size_t)ne] = q[3];
        }
        for (int ii = 0; ii < ne; ii++) {
            const int g = list[ii];
            const int a = pn2[m->bface[3 * g + 1]], b = pn2[m->bface[3 * g + 2]];
            for (int k = 0; k < NV; k++) {
                if (a >= 0) Dp[(size_t)k * (ne + 1) + a] += D[(size_t)k * ne + ii];
                if (b >= 0) Dp[(size_t)k * (ne + 1) + b] += D[(size_t)k * ne + ii];
            }
            if (a >= 0) cnt[a] += 1;
            if (b >= 0) cnt[b] += 1;
        }
        for (int ii = 0; ii <= ne; ii++) {
            fprintf(f, "% .10e,", m->node_xy[2 * pn[ii]]);
            fprintf(f, "% .10e,", m->node_xy[2 * pn[ii] + 1]);
            for (int k = 0; k < NV; k++) fprintf(f, "% .10e,", Dp[(size_t)k * (ne + 1) + ii] / cnt[ii]);
            fprintf(f, "\n");
        }
    }
    fclose(f);
    free(list); free(pn); free(cnt); free(pn2); free(D); free(Dp);
    return CSU_OK;
}
