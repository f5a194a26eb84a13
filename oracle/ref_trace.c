/*
 * ref_trace.c -- trace driver for the reference solver (test infrastructure).
 *
 * Linked against the reference's own objects (every reference source but
 * main.c, built by oracle/Makefile into oracle/_ref/).  It calls the
 * reference's functions unchanged and writes full-precision binary traces that
 * the reference's own outputs cannot give (its convergence.csv prints
 * residuals at %+.4e, solver.c:1902):
 *
 *   csu_ref_trace <case/> <out/> mesh
 *       con.i32    [Ncon][4]  e0 e1 p0 p1 (mesh.c:1172-1218 hash order)
 *       cellv.i32  [Nelem][4] vertex ids after the CCW fix (-1 pad)
 *       nbr.i32    [Nelem][4] neighbour list in neiL order; a boundary
 *                             neighbour is -(1+g), g = global boundary index
 *                             (markers in file order, elements in order)
 *       cface.i32  [Nelem][4] signed 1-based face ids (mesh.c:1011-1037)
 *       bface.i32  [Nb][3]    e0 p0 p1 after the border orientation fix
 *       bmark.i32  [Nb]       marker index of each boundary face
 *       dist.f64   [Nelem]    wall distance (SA/SST cases only)
 *   csu_ref_trace <case/> <out/> resid [U.f64]
 *       one solverCalcR (solver.c:685) on the initial state or on U.f64
 *       (SoA [Nvar][Nelem] doubles); writes R, dPx, dPy, phi ([Nvar][Nelem]),
 *       P ([Nelem][Nvar+1]), Pb ([Nb][Nvar+1]), U (after the clamps)
 *   csu_ref_trace <case/> <out/> hist N [U.f64]
 *       N iterations of the solverSolve loop body (solver.c:1864-1890);
 *       writes res.f64 [N][Nvar], coef.f64 [N][2] (solverCalcCoeff Cx, Cy),
 *       U.f64 (final state), dtl.f64 (last local time step, when present)
 *   csu_ref_trace <case/> <out/> tube [N]
 *       the tube branch of solverSolve (solver.c:1912-1936) to tmax (or N
 *       steps at most) with the reference's solverCalcDt and solverStepRK;
 *       writes U.f64 and t.f64 = [t, steps]
 *
 * Everything the reference prints goes to stdout; callers discard it.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdbool.h>
#include "utils.h"
#include "input.h"
#include "mesh.h"
#include "solver.h"
#include "flux.h"
#include "boundary.h"
#include "implicit.h"

static char outdir[512];

static void write_bin(const char* name, const void* data, size_t bytes)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s%s", outdir, name);
    FILE* f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "ref_trace: cannot write %s\n", path); exit(2); }
    if (bytes && fwrite(data, 1, bytes, f) != bytes) { fprintf(stderr, "ref_trace: short write\n"); exit(2); }
    fclose(f);
}

static void write_soa(const char* name, double** A, int nrow, int ncol)
{
    double* buf = malloc(sizeof(double) * (size_t)nrow * (size_t)ncol);
    for (int k = 0; k < nrow; k++)
        memcpy(buf + (size_t)k * ncol, A[k], sizeof(double) * (size_t)ncol);
    write_bin(name, buf, sizeof(double) * (size_t)nrow * (size_t)ncol);
    free(buf);
}

static void load_U(SOLVER* s, const char* path)
{
    int nc = s->mesh->Nelem;
    FILE* f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "ref_trace: cannot read %s\n", path); exit(2); }
    for (int k = 0; k < s->Nvar; k++)
        if (fread(s->U[k], sizeof(double), (size_t)nc, f) != (size_t)nc) {
            fprintf(stderr, "ref_trace: short state file\n"); exit(2);
        }
    fclose(f);
}

/* pointer -> global boundary index, by linear search over markers (setup only) */
static int bc_global_index(MESH* m, ELEMENT* e)
{
    int off = 0;
    for (int b = 0; b < m->Nmark; b++) {
        MESHBC* bc = m->bc[b];
        for (int j = 0; j < bc->Nelem; j++)
            if (bc->elemL[j] == e) return off + j;
        off += bc->Nelem;
    }
    return -1;
}

static int n_bfaces(MESH* m)
{
    int nb = 0;
    for (int b = 0; b < m->Nmark; b++) nb += m->bc[b]->Nelem;
    return nb;
}

static void dump_mesh(SOLVER* s)
{
    MESH* m = s->mesh;
    int nc = m->Nelem, nb = n_bfaces(m);
    int* con = malloc(sizeof(int) * 4 * (size_t)m->Ncon);
    for (int i = 0; i < m->Ncon; i++)
        for (int k = 0; k < 4; k++) con[4 * i + k] = m->con[i][k];
    write_bin("con.i32", con, sizeof(int) * 4 * (size_t)m->Ncon);
    free(con);

    int* cellv = malloc(sizeof(int) * 4 * (size_t)nc);
    int* nbr = malloc(sizeof(int) * 4 * (size_t)nc);
    int* cface = malloc(sizeof(int) * 4 * (size_t)nc);
    /* interior neighbours are elemL entries; boundary ones have Np == 2 */
    for (int i = 0; i < nc; i++) {
        ELEMENT* E = m->elemL[i];
        for (int k = 0; k < 4; k++) {
            cellv[4 * i + k] = k < E->Np ? E->p[k] : -1;
            cface[4 * i + k] = k < E->Np ? E->f[k] : 0;
            nbr[4 * i + k] = 0x7fffffff;
            if (k < E->Np) {
                ELEMENT* N = E->neiL[k];
                if (N->Np == 2) nbr[4 * i + k] = -(1 + bc_global_index(m, N));
                else nbr[4 * i + k] = N->ii;
            }
        }
    }
    write_bin("cellv.i32", cellv, sizeof(int) * 4 * (size_t)nc);
    write_bin("nbr.i32", nbr, sizeof(int) * 4 * (size_t)nc);
    write_bin("cface.i32", cface, sizeof(int) * 4 * (size_t)nc);
    free(cellv); free(nbr); free(cface);

    int* bface = malloc(sizeof(int) * 3 * (size_t)(nb > 0 ? nb : 1));
    int* bmark = malloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
    int g = 0;
    for (int b = 0; b < m->Nmark; b++)
        for (int j = 0; j < m->bc[b]->Nelem; j++, g++) {
            ELEMENT* B = m->bc[b]->elemL[j];
            bface[3 * g + 0] = B->neiL[0]->ii;
            bface[3 * g + 1] = B->p[0];
            bface[3 * g + 2] = B->p[1];
            bmark[g] = b;
        }
    write_bin("bface.i32", bface, sizeof(int) * 3 * (size_t)nb);
    write_bin("bmark.i32", bmark, sizeof(int) * (size_t)nb);
    free(bface); free(bmark);

    if (s->sa || s->sstFlag) {
        double* d = malloc(sizeof(double) * (size_t)nc);
        for (int i = 0; i < nc; i++) d[i] = m->elemL[i]->d;
        write_bin("dist.f64", d, sizeof(double) * (size_t)nc);
        free(d);
    }
}

static void dump_P(SOLVER* s)
{
    MESH* m = s->mesh;
    int nc = m->Nelem, nv1 = s->Nvar + 1, nb = n_bfaces(m);
    double* P = malloc(sizeof(double) * (size_t)nc * nv1);
    for (int i = 0; i < nc; i++)
        for (int k = 0; k < nv1; k++) P[(size_t)i * nv1 + k] = m->elemL[i]->P[k];
    write_bin("P.f64", P, sizeof(double) * (size_t)nc * nv1);
    free(P);
    double* Pb = malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1) * nv1);
    int g = 0;
    for (int b = 0; b < m->Nmark; b++)
        for (int j = 0; j < m->bc[b]->Nelem; j++, g++)
            for (int k = 0; k < nv1; k++) Pb[(size_t)g * nv1 + k] = m->bc[b]->elemL[j]->P[k];
    write_bin("Pb.f64", Pb, sizeof(double) * (size_t)nb * nv1);
    free(Pb);
}

/* one pass of the solverSolve loop body, solver.c:1866-1889 */
static void iterate_once(SOLVER* s, int ii)
{
    if (s->timeScheme == 0) {
        solverCalcDt(s);
        solverStepRK(s);
    } else if (s->timeScheme == 1) {
        solverCalcR(s, s->U);
        implicitCalcD(s);
        implicitLUSGS_L(s);
        implicitLUSGS_U(s);
        solverUpdateUImplicit(s);
    } else if (s->timeScheme == 2) {
        solverCalcR(s, s->U);
        implicitCalcDPLUR(s);
        solverUpdateUImplicit(s);
    }
    if (ii == s->dtLocalN) s->dtLocal = 0;
}

int main(int argc, char** argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: %s <case/> <out/> mesh|resid [U.f64]|hist N [U.f64]|tube [N]\n", argv[0]);
        return 2;
    }
    snprintf(outdir, sizeof(outdir), "%s", argv[2]);
    const char* mode = argv[3];

    SOLVER* s = solverInit(argv[1]);
    int nc = s->mesh->Nelem;

    if (strcmp(mode, "mesh") == 0) {
        dump_mesh(s);
    } else if (strcmp(mode, "resid") == 0) {
        if (argc > 4) load_U(s, argv[4]);
        solverCalcR(s, s->U);
        write_soa("R.f64", s->R, s->Nvar, nc);
        write_soa("dPx.f64", s->dPx, s->Nvar, nc);
        write_soa("dPy.f64", s->dPy, s->Nvar, nc);
        write_soa("phi.f64", s->phi, s->Nvar, nc);
        write_soa("U.f64", s->U, s->Nvar, nc);
        dump_P(s);
    } else if (strcmp(mode, "hist") == 0) {
        if (argc < 5) { fprintf(stderr, "hist needs N\n"); return 2; }
        int N = atoi(argv[4]);
        if (argc > 5) load_U(s, argv[5]);
        double* res = malloc(sizeof(double) * (size_t)(N > 0 ? N : 1) * s->Nvar);
        double* coef = malloc(sizeof(double) * (size_t)(N > 0 ? N : 1) * 2);
        for (int ii = 0; ii < N; ii++) {
            iterate_once(s, ii);
            solverCalcRes(s);
            double Cx, Cy;
            solverCalcCoeff(s, &Cx, &Cy);
            for (int k = 0; k < s->Nvar; k++) res[(size_t)ii * s->Nvar + k] = s->res[k];
            coef[2 * (size_t)ii + 0] = Cx;
            coef[2 * (size_t)ii + 1] = Cy;
        }
        write_bin("res.f64", res, sizeof(double) * (size_t)N * s->Nvar);
        write_bin("coef.f64", coef, sizeof(double) * (size_t)N * 2);
        write_soa("U.f64", s->U, s->Nvar, nc);
        write_soa("R.f64", s->R, s->Nvar, nc);
        if (s->dtLocal == 1 || s->timeScheme != 0)
            write_bin("dtl.f64", s->dtL, sizeof(double) * (size_t)nc);
        free(res); free(coef);
    } else if (strcmp(mode, "tube") == 0) {
        double tmax = strtod(inputGetValue(s->input, "tmax"), NULL);
        double t = 0.0;
        int stop = 0, n = 0, cap = argc > 4 ? atoi(argv[4]) : 0;
        while (!stop && !(cap > 0 && n >= cap)) {
            solverCalcDt(s);
            if (t + s->dt > tmax) {
                s->dt = tmax - t;
                stop = 1;
            }
            solverStepRK(s);
            t += s->dt * s->stages / 2.0;
            n++;
        }
        double tn[2] = {t, (double)n};
        write_bin("t.f64", tn, sizeof tn);
        write_soa("U.f64", s->U, s->Nvar, nc);
    } else {
        fprintf(stderr, "unknown mode %s\n", mode);
        return 2;
    }
    fflush(stdout);
    /* the reference's solverFree has latent double frees (dtL with dtLocal and
       an implicit scheme); the process exits without it */
    return 0;
}
