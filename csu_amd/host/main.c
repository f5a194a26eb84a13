/*
 * main.c -- the drop-in `executable`: ./executable <case>/
 *
 * Same process contract as the reference main.c:18-45 + solverSolve
 * (solver.c:1836-1911): reads <case>/input.dat, mesh.su2 (and restart.csv
 * when restart=1), runs Nmax iterations, prints one line per iteration and
 * appends it to convergence.csv, then writes solution.csv, solution2.csv,
 * restart.csv, friction.csv and, for SA and SST runs, surfData.csv.  With
 * tube=1 it runs the time-accurate branch to tmax instead
 * (solver.c:1912-1941) and writes solution.csv, solution2.csv and
 * restart.csv.  The iterations run on the GPU through csu.h; the host only
 * formats the per-iteration records, which come back in blocks.
 *
 * As the reference, input errors print a message and exit with status 0
 * (input.c:156-160); device errors print a message and exit with status 1.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include "csu.h"
#include "csu_host.h"

static double seconds(struct timeval a, struct timeval b)
{
    return ((double)b.tv_sec + b.tv_usec / 1000000.) - ((double)a.tv_sec + a.tv_usec / 1000000.);
}

int main(int argc, char** argv)
{
    struct timeval start, stop;
    gettimeofday(&start, NULL);
    if (argc < 2) {
        printf("usage: %s <case directory/>\n", argv[0]);
        return 0;
    }
    const char* wd = argv[1];
    csu_case* c = NULL;
    int rc = csu_case_load(wd, &c);
    if (rc == CSU_ERR_INPUT) {
        printf("Error: Value not found in the input file: %s.\n", csu_host_missing_key());
        return 0;
    }
    if (rc) {
        printf("Error: cannot load case %s (code %d).\n", wd, rc);
        return 0;
    }
    printf("mesh: %d cells, %d interior faces, %d boundary faces\n",
           c->mesh.ncell, c->mesh.nface, c->mesh.nbface);

    csu_dev* d = NULL;
    rc = csu_dev_create(&c->mesh, &c->prm, &d);
    if (!rc) rc = csu_dev_set_state(d, c->U);
    if (rc) {
        printf("Error: device setup failed: %s\n", csu_last_error());
        csu_case_free(c);
        return 1;
    }

    char path[1200];
    if (c->tube) {
        /* tube branch of solverSolve (solver.c:1912-1941): no convergence
           file, no wall shear; the run ends at tmax */
        printf("\nmain: running solution:\n");
        double t = 0.0;
        int32_t n = 0, fin = 0;
        rc = csu_dev_advance_time(d, c->tmax, 0, &t, &n, &fin);
        if (rc) { printf("Error: time-accurate run failed: %s\n", csu_last_error()); return 1; }
        printf("%i steps\n", n);
        printf("time %f s\n", t);
        goto write;
    }
    snprintf(path, sizeof path, "%sconvergence.csv", wd);
    FILE* conv = fopen(path, c->restart ? "a" : "w");
    if (!conv) { printf("Error: cannot open %s\n", path); return 1; }

    const int nv = c->prm.nvar, block = 200;
    double* res = malloc(sizeof(double) * (size_t)block * nv);
    double* coef = malloc(sizeof(double) * (size_t)block * 4);
    printf("\nmain: running solution:\n");
    for (int ii = 0; ii < c->nmax; ii += block) {
        int n = c->nmax - ii < block ? c->nmax - ii : block;
        rc = csu_dev_iterate(d, n, res, coef);
        if (rc) { printf("Error: iteration failed: %s\n", csu_last_error()); return 1; }
        for (int k = 0; k < n; k++) {
            const double* r = res + (size_t)k * nv;
            const double* q = coef + (size_t)k * 4;
            printf("%i, ", ii + k);
            for (int v = 0; v < nv; v++) printf(" %+.4e,", r[v]);
            printf("\n");
            printf("%f, %f\n", q[0] + q[1], q[2] + q[3]);
            fprintf(conv, "%i,", ii + k);
            for (int v = 0; v < nv; v++) fprintf(conv, " %+.4e,", r[v]);
            fprintf(conv, " %e, %e, %e, %e,", q[0], q[1], q[2], q[3]);
            fprintf(conv, "\n");
        }
    }
    fclose(conv);
    free(res); free(coef);

    /* friction.csv (solverCalcCoeff2, solver.c:1303-1340): per wall / wallT
       face in marker order, x, txx, txy, tyy, yp; empty for inviscid runs */
    snprintf(path, sizeof path, "%sfriction.csv", wd);
    FILE* fr = fopen(path, "w");
    if (fr) {
        const csu_mesh_desc* m = &c->mesh;
        if (c->prm.laminar || c->prm.sa || c->prm.sst) {
            double* wt = malloc(sizeof(double) * 4 * (size_t)(m->nbface ? m->nbface : 1));
            if (wt && csu_dev_wall_tensor(d, wt) == CSU_OK) {
                for (int g = 0; g < m->nbface; g++) {
                    int t = m->mark_type[m->bface_mark[g]];
                    if (t != CSU_BC_WALL && t != CSU_BC_WALLT) continue;
                    const int32_t* q = m->bface + 3 * g;
                    double x = (m->node_xy[2 * q[1]] + m->node_xy[2 * q[2]]) / 2;
                    const double* w = wt + 4 * (size_t)g;
                    fprintf(fr, "%e, %e, %e, %e, %e,\n", x, w[0], w[1], w[2], w[3]);
                }
            }
            free(wt);
        }
        fclose(fr);
    }

write:
    rc = csu_dev_get_state(d, c->U);
    if (rc) { printf("Error: %s\n", csu_last_error()); return 1; }
    printf("main: saving the solution.\n");
    snprintf(path, sizeof path, "%ssolution.csv", wd);
    csu_write_solution(c, c->U, path);
    /* solverWriteSolution2 clamps U in place (solverCalcPrimitive) and
       writes the node averages of P; restart.csv then holds the clamped U */
    double* Pc = malloc(sizeof(double) * (size_t)(c->prm.nvar + 1) * (size_t)c->mesh.ncell);
    if (!Pc) { printf("Error: out of memory\n"); return 1; }
    rc = csu_dev_primitives(d, Pc);
    if (!rc) rc = csu_dev_get_state(d, c->U);
    if (rc) { printf("Error: %s\n", csu_last_error()); return 1; }
    printf("main: saving the solution 2.\n");
    snprintf(path, sizeof path, "%ssolution2.csv", wd);
    double* X = NULL;
    if (c->prm.sst) {
        X = malloc(sizeof(double) * 5 * (size_t)c->mesh.ncell);
        if (!X) { printf("Error: out of memory\n"); return 1; }
        rc = csu_dev_sst_fields(d, X);
        if (rc) { printf("Error: %s\n", csu_last_error()); return 1; }
    }
    csu_write_solution2(c, Pc, X, path);
    free(X);
    free(Pc);
    printf("main: saving the restart file.\n");
    snprintf(path, sizeof path, "%srestart.csv", wd);
    csu_write_restart(c, c->U, path);
    if (c->prm.sa || c->prm.sst) {   /* solverWriteSurf (solver.c:1984-2007) */
        double* F = malloc(sizeof(double) * 11 * (size_t)(c->mesh.nbface ? c->mesh.nbface : 1));
        if (!F) { printf("Error: out of memory\n"); return 1; }
        rc = csu_dev_surf_faces(d, F);
        if (rc) { printf("Error: %s\n", csu_last_error()); return 1; }
        snprintf(path, sizeof path, "%ssurfData.csv", wd);
        csu_write_surf(c, F, path);
        free(F);
    }

    csu_dev_destroy(d);
    csu_case_free(c);
    gettimeofday(&stop, NULL);
    printf("\nDuration %f s\n", seconds(start, stop));
    return 0;
}
