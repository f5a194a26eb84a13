/*
 * csu_host.h -- host-side C of the drop-in: case files in, case files out.
 *
 * These calls replace the reference's setup and output code around the hot
 * path (input.c, mesh.c, readTables.c, solverSetData/solverInitDomain and the
 * writers in solver.c) with the same file formats.  They feed csu_dev_* in
 * csu.h.  Paths are case-directory prefixes ending in '/', as the reference
 * takes argv[1] (main.c:25, solver.c:1734-1739).
 */
#ifndef CSU_HOST_H
#define CSU_HOST_H

#include <stdint.h>
#include "csu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* input.dat key table: reference input.c:7-115 grammar. */
typedef struct csu_input csu_input;
int         csu_input_read(const char* path, csu_input** out);
void        csu_input_free(csu_input* in);
int         csu_input_count(const csu_input* in);
/* last duplicate wins (input.c:142-163); NULL when absent */
const char* csu_input_get(const csu_input* in, const char* name);
const char* csu_input_name(const csu_input* in, int i);
const char* csu_input_value(const csu_input* in, int i);

/* A loaded case: mesh after the reference's preprocessing, options, state. */
typedef struct csu_case {
    csu_mesh_desc mesh;        /* views into the arrays below              */
    csu_params    prm;
    double*  U;                /* [nvar][ncell] initial state              */
    int32_t  nmax;             /* Nmax                                     */
    int32_t  threads;          /* threads (host CPU paths only)            */
    int32_t  restart;
    int32_t  tube;             /* 1: unsteady shock-tube mode to t = tmax  */
    double   tmax;             /* tube: final time                         */
    char     write_surf[64];
    char     wd[1024];
    char   (*mark_name)[64];   /* [nmark]                                  */
    /* owned storage */
    double*  node_xy;
    int32_t* cell_vert;
    int32_t* face;
    int32_t* bface;
    int32_t* bface_mark;
    int32_t* mark_type;
    int32_t* cell_nbr;
    int32_t* cell_face;
    double*  wall_dist;
    csu_input* input;
} csu_case;

/* Read <wd>input.dat and <wd>mesh.su2 (and <wd>restart.csv when restart=1),
 * build the connectivity in the reference's order and the initial state. */
int  csu_case_load(const char* wd, csu_case** out);
void csu_case_free(csu_case* c);

/* Mesh only (no input.dat): SU2 file -> connectivity.  nvar only sizes the
 * state; axi selects the axisymmetric volume used by the orientation fix. */
int  csu_mesh_read_su2(const char* path, int32_t axi, csu_case* c);

/* Wall distance: min point-to-segment distance from each cell's vertex
 * average to the wall/wallT boundary faces (reference sa.c:703-793). */
int  csu_wall_distance(csu_case* c);

/* Writers, formats of the reference: restart.csv (solver.c:327-355),
 * solution.csv (solver.c:222-325). */
int  csu_write_restart(const csu_case* c, const double* U, const char* path);
int  csu_read_restart(const csu_case* c, double* U, const char* path);
int  csu_write_solution(const csu_case* c, const double* U, const char* path);
/* solution2.csv (solverWriteSolution2, solver.c:2009-2220) from the clamped
 * primitive state P [ncell][nvar+1] (csu_dev_primitives); SST runs also
 * take sst [ncell][5] (csu_dev_sst_fields), NULL for the other models. */
int  csu_write_solution2(const csu_case* c, const double* P, const double* sst, const char* path);
/* surfData.csv (saSolverWriteSurf sa.c:854, saCC_SolverWriteSurf
 * saCC.c:599, sstSolverWriteSurf sst.c:820) for SA and SST runs, from the
 * per-face data of csu_dev_surf_faces ([nbface][11]); the marker is the
 * writeSurf key's (default "wall").  Laminar runs: CSU_ERR_UNSUPPORTED. */
int  csu_write_surf(const csu_case* c, const double* faces, const char* path);

/* Greedy colouring for the colour/RCM renumbering (csu_amd/renumber.py):
 * cells visited in order[0..n) each take the smallest colour no coloured
 * neighbour (CSR indptr/indices) has.  No counterpart in the reference. */
int  csu_greedy_colour(int32_t n, const int64_t* indptr, const int32_t* indices, const int64_t* order,
                       int32_t* colour);

/* Name of the input.dat key whose absence (or bad value) made the last
 * csu_case_load fail with CSU_ERR_INPUT. */
const char* csu_host_missing_key(void);

#ifdef __cplusplus
}
#endif
#endif /* CSU_HOST_H */
