/*
 * csu.h -- C ABI of the MI355X residual-and-update path (libcsu.so).
 *
 * The library replaces the hot loop of the reference solver: everything
 * below solverSolve's time loop (reference solver.c:1864-1907).  The host
 * keeps the reference's process contract (input.dat, mesh.su2, restart.csv,
 * solution.csv, convergence.csv); this header is the thin layer between
 * that host C and the HIP kernels.  Plain C types only; every call returns
 * CSU_OK (0) or a CSU_ERR_* code and leaves a message for csu_last_error().
 * Unlike the reference (printf + exit(0), e.g. input.c:156-160,
 * solver.c:1973-1977) nothing here terminates the process.
 *
 * Layout conventions (reference solver.h:86-170, mesh.h:2-52):
 *   state U, residual R, gradients, limiter   SoA [nvar][ncell] double
 *   interior faces (reference mesh->con)      [nface][4] = e0 e1 p0 p1
 *   per-cell signed face ids (ELEMENT->f)     [ncell][4], 1-based, 0 = none
 *   per-cell neighbour list (ELEMENT->neiL)   [ncell][4], >= 0 cell id,
 *                                             < 0 boundary face -(1+g)
 * Variable order: rho, rho*u, rho*v, rho*E [, rho*nu~ | rho*k, rho*omega].
 */
#ifndef CSU_H
#define CSU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSU_OK              0
#define CSU_ERR_ARG         1   /* bad argument / inconsistent mesh       */
#define CSU_ERR_HIP         2   /* HIP runtime failure                    */
#define CSU_ERR_UNSUPPORTED 3   /* option combination not implemented     */
#define CSU_ERR_NOMEM       4
#define CSU_ERR_IO          5   /* host file input/output                 */
#define CSU_ERR_INPUT       6   /* input.dat key missing or invalid       */
#define CSU_ERR_DIVERGED    7   /* residual no longer finite (server)     */

/* boundary types, reference boundary.c:479-511 */
enum { CSU_BC_SYMMETRY = 0, CSU_BC_INLET = 1, CSU_BC_OUTLET = 2,
       CSU_BC_WALL = 3, CSU_BC_WALLT = 4 };
/* convective fluxes, reference flux.c:14-46 */
enum { CSU_FLUX_ROE = 0, CSU_FLUX_AUSM = 1, CSU_FLUX_AUSMDV = 2,
       CSU_FLUX_AUSMPUP = 3, CSU_FLUX_AUSMPUP2 = 4 };
/* time schemes, reference solver.c:1957-1981 */
enum { CSU_TIME_RK = 0, CSU_TIME_LUSGS = 1, CSU_TIME_DPLUR = 2 };
/* arithmetic of the residual path */
enum { CSU_PREC_FP64 = 0,   /* everything FP64; least-squares gradients
                               from per-cell weights precomputed in long
                               double (parity path)                      */
       CSU_PREC_MIXED = 1,  /* FP32 reconstruction/gradient/limiter/flux,
                               FP64 state, accumulation, update, norms   */
       CSU_PREC_FP64_LD = 2 };/* FP64 with the reference's long double
                               least-squares sums emulated per call
                               (double-double), solver.c:1041-1073       */

/* Mesh after the reference's preprocessing (mesh.c:115-266): cells CCW,
 * interior faces in the reference's enumeration order, boundary faces
 * oriented CCW with respect to their cell.  Built by csu_case_load(). */
typedef struct csu_mesh_desc {
    int32_t ncell, nnode, nface, nbface, nmark, axi;
    const double*  node_xy;     /* [nnode][2]                              */
    const int32_t* cell_vert;   /* [ncell][4] vertex ids, -1 pads triangles */
    const int32_t* face;        /* [nface][4] e0 e1 p0 p1                   */
    const int32_t* bface;       /* [nbface][3] e0 p0 p1                     */
    const int32_t* bface_mark;  /* [nbface] marker index                    */
    const int32_t* mark_type;   /* [nmark] CSU_BC_*                         */
    const int32_t* cell_nbr;    /* [ncell][4] see above; unused slots INT32_MAX */
    const int32_t* cell_face;   /* [ncell][4] see above                     */
    const double*  wall_dist;   /* [ncell] wall distance (SA/SST) or NULL   */
} csu_mesh_desc;

/* Solver options: the input.dat keys of the reference (solver.c:1403-1658,
 * 1748-1798) after defaults, plus the free-stream state conditionState
 * derives from them (solver.c:40-89). */
typedef struct csu_params {
    int32_t nvar;          /* 4 Euler/laminar, 5 SA, 6 SST                */
    int32_t order;         /* 1 or 2                                      */
    int32_t flux;          /* CSU_FLUX_*                                  */
    int32_t stages;        /* RK stages 1,3,4,5 (2 is a no-op, as in ref) */
    int32_t time_scheme;   /* CSU_TIME_*                                  */
    int32_t dt_local;      /* 1: local time step                          */
    int32_t dt_local_n;    /* iteration after which dt_local switches off */
    int32_t laminar, sa, sst, sst_trans, tp;
    int32_t limiter;       /* 0 or 1 (limiter.c:45-94); 2: Barth-Jespersen
                              (limiterBJ, solver.c:1132, which the
                              reference never calls): an extension      */
    int32_t turb1order, visc_blazek, nlinear;
    int32_t precision;     /* CSU_PREC_*                                  */
    int32_t sa_cc;         /* SA: the saCC.c form (density-gradient terms) */
    double cfl, rlim, pout, twall, lim_k, efix, pr, pr_t, sref, wimp, turb_ratio;
    double gamma, gas_r;
    double mach_inf;       /* inlet Mach (AUSM+up reference Mach)         */
    double pin[7];         /* inlet primitive state r u v p T [nu~ | k om] */
    double uin[6];         /* inlet conservative state                    */
    double sst_kfactor, sst_ofactor, sst_owallfactor, sst_l;
    int32_t gradient;      /* 0: least squares (solver.c:1031-1074);
                              1: Green-Gauss (an extension: planar runs,
                              precision 0 or 1)                          */
} csu_params;

typedef struct csu_dev csu_dev;

/* Upload mesh + options, allocate all device state on the current device.
 * The initial state is the free stream prm->uin. */
int  csu_dev_create(const csu_mesh_desc* mesh, const csu_params* prm, csu_dev** out);
void csu_dev_destroy(csu_dev* dev);

/* Wall distance of every cell, reference saCalcD (sa.c:703-763): the
 * minimum over wall / wallT boundary faces of the distance from the cell's
 * vertex average to the face segment, computed on the current device.
 * wd: [ncell] host output.  csu_dev_create runs the same kernel itself when
 * SA or SST is on and mesh->wall_dist is NULL. */
int  csu_wall_distance_dev(const csu_mesh_desc* mesh, double* wd);

/* Host <-> device state, SoA [nvar][ncell] FP64. */
int  csu_dev_set_state(csu_dev* dev, const double* U);
int  csu_dev_get_state(csu_dev* dev, double* U);

/* One residual evaluation on the current state: reference solverCalcR
 * (solver.c:685-739), including the positivity clamps it writes back into U.
 * Any output pointer may be NULL.  R, dPx, dPy, phi: [nvar][ncell];
 * P: [ncell][nvar+1] primitive; Pb: [nbface][nvar+1] boundary ghost states. */
int  csu_dev_residual(csu_dev* dev, double* R, double* dPx, double* dPy,
                      double* phi, double* P, double* Pb);

/* niter iterations of the solverSolve loop body (solver.c:1866-1889),
 * with the residual norm (solverCalcRes, solver.c:867-892) and the force
 * coefficients (solverCalcCoeff3, solver.c:1342-1401) after each one.
 * res: [niter][nvar] or NULL; coef: [niter][4] = Cx_p Cx_v Cy_p Cy_v or NULL.
 * Iterations are counted across calls (the dt_local_n switch uses it). */
int  csu_dev_iterate(csu_dev* dev, int32_t niter, double* res, double* coef);

/* Change the CFL number between iterations (CFL ramps).  The next
 * iteration computes its time step from the current state with the new
 * number, as the reference does when a run is restarted from that state
 * with a new CFL in input.dat (its way of ramping the CFL). */
int  csu_dev_set_cfl(csu_dev* dev, double cfl);

/* Arithmetic of the following iterations (CSU_PREC_*), e.g. mixed
 * precision until the residual has dropped far enough and FP64 from there
 * on (mixed precision's FP32 fluxes level the residual off near a 1e-6
 * drop; FP64 continues to the FP64 steady state).  The state carries over;
 * works on a batch too (all its solves).  No counterpart in the
 * reference, which runs in FP64 throughout. */
int  csu_dev_set_precision(csu_dev* dev, int32_t precision);

/* solverCalcPrimitive(solver, solver->U) (solver.c:1153-1218): the
 * positivity clamps applied to the device state in place, and the boundary
 * ghost states recomputed from it, as the reference's solverWriteSolution2
 * does before solverWriteReestart writes restart.csv (main.c:29-31);
 * P: [ncell][nvar+1] = r u v p T [nu~ | k om], or NULL. */
int  csu_dev_primitives(csu_dev* dev, double* P);

/* Time-accurate run: the tube branch of solverSolve (solver.c:1912-1941).
 * Each step takes the global time step (CFL x the smallest local one), cut
 * so that t + dt does not pass tmax; t advances by dt*stages/2 per step and
 * the run ends after the cut step.  *t: start time in, end time out (the
 * reference starts at 0).  max_iter > 0 caps the number of steps (0: no
 * cap); *niter receives the steps taken and *finished 1 when the cut step
 * was taken (either may be NULL).  Needs the RK scheme and dt_local 0. */
int  csu_dev_advance_time(csu_dev* dev, double tmax, int32_t max_iter, double* t,
                          int32_t* niter, int32_t* finished);

/* Asynchronous form of csu_dev_iterate for callers that keep several
 * contexts busy at once: enqueue niter iterations on the context's own
 * stream and return; csu_dev_ready reports 1 once they (and the copy of
 * their history) are done, 0 while running; csu_dev_async_result waits and
 * returns the histories of that call ([niter][nvar], [niter][4]; NULL
 * allowed).  No counterpart in the reference (one solve per process). */
int  csu_dev_iterate_async(csu_dev* dev, int32_t niter);
int  csu_dev_ready(csu_dev* dev);
int  csu_dev_async_result(csu_dev* dev, double* res, double* coef);

/* Device bytes held by a context, and an upper estimate of what
 * csu_dev_create would allocate for a mesh and options with a history ring
 * of hist iterations (used to admit solves by HBM footprint). */
int64_t csu_dev_bytes(const csu_dev* dev);
int64_t csu_dev_footprint(const csu_mesh_desc* mesh, const csu_params* prm, int32_t hist);

/* Wall shear of the last residual evaluation, per boundary face
 * ([nbface][4] = txx txy tyy yp; rows of wall / wallT faces are set), as
 * solverCalcCoeff2 writes it to friction.csv (solver.c:1303-1340):
 * boundaryCalcTensorWall (boundary.c:744-793) for laminar runs,
 * saCalcTensorWall (sa.c:795-852) for SA and SST.  Viscous runs only. */
int  csu_dev_wall_tensor(csu_dev* dev, double* out);

/* SST runs: the per-cell fields sstInterMiT (sst.c:964-1007) computes for
 * solverWriteSolution2 (solver.c:2132), out [ncell][5] = eddy viscosity,
 * F1, F2, wall distance, omega: sstSources on the current primitives (call
 * csu_dev_primitives first, as the reference does) and the gradients of
 * the last residual evaluation. */
int  csu_dev_sst_fields(csu_dev* dev, double* out);

/* SA and SST runs: per boundary face, what the surfData.csv writers take
 * (saSolverWriteSurf sa.c:854, saCC_SolverWriteSurf saCC.c:599,
 * sstSolverWriteSurf sst.c:820), out [nbface][11] = the boundary viscous
 * flux x, y, energy and the eddy viscosity of the boundary viscous-flux
 * routine on the current primitives and the last residual's gradients,
 * the ghost state of the last residual (r u v p T, nu~ or k) and the wall
 * distance of the face's cell. */
int  csu_dev_surf_faces(csu_dev* dev, double* out);

/* Allocation-free form of csu_dev_surf_faces: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): the
 * per-face data is staged in a fixed device array of the library
 * (11 x CSU_SURF_NOALLOC_MAX_NB doubles = 5.8 MB, one per GPU, held until
 * the process ends) instead of a buffer allocated per call; nothing is
 * allocated.  Same output as csu_dev_surf_faces.  Contexts with more than
 * CSU_SURF_NOALLOC_MAX_NB boundary faces are refused (CSU_ERR_ARG).  All
 * calls on one GPU share that array: calls that overlap on one GPU (from
 * several host threads) race. */
#define CSU_SURF_NOALLOC_MAX_NB 65536
int  csu_dev_surf_faces_noalloc(csu_dev* dev, double* out);

/* ---- Batches: many independent steady solves in one context ----------
 * No counterpart in the reference, which runs one solve per process
 * (solverSolve, solver.c:1836-1943); each solve of a batch follows that
 * loop's semantics and gives, bit for bit, the results it gives alone.
 * The meshes are concatenated into one set of arrays (each solve's cells,
 * faces and boundary faces starting on a multiple of 256), so one set of
 * kernel launches -- captured in one graph per iteration -- covers every
 * solve.  Per solve: free stream, outlet pressure, wall temperature,
 * AUSM+up reference Mach, limiter reference values, CFL number, reference
 * length, and every reduction (residual norm, forces, global time step,
 * limiter range, DPLUR stop test).  Options that select the kernels
 * (model, flux, order, limiter, time scheme, stages, precision, gas, ...)
 * must be the same for all solves (CSU_ERR_ARG otherwise): group unlike
 * solves into separate batches.  csu_dev_iterate, csu_dev_iterate_async,
 * csu_dev_async_result, csu_dev_sync, csu_dev_set_graphs and the profiling
 * calls work on a batch; their histories then hold [niter][n][nvar] norms
 * and [niter][n][4] forces.  The whole-mesh calls (set/get_state,
 * residual, primitives, set_cfl, advance_time, wall tensor, SST fields,
 * surface data) refuse a batch; a tube (time-accurate) run is not batched. */
int  csu_batch_create(int32_t n, const csu_mesh_desc* const* meshes, const csu_params* const* prms,
                      csu_dev** out);
/* 1 if two solves can share a batch (the options that select the kernels
 * match), 0 if not, < 0 on a null argument.  Host arithmetic only. */
int  csu_batch_compatible(const csu_mesh_desc* mesh_a, const csu_params* prm_a, const csu_mesh_desc* mesh_b,
                          const csu_params* prm_b);
int  csu_batch_size(const csu_dev* dev);                    /* solves, or -1 */
int  csu_batch_ncell(const csu_dev* dev, int32_t solve);    /* cells of a solve, or -1 */
/* State of one solve, SoA [nvar][ncell of the solve] FP64. */
int  csu_batch_set_state(csu_dev* dev, int32_t solve, const double* U);
int  csu_batch_get_state(csu_dev* dev, int32_t solve, double* U);
/* CFL number of one solve, as csu_dev_set_cfl for a single solve. */
int  csu_batch_set_cfl(csu_dev* dev, int32_t solve, double cfl);
/* Hold (on = 0) or resume a solve: a held solve's state stays as it is
 * while the others iterate (its history rows are not meaningful). */
int  csu_batch_set_active(csu_dev* dev, int32_t solve, int32_t on);

typedef struct csu_batch_result {
    int32_t status;        /* 0, CSU_ERR_DIVERGED, or -1: held by the caller */
    int32_t converged;     /* reached the drop                            */
    int32_t iters;         /* iterations up to the stopping one           */
    int32_t iter_conv;     /* iteration that reached the drop, or -1      */
    int32_t iters_run;     /* iterations run (whole chunks)               */
    int32_t iter_switch;   /* mixed precision: iterations before FP64, or -1 */
    double res0, res_last; /* density residual of the first / stopping iteration */
    double res_peak;       /* largest density residual                    */
    double coef_last[4];   /* Cx_p Cx_v Cy_p Cy_v of the stopping iteration */
    double cfl;            /* CFL number at the end                       */
} csu_batch_result;

/* Run every active solve until its density residual (solverCalcRes,
 * solver.c:867-892) has dropped by `drop` from its first iteration
 * (peak_ref 0) or from its largest so far (peak_ref 1), is no longer
 * finite, or max_iter iterations are done, `chunk` iterations at a time; a
 * solve is held from the end of the chunk in which it stopped, so its
 * state is that of its run alone after iters_run iterations.  With
 * growth > 1 each running solve's CFL is multiplied by growth after every
 * chunk, up to cfl_max[solve] (cfl_max NULL: no ramp).  A mixed-precision
 * batch with switch_drop > 0 continues in FP64 as soon as one running solve
 * has reached that drop or stalled (no 10 % fall of its residual in four
 * chunks once below a tenth of its peak).  out: [n]. */
int  csu_batch_run(csu_dev* dev, double drop, int32_t max_iter, int32_t chunk,
                   const double* cfl_max, double growth, int32_t peak_ref, double switch_drop,
                   csu_batch_result* out);

/* Launch-shape autotuner (no counterpart in the reference).  Before the
 * first iteration, time every candidate -- workgroup size of the gradient
 * and face kernels (64, 128, 256), of the LU-SGS level launches (64, 128,
 * 256), XCD-contiguous or plain tile order, and for inviscid second-order
 * RK runs the tile form (the face fluxes of each 256-cell tile evaluated
 * into LDS by the update kernel, no face-flux kernel or array) -- over
 * `iters` iterations (the better of two runs) and keep the fastest; the
 * context then continues from the state it had, and later contexts of the
 * same kind of run and mesh-size bucket (log2 of the cell count) in the
 * process take the same choice without timing.  Results do not depend on
 * these choices (an untuned context takes the tile form from 200k cells
 * on).  choice: [4] (block, LU-SGS block, XCD map, tile) or NULL.
 * csu_dev_launch_shape reports the shapes in use (returns 1 if tuned). */
int  csu_dev_autotune(csu_dev* dev, int32_t iters, int32_t* choice);
int  csu_dev_launch_shape(const csu_dev* dev, int32_t* choice);
/* Set the launch shapes explicitly (choice [4] as above: 64, 128 or 256;
 * 64, 128 or 256; 0 or 1; 0 or 1, 1 only where the tile form exists), e.g.
 * one the autotuner chose in an earlier process, or a fixed one for
 * profiling runs.  Results do not change. */
int  csu_dev_set_launch_shape(csu_dev* dev, const int32_t* choice);

/* Read + write rate (GB/s) of a plain FP64 streaming kernel over two
 * arrays of `bytes` each on the current device, best of `reps` launches:
 * the practical HBM ceiling the stage kernels are measured against (no
 * counterpart in the reference). */
int  csu_stream_rate(int64_t bytes, int32_t reps, double* gbps);

/* Iterations done so far; local time step [ncell] of the last iteration. */
int  csu_dev_iteration(const csu_dev* dev);
int  csu_dev_get_dt(csu_dev* dev, double* dtl);

/* Iterations are captured into HIP graphs and replayed (two RK iterations
 * or one implicit iteration per graph) unless profiling is on or this is
 * switched off; results are the same either way.  Default: on. */
int  csu_dev_set_graphs(csu_dev* dev, int32_t on);

/* Block until all queued device work of this context is done. */
int  csu_dev_sync(csu_dev* dev);

/* Average duration (ms) of each kernel of the last csu_dev_iterate call,
 * measured with HIP events on the library's stream when profiling is on.
 * names: receives up to max entries of static strings. Returns the count. */
int  csu_dev_set_profiling(csu_dev* dev, int32_t on);
int  csu_dev_kernel_times(csu_dev* dev, const char** names, double* avg_ms,
                          int32_t* launches, int32_t max);

const char* csu_last_error(void);
const char* csu_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CSU_H */
