/* csu_serve.h -- continuous-batching solve server over csu_dev contexts.
 *
 * No counterpart in the reference, which runs one solve per process
 * (solver.c:1836-1943).  Each admitted solve is an independent device
 * context with its own stream (or, with batching on, a batch of solves in
 * one context); the server keeps several in flight so their kernels
 * overlap on the GPU, admits queued solves as slots and HBM budget
 * free up, and retires a solve when its density residual has dropped by the
 * requested factor from its first iteration (or its iteration cap is hit).
 * Because contexts share nothing and every reduction is a fixed tree, a
 * solve's results are bit-identical whether it runs alone or in the mix.
 *
 * Thread model: one host thread drives a server.
 */
#ifndef CSU_SERVE_H
#define CSU_SERVE_H
#include "csu.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct csu_server csu_server;

typedef struct {
    int32_t id;            /* submit order, from 0                                */
    int32_t status;        /* CSU_OK, CSU_ERR_DIVERGED (non-finite residual: the  */
                           /* solve is stopped; its state is kept like a finished */
                           /* one's), or the error of create / iterate            */
    int32_t converged;     /* 1: drop reached; 0: iteration cap reached           */
    int32_t iters;         /* iterations run                                       */
    int32_t iter_conv;     /* first iteration (1-based) at the drop, -1 if none    */
    int32_t ncell;
    double t_submit, t_admit, t_done;   /* seconds on the server clock             */
    double res0, res_last; /* density residual of the first and last iteration     */
    double res_peak;       /* largest density residual                             */
    double coef_last[4];   /* Cx_p Cx_v Cy_p Cy_v of the last iteration            */
    int64_t bytes;         /* device bytes the solve held                          */
    int32_t iter_switch;   /* iterations run in mixed precision before FP64, or -1 */
    int32_t pad_;
} csu_job_result;

/* max_slots: solves in flight at once; hbm_budget: device bytes the
 * server's contexts may hold together; chunk: iterations enqueued per
 * solve between convergence checks. */
int  csu_server_create(int32_t max_slots, int64_t hbm_budget, int32_t chunk, csu_server** out);
void csu_server_destroy(csu_server* s);

/* Batched units (no counterpart in the reference).  max_solves > 1: each
 * slot holds a batch context (csu_batch_create) instead of one solve: at
 * admission the oldest queued solve and every later queued solve that can
 * share its kernels (csu_batch_compatible) are packed into one batch, up to
 * max_solves solves and max_cells cells (0: no cell limit) within the HBM
 * budget, and run with one set of launches per iteration.  A solve of a
 * batch is held once it stops and its result is reported at once; the
 * batch's memory is freed when its last solve stops.  Results are those of
 * the solve alone, bit for bit, with two exceptions: a batch runs whole
 * chunks, so a solve stopped by its iteration cap keeps the state of the
 * end of that chunk; and a mixed-precision batch switches to FP64 as a
 * whole, as soon as one running solve has reached its switch drop or
 * stalled (then the converged results agree to the converged tolerance,
 * not bit for bit).  max_solves 1 (the default): one context per solve.  Set
 * before the first csu_server_step. */
int  csu_server_set_batching(csu_server* s, int32_t max_solves, int64_t max_cells);

/* Queue a solve.  mesh and prm (and U0 if not NULL: initial state
 * [nvar][ncell], otherwise the free stream) must stay valid until the solve
 * is admitted.  drop: residual drop that counts as converged (1e-8);
 * max_iter: iteration cap; keep_state: keep the final state for
 * csu_server_state.  *id receives the job id. */
int  csu_server_submit(csu_server* s, const csu_mesh_desc* mesh, const csu_params* prm, const double* U0,
                       double drop, int32_t max_iter, int32_t keep_state, int32_t* id);

/* Admit what fits, keep every admitted solve's stream fed, wait up to
 * max_wait seconds for at least one chunk to finish and retire finished
 * solves.  Returns the number of solves retired by this call, or < 0. */
int  csu_server_step(csu_server* s, double max_wait);

/* CFL ramp for a job still in the queue: after each chunk its CFL is
 * multiplied by growth (>= 1) up to cfl_max (csu_dev_set_cfl), the
 * device-side form of the reference's chain of restarts at rising CFL. */
int  csu_server_set_ramp(csu_server* s, int32_t id, double cfl_max, double growth);

/* Options of a job still in the queue.  peak_ref 1: residual drops are
 * measured from the largest density residual so far instead of the first
 * one (a solve started from a uniform free stream past a wall or an airfoil
 * has a first density residual near rounding level, which makes drops from
 * it meaningless).  switch_drop > 0 (a mixed-precision job): once the drop
 * reaches it, or the residual has not fallen by 10 % in four chunks once
 * below a tenth of its peak (the level where FP32 flux rounding dominates
 * it), the job continues in FP64 (csu_dev_set_precision), so it converges
 * to the FP64 steady state. */
int  csu_server_set_options(csu_server* s, int32_t id, int32_t peak_ref, double switch_drop);

/* Move up to max finished results into out; returns the count. */
int  csu_server_results(csu_server* s, csu_job_result* out, int32_t max);

/* Queued (not yet admitted) and active solves. */
int  csu_server_pending(const csu_server* s, int32_t* queued, int32_t* active);

/* Final state of a finished job submitted with keep_state. */
int  csu_server_state(const csu_server* s, int32_t id, double* U);

/* Seconds since csu_server_create. */
double csu_server_clock(const csu_server* s);

#ifdef __cplusplus
}
#endif
#endif /* CSU_SERVE_H */
