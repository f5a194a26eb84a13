/*
 * csu_oracle.h -- CPU restatement of the reference's residual-and-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load liboracle (oracle/lib/); the product
 * (libcsu.so) never does.  It takes the same csu_mesh_desc / csu_params as
 * the device library and restates, in plain C with x87 long double where the
 * reference uses it, the reference functions cited in csu_oracle.c.
 */
#ifndef CSU_ORACLE_H
#define CSU_ORACLE_H
#include "csu.h"

typedef struct csu_oracle csu_oracle;

int  oracle_create(const csu_mesh_desc* m, const csu_params* p, csu_oracle** out);
void oracle_destroy(csu_oracle* o);
int  oracle_set_state(csu_oracle* o, const double* U);
int  oracle_get_state(csu_oracle* o, double* U);
int  oracle_get_dt(csu_oracle* o, double* dtl);
/* R, dPx, dPy, phi [nvar][ncell]; P [ncell][nvar+1]; Pb [nbface][nvar+1] */
int  oracle_residual(csu_oracle* o, double* R, double* dPx, double* dPy, double* phi,
                     double* P, double* Pb);
/* res [niter][nvar], coef [niter][4] (Cx_p Cx_v Cy_p Cy_v) */
int  oracle_iterate(csu_oracle* o, int niter, double* res, double* coef);
int  oracle_set_cfl(csu_oracle* o, double cfl);
int  oracle_advance_time(csu_oracle* o, double tmax, int max_iter, double* t, int* niter, int* finished);
/* P [ncell][nvar+1]: the primitive pass alone (clamps U, keeps the gradients) */
int  oracle_primitives(csu_oracle* o, double* P);
/* SST: [ncell][5] eddy viscosity, F1, F2, wall distance, omega */
int  oracle_sst_fields(const csu_oracle* o, double* out);
/* SA/SST: [nbface][11] wall viscous flux x y e, eddy viscosity, ghost
   r u v p T (nu~ | k), wall distance of the face's cell */
int  oracle_surf_faces(const csu_oracle* o, double* out);
int  oracle_wall_tensor(const csu_oracle* o, double* out);   /* [nbface][4] txx txy tyy yp */
int  oracle_set_threads(int n);
/* DPLUR test hook: D [nc], A [nc][4][nv*nv], nb [nc][4], nrm [nc][4][3],
   dq [6][nc], R [nv][nc], dW [nv][nc] of one implicit step */
int  oracle_dplur_system(csu_oracle* o, double* D, double* A, int* nb, double* nrm, double* dq, double* R,
                         double* dW);
/* Face-flux test hook: the convective flux of the oracle's options between
   face-frame states L and R ([nv]: r, u_n, u_t, p, scalars) into f [nv] */
int  oracle_face_flux(const csu_oracle* o, const double* L, const double* R, double* f);

#endif
