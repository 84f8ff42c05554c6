/*
 * sd2d_oracle.h - CPU restatement of the reference SD2D-MHD residual and
 * Runge-Kutta path, used ONLY as the checker (tests, smoke check) and as
 * the CPU baseline of bench.py.  Nothing in the product links it.
 *
 * Data are kept exactly as the reference's module setup2d keeps them
 * (Fortran column-major, 1-based tables) and every routine follows the
 * reference routine it names, operation by operation, so that when it is
 * compiled without FMA contraction its results equal the reference's
 * bit for bit on the same inputs.
 */
#ifndef SD2D_ORACLE_H
#define SD2D_ORACLE_H

#include "../include/sd2d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const sd2d_mesh *mesh, const sd2d_params *params);
void orc_destroy(orc_ctx *o);
int orc_set_state(orc_ctx *o, const double *q, const sd2d_maxima *mx);
int orc_get_state(orc_ctx *o, double *q, sd2d_maxima *mx);
void orc_set_time(orc_ctx *o, double ctime, int iter);
void orc_get_time(orc_ctx *o, double *ctime, int *iter);
/* compresid: outputs may be NULL */
int orc_compresid(orc_ctx *o, double *resid, double *nablaqs);
/* nsteps calls of RK_S5P3 / FiveStageRk (by k_stage); norm_every as in
 * sd2d_step (0 = never) */
int orc_step(orc_ctx *o, int nsteps, int norm_every);
/* last c_h, epsilon0, norms */
void orc_diag(orc_ctx *o, double *c_h, double *epsilon0, double *resid_rms,
              double *divb_rms);
int orc_get_av(orc_ctx *o, double *sm_ind, double *muav);
/* number of OpenMP threads the oracle uses (1 if built without OpenMP) */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
