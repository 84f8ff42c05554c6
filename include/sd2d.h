/*
 * sd2d.h - C ABI of the MI355X (gfx950) SD2D-MHD residual and time-step path.
 *
 * The reference solver keeps the whole discretisation in the Fortran module
 * `setup2d` and calls an argument-less `subroutine compresid`
 * (compresid.f90:1) five times per step from RK_S5P3 (iterations.f90:183-209).
 * This ABI replaces that call.  The host (Fortran through ISO_C_BINDING, or
 * any C caller) hands over the arrays that `setup2d` holds after INIT_SETUP,
 * in the reference's own Fortran layouts and 1-based index conventions, once;
 * afterwards the state lives in HBM and each call has the per-call contract
 * of `compresid`.
 *
 * Conventions
 *   - N is the number of solution points per direction (p+1), 3 <= N <= 5.
 *   - Every array pointer is a plain host pointer to column-major (Fortran)
 *     data; integer tables are 1-based exactly as the reference stores them.
 *   - Every entry point returns SD2D_OK (0) or one of the SD2D_E* codes;
 *     sd2d_last_error() gives a message.  Nothing exits or throws.
 *   - One context per GPU.  Calls on one context are not re-entrant; all
 *     device work runs on the context's own HIP stream and every call
 *     returns after that work has completed.
 */
#ifndef SD2D_H
#define SD2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SD2D_OK = 0,
  SD2D_EINVAL = 1,       /* bad argument, inconsistent mesh table          */
  SD2D_EHIP = 2,         /* a HIP runtime call failed                      */
  SD2D_ENOMEM = 3,       /* device or host allocation failed               */
  SD2D_ENONFINITE = 4,   /* opt-in finiteness check found a NaN/Inf        */
  SD2D_EUNSUPPORTED = 5  /* N or mode outside what the kernels implement   */
};

/* Opaque per-GPU context. */
typedef struct sd2d_ctx sd2d_ctx;

/* Precision of the device path.  SD2D_MIXED: interpolation, fluxes,
 * Rusanov and derivatives in FP32 (operators and metrics rounded to FP32);
 * the state, the Jacobian scaling, the sources, the RK combination, the
 * norms and the AV sensor in FP64.  It applies to sd2d_step;
 * sd2d_compresid always evaluates in FP64. */
enum { SD2D_FP64 = 0, SD2D_MIXED = 1 };

/* Tile shape: the device processes cells in tiles of 12/16/20 (N = 5/4/3),
 * one workgroup per tile.  SD2D_TILE_STRIPS takes consecutive cells of the
 * mesh numbering (on a mesh numbered row by row, each row padded to a
 * multiple of the tile when that adds at most 1/16, so that the tiles of
 * consecutive rows line up); SD2D_TILE_BLOCKS groups compact patches of cells
 * (recursive bisection of the cell adjacency graph), so fewer neighbour
 * traces come from other tiles, but those that do are scattered over more
 * tiles.  SD2D_TILE_ROWPAIRS takes, on a mesh numbered row by row, two
 * consecutive rows of 6/8/10 cells each, so that the faces between the two
 * rows lie inside the tile and every line across the rows gathers at most
 * one end from another tile (a last odd row is cut into strips; a mesh not
 * numbered row by row gets strips).  SD2D_TILE_DEFAULT is
 * SD2D_TILE_ROWPAIRS (measured fastest on structured O-grid and periodic
 * meshes: 1-2 % over strips, which beat blocks).  Results do not depend on
 * the shape (every face is evaluated with the same operands either way);
 * only the speed does.  No counterpart in the reference, whose cells are
 * processed in file order. */
enum { SD2D_TILE_DEFAULT = 0, SD2D_TILE_STRIPS = 1, SD2D_TILE_BLOCKS = 2, SD2D_TILE_ROWPAIRS = 3 };

/*
 * Mesh, operators and metrics exactly as module setup2d holds them after
 * INIT_SETUP (initsetup.f90:1-169) on one rank.
 *   if2c      IF2C(NFACE,4)       connectivity.f90:156-216 (left cell, right
 *                                 cell, left local face, right local face)
 *   ibf*      IBFINTER/IBFINL/IBFOUT/IBFWAL/IBFSYMP/IBFREE face lists
 *                                 (connect_bdry.f90:114-244)
 *   cyc2fpair CYC2FPAIR(NPAIR,2)  (map_cyc_loc.f90:10-35)
 *   lmat,mmat Lmat(N+1,N), Mmat(N+1,N) (initsetup.f90:262-319)
 *   xs,xs_low Xs(N), Xs_low(N-1)  (initsetup.f90:171-259, :43-44)
 *   jac,s1,s2 Jac(N,N,NCELL), S1(2,2,N+1,N,NCELL), S2(2,2,N,N+1,NCELL)
 *                                 (calcjacob.f90:55,85-89,121-125), from
 *                                 4-node cells or, CURVE_WALL = 1, from the
 *                                 8 nodes of every cell (calcjacob.f90:26-53)
 * The arrays are copied; the caller may free them after sd2d_create.
 * The two cells of an interior or periodic face must give its face points
 * one normal up to sign, to 1e-9 of its length: the reference evaluates
 * the face with the left cell's normal (interfaceflux_all.f90:47,54,
 * cyclocflux_all.f90:47,54).  Where the right cell's own S1/S2 vector
 * differs from it by rounding (8-node cells), the device uses the left's
 * at that face point; a larger difference (a non-conforming face) is
 * refused with SD2D_EUNSUPPORTED.
 */
typedef struct sd2d_mesh {
  int32_t N;
  int32_t ncell;        /* NCELL (== NCELLGLOB on one rank)                 */
  int32_t nface;        /* NFACE                                            */
  const int32_t *if2c;
  int32_t ninter;  const int32_t *ibfinter;
  int32_t ninlet;  const int32_t *ibfinl;
  int32_t noutlet; const int32_t *ibfout;
  int32_t nwall;   const int32_t *ibfwal;
  int32_t nsymp;   const int32_t *ibfsymp;
  int32_t nfree;   const int32_t *ibfree;
  int32_t npair;   const int32_t *cyc2fpair;
  const double *lmat;
  const double *mmat;
  const double *xs;
  const double *xs_low;
  const double *jac;
  const double *s1;
  const double *s2;
} sd2d_mesh;

/* Run parameters: the QUAD.INP records the path reads (readinput.f90:6-43)
 * plus the hard-coded GLM alpha (initsetup.f90:165). */
typedef struct sd2d_params {
  int32_t vismode;      /* 0 or 1 (bcflux.f90:207-227 wall ghost)          */
  int32_t artiv;        /* 1: evaluate the AV smoothness sensor (AV.f90)   */
  double mu, eta0;      /* viscosity, resistivity                          */
  double alpha;         /* GLM damping, 2.0 in the reference               */
  double rinf, pinf, uinf, vinf, bxinf, byinf;  /* free stream             */
  double dxcycl, dycycl;                        /* periodic lengths        */
  int32_t ncellglob;    /* NCELLGLOB, used by the AV length scale          */
  int32_t k_stage;      /* 3: RK_S5P3, 5: FiveStageRk (iterations.f90:22)  */
  double dt;            /* fixed time step                                 */
  int32_t precision;    /* SD2D_FP64 or SD2D_MIXED                         */
  int32_t tile_shape;   /* SD2D_TILE_*: which cells share a workgroup      */
} sd2d_params;

/* Running maxima of the GLM wave speed (F1 in the path description):
 * eigvmax/umax are the local running maxima, eigvmax_g/umax_g the values
 * after the last MPI_ALLREDUCE (compflux.f90:10-11).  They are part of the
 * state: c_h of the next call is sqrt(eigvmax_g*(eigvmax_g-umax_g)). */
typedef struct sd2d_maxima {
  double eigvmax, umax, eigvmax_g, umax_g;
} sd2d_maxima;

/* Diagnostics after the most recent residual evaluation. */
typedef struct sd2d_diag {
  double resid_rms;     /* sqrt(sum resid^2 /(NCELLGLOB*N*N*9))  iterations.f90:61 */
  double divb_rms;      /* sqrt(sum divB^2  /(NCELLGLOB*N*N))    iterations.f90:62 */
  double eigvmax_g, umax_g;
  double c_h;           /* c_h used by the last residual evaluation        */
  double epsilon0;      /* AV.f90:126-127                                  */
  double ctime;
  int32_t iter;
  int32_t norms_valid;  /* 1 if resid_rms/divb_rms belong to the last step */
} sd2d_diag;

/* Create a context on the current HIP device.  Replaces the allocation and
 * set-up of the path's module arrays in INIT_SETUP (initsetup.f90:12-80). */
int sd2d_create(const sd2d_mesh *mesh, const sd2d_params *params,
                sd2d_ctx **out);

/* Free every device and host resource of the context. */
void sd2d_destroy(sd2d_ctx *ctx);

/* Message for the last error on this context (or a global one if ctx is
 * NULL); never NULL. */
const char *sd2d_last_error(const sd2d_ctx *ctx);

/* Upload Q(9,N,N,NCELL) (setup2d.f90:54) and, if not NULL, the running
 * maxima.  Replaces the IC/restart load of INIT_SETUP
 * (initsetup.f90:154-167). */
int sd2d_set_state(sd2d_ctx *ctx, const double *q,
                   const sd2d_maxima *maxima);

/* Download Q(9,N,N,NCELL) and/or the running maxima (either may be NULL). */
int sd2d_get_state(sd2d_ctx *ctx, double *q, sd2d_maxima *maxima);

/* Simulation clock carried with the state (setup2d.f90:78-79). */
int sd2d_set_time(sd2d_ctx *ctx, double ctime, int32_t iter);

/*
 * The per-call contract of `compresid` (compresid.f90:1-80, with
 * compflux.f90:1-33 and everything it calls): from the device state Q and
 * the lagged maxima, compute resid(9,N,N,NCELL) and
 * nablaQs(9,2,N,N,NCELL) and update the running maxima and the AV sensor.
 * resid_out / nablaqs_out are host pointers in Fortran layout; either may
 * be NULL, in which case that result is not copied back.
 */
int sd2d_compresid(sd2d_ctx *ctx, double *resid_out,
                   double *nablaqs_out);

/* Convenience for a Fortran host that keeps Q in its own module array:
 * sd2d_set_state(ctx, q, NULL) followed by sd2d_compresid. */
int sd2d_compresid_q(sd2d_ctx *ctx, const double *q,
                     double *resid_out, double *nablaqs_out);

/*
 * sd2d_compresid without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc), since sd2d_compresid allocates its nablaQs staging on its
 * first call.  Same contract and bitwise the same results as
 * sd2d_compresid; the nablaQs staging is a static device buffer of
 * SD2D_NOALLOC_MAX_DOUBLES doubles (64 MiB) on each GPU, held until the
 * process ends, so 2*9*N*N*NCELL must not exceed it (SD2D_EINVAL
 * otherwise; at N = 5 up to 18,641 cells).  Every call on one GPU uses
 * the same buffer: calls that overlap on one GPU (contexts driven from
 * several host threads) race.  Profiling (sd2d_set_profiling) must be off.
 */
#define SD2D_NOALLOC_MAX_DOUBLES (8 * 1024 * 1024)
int sd2d_compresid_noalloc(sd2d_ctx *ctx, double *resid_out,
                           double *nablaqs_out);

/*
 * Device-resident time stepping: `nsteps` calls of RK_S5P3
 * (iterations.f90:158-213) or FiveStageRk (iterations.f90:88-156) as chosen
 * by params.k_stage, including the ctime bookkeeping.  When norm_every > 0,
 * the residual and divB RMS norms of iterations.f90:34-62 are evaluated on
 * the steps with mod(iter, norm_every) == 1 and the latest is available
 * from sd2d_get_diag.
 */
int sd2d_step(sd2d_ctx *ctx, int32_t nsteps, int32_t norm_every);

/*
 * One stage (1..5) of the current step of the configured scheme, for hosts
 * that drive the stages themselves (the calls of compresid inside RK_S5P3,
 * iterations.f90:183-209, and FiveStageRk, iterations.f90:107-144).
 * Stages run in order: `stage` must be the next one of the step
 * (SD2D_EINVAL otherwise); stage 1 starts a step (iter + 1) and stage 5
 * completes it (the new state becomes the one sd2d_get_state returns, and
 * ctime advances by dt).  Five calls give bitwise the result of one
 * sd2d_step(ctx, 1, 0).  No norms are formed.  While a step is in
 * progress (between stage 1 and stage 5) the state buffer holds a stage's
 * working values, so sd2d_step, sd2d_compresid, sd2d_get_state with a
 * state pointer, sd2d_write_restart, sd2d_fwd, sd2d_vjp and
 * sd2d_compresid_vjp return SD2D_EINVAL; sd2d_get_state with only the
 * maxima pointer is allowed, and sd2d_set_state with a state abandons the
 * step.
 */
int sd2d_rk_stage(sd2d_ctx *ctx, int32_t stage);

/* Diagnostics of the latest stage / step. */
int sd2d_get_diag(sd2d_ctx *ctx, sd2d_diag *diag);

/*
 * AV sensor output of the latest residual evaluation:
 * SM_IND(2,NCELL) and muAV(2,NCELL) (AV.f90:83,129-140) and epsilon0
 * (AV.f90:126-127).  Any pointer may be NULL.  Requires params.artiv == 1.
 */
int sd2d_get_av(sd2d_ctx *ctx, double *sm_ind, double *muav,
                double *epsilon0);

/*
 * Nodal values of the Tecplot output (tecplotter.f90:17-60) from the device
 * state: plotq(9,nnode) = rho, u, v, w, p, Bx, By, Bz, psi at every node of
 * a plot mesh refined NRE x NRE per cell, and plotav(2,nnode) = muAV,
 * etaAV, each the average over the cells that hold the node, in the
 * reference's arithmetic (bitwise its values for the same state).  The plot
 * mesh comes from the host as CSR: the occurrences of node i are
 * occ[xadj[i] .. xadj[i+1]-1], ascending in cell, each coded
 * cell*(nre+1)^2 + ii + (nre+1)*jj (0-based mesh cell and node position
 * (ii/nre, jj/nre) in the cell); muav(2,NCELL) is the per-cell muAV of
 * sd2d_get_av (plotav may be NULL, then muav too).  The mesh numbering of
 * tecplotter2dsetup.f90 and the #!TDV111 file are the host's part
 * (sddc_amd/tecplot.py).
 */
int sd2d_plot_nodes(sd2d_ctx *ctx, int32_t nre, int32_t nnode, const int32_t *xadj,
                    const int32_t *occ, const double *muav, double *plotq, double *plotav);

/*
 * Restart files in the reference's format (WRITE_ALL_DATA_BINARY_MPI,
 * write_data.f90:44-74; READ_ALL_DATA_BINARY, read_data_restart.f90:35-67):
 * int32 iter, float64 ctime, then one record per cell: int32 global cell
 * number, float64 Q(9,N,N) of that cell; native byte order.
 *
 * sd2d_write_restart writes the device state with the context's iter and
 * ctime.  icg[ic] is the global number of local cell ic (1-based, as
 * IC2ICG); NULL means 1..NCELL.
 *
 * sd2d_read_restart loads such a file: every record's cell goes to local
 * cell ICG-1 (one rank holds the whole mesh, so NCELLGLOB == NCELL), iter
 * and ctime are taken from the header, and the running maxima are reset to
 * 0 as INIT_SETUP does after a restart (initsetup.f90:154-167).  A file of
 * the wrong size or a cell number outside 1..NCELL is SD2D_EINVAL.
 */
int sd2d_write_restart(sd2d_ctx *ctx, const char *path, const int32_t *icg);
int sd2d_read_restart(sd2d_ctx *ctx, const char *path);

/*
 * Checkpoints that resume bit for bit, no counterpart in the reference.
 * The reference's restart file holds only iter, ctime and Q, and a
 * restarted run begins with the running maxima at 0 (initsetup.f90:154-167),
 * so its first c_h differs from the continuous run's.  sd2d_write_checkpoint
 * writes the restart file exactly as sd2d_write_restart and, beside it, the
 * sidecar <path>.max: 8 bytes "SD2DMAX1", int32 iter, float64 ctime (those
 * of the restart file), float64 eigvmax, umax, eigvmax_g, umax_g (native
 * byte order).  sd2d_read_checkpoint reads both and restores the maxima, so
 * stepping on gives bitwise the continuous run; a missing sidecar or one of
 * another step is SD2D_EINVAL.  The restart file alone still reads with
 * sd2d_read_restart (and by the reference program).
 */
int sd2d_write_checkpoint(sd2d_ctx *ctx, const char *path, const int32_t *icg);
int sd2d_read_checkpoint(sd2d_ctx *ctx, const char *path);

/*
 * Measurement aid, no counterpart in the reference.  While profiling is on,
 * every stage-kernel launch of sd2d_step / sd2d_compresid is bracketed by
 * HIP events on the context's stream; sd2d_get_profile returns the summed
 * stage-kernel time in milliseconds and the number of launches since
 * profiling was last switched on.
 */
int sd2d_set_profiling(sd2d_ctx *ctx, int32_t on);
int sd2d_get_profile(sd2d_ctx *ctx, double *stage_ms, int32_t *nlaunch);

/*
 * Opt-in finiteness check, no counterpart in the reference (whose NaNs
 * propagate silently; the running maxima never adopt one, getrusanov.f90:
 * 116-123).  With every > 0, sd2d_step scans the whole state for a NaN or
 * an infinity after each step whose iter is a multiple of `every` and after
 * the last step of the call, and sd2d_rk_stage after stage 5 of such a
 * step; a hit returns SD2D_ENONFINITE and stops the call there, with the
 * step kept (iter, ctime and the norms include it) so the state can be
 * read or written.  every = 0 (the default) switches the check off.  A
 * scan reads the state once (about 0.12 ms for the 512 x 1024, N=5 mesh).
 */
int sd2d_set_check_finite(sd2d_ctx *ctx, int32_t every);

/*
 * Ensembles: many independent runs of the path stepped together, each with
 * its own mesh, order N, free stream, time step, state, running maxima and
 * time.  No counterpart in the reference, which runs one case per program;
 * each member follows the per-step contract of sd2d_step (RK_S5P3 only,
 * inviscid path, one precision and ARTIV/vismode for all members).
 * Members of one order share the launches of every RK stage; one whole
 * step of all members is a HIP graph replayed per sd2d_ens_step.  The
 * mesh and params arrays passed to sd2d_ens_create are read again by
 * sd2d_ens_tune and must stay valid until the ensemble is destroyed.
 *
 *   sd2d_ens_create       nmem members: meshes[i], params[i] (tile_shape of
 *                         the first member of each order is that order's
 *                         starting shape)
 *   sd2d_ens_set_state / sd2d_ens_get_state / sd2d_ens_set_time /
 *   sd2d_ens_get_time     member i, as sd2d_set_state / sd2d_get_state /
 *                         sd2d_set_time / sd2d_get_diag's ctime and iter
 *   sd2d_ens_step         nsteps RK_S5P3 steps of every member
 *   sd2d_ens_tune         autotuner: for each order, steps every tile shape
 *                         (SD2D_TILE_ROWPAIRS, SD2D_TILE_STRIPS,
 *                         SD2D_TILE_BLOCKS) trial_steps
 *                         times from the current states, keeps the fastest
 *                         and restores the states; shapes[N-3] returns the
 *                         choice for each order present (others untouched)
 */
typedef struct sd2d_ens sd2d_ens;
int sd2d_ens_create(int32_t nmem, const sd2d_mesh *meshes, const sd2d_params *params,
                    sd2d_ens **out);
void sd2d_ens_destroy(sd2d_ens *ens);
const char *sd2d_ens_last_error(const sd2d_ens *ens);
int sd2d_ens_set_state(sd2d_ens *ens, int32_t member, const double *q, const sd2d_maxima *maxima);
int sd2d_ens_get_state(sd2d_ens *ens, int32_t member, double *q, sd2d_maxima *maxima);
int sd2d_ens_set_time(sd2d_ens *ens, int32_t member, double ctime, int32_t iter);
int sd2d_ens_get_time(sd2d_ens *ens, int32_t member, double *ctime, int32_t *iter);
/* A member's checkpoint (restart file of its state, iter and ctime, plus
 * the running-maxima sidecar), with the format and contract of
 * sd2d_write_checkpoint / sd2d_read_checkpoint. */
int sd2d_ens_write_checkpoint(sd2d_ens *ens, int32_t member, const char *path, const int32_t *icg);
int sd2d_ens_read_checkpoint(sd2d_ens *ens, int32_t member, const char *path);
int sd2d_ens_step(sd2d_ens *ens, int32_t nsteps);
int sd2d_ens_tune(sd2d_ens *ens, int32_t trial_steps, int32_t *shapes);
/* The finiteness check of sd2d_set_check_finite for every member: with
 * every > 0, sd2d_ens_step scans every member's state after each
 * `every`-th step of the call and after its last step, and returns
 * SD2D_ENONFINITE when some member holds a NaN or an infinity (the step is
 * kept for all members).  sd2d_ens_nonfinite then gives flags[i] = 1 for
 * those members and 0 for the others (nmem entries; all 0 before any
 * check). */
int sd2d_ens_set_check_finite(sd2d_ens *ens, int32_t every);
int sd2d_ens_nonfinite(sd2d_ens *ens, int32_t *flags);

/*
 * Adjoint (vector-Jacobian products), no counterpart in the reference: the
 * FP64 inviscid path with RK_S5P3, one run (not ensembles).
 *
 * sd2d_fwd      nsteps RK_S5P3 steps like sd2d_step (norms off), keeping a
 *               checkpoint (state and running maxima) every ckpt steps and
 *               returning the objective J = (1/nsteps) sum over the steps'
 *               outputs of the wall drag D(Q) = sum over WALL face points
 *               of w_i p (S_out)_x: gas pressure of the cell's trace, the
 *               outward metric vector of the face point, and w_i the
 *               interpolatory quadrature weights of the solution-point
 *               positions Xs along the face.
 * sd2d_vjp      after sd2d_fwd: for the cotangents qbar_final (of the final
 *               state, Fortran layout, or NULL) and jbar (of J) returns
 *               qbar_init (of the initial state, Fortran layout, or NULL)
 *               and pbar[6] (of rinf, pinf, uinf, vinf, bxinf, byinf, which
 *               the boundary ghosts read, or NULL).  The lagged GLM speed
 *               c_h of every evaluation is differentiated through the
 *               running maxima it comes from, back to the flux point that
 *               attained each maximum; chbar (5*nsteps entries, or NULL)
 *               returns the cotangent of each evaluation's c_h.  The
 *               context returns to the final state of sd2d_fwd.
 * sd2d_compresid_vjp  for the residual of the current state with the
 *               current c_h (sd2d_compresid's contract): qbar = (dR/dQ)^T
 *               rbar, *chbar = (dR/dc_h)^T rbar and pbar[6] (Fortran
 *               layouts; outputs may be NULL).
 * sd2d_adj_save / sd2d_adj_load  the checkpoints of the last sd2d_fwd as
 *               files, so that the backward pass can run in another
 *               context or process, or after a crash (no counterpart in
 *               the reference, whose restart files write_data.f90:44-74
 *               they follow): <prefix>.NNNN.dat for
 *               checkpoint NNNN = 0..ceil(nsteps/ckpt), each a reference
 *               restart file with the maxima sidecar of
 *               sd2d_write_checkpoint, and <prefix>.adj (8 bytes "SD2DADJ1",
 *               int32 N, NCELL, nsteps, ckpt, final iter, float64 J, final
 *               ctime).  sd2d_adj_load leaves the context as sd2d_fwd does
 *               (the final state, maxima, ctime and iter), ready for
 *               sd2d_vjp, whose results are then bitwise those of the
 *               context that ran the forward; a mesh or order other than
 *               the files' is SD2D_EINVAL.
 */
int sd2d_fwd(sd2d_ctx *ctx, int32_t nsteps, int32_t ckpt, double *objective);
int sd2d_vjp(sd2d_ctx *ctx, const double *qbar_final, double jbar, double *qbar_init,
             double *pbar, double *chbar);
int sd2d_compresid_vjp(sd2d_ctx *ctx, const double *rbar, double *qbar, double *chbar,
                       double *pbar);
int sd2d_adj_save(sd2d_ctx *ctx, const char *prefix);
int sd2d_adj_load(sd2d_ctx *ctx, const char *prefix);

#ifdef __cplusplus
}
#endif

#endif /* SD2D_H */
