/*
 * csrst_hip.h -- C ABI of the MI355X (gfx950) implementation of CSRST's
 * particle-in-cell time step.
 *
 * Plain pointers and sizes only; every call returns 0 on success or a
 * negative CSRST_E* code, and csrst_last_error() returns the message.  No
 * exception crosses the ABI.  A context is used from one host thread; its
 * GPU work is ordered on the context's stream.
 *
 * Two steppers sit behind this ABI:
 *  - csrst_ctx: the reference's two-grid step (global masked chamber grid +
 *    bunch-following window), FP64, bit-identical to the reference
 *    arithmetic.  It replaces TETMSolver::updateField_firstStep/secondStep
 *    (TETMSolver.hh:110-266) with PartBunch::push_particles
 *    (PartBunch.cpp:215-519), the ShapeFunction<IntCIC|IntS1|IntS2> deposits
 *    and the window bookkeeping of main.cpp:390-506, 562-652.
 *  - csrst_mw: the production moving-window step (one W x H grid that
 *    follows the bunch and carries the conformal chamber walls), FP64 or
 *    FP32 with split-precision particle positions.
 */
#ifndef CSRST_HIP_H
#define CSRST_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CSRST_OK 0
#define CSRST_EINVAL (-1)
#define CSRST_ENOMEM (-2)
#define CSRST_EHIP (-3)
#define CSRST_EWINDOW (-4)   /* bunch left the window / window move outside {0,1}x{-1,0,1} */
#define CSRST_EGEOM (-5)     /* geometry set-up failed (crossBoundary out of range) */
#define CSRST_ESTATE (-6)    /* call order violated (e.g. step before init) */

typedef struct csrst_ctx csrst_ctx;
typedef struct csrst_mw csrst_mw;

/* Commands input of main.cpp:111-136 that shapes the step. */
typedef struct {
    int Nx, Ny;              /* -x/--numgridpx, -y/--numgridpy: global grid */
    double lenBefore;        /* --lengthbefore [m] */
    double lenAfter;         /* -l/--length [m] */
    double width;            /* -w/--width [m] */
    double phiDeg;           /* -P/--phi [deg] */
    double ekinMeV;          /* -E/--Ekin [MeV] */
    double bz;               /* -B/--Bz [T] */
    int order;               /* shape function: 0 IntCIC (defs.hh:55), 1 IntS1, 2 IntS2 */
    int s2bug;               /* 1: IntS2 gather reads (I+1, J+2) as at ScatterGatherImpl.hh:467,525 */
    int statModulus;         /* --statmodulus: wake sampled every statModulus steps */
} csrst_params;

const char *csrst_version(void);
const char *csrst_last_error(const void *ctx_or_null);
int csrst_device_count(int *n);

/* ---- two-grid reference step ------------------------------------------ */

/* Allocates the global fields (AllocateFields.cpp:57-120), builds the Bend
 * geometry and the conformal boundary cells (Bend.cpp:85-353,
 * SCBoundaryCondition.hh:96-137).  main.cpp:168-197. */
int csrst_create(const csrst_params *p, csrst_ctx **out);
int csrst_destroy(csrst_ctx *c);

/* Mesh spacing, origin and time step: out[0..1] hr, out[2..3] origin,
 * out[4] dt = hr0/c (main.cpp:187), out[5] R, out[6] phi. */
int csrst_geometry(csrst_ctx *c, double *out);
/* Inside mask (Nx*Ny bytes) and the boundary-cell list: ij (2 ints per cell),
 * data (lambda0, lambda1, area, lambdaN0, lambdaN1 per cell). */
int csrst_boundary_cells(csrst_ctx *c, uint8_t *mask, int *ij, double *data, int cap, int *count);

/* Bunch set-up of the -I start: window = bunch box +- 5 cells
 * (main.cpp:199-217), PartBunch sort (PartBunch.cpp:1817-1844),
 * makeReady(dt/2) (PartBunch.cpp:164-177), global->window copy and
 * boundary exchange (main.cpp:335-347).  R, P: 2 doubles per particle. */
int csrst_init_bunch(csrst_ctx *c, const double *R, const double *P, const double *Q, int n);
/* Restart (main.cpp:147-153,226-228; LoadH5Part.cpp:161-321): as
 * csrst_init_bunch with oldR given instead of makeReady, the global E and H
 * as last uploaded with csrst_upload_field (restoreFields) copied into the
 * window, global J zeroed (main.cpp:388), and the step counter, which phases
 * the wake samples (iteration % statModulus), set to step. */
int csrst_restore_bunch(csrst_ctx *c, const double *R, const double *oldR, const double *P, const double *Q, int n,
                        int step);

/* Caller geometry in place of the one csrst_create computed from the Bend,
 * before csrst_init_bunch: the inside mask (Nx*Ny bytes, 0 or 1, the cells
 * the masked sweeps update, TETMSolver.hh:142-249) and the conformal cell
 * list of SCBoundaryCondition::getBoundaryCells (SCBoundaryCondition.hh:39-84,
 * BoundaryCell.hh:28-33): ij 2 ints per cell, data (lambda0, lambda1, area,
 * lambdaN0, lambdaN1) per cell, areas after limitArea.  The layout is that of
 * csrst_boundary_cells.  The Poisson start uses the same cells. */
int csrst_set_mask(csrst_ctx *c, const uint8_t *mask);
int csrst_set_boundary_cells(csrst_ctx *c, const int *ij, const double *data, int count);
/* The context's GPU work goes to `stream` (a hipStream_t; NULL: the
 * context's own stream) from here on; the call first waits for the work
 * already queued.  The caller keeps the stream alive while it is set. */
int csrst_set_stream(csrst_ctx *c, void *stream);

/* Whole steps (main.cpp:390-506). */
int csrst_step(csrst_ctx *c, int nsteps);
/* Fine-grained parts of one step, for bisection against the CPU oracle. */
int csrst_first_step(csrst_ctx *c);            /* TETMSolver.hh:110-211 */
int csrst_second_step(csrst_ctx *c);           /* TETMSolver.hh:213-266 */
int csrst_move_window(csrst_ctx *c, int *moveI, int *moveJ);   /* main.cpp:494-501 */
/* copyFields(EFD, HFD, JFD, miniMesh E/H/J, -ddim.min, true, copyCurrent)
 * (main.cpp:534-548, Communicator.hh:212-268): the window's E, H (and J when
 * copyCurrent != 0) overwrite the global fields under the window, clipped to
 * the global domain.  The reference calls it before every field plot
 * (main.cpp:417-419) and line dump (main.cpp:472-475). */
int csrst_copy_fields(csrst_ctx *c, int copyCurrent);
/* One part of a step, in the order of TETMSolver.hh:110-266 and main.cpp:494-501:
 * 0 push + deposit Jx, 1 push + deposit Jy (PartBunch::push_particles),
 * 2 H half step, 3 Ex, 4 Ey (global and window sweeps), 5/6/7 exchange H/Ex/Ey
 * (exchangeBoundaries), 8/9/10 conformal Hz/Ex/Ey (SCBoundaryCondition),
 * 11 drift -dt/2, 12 wake sample, 13 J = 0, 14 moveMesh. */
int csrst_substep(csrst_ctx *c, int op);

/* which: 0 global E, 1 global H, 2 window E, 3 window H, 4 window J, 5 global J.
 * Planar layout [2][m+2][n+2] over [-1, n] x [-1, m] (guard layer included).
 * *count receives the number of doubles. */
int csrst_download_field(csrst_ctx *c, int which, double *out, int cap, int *count);
int csrst_upload_field(csrst_ctx *c, int which, const double *in, int count);
int csrst_download_particles(csrst_ctx *c, double *R, double *P, double *oldR, int n);
/* the charges (PartBunch::Q) in the order of csrst_download_particles */
int csrst_download_charges(csrst_ctx *c, double *Q, int n);
/* info: dt, window origin x, y, window position x, y, W, H, steps, wake samples */
int csrst_info(csrst_ctx *c, double *info);
/* CSR wake accumulator of IntegrateElongCommand (IntegrateElong.cpp:38-89). */
int csrst_wake(csrst_ctx *c, double *out, int cap, int *count);
/* IntegrateElongCommand::reset (IntegrateElong.hh:59-62): the accumulator to zero,
 * as main.cpp:445-463 does after writing it at the bend's entry and exit. */
int csrst_wake_reset(csrst_ctx *c);
/* Bend::isInsideBend (Bend.hh:307-323) at R[0..1]: *inside = 1 or 0. */
int csrst_is_inside_bend(csrst_ctx *c, const double *R, int *inside);
/* rmean(2), pmean(2), qtotal, field energy of the global grid */
int csrst_diag(csrst_ctx *c, double *out);
/* Beam parameters of PartBunch::calcBeamParameters / calcMoments
 * (PartBunch.cpp:584-756) as DataSinkCommand prints them (DataSink.cpp:91-148):
 * out[CSRST_BEAM_STATS] = Ekin (MeV), rrms x y, prms x y, normalised emittance
 * x y, rmean x y, pmean x y, Dy, DDy, dE (MeV), rmin x y, rmax x y, rprms x y.
 * Moments of the positions and momenta centred on the centroid and rotated
 * into the frame of the mean momentum; double-double sums. */
#define CSRST_BEAM_STATS 20
int csrst_beam_stats(csrst_ctx *c, double *out);
/* SaveEMFieldsCommand::execute (SaveEMFields.cpp:34-160): the fields on the
 * line through the bunch centroid along its mean momentum, at the global
 * grid's columns from 1.5 min - 0.5 max to 2 max - min of the bunch's cell
 * range, gathered from the global E, H, J with the run's shape
 * (ScatterGatherImpl semantics; the reference's IPPL gather is not
 * available to pin them).  Rows of 7 doubles: s (signed distance from the
 * centroid), E along P, E across P, Hzx + Hzy, J along P, x, y -- the
 * columns of Data/FieldOnLine_NNNN.dat.  The reference copies the window
 * into the global fields first (csrst_copy_fields(c, 0), main.cpp:468-476).
 * out == NULL: *count receives the row count. */
int csrst_line_fields(csrst_ctx *c, double *out, int cap, int *count);
/* Gauss's law on the window: per window vertex (i, j), row-major [H][W],
 * (Ex(i,j) - Ex(i-1,j))/hx + (Ey(i,j) - Ey(i,j-1))/hy - rho(i,j)/eps0, with
 * rho the bunch's getChargeDensity of the run's order at R (the divergence is
 * formed as PartBunch::calcChargeConservation forms div J, PartBunch.cpp:540-582).
 * Its change from step to step is the charge-conservation drift of the field
 * solve.  out == NULL: *count receives W*H. */
int csrst_gauss_residual(csrst_ctx *c, double *out, int cap, int *count);
/* The field sums of CalcEnergyCommand::execute (CalcEnergy.cpp:60-124) over the
 * global grid plus the window minus the global nodes under the window, times
 * the cell area: out[0] = sum E.E, out[1] = sum (Hzx + Hzy)^2.  The command's
 * energy, 1/4 eps0 (out[0] + previous out[0]) + 1/2 mu0 out[1], keeps the
 * previous call's out[0], which the caller holds. */
int csrst_field_energy(csrst_ctx *c, double *out);

/* ---- moving-window production step -------------------------------------- */

typedef struct {
    int W, H;                /* window cells */
    double dx, dy;           /* mesh spacing [m] */
    double lenBefore, lenAfter, width, phiDeg, ekinMeV, bz;   /* chamber and bend */
    int order;               /* 0, 1, 2 */
    int precision;           /* 0: FP64, 1: FP32 fields + int32 cell / FP32 offset positions */
    int sortEvery;           /* cell sort period in steps (0: never) */
    int statModulus;         /* wake sampling period */
    double cfl;              /* c*dt/dx; 1.0 is the reference's dt = hx/c */
    int pml;                 /* absorbing-layer columns at each window end (0: none, as the
                                reference, whose PML coefficients are zeroed, TETMSolver.cpp:59-62) */
    int wakeGrid;            /* wake samples: 0 at fixed offsets (k - ns/2) dx from the centroid
                                along the mean momentum; 1 at the window's grid columns k + 2, the
                                points IntegrateElong.cpp:59-61 takes */
} csrst_mw_params;

int csrst_mw_create(const csrst_mw_params *p, csrst_mw **out);
int csrst_mw_destroy(csrst_mw *m);
/* Positions R in metres (global chamber frame), P = beta*gamma, Q [C]. */
int csrst_mw_init_bunch(csrst_mw *m, const double *R, const double *P, const double *Q, int n);
int csrst_mw_step(csrst_mw *m, int nsteps);
/* csrst_set_stream for the moving-window step. */
int csrst_mw_set_stream(csrst_mw *m, void *stream);
/* One push of the step on the fields as they stand (PartBunch::push_particles,
 * PartBunch.cpp:215-519, with its deposit of J[comp], comp 0 = Jx, 1 = Jy,
 * added into the window's J), followed by the drift by -dt/2 that the step
 * makes after it (TETMSolver.hh:210, 265).  No field update, no window move. */
int csrst_mw_push(csrst_mw *m, int comp);
/* Field-only start: window logical cell (0, 0) at global cell (wx0, wy0),
 * geometry computed, fields zero, no particles (steps then run the field
 * sweeps alone and the window stays in place). */
int csrst_mw_place(csrst_mw *m, int wx0, int wy0);
/* planar [6][H+2][W+2] doubles as csrst_mw_download_fields; the two Hz
 * planes are summed (the step keeps Hz = Hzx + Hzy). */
int csrst_mw_upload_fields(csrst_mw *m, const double *in, int count);
/* planar [6][H+2][W+2] doubles: Ex, Ey, Hzx, Hzy, Jx, Jy (guards incl.) */
int csrst_mw_download_fields(csrst_mw *m, double *out, int cap, int *count);
int csrst_mw_download_particles(csrst_mw *m, double *R, double *P, int n);
/* csrst_beam_stats for the moving-window bunch (PartBunch.cpp:584-756) */
int csrst_mw_beam_stats(csrst_mw *m, double *out);
/* info: dt, origin x, y, column offset, row offset, W, H, steps, wake samples, n */
int csrst_mw_info(csrst_mw *m, double *info);
int csrst_mw_wake(csrst_mw *m, double *out, int cap, int *count);
/* Average device time of each kernel family over the last csrst_mw_step call:
 * [0] push+deposit, [1] field sweeps, [2] window move, [3] sort, [4] wake (ms). */
int csrst_mw_timing(csrst_mw *m, double *ms, int cap);

/* Poisson start of the moving window: after csrst_mw_init_bunch, before the
 * first step, the three solves of PoissonSolver::computeField
 * (PoissonSolver.cpp:85-201) on the window grid, fillFields and
 * resetOutsideDomain.  The window edges are Dirichlet (the reference's grid
 * is the whole chamber; its boosted field decays over width/gamma, far inside
 * any window that holds the bunch).  Multigrid in FP64; the fields are stored
 * in the window's precision. */
int csrst_mw_poisson(csrst_mw *m, double tol, int maxit, int *cycles, double *resid);

/* ---- ShapeFunction and ScatterGatherImpl on caller memory --------------------
 * Stateless FP64 kernels with the reference's arithmetic, bit-identical to a
 * CPU loop over the particles in their order (sums over particles are
 * ordered by a stable sort of the contributions).  Every array is a device
 * pointer; positions R and oldR are (x, y) pairs per particle.  Fields are
 * node patches [ny][nx], x fastest.  `stream` is a hipStream_t (NULL: the
 * null stream); each call returns once its work has finished.  A stencil
 * that leaves the patch returns CSRST_EWINDOW (its values dropped or read as
 * zero); the other nodes are still written.
 *
 * ScatterGatherImpl<IntOp,double>::gather / ::scatter
 * (ScatterGatherImpl.hh:27-64; IntCIC 70-194, IntS1 202-392, IntS2 407-641):
 * order 0 IntCIC, 1 IntS1, 2 IntS2; kind 0 the edge field Ex (at
 * (i+1/2, j)), 1 the edge field Ey (at (i, j+1/2)), 2 a cell field (at
 * (i+1/2, j+1/2)); node (0, 0) of the patch sits at `origin`.  s2bug: the
 * IntS2 gather's (I+1, J+2) read of ScatterGatherImpl.hh:467,525.  Scatter
 * adds val[k] times the weights into the field. */
int csrst_sg_gather(int order, int s2bug, int kind, const double *field, int nx, int ny, const double origin[2],
                    const double dx[2], const double *R, int n, double *out, void *stream);
int csrst_sg_scatter(int order, int kind, double *field, int nx, int ny, const double origin[2], const double dx[2],
                     const double *R, const double *val, int n, void *stream);
/* ShapeFunction<IntOp>::getCurrentDensity{X,Y} with its Base part
 * (ShapeFunction.hh:30-508; ZerothOrderShapeFunction.cpp:202-294,
 * FirstOrderShapeFunction.cpp:226-329,540-683,
 * SecondOrderShapeFunction.cpp:222-345,583-826): the charge-conserving
 * current of the trajectories oldR -> R over dt, component comp (0 Jx, 1 Jy),
 * onto the node patch [i0, i0+nx) x [j0, j0+ny) of the mesh whose node
 * (0, 0) is at `origin`; the patch is cleared first. */
int csrst_sf_current_density(int order, int comp, double *J, int i0, int j0, int nx, int ny, const double origin[2],
                             const double dx[2], double dt, const double *oldR, const double *R, const double *Q,
                             int n, void *stream);
/* csrst_sf_current_density without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (sort keys and values, the sort's temporary storage, an error word)
 * is a static 32 MiB block of device memory per GPU, held until the process
 * ends; inputs needing more (about 27k particles of S2, 55k of S1, 160k of
 * S0) return CSRST_EINVAL.  Calls that overlap on one GPU share that block
 * and race: issue them one after the other.  Results equal
 * csrst_sf_current_density's bit for bit. */
int csrst_sf_current_density_noalloc(int order, int comp, double *J, int i0, int j0, int nx, int ny,
                                     const double origin[2], const double dx[2], double dt, const double *oldR,
                                     const double *R, const double *Q, int n, void *stream);
/* ShapeFunction<IntOp>::getChargeDensity and getChargeDensityDiff
 * (ZerothOrderShapeFunction.cpp:296-459, FirstOrderShapeFunction.cpp:331-538,
 * SecondOrderShapeFunction.cpp:347-581): vertex charge density over hx hy on
 * the node patch (cleared first); Diff: rho(R) - rho(oldR), each particle's
 * two stencils summed in that order. */
int csrst_sf_charge_density(int order, double *rho, int i0, int j0, int nx, int ny, const double origin[2],
                            const double dx[2], const double *R, const double *Q, int n, void *stream);
int csrst_sf_charge_density_diff(int order, double *rho, int i0, int j0, int nx, int ny, const double origin[2],
                                 const double dx[2], const double *R, const double *oldR, const double *Q, int n,
                                 void *stream);

/* ---- Poisson initialisation ------------------------------------------------
 * Replaces PoissonSolver::computeField (PoissonSolver.cpp:85-201), whose
 * Belos block-CG + ML AMG solves the cutoffStencil system
 * (PoissonSolver.cpp:678-766); here FP64 geometric multigrid on the GPU.
 *
 * csrst_poisson_solve: one solve on host arrays over an nx x ny node grid
 * (row-major, x fastest).  cls per node: 0 = identity row (phi = b, not
 * coupled), 1 = unknown (5-point row), 2 = identity row coupled to unknown
 * neighbours (its b acts as a boundary value).  hx is the x spacing already
 * scaled by gamma (PoissonSolver.cpp:71).  phi: in = initial guess, out =
 * solution.  Stops when ||b - A phi|| <= tol ||b|| over unknown rows or after
 * maxit V-cycles; reports the cycles used and the final relative residual. */
/* csrst_poisson_compute_field: after csrst_init_bunch and before the first
 * step, the three solves of PoissonSolver::computeField on the global grid
 * (Ey from the charge at t + dt/2, Hz with the bunch half a cell back, Ex),
 * fillFields, resetOutsideDomain, then the window copy and exchanges as in
 * main.cpp:229-341.  cycles[3], resid[3]: V-cycles and relative residual of
 * each solve (either may be NULL).  The reference runs solve 1 to 1e-12 and
 * solves 2 and 3 to 1e-6 from a random start; here every solve runs to tol. */
int csrst_poisson_compute_field(csrst_ctx *c, double tol, int maxit, int *cycles, double *resid);
int csrst_poisson_solve(int nx, int ny, double hx, double hy, const uint8_t *cls, const double *b,
                        double *phi, double tol, int maxit, int *cycles, double *resid);

/* ---- learned warm start for the Poisson multigrid ---------------------------
 * No counterpart in the reference (its Poisson start iterates from a random
 * guess, PoissonSolver.cpp:158).  A 4-level convolutional U-Net, 32 base
 * channels, BF16 weights and activations, FP32 accumulation on MFMA, mapping
 * a charge density on an H x W node grid to a potential that seeds
 * csrst_poisson_solve; the multigrid still iterates to its tolerance.
 * Arrays are [B][H][W] floats (host); `scale` multiplies the input density
 * (the network is trained on densities of order one and the Poisson problem
 * is linear, so callers scale phi back by 1/scale). */
typedef struct csrst_warmstart csrst_warmstart;
int csrst_warmstart_create(int H, int W, int maxBatch, csrst_warmstart **out);
int csrst_warmstart_destroy(csrst_warmstart *u);
int csrst_warmstart_num_params(csrst_warmstart *u, int *n);
int csrst_warmstart_get_weights(csrst_warmstart *u, float *w, int n);
int csrst_warmstart_set_weights(csrst_warmstart *u, const float *w, int n);
int csrst_warmstart_infer(csrst_warmstart *u, const float *rho, float *out, int B, float scale);
/* mean-square loss against phi and the FP32 gradient of every parameter */
int csrst_warmstart_gradients(csrst_warmstart *u, const float *rho, const float *phi, int B, float scale,
                              float *loss, float *grads, int n);
/* one Adam step (beta 0.9 / 0.999); *loss is the batch loss before the update */
int csrst_warmstart_train_step(csrst_warmstart *u, const float *rho, const float *phi, int B, float scale,
                               float lr, float *loss);
/* The network between V-cycles (no counterpart in the reference either).  The
 * multigrid of csrst_poisson_solve on the network's grid (nx = W, ny = H),
 * kept on the device while the caller passes the same hx, hy and cls.  At an
 * iterate phi_k the residual r_k, scaled to unit RMS over the unknowns, goes
 * through the network; its output c is a direction for the error and the
 * iterate moves by alpha c with alpha = (c . r_k) / (c . A c), the step that
 * minimises the energy norm of the error along c; a V-cycle follows.
 * mode 0: plain V-cycles from phi; 1: the network's step before every V-cycle;
 * 2: before the first V-cycle only; 3: before every V-cycle the energy-norm
 * minimiser over span{c, phi_k - phi_k-1} (the network's direction and the
 * previous iteration's change); 4: over the previous change alone.
 * b, phi: [H][W] doubles (host); phi in: the guess, out: the solution;
 * cycles and resid as csrst_poisson_solve. */
int csrst_warmstart_solve(csrst_warmstart *u, double hx, double hy, const uint8_t *cls, const double *b,
                          double *phi, double tol, int maxit, int mode, int *cycles, double *resid);
/* Training on the solver's own iterates: for each of the nrhs right-hand sides
 * b[nrhs][H][W], phi* from a solve to 1e-13, then pairs (r_k / rms(r_k),
 * (phi* - phi_k) / rms(phi* - phi_k)) at iterates k < kmax of the loop (with
 * the network's step when inloop is 1), one Adam step of rate lr per maxBatch
 * pairs.  *loss: mean loss of those steps; *pairs: pairs used. */
int csrst_warmstart_train_iterates(csrst_warmstart *u, double hx, double hy, const uint8_t *cls, const double *b,
                                   int nrhs, int kmax, int inloop, float lr, float *loss, int *pairs);

#ifdef __cplusplus
}
#endif
#endif
