/*
 * csrst_oracle.h -- CPU FP64 oracle for the CSRST particle-in-cell time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain C++ restatement of the
 * reference algorithm (kreischtauf/CSRST, src/) and exists to check the HIP
 * product path.  Only tests/, __graft_entry__.smoke() and the CPU-baseline leg
 * of bench.py may load it.  The product library never links, loads or calls it.
 *
 * Build: g++ -O3 -ffp-contract=off (no FMA contraction, like the reference's
 * ISO C++ build, CMakeLists.txt:85-87).
 *
 * What pins each function (tests/test_oracle.py, tests/test_oracle_pins.py;
 * tests/test_oracle_mutants.py plants one mistake per pin and requires the pin
 * to fail on it):
 *   gather / scatter weights, IntS2 bug   the reference's golden vectors
 *                                         (test/scattergather.cpp:90-416)
 *   crossings                             brute-force segment sampling
 *   S0/S1/S2 deposits                     discrete continuity identity on the
 *                                         reference's test trajectories and
 *                                         random ones; a uniform current
 *   Q_rsqrt                               error bound against 1/sqrt(x)
 *   Yee split step (bulk)                 plane waves: |lambda| = 1, fields
 *                                         follow the mode to 1e-11
 *   conformal cells, limitArea            walls through vertices reduce to
 *                                         PEC; a wave beside a fractional wall
 *   push, rotation, external field        the rotation map in the field strip;
 *                                         uncharged orbits on their circle
 *   exchange, moveMesh                    window = global grid under it
 *   Poisson operator and solve            manufactured solutions; line charge
 *                                         between plates (closed form)
 *   beam statistics                       4-particle closed form, rotation
 *                                         invariance
 *   line fields (SaveEMFields)            uniform fields, line geometry
 *   CSR wake (IntegrateElong)             parity unpinned at the IPPL IntCIC
 *                                         gather (IPPL absent)
 *   Distribution (GSL Halton)             parity unpinned (GSL absent)
 */
#ifndef CSRST_ORACLE_H
#define CSRST_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- primitives ------------------------------------------------------- */
float orc_q_rsqrt(float x);                                       /* utils.hh:491-504 */
int   orc_crossings(int dual, const double oldR[2], const double newR[2],
                    const double dx[2], const double origin[2],
                    double *frac, double *sub);                   /* utils.hh:364-489 */

/* ScatterGatherImpl on a dense patch (row-major, pnx fastest).
 * order: 0 = IntCIC, 1 = IntS1, 2 = IntS2.  s2bug selects the reference's
 * (I+1, J+2) index for the last IntS2 gather term (ScatterGatherImpl.hh:467). */
void orc_gather_edge(int order, int comp, int s2bug, const double *patch, int pnx, int pny,
                     const double origin[2], const double dx[2],
                     const double *R, int n, double *out);
void orc_gather_cell(int order, int s2bug, const double *patch, int pnx, int pny,
                     const double origin[2], const double dx[2],
                     const double *R, int n, double *out);
void orc_scatter_edge(int order, int comp, double *patch, int pnx, int pny,
                      const double origin[2], const double dx[2],
                      const double *R, const double *val, int n);
void orc_scatter_cell(int order, double *patch, int pnx, int pny,
                      const double origin[2], const double dx[2],
                      const double *R, const double *val, int n);

/* Charge-conserving current deposition onto an absolute-index patch
 * [i0, i0+pnx) x [j0, j0+pny) of a mesh with the given origin/spacing.
 * Returns -1 if a contribution falls outside the patch. */
int orc_deposit(int order, int comp, double *J, int i0, int j0, int pnx, int pny,
                const double origin[2], const double dx[2], double dt,
                const double *oldR, const double *R, const double *Q, int n);
/* Vertex charge density (divided by hx*hy); same patch convention. */
/* PartBunch::calcMoments + calcBeamParameters (PartBunch.cpp:584-756): out[20] =
 * Ekin, rrms(2), prms(2), emittance(2), rmean(2), pmean(2), Dy, DDy, dE, rmin(2), rmax(2), rprms(2) */
int orc_beam_parameters(const double *R, const double *P, int n, double *out);
int orc_charge_density(int order, double *rho, int i0, int j0, int pnx, int pny,
                       const double origin[2], const double dx[2],
                       const double *R, const double *Q, int n);

/* ---- geometry (Bend.cpp) --------------------------------------------- */
void *orc_bend_create(double lenBefore, double lenAfter, double width, double phiDeg,
                      double ekinMeV, double bz, int Nx, int Ny);
void  orc_bend_destroy(void *b);
/* out: hr0 hr1 origin0 origin1 R phi center0 center1 domain_length domain_width */
void  orc_bend_info(void *b, double *out);
int   orc_bend_is_inside(void *b, double x, double y);
void  orc_bend_mask(void *b, uint8_t *mask);                       /* Nx*Ny, row-major */
int   orc_bend_num_bcells(void *b);
/* ij: 2 ints per cell; data: lambda0 lambda1 area lN0 lN1 per cell (after limitArea) */
void  orc_bend_bcells(void *b, int *ij, double *data);
void  orc_bend_external_field(void *b, const double *R, int n, double *out);

/* ---- two-grid time step (main.cpp:390-506) -------------------------- */
typedef struct {
    int Nx, Ny;                 /* global grid (AllocateFields) */
    double lenBefore, lenAfter, width, phiDeg, ekinMeV, bz;   /* Bend */
    int order;                  /* 0 CIC, 1 S1, 2 S2 */
    int s2bug;                  /* reproduce ScatterGatherImpl.hh:467 */
    int statModulus;            /* wake sampling period (steps) */
} orc_sim_params;

void *orc_sim_create(const orc_sim_params *p);
void  orc_sim_destroy(void *s);
/* R, P interleaved (x,y) per particle; Q per particle.  Performs the -I start:
 * window sizing, PartBunch sort, makeReady(dt/2), copy global->window, exchange. */
int   orc_sim_init(void *s, const double *R, const double *P, const double *Q, int n);
/* The restart: oldR as given (no makeReady), global E/H as set with
   orc_sim_set_field copied into the window, step counter = step. */
int   orc_sim_restore(void *s, const double *R, const double *oldR, const double *P, const double *Q, int n,
                      int step);
int   orc_sim_step(void *s, int nsteps);
/* info: dt, window origin x/y, ddim first x/y, W, H, steps done, wake samples */
void  orc_sim_info(void *s, double *info);
/* which: 0 global E, 1 global H, 2 window E, 3 window H, 4 window J.
 * Copies the interleaved 2-component field over [-1, n] x [-1, m] (guards incl.)
 * Returns number of doubles written. */
int   orc_sim_get_field(void *s, int which, double *out);
void  orc_sim_get_particles(void *s, double *R, double *P, double *oldR);
int   orc_sim_get_wake(void *s, double *out);
/* compensated diagnostics: rmean(2) pmean(2) qtotal field_energy */
/* CalcEnergyCommand sums (CalcEnergy.cpp:60-124): out[2] = sum E.E, sum (Hzx+Hzy)^2, times cell area */
void  orc_sim_field_energy(void *s, double *out);
void  orc_sim_diag(void *s, double *out);
/* set a field (same layout as get) */
int   orc_sim_set_field(void *s, int which, const double *in);
/* Global-grid sweeps only (no particles, no window): per step
 * H/2, SC Hz, Ex, SC Ex, H/2, SC Hz, Ey, SC Ey with J as set. */
int   orc_sim_field_steps(void *s, int nsteps);
/* one part of the step (codes as csrst_substep of include/csrst_hip.h) */
int   orc_sim_substep(void *s, int op);

/* The particle loop of PartBunch::push_particles (PartBunch.cpp:215-383) on
 * given fields: half drift, gather of Ex, Ey and Hz = Hzx + Hzy, kick and
 * rotation with the bend's external field, half drift; then, if driftBack,
 * the drift by -dt/2 of TETMSolver.hh:210.  The three fields share one patch
 * [pnx x pny] whose node (0, 0) sits at `origin` (the staggering offsets are
 * applied by the gathers).  oldR and Rmid (either may be NULL) receive R
 * before the push and after it (before the drift back: the deposit's end). */
int   orc_push_patch(int order, int s2bug, void *bend, double dt, const double *ex, const double *ey,
                     const double *hz, int pnx, int pny, const double origin[2], const double dx[2],
                     double *R, double *P, double *oldR, double *Rmid, int n, int driftBack);

/* SaveEMFieldsCommand::execute (SaveEMFields.cpp:34-160) on the Sim's
 * global E, H, J: rows of 7 (s, E along P, E across P, Hz, J along P, x, y);
 * returns the row count, -1 if cap is too small. */
int   orc_sim_line_fields(void *s, double *out, int cap);

/* ---- Poisson initialisation (PoissonSolver.cpp:85-201, 678-766) ------ */
/* Node classes of cutoffStencil on the bend's Nx x Ny grid (row-major);
 * returns maxX = lengthStraightSection in cells. */
int   orc_poisson_classes(void *bend, uint8_t *flag, uint8_t *active);
/* Direct solve of the cutoffStencil system with spacings hx (already
 * gamma-scaled), hy; identity rows give phi = b.  0 ok, -1 not SPD. */
int   orc_poisson_solve(int nx, int ny, double hx, double hy, const uint8_t *flag,
                        const uint8_t *active, const double *b, double *phi);
/* After orc_sim_init: the three Poisson solves into the global E and H,
 * resetOutsideDomain, then the window copy and exchanges.  gamma <= 0 takes
 * it from the mean momentum. */
int   orc_sim_poisson(void *s, double gamma);

#ifdef __cplusplus
}
#endif
#endif
