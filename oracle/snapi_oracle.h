/*
 * FP64 CPU restatement of SNAPI's hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the oracle the HIP path is checked against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or the timed CPU baseline; the product never calls it.
 *
 * Each function cites the reference file:line whose semantics it restates.
 * How each one is pinned (golden vectors from the reference build, closed
 * forms, invariants) is listed in its comment; functions with no such check
 * say "parity unpinned".
 */
#ifndef SNAPI_ORACLE_H
#define SNAPI_ORACLE_H

#include "../include/snapi_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Humlicek W4 Voigt/Faraday-Voigt, cmlib/profiles.cc:7-37 (values only).
 * Pinned: golden op/em of the reference build (every line evaluation goes
 * through it), symmetry H(-v)=H(v), F(-v)=-F(v), and scipy's wofz to the
 * published 1e-4 accuracy of W4. */
void orc_voigt(double v, double a, double *H, double *F);

/* Zeeman pattern of one transition, atom::zeeman_setup atom.cc:1620-1773.
 * j2u, j2l: 2J of the upper/lower level.  Returns split flag; fills nm[3],
 * shift[3][nm], str[3][nm] (rows: pi, sigma_b, sigma_r).
 * Pinned: the reference's dumped pattern (tests/golden/cfg0_synth.npz 'zc'),
 * strengths summing to 1 per group, and g_eff of the pattern. */
int orc_zeeman(int j2u, int j2l, double gu, double gl, int nm[3], double shift[3][32], double str[3][32]);

/* Planck function B_lambda, cmlib/mathtools.cc:1150-1152. */
double orc_planck(double lam, double T);

/* op[nd][4][4], em[nd][4] of one column at one vacuum wavelength:
 * atmosphere::op_em_vector (atmos/atmos_rad_synthesis.cc:20-72) with the
 * background of the column state, H- (mol.cc:379-412, 516-545, 609-632), and
 * every line of the table (atom_rad_synthesis.cc:396-505), then the matrix
 * symmetrisation of atmos_rad_synthesis.cc:56-68, then (tau_grid) division by
 * op_ref (atmos/atmos_rad.cc:175-185).  state is one column [nfield][nd].
 * Pinned: golden 'opem' (reference op/em at every 20th wavelength, all depths)
 * and 'spop' (H- per wavelength). */
void orc_opem(const snapi_line_table *lt, const snapi_ray *ray, int nd, const double *state, double lam,
              int normalise, double *op, double *em);

/* Polarized formal solution atmos_ppbez::formal, plpar/atmos_ppbez_rts.cc:
 * 918-1218, with the reference's linear source-function weights (the w2 term
 * vanishes for any step below 1e10) and the last point copied.
 * op[nd][16] em[nd][4] grid[nd]; S[nd][4] out.  Returns 0, or -3 when a
 * forward step exceeds the 1e10 switch of the quadratic branch.
 * Pinned: golden 'stokes' and 'sdepth' (per-depth solution) of the reference
 * build; constant source function gives I = S; linear source function on a
 * semi-infinite grid gives Eddington-Barbier exactly. */
int orc_formal_pol(int nd, const double *grid, const double *op, const double *em, double theta, double boundary,
                   double *S);

/* Scalar formal solution with diagonal Lambda*, plpar/atmos_ppbez_rts.cc:21-219.
 * Pinned: golden 'sl' of the reference build; Lambda* equals the local
 * weight w1; I = S for constant S. */
int orc_formal_scalar(int nd, const double *grid, const double *op, const double *em, double theta, double boundary,
                      double *I, double *L);

/* obs_stokes for a batch of columns: state [ncol][nfield][nd], stokes [ncol][nlambda][4]. */
int orc_synth(const snapi_line_table *lt, const snapi_ray *ray, int ncol, int nd, const double *state, int nlambda,
              const double *lam, double *stokes, int nthreads);

/* Per-depth responses (see snapi_stokes_responses), FP64 forward mode.
 * resp [ncol][7][nd][nlambda][4]. */
int orc_responses(const snapi_line_table *lt, const snapi_ray *ray, int ncol, int nd, const double *state,
                  const double *dstate, int nlambda, const double *lam, double *stokes, double *resp, int nthreads);

/* ALI for one NLTE atom (snapi_oracle_ali.c); single column for the first
 * two, ncol columns for the solve.  Layouts as the snapi_ali_* entry points. */
int orc_ali_accumulate(const snapi_ali_atom *at, const snapi_ali_grid *g, int nd, const double *colstate,
                       const double *vproj, const double *gamma, const double *bgop, const double *bgem,
                       const double *pops, double *acc);
double orc_ali_newpops(const snapi_ali_atom *at, int nd, const double *colstate, const double *coll,
                       const double *pops, const double *acc, int alo, double *newpops);
int orc_ali_solve(const snapi_ali_atom *at, const snapi_ali_grid *g, int ncol, int nd, const double *colstate,
                  const double *vproj, const double *gamma, const double *bgop, const double *bgem,
                  const double *coll, double tol, int max_iter, double *pops, int *iterations, double *relchange,
                  int nthreads);

/* Chemical equilibrium and LTE populations (snapi_oracle_chem.c); layouts as
 * snapi_chemeq.  orc_chemeq_point returns the Newton steps taken or -1. */
void orc_chem_partition(const snapi_partition *pf, double T, double ne, double *U, double *dU);
void orc_chem_ionfrc(const snapi_chem_element *el, double T, double ne, double *fk, double *df);
double orc_chem_dissoc(const snapi_chem_molecule *m, double T);
int orc_chemeq_point(const snapi_chem_table *tb, double T, double Nt, double *ne, double *stage, double *mol,
                     double *lev, double *opref);
double orc_chem_opref(const snapi_chem_table *tb, double T, double ne, const double *stage_pop);
int orc_eos_state(const snapi_chem_table *tb, const snapi_state_map *map, int nd, double *state, double *resp,
                  double *dstate);
int orc_eos_state_batch(const snapi_chem_table *tb, const snapi_state_map *map, int ncol, int nd, double *state,
                        double *resp, double *dstate, int nthreads);
int orc_hequilibrium(const snapi_chem_table *tb, int nd, const double *T, const double *logtau, double *Nt,
                     double *ne, double *opref, double *rho, int *iters);
int orc_hequilibrium_batch(const snapi_chem_table *tb, int ncol, int nd, const double *T, const double *logtau,
                           double *Nt, double *ne, double *opref, double *rho, int *iters, int nthreads);
int orc_hequilibrium_tol(const snapi_chem_table *tb, int nd, const double *T, const double *logtau, double *Nt,
                         double tol, int max_iter);
int orc_he_tangent(const snapi_chem_table *tb, int nd, const double *T, const double *logtau, const double *Nt,
                   int ndir, const double *dT, double *dNt);
int orc_chemeq(const snapi_chem_table *tb, int ncol, int nd, const double *T, const double *Nt, double *ne,
               double *stage, double *mol, double *lev, double *opref, int *iterations, int nthreads);

/* NLTE population responses (snapi_oracle_nlteresp.c): formal solution with
 * the responses of every intensity to every opacity and emissivity of one
 * ray (Rop, Rem [nd][nd]); and the pivot-free Crout solve of an n x n system
 * for nrhs right-hand sides (B, X [nrhs][n]). */
int orc_formal_responses(int nd, const double *grid, const double *op, const double *em, double theta,
                         double boundary, double *I, double *L, double *Rop, double *Rem);
int orc_response_solve(int n, int nrhs, const double *M, const double *B, double *X);

/* NLTE population-response system (snapi_oracle_nlteasm.c): the assembly of
 * the reference's response matrix and right-hand sides for one column, with
 * the layouts of snapi_nlte_response_system; and two of its pieces, the
 * profile norms of one ray and the line profiles at one wavelength. */
int orc_nlte_response_system(const snapi_nlte_atom *atom, int nd, int nray, int nlam, int tau_grid, double boundary,
                             const double *ray_theta, const double *ray_weight, const double *lambda,
                             const double *lambda_weight, const double *sigma, const double *depth,
                             const double *pops, const double *damp, const double *coll, const double *vr,
                             const double *opem, double *M, double *beta);
int orc_nlte_profile_norm(const snapi_nlte_atom *atom, int nd, int nlam, const double *lambda,
                          const double *lambda_weight, const double *depth, const double *damp, const double *vr,
                          double *norm, double *normd);
int orc_nlte_line_profiles(const snapi_nlte_atom *atom, int nd, double lambda, const double *depth,
                           const double *damp, const double *vr, double *prof);

#ifdef __cplusplus
}
#endif
#endif
