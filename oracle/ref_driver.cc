// Test driver around the reference SNAPI library (oracle/_ref/libsnapi_ref.so).
//
// It plays the part of the reference worker (slave/slave.cc:117-146): build an
// atmosphere from an `atmos{}` config block (atmos/acfg.cc:16), read its SPINOR
// column (atmos/atmos_fio.cc:25), and call the reference's own entry points.
// Besides the final observables it writes the intermediate state that the
// reference holds between `nltepops()` and `formal()` so that the oracle and
// the HIP path can be fed exactly the same inputs.
//
// Modes (all output is plain text on stdout, one `key value...` record a line):
//   synth  <atmos.cfg> <lambda.txt> <el> <az> [every]
//          per-depth state, per-line atomic data, op/em at every `every`-th
//          wavelength, and the Stokes vector of atmosphere::obs_stokes
//          (atmos/atmos_obs.cc:16-87).
//   resp   <atmos.cfg> <lambda.txt> <el> <az>
//          atmosphere::obs_stokes_responses (atmos/atmos_rad_synthesis.cc:166-399):
//          Stokes plus d(IQUV)/d(q_p(z)) for the 7 physical parameters.
//   lmstep <nrow> <npar> <lm> <seed>
//          one damped normal-equation solve through the reference's cmlib
//          routines (multiply_with_transpose, solve_svd) on a seeded J and r,
//          in the order atmos/atmos_fit.cc:129-160 uses them.
//   formal_scalar <atmos.cfg> <lambda.txt> <el> <az>
//          unpolarised formal solution with approximate Lambda operator
//          (plpar/atmos_ppbez_rts.cc:21-219) on the reference's continuum.
//   nlte   <atmos.cfg> [max_iter]
//          atmosphere::nltepops (atmos/atmos_pop.cc:67-240) for the NLTE atom:
//          the inputs of one ALI iteration, its J/Lambda*/norm sums and new
//          populations, and the converged populations.
//   interp <nnode> <is_temp> <seed>
//          atmospheric_interpolation on a seeded node set.
//   chemeq <atmos.cfg>
//          the species table, then chemical equilibrium and LTE populations
//          (atmosphere::ltepops, atmos/atmos_pop.cc:30-45) on every depth.
//   nlteresp <atmos.cfg> [every]
//          NLTE population responses of the active atom: its linear system,
//          right-hand sides and solution, and formal_with_responses inputs
//          and outputs at selected wavelengths.
//   nodes  <atmos.cfg> <lambda.txt> <el> <az> <model.txt>
//          atmos_pp::build_from_nodes (plpar/atmos_pp.cc:58-248) at the node
//          values of model.txt, the ray projection of the field
//          (atmos/atmos_rad.cc:15-51), the synthesis-path Stokes vector of
//          forward_evaluate (atmos/atmos_fit.cc:717-737), and
//          obs_stokes_responses_to_nodes (atmos/atmos_fit.cc:1205-1352): the
//          response-path Stokes vector, the node Jacobian and the model's
//          node responses dq_p(z)/dnode.
//   time_synth <atmos.cfg> <lambda.txt> <el> <az> <reps>
//          wall time of obs_stokes per call on the column.
//   psf    <nl> <npsf> <seed>
//          convolve_spectra_with_psf (cmlib/mathtools.cc:1429-1465) on seeded
//          spectra and a seeded PSF of npsf samples p[0..npsf-1], handed to
//          the routine as the array it indexes: element 0 = p[0], elements
//          1..npsf-1 = p[1..npsf-1] and element npsf = p[0], so that its sum
//          over 1..npsf is the sum of the samples.
//   reg    <atmos.cfg> <model.txt>
//          atmosphere::regularize_hessian (atmos/atmos_fit.cc:1371-1480) on a
//          zero J^T J for the model's regularisation: the matrix it adds.
//   fit    <atmos.cfg> <lambda.txt> <el> <az> <fit.txt> [last]
//          atmosphere::stokes_lm_fit (atmos/atmos_fit.cc:18-227) from the start
//          model of fit.txt on its observation, run with 1, 2, ..., kmax
//          iterations: the nodes each run returns (after polish_angles) and
//          the chi^2 of the spectrum it returns, which is the fit's
//          trajectory, the algorithm being deterministic.
//
// This file is test infrastructure: it is compiled only against the reference
// sources (oracle/ref.mk) and never linked into the product.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <ctime>
#include <unistd.h>

// The driver reads the reference's protected state; it is compiled on its own
// and only ever inspects, never changes, that state.
#define protected public
#define private public
#include "types.h"
#include "io.h"
#include "atomcfg.h"
// the partition-function classes keep their parameters in the default
// (private) section of a `class`; read them as struct members
#define class struct
#include "partf.h"
#undef class
#include "mem.h"
#include "mathtools.h"
#include "uts.h"
#include "atmos.h"
#include "acfg.h"
#include "atmol/atmol.h"
#include "atmol/atom/atom.h"
#include "atmol/mol/mol.h"
#include "plpar/atmos_ppbez.h"
#include "ana_io.h"
#include "const.h"  // last: it defines one-letter macros (c, h, k, R, ...)
#undef protected
#undef private

static char *slurp(const char *fn) {
  FILE *f = fopen(fn, "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", fn); exit(2); }
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *b = new char[n + 1];
  size_t got = fread(b, 1, n, f);
  b[got] = 0;
  fclose(f);
  return b;
}

// lambda file: one air wavelength per line, in Angstrom.
static fp_t *read_lambda(const char *fn, int &n) {
  FILE *f = fopen(fn, "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", fn); exit(2); }
  fp_t buf[100000];
  n = 0;
  double v;
  while (fscanf(f, "%lf", &v) == 1) buf[n++] = v * 1e-8;  // Angstrom -> cm
  fclose(f);
  fp_t *lam = new fp_t[n] - 1;  // 1-based, as jsub/cfg.cc:209-212 hands it over
  for (int i = 1; i <= n; ++i) lam[i] = buf[i - 1];
  return lam;
}

static atmosphere *make_atmos(const char *cfgfile, io_class &io) {
  char *txt = slurp(cfgfile);
  // the config file holds the body of an atmos{} block
  acfg *cfg = new acfg(txt, io);
  atmosphere *atm = atmos_new(cfg, io);
  char wd[4096];
  if (!getcwd(wd, sizeof(wd))) strcpy(wd, ".");
  atm->init(wd, &io);
  // Hand the column over the way the master hands it to a worker: pack it and
  // rebuild it from the buffer (master/job.cc:63-86, slave/slave.cc:107-110).
  int32_t sz = atm->size(io);
  uint08_t *buf = new uint08_t[sz];
  atm->pack(buf, 0, io);
  int32_t offs = 0;
  atmosphere *w = atmos_new(buf, offs, 0, io);
  delete[] buf;
  delete atm;
  atm = w;
  // SNAPI_REF_HE=1: re-integrate hydrostatic equilibrium on the tau grid with
  // the reference's own equation of state (plpar/atmos_pp.cc:373-455) before
  // anything else, so that Pg follows from T(tau) and log g alone.
  if (getenv("SNAPI_REF_HE") && atoi(getenv("SNAPI_REF_HE"))) atm->enforce_hequilibrium();
  return atm;
}

static void dump_state(atmosphere *a, fp_t ***Vr, fp_t ****Bv) {
  int x1 = a->x1l, x2 = a->x2l;
  printf("nd %d\n", a->x3h - a->x3l + 1);
  printf("natm %d\n", a->natm);
  printf("tau_grid %d\n", a->tau_grid);
  for (int z = a->x3l; z <= a->x3h; ++z) {
    printf("depth %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
           z, a->T[x1][x2][z], a->Ne[x1][x2][z], a->Nt[x1][x2][z], a->Vt[x1][x2][z],
           Vr[x1][x2][z], Bv[1][x1][x2][z], Bv[2][x1][x2][z], Bv[3][x1][x2][z],
           a->op_referent[x1][x2][z], a->tau_referent[x1][x2][z], a->rt_grid[z],
           a->get_pop(x1, x2, z, 0, 0), a->x3[z]);
  }
  // per-atom line data and populations
  for (int s = 0; s < a->natm; ++s) {
    atom *at = dynamic_cast<atom *>(a->atml[s]);
    if (!at) {
      printf("species %d mol %s\n", s, a->atml[s]->id);
      continue;
    }
    printf("species %d atom %s Z %d mass %.17g nlte %d\n", s, at->id, (int)at->Z, at->mass, (int)at->NLTE);
    for (int zi = 0; zi <= at->Z; ++zi)
      for (int l = 0; l < at->nl[zi]; ++l)
        printf("level %d %d %d %.17g g %d j %d lqn %d glande %.17g\n", s, zi, l, at->ee[zi][l], (int)at->g[zi][l],
               (int)at->j_qn[zi][l], (int)at->l_qn[zi][l], at->g_lande[zi][l]);
    for (int zi = 0; zi <= at->Z; ++zi)
      for (int i = 1; i < at->nl[zi]; ++i)
        for (int ii = 0; ii < i; ++ii) {
          if (!(at->A[zi][i][ii] > 1E0)) continue;
          int tr = at->tmap[zi][i][ii];
          fp_t lam = h * c / (at->ee[zi][i] - at->ee[zi][ii]);
          printf("line %d %d %d %d tr %d lam %.17g A %.17g Blu %.17g Bul %.17g grad %.17g sigma %.17g alpha %.17g split %d\n",
                 s, zi, ii, i, tr, lam, at->A[zi][i][ii], at->B[zi][ii][i], at->B[zi][i][ii],
                 at->damp_rad(at->A[zi], ii, i), at->col_dam_cross_section[zi][i][ii],
                 at->alpha_col_dam[zi][i][ii], at->split ? at->split[tr] : -1);
          if (at->split && at->split[tr]) {
            for (int m = 0; m < at->nm[tr][0]; ++m)
              printf("zc %d %d p %.17g %.17g\n", s, tr, at->delta_lambda_p[tr][m], at->S_p[tr][m]);
            for (int m = 0; m < at->nm[tr][1]; ++m)
              printf("zc %d %d b %.17g %.17g\n", s, tr, at->delta_lambda_b[tr][m], at->S_b[tr][m]);
            for (int m = 0; m < at->nm[tr][2]; ++m)
              printf("zc %d %d r %.17g %.17g\n", s, tr, at->delta_lambda_r[tr][m], at->S_r[tr][m]);
          }
          for (int z = a->x3l; z <= a->x3h; ++z) {
            fp_t T = a->T[x1][x2][z];
            fp_t dld = at->broad_dop(lam, T, a->Vt[x1][x2][z]);
            fp_t gc = at->damp_col(x1, x2, z, zi, i, ii, T, a->Ne[x1][x2][z], lam);
            printf("lpop %d %d %d %.17g %.17g %.17g %.17g\n", s, tr, z, at->pop[x1][x2][z].n[zi][ii],
                   at->pop[x1][x2][z].n[zi][i], dld, gc);
          }
        }
  }
}

static void dump_opem(atmosphere *a, int nlambda, fp_t *lam_vac, fp_t ******op, fp_t *****em, int every) {
  int x1 = a->x1l, x2 = a->x2l;
  for (int l = 1; l <= nlambda; l += every)
    for (int z = a->x3l; z <= a->x3h; ++z) {
      printf("opem %d %d %.17g", l, z, lam_vac[l]);
      for (int s = 1; s <= 4; ++s)
        for (int sp = 1; sp <= 4; ++sp) printf(" %.17g", op[l][x1][x2][z][s][sp]);
      for (int s = 1; s <= 4; ++s) printf(" %.17g", em[l][x1][x2][z][s]);
      printf("\n");
    }
}

// Background (every species but the line transitions) per depth at lambda index l:
// atom bound-free and H- by calling each species with its lines removed is not
// possible without editing the reference, so the driver records each species'
// whole contribution separately instead.
static void dump_species_opem(atmosphere *a, fp_t ***Vr, fp_t ****Bv, fp_t th, fp_t ph, int nlambda,
                              fp_t *lam_vac, int every) {
  int x1 = a->x1l, x2 = a->x2l;
  for (int s = 0; s < a->natm; ++s) {
    fp_t ******op = ft6dim(1, nlambda, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h, 1, 4, 1, 4);
    fp_t *****em = ft5dim(1, nlambda, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h, 1, 4);
    memset(op[1][x1][x2][a->x3l][1] + 1, 0, nlambda * (a->x3h - a->x3l + 1) * 16 * sizeof(fp_t));
    memset(em[1][x1][x2][a->x3l] + 1, 0, nlambda * (a->x3h - a->x3l + 1) * 4 * sizeof(fp_t));
    a->atml[s]->op_em_vector(a->T, a->Ne, Vr, a->Vt, Bv, th, ph, lam_vac, nlambda, op, em);
    for (int l = 1; l <= nlambda; l += every)
      for (int z = a->x3l; z <= a->x3h; ++z)
        printf("spop %d %d %d %.17g %.17g\n", s, l, z, op[l][x1][x2][z][1][1], em[l][x1][x2][z][1]);
    // bound-free part of an atom on its own (atom_rad_synthesis.cc:226-262)
    if (atom *at = dynamic_cast<atom *>(a->atml[s])) {
      memset(op[1][x1][x2][a->x3l][1] + 1, 0, nlambda * (a->x3h - a->x3l + 1) * 16 * sizeof(fp_t));
      memset(em[1][x1][x2][a->x3l] + 1, 0, nlambda * (a->x3h - a->x3l + 1) * 4 * sizeof(fp_t));
      at->boundfree_op_em_vector(a->T, a->Ne, Vr, th, ph, lam_vac, nlambda, op, em);
      for (int l = 1; l <= nlambda; l += nlambda - 1)
        for (int z = a->x3l; z <= a->x3h; ++z)
          printf("bfop %d %d %d %.17g %.17g\n", s, l, z, op[l][x1][x2][z][1][1], em[l][x1][x2][z][1]);
      // each transition's edge and cross-section at the last wavelength, where
      // the background of a window below every edge is evaluated
      for (int zi = 0; zi < at->Z; ++zi)
        for (int l = 0; l < at->nl[zi]; ++l)
          if (at->bf[zi][l])
            printf("bfsig %d %d %d %.17g %.17g %.17g\n", s, zi, l, at->bf[zi][l]->getlambda_crit(),
                   lam_vac[nlambda], at->bf[zi][l]->U(lam_vac[nlambda]));
    }
    del_ft6dim(op, 1, nlambda, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h, 1, 4, 1, 4);
    del_ft5dim(em, 1, nlambda, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h, 1, 4);
  }
}

static int mode_synth(int argc, char **argv, io_class &io) {
  if (argc < 6) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  int nlambda;
  fp_t *lambda = read_lambda(argv[3], nlambda);
  fp_t el = atof(argv[4]), az = atof(argv[5]);
  int every = (argc > 6) ? atoi(argv[6]) : 10;
  int x1 = a->x1l, x2 = a->x2l;

  // The preamble of atmosphere::obs_stokes (atmos/atmos_obs.cc:23-58), step by
  // step, so the state in between can be recorded.
  a->set_grid(a->tau_grid);
  a->boundary_condition_for_rt = -1;
  a->popsetup();
  a->compute_op_referent();
  if (!a->tau_grid) a->compute_tau_referent();
  a->nltepops();
  a->atm_pop_setup();
  a->atm_pop_fill();
  fp_t ***Vr = a->project(a->Vx, a->Vy, a->Vz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t ****Bv = a->transform(a->Bx, a->By, a->Bz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  for (int s = 0; s < a->natm; ++s) {
    a->atml[s]->rtsetup(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    a->atml[s]->zeeman_setup();
  }
  fp_t *lam_vac = airtovac(lambda + 1, nlambda) - 1;
  printf("nlambda %d\n", nlambda);
  for (int l = 1; l <= nlambda; ++l) printf("lambda %d %.17g %.17g\n", l, lambda[l], lam_vac[l]);
  dump_state(a, Vr, Bv);
  dump_species_opem(a, Vr, Bv, el, az, nlambda, lam_vac, every);

  fp_t ******op = ft6dim(1, nlambda, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h, 1, 4, 1, 4);
  fp_t *****em = ft5dim(1, nlambda, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h, 1, 4);
  a->op_em_vector(Vr, Bv, el, az, lam_vac, nlambda, op, em);
  dump_opem(a, nlambda, lam_vac, op, em, every);
  fp_t ****S = ft4dim(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h, 1, 4);
  for (int l = 1; l <= nlambda; ++l) {
    if (a->tau_grid) a->normalize_to_referent_opacity(op[l], em[l]);
    a->formal(a->rt_grid, S, 0, op[l], em[l], el, az, a->boundary_condition_for_rt);
    printf("stokes_steps %d %.17g %.17g %.17g %.17g\n", l, S[x1][x2][a->x3l][1], S[x1][x2][a->x3l][2],
           S[x1][x2][a->x3l][3], S[x1][x2][a->x3l][4]);
    if (l == 1 || l == nlambda / 2 || (every > 0 && l % (every * 4) == 1))
      for (int z = a->x3l; z <= a->x3h; ++z)
        printf("sdepth %d %d %.17g %.17g %.17g %.17g\n", l, z, S[x1][x2][z][1], S[x1][x2][z][2], S[x1][x2][z][3],
               S[x1][x2][z][4]);
  }
  for (int s = 0; s < a->natm; ++s) {
    a->atml[s]->rtclean(0, 0, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    a->atml[s]->zeeman_clear();
  }
  a->popclean();

  // And the entry point itself, end to end, for the final observable.
  observable *o = a->obs_stokes(el, az, lambda, nlambda);
  fp_t **Sv = o->get_S(1, 1);
  for (int l = 1; l <= nlambda; ++l)
    printf("stokes %d %.17g %.17g %.17g %.17g\n", l, Sv[1][l], Sv[2][l], Sv[3][l], Sv[4][l]);
  delete o;
  return 0;
}

// The state and thermodynamic derivatives the response path works with,
// recorded right where obs_stokes_responses (atmos/atmos_rad_synthesis.cc:
// 166-399) builds its opacities: the same preamble (:173-190), then the
// per-depth state, d/dT and d/dNt (p = 1, 2) of Ne, neutral H, the referent
// opacity and the populations of every line level, and the bound-free
// background.  ne_lte_derivatives (atmos/atmos_responses.cc:364-423) leaves T
// raised by delta_T/2 and Ne, populations re-solved there, so the state is
// recorded after it, as the response path sees it.
static void dump_resp_state(atmosphere *a, fp_t el, fp_t az, fp_t *lambda, int nlambda) {
  int x1 = a->x1l, x2 = a->x2l;
  a->boundary_condition_for_rt = -1;
  a->set_grid(a->tau_grid);
  a->popsetup();
  for (int s = 0; s < a->natm; ++s) a->atml[s]->set_parent_atmosphere(a);
  if (a->tau_grid) {
    a->compute_op_referent();
    a->compute_op_referent_derivative();
  }
  a->nltepops();
  a->respsetup();
  a->ne_lte_derivatives();
  a->compute_nlte_population_responses(0);
  fp_t ***Vr = a->project(a->Vx, a->Vy, a->Vz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t ****Bv = a->transform(a->Bx, a->By, a->Bz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  for (int s = 0; s < a->natm; ++s) {
    a->atml[s]->rtsetup(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    a->atml[s]->zeeman_setup();
  }
  fp_t *lam_vac = airtovac(lambda + 1, nlambda) - 1;
  for (int l = 1; l <= nlambda; ++l) printf("lambda %d %.17g %.17g\n", l, lambda[l], lam_vac[l]);
  dump_state(a, Vr, Bv);
  // the field in the atmosphere's frame, which transform_responses reads
  for (int z = a->x3l; z <= a->x3h; ++z)
    printf("bvec %d %.17g %.17g %.17g\n", z, a->Bx[x1][x2][z], a->By[x1][x2][z], a->Bz[x1][x2][z]);
  for (int z = a->x3l; z <= a->x3h; ++z) {
    printf("thermo %d %.17g", z, a->Nt[x1][x2][z]);
    for (int p = 1; p <= 2; ++p)
      printf(" %.17g %.17g %.17g", a->Ne_lte_der[p][x1][x2][z], a->get_neutral_H_derivative_lte(p, x1, x2, z),
             a->tau_grid ? a->op_referent_derivative[p][z][x1][x2][z] : 0.0);
    printf("\n");
  }
  for (int s = 0; s < a->natm; ++s) {
    atom *at = dynamic_cast<atom *>(a->atml[s]);
    if (!at) continue;
    for (int zi = 0; zi <= at->Z; ++zi)
      for (int i = 1; i < at->nl[zi]; ++i)
        for (int ii = 0; ii < i; ++ii) {
          if (!(at->A[zi][i][ii] > 1E0)) continue;
          for (int z = a->x3l; z <= a->x3h; ++z) {
            printf("dpop %d %d %d %d %d", s, zi, ii, i, z);
            for (int p = 1; p <= 2; ++p)
              printf(" %.17g %.17g", at->level_responses[p][(z - a->x3l) * at->nmap + at->rmap[zi][ii] + 1][z],
                     at->level_responses[p][(z - a->x3l) * at->nmap + at->rmap[zi][i] + 1][z]);
            printf("\n");
          }
        }
  }
  dump_species_opem(a, Vr, Bv, el, az, nlambda, lam_vac, nlambda > 1 ? nlambda - 1 : 1);
  for (int s = 0; s < a->natm; ++s) {
    a->atml[s]->zeeman_clear();
    a->atml[s]->rtclean(1, nlambda, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  }
  a->respclean();
  a->popclean();
  if (a->tau_grid) a->delete_op_referent_derivative();
}

static int mode_resp(int argc, char **argv, io_class &io) {
  if (argc < 6) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  int nlambda;
  fp_t *lambda = read_lambda(argv[3], nlambda);
  fp_t el = atof(argv[4]), az = atof(argv[5]);
  dump_resp_state(a, el, az, lambda, nlambda);
  // the entry point itself runs on a fresh copy of the same column
  delete a;
  a = make_atmos(argv[2], io);
  a->set_grid(a->tau_grid);
  int nd = a->x3h - a->x3l + 1;
  fp_t ****rsp = ft4dim(1, 7, a->x3l, a->x3h, 1, nlambda, 1, 4);
  // The response entry point of the fit path (atmos/atmos_rad_synthesis.cc:
  // 166-399) with no node model: it fills the per-depth responses to the
  // seven physical parameters T, Nt, vt, vLOS, |B|, incl, azi.
  observable *o = a->obs_stokes_responses(el, az, lambda, nlambda, (fp_t ***)0, (model *)0, rsp);
  fp_t **Sv = o->get_S(1, 1);
  printf("nlambda %d nd %d\n", nlambda, nd);
  for (int l = 1; l <= nlambda; ++l)
    printf("stokes %d %.17g %.17g %.17g %.17g\n", l, Sv[1][l], Sv[2][l], Sv[3][l], Sv[4][l]);
  for (int p = 1; p <= 7; ++p)
    for (int z = a->x3l; z <= a->x3h; ++z)
      for (int l = 1; l <= nlambda; ++l)
        printf("resp %d %d %d %.17g %.17g %.17g %.17g\n", p, z, l, rsp[p][z][l][1], rsp[p][z][l][2], rsp[p][z][l][3],
               rsp[p][z][l][4]);
  return 0;
}

// xorshift-based deterministic generator shared with the Python side of the
// tests (tests/golden/make_golden.py restates it).
static unsigned long long g_state;
static double urand() {
  g_state ^= g_state << 13;
  g_state ^= g_state >> 7;
  g_state ^= g_state << 17;
  return (double)(g_state >> 11) * (1.0 / 9007199254740992.0);
}

static int mode_lmstep(int argc, char **argv, io_class &io) {
  if (argc < 6) return 2;
  int nrow = atoi(argv[2]), npar = atoi(argv[3]);
  fp_t lm = atof(argv[4]);
  g_state = strtoull(argv[5], 0, 10) * 2654435761ULL + 88172645463325252ULL;
  fp_t **J = ft2dim(1, nrow, 1, npar);
  fp_t *r = new fp_t[nrow] - 1;
  for (int i = 1; i <= nrow; ++i)
    for (int p = 1; p <= npar; ++p) J[i][p] = 2.0 * urand() - 1.0;
  // make the last column nearly dependent on the first so the SVD cut matters
  for (int i = 1; i <= nrow; ++i) J[i][npar] = J[i][1] * (1.0 + 1e-9 * (2.0 * urand() - 1.0));
  for (int i = 1; i <= nrow; ++i) r[i] = 2.0 * urand() - 1.0;
  printf("nrow %d npar %d lm %.17g\n", nrow, npar, lm);
  for (int i = 1; i <= nrow; ++i) {
    printf("J %d", i);
    for (int p = 1; p <= npar; ++p) printf(" %.17g", J[i][p]);
    printf(" r %.17g\n", r[i]);
  }
  // J^T J and J^T r exactly as atmos/atmos_fit.cc:129-146 forms them
  fp_t **JT = transpose(J, nrow, npar);
  fp_t **JTJ = multiply_with_transpose(J, nrow, npar);
  fp_t *rhs = multiply_vector(JT, r, npar, nrow);
  for (int i = 1; i <= npar; ++i) {
    printf("JTJ %d", i);
    for (int j = 1; j <= npar; ++j) printf(" %.17g", JTJ[i][j]);
    printf(" rhs %.17g\n", rhs[i]);
  }
  for (int i = 1; i <= npar; ++i) JTJ[i][i] *= (lm + 1.0);
  fp_t **corr = solve_svd(JTJ, &rhs - 1, 1, npar, 1);
  printf("corr");
  for (int i = 1; i <= npar; ++i) printf(" %.17g", corr[1][i]);
  printf("\n");
  return 0;
}

static int mode_formal_scalar(int argc, char **argv, io_class &io) {
  if (argc < 6) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  int nlambda;
  fp_t *lambda = read_lambda(argv[3], nlambda);
  fp_t el = atof(argv[4]), az = atof(argv[5]);
  int x1 = a->x1l, x2 = a->x2l;
  a->set_grid(a->tau_grid);
  a->boundary_condition_for_rt = -1;
  a->popsetup();
  a->compute_op_referent();
  if (!a->tau_grid) a->compute_tau_referent();
  a->nltepops();
  fp_t ***Vr = a->project(a->Vx, a->Vy, a->Vz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t ****Bv = a->transform(a->Bx, a->By, a->Bz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t ***S = ft3dim(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t ***L = ft3dim(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  printf("nd %d\n", a->x3h - a->x3l + 1);
  for (int z = a->x3l; z <= a->x3h; ++z) printf("grid %d %.17g\n", z, a->rt_grid[z]);
  // the background opacity of the ALI loop (atmos/atmos_pop.cc:148-160,
  // atmos/atmos_rad.cc:225-246) needs the transition maps of rtsetup
  for (int s = 0; s < a->natm; ++s) a->atml[s]->rtsetup(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  for (int l = 1; l <= nlambda; ++l) {
    fp_t lv = airtovac(lambda[l]);
    for (int s = 0; s < a->natm; ++s) a->atml[s]->prof_init();
    fp_t ***op = a->opacity_lte(a->T, a->Ne, Vr, a->Vt, Bv, el, az, lv);
    fp_t ***em = a->emissivity_lte(a->T, a->Ne, Vr, a->Vt, Bv, el, az, lv);
    if (a->tau_grid) a->normalize_to_referent_opacity(op, em);
    for (int z = a->x3l; z <= a->x3h; ++z)
      printf("opem1 %d %d %.17g %.17g\n", l, z, op[x1][x2][z], em[x1][x2][z]);
    a->formal(a->rt_grid, S, L, op, em, el, az, a->boundary_condition_for_rt);
    for (int z = a->x3l; z <= a->x3h; ++z)
      printf("sl %d %d %.17g %.17g\n", l, z, S[x1][x2][z], L[x1][x2][z]);
    del_ft3dim(op, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    del_ft3dim(em, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  }
  for (int s = 0; s < a->natm; ++s) a->atml[s]->rtclean(0, 0, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  a->popclean();
  return 0;
}

// NLTE populations of the active atom (atmosphere::nltepops, atmos/
// atmos_pop.cc:67-240).  The ALI loop is driven here with the reference's own
// functions so that every input of one iteration (frequency grid, rays,
// background opacity, profiles' damping, collision rates, cross-sections)
// and its results (J, Lambda*, norm sums after atom::add; the populations of
// atom::newpops) can be written; the converged populations of the
// reference's own nltepops() on a fresh atmosphere are written last.
static int mode_nlte(int argc, char **argv, io_class &io) {
  if (argc < 3) return 2;
  int maxit = (argc > 3) ? atoi(argv[3]) : 100;
  atmosphere *a = make_atmos(argv[2], io);
  const int x1 = a->x1l, x2 = a->x2l;
  a->set_grid(a->tau_grid);
  a->boundary_condition_for_rt = -1;
  a->popsetup();
  a->compute_op_referent();
  if (!a->tau_grid) a->compute_tau_referent();
  // --- nltepops set-up (atmos_pop.cc:71-131)
  a->ltepops();
  fp_t *lambda = 0;
  int32_t nlambda = 0;
  for (int s = 0; s < a->natm; ++s) a->atml[s]->rtsetup(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  for (int s = 0; s < a->natm; ++s)
    lambda = a->atml[s]->getlambda(lambda, nlambda, a->T[x1][x2][a->x3h], a->Nt[x1][x2][a->x3h], a->Ne[x1][x2][a->x3h]);
  int ntp;
  fp_t *th, *ph;
  fp_t *bin = a->anglesetup(th, ph, ntp);
  fp_t *lambda_w = new fp_t[nlambda];
  if (nlambda >= 2) {
    lambda_w[0] = (lambda[1] - lambda[0]) * 0.5;
    lambda_w[nlambda - 1] = (lambda[nlambda - 1] - lambda[nlambda - 2]) * 0.5;
    for (int l = 1; l < nlambda - 1; ++l) lambda_w[l] = (lambda[l + 1] - lambda[l - 1]) * 0.5;
  } else if (nlambda)
    lambda_w[0] = 1.0;
  if (a->tau_grid) a->compute_op_referent();
  for (int s = 0; s < a->natm; ++s) a->atml[s]->compute_active_population(a->T, a->Ne);
  atom *at = 0;
  int sa = -1;
  for (int s = 0; s < a->natm; ++s) {
    atom *t = dynamic_cast<atom *>(a->atml[s]);
    if (t && t->NLTE) {
      at = t;
      sa = s;
    }
  }
  if (!at) return 3;
  const int nd = a->x3h - a->x3l + 1;
  printf("nd %d\nntp %d\nnlam %d\nboundary %.17g\ntau_grid %d\n", nd, ntp, nlambda, a->boundary_condition_for_rt,
         a->tau_grid);
  for (int tp = 1; tp <= ntp; ++tp) printf("ray %d %.17g %.17g\n", tp, th[tp], bin[tp]);
  for (int l = 0; l < nlambda; ++l)
    printf("lam %d %.17g %.17g %.17g\n", l, lambda[l], lambda_w[l], a->get_opacity_fudge(lambda[l]));
  printf("atom %d nmap %d ntr %d Z %d mass %.17g\n", sa, at->nmap, at->ntr, (int)at->Z, at->mass);
  for (int i = 0; i < at->nmap; ++i)
    printf("level %d %d %d %.17g %d\n", i, (int)at->zmap[i], (int)at->lmap[i], at->ee[at->zmap[i]][at->lmap[i]],
           (int)at->g[at->zmap[i]][at->lmap[i]]);
  for (int tr = 1; tr <= (int)at->ntr; ++tr) {
    const int z = at->inverse_tmap[tr][2], lo = at->inverse_tmap[tr][3], up = at->inverse_tmap[tr][4];
    if (up < at->nl[z]) {
      const fp_t lam0 = h * c / (at->ee[z][up] - at->ee[z][lo]);
      printf("trans %d line %d %d %.17g %.17g %.17g %.17g %.17g\n", tr, (int)at->rmap[z][lo], (int)at->rmap[z][up],
             lam0, at->A[z][up][lo], at->B[z][lo][up], at->B[z][up][lo], at->ee[z][up] - at->ee[z][lo]);
    } else {
      printf("trans %d bf %d %d %.17g %.17g\n", tr, (int)at->rmap[z][lo], (int)at->rmap[z + 1][0],
             (fp_t)at->g[z][lo] / (fp_t)at->g[z + 1][0], at->ip[z] - at->ee[z][lo]);
      for (int l = 0; l < nlambda; ++l)
        printf("sigma %d %d %.17g\n", tr, l, at->bf[z][lo] ? at->bf[z][lo]->U(lambda[l]) : 0.0);
    }
  }
  for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) {
    const fp_t T = a->T[x1][x2][x3i], ne = a->Ne[x1][x2][x3i];
    printf("depth %d %.17g %.17g %.17g %.17g %.17g %.17g\n", x3i, T, a->Vt[x1][x2][x3i], ne,
           a->op_referent[x1][x2][x3i], a->rt_grid[x3i], at->pop[x1][x2][x3i].Na);
    for (int i = 0; i < at->nmap; ++i)
      printf("pop0 %d %d %.17g\n", x3i, i, at->pop[x1][x2][x3i].n[at->zmap[i]][at->lmap[i]]);
    for (int tr = 1; tr <= (int)at->ntr; ++tr) {
      const int z = at->inverse_tmap[tr][2], lo = at->inverse_tmap[tr][3], up = at->inverse_tmap[tr][4];
      if (up < at->nl[z]) {
        const fp_t lam0 = h * c / (at->ee[z][up] - at->ee[z][lo]);
        const fp_t ar = at->damp_rad(at->A[z], lo, up);
        const fp_t gc = at->damp_col(x1, x2, x3i, z, up, lo, T, ne, lam0);
        printf("gamma %d %d %.17g\n", tr, x3i, ar + gc * turn_on_damping);
      }
    }
    // collision rates between mapped levels, as atom::newpops uses them (atom_pop.cc:51-105)
    for (int i = 0; i < at->nmap - 1; ++i) {
      const int l = at->lmap[i], z = at->zmap[i];
      for (int ll = 0; ll < at->nl[z]; ++ll) {
        if (ll == l) continue;
        printf("coll %d %d %d %.17g\n", x3i, i, (int)at->rmap[z][ll], at->C_ij(z, l, ll, T, ne));
      }
      const int ic = at->rmap[z + 1][0];
      printf("coll %d %d %d %.17g\n", x3i, i, ic, at->C_i_cont(z, l, T, ne));
      printf("coll %d %d %d %.17g\n", x3i, ic, i, at->C_cont_i(z, l, T, ne));
    }
  }
  // --- the ALI loop (atmos_pop.cc:133-205)
  fp_t ***S = ft3dim(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t ***L = ft3dim(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t relative_change = 1.0;
  int iter;
  for (iter = 0; iter < maxit; ++iter) {
    for (int s = 0; s < a->natm; ++s) a->atml[s]->rtinit();
    for (int tp = 1; tp <= ntp; ++tp) {
      fp_t ***Vr = a->project(a->Vx, a->Vy, a->Vz, th[tp], ph[tp], a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      fp_t ****Bv = a->transform(a->Bx, a->By, a->Bz, th[tp], ph[tp], a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      if (iter == 0)
        for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) printf("vproj %d %d %.17g\n", tp, x3i, Vr[x1][x2][x3i]);
      for (int l = 0; l < nlambda; ++l) {
        for (int s = 0; s < a->natm; ++s) a->atml[s]->prof_init();
        fp_t ***op = a->opacity_lte(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
        fp_t ***em = a->emissivity_lte(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
        if (iter == 0) {
          // the active atom's population-independent part (free-free, Rayleigh)
          fp_t ***ff = at->freefree_op(a->T, a->Ne, Vr, lambda[l]);
          fp_t ***ry = at->rayleigh_op(lambda[l]);
          for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) {
            const fp_t p = ff[x1][x2][x3i] + ry[x1][x2][x3i];
            printf("bg %d %d %d %.17g %.17g\n", tp, l, x3i, op[x1][x2][x3i] + p,
                   em[x1][x2][x3i] + p * Planck_f(lambda[l], a->T[x1][x2][x3i]));
          }
          del_ft3dim(ff, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
          del_ft3dim(ry, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
        }
        op = a->add(a->opacity_active_only(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]), op, a->x1l, a->x1h,
                    a->x2l, a->x2h, a->x3l, a->x3h);
        em = a->add(a->emissivity_active_only(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]), em, a->x1l,
                    a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
        if (a->tau_grid) a->normalize_to_referent_opacity(op, em);
        a->formal(a->rt_grid, S, L, op, em, th[tp], ph[tp], a->boundary_condition_for_rt);
        if (a->tau_grid) a->de_normalize(op, em);
        for (int s = 0; s < a->natm; ++s) a->atml[s]->add(S, L, op, lambda[l], lambda_w[l], bin[tp]);
        del_ft3dim(em, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
        del_ft3dim(op, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      }
      del_ft4dim(Bv, 1, 3, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      del_ft3dim(Vr, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    }
    if (iter == 0) {
      for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) {
        for (int tr = 1; tr <= (int)at->ntr; ++tr)
          printf("acc %d %d %.17g %.17g %.17g %.17g\n", x3i, tr, at->Jb[x1][x2][x3i][tr], at->Ju[x1][x2][x3i][tr],
                 at->Ls[x1][x2][x3i][tr], at->norm[x1][x2][x3i][tr]);
        fp_t *np = at->newpops(x1, x2, x3i, 1);
        for (int i = 0; i < at->nmap; ++i) printf("newpop %d %d %.17g\n", x3i, i, np[i]);
        delete[] np;
      }
    }
    relative_change = a->newpops(a->T, a->Nt, a->Ne, lambda, nlambda, 1);
    int alo = 1;
    if (relative_change < 0.0) {
      alo = 0;
      relative_change = a->newpops(a->T, a->Nt, a->Ne, lambda, nlambda, 0);
    }
    printf("iter %d %.17g %d\n", iter, relative_change, alo);
    if (relative_change < 0.0) break;
    if (relative_change < 1E-2) break;
  }
  for (int x3i = a->x3l; x3i <= a->x3h; ++x3i)
    for (int i = 0; i < at->nmap; ++i)
      printf("final %d %d %.17g\n", x3i, i, at->pop[x1][x2][x3i].n[at->zmap[i]][at->lmap[i]]);
  // --- the reference's own nltepops on a fresh atmosphere
  atmosphere *b = make_atmos(argv[2], io);
  b->set_grid(b->tau_grid);
  b->boundary_condition_for_rt = -1;
  b->popsetup();
  b->compute_op_referent();
  if (!b->tau_grid) b->compute_tau_referent();
  int rv = b->nltepops();
  atom *bt = dynamic_cast<atom *>(b->atml[sa]);
  printf("ref_rv %d\n", rv);
  for (int x3i = b->x3l; x3i <= b->x3h; ++x3i)
    for (int i = 0; i < bt->nmap; ++i)
      printf("ref_final %d %d %.17g\n", x3i, i, bt->pop[x1][x2][x3i].n[bt->zmap[i]][bt->lmap[i]]);
  return 0;
}

// Node interpolation (cmlib/mathtools.cc:1208-1315) for a seeded node set:
// interp <nnode> <is_temp> <seed>; grid of 71 log tau points from -6 to 1.
static int mode_interp(int argc, char **argv, io_class &io) {
  if (argc < 5) return 2;
  int nn = atoi(argv[2]), is_temp = atoi(argv[3]);
  g_state = strtoull(argv[4], 0, 10) * 2654435761ULL + 88172645463325252ULL;
  const int nd = 71;
  fp_t *nt = new fp_t[nn] - 1, *nv = new fp_t[nn] - 1;
  fp_t *grid = new fp_t[nd] - 1, *q = new fp_t[nd] - 1;
  // ascending node positions inside [-5.5, 0.5], values of either sign
  fp_t pos = -5.5;
  for (int i = 1; i <= nn; ++i) {
    pos += (6.0 / nn) * (0.3 + 0.7 * urand());
    nt[i] = pos;
    nv[i] = is_temp ? 4000.0 + 3000.0 * urand() : 2.0 * urand() - 1.0;
  }
  // optional: the deepest node's position (the temperature's bottom slope
  // comes from a table in that position)
  if (argc > 5) nt[nn] = atof(argv[5]);
  for (int z = 1; z <= nd; ++z) grid[z] = -6.0 + 7.0 * (z - 1) / (nd - 1);
  atmospheric_interpolation(nt, nv, nn, grid, q, 1, nd, is_temp);
  printf("nodes");
  for (int i = 1; i <= nn; ++i) printf(" %.17g %.17g", nt[i], nv[i]);
  printf("\n");
  for (int z = 1; z <= nd; ++z) printf("q %.17g %.17g\n", grid[z], q[z]);
  return 0;
}

// Gaussian spectral convolution (convolve_spectra_with_gauss, cmlib/
// mathtools.cc:1350-1427) of a seeded four-component spectrum:
//   conv <nlambda> <width [A]> <seed> <uneven>
// wavelengths around 6302 A, evenly spaced 10 mA apart or (uneven = 1) with
// seeded gaps of 5..15 mA; the components are two absorption profiles and
// their derivatives.  Prints the grid, the input and the convolved spectra.
static int mode_conv(int argc, char **argv, io_class &io) {
  if (argc < 6) return 2;
  int nl = atoi(argv[2]);
  fp_t width = atof(argv[3]) * 1E-8;
  g_state = strtoull(argv[4], 0, 10) * 2654435761ULL + 88172645463325252ULL;
  int uneven = atoi(argv[5]);
  fp_t *lam = new fp_t[nl] - 1;
  fp_t **S = ft2dim(1, 4, 1, nl);
  fp_t x = 6300.5;
  for (int l = 1; l <= nl; ++l) {
    lam[l] = x * 1E-8;
    x += uneven ? 0.005 + 0.01 * urand() : 0.01;
  }
  const fp_t c1 = 6301.5, c2 = 6302.5, d = 0.08 + 0.04 * urand();
  for (int l = 1; l <= nl; ++l) {
    const fp_t a = lam[l] * 1E8;
    const fp_t g1 = exp(-(a - c1) * (a - c1) / (d * d)), g2 = exp(-(a - c2) * (a - c2) / (d * d));
    S[1][l] = 1.0 - 0.6 * g1 - 0.4 * g2;
    S[2][l] = 0.05 * g1 * (a - c1) / d;
    S[3][l] = -0.03 * g2;
    S[4][l] = 0.2 * (a - c1) / d * g1 - 0.1 * (a - c2) / d * g2;
  }
  for (int l = 1; l <= nl; ++l)
    printf("in %.17g %.17g %.17g %.17g %.17g\n", lam[l], S[1][l], S[2][l], S[3][l], S[4][l]);
  convolve_spectra_with_gauss(S, lam, nl, width);
  for (int l = 1; l <= nl; ++l)
    printf("out %.17g %.17g %.17g %.17g %.17g\n", lam[l], S[1][l], S[2][l], S[3][l], S[4][l]);
  return 0;
}

static int mode_psf(int argc, char **argv, io_class &io) {
  if (argc < 5) return 2;
  const int nl = atoi(argv[2]), n = atoi(argv[3]);
  if (nl < 1 || n < 1) return 2;
  g_state = strtoull(argv[4], 0, 10) * 2654435761ULL + 88172645463325252ULL;
  fp_t **S = ft2dim(1, 4, 1, nl);
  for (int l = 1; l <= nl; ++l)
    for (int s = 1; s <= 4; ++s) S[s][l] = (s == 1 ? 1.0 : 0.0) + 0.3 * (2.0 * urand() - 1.0);
  fp_t *p = new fp_t[n];
  for (int j = 0; j < n; ++j) p[j] = 0.1 + urand();
  fp_t *arr = new fp_t[n + 1];
  arr[0] = p[0];
  for (int j = 1; j < n; ++j) arr[j] = p[j];
  arr[n] = p[0];
  printf("psf");
  for (int j = 0; j < n; ++j) printf(" %.17g", p[j]);
  printf("\n");
  for (int l = 1; l <= nl; ++l) printf("in %.17g %.17g %.17g %.17g\n", S[1][l], S[2][l], S[3][l], S[4][l]);
  convolve_spectra_with_psf(S, 0, nl, n, arr);
  for (int l = 1; l <= nl; ++l) printf("out %.17g %.17g %.17g %.17g\n", S[1][l], S[2][l], S[3][l], S[4][l]);
  return 0;
}

static model *read_model(FILE *f);

static int mode_reg(int argc, char **argv, io_class &io) {
  if (argc < 4) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  FILE *f = fopen(argv[3], "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", argv[3]); return 2; }
  model *m = read_model(f);
  fclose(f);
  const int np = m->get_N_nodes_total();
  fp_t **JTJ = ft2dim(1, np, 1, np);
  memset(JTJ[1] + 1, 0, np * np * sizeof(fp_t));
  fp_t *rhs = new fp_t[np] - 1;
  memset(rhs + 1, 0, np * sizeof(fp_t));
  a->regularize_hessian(JTJ, rhs, m);
  printf("np %d\n", np);
  for (int i = 1; i <= np; ++i) {
    printf("G %d", i);
    for (int j = 1; j <= np; ++j) printf(" %.17g", JTJ[i][j]);
    printf(" rhs %.17g\n", rhs[i]);
  }
  return 0;
}

// ANA f0 files as the reference writes them (ana_fzwrite, cmlib/
// ana_write.cc:120-171, through write_file, cmlib/fileio.cc:273-330):
//   ana <out_prefix>
// writes <out_prefix>_3d.f0 (write_file of a 3-D array, FLOAT32, the
// nodes-cube layout), <out_prefix>_4d.f0 (write_file of a 4-D array,
// FLOAT64, the spectra-cube layout) with seeded values, and prints the values
// in the reference's index order.
static int mode_ana(int argc, char **argv, io_class &io) {
  if (argc < 3) return 2;
  g_state = 12345ULL * 2654435761ULL + 88172645463325252ULL;
  const int n1 = 3, n2 = 4, n3 = 5;
  fp_t ***a = ft3dim(1, n1, 1, n2, 1, n3);
  for (int i = 1; i <= n1; ++i)
    for (int j = 1; j <= n2; ++j)
      for (int kk = 1; kk <= n3; ++kk) {
        a[i][j][kk] = 1000.0 * urand() - 500.0;
        printf("a3 %d %d %d %.9g\n", i, j, kk, (double)(float)a[i][j][kk]);
      }
  // write_file's 3-D layout (fileio.cc:273-286; fileio.cc itself needs
  // cfitsio and is not built): float, dims (n1, n2, n3), n1 fastest
  char name[1024];
  snprintf(name, sizeof(name), "%s_3d.f0", argv[2]);
  {
    int ds[4] = {0, n1, n2, n3};
    float *q = new float[n1 * n2 * n3];
    for (int i = 1; i <= n1; ++i)
      for (int j = 1; j <= n2; ++j)
        for (int kk = 1; kk <= n3; ++kk) q[((kk - 1) * n2 + (j - 1)) * n1 + (i - 1)] = (float)a[i][j][kk];
    ana_fzwrite((byte *)q, name, ds, 3, 0, FLOAT32, io);
    delete[] q;
  }
  const int m1 = 2, m2 = 3, m3 = 4, m4 = 6;
  fp_t ****b = ft4dim(1, m1, 1, m2, 1, m3, 1, m4);
  for (int i = 1; i <= m1; ++i)
    for (int j = 1; j <= m2; ++j)
      for (int kk = 1; kk <= m3; ++kk)
        for (int l = 1; l <= m4; ++l) {
          b[i][j][kk][l] = urand() * 1e14;
          printf("a4 %d %d %d %d %.17g\n", i, j, kk, l, b[i][j][kk][l]);
        }
  // write_file's 4-D double layout (fileio.cc:320-330): dims (m4, m3, m2, m1),
  // the array as it lies in memory
  snprintf(name, sizeof(name), "%s_4d.f0", argv[2]);
  {
    int ds[5] = {0, m4, m3, m2, m1};
    ana_fzwrite((byte *)(b[1][1][1] + 1), name, ds, 4, 0, FLOAT64, io);
  }
  return 0;
}

// Chemical equilibrium (atmos/atmos_chm.cc:28-363) and LTE level populations
// (atmos/atmol/atom/atom.cc:763-787) on every depth point of the column:
//   chemeq <atmos.cfg>
// The species table (elements with ionisation energies, partition-function
// parameters and levels; molecules with components and dissociation-constant
// forms) is written first, then per depth: T, Nt, the incoming Ne, partition
// functions and ionisation fractions at that Ne, dissociation constants, and
// after atmosphere::ltepops the final Ne, stage and level populations and the
// H- density.
static int mode_chemeq(int argc, char **argv, io_class &io) {
  if (argc < 3) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  const int x1 = a->x1l, x2 = a->x2l;
  // the hydrostatic integration's result when make_atmos ran it
  // (atmos_pp::enforce_hequilibrium, plpar/atmos_pp.cc:372-455)
  if (getenv("SNAPI_REF_HE") && atoi(getenv("SNAPI_REF_HE")))
    for (int x3i = a->x3l; x3i <= a->x3h; ++x3i)
      printf("he %d %.17g %.17g %.17g %.17g %.17g %.17g\n", x3i, a->T[x1][x2][x3i],
             log10(-a->tau_referent[x1][x2][x3i]), a->Nt[x1][x2][x3i], a->Ne[x1][x2][x3i],
             a->op_referent[x1][x2][x3i], a->rho[x1][x2][x3i]);
  a->popsetup();
  atom **el = new atom *[a->natm];
  int nel = 0;
  for (int s = 0; s < a->natm; ++s) {
    int nc;
    if (!a->atml[s]->components(nc)) el[nel++] = dynamic_cast<atom *>(a->atml[s]);
  }
  printf("nd %d\nnelement %d\n", a->x3h - a->x3l + 1, nel);
  for (int ie = 0; ie < nel; ++ie) {
    atom *t = el[ie];
    printf("element %d %d %.17g\n", ie + 1, (int)t->Z, t->abund);
    for (int z = 0; z < t->Z; ++z) printf("eion %d %d %.17g\n", ie + 1, z, t->ip[z]);
    for (int z = 0; z <= t->Z; ++z) {
      pf *p = t->partf[z];
      printf("pf %d %d %d", ie + 1, z, (int)p->type);
      if (p->type == PF_TYPE_CONST) printf(" %.17g", ((cpf *)p)->val);
      if (p->type == PF_TYPE_IRWIN) {
        irwin *w = (irwin *)p;
        printf(" %d", w->n);
        for (int i = 0; i < w->n; ++i) printf(" %.17g", w->a[i]);
      }
      if (p->type == PF_TYPE_TRAV) {
        trav *w = (trav *)p;
        printf(" %.17g %.17g %d", w->z, w->g0, w->n);
      }
      printf("\n");
      if (p->type == PF_TYPE_TRAV) {
        trav *w = (trav *)p;
        for (int i = 0; i < w->n; ++i) {
          printf("term %d %d %d %.17g %.17g %.17g %d %d", ie + 1, z, i, w->g2[i], w->ee[i], w->ll[i], (int)w->asym[i],
                 w->m[i]);
          for (int l = 0; l < w->m[i]; ++l) printf(" %.17g %.17g", w->g[i][l], w->a[i][l]);
          printf("\n");
        }
      }
      for (int l = 0; l < t->nl[z]; ++l)
        printf("lvl %d %d %d %d %.17g\n", ie + 1, z, l, (int)t->g[z][l], t->ee[z][l]);
    }
    // bound-free cross-sections at the referent wavelength (atom::opacity_continuum,
    // atmos/atmol/atom/atom_rad.cc:151-187)
    const fp_t lr = airtovac(lambda_referent);
    for (int z = 0; z < t->Z; ++z)
      for (int l = 0; l < t->nl[z]; ++l) {
        const fp_t sg = (t->bf[z][l]) ? t->bf[z][l]->U(lr) : 0.0, sg0 = (t->bf[z][l]) ? t->bf[z][l]->U(lambda_referent) : 0.0;
        if (sg != 0.0 || sg0 != 0.0)
          printf("bf %d %d %d %d %.17g %d %.17g %.17g %.17g\n", ie + 1, z, l, (int)t->g[z][l], t->ee[z][l],
                 (int)t->g[z + 1][0], t->ee[z + 1][0], sg, sg0);
      }
  }
  printf("lref %.17g %.17g\n", airtovac(lambda_referent), (fp_t)lambda_referent);
  int nmol = 0;
  for (int s = 0; s < a->natm; ++s) {
    int nc;
    ion *cmp = a->atml[s]->components(nc);
    if (!cmp) continue;
    mol *m = (mol *)a->atml[s];
    eqc *K = m->K[0];
    printf("mol %d %d %d", ++nmol, nc, (int)K->type);
    if (K->type == EQC_TYPE_CNST) printf(" 1 %.17g", ((eqc_cnst *)K)->coef);
    if (K->type == EQC_TYPE_POLY) {
      eqc_poly *pk = (eqc_poly *)K;
      printf(" %d", (int)pk->nc);
      for (int i = 0; i < pk->nc; ++i) printf(" %.17g", pk->cfs[i]);
    }
    if (K->type == EQC_TYPE_HMIN) printf(" 0");
    printf("\n");
    for (int i = 0; i < nc; ++i) {
      int idx = 0;
      for (int ie = 0; ie < nel; ++ie)
        if (cmp[i].atom == el[ie]) idx = ie + 1;
      printf("comp %d %d %d %d\n", nmol, i, idx, (int)cmp[i].z);
    }
  }
  fp_t *fk = 0, *df = 0;
  int nk = 0;
  for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) {
    const fp_t T = a->T[x1][x2][x3i], Nt = a->Nt[x1][x2][x3i], ne = a->Ne[x1][x2][x3i];
    printf("point %d %.17g %.17g %.17g\n", x3i, T, Nt, ne);
    for (int ie = 0; ie < nel; ++ie) {
      atom *t = el[ie];
      for (int z = 0; z <= t->Z; ++z)
        printf("U %d %d %d %.17g %.17g\n", x3i, ie + 1, z, t->partf[z]->U(T, ne, io), t->partf[z]->dU(T, ne, io));
      t->ionfrc(T, ne, fk, df, nk);
      for (int z = 0; z < nk; ++z) printf("frc %d %d %d %.17g %.17g\n", x3i, ie + 1, z, fk[z], df[z]);
      delete[] fk;
      delete[] df;
      fk = df = 0;
    }
    int m = 0;
    for (int s = 0; s < a->natm; ++s) {
      int nc;
      if (a->atml[s]->components(nc)) printf("K %d %d %.17g\n", x3i, ++m, ((mol *)a->atml[s])->dissoc(T, ne));
    }
  }
  a->ltepops();
  for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) {
    printf("ne %d %.17g\n", x3i, a->Ne[x1][x2][x3i]);
    for (int ie = 0; ie < nel; ++ie) {
      atom *t = el[ie];
      pps &p = t->pop[x1][x2][x3i];
      printf("Nt %d %d %.17g\n", x3i, ie + 1, p.Nt);
      for (int z = 0; z <= t->Z; ++z) {
        printf("N %d %d %d %.17g\n", x3i, ie + 1, z, p.N[z]);
        for (int l = 0; l < t->nl[z]; ++l) printf("n %d %d %d %d %.17g\n", x3i, ie + 1, z, l, p.n[z][l]);
      }
    }
    for (int s = 0; s < a->natm; ++s) {
      h_minus_mol *hm = dynamic_cast<h_minus_mol *>(a->atml[s]);
      if (hm) printf("hmin %d %.17g\n", x3i, hm->N[x1][x2][x3i]);
    }
  }
  // referent opacity after chemeq + LTE at every depth (atmos_ppbez::compute_op_referent,
  // plpar/atmos_ppbez_rts.cc:268-281)
  a->compute_op_referent();
  for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) printf("opref %d %.17g\n", x3i, a->op_referent[x1][x2][x3i]);
  return 0;
}

// NLTE population responses of the active atom (atmosphere::
// compute_nlte_population_responses, atmos/atmos_responses.cc:13-157, as
// obs_scalar_responses prepares it, atmos_obs.cc:168-181):
//   nlteresp <atmos.cfg> [every]
// First the atom's assembled linear system, its three right-hand sides per
// perturbed depth and the level responses it solves for; then, at every
// `every`-th wavelength of its NLTE grid and for both rays, the inputs and
// outputs of atmos_ppbez::formal_with_responses (plpar/atmos_ppbez_rts.cc:
// 349-778): normalised opacity and emissivity, intensity and the responses
// of every intensity to every opacity and emissivity.
static int mode_nlteresp(int argc, char **argv, io_class &io) {
  if (argc < 3) return 2;
  const int every = (argc > 3) ? atoi(argv[3]) : 40;
  atmosphere *a = make_atmos(argv[2], io);
  const int x1 = a->x1l, x2 = a->x2l;
  a->set_grid(a->tau_grid);
  a->boundary_condition_for_rt = -1;
  a->popsetup();
  for (int s = 0; s < a->natm; ++s) a->atml[s]->set_parent_atmosphere(a);
  if (a->tau_grid) a->compute_op_referent_derivative();
  a->nltepops();
  a->respsetup();
  a->ne_lte_derivatives();
  a->compute_nlte_population_responses(0);
  atom *at = 0;
  for (int s = 0; s < a->natm; ++s) {
    atom *t = dynamic_cast<atom *>(a->atml[s]);
    if (t && t->NLTE) at = t;
  }
  if (!at) return 3;
  const int nd = a->x3h - a->x3l + 1, nm = at->nmap, n = nd * nm;
  printf("nd %d\nnmap %d\ntau_grid %d\nboundary %.17g\n", nd, nm, a->tau_grid, a->boundary_condition_for_rt);
  for (int r = 1; r <= n; ++r)
    for (int cc = 1; cc <= n; ++cc)
      if (at->response_matrix[r][cc] != 0.0) printf("M %d %d %.17g\n", r - 1, cc - 1, at->response_matrix[r][cc]);
  for (int kk = 1; kk <= nd; ++kk)
    for (int r = 1; r <= n; ++r)
      printf("beta %d %d %.17g %.17g %.17g\n", kk - 1, r - 1, at->beta_Temp[kk][r], at->beta_density[kk][r],
             at->beta_v_micro[kk][r]);
  for (int p = 1; p <= 3; ++p)
    for (int r = 1; r <= n; ++r)
      for (int kk = 1; kk <= nd; ++kk) printf("resp %d %d %d %.17g\n", p, r - 1, kk - 1, at->level_responses[p][r][kk]);
  // formal_with_responses at selected wavelengths of the NLTE grid
  fp_t *lambda = 0;
  int32_t nlambda = 0;
  for (int s = 0; s < a->natm; ++s) a->atml[s]->rtsetup(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  for (int s = 0; s < a->natm; ++s)
    lambda = a->atml[s]->getlambda(lambda, nlambda, a->T[x1][x2][a->x3h], a->Nt[x1][x2][a->x3h], a->Ne[x1][x2][a->x3h]);
  for (int s = 0; s < a->natm; ++s) a->atml[s]->rtinit();
  int ntp;
  fp_t *th, *ph;
  a->anglesetup(th, ph, ntp);
  fp_t ***S = ft3dim(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t **Rop = ft2dim(a->x3l, a->x3h, a->x3l, a->x3h);
  fp_t **Rem = ft2dim(a->x3l, a->x3h, a->x3l, a->x3h);
  for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) printf("grid %d %.17g\n", x3i - a->x3l, a->rt_grid[x3i]);
  int prob = 0;
  for (int tp = 1; tp <= ntp; ++tp) {
    fp_t ***Vr = a->project(a->Vx, a->Vy, a->Vz, th[tp], ph[tp], a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    fp_t ****Bv = a->transform(a->Bx, a->By, a->Bz, th[tp], ph[tp], a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    for (int l = 0; l < nlambda; l += every) {
      for (int s = 0; s < a->natm; ++s) a->atml[s]->prof_init();
      fp_t ***op = a->opacity(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
      fp_t ***em = a->emissivity(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
      if (a->tau_grid) a->normalize_to_referent_opacity(op, em);
      a->formal_with_responses(a->rt_grid, S, 0, Rop, Rem, op, em, th[tp], ph[tp], a->boundary_condition_for_rt);
      printf("prob %d %d %d %.17g %.17g\n", prob, tp, l, th[tp], lambda[l]);
      for (int x3i = a->x3l; x3i <= a->x3h; ++x3i)
        printf("opem %d %d %.17g %.17g %.17g\n", prob, x3i - a->x3l, op[x1][x2][x3i], em[x1][x2][x3i], S[x1][x2][x3i]);
      for (int z = a->x3l; z <= a->x3h; ++z)
        for (int zp = a->x3l; zp <= a->x3h; ++zp)
          printf("R %d %d %d %.17g %.17g\n", prob, z - a->x3l, zp - a->x3l, Rop[z][zp], Rem[z][zp]);
      del_ft3dim(em, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      del_ft3dim(op, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      ++prob;
    }
    del_ft4dim(Bv, 1, 3, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    del_ft3dim(Vr, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  }
  return 0;
}

// nlterin <atmos.cfg>: the inputs of atom::add_response_contributions_new and
// of the rate part of atom::compute_nlte_population_responses
// (atmos/atmol/atom/atom_responses.cc:26-200, 668-1128), taken as
// atmosphere::compute_nlte_population_responses (atmos/atmos_responses.cc:
// 13-157) forms them after nltepops: per depth the state, the electron-density
// and referent-opacity derivatives, the derivatives of the atom's total
// population; per level the populations; per line the damping and its
// temperature derivative; per level pair the collision rates and their full
// derivatives; per ray and wavelength the projected velocity, the total
// opacity and emissivity and the LTE background's perturbations.  At every
// 40th wavelength also the atom's profiles and profile norms for checking.
static int mode_nlterin(int argc, char **argv, io_class &io) {
  if (argc < 3) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  const int x1 = a->x1l, x2 = a->x2l;
  a->set_grid(a->tau_grid);
  a->boundary_condition_for_rt = -1;
  a->popsetup();
  for (int s = 0; s < a->natm; ++s) a->atml[s]->set_parent_atmosphere(a);
  if (a->tau_grid) a->compute_op_referent_derivative();
  a->nltepops();
  a->respsetup();
  a->ne_lte_derivatives();
  atom *at = 0;
  for (int s = 0; s < a->natm; ++s) {
    atom *t = dynamic_cast<atom *>(a->atml[s]);
    if (t && t->NLTE) at = t;
  }
  if (!at) return 3;
  fp_t *lambda = 0;
  int32_t nlambda = 0;
  for (int s = 0; s < a->natm; ++s) a->atml[s]->rtsetup(a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  for (int s = 0; s < a->natm; ++s)
    lambda = a->atml[s]->getlambda(lambda, nlambda, a->T[x1][x2][a->x3h], a->Nt[x1][x2][a->x3h], a->Ne[x1][x2][a->x3h]);
  for (int s = 0; s < a->natm; ++s) a->atml[s]->rtinit();
  // as atmos_responses.cc:35-44: the LTE atoms' population responses enter
  // their opacity perturbations
  for (int s = 0; s < a->natm; ++s)
    if (a->atml[s]->check_if_nlte()) a->atml[s]->responses_init();
  for (int s = 0; s < a->natm; ++s)
    if (!a->atml[s]->check_if_nlte()) {
      a->atml[s]->responses_init();
      a->atml[s]->compute_lte_population_responses_analytical(a->T, a->Ne);
    }
  int ntp;
  fp_t *th, *ph;
  fp_t *bin = a->anglesetup(th, ph, ntp);
  fp_t *lambda_w = new fp_t[nlambda];
  if (nlambda >= 2) {
    lambda_w[0] = (lambda[1] - lambda[0]) * 0.5;
    lambda_w[nlambda - 1] = (lambda[nlambda - 1] - lambda[nlambda - 2]) * 0.5;
    for (int l = 1; l < nlambda - 1; ++l) lambda_w[l] = (lambda[l + 1] - lambda[l - 1]) * 0.5;
  } else if (nlambda)
    lambda_w[0] = 1.0;
  const int nd = a->x3h - a->x3l + 1;
  printf("nd %d\nntp %d\nnlam %d\nboundary %.17g\ntau_grid %d\n", nd, ntp, nlambda, a->boundary_condition_for_rt,
         a->tau_grid);
  printf("atom nmap %d ntr %d Z %d mass %.17g saha %.17g dT %.17g dN %.17g damp %d\n", at->nmap, at->ntr, (int)at->Z,
         at->mass, saha_const, delta_T, delta_Nt_frac, (int)turn_on_damping);
  for (int tp = 1; tp <= ntp; ++tp) printf("ray %d %.17g %.17g\n", tp, th[tp], bin[tp]);
  for (int l = 0; l < nlambda; ++l) printf("lam %d %.17g %.17g\n", l, lambda[l], lambda_w[l]);
  for (int i = 0; i < at->nmap; ++i) {
    const int z = at->zmap[i], lv = at->lmap[i];
    printf("level %d %d %d %.17g %d %.17g\n", i, z, lv, at->ee[z][lv], (int)at->g[z][lv], z < at->Z ? at->ip[z] : 0.0);
  }
  for (int tr = 1; tr <= (int)at->ntr; ++tr) {
    const int z = at->inverse_tmap[tr][2], lo = at->inverse_tmap[tr][3], up = at->inverse_tmap[tr][4];
    if (up < at->nl[z]) {
      printf("trans %d line %d %d %d %d %.17g %.17g %.17g\n", tr, z, lo, up, (int)at->rmap[z][up], at->A[z][up][lo],
             at->B[z][lo][up], at->B[z][up][lo]);
    } else {
      printf("trans %d bf %d %d %d %d %.17g\n", tr, z, lo, up, (int)at->rmap[z + 1][0],
             at->bf[z][lo] ? at->bf[z][lo]->getlambda_crit() : 0.0);
      for (int l = 0; l < nlambda; ++l)
        printf("sigma %d %d %.17g\n", tr, l, at->bf[z][lo] ? at->bf[z][lo]->U(lambda[l]) : 0.0);
    }
  }
  for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) {
    const fp_t T = a->T[x1][x2][x3i], ne = a->Ne[x1][x2][x3i];
    const fp_t dT = at->derivative_active_population(x1, x2, x3i);
    const fp_t dN = at->derivative_active_population_density(x1, x2, x3i);
    printf("depth %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", x3i - a->x3l,
           T, a->Vt[x1][x2][x3i], ne, a->Nt[x1][x2][x3i], a->tau_grid ? a->op_referent[x1][x2][x3i] : 1.0,
           a->rt_grid[x3i], a->get_ne_lte_derivative(1, x1, x2, x3i), a->get_ne_lte_derivative(2, x1, x2, x3i),
           a->tau_grid ? a->get_op_referent_der(1, x3i, x1, x2, x3i) : 0.0,
           a->tau_grid ? a->get_op_referent_der(2, x3i, x1, x2, x3i) : 0.0,
           a->tau_grid ? a->get_op_referent_der(3, x3i, x1, x2, x3i) : 0.0, dT, dN,
           at->fetch_population(x1, x2, x3i, 0, 0));
    for (int i = 0; i < at->nmap; ++i)
      printf("pop %d %d %.17g\n", x3i - a->x3l, i, at->pop[x1][x2][x3i].n[at->zmap[i]][at->lmap[i]]);
    for (int tr = 1; tr <= (int)at->ntr; ++tr) {
      const int z = at->inverse_tmap[tr][2], lo = at->inverse_tmap[tr][3], up = at->inverse_tmap[tr][4];
      if (up < at->nl[z]) {
        const fp_t lam0 = fabs(h * c / (at->ee[z][up] - at->ee[z][lo]));
        printf("damp %d %d %.17g %.17g %.17g\n", tr, x3i - a->x3l, at->damp_rad(at->A[z], lo, up),
               at->damp_col(x1, x2, x3i, z, up, lo, T, ne, lam0), at->damp_col_der_T(x1, x2, x3i, z, up, lo, T, ne, lam0));
      }
    }
    for (int i = 0; i < at->nmap; ++i)
      for (int j = 0; j < at->nmap; ++j)
        printf("coll %d %d %d %.17g %.17g %.17g\n", x3i - a->x3l, i, j, at->collisional_rates(x1, x2, x3i, i, j),
               at->derivative_collisions_full_temp(x1, x2, x3i, i, j),
               at->derivative_collisions_full_density(x1, x2, x3i, i, j));
  }
  // per ray and wavelength, in the order of the responses' loop
  for (int tp = 1; tp <= ntp; ++tp) {
    fp_t ***Vr = a->project(a->Vx, a->Vy, a->Vz, th[tp], ph[tp], a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    fp_t ****Bv = a->transform(a->Bx, a->By, a->Bz, th[tp], ph[tp], a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    for (int x3i = a->x3l; x3i <= a->x3h; ++x3i) printf("vr %d %d %.17g\n", tp, x3i - a->x3l, Vr[x1][x2][x3i]);
    at->compute_profile_norm(th[tp], ph[tp], lambda, lambda_w, Vr, nlambda);
    for (int x3i = a->x3l; x3i <= a->x3h; ++x3i)
      for (int tr = 1; tr <= (int)at->ntr; ++tr)
        printf("norm %d %d %d %.17g %.17g %.17g %.17g\n", tp, x3i - a->x3l, tr, at->norm[x1][x2][x3i][tr],
               at->norm_derivative[1][x1][x2][x3i][tr], at->norm_derivative[2][x1][x2][x3i][tr],
               at->norm_derivative[3][x1][x2][x3i][tr]);
    for (int l = 0; l < nlambda; ++l) {
      for (int s = 0; s < a->natm; ++s) a->atml[s]->prof_init();
      fp_t ***op = a->opacity(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
      fp_t ***em = a->emissivity(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
      fp_t *****opp = a->opacity_pert_lte(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
      fp_t *****emp = a->emissivity_pert_lte(a->T, a->Ne, Vr, a->Vt, Bv, th[tp], ph[tp], lambda[l]);
      for (int x3i = a->x3l; x3i <= a->x3h; ++x3i)
        printf("pp %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", tp, l, x3i - a->x3l,
               op[x1][x2][x3i], em[x1][x2][x3i], opp[1][x3i][x1][x2][x3i], opp[2][x3i][x1][x2][x3i],
               opp[3][x3i][x1][x2][x3i], emp[1][x3i][x1][x2][x3i], emp[2][x3i][x1][x2][x3i], emp[3][x3i][x1][x2][x3i]);
      if (l % 40 == 0)
        for (int x3i = a->x3l; x3i <= a->x3h; ++x3i)
          for (int tr = 1; tr <= (int)at->ntr; ++tr)
            printf("prof %d %d %d %d %.17g\n", tp, l, x3i - a->x3l, tr, at->current_profile[x1][x2][x3i][tr]);
      del_ft3dim(em, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      del_ft3dim(op, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      del_ft5dim(opp, 1, 7, a->x3l, a->x3h, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
      del_ft5dim(emp, 1, 7, a->x3l, a->x3h, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    }
    del_ft4dim(Bv, 1, 3, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
    del_ft3dim(Vr, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  }
  // the system itself, for the record
  a->compute_nlte_population_responses(0);
  const int n = nd * at->nmap;
  for (int r = 1; r <= n; ++r)
    for (int cc = 1; cc <= n; ++cc)
      if (at->response_matrix[r][cc] != 0.0) printf("M %d %d %.17g\n", r - 1, cc - 1, at->response_matrix[r][cc]);
  for (int kk = 1; kk <= nd; ++kk)
    for (int r = 1; r <= n; ++r)
      printf("beta %d %d %.17g %.17g %.17g\n", kk - 1, r - 1, at->beta_Temp[kk][r], at->beta_density[kk][r],
             at->beta_v_micro[kk][r]);
  return 0;
}

// Node model from a text stream: "tau_min tau_max", the node count, then one
// "code logtau value" line per node (code 1 T, 2 vt, 3 vLOS, 4 |B|,
// 5 inclination [rad], 6 azimuth [rad]), grouped in that order as the
// reference's model holds them (atmos/atmos.h:377-531), then the count of
// regularisation entries and one "code type alpha" line each (the model's
// *_reg_type and *_reg_alpha; the rest off).
static model *read_model(FILE *f) {
  double tmin, tmax;
  int n;
  if (fscanf(f, "%lf %lf %d", &tmin, &tmax, &n) != 3 || n < 1) { fprintf(stderr, "bad model header\n"); exit(2); }
  int cnt[7] = {0, 0, 0, 0, 0, 0, 0};
  int *code = new int[n];
  double *lt = new double[n], *val = new double[n];
  for (int i = 0; i < n; ++i) {
    if (fscanf(f, "%d %lf %lf", &code[i], &lt[i], &val[i]) != 3 || code[i] < 1 || code[i] > 6) {
      fprintf(stderr, "bad model node\n");
      exit(2);
    }
    if (i && code[i] < code[i - 1]) { fprintf(stderr, "model nodes must be grouped by parameter\n"); exit(2); }
    cnt[code[i]]++;
  }
  model *m = model_new(cnt[1], cnt[2], cnt[3], cnt[4], cnt[5], cnt[6]);
  m->set_tau_min(tmin);
  m->set_tau_max(tmax);
  int idx = 0;
  for (int pc = 1; pc <= 6; ++pc) {
    if (!cnt[pc]) continue;
    fp_t *t = new fp_t[cnt[pc]] - 1, *v = new fp_t[cnt[pc]] - 1;
    for (int j = 1; j <= cnt[pc]; ++j, ++idx) {
      t[j] = lt[idx];
      v[j] = val[idx];
    }
    switch (pc) {
      case 1: m->set_temp_nodes(t, v); break;
      case 2: m->set_vt_nodes(t, v); break;
      case 3: m->set_vs_nodes(t, v); break;
      case 4: m->set_B_nodes(t, v); break;
      case 5: m->set_theta_nodes(t, v); break;
      case 6: m->set_phi_nodes(t, v); break;
    }
    delete[](t + 1);
    delete[](v + 1);
  }
  int rtype[7] = {0, 0, 0, 0, 0, 0, 0};
  double ralpha[7] = {0, 0, 0, 0, 0, 0, 0};
  int nreg = 0;
  if (fscanf(f, "%d", &nreg) != 1 || nreg < 0 || nreg > 6) { fprintf(stderr, "bad regularisation count\n"); exit(2); }
  for (int i = 0; i < nreg; ++i) {
    int pc, t;
    double al;
    if (fscanf(f, "%d %d %lf", &pc, &t, &al) != 3 || pc < 1 || pc > 6) {
      fprintf(stderr, "bad regularisation entry\n");
      exit(2);
    }
    rtype[pc] = t;
    ralpha[pc] = al;
  }
  m->set_temp_reg(rtype[1], ralpha[1]);
  m->set_vt_reg(rtype[2], ralpha[2]);
  m->set_vs_reg(rtype[3], ralpha[3]);
  m->set_B_reg(rtype[4], ralpha[4]);
  m->set_theta_reg(rtype[5], ralpha[5]);
  m->set_phi_reg(rtype[6], ralpha[6]);
  delete[] code;
  delete[] lt;
  delete[] val;
  return m;
}

static void print_nodes(const char *tag, int step, model *m) {
  printf("%s %d", tag, step);
  for (int i = 1; i <= m->get_N_nodes_total(); ++i) printf(" %.17g", m->get_parameter(i));
  printf("\n");
}

static int mode_nodes(int argc, char **argv, io_class &io) {
  if (argc < 7) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  int nlambda;
  fp_t *lambda = read_lambda(argv[3], nlambda);
  fp_t el = atof(argv[4]), az = atof(argv[5]);
  FILE *f = fopen(argv[6], "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", argv[6]); return 2; }
  model *m = read_model(f);
  fclose(f);
  int x1 = a->x1l, x2 = a->x2l;
  const int np = m->get_N_nodes_total();
  a->set_grid(1);
  a->build_from_nodes(m);
  printf("nd %d np %d nlambda %d\n", a->x3h - a->x3l + 1, np, nlambda);
  fp_t ***Vr = a->project(a->Vx, a->Vy, a->Vz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  fp_t ****Bv = a->transform(a->Bx, a->By, a->Bz, el, az, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  for (int z = a->x3l; z <= a->x3h; ++z) {
    printf("bfn %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", z - a->x3l,
           a->tau_referent[x1][x2][z], a->T[x1][x2][z], a->Nt[x1][x2][z], a->Ne[x1][x2][z],
           a->op_referent[x1][x2][z], a->rho[x1][x2][z], a->Vt[x1][x2][z], a->Vz[x1][x2][z], a->Bx[x1][x2][z],
           a->By[x1][x2][z], a->Bz[x1][x2][z]);
    printf("proj %d %.17g %.17g %.17g %.17g\n", z - a->x3l, Vr[x1][x2][z], Bv[1][x1][x2][z], Bv[2][x1][x2][z],
           Bv[3][x1][x2][z]);
  }
  del_ft4dim(Bv, 1, 3, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  del_ft3dim(Vr, a->x1l, a->x1h, a->x2l, a->x2h, a->x3l, a->x3h);
  observable *fw = a->forward_evaluate(el, az, lambda, nlambda, 0.0, 1.0, 0.0, 0, 0);
  fp_t **Sf = fw->get_S(1, 1);
  for (int l = 1; l <= nlambda; ++l) printf("fwd %d %.17g %.17g %.17g %.17g\n", l, Sf[1][l], Sf[2][l], Sf[3][l], Sf[4][l]);
  del_ft2dim(Sf, 1, 4, 1, nlambda);
  delete fw;
  fp_t ***J = ft3dim(1, np, 1, nlambda, 1, 4);
  observable *o = a->obs_stokes_responses_to_nodes(m, el, az, lambda, nlambda, J, 0);
  fp_t **Sv = o->get_S(1, 1);
  for (int l = 1; l <= nlambda; ++l)
    printf("stokes %d %.17g %.17g %.17g %.17g\n", l, Sv[1][l], Sv[2][l], Sv[3][l], Sv[4][l]);
  for (int i = 1; i <= np; ++i)
    for (int l = 1; l <= nlambda; ++l)
      printf("jac %d %d %.17g %.17g %.17g %.17g\n", i, l, J[i][l][1], J[i][l][2], J[i][l][3], J[i][l][4]);
  // the state the response path left behind (T raised by delta_T/2 by
  // ne_lte_derivatives, atmos/atmos_responses.cc:364-423)
  for (int z = a->x3l; z <= a->x3h; ++z)
    printf("post %d %.17g %.17g %.17g %.17g\n", z - a->x3l, a->T[x1][x2][z], a->Nt[x1][x2][z], a->Ne[x1][x2][z],
           a->op_referent[x1][x2][z]);
  fp_t ***NR = m->get_response_to_parameters();
  const int nd = a->x3h - a->x3l + 1;
  for (int i = 1; i <= np; ++i)
    for (int q = 1; q <= 7; ++q)
      for (int z = 1; z <= nd; ++z)
        if (NR[i][q][z] != 0.0) printf("nmap %d %d %d %.17g\n", i, q, z - 1, NR[i][q][z]);
  print_nodes("model", 0, m);
  return 0;
}

static int mode_fit(int argc, char **argv, io_class &io) {
  if (argc < 7) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  int nlambda;
  fp_t *lambda = read_lambda(argv[3], nlambda);
  fp_t el = atof(argv[4]), az = atof(argv[5]);
  FILE *f = fopen(argv[6], "r");
  if (!f) { fprintf(stderr, "cannot open %s\n", argv[6]); return 2; }
  model *start = read_model(f);
  fp_t ws[4], lm_start;
  int kmax;
  if (fscanf(f, "%lf %lf %lf %lf %lf %d", &ws[0], &ws[1], &ws[2], &ws[3], &lm_start, &kmax) != 6) return 2;
  fp_t *mask = new fp_t[nlambda] - 1;
  for (int l = 1; l <= nlambda; ++l)
    if (fscanf(f, "%lf", &mask[l]) != 1) return 2;
  fp_t ****S = ft4dim(1, 1, 1, 1, 1, 4, 1, nlambda);
  for (int l = 1; l <= nlambda; ++l)
    for (int s = 1; s <= 4; ++s)
      if (fscanf(f, "%lf", &S[1][1][s][l]) != 1) return 2;
  fclose(f);
  const int np = start->get_N_nodes_total();
  printf("np %d nlambda %d kmax %d\n", np, nlambda, kmax);
  print_nodes("start", 0, start);
  // a seventh argument "last": only the run with kmax iterations
  const int first = (argc > 7 && !strcmp(argv[7], "last")) ? kmax : 1;
  for (int it = first; it <= kmax; ++it) {
    model *m = clone(start);
    observable *o = new observable(1, 1, 4, nlambda);
    o->set(S);
    o->set_lambda(lambda);
    o->set_mask(mask);
    o->set_w_stokes(ws);
    o->set_inv_parameters(0.0, 0.0, 1.0, 1.0);
    o->set_no_iterations(it);
    o->set_start_lambda(lm_start);
    o->set_stopping_chisq(0.0);
    // the fit reports the current chi^2 of every iteration on stderr; the
    // marker tells the runs apart
    fprintf(stderr, "fitrun %d\n", it);
    fflush(stderr);
    observable *r = a->stokes_lm_fit(o, el, az, m);
    fflush(stderr);
    fp_t **Sr = r->get_S(1, 1);
    fp_t chi = 0.0;
    for (int l = 1; l <= nlambda; ++l)
      for (int s = 1; s <= 4; ++s) {
        fp_t d = (S[1][1][s][l] - Sr[s][l]) * ws[s - 1] * mask[l];
        chi += d * d;
      }
    print_nodes("fitk", it, m);
    printf("fitchi %d %.17g\n", it, chi);
    if (it == kmax)
      for (int l = 1; l <= nlambda; ++l)
        printf("fitS %d %.17g %.17g %.17g %.17g\n", l, Sr[1][l], Sr[2][l], Sr[3][l], Sr[4][l]);
    fflush(stdout);
    del_ft2dim(Sr, 1, 4, 1, nlambda);
    delete r;
    delete o;
    delete m;
  }
  return 0;
}

// time_synth <atmos.cfg> <lambda.txt> <el> <az> <reps>: wall time of
// atmosphere::obs_stokes (atmos/atmos_obs.cc:16-87) on the column, one call
// untimed, then reps timed calls.
static int mode_time_synth(int argc, char **argv, io_class &io) {
  if (argc < 7) return 2;
  atmosphere *a = make_atmos(argv[2], io);
  int nlambda;
  fp_t *lambda = read_lambda(argv[3], nlambda);
  fp_t el = atof(argv[4]), az = atof(argv[5]);
  const int reps = atoi(argv[6]);
  delete a->obs_stokes(el, az, lambda, nlambda);
  timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int r = 0; r < reps; ++r) delete a->obs_stokes(el, az, lambda, nlambda);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  const double sec = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
  printf("time_synth nd %d nlambda %d reps %d seconds_per_call %.6f lambda_points_per_s %.6g\n",
         a->x3h - a->x3l + 1, nlambda, reps, sec / reps, reps * (double)nlambda / sec);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: ref_driver synth|resp|lmstep|formal_scalar ...\n");
    return 2;
  }
  io_class io;
  io.reconf(0);
  if (!strcmp(argv[1], "synth")) return mode_synth(argc, argv, io);
  if (!strcmp(argv[1], "resp")) return mode_resp(argc, argv, io);
  if (!strcmp(argv[1], "lmstep")) return mode_lmstep(argc, argv, io);
  if (!strcmp(argv[1], "formal_scalar")) return mode_formal_scalar(argc, argv, io);
  if (!strcmp(argv[1], "interp")) return mode_interp(argc, argv, io);
  if (!strcmp(argv[1], "conv")) return mode_conv(argc, argv, io);
  if (!strcmp(argv[1], "ana")) return mode_ana(argc, argv, io);
  if (!strcmp(argv[1], "nlte")) return mode_nlte(argc, argv, io);
  if (!strcmp(argv[1], "chemeq")) return mode_chemeq(argc, argv, io);
  if (!strcmp(argv[1], "nlteresp")) return mode_nlteresp(argc, argv, io);
  if (!strcmp(argv[1], "nodes")) return mode_nodes(argc, argv, io);
  if (!strcmp(argv[1], "fit")) return mode_fit(argc, argv, io);
  if (!strcmp(argv[1], "psf")) return mode_psf(argc, argv, io);
  if (!strcmp(argv[1], "nlterin")) return mode_nlterin(argc, argv, io);
  if (!strcmp(argv[1], "reg")) return mode_reg(argc, argv, io);
  if (!strcmp(argv[1], "time_synth")) return mode_time_synth(argc, argv, io);
  fprintf(stderr, "unknown mode %s\n", argv[1]);
  return 2;
}
