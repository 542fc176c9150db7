/*
 * FP64 CPU restatement of the assembly of the reference's NLTE
 * population-response system -- TEST INFRASTRUCTURE ONLY (see snapi_oracle.h).
 *
 *   driver            atmosphere::compute_nlte_population_responses,
 *                     atmos/atmos_responses.cc:13-157
 *   profile norms     atom::compute_profile_norm, atmos/atmol/atom/atom_rad.cc:885-930,
 *                     with voigt_der_T / _density / _vt, atom_rad.cc:666-797
 *   contributions     atom::add_response_contributions_new,
 *                     atmos/atmol/atom/atom_responses.cc:668-1128, and its
 *                     helpers compute_xa_der_scalar (atom_rad.cc:611-664),
 *                     op/em_derivative_to_level, op/em_derivative_explicit,
 *                     bb_op/em_derivative_to_level, bb_op/em_derivative_explicit,
 *                     dR_i_cont_dqk, dR_cont_i_dqk, dR_i_cont_dI, dR_cont_i_dI
 *                     (atom_responses.cc:1779-2012)
 *   rate part         atom::compute_nlte_population_responses,
 *                     atom_responses.cc:26-170, with R_ij, R_i_cont, R_cont_i
 *                     (atom_rad.cc:1727-1854)
 *   profiles          fvoigt / fvoigtn, cmlib/profiles.cc:39-112; the line
 *                     profile of atom::boundbound_op, atom_rad.cc:502-537
 *
 * Every expression keeps the reference's operands and their order of
 * evaluation (left to right, no contraction), and the contributions are
 * added in the reference's loop order, so that the assembled system equals
 * the reference's to rounding of the transcendental functions.
 *
 * Reference behaviour kept as it is (each one a line of the reference):
 *  - the rate part divides each line's mean intensity by the profile norm of
 *    the LAST ray (norm is overwritten per ray, atom_responses.cc:98);
 *  - the local J_chunk term of the lines indexes its rows with the levels'
 *    indices within their stage, not their map indices (atom_responses.cc:
 *    1016-1017); equal for an atom whose first stage starts the map;
 *  - em_derivative_explicit uses 2.07E-16 for the Saha constant
 *    (atom_responses.cc:1873, 1878);
 *  - only lines with A > 1 s^-1 and bound-free transitions at or below their
 *    edge wavelength are assembled; every line enters the mean intensities
 *    and the level derivatives of the bound-free terms;
 *  - beta_v_micro keeps its last-level rows (only beta_Temp and beta_density
 *    are overwritten there, atom_responses.cc:52-58, 75-81).
 *
 * Pinned: tests/golden/nlterin.npz, the inputs the reference build forms for
 * the Ca II atom of the cfg0 column and the system (M, beta) it assembles
 * from them, and its profile and profile-norm samples.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "snapi_oracle.h"

#define C_LIGHT 2.99792458E10
#define H_PLANCK 6.6260687652E-27
#define K_BOLTZ 1.3806503E-16
#define PI_D 3.141592653589793238462643383279502884197169399375105
#define SAHA 2.0706615382E-16
#define DELTA_T 5.0
#define DELTA_NT_FRAC 1E-3
#define MAXL 16

typedef struct {
  double re, im;
} cx;
static cx mk(double r, double i) { cx z = {r, i}; return z; }
static cx add(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static cx sub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
static cx mul(cx a, cx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cx dvd(cx a, cx b) {
  double f = b.re * b.re + b.im * b.im;
  return mk((a.re * b.re + a.im * b.im) / f, (a.im * b.re - a.re * b.im) / f);
}
static cx scl(cx a, double s) { return mk(a.re * s, a.im * s); } /* complex * real, real * complex */
static cx radd(double s, cx a) { return mk(a.re + s, a.im); }  /* real + complex */
static cx rsub(double s, cx a) { return mk(s - a.re, -a.im); } /* real - complex */

/* fvoigt(v, a): H of the Humlicek W4 approximant (profiles.cc:91-112). */
static double fvoigt2(double v, double a) {
  double H, F;
  orc_voigt(v, a, &H, &F);
  return H;
}

/* fvoigtn (profiles.cc:39-89): H, F and their derivatives to v of the W4
 * approximant itself. */
static void fvoigtn(double v, double a, double *hh, double *ff, double *dh, double *df) {
  const double s = fabs(v) + a;
  const cx t = mk(a, -v), dt = mk(0, -1.0);
  const cx u = mk(a * a - v * v, -2.0 * a * v), du = mk(-2.0 * v, -2.0 * a);
  cx w4, dw4;
  if (s >= 15.0) {
    const cx q = radd(0.5, u);
    w4 = dvd(scl(t, 0.5641896), q);
    dw4 = sub(dvd(scl(dt, 0.5641896), q), mul(dvd(dvd(scl(t, 0.5641896), q), q), du));
  } else if (s > 5.5) {
    const cx num = radd(1.410474, scl(u, 0.5641896));
    const cx den = radd(0.75, mul(u, radd(3.0, u)));
    w4 = dvd(mul(t, num), den);
    dw4 = add(dvd(mul(dt, num), den), dvd(scl(mul(t, du), 0.5641896), den));
    dw4 = sub(dw4, mul(dvd(dvd(mul(mul(t, du), num), den), den), radd(3.0, scl(u, 2.0))));
  } else if (a > (0.195 * fabs(v) - 0.176)) {
    const cx upper = radd(16.4955, mul(t, radd(20.20933, mul(t, radd(11.96482, mul(t, radd(3.778987,
                                                                                          scl(t, 0.5642236))))))));
    const cx lower = radd(16.4955, mul(t, radd(38.82363, mul(t, radd(39.27121, mul(t, radd(21.69274,
                                                                                          mul(t, radd(6.699398, t)))))))));
    w4 = dvd(upper, lower);
    cx du_ = radd(20.20933, scl(t, 11.96482 * 2.0));
    du_ = add(du_, mul(scl(t, 3.778987 * 3.0), t));
    du_ = add(du_, mul(mul(scl(t, 0.5642236 * 4.0), t), t));
    cx dl_ = radd(38.82363, scl(t, 39.27121 * 2.0));
    dl_ = add(dl_, mul(scl(t, 21.69274 * 3.0), t));
    dl_ = add(dl_, mul(mul(scl(t, 6.699398 * 4.0), t), t));
    dl_ = add(dl_, mul(mul(mul(scl(t, 5.0), t), t), t));
    dw4 = dvd(dvd(mul(dt, sub(mul(du_, lower), mul(dl_, upper))), lower), lower);
  } else {
    cx p = rsub(1.320522, scl(u, 0.56419));
    p = rsub(35.76683, mul(u, p));
    p = rsub(219.0313, mul(u, p));
    p = rsub(1540.787, mul(u, p));
    p = rsub(3321.9905, mul(u, p));
    p = rsub(36183.31, mul(u, p));
    w4 = mul(t, p);
    dw4 = mul(dt, p);
    cx pp = rsub(1.320522 * 5.0, scl(scl(u, 0.56419), 6.0));
    pp = rsub(35.76683 * 4.0, mul(u, pp));
    pp = rsub(219.0313 * 3.0, mul(u, pp));
    pp = rsub(1540.787 * 2.0, mul(u, pp));
    pp = rsub(3321.9905, mul(u, pp));
    dw4 = sub(dw4, mul(mul(t, du), pp));
    cx q = rsub(1.841439, u);
    q = rsub(61.57037, mul(u, q));
    q = rsub(364.2191, mul(u, q));
    q = rsub(2186.181, mul(u, q));
    q = rsub(9022.228, mul(u, q));
    q = rsub(24322.84, mul(u, q));
    const cx v4 = rsub(32066.6, mul(u, q));
    cx qq = rsub(1.841439 * 6.0, scl(u, 7.0));
    qq = rsub(61.57037 * 5.0, mul(u, qq));
    qq = rsub(364.2191 * 4.0, mul(u, qq));
    qq = rsub(2186.181 * 3.0, mul(u, qq));
    qq = rsub(9022.228 * 2.0, mul(u, qq));
    qq = rsub(24322.84, mul(u, qq));
    const cx dv4 = mul(scl(du, -1.0), qq);
    const double su = sin(u.im);
    const double dsu = cos(u.im) * du.im;
    const double eu = exp(u.re);
    cx full = add(scl(mk(sqrt(1.0 - su * su), su), eu * du.re), scl(mk(-su * du.im, dsu), eu));
    full = sub(full, dvd(dvd(sub(mul(dw4, v4), mul(dv4, w4)), v4), v4));
    dw4 = full;
    w4 = sub(scl(mk(sqrt(1.0 - su * su), su), eu), dvd(w4, v4));
  }
  *hh = w4.re / 1.77245386399759;
  *ff = w4.im / 1.77245386399759;
  *dh = dw4.re / 1.77245386399759;
  *df = dw4.im / 1.77245386399759;
}

static double planck(double lam, double T) { return orc_planck(lam, T); }
static double planck_der(double lam, double T) {
  double e = (1.4387769 / lam / T);
  double t = (e > 5.0) ? 1.0 : exp(e) / (exp(e) - 1.0);
  return planck(lam, T) * t * 1.4387769 / lam / T / T;
}

/* The atom seen through the reference's stage/level arrays. */
typedef struct {
  const snapi_nlte_atom *a;
  int nmap, ntr, Z;
  int first[MAXL], nl[MAXL];
  int tmap[MAXL][MAXL]; /* pair of map levels -> line, -1 */
  int bftr[MAXL];       /* map level -> its bound-free transition, -1 */
  double A[MAXL][MAXL], B[MAXL][MAXL];
} atomv;

static int atom_view(const snapi_nlte_atom *a, atomv *v) {
  memset(v, 0, sizeof(*v));
  v->a = a;
  v->nmap = a->nmap;
  v->ntr = a->ntr;
  v->Z = a->nstage - 1;
  if (a->nmap < 2 || a->nmap > MAXL || a->nstage < 2 || a->nstage > MAXL) return -1;
  for (int z = 0; z < a->nstage; ++z) v->first[z] = -1;
  for (int i = 0; i < a->nmap; ++i) {
    const int z = a->level_stage[i];
    if (z < 0 || z >= a->nstage) return -1;
    if (v->first[z] < 0) v->first[z] = i;
    if (a->level_index[i] != i - v->first[z]) return -1;
    v->nl[z]++;
  }
  for (int i = 0; i < MAXL; ++i) {
    v->bftr[i] = -1;
    for (int j = 0; j < MAXL; ++j) v->tmap[i][j] = -1;
  }
  for (int t = 0; t < a->ntr; ++t) {
    const int lo = a->tr_lower[t], up = a->tr_upper[t];
    if (a->tr_type[t] == SNAPI_TR_LINE) {
      v->tmap[lo][up] = v->tmap[up][lo] = t;
      v->A[up][lo] = a->tr_A[t];
      v->B[lo][up] = a->tr_Blu[t];
      v->B[up][lo] = a->tr_Bul[t];
    } else {
      v->bftr[lo] = t;
    }
  }
  return 0;
}

static double broad_dop(const atomv *v, double lam0, double T, double vt) {
  return sqrt(2.0 * K_BOLTZ * T / v->a->mass + vt * vt) * lam0 / C_LIGHT;
}

typedef struct {
  int nd;
  const double *depth, *pops, *damp;
} colv;
#define F(c, f, d) ((c)->depth[(size_t)(f) * (c)->nd + (d)])
#define POP(c, i, d) ((c)->pops[(size_t)(i) * (c)->nd + (d)])
#define DCOL(c, t, d) ((c)->damp[(size_t)(t) * (c)->nd + (d)])
#define DCOLT(c, v, t, d) ((c)->damp[((size_t)(v)->ntr + (t)) * (c)->nd + (d)])

/* voigt_der_T, voigt_der_density, voigt_der_vt (atom_rad.cc:666-797) at
 * depth d for line t at wavelength lambda. */
static void voigt_ders(const atomv *v, const colv *c, int t, int d, double lambda, double vr, double out[3]) {
  const snapi_nlte_atom *a = v->a;
  const int up = a->tr_upper[t], lo = a->tr_lower[t];
  const double lam = fabs(H_PLANCK * C_LIGHT / (a->level_energy[up] - a->level_energy[lo]));
  const double T = F(c, SNAPI_NLTE_T, d), vt = F(c, SNAPI_NLTE_VT, d);
  const double dld = broad_dop(v, lam, T, vt);
  const double x = (lam - lambda * (1.0 + vr / C_LIGHT)) / dld;
  const double dampcol = DCOL(c, t, d);
  const double damping = a->tr_damp_rad[t] + dampcol * a->damping;
  const double aa = damping / dld * (lam * lam) / (4.0 * PI_D * C_LIGHT);
  const double dvd_ = dld / lam * C_LIGHT;
  /* temperature: central difference of the profile along (dx, da, d dld) */
  {
    const double d_dld = lam / (C_LIGHT * dvd_) * K_BOLTZ / a->mass;
    const double dx = -x / dld * d_dld;
    double da = -aa / dld * d_dld;
    da += DCOLT(c, v, t, d) / dld * (lam * lam) / (4.0 * PI_D * C_LIGHT) * a->damping;
    out[0] = (fvoigt2(x + dx * 0.5 * DELTA_T, aa + da * 0.5 * DELTA_T) / (dld + 0.5 * d_dld * DELTA_T) -
              fvoigt2(x - dx * 0.5 * DELTA_T, aa - da * 0.5 * DELTA_T) / (dld - 0.5 * d_dld * DELTA_T)) /
             DELTA_T;
  }
  /* density: central difference in the collisional damping */
  {
    const double Nt = F(c, SNAPI_NLTE_NT, d);
    const double delta_Nt = DELTA_NT_FRAC * Nt;
    const double delta_damp = dampcol * DELTA_NT_FRAC * a->damping;
    const double delta_a = delta_damp / dld * (lam * lam) / (4.0 * PI_D * C_LIGHT);
    const double der = (fvoigt2(x, aa + delta_a * 0.5) - fvoigt2(x, aa - delta_a * 0.5)) / dld / delta_Nt;
    out[1] = der * a->damping;
  }
  /* microturbulence: analytic through fvoigtn */
  {
    const double d_dld = lam / (C_LIGHT * dvd_) * vt;
    const double dx = -x / dld * d_dld;
    const double da = -aa / dld * d_dld;
    double H, Fv, dH, dF;
    fvoigtn(x, aa, &H, &Fv, &dH, &dF);
    const double dH_dT = dH / dld * dx - dF / dld * da;
    out[2] = dH_dT - H / dld / dld * d_dld;
  }
}

/* compute_profile_norm (atom_rad.cc:885-930) for one ray: norm[nd][ntr],
 * normd[3][nd][ntr] (zero for bound-free transitions). */
static void profile_norm(const atomv *v, const colv *c, const double *vr, int nlam, const double *lambda,
                         const double *lw, double *norm, double *normd) {
  const snapi_nlte_atom *a = v->a;
  const int nd = c->nd, ntr = v->ntr;
  memset(norm, 0, sizeof(double) * nd * ntr);
  memset(normd, 0, sizeof(double) * 3 * nd * ntr);
  for (int t = 0; t < ntr; ++t) {
    if (a->tr_type[t] != SNAPI_TR_LINE) continue;
    const int up = a->tr_upper[t], lo = a->tr_lower[t];
    const double lam = H_PLANCK * C_LIGHT / (a->level_energy[up] - a->level_energy[lo]);
    const double ar = a->tr_damp_rad[t];
    const double sf = (lam * lam) / (4.0 * PI_D * C_LIGHT);
    for (int d = 0; d < nd; ++d)
      for (int l = 0; l < nlam; ++l) {
        const double T = F(c, SNAPI_NLTE_T, d), vturb = F(c, SNAPI_NLTE_VT, d);
        const double dld = broad_dop(v, lambda[l], T, vturb);
        const double gc = DCOL(c, t, d);
        const double aa = sf * (ar + gc * a->damping) / dld;
        const double x = (lam - lambda[l] * (1.0 + vr[d] / C_LIGHT)) / dld;
        const double profile = 1.0 / dld * fvoigt2(x, aa);
        double der[3];
        voigt_ders(v, c, t, d, lambda[l], vr[d], der);
        norm[d * ntr + t] += profile * lw[l];
        for (int q = 0; q < 3; ++q) normd[((size_t)q * nd + d) * ntr + t] += der[q] * lw[l];
      }
  }
}

/* The line profiles of atom::boundbound_op at one wavelength: prof[nd][ntr]. */
static void line_profiles(const atomv *v, const colv *c, const double *vr, double lambda, double *prof) {
  const snapi_nlte_atom *a = v->a;
  const int nd = c->nd, ntr = v->ntr;
  memset(prof, 0, sizeof(double) * nd * ntr);
  for (int t = 0; t < ntr; ++t) {
    if (a->tr_type[t] != SNAPI_TR_LINE) continue;
    const int up = a->tr_upper[t], lo = a->tr_lower[t];
    const double lam = H_PLANCK * C_LIGHT / (a->level_energy[up] - a->level_energy[lo]);
    const double ar = a->tr_damp_rad[t];
    const double sf = (lam * lam) / (4.0 * PI_D * C_LIGHT);
    for (int d = 0; d < nd; ++d) {
      const double dld = broad_dop(v, lam, F(c, SNAPI_NLTE_T, d), F(c, SNAPI_NLTE_VT, d));
      const double gc = DCOL(c, t, d);
      const double aa = sf * (ar + gc * a->damping) / dld;
      const double x = (lam - lambda * (1.0 + vr[d] / C_LIGHT)) / dld;
      prof[d * ntr + t] = 1.0 / dld * fvoigt2(x, aa);
    }
  }
}

/* Profile derivatives of compute_xa_der_scalar + fvoigtn
 * (atom_responses.cc:733-768) for the lines with A > 1: pd[3][nd][ntr]. */
static void profile_derivatives(const atomv *v, const colv *c, const double *vr, double lambda, double *pd) {
  const snapi_nlte_atom *a = v->a;
  const int nd = c->nd, ntr = v->ntr;
  memset(pd, 0, sizeof(double) * 3 * nd * ntr);
  for (int d = 0; d < nd; ++d)
    for (int t = 0; t < ntr; ++t) {
      if (a->tr_type[t] != SNAPI_TR_LINE || !(a->tr_A[t] > 1.0)) continue;
      const int up = a->tr_upper[t], lo = a->tr_lower[t];
      const double lam = fabs(H_PLANCK * C_LIGHT / (a->level_energy[up] - a->level_energy[lo]));
      const double T = F(c, SNAPI_NLTE_T, d), vt = F(c, SNAPI_NLTE_VT, d);
      const double dld = broad_dop(v, lam, T, vt);
      const double x = (lam - lambda * (1.0 + vr[d] / C_LIGHT)) / dld;
      const double dampcol = DCOL(c, t, d);
      const double damping = a->tr_damp_rad[t] + dampcol * a->damping;
      const double aa = damping / dld * (lam * lam) / (4.0 * PI_D * C_LIGHT);
      const double dvd_ = sqrt(2.0 * K_BOLTZ * T / a->mass + vt * vt);
      double xd[3], ad[3], ld[3];
      ld[0] = lam / (C_LIGHT * dvd_) * K_BOLTZ / a->mass;
      xd[0] = -x / dld * ld[0];
      xd[1] = 0.0;
      ld[1] = 0.0;
      ld[2] = lam / (C_LIGHT * dvd_) * vt;
      xd[2] = -x / dld * ld[2];
      ad[0] = -aa / dld * ld[0];
      ad[0] += DCOLT(c, v, t, d) / dld * (lam * lam) / (4.0 * PI_D * C_LIGHT) * a->damping;
      const double Nt = F(c, SNAPI_NLTE_NT, d);
      const double delta_Nt = DELTA_NT_FRAC * Nt;
      const double delta_damp = dampcol * DELTA_NT_FRAC;
      const double delta_a = delta_damp / dld * (lam * lam) / (4.0 * PI_D * C_LIGHT);
      ad[1] = delta_a / delta_Nt * a->damping;
      ad[2] = -aa / dld * ld[2];
      double H, Fv, dH, dF;
      fvoigtn(x, aa, &H, &Fv, &dH, &dF);
      const double dld2 = broad_dop(v, lam, T, vt);
      for (int q = 0; q < 3; ++q)
        pd[((size_t)q * nd + d) * ntr + t] = dH * xd[q] / dld2 - dF * ad[q] / dld2 - H / dld2 / dld2 * ld[q];
    }
}

/* One (ray, wavelength) problem: add_response_contributions_new
 * (atom_responses.cc:668-1128).  opp, emm: the total opacity and emissivity
 * as the reference hands them over (normalised and back); pert[6][nd]: LTE
 * perturbations (op T, Nt, vt; em T, Nt, vt); sg[ntr] the bound-free
 * cross-sections at this wavelength. */
static void add_contributions(const atomv *v, const colv *c, const double *vr, double lambda, double lambda_w,
                              double angular_weight, const double *sg, const double *norm, const double *normd,
                              const double *opp, const double *emm, const double *pert, const double *I,
                              const double *Rop, const double *Rem, int tau_grid, double *Jm, double *Jn, double *M,
                              double *beta, double *prof, double *pd) {
  const snapi_nlte_atom *a = v->a;
  const int nd = c->nd, ntr = v->ntr, nmap = v->nmap, n = nd * nmap;
  const double *pert_op[3] = {pert, pert + nd, pert + 2 * nd};
  const double *pert_em[3] = {pert + 3 * nd, pert + 4 * nd, pert + 5 * nd};
  line_profiles(v, c, vr, lambda, prof);
#define PROF(d, t) prof[(d) * ntr + (t)]
#define NRM(d, t) norm[(d) * ntr + (t)]
#define SIGL(i) (v->bftr[i] >= 0 ? sg[v->bftr[i]] : 0.0)
  /* radiative rates: the mean intensities of every transition */
  for (int d = 0; d < nd; ++d)
    for (int t = 0; t < ntr; ++t) {
      const int lo = a->tr_lower[t], up = a->tr_upper[t];
      if (a->tr_type[t] == SNAPI_TR_LINE) {
        const double line_energy = a->level_energy[up] - a->level_energy[lo];
        const double line_opacity = (POP(c, lo, d) * v->B[lo][up] - POP(c, up, d) * v->B[up][lo]) * PROF(d, t) *
                                    line_energy * 0.25 / PI_D;
        const double line_factor = line_opacity / opp[d];
        const double el = lambda_w * angular_weight * I[d] * PROF(d, t) * line_factor * 0.25 / PI_D;
        Jm[d * ntr + t] += el;
        Jn[d * ntr + t] += el;
      } else {
        const double sigma = sg[t];
        Jm[d * ntr + t] += lambda_w * angular_weight * sigma * lambda / H_PLANCK / C_LIGHT * I[d];
        Jn[d * ntr + t] += lambda_w * angular_weight * sigma * lambda / H_PLANCK / C_LIGHT *
                           (I[d] + 2.0 * H_PLANCK * C_LIGHT * C_LIGHT / pow(lambda, 5.0)) *
                           exp(-H_PLANCK * C_LIGHT / lambda / K_BOLTZ / F(c, SNAPI_NLTE_T, d));
      }
    }
  profile_derivatives(v, c, vr, lambda, pd);
#define PD(q, d, t) pd[((size_t)(q) * nd + (d)) * ntr + (t)]
  /* explicit derivatives of the bound-free and bound-bound op and em */
  double *bfop = calloc((size_t)nd * 3, sizeof(double)), *bfem = calloc((size_t)nd * 3, sizeof(double));
  double *bbop = calloc((size_t)nd * 3, sizeof(double)), *bbem = calloc((size_t)nd * 3, sizeof(double));
  double *dop = calloc((size_t)nd * nmap, sizeof(double)), *dem = calloc((size_t)nd * nmap, sizeof(double));
  double *dopb = calloc((size_t)nd * nmap, sizeof(double)), *demb = calloc((size_t)nd * nmap, sizeof(double));
  for (int d = 0; d < nd; ++d) {
    const double T = F(c, SNAPI_NLTE_T, d), ne = F(c, SNAPI_NLTE_NE, d);
    const double dneT = F(c, SNAPI_NLTE_DNE_DT, d), dneN = F(c, SNAPI_NLTE_DNE_DNT, d);
    /* op_derivative_explicit, em_derivative_explicit (atom_responses.cc:1821-1891) */
    for (int z = 0; z < v->Z; ++z)
      for (int l = 0; l < v->nl[z]; ++l) {
        const int i = v->first[z] + l, cont = v->first[z + 1];
        const double sigma = SIGL(i);
        if (sigma) {
          const double gl = a->level_g[i], gc = a->level_g[cont], nc = POP(c, cont, d);
          {
            const double delta_E = (a->stage_ip[z] - a->level_energy[i] - H_PLANCK * C_LIGHT / lambda);
            const double cp = -nc * SAHA * gl / gc * sigma;
            bfop[d * 3 + 0] += cp * (dneT * pow(T, -1.5) * exp(delta_E / K_BOLTZ / T));
            bfop[d * 3 + 0] += cp * ne * (-1.5) * pow(T, -2.5) * exp(delta_E / K_BOLTZ / T);
            bfop[d * 3 + 0] += cp * ne * pow(T, -1.5) * exp(delta_E / K_BOLTZ / T) * (-delta_E / K_BOLTZ / T / T);
            bfop[d * 3 + 1] += cp * (dneN * pow(T, -1.5) * exp(delta_E / K_BOLTZ / T));
          }
          {
            const double B_planck = planck(lambda, T);
            const double delta_E = (a->stage_ip[z] - a->level_energy[i]);
            const double pop_mod = nc * ne * 2.07E-16 * pow(T, -1.5) * gl / gc * exp(delta_E / K_BOLTZ / T);
            bfem[d * 3 + 0] += sigma *
                               (-pop_mod * exp(-H_PLANCK * C_LIGHT / lambda / K_BOLTZ / T) * H_PLANCK * C_LIGHT /
                                lambda / K_BOLTZ / T / T) *
                               B_planck;
            const double cp = nc * 2.07E-16 * gl / gc * B_planck * sigma *
                              (1.0 - exp(-H_PLANCK * C_LIGHT / lambda / K_BOLTZ / T));
            bfem[d * 3 + 0] += cp * (dneT * pow(T, -1.5) * exp(delta_E / K_BOLTZ / T) +
                                     ne * (-1.5 * pow(T, -2.5)) * exp(delta_E / K_BOLTZ / T) +
                                     ne * pow(T, -1.5) * exp(delta_E / K_BOLTZ / T) * -delta_E / K_BOLTZ / T / T);
            bfem[d * 3 + 0] += pop_mod * sigma * (1.0 - exp(-H_PLANCK * C_LIGHT / lambda / K_BOLTZ / T)) *
                               planck_der(lambda, T);
            bfem[d * 3 + 1] += cp * (dneN * pow(T, -1.5) * exp(delta_E / K_BOLTZ / T));
          }
        }
      }
    /* bb_op_derivative_explicit, bb_em_derivative_explicit (atom_responses.cc:1927-1966) */
    for (int z = 0; z <= v->Z; ++z)
      for (int il = 1; il < v->nl[z]; ++il)
        for (int iil = 0; iil < il; ++iil) {
          const int i = v->first[z] + il, ii = v->first[z] + iil;
          const int t = v->tmap[i][ii];
          if (t < 0) continue;
          const double lam = H_PLANCK * C_LIGHT / (a->level_energy[i] - a->level_energy[ii]);
          const double X = (POP(c, ii, d) * v->B[ii][i] - POP(c, i, d) * v->B[i][ii]);
          for (int q = 0; q < 3; ++q) {
            bbop[d * 3 + q] += PD(q, d, t) * X * H_PLANCK * C_LIGHT / lam / 4.0 / PI_D;
            bbem[d * 3 + q] += PD(q, d, t) * POP(c, i, d) * v->A[i][ii] * H_PLANCK * C_LIGHT / lam / 4.0 / PI_D;
          }
        }
    /* derivatives to the level populations (atom_responses.cc:1779-1925) */
    for (int i = 0; i < nmap; ++i) {
      const int z = a->level_stage[i], l = a->level_index[i];
      double der = SIGL(i);
      if (z > 0 && l == 0)
        for (int ll = 0; ll < v->nl[z - 1]; ++ll) {
          const int j = v->first[z - 1] + ll;
          const double sigmaa = SIGL(j);
          const double delta_E = a->stage_ip[z - 1] - a->level_energy[j] - H_PLANCK * C_LIGHT / lambda;
          der -= sigmaa * ne * SAHA * pow(T, -1.5) * a->level_g[j] / a->level_g[i] * exp(delta_E / K_BOLTZ / T);
        }
      dop[d * nmap + i] = der;
      der = 0;
      if (z > 0 && l == 0)
        for (int ll = 0; ll < v->nl[z - 1]; ++ll) {
          const int j = v->first[z - 1] + ll;
          const double sigmaa = SIGL(j);
          const double delta_E = a->stage_ip[z - 1] - a->level_energy[j];
          der += sigmaa * ne * SAHA * pow(T, -1.5) * a->level_g[j] / a->level_g[i] * exp(delta_E / K_BOLTZ / T) *
                 (1.0 - exp(-H_PLANCK * C_LIGHT / lambda / K_BOLTZ / T)) * planck(lambda, T);
        }
      dem[d * nmap + i] = der;
      der = 0.0;
      for (int ll = l + 1; ll < v->nl[z]; ++ll) {
        const int j = v->first[z] + ll, t = v->tmap[i][j];
        der += v->B[i][j] * H_PLANCK * C_LIGHT / lambda * 0.25 / PI_D * (t >= 0 ? PROF(d, t) : 0.0);
      }
      for (int ll = 0; ll < l; ++ll) {
        const int j = v->first[z] + ll, t = v->tmap[i][j];
        der -= v->B[i][j] * H_PLANCK * C_LIGHT / lambda * 0.25 / PI_D * (t >= 0 ? PROF(d, t) : 0.0);
      }
      dopb[d * nmap + i] = der;
      der = 0.0;
      for (int ll = 0; ll < l; ++ll) {
        const int j = v->first[z] + ll, t = v->tmap[i][j];
        der += v->A[i][j] * H_PLANCK * C_LIGHT / lambda * 0.25 / PI_D * (t >= 0 ? PROF(d, t) : 0.0);
      }
      demb[d * nmap + i] = der;
    }
  }
#define MM(r, cc) M[(size_t)(r) * n + (cc)]
#define BETA(q, k, r) beta[((size_t)(q) * nd + (k)) * n + (r)]
  for (int l = 0; l < nd; ++l)
    for (int t = 0; t < ntr; ++t) {
      const int lo = a->tr_lower[t], up = a->tr_upper[t];
      int type = 0;
      if (a->tr_type[t] == SNAPI_TR_BF) {
        if (lambda <= a->tr_lambda_crit[t]) type = 1;
      } else if (v->A[up][lo] > 1.0) {
        type = 2;
      }
      if (type == 2) {
        const int ri = up, rii = lo;                   /* map indices */
        const int i_raw = a->level_index[up], ii_raw = a->level_index[lo]; /* stage indices */
        const double Blu = v->B[lo][up], Bul = v->B[up][lo], A = v->A[up][lo];
        const double line_energy = a->level_energy[up] - a->level_energy[lo];
        const double lam = H_PLANCK * C_LIGHT / line_energy;
        const double line_opacity = (POP(c, lo, l) * Blu - POP(c, up, l) * Bul) * line_energy * 0.25 / PI_D;
        const double line_factor = line_opacity * PROF(l, t) / opp[l];
        const double J_chunk = (Blu * POP(c, lo, l) - Bul * POP(c, up, l)) * I[l] * lambda_w * angular_weight *
                               PROF(l, t) / NRM(l, t) * line_factor * 0.25 / PI_D;
        for (int q = 0; q < 3; ++q) {
          const double derivative = PD(q, l, t);
          double derived_part = 0.0;
          derived_part += derivative / PROF(l, t);
          derived_part -= normd[((size_t)q * nd + l) * ntr + t] / NRM(l, t);
          derived_part += (1.0 - line_factor) / line_factor * line_opacity / opp[l] * derivative;
          if (q < 2) derived_part -= (pert_op[q][l] + bfop[l * 3 + q]) / opp[l];
          BETA(q, l, l * nmap + ri) -= J_chunk * derived_part;
          BETA(q, l, l * nmap + rii) += J_chunk * derived_part;
        }
        const double dR_dI = (Blu * POP(c, lo, l) - Bul * POP(c, up, l)) * lambda_w * angular_weight * PROF(l, t) /
                             NRM(l, t) * line_factor * 0.25 / PI_D;
        for (int ll = 0; ll < nd; ++ll) {
          const double op_ref = tau_grid ? F(c, SNAPI_NLTE_OPREF, ll) : 1.0;
          const double ro = Rop[(size_t)l * nd + ll], re = Rem[(size_t)l * nd + ll];
          const double d_op_d_profile =
              (POP(c, lo, ll) * Blu - POP(c, up, ll) * Bul) * H_PLANCK * C_LIGHT / lam * 0.25 / PI_D;
          const double d_em_d_profile = POP(c, up, ll) * A * H_PLANCK * C_LIGHT / lam * 0.25 / PI_D;
          for (int q = 0; q < 3; ++q) {
            const double derivative = PD(q, ll, t);
            double d_op_d_qk = derivative * d_op_d_profile;
            double d_em_d_qk = derivative * d_em_d_profile;
            if (q < 2) {
              d_op_d_qk += bfop[ll * 3 + q] + pert_op[q][ll];
              d_em_d_qk += bfem[ll * 3 + q] + pert_em[q][ll];
            } else {
              d_op_d_qk += pert_op[2][ll];
              d_em_d_qk += pert_em[2][ll];
            }
            const double remote = dR_dI * (ro * d_op_d_qk + re * d_em_d_qk);
            BETA(q, ll, l * nmap + ri) -= remote / op_ref;
            BETA(q, ll, l * nmap + rii) += remote / op_ref;
            if (tau_grid) {
              const double opd = F(c, SNAPI_NLTE_DOPREF_DT + q, ll);
              BETA(q, ll, l * nmap + ri) += dR_dI * (ro * opd * opp[ll] + re * opd * emm[ll]) / op_ref / op_ref;
              BETA(q, ll, l * nmap + rii) -= dR_dI * (ro * opd * opp[ll] + re * opd * emm[ll]) / op_ref / op_ref;
            }
          }
          double der_i = line_energy * 0.25 / PI_D * PROF(ll, t) * (A * re - Bul * ro);
          double der_ii = line_energy * 0.25 / PI_D * PROF(ll, t) * (Blu * ro);
          der_i /= op_ref;
          der_ii /= op_ref;
          if (l == ll) {
            der_i -= (1.0 - line_factor) / opp[l] * Bul * I[l] * line_energy * 0.25 / PI_D * PROF(l, t) / line_factor;
            der_ii += (1.0 - line_factor) / opp[l] * Blu * I[l] * line_energy * 0.25 / PI_D * PROF(l, t) / line_factor;
          }
          const double nd_ = der_ii * dR_dI;
          const double dd = der_i * dR_dI;
          MM(l * nmap + ri, ll * nmap + rii) += nd_;
          MM(l * nmap + ri, ll * nmap + ri) += dd;
          MM(l * nmap + rii, ll * nmap + ri) -= dd;
          MM(l * nmap + rii, ll * nmap + rii) -= nd_;
          for (int iii = 0; iii < nmap; ++iii) {
            MM(l * nmap + ri, ll * nmap + iii) += dR_dI / op_ref * (ro * dop[ll * nmap + iii] + re * dem[ll * nmap + iii]);
            MM(l * nmap + rii, ll * nmap + iii) -= dR_dI / op_ref * (ro * dop[ll * nmap + iii] + re * dem[ll * nmap + iii]);
            if (l == ll) {
              MM(l * nmap + i_raw, ll * nmap + iii) -= J_chunk * dop[l * nmap + iii] / opp[l];
              MM(l * nmap + ii_raw, ll * nmap + iii) += J_chunk * dop[l * nmap + iii] / opp[l];
            }
          }
        }
      } else if (type == 1) {
        const int z = a->level_stage[lo], rc = up, rii = lo;
        const double T = F(c, SNAPI_NLTE_T, l), n_e = F(c, SNAPI_NLTE_NE, l);
        const double sigma = SIGL(lo);
        const double gi = a->level_g[lo], gc = a->level_g[up];
        const double delta_E = a->stage_ip[z] - a->level_energy[lo];
        const double dRic_dI = sigma * lambda / H_PLANCK / C_LIGHT;
        const double dRci_dI = sigma * lambda / H_PLANCK / C_LIGHT *
                               exp((delta_E - H_PLANCK * C_LIGHT / lambda) / K_BOLTZ / T) * gi / gc * pow(T, -1.5) *
                               n_e * SAHA;
        double dRic_dq[2] = {0.0, 0.0}, dRci_dq[2];
        {
          const double cp = (I[l] + 2.0 * H_PLANCK * C_LIGHT * C_LIGHT / pow(lambda, 5.0)) * sigma * lambda /
                            H_PLANCK / C_LIGHT * gi / gc * SAHA;
          const double ex = exp((delta_E - H_PLANCK * C_LIGHT / lambda) / K_BOLTZ / T);
          dRci_dq[0] = ex * pow(T, -1.5) * F(c, SNAPI_NLTE_DNE_DT, l);
          dRci_dq[0] += ex * (-1.5) * pow(T, -2.5) * n_e;
          dRci_dq[0] += ex * (-delta_E + H_PLANCK * C_LIGHT / lambda) / K_BOLTZ / T / T * pow(T, -1.5) * n_e;
          dRci_dq[0] *= cp;
          dRci_dq[1] = cp * ex * pow(T, -1.5) * F(c, SNAPI_NLTE_DNE_DNT, l);
        }
        const double nc = POP(c, rc, l), ni = POP(c, rii, l);
        for (int q = 0; q < 2; ++q) {
          BETA(q, l, l * nmap + rc) += nc * dRci_dq[q] * angular_weight * lambda_w;
          BETA(q, l, l * nmap + rii) -= nc * dRci_dq[q] * angular_weight * lambda_w;
          BETA(q, l, l * nmap + rc) -= ni * dRic_dq[q] * angular_weight * lambda_w;
          BETA(q, l, l * nmap + rii) += ni * dRic_dq[q] * angular_weight * lambda_w;
        }
        for (int ll = 0; ll < nd; ++ll) {
          const double op_ref = tau_grid ? F(c, SNAPI_NLTE_OPREF, ll) : 1.0;
          const double ro = Rop[(size_t)l * nd + ll], re = Rem[(size_t)l * nd + ll];
          double dI[2];
          for (int q = 0; q < 2; ++q) {
            dI[q] = ro * pert_op[q][ll] + re * pert_em[q][ll];
            dI[q] += ro * bbop[ll * 3 + q] + re * bbem[ll * 3 + q];
            dI[q] += ro * bfop[ll * 3 + q] + re * bfem[ll * 3 + q];
          }
          for (int q = 0; q < 2; ++q) {
            BETA(q, ll, l * nmap + rc) += nc * dRci_dI * dI[q] * angular_weight * lambda_w / op_ref;
            BETA(q, ll, l * nmap + rii) -= nc * dRci_dI * dI[q] * angular_weight * lambda_w / op_ref;
          }
          for (int q = 0; q < 2; ++q) {
            BETA(q, ll, l * nmap + rc) -= ni * dRic_dI * dI[q] * angular_weight * lambda_w / op_ref;
            BETA(q, ll, l * nmap + rii) += ni * dRic_dI * dI[q] * angular_weight * lambda_w / op_ref;
          }
          if (tau_grid)
            for (int q = 0; q < 2; ++q) {
              const double opd = F(c, SNAPI_NLTE_DOPREF_DT + q, ll);
              const double X = (ro * opd * opp[ll] + re * opd * emm[ll]);
              BETA(q, ll, l * nmap + rc) -= nc * dRci_dI * angular_weight * lambda_w * X / op_ref / op_ref;
              BETA(q, ll, l * nmap + rii) += nc * dRci_dI * angular_weight * lambda_w * X / op_ref / op_ref;
              BETA(q, ll, l * nmap + rc) += ni * dRic_dI * angular_weight * lambda_w * X / op_ref / op_ref;
              BETA(q, ll, l * nmap + rii) -= ni * dRic_dI * angular_weight * lambda_w * X / op_ref / op_ref;
            }
          for (int iii = 0; iii < nmap; ++iii) {
            const double d_op_d_n = dopb[ll * nmap + iii] + dop[ll * nmap + iii];
            const double d_em_d_n = demb[ll * nmap + iii] + dem[ll * nmap + iii];
            const double X = (ro * d_op_d_n + re * d_em_d_n);
            MM(l * nmap + rc, ll * nmap + iii) += ni * dRic_dI * X * lambda_w * angular_weight / op_ref;
            MM(l * nmap + rii, ll * nmap + iii) -= ni * dRic_dI * X * lambda_w * angular_weight / op_ref;
            MM(l * nmap + rc, ll * nmap + iii) -= nc * dRci_dI * X * lambda_w * angular_weight / op_ref;
            MM(l * nmap + rii, ll * nmap + iii) += nc * dRci_dI * X * lambda_w * angular_weight / op_ref;
          }
        }
      }
    }
  free(bfop);
  free(bfem);
  free(bbop);
  free(bbem);
  free(dop);
  free(dem);
  free(dopb);
  free(demb);
}

/* The rate part of atom::compute_nlte_population_responses
 * (atom_responses.cc:26-170). */
static void rate_part(const atomv *v, const colv *c, const double *coll, const double *Jm, const double *Jn,
                      const double *norm_last, double *M, double *beta) {
  const snapi_nlte_atom *a = v->a;
  const int nd = c->nd, nmap = v->nmap, ntr = v->ntr, n = nd * nmap;
#define COLL(q, i, j, d) coll[(((size_t)(q) * nmap + (i)) * nmap + (j)) * nd + (d)]
  for (int q = 0; q < 2; ++q) {
    for (int d = 0; d < nd; ++d)
      for (int i = 0; i < nmap; ++i)
        for (int ii = 0; ii < nmap; ++ii)
          BETA(q, d, d * nmap + i) += (POP(c, i, d) * COLL(1 + q, i, ii, d) - POP(c, ii, d) * COLL(1 + q, ii, i, d));
    for (int k = 0; k < nd; ++k)
      for (int d = 0; d < nd; ++d)
        BETA(q, k, d * nmap + nmap - 1) = (k == d) ? F(c, SNAPI_NLTE_DNACT_DT + q, k) : 0.0;
  }
  for (int l = 0; l < nd; ++l) {
    const double Temp = F(c, SNAPI_NLTE_T, l), ne = F(c, SNAPI_NLTE_NE, l);
    for (int i0 = 0; i0 < nmap - 1; ++i0) {
      const int z = a->level_stage[i0];
      for (int lii = 0; lii < v->nl[z]; ++lii) {
        const int j = v->first[z] + lii, t = v->tmap[i0][j];
        const double JJ = (t >= 0) ? Jm[l * ntr + t] / norm_last[l * ntr + t] : -1.0;
        double R = 1.0 * (v->A[i0][j] + v->B[i0][j] * JJ);
        double C = 1.0 * COLL(0, i0, j, l);
        MM(l * nmap + i0, l * nmap + i0) -= (R + C);
        R = 1.0 * (v->A[j][i0] + v->B[j][i0] * JJ);
        C = 1.0 * COLL(0, j, i0, l);
        MM(l * nmap + i0, l * nmap + j) += (R + C);
      }
      const int cont = v->first[z + 1], t = v->bftr[i0];
      double JJ = (t >= 0) ? Jm[l * ntr + t] : -1.0;
      double R = 1.0 * JJ;
      double C = COLL(0, i0, cont, l);
      MM(l * nmap + i0, l * nmap + i0) -= (R + C);
      MM(l * nmap + cont, l * nmap + i0) += (R + C);
      JJ = (t >= 0) ? Jn[l * ntr + t] : -1.0;
      const double dE = a->stage_ip[z] - a->level_energy[i0];
      R = 1.0 * (JJ * a->level_g[i0] / a->level_g[cont] * exp(dE / K_BOLTZ / Temp) * pow(Temp, -1.5) * ne * SAHA);
      C = COLL(0, cont, i0, l);
      MM(l * nmap + i0, l * nmap + cont) += (R + C);
      MM(l * nmap + cont, l * nmap + cont) -= (R + C);
    }
  }
  for (int l = 0; l < nd; ++l)
    for (int ll = 0; ll < nd; ++ll)
      for (int ii = 0; ii < nmap; ++ii) MM(l * nmap + nmap - 1, ll * nmap + ii) = (l == ll) ? 1.0 : 0.0;
}

int orc_nlte_response_system(const snapi_nlte_atom *atom, int nd, int nray, int nlam, int tau_grid, double boundary,
                             const double *ray_theta, const double *ray_weight, const double *lambda,
                             const double *lambda_weight, const double *sigma, const double *depth,
                             const double *pops, const double *damp, const double *coll, const double *vr,
                             const double *opem, double *M, double *beta) {
  atomv v;
  if (atom_view(atom, &v)) return -1;
  if (v.nl[v.Z] != 1 || v.first[v.Z] != v.nmap - 1) return -1;
  const int nmap = v.nmap, ntr = v.ntr, n = nd * nmap;
  colv c = {nd, depth, pops, damp};
  memset(M, 0, sizeof(double) * (size_t)n * n);
  memset(beta, 0, sizeof(double) * (size_t)3 * nd * n);
  double *Jm = calloc((size_t)nd * ntr, sizeof(double)), *Jn = calloc((size_t)nd * ntr, sizeof(double));
  double *norm = calloc((size_t)nd * ntr, sizeof(double)), *normd = calloc((size_t)3 * nd * ntr, sizeof(double));
  double *prof = malloc(sizeof(double) * nd * ntr), *pd = malloc(sizeof(double) * 3 * nd * ntr);
  double *opn = malloc(sizeof(double) * nd), *emn = malloc(sizeof(double) * nd);
  double *opp = malloc(sizeof(double) * nd), *emm = malloc(sizeof(double) * nd);
  double *I = malloc(sizeof(double) * nd), *L = malloc(sizeof(double) * nd);
  double *Rop = malloc(sizeof(double) * nd * nd), *Rem = malloc(sizeof(double) * nd * nd);
  double *sg = malloc(sizeof(double) * (ntr > 0 ? ntr : 1));
  int rc = 0;
  for (int tp = 0; tp < nray && !rc; ++tp) {
    const double *vrr = vr + (size_t)tp * nd;
    profile_norm(&v, &c, vrr, nlam, lambda, lambda_weight, norm, normd);
    for (int l = 0; l < nlam; ++l) {
      const double *b = opem + ((size_t)tp * nlam + l) * 8 * nd;
      for (int d = 0; d < nd; ++d) {
        opn[d] = b[d];
        emn[d] = b[nd + d];
        if (tau_grid) {
          opn[d] /= F(&c, SNAPI_NLTE_OPREF, d);
          emn[d] /= F(&c, SNAPI_NLTE_OPREF, d);
        }
      }
      if (orc_formal_responses(nd, depth + (size_t)SNAPI_NLTE_GRID * nd, opn, emn, ray_theta[tp], boundary, I, L,
                               Rop, Rem)) {
        rc = -2;
        break;
      }
      for (int d = 0; d < nd; ++d) {
        opp[d] = opn[d];
        emm[d] = emn[d];
        if (tau_grid) {
          opp[d] *= F(&c, SNAPI_NLTE_OPREF, d);
          emm[d] *= F(&c, SNAPI_NLTE_OPREF, d);
        }
      }
      for (int t = 0; t < ntr; ++t) sg[t] = sigma[(size_t)t * nlam + l];
      add_contributions(&v, &c, vrr, lambda[l], lambda_weight[l], ray_weight[tp], sg, norm, normd, opp, emm,
                        b + 2 * nd, I, Rop, Rem, tau_grid, Jm, Jn, M, beta, prof, pd);
    }
  }
  if (!rc) rate_part(&v, &c, coll, Jm, Jn, norm, M, beta);
  free(Jm);
  free(Jn);
  free(norm);
  free(normd);
  free(prof);
  free(pd);
  free(opn);
  free(emn);
  free(opp);
  free(emm);
  free(I);
  free(L);
  free(Rop);
  free(Rem);
  free(sg);
  return rc;
}

/* The pieces the system is built from, for checking: the profile norms of
 * one ray (norm [nd][ntr], normd [3][nd][ntr]) and the line profiles at one
 * wavelength (prof [nd][ntr]). */
int orc_nlte_profile_norm(const snapi_nlte_atom *atom, int nd, int nlam, const double *lambda,
                          const double *lambda_weight, const double *depth, const double *damp, const double *vr,
                          double *norm, double *normd) {
  atomv v;
  if (atom_view(atom, &v)) return -1;
  colv c = {nd, depth, NULL, damp};
  profile_norm(&v, &c, vr, nlam, lambda, lambda_weight, norm, normd);
  return 0;
}

int orc_nlte_line_profiles(const snapi_nlte_atom *atom, int nd, double lambda, const double *depth,
                           const double *damp, const double *vr, double *prof) {
  atomv v;
  if (atom_view(atom, &v)) return -1;
  colv c = {nd, depth, NULL, damp};
  line_profiles(&v, &c, vr, lambda, prof);
  return 0;
}
