/*
 * FP64 CPU restatement of the reference's chemical equilibrium and LTE
 * populations -- TEST INFRASTRUCTURE ONLY (see snapi_oracle.h).
 *
 *   partition functions   atmos/atmol/atom/partf.cc:14-358 (pf, cpf, trav, irwin)
 *   ionisation fractions  atom::ionfrc, atmos/atmol/atom/atom.cc:892-924
 *   dissociation consts   eqc_cnst/eqc_poly/eqc_hmin, atmos/atmol/mol/moleqc.cc;
 *                         mol::dissoc, atmos/atmol/mol/mol.cc:159-162
 *   equilibrium           atmosphere::chemeq, atmos/atmos_chm.cc:28-363
 *   level populations     atom::lte, atmos/atmol/atom/atom.cc:779-787
 *   referent opacity      atmosphere::opacity_continuum, atmos/atmos_rad.cc:286-294
 *   hydrostatic eq.       atmos_pp::enforce_hequilibrium, atmos/plpar/atmos_pp.cc:372-455
 *   hydrostatic tangent   no reference counterpart (the reference differences
 *                         two rebuilds, atmos_fit.cc:1225-1248)
 *
 * The reference solves each Newton system by SVD (cmlib svdcmp) and exits
 * when a singular value falls below 1e-30 of the largest; for every system it
 * does solve, the SVD solution is the unique solution of A v = d, which the
 * Gaussian elimination with partial pivoting used here computes to rounding.
 *
 * Pinned: the reference build's own partition functions, ionisation
 * fractions, dissociation constants, electron densities, stage and level
 * populations and H- densities on 133 depth points
 * (tests/golden/chemeq.npz); the pure-hydrogen Saha limit in closed form;
 * charge and particle conservation at the result; the referent opacity
 * against the reference's compute_op_referent on the same points and the LTE
 * detailed-balance form of its bound-free term; the hydrostatic march against
 * the reference's enforce_hequilibrium on the cfg0 column (Nt, Ne, opacity and
 * density at every depth) and against the pressure balance it integrates;
 * the hydrostatic tangent against central differences of the march solved to
 * rounding (orc_hequilibrium_tol).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "snapi_oracle.h"

/* CGS constants as cmlib/const.h states them. */
#define C_PI 3.141592653589793238462643383279502884197169399375105
#define C_H 6.6260687652E-27
#define C_K 1.3806503E-16
#define C_ME 9.1093818872E-28
#define C_E 4.8032042510E-10
#define C_RYD 1.097373156853955E5
#define C_E0 7.957747154594775607E-02
#define C_C 2.99792458E10
#define C_SAHA 2.0706615382E-16

#define MAXE SNAPI_CHEM_MAX_ELEMENTS
#define MAXS SNAPI_CHEM_MAX_STAGES
#define MAXM SNAPI_CHEM_MAX_MOLECULES

/* Fischel & Sparks (1971) simplified eqs. (26)/(27) for the asymptotic sum
 * and its derivative with respect to p: trav::qas_f / trav::dqas_f,
 * partf.cc:261-288. */
static double qas_f(double l, double p, double dz) {
  if (p > l) {
    const double llm = l - 1.0;
    return ((1.3333333 * p + 0.5) * p + 0.16666667 + 1.33333333 * dz) * p - 0.4 * dz * dz / p -
           ((0.33333333 * llm - 0.5) * llm - 0.16666667 - dz) * llm + 0.5 * dz * dz / l;
  }
  const double axl2 = dz / (l * l);
  return (1.0 + axl2 * (0.33333333 + 0.1 * axl2)) * p * p * p * p / l;
}

static double dqas_f(double l, double p, double dz) {
  if (p > l) return ((3.0 * 1.3333333 * p + 1.0) * p + 0.16666667 + 1.33333333 * dz) + 0.4 * dz * dz / (p * p);
  const double axl2 = dz / (l * l);
  return 4.0 * (1.0 + axl2 * (0.33333333 + 0.1 * axl2)) * p * p * p / l;
}

/* Highest-level constant of trav::U, partf.cc:222: 1.405169521381833474e5
 * (4 pi e0) / (e Ryd^2)^2 in the CGS units of const.h. */
static double trav_f(void) {
  const double er = C_E * C_RYD * C_RYD;
  return 1.405169521381833474E+05 * ((4.0 * C_PI * C_E0) / (er * er));
}

/* U and dU/dne of one stage: pf::U (1), cpf::U, trav::U/dU (partf.cc:217-250,
 * including dU keeping only the last term's contribution), irwin::U/dU
 * (partf.cc:345-358). */
void orc_chem_partition(const snapi_partition *pf, double T, double ne, double *U, double *dU) {
  *U = 1.0;
  *dU = 0.0;
  if (pf->type == SNAPI_PF_CONST) {
    *U = pf->value;
  } else if (pf->type == SNAPI_PF_IRWIN) {
    const double lnT = log(T);
    double s = 0.0;
    for (int i = 0; i < pf->nterm; ++i) s += pf->irwin[i] * pow(lnT, (double)i);
    *U = exp(s);
  } else if (pf->type == SNAPI_PF_TRAV) {
    const double x = 5040.0 / T, f = trav_f(), z = pf->charge;
    const double p = z * f / pow(ne, 1.0 / 6.0);
    const double alpha = 31.321 * z * z * x;
    double sum = pf->g0;
    for (int i = 0; i < pf->nterm; ++i) {
      for (int l = 0; l < pf->ncoef[i]; ++l) sum += pf->alpha[i][l] * exp(-2.302585 * pf->gamma[i][l] * x);
      sum += pf->g2[i] * exp(-2.302585 * pf->xion[i] * x) * qas_f(pf->xl[i], p, alpha);
    }
    *U = sum;
    double dsum = 0.0;
    for (int i = 0; i < pf->nterm; ++i)
      dsum = -pf->g2[i] * exp(-2.302585 * pf->xion[i] * x) * dqas_f(pf->xl[i], p, alpha) * z * f /
             (6.0 * pow(ne, 7.0 / 6.0));
    *dU = dsum;
  }
}

/* atom::ionfrc, atom.cc:892-924: fractions fk[z] = N_z/N and dfk/dne. */
void orc_chem_ionfrc(const snapi_chem_element *el, double T, double ne, double *fk, double *df) {
  const int Z = el->nstage - 1;
  const double fct = 2.0 * pow((2.0 * C_PI * C_ME * C_K * T) / (C_H * C_H), 1.5);
  double Up, dUp;
  orc_chem_partition(&el->pf[0], T, ne, &Up, &dUp);
  dUp /= Up;
  double fsum = (fk[0] = 1.0);
  double dfsum = (df[0] = 0.0);
  for (int z = 0; z < Z; ++z) {
    double Un, dUn;
    orc_chem_partition(&el->pf[z + 1], T, ne, &Un, &dUn);
    dUn /= Un;
    const double f = fct * (Un / Up) * exp(-el->eion[z] / (C_K * T)) / ne;
    const double fd = f * ((dUn) - (dUp * ne + 1.0) / ne);
    fsum += (fk[z + 1] = fk[z] * f);
    dfsum += (df[z + 1] = (df[z] * f + fk[z] * fd));
    Up = Un;
    dUp = dUn;
  }
  for (int z = 0; z <= Z; ++z) {
    df[z] = (df[z] - dfsum * fk[z] / fsum) / fsum;
    fk[z] /= fsum;
  }
}

/* Dissociation constant: eqc_cnst::K, eqc_poly::K (10^poly(5040/T)) and
 * eqc_hmin::K (moleqc.cc); eqc (NONE) returns 0. */
double orc_chem_dissoc(const snapi_chem_molecule *m, double T) {
  const double x = 5040.0 / T;
  if (m->eqc == SNAPI_EQC_CONST) return m->coef[0];
  if (m->eqc == SNAPI_EQC_POLY) {
    double s = m->coef[m->ncoef - 1];
    for (int i = m->ncoef - 2; i >= 0; --i) s = m->coef[i] + x * s;
    return exp(2.302585093 * s);
  }
  if (m->eqc == SNAPI_EQC_HMIN) {
    const double s = 0.1249 + (2.5 * log(T) / 2.302585) - 0.747 * x;
    return exp(2.302585093 * s);
  }
  return 0.0;
}

/* Solves A v = d (n x n, row-major, destroyed) by Gaussian elimination with
 * partial pivoting; returns -1 for a zero or non-finite pivot. */
static int solve_pivot(int n, double *A, double *d, double *v) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int r = k + 1; r < n; ++r)
      if (fabs(A[r * n + k]) > fabs(A[p * n + k])) p = r;
    if (!(fabs(A[p * n + k]) > 0.0) || !isfinite(A[p * n + k])) return -1;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      double t = d[k];
      d[k] = d[p];
      d[p] = t;
    }
    for (int r = k + 1; r < n; ++r) {
      const double l = A[r * n + k] / A[k * n + k];
      for (int j = k + 1; j < n; ++j) A[r * n + j] -= l * A[k * n + j];
      d[r] -= l * d[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = d[k];
    for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * v[j];
    v[k] = s / A[k * n + k];
  }
  return 0;
}

/* LTE population of a level (atom::lte, atom.cc:779-787). */
static double lte_level(const snapi_chem_element *el, int stage, double g, double energy, double T, double ne,
                        double Nstage) {
  double U, dU;
  orc_chem_partition(&el->pf[stage], T, ne, &U, &dU);
  return Nstage * ((g != 0.0) ? g * exp(-energy / (C_K * T)) / U : 1.0);
}

/* Continuum opacity at the table's referent wavelength once chemeq and LTE
 * are done: atmosphere::opacity_continuum (atmos_rad.cc:286-294) -- Thomson
 * 6.65e-25 Ne, every element's bound-free transitions with stimulated
 * emission in table order (atom::opacity_continuum, atom_rad.cc:151-187,
 * whose free-free term is computed and then discarded), then H- bound-free
 * and free-free of element 1's neutral stage (h_minus_mol::opacity_continuum,
 * mol.cc:669-695).  stage_pop: every stage of every element, table order. */
double orc_chem_opref(const snapi_chem_table *tb, double T, double ne, const double *stage_pop) {
  const double lam = tb->lambda_ref;
  double op = 6.65E-25 * ne;
  int off[MAXE + 1];
  off[1] = 0;
  for (int a = 1; a < tb->nelement; ++a) off[a + 1] = off[a] + tb->element[a - 1].nstage;
  for (int a = 1; a <= tb->nelement; ++a) {
    const snapi_chem_element *el = &tb->element[a - 1];
    double part = 0.0;
    for (int j = 0; j < tb->nbf; ++j) {
      const snapi_chem_bf *b = &tb->bf[j];
      if (b->element != a || b->sigma == 0.0) continue;
      const int z = b->stage;
      const double nl = lte_level(el, z, b->g, b->energy, T, ne, stage_pop[off[a] + z]);
      const double nu = lte_level(el, z + 1, b->g_up, b->energy_up, T, ne, stage_pop[off[a] + z + 1]);
      const double pop_mod = nu * ne * C_SAHA * pow(T, -1.5) * b->g / b->g_up * exp((el->eion[z] - b->energy) / C_K / T);
      part += b->sigma * (nl - pop_mod * exp(-C_H * C_C / lam / C_K / T));
    }
    op += part;
  }
  if (tb->hminus) {
    const double NH = stage_pop[0];
    static const double a[7] = {1.99654, -1.18267E-5, 2.64243E-6, -4.40524E-10, 3.23992E-14, -1.39568E-18,
                                2.78701E-23};
    double alpha = 0.0;
    for (int i = 0; i < 7; ++i) alpha += a[i] * pow(lam * 1E8, i);
    double T_P = 5040. / T;
    const double op_bf = 4.158E-28 * alpha * ne * C_K * T * pow(T_P, 2.5) * pow(10.0, 0.754 * T_P) * NH *
                         (1.0 - exp(-C_H * C_C / lam / C_K / T));
    const double l = log10(lam * 1E8);
    const double f_0 = -2.2763 - 1.6850 * l + 0.76661 * l * l - 0.053346 * l * l * l;
    const double f_1 = 15.2827 - 9.2846 * l + 1.99381 * l * l - 0.142631 * l * l * l;
    const double f_2 = -197.789 + 190.266 * l - 67.9775 * l * l + 10.6913 * l * l * l - 0.625151 * l * l * l * l;
    T_P = 5040. / T;
    const double op_ff =
        1E-26 * ne * C_K * T * pow(10.0, f_0 + f_1 * log10(T_P) + f_2 * log10(T_P) * log10(T_P)) * NH;
    op += op_bf + op_ff;
  }
  return op;
}

/* One depth point: atmosphere::chemeq (atmos_chm.cc:28-363) then atom::lte
 * (atom.cc:779-787).  Outputs as snapi_chemeq for one point (stage, mol,
 * lev may be NULL).  Returns the Newton steps taken, or -1 (singular). */
static int chemeq_point(const snapi_chem_table *tb, double T, double Nt, double *ne_out, double *stage,
                        double *mol, double *lev, double *opref, double *n1_total) {
  const int na = tb->nelement, nm = tb->nmolecule, n1 = na + 1;
  const double kT = C_K * T;
  double abund[MAXE + 1], N[MAXE + 1], d[MAXE + 1], v[MAXE + 1], A[(MAXE + 1) * (MAXE + 1)];
  double fk[MAXE + 1][MAXS], df[MAXE + 1][MAXS];
  double K[MAXM], q[MAXM], nmd[MAXM], dm[MAXM][MAXE + 1];
  int mul[MAXM][MAXE + 1], nk[MAXE + 1];
  for (int a = 1; a <= na; ++a) {
    abund[a] = tb->element[a - 1].abundance;
    nk[a] = tb->element[a - 1].nstage;
  }
  nk[0] = 1;
  fk[0][0] = 1.0;
  df[0][0] = 0.0;
  for (int m = 0; m < nm; ++m) {  // atmos_chm.cc:79-95
    const snapi_chem_molecule *M = &tb->molecule[m];
    q[m] = 0.0;
    for (int a = 0; a <= na; ++a) mul[m][a] = 0;
    for (int i = 0; i < M->ncomp; ++i) {
      mul[m][M->element[i]] += 1;
      q[m] += (double)M->stage[i];
    }
    K[m] = orc_chem_dissoc(M, T);
  }
  // starting values: ionisation only, Ne = 1 and Ne = Nt (atmos_chm.cc:127-154)
  double F1 = 0.0, S1 = 0.0, F3 = 0.0, S3 = 0.0;
  for (int a = 1; a <= na; ++a) {
    orc_chem_ionfrc(&tb->element[a - 1], T, 1.0, fk[a], df[a]);
    double sum = 0.0;
    for (int i = 0; i < nk[a]; ++i) sum += i * fk[a][i];
    F1 += abund[a] * sum;
    S1 += abund[a];
  }
  for (int a = 1; a <= na; ++a) {
    orc_chem_ionfrc(&tb->element[a - 1], T, Nt, fk[a], df[a]);
    double sum = 0.0;
    for (int i = 0; i < nk[a]; ++i) sum += i * fk[a][i];
    F3 += abund[a] * sum;
    S3 += abund[a];
  }
  const double aa = F3 / (S3 * Nt);
  double bb = F1 / S1 - aa;
  const double cc = -bb * Nt;
  bb += 1.0 - aa * Nt;
  const double root = (-bb + sqrt(bb * bb - 4.0 * aa * cc)) / (2.0 * aa);
  N[0] = (root > 1.0) ? root : 1.0;  // cmlib/uts.h max(): a NaN root gives 1
  // averaging until the electron count agrees to 1% (atmos_chm.cc:157-169)
  double esum;
  do {
    esum = 0.0;
    for (int a = 1; a <= na; ++a) {
      orc_chem_ionfrc(&tb->element[a - 1], T, N[0], fk[a], df[a]);
      N[a] = (Nt - N[0]) * abund[a] / S3;
      for (int i = 0; i < nk[a]; ++i) esum += fk[a][i] * N[a] * (double)i;
    }
    N[0] = 0.5 * (N[0] + esum);
  } while ((1E-2 * (N[0] + esum)) < fabs(N[0] - esum));
  // Newton-Raphson (atmos_chm.cc:179-305)
  int steps = 0, iter = 0;
  for (;;) {
    for (int a = 1; a <= na; ++a) orc_chem_ionfrc(&tb->element[a - 1], T, N[0], fk[a], df[a]);
    for (int m = 0; m < nm; ++m) {
      const snapi_chem_molecule *M = &tb->molecule[m];
      nmd[m] = 1.0 / (K[m] * kT);
      for (int a = 0; a <= na; ++a) dm[m][a] = 0.0;
      for (int i = 0; i < M->ncomp; ++i) {
        const int a = M->element[i], z = M->stage[i] > 0 ? M->stage[i] : 0;
        nmd[m] *= fk[a][z] * N[a] * kT;
        dm[m][a] += 1.0 / N[a];
        dm[m][0] += df[a][z] / fk[a][z];
      }
      for (int a = 0; a <= na; ++a) dm[m][a] *= nmd[m];
    }
    // defects
    d[0] = -N[0];
    for (int a = 1; a <= na; ++a)
      for (int i = 0; i < nk[a]; ++i) d[0] += N[a] * fk[a][i] * (double)i;
    for (int m = 0; m < nm; ++m) d[0] += q[m] * nmd[m];
    d[1] = N[0] - Nt;
    for (int a = 1; a <= na; ++a) d[1] += N[a];
    for (int m = 0; m < nm; ++m) d[1] += nmd[m];
    double Nr = N[1];
    for (int m = 0; m < nm; ++m) Nr += nmd[m] * (double)mul[m][1];
    for (int a = 2; a <= na; ++a) {
      double Nc = N[a];
      for (int m = 0; m < nm; ++m) Nc += nmd[m] * (double)mul[m][a];
      d[a] = abund[a] * Nr - abund[1] * Nc;
    }
    // Jacobian, row-major A[row * n1 + col]
    A[0] = -1.0;
    for (int a = 1; a <= na; ++a)
      for (int i = 1; i < nk[a]; ++i) A[0] += N[a] * df[a][i] * (double)i;
    for (int m = 0; m < nm; ++m) A[0] += q[m] * dm[m][0];
    A[n1] = 1.0;
    for (int m = 0; m < nm; ++m) A[n1] += dm[m][0];
    double dNr = 0.0;
    for (int m = 0; m < nm; ++m) dNr += dm[m][0] * (double)mul[m][1];
    for (int a = 2; a <= na; ++a) {
      double dNc = 0.0;
      for (int m = 0; m < nm; ++m) dNc += dm[m][0] * (double)mul[m][a];
      A[a * n1] = abund[a] * dNr - abund[1] * dNc;
    }
    for (int b = 1; b <= na; ++b) {
      A[b] = 0.0;
      for (int i = 0; i < nk[b]; ++i) A[b] += fk[b][i] * (double)i;
      for (int m = 0; m < nm; ++m) A[b] += q[m] * dm[m][b];
      A[n1 + b] = 1.0;
      for (int m = 0; m < nm; ++m) A[n1 + b] += dm[m][b];
      dNr = 0.0;
      for (int m = 0; m < nm; ++m) dNr += dm[m][b] * (double)mul[m][1];
      if (b == 1) dNr += 1.0;
      for (int a = 2; a <= na; ++a) {
        double dNc = 0.0;
        for (int m = 0; m < nm; ++m) dNc += dm[m][b] * (double)mul[m][a];
        if (b == a) dNc += 1.0;
        A[a * n1 + b] = abund[a] * dNr - abund[1] * dNc;
      }
    }
    double dd[MAXE + 1];
    memcpy(dd, d, sizeof(double) * n1);
    if (solve_pivot(n1, A, dd, v)) {
      steps = -1;
      break;
    }
    for (int a = 0; a <= na; ++a) {
      N[a] -= v[a];
      if (N[a] < 1E-2) N[a] = 1E-2;
    }
    ++steps;
    // the reference counts two per step (one inside a log call) and stops above 100
    iter += 2;
    if (iter > 100) break;
    int conv = 1;
    for (int i = 0; i <= na; ++i)
      if (fabs(d[i] / N[i]) > 1E-4) conv = 0;
    if (conv) break;
  }
  *ne_out = N[0];
  // stage populations: the final N with the last iteration's fractions
  // (atom::pupdate, atom.cc:1449-1453)
  double st_all[MAXE * MAXS];
  {
    int s = 0;
    for (int a = 1; a <= na; ++a)
      for (int z = 0; z < nk[a]; ++z) st_all[s++] = N[a] * fk[a][z];
    if (stage) memcpy(stage, st_all, sizeof(double) * s);
  }
  if (opref) *opref = orc_chem_opref(tb, T, N[0], st_all);
  if (n1_total) *n1_total = N[1];
  if (mol) {
    for (int m = 0; m < nm; ++m) {
      const snapi_chem_molecule *M = &tb->molecule[m];
      double x = 1.0 / (K[m] * kT);
      for (int i = 0; i < M->ncomp; ++i) {
        const int a = M->element[i], z = M->stage[i] > 0 ? M->stage[i] : 0;
        x *= fk[a][z] * N[a] * kT;
      }
      mol[m] = x;
    }
  }
  if (lev) {
    for (int l = 0; l < tb->nlevel; ++l) {
      const snapi_chem_level *L = &tb->level[l];
      const snapi_chem_element *el = &tb->element[L->element - 1];
      const double Nz = N[L->element] * fk[L->element][L->stage];
      double U, dU;
      orc_chem_partition(&el->pf[L->stage], T, N[0], &U, &dU);
      lev[l] = Nz * ((L->g != 0.0) ? L->g * exp(-L->energy / kT) / U : 1.0);
    }
  }
  return steps;
}

int orc_chemeq_point(const snapi_chem_table *tb, double T, double Nt, double *ne_out, double *stage, double *mol,
                     double *lev, double *opref) {
  return chemeq_point(tb, T, Nt, ne_out, stage, mol, lev, opref, NULL);
}

int orc_chemeq(const snapi_chem_table *tb, int ncol, int nd, const double *T, const double *Nt, double *ne,
               double *stage, double *mol, double *lev, double *opref, int *iterations, int nthreads) {
  int nst = 0;
  for (int a = 0; a < tb->nelement; ++a) nst += tb->element[a].nstage;
  const int nm = tb->nmolecule, nl = tb->nlevel;
  const long np = (long)ncol * nd;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (long p = 0; p < np; ++p) {
    const long c = p / nd, z = p % nd;
    double st[MAXE * MAXS], mo[MAXM], le[SNAPI_CHEM_MAX_LEVELS], e, kr;
    const int it = orc_chemeq_point(tb, T[p], Nt[p], &e, st, mo, le, &kr);
    ne[p] = e;
    if (opref) opref[p] = kr;
    if (iterations) iterations[p] = it;
    if (stage)
      for (int s = 0; s < nst; ++s) stage[(c * nst + s) * nd + z] = st[s];
    if (mol)
      for (int m = 0; m < nm; ++m) mol[(c * nm + m) * nd + z] = mo[m];
    if (lev)
      for (int l = 0; l < nl; ++l) lev[(c * nl + l) * nd + z] = le[l];
  }
  return 0;
}

/* Hydrostatic equilibrium of one column on its optical-depth grid
 * (atmos_pp::enforce_hequilibrium, atmos/plpar/atmos_pp.cc:372-455): the top
 * point from a 0.3 dyn/cm^2 guess, iterated on Nt = 10^logtau rho g / kappa / kT;
 * then downwards, each point's pressure step g / kappa_mean * dtau with
 * kappa_mean the geometric mean of opacity per mass of the point and the one
 * above, iterated up to 20 times and left when the correction falls below
 * 1e-2 of the TOP point's Nt; every point's opacity and density are those of
 * its last iteration (before its last correction), and the final state is
 * chemical equilibrium at the final Nt.  rho = 1.4 * 1.67e-24 times the total
 * (free) density of element 1.  iters[z]: iterations run at depth z. */
#define C_GRAV 274.88E2
static int hequilibrium_march(const snapi_chem_table *tb, int nd, const double *T, const double *logtau, double *Nt,
                              double *ne, double *opref, double *rho, int *iters, double break_me, int max_iter) {
  if (nd < 1 || !(tb->lambda_ref > 0.0)) return -1;
  double e, op_prev = 0.0, rho_prev = 0.0;
  {
    Nt[0] = 0.3 / C_K / T[0];
    int it = 0;
    for (; it < max_iter;) {
      double op, nh;
      chemeq_point(tb, T[0], Nt[0], &e, NULL, NULL, NULL, &op, &nh);
      const double r = nh * 1.4 * 1.67E-24;
      op_prev = op;
      rho_prev = r;
      const double dN = (pow(10.0, logtau[0]) * r * C_GRAV / op) / C_K / T[0] - Nt[0];
      Nt[0] += dN;
      ++it;
      if (fabs(dN / Nt[0]) < break_me) break;
    }
    if (iters) iters[0] = it;
  }
  for (int z = 1; z < nd; ++z) {
    const double delta_tau = pow(10, logtau[z]) - pow(10, logtau[z - 1]);
    double delta_p = rho_prev * C_GRAV / op_prev * delta_tau;
    Nt[z] = (Nt[z - 1] * C_K * T[z - 1] + delta_p) / C_K / T[z];
    double op = op_prev, r = rho_prev;
    int it = 0;
    for (; it < max_iter;) {
      double nh;
      chemeq_point(tb, T[z], Nt[z], &e, NULL, NULL, NULL, &op, &nh);
      r = nh * 1.4 * 1.67E-24;
      const double kappa_mean = sqrt(op * op_prev / r / rho_prev);
      delta_p = C_GRAV / kappa_mean * delta_tau;
      const double dN = (Nt[z - 1] * C_K * T[z - 1] + delta_p) / C_K / T[z] - Nt[z];
      Nt[z] += dN;
      ++it;
      if (fabs(dN / Nt[0]) < break_me) break;
    }
    op_prev = op;
    rho_prev = r;
    if (iters) iters[z] = it;
  }
  for (int z = 0; z < nd; ++z) {
    double op, nh;
    chemeq_point(tb, T[z], Nt[z], &e, NULL, NULL, NULL, &op, &nh);
    if (ne) ne[z] = e;
    if (opref) opref[z] = op;
    if (rho) rho[z] = nh * 1.4 * 1.67E-24;
  }
  return 0;
}

int orc_hequilibrium(const snapi_chem_table *tb, int nd, const double *T, const double *logtau, double *Nt,
                     double *ne, double *opref, double *rho, int *iters) {
  return hequilibrium_march(tb, nd, T, logtau, Nt, ne, opref, rho, iters, 1E-2, 20);
}

/* The same march with another stopping rule (correction below tol of the
 * top point's Nt, at most max_iter iterations per depth): with a tight tol
 * it solves the hydrostatic equations to rounding, the reference for the
 * tangent below. */
int orc_hequilibrium_tol(const snapi_chem_table *tb, int nd, const double *T, const double *logtau, double *Nt,
                         double tol, int max_iter) {
  return hequilibrium_march(tb, nd, T, logtau, Nt, NULL, NULL, NULL, NULL, tol, max_iter);
}

/* Tangent of the hydrostatic equilibrium (no counterpart in the reference,
 * which differences two rebuilds): with P = Nt k T and r = rho / kappa the
 * mass per unit opacity, the converged march satisfies
 *   P_0 = tau_0 g r_0,   P_z = P_(z-1) + g (tau_z - tau_(z-1)) sqrt(r_z r_(z-1)).
 * Differentiated along a temperature perturbation dT[z], with d ln r / dT and
 * d ln r / dNt at each point by central differences of the chemical
 * equilibrium (T +- delta_T/2, Nt +- delta_Nt_frac Nt / 2, cmlib/const.h:
 * 27-29), every dNt[z] follows from the point above it:
 *   dNt_0 (k T_0 - Q_0 a_N0) = (Q_0 a_T0 - k Nt_0) dT_0,  Q_0 = tau_0 g r_0;
 *   dNt_z (k T_z - Q_z a_Nz / 2) = -k Nt_z dT_z + k (T_(z-1) dNt_(z-1) + Nt_(z-1) dT_(z-1))
 *        + Q_z / 2 (a_Tz dT_z + a_T(z-1) dT_(z-1) + a_N(z-1) dNt_(z-1)),
 *   Q_z = g (tau_z - tau_(z-1)) sqrt(r_z r_(z-1)).
 * T, logtau, Nt [nd]: a converged column; dT, dNt [ndir][nd]. */
int orc_he_tangent(const snapi_chem_table *tb, int nd, const double *T, const double *logtau, const double *Nt,
                   int ndir, const double *dT, double *dNt) {
  if (nd < 1 || ndir < 0 || !(tb->lambda_ref > 0.0)) return -1;
  double *r = (double *)malloc(sizeof(double) * 3 * nd), *aT = r + nd, *aN = r + 2 * nd;
  for (int z = 0; z < nd; ++z) {
    double e, op, nh, rv[5];
    const double dNs = 1E-3 * Nt[z];
    const double Tv[5] = {T[z], T[z] + 2.5, T[z] - 2.5, T[z], T[z]};
    const double Nv[5] = {Nt[z], Nt[z], Nt[z], Nt[z] + 0.5 * dNs, Nt[z] - 0.5 * dNs};
    for (int v = 0; v < 5; ++v) {
      chemeq_point(tb, Tv[v], Nv[v], &e, NULL, NULL, NULL, &op, &nh);
      rv[v] = nh * 1.4 * 1.67E-24 / op;
    }
    r[z] = rv[0];
    aT[z] = (rv[1] - rv[2]) / 5.0 / rv[0];
    aN[z] = (rv[3] - rv[4]) / dNs / rv[0];
  }
  for (int k = 0; k < ndir; ++k) {
    const double *dt = dT + (size_t)k * nd;
    double *dn = dNt + (size_t)k * nd;
    double tau_prev = pow(10.0, logtau[0]);
    const double Q0 = tau_prev * C_GRAV * r[0];
    dn[0] = (Q0 * aT[0] - C_K * Nt[0]) * dt[0] / (C_K * T[0] - Q0 * aN[0]);
    for (int z = 1; z < nd; ++z) {
      const double tau = pow(10.0, logtau[z]);
      const double Q = C_GRAV * (tau - tau_prev) * sqrt(r[z] * r[z - 1]);
      const double rhs = -C_K * Nt[z] * dt[z] + C_K * (T[z - 1] * dn[z - 1] + Nt[z - 1] * dt[z - 1]) +
                         0.5 * Q * (aT[z] * dt[z] + aT[z - 1] * dt[z - 1] + aN[z - 1] * dn[z - 1]);
      dn[z] = rhs / (C_K * T[z] - 0.5 * Q * aN[z]);
      tau_prev = tau;
    }
  }
  free(r);
  return 0;
}

int orc_hequilibrium_batch(const snapi_chem_table *tb, int ncol, int nd, const double *T, const double *logtau,
                           double *Nt, double *ne, double *opref, double *rho, int *iters, int nthreads) {
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1) reduction(| : bad)
  for (int c = 0; c < ncol; ++c) {
    const long o = (long)c * nd;
    bad |= orc_hequilibrium(tb, nd, T + o, logtau + o, Nt + o, ne ? ne + o : NULL, opref ? opref + o : NULL,
                            rho ? rho + o : NULL, iters ? iters + o : NULL) != 0;
  }
  return bad ? -1 : 0;
}

/* ---- the per-depth state of the synthesis and response paths ---- */

#define C_DELTA_T 5.0     /* delta_T, cmlib/const.h:27 */
#define C_DELTA_NTF 1E-3  /* delta_Nt_frac, cmlib/const.h:29 */

/* Planck_f (cmlib/mathtools.cc:1150-1152) with the reference's constants. */
static double planck_f(double lam, double T) {
  return 2.0 * 6.62606957 * 8.9875517E-7 / lam / lam / lam / lam / lam / (exp(1.4387769 / lam / T) - 1.0);
}

static void stage_offsets(const snapi_chem_table *tb, int *soff) {
  soff[1] = 0;
  for (int a = 1; a < tb->nelement; ++a) soff[a + 1] = soff[a] + tb->element[a - 1].nstage;
}

/* Thomson scattering at ne, T (atmosphere::op_em_vector, atmos_rad_synthesis.cc:
 * 20-50: emissivity at lambda_mid) plus the bound-free background of the
 * map's transitions at lambda_bg (atom::boundfree_op_em_vector,
 * atom_rad_synthesis.cc:226-262), whose level populations come from the stage
 * populations st at (Tl, nel), as atom::lte left them. */
static void background(const snapi_chem_table *tb, const snapi_state_map *map, const int *soff, double T, double ne,
                       const double *st, double Tl, double nel, double *op, double *em) {
  double o = ne * 6.65E-25;
  double e = o * planck_f(map->lambda_mid, T);
  const double lam = map->lambda_bg;
  for (int j = 0; j < tb->nbf; ++j) {
    const double sig = map->sigma_bg[j];
    if (sig == 0.0) continue;
    const snapi_chem_bf *b = &tb->bf[j];
    const snapi_chem_element *el = &tb->element[b->element - 1];
    const int z = b->stage;
    const double nl = lte_level(el, z, b->g, b->energy, Tl, nel, st[soff[b->element] + z]);
    const double nu = lte_level(el, z + 1, b->g_up, b->energy_up, Tl, nel, st[soff[b->element] + z + 1]);
    const double pop_mod = nu * ne * C_SAHA * pow(T, -1.5) * b->g / b->g_up * exp((el->eion[z] - b->energy) / C_K / T);
    const double stim = exp(-C_H * C_C / lam / C_K / T);
    o += sig * (nl - pop_mod * stim);
    e += sig * pop_mod * (1.0 - stim) * planck_f(lam, T);
  }
  *op = o;
  *em = e;
}

static double pf_U(const snapi_partition *pf, double T, double ne) {
  double U, dU;
  orc_chem_partition(pf, T, ne, &U, &dU);
  return U;
}

/* One column's state as the reference's synthesis path holds it after popsetup
 * and compute_op_referent, and -- with resp/dstate -- as its response path
 * holds it (obs_stokes_responses, atmos_rad_synthesis.cc:173-190):
 *   compute_op_referent_derivative (atmos_ppbez_rts.cc:283-325): central
 *     differences of the referent opacity, T +- delta_T/2 and Nt +- delta_Nt/2;
 *   ne_lte_derivatives (atmos_responses.cc:364-423): Ne and stage populations
 *     by a forward difference in T (T + delta_T against T) and a central one
 *     in Nt at T + delta_T/2, after which T stays raised by delta_T/2 with Ne
 *     and the stage populations re-solved there, while the level populations
 *     and the referent opacity keep their values at T;
 *   atom::compute_lte_population_responses_analytical (atom_responses.cc:
 *     453-501): level responses from the stage responses and the Boltzmann
 *     factor, with U differenced at +- delta_T/2 and +- 1e-3 Ne / 2.
 * state [nfield][nd]: T and NT in; NE, NHI, OPREF, BGOP, BGEM and the line
 * level populations out.  resp [nfield][nd] (may be NULL): the whole state as
 * the response path sees it.  dstate [2][nfield][nd] (may be NULL): d/dT and
 * d/dNt of NE, NHI, OPREF and the line level populations, zero elsewhere. */
int orc_eos_state(const snapi_chem_table *tb, const snapi_state_map *map, int nd, double *state, double *resp,
                  double *dstate) {
  const int nf = SNAPI_NF_BASE + 2 * map->nline;
  int soff[MAXE + 2];
  stage_offsets(tb, soff);
  const int ih = soff[map->nhi_element];
  if (dstate) memset(dstate, 0, sizeof(double) * 2 * (size_t)nf * nd);
  for (int zd = 0; zd < nd; ++zd) {
    const double T0 = state[SNAPI_F_T * nd + zd], Nt0 = state[SNAPI_F_NT * nd + zd];
    double st0[MAXE * MAXS], ne0, op0;
    chemeq_point(tb, T0, Nt0, &ne0, st0, NULL, NULL, &op0, NULL);
    double bo, be;
    background(tb, map, soff, T0, ne0, st0, T0, ne0, &bo, &be);
    state[SNAPI_F_NE * nd + zd] = ne0;
    state[SNAPI_F_NHI * nd + zd] = st0[ih];
    state[SNAPI_F_OPREF * nd + zd] = op0;
    state[SNAPI_F_BGOP * nd + zd] = bo;
    state[SNAPI_F_BGEM * nd + zd] = be;
    for (int i = 0; i < map->nline; ++i)
      for (int u = 0; u < 2; ++u) {
        const snapi_chem_level *L = &tb->level[map->line_level[i][u]];
        state[(SNAPI_NF_BASE + 2 * i + u) * nd + zd] =
            lte_level(&tb->element[L->element - 1], L->stage, L->g, L->energy, T0, ne0, st0[soff[L->element] + L->stage]);
      }
    if (resp)
      for (int f = 0; f < nf; ++f) resp[f * nd + zd] = state[f * nd + zd];
    if (!resp && !dstate) continue;
    double T = T0, Nt = Nt0, e, op_a, op_b;
    /* compute_op_referent_derivative */
    T += C_DELTA_T * 0.5;
    chemeq_point(tb, T, Nt, &e, NULL, NULL, NULL, &op_a, NULL);
    T -= C_DELTA_T;
    chemeq_point(tb, T, Nt, &e, NULL, NULL, NULL, &op_b, NULL);
    const double dop_dT = (op_a - op_b) / C_DELTA_T;
    T += C_DELTA_T * 0.5;
    const double dNt0 = Nt * C_DELTA_NTF;
    Nt += dNt0 * 0.5;
    chemeq_point(tb, T, Nt, &e, NULL, NULL, NULL, &op_a, NULL);
    Nt -= dNt0;
    chemeq_point(tb, T, Nt, &e, NULL, NULL, NULL, &op_b, NULL);
    const double dop_dNt = (op_a - op_b) / dNt0;
    Nt += dNt0 * 0.5;
    /* ne_lte_derivatives */
    double sa[MAXE * MAXS], sb[MAXE * MAXS], dsT[MAXE * MAXS], dsN[MAXE * MAXS], ne_a, ne_b;
    const int nst = soff[tb->nelement] + tb->element[tb->nelement - 1].nstage;
    T += C_DELTA_T;
    chemeq_point(tb, T, Nt, &ne_a, sa, NULL, NULL, NULL, NULL);
    T -= C_DELTA_T;
    chemeq_point(tb, T, Nt, &ne_b, sb, NULL, NULL, NULL, NULL);
    const double dne_dT = (ne_a - ne_b) / C_DELTA_T;
    for (int s = 0; s < nst; ++s) dsT[s] = (sa[s] - sb[s]) / C_DELTA_T;
    T += C_DELTA_T * 0.5;
    const double dNt1 = C_DELTA_NTF * Nt;
    Nt += dNt1 * 0.5;
    chemeq_point(tb, T, Nt, &ne_a, sa, NULL, NULL, NULL, NULL);
    Nt -= dNt1;
    chemeq_point(tb, T, Nt, &ne_b, sb, NULL, NULL, NULL, NULL);
    const double dne_dNt = (ne_a - ne_b) / dNt1;
    for (int s = 0; s < nst; ++s) dsN[s] = (sa[s] - sb[s]) / dNt1;
    Nt += dNt1 * 0.5;
    double ste[MAXE * MAXS], ne_e;
    chemeq_point(tb, T, Nt, &ne_e, ste, NULL, NULL, NULL, NULL);
    if (resp) {
      background(tb, map, soff, T, ne_e, st0, T0, ne0, &bo, &be);
      resp[SNAPI_F_T * nd + zd] = T;
      resp[SNAPI_F_NT * nd + zd] = Nt;
      resp[SNAPI_F_NE * nd + zd] = ne_e;
      resp[SNAPI_F_NHI * nd + zd] = ste[ih];
      resp[SNAPI_F_BGOP * nd + zd] = bo;
      resp[SNAPI_F_BGEM * nd + zd] = be;
    }
    if (!dstate) continue;
    double *dT = dstate, *dN = dstate + (size_t)nf * nd;
    dT[SNAPI_F_NE * nd + zd] = dne_dT;
    dN[SNAPI_F_NE * nd + zd] = dne_dNt;
    dT[SNAPI_F_NHI * nd + zd] = dsT[ih];
    dN[SNAPI_F_NHI * nd + zd] = dsN[ih];
    dT[SNAPI_F_OPREF * nd + zd] = dop_dT;
    dN[SNAPI_F_OPREF * nd + zd] = dop_dNt;
    /* compute_lte_population_responses_analytical at the raised T */
    for (int i = 0; i < map->nline; ++i)
      for (int u = 0; u < 2; ++u) {
        const snapi_chem_level *L = &tb->level[map->line_level[i][u]];
        const snapi_partition *pf = &tb->element[L->element - 1].pf[L->stage];
        const int s = soff[L->element] + L->stage;
        const double U = pf_U(pf, T, ne_e);
        double dU_dT = pf_U(pf, T + 0.5 * C_DELTA_T, ne_e);
        dU_dT -= pf_U(pf, T - 0.5 * C_DELTA_T, ne_e);
        dU_dT /= C_DELTA_T;
        const double dne = C_DELTA_NTF * ne_e;
        double dU_dne = pf_U(pf, T, ne_e + dne * 0.5);
        dU_dne -= pf_U(pf, T, ne_e - dne * 0.5);
        dU_dne /= dne;
        const double bz = L->g * exp(-L->energy / (C_K * T));
        double rT = dsT[s] * bz / U;
        rT -= ste[s] * bz / U / U * (dU_dT + dU_dne * dne_dT);
        rT += ste[s] * bz / U * L->energy / C_K / T / T;
        double rN = dsN[s] * bz / U;
        rN -= ste[s] * bz / U / U * (dU_dne * dne_dNt);
        dT[(SNAPI_NF_BASE + 2 * i + u) * nd + zd] = rT;
        dN[(SNAPI_NF_BASE + 2 * i + u) * nd + zd] = rN;
      }
  }
  return 0;
}

int orc_eos_state_batch(const snapi_chem_table *tb, const snapi_state_map *map, int ncol, int nd, double *state,
                        double *resp, double *dstate, int nthreads) {
  const size_t nf = SNAPI_NF_BASE + 2 * (size_t)map->nline;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int c = 0; c < ncol; ++c)
    orc_eos_state(tb, map, nd, state + c * nf * nd, resp ? resp + c * nf * nd : NULL,
                  dstate ? dstate + c * 2 * nf * nd : NULL);
  return 0;
}
