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
 * charge and particle conservation at the result.
 */
#include <math.h>
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

/* One depth point: atmosphere::chemeq (atmos_chm.cc:28-363) then atom::lte
 * (atom.cc:779-787).  Outputs as snapi_chemeq for one point (stage, mol,
 * lev may be NULL).  Returns the Newton steps taken, or -1 (singular). */
int orc_chemeq_point(const snapi_chem_table *tb, double T, double Nt, double *ne_out, double *stage, double *mol,
                     double *lev) {
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
  if (stage) {
    int s = 0;
    for (int a = 1; a <= na; ++a)
      for (int z = 0; z < nk[a]; ++z) stage[s++] = N[a] * fk[a][z];
  }
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

int orc_chemeq(const snapi_chem_table *tb, int ncol, int nd, const double *T, const double *Nt, double *ne,
               double *stage, double *mol, double *lev, int *iterations, int nthreads) {
  int nst = 0;
  for (int a = 0; a < tb->nelement; ++a) nst += tb->element[a].nstage;
  const int nm = tb->nmolecule, nl = tb->nlevel;
  const long np = (long)ncol * nd;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (long p = 0; p < np; ++p) {
    const long c = p / nd, z = p % nd;
    double st[MAXE * MAXS], mo[MAXM], le[SNAPI_CHEM_MAX_LEVELS], e;
    const int it = orc_chemeq_point(tb, T[p], Nt[p], &e, st, mo, le);
    ne[p] = e;
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
