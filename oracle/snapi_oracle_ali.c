/*
 * FP64 CPU restatement of the reference's ALI iteration for one NLTE atom:
 * the checker of the HIP kernels (test infrastructure only; never linked into
 * the product).
 *
 *   orc_ali_accumulate  atmosphere::nltepops's ray/wavelength loop
 *                       (atmos/atmos_pop.cc:133-195) with the atom's scalar
 *                       opacity (atom_rad.cc:34-60: bound-free 312-437,
 *                       bound-bound 439-537) and atom::add (atom.cc:1865-1927)
 *   orc_ali_newpops     atom::newpops (atom_pop.cc:27-135) with R_ij,
 *                       R_ij_local_ALO, R_i_cont, R_cont_i
 *                       (atom_rad.cc:1727-1855) and the pivot-free Crout
 *                       solve (cmlib/mathtools.cc:487-514)
 *   orc_ali_solve       the iteration with the Lambda-iteration fallback and
 *                       the relative-change stop (atmos_pop.cc:198-214,
 *                       atmosphere::newpops atmos_pop.cc:339-412)
 * The population-independent opacity (LTE species, the atom's free-free and
 * Rayleigh terms), the line damping and the collision rates are inputs, as
 * the reference computes them once per column.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "snapi_oracle.h"

#define C_LIGHT 2.99792458E10
#define H_PLANCK 6.6260687652E-27
#define K_BOLTZ 1.3806503E-16
#define PI_D 3.141592653589793238462643383279502884
#define SAHA_CONST 2.0706615382E-16

/* one column; see the header for the layouts */
int orc_ali_accumulate(const snapi_ali_atom *at, const snapi_ali_grid *g, int nd, const double *cs,
                       const double *vproj, const double *gamma, const double *bgop, const double *bgem,
                       const double *pops, double *acc) {
  const int nl = g->nlam, nt = at->ntrans, nlev = at->nlevel;
  double *op = (double *)malloc(sizeof(double) * nd), *em = (double *)malloc(sizeof(double) * nd);
  double *opn = (double *)malloc(sizeof(double) * nd), *emn = (double *)malloc(sizeof(double) * nd);
  double *I = (double *)malloc(sizeof(double) * nd), *L = (double *)malloc(sizeof(double) * nd);
  double *prof = (double *)malloc(sizeof(double) * nd * nt);
  const double *T = cs + SNAPI_A_T * nd, *vt = cs + SNAPI_A_VT * nd, *ne = cs + SNAPI_A_NE * nd;
  const double *opref = cs + SNAPI_A_OPREF * nd, *grid = cs + SNAPI_A_GRID * nd;
  memset(acc, 0, sizeof(double) * nd * nt * 4);
  for (int r = 0; r < g->nray; ++r) {
    const double *bo = bgop + (size_t)(g->bg_per_ray ? r : 0) * nd * nl;
    const double *be = bgem + (size_t)(g->bg_per_ray ? r : 0) * nd * nl;
    const double *vr = vproj + (size_t)r * nd;
    for (int l = 0; l < nl; ++l) {
      const double lam = g->lam[l], fud = g->fudge[l];
      for (int z = 0; z < nd; ++z) {
        const double *n = pops + (size_t)z * nlev;
        double o = bo[(size_t)z * nl + l], e = be[(size_t)z * nl + l];
        const double Tz = T[z];
        for (int t = 0; t < nt; ++t) {
          const snapi_ali_transition *tr = &at->tr[t];
          prof[(size_t)z * nt + t] = 0.0;
          if (tr->kind == SNAPI_TR_BF) {
            const double sig = g->sigma[(size_t)t * nl + l];
            if (sig == 0.0) continue;
            const double base = n[tr->upper] * ne[z] * SAHA_CONST * pow(Tz, -1.5) * tr->gratio;
            const double pm = base * exp((tr->de - H_PLANCK * C_LIGHT / lam) / K_BOLTZ / Tz);
            o += sig * (n[tr->lower] - pm) * fud;
            const double pm2 = base * exp(tr->de / K_BOLTZ / Tz);
            e += sig * pm2 * (1.0 - exp(-H_PLANCK * C_LIGHT / lam / K_BOLTZ / Tz)) * orc_planck(lam, Tz) * fud;
          } else {
            const double l0 = tr->lam0;
            const double dld = sqrt(2.0 * K_BOLTZ * Tz / at->mass + vt[z] * vt[z]) * l0 / C_LIGHT;
            const double sf = l0 * l0 / (4.0 * PI_D * C_LIGHT);
            const double a = sf * gamma[(size_t)t * nd + z] / dld;
            const double x = (l0 - lam * (1.0 + vr[z] / C_LIGHT)) / dld;
            double H, F;
            orc_voigt(x, a, &H, &F);
            const double p = 1.0 / dld * H;
            prof[(size_t)z * nt + t] = p;
            const double nlo = n[tr->lower], nup = n[tr->upper];
            o += p * nlo * tr->blu * H_PLANCK * C_LIGHT / lam / 4.0 / PI_D * (1.0 - nup / nlo * tr->bul / tr->blu);
            e += p * nup * tr->aul * H_PLANCK * C_LIGHT / lam / 4.0 / PI_D;
          }
        }
        op[z] = o;
        em[z] = e;
        opn[z] = g->tau_grid ? o / opref[z] : o;
        emn[z] = g->tau_grid ? e / opref[z] : e;
      }
      orc_formal_scalar(nd, grid, opn, emn, g->theta[r], g->boundary, I, L);
      const double ww = g->lamw[l] * g->wray[r];
      for (int z = 0; z < nd; ++z) {
        const double *n = pops + (size_t)z * nlev;
        const double opd = g->tau_grid ? opn[z] * opref[z] : op[z];
        for (int t = 0; t < nt; ++t) {
          const snapi_ali_transition *tr = &at->tr[t];
          double *A = acc + ((size_t)z * nt + t) * 4;
          const double p = prof[(size_t)z * nt + t];
          A[3] += ww * p * 0.25 / PI_D;
          if (tr->kind == SNAPI_TR_LINE) {
            const double lop = (n[tr->lower] * tr->blu - n[tr->upper] * tr->bul) * p * tr->energy * 0.25 / PI_D;
            const double f = lop / opd;
            const double c = ww * I[z] * p * f * 0.25 / PI_D;
            A[0] += c;
            A[1] += c;
            A[2] += ww * L[z] * p * f * 0.25 / PI_D;
          } else {
            const double sig = g->sigma[(size_t)t * nl + l];
            const double lam = g->lam[l];
            A[0] += ww * sig * lam / H_PLANCK / C_LIGHT * I[z];
            A[1] += ww * sig * lam / H_PLANCK / C_LIGHT *
                    (I[z] + 2.0 * H_PLANCK * C_LIGHT * C_LIGHT / pow(lam, 5.0)) *
                    exp(-H_PLANCK * C_LIGHT / lam / K_BOLTZ / T[z]);
          }
        }
      }
    }
  }
  free(op); free(em); free(opn); free(emn); free(I); free(L); free(prof);
  return 0;
}

/* pivot-free Crout LU (cmlib/mathtools.cc:487-514: unit upper triangle,
 * diagonal in L) and the solve, for an n x n row-major matrix */
static void crout_solve(int n, const double *A, const double *b, double *x) {
  double *D = (double *)malloc(sizeof(double) * n * n), *y = (double *)malloc(sizeof(double) * n);
  for (int k = 0; k < n; ++k) {
    for (int i = k; i < n; ++i) {
      double sum = 0.0;
      for (int p = 0; p < k; ++p) sum += D[i * n + p] * D[p * n + k];
      D[i * n + k] = A[i * n + k] - sum;
    }
    for (int j = k + 1; j < n; ++j) {
      double sum = 0.0;
      for (int p = 0; p < k; ++p) sum += D[k * n + p] * D[p * n + j];
      D[k * n + j] = (A[k * n + j] - sum) / D[k * n + k];
    }
  }
  for (int i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int p = 0; p < i; ++p) sum += D[i * n + p] * y[p];
    y[i] = (b[i] - sum) / D[i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double sum = 0.0;
    for (int p = i + 1; p < n; ++p) sum += D[i * n + p] * x[p];
    x[i] = y[i] - sum;
  }
  free(D);
  free(y);
}

/* line transition between mapped levels i and j, or -1 */
static int line_between(const snapi_ali_atom *at, int i, int j) {
  for (int t = 0; t < at->ntrans; ++t)
    if (at->tr[t].kind == SNAPI_TR_LINE &&
        ((at->tr[t].lower == i && at->tr[t].upper == j) || (at->tr[t].lower == j && at->tr[t].upper == i)))
      return t;
  return -1;
}

/* radiative rate from -> to of line t: R_ij (alo = 0) or R_ij_local_ALO */
static double line_rate(const snapi_ali_atom *at, int t, int from, int to, double JJ, double LL, double nfrom,
                        double nto, int alo) {
  const snapi_ali_transition *tr = &at->tr[t];
  const int down = (from == tr->upper);
  const double A = down ? tr->aul : 0.0, B = down ? tr->bul : tr->blu;
  if (!alo) return A + B * JJ;
  const double nu = down ? nfrom : nto, nl = down ? nto : nfrom;
  const double S = nu * tr->aul / (nl * tr->blu - nu * tr->bul);
  return A * (1.0 - LL) + B * (JJ - LL * S);
}

/* new populations of one column (atom::newpops at every depth); returns the
 * largest |new/old - 1| or -1 if a population came out negative */
double orc_ali_newpops(const snapi_ali_atom *at, int nd, const double *cs, const double *coll, const double *pops,
                       const double *acc, int alo, double *newp) {
  const int n = at->nlevel, nt = at->ntrans;
  double *M = (double *)malloc(sizeof(double) * n * n), *b = (double *)malloc(sizeof(double) * n);
  int *bf_of = (int *)malloc(sizeof(int) * n), *stage = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) bf_of[i] = -1;
  for (int t = 0; t < nt; ++t)
    if (at->tr[t].kind == SNAPI_TR_BF) bf_of[at->tr[t].lower] = t;
  /* stage index: levels up to and including the continuum target share a stage */
  {
    int s = 0;
    for (int i = 0; i < n; ++i) {
      stage[i] = s;
      if (i + 1 < n && bf_of[i] >= 0 && at->tr[bf_of[i]].upper == i + 1) ++s;
    }
  }
  double maxrc = 0.0;
  int negative = 0;
  for (int z = 0; z < nd; ++z) {
    const double *p = pops + (size_t)z * n;
    const double *A4 = acc + (size_t)z * nt * 4;
    const double *C = coll + (size_t)z * n * n;
    const double T = cs[SNAPI_A_T * nd + z], ne = cs[SNAPI_A_NE * nd + z];
    memset(M, 0, sizeof(double) * n * n);
    memset(b, 0, sizeof(double) * n);
    for (int i = 0; i < n - 1; ++i) {
      for (int j = 0; j < n; ++j) {
        if (stage[j] != stage[i] || j == i) continue;
        const int t = line_between(at, i, j);
        double Rout = 0.0, Rin = 0.0;
        if (t >= 0) {
          const double JJ = A4[t * 4 + 0] / A4[t * 4 + 3], LL = A4[t * 4 + 2] / A4[t * 4 + 3];
          Rout = line_rate(at, t, i, j, JJ, LL, p[i], p[j], alo);
          Rin = line_rate(at, t, j, i, JJ, LL, p[j], p[i], alo);
        }
        M[i * n + i] -= (Rout + C[i * n + j]);
        M[i * n + j] += (Rin + C[j * n + i]);
      }
      const int t = bf_of[i];
      if (t >= 0) {
        const int ic = at->tr[t].upper;
        double R = A4[t * 4 + 0];
        M[i * n + i] -= (R + C[i * n + ic]);
        M[ic * n + i] += (R + C[i * n + ic]);
        R = A4[t * 4 + 1] * at->tr[t].gratio * exp(at->tr[t].de / K_BOLTZ / T) * pow(T, -1.5) * ne * SAHA_CONST;
        M[i * n + ic] += (R + C[ic * n + i]);
        M[ic * n + ic] -= (R + C[ic * n + i]);
      }
    }
    int rep = n - 1;
    double maxpop = p[rep];
    for (int i = 0; i < n - 1; ++i)
      if (p[i] > maxpop) {
        maxpop = p[i];
        rep = i;
      }
    for (int j = 0; j < n; ++j) M[rep * n + j] = 1.0;
    b[rep] = cs[SNAPI_A_NA * nd + z];
    crout_solve(n, M, b, newp + (size_t)z * n);
    for (int i = 0; i < n; ++i) {
      const double rc = newp[(size_t)z * n + i] / p[i] - 1.0;
      if (rc < -1.0) negative = 1;
      if (fabs(rc) > maxrc) maxrc = fabs(rc);
    }
  }
  free(M); free(b); free(bf_of); free(stage);
  return negative ? -1.0 : maxrc;
}

int orc_ali_solve(const snapi_ali_atom *at, const snapi_ali_grid *g, int ncol, int nd, const double *colstate,
                  const double *vproj, const double *gamma, const double *bgop, const double *bgem,
                  const double *coll, double tol, int max_iter, double *pops, int *iterations, double *relchange,
                  int nthreads) {
  const int n = at->nlevel, nt = at->ntrans, nl = g->nlam;
  const size_t bgs = (size_t)(g->bg_per_ray ? g->nray : 1) * nd * nl;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (int col = 0; col < ncol; ++col) {
    const double *cs = colstate + (size_t)col * SNAPI_A_NF * nd;
    double *p = pops + (size_t)col * nd * n;
    double *acc = (double *)malloc(sizeof(double) * nd * nt * 4), *np = (double *)malloc(sizeof(double) * nd * n);
    double rc = 1.0;
    int it;
    for (it = 0; it < max_iter; ++it) {
      orc_ali_accumulate(at, g, nd, cs, vproj + (size_t)col * g->nray * nd, gamma + (size_t)col * nt * nd,
                         bgop + (size_t)col * bgs, bgem + (size_t)col * bgs, p, acc);
      rc = orc_ali_newpops(at, nd, cs, coll + (size_t)col * nd * n * n, p, acc, 1, np);
      if (rc < 0.0) rc = orc_ali_newpops(at, nd, cs, coll + (size_t)col * nd * n * n, p, acc, 0, np);
      if (rc < 0.0) break;
      memcpy(p, np, sizeof(double) * nd * n);
      if (rc < tol) break;
    }
    iterations[col] = it + 1 > max_iter ? max_iter : it + 1;
    relchange[col] = rc;
    free(acc);
    free(np);
  }
  return 0;
}
