/*
 * FP64 CPU restatement of the two numerical kernels of the reference's NLTE
 * population responses -- TEST INFRASTRUCTURE ONLY (see snapi_oracle.h).
 *
 *   formal_with_responses  atmos_ppbez::formal_with_responses,
 *                          atmos/plpar/atmos_ppbez_rts.cc:349-778
 *   response system solve  the Crout factorisation and solves of
 *                          atom::compute_nlte_population_responses,
 *                          atmos/atmol/atom/atom_responses.cc:171-200, with
 *                          Crout/solveCrout of cmlib/mathtools.cc:487-514
 *
 * The reference's formal_with_responses runs with bezier_delimiter = 1e10
 * (atmos_ppbez_rts.cc:13): the forward optical-depth step never exceeds it
 * on a tau grid, so the weights are the linear ones, w2 = 0 and C = 0, and
 * every term that w2 or C multiplies is zero; those terms are left out here.
 *
 * Pinned: the reference build's own intensities, responses to opacity and
 * emissivity (12 wavelength/ray problems of the Ca II 5-level atom in the
 * cfg0 column), and its assembled 366 x 366 system, right-hand sides and
 * level responses (tests/golden/nlteresp.npz); a constant source function
 * gives I = S with zero response to uniform scaling of op and em together;
 * the responses equal central differences of the formal solution.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "snapi_oracle.h"

/* d(dtau)/d(op) and the optical-depth steps of one ray, as the reference
 * forms them (atmos_ppbez_rts.cc:415-510).  dX is nd x nd. */
static void delta_tau_and_derivative(int nd, const double *g, const double *oo, double ct, int zl, int zh, int dz,
                                     double *dtau, double *dX) {
  double *D = calloc(nd, sizeof(double));
  double *dD = calloc((size_t)nd * nd, sizeof(double));
  D[zl] = (oo[zl + dz] - oo[zl]) / (g[zl + dz] - g[zl]) * (-ct);
  D[zh] = (oo[zh] - oo[zh - dz]) / (g[zh] - g[zh - dz]) * (-ct);
  for (int z = zl + dz; (z - zh) * dz < 0; z += dz) {
    D[z] = 0.0;
    const double h2 = (g[z + dz] - g[z]) / (-ct);
    const double h1 = (g[z] - g[z - dz]) / (-ct);
    const double d2 = (oo[z + dz] - oo[z]) / h2;
    const double d1 = (oo[z] - oo[z - dz]) / h1;
    if (d1 * d2 > 0.0) {
      const double alpha = (1.0 + h1 / (h1 + h2)) / 3.0;
      D[z] = d1 * d2 / (alpha * d1 + (1.0 - alpha) * d2);
    }
  }
  dtau[zl] = 0.0;
  for (int z = zl + dz; (z - zh - dz) * dz < 0; z += dz) {
    const double C0 = oo[z] - (g[z] - g[z - dz]) / (-ct) / 2.0 * D[z];
    const double C1 = oo[z - dz] + (g[z] - g[z - dz]) / (-ct) / 2.0 * D[z - dz];
    const double C = 0.5 * (C0 + C1);
    dtau[z] = (g[z] - g[z - dz]) / (-ct) / 3.0 * (oo[z - dz] + oo[z] + C);
  }
  // derivative of the Fritsch-Butland derivative with respect to the opacities
  dD[zl * nd + zl] = -1.0 / (g[zl + dz] - g[zl]) * (-ct);
  dD[zl * nd + zl + dz] = 1.0 / (g[zl + dz] - g[zl]) * (-ct);
  dD[zh * nd + zh - dz] = -1.0 / (g[zh] - g[zh - dz]) * (-ct);
  dD[zh * nd + zh] = 1.0 / (g[zh] - g[zh - dz]) * (-ct);
  for (int i = zl + dz; (i - zh) * dz < 0; i += dz) {
    const double h2 = (g[i + dz] - g[i]) / (-ct);
    const double h1 = (g[i] - g[i - dz]) / (-ct);
    const double d2 = (oo[i + dz] - oo[i]) / h2;
    const double d1 = (oo[i] - oo[i - dz]) / h1;
    const double alpha = (1.0 + h1 / (h1 + h2)) / 3.0;
    if (d1 * d2 > 0.0) {
      const double den = (alpha * d1 + (1.0 - alpha) * d2);
      dD[i * nd + i - dz] = (-1.0 / h1 * d2 * den + alpha / h1 * d1 * d2) / den / den;
      dD[i * nd + i] = (1.0 / h1 * d2 * den - 1.0 / h2 * d1 * den - alpha / h1 * d1 * d2 + (1.0 - alpha) / h2 * d1 * d2) /
                       den / den;
      dD[i * nd + i + dz] = (1.0 / h2 * d1 * den - (1.0 - alpha) / h2 * d1 * d2) / den / den;
    }
  }
  memset(dX, 0, sizeof(double) * (size_t)nd * nd);
  for (int i = zl + dz; (i - zh - dz) * dz < 0; i += dz) {
    const double hi = (g[i] - g[i - dz]) / (-ct);
    dX[i * nd + i] = 1.0 / 2.0 * hi;
    dX[i * nd + i - dz] = 1.0 / 2.0 * hi;
    for (int j = zl; (j - zh - dz) * dz < 0; j += dz)
      dX[i * nd + j] += hi * hi / 12.0 * dD[(i - dz) * nd + j] - hi * hi / 12.0 * dD[i * nd + j];
  }
  free(D);
  free(dD);
}

/* Linear short-characteristics weights and their derivatives with respect to
 * the step (atmos_ppbez_rts.cc:615-636, 705-717). */
static void linear_weights(double tb, double *etb, double *w0, double *w1, double *dw0, double *dw1) {
  *etb = (tb < 1E-2) ? (1.0 - tb + tb * tb * 0.5 - tb * tb * tb / 6.0) : exp(-tb);
  if (tb < 1E-2) {
    *w0 = tb * (0.5 - tb * (1.0 / 3.0 - tb / 8.0));
    *dw0 = 1.0 / 2.0 - tb * (2.0 / 3.0 - 3.0 * tb / 8.0);
    *w1 = 0.5 * tb * (1.0 - tb / 3.0 * (1.0 - tb / 4.0));
    *dw1 = 0.5 - tb * (1.0 / 3.0 - tb / 8.0);
  } else {
    *w0 = 1.0 / tb - *etb * (1.0 + 1.0 / tb);
    *dw0 = -*w0 / tb + *etb;
    *w1 = 1.0 - 1.0 / tb + *etb / tb;
    *dw1 = (1.0 - *w1) / tb - *etb / tb;
  }
}

int orc_formal_responses(int nd, const double *g, const double *oo, const double *ee, double theta, double boundary,
                         double *I, double *L, double *Rop, double *Rem) {
  if (nd < 2) return -1;
  const double ct = cos(theta);
  const int dz = (ct > 0) ? 1 : -1, zl = (ct > 0) ? 0 : nd - 1, zh = (ct > 0) ? nd - 1 : 0;
  memset(Rop, 0, sizeof(double) * (size_t)nd * nd);
  memset(Rem, 0, sizeof(double) * (size_t)nd * nd);
  I[zl] = (ct > 0) ? 0.0 : ee[zl] / oo[zl];
  if (ct < 0 && boundary >= 0) I[zl] = boundary;
  if (ct >= 0) L[zl] = 0.0;
  if (ct <= 0) {
    L[zl] = (boundary == -1) ? 1.0 : 0.0;
    if (boundary == -1) {
      Rop[zl * nd + zl] = -ee[zl] / oo[zl] / oo[zl];
      Rem[zl * nd + zl] = 1.0 / oo[zl];
    }
  }
  double *dtau = malloc(sizeof(double) * nd);
  double *dX = malloc(sizeof(double) * (size_t)nd * nd);
  delta_tau_and_derivative(nd, g, oo, ct, zl, zh, dz, dtau, dX);
  for (int z = zl + dz; (z - zh - dz) * dz < 0; z += dz) {
    const double tb = dtau[z];
    double etb, w0, w1, dw0, dw1;
    linear_weights(tb, &etb, &w0, &w1, &dw0, &dw1);
    const double s0 = ee[z - dz] / oo[z - dz], s1 = ee[z] / oo[z];
    I[z] = I[z - dz] * etb + w0 * s0 + w1 * s1;
    double *re = Rem + (size_t)z * nd, *ro = Rop + (size_t)z * nd;
    const double *pe = Rem + (size_t)(z - dz) * nd, *po = Rop + (size_t)(z - dz) * nd;
    re[z] = w1 / oo[z];
    re[z - dz] = w0 / oo[z - dz];
    for (int zp = 0; zp < nd; ++zp) re[zp] += pe[zp] * etb;
    ro[z] = -w1 * ee[z] / oo[z] / oo[z];
    ro[z - dz] = -w0 * ee[z - dz] / oo[z - dz] / oo[z - dz];
    for (int zp = 0; zp < nd; ++zp) ro[zp] += po[zp] * etb;
    for (int zp = 0; zp < nd; ++zp) {
      ro[zp] += -I[z - dz] * etb * dX[(size_t)z * nd + zp];
      ro[zp] += (s0 * dw0 + s1 * dw1) * dX[(size_t)z * nd + zp];
    }
    L[z] = w1;
  }
  free(dtau);
  free(dX);
  return 0;
}

/* Crout factorisation without pivoting and its solves, as cmlib/mathtools.cc:
 * 487-514 write them (lower triangle with the diagonal, unit upper triangle). */
static void crout(int d, const double *S, double *D) {
  for (int k = 0; k < d; ++k) {
    for (int i = k; i < d; ++i) {
      double sum = 0.;
      for (int p = 0; p < k; ++p) sum += D[(size_t)i * d + p] * D[(size_t)p * d + k];
      D[(size_t)i * d + k] = S[(size_t)i * d + k] - sum;
    }
    for (int j = k + 1; j < d; ++j) {
      double sum = 0.;
      for (int p = 0; p < k; ++p) sum += D[(size_t)k * d + p] * D[(size_t)p * d + j];
      D[(size_t)k * d + j] = (S[(size_t)k * d + j] - sum) / D[(size_t)k * d + k];
    }
  }
}

static void solve_crout(int d, const double *LU, const double *b, double *x, double *y) {
  for (int i = 0; i < d; ++i) {
    double sum = 0.;
    for (int k = 0; k < i; ++k) sum += LU[(size_t)i * d + k] * y[k];
    y[i] = (b[i] - sum) / LU[(size_t)i * d + i];
  }
  for (int i = d - 1; i >= 0; --i) {
    double sum = 0.;
    for (int k = i + 1; k < d; ++k) sum += LU[(size_t)i * d + k] * x[k];
    x[i] = (y[i] - sum);
  }
}

int orc_response_solve(int n, int nrhs, const double *M, const double *B, double *X) {
  double *LU = malloc(sizeof(double) * (size_t)n * n);
  double *y = malloc(sizeof(double) * n);
  crout(n, M, LU);
  for (int r = 0; r < nrhs; ++r) solve_crout(n, LU, B + (size_t)r * n, X + (size_t)r * n, y);
  free(LU);
  free(y);
  return 0;
}
