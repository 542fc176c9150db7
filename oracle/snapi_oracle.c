/*
 * FP64 CPU restatement of SNAPI's hot path -- TEST INFRASTRUCTURE ONLY.
 * See snapi_oracle.h: the product never links or calls this file; tests/,
 * the smoke check and bench.py's CPU-baseline leg use it as the checker.
 */
#include "snapi_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* CGS constants as cmlib/const.h:10-24 states them. */
#define C_LIGHT 2.99792458E10
#define H_PLANCK 6.6260687652E-27
#define K_BOLTZ 1.3806503E-16
#define PI_D 3.141592653589793238462643383279502884197169399375105

/* ---------------------------------------------------------------- complex */
typedef struct {
  double re, im;
} cpx;
static inline cpx cmk(double r, double i) { cpx z = {r, i}; return z; }
static inline cpx cadd(cpx a, cpx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cpx csub(cpx a, cpx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cpx cmul(cpx a, cpx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cpx cscl(cpx a, double s) { return cmk(a.re * s, a.im * s); }
static inline cpx caddr(cpx a, double s) { return cmk(a.re + s, a.im); }
static inline cpx crsub(double s, cpx a) { return cmk(s - a.re, -a.im); }
static inline cpx cdiv(cpx a, cpx b) {
  double d = b.re * b.re + b.im * b.im;
  return cmk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

/* Humlicek (1982, JQSRT 27, 437) W4 rational approximation of the Faddeeva
 * function in its four regions, as cmlib/profiles.cc:7-37 evaluates it
 * (including cos(Im u) taken as +sqrt(1 - sin^2)). */
void orc_voigt(double v, double a, double *H, double *F) {
  double s = fabs(v) + a;
  cpx t = cmk(a, -v);
  cpx u = cmk(a * a - v * v, -2.0 * a * v);
  cpx w;
  if (s >= 15.0) {
    w = cdiv(cscl(t, 0.5641896), caddr(u, 0.5));
  } else if (s > 5.5) {
    cpx num = cmul(t, caddr(cscl(u, 0.5641896), 1.410474));
    cpx den = caddr(cmul(u, caddr(u, 3.0)), 0.75);
    w = cdiv(num, den);
  } else if (a > (0.195 * fabs(v) - 0.176)) {
    cpx num = caddr(cmul(t, caddr(cmul(t, caddr(cmul(t, caddr(cscl(t, 0.5642236), 3.778987)), 11.96482)), 20.20933)),
                16.4955);
    cpx den = caddr(
        cmul(t, caddr(cmul(t, caddr(cmul(t, caddr(cmul(t, caddr(t, 6.699398)), 21.69274)), 39.27121)), 38.82363)),
        16.4955);
    w = cdiv(num, den);
  } else {
    cpx p = crsub(1.320522, cscl(u, 0.56419));
    p = crsub(35.76683, cmul(u, p));
    p = crsub(219.0313, cmul(u, p));
    p = crsub(1540.787, cmul(u, p));
    p = crsub(3321.9905, cmul(u, p));
    p = crsub(36183.31, cmul(u, p));
    cpx w4 = cmul(t, p);
    cpx q = crsub(1.841439, u);
    q = crsub(61.57037, cmul(u, q));
    q = crsub(364.2191, cmul(u, q));
    q = crsub(2186.181, cmul(u, q));
    q = crsub(9022.228, cmul(u, q));
    q = crsub(24322.84, cmul(u, q));
    q = crsub(32066.6, cmul(u, q));
    double su = sin(u.im);
    double e = exp(u.re);
    w = csub(cmk(e * sqrt(1.0 - su * su), e * su), cdiv(w4, q));
  }
  *H = w.re / 1.77245386399759;
  *F = w.im / 1.77245386399759;
}

double orc_planck(double lam, double T) {
  return 2.0 * 6.62606957 * 8.9875517E-7 / lam / lam / lam / lam / lam / (exp(1.4387769 / lam / T) - 1.0);
}

/* atom::zeeman_setup (atom.cc:1620-1773): pattern of a J_u -> J_l transition
 * in the LS/anomalous Zeeman picture, strengths normalised per group. */
int orc_zeeman(int j2u, int j2l, double gu, double gl, int nm[3], double shift[3][32], double str[3][32]) {
  int split = (j2u - j2l == 2) || (j2u - j2l == -2) || ((j2u == j2l) && (j2u != 0));
  nm[0] = nm[1] = nm[2] = 0;
  if (!split) return 0;
  int j2min = j2u < j2l ? j2u : j2l;
  nm[0] = j2min + 1;
  nm[1] = nm[2] = (j2u + j2l) / 2;
  double ju = j2u / 2.0, jl = j2l / 2.0, jmin = j2min / 2.0;
  double np = 0, nb = 0, nr = 0;
  if (j2u == j2l) {
    for (int i = 0; i < nm[0]; ++i) {
      double m = i - ju;
      str[0][i] = m * m;
      shift[0][i] = m * (gl - gu);
      np += str[0][i];
    }
    for (int i = 0; i < nm[1]; ++i) {
      double m = i - ju + 1.0;
      str[1][i] = 0.5 * (ju + m) * (ju - m + 1.0);
      shift[1][i] = m * (gl - gu) - gl;
      nb += str[1][i];
      m = i - ju;
      str[2][i] = 0.5 * (ju - m) * (ju + m + 1.0);
      shift[2][i] = m * (gl - gu) + gl;
      nr += str[2][i];
    }
  } else {
    int off = (j2l > j2u) ? 1 : 0;
    for (int i = 0; i < nm[0]; ++i) {
      double m = i - jmin;
      if (jl < ju) {
        str[1][i] = 0.5 * (jmin + m + 1) * (jmin + m + 2);
        str[2][i] = 0.5 * (jmin - m + 1) * (jmin - m + 2);
      } else {
        str[1][i] = 0.5 * (jmin - m + 1) * (jmin - m + 2);
        str[2][i] = 0.5 * (jmin + m + 1) * (jmin + m + 2);
      }
      str[0][i] = (jmin + 1) * (jmin + 1) - m * m;
      np += str[0][i];
      nb += str[1][i];
      nr += str[2][i];
      shift[0][i] = m * (gl - gu);
      shift[1][i] = (m - off) * (gl - gu) - gu;
      shift[2][i] = (m + off) * (gl - gu) + gu;
    }
  }
  for (int i = 0; i < nm[0]; ++i) str[0][i] /= np;
  for (int i = 0; i < nm[1]; ++i) str[1][i] /= nb;
  for (int i = 0; i < nm[2]; ++i) str[2][i] /= nr;
  return 1;
}

#define ST(f, z) state[(size_t)(f) * nd + (z)]

/* H- bound-free and free-free after Gray (mol.cc:249-251, 516-545, 609-632);
 * opacity fudge 1. */
static void hminus(double lam, double T, double ne, double nhi, double *op, double *em) {
  static const double ac[7] = {1.99654, -1.18267E-5, 2.64243E-6, -4.40524E-10, 3.23992E-14, -1.39568E-18,
                               2.78701E-23};
  double la = lam * 1E8;
  double alpha = 0.0;
  for (int i = 0; i < 7; ++i) alpha += ac[i] * pow(la, i);
  double TP = 5040. / T;
  double opbf = 4.158E-28 * alpha * ne * K_BOLTZ * T * pow(TP, 2.5) * pow(10.0, 0.754 * TP) * nhi *
                (1.0 - exp(-H_PLANCK * C_LIGHT / lam / K_BOLTZ / T));
  double l = log10(la);
  double f0 = -2.2763 - 1.6850 * l + 0.76661 * l * l - 0.053346 * l * l * l;
  double f1 = 15.2827 - 9.2846 * l + 1.99381 * l * l - 0.142631 * l * l * l;
  double f2 = -197.789 + 190.266 * l - 67.9775 * l * l + 10.6913 * l * l * l - 0.625151 * l * l * l * l;
  double lt = log10(TP);
  double opff = 1E-26 * ne * K_BOLTZ * T * pow(10.0, f0 + f1 * lt + f2 * lt * lt) * nhi;
  *op = opbf + opff;
  *em = (*op) * orc_planck(lam, T);
}

void orc_opem(const snapi_line_table *lt, const snapi_ray *ray, int nd, const double *state, double lam,
              int normalise, double *op, double *em) {
  memset(op, 0, sizeof(double) * 16 * nd);
  memset(em, 0, sizeof(double) * 4 * nd);
  for (int z = 0; z < nd; ++z) {
    double *K = op + 16 * z, *E = em + 4 * z;
    double T = ST(SNAPI_F_T, z);
    /* Thomson and atomic bound-free: wavelength independent within the window */
    K[0] = ST(SNAPI_F_BGOP, z);
    E[0] = ST(SNAPI_F_BGEM, z);
    if (ray->hminus) {
      double o, e;
      hminus(ray->hminus_mid ? ray->lam_mid : lam, T, ST(SNAPI_F_NE, z), ST(SNAPI_F_NHI, z), &o, &e);
      K[0] += o;
      E[0] += e;
    }
    double vt = ST(SNAPI_F_VT, z), vl = ST(SNAPI_F_VLOS, z), bm = ST(SNAPI_F_BMAG, z);
    double thb = ST(SNAPI_F_BINC, z) * PI_D / 180.0, phb = ST(SNAPI_F_BAZI, z) * PI_D / 180.0;
    if (bm < 0.001) thb = phb = 0.0;
    double sp = sin(2.0 * phb), cp = cos(2.0 * phb), st = sin(thb), ct = cos(thb);
    double nhi = ST(SNAPI_F_NHI, z);
    for (int il = 0; il < lt->nline; ++il) {
      const snapi_line *L = &lt->line[il];
      double l0 = L->lam0;
      double nl = ST(SNAPI_NF_BASE + 2 * il, z), nu = ST(SNAPI_NF_BASE + 2 * il + 1, z);
      /* broad_dop atom_rad.cc:879-883, damp_col atom.cc:939-960 */
      double dld = sqrt(2.0 * K_BOLTZ * T / L->mass + vt * vt) * l0 / C_LIGHT;
      double red = 1.66E-24 / (1.0 / 1.008 + 1.66E-24 / L->mass);
      double vm = sqrt(8.0 * K_BOLTZ * T / PI_D / red);
      double gc = nhi * vm * pow(vm / 1E6, -L->alpha) * L->sigma;
      double sf = (l0 * l0) / (4.0 * PI_D * C_LIGHT);
      double a = sf * (L->grad + gc) / dld;
      double lB = 4.67E-5 * l0 * l0 * bm;
      double opl = (nl * L->blu - nu * L->bul) * H_PLANCK * C_LIGHT / l0 / 4.0 / PI_D;
      double eml = nu * L->aul * H_PLANCK * C_LIGHT / l0 / 4.0 / PI_D;
      double lshift = lam * (1.0 + vl / C_LIGHT);
      if (!L->split) {
        double Hh, Ff;
        orc_voigt((l0 - lshift) / dld, a, &Hh, &Ff);
        K[0] += opl * Hh / dld;
        E[0] += eml * Hh / dld;
        continue;
      }
      double Hs[3] = {0, 0, 0}, Fs[3] = {0, 0, 0};
      int c = L->comp0;
      for (int g = 0; g < 3; ++g)
        for (int m = 0; m < L->ncomp[g]; ++m, ++c) {
          double Hh, Ff;
          orc_voigt((l0 + lt->comp[c].shift * lB - lshift) / dld, a, &Hh, &Ff);
          Hs[g] += Hh * lt->comp[c].strength;
          Fs[g] += Ff * lt->comp[c].strength;
        }
      for (int g = 0; g < 3; ++g) {
        Hs[g] /= dld;
        Fs[g] /= dld;
      }
      double Hp = Hs[0], Hb = Hs[1], Hr = Hs[2], Fp = Fs[0], Fb = Fs[1], Fr = Fs[2];
      double eI = 0.5 * (Hp * st * st + 0.5 * (Hr + Hb) * (1.0 + ct * ct));
      double eQ = 0.5 * (Hp - 0.5 * (Hr + Hb)) * st * st * cp;
      double eU = 0.5 * (Hp - 0.5 * (Hr + Hb)) * st * st * sp;
      double eV = 0.5 * (Hr - Hb) * ct;
      double rQ = 0.5 * (Fp - 0.5 * (Fr + Fb)) * st * st * cp;
      double rU = 0.5 * (Fp - 0.5 * (Fr + Fb)) * st * st * sp;
      double rV = 0.5 * (Fr - Fb) * ct;
      K[0] += eI * opl;
      K[1] += eQ * opl;
      K[2] += eU * opl;
      K[3] += eV * opl;
      K[2 * 4 + 3] += rQ * opl; /* [3][4] */
      K[3 * 4 + 1] += rU * opl; /* [4][2] */
      K[1 * 4 + 2] += rV * opl; /* [2][3] */
      E[0] += eI * eml;
      E[1] += eQ * eml;
      E[2] += eU * eml;
      E[3] += eV * eml;
    }
    /* Thomson (atmos_rad_synthesis.cc:41-50) is part of BGOP/BGEM.
     * symmetrise (atmos_rad_synthesis.cc:56-68) */
    K[5] = K[10] = K[15] = K[0];
    K[4] = K[1];
    K[8] = K[2];
    K[12] = K[3];
    K[2 * 4 + 1] = -K[1 * 4 + 2];
    K[1 * 4 + 3] = -K[3 * 4 + 1];
    K[3 * 4 + 2] = -K[2 * 4 + 3];
    if (normalise) {
      double r = ST(SNAPI_F_OPREF, z);
      for (int i = 0; i < 16; ++i) K[i] /= r;
      for (int i = 0; i < 4; ++i) E[i] /= r;
    }
  }
}

/* Crout LU without pivoting and its solve, cmlib/mathtools.cc:487-514
 * (unit diagonal on U, L keeps the diagonal). */
static void crout4(const double *S, double *D) {
  const int d = 4;
  for (int k = 0; k < d; ++k) {
    for (int i = k; i < d; ++i) {
      double sum = 0.;
      for (int p = 0; p < k; ++p) sum += D[i * d + p] * D[p * d + k];
      D[i * d + k] = S[i * d + k] - sum;
    }
    for (int j = k + 1; j < d; ++j) {
      double sum = 0.;
      for (int p = 0; p < k; ++p) sum += D[k * d + p] * D[p * d + j];
      D[k * d + j] = (S[k * d + j] - sum) / D[k * d + k];
    }
  }
}
static void crout4_solve(const double *LU, const double *b, double *x) {
  const int d = 4;
  double y[4];
  for (int i = 0; i < d; ++i) {
    double sum = 0.;
    for (int k = 0; k < i; ++k) sum += LU[i * d + k] * y[k];
    y[i] = (b[i] - sum) / LU[i * d + i];
  }
  for (int i = d - 1; i >= 0; --i) {
    double sum = 0.;
    for (int k = i + 1; k < d; ++k) sum += LU[i * d + k] * x[k];
    x[i] = y[i] - sum;
  }
}

/* Fritsch-Butland monotone derivative of the opacity along the ray and the
 * cubic-Hermite optical-depth steps (atmos_ppbez_rts.cc:961-988).
 * ord[k] is the depth index of the k-th point along the ray. */
static void delta_tau(int nd, const int *ord, const double *grid, const double *oo, int ostride, double ct,
                      double *dtau) {
  double *D = (double *)malloc(sizeof(double) * nd);
#define OO(k) oo[(size_t)ord[k] * ostride]
#define GR(k) grid[ord[k]]
  D[0] = (OO(1) - OO(0)) / (GR(1) - GR(0)) * (-ct);
  D[nd - 1] = (OO(nd - 1) - OO(nd - 2)) / (GR(nd - 1) - GR(nd - 2)) * (-ct);
  for (int k = 1; k < nd - 1; ++k) {
    D[k] = 0.0;
    double h2 = (GR(k + 1) - GR(k)) / (-ct);
    double h1 = (GR(k) - GR(k - 1)) / (-ct);
    double d2 = (OO(k + 1) - OO(k)) / h2;
    double d1 = (OO(k) - OO(k - 1)) / h1;
    if (d1 * d2 > 0) {
      double al = (1.0 + h1 / (h1 + h2)) / 3.0;
      D[k] = d1 * d2 / (al * d1 + (1.0 - al) * d2);
    }
  }
  dtau[0] = 0.0;
  for (int k = 1; k < nd; ++k) {
    double hk = (GR(k) - GR(k - 1)) / (-ct);
    double C0 = OO(k) - hk / 2.0 * D[k];
    double C1 = OO(k - 1) + hk / 2.0 * D[k - 1];
    double C = 0.5 * (C0 + C1);
    dtau[k] = hk / 3.0 * (OO(k - 1) + OO(k) + C);
  }
#undef OO
#undef GR
  free(D);
}

/* linear (w2 = 0) short-characteristics weights, atmos_ppbez_rts.cc:1067-1096 */
static void weights(double tb, double *etb, double *w0, double *w1) {
  const double expansion_delimiter = 1E-2;
  *etb = (tb < expansion_delimiter) ? (1.0 - tb + tb * tb * 0.5 - tb * tb * tb / 6.0) : exp(-tb);
  if (tb < expansion_delimiter) {
    *w0 = tb * (0.5 - tb * (1.0 / 3.0 - tb / 8.0));
    *w1 = 0.5 * tb * (1.0 - tb / 3.0 * (1.0 - tb / 4.0));
  } else {
    *w0 = 1.0 / tb - *etb * (1.0 + 1.0 / tb);
    *w1 = 1.0 - 1.0 / tb + *etb / tb;
  }
}

int orc_formal_pol(int nd, const double *grid, const double *op, const double *em, double theta, double boundary,
                   double *S) {
  double ct = cos(theta);
  int dz = (ct > 0) ? 1 : -1;
  int *ord = (int *)malloc(sizeof(int) * nd);
  for (int k = 0; k < nd; ++k) ord[k] = (ct > 0) ? k : nd - 1 - k;
  double *dtau = (double *)malloc(sizeof(double) * nd);
  delta_tau(nd, ord, grid, op, 16, ct, dtau);
  int rc = 0;
  (void)dz;
  /* boundary (atmos_ppbez_rts.cc:944-949) */
  int z0 = ord[0];
  S[4 * z0 + 0] = (ct > 0) ? 0.0 : em[4 * z0] / op[16 * z0];
  if (ct < 0 && boundary >= 0) S[4 * z0] = boundary;
  S[4 * z0 + 1] = S[4 * z0 + 2] = S[4 * z0 + 3] = 0.0;
  for (int k = 1; k < nd - 1; ++k) {
    int z = ord[k], zp = ord[k - 1];
    double tb = dtau[k], tf = dtau[k + 1];
    if (tf > 1E10) rc = SNAPI_ERR_UNSUPPORTED;
    double etb, w0, w1;
    weights(tb, &etb, &w0, &w1);
    const double *Kz = op + 16 * z, *Kp = op + 16 * zp;
    double Kr[16], Kpr[16], Sz[4], Sp[4];
    for (int s = 0; s < 4; ++s) {
      Sz[s] = em[4 * z + s] / Kz[0];
      Sp[s] = em[4 * zp + s] / Kp[0];
      for (int q = 0; q < 4; ++q) {
        Kr[4 * s + q] = Kz[4 * s + q] / Kz[0];
        Kpr[4 * s + q] = Kp[4 * s + q] / Kp[0];
      }
      Kr[5 * s] -= 1.0;
      Kpr[5 * s] -= 1.0;
    }
    double A[16], B[4], LU[16];
    const double *Ip = S + 4 * zp;
    for (int s = 0; s < 4; ++s) {
      for (int q = 0; q < 4; ++q) A[4 * s + q] = w1 * Kr[4 * s + q];
      A[5 * s] += 1.0;
      double kv = 0.0;
      for (int q = 0; q < 4; ++q) kv += Kpr[4 * s + q] * Ip[q];
      B[s] = -w0 * kv + w0 * Sp[s] + w1 * Sz[s] + Ip[s] * etb;
    }
    crout4(A, LU);
    crout4_solve(LU, B, S + 4 * z);
  }
  if (nd > 1) {
    int zh = ord[nd - 1], zl1 = ord[nd - 2];
    for (int s = 0; s < 4; ++s) S[4 * zh + s] = S[4 * zl1 + s];
  }
  free(ord);
  free(dtau);
  return rc;
}

/* Scalar formal solution and diagonal Lambda*, atmos_ppbez_rts.cc:21-219. */
int orc_formal_scalar(int nd, const double *grid, const double *op, const double *em, double theta, double boundary,
                      double *I, double *L) {
  double ct = cos(theta);
  int *ord = (int *)malloc(sizeof(int) * nd);
  for (int k = 0; k < nd; ++k) ord[k] = (ct > 0) ? k : nd - 1 - k;
  double *dtau = (double *)malloc(sizeof(double) * nd);
  delta_tau(nd, ord, grid, op, 1, ct, dtau);
  int z0 = ord[0];
  I[z0] = (ct > 0) ? 0.0 : em[z0] / op[z0];
  if (ct < 0 && boundary >= 0) I[z0] = boundary;
  if (L) {
    if (ct >= 0) L[z0] = 0.0;
    if (ct <= 0) L[z0] = (boundary == -1) ? 1.0 : 0.0;
  }
  for (int k = 1; k < nd; ++k) {
    int z = ord[k], zp = ord[k - 1];
    double tb = dtau[k];
    double etb, w0, w1;
    weights(tb, &etb, &w0, &w1);
    I[z] = I[zp] * etb + w0 * em[zp] / op[zp] + w1 * em[z] / op[z];
    if (L) L[z] = w1;
  }
  free(ord);
  free(dtau);
  return 0;
}

int orc_synth(const snapi_line_table *lt, const snapi_ray *ray, int ncol, int nd, const double *state, int nlambda,
              const double *lam, double *stokes, int nthreads) {
  int nf = SNAPI_NF_BASE + 2 * lt->nline;
  int rc = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic) reduction(min : rc)
#endif
  for (long long job = 0; job < (long long)ncol * nlambda; ++job) {
    int col = (int)(job / nlambda), l = (int)(job % nlambda);
    const double *st = state + (size_t)col * nf * nd;
    double *op = (double *)malloc(sizeof(double) * 16 * nd);
    double *em = (double *)malloc(sizeof(double) * 4 * nd);
    double *S = (double *)malloc(sizeof(double) * 4 * nd);
    orc_opem(lt, ray, nd, st, lam[l], ray->tau_grid, op, em);
    int r = orc_formal_pol(nd, st + (size_t)SNAPI_F_GRID * nd, op, em, ray->theta, -1.0, S);
    if (r < rc) rc = r;
    double *out = stokes + ((size_t)col * nlambda + l) * 4;
    for (int s = 0; s < 4; ++s) out[s] = S[s]; /* emergent: depth index 0 (top) */
    free(op);
    free(em);
    free(S);
  }
  return rc;
}
