/*
 * Per-depth response functions, FP64 restatement of the reference's
 * algorithm -- TEST INFRASTRUCTURE ONLY (see snapi_oracle.h).
 *
 * atmosphere::obs_stokes_responses (atmos/atmos_rad_synthesis.cc:166-399)
 * with no node model:
 *   1. opacity/emissivity and their local derivatives along the 7 physical
 *      parameters (T, Nt, vt, vLOS, |B|, inclination, azimuth):
 *      op_em_vector_plus_pert (atmos/atmos_rad_synthesis.cc:76-162),
 *      atom::boundbound_op_em_vector_plus_pert (atom_rad_synthesis.cc:771-1026),
 *      h_minus_mol::op_em_vector_plus_pert (mol.cc:414-469, H- at the middle
 *      wavelength), bound-free derivatives added as zero (:384-385);
 *   2. normalisation by the referent opacity with its derivative
 *      (atmos/atmos_rad.cc:187-201);
 *   3. propagation through the solver by the central difference
 *      I(op + d/2) - I(op - d/2) of formal_pert_numerical
 *      (plpar/atmos_ppbez_rts.cc:1708-1776), Nt steps scaled by 1e-3 Nt and
 *      angle steps by 1e-3 rad as the caller does (:275-305).
 * The response of the Voigt profile uses fvoigtn (cmlib/profiles.cc:39-89),
 * the derivative of each W4 region.
 * Pinned: tests/golden/cfg0_resp.npz (the reference's responses at 31
 * wavelengths, 61 depths, all 7 parameters).
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
#define PI_D 3.141592653589793238462643383279502884197169399375105
#define DELTA_NT_FRAC 1E-3
#define DELTA_ANGLE 0.001

typedef struct {
  double re, im;
} cx;
static cx mk(double r, double i) { cx z = {r, i}; return z; }
static cx add(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static cx sub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
static cx mul(cx a, cx b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static cx scl(cx a, double s) { return mk(a.re * s, a.im * s); }
static cx addr(cx a, double s) { return mk(a.re + s, a.im); }
static cx rsub(double s, cx a) { return mk(s - a.re, -a.im); }
static cx dvd(cx a, cx b) {
  double d = b.re * b.re + b.im * b.im;
  return mk((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}

/* fvoigtn, cmlib/profiles.cc:39-89: W4 values and the derivative of each
 * region's expression with respect to v (dH, dF). */
void orc_voigtn(double v, double a, double *H, double *F, double *dH, double *dF) {
  double s = fabs(v) + a;
  cx t = mk(a, -v), dt = mk(0, -1.0);
  cx u = mk(a * a - v * v, -2.0 * a * v), du = mk(-2.0 * v, -2.0 * a);
  cx w4, dw4;
  if (s >= 15.0) {
    cx den = addr(u, 0.5);
    w4 = dvd(scl(t, 0.5641896), den);
    dw4 = sub(dvd(scl(dt, 0.5641896), den), mul(dvd(dvd(scl(t, 0.5641896), den), den), du));
  } else if (s > 5.5) {
    cx num = addr(scl(u, 0.5641896), 1.410474);
    cx den = addr(mul(u, addr(u, 3.0)), 0.75);
    w4 = dvd(mul(t, num), den);
    dw4 = add(dvd(mul(dt, num), den), dvd(scl(mul(t, du), 0.5641896), den));
    dw4 = sub(dw4, mul(dvd(dvd(mul(mul(t, du), num), den), den), addr(scl(u, 2.0), 3.0)));
  } else if (a > (0.195 * fabs(v) - 0.176)) {
    cx up = addr(mul(t, addr(mul(t, addr(mul(t, addr(scl(t, 0.5642236), 3.778987)), 11.96482)), 20.20933)), 16.4955);
    cx lo = addr(mul(t, addr(mul(t, addr(mul(t, addr(mul(t, addr(t, 6.699398)), 21.69274)), 39.27121)), 38.82363)),
                 16.4955);
    w4 = dvd(up, lo);
    cx t2 = mul(t, t), t3 = mul(t2, t), t4 = mul(t3, t);
    cx dup = add(add(addr(scl(t, 11.96482 * 2.0), 20.20933), scl(t2, 3.778987 * 3.0)), scl(t3, 0.5642236 * 4.0));
    cx dlo = add(add(add(addr(scl(t, 39.27121 * 2.0), 38.82363), scl(t2, 21.69274 * 3.0)), scl(t3, 6.699398 * 4.0)),
                 scl(t4, 5.0));
    dw4 = mul(dt, dvd(dvd(sub(mul(dup, lo), mul(dlo, up)), lo), lo));
  } else {
    cx P = rsub(1.320522, scl(u, 0.56419));
    P = rsub(35.76683, mul(u, P));
    P = rsub(219.0313, mul(u, P));
    P = rsub(1540.787, mul(u, P));
    P = rsub(3321.9905, mul(u, P));
    P = rsub(36183.31, mul(u, P));
    cx w = mul(t, P);
    cx dP = rsub(1.320522 * 5.0, scl(u, 0.56419 * 6.0));
    dP = rsub(35.76683 * 4.0, mul(u, dP));
    dP = rsub(219.0313 * 3.0, mul(u, dP));
    dP = rsub(1540.787 * 2.0, mul(u, dP));
    dP = rsub(3321.9905, mul(u, dP));
    cx dw = sub(mul(dt, P), mul(mul(t, du), dP));
    cx Q = rsub(1.841439, u);
    Q = rsub(61.57037, mul(u, Q));
    Q = rsub(364.2191, mul(u, Q));
    Q = rsub(2186.181, mul(u, Q));
    Q = rsub(9022.228, mul(u, Q));
    Q = rsub(24322.84, mul(u, Q));
    Q = rsub(32066.6, mul(u, Q));
    cx dQ = rsub(1.841439 * 6.0, scl(u, 7.0));
    dQ = rsub(61.57037 * 5.0, mul(u, dQ));
    dQ = rsub(364.2191 * 4.0, mul(u, dQ));
    dQ = rsub(2186.181 * 3.0, mul(u, dQ));
    dQ = rsub(9022.228 * 2.0, mul(u, dQ));
    dQ = rsub(24322.84, mul(u, dQ));
    dQ = scl(mul(du, dQ), -1.0);
    double su = sin(u.im), dsu = cos(u.im) * du.im;
    double e = exp(u.re);
    cx cs = mk(sqrt(1.0 - su * su), su);
    cx full = add(scl(cs, e * du.re), scl(mk(-su * du.im, dsu), e));
    full = sub(full, dvd(sub(mul(dw, Q), mul(dQ, w)), mul(Q, Q)));
    dw4 = full;
    w4 = sub(scl(cs, e), dvd(w, Q));
  }
  *H = w4.re / 1.77245386399759;
  *F = w4.im / 1.77245386399759;
  *dH = dw4.re / 1.77245386399759;
  *dF = dw4.im / 1.77245386399759;
}

/* Planck_f_derivative, cmlib/mathtools.cc:1154-1158 */
static double planck_dT(double lam, double T) {
  double x = 1.4387769 / lam / T;
  double tmp = (x > 5.0) ? 1.0 : exp(x) / (exp(x) - 1.0);
  return orc_planck(lam, T) * tmp * 1.4387769 / lam / T / T;
}

#define ST(f, z) state[(size_t)(f) * nd + (z)]
#define DS(p, f, z) dstate[((size_t)(p) * nf + (f)) * nd + (z)]

/* op[nd][16], em[nd][4] and local derivatives dop[7][nd][16], dem[7][nd][4]
 * (un-normalised, parameter index 0..6 = T, Nt, vt, vLOS, |B|, incl, azi). */
static void opem_pert(const snapi_line_table *lt, const snapi_ray *ray, int nd, int nf, const double *state,
                      const double *dstate, double lam, double *op, double *em, double *dop, double *dem) {
  memset(op, 0, sizeof(double) * 16 * nd);
  memset(em, 0, sizeof(double) * 4 * nd);
  memset(dop, 0, sizeof(double) * 7 * 16 * nd);
  memset(dem, 0, sizeof(double) * 7 * 4 * nd);
  const double lm = ray->lam_mid;
#define DOP(p, z, i) dop[((size_t)(p) * nd + (z)) * 16 + (i)]
#define DEM(p, z, i) dem[((size_t)(p) * nd + (z)) * 4 + (i)]
  for (int z = 0; z < nd; ++z) {
    double *K = op + 16 * z, *E = em + 4 * z;
    const double T = ST(SNAPI_F_T, z), Ne = ST(SNAPI_F_NE, z), nhi = ST(SNAPI_F_NHI, z), Nt = ST(SNAPI_F_NT, z);
    const double dNe[2] = {DS(0, SNAPI_F_NE, z), DS(1, SNAPI_F_NE, z)};
    const double dNH[2] = {DS(0, SNAPI_F_NHI, z), DS(1, SNAPI_F_NHI, z)};
    /* Thomson at the middle wavelength (atmos_rad_synthesis.cc:99-123) */
    const double opth = Ne * 6.65E-25;
    const double emth = opth * orc_planck(lm, T);
    K[0] += opth;
    E[0] = emth;
    for (int p = 0; p < 2; ++p) {
      const double o = dNe[p] * 6.65E-25;
      DOP(p, z, 0) += o;
      DEM(p, z, 0) += emth / opth * o + (p == 0 ? Ne * 6.65E-25 * planck_dT(lm, T) : 0.0);
    }
    /* atomic bound-free: opacity added, derivatives added as zero */
    K[0] += ST(SNAPI_F_BGOP, z) - opth;
    E[0] += ST(SNAPI_F_BGEM, z) - Ne * 6.65E-25 * orc_planck(lm, T);
    /* H- at lam_h (mol.cc:414-469, 516-632) */
    if (ray->hminus) {
      const double lh = ray->hminus_mid ? lm : lam;
      static const double ac[7] = {1.99654, -1.18267E-5, 2.64243E-6, -4.40524E-10, 3.23992E-14, -1.39568E-18,
                                   2.78701E-23};
      double alpha = 0.0;
      for (int i = 0; i < 7; ++i) alpha += ac[i] * pow(lh * 1E8, i);
      const double TP = 5040. / T;
      const double stim = exp(-H_PLANCK * C_LIGHT / lh / K_BOLTZ / T);
      const double obf = 4.158E-28 * alpha * Ne * K_BOLTZ * T * pow(TP, 2.5) * pow(10.0, 0.754 * TP) * nhi * (1.0 - stim);
      const double l = log10(lh * 1E8);
      const double f0 = -2.2763 - 1.6850 * l + 0.76661 * l * l - 0.053346 * l * l * l;
      const double f1 = 15.2827 - 9.2846 * l + 1.99381 * l * l - 0.142631 * l * l * l;
      const double f2 = -197.789 + 190.266 * l - 67.9775 * l * l + 10.6913 * l * l * l - 0.625151 * l * l * l * l;
      const double ltp = log10(TP);
      const double ef = pow(10.0, f0 + f1 * ltp + f2 * ltp * ltp);
      const double off = 1E-26 * Ne * K_BOLTZ * T * ef * nhi;
      const double o = obf + off;
      const double B = orc_planck(lh, T);
      K[0] += o;
      E[0] += o * B;
      double dbf[2], dff[2];
      /* bound-free (mol.cc:652-663) */
      dbf[0] = obf * (1.0 / T + 2.5 * 1.0 / TP * (-5040.0 / T / T));
      dbf[0] += obf * dNe[0] / Ne;
      dbf[0] += obf * dNH[0] / nhi;
      dbf[0] += obf * (-2.302585 * 0.754 * 5040.0 / T / T);
      dbf[0] += obf * (-stim * H_PLANCK * C_LIGHT / lh / K_BOLTZ / T / T) / (1.0 - stim);
      dbf[1] = obf * dNe[1] / Ne + obf * dNH[1] / nhi;
      /* free-free (mol.cc:560-592) */
      const double dTP = -TP / T;
      const double dltp = 1.0 / 2.302585 / TP * dTP;
      const double def = ef * 2.302585 * (f1 * dltp + 2.0 * f2 * ltp * dltp);
      const double cf = 1E-26 * K_BOLTZ;
      dff[0] = cf * Ne * ef * nhi + cf * T * Ne * nhi * def + cf * T * ef * (Ne * dNH[0] + nhi * dNe[0]);
      dff[1] = cf * T * ef * (Ne * dNH[1] + nhi * dNe[1]);
      for (int p = 0; p < 2; ++p) {
        const double d = dbf[p] + dff[p];
        DOP(p, z, 0) += d;
        DEM(p, z, 0) += B * d + (p == 0 ? o * planck_dT(lh, T) : 0.0);
      }
    }
    /* lines */
    const double vt = ST(SNAPI_F_VT, z), vl = ST(SNAPI_F_VLOS, z), bm = ST(SNAPI_F_BMAG, z);
    double thb = ST(SNAPI_F_BINC, z) * PI_D / 180.0, phb = ST(SNAPI_F_BAZI, z) * PI_D / 180.0;
    if (bm < 0.001) thb = phb = 0.0;
    const double sp = sin(2.0 * phb), cp = cos(2.0 * phb), st = sin(thb), ct = cos(thb);
    for (int il = 0; il < lt->nline; ++il) {
      const snapi_line *L = &lt->line[il];
      const double l0 = L->lam0;
      const double nl = ST(SNAPI_NF_BASE + 2 * il, z), nu = ST(SNAPI_NF_BASE + 2 * il + 1, z);
      const double cfac = H_PLANCK * C_LIGHT / l0 / 4.0 / PI_D;
      const double dvd_ = sqrt(2.0 * K_BOLTZ * T / L->mass + vt * vt);
      const double dld = dvd_ * l0 / C_LIGHT;
      const double red = 1.66E-24 / (1.0 / 1.008 + 1.66E-24 / L->mass);
      const double vm = sqrt(8.0 * K_BOLTZ * T / PI_D / red);
      const double gc = nhi * vm * pow(vm / 1E6, -L->alpha) * L->sigma;
      const double sf = (l0 * l0) / (4.0 * PI_D * C_LIGHT);
      const double a = sf * (L->grad + gc) / dld;
      const double opl = (nl * L->blu - nu * L->bul) * cfac;
      const double eml = nu * L->aul * cfac;
      double dopl[2], deml[2];
      for (int p = 0; p < 2; ++p) {
        dopl[p] = (DS(p, SNAPI_NF_BASE + 2 * il, z) * L->blu - DS(p, SNAPI_NF_BASE + 2 * il + 1, z) * L->bul) * cfac;
        deml[p] = DS(p, SNAPI_NF_BASE + 2 * il + 1, z) * L->aul * cfac;
      }
      /* derivatives of dld and a (atom_rad.cc:611-664, 836-875; damp_col_der_T atom.cc:969-985) */
      const double ddld_T = l0 / (C_LIGHT * dvd_) * K_BOLTZ / L->mass;
      const double ddld_vt = l0 / (C_LIGHT * dvd_) * vt;
      const double vmd = 0.5 / vm * 8.0 * K_BOLTZ / PI_D / red;
      double dgc_T = nhi * L->sigma * (vmd * pow(vm / 1E6, -L->alpha) -
                                       vm * L->alpha * pow(vm / 1E6, -L->alpha - 1.0) * vmd / 1E6);
      dgc_T += dNH[0] * vm * pow(vm / 1E6, -L->alpha) * L->sigma;
      double a_der[7] = {0, 0, 0, 0, 0, 0, 0};
      a_der[0] = -a / dld * ddld_T + dgc_T / dld * sf;
      a_der[1] = (gc * DELTA_NT_FRAC / dld * sf) / (DELTA_NT_FRAC * Nt);
      a_der[2] = -a / dld * ddld_vt;
      const double lshift = lam * (1.0 + vl / C_LIGHT);
      if (!L->split) {
        double Hh, Ff, dH, dF;
        const double x = (l0 - lshift) / dld;
        orc_voigtn(x, a, &Hh, &Ff, &dH, &dF);
        K[0] += opl * Hh / dld;
        E[0] += eml * Hh / dld;
        double x_der[4] = {-x / dld * ddld_T, 0.0, -x / dld * ddld_vt, -lam * (1.0 / C_LIGHT) / dld};
        double dld_der[4] = {ddld_T, 0.0, ddld_vt, 0.0};
        for (int p = 0; p < 4; ++p) {
          const double g = (dH * x_der[p] - dF * a_der[p] - Hh / dld * dld_der[p]) / dld;
          DOP(p, z, 0) += opl * g;
          DEM(p, z, 0) += eml * g;
        }
        for (int p = 0; p < 2; ++p) {
          DOP(p, z, 0) += dopl[p] * Hh / dld;
          DEM(p, z, 0) += deml[p] * Hh / dld;
        }
        continue;
      }
      const double lB = 4.67E-5 * l0 * l0 * bm;
      double Hs[3] = {0, 0, 0}, Fs[3] = {0, 0, 0}, Hd[3][7], Fd[3][7];
      memset(Hd, 0, sizeof(Hd));
      memset(Fd, 0, sizeof(Fd));
      int c = L->comp0;
      for (int g = 0; g < 3; ++g)
        for (int m = 0; m < L->ncomp[g]; ++m, ++c) {
          const double sh = lt->comp[c].shift, sw = lt->comp[c].strength;
          const double x = (l0 + sh * lB - lshift) / dld;
          double Hh, Ff, dH, dF;
          orc_voigtn(x, a, &Hh, &Ff, &dH, &dF);
          Hs[g] += Hh * sw;
          Fs[g] += Ff * sw;
          double xd[7] = {-x / dld * ddld_T, 0.0, -x / dld * ddld_vt, -lam * (1.0 / C_LIGHT) / dld,
                          (sh * 4.67E-5 * l0 * l0) / dld, 0.0, 0.0};
          for (int p = 0; p < 7; ++p) {
            Hd[g][p] += (dH * xd[p] - dF * a_der[p]) * sw;
            Fd[g][p] += (dF * xd[p] + dH * a_der[p]) * sw;
          }
        }
      for (int g = 0; g < 3; ++g) {
        Hs[g] /= dld;
        Fs[g] /= dld;
      }
      /* geometry of the Zeeman pattern (atom_rad_synthesis.cc:486-497) */
#define GEOM(Hp, Hb, Hr, Fp, Fb, Fr, fac, KK, EE, withF)                                \
  do {                                                                                  \
    const double eI = 0.5 * ((Hp) * st * st + 0.5 * ((Hr) + (Hb)) * (1.0 + ct * ct));   \
    const double eQ = 0.5 * ((Hp) - 0.5 * ((Hr) + (Hb))) * st * st * cp;                \
    const double eU = 0.5 * ((Hp) - 0.5 * ((Hr) + (Hb))) * st * st * sp;                \
    const double eV = 0.5 * ((Hr) - (Hb)) * ct;                                         \
    (KK)[0] += eI * (fac)[0];                                                           \
    (KK)[1] += eQ * (fac)[0];                                                           \
    (KK)[2] += eU * (fac)[0];                                                           \
    (KK)[3] += eV * (fac)[0];                                                           \
    if (withF) {                                                                        \
      (KK)[2 * 4 + 3] += 0.5 * ((Fp) - 0.5 * ((Fr) + (Fb))) * st * st * cp * (fac)[0];  \
      (KK)[3 * 4 + 1] += 0.5 * ((Fp) - 0.5 * ((Fr) + (Fb))) * st * st * sp * (fac)[0];  \
      (KK)[1 * 4 + 2] += 0.5 * ((Fr) - (Fb)) * ct * (fac)[0];                          \
    }                                                                                   \
    (EE)[0] += eI * (fac)[1];                                                           \
    (EE)[1] += eQ * (fac)[1];                                                           \
    (EE)[2] += eU * (fac)[1];                                                           \
    (EE)[3] += eV * (fac)[1];                                                           \
  } while (0)
      {
        const double f[2] = {opl, eml};
        GEOM(Hs[0], Hs[1], Hs[2], Fs[0], Fs[1], Fs[2], f, K, E, 1);
      }
      for (int p = 0; p < 2; ++p) {
        const double f[2] = {dopl[p], deml[p]};
        GEOM(Hs[0], Hs[1], Hs[2], Fs[0], Fs[1], Fs[2], f, &DOP(p, z, 0), &DEM(p, z, 0), 1);
      }
      for (int p = 0; p < 7; ++p)
        for (int g = 0; g < 3; ++g) {
          Hd[g][p] /= dld;
          Fd[g][p] /= dld;
        }
      for (int g = 0; g < 3; ++g) {
        Hd[g][0] -= ddld_T / dld * Hs[g];
        Hd[g][2] -= ddld_vt / dld * Hs[g];
        Fd[g][0] -= ddld_T / dld * Fs[g];
        Fd[g][2] -= ddld_vt / dld * Fs[g];
      }
      for (int p = 0; p < 5; ++p) {
        const double f[2] = {opl, eml};
        GEOM(Hd[0][p], Hd[1][p], Hd[2][p], Fd[0][p], Fd[1][p], Fd[2][p], f, &DOP(p, z, 0), &DEM(p, z, 0), 1);
      }
      /* inclination (with the reference's rho_V term, :999) and azimuth */
      {
        const double Hp = Hs[0], Hb = Hs[1], Hr = Hs[2], Fp = Fs[0], Fb = Fs[1], Fr = Fs[2];
        double *k6 = &DOP(5, z, 0), *e6 = &DEM(5, z, 0), *k7 = &DOP(6, z, 0), *e7 = &DEM(6, z, 0);
        const double hl = Hp - 0.5 * (Hr + Hb), fl = Fp - 0.5 * (Fr + Fb);
        k6[0] += 0.5 * (Hp * 2.0 * st * ct - 0.5 * (Hr + Hb) * (2.0 * ct * st)) * opl;
        k6[1] += 0.5 * hl * 2.0 * st * ct * cp * opl;
        k6[2] += 0.5 * hl * 2.0 * st * ct * sp * opl;
        k6[3] += -0.5 * (Hr - Hb) * st * opl;
        k6[2 * 4 + 3] += 0.5 * fl * 2.0 * ct * st * cp * opl;
        k6[3 * 4 + 1] += 0.5 * fl * 2.0 * ct * st * sp * opl;
        k6[1 * 4 + 2] += 0.5 * (Fr - Fb) * ct * opl;
        e6[0] += 0.5 * (Hp * 2.0 * st * ct - 0.5 * (Hr + Hb) * (2.0 * ct * st)) * eml;
        e6[1] += 0.5 * hl * 2.0 * st * ct * cp * eml;
        e6[2] += 0.5 * hl * 2.0 * st * ct * sp * eml;
        e6[3] += -0.5 * (Hr - Hb) * st * eml;
        k7[1] += -0.5 * hl * st * st * 2.0 * sp * opl;
        k7[2] += 0.5 * hl * st * st * 2.0 * cp * opl;
        k7[2 * 4 + 3] += -0.5 * fl * st * st * 2.0 * sp * opl;
        k7[3 * 4 + 1] += 0.5 * fl * st * st * 2.0 * cp * opl;
        e7[1] += -0.5 * hl * st * st * 2.0 * sp * eml;
        e7[2] += 0.5 * hl * st * st * 2.0 * cp * eml;
      }
    }
  }
  /* symmetrise the matrices (atmos_rad_synthesis.cc:131-160) */
  for (int q = -1; q < 7; ++q)
    for (int z = 0; z < nd; ++z) {
      double *Km = (q < 0) ? op + 16 * z : &DOP(q, z, 0);
      Km[5] = Km[10] = Km[15] = Km[0];
      Km[4] = Km[1];
      Km[8] = Km[2];
      Km[12] = Km[3];
      Km[2 * 4 + 1] = -Km[1 * 4 + 2];
      Km[1 * 4 + 3] = -Km[3 * 4 + 1];
      Km[3 * 4 + 2] = -Km[2 * 4 + 3];
    }
  /* normalisation with the referent-opacity derivative (atmos_rad.cc:187-201) */
  if (ray->tau_grid)
    for (int z = 0; z < nd; ++z) {
      const double r = ST(SNAPI_F_OPREF, z);
      for (int p = 0; p < 7; ++p) {
        const double dr = (p < 2) ? DS(p, SNAPI_F_OPREF, z) : 0.0;
        for (int i = 0; i < 16; ++i) DOP(p, z, i) = DOP(p, z, i) / r - op[16 * z + i] * dr / r / r;
        for (int i = 0; i < 4; ++i) DEM(p, z, i) = DEM(p, z, i) / r - em[4 * z + i] * dr / r / r;
      }
      for (int i = 0; i < 16; ++i) op[16 * z + i] /= r;
      for (int i = 0; i < 4; ++i) em[4 * z + i] /= r;
    }
#undef DOP
#undef DEM
}

/* Step multiplier of the central differences (1 = the reference's steps);
 * smaller values approach the exact derivative, for checks only. */
static double g_fd_scale = 1.0;
void orc_set_fd_scale(double s) { g_fd_scale = s; }

int orc_responses(const snapi_line_table *lt, const snapi_ray *ray, int ncol, int nd, const double *state_all,
                  const double *dstate_all, int nlambda, const double *lam, double *stokes, double *resp,
                  int nthreads) {
  const int nf = SNAPI_NF_BASE + 2 * lt->nline;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (long long job = 0; job < (long long)ncol * nlambda; ++job) {
    const int col = (int)(job / nlambda), l = (int)(job % nlambda);
    const double *state = state_all + (size_t)col * nf * nd;
    const double *dstate = dstate_all + (size_t)col * 2 * nf * nd;
    double *op = malloc(sizeof(double) * 16 * nd), *em = malloc(sizeof(double) * 4 * nd);
    double *dop = malloc(sizeof(double) * 7 * 16 * nd), *dem = malloc(sizeof(double) * 7 * 4 * nd);
    double *op2 = malloc(sizeof(double) * 16 * nd), *em2 = malloc(sizeof(double) * 4 * nd);
    double *S = malloc(sizeof(double) * 4 * nd), *Sp = malloc(sizeof(double) * 4 * nd);
    opem_pert(lt, ray, nd, nf, state, dstate, lam[l], op, em, dop, dem);
    const double *grid = state + (size_t)SNAPI_F_GRID * nd;
    orc_formal_pol(nd, grid, op, em, ray->theta, -1.0, S);
    for (int s = 0; s < 4; ++s) stokes[((size_t)col * nlambda + l) * 4 + s] = S[s];
    for (int p = 0; p < 7; ++p)
      for (int zk = 0; zk < nd; ++zk) {
        /* the caller scales Nt and angle perturbations before the difference
         * and divides the result by the same factor */
        double sc = g_fd_scale;
        if (p == 1) sc *= ST(SNAPI_F_NT, zk) * DELTA_NT_FRAC;
        if (p == 5 || p == 6) sc *= DELTA_ANGLE;
        const double *dk = dop + ((size_t)p * nd + zk) * 16, *de = dem + ((size_t)p * nd + zk) * 4;
        double res[4];
        for (int sgn = 0; sgn < 2; ++sgn) {
          const double f = sgn == 0 ? 0.5 * sc : -0.5 * sc;
          memcpy(op2, op, sizeof(double) * 16 * nd);
          memcpy(em2, em, sizeof(double) * 4 * nd);
          for (int i = 0; i < 16; ++i) op2[16 * zk + i] += f * dk[i];
          for (int i = 0; i < 4; ++i) em2[4 * zk + i] += f * de[i];
          orc_formal_pol(nd, grid, op2, em2, ray->theta, -1.0, Sp);
          for (int s = 0; s < 4; ++s) res[s] = (sgn == 0) ? Sp[s] : res[s] - Sp[s];
        }
        for (int s = 0; s < 4; ++s)
          resp[((((size_t)col * 7 + p) * nd + zk) * nlambda + l) * 4 + s] = res[s] / sc;
      }
    free(op);
    free(em);
    free(dop);
    free(dem);
    free(op2);
    free(em2);
    free(S);
    free(Sp);
  }
  return 0;
}
