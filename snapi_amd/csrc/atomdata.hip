// Host-side atomic data of the bound-bound transitions: the line table the
// synthesis kernels read, built from level data as the reference's atom
// class builds it.
//   Einstein coefficients        atom::atom (atmos/atmol/atom/atom.cc:233-253)
//   radiative damping            atom::damp_rad (atom.cc:926-935)
//   van der Waals cross-section  atom::compute_damp_col and
//                                interpolate_col_damp_sp/_pd/_df (atom.cc:987-1308)
//   Zeeman pattern               atom::zeeman_setup (atom.cc:1620-1773)
// No device code: these are per-transition constants computed once per
// table.
#include <hip/hip_runtime.h>

#include <cmath>

#include "abo_tables.h"
#include "rt_common.h"

namespace snapi {

// interpol_1d (cmlib/mathtools.cc:1057-1117): quadratic Bezier between the
// bracketing points, control point the mean of the two ends' Fritsch-Butland
// tangents (one-sided at the ends of the table), the end values outside.
static double bezier_interp_1d(const double *f, const double *x, int n, double xi) {
  if (!(xi < x[n - 1] && xi > x[0])) return xi >= x[n - 1] ? f[n - 1] : f[0];
  int i = 0;
  while (i <= n - 2 && !(xi < x[i + 1])) ++i;
  auto fb = [](double d1, double d2, double h1, double h2) {
    const double a = (1.0 + h2 / (h1 + h2)) / 3.0;
    return d1 * d2 > 0.0 ? d1 * d2 / (a * d2 + (1.0 - a) * d1) : 0.0;
  };
  double h1 = 0.0, h2 = 0.0, t0;
  if (i > 0) {
    h1 = x[i] - x[i - 1];
    h2 = x[i + 1] - x[i];
    t0 = fb((f[i] - f[i - 1]) / h1, (f[i + 1] - f[i]) / h2, h1, h2);
  } else {
    t0 = (f[1] - f[0]) / (x[1] - x[0]);
  }
  // the left end's control value uses the step to its right neighbour (zero
  // in the first interval), as the reference's running h2 does
  const double c0 = f[i] + h2 * 0.5 * t0;
  double t1;
  if (i + 1 < n - 1) {
    h1 = x[i + 1] - x[i];
    h2 = x[i + 2] - x[i + 1];
    t1 = fb((f[i + 1] - f[i]) / h1, (f[i + 2] - f[i + 1]) / h2, h1, h2);
  } else {
    t1 = (f[n - 1] - f[n - 2]) / (x[n - 1] - x[n - 2]);
  }
  const double c1 = f[i + 1] - h1 * 0.5 * t1;
  const double cm = 0.5 * (c0 + c1);
  const double u = (xi - x[i]) / (x[i + 1] - x[i]);
  return (1.0 - u) * (1.0 - u) * f[i] + u * u * f[i + 1] + 2.0 * u * (1.0 - u) * cm;
}

// interpol_2d (cmlib/mathtools.cc:1006-1041): interpol_1d along the columns
// of the three rows around x, then along those three values.
static double bezier_interp_2d(const double *f, const double *x, const double *y, int nx, int ny, double xi,
                               double yi) {
  int i = 0;
  for (i = 0; i <= nx - 2; ++i)
    if (xi < x[i + 1]) break;
  int r0 = i - 1;
  if (i == 0) r0 = 0;
  else if (i >= nx - 1) r0 = i - 2;
  double v[3], xs[3];
  for (int k = 0; k < 3; ++k) {
    v[k] = bezier_interp_1d(f + (size_t)(r0 + k) * ny, y, ny, yi);
    xs[k] = x[r0 + k];
  }
  return bezier_interp_1d(v, xs, 3, xi);
}

// One of the three tables at (n* of the lower-l state, n* of the upper-l
// state): rows start at n* = row0, columns at col0, both in steps of 0.1.
static void abo_lookup(const double *sig, const double *alp, int nr, int nc, double row0, double col0, double nrow,
                       double ncol, double &sigma, double &alpha) {
  double xr[32], xc[32];
  for (int i = 0; i < nr; ++i) xr[i] = row0 + i * 0.1;
  for (int i = 0; i < nc; ++i) xc[i] = col0 + i * 0.1;
  const double s = bezier_interp_2d(sig, xr, xc, nr, nc, nrow, ncol);
  alpha = bezier_interp_2d(alp, xr, xc, nr, nc, nrow, ncol);
  // atomic units at 10 km/s to the reference's cross-section of
  // w = N_H v (v / 1e6)^-alpha sigma (atom.cc:939-960)
  sigma = s * 5.62E-17 * pow(4.0 / kPi, alpha / 2.0) * tgamma(2.0 - alpha / 2.0);
}

}  // namespace snapi

using namespace snapi;

extern "C" int snapi_zeeman_pattern(int j2_lower, int j2_upper, double g_lower, double g_upper, int *split,
                                    int *ncomp, snapi_component *comp, int max_comp) {
  if (!split || !ncomp || j2_lower < 0 || j2_upper < 0 || (max_comp > 0 && !comp)) return SNAPI_ERR_ARG;
  const int dj = j2_upper - j2_lower;
  *split = (dj == 2 || dj == -2 || (dj == 0 && j2_upper != 0)) ? 1 : 0;
  ncomp[0] = ncomp[1] = ncomp[2] = 0;
  if (!*split) return SNAPI_OK;
  const int j2min = j2_upper < j2_lower ? j2_upper : j2_lower;
  const int np = j2min + 1, ns = (j2_upper + j2_lower) / 2;
  if (np + 2 * ns > max_comp) return SNAPI_ERR_ARG;
  ncomp[0] = np;
  ncomp[1] = ncomp[2] = ns;
  snapi_component *P = comp, *Bl = comp + np, *Rd = comp + np + ns;
  const double ju = j2_upper / 2.0, jl = j2_lower / 2.0, jmin = j2min / 2.0, dg = g_lower - g_upper;
  double sp = 0.0, sb = 0.0, sr = 0.0;
  if (dj == 0) {
    // Delta J = 0: pi from m = -J..J, sigma over the J + J transitions
    for (int i = 0; i < np; ++i) {
      const double m = i - ju;
      P[i] = snapi_component{m * dg, m * m};
      sp += m * m;
    }
    for (int i = 0; i < ns; ++i) {
      const double mb = i - ju + 1.0, mr = i - ju;
      Bl[i] = snapi_component{mb * dg - g_lower, 0.5 * (ju + mb) * (ju - mb + 1.0)};
      Rd[i] = snapi_component{mr * dg + g_lower, 0.5 * (ju - mr) * (ju + mr + 1.0)};
      sb += Bl[i].strength;
      sr += Rd[i].strength;
    }
  } else {
    // Delta J = +-1: 2 Jmin + 1 components in each group
    const int off = j2_lower > j2_upper ? 1 : 0;
    for (int i = 0; i < np; ++i) {
      const double m = i - jmin;
      const double up = 0.5 * (jmin + m + 1.0) * (jmin + m + 2.0), dn = 0.5 * (jmin - m + 1.0) * (jmin - m + 2.0);
      P[i] = snapi_component{m * dg, (jmin + 1.0) * (jmin + 1.0) - m * m};
      Bl[i] = snapi_component{(m - off) * dg - g_upper, jl < ju ? up : dn};
      Rd[i] = snapi_component{(m + off) * dg + g_upper, jl < ju ? dn : up};
      sp += P[i].strength;
      sb += Bl[i].strength;
      sr += Rd[i].strength;
    }
  }
  for (int i = 0; i < np; ++i) P[i].strength /= sp;
  for (int i = 0; i < ns; ++i) {
    Bl[i].strength /= sb;
    Rd[i].strength /= sr;
  }
  return SNAPI_OK;
}

extern "C" int snapi_vdw_damping(int stage, double e_ion, double e_lower, double e_upper, int l_lower, int l_upper,
                                 double *sigma, double *alpha) {
  if (!sigma || !alpha || stage < 0 || !(e_ion > e_lower) || !(e_ion > e_upper)) return SNAPI_ERR_ARG;
  const double zc = stage + 1.0;
  // effective principal quantum numbers from the binding energies (Rydberg 2.18e-11 erg)
  const double nl = sqrt(zc * zc * 2.18E-11 / (e_ion - e_lower));
  const double nu = sqrt(zc * zc * 2.18E-11 / (e_ion - e_upper));
  const int ls = l_lower + l_upper;
  if (stage == 0 && ls == 1) {  // s-p
    const double ns_ = l_upper == 1 ? nl : nu, np_ = l_upper == 1 ? nu : nl;
    abo_lookup(kAbo_sp_sigma, kAbo_sp_alpha, kAbo_sp_rows, kAbo_sp_cols, 1.0, 1.3, ns_, np_, *sigma, *alpha);
  } else if (stage == 0 && ls == 3) {  // p-d
    const double nd_ = l_upper == 2 ? nu : nl, np_ = l_upper == 2 ? nl : nu;
    abo_lookup(kAbo_pd_sigma, kAbo_pd_alpha, kAbo_pd_rows, kAbo_pd_cols, 1.3, 2.3, np_, nd_, *sigma, *alpha);
  } else if (stage == 0 && ls == 5) {  // d-f
    const double nf_ = l_upper == 3 ? nu : nl, nd_ = l_upper == 3 ? nl : nu;
    abo_lookup(kAbo_df_sigma, kAbo_df_alpha, kAbo_df_rows, kAbo_df_cols, 1.3, 2.3, nd_, nf_, *sigma, *alpha);
  } else {
    // Unsold's hydrogenic <r^2> and C6 (Gray, The observation and analysis of
    // stellar photospheres, 3rd ed., eq. 11.29), alpha = 0.4
    const double ru = nu * nu / 2.0 / (zc * zc) * (5.0 * nu * nu + 1.0);
    const double rl = nl * nl / 2.0 / (zc * zc) * (5.0 * nl * nl + 1.0);
    const double c6 = 1.41 * 6.46E-34 * (ru - rl);
    *alpha = 0.4;
    *sigma = 17.0 * pow(c6, 0.4) / 251.0;
  }
  return SNAPI_OK;
}

extern "C" int snapi_radiative_damping(int nlevel, const double *A, int lower, int upper, double *grad) {
  if (!A || !grad || nlevel < 2 || lower < 0 || upper <= lower || upper >= nlevel) return SNAPI_ERR_ARG;
  double gu = 0.0, gl = 0.0;
  for (int l = 0; l < upper; ++l) gu += A[(size_t)upper * nlevel + l];
  for (int l = 0; l < lower; ++l) gl += A[(size_t)lower * nlevel + l];
  *grad = gu + gl;
  return SNAPI_OK;
}

extern "C" int snapi_line_add(snapi_line_table *t, double mass, int stage, double e_ion, const snapi_level *lower,
                              const snapi_level *upper, double aul, double grad) {
  if (!t || !lower || !upper || !(mass > 0.0) || !(upper->energy > lower->energy) || !(lower->g > 0.0) ||
      !(upper->g > 0.0) || aul < 0.0)
    return SNAPI_ERR_ARG;
  if (t->nline < 0 || t->nline >= SNAPI_MAX_LINES || t->ncomp_total < 0) return SNAPI_ERR_ARG;
  snapi_line &L = t->line[t->nline];
  L = snapi_line{};
  const double lam = kH * kC / (upper->energy - lower->energy);
  L.lam0 = lam;
  L.aul = aul;
  // B_ul = lambda^3 / (2 h c) A_ul, B_lu = g_u / g_l B_ul, both per unit
  // wavelength (x lambda^2 / c), the reference's units
  const double bul = lam * lam * lam / (2.0 * kH * kC) * aul;
  L.bul = bul * lam * lam / kC;
  L.blu = upper->g / lower->g * bul * lam * lam / kC;
  L.grad = grad;
  L.mass = mass;
  int rc = snapi_vdw_damping(stage, e_ion, lower->energy, upper->energy, lower->l, upper->l, &L.sigma, &L.alpha);
  if (rc) return rc;
  int split = 0, nc[3] = {0, 0, 0};
  rc = snapi_zeeman_pattern(lower->j2, upper->j2, lower->g_lande, upper->g_lande, &split, nc, t->comp + t->ncomp_total,
                            SNAPI_MAX_COMPONENTS - t->ncomp_total);
  if (rc) return rc;
  L.split = split;
  L.ncomp[0] = nc[0];
  L.ncomp[1] = nc[1];
  L.ncomp[2] = nc[2];
  L.comp0 = t->ncomp_total;
  t->ncomp_total += nc[0] + nc[1] + nc[2];
  t->nline += 1;
  return SNAPI_OK;
}
