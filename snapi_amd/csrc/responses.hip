// Per-depth response functions and the chi^2 gradient of the polarized
// synthesis, by a hand-written reverse (adjoint) sweep of the formal
// solution on gfx950.  One work-item per (column, wavelength):
//   1. forward: opacity build + recurrence as in synthesis (synth_kernel.h),
//      K, E and I of every depth point kept in a per-work-item scratch area;
//   2. reverse: adjoints of I along the ray, back through the 4x4 solves, the
//      short-characteristics weights, the optical-depth steps and the
//      Fritsch-Butland derivative, giving dI_out/d(K, E) at every depth for
//      up to four seeds (the four emergent Stokes components, or one seed
//      w_s (S - S_obs) for the gradient);
//   3. contraction: at every depth the local derivatives of (K, E) along the
//      seven physical parameters (T, Nt, vt, vLOS, |B|, inclination,
//      azimuth), restating the analytic part of the reference's
//      op_em_vector_plus_pert (atmos/atmos_rad_synthesis.cc:76-162,
//      atmos/atmol/atom/atom_rad_synthesis.cc:771-1026, atmos/atmol/mol/
//      mol.cc:414-469, profile derivatives atom_rad.cc:611-664, 795-875) and
//      the referent-opacity normalisation (atmos/atmos_rad.cc:187-201).
// The reference propagates those local derivatives through the solver by a
// central difference (formal_pert_numerical, plpar/atmos_ppbez_rts.cc:
// 1708-1776); the adjoint gives the same derivative without the 2 x 7 x ND
// extra formal solutions per wavelength.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "rt_common.h"
#include "synth_kernel.h"
#include "resp_kernel.h"
#include "deriv_kernel.h"

namespace snapi {

// scratch: [depth k][field][work-item]
template <typename R> struct Scratch {
  R *p;
  long long W;
  int nd;
  __device__ R &fwd(int k, int f, long long w) const { return p[((long long)k * kFwd + f) * W + w]; }
  __device__ R &adj(int k, int f, int s, long long w) const {
    return p[(long long)nd * kFwd * W + (((long long)k * kAdj + f) * 4 + s) * W + w];
  }
};

// Local derivatives of the normalised opacity/emissivity of depth z along
// the seven physical parameters.  dK[p][7] in the compact order
// (eta_I, eta_Q, eta_U, eta_V, rho_Q, rho_U, rho_V), dE[p][4].
template <typename S>
__device__ void local_derivs(const KLines &L, const SynthArgs &a, const S *__restrict__ st, const S *__restrict__ dst,
                             int z, double lam, double dK[7][7], double dE[7][4]) {
  const int nd = a.nd, nf = a.nfield;
#define S_(f) ((double)st[(size_t)(f)*nd + z])
#define D_(p, f) ((double)dst[((size_t)(p)*nf + (f)) * nd + z])
  for (int p = 0; p < 7; ++p) {
    for (int i = 0; i < 7; ++i) dK[p][i] = 0.0;
    for (int i = 0; i < 4; ++i) dE[p][i] = 0.0;
  }
  double K[7] = {0, 0, 0, 0, 0, 0, 0}, E[4] = {0, 0, 0, 0};
  const double T = S_(SNAPI_F_T), Ne = S_(SNAPI_F_NE), nhi = S_(SNAPI_F_NHI), Nt = S_(SNAPI_F_NT);
  const double dNe[2] = {D_(0, SNAPI_F_NE), D_(1, SNAPI_F_NE)};
  const double dNH[2] = {D_(0, SNAPI_F_NHI), D_(1, SNAPI_F_NHI)};
  const double lm = a.lam_mid;
  auto planck = [](double l, double t) {
    return 2.0 * 6.62606957 * 8.9875517E-7 / l / l / l / l / l / (exp(1.4387769 / l / t) - 1.0);
  };
  auto planck_dT = [&](double l, double t) {
    const double x = 1.4387769 / l / t;
    const double tmp = (x > 5.0) ? 1.0 : exp(x) / (exp(x) - 1.0);
    return planck(l, t) * tmp * 1.4387769 / l / t / t;
  };
  // Thomson at the middle wavelength and the atomic bound-free background
  const double opth = Ne * kThomson;
  K[0] = S_(SNAPI_F_BGOP);
  E[0] = S_(SNAPI_F_BGEM);
  for (int p = 0; p < 2; ++p) {
    const double o = dNe[p] * kThomson;
    dK[p][0] += o;
    dE[p][0] += planck(lm, T) * o + (p == 0 ? opth * planck_dT(lm, T) : 0.0);
  }
  if (a.hminus) {
    const double lh = a.hminus_mid ? lm : lam;
    const double ac[7] = {1.99654, -1.18267E-5, 2.64243E-6, -4.40524E-10, 3.23992E-14, -1.39568E-18, 2.78701E-23};
    double alpha = 0.0, pw = 1.0;
    for (int i = 0; i < 7; ++i) {
      alpha += ac[i] * pw;
      pw *= lh * 1E8;
    }
    const double TP = 5040. / T;
    const double stim = exp(-kH * kC / lh / kK / T);
    const double obf = 4.158E-28 * alpha * Ne * kK * T * pow(TP, 2.5) * pow(10.0, 0.754 * TP) * nhi * (1.0 - stim);
    const double l = log10(lh * 1E8);
    const double f0 = -2.2763 - 1.6850 * l + 0.76661 * l * l - 0.053346 * l * l * l;
    const double f1 = 15.2827 - 9.2846 * l + 1.99381 * l * l - 0.142631 * l * l * l;
    const double f2 = -197.789 + 190.266 * l - 67.9775 * l * l + 10.6913 * l * l * l - 0.625151 * l * l * l * l;
    const double ltp = log10(TP);
    const double ef = pow(10.0, f0 + f1 * ltp + f2 * ltp * ltp);
    const double off = 1E-26 * Ne * kK * T * ef * nhi;
    const double o = obf + off;
    const double B = planck(lh, T);
    K[0] += o;
    E[0] += o * B;
    double d[2];
    d[0] = obf * (1.0 / T + 2.5 * 1.0 / TP * (-5040.0 / T / T)) + obf * dNe[0] / Ne + obf * dNH[0] / nhi +
           obf * (-2.302585 * 0.754 * 5040.0 / T / T) + obf * (-stim * kH * kC / lh / kK / T / T) / (1.0 - stim);
    d[1] = obf * dNe[1] / Ne + obf * dNH[1] / nhi;
    const double dTP = -TP / T;
    const double dltp = 1.0 / 2.302585 / TP * dTP;
    const double def = ef * 2.302585 * (f1 * dltp + 2.0 * f2 * ltp * dltp);
    const double cf = 1E-26 * kK;
    d[0] += cf * Ne * ef * nhi + cf * T * Ne * nhi * def + cf * T * ef * (Ne * dNH[0] + nhi * dNe[0]);
    d[1] += cf * T * ef * (Ne * dNH[1] + nhi * dNe[1]);
    for (int p = 0; p < 2; ++p) {
      dK[p][0] += d[p];
      dE[p][0] += B * d[p] + (p == 0 ? o * planck_dT(lh, T) : 0.0);
    }
  }
  const double vt = S_(SNAPI_F_VT), vl = S_(SNAPI_F_VLOS), bm = S_(SNAPI_F_BMAG);
  double thb = S_(SNAPI_F_BINC) * kPi / 180.0, phb = S_(SNAPI_F_BAZI) * kPi / 180.0;
  if (bm < 0.001) thb = phb = 0.0;
  const double sp = sin(2.0 * phb), cp = cos(2.0 * phb), sn = sin(thb), cs = cos(thb);
  for (int il = 0; il < L.nline; ++il) {
    const KLine &ln = L.l[il];
    const double l0 = ln.lam0;
    const double nlo = S_(SNAPI_NF_BASE + 2 * il), nup = S_(SNAPI_NF_BASE + 2 * il + 1);
    const double dvd = sqrt(ln.twok_m * T + vt * vt);
    const double dld = dvd * l0 / kC;
    const double vm = sqrt(ln.vmfac * T);
    const double vpow = pow(vm / 1E6, -ln.alpha);
    const double gc = nhi * vm * vpow * ln.sigma;
    const double av = ln.sf * (ln.grad + gc) / dld;
    const double opl = nlo * ln.opf_l - nup * ln.opf_u;
    const double eml = nup * ln.emf;
    double dopl[2], deml[2];
    for (int p = 0; p < 2; ++p) {
      dopl[p] = D_(p, SNAPI_NF_BASE + 2 * il) * ln.opf_l - D_(p, SNAPI_NF_BASE + 2 * il + 1) * ln.opf_u;
      deml[p] = D_(p, SNAPI_NF_BASE + 2 * il + 1) * ln.emf;
    }
    const double ddld_T = l0 / (kC * dvd) * (0.5 * ln.twok_m);
    const double ddld_vt = l0 / (kC * dvd) * vt;
    const double vmd = 0.5 / vm * ln.vmfac;
    double dgc_T = nhi * ln.sigma * (vmd * vpow - vm * ln.alpha * pow(vm / 1E6, -ln.alpha - 1.0) * vmd / 1E6);
    dgc_T += dNH[0] * vm * vpow * ln.sigma;
    const double a_T = -av / dld * ddld_T + dgc_T / dld * ln.sf;
    const double a_N = (gc * 1E-3 / dld * ln.sf) / (1E-3 * Nt);
    const double a_vt = -av / dld * ddld_vt;
    const double lshift = lam * (1.0 + vl / kC);
    if (!ln.split) {
      double H, F, dH, dF;
      const double x = (l0 - lshift) / dld;
      voigt_d<double>(x, av, H, F, dH, dF);
      K[0] += opl * H / dld;
      E[0] += eml * H / dld;
      const double xd[4] = {-x / dld * ddld_T, 0.0, -x / dld * ddld_vt, -lam / kC / dld};
      const double ad[4] = {a_T, a_N, a_vt, 0.0};
      const double dd[4] = {ddld_T, 0.0, ddld_vt, 0.0};
      for (int p = 0; p < 4; ++p) {
        const double g = (dH * xd[p] - dF * ad[p] - H / dld * dd[p]) / dld;
        dK[p][0] += opl * g;
        dE[p][0] += eml * g;
      }
      for (int p = 0; p < 2; ++p) {
        dK[p][0] += dopl[p] * H / dld;
        dE[p][0] += deml[p] * H / dld;
      }
      continue;
    }
    const double lB = ln.lbfac * bm;
    double Hs[3] = {0, 0, 0}, Fs[3] = {0, 0, 0}, Hd[3][5], Fd[3][5];
    for (int g = 0; g < 3; ++g)
      for (int p = 0; p < 5; ++p) Hd[g][p] = Fd[g][p] = 0.0;
    int c = ln.c0;
    for (int g = 0; g < 3; ++g)
      for (int m = 0; m < ln.nc[g]; ++m, ++c) {
        const double sh = L.shift[c], sw = L.str[c];
        const double x = (l0 + sh * lB - lshift) / dld;
        double H, F, dH, dF;
        voigt_d<double>(x, av, H, F, dH, dF);
        Hs[g] += H * sw;
        Fs[g] += F * sw;
        const double xd[5] = {-x / dld * ddld_T, 0.0, -x / dld * ddld_vt, -lam / kC / dld, sh * ln.lbfac / dld};
        const double ad[5] = {a_T, a_N, a_vt, 0.0, 0.0};
        for (int p = 0; p < 5; ++p) {
          Hd[g][p] += (dH * xd[p] - dF * ad[p]) * sw;
          Fd[g][p] += (dF * xd[p] + dH * ad[p]) * sw;
        }
      }
    for (int g = 0; g < 3; ++g) {
      Hs[g] /= dld;
      Fs[g] /= dld;
      for (int p = 0; p < 5; ++p) {
        Hd[g][p] /= dld;
        Fd[g][p] /= dld;
      }
      Hd[g][0] -= ddld_T / dld * Hs[g];
      Hd[g][2] -= ddld_vt / dld * Hs[g];
      Fd[g][0] -= ddld_T / dld * Fs[g];
      Fd[g][2] -= ddld_vt / dld * Fs[g];
    }
    // pattern geometry (atom_rad_synthesis.cc:486-497) applied to (H, F)
    auto geom = [&](const double *H3, const double *F3, double fo, double fe, double *KK, double *EE) {
      const double eI = 0.5 * (H3[0] * sn * sn + 0.5 * (H3[2] + H3[1]) * (1.0 + cs * cs));
      const double hl = H3[0] - 0.5 * (H3[2] + H3[1]), fl = F3[0] - 0.5 * (F3[2] + F3[1]);
      const double eQ = 0.5 * hl * sn * sn * cp, eU = 0.5 * hl * sn * sn * sp, eV = 0.5 * (H3[2] - H3[1]) * cs;
      KK[0] += eI * fo;
      KK[1] += eQ * fo;
      KK[2] += eU * fo;
      KK[3] += eV * fo;
      KK[4] += 0.5 * fl * sn * sn * cp * fo;
      KK[5] += 0.5 * fl * sn * sn * sp * fo;
      KK[6] += 0.5 * (F3[2] - F3[1]) * cs * fo;
      EE[0] += eI * fe;
      EE[1] += eQ * fe;
      EE[2] += eU * fe;
      EE[3] += eV * fe;
    };
    geom(Hs, Fs, opl, eml, K, E);
    for (int p = 0; p < 2; ++p) geom(Hs, Fs, dopl[p], deml[p], dK[p], dE[p]);
    for (int p = 0; p < 5; ++p) {
      const double H3[3] = {Hd[0][p], Hd[1][p], Hd[2][p]}, F3[3] = {Fd[0][p], Fd[1][p], Fd[2][p]};
      geom(H3, F3, opl, eml, dK[p], dE[p]);
    }
    // inclination (rho_V term as the reference writes it, :999) and azimuth
    const double Hp = Hs[0], Hb = Hs[1], Hr = Hs[2], Fp = Fs[0], Fb = Fs[1], Fr = Fs[2];
    const double hl = Hp - 0.5 * (Hr + Hb), fl = Fp - 0.5 * (Fr + Fb);
    const double gI = 0.5 * (Hp * 2.0 * sn * cs - 0.5 * (Hr + Hb) * (2.0 * cs * sn));
    dK[5][0] += gI * opl;
    dK[5][1] += 0.5 * hl * 2.0 * sn * cs * cp * opl;
    dK[5][2] += 0.5 * hl * 2.0 * sn * cs * sp * opl;
    dK[5][3] += -0.5 * (Hr - Hb) * sn * opl;
    dK[5][4] += 0.5 * fl * 2.0 * cs * sn * cp * opl;
    dK[5][5] += 0.5 * fl * 2.0 * cs * sn * sp * opl;
    dK[5][6] += 0.5 * (Fr - Fb) * cs * opl;
    dE[5][0] += gI * eml;
    dE[5][1] += 0.5 * hl * 2.0 * sn * cs * cp * eml;
    dE[5][2] += 0.5 * hl * 2.0 * sn * cs * sp * eml;
    dE[5][3] += -0.5 * (Hr - Hb) * sn * eml;
    dK[6][1] += -0.5 * hl * sn * sn * 2.0 * sp * opl;
    dK[6][2] += 0.5 * hl * sn * sn * 2.0 * cp * opl;
    dK[6][4] += -0.5 * fl * sn * sn * 2.0 * sp * opl;
    dK[6][5] += 0.5 * fl * sn * sn * 2.0 * cp * opl;
    dE[6][1] += -0.5 * hl * sn * sn * 2.0 * sp * eml;
    dE[6][2] += 0.5 * hl * sn * sn * 2.0 * cp * eml;
  }
  if (a.tau_grid) {
    const double r = S_(SNAPI_F_OPREF);
    const double dr[2] = {D_(0, SNAPI_F_OPREF), D_(1, SNAPI_F_OPREF)};
    for (int p = 0; p < 7; ++p) {
      const double drp = p < 2 ? dr[p] : 0.0;
      for (int i = 0; i < 7; ++i) dK[p][i] = dK[p][i] / r - K[i] * drp / r / r;
      for (int i = 0; i < 4; ++i) dE[p][i] = dE[p][i] / r - E[i] * drp / r / r;
    }
  }
#undef S_
#undef D_
}

// derivatives of the linear weights with respect to the step (both branches)
template <typename R> __device__ __forceinline__ void sc_weights_d(R tb, R &detb, R &dw0, R &dw1) {
  if (tb < R(1E-2)) {
    detb = R(-1) + tb - R(0.5) * tb * tb;
    dw0 = R(0.5) - R(2.0 / 3.0) * tb + R(3.0 / 8.0) * tb * tb;
    dw1 = R(0.5) - tb * R(1.0 / 3.0) + tb * tb * R(0.125);
  } else {
    const R e = rexp<R>(-tb), i = R(1) / tb;
    detb = -e;
    dw0 = -i * i + e * (R(1) + i) + e * i * i;
    dw1 = i * i - e * i - e * i * i;
  }
}

// Fritsch-Butland derivative and its partials with respect to (om, o0, op)
template <typename R>
__device__ __forceinline__ R fb_deriv_d(R om, R o0, R op, R h1, R h2, R &gm, R &g0, R &gp) {
  const R d2 = (op - o0) / h2, d1 = (o0 - om) / h1;
  gm = g0 = gp = R(0);
  if (d1 * d2 > R(0)) {
    const R al = (R(1) + h1 / (h1 + h2)) * R(1.0 / 3.0);
    const R den = al * d1 + (R(1) - al) * d2;
    const R D = d1 * d2 / den;
    const R dD1 = d2 * d2 * (R(1) - al) / (den * den);
    const R dD2 = d1 * d1 * al / (den * den);
    gm = -dD1 / h1;
    g0 = dD1 / h1 - dD2 / h2;
    gp = dD2 / h2;
    return D;
  }
  return R(0);
}

// adjoint of y = K~ x with respect to the six reduced-matrix elements
template <typename R> __device__ __forceinline__ void kmul_adj(const R yb[4], const R x[4], R kb[6]) {
  kb[0] += yb[0] * x[1] + yb[1] * x[0];   // a
  kb[1] += yb[0] * x[2] + yb[2] * x[0];   // b
  kb[2] += yb[0] * x[3] + yb[3] * x[0];   // c
  kb[3] += yb[1] * x[2] - yb[2] * x[1];   // d
  kb[4] += -yb[1] * x[3] + yb[3] * x[1];  // e
  kb[5] += yb[2] * x[3] - yb[3] * x[2];   // f
}

template <typename R, int NS, bool STAGED>
__global__ __launch_bounds__(kRespThreads) void resp_kernel(const KLines lines, const SynthArgs args,
                                                           const RespArgs ra, const R *__restrict__ state,
                                                           const R *__restrict__ dstate,
                                                           const double *__restrict__ lam, Scratch<R> scr,
                                                           R *__restrict__ stokes, R *__restrict__ out,
                                                           int *__restrict__ status) {
  extern __shared__ unsigned char smem_raw[];
  R *sm = reinterpret_cast<R *>(smem_raw);
  const int nd = args.nd, nl = args.nlambda, nf = args.nfield;
  const int cl = blockIdx.x;  // column within the launch
  const int col = ra.col0 + cl;
  if (ra.active && !ra.active[col]) return;  // uniform over the workgroup
  const R *st = state + (size_t)col * nf * nd;
  const R *dst = dstate + (size_t)col * 2 * nf * nd;
  ColumnTile<R> tile = ColumnTile<R>::carve(sm, nd, lines.nline);
  // the wavelength-independent derivative factors, staged once per column
  // (when they fit and the background's T and Nt derivatives do not depend
  // on the wavelength: H- at the middle wavelength, as the response path
  // takes it); otherwise local_derivs evaluates everything per wavelength
  R *sd = sm + ColumnTile<R>::bytes(nd, lines.nline) / sizeof(R);
  DerivTile<R> dtile = DerivTile<R>::carve(sd, nd, lines.nline, 0);
  ThermoTile<R> ttile = ThermoTile<R>::carve(sd + DerivTile<R>::bytes(nd, lines.nline, 0, false) / sizeof(R), nd,
                                             lines.nline);
  stage_column<R, R>(lines, args, st, tile);
  if (STAGED) {
    stage_deriv<R, R>(lines, args, st, dtile);
    stage_thermo<R, R>(lines, args, st, dst, ttile);
  }
  __syncthreads();
  const bool up = args.ct < 0.0;
  const R mct = (R)(-args.ct);
  constexpr int ns = NS;  // seeds: 4 (the emergent Stokes components) or 1 (the gradient)
  auto zof = [&](int k) { return up ? nd - 1 - k : k; };
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    const long long w = (long long)cl * nl + l;
    const double lm = lam[l];
    LamConst<R> c;
    lambda_consts<R>(lines, args, lm, c);
    // ---------------- forward: opacity of every depth, then the recurrence
    for (int k = 0; k < nd; ++k) {
      R K[7], E[4];
      opacity_at<R>(lines, args, tile, c, zof(k), K, E);
      for (int i = 0; i < 7; ++i) scr.fwd(k, i, w) = K[i];
      for (int i = 0; i < 4; ++i) scr.fwd(k, 7 + i, w) = E[i];
      for (int i = 0; i < kAdj; ++i)
        for (int s = 0; s < ns; ++s) scr.adj(k, i, s, w) = R(0);
    }
    auto load = [&](int k, R K[7], R E[4]) {
      for (int i = 0; i < 7; ++i) K[i] = scr.fwd(k, i, w);
      for (int i = 0; i < 4; ++i) E[i] = scr.fwd(k, 7 + i, w);
    };
    auto hof = [&](int k) { return (tile.grid[zof(k)] - tile.grid[zof(k - 1)]) / mct; };
    auto ooof = [&](int k) { return scr.fwd(k, 0, w); };
    auto Dof = [&](int k) {  // D_k, 0 <= k <= nd-2
      if (k == 0) return (ooof(1) - ooof(0)) / hof(1);
      return fb_deriv<R>(ooof(k - 1), ooof(k), ooof(k + 1), hof(k), hof(k + 1));
    };
    R Io[4];
    {
      R K0[7], E0[4];
      load(0, K0, E0);
      R I[4] = {up ? E0[0] / K0[0] : R(0), R(0), R(0), R(0)};
      for (int s = 0; s < 4; ++s) scr.fwd(0, 11 + s, w) = I[s];
      for (int k = 1; k <= nd - 2; ++k) {
        R Kp[7], Ep[4], Kc[7], Ec[4];
        load(k - 1, Kp, Ep);
        load(k, Kc, Ec);
        Kred<R> kp, kc;
        R Sp[4], Sc[4];
        reduce_ks<R>(Kp, Ep, kp, Sp);
        reduce_ks<R>(Kc, Ec, kc, Sc);
        const R h = hof(k);
        const R tb = h * R(0.5) * (Kp[0] + Kc[0]) + h * h * R(1.0 / 12.0) * (Dof(k - 1) - Dof(k));
        R etb, w0, w1;
        sc_weights<R>(tb, etb, w0, w1);
        R kv[4], B[4], In[4];
        kmul<R>(kp, I, kv);
        for (int s = 0; s < 4; ++s) B[s] = -w0 * kv[s] + w0 * Sp[s] + w1 * Sc[s] + I[s] * etb;
        solve_1pk<R>(kc, w1, B, In);
        for (int s = 0; s < 4; ++s) {
          I[s] = In[s];
          scr.fwd(k, 11 + s, w) = In[s];
        }
      }
      const int ko = nd >= 3 ? nd - 2 : 0;
      for (int s = 0; s < 4; ++s) Io[s] = scr.fwd(ko, 11 + s, w);
      for (int s = 0; s < 4; ++s) stokes[((size_t)col * nl + l) * 4 + s] = Io[s];
      if (!(fabs((double)Io[0]) < 1e300)) atomicOr(status, 2);
    }
    // ---------------- reverse sweep
    R Ib[4][4];  // [seed][component]
    for (int s = 0; s < 4; ++s)
      for (int q = 0; q < 4; ++q) Ib[s][q] = R(0);
    if (ra.sink == SINK_GRAD) {
      const R *ob = static_cast<const R *>(ra.obs) + ((size_t)col * nl + l) * 4;
      for (int q = 0; q < 4; ++q) Ib[0][q] = (R)ra.ws[q] * (Io[q] - ob[q]);
    } else {
      for (int s = 0; s < 4; ++s) Ib[s][s] = R(1);
    }
    // carried adjoints of the reduced quantities of depth k (from step k+1)
    R Kr_b[4][6], S_b[4][4], D_b[4];
    for (int s = 0; s < 4; ++s) {
      for (int q = 0; q < 6; ++q) Kr_b[s][q] = R(0);
      for (int q = 0; q < 4; ++q) S_b[s][q] = R(0);
      D_b[s] = R(0);
    }
    // converts the adjoint of (K~, S) of depth k into raw (K, E) adjoints
    auto flush_reduced = [&](int k, const R Kr[4][6], const R Sb[4][4]) {
      R K[7], E[4];
      load(k, K, E);
      const R iI = R(1) / K[0];
      const int map[6] = {1, 2, 3, 6, 5, 4};  // a..f <- K[1], K[2], K[3], K[6], K[5], K[4]
      for (int s = 0; s < ns; ++s) {
        R kI = R(0);
        for (int q = 0; q < 6; ++q) {
          scr.adj(k, map[q], s, w) += Kr[s][q] * iI;
          kI -= Kr[s][q] * K[map[q]] * iI * iI;
        }
        for (int q = 0; q < 4; ++q) {
          scr.adj(k, 7 + q, s, w) += Sb[s][q] * iI;
          kI -= Sb[s][q] * E[q] * iI * iI;
        }
        scr.adj(k, 0, s, w) += kI;
      }
    };
    // propagates the adjoint of D_k into the opacities it is made of
    auto flush_D = [&](int k, const R Db[4]) {
      if (k == 0) {
        const R h1 = hof(1);
        for (int s = 0; s < ns; ++s) {
          scr.adj(1, 0, s, w) += Db[s] / h1;
          scr.adj(0, 0, s, w) -= Db[s] / h1;
        }
        return;
      }
      R gm, g0, gp;
      fb_deriv_d<R>(ooof(k - 1), ooof(k), ooof(k + 1), hof(k), hof(k + 1), gm, g0, gp);
      for (int s = 0; s < ns; ++s) {
        scr.adj(k - 1, 0, s, w) += Db[s] * gm;
        scr.adj(k, 0, s, w) += Db[s] * g0;
        scr.adj(k + 1, 0, s, w) += Db[s] * gp;
      }
    };
    for (int k = nd - 2; k >= 1; --k) {
      R Kp[7], Ep[4], Kc[7], Ec[4];
      load(k - 1, Kp, Ep);
      load(k, Kc, Ec);
      Kred<R> kp, kc;
      R Sp[4], Sc[4];
      reduce_ks<R>(Kp, Ep, kp, Sp);
      reduce_ks<R>(Kc, Ec, kc, Sc);
      const R h = hof(k);
      const R Dm = Dof(k - 1), D0 = Dof(k);
      const R tb = h * R(0.5) * (Kp[0] + Kc[0]) + h * h * R(1.0 / 12.0) * (Dm - D0);
      R etb, w0, w1, detb, dw0, dw1;
      sc_weights<R>(tb, etb, w0, w1);
      sc_weights_d<R>(tb, detb, dw0, dw1);
      R Ip[4], Ik[4];
      for (int q = 0; q < 4; ++q) {
        Ip[q] = scr.fwd(k - 1, 11 + q, w);
        Ik[q] = scr.fwd(k, 11 + q, w);
      }
      R kvp[4], kvc[4];
      kmul<R>(kp, Ip, kvp);
      kmul<R>(kc, Ik, kvc);
      // transposed system: (1 + w1 K~)^T = 1 + w1 K~' with d, e, f negated
      const Kred<R> kcT = {kc.a, kc.b, kc.c, -kc.d, -kc.e, -kc.f};
      const Kred<R> kpT = {kp.a, kp.b, kp.c, -kp.d, -kp.e, -kp.f};
      R Krc_b[4][6], Sc_b[4][4], Krp_b[4][6], Sp_b[4][4], D0_b[4], Dm_b[4];
      for (int s = 0; s < ns; ++s) {
        R Bb[4];
        solve_1pk<R>(kcT, w1, Ib[s], Bb);
        // A = 1 + w1 K~_k :  Abar = -Bb (x) I_k
        R kb[6] = {0, 0, 0, 0, 0, 0};
        const R nb[4] = {-w1 * Bb[0], -w1 * Bb[1], -w1 * Bb[2], -w1 * Bb[3]};
        kmul_adj<R>(nb, Ik, kb);
        R w1b = -(Bb[0] * kvc[0] + Bb[1] * kvc[1] + Bb[2] * kvc[2] + Bb[3] * kvc[3]);
        w1b += Bb[0] * Sc[0] + Bb[1] * Sc[1] + Bb[2] * Sc[2] + Bb[3] * Sc[3];
        R w0b = R(0), etbb = R(0);
        for (int q = 0; q < 4; ++q) {
          w0b += Bb[q] * (Sp[q] - kvp[q]);
          etbb += Bb[q] * Ip[q];
        }
        for (int q = 0; q < 6; ++q) Krc_b[s][q] = Kr_b[s][q] + kb[q];
        for (int q = 0; q < 4; ++q) Sc_b[s][q] = S_b[s][q] + w1 * Bb[q];
        // B = -w0 K~_{k-1} I_{k-1} + w0 S_{k-1} + w1 S_k + etb I_{k-1}
        const R vb[4] = {-w0 * Bb[0], -w0 * Bb[1], -w0 * Bb[2], -w0 * Bb[3]};
        R kbp[6] = {0, 0, 0, 0, 0, 0};
        kmul_adj<R>(vb, Ip, kbp);
        R Ipb[4];
        kmul<R>(kpT, vb, Ipb);
        for (int q = 0; q < 4; ++q) Ib[s][q] = Ipb[q] + etb * Bb[q];
        for (int q = 0; q < 6; ++q) Krp_b[s][q] = kbp[q];
        for (int q = 0; q < 4; ++q) Sp_b[s][q] = w0 * Bb[q];
        const R tbb = w0b * dw0 + w1b * dw1 + etbb * detb;
        // tb = h/2 (oo_{k-1} + oo_k) + h^2/12 (D_{k-1} - D_k)
        scr.adj(k - 1, 0, s, w) += tbb * h * R(0.5);
        scr.adj(k, 0, s, w) += tbb * h * R(0.5);
        Dm_b[s] = tbb * h * h * R(1.0 / 12.0);
        D0_b[s] = D_b[s] - tbb * h * h * R(1.0 / 12.0);
      }
      // depth k is complete for K~, S; D_k is complete
      flush_reduced(k, Krc_b, Sc_b);
      flush_D(k, D0_b);
      for (int s = 0; s < ns; ++s) {
        for (int q = 0; q < 6; ++q) Kr_b[s][q] = Krp_b[s][q];
        for (int q = 0; q < 4; ++q) S_b[s][q] = Sp_b[s][q];
        D_b[s] = Dm_b[s];
      }
    }
    // bottom point: boundary I_0 = (E_0[0]/K_0[0], 0, 0, 0) for the outgoing ray
    if (nd >= 3) flush_D(0, D_b);
    flush_reduced(0, Kr_b, S_b);
    if (up) {
      R K0b[7], E0b[4];
      load(0, K0b, E0b);
      for (int s = 0; s < ns; ++s) {
        scr.adj(0, 7, s, w) += Ib[s][0] / K0b[0];
        scr.adj(0, 0, s, w) -= Ib[s][0] * E0b[0] / (K0b[0] * K0b[0]);
      }
    }
    // ---------------- contraction with the local derivatives; node sinks
    // keep g[s][p] of every depth in the adjoint area of the scratch
    for (int k = 0; k < nd; ++k) {
      const int z = zof(k);
      if constexpr (STAGED) {
        R Kz[7], Ez[4], G[NQ_TH][11];
        deriv_at<R, true, kRowsAll>(lines, args, tile, dtile, ttile, c, z, kRowsAll, Kz, Ez, G);
        for (int s = 0; s < ns; ++s) {
          R Kb[7], Eb[4];
          for (int i = 0; i < 7; ++i) Kb[i] = scr.adj(k, i, s, w);
          for (int i = 0; i < 4; ++i) Eb[i] = scr.adj(k, 7 + i, s, w);
#pragma unroll
          for (int p = 0; p < 7; ++p) {
            const int q = qrow(p);
            R g = R(0);
#pragma unroll
            for (int i = 0; i < 7; ++i) g += Kb[i] * G[q][i];
#pragma unroll
            for (int i = 0; i < 4; ++i) g += Eb[i] * G[q][7 + i];
            if (ra.sink == SINK_RESP)
              out[((((size_t)col * 7 + p) * nd + z) * nl + l) * 4 + s] = g;
            else if (ra.sink == SINK_GRAD)
              atomicAdd(&out[((size_t)col * 7 + p) * nd + z], g);
            else
              scr.adj(k, p, s, w) = g;  // entries 0..6 of depth k are no longer read
          }
        }
      } else {
      double dK[7][7], dE[7][4];
      local_derivs<R>(lines, args, st, dst, z, lm, dK, dE);
      for (int s = 0; s < ns; ++s) {
        double Kb[7], Eb[4];
        for (int i = 0; i < 7; ++i) Kb[i] = (double)scr.adj(k, i, s, w);
        for (int i = 0; i < 4; ++i) Eb[i] = (double)scr.adj(k, 7 + i, s, w);
        for (int p = 0; p < 7; ++p) {
          double g = 0.0;
          for (int i = 0; i < 7; ++i) g += Kb[i] * dK[p][i];
          for (int i = 0; i < 4; ++i) g += Eb[i] * dE[p][i];
          if (ra.sink == SINK_RESP)
            out[((((size_t)col * 7 + p) * nd + z) * nl + l) * 4 + s] = (R)g;
          else if (ra.sink == SINK_GRAD)
            atomicAdd(&out[((size_t)col * 7 + p) * nd + z], (R)g);
          else
            scr.adj(k, p, s, w) = (R)g;  // entries 0..6 of depth k are no longer read
        }
      }
      }
    }
    if (ra.sink == SINK_NODE) {
      const R *nm = static_cast<const R *>(ra.nmap) + (size_t)col * ra.nnode * nd;
      const R mk = ra.mask ? static_cast<const R *>(ra.mask)[(size_t)col * nl + l] : R(1);
      for (int s = 0; s < 4; ++s)
        for (int n = 0; n < ra.nnode; ++n) {
          const int p = ra.pidx[n];
          R acc = R(0);
          for (int k = 0; k < nd; ++k) acc += scr.adj(k, p, s, w) * nm[(size_t)n * nd + zof(k)];
          if (ra.nmap2 && ra.pidx2[n] >= 0) {
            const R *nm2 = static_cast<const R *>(ra.nmap2) + (size_t)col * ra.nnode * nd;
            const int p2 = ra.pidx2[n];
            for (int k = 0; k < nd; ++k) acc += scr.adj(k, p2, s, w) * nm2[(size_t)n * nd + zof(k)];
          }
          out[((size_t)col * nl * 4 + (size_t)l * 4 + s) * ra.nnode + n] = acc * (R)(ra.nscale[n] * ra.ws[s]) * mk;
        }
    }
  }
}

template <typename R>
int launch_resp(const KLines &kl, const SynthArgs &a, RespArgs ra, int ncol, const void *ds_, const void *dd_,
                       const double *dl, void *dstk_, void *dout_, void *dscr_, int chunk, hipStream_t stream) {
  int threads = ((a.nlambda + 63) / 64) * 64;
  if (threads > kRespThreads) threads = kRespThreads;
  size_t smem = ColumnTile<R>::bytes(a.nd, kl.nline);
  const size_t sderiv = DerivTile<R>::bytes(a.nd, kl.nline, 0, false) + ThermoTile<R>::bytes(a.nd, kl.nline);
  // the staged T and Nt rows take H- at the middle wavelength; with H- at
  // every wavelength they hold only when no output reads those rows (a node
  // Jacobian of kinematic and magnetic nodes)
  bool thermo_rows = ra.sink != SINK_NODE;
  for (int n = 0; n < ra.nnode && !thermo_rows; ++n)
    thermo_rows = ra.pidx[n] < 2 || (ra.nmap2 && ra.pidx2[n] >= 0);
  ra.staged_derivs = (!a.hminus || a.hminus_mid || !thermo_rows) && smem + sderiv <= 65536;
  if (ra.staged_derivs) smem += sderiv;
  int *status = device_status_slot();
  if (!status) return SNAPI_ERR_DEVICE;
  for (int c0 = 0; c0 < ncol; c0 += chunk) {
    const int nc = (ncol - c0 < chunk) ? ncol - c0 : chunk;
    ra.col0 = c0;
    ra.nwork = (long long)nc * a.nlambda;
    Scratch<R> scr{static_cast<R *>(dscr_), ra.nwork, a.nd};
    auto kern = ra.staged_derivs ? (ra.nseed == 1 ? resp_kernel<R, 1, true> : resp_kernel<R, 4, true>)
                                 : (ra.nseed == 1 ? resp_kernel<R, 1, false> : resp_kernel<R, 4, false>);
    hipLaunchKernelGGL(kern, dim3(nc), dim3(threads), smem, stream, kl, a, ra, static_cast<const R *>(ds_),
                       static_cast<const R *>(dd_), dl, scr, static_cast<R *>(dstk_), static_cast<R *>(dout_),
                       status);
    if (hipGetLastError() != hipSuccess) return SNAPI_ERR_DEVICE;
  }
  return SNAPI_OK;
}

// bytes of scratch per column
size_t resp_scratch_per_col(int nlambda, int nd, size_t esz) {
  return (size_t)nlambda * nd * (kFwd + kAdj * 4) * esz;
}

template int launch_resp<double>(const KLines &, const SynthArgs &, RespArgs, int, const void *, const void *,
                                 const double *, void *, void *, void *, int, hipStream_t);
template int launch_resp<float>(const KLines &, const SynthArgs &, RespArgs, int, const void *, const void *,
                                const double *, void *, void *, void *, int, hipStream_t);

// Angle and velocity responses from the ray's frame to the atmosphere's
// (atmosphere::transform_responses, atmos/atmos_rad.cc:53-101): per depth,
// with the atmosphere-frame field (Bx, By, Bz), the derivatives of the ray
// frame's inclination and azimuth (acos(B'_z/|B|), atan(B'_y/B'_x) of the
// rotated field) with respect to the atmosphere frame's inclination and
// azimuth (acos(Bz/|B|), atan(By/Bx)); the velocity response is multiplied
// by cos(phi) as the reference does (:89).  Where the field vanishes or
// points along the ray the reference divides by zero; the angle responses
// are zero there.  One work-item per (column, depth, wavelength).
template <typename R>
__global__ __launch_bounds__(256) void transform_resp_kernel(int ncol, int nd, int nl, const double *__restrict__ bvec,
                                                             double theta, double phi, R *__restrict__ resp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = (long long)ncol * nd * nl;
  if (i >= n) return;
  const int l = (int)(i % nl);
  const int z = (int)((i / nl) % nd);
  const int col = (int)(i / ((long long)nl * nd));
  const double cp = cos(phi), sp = sin(phi), ct = cos(kPi - theta), st = sin(kPi - theta);
  const double *b = bvec + (size_t)col * 3 * nd;
  const double Bx = b[z], By = b[nd + z], Bz = b[2 * nd + z];
  const double nx = ct * cp * Bx + sp * By - st * cp * Bz;
  const double ny = -ct * sp * Bx + cp * By + st * sp * Bz;
  const double nz = st * Bx + ct * Bz;
  const double bm = sqrt(Bx * Bx + By * By + Bz * Bz);
  double ii_t = 0.0, ii_p = 0.0, az_t = 0.0, az_p = 0.0;
  if (bm > 0.0 && nx * nx + ny * ny > 0.0 && (Bx != 0.0 || By != 0.0)) {
    const double q = -1.0 / sqrt(1.0 - nz * nz / bm / bm);
    const double dix = (st * bm - Bx / bm * nz) / bm / bm * q;
    const double diy = (-By / bm * nz) / bm / bm * q;
    const double diz = (ct * bm - Bz / bm * nz) / bm / bm * q;
    const double r = 1.0 / (nx * nx) / (1.0 + ny * ny / nx / nx);
    const double dax = (-ct * sp * nx - ct * cp * ny) * r;
    const double day = (cp * nx - sp * ny) * r;
    const double daz = (st * sp * nx + st * cp * ny) * r;
    const double tb = acos(Bz / bm), pb = atan(By / Bx);
    const double bxt = bm * cos(tb) * cos(pb), bxp = -bm * sin(tb) * sin(pb);
    const double byt = bm * cos(tb) * sin(pb), byp = bm * sin(tb) * cos(pb);
    const double bzt = -bm * sin(tb);
    ii_t = dix * bxt + diy * byt + diz * bzt;
    ii_p = dix * bxp + diy * byp;
    az_t = dax * bxt + day * byt + daz * bzt;
    az_p = dax * bxp + day * byp;
  }
  auto at = [&](int p) { return resp + (((((size_t)col * 7 + p) * nd + z) * nl + l) * 4); };
  R *v = at(3), *ri = at(5), *ra = at(6);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double to_in = (double)ri[s], to_az = (double)ra[s];
    v[s] = (R)((double)v[s] * cp);
    ri[s] = (R)(to_in * ii_t + to_az * az_t);
    ra[s] = (R)(to_in * ii_p + to_az * az_p);
  }
}

template <typename R>
static int launch_transform(int ncol, int nd, int nl, const double *bvec, double theta, double phi, void *resp,
                            hipStream_t stream) {
  const long long n = (long long)ncol * nd * nl;
  const long long blocks = (n + 255) / 256;
  if (blocks > 0x7fffffffLL) return SNAPI_ERR_ARG;
  hipLaunchKernelGGL(transform_resp_kernel<R>, dim3((unsigned)blocks), dim3(256), 0, stream, ncol, nd, nl, bvec,
                     theta, phi, static_cast<R *>(resp));
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

static int run_resp(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd,
                    const void *state, const void *dstate, int nlambda, const double *lam, void *stokes,
                    void *out_host, int sink, const void *obs, const double *ws, const double *bvec = nullptr,
                    double phi = 0.0) {
  KLines kl;
  SynthArgs a;
  int rc = prepare_synth(lines, ray, ncol, nd, nlambda, kl, a);
  if (rc) return rc;
  if (!state || !dstate || !lam || !stokes || !out_host) return SNAPI_ERR_ARG;
  if (sink == SINK_GRAD && (!obs || !ws)) return SNAPI_ERR_ARG;
  // per-depth responses are formed in the atmosphere's frame only on the
  // mu = 1 ray (theta = pi), where atmosphere::transform_responses
  // (atmos/atmos_rad.cc:53-101) is the identity
  // (with the atmosphere-frame field given, any ray: the responses are
  // transformed as the reference transforms them)
  if (!bvec && !(cos(ray->theta) < -0.999999)) return SNAPI_ERR_UNSUPPORTED;
  if (bvec && sink != SINK_RESP) return SNAPI_ERR_ARG;
  const size_t es = precision == SNAPI_FP64 ? 8 : 4;
  const size_t nf = a.nfield;
  const size_t sb = (size_t)ncol * nf * nd * es, db = 2 * sb, lb = (size_t)nlambda * 8;
  const size_t stb = (size_t)ncol * nlambda * 4 * es;
  const size_t ob = (sink == SINK_RESP ? (size_t)ncol * 7 * nd * nlambda * 4 : (size_t)ncol * 7 * nd) * es;
  const size_t per_col = resp_scratch_per_col(nlambda, nd, es);
  size_t chunk_s = ((size_t)2 << 30) / (per_col > 0 ? per_col : 1);
  if (chunk_s < 1) chunk_s = 1;
  if (chunk_s > (size_t)ncol) chunk_s = (size_t)ncol;
  const int chunk = (int)chunk_s;
  void *ds = nullptr, *dd = nullptr, *dstk = nullptr, *dout = nullptr, *dscr = nullptr, *dobs = nullptr;
  double *dl = nullptr, *dbv = nullptr;
  const size_t bvb = bvec ? (size_t)ncol * 3 * nd * sizeof(double) : 0;
  rc = SNAPI_OK;
  if (hipMalloc(&ds, sb) != hipSuccess || hipMalloc(&dd, db) != hipSuccess ||
      hipMalloc((void **)&dl, lb) != hipSuccess || hipMalloc(&dstk, stb) != hipSuccess ||
      hipMalloc(&dout, ob) != hipSuccess || hipMalloc(&dscr, per_col * chunk) != hipSuccess ||
      (sink == SINK_GRAD && hipMalloc(&dobs, stb) != hipSuccess) ||
      (bvec && hipMalloc((void **)&dbv, bvb) != hipSuccess)) {
    rc = SNAPI_ERR_DEVICE;
  } else {
    if (hipMemcpy(ds, state, sb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dd, dstate, db, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dl, lam, lb, hipMemcpyHostToDevice) != hipSuccess || hipMemset(dout, 0, ob) != hipSuccess ||
        (dobs && hipMemcpy(dobs, obs, stb, hipMemcpyHostToDevice) != hipSuccess) ||
        (dbv && hipMemcpy(dbv, bvec, bvb, hipMemcpyHostToDevice) != hipSuccess))
      rc = SNAPI_ERR_DEVICE;
    else
      rc = reset_status();
  }
  if (rc == SNAPI_OK) {
    RespArgs ra = {};
    ra.nseed = sink == SINK_GRAD ? 1 : 4;
    ra.sink = sink;
    ra.obs = dobs;
    for (int s = 0; s < 4; ++s) ra.ws[s] = ws ? ws[s] : 1.0;
    if (precision == SNAPI_FP64)
      rc = launch_resp<double>(kl, a, ra, ncol, ds, dd, dl, dstk, dout, dscr, chunk, 0);
    else
      rc = launch_resp<float>(kl, a, ra, ncol, ds, dd, dl, dstk, dout, dscr, chunk, 0);
    if (rc == SNAPI_OK && dbv)
      rc = precision == SNAPI_FP64 ? launch_transform<double>(ncol, nd, nlambda, dbv, ray->theta, phi, dout, 0)
                                   : launch_transform<float>(ncol, nd, nlambda, dbv, ray->theta, phi, dout, 0);
    if (rc == SNAPI_OK && hipDeviceSynchronize() != hipSuccess) rc = SNAPI_ERR_DEVICE;
    if (rc == SNAPI_OK) {
      if (hipMemcpy(stokes, dstk, stb, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(out_host, dout, ob, hipMemcpyDeviceToHost) != hipSuccess)
        rc = SNAPI_ERR_DEVICE;
      else
        rc = read_status();
    }
  }
  hipFree(ds);
  hipFree(dd);
  hipFree(dl);
  hipFree(dstk);
  hipFree(dout);
  hipFree(dscr);
  hipFree(dobs);
  hipFree(dbv);
  return rc;
}

}  // namespace snapi

using namespace snapi;

extern "C" int snapi_stokes_responses(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol,
                                      int nd, const void *state, const void *dstate, int nlambda, const double *lam,
                                      void *stokes, void *resp) {
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  return run_resp(precision, lines, ray, ncol, nd, state, dstate, nlambda, lam, stokes, resp, SINK_RESP, nullptr,
                  nullptr);
}

extern "C" int snapi_stokes_adjoint(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol,
                                    int nd, const void *state, const void *dstate, int nlambda, const double *lam,
                                    const void *obs, const double *wstokes, void *grad) {
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  if (!grad || ncol <= 0 || nlambda <= 0 || nd < 2 || nd > SNAPI_MAX_ND) return SNAPI_ERR_ARG;
  // the emergent spectrum is computed on the way and not returned
  std::vector<double> tmp;
  try {
    tmp.resize((size_t)ncol * nlambda * 4);
  } catch (...) {
    return SNAPI_ERR_ARG;
  }
  return run_resp(precision, lines, ray, ncol, nd, state, dstate, nlambda, lam, tmp.data(), grad, SINK_GRAD, obs,
                  wstokes);
}

extern "C" int snapi_stokes_responses_atm(int precision, const snapi_line_table *lines, const snapi_ray *ray,
                                          double phi, const double *bvec, int ncol, int nd, const void *state,
                                          const void *dstate, int nlambda, const double *lam, void *stokes,
                                          void *resp) {
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  if (!bvec) return SNAPI_ERR_ARG;
  return run_resp(precision, lines, ray, ncol, nd, state, dstate, nlambda, lam, stokes, resp, SINK_RESP, nullptr,
                  nullptr, bvec, phi);
}

extern "C" int snapi_transform_responses_dev(int precision, int ncol, int nd, int nlambda, const double *bvec_dev,
                                             double theta, double phi, void *resp_dev, void *stream) {
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  if (ncol <= 0 || nd <= 0 || nlambda <= 0 || !bvec_dev || !resp_dev) return SNAPI_ERR_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return precision == SNAPI_FP64 ? launch_transform<double>(ncol, nd, nlambda, bvec_dev, theta, phi, resp_dev, st)
                                 : launch_transform<float>(ncol, nd, nlambda, bvec_dev, theta, phi, resp_dev, st);
}

extern "C" int snapi_transform_responses(int precision, int ncol, int nd, int nlambda, const double *bvec,
                                         double theta, double phi, void *resp) {
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  if (ncol <= 0 || nd <= 0 || nlambda <= 0 || !bvec || !resp) return SNAPI_ERR_ARG;
  const size_t es = precision == SNAPI_FP64 ? 8 : 4;
  const size_t rb = (size_t)ncol * 7 * nd * nlambda * 4 * es, bb = (size_t)ncol * 3 * nd * sizeof(double);
  void *dr = nullptr;
  double *db = nullptr;
  int rc = SNAPI_OK;
  if (hipMalloc(&dr, rb) != hipSuccess || hipMalloc((void **)&db, bb) != hipSuccess) rc = SNAPI_ERR_DEVICE;
  else if (hipMemcpy(dr, resp, rb, hipMemcpyHostToDevice) != hipSuccess ||
           hipMemcpy(db, bvec, bb, hipMemcpyHostToDevice) != hipSuccess)
    rc = SNAPI_ERR_DEVICE;
  if (rc == SNAPI_OK) rc = snapi_transform_responses_dev(precision, ncol, nd, nlambda, db, theta, phi, dr, nullptr);
  if (rc == SNAPI_OK && (hipDeviceSynchronize() != hipSuccess ||
                         hipMemcpy(resp, dr, rb, hipMemcpyDeviceToHost) != hipSuccess))
    rc = SNAPI_ERR_DEVICE;
  if (dr) (void)hipFree(dr);
  if (db) (void)hipFree(db);
  return rc;
}
