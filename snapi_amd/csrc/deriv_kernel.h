// Local derivatives of the opacity and emissivity of one depth point along
// the seven physical parameters, from per-column tiles staged in LDS: shared
// by the forward-mode (tangent.hip) and reverse-mode (responses.hip) sweeps.
// They restate the analytic part of op_em_vector_plus_pert
// (atmos/atmos_rad_synthesis.cc:76-162, atmos/atmol/atom/
// atom_rad_synthesis.cc:771-1026, atmos/atmol/mol/mol.cc:414-469) and the
// referent-opacity normalisation (atmos/atmos_rad.cc:187-201).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "rt_common.h"
#include "synth_kernel.h"

namespace snapi {

// parameter rows of the local derivatives: the kinematic and magnetic ones,
// then (kernels with thermodynamic directions) temperature and density
enum { Q_VT = 0, Q_VL = 1, Q_B = 2, Q_TH = 3, Q_PH = 4, Q_T = 5, Q_N = 6, NQ = 5, NQ_TH = 7 };

// A set of parameter rows as a bit mask (bit q: row q above), and the
// position of row q among the rows of the set: kernels carry only the rows
// of the set their node directions read, in that compact order.
constexpr unsigned kRowsKin = 0x1Fu;  // vt, vLOS, |B|, inclination, azimuth
constexpr unsigned kRowsAll = 0x7Fu;  // and temperature, density
__host__ __device__ constexpr int nrows(unsigned m) { return m ? (int)(m & 1u) + nrows(m >> 1) : 0; }
__host__ __device__ constexpr int rslot(unsigned m, int q) { return nrows(m & ((1u << q) - 1u)); }
__host__ __device__ constexpr bool has_row(unsigned m, int q) { return (m >> q) & 1u; }
// the row at position `slot` of the set
__host__ __device__ constexpr int row_of(unsigned m, int slot, int q = 0) {
  return q > 31 ? -1 : has_row(m, q) ? (slot == 0 ? q : row_of(m, slot - 1, q + 1)) : row_of(m, slot, q + 1);
}

// parameter row of a per-depth response index (0 T, 1 Nt, 2 vt, 3 vLOS,
// 4 |B|, 5 inclination, 6 azimuth)
__host__ __device__ inline int qrow(int pidx) { return pidx >= 2 ? pidx - 2 : (pidx == 0 ? Q_T : Q_N); }

// LDS tile of the derivative factors of one column, after the ColumnTile.
template <typename R> struct DerivTile {
  R *sct, *dqt, *dut, *mhs;  // per depth: sin(th)cos(th), d qf/d th, d uf/d th, -sin(th)/2
  R *gv;                     // per line and depth: d ln(Doppler width) / d vt
  R *nm;                     // [node of the pass][z]: node map d q(z) / d node
  R *nm2;                    // [node of the pass][z]: the node's second row (Nt of a temperature node)
  // nn nodes in the pass; two: the second node map is held (passes with
  // temperature or density rows)
  static __host__ __device__ size_t bytes(int nd, int nline, int nn, bool two) {
    return sizeof(R) * (size_t)nd * (4 + nline + (two ? 2 : 1) * nn);
  }
  static __device__ DerivTile carve(R *p, int nd, int nline, int nn) {
    DerivTile t;
    t.sct = p;
    t.dqt = p + nd;
    t.dut = p + 2 * nd;
    t.mhs = p + 3 * nd;
    t.gv = p + 4 * nd;
    t.nm = p + (size_t)(4 + nline) * nd;
    t.nm2 = t.nm + (size_t)nn * nd;
    return t;
  }
};

// LDS tile of the temperature and density derivatives of one column (the
// response path's: the per-depth derivative fields of dstate), after the
// DerivTile.
template <typename R> struct ThermoTile {
  // per line and depth: d ln(Doppler width)/dT; the damping's T derivative
  // beyond its Doppler-width part; da/dNt; the line's normalised opacity and
  // emissivity factors differentiated through the level populations (d/dT,
  // d/dNt)
  R *gT, *adT, *aN, *pT, *eT, *pN, *eN;
  // per depth: d/dT, d/dNt of the background eta_I and eps_I (Thomson and H-
  // at the middle wavelength; the bound-free part the reference adds as 0),
  // normalised; d ln(kappa_ref)/dT, d ln(kappa_ref)/dNt
  R *bKT, *bET, *bKN, *bEN, *rT, *rN;
  static __host__ __device__ size_t bytes(int nd, int nline) { return sizeof(R) * (size_t)nd * (6 + 7 * nline); }
  static __device__ ThermoTile carve(R *p, int nd, int nline) {
    ThermoTile t;
    R **f[6] = {&t.bKT, &t.bET, &t.bKN, &t.bEN, &t.rT, &t.rN};
    for (int i = 0; i < 6; ++i) {
      *f[i] = p;
      p += nd;
    }
    R **g[7] = {&t.gT, &t.adT, &t.aN, &t.pT, &t.eT, &t.pN, &t.eN};
    for (int i = 0; i < 7; ++i) {
      *g[i] = p;
      p += (size_t)nd * nline;
    }
    return t;
  }
};

// The wavelength-independent parts of the T and Nt derivatives of the
// opacity of op_em_vector_plus_pert (atom_rad_synthesis.cc:771-1026, mol.cc:
// 414-469, Thomson and the referent-opacity normalisation of
// atmos_rad.cc:187-201), in double from the state and dstate
// [2][nfield][nd] (d/dT, d/dNt of NE, NHI, OPREF and the level populations).
// H- is taken at the middle wavelength, as that path does.
template <typename R, typename S>
__device__ void stage_thermo(const KLines &L, const SynthArgs &a, const S *__restrict__ st, const S *__restrict__ dst,
                             ThermoTile<R> &t) {
  const int nd = a.nd, nf = a.nfield;
  const double lm = a.lam_mid;
  auto planck = [](double l, double T) {
    return 2.0 * 6.62606957 * 8.9875517E-7 / l / l / l / l / l / (exp(1.4387769 / l / T) - 1.0);
  };
  auto planck_dT = [&](double l, double T) {
    const double x = 1.4387769 / l / T;
    const double tmp = (x > 5.0) ? 1.0 : exp(x) / (exp(x) - 1.0);
    return planck(l, T) * tmp * 1.4387769 / l / T / T;
  };
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    const double T = (double)st[SNAPI_F_T * nd + z], Ne = (double)st[SNAPI_F_NE * nd + z];
    const double nhi = (double)st[SNAPI_F_NHI * nd + z];
    double dNe[2], dNH[2], dr[2];
    for (int p = 0; p < 2; ++p) {
      dNe[p] = (double)dst[((size_t)p * nf + SNAPI_F_NE) * nd + z];
      dNH[p] = (double)dst[((size_t)p * nf + SNAPI_F_NHI) * nd + z];
      dr[p] = (double)dst[((size_t)p * nf + SNAPI_F_OPREF) * nd + z];
    }
    const double r = a.tau_grid ? (double)st[SNAPI_F_OPREF * nd + z] : 1.0;
    const double iref = 1.0 / r;
    double dK[2], dE[2];
    const double opth = Ne * kThomson;
    for (int p = 0; p < 2; ++p) {
      const double o = dNe[p] * kThomson;
      dK[p] = o;
      dE[p] = planck(lm, T) * o + (p == 0 ? opth * planck_dT(lm, T) : 0.0);
    }
    if (a.hminus) {
      const double ac[7] = {1.99654, -1.18267E-5, 2.64243E-6, -4.40524E-10, 3.23992E-14, -1.39568E-18, 2.78701E-23};
      double alpha = 0.0, pw = 1.0;
      for (int i = 0; i < 7; ++i) {
        alpha += ac[i] * pw;
        pw *= lm * 1E8;
      }
      const double TP = 5040. / T;
      const double stim = exp(-kH * kC / lm / kK / T);
      const double obf = 4.158E-28 * alpha * Ne * kK * T * pow(TP, 2.5) * pow(10.0, 0.754 * TP) * nhi * (1.0 - stim);
      const double l = log10(lm * 1E8);
      const double f0 = -2.2763 - 1.6850 * l + 0.76661 * l * l - 0.053346 * l * l * l;
      const double f1 = 15.2827 - 9.2846 * l + 1.99381 * l * l - 0.142631 * l * l * l;
      const double f2 = -197.789 + 190.266 * l - 67.9775 * l * l + 10.6913 * l * l * l - 0.625151 * l * l * l * l;
      const double ltp = log10(TP);
      const double ef = pow(10.0, f0 + f1 * ltp + f2 * ltp * ltp);
      const double off = 1E-26 * Ne * kK * T * ef * nhi;
      const double o = obf + off;
      const double B = planck(lm, T);
      double d[2];
      d[0] = obf * (1.0 / T + 2.5 * 1.0 / TP * (-5040.0 / T / T)) + obf * dNe[0] / Ne + obf * dNH[0] / nhi +
             obf * (-2.302585 * 0.754 * 5040.0 / T / T) + obf * (-stim * kH * kC / lm / kK / T / T) / (1.0 - stim);
      d[1] = obf * dNe[1] / Ne + obf * dNH[1] / nhi;
      const double dltp = 1.0 / 2.302585 / TP * (-TP / T);
      const double def = ef * 2.302585 * (f1 * dltp + 2.0 * f2 * ltp * dltp);
      const double cf = 1E-26 * kK;
      d[0] += cf * Ne * ef * nhi + cf * T * Ne * nhi * def + cf * T * ef * (Ne * dNH[0] + nhi * dNe[0]);
      d[1] += cf * T * ef * (Ne * dNH[1] + nhi * dNe[1]);
      for (int p = 0; p < 2; ++p) {
        dK[p] += d[p];
        dE[p] += B * d[p] + (p == 0 ? o * planck_dT(lm, T) : 0.0);
      }
    }
    t.bKT[z] = (R)(dK[0] * iref);
    t.bET[z] = (R)(dE[0] * iref);
    t.bKN[z] = (R)(dK[1] * iref);
    t.bEN[z] = (R)(dE[1] * iref);
    t.rT[z] = (R)(a.tau_grid ? dr[0] / r : 0.0);
    t.rN[z] = (R)(a.tau_grid ? dr[1] / r : 0.0);
  }
  for (int i = threadIdx.x; i < nd * L.nline; i += blockDim.x) {
    const int il = i / nd, z = i - il * nd;
    const KLine &ln = L.l[il];
    const double T = (double)st[SNAPI_F_T * nd + z], vt = (double)st[SNAPI_F_VT * nd + z];
    const double nhi = (double)st[SNAPI_F_NHI * nd + z], Nt = (double)st[SNAPI_F_NT * nd + z];
    const double iref = a.tau_grid ? 1.0 / (double)st[SNAPI_F_OPREF * nd + z] : 1.0;
    const double dNH0 = (double)dst[((size_t)0 * nf + SNAPI_F_NHI) * nd + z];
    const double dvd2 = ln.twok_m * T + vt * vt;
    const double dld = sqrt(dvd2) * ln.lam0 / kC;
    const double vm = sqrt(ln.vmfac * T);
    const double vpow = pow(vm / 1E6, -ln.alpha);
    const double gc = nhi * vm * vpow * ln.sigma;
    const double vmd = 0.5 / vm * ln.vmfac;
    double dgc_T = nhi * ln.sigma * (vmd * vpow - vm * ln.alpha * pow(vm / 1E6, -ln.alpha - 1.0) * vmd / 1E6);
    dgc_T += dNH0 * vm * vpow * ln.sigma;
    t.gT[i] = (R)(0.5 * ln.twok_m / dvd2);
    t.adT[i] = (R)(ln.sf * dgc_T / dld);
    // the reference's density derivative of the damping (a_der for Nt)
    t.aN[i] = (R)((gc * 1E-3 / dld * ln.sf) / (1E-3 * Nt));
    const int fl = SNAPI_NF_BASE + 2 * il, fu = fl + 1;
    for (int p = 0; p < 2; ++p) {
      const double dlo = (double)dst[((size_t)p * nf + fl) * nd + z], dup = (double)dst[((size_t)p * nf + fu) * nd + z];
      const R po = (R)((dlo * ln.opf_l - dup * ln.opf_u) / dld * iref);
      const R pe = (R)(dup * ln.emf / dld * iref);
      if (p == 0) {
        t.pT[i] = po;
        t.eT[i] = pe;
      } else {
        t.pN[i] = po;
        t.eN[i] = pe;
      }
    }
  }
}

template <typename R, typename S>
__device__ void stage_deriv(const KLines &L, const SynthArgs &a, const S *__restrict__ st, DerivTile<R> &t) {
  const int nd = a.nd;
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    const double bm = (double)st[SNAPI_F_BMAG * nd + z];
    double thb = (double)st[SNAPI_F_BINC * nd + z] * kPi / 180.0;
    double phb = (double)st[SNAPI_F_BAZI * nd + z] * kPi / 180.0;
    if (bm < 0.001) thb = phb = 0.0;
    const double sn = sin(thb), cs = cos(thb), sp = sin(2.0 * phb), cp = cos(2.0 * phb);
    t.sct[z] = (R)(sn * cs);
    t.dqt[z] = (R)(sn * cs * cp);
    t.dut[z] = (R)(sn * cs * sp);
    t.mhs[z] = (R)(-0.5 * sn);
  }
  for (int i = threadIdx.x; i < nd * L.nline; i += blockDim.x) {
    const int il = i / nd, z = i - il * nd;
    const double T = (double)st[SNAPI_F_T * nd + z];
    const double vt = (double)st[SNAPI_F_VT * nd + z];
    t.gv[i] = (R)(vt / (L.l[il].twok_m * T + vt * vt));
  }
}

// Opacity (K, E) of depth z and its derivatives G[q] = d(K, E)/dq along the
// parameters in qmask, entries 0..6 = K, 7..10 = E.  FULL = false computes
// only the eta_I entry G[q][0] (the look-ahead of the optical-depth step).
template <typename R, bool FULL, unsigned ROWS>
__device__ __forceinline__ void deriv_at(const KLines &L, const SynthArgs &a, const ColumnTile<R> &t,
                                         const DerivTile<R> &d, const ThermoTile<R> &th, const LamConst<R> &c, int z,
                                         unsigned qmask, R K[7], R E[4], R G[nrows(ROWS)][11]) {
  R KI = t.bgop[z], EI = t.bgem[z];
  if (a.hminus) {
    const R tinv = t.tinv[z];
    const R stim = R(1) - rexp<R>(-c.hckl * tinv);
    const R Lt = t.ltp[z];
    const R o = c.hbf * t.cbf[z] * stim + t.cff[z] * rexp2<R>(R(kLog2_10) * (c.f0 + c.f1 * Lt + c.f2 * Lt * Lt));
    const R B = c.plc * rrcp<R>(rexp<R>(c.ple * tinv) - R(1));
    KI += o;
    EI += o * B;
  }
  constexpr int NR = nrows(ROWS);
#pragma unroll
  for (int q = 0; q < NR; ++q)
#pragma unroll
    for (int i = 0; i < 11; ++i) G[q][i] = R(0);
  // each row's place in G (0 for a row outside the set, whose code is dead)
  constexpr int sVT = has_row(ROWS, Q_VT) ? rslot(ROWS, Q_VT) : 0;
  constexpr int sVL = has_row(ROWS, Q_VL) ? rslot(ROWS, Q_VL) : 0;
  constexpr int sB = has_row(ROWS, Q_B) ? rslot(ROWS, Q_B) : 0;
  constexpr int sTH = has_row(ROWS, Q_TH) ? rslot(ROWS, Q_TH) : 0;
  constexpr int sPH = has_row(ROWS, Q_PH) ? rslot(ROWS, Q_PH) : 0;
  constexpr int sT = has_row(ROWS, Q_T) ? rslot(ROWS, Q_T) : 0;
  constexpr int sN = has_row(ROWS, Q_N) ? rslot(ROWS, Q_N) : 0;
  const bool wvt = has_row(ROWS, Q_VT) && (qmask & (1u << Q_VT));
  const bool wvl = has_row(ROWS, Q_VL) && (qmask & (1u << Q_VL));
  const bool wb = has_row(ROWS, Q_B) && (qmask & (1u << Q_B));
  const bool wth = has_row(ROWS, Q_TH) && (qmask & (1u << Q_TH));
  const bool wph = has_row(ROWS, Q_PH) && (qmask & (1u << Q_PH));
  constexpr bool kTh = has_row(ROWS, Q_T) || has_row(ROWS, Q_N);
  const bool wT = has_row(ROWS, Q_T) && (qmask & (1u << Q_T));
  const bool wN = has_row(ROWS, Q_N) && (qmask & (1u << Q_N));
  R KQ = 0, KU = 0, KV = 0, RQ = 0, RU = 0, RV = 0, EQ = 0, EU = 0, EV = 0;
  const int nd = a.nd;
  const R s2h = t.s2h[z], c2q = t.c2q[z], qf = t.qf[z], uf = t.uf[z], vf = t.vf[z];
  for (int il = 0; il < L.nline; ++il) {
    const KLine &ln = L.l[il];
    const int i = il * nd + z;
    const R idld = t.idld[i];
    const R x0 = (R)(ln.lam0 - c.lam) * idld - c.lamR * t.vcd[i];
    const R av = t.a[i];
    const R opl = t.opl[i], eml = t.eml[i];
    const R gv = d.gv[i];
    const R cvl = -c.lamR * idld * R(1.0 / kC);  // d x / d vLOS
    if (!ln.split) {
      R H, F, Hx, Fx;
      voigt_d<R>(x0, av, H, F, Hx, Fx);
      KI += opl * H;
      EI += eml * H;
      const R gt = -gv * (H + x0 * Hx - av * Fx), gl = cvl * Hx;
      if (wvt) {
        G[sVT][0] += opl * gt;
        if (FULL) G[sVT][7] += eml * gt;
      }
      if (wvl) {
        G[sVL][0] += opl * gl;
        if (FULL) G[sVL][7] += eml * gl;
      }
      if constexpr (kTh) {
        if (wT) {
          const R gT = th.gT[i], adT = th.adT[i];
          const R gtt = -gT * (H + x0 * Hx - av * Fx) - adT * Fx;
          G[sT][0] += opl * gtt + th.pT[i] * H;
          if (FULL) G[sT][7] += eml * gtt + th.eT[i] * H;
        }
        if (wN) {
          const R gnn = -th.aN[i] * Fx;
          G[sN][0] += opl * gnn + th.pN[i] * H;
          if (FULL) G[sN][7] += eml * gnn + th.eN[i] * H;
        }
      }
      continue;
    }
    const R lbd = t.lbd[i];
    const R cB = (R)ln.lbfac * idld;  // d x / d|B| per unit shift
    R HS[3] = {0, 0, 0}, FS[3] = {0, 0, 0}, HX[3] = {0, 0, 0}, FX[3] = {0, 0, 0}, HXS[3] = {0, 0, 0},
      FXS[3] = {0, 0, 0};
    int cidx = ln.c0;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      for (int m = 0; m < ln.nc[g]; ++m, ++cidx) {
        const R sh = (R)L.shift[cidx], sw = (R)L.str[cidx];
        R H, F, Hx, Fx;
        voigt_d<R>(x0 + sh * lbd, av, H, F, Hx, Fx);
        HS[g] += H * sw;
        FS[g] += F * sw;
        HX[g] += Hx * sw;
        FX[g] += Fx * sw;
        HXS[g] += Hx * sw * sh;
        FXS[g] += Fx * sw * sh;
      }
    }
    // pattern geometry (atom_rad_synthesis.cc:486-497) of a profile triplet
    auto addf = [&](const R h[3], const R f[3], R *Kq, bool full, R fo, R fe) {
      const R hbr = h[2] + h[1], fbr = f[2] + f[1];
      const R eI = s2h * h[0] + c2q * hbr;
      Kq[0] += eI * fo;
      if (!full) return;
      const R hl = h[0] - R(0.5) * hbr, fl = f[0] - R(0.5) * fbr;
      const R eQ = qf * hl, eU = uf * hl, eV = vf * (h[2] - h[1]);
      Kq[1] += eQ * fo;
      Kq[2] += eU * fo;
      Kq[3] += eV * fo;
      Kq[4] += qf * fl * fo;
      Kq[5] += uf * fl * fo;
      Kq[6] += vf * (f[2] - f[1]) * fo;
      Kq[7] += eI * fe;
      Kq[8] += eQ * fe;
      Kq[9] += eU * fe;
      Kq[10] += eV * fe;
    };
    auto add = [&](const R h[3], const R f[3], R *Kq, bool full) { addf(h, f, Kq, full, opl, eml); };
    {
      R kk[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
      add(HS, FS, kk, true);
      KI += kk[0]; KQ += kk[1]; KU += kk[2]; KV += kk[3];
      RQ += kk[4]; RU += kk[5]; RV += kk[6];
      EI += kk[7]; EQ += kk[8]; EU += kk[9]; EV += kk[10];
    }
    if (wvt) {
      R h[3], f[3];
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        h[g] = -gv * (HS[g] + x0 * HX[g] + lbd * HXS[g] - av * FX[g]);
        f[g] = -gv * (FS[g] + x0 * FX[g] + lbd * FXS[g] + av * HX[g]);
      }
      add(h, f, G[sVT], FULL);
    }
    if (wvl) {
      R h[3], f[3];
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        h[g] = cvl * HX[g];
        f[g] = cvl * FX[g];
      }
      add(h, f, G[sVL], FULL);
    }
    if (wb) {
      R h[3], f[3];
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        h[g] = cB * HXS[g];
        f[g] = cB * FXS[g];
      }
      add(h, f, G[sB], FULL);
    }
    if constexpr (kTh) {
      if (wT) {
        const R gT = th.gT[i], adT = th.adT[i];
        R h[3], f[3];
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          h[g] = -gT * (HS[g] + x0 * HX[g] + lbd * HXS[g] - av * FX[g]) - adT * FX[g];
          f[g] = -gT * (FS[g] + x0 * FX[g] + lbd * FXS[g] + av * HX[g]) + adT * HX[g];
        }
        add(h, f, G[sT], FULL);
        addf(HS, FS, G[sT], FULL, th.pT[i], th.eT[i]);
      }
      if (wN) {
        const R aN = th.aN[i];
        R h[3], f[3];
#pragma unroll
        for (int g = 0; g < 3; ++g) {
          h[g] = -aN * FX[g];
          f[g] = aN * HX[g];
        }
        add(h, f, G[sN], FULL);
        addf(HS, FS, G[sN], FULL, th.pN[i], th.eN[i]);
      }
    }
    const R hl = HS[0] - R(0.5) * (HS[2] + HS[1]), fl = FS[0] - R(0.5) * (FS[2] + FS[1]);
    if (wth) {
      // d/d inclination; rho_V as the reference writes it (cos, :999)
      const R sct = d.sct[z], dqt = d.dqt[z], dut = d.dut[z], mhs = d.mhs[z];
      const R eI = sct * hl;
      G[sTH][0] += eI * opl;
      if (FULL) {
        const R eQ = dqt * hl, eU = dut * hl, eV = mhs * (HS[2] - HS[1]);
        G[sTH][1] += eQ * opl;
        G[sTH][2] += eU * opl;
        G[sTH][3] += eV * opl;
        G[sTH][4] += dqt * fl * opl;
        G[sTH][5] += dut * fl * opl;
        G[sTH][6] += vf * (FS[2] - FS[1]) * opl;
        G[sTH][7] += eI * eml;
        G[sTH][8] += eQ * eml;
        G[sTH][9] += eU * eml;
        G[sTH][10] += eV * eml;
      }
    }
    if (wph && FULL) {
      // d/d azimuth: d qf = -2 uf, d uf = 2 qf
      G[sPH][1] += R(-2) * uf * hl * opl;
      G[sPH][2] += R(2) * qf * hl * opl;
      G[sPH][4] += R(-2) * uf * fl * opl;
      G[sPH][5] += R(2) * qf * fl * opl;
      G[sPH][8] += R(-2) * uf * hl * eml;
      G[sPH][9] += R(2) * qf * hl * eml;
    }
  }
  K[0] = KI; K[1] = KQ; K[2] = KU; K[3] = KV; K[4] = RQ; K[5] = RU; K[6] = RV;
  E[0] = EI; E[1] = EQ; E[2] = EU; E[3] = EV;
  if constexpr (kTh) {
    // the background's T and Nt terms, then the referent-opacity
    // normalisation: d(K / r) = dK / r - (K / r) d ln r
    if (wT) {
      G[sT][0] += th.bKT[z];
      if (FULL) G[sT][7] += th.bET[z];
      const R rt = th.rT[z];
      G[sT][0] -= rt * KI;
      if (FULL) {
#pragma unroll
        for (int k = 1; k < 7; ++k) G[sT][k] -= rt * K[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) G[sT][7 + k] -= rt * E[k];
      }
    }
    if (wN) {
      G[sN][0] += th.bKN[z];
      if (FULL) G[sN][7] += th.bEN[z];
      const R rn = th.rN[z];
      G[sN][0] -= rn * KI;
      if (FULL) {
#pragma unroll
        for (int k = 1; k < 7; ++k) G[sN][k] -= rn * K[k];
#pragma unroll
        for (int k = 0; k < 4; ++k) G[sN][7 + k] -= rn * E[k];
      }
    }
  }
}

}  // namespace snapi
