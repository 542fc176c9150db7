// Per-column staging and the per-wavelength opacity + formal-solution
// recurrence, shared by the synthesis kernel and the response/adjoint
// kernels.  See synth.hip for the reference functions restated.
#pragma once
#include <hip/hip_runtime.h>

#include "rt_common.h"

namespace snapi {

constexpr int kSynthThreads = 512;
constexpr double kLog2_10 = 3.321928094887362347870319429489390175864831393;

struct SynthArgs {
  int ncol, nd, nlambda, nfield;
  double ct;       // cos(theta) of the ray
  double lam_mid;  // H- wavelength when hminus_mid is set
  int tau_grid, hminus, hminus_mid;
};

// LDS tile of one column: per-depth quantities that do not depend on the
// wavelength, computed once per column (in double) and shared by every
// wavelength of the block.
template <typename R> struct ColumnTile {
  // per depth
  R *grid, *bgop, *bgem, *cbf, *cff, *ltp, *tinv, *s2h, *c2q, *qf, *uf, *vf;
  // per line and depth: [line * nd + z]
  R *idld, *vcd, *lbd, *a, *opl, *eml;
  static constexpr int kPerDepth = 12;
  static constexpr int kPerLine = 6;
  static __host__ __device__ size_t bytes(int nd, int nline) {
    return sizeof(R) * (size_t)nd * (kPerDepth + kPerLine * nline);
  }
  static __device__ ColumnTile carve(R *p, int nd, int nline) {
    ColumnTile t;
    R **f[kPerDepth] = {&t.grid, &t.bgop, &t.bgem, &t.cbf, &t.cff, &t.ltp,
                        &t.tinv, &t.s2h,  &t.c2q,  &t.qf,  &t.uf,  &t.vf};
    for (int i = 0; i < kPerDepth; ++i) {
      *f[i] = p;
      p += nd;
    }
    R **g[kPerLine] = {&t.idld, &t.vcd, &t.lbd, &t.a, &t.opl, &t.eml};
    for (int i = 0; i < kPerLine; ++i) {
      *g[i] = p;
      p += (size_t)nd * nline;
    }
    return t;
  }
};

// Per-depth part of the opacity build (reference functions: broad_dop
// atom_rad.cc:879-883, damp_rad atom.cc:926-935, damp_col atom.cc:939-960,
// the line-strength factors of atom_rad_synthesis.cc:432-433, the
// magnetic-field geometry of :416-440, H- depth factors of mol.cc:527-541 and
// 621-627), evaluated in double from the column state.
// kRaw: the line factors opl and eml carry the 1/sqrt(pi) of H and F, for
// the opacity build on the unscaled W4 value (voigt_raw).
template <typename R, typename S, bool kRaw = false>
__device__ void stage_column(const KLines &L, const SynthArgs &a, const S *__restrict__ st, ColumnTile<R> &t) {
  const int nd = a.nd;
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    const double T = (double)st[SNAPI_F_T * nd + z];
    const double ne = (double)st[SNAPI_F_NE * nd + z];
    const double nhi = (double)st[SNAPI_F_NHI * nd + z];
    const double iref = a.tau_grid ? 1.0 / (double)st[SNAPI_F_OPREF * nd + z] : 1.0;
    t.grid[z] = (R)(double)st[SNAPI_F_GRID * nd + z];
    t.bgop[z] = (R)((double)st[SNAPI_F_BGOP * nd + z] * iref);
    t.bgem[z] = (R)((double)st[SNAPI_F_BGEM * nd + z] * iref);
    const double TP = 5040.0 / T;
    t.cbf[z] = (R)(4.158E-28 * ne * kK * T * pow(TP, 2.5) * pow(10.0, 0.754 * TP) * nhi * iref);
    t.cff[z] = (R)(1E-26 * ne * kK * T * nhi * iref);
    t.ltp[z] = (R)log10(TP);
    t.tinv[z] = (R)(1.0 / T);
    const double bm = (double)st[SNAPI_F_BMAG * nd + z];
    double thb = (double)st[SNAPI_F_BINC * nd + z] * kPi / 180.0;
    double phb = (double)st[SNAPI_F_BAZI * nd + z] * kPi / 180.0;
    if (bm < 0.001) thb = phb = 0.0;
    const double sp = sin(2.0 * phb), cp = cos(2.0 * phb), sn = sin(thb), cs = cos(thb);
    t.s2h[z] = (R)(0.5 * sn * sn);
    t.c2q[z] = (R)(0.25 * (1.0 + cs * cs));
    t.qf[z] = (R)(0.5 * sn * sn * cp);
    t.uf[z] = (R)(0.5 * sn * sn * sp);
    t.vf[z] = (R)(0.5 * cs);
  }
  const int nl = L.nline;
  for (int i = threadIdx.x; i < nd * nl; i += blockDim.x) {
    const int il = i / nd, z = i - il * nd;
    const KLine &K = L.l[il];
    const double T = (double)st[SNAPI_F_T * nd + z];
    const double vt = (double)st[SNAPI_F_VT * nd + z];
    const double vl = (double)st[SNAPI_F_VLOS * nd + z];
    const double bm = (double)st[SNAPI_F_BMAG * nd + z];
    const double nhi = (double)st[SNAPI_F_NHI * nd + z];
    const double iref = a.tau_grid ? 1.0 / (double)st[SNAPI_F_OPREF * nd + z] : 1.0;
    const double nlo = (double)st[(SNAPI_NF_BASE + 2 * il) * nd + z];
    const double nup = (double)st[(SNAPI_NF_BASE + 2 * il + 1) * nd + z];
    const double dld = sqrt(K.twok_m * T + vt * vt) * K.lam0 / kC;
    const double vm = sqrt(K.vmfac * T);
    const double gc = nhi * vm * pow(vm / 1E6, -K.alpha) * K.sigma;
    const double idld = 1.0 / dld;
    t.idld[i] = (R)idld;
    t.vcd[i] = (R)(vl / kC * idld);
    t.lbd[i] = (R)(K.lbfac * bm * idld);
    t.a[i] = (R)(K.sf * (K.grad + gc) * idld);
    const double ws = kRaw ? 1.0 / kSqrtPi : 1.0;
    t.opl[i] = (R)((nlo * K.opf_l - nup * K.opf_u) * idld * iref * ws);
    t.eml[i] = (R)(nup * K.emf * idld * iref * ws);
  }
}

// Optical-path steps along the ray, index k = position along the ray:
// h_k = (grid_k - grid_{k-1}) / (-cos theta) for k >= 1, its reciprocal, and
// the Fritsch-Butland weight al_k = (1 + h_k / (h_k + h_{k+1})) / 3 at the
// interior points (plpar/atmos_ppbez_rts.cc:961-988).  They depend on the
// column's grid only, so every wavelength shares them.
template <typename R> struct RayGeo {
  R *h, *ih, *al;
  static __host__ __device__ size_t bytes(int nd) { return sizeof(R) * 3 * (size_t)nd; }
  static __device__ RayGeo carve(R *p, int nd) { return {p, p + nd, p + 2 * nd}; }
};
template <typename R, typename S>
__device__ void stage_ray(const SynthArgs &a, const S *__restrict__ st, RayGeo<R> &g) {
  const int nd = a.nd;
  const bool up = a.ct < 0.0;
  const double mct = -a.ct;
  const S *grid = st + SNAPI_F_GRID * nd;
  auto gr = [&](int k) { return (double)grid[up ? nd - 1 - k : k]; };
  for (int k = threadIdx.x; k < nd; k += blockDim.x) {
    if (k == 0) {
      g.h[0] = g.ih[0] = g.al[0] = R(0);
      continue;
    }
    const double h = (gr(k) - gr(k - 1)) / mct;
    g.h[k] = (R)h;
    g.ih[k] = (R)(1.0 / h);
    g.al[k] = R(0);
    if (k <= nd - 2) {
      const double h2 = (gr(k + 1) - gr(k)) / mct;
      g.al[k] = (R)((1.0 + h / (h + h2)) / 3.0);
    }
  }
}

// Wavelength-dependent constants of one work-item.
template <typename R> struct LamConst {
  double lam;           // lambda
  R lamR;
  R hbf, f0, f1, f2;    // H- bound-free alpha(lambda), free-free exponent coefficients
  R hckl;               // h c / (k lambda)
  R plc, ple;           // Planck: 2 h c^2 / lambda^5 and hc/(k lambda) as the reference writes them
};

template <typename R>
__device__ __forceinline__ void lambda_consts(const KLines &L, const SynthArgs &a, double lam, LamConst<R> &c) {
  c.lam = lam;
  c.lamR = (R)lam;
  // the response path of the reference evaluates H- once, at the middle of
  // the wavelength range (atmos/atmol/mol/mol.cc:420-423)
  const double lh = a.hminus_mid ? a.lam_mid : lam;
  const double la = lh * 1E8;
  // H- bound-free (mol.cc:612-617) and free-free (mol.cc:524-527), Gray
  const double ac[7] = {1.99654, -1.18267E-5, 2.64243E-6, -4.40524E-10, 3.23992E-14, -1.39568E-18, 2.78701E-23};
  double alpha = 0.0, p = 1.0;
  for (int i = 0; i < 7; ++i) {
    alpha += ac[i] * p;
    p *= la;
  }
  const double l = log10(la);
  c.hbf = (R)alpha;
  c.f0 = (R)(-2.2763 - 1.6850 * l + 0.76661 * l * l - 0.053346 * l * l * l);
  c.f1 = (R)(15.2827 - 9.2846 * l + 1.99381 * l * l - 0.142631 * l * l * l);
  c.f2 = (R)(-197.789 + 190.266 * l - 67.9775 * l * l + 10.6913 * l * l * l - 0.625151 * l * l * l * l);
  c.hckl = (R)(kH * kC / lh / kK);
  c.plc = (R)(2.0 * 6.62606957 * 8.9875517E-7 / lh / lh / lh / lh / lh);
  c.ple = (R)(1.4387769 / lh);
}

// Opacity (7 independent matrix elements) and emissivity of one depth point,
// normalised by the referent opacity when the tau grid is used.
// K = {eta_I, eta_Q, eta_U, eta_V, rho_Q, rho_U, rho_V}.
// kRaw: the tile was staged with kRaw (1/sqrt(pi) in opl and eml) and the
// unscaled W4 value is accumulated.
template <typename R, bool kRaw = false>
__device__ __forceinline__ void opacity_at(const KLines &L, const SynthArgs &a, const ColumnTile<R> &t,
                                           const LamConst<R> &c, int z, R K[7], R E[4]) {
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
  R KQ = 0, KU = 0, KV = 0, RQ = 0, RU = 0, RV = 0, EQ = 0, EU = 0, EV = 0;
  const int nd = a.nd;
  for (int il = 0; il < L.nline; ++il) {
    const KLine &ln = L.l[il];
    const int i = il * nd + z;
    // (lam0 - lambda) formed in double: in float the difference of two
    // nearly equal wavelengths would lose the Doppler-width scale
    const R x0 = (R)(ln.lam0 - c.lam) * t.idld[i] - c.lamR * t.vcd[i];
    const R av = t.a[i];
    const R opl = t.opl[i], eml = t.eml[i];
    if (!ln.split) {
      R H, F;
      if (kRaw) voigt_raw<R>(x0, av, H, F);
      else voigt<R>(x0, av, H, F);
      KI += opl * H;
      EI += eml * H;
      continue;
    }
    const R lbd = t.lbd[i];
    R Hs[3] = {0, 0, 0}, Fs[3] = {0, 0, 0};
    int cidx = ln.c0;
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      for (int m = 0; m < ln.nc[g]; ++m, ++cidx) {
        R H, F;
        if (kRaw) voigt_raw<R>(x0 + (R)L.shift[cidx] * lbd, av, H, F);
        else voigt<R>(x0 + (R)L.shift[cidx] * lbd, av, H, F);
        const R sw = (R)L.str[cidx];
        Hs[g] += H * sw;
        Fs[g] += F * sw;
      }
    }
    const R hbr = Hs[2] + Hs[1], fbr = Fs[2] + Fs[1];
    const R eI = t.s2h[z] * Hs[0] + t.c2q[z] * hbr;
    const R hl = Hs[0] - R(0.5) * hbr, fl = Fs[0] - R(0.5) * fbr;
    const R eQ = t.qf[z] * hl, eU = t.uf[z] * hl, eV = t.vf[z] * (Hs[2] - Hs[1]);
    const R rQ = t.qf[z] * fl, rU = t.uf[z] * fl, rV = t.vf[z] * (Fs[2] - Fs[1]);
    KI += eI * opl;
    KQ += eQ * opl;
    KU += eU * opl;
    KV += eV * opl;
    RQ += rQ * opl;
    RU += rU * opl;
    RV += rV * opl;
    EI += eI * eml;
    EQ += eQ * eml;
    EU += eU * eml;
    EV += eV * eml;
  }
  K[0] = KI; K[1] = KQ; K[2] = KU; K[3] = KV; K[4] = RQ; K[5] = RU; K[6] = RV;
  E[0] = EI; E[1] = EQ; E[2] = EU; E[3] = EV;
}

template <typename R> __device__ __forceinline__ void reduce_ks(const R K[7], const R E[4], Kred<R> &k, R S[4]) {
  const R i = rrcp<R>(K[0]);
  k.a = K[1] * i; k.b = K[2] * i; k.c = K[3] * i;
  k.d = K[6] * i; k.e = K[5] * i; k.f = K[4] * i;
  S[0] = E[0] * i; S[1] = E[1] * i; S[2] = E[2] * i; S[3] = E[3] * i;
}

// One wavelength: opacity build streamed along the ray with one point of
// look-ahead (the Fritsch-Butland derivative at k needs eta_I at k+1), and
// the linear DELO-type recurrence
//   (1 + w1 K~_k) I_k = e^{-dtau} I_{k-1} + w0 (S_{k-1} - K~_{k-1} I_{k-1}) + w1 S_k
// of plpar/atmos_ppbez_rts.cc:1055-1197.  X = S_{k-1} - K~_{k-1} I_{k-1} is
// formed as soon as I_{k-1} is known, so that the opacity build of the next
// point holds 4 values of the previous point instead of 14.  The last point
// along the ray is the copy of the one before it (:1201-1205), so the
// emergent vector is I_{nd-2}.  The tile is staged with kRaw.  Returns a
// snapi_status bit pattern (0 = fine).
template <typename R>
__device__ int solve_lambda(const KLines &L, const SynthArgs &a, const ColumnTile<R> &t, const RayGeo<R> &g,
                            double lam, R out[4]) {
  LamConst<R> c;
  lambda_consts<R>(L, a, lam, c);
  const int nd = a.nd;
  const bool up = a.ct < 0.0;  // outgoing ray: from the bottom (z = nd-1) to the top
  auto zof = [&](int k) { return up ? nd - 1 - k : k; };
  Kred<R> kc;
  R Sc[4], Ip[4], X[4], oom, oo0;
  {
    R K0[7], E0[4], Sp[4], kv[4];
    Kred<R> kp;
    opacity_at<R, true>(L, a, t, c, zof(0), K0, E0);
    oom = K0[0];
    Ip[0] = up ? E0[0] / K0[0] : R(0);
    Ip[1] = Ip[2] = Ip[3] = R(0);
    reduce_ks<R>(K0, E0, kp, Sp);
    kmul<R>(kp, Ip, kv);
#pragma unroll
    for (int s = 0; s < 4; ++s) X[s] = Sp[s] - kv[s];
  }
  {
    R K1[7], E1[4];
    opacity_at<R, true>(L, a, t, c, zof(1), K1, E1);
    oo0 = K1[0];
    reduce_ks<R>(K1, E1, kc, Sc);
  }
  R h0 = g.h[1];
  R Dm = (oo0 - oom) * g.ih[1];  // one-sided at the first point
  int rc = 0;
  for (int k = 1; k <= nd - 2; ++k) {
    R Kn[7], En[4];
    opacity_at<R, true>(L, a, t, c, zof(k + 1), Kn, En);
    const R oop = Kn[0];
    const R d1 = (oo0 - oom) * g.ih[k], d2 = (oop - oo0) * g.ih[k + 1];
    const R al = g.al[k];
    const R D0 = d1 * d2 > R(0) ? d1 * d2 * rrcp<R>(al * d1 + (R(1) - al) * d2) : R(0);
    const R C = R(0.5) * ((oo0 - h0 * R(0.5) * D0) + (oom + h0 * R(0.5) * Dm));
    const R tb = h0 * R(1.0 / 3.0) * (oom + oo0 + C);
    if (!(tb < R(1E10))) rc |= 1;
    R etb, w0, w1;
    sc_weights<R>(tb, etb, w0, w1);
    R B[4], In[4], kv[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) B[s] = w0 * X[s] + w1 * Sc[s] + Ip[s] * etb;
    solve_1pk<R>(kc, w1, B, In);
    kmul<R>(kc, In, kv);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      Ip[s] = In[s];
      X[s] = Sc[s] - kv[s];
    }
    reduce_ks<R>(Kn, En, kc, Sc);
    oom = oo0;
    oo0 = oop;
    Dm = D0;
    h0 = g.h[k + 1];
  }
  out[0] = Ip[0];
  out[1] = Ip[1];
  out[2] = Ip[2];
  out[3] = Ip[3];
  if (!(fabs((double)Ip[0]) < 1e300) || Ip[0] != Ip[0]) rc |= 2;
  return rc;
}

// Host side: convert the public line table into kernel constants and check
// the sizes the kernels support.
int prepare_synth(const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd, int nlambda, KLines &kl,
                  SynthArgs &a);
// Synthesis restricted to the columns with active[col / active_div] set.
int synth_dev_masked(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd,
                     const void *state_dev, int nlambda, const double *lam_dev, void *stokes_dev, const int *active,
                     int active_div, hipStream_t stream);
// The int the kernels OR their status bits into: the calling thread's bound
// context (snapi_context_bind) on the current device, else the device's own.
int *device_status_slot();
int reset_status();
int read_status();

}  // namespace snapi
