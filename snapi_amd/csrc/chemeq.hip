// Chemical equilibrium and LTE level populations on gfx950, in double.
//
// Restates atmosphere::ltepops (atmos/atmos_pop.cc:30-45) for every depth
// point: atmosphere::chemeq (atmos/atmos_chm.cc:28-363) -- a starting
// electron density from ionisation alone, averaging to 1 % agreement, then
// Newton steps on (Ne, N_1..N_n) with ionisation fractions from
// atom::ionfrc (atmos/atmol/atom/atom.cc:892-924), partition functions
// (atmos/atmol/atom/partf.cc:14-358) and molecules in dissociation
// equilibrium (atmos/atmol/mol/moleqc.cc) -- then atom::lte
// (atmos/atmol/atom/atom.cc:779-787) for the requested levels.
//
// Layout: a team of 16 lanes (a quarter wavefront) owns one depth point and
// keeps the whole Newton system in registers.  Lane j <= n owns unknown j
// (0: electrons, a: element a) and column j of the Jacobian; lane a also
// evaluates element a's partition functions and ionisation fractions; lane
// n+1 holds the right-hand side; molecule m is evaluated on lane m mod 16.
// Sums over elements and molecules are gathered with team shuffles in the
// reference's order, and the system is solved by Gaussian elimination with
// partial pivoting across the team's columns (the reference's SVD solve
// gives the same solution for every system it accepts; it exits on a
// singular one, here the point reports -1 steps).  The temperature-only
// parts of the partition functions and Saha factors are evaluated once per
// point.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

#include "rt_common.h"
#include "chem_device.h"

namespace {

constexpr int kTeam = 16;
constexpr int kChemThreads = 256;
constexpr int kMaxAverage = 1000;  // passes of the starting-value averaging
constexpr int kMaxSched = 12;      // Traving stage evaluations per lane
constexpr int kMaxS = SNAPI_CHEM_MAX_STAGES;
constexpr int kMaxE = SNAPI_CHEM_MAX_ELEMENTS;
constexpr int kMaxM = SNAPI_CHEM_MAX_MOLECULES;
constexpr int kMaxC = SNAPI_CHEM_MAX_COMPONENTS;
static_assert(kMaxE + 2 <= kTeam, "unknowns and right-hand side must fit a team");
static_assert(kMaxM <= 2 * kTeam, "two molecule rounds per team");

// CGS constants as cmlib/const.h states them.
constexpr double kPi = 3.141592653589793238462643383279502884197169399375105;
constexpr double kH = 6.6260687652E-27;
constexpr double kK = 1.3806503E-16;
constexpr double kMe = 9.1093818872E-28;
constexpr double kE = 4.8032042510E-10;
constexpr double kRyd = 1.097373156853955E5;
constexpr double kE0 = 7.957747154594775607E-02;
constexpr double kC = 2.99792458E10;
constexpr double kSaha = 2.0706615382E-16;  // h^3 / (2 pi me k)^1.5 / 2

// Table-derived constants, computed on the host.
// Stage descriptor of a Traving partition function: its asymptotic terms
// are nterm consecutive entries from toff of the element's term list.
struct StageDesc {
  int nterm;     // -1: the whole U is temperature-only (NONE, CONST, IRWIN)
  int toff;
  double zf;     // effective charge times the highest-level constant
  double rzf;    // its reciprocal
  double z2;     // 31.321 times the effective charge squared
};

struct ChemDerived {
  int nst;                            // ionisation stages over all elements
  int soff[kMaxE];                    // first stage of each element in the stage output
  int roff[kMaxE];                    // element's slots in a point's LDS region
  int uoff[kMaxE];                    // U, 1/U and dlnU/dne tables within them
  int nslot;                          // doubles of a point's LDS region
  int nsched[kTeam];                  // Traving stages evaluated by each lane of a team
  int sched[kTeam][kMaxSched];        // element * kMaxS + stage
  double travf;                       // highest-level constant of trav::U (partf.cc:222)
  double hm[4];                       // H- at lambda_ref: Gray's alpha and the free-free f_0, f_1, f_2
  StageDesc sd[kMaxE][kMaxS];
  double xl[kMaxE][kMaxS * SNAPI_CHEM_MAX_TERMS];  // first quantum number of each term
  double rxl[kMaxE][kMaxS * SNAPI_CHEM_MAX_TERMS]; // its reciprocal
  double q[kMaxM];                    // molecular charge: sum of component charge indices
  double mul[kMaxM][kTeam];           // multiplicity of unknown j (0: electrons) in molecule m
  int celem[kMaxM][kMaxC];            // component unknown index
  int cstage[kMaxM][kMaxC];           // component stage clamped at 0
};

__device__ __forceinline__ double tshfl(double v, int src) { return __shfl(v, src, kTeam); }
__device__ __forceinline__ int tshfl(int v, int src) { return __shfl(v, src, kTeam); }

// Fischel & Sparks (1971) simplified eqs. (26)/(27) and the derivative in p
// (trav::qas_f, trav::dqas_f, partf.cc:261-288), with the divisions by p and
// l taken as products with their reciprocals rp and rl.
__device__ __forceinline__ double qas_f(double l, double rl, double p, double rp, double dz) {
  if (p > l) {
    const double llm = l - 1.0;
    return ((1.3333333 * p + 0.5) * p + 0.16666667 + 1.33333333 * dz) * p - 0.4 * dz * dz * rp -
           ((0.33333333 * llm - 0.5) * llm - 0.16666667 - dz) * llm + 0.5 * dz * dz * rl;
  }
  const double axl2 = dz * (rl * rl);
  return (1.0 + axl2 * (0.33333333 + 0.1 * axl2)) * p * p * p * p * rl;
}

__device__ __forceinline__ double dqas_f(double l, double rl, double p, double rp, double dz) {
  if (p > l) return ((3.0 * 1.3333333 * p + 1.0) * p + 0.16666667 + 1.33333333 * dz) + 0.4 * dz * dz * (rp * rp);
  const double axl2 = dz * (rl * rl);
  return 4.0 * (1.0 + axl2 * (0.33333333 + 0.1 * axl2)) * p * p * p * rl;
}

// Temperature-only part of one stage's partition function: the whole U for
// NONE (1), CONST and IRWIN; g0 plus the explicit levels for TRAV.
__device__ double pf_base(const snapi_partition &pf, double T, double x) {
  if (pf.type == SNAPI_PF_CONST) return pf.value;
  if (pf.type == SNAPI_PF_IRWIN) {
    const double lnT = log(T);
    double s = 0.0;
    for (int i = 0; i < pf.nterm; ++i) s += pf.irwin[i] * pow(lnT, (double)i);
    return exp(s);
  }
  if (pf.type != SNAPI_PF_TRAV) return 1.0;
  double sum = pf.g0;
  for (int i = 0; i < pf.nterm; ++i)
    for (int l = 0; l < pf.ncoef[i]; ++l) sum += pf.alpha[i][l] * exp(-2.302585 * pf.gamma[i][l] * x);
  return sum;
}

// Electron-density powers and reciprocals shared by every stage of one
// evaluation: ne^(1/6) enters the highest bound level p = z f / ne^(1/6),
// ne^(7/6) the derivative of trav::U.
struct NePow {
  double ne, rne, r6, rr6, k76;  // ne, 1/ne, ne^(1/6), ne^(-1/6), 1/(6 ne^(7/6))
};

__device__ __forceinline__ NePow ne_powers(double ne) {
  NePow w;
  w.ne = ne;
  w.rne = 1.0 / ne;
  w.r6 = cbrt(sqrt(ne));
  w.rr6 = 1.0 / w.r6;
  w.k76 = w.rne * w.rr6 * (1.0 / 6.0);
  return w;
}

// U (and with deriv dU/dne) of one stage from its temperature-only part
// `base` and, for a Traving stage, the temperature-only weights c of its
// asymptotic terms; trav::dU keeps only the last term (partf.cc:236-250).
template <bool kDeriv>
__device__ __forceinline__ void pf_eval(const StageDesc &d, const double *xl, const double *rxl, double base,
                                        const double *c, double x, const NePow &w, double &U, double &dU) {
  U = base;
  dU = 0.0;
  const int n = d.nterm;
  if (n < 0) return;
  const double p = d.zf * w.rr6, rp = w.r6 * d.rzf;
  const double alpha = d.z2 * x;
  for (int i = 0; i < n; ++i) {
    const double l = xl[d.toff + i], rl = rxl[d.toff + i], ci = c[d.toff + i];
    U += ci * qas_f(l, rl, p, rp, alpha);
    if (kDeriv && i == n - 1) dU = -ci * dqas_f(l, rl, p, rp, alpha) * d.zf * w.k76;
  }
}

// Raw sums of atom::ionfrc (atom.cc:892-924) over one element's stages:
// with the unnormalised fractions fk_z (fk_0 = 1) and derivatives df_z,
// fsum = sum fk, dfsum = sum df, ez = sum z fk, dz = sum z df, and fk_1,
// df_1.  The normalised values follow as fk/fsum and (df - dfsum fk/fsum)/fsum.
struct IonSums {
  double fsum, dfsum, ez, dz, f1, d1;
};

// Makes one lane's LDS writes visible to the other lanes of its wavefront.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The lanes of a team evaluate the Traving stages of all elements (the
// schedule spreads their terms evenly over the 16 lanes) into the U tables.
template <bool kDeriv>
__device__ __forceinline__ void eval_stages(const ChemDerived *dv, const StageDesc (*sd)[kMaxS],
                                            const double (*xl)[kMaxS * SNAPI_CHEM_MAX_TERMS],
                                            const double (*rxl)[kMaxS * SNAPI_CHEM_MAX_TERMS], const int *nsched,
                                            const int (*sched)[kMaxSched], const int *ns_of, int lane, double *team,
                                            double x, const NePow &w) {
  const int n = nsched[lane];
  for (int k = 0; k < n; ++k) {
    const int a = sched[lane][k] / kMaxS, z = sched[lane][k] % kMaxS, ns = ns_of[a];
    double *Ra = team + dv->roff[a];
    double U, dU;
    pf_eval<kDeriv>(sd[a][z], xl[a], rxl[a], Ra[z], Ra + 2 * ns - 1, x, w, U, dU);
    double *Ut = Ra + dv->uoff[a];
    const double rU = 1.0 / U;
    Ut[z] = U;
    Ut[ns + z] = rU;
    if (kDeriv) Ut[2 * ns + z] = dU * rU;
  }
}

// atom::ionfrc (atom.cc:892-924) from the U tables: the unnormalised
// fractions' recursion over the stages.  Without kDeriv only fsum, ez and
// f1 are formed; with store, fk_{z+1} is kept in frac[z].
template <bool kDeriv>
__device__ IonSums ionfrc(int ns, double *R, const double *Ut, double *frac, double fct, const NePow &w, bool store) {
  double *sx = R + ns;
  double rUp = Ut[ns], dUp = kDeriv ? Ut[2 * ns] : 0.0;
  IonSums s{1.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double fk = 1.0, dfv = 0.0;
  for (int z = 0; z < ns - 1; ++z) {
    const double Un = Ut[z + 1];
    const double fz = fct * (Un * rUp) * sx[z] * w.rne;
    if (kDeriv) {
      const double dUn = Ut[2 * ns + z + 1];
      const double fd = fz * ((dUn) - (dUp * w.ne + 1.0) * w.rne);
      dfv = dfv * fz + fk * fd;
      s.dfsum += dfv;
      s.dz += (z + 1) * dfv;
      dUp = dUn;
    }
    fk = fk * fz;
    s.fsum += fk;
    s.ez += (z + 1) * fk;
    if (z == 0) {
      s.f1 = fk;
      s.d1 = dfv;
    }
    if (store) frac[z] = fk;
    rUp = Ut[ns + z + 1];
  }
  return s;
}

// Dissociation constant (eqc_cnst / eqc_poly / eqc_hmin, moleqc.cc).
__device__ double dissoc(const snapi_chem_molecule &m, double T) {
  const double x = 5040.0 / T;
  if (m.eqc == SNAPI_EQC_CONST) return m.coef[0];
  if (m.eqc == SNAPI_EQC_POLY) {
    double s = m.coef[m.ncoef - 1];
    for (int i = m.ncoef - 2; i >= 0; --i) s = m.coef[i] + x * s;
    return exp(2.302585093 * s);
  }
  if (m.eqc == SNAPI_EQC_HMIN) {
    const double s = 0.1249 + (2.5 * log(T) / 2.302585) - 0.747 * x;
    return exp(2.302585093 * s);
  }
  return 0.0;
}

// The wavelength-only parts of the H- opacity (h_minus_mol::opacity_continuum,
// mol.cc:669-695): Gray's bound-free cross-section polynomial alpha(lambda)
// and the coefficients f_0, f_1, f_2 of the free-free fit, lam in cm.
void hminus_lambda_terms(double lam, double hm[4]) {
  const double a[7] = {1.99654, -1.18267E-5, 2.64243E-6, -4.40524E-10, 3.23992E-14, -1.39568E-18, 2.78701E-23};
  double alpha = 0.0;
  for (int i = 0; i < 7; ++i) alpha += a[i] * pow(lam * 1E8, (double)i);
  const double l = log10(lam * 1E8);
  hm[0] = alpha;
  hm[1] = -2.2763 - 1.6850 * l + 0.76661 * l * l - 0.053346 * l * l * l;
  hm[2] = 15.2827 - 9.2846 * l + 1.99381 * l * l - 0.142631 * l * l * l;
  hm[3] = -197.789 + 190.266 * l - 67.9775 * l * l + 10.6913 * l * l * l - 0.625151 * l * l * l * l;
}

// H- bound-free (the cross-section times the Saha ratio, with stimulated
// emission factor stim_1 = 1 - exp(-hc/lambda kT)) and free-free opacity on
// NH neutral hydrogen atoms (mol.cc:669-695), from the wavelength terms hm.
__device__ double hminus_opacity(double T, double ne, double NH, const double *hm, double stim_1) {
  const double T_P = 5040. / T;
  const double op_bf = 4.158E-28 * hm[0] * ne * kK * T * pow(T_P, 2.5) * pow(10.0, 0.754 * T_P) * NH * stim_1;
  const double lt = log10(T_P);
  const double op_ff = 1E-26 * ne * kK * T * pow(10.0, hm[1] + hm[2] * lt + hm[3] * lt * lt) * NH;
  return op_bf + op_ff;
}

// Phases of one point's computation; every phase starts with the ionisation
// fractions of all elements at the phase's electron density.
enum : int { kAtOne = 0, kAtNt = 1, kAverage = 2, kNewton = 3, kFinal = 4 };

// Table descriptors that every workgroup copies into LDS.
struct ChemShared {
  StageDesc sd[kMaxE][kMaxS];
  double xl[kMaxE][kMaxS * SNAPI_CHEM_MAX_TERMS];
  double rxl[kMaxE][kMaxS * SNAPI_CHEM_MAX_TERMS];
  int nsched[kTeam], sched[kTeam][kMaxSched], ns[kMaxE];
};

__device__ __forceinline__ void load_shared(ChemShared &sh, const snapi_chem_table *__restrict__ tb,
                                            const ChemDerived *__restrict__ dv, int tid) {
  const int na = tb->nelement;
  for (int i = tid; i < na * kMaxS; i += kChemThreads) sh.sd[i / kMaxS][i % kMaxS] = dv->sd[i / kMaxS][i % kMaxS];
  for (int i = tid; i < na * kMaxS * SNAPI_CHEM_MAX_TERMS; i += kChemThreads) {
    const int a = i / (kMaxS * SNAPI_CHEM_MAX_TERMS), t = i % (kMaxS * SNAPI_CHEM_MAX_TERMS);
    sh.xl[a][t] = dv->xl[a][t];
    sh.rxl[a][t] = dv->rxl[a][t];
  }
  if (tid < kTeam) {
    sh.nsched[tid] = dv->nsched[tid];
    for (int k = 0; k < kMaxSched; ++k) sh.sched[tid][k] = dv->sched[tid][k];
  }
  if (tid < na) sh.ns[tid] = tb->element[tid].nstage;
  __syncthreads();
}

// One lane's place in its team: lane 0 the electrons, lanes 1..na the
// elements, lane na+1 the right-hand side of the Newton system; an element
// lane's slots in the team's LDS region: base[ns] (temperature-only part of
// U per stage), sx[ns-1] (Saha factor exp(-chi/kT) per stage), the Traving
// term weights, then from uoff the tables U[ns], 1/U[ns], dlnU/dne[ns] at
// the current electron density and the final unnormalised fractions
// fk_1..fk_(ns-1).
struct ChemTeam {
  int tid, lane, na, nm, ea, ns;
  bool is_el, is_unk;
  double ab, ab1, S;
  double *team, *R, *Ut, *frac;
  const StageDesc *sd;
  const double *xl, *rxl;
};

__device__ __forceinline__ ChemTeam make_team(const snapi_chem_table *__restrict__ tb,
                                              const ChemDerived *__restrict__ dv, const ChemShared &sh,
                                              double *region, int tid) {
  ChemTeam cx;
  cx.tid = tid;
  cx.lane = tid & (kTeam - 1);
  cx.na = tb->nelement;
  cx.nm = tb->nmolecule;
  cx.is_el = cx.lane >= 1 && cx.lane <= cx.na;
  cx.is_unk = cx.lane <= cx.na;
  cx.ea = cx.is_el ? cx.lane - 1 : 0;
  cx.ns = cx.is_el ? tb->element[cx.ea].nstage : 1;
  cx.ab = cx.is_el ? tb->element[cx.ea].abundance : 0.0;
  cx.ab1 = tb->element[0].abundance;
  double S = 0.0;
  for (int a = 0; a < cx.na; ++a) S += tb->element[a].abundance;
  cx.S = S;
  cx.sd = sh.sd[cx.ea];
  cx.xl = sh.xl[cx.ea];
  cx.rxl = sh.rxl[cx.ea];
  cx.team = region + (tid / kTeam) * dv->nslot;
  cx.R = cx.team + dv->roff[cx.ea];
  cx.Ut = cx.R + dv->uoff[cx.ea];
  cx.frac = cx.Ut + 3 * cx.ns;
  return cx;
}

// Temperature-dependent state of a team's point.
struct ChemTemp {
  double T, x, kT, fct, Km[2];
};

// Writes the temperature-only parts of every element's partition functions
// and Saha factors into the team's LDS region; the dissociation constants of
// the lane's molecules (m = lane, lane + 16) stay in registers.
__device__ __forceinline__ ChemTemp set_temperature(const snapi_chem_table *__restrict__ tb, const ChemTeam &cx,
                                                    double T) {
  ChemTemp ct;
  ct.T = T;
  const double x = 5040.0 / T;
  ct.x = x;
  ct.kT = kK * T;
  ct.fct = 2.0 * pow((2.0 * kPi * kMe * kK * T) / (kH * kH), 1.5);
  const int lane = cx.lane, ns = cx.ns, nm = cx.nm;
  const bool is_el = cx.is_el;
  const snapi_chem_element &el = tb->element[cx.ea];
  const StageDesc *sd = cx.sd;
  double *R = cx.R, *Ut = cx.Ut;
  double *Km = ct.Km;
  if (is_el) {
    // temperature-only parts: explicit levels of every stage, Saha factors,
    // and the weights of the Traving asymptotic terms
    for (int z = 0; z < ns; ++z) {
      const snapi_partition &pf = el.pf[z];
      R[z] = pf_base(pf, T, x);
      if (z < ns - 1) R[ns + z] = exp(-el.eion[z] / (kK * T));
      if (pf.type == SNAPI_PF_TRAV) {
        for (int i = 0; i < pf.nterm; ++i) R[2 * ns - 1 + sd[z].toff + i] = pf.g2[i] * exp(-2.302585 * pf.xion[i] * x);
      } else {
        // U of the other stages depends on T alone
        Ut[z] = R[z];
        Ut[ns + z] = 1.0 / R[z];
        Ut[2 * ns + z] = 0.0;
      }
    }
  }
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int m = r * kTeam + lane;
    Km[r] = (m < nm) ? dissoc(tb->molecule[m], T) : 1.0;
  }
  return ct;
}

// Result of one point's equilibrium.
struct ChemSol {
  double N0, Nown, rf;  // electron density, the lane's unknown, 1 / sum of its fractions
  IonSums s;
  int steps;
};

// atmosphere::chemeq (atmos_chm.cc:28-363) at one point: starting value from
// ionisation alone, averaging, then Newton steps; the last phase evaluates
// the fractions at the final electron density's predecessor and stores them.
__device__ __forceinline__ ChemSol solve_point(const snapi_chem_table *__restrict__ tb,
                                               const ChemDerived *__restrict__ dv, const ChemShared &sh,
                                               const ChemTeam &cx, const ChemTemp &ct, double Nt) {
  const int tid = cx.tid, lane = cx.lane, na = cx.na, nm = cx.nm, ns = cx.ns;
  const bool is_el = cx.is_el, is_unk = cx.is_unk;
  const double ab = cx.ab, ab1 = cx.ab1, S = cx.S;
  const double x = ct.x, kT = ct.kT, fct = ct.fct;
  const double *Km = ct.Km;
  double *team = cx.team, *R = cx.R, *Ut = cx.Ut, *frac = cx.frac;
  const auto &sd_s = sh.sd;
  const auto &xl_s = sh.xl;
  const auto &rxl_s = sh.rxl;
  const auto &nsched_s = sh.nsched;
  const auto &sched_s = sh.sched;
  const auto &ns_s = sh.ns;
  const int rhs = na + 1;
  int phase = kAtOne, pass = 0, steps = 0, iter = 0;
  double ne_eval = 1.0, F1 = 0.0, N0 = 0.0, Nown = 0.0;
  IonSums s{1.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (;;) {
    // The four teams of a wavefront run their Newton steps together: a team
    // that has its starting value waits while another team of its wavefront
    // is still averaging, so that the Newton code is issued once per step
    // for all of them rather than once per team.
    const bool before = __ballot(phase < kNewton) != 0;
    if (phase == kNewton && before) continue;
    {
      const NePow w = ne_powers(ne_eval);
      if (phase == kNewton) {
        eval_stages<true>(dv, sd_s, xl_s, rxl_s, nsched_s, sched_s, ns_s, lane, team, x, w);
        wave_lds_sync();
        if (is_el) s = ionfrc<true>(ns, R, Ut, frac, fct, w, false);
      } else {
        eval_stages<false>(dv, sd_s, xl_s, rxl_s, nsched_s, sched_s, ns_s, lane, team, x, w);
        wave_lds_sync();
        if (is_el) s = ionfrc<false>(ns, R, Ut, frac, fct, w, phase == kFinal);
      }
      wave_lds_sync();
    }
    if (phase == kFinal) break;
    const double rf = 1.0 / s.fsum;
    if (phase == kAtOne || phase == kAtNt) {
      // starting electron density from ionisation alone (atmos_chm.cc:127-154)
      const double v = ab * (s.ez * rf);
      double F = 0.0;
      for (int a = 1; a <= na; ++a) F += tshfl(v, a);
      if (phase == kAtOne) {
        F1 = F;
        ne_eval = Nt;
        phase = kAtNt;
        continue;
      }
      const double aa = F / (S * Nt);
      double bb = F1 / S - aa;
      const double cc = -bb * Nt;
      bb += 1.0 - aa * Nt;
      const double root = (-bb + sqrt(bb * bb - 4.0 * aa * cc)) / (2.0 * aa);
      N0 = (root > 1.0) ? root : 1.0;
      ne_eval = N0;
      phase = kAverage;
      continue;
    }
    if (phase == kAverage) {
      // average until the electron count agrees to 1 % (atmos_chm.cc:157-169);
      // the reference has no limit, a point still outside 1 % after
      // kMaxAverage passes continues from the current estimate
      Nown = is_el ? (Nt - N0) * ab / S : 0.0;
      const double sl = (s.ez * rf) * Nown;
      double esum = 0.0;
      for (int a = 1; a <= na; ++a) esum += tshfl(sl, a);
      N0 = 0.5 * (N0 + esum);
      if (!((1E-2 * (N0 + esum)) < fabs(N0 - esum)) || ++pass >= kMaxAverage) {
        phase = kNewton;
        if (lane == 0) Nown = N0;
      }
      ne_eval = N0;
      continue;
    }

    // one Newton-Raphson step at Ne = N0 (atmos_chm.cc:179-305)
    N0 = ne_eval;
    const double N1 = tshfl(Nown, 1);
    const double eown = s.ez * rf;
    const double tdef = Nown * eown;
    const double uder = Nown * ((s.dz - s.dfsum * s.ez * rf) * rf);
    const double fk0 = is_el ? rf : 1.0, fk1 = is_el ? s.f1 * rf : 0.0;
    const double df0 = is_el ? (0.0 - s.dfsum * rf) * rf : 0.0;
    const double df1 = is_el ? (s.d1 - s.dfsum * s.f1 * rf) * rf : 0.0;
    // molecules: number densities and the electron-density factor of dm[m][0]
    double nmr[2], dmr[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int m = r * kTeam + lane;
      const bool ok = m < nm;
      const int mm = ok ? m : 0;
      double nmv = 1.0 / (Km[r] * kT), d0f = 0.0;
#pragma unroll
      for (int c = 0; c < kMaxC; ++c) {
        const bool use = ok && c < tb->molecule[mm].ncomp;
        const int a = use ? dv->celem[mm][c] : 0;
        const int zc = use ? dv->cstage[mm][c] : 0;
        const double Na = tshfl(Nown, a);
        const double f0 = tshfl(fk0, a), f1 = tshfl(fk1, a), g0 = tshfl(df0, a), g1 = tshfl(df1, a);
        if (use) {
          const double fz = zc ? f1 : f0, gz = zc ? g1 : g0;
          nmv *= fz * Na * kT;
          if (a == 0) d0f += 1.0 / Na;
          d0f += gz / fz;
        }
      }
      nmr[r] = ok ? nmv : 0.0;
      dmr[r] = ok ? nmv * d0f : 0.0;
    }
    // defects
    double d0 = -N0;
    for (int a = 1; a <= na; ++a) d0 += tshfl(tdef, a);
    double d1 = N0 - Nt;
    for (int a = 1; a <= na; ++a) d1 += tshfl(Nown, a);
    double Nr = N1, Nc = Nown;
    for (int m = 0; m < nm; ++m) {
      const double nmv = tshfl(m < kTeam ? nmr[0] : nmr[1], m & (kTeam - 1));
      d0 += dv->q[m] * nmv;
      d1 += nmv;
      Nr += nmv * dv->mul[m][1];
      Nc += nmv * dv->mul[m][lane];
    }
    const double down = (lane == 0) ? d0 : (lane == 1) ? d1 : ab * Nr - ab1 * Nc;
    // Jacobian column of this lane (rows 0..n), or the right-hand side
    double col[kTeam];
    {
      double A00 = -1.0;
      for (int a = 1; a <= na; ++a) A00 += tshfl(uder, a);
      double acc0 = (lane == 0) ? A00 : eown, acc1 = 1.0, dNr = 0.0;
      double dNc[kTeam];
#pragma unroll
      for (int a = 0; a < kTeam; ++a) dNc[a] = 0.0;
      const double inv = 1.0 / Nown;
      for (int m = 0; m < nm; ++m) {
        const double nmv = tshfl(m < kTeam ? nmr[0] : nmr[1], m & (kTeam - 1));
        const double d0m = tshfl(m < kTeam ? dmr[0] : dmr[1], m & (kTeam - 1));
        double dmj;
        if (lane == 0) {
          dmj = d0m;
        } else {
          double sj = 0.0;
          for (int c = 0; c < tb->molecule[m].ncomp; ++c)
            if (dv->celem[m][c] == lane) sj += inv;
          dmj = sj * nmv;
        }
        acc0 += dv->q[m] * dmj;
        acc1 += dmj;
        dNr += dmj * dv->mul[m][1];
#pragma unroll
        for (int a = 2; a < kTeam; ++a) dNc[a] += dmj * dv->mul[m][a];
      }
      if (lane == 1) dNr += 1.0;
      col[0] = acc0;
      col[1] = acc1;
#pragma unroll
      for (int a = 2; a < kTeam; ++a) {
        const double dn = dNc[a] + ((lane == a) ? 1.0 : 0.0);
        const double aba = (a <= na) ? tb->element[(a <= na ? a : 1) - 1].abundance : 0.0;
        col[a] = aba * dNr - ab1 * dn;
      }
      // the right-hand side lane collects the defects
#pragma unroll
      for (int a = 0; a < kTeam; ++a) {
        const double da = tshfl(down, a);
        if (lane == rhs) col[a] = (a <= na) ? da : 0.0;
      }
      if (lane > rhs) {
#pragma unroll
        for (int a = 0; a < kTeam; ++a) col[a] = 0.0;
      }
    }
    // Gaussian elimination with partial pivoting over the team's columns
    bool singular = false;
#pragma unroll
    for (int k = 0; k < kTeam - 1; ++k) {
      if (k <= na) {
        int p = k;
        double best = fabs(col[k]);
#pragma unroll
        for (int r = k + 1; r < kTeam; ++r)
          if (r <= na && fabs(col[r]) > best) {
            best = fabs(col[r]);
            p = r;
          }
        p = tshfl(p, k);
#pragma unroll
        for (int r = k + 1; r < kTeam; ++r)
          if (r == p) {
            const double t = col[k];
            col[k] = col[r];
            col[r] = t;
          }
        const double piv = tshfl(col[k], k);
        if (!(fabs(piv) > 0.0) || !isfinite(piv)) singular = true;
#pragma unroll
        for (int r = k + 1; r < kTeam; ++r) {
          if (r <= na) {
            const double l = tshfl(col[r], k) / piv;
            if (lane > k) col[r] -= l * col[k];
          }
        }
      }
    }
    if (singular) {
      steps = -1;
      phase = kFinal;
      continue;
    }
    // back substitution: v_k = (rhs_k - sum_{j>k} U_kj v_j) / U_kk
    double myv = 0.0;
#pragma unroll
    for (int k = kTeam - 2; k >= 0; --k) {
      if (k <= na) {
        double sk = col[k];  // meaningful on the right-hand side lane
#pragma unroll
        for (int j = k + 1; j < kTeam - 1; ++j)
          if (j <= na) sk -= tshfl(col[k] * myv, j);
        const double ukk = tshfl(col[k], k);
        const double vk = tshfl(sk / ukk, rhs);
        if (lane == k) myv = vk;
      }
    }
    if (is_unk) {
      Nown -= myv;
      if (Nown < 1E-2) Nown = 1E-2;
    }
    ++steps;
    // the reference counts two per step (one inside a log call) and stops above 100
    iter += 2;
    const bool bad = is_unk && (fabs(down / Nown) > 1E-4);
    const unsigned long long b = __ballot(bad);
    const unsigned long long team = 0xFFFFull << (tid & 63 & ~(kTeam - 1));
    if (iter > 100 || !(b & team)) {
      // outputs use the fractions of this step's electron density
      phase = kFinal;
      continue;
    }
    ne_eval = tshfl(Nown, 0);
  }

  ChemSol sol;
  sol.N0 = tshfl(Nown, 0);
  sol.Nown = Nown;
  sol.rf = 1.0 / s.fsum;
  sol.s = s;
  sol.steps = steps;
  return sol;
}

// Continuum opacity at the referent wavelength once a point is solved
// (atmosphere::opacity_continuum, atmos_rad.cc:286-294): Thomson scattering,
// each element's bound-free transitions with stimulated emission
// (atom::opacity_continuum, atom_rad.cc:151-187; its free-free part is
// discarded there), then H- on element 1's neutral stage.  Every lane of the
// team returns the sum.
__device__ __forceinline__ double referent_opacity(const snapi_chem_table *__restrict__ tb,
                                                   const ChemDerived *__restrict__ dv, const ChemTeam &cx,
                                                   const ChemTemp &ct, const ChemSol &sol) {
  const double lam = tb->lambda_ref, T = ct.T, kT = ct.kT, N0 = sol.N0, Nown = sol.Nown, rf = sol.rf;
  const int lane = cx.lane, ns = cx.ns;
  const double *R = cx.R, *frac = cx.frac;
  double part = 0.0;
  const double stim = exp(-kH * kC / lam / kK / T);
  if (cx.is_el) {
    const NePow w = ne_powers(N0);
    const double tp = pow(T, -1.5);
    const snapi_chem_element &el = tb->element[cx.ea];
    for (int j = 0; j < tb->nbf; ++j) {
      const snapi_chem_bf &b = tb->bf[j];
      if (b.element != lane || b.sigma == 0.0) continue;
      const int z = b.stage;
      double Ul, Uu, dU;
      pf_eval<false>(cx.sd[z], cx.xl, cx.rxl, R[z], R + 2 * ns - 1, ct.x, w, Ul, dU);
      pf_eval<false>(cx.sd[z + 1], cx.xl, cx.rxl, R[z + 1], R + 2 * ns - 1, ct.x, w, Uu, dU);
      const double nl = Nown * ((z ? frac[z - 1] : 1.0) * rf) * ((b.g != 0.0) ? b.g * exp(-b.energy / kT) / Ul : 1.0);
      const double nu = Nown * (frac[z] * rf) * ((b.g_up != 0.0) ? b.g_up * exp(-b.energy_up / kT) / Uu : 1.0);
      const double pop_mod = nu * N0 * kSaha * tp * b.g / b.g_up * exp((el.eion[z] - b.energy) / kK / T);
      part += b.sigma * (nl - pop_mod * stim);
    }
  }
  double op = 6.65E-25 * N0;
#pragma unroll
  for (int a = 1; a < kTeam; ++a) {
    const double pa = tshfl(part, a);
    if (a <= cx.na) op += pa;
  }
  const double NH = tshfl(Nown * rf, 1);  // neutral hydrogen: element 1, stage 0
  if (tb->hminus) op += hminus_opacity(T, N0, NH, dv->hm, 1.0 - stim);
  return op;
}

// Mass density from the total density of element 1, 1.4 hydrogen masses per
// hydrogen atom (atmos_pp.cc:404), in the reference's operation order.
__device__ __forceinline__ double mass_density(double NH) { return NH * 1.4 * 1.67E-24; }

__global__ __launch_bounds__(kChemThreads) void chemeq_kernel(const snapi_chem_table *__restrict__ tb,
                                                              const ChemDerived *__restrict__ dv, long npt,
                                                              const double *__restrict__ Tin,
                                                              const double *__restrict__ Ntin, int nd,
                                                              double *__restrict__ ne_out,
                                                              double *__restrict__ stage_out,
                                                              double *__restrict__ mol_out,
                                                              double *__restrict__ lev_out,
                                                              double *__restrict__ opref_out,
                                                              double *__restrict__ rho_out, int *__restrict__ it_out) {
  __shared__ ChemShared sh;
  extern __shared__ double region_s[];
  const int tid = threadIdx.x;
  load_shared(sh, tb, dv, tid);
  const ChemTeam cx = make_team(tb, dv, sh, region_s, tid);
  const long pt = ((long)blockIdx.x * blockDim.x + tid) / kTeam;
  if (pt >= npt) return;  // the whole team leaves together; no barrier follows
  const ChemTemp ct = set_temperature(tb, cx, Tin[pt]);
  const ChemSol sol = solve_point(tb, dv, sh, cx, ct, Ntin[pt]);
  const int lane = cx.lane, ns = cx.ns, nm = cx.nm;
  const bool is_el = cx.is_el;
  const double x = ct.x, kT = ct.kT;
  const double *Km = ct.Km, *R = cx.R, *frac = cx.frac;
  const StageDesc *sd = cx.sd;
  const double *xl = cx.xl, *rxl = cx.rxl;
  // outputs: the final N with the last step's fractions (atom::pupdate, atom.cc:1449-1453)
  const long c = pt / nd, zd = pt % nd;
  const double N0 = sol.N0, Nown = sol.Nown, rf = sol.rf;
  const IonSums &s = sol.s;
  if (lane == 0) {
    if (ne_out) ne_out[pt] = N0;
    if (it_out) it_out[pt] = sol.steps;
  }
  if (stage_out && is_el) {
    const int off = dv->soff[lane - 1], nst = dv->nst;
    for (int z = 0; z < ns; ++z)
      stage_out[((long)c * nst + off + z) * nd + zd] = Nown * ((z ? frac[z - 1] : 1.0) * rf);
  }
  if (mol_out) {
    const double fk0 = is_el ? rf : 1.0, fk1 = is_el ? s.f1 * rf : 0.0;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int m = r * kTeam + lane;
      const bool ok = m < nm;
      const int mm = ok ? m : 0;
      double nmv = 1.0 / (Km[r] * kT);
#pragma unroll
      for (int cc2 = 0; cc2 < kMaxC; ++cc2) {
        const bool use = ok && cc2 < tb->molecule[mm].ncomp;
        const int a = use ? dv->celem[mm][cc2] : 0;
        const int zc = use ? dv->cstage[mm][cc2] : 0;
        const double Na = tshfl(Nown, a), f0 = tshfl(fk0, a), f1 = tshfl(fk1, a);
        if (use) nmv *= (zc ? f1 : f0) * Na * kT;
      }
      if (ok) mol_out[((long)c * nm + m) * nd + zd] = nmv;
    }
  }
  if (lev_out && is_el) {
    const int nl = tb->nlevel;
    const NePow w = ne_powers(N0);
    for (int l = 0; l < nl; ++l) {
      const snapi_chem_level &L = tb->level[l];
      if (L.element != lane) continue;
      const int z = L.stage;
      const double Nz = Nown * ((z ? frac[z - 1] : 1.0) * rf);
      double U, dU;
      pf_eval<false>(sd[z], xl, rxl, R[z], R + 2 * ns - 1, x, w, U, dU);
      lev_out[((long)c * nl + l) * nd + zd] = Nz * ((L.g != 0.0) ? L.g * exp(-L.energy / kT) / U : 1.0);
    }
  }
  if (opref_out) {
    const double op = referent_opacity(tb, dv, cx, ct, sol);
    if (lane == 0) opref_out[pt] = op;
  }
  if (rho_out && lane == 1) rho_out[pt] = mass_density(Nown);
}

constexpr double kGrav = 274.88E2;  // gravitational acceleration at the solar surface [cm/s^2]
constexpr int kHeMaxIter = 20;      // iterations per depth (atmos_pp.cc:385)

// Hydrostatic equilibrium of a batch of columns, one team per column
// (atmos_pp::enforce_hequilibrium, atmos_pp.cc:372-455).  The top point
// starts from 0.3 dyn/cm^2 and iterates Nt = 10^logtau rho g / kappa / kT
// until a correction is below 1e-2 of Nt; each deeper point starts from the
// pressure step rho g / kappa * dtau of the point above and iterates the step
// g / kappa_mean * dtau, kappa_mean the geometric mean of the two points'
// opacity per unit mass, until a correction is below 1e-2 of the TOP point's
// Nt, for at most 20 iterations each.  A point's opacity and density for the
// next step are those of its last chemical equilibrium, evaluated before its
// last correction.  The temperature-only parts are set once per depth and kept
// in LDS through that depth's iterations.  Writes the final Nt and the
// iterations per depth; chemical equilibrium at the final Nt follows in
// chemeq_kernel.
__global__ __launch_bounds__(kChemThreads, 2) void hequilibrium_kernel(const snapi_chem_table *__restrict__ tb,
                                                                    const ChemDerived *__restrict__ dv, long ncol,
                                                                    int nd, const double *__restrict__ Tin,
                                                                    const double *__restrict__ logtau,
                                                                    double *__restrict__ Nt_out,
                                                                    int *__restrict__ it_out,
                                                                    const int *__restrict__ active, int group) {
  __shared__ ChemShared sh;
  extern __shared__ double region_s[];
  const int tid = threadIdx.x;
  load_shared(sh, tb, dv, tid);
  const ChemTeam cx = make_team(tb, dv, sh, region_s, tid);
  const long col = ((long)blockIdx.x * blockDim.x + tid) / kTeam;
  if (col >= ncol) return;  // the whole team leaves together; no barrier follows
  if (active && !active[col / group]) return;
  const double *T = Tin + col * nd, *lt = logtau + col * nd;
  double Ntop = 0.0, Nprev = 0.0, Tprev = 0.0, tprev = 0.0, op_prev = 0.0, rho_prev = 0.0;
  for (int z = 0; z < nd; ++z) {
    // every value below is the same on all lanes of the team
    const double Tz = T[z];
    const ChemTemp ct = set_temperature(tb, cx, Tz);
    const double tau = pow(10.0, lt[z]);
    const double dtau = z ? tau - tprev : 0.0;
    double Nt = z ? (Nprev * kK * Tprev + rho_prev * kGrav / op_prev * dtau) / kK / Tz : 0.3 / kK / Tz;
    double op = 0.0, rho = 0.0;
    int it = 0;
    while (it < kHeMaxIter) {
      const ChemSol sol = solve_point(tb, dv, sh, cx, ct, Nt);
      op = referent_opacity(tb, dv, cx, ct, sol);
      rho = mass_density(tshfl(sol.Nown, 1));
      double dN;
      if (z == 0) {
        dN = (tau * rho * kGrav / op) / kK / Tz - Nt;
      } else {
        const double kappa_mean = sqrt(op * op_prev / rho / rho_prev);
        const double delta_p = kGrav / kappa_mean * dtau;
        dN = (Nprev * kK * Tprev + delta_p) / kK / Tz - Nt;
      }
      Nt += dN;
      ++it;
      if (fabs(dN / (z ? Ntop : Nt)) < 1E-2) break;
    }
    if (z == 0) Ntop = Nt;
    if (cx.lane == 0) {
      Nt_out[col * nd + z] = Nt;
      if (it_out) it_out[col * nd + z] = it;
    }
    Nprev = Nt;
    Tprev = Tz;
    tprev = tau;
    op_prev = op;
    rho_prev = rho;
  }
}

// ---- the per-depth state of the synthesis and response paths ----

constexpr double kDeltaT = 5.0;     // delta_T, cmlib/const.h:27
constexpr double kDeltaNtf = 1E-3;  // delta_Nt_frac, cmlib/const.h:29
constexpr int kEosVariants = 10;    // chemical equilibria per depth point with derivatives
constexpr int kMaxWant = 64;        // stage populations recorded per equilibrium

// Stage populations an equilibrium records: idx[a][z] >= 0 is the slot of
// stage z of element a (0-based) in a variant's record.
struct EosWant {
  int nwant;
  signed char idx[kMaxE][kMaxS];
};

// (T, Nt) of variant v at a point, with the reference's arithmetic order:
// v0 the point itself (popsetup); v1..v4 compute_op_referent_derivative
// (atmos_ppbez_rts.cc:283-325, T +- delta_T/2 then Nt +- delta_Nt/2); v5..v9
// ne_lte_derivatives (atmos_responses.cc:364-423: T + delta_T and T, then at
// T + delta_T/2: Nt +- delta_Nt/2 and Nt).  Also the step sizes.
struct EosSeq {
  double T[kEosVariants], Nt[kEosVariants], dNt0, dNt1;
};

__host__ __device__ inline EosSeq eos_sequence(double T0, double Nt0) {
  EosSeq q;
  double T = T0, Nt = Nt0;
  q.T[0] = T, q.Nt[0] = Nt;
  T += kDeltaT * 0.5;
  q.T[1] = T, q.Nt[1] = Nt;
  T -= kDeltaT;
  q.T[2] = T, q.Nt[2] = Nt;
  T += kDeltaT * 0.5;
  q.dNt0 = Nt * kDeltaNtf;
  Nt += q.dNt0 * 0.5;
  q.T[3] = T, q.Nt[3] = Nt;
  Nt -= q.dNt0;
  q.T[4] = T, q.Nt[4] = Nt;
  Nt += q.dNt0 * 0.5;
  T += kDeltaT;
  q.T[5] = T, q.Nt[5] = Nt;
  T -= kDeltaT;
  q.T[6] = T, q.Nt[6] = Nt;
  T += kDeltaT * 0.5;
  q.dNt1 = kDeltaNtf * Nt;
  Nt += q.dNt1 * 0.5;
  q.T[7] = T, q.Nt[7] = Nt;
  Nt -= q.dNt1;
  q.T[8] = T, q.Nt[8] = Nt;
  Nt += q.dNt1 * 0.5;
  q.T[9] = T, q.Nt[9] = Nt;
  return q;
}

// One team per (depth point, variant): chemical equilibrium at the variant's
// (T, Nt); the record holds Ne, the referent opacity and the wanted stage
// populations.  state [ncol][nf][nd] supplies T and NT.
// With he_rec, the first five variants (the point, T +- 2.5 K, Nt +- half the
// Nt step: the hydrostatic tangent's points, up to the rounding of the
// reference's in-place steps) also give rho / kappa_ref at the referent
// wavelength of the second table tb2 (the hydrostatic march's), in
// he_partials_kernel's record layout [point][5].
__global__ __launch_bounds__(kChemThreads) void eos_variants_kernel(const snapi_chem_table *__restrict__ tb,
                                                                    const ChemDerived *__restrict__ dv, long npt,
                                                                    int nd, int nf, int nvar,
                                                                    const double *__restrict__ state, EosWant want,
                                                                    double *__restrict__ rec,
                                                                    const int *__restrict__ active,
                                                                    const snapi_chem_table *__restrict__ tb2,
                                                                    const ChemDerived *__restrict__ dv2,
                                                                    double *__restrict__ he_rec) {
  __shared__ ChemShared sh;
  extern __shared__ double region_s[];
  const int tid = threadIdx.x;
  load_shared(sh, tb, dv, tid);
  const ChemTeam cx = make_team(tb, dv, sh, region_s, tid);
  const long t = ((long)blockIdx.x * blockDim.x + tid) / kTeam;
  if (t >= npt * nvar) return;  // the whole team leaves together; no barrier follows
  const long pt = t / nvar;
  const int v = (int)(t % nvar);
  const long c = pt / nd, zd = pt % nd;
  if (active && !active[c]) return;
  const double *st = state + c * nf * nd + zd;
  const EosSeq q = eos_sequence(st[SNAPI_F_T * nd], st[SNAPI_F_NT * nd]);
  double Tv = q.T[0], Ntv = q.Nt[0];
#pragma unroll
  for (int k = 1; k < kEosVariants; ++k)  // selects in registers, no indexed stack array
    if (k == v) Tv = q.T[k], Ntv = q.Nt[k];
  const ChemTemp ct = set_temperature(tb, cx, Tv);
  const ChemSol sol = solve_point(tb, dv, sh, cx, ct, Ntv);
  const int nrec = 2 + want.nwant;
  double *r = rec + t * nrec;
  const double op = (v <= 4) ? referent_opacity(tb, dv, cx, ct, sol) : 0.0;
  const bool he = he_rec && v <= 4;  // uniform over the team
  const double op2 = he ? referent_opacity(tb2, dv2, cx, ct, sol) : 1.0;
  const double rho = mass_density(tshfl(sol.Nown, 1));
  if (cx.lane == 0) {
    r[0] = sol.N0;
    r[1] = op;
    if (he) he_rec[pt * 5 + v] = rho / op2;
  }
  if (cx.is_el)
    for (int z = 0; z < cx.ns; ++z) {
      const int k = want.idx[cx.ea][z];
      if (k >= 0) r[2 + k] = sol.Nown * ((z ? cx.frac[z - 1] : 1.0) * sol.rf);
    }
}

// Tangent of the hydrostatic equilibrium.  With P = Nt k T and r = rho /
// kappa the mass per unit opacity, a converged march satisfies
//   P_0 = tau_0 g r_0,   P_z = P_(z-1) + g (tau_z - tau_(z-1)) sqrt(r_z r_(z-1)),
// so a temperature perturbation dT moves Nt by the solution of the
// linearised equations, from the top down.  d ln r / dT and d ln r / dNt at
// each point come from central differences of the chemical equilibrium with
// the reference's steps (T +- delta_T/2, Nt +- delta_Nt_frac Nt / 2).  The
// reference instead differences two hydrostatic rebuilds per temperature
// node (atmos_fit.cc:1225-1248); this needs five equilibria per point for
// all nodes together.

// One team per (point, variant): r at v0 (T, Nt), v1/v2 T +- delta_T/2,
// v3/v4 Nt +- delta_Nt_frac Nt / 2.  rec [npt][5].
__global__ __launch_bounds__(kChemThreads) void he_partials_kernel(const snapi_chem_table *__restrict__ tb,
                                                                   const ChemDerived *__restrict__ dv, long npt,
                                                                   int nd, const double *__restrict__ Tin,
                                                                   const double *__restrict__ Ntin,
                                                                   double *__restrict__ rec,
                                                                   const int *__restrict__ active) {
  __shared__ ChemShared sh;
  extern __shared__ double region_s[];
  const int tid = threadIdx.x;
  load_shared(sh, tb, dv, tid);
  const ChemTeam cx = make_team(tb, dv, sh, region_s, tid);
  const long t = ((long)blockIdx.x * blockDim.x + tid) / kTeam;
  if (t >= npt * 5) return;  // the whole team leaves together; no barrier follows
  const long pt = t / 5;
  const int v = (int)(t % 5);
  if (active && !active[pt / nd]) return;
  const double T = Tin[pt], Nt = Ntin[pt], dNs = kDeltaNtf * Nt;
  const double Tv = v == 1 ? T + 0.5 * kDeltaT : v == 2 ? T - 0.5 * kDeltaT : T;
  const double Nv = v == 3 ? Nt + 0.5 * dNs : v == 4 ? Nt - 0.5 * dNs : Nt;
  const ChemTemp ct = set_temperature(tb, cx, Tv);
  const ChemSol sol = solve_point(tb, dv, sh, cx, ct, Nv);
  const double op = referent_opacity(tb, dv, cx, ct, sol);
  const double rho = mass_density(tshfl(sol.Nown, 1));
  if (cx.lane == 0) rec[t] = rho / op;
}

// One thread per (column, direction): the linearised march.  T, logtau, Nt
// [ncol][nd]; dT, dN [ncol][ndir][nd].
__global__ void he_tangent_kernel(long ncol, int nd, int ndir, const double *__restrict__ T,
                                  const double *__restrict__ logtau, const double *__restrict__ Nt,
                                  const double *__restrict__ rec, const double *__restrict__ dT,
                                  double *__restrict__ dN, const int *__restrict__ active) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol * ndir) return;
  const long c = i / ndir;
  if (active && !active[c]) return;
  const double *Tc = T + c * nd, *lt = logtau + c * nd, *Nc = Nt + c * nd, *rc = rec + c * nd * 5;
  const double *dt = dT + i * nd;
  double *dn = dN + i * nd;
  double r_prev = 0.0, aT_prev = 0.0, aN_prev = 0.0, tau_prev = 0.0;
  for (int z = 0; z < nd; ++z) {
    const double *q = rc + z * 5;
    const double r = q[0], aT = (q[1] - q[2]) / kDeltaT / r, aN = (q[3] - q[4]) / (kDeltaNtf * Nc[z]) / r;
    const double tau = pow(10.0, lt[z]);
    double d;
    if (z == 0) {
      const double Q = tau * kGrav * r;
      d = (Q * aT - kK * Nc[0]) * dt[0] / (kK * Tc[0] - Q * aN);
    } else {
      const double Q = kGrav * (tau - tau_prev) * sqrt(r * r_prev);
      const double rhs = -kK * Nc[z] * dt[z] + kK * (Tc[z - 1] * dn[z - 1] + Nc[z - 1] * dt[z - 1]) +
                         0.5 * Q * (aT * dt[z] + aT_prev * dt[z - 1] + aN_prev * dn[z - 1]);
      d = rhs / (kK * Tc[z] - 0.5 * Q * aN);
    }
    dn[z] = d;
    r_prev = r;
    aT_prev = aT;
    aN_prev = aN;
    tau_prev = tau;
  }
}

// U of stage z of element a (0-based) at (T, ne), as the team evaluation
// forms it from the temperature-only part and the Traving asymptotic terms.
__device__ double pf_direct(const snapi_chem_table *__restrict__ tb, const ChemDerived *__restrict__ dv, int a,
                            int z, double T, double ne) {
  const snapi_partition &pf = tb->element[a].pf[z];
  const double x = 5040.0 / T;
  double U = pf_base(pf, T, x);
  const StageDesc &d = dv->sd[a][z];
  if (d.nterm < 0) return U;
  const NePow w = ne_powers(ne);
  const double p = d.zf * w.rr6, rp = w.r6 * d.rzf, alpha = d.z2 * x;
  for (int i = 0; i < d.nterm; ++i) {
    const double ci = pf.g2[i] * exp(-2.302585 * pf.xion[i] * x);
    U += ci * qas_f(dv->xl[a][d.toff + i], dv->rxl[a][d.toff + i], p, rp, alpha);
  }
  return U;
}

// LTE population of a level (atom::lte, atom.cc:779-787).
__device__ double lte_pop(const snapi_chem_table *__restrict__ tb, const ChemDerived *__restrict__ dv, int a, int z,
                          double g, double energy, double T, double ne, double Nz) {
  return Nz * ((g != 0.0) ? g * exp(-energy / (kK * T)) / pf_direct(tb, dv, a, z, T, ne) : 1.0);
}

// Planck_f (cmlib/mathtools.cc:1150-1152) with the reference's constants.
__device__ double planck_f(double lam, double T) {
  return 2.0 * 6.62606957 * 8.9875517E-7 / lam / lam / lam / lam / lam / (exp(1.4387769 / lam / T) - 1.0);
}

// Thomson scattering at (T, ne) with its emissivity at lambda_mid
// (atmosphere::op_em_vector, atmos_rad_synthesis.cc:20-50) plus the map's
// bound-free background at lambda_bg (atom::boundfree_op_em_vector,
// atom_rad_synthesis.cc:226-262) with level populations from the stage
// populations stp (record slots) at (Tl, nel).
__device__ void eos_background(const snapi_chem_table *__restrict__ tb, const ChemDerived *__restrict__ dv,
                               const snapi_state_map *__restrict__ map, const EosWant &want, double T, double ne,
                               const double *stp, double Tl, double nel, double &op, double &em) {
  double o = ne * 6.65E-25;
  double e = o * planck_f(map->lambda_mid, T);
  const double lam = map->lambda_bg;
  for (int j = 0; j < tb->nbf; ++j) {
    const double sig = map->sigma_bg[j];
    if (sig == 0.0) continue;
    const snapi_chem_bf &b = tb->bf[j];
    const int a = b.element - 1, z = b.stage;
    const double nl = lte_pop(tb, dv, a, z, b.g, b.energy, Tl, nel, stp[want.idx[a][z]]);
    const double nu = lte_pop(tb, dv, a, z + 1, b.g_up, b.energy_up, Tl, nel, stp[want.idx[a][z + 1]]);
    const double pop_mod = nu * ne * kSaha * pow(T, -1.5) * b.g / b.g_up * exp((tb->element[a].eion[z] - b.energy) / kK / T);
    const double stim = exp(-kH * kC / lam / kK / T);
    o += sig * (nl - pop_mod * stim);
    e += sig * pop_mod * (1.0 - stim) * planck_f(lam, T);
  }
  op = o;
  em = e;
}

// One thread per depth point: the synthesis path's state from variant 0 and,
// with resp/dstate, the response path's state and the thermodynamic
// derivatives from the other variants: referent opacity by central
// differences (v1 - v2, v3 - v4), Ne and the stage populations by a forward
// difference in T (v5 - v6) and a central one in Nt (v7 - v8), the response
// state at v9 (T raised by delta_T/2, level populations and referent opacity
// kept from v0), level responses by
// atom::compute_lte_population_responses_analytical (atom_responses.cc:453-501).
__global__ void eos_combine_kernel(const snapi_chem_table *__restrict__ tb, const ChemDerived *__restrict__ dv,
                                   const snapi_state_map *__restrict__ map, long npt, int nd, int nf, int nvar,
                                   EosWant want, const double *__restrict__ rec, double *__restrict__ state,
                                   double *__restrict__ resp, double *__restrict__ dstate,
                                   const int *__restrict__ active) {
  const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (pt >= npt) return;
  const long c = pt / nd, zd = pt % nd;
  if (active && !active[c]) return;
  double *st = state + c * nf * nd + zd;
  const double T0 = st[SNAPI_F_T * nd];
  const int nrec = 2 + want.nwant;
  const double *r0 = rec + pt * nvar * nrec;
  const double ne0 = r0[0], *s0 = r0 + 2;
  const int ih = want.idx[map->nhi_element - 1][0];
  double bo, be;
  eos_background(tb, dv, map, want, T0, ne0, s0, T0, ne0, bo, be);
  st[SNAPI_F_NE * nd] = ne0;
  st[SNAPI_F_NHI * nd] = s0[ih];
  st[SNAPI_F_OPREF * nd] = r0[1];
  st[SNAPI_F_BGOP * nd] = bo;
  st[SNAPI_F_BGEM * nd] = be;
  for (int i = 0; i < map->nline; ++i)
    for (int u = 0; u < 2; ++u) {
      const snapi_chem_level &L = tb->level[map->line_level[i][u]];
      const int a = L.element - 1;
      st[(SNAPI_NF_BASE + 2 * i + u) * nd] = lte_pop(tb, dv, a, L.stage, L.g, L.energy, T0, ne0, s0[want.idx[a][L.stage]]);
    }
  if (!resp && !dstate) return;
  const EosSeq q = eos_sequence(T0, st[SNAPI_F_NT * nd]);
  const double *rv[kEosVariants];
  for (int v = 0; v < kEosVariants; ++v) rv[v] = r0 + v * nrec;
  const double T = q.T[9], ne_e = rv[9][0], *se = rv[9] + 2;
  if (resp) {
    double *rs = resp + c * nf * nd + zd;
    for (int f = 0; f < nf; ++f) rs[f * nd] = st[f * nd];
    eos_background(tb, dv, map, want, T, ne_e, s0, T0, ne0, bo, be);
    rs[SNAPI_F_T * nd] = T;
    rs[SNAPI_F_NT * nd] = q.Nt[9];
    rs[SNAPI_F_NE * nd] = ne_e;
    rs[SNAPI_F_NHI * nd] = se[ih];
    rs[SNAPI_F_BGOP * nd] = bo;
    rs[SNAPI_F_BGEM * nd] = be;
  }
  if (!dstate) return;
  double *dT = dstate + c * 2 * nf * nd + zd, *dN = dT + (size_t)nf * nd;
  for (int f = 0; f < nf; ++f) dT[f * nd] = dN[f * nd] = 0.0;
  const double dne_dT = (rv[5][0] - rv[6][0]) / kDeltaT, dne_dNt = (rv[7][0] - rv[8][0]) / q.dNt1;
  dT[SNAPI_F_NE * nd] = dne_dT;
  dN[SNAPI_F_NE * nd] = dne_dNt;
  dT[SNAPI_F_NHI * nd] = (rv[5][2 + ih] - rv[6][2 + ih]) / kDeltaT;
  dN[SNAPI_F_NHI * nd] = (rv[7][2 + ih] - rv[8][2 + ih]) / q.dNt1;
  dT[SNAPI_F_OPREF * nd] = (rv[1][1] - rv[2][1]) / kDeltaT;
  dN[SNAPI_F_OPREF * nd] = (rv[3][1] - rv[4][1]) / q.dNt0;
  for (int i = 0; i < map->nline; ++i)
    for (int u = 0; u < 2; ++u) {
      const snapi_chem_level &L = tb->level[map->line_level[i][u]];
      const int a = L.element - 1, z = L.stage, k = 2 + want.idx[a][z];
      const double dsT = (rv[5][k] - rv[6][k]) / kDeltaT, dsN = (rv[7][k] - rv[8][k]) / q.dNt1;
      const double U = pf_direct(tb, dv, a, z, T, ne_e);
      double dU_dT = pf_direct(tb, dv, a, z, T + 0.5 * kDeltaT, ne_e);
      dU_dT -= pf_direct(tb, dv, a, z, T - 0.5 * kDeltaT, ne_e);
      dU_dT /= kDeltaT;
      const double dne = kDeltaNtf * ne_e;
      double dU_dne = pf_direct(tb, dv, a, z, T, ne_e + dne * 0.5);
      dU_dne -= pf_direct(tb, dv, a, z, T, ne_e - dne * 0.5);
      dU_dne /= dne;
      const double bz = L.g * exp(-L.energy / (kK * T)), Nz = rv[9][k];
      double rT = dsT * bz / U;
      rT -= Nz * bz / U / U * (dU_dT + dU_dne * dne_dT);
      rT += Nz * bz / U * L.energy / kK / T / T;
      double rN = dsN * bz / U;
      rN -= Nz * bz / U / U * (dU_dne * dne_dNt);
      dT[(SNAPI_NF_BASE + 2 * i + u) * nd] = rT;
      dN[(SNAPI_NF_BASE + 2 * i + u) * nd] = rN;
    }
}

// Checks the table against the kernel's limits and derives its constants.
int derive(const snapi_chem_table *tb, ChemDerived &d) {
  if (tb->nelement < 1 || tb->nelement > kMaxE || tb->nmolecule < 0 || tb->nmolecule > kMaxM ||
      tb->nlevel < 0 || tb->nlevel > SNAPI_CHEM_MAX_LEVELS)
    return SNAPI_ERR_UNSUPPORTED;
  memset(&d, 0, sizeof(d));
  int nst = 0;
  for (int a = 0; a < tb->nelement; ++a) {
    const snapi_chem_element &el = tb->element[a];
    if (el.nstage < 1 || el.nstage > kMaxS || !(el.abundance >= 0.0)) return SNAPI_ERR_UNSUPPORTED;
    for (int z = 0; z < el.nstage - 1; ++z)
      if (!std::isfinite(el.eion[z])) return SNAPI_ERR_UNSUPPORTED;
    for (int z = 0; z < el.nstage; ++z) {
      const snapi_partition &pf = el.pf[z];
      if (pf.type < SNAPI_PF_NONE || pf.type > SNAPI_PF_IRWIN) return SNAPI_ERR_UNSUPPORTED;
      if (pf.type == SNAPI_PF_TRAV) {
        if (pf.nterm < 0 || pf.nterm > SNAPI_CHEM_MAX_TERMS) return SNAPI_ERR_UNSUPPORTED;
        for (int i = 0; i < pf.nterm; ++i)
          if (pf.ncoef[i] < 0 || pf.ncoef[i] > SNAPI_CHEM_MAX_COEFS) return SNAPI_ERR_UNSUPPORTED;
      }
      if (pf.type == SNAPI_PF_IRWIN && (pf.nterm < 0 || pf.nterm > SNAPI_CHEM_MAX_COEFS))
        return SNAPI_ERR_UNSUPPORTED;
    }
    d.soff[a] = nst;
    nst += el.nstage;
  }
  d.nst = nst;
  const double er = kE * kRyd * kRyd;
  d.travf = 1.405169521381833474E+05 * ((4.0 * kPi * kE0) / (er * er));
  if (tb->lambda_ref > 0.0) hminus_lambda_terms(tb->lambda_ref, d.hm);
  int slot = 0;
  for (int a = 0; a < tb->nelement; ++a) {
    const snapi_chem_element &el = tb->element[a];
    int nt = 0;
    for (int z = 0; z < el.nstage; ++z) {
      const snapi_partition &pf = el.pf[z];
      StageDesc &sd = d.sd[a][z];
      sd.nterm = -1;
      if (pf.type != SNAPI_PF_TRAV) continue;
      sd.nterm = pf.nterm;
      sd.toff = nt;
      sd.zf = pf.charge * d.travf;
      sd.rzf = 1.0 / sd.zf;
      sd.z2 = 31.321 * pf.charge * pf.charge;
      for (int i = 0; i < pf.nterm; ++i) {
        d.xl[a][nt + i] = pf.xl[i];
        d.rxl[a][nt + i] = 1.0 / pf.xl[i];
      }
      nt += pf.nterm;
    }
    d.roff[a] = slot;
    d.uoff[a] = 2 * el.nstage - 1 + nt;
    slot += 6 * el.nstage - 2 + nt;
  }
  d.nslot = slot;
  // Traving stages to lanes, longest first onto the least-loaded lane
  std::vector<std::pair<int, int>> st;  // (terms, element * kMaxS + stage)
  for (int a = 0; a < tb->nelement; ++a)
    for (int z = 0; z < tb->element[a].nstage; ++z)
      if (d.sd[a][z].nterm >= 0) st.push_back({d.sd[a][z].nterm + 1, a * kMaxS + z});
  std::stable_sort(st.begin(), st.end(), [](const std::pair<int, int> &u, const std::pair<int, int> &v) {
    return u.first > v.first;
  });
  int load[kTeam] = {0};
  for (const auto &e : st) {
    int best = 0;
    for (int l = 1; l < kTeam; ++l)
      if (load[l] < load[best]) best = l;
    if (d.nsched[best] >= kMaxSched) return SNAPI_ERR_UNSUPPORTED;
    d.sched[best][d.nsched[best]++] = e.second;
    load[best] += e.first;
  }
  if (!(tb->element[0].abundance > 0.0)) return SNAPI_ERR_UNSUPPORTED;
  for (int m = 0; m < tb->nmolecule; ++m) {
    const snapi_chem_molecule &M = tb->molecule[m];
    if (M.ncomp < 1 || M.ncomp > kMaxC) return SNAPI_ERR_UNSUPPORTED;
    if (M.eqc == SNAPI_EQC_POLY && (M.ncoef < 1 || M.ncoef > SNAPI_CHEM_MAX_COEFS)) return SNAPI_ERR_UNSUPPORTED;
    if (M.eqc < SNAPI_EQC_NONE || M.eqc > SNAPI_EQC_HMIN) return SNAPI_ERR_UNSUPPORTED;
    for (int i = 0; i < M.ncomp; ++i) {
      const int e = M.element[i], z = M.stage[i] > 0 ? M.stage[i] : 0;
      if (e < 0 || e > tb->nelement || z >= SNAPI_CHEM_MAX_COMP_STAGE) return SNAPI_ERR_UNSUPPORTED;
      if (e > 0 && z >= tb->element[e - 1].nstage) return SNAPI_ERR_UNSUPPORTED;
      if (e == 0 && z != 0) return SNAPI_ERR_UNSUPPORTED;
      d.celem[m][i] = e;
      d.cstage[m][i] = z;
      d.mul[m][e] += 1.0;
      d.q[m] += (double)M.stage[i];
    }
  }
  for (int l = 0; l < tb->nlevel; ++l) {
    const snapi_chem_level &L = tb->level[l];
    if (L.element < 1 || L.element > tb->nelement || L.stage < 0 ||
        L.stage >= tb->element[L.element - 1].nstage)
      return SNAPI_ERR_UNSUPPORTED;
  }
  if (tb->nbf < 0 || tb->nbf > SNAPI_CHEM_MAX_BF) return SNAPI_ERR_UNSUPPORTED;
  for (int j = 0; j < tb->nbf; ++j) {
    const snapi_chem_bf &b = tb->bf[j];
    // a transition needs the stage above it
    if (b.element < 1 || b.element > tb->nelement || b.stage < 0 ||
        b.stage + 1 >= tb->element[b.element - 1].nstage)
      return SNAPI_ERR_UNSUPPORTED;
  }
  return SNAPI_OK;
}

// Workgroups for n teams and the dynamic LDS of their regions; false when
// either exceeds what a launch takes.
bool team_grid(long n, int nslot, long &blocks, size_t &lds) {
  blocks = (n * kTeam + kChemThreads - 1) / kChemThreads;
  lds = (size_t)(kChemThreads / kTeam) * nslot * sizeof(double);
  return blocks <= 0x7fffffffL && lds <= 44 * 1024;  // with the static tables, within 64 KiB
}

int launch_chemeq(const snapi_chem_table *dtb, const ChemDerived *ddv, int nslot, long npt, int nd, const double *T,
                  const double *Nt, double *ne, double *stage, double *mol, double *lev, double *opref, double *rho,
                  int *it, hipStream_t st) {
  long blocks;
  size_t lds;
  if (!team_grid(npt, nslot, blocks, lds)) return SNAPI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(chemeq_kernel, dim3((unsigned)blocks), dim3(kChemThreads), lds, st, dtb, ddv, npt, T, Nt, nd,
                     ne, stage, mol, lev, opref, rho, it);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

int launch_he_tangent(const snapi_chem_table *dtb, const ChemDerived *ddv, int nslot, long ncol, int nd,
                      const double *T, const double *logtau, const double *Nt, int ndir, const double *dT, double *dN,
                      double *rec, hipStream_t st, const int *active = nullptr) {
  long blocks;
  size_t lds;
  if (!team_grid(ncol * nd * 5, nslot, blocks, lds)) return SNAPI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(he_partials_kernel, dim3((unsigned)blocks), dim3(kChemThreads), lds, st, dtb, ddv, ncol * nd, nd,
                     T, Nt, rec, active);
  const long b2 = (ncol * ndir + 127) / 128;
  if (b2 > 0x7fffffffL) return SNAPI_ERR_UNSUPPORTED;
  if (b2 > 0)
    hipLaunchKernelGGL(he_tangent_kernel, dim3((unsigned)b2), dim3(128), 0, st, ncol, nd, ndir, T, logtau, Nt, rec, dT,
                       dN, active);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

int launch_hequilibrium(const snapi_chem_table *dtb, const ChemDerived *ddv, int nslot, long ncol, int nd,
                        const double *T, const double *logtau, double *Nt, int *it, hipStream_t st,
                        const int *active = nullptr, int group = 1) {
  long blocks;
  size_t lds;
  if (!team_grid(ncol, nslot, blocks, lds)) return SNAPI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(hequilibrium_kernel, dim3((unsigned)blocks), dim3(kChemThreads), lds, st, dtb, ddv, ncol, nd, T,
                     logtau, Nt, it, active, group);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

#define CCHK(x)              \
  do {                       \
    if ((x) != hipSuccess) { \
      rc = SNAPI_ERR_DEVICE; \
      goto done;             \
    }                        \
  } while (0)

// Copies the table and its derived constants into one device buffer.
int upload_table(const snapi_chem_table *tb, const ChemDerived &d, void **buf, hipStream_t st) {
  *buf = nullptr;
  const size_t n = sizeof(snapi_chem_table) + sizeof(ChemDerived);
  if (hipMalloc(buf, n) != hipSuccess) return SNAPI_ERR_DEVICE;
  if (hipMemcpyAsync(*buf, tb, sizeof(snapi_chem_table), hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync((char *)*buf + sizeof(snapi_chem_table), &d, sizeof(ChemDerived), hipMemcpyHostToDevice, st) !=
          hipSuccess) {
    hipFree(*buf);
    *buf = nullptr;
    return SNAPI_ERR_DEVICE;
  }
  return SNAPI_OK;
}

// Checks the state map against the table and lists the stage populations
// the equilibria must record.
int eos_want(const snapi_chem_table *tb, const snapi_state_map *map, EosWant &w) {
  if (map->nline < 0 || map->nline > SNAPI_MAX_LINES || map->nhi_element < 1 || map->nhi_element > tb->nelement ||
      !(map->lambda_mid > 0.0))
    return SNAPI_ERR_ARG;
  memset(&w, -1, sizeof(w));
  w.nwant = 0;
  auto add = [&](int a, int z) {
    if (w.idx[a][z] < 0) {
      if (w.nwant >= kMaxWant) return false;
      w.idx[a][z] = (signed char)w.nwant++;
    }
    return true;
  };
  if (!add(map->nhi_element - 1, 0)) return SNAPI_ERR_UNSUPPORTED;
  for (int i = 0; i < map->nline; ++i)
    for (int u = 0; u < 2; ++u) {
      const int l = map->line_level[i][u];
      if (l < 0 || l >= tb->nlevel) return SNAPI_ERR_ARG;
      if (!add(tb->level[l].element - 1, tb->level[l].stage)) return SNAPI_ERR_UNSUPPORTED;
    }
  for (int j = 0; j < tb->nbf; ++j) {
    if (map->sigma_bg[j] == 0.0) continue;
    if (!(map->lambda_bg > 0.0)) return SNAPI_ERR_ARG;
    const snapi_chem_bf &b = tb->bf[j];
    if (!add(b.element - 1, b.stage) || !add(b.element - 1, b.stage + 1)) return SNAPI_ERR_UNSUPPORTED;
  }
  return SNAPI_OK;
}

int launch_eos(const snapi_chem_table *dtb, const ChemDerived *ddv, const snapi_state_map *dmap, int nslot, long npt,
               int nd, int nf, int nvar, const EosWant &want, double *rec, double *state, double *resp,
               double *dstate, hipStream_t st, const int *active = nullptr,
               const snapi_chem_table *dtb2 = nullptr, const ChemDerived *ddv2 = nullptr, double *he_rec = nullptr) {
  long blocks;
  size_t lds;
  if (!team_grid(npt * nvar, nslot, blocks, lds)) return SNAPI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(eos_variants_kernel, dim3((unsigned)blocks), dim3(kChemThreads), lds, st, dtb, ddv, npt, nd, nf,
                     nvar, state, want, rec, active, dtb2, ddv2, he_rec);
  if (hipGetLastError() != hipSuccess) return SNAPI_ERR_DEVICE;
  const long b2 = (npt + 255) / 256;
  if (b2 > 0x7fffffffL) return SNAPI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(eos_combine_kernel, dim3((unsigned)b2), dim3(256), 0, st, dtb, ddv, dmap, npt, nd, nf, nvar, want,
                     rec, state, resp, dstate, active);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

}  // namespace

namespace snapi {

// Device copies of a species table (with its derived constants) and,
// optionally, a state map, for callers that run the hydrostatic march and
// the equation-of-state pass many times (the inversion).
int chem_upload(const snapi_chem_table *tb, const snapi_state_map *map, ChemDevice &cd, hipStream_t st) {
  cd = ChemDevice{};
  if (!(tb->lambda_ref > 0.0)) return SNAPI_ERR_ARG;
  ChemDerived d;
  int rc = derive(tb, d);
  if (rc) return rc;
  cd.nslot = d.nslot;
  if (map) {
    EosWant w;
    rc = eos_want(tb, map, w);
    if (rc) return rc;
    cd.nrec = 2 + w.nwant;
    cd.nf = SNAPI_NF_BASE + 2 * map->nline;
    static_assert(sizeof(EosWant) <= sizeof(cd.want), "want table fits");
    memcpy(cd.want, &w, sizeof(w));
    if (hipMalloc(&cd.map, sizeof(snapi_state_map)) != hipSuccess ||
        hipMemcpyAsync(cd.map, map, sizeof(snapi_state_map), hipMemcpyHostToDevice, st) != hipSuccess) {
      chem_free(cd);
      return SNAPI_ERR_DEVICE;
    }
  }
  rc = upload_table(tb, d, &cd.buf, st);
  if (rc) chem_free(cd);
  return rc;
}

void chem_free(ChemDevice &cd) {
  hipFree(cd.buf);
  hipFree(cd.map);
  cd.buf = cd.map = nullptr;
}

size_t chem_eos_scratch(const ChemDevice &cd, long npt, bool derivatives) {
  return (size_t)npt * (derivatives ? kEosVariants : 1) * cd.nrec * sizeof(double);
}

int chem_hequilibrium(const ChemDevice &cd, long ncol, int nd, const double *T, const double *logtau, double *Nt,
                      const int *active, int group, hipStream_t st) {
  return launch_hequilibrium((const snapi_chem_table *)cd.buf,
                             (const ChemDerived *)((const char *)cd.buf + sizeof(snapi_chem_table)), cd.nslot, ncol,
                             nd, T, logtau, Nt, nullptr, st, active, group);
}

size_t chem_he_tangent_scratch(long ncol, int nd) { return (size_t)ncol * nd * 5 * sizeof(double); }

int chem_he_tangent(const ChemDevice &cd, long ncol, int nd, const double *T, const double *logtau, const double *Nt,
                    int ndir, const double *dT, double *dN, double *scratch, const int *active, hipStream_t st) {
  return launch_he_tangent((const snapi_chem_table *)cd.buf,
                           (const ChemDerived *)((const char *)cd.buf + sizeof(snapi_chem_table)), cd.nslot, ncol, nd,
                           T, logtau, Nt, ndir, dT, dN, scratch, st, active);
}

int chem_eos_he_tangent(const ChemDevice &cd, const ChemDevice &cd_he, long ncol, int nd, double *state,
                        double *resp, double *dstate, double *scratch, const double *T, const double *logtau,
                        const double *Nt, int ndir, const double *dT, double *dN, double *he_scratch,
                        const int *active, hipStream_t st) {
  if (!resp && !dstate) return SNAPI_ERR_ARG;
  // the two tables must hold the same species (they differ in the referent
  // wavelength): the team layout of the first serves the second's opacity
  if (cd.nslot != cd_he.nslot) return SNAPI_ERR_ARG;
  EosWant w;
  memcpy(&w, cd.want, sizeof(w));
  const auto *tb2 = (const snapi_chem_table *)cd_he.buf;
  const auto *dv2 = (const ChemDerived *)((const char *)cd_he.buf + sizeof(snapi_chem_table));
  int rc = launch_eos((const snapi_chem_table *)cd.buf,
                      (const ChemDerived *)((const char *)cd.buf + sizeof(snapi_chem_table)),
                      (const snapi_state_map *)cd.map, cd.nslot, ncol * nd, nd, cd.nf, kEosVariants, w, scratch, state,
                      resp, dstate, st, active, tb2, dv2, he_scratch);
  if (rc) return rc;
  const long b2 = (ncol * ndir + 127) / 128;
  if (b2 > 0x7fffffffL) return SNAPI_ERR_UNSUPPORTED;
  if (b2 > 0)
    hipLaunchKernelGGL(he_tangent_kernel, dim3((unsigned)b2), dim3(128), 0, st, ncol, nd, ndir, T, logtau, Nt,
                       he_scratch, dT, dN, active);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

int chem_eos(const ChemDevice &cd, long ncol, int nd, double *state, double *resp, double *dstate, double *scratch,
             const int *active, hipStream_t st) {
  EosWant w;
  memcpy(&w, cd.want, sizeof(w));
  const int nvar = (resp || dstate) ? kEosVariants : 1;
  return launch_eos((const snapi_chem_table *)cd.buf,
                    (const ChemDerived *)((const char *)cd.buf + sizeof(snapi_chem_table)),
                    (const snapi_state_map *)cd.map, cd.nslot, ncol * nd, nd, cd.nf, nvar, w, scratch, state, resp,
                    dstate, st, active);
}

}  // namespace snapi

extern "C" int snapi_chemeq_dev(const snapi_chem_table *table, int ncol, int nd, const double *T_dev,
                                const double *Nt_dev, double *ne_dev, double *stage_dev, double *molecule_dev,
                                double *level_dev, double *opref_dev, int *iterations_dev, void *stream) {
  if (!table || ncol <= 0 || nd <= 0 || !T_dev || !Nt_dev || !ne_dev) return SNAPI_ERR_ARG;
  if (opref_dev && !(table->lambda_ref > 0.0)) return SNAPI_ERR_ARG;
  ChemDerived d;
  int rc = derive(table, d);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  void *buf = nullptr;
  rc = upload_table(table, d, &buf, st);
  if (rc) return rc;
  rc = launch_chemeq((const snapi_chem_table *)buf,
                     (const ChemDerived *)((const char *)buf + sizeof(snapi_chem_table)), d.nslot, (long)ncol * nd, nd, T_dev,
                     Nt_dev, ne_dev, stage_dev, table->nmolecule ? molecule_dev : nullptr,
                     table->nlevel ? level_dev : nullptr, opref_dev, nullptr, iterations_dev, st);
  // the table buffer is freed once the kernel that reads it has finished
  if (hipStreamSynchronize(st) != hipSuccess && rc == SNAPI_OK) rc = SNAPI_ERR_DEVICE;
  hipFree(buf);
  return rc;
}

extern "C" int snapi_chemeq(const snapi_chem_table *table, int ncol, int nd, const double *T, const double *Nt,
                            double *ne, double *stage, double *molecule, double *level, double *opref,
                            int *iterations) {
  if (!table || ncol <= 0 || nd <= 0 || !T || !Nt || !ne) return SNAPI_ERR_ARG;
  if (opref && !(table->lambda_ref > 0.0)) return SNAPI_ERR_ARG;
  ChemDerived d;
  int rc = derive(table, d);
  if (rc) return rc;
  const size_t np = (size_t)ncol * nd;
  const size_t nst = (size_t)d.nst, nm = (size_t)table->nmolecule, nl = (size_t)table->nlevel;
  double *dT = nullptr, *dNt = nullptr, *dne = nullptr, *dst = nullptr, *dmo = nullptr, *dlv = nullptr,
         *dkr = nullptr;
  int *dit = nullptr;
  void *buf = nullptr;
  CCHK(hipMalloc((void **)&dT, np * 8));
  CCHK(hipMalloc((void **)&dNt, np * 8));
  CCHK(hipMalloc((void **)&dne, np * 8));
  if (stage) CCHK(hipMalloc((void **)&dst, np * nst * 8));
  if (molecule && nm) CCHK(hipMalloc((void **)&dmo, np * nm * 8));
  if (level && nl) CCHK(hipMalloc((void **)&dlv, np * nl * 8));
  if (opref) CCHK(hipMalloc((void **)&dkr, np * 8));
  if (iterations) CCHK(hipMalloc((void **)&dit, np * sizeof(int)));
  CCHK(hipMemcpy(dT, T, np * 8, hipMemcpyHostToDevice));
  CCHK(hipMemcpy(dNt, Nt, np * 8, hipMemcpyHostToDevice));
  rc = upload_table(table, d, &buf, nullptr);
  if (rc) goto done;
  rc = launch_chemeq((const snapi_chem_table *)buf,
                     (const ChemDerived *)((const char *)buf + sizeof(snapi_chem_table)), d.nslot, (long)np, nd, dT, dNt, dne,
                     dst, dmo, dlv, dkr, nullptr, dit, nullptr);
  if (rc) goto done;
  CCHK(hipDeviceSynchronize());
  CCHK(hipMemcpy(ne, dne, np * 8, hipMemcpyDeviceToHost));
  if (dst) CCHK(hipMemcpy(stage, dst, np * nst * 8, hipMemcpyDeviceToHost));
  if (dmo) CCHK(hipMemcpy(molecule, dmo, np * nm * 8, hipMemcpyDeviceToHost));
  if (dlv) CCHK(hipMemcpy(level, dlv, np * nl * 8, hipMemcpyDeviceToHost));
  if (dkr) CCHK(hipMemcpy(opref, dkr, np * 8, hipMemcpyDeviceToHost));
  if (dit) CCHK(hipMemcpy(iterations, dit, np * sizeof(int), hipMemcpyDeviceToHost));
done:
  hipFree(dT); hipFree(dNt); hipFree(dne); hipFree(dst); hipFree(dmo); hipFree(dlv); hipFree(dkr); hipFree(dit); hipFree(buf);
  return rc;
}

extern "C" int snapi_hequilibrium_dev(const snapi_chem_table *table, int ncol, int nd, const double *T_dev,
                                      const double *logtau_dev, double *Nt_dev, double *ne_dev, double *opref_dev,
                                      double *rho_dev, double *stage_dev, double *molecule_dev, double *level_dev,
                                      int *iterations_dev, void *stream) {
  if (!table || ncol <= 0 || nd <= 0 || !T_dev || !logtau_dev || !Nt_dev) return SNAPI_ERR_ARG;
  if (!(table->lambda_ref > 0.0)) return SNAPI_ERR_ARG;
  ChemDerived d;
  int rc = derive(table, d);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  void *buf = nullptr;
  rc = upload_table(table, d, &buf, st);
  if (rc) return rc;
  const snapi_chem_table *dtb = (const snapi_chem_table *)buf;
  const ChemDerived *ddv = (const ChemDerived *)((const char *)buf + sizeof(snapi_chem_table));
  rc = launch_hequilibrium(dtb, ddv, d.nslot, ncol, nd, T_dev, logtau_dev, Nt_dev, iterations_dev, st);
  // chemical equilibrium, LTE populations, opacity and density at the final Nt (atmos_pp.cc:449-454)
  if (rc == SNAPI_OK && (ne_dev || opref_dev || rho_dev || stage_dev || molecule_dev || level_dev))
    rc = launch_chemeq(dtb, ddv, d.nslot, (long)ncol * nd, nd, T_dev, Nt_dev, ne_dev, stage_dev,
                       table->nmolecule ? molecule_dev : nullptr, table->nlevel ? level_dev : nullptr, opref_dev,
                       rho_dev, nullptr, st);
  if (hipStreamSynchronize(st) != hipSuccess && rc == SNAPI_OK) rc = SNAPI_ERR_DEVICE;
  hipFree(buf);
  return rc;
}

extern "C" int snapi_hequilibrium(const snapi_chem_table *table, int ncol, int nd, const double *T,
                                  const double *logtau, double *Nt, double *ne, double *opref, double *rho,
                                  double *stage, double *molecule, double *level, int *iterations) {
  if (!table || ncol <= 0 || nd <= 0 || !T || !logtau || !Nt) return SNAPI_ERR_ARG;
  if (!(table->lambda_ref > 0.0)) return SNAPI_ERR_ARG;
  ChemDerived d;
  int rc = derive(table, d);
  if (rc) return rc;
  const size_t np = (size_t)ncol * nd;
  const size_t nst = (size_t)d.nst, nm = (size_t)table->nmolecule, nl = (size_t)table->nlevel;
  double *dT = nullptr, *dlt = nullptr, *dNt = nullptr, *dne = nullptr, *dkr = nullptr, *drho = nullptr,
         *dst = nullptr, *dmo = nullptr, *dlv = nullptr;
  int *dit = nullptr;
  CCHK(hipMalloc((void **)&dT, np * 8));
  CCHK(hipMalloc((void **)&dlt, np * 8));
  CCHK(hipMalloc((void **)&dNt, np * 8));
  if (ne) CCHK(hipMalloc((void **)&dne, np * 8));
  if (opref) CCHK(hipMalloc((void **)&dkr, np * 8));
  if (rho) CCHK(hipMalloc((void **)&drho, np * 8));
  if (stage) CCHK(hipMalloc((void **)&dst, np * nst * 8));
  if (molecule && nm) CCHK(hipMalloc((void **)&dmo, np * nm * 8));
  if (level && nl) CCHK(hipMalloc((void **)&dlv, np * nl * 8));
  if (iterations) CCHK(hipMalloc((void **)&dit, np * sizeof(int)));
  CCHK(hipMemcpy(dT, T, np * 8, hipMemcpyHostToDevice));
  CCHK(hipMemcpy(dlt, logtau, np * 8, hipMemcpyHostToDevice));
  rc = snapi_hequilibrium_dev(table, ncol, nd, dT, dlt, dNt, dne, dkr, drho, dst, dmo, dlv, dit, nullptr);
  if (rc) goto done;
  CCHK(hipMemcpy(Nt, dNt, np * 8, hipMemcpyDeviceToHost));
  if (dne) CCHK(hipMemcpy(ne, dne, np * 8, hipMemcpyDeviceToHost));
  if (dkr) CCHK(hipMemcpy(opref, dkr, np * 8, hipMemcpyDeviceToHost));
  if (drho) CCHK(hipMemcpy(rho, drho, np * 8, hipMemcpyDeviceToHost));
  if (dst) CCHK(hipMemcpy(stage, dst, np * nst * 8, hipMemcpyDeviceToHost));
  if (dmo) CCHK(hipMemcpy(molecule, dmo, np * nm * 8, hipMemcpyDeviceToHost));
  if (dlv) CCHK(hipMemcpy(level, dlv, np * nl * 8, hipMemcpyDeviceToHost));
  if (dit) CCHK(hipMemcpy(iterations, dit, np * sizeof(int), hipMemcpyDeviceToHost));
done:
  hipFree(dT); hipFree(dlt); hipFree(dNt); hipFree(dne); hipFree(dkr); hipFree(drho);
  hipFree(dst); hipFree(dmo); hipFree(dlv); hipFree(dit);
  return rc;
}

extern "C" int snapi_hequilibrium_tangent_dev(const snapi_chem_table *table, int ncol, int nd, const double *T_dev,
                                              const double *logtau_dev, const double *Nt_dev, int ndir,
                                              const double *dT_dev, double *dNt_dev, void *stream) {
  if (!table || ncol <= 0 || nd <= 0 || ndir < 0 || !T_dev || !logtau_dev || !Nt_dev) return SNAPI_ERR_ARG;
  if (ndir > 0 && (!dT_dev || !dNt_dev)) return SNAPI_ERR_ARG;
  if (!(table->lambda_ref > 0.0)) return SNAPI_ERR_ARG;
  ChemDerived d;
  int rc = derive(table, d);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  void *buf = nullptr;
  double *rec = nullptr;
  rc = upload_table(table, d, &buf, st);
  if (rc) return rc;
  if (hipMalloc((void **)&rec, (size_t)ncol * nd * 5 * sizeof(double)) != hipSuccess) {
    hipFree(buf);
    return SNAPI_ERR_DEVICE;
  }
  rc = launch_he_tangent((const snapi_chem_table *)buf,
                         (const ChemDerived *)((const char *)buf + sizeof(snapi_chem_table)), d.nslot, ncol, nd, T_dev,
                         logtau_dev, Nt_dev, ndir, dT_dev, dNt_dev, rec, st);
  if (hipStreamSynchronize(st) != hipSuccess && rc == SNAPI_OK) rc = SNAPI_ERR_DEVICE;
  hipFree(rec);
  hipFree(buf);
  return rc;
}

extern "C" int snapi_hequilibrium_tangent(const snapi_chem_table *table, int ncol, int nd, const double *T,
                                          const double *logtau, const double *Nt, int ndir, const double *dT,
                                          double *dNt) {
  if (!table || ncol <= 0 || nd <= 0 || ndir < 0 || !T || !logtau || !Nt) return SNAPI_ERR_ARG;
  if (ndir > 0 && (!dT || !dNt)) return SNAPI_ERR_ARG;
  const size_t np = (size_t)ncol * nd, nq = np * ndir;
  double *dT_ = nullptr, *dlt = nullptr, *dN_ = nullptr, *ddT = nullptr, *ddN = nullptr;
  int rc = SNAPI_OK;
  CCHK(hipMalloc((void **)&dT_, np * 8));
  CCHK(hipMalloc((void **)&dlt, np * 8));
  CCHK(hipMalloc((void **)&dN_, np * 8));
  if (nq) {
    CCHK(hipMalloc((void **)&ddT, nq * 8));
    CCHK(hipMalloc((void **)&ddN, nq * 8));
    CCHK(hipMemcpy(ddT, dT, nq * 8, hipMemcpyHostToDevice));
  }
  CCHK(hipMemcpy(dT_, T, np * 8, hipMemcpyHostToDevice));
  CCHK(hipMemcpy(dlt, logtau, np * 8, hipMemcpyHostToDevice));
  CCHK(hipMemcpy(dN_, Nt, np * 8, hipMemcpyHostToDevice));
  rc = snapi_hequilibrium_tangent_dev(table, ncol, nd, dT_, dlt, dN_, ndir, ddT, ddN, nullptr);
  if (rc) goto done;
  if (nq) CCHK(hipMemcpy(dNt, ddN, nq * 8, hipMemcpyDeviceToHost));
done:
  hipFree(dT_); hipFree(dlt); hipFree(dN_); hipFree(ddT); hipFree(ddN);
  return rc;
}

extern "C" int snapi_eos_state_dev(const snapi_chem_table *table, const snapi_state_map *map, int ncol, int nd,
                                   double *state_dev, double *resp_dev, double *dstate_dev, void *stream) {
  if (!table || !map || ncol <= 0 || nd <= 0 || !state_dev) return SNAPI_ERR_ARG;
  if (!(table->lambda_ref > 0.0)) return SNAPI_ERR_ARG;
  ChemDerived d;
  int rc = derive(table, d);
  if (rc) return rc;
  EosWant want;
  rc = eos_want(table, map, want);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const int nf = SNAPI_NF_BASE + 2 * map->nline;
  const long npt = (long)ncol * nd;
  const int nvar = (resp_dev || dstate_dev) ? kEosVariants : 1;
  void *buf = nullptr, *mbuf = nullptr, *rec = nullptr;
  rc = upload_table(table, d, &buf, st);
  if (rc) return rc;
  if (hipMalloc(&mbuf, sizeof(snapi_state_map)) != hipSuccess ||
      hipMalloc(&rec, (size_t)npt * nvar * (2 + want.nwant) * sizeof(double)) != hipSuccess ||
      hipMemcpyAsync(mbuf, map, sizeof(snapi_state_map), hipMemcpyHostToDevice, st) != hipSuccess)
    rc = SNAPI_ERR_DEVICE;
  if (rc == SNAPI_OK)
    rc = launch_eos((const snapi_chem_table *)buf, (const ChemDerived *)((const char *)buf + sizeof(snapi_chem_table)),
                    (const snapi_state_map *)mbuf, d.nslot, npt, nd, nf, nvar, want, (double *)rec, state_dev,
                    resp_dev, dstate_dev, st);
  if (hipStreamSynchronize(st) != hipSuccess && rc == SNAPI_OK) rc = SNAPI_ERR_DEVICE;
  hipFree(buf);
  hipFree(mbuf);
  hipFree(rec);
  return rc;
}

extern "C" int snapi_eos_state(const snapi_chem_table *table, const snapi_state_map *map, int ncol, int nd,
                               double *state, double *resp, double *dstate) {
  if (!table || !map || ncol <= 0 || nd <= 0 || !state) return SNAPI_ERR_ARG;
  if (map->nline < 0 || map->nline > SNAPI_MAX_LINES) return SNAPI_ERR_ARG;
  const size_t n = (size_t)ncol * (SNAPI_NF_BASE + 2 * map->nline) * nd;
  double *ds = nullptr, *dr = nullptr, *dd = nullptr;
  int rc = SNAPI_OK;
  CCHK(hipMalloc((void **)&ds, n * 8));
  if (resp) CCHK(hipMalloc((void **)&dr, n * 8));
  if (dstate) CCHK(hipMalloc((void **)&dd, 2 * n * 8));
  CCHK(hipMemcpy(ds, state, n * 8, hipMemcpyHostToDevice));
  rc = snapi_eos_state_dev(table, map, ncol, nd, ds, dr, dd, nullptr);
  if (rc) goto done;
  CCHK(hipMemcpy(state, ds, n * 8, hipMemcpyDeviceToHost));
  if (dr) CCHK(hipMemcpy(resp, dr, n * 8, hipMemcpyDeviceToHost));
  if (dd) CCHK(hipMemcpy(dstate, dd, 2 * n * 8, hipMemcpyDeviceToHost));
done:
  hipFree(ds); hipFree(dr); hipFree(dd);
  return rc;
}
