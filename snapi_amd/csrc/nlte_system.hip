// The NLTE population-response system of one atom on gfx950, in double:
// atmosphere::compute_nlte_population_responses (atmos/atmos_responses.cc:
// 13-157) for a batch of columns.
//
//   norm_kernel      atom::compute_profile_norm (atmos/atmol/atom/atom_rad.cc:
//                    885-930): per (column, ray, transition, depth) the
//                    wavelength-integrated profile and its T, Nt, vt
//                    derivatives (voigt_der_T / _density / _vt,
//                    atom_rad.cc:666-797).
//   prep_kernel      the opacity and emissivity normalised by the referent
//                    opacity as the formal solver takes them
//                    (atmosphere::normalize_to_referent_opacity).
//   (formal_resp_kernel of nlte_resp.hip: intensity and dI/dop, dI/dem.)
//   coef_kernel      everything of atom::add_response_contributions_new
//                    (atom_responses.cc:668-1128) that depends on one depth
//                    of one (ray, wavelength) problem: line profiles and
//                    their derivatives, the mean-intensity increments, the
//                    explicit and level derivatives of op and em, the local
//                    terms, and two compact records: the depth as the
//                    perturbed row l (the rate weights gamma of each
//                    assembled transition, summed into two level vectors)
//                    and as the perturbing depth ll (what multiplies
//                    dI_l/dop_ll and dI_l/dem_ll).
//   assemble_kernel  the non-local sums: for every pair (l, ll) the sum over
//                    problems of dI_l/dop_ll and dI_l/dem_ll times the two
//                    records, into the (l, ll) block of the matrix and the
//                    ll right-hand sides.  Each line's own terms reduce to
//                    two sums per (l, ll) (X = sum gamma s dI/dop,
//                    Y = sum gamma s dI/dem) that enter four entries at the
//                    end.
//   sum_kernel       the per-problem mean intensities and local terms summed
//                    over problems, in problem order.
//   rate_kernel      the rate part of atom::compute_nlte_population_responses
//                    (atom_responses.cc:26-170): radiative and collisional
//                    rates of the diagonal blocks, the collisional terms of
//                    the right-hand sides and the particle-conservation rows.
//
// Layouts per column: depth [SNAPI_NLTE_NFIELD][nd], pops [nmap][nd], damp
// [2][ntr][nd], coll [3][nmap][nmap][nd], vr [nray][nd], opem
// [nray][nlam][8][nd]; problems p = ray * nlam + wavelength, in the
// reference's loop order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nlte_device.h"
#include "rt_common.h"

using namespace snapi;

namespace {

constexpr int kMaxLev = 8;     // levels of the atom
constexpr int kMaxTr = 32;     // transitions
constexpr int kMaxRay = 16;    // rays of the angular quadrature
constexpr int kRows = 8;       // rows l per assembly workgroup (one wavefront each)
constexpr double kSaha = 2.0706615382E-16;  // cmlib/const.h:24
constexpr double kDeltaT = 5.0;             // cmlib/const.h:27
constexpr double kDeltaNtFrac = 1E-3;       // cmlib/const.h:29

// The atom and the quadrature, by value in the kernel arguments.
struct NAtom {
  int nmap, ntr, nstage, damping, nla, nray, tau_grid;
  double mass;
  int stage[kMaxLev], lidx[kMaxLev], first[kMaxLev], nl[kMaxLev], bftr[kMaxLev];
  double ee[kMaxLev], g[kMaxLev], ip[kMaxLev];
  int type[kMaxTr], lo[kMaxTr], up[kMaxTr];
  double A[kMaxTr], Blu[kMaxTr], Bul[kMaxTr], ar[kMaxTr], lcrit[kMaxTr];
  int la[kMaxTr];                      // assembled lines (A > 1), in the reference's pair order
  signed char tmap[kMaxLev][kMaxLev];  // line between two levels, -1
  double wray[kMaxRay];
};

// Record sizes: row record [nla + 2 nmap]; column record [4 nmap + 10 + 7 nla];
// per-problem sums [2 ntr + nmap^2 + 3 nmap].
__host__ __device__ inline int nrow_rec(const NAtom &a) { return a.nla + 2 * a.nmap; }
__host__ __device__ inline int ncol_rec(const NAtom &a) { return 4 * a.nmap + 10 + 7 * a.nla; }
__host__ __device__ inline int nsum_rec(const NAtom &a) { return 2 * a.ntr + a.nmap * a.nmap + 3 * a.nmap; }

__device__ __forceinline__ double broad_dop(double lam0, double T, double vt, double mass) {
  return sqrt(2.0 * kK * T / mass + vt * vt) * lam0 / kC;
}
__device__ __forceinline__ double fvoigt(double v, double a) {
  double H, F;
  voigt<double>(v, a, H, F);
  return H;
}
__device__ __forceinline__ double planck_f(double lam, double T) {
  return 2.0 * 6.62606957 * 8.9875517E-7 / lam / lam / lam / lam / lam / (exp(1.4387769 / lam / T) - 1.0);
}
__device__ __forceinline__ double planck_fd(double lam, double T) {
  const double e = (1.4387769 / lam / T);
  const double t = (e > 5.0) ? 1.0 : exp(e) / (exp(e) - 1.0);
  return planck_f(lam, T) * t * 1.4387769 / lam / T / T;
}

#pragma clang fp contract(off)

// ---------------------------------------------------------------- norms
// grid (ceil(nd/64), ntr, ncol * nray); norm [col][ray][nd][ntr],
// normd [col][ray][3][nd][ntr].
__global__ __launch_bounds__(64) void norm_kernel(NAtom at, int nd, int nlam, const double *__restrict__ lam,
                                                  const double *__restrict__ lamw, const double *__restrict__ depth,
                                                  const double *__restrict__ damp, const double *__restrict__ vr,
                                                  double *__restrict__ norm, double *__restrict__ normd) {
  const int d = blockIdx.x * 64 + threadIdx.x;
  const int t = blockIdx.y;
  const int cr = blockIdx.z;  // column * nray + ray
  const int col = cr / at.nray;
  if (d >= nd) return;
  double n0 = 0.0, n1 = 0.0, n2 = 0.0, n3 = 0.0;
  if (at.type[t] == SNAPI_TR_LINE) {
    const double *dp = depth + (size_t)col * SNAPI_NLTE_NFIELD * nd;
    const double T = dp[SNAPI_NLTE_T * nd + d], vt = dp[SNAPI_NLTE_VT * nd + d], Nt = dp[SNAPI_NLTE_NT * nd + d];
    const double v = vr[(size_t)cr * nd + d];
    const double *dm = damp + (size_t)col * 2 * at.ntr * nd;
    const double gc = dm[(size_t)t * nd + d], gcT = dm[((size_t)at.ntr + t) * nd + d];
    const int up = at.up[t], lo = at.lo[t];
    const double lam0 = kH * kC / (at.ee[up] - at.ee[lo]);
    const double lamc = fabs(kH * kC / (at.ee[up] - at.ee[lo]));
    const double ar = at.ar[t];
    const double sf = (lam0 * lam0) / (4.0 * kPi * kC);
    // wavelength-independent parts of voigt_der_T / _density / _vt
    const double dld = broad_dop(lamc, T, vt, at.mass);
    const double damping = ar + gc * at.damping;
    const double aa = damping / dld * (lamc * lamc) / (4.0 * kPi * kC);
    const double dvd = dld / lamc * kC;
    const double ddT = lamc / (kC * dvd) * kK / at.mass;
    double daT = -aa / dld * ddT;
    daT += gcT / dld * (lamc * lamc) / (4.0 * kPi * kC) * at.damping;
    const double delta_Nt = kDeltaNtFrac * Nt;
    const double delta_a = gc * kDeltaNtFrac * at.damping / dld * (lamc * lamc) / (4.0 * kPi * kC);
    const double ddv = lamc / (kC * dvd) * vt;
    const double dav = -aa / dld * ddv;
    for (int l = 0; l < nlam; ++l) {
      const double lm = lam[l], w = lamw[l];
      const double dl = broad_dop(lm, T, vt, at.mass);
      const double a = sf * (ar + gc * at.damping) / dl;
      const double x = (lam0 - lm * (1.0 + v / kC)) / dl;
      n0 += 1.0 / dl * fvoigt(x, a) * w;
      const double xc = (lamc - lm * (1.0 + v / kC)) / dld;
      const double dxT = -xc / dld * ddT;
      const double dT = (fvoigt(xc + dxT * 0.5 * kDeltaT, aa + daT * 0.5 * kDeltaT) / (dld + 0.5 * ddT * kDeltaT) -
                         fvoigt(xc - dxT * 0.5 * kDeltaT, aa - daT * 0.5 * kDeltaT) / (dld - 0.5 * ddT * kDeltaT)) /
                        kDeltaT;
      n1 += dT * w;
      const double dN = (fvoigt(xc, aa + delta_a * 0.5) - fvoigt(xc, aa - delta_a * 0.5)) / dld / delta_Nt;
      n2 += dN * at.damping * w;
      double H, F, dH, dF;
      voigt_d<double>(xc, aa, H, F, dH, dF);
      const double dxv = -xc / dld * ddv;
      n3 += (dH / dld * dxv - dF / dld * dav - H / dld / dld * ddv) * w;
    }
  }
  const size_t o = (size_t)cr * nd * at.ntr;
  norm[o + (size_t)d * at.ntr + t] = n0;
  const size_t od = (size_t)cr * 3 * nd * at.ntr;
  normd[od + (size_t)d * at.ntr + t] = n1;
  normd[od + ((size_t)nd + d) * at.ntr + t] = n2;
  normd[od + ((size_t)2 * nd + d) * at.ntr + t] = n3;
}

// ---------------------------------------------------------------- prep
// grid (ceil(nd/64), nprob, ncol): opn, emn [col][p][nd]; grid [col][nd].
__global__ __launch_bounds__(64) void prep_kernel(int nd, int nprob, int tau_grid, const double *__restrict__ depth,
                                                  const double *__restrict__ opem, double *__restrict__ opn,
                                                  double *__restrict__ emn, double *__restrict__ grid) {
  const int d = blockIdx.x * 64 + threadIdx.x, p = blockIdx.y, col = blockIdx.z;
  if (d >= nd) return;
  const double *dp = depth + (size_t)col * SNAPI_NLTE_NFIELD * nd;
  const double *b = opem + ((size_t)col * nprob + p) * 8 * nd;
  double o = b[d], e = b[nd + d];
  if (tau_grid) {
    o /= dp[SNAPI_NLTE_OPREF * nd + d];
    e /= dp[SNAPI_NLTE_OPREF * nd + d];
  }
  const size_t q = ((size_t)col * nprob + p) * nd + d;
  opn[q] = o;
  emn[q] = e;
  if (p == 0) grid[(size_t)col * nd + d] = dp[SNAPI_NLTE_GRID * nd + d];
}

// ---------------------------------------------------------------- coefficients
// grid (ceil(nd/64), nprob, ncol), 64 work-items, one depth each.
// Per-lane scratch in LDS, [item][64]: phi [ntr], pd [3][nla], level
// derivatives [4][nmap]; the local terms accumulate in their slots of sums,
// which only this work-item touches.
struct CoefOut {
  double *rrec;  // [col][p][nd][nrow]
  double *crec;  // [col][p][ncolrec][nd]
  double *sums;  // [col][p][nsum][nd]
};

__global__ __launch_bounds__(64) void coef_kernel(NAtom at, int nd, int nlam, const double *__restrict__ lam,
                                                  const double *__restrict__ lamw, const double *__restrict__ sigma,
                                                  const double *__restrict__ depth, const double *__restrict__ pops,
                                                  const double *__restrict__ damp, const double *__restrict__ vr,
                                                  const double *__restrict__ opem, const double *__restrict__ Iall,
                                                  const double *__restrict__ norm, const double *__restrict__ normd,
                                                  CoefOut out) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const int d = blockIdx.x * 64 + lane, p = blockIdx.y, col = blockIdx.z;
  const int nprob = at.nray * nlam;
  const int tp = p / nlam, l = p - tp * nlam;
  const int NM = at.nmap, NT = at.ntr, NLA = at.nla;
  double *phi = lds;                      // [NT][64]
  double *pd = phi + NT * 64;             // [3][NLA][64]
  double *dop = pd + 3 * NLA * 64;        // [NM][64] each
  double *dem = dop + NM * 64;
  double *dopb = dem + NM * 64;
  double *demb = dopb + NM * 64;
#define LD(arr, i) arr[(i) * 64 + lane]
  if (d >= nd) return;
  const double *dp = depth + (size_t)col * SNAPI_NLTE_NFIELD * nd;
  const double T = dp[SNAPI_NLTE_T * nd + d], vt = dp[SNAPI_NLTE_VT * nd + d], ne = dp[SNAPI_NLTE_NE * nd + d];
  const double Nt = dp[SNAPI_NLTE_NT * nd + d];
  const double dneT = dp[SNAPI_NLTE_DNE_DT * nd + d], dneN = dp[SNAPI_NLTE_DNE_DNT * nd + d];
  const double opref = at.tau_grid ? dp[SNAPI_NLTE_OPREF * nd + d] : 1.0;
  double opd[3];
  for (int q = 0; q < 3; ++q) opd[q] = at.tau_grid ? dp[(SNAPI_NLTE_DOPREF_DT + q) * nd + d] : 0.0;
  const double *pp = pops + (size_t)col * NM * nd;
#define POP(i) pp[(size_t)(i) * nd + d]
  const double v = vr[((size_t)col * at.nray + tp) * nd + d];
  const double *dm = damp + (size_t)col * 2 * NT * nd;
  const double *b = opem + ((size_t)col * nprob + p) * 8 * nd;
  double opp = b[d], emm = b[nd + d];
  if (at.tau_grid) {  // normalised for the formal solution and back, as the reference hands them over
    opp /= dp[SNAPI_NLTE_OPREF * nd + d];
    emm /= dp[SNAPI_NLTE_OPREF * nd + d];
    opp *= dp[SNAPI_NLTE_OPREF * nd + d];
    emm *= dp[SNAPI_NLTE_OPREF * nd + d];
  }
  double pert_op[3], pert_em[3];
  for (int q = 0; q < 3; ++q) {
    pert_op[q] = b[(2 + q) * nd + d];
    pert_em[q] = b[(5 + q) * nd + d];
  }
  const double I = Iall[((size_t)col * nprob + p) * nd + d];
  const double lambda = lam[l], lambda_w = lamw[l], angular_weight = at.wray[tp];
  const double *nrm = norm + (((size_t)col * at.nray + tp) * nd + d) * NT;
  const double *nrmd = normd + ((size_t)col * at.nray + tp) * 3 * nd * NT;
#define NRMD(q, t) nrmd[((size_t)(q) * nd + d) * NT + (t)]
#define SIGT(t) sigma[(size_t)(t) * nlam + l]
#define SIGL(i) (at.bftr[i] >= 0 ? SIGT(at.bftr[i]) : 0.0)
  const size_t pbase = (size_t)col * nprob + p;
  double *sums = out.sums + pbase * nsum_rec(at) * nd;
  // 1. line profiles (atom::boundbound_op) and the mean-intensity increments
  for (int t = 0; t < NT; ++t) {
    double jm = 0.0, jn = 0.0;
    if (at.type[t] == SNAPI_TR_LINE) {
      const int up = at.up[t], lo = at.lo[t];
      const double lam0 = kH * kC / (at.ee[up] - at.ee[lo]);
      const double sf = (lam0 * lam0) / (4.0 * kPi * kC);
      const double dld = broad_dop(lam0, T, vt, at.mass);
      const double a = sf * (at.ar[t] + dm[(size_t)t * nd + d] * at.damping) / dld;
      const double x = (lam0 - lambda * (1.0 + v / kC)) / dld;
      const double ph = 1.0 / dld * fvoigt(x, a);
      LD(phi, t) = ph;
      const double line_energy = at.ee[up] - at.ee[lo];
      const double line_opacity = (POP(lo) * at.Blu[t] - POP(up) * at.Bul[t]) * ph * line_energy * 0.25 / kPi;
      const double line_factor = line_opacity / opp;
      jm = jn = lambda_w * angular_weight * I * ph * line_factor * 0.25 / kPi;
    } else {
      LD(phi, t) = 0.0;
      const double sg = SIGT(t);
      jm = lambda_w * angular_weight * sg * lambda / kH / kC * I;
      jn = lambda_w * angular_weight * sg * lambda / kH / kC * (I + 2.0 * kH * kC * kC / pow(lambda, 5.0)) *
           exp(-kH * kC / lambda / kK / T);
    }
    sums[(size_t)t * nd + d] = jm;
    sums[((size_t)NT + t) * nd + d] = jn;
  }
  // 2. profile derivatives of the assembled lines (compute_xa_der_scalar + fvoigtn)
  for (int j = 0; j < NLA; ++j) {
    const int t = at.la[j], up = at.up[t], lo = at.lo[t];
    const double lamc = fabs(kH * kC / (at.ee[up] - at.ee[lo]));
    const double dld = broad_dop(lamc, T, vt, at.mass);
    const double x = (lamc - lambda * (1.0 + v / kC)) / dld;
    const double dampcol = dm[(size_t)t * nd + d];
    const double damping = at.ar[t] + dampcol * at.damping;
    const double a = damping / dld * (lamc * lamc) / (4.0 * kPi * kC);
    const double dvd = sqrt(2.0 * kK * T / at.mass + vt * vt);
    const double ld0 = lamc / (kC * dvd) * kK / at.mass, ld2 = lamc / (kC * dvd) * vt;
    const double xd0 = -x / dld * ld0, xd2 = -x / dld * ld2;
    double ad0 = -a / dld * ld0;
    ad0 += dm[((size_t)NT + t) * nd + d] / dld * (lamc * lamc) / (4.0 * kPi * kC) * at.damping;
    const double delta_Nt = kDeltaNtFrac * Nt;
    const double delta_a = dampcol * kDeltaNtFrac / dld * (lamc * lamc) / (4.0 * kPi * kC);
    const double ad1 = delta_a / delta_Nt * at.damping;
    const double ad2 = -a / dld * ld2;
    double H, F, dH, dF;
    voigt_d<double>(x, a, H, F, dH, dF);
    LD(pd, j) = dH * xd0 / dld - dF * ad0 / dld - H / dld / dld * ld0;
    LD(pd, NLA + j) = dH * 0.0 / dld - dF * ad1 / dld - H / dld / dld * 0.0;
    LD(pd, 2 * NLA + j) = dH * xd2 / dld - dF * ad2 / dld - H / dld / dld * ld2;
  }
  // 3. explicit derivatives of the bound-free op and em (op/em_derivative_explicit)
  double bfop[2] = {0.0, 0.0}, bfem[2] = {0.0, 0.0};
  for (int i = 0; i < NM; ++i) {
    const int z = at.stage[i];
    if (z >= at.nstage - 1) continue;
    const double sigma_i = SIGL(i);
    if (!sigma_i) continue;
    const int cont = at.first[z + 1];
    const double gl = at.g[i], gcn = at.g[cont], nc = POP(cont);
    {
      const double delta_E = (at.ip[z] - at.ee[i] - kH * kC / lambda);
      const double cp = -nc * kSaha * gl / gcn * sigma_i;
      bfop[0] += cp * (dneT * pow(T, -1.5) * exp(delta_E / kK / T));
      bfop[0] += cp * ne * (-1.5) * pow(T, -2.5) * exp(delta_E / kK / T);
      bfop[0] += cp * ne * pow(T, -1.5) * exp(delta_E / kK / T) * (-delta_E / kK / T / T);
      bfop[1] += cp * (dneN * pow(T, -1.5) * exp(delta_E / kK / T));
    }
    {
      const double B_planck = planck_f(lambda, T);
      const double delta_E = (at.ip[z] - at.ee[i]);
      const double pop_mod = nc * ne * 2.07E-16 * pow(T, -1.5) * gl / gcn * exp(delta_E / kK / T);
      bfem[0] += sigma_i * (-pop_mod * exp(-kH * kC / lambda / kK / T) * kH * kC / lambda / kK / T / T) * B_planck;
      const double cp = nc * 2.07E-16 * gl / gcn * B_planck * sigma_i * (1.0 - exp(-kH * kC / lambda / kK / T));
      bfem[0] += cp * (dneT * pow(T, -1.5) * exp(delta_E / kK / T) + ne * (-1.5 * pow(T, -2.5)) * exp(delta_E / kK / T) +
                       ne * pow(T, -1.5) * exp(delta_E / kK / T) * -delta_E / kK / T / T);
      bfem[0] += pop_mod * sigma_i * (1.0 - exp(-kH * kC / lambda / kK / T)) * planck_fd(lambda, T);
      bfem[1] += cp * (dneN * pow(T, -1.5) * exp(delta_E / kK / T));
    }
  }
  // 4. explicit derivatives of the bound-bound op and em (bb_op/em_derivative_explicit):
  // only the assembled lines have profile derivatives
  double bbop[3] = {0.0, 0.0, 0.0}, bbem[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < NLA; ++j) {
    const int t = at.la[j], up = at.up[t], lo = at.lo[t];
    const double lm = kH * kC / (at.ee[up] - at.ee[lo]);
    const double X = (POP(lo) * at.Blu[t] - POP(up) * at.Bul[t]);
    for (int q = 0; q < 3; ++q) {
      bbop[q] += LD(pd, q * NLA + j) * X * kH * kC / lm / 4.0 / kPi;
      bbem[q] += LD(pd, q * NLA + j) * POP(up) * at.A[t] * kH * kC / lm / 4.0 / kPi;
    }
  }
  // 5. derivatives of op and em to each level population (atom_responses.cc:1779-1925)
  for (int i = 0; i < NM; ++i) {
    const int z = at.stage[i], li = at.lidx[i];
    double der = SIGL(i), dere = 0.0;
    if (z > 0 && li == 0)
      for (int k = 0; k < at.nl[z - 1]; ++k) {
        const int j = at.first[z - 1] + k;
        const double sj = SIGL(j);
        const double dE1 = at.ip[z - 1] - at.ee[j] - kH * kC / lambda;
        der -= sj * ne * kSaha * pow(T, -1.5) * at.g[j] / at.g[i] * exp(dE1 / kK / T);
        const double dE2 = at.ip[z - 1] - at.ee[j];
        dere += sj * ne * kSaha * pow(T, -1.5) * at.g[j] / at.g[i] * exp(dE2 / kK / T) *
                (1.0 - exp(-kH * kC / lambda / kK / T)) * planck_f(lambda, T);
      }
    LD(dop, i) = der;
    LD(dem, i) = dere;
    double db = 0.0, deb = 0.0;
    for (int k = li + 1; k < at.nl[z]; ++k) {
      const int j = at.first[z] + k, t = at.tmap[i][j];
      if (t >= 0) db += at.Blu[t] * kH * kC / lambda * 0.25 / kPi * LD(phi, t);
    }
    for (int k = 0; k < li; ++k) {
      const int j = at.first[z] + k, t = at.tmap[i][j];
      if (t >= 0) {
        db -= at.Bul[t] * kH * kC / lambda * 0.25 / kPi * LD(phi, t);
        deb += at.A[t] * kH * kC / lambda * 0.25 / kPi * LD(phi, t);
      }
    }
    LD(dopb, i) = db;
    LD(demb, i) = deb;
  }
  // 6. the depth as the perturbed row l: rate weights and local terms
  const int nloc = NM * NM + 3 * NM;
  double *loc = sums + (size_t)2 * NT * nd + d;
  for (int k = 0; k < nloc; ++k) loc[(size_t)k * nd] = 0.0;
#define LM(a, c) loc[(size_t)((a) * NM + (c)) * nd]
#define LB(q, a) loc[(size_t)(NM * NM + (q) * NM + (a)) * nd]
  double *rr = out.rrec + (pbase * nd + d) * nrow_rec(at);
  for (int k = 0; k < 2 * NM; ++k) rr[NLA + k] = 0.0;
  for (int j = 0; j < NLA; ++j) {
    const int t = at.la[j], ri = at.up[t], rii = at.lo[t];
    const int i_raw = at.lidx[ri], ii_raw = at.lidx[rii];
    const double Blu = at.Blu[t], Bul = at.Bul[t], ph = LD(phi, t), nr = nrm[t];
    const double line_energy = at.ee[ri] - at.ee[rii];
    const double line_opacity = (POP(rii) * Blu - POP(ri) * Bul) * line_energy * 0.25 / kPi;
    const double line_factor = line_opacity * ph / opp;
    const double J_chunk = (Blu * POP(rii) - Bul * POP(ri)) * I * lambda_w * angular_weight * ph / nr * line_factor *
                           0.25 / kPi;
    const double dR_dI = (Blu * POP(rii) - Bul * POP(ri)) * lambda_w * angular_weight * ph / nr * line_factor * 0.25 / kPi;
    rr[j] = dR_dI;
    rr[NLA + ri] += dR_dI;
    rr[NLA + rii] -= dR_dI;
    for (int q = 0; q < 3; ++q) {
      const double derivative = LD(pd, q * NLA + j);
      double dpart = 0.0;
      dpart += derivative / ph;
      dpart -= NRMD(q, t) / nr;
      dpart += (1.0 - line_factor) / line_factor * line_opacity / opp * derivative;
      if (q < 2) dpart -= (pert_op[q] + bfop[q]) / opp;
      LB(q, ri) -= J_chunk * dpart;
      LB(q, rii) += J_chunk * dpart;
    }
    const double dil = -((1.0 - line_factor) / opp * Bul * I * line_energy * 0.25 / kPi * ph / line_factor);
    const double diil = (1.0 - line_factor) / opp * Blu * I * line_energy * 0.25 / kPi * ph / line_factor;
    LM(ri, rii) += diil * dR_dI;
    LM(ri, ri) += dil * dR_dI;
    LM(rii, ri) -= dil * dR_dI;
    LM(rii, rii) -= diil * dR_dI;
    for (int k = 0; k < NM; ++k) {
      LM(i_raw, k) -= J_chunk * LD(dop, k) / opp;
      LM(ii_raw, k) += J_chunk * LD(dop, k) / opp;
    }
  }
  for (int t = 0; t < NT; ++t) {
    if (at.type[t] != SNAPI_TR_BF || !(lambda <= at.lcrit[t])) continue;
    const int rii = at.lo[t], rc = at.up[t], z = at.stage[rii];
    const double sg = SIGL(rii);
    const double gi = at.g[rii], gcn = at.g[rc];
    const double delta_E = at.ip[z] - at.ee[rii];
    const double dRic_dI = sg * lambda / kH / kC;
    const double ex = exp((delta_E - kH * kC / lambda) / kK / T);
    const double dRci_dI = sg * lambda / kH / kC * ex * gi / gcn * pow(T, -1.5) * ne * kSaha;
    const double cp = (I + 2.0 * kH * kC * kC / pow(lambda, 5.0)) * sg * lambda / kH / kC * gi / gcn * kSaha;
    double dq0 = ex * pow(T, -1.5) * dneT;
    dq0 += ex * (-1.5) * pow(T, -2.5) * ne;
    dq0 += ex * (-delta_E + kH * kC / lambda) / kK / T / T * pow(T, -1.5) * ne;
    dq0 *= cp;
    const double dq1 = cp * ex * pow(T, -1.5) * dneN;
    const double nc = POP(rc), ni = POP(rii);
    LB(0, rc) += nc * dq0 * angular_weight * lambda_w;
    LB(0, rii) -= nc * dq0 * angular_weight * lambda_w;
    LB(1, rc) += nc * dq1 * angular_weight * lambda_w;
    LB(1, rii) -= nc * dq1 * angular_weight * lambda_w;
    const double gbf = (ni * dRic_dI - nc * dRci_dI) * lambda_w * angular_weight;
    rr[NLA + NM + rc] += gbf;
    rr[NLA + NM + rii] -= gbf;
  }
  // 7. the depth as the perturbing depth ll
  double *cr = out.crec + pbase * ncol_rec(at) * nd + d;
#define CR(f) cr[(size_t)(f) * nd]
  for (int i = 0; i < NM; ++i) {
    CR(i) = LD(dop, i) / opref;
    CR(NM + i) = LD(dem, i) / opref;
    CR(2 * NM + i) = (LD(dopb, i) + LD(dop, i)) / opref;
    CR(3 * NM + i) = (LD(demb, i) + LD(dem, i)) / opref;
  }
  const int s0 = 4 * NM;
  for (int q = 0; q < 3; ++q) {
    const double so = (q < 2 ? bfop[q] + pert_op[q] : pert_op[2]);
    const double se = (q < 2 ? bfem[q] + pert_em[q] : pert_em[2]);
    CR(s0 + q) = so / opref - opd[q] * opp / opref / opref;
    CR(s0 + 3 + q) = se / opref - opd[q] * emm / opref / opref;
  }
  for (int q = 0; q < 2; ++q) {
    CR(s0 + 6 + q) = (pert_op[q] + bbop[q] + bfop[q]) / opref - opd[q] * opp / opref / opref;
    CR(s0 + 8 + q) = (pert_em[q] + bbem[q] + bfem[q]) / opref - opd[q] * emm / opref / opref;
  }
  const int l0 = s0 + 10;
  for (int j = 0; j < NLA; ++j) {
    const int t = at.la[j], up = at.up[t], lo = at.lo[t];
    const double line_energy = at.ee[up] - at.ee[lo];
    const double lm = kH * kC / line_energy;
    CR(l0 + 7 * j) = line_energy * 0.25 / kPi * LD(phi, t) / opref;
    const double dopr = (POP(lo) * at.Blu[t] - POP(up) * at.Bul[t]) * kH * kC / lm * 0.25 / kPi;
    const double demr = POP(up) * at.A[t] * kH * kC / lm * 0.25 / kPi;
    for (int q = 0; q < 3; ++q) {
      CR(l0 + 7 * j + 1 + q) = LD(pd, q * NLA + j) * dopr / opref;
      CR(l0 + 7 * j + 4 + q) = LD(pd, q * NLA + j) * demr / opref;
    }
  }
#undef CR
#undef LM
#undef LB
#undef LD
#undef POP
#undef NRMD
#undef SIGT
#undef SIGL
}

// ---------------------------------------------------------------- sums over problems
// grid (ceil(nd/64), nsum, ncol): out [col][nsum][nd].
__global__ __launch_bounds__(64) void sum_kernel(int nd, int nprob, int nsum, const double *__restrict__ sums,
                                                 double *__restrict__ out) {
  const int d = blockIdx.x * 64 + threadIdx.x, k = blockIdx.y, col = blockIdx.z;
  if (d >= nd) return;
  double s = 0.0;
  const double *src = sums + (size_t)col * nprob * nsum * nd + (size_t)k * nd + d;
  for (int p = 0; p < nprob; ++p) s += src[(size_t)p * nsum * nd];
  out[((size_t)col * nsum + k) * nd + d] = s;
}

// ---------------------------------------------------------------- assembly
// grid (ceil(nd/kRows), ncol), block (64, kRows): lane ll, wavefront l.  The
// column record of one problem ([ncolrec][nd], the same for every row l) is
// staged in LDS once per workgroup; each work-item prefetches its share of
// the next problem's record into registers while it works on the current
// one.  line0: first assembled line of this pass (NL lines per pass); dense:
// this pass also forms the level-vector terms and writes the blocks (the
// first pass), later passes add their lines' terms to what is there.
constexpr int kPre = 9;  // prefetched doubles per work-item (records of up to 9 * 64 * kRows doubles)

template <int NM, int NL>
__global__ __launch_bounds__(64 * kRows) void assemble_kernel(NAtom at, int nd, int nprob, int line0, int dense,
                                                              const double *__restrict__ Rop,
                                                              const double *__restrict__ Rem,
                                                              const double *__restrict__ rrec,
                                                              const double *__restrict__ crec, double *__restrict__ M,
                                                              double *__restrict__ beta) {
#pragma clang fp contract(fast)
  extern __shared__ double cbuf[];  // [ncolrec][nd]
  const int ll = threadIdx.x;
  const int tid = threadIdx.y * 64 + threadIdx.x;
  const int l = __builtin_amdgcn_readfirstlane(blockIdx.x * kRows + threadIdx.y);
  const int col = blockIdx.y;
  const bool row_ok = l < nd;  // no early return: the workgroup meets at barriers
  const int llc = ll < nd ? ll : nd - 1;
  const int lc = row_ok ? l : nd - 1;
  const int NLA = at.nla;
  const int nr = nrow_rec(at), nc = ncol_rec(at);
  const int ncd = nc * nd;
  const int nline = min(NL, NLA - line0);
  const double *src = crec + (size_t)col * nprob * ncd;
  double Mc[NM][NM], Bc[3][NM];
  double X[NL], Y[NL], Z[3][NL];
#pragma unroll
  for (int a = 0; a < NM; ++a) {
#pragma unroll
    for (int c = 0; c < NM; ++c) Mc[a][c] = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) Bc[q][a] = 0.0;
  }
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    X[j] = Y[j] = 0.0;
    Z[0][j] = Z[1][j] = Z[2][j] = 0.0;
  }
  double pre[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int i = tid + k * 64 * kRows;
    pre[k] = i < ncd ? src[i] : 0.0;
  }
  const int s0 = 4 * NM, l0 = s0 + 10;
  for (int p = 0; p < nprob; ++p) {
    __syncthreads();  // the previous problem's record is no longer read
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
      const int i = tid + k * 64 * kRows;
      if (i < ncd) cbuf[i] = pre[k];
    }
    for (int i = tid + kPre * 64 * kRows; i < ncd; i += 64 * kRows) cbuf[i] = src[(size_t)p * ncd + i];
    __syncthreads();
    if (p + 1 < nprob) {
#pragma unroll
      for (int k = 0; k < kPre; ++k) {
        const int i = tid + k * 64 * kRows;
        if (i < ncd) pre[k] = src[(size_t)(p + 1) * ncd + i];
      }
    }
    const size_t pb = (size_t)col * nprob + p;
    const double *R = rrec + (pb * nd + lc) * nr;
    const size_t ro_off = (pb * nd + lc) * nd + llc;
    const double ro = Rop[ro_off], re = Rem[ro_off];
    const double *C = cbuf + llc;
    if (dense) {
      double wl[NM], wb[NM];
#pragma unroll
      for (int c = 0; c < NM; ++c) {
        wl[c] = ro * C[c * nd] + re * C[(NM + c) * nd];
        wb[c] = ro * C[(2 * NM + c) * nd] + re * C[(3 * NM + c) * nd];
      }
      double tl[3], tb[2];
#pragma unroll
      for (int q = 0; q < 3; ++q) tl[q] = ro * C[(s0 + q) * nd] + re * C[(s0 + 3 + q) * nd];
#pragma unroll
      for (int q = 0; q < 2; ++q) tb[q] = ro * C[(s0 + 6 + q) * nd] + re * C[(s0 + 8 + q) * nd];
#pragma unroll
      for (int a = 0; a < NM; ++a) {
        const double gl = R[NLA + a], gb = R[NLA + NM + a];
#pragma unroll
        for (int c = 0; c < NM; ++c) Mc[a][c] += gl * wl[c] + gb * wb[c];
        Bc[0][a] -= gl * tl[0] + gb * tb[0];
        Bc[1][a] -= gl * tl[1] + gb * tb[1];
        Bc[2][a] -= gl * tl[2];
      }
    }
#pragma unroll
    for (int j = 0; j < NL; ++j) {
      if (j < nline) {
        const int jj = line0 + j;
        const double g = R[jj];
        const double *Cj = C + (l0 + 7 * jj) * nd;
        const double h = g * Cj[0];
        X[j] += h * ro;
        Y[j] += h * re;
#pragma unroll
        for (int q = 0; q < 3; ++q) Z[q][j] += g * (ro * Cj[(1 + q) * nd] + re * Cj[(4 + q) * nd]);
      }
    }
  }
  if (!row_ok || ll >= nd) return;
  const int n = nd * NM;
  double *Mrow = M + (size_t)col * n * n + (size_t)(l * NM) * n + ll * NM;  // row (l, a), column (ll, b)
  double *Bq = beta + (size_t)col * 3 * nd * n;                             // [q][ll][(l, a)]
  if (dense) {
#pragma unroll
    for (int a = 0; a < NM; ++a) {
#pragma unroll
      for (int c = 0; c < NM; ++c) Mrow[(size_t)a * n + c] = Mc[a][c];
#pragma unroll
      for (int q = 0; q < 3; ++q) Bq[((size_t)q * nd + ll) * n + l * NM + a] = Bc[q][a];
    }
  }
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    if (j >= nline) break;
    const int t = at.la[line0 + j], ri = at.up[t], rii = at.lo[t];
    const double dd = at.A[t] * Y[j] - at.Bul[t] * X[j];
    const double nn = at.Blu[t] * X[j];
    Mrow[(size_t)ri * n + ri] += dd;
    Mrow[(size_t)ri * n + rii] += nn;
    Mrow[(size_t)rii * n + ri] -= dd;
    Mrow[(size_t)rii * n + rii] -= nn;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      Bq[((size_t)q * nd + ll) * n + l * NM + ri] -= Z[q][j];
      Bq[((size_t)q * nd + ll) * n + l * NM + rii] += Z[q][j];
    }
  }
}

// ---------------------------------------------------------------- rate part
// grid (ceil(nd/64), ncol): one work-item per depth l.  tot [col][nsum][nd]
// holds the mean intensities and the local terms summed over problems.
__global__ __launch_bounds__(64) void rate_kernel(NAtom at, int nd, const double *__restrict__ depth,
                                                  const double *__restrict__ pops, const double *__restrict__ coll,
                                                  const double *__restrict__ norm_last_ray, int norm_stride,
                                                  const double *__restrict__ tot, double *__restrict__ M,
                                                  double *__restrict__ beta) {
  const int l = blockIdx.x * 64 + threadIdx.x, col = blockIdx.y;
  if (l >= nd) return;
  const int NM = at.nmap, NT = at.ntr, n = nd * NM;
  double *Mc = M + (size_t)col * n * n;
  double *B = beta + (size_t)col * 3 * nd * n;
  const double *tt = tot + (size_t)col * nsum_rec(at) * nd;
  const double *dp = depth + (size_t)col * SNAPI_NLTE_NFIELD * nd;
  const double *pp = pops + (size_t)col * NM * nd;
  const double *cl = coll + (size_t)col * 3 * NM * NM * nd;
  const double *nrm = norm_last_ray + (size_t)col * norm_stride + (size_t)l * NT;
#define MM(r, c) Mc[(size_t)(r) * n + (c)]
#define BETA(q, k, r) B[((size_t)(q) * nd + (k)) * n + (r)]
#define COLL(q, i, j) cl[(((size_t)(q) * NM + (i)) * NM + (j)) * nd + l]
#define POP(i) pp[(size_t)(i) * nd + l]
#define TOT(k) tt[(size_t)(k) * nd + l]
  // local terms of the problems
  for (int a = 0; a < NM; ++a) {
    for (int c = 0; c < NM; ++c) MM(l * NM + a, l * NM + c) += TOT(2 * NT + a * NM + c);
    for (int q = 0; q < 3; ++q) BETA(q, l, l * NM + a) += TOT(2 * NT + NM * NM + q * NM + a);
  }
  // collisional terms of the right-hand sides, then the conservation rows
  for (int q = 0; q < 2; ++q) {
    for (int i = 0; i < NM; ++i) {
      double s = BETA(q, l, l * NM + i);
      for (int ii = 0; ii < NM; ++ii) s += (POP(i) * COLL(1 + q, i, ii) - POP(ii) * COLL(1 + q, ii, i));
      BETA(q, l, l * NM + i) = s;
    }
    for (int k = 0; k < nd; ++k) BETA(q, k, l * NM + NM - 1) = (k == l) ? dp[(SNAPI_NLTE_DNACT_DT + q) * nd + l] : 0.0;
  }
  // rates of the diagonal block
  const double Temp = dp[SNAPI_NLTE_T * nd + l], ne = dp[SNAPI_NLTE_NE * nd + l];
  for (int i0 = 0; i0 < NM - 1; ++i0) {
    const int z = at.stage[i0];
    for (int k = 0; k < at.nl[z]; ++k) {
      const int j = at.first[z] + k, t = at.tmap[i0][j];
      const double JJ = (t >= 0) ? TOT(t) / nrm[t] : -1.0;
      // R_ij (atom_rad.cc:1727-1730) with A and B of the pair (zero without a line)
      double Aij = 0.0, Bij = 0.0, Aji = 0.0, Bji = 0.0;
      if (t >= 0) {
        if (at.up[t] == i0) {
          Aij = at.A[t];
          Bij = at.Bul[t];
          Bji = at.Blu[t];
        } else {
          Aji = at.A[t];
          Bij = at.Blu[t];
          Bji = at.Bul[t];
        }
      }
      double R = 1.0 * (Aij + Bij * JJ);
      double C = 1.0 * COLL(0, i0, j);
      MM(l * NM + i0, l * NM + i0) -= (R + C);
      R = 1.0 * (Aji + Bji * JJ);
      C = 1.0 * COLL(0, j, i0);
      MM(l * NM + i0, l * NM + j) += (R + C);
    }
    const int cont = at.first[z + 1], t = at.bftr[i0];
    double JJ = (t >= 0) ? TOT(t) : -1.0;
    double R = 1.0 * JJ;
    double C = COLL(0, i0, cont);
    MM(l * NM + i0, l * NM + i0) -= (R + C);
    MM(l * NM + cont, l * NM + i0) += (R + C);
    JJ = (t >= 0) ? TOT(NT + t) : -1.0;
    const double dE = at.ip[z] - at.ee[i0];
    R = 1.0 * (JJ * at.g[i0] / at.g[cont] * exp(dE / kK / Temp) * pow(Temp, -1.5) * ne * kSaha);
    C = COLL(0, cont, i0);
    MM(l * NM + i0, l * NM + cont) += (R + C);
    MM(l * NM + cont, l * NM + cont) -= (R + C);
  }
  for (int ll = 0; ll < nd; ++ll)
    for (int ii = 0; ii < NM; ++ii) MM(l * NM + NM - 1, ll * NM + ii) = (l == ll) ? 1.0 : 0.0;
#undef MM
#undef BETA
#undef COLL
#undef POP
#undef TOT
}

// X [col][3 nd][n] (solution per right-hand side) -> resp [col][3][n][nd].
__global__ __launch_bounds__(256) void resp_transpose_kernel(int nd, int n, const double *__restrict__ X,
                                                             double *__restrict__ resp) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t per = (size_t)3 * nd * n;
  const size_t col = blockIdx.y;
  if (i >= per) return;
  const int q = (int)(i / ((size_t)nd * n));
  const int rem = (int)(i - (size_t)q * nd * n);
  const int k = rem / n, r = rem - k * n;
  resp[col * per + ((size_t)q * n + r) * nd + k] = X[col * per + i];
}

// ---------------------------------------------------------------- host side
int build_atom(const snapi_nlte_atom *a, int nray, const double *wray, int tau_grid, NAtom &v) {
  std::memset(&v, 0, sizeof(v));
  if (!a || !a->level_stage || !a->level_index || !a->level_energy || !a->level_g || !a->stage_ip || !a->tr_type ||
      !a->tr_lower || !a->tr_upper || !a->tr_A || !a->tr_Blu || !a->tr_Bul || !a->tr_damp_rad || !a->tr_lambda_crit)
    return SNAPI_ERR_ARG;
  if (a->nmap < 2 || a->nmap > kMaxLev || a->ntr < 0 || a->ntr > kMaxTr || a->nstage < 2 || a->nstage > kMaxLev ||
      nray < 1 || nray > kMaxRay || !(a->mass > 0.0))
    return SNAPI_ERR_UNSUPPORTED;
  v.nmap = a->nmap;
  v.ntr = a->ntr;
  v.nstage = a->nstage;
  v.damping = a->damping;
  v.mass = a->mass;
  v.nray = nray;
  v.tau_grid = tau_grid ? 1 : 0;
  for (int r = 0; r < nray; ++r) v.wray[r] = wray[r];
  for (int z = 0; z < kMaxLev; ++z) v.first[z] = -1;
  for (int i = 0; i < a->nmap; ++i) {
    const int z = a->level_stage[i];
    if (z < 0 || z >= a->nstage || (i > 0 && z < a->level_stage[i - 1])) return SNAPI_ERR_UNSUPPORTED;
    if (v.first[z] < 0) v.first[z] = i;
    if (a->level_index[i] != i - v.first[z]) return SNAPI_ERR_UNSUPPORTED;
    v.nl[z]++;
    v.stage[i] = z;
    v.lidx[i] = a->level_index[i];
    v.ee[i] = a->level_energy[i];
    v.g[i] = a->level_g[i];
    v.bftr[i] = -1;
  }
  // the top stage holds one level, the last; every stage from the first
  // level's up has levels
  if (v.nl[a->nstage - 1] != 1 || v.first[a->nstage - 1] != a->nmap - 1) return SNAPI_ERR_UNSUPPORTED;
  for (int z = v.stage[0]; z < a->nstage; ++z)
    if (v.nl[z] < 1) return SNAPI_ERR_UNSUPPORTED;
  for (int z = 0; z < a->nstage; ++z) v.ip[z] = a->stage_ip[z];
  for (int i = 0; i < kMaxLev; ++i)
    for (int j = 0; j < kMaxLev; ++j) v.tmap[i][j] = -1;
  for (int t = 0; t < a->ntr; ++t) {
    const int lo = a->tr_lower[t], up = a->tr_upper[t];
    if (lo < 0 || lo >= a->nmap || up < 0 || up >= a->nmap) return SNAPI_ERR_ARG;
    v.type[t] = a->tr_type[t];
    v.lo[t] = lo;
    v.up[t] = up;
    v.A[t] = a->tr_A[t];
    v.Blu[t] = a->tr_Blu[t];
    v.Bul[t] = a->tr_Bul[t];
    v.ar[t] = a->tr_damp_rad[t];
    v.lcrit[t] = a->tr_lambda_crit[t];
    if (v.type[t] == SNAPI_TR_LINE) {
      if (v.stage[lo] != v.stage[up] || !(v.ee[up] > v.ee[lo]) || v.tmap[lo][up] >= 0) return SNAPI_ERR_UNSUPPORTED;
      v.tmap[lo][up] = v.tmap[up][lo] = (signed char)t;
    } else if (v.type[t] == SNAPI_TR_BF) {
      const int z = v.stage[lo];
      if (z + 1 >= a->nstage || up != v.first[z + 1] || v.bftr[lo] >= 0) return SNAPI_ERR_UNSUPPORTED;
      v.bftr[lo] = t;
    } else {
      return SNAPI_ERR_ARG;
    }
  }
  // assembled lines in the order of the reference's pair loops: stage,
  // upper level, lower level
  std::vector<int> la;
  for (int z = 0; z < a->nstage; ++z)
    for (int iu = 1; iu < v.nl[z]; ++iu)
      for (int il = 0; il < iu; ++il) {
        const int t = v.tmap[v.first[z] + iu][v.first[z] + il];
        if (t >= 0 && v.A[t] > 1.0) la.push_back(t);
      }
  v.nla = (int)la.size();
  for (int j = 0; j < v.nla; ++j) v.la[j] = la[j];
  return SNAPI_OK;
}

template <int NM, int NL>
int launch_assemble(const NAtom &at, int ncol, int nd, int nprob, const double *Rop, const double *Rem,
                    const double *rrec, const double *crec, double *M, double *beta, hipStream_t st) {
  const dim3 grid((nd + kRows - 1) / kRows, ncol), block(64, kRows);
  const size_t lds = (size_t)ncol_rec(at) * nd * sizeof(double);
  int line0 = 0, dense = 1;
  do {
    hipLaunchKernelGGL((assemble_kernel<NM, NL>), grid, block, lds, st, at, nd, nprob, line0, dense, Rop, Rem, rrec,
                       crec, M, beta);
    if (hipGetLastError() != hipSuccess) return SNAPI_ERR_DEVICE;
    line0 += NL;
    dense = 0;
  } while (line0 < at.nla);
  return SNAPI_OK;
}

template <int NM>
int launch_assemble_nm(const NAtom &at, int ncol, int nd, int nprob, const double *Rop, const double *Rem,
                       const double *rrec, const double *crec, double *M, double *beta, hipStream_t st) {
  // lines per pass: the assembled lines' count rounded up to an even number, at most 8
  if (at.nla <= 2) return launch_assemble<NM, 2>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
  if (at.nla <= 4) return launch_assemble<NM, 4>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
  if (at.nla <= 6) return launch_assemble<NM, 6>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
  return launch_assemble<NM, 8>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
}

int dispatch_assemble(const NAtom &at, int ncol, int nd, int nprob, const double *Rop, const double *Rem,
                      const double *rrec, const double *crec, double *M, double *beta, hipStream_t st) {
  if ((size_t)ncol_rec(at) * nd * sizeof(double) > 64 * 1024) return SNAPI_ERR_UNSUPPORTED;
  switch (at.nmap) {
    case 2: return launch_assemble_nm<2>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
    case 3: return launch_assemble_nm<3>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
    case 4: return launch_assemble_nm<4>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
    case 5: return launch_assemble_nm<5>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
    case 6: return launch_assemble_nm<6>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
    case 7: return launch_assemble_nm<7>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
    case 8: return launch_assemble_nm<8>(at, ncol, nd, nprob, Rop, Rem, rrec, crec, M, beta, st);
  }
  return SNAPI_ERR_UNSUPPORTED;
}

// Device bytes per column of the work area.
size_t per_column_bytes(const NAtom &at, int nd, int nprob) {
  const size_t P = (size_t)nprob * nd;
  size_t b = 0;
  b += 4 * P;                                 // opn, emn, I, L
  b += 2 * P * nd;                            // Rop, Rem
  b += P * (nrow_rec(at) + ncol_rec(at) + nsum_rec(at));
  b += (size_t)at.nray * nd * at.ntr * 4;     // norm, normd
  b += (size_t)nsum_rec(at) * nd + nd;        // sums, grid
  return b * sizeof(double);
}

size_t solve_bytes(int nd, int nmap) {
  const size_t n = (size_t)nd * nmap;
  return (2 * n * n + 3 * (size_t)nd * n) * sizeof(double);  // LU, LT, X
}

#define NCHK(x)              \
  do {                       \
    if ((x) != hipSuccess) { \
      rc = SNAPI_ERR_DEVICE; \
      goto done;             \
    }                        \
  } while (0)

int system_dev(const NAtom &at, int ncol, int nd, int nlam, double boundary, const double *ray_theta,
               const double *lam, const double *lamw, const double *sigma, const double *depth, const double *pops,
               const double *damp, const double *coll, const double *vr, const double *opem, double *M, double *beta,
               double *resp, hipStream_t st) {
  const int nprob = at.nray * nlam;
  const int n = nd * at.nmap;
  const int NT = at.ntr, nr = nrow_rec(at), nc = ncol_rec(at), ns = nsum_rec(at);
  const size_t per = per_column_bytes(at, nd, nprob) + (resp ? solve_bytes(nd, at.nmap) : 0);
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return SNAPI_ERR_DEVICE;
  const size_t budget = freeb / 2;
  if (per > budget) return SNAPI_ERR_CAPACITY;
  int chunk = (int)std::min<size_t>((size_t)ncol, std::max<size_t>(1, budget / per));
  // SNAPI_NLTE_COLUMNS_PER_PASS caps the columns of one pass (a smaller work area)
  if (const char *e = std::getenv("SNAPI_NLTE_COLUMNS_PER_PASS")) {
    const int cap = std::atoi(e);
    if (cap > 0 && cap < chunk) chunk = cap;
  }
  // LDS of the coefficient kernel
  const size_t lds = (size_t)64 * sizeof(double) * (NT + 3 * at.nla + 4 * at.nmap);
  if (lds > 64 * 1024) return SNAPI_ERR_UNSUPPORTED;
  int rc = SNAPI_OK;
  double *w = nullptr;
  std::vector<double> th(nprob);
  for (int r = 0; r < at.nray; ++r)
    for (int l = 0; l < nlam; ++l) th[(size_t)r * nlam + l] = ray_theta[r];
  const size_t P = (size_t)chunk * nprob * nd;
  double *opn, *emn, *I, *L, *Ro, *Re, *rrec, *crec, *sums, *nrm, *nrmd, *tot, *grid, *th_d, *LU = nullptr, *Y = nullptr,
                                                                                           *X = nullptr;
  size_t total = (4 * P + 2 * P * nd + P * (nr + nc + ns) + (size_t)chunk * at.nray * nd * NT * 4 +
                  (size_t)chunk * (ns * nd + nd) + nprob) *
                 sizeof(double);
  if (resp) total += (size_t)chunk * solve_bytes(nd, at.nmap);
  NCHK(hipMalloc((void **)&w, total));
  {
    double *q = w;
    opn = q, q += P;
    emn = q, q += P;
    I = q, q += P;
    L = q, q += P;
    Ro = q, q += P * nd;
    Re = q, q += P * nd;
    rrec = q, q += P * nr;
    crec = q, q += P * nc;
    sums = q, q += P * ns;
    nrm = q, q += (size_t)chunk * at.nray * nd * NT;
    nrmd = q, q += (size_t)chunk * at.nray * nd * NT * 3;
    tot = q, q += (size_t)chunk * ns * nd;
    grid = q, q += (size_t)chunk * nd;
    th_d = q, q += nprob;
    if (resp) {
      LU = q, q += (size_t)chunk * n * n;
      Y = q, q += (size_t)chunk * n * n;
      X = q, q += (size_t)chunk * 3 * nd * n;
    }
  }
  NCHK(hipMemcpyAsync(th_d, th.data(), sizeof(double) * nprob, hipMemcpyHostToDevice, st));
  for (int c0 = 0; c0 < ncol; c0 += chunk) {
    const int cc = std::min(chunk, ncol - c0);
    const double *dp = depth + (size_t)c0 * SNAPI_NLTE_NFIELD * nd;
    const double *pp = pops + (size_t)c0 * at.nmap * nd;
    const double *dm = damp + (size_t)c0 * 2 * NT * nd;
    const double *cl = coll + (size_t)c0 * 3 * at.nmap * at.nmap * nd;
    const double *vv = vr + (size_t)c0 * at.nray * nd;
    const double *oe = opem + (size_t)c0 * nprob * 8 * nd;
    double *Mc = M + (size_t)c0 * n * n;
    double *Bc = beta + (size_t)c0 * 3 * nd * n;
    const unsigned gx = (unsigned)((nd + 63) / 64);
    if (NT > 0) {
      hipLaunchKernelGGL(norm_kernel, dim3(gx, NT, cc * at.nray), dim3(64), 0, st, at, nd, nlam, lam, lamw, dp, dm, vv,
                         nrm, nrmd);
      NCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(prep_kernel, dim3(gx, nprob, cc), dim3(64), 0, st, nd, nprob, at.tau_grid, dp, oe, opn, emn,
                       grid);
    NCHK(hipGetLastError());
    rc = launch_formal_resp(cc, nprob, nd, grid, opn, emn, th_d, boundary, I, L, Ro, Re, st);
    if (rc) goto done;
    hipLaunchKernelGGL(coef_kernel, dim3(gx, nprob, cc), dim3(64), lds, st, at, nd, nlam, lam, lamw, sigma, dp, pp,
                       dm, vv, oe, I, nrm, nrmd, CoefOut{rrec, crec, sums});
    NCHK(hipGetLastError());
    hipLaunchKernelGGL(sum_kernel, dim3(gx, ns, cc), dim3(64), 0, st, nd, nprob, ns, sums, tot);
    NCHK(hipGetLastError());
    rc = dispatch_assemble(at, cc, nd, nprob, Ro, Re, rrec, crec, Mc, Bc, st);
    if (rc) goto done;
    hipLaunchKernelGGL(rate_kernel, dim3(gx, cc), dim3(64), 0, st, at, nd, dp, pp, cl,
                       nrm + (size_t)(at.nray - 1) * nd * NT, at.nray * nd * NT, tot, Mc, Bc);
    NCHK(hipGetLastError());
    if (resp) {
      rc = launch_response_solve(cc, n, 3 * nd, Mc, Bc, LU, Y, X, st);
      if (rc) goto done;
      const size_t per_col = (size_t)3 * nd * n;
      hipLaunchKernelGGL(resp_transpose_kernel, dim3((unsigned)((per_col + 255) / 256), cc), dim3(256), 0, st, nd, n,
                         X, resp + (size_t)c0 * per_col);
      NCHK(hipGetLastError());
    }
  }
  NCHK(hipStreamSynchronize(st));
done:
  if (w) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(w);
  }
  return rc;
}

int check_sizes(int ncol, int nd, int nray, int nlam) {
  if (ncol <= 0 || nd <= 0 || nray <= 0 || nlam <= 0) return SNAPI_ERR_ARG;
  if (nd < 2 || nd > 256 || nray > kMaxRay) return SNAPI_ERR_UNSUPPORTED;
  if ((long)nray * nlam > 65535) return SNAPI_ERR_UNSUPPORTED;  // problems on one grid axis
  return SNAPI_OK;
}

}  // namespace

extern "C" int snapi_nlte_response_system_dev(const snapi_nlte_atom *atom, int ncol, int nd, int nray, int nlam,
                                              int tau_grid, double boundary, const double *ray_theta,
                                              const double *ray_weight, const double *lambda_dev,
                                              const double *lambda_weight_dev, const double *sigma_dev,
                                              const double *depth_dev, const double *pops_dev, const double *damp_dev,
                                              const double *coll_dev, const double *vr_dev, const double *opem_dev,
                                              double *M_dev, double *beta_dev, double *resp_dev, void *stream) {
  int rc = check_sizes(ncol, nd, nray, nlam);
  if (rc) return rc;
  if (!ray_theta || !ray_weight || !lambda_dev || !lambda_weight_dev || !sigma_dev || !depth_dev || !pops_dev ||
      !damp_dev || !coll_dev || !vr_dev || !opem_dev || !M_dev || !beta_dev)
    return SNAPI_ERR_ARG;
  NAtom at;
  rc = build_atom(atom, nray, ray_weight, tau_grid, at);
  if (rc) return rc;
  return system_dev(at, ncol, nd, nlam, boundary, ray_theta, lambda_dev, lambda_weight_dev, sigma_dev, depth_dev,
                    pops_dev, damp_dev, coll_dev, vr_dev, opem_dev, M_dev, beta_dev, resp_dev, (hipStream_t)stream);
}

extern "C" int snapi_nlte_response_system(const snapi_nlte_atom *atom, int ncol, int nd, int nray, int nlam,
                                          int tau_grid, double boundary, const double *ray_theta,
                                          const double *ray_weight, const double *lambda, const double *lambda_weight,
                                          const double *sigma, const double *depth, const double *pops,
                                          const double *damp, const double *coll, const double *vr, const double *opem,
                                          double *M, double *beta, double *resp) {
  int rc = check_sizes(ncol, nd, nray, nlam);
  if (rc) return rc;
  if (!ray_theta || !ray_weight || !lambda || !lambda_weight || !sigma || !depth || !pops || !damp || !coll || !vr ||
      !opem || !M || !beta)
    return SNAPI_ERR_ARG;
  NAtom at;
  rc = build_atom(atom, nray, ray_weight, tau_grid, at);
  if (rc) return rc;
  const int NM = at.nmap, NT = at.ntr, n = nd * NM;
  const size_t sz[11] = {(size_t)nlam,
                         (size_t)nlam,
                         (size_t)(NT > 0 ? NT : 1) * nlam,
                         (size_t)ncol * SNAPI_NLTE_NFIELD * nd,
                         (size_t)ncol * NM * nd,
                         (size_t)ncol * 2 * (NT > 0 ? NT : 1) * nd,
                         (size_t)ncol * 3 * NM * NM * nd,
                         (size_t)ncol * nray * nd,
                         (size_t)ncol * nray * nlam * 8 * nd,
                         (size_t)ncol * n * n,
                         (size_t)ncol * 3 * nd * n};
  const double *src[9] = {lambda, lambda_weight, sigma, depth, pops, damp, coll, vr, opem};
  double *dv[12] = {nullptr};
  for (int i = 0; i < 11; ++i) NCHK(hipMalloc((void **)&dv[i], sz[i] * sizeof(double)));
  if (resp) NCHK(hipMalloc((void **)&dv[11], (size_t)ncol * 3 * n * nd * sizeof(double)));
  for (int i = 0; i < 9; ++i)
    if (i != 2 || NT > 0) NCHK(hipMemcpy(dv[i], src[i], sz[i] * sizeof(double), hipMemcpyHostToDevice));
  rc = system_dev(at, ncol, nd, nlam, boundary, ray_theta, dv[0], dv[1], dv[2], dv[3], dv[4], dv[5], dv[6], dv[7],
                  dv[8], dv[9], dv[10], dv[11], nullptr);
  if (rc) goto done;
  NCHK(hipMemcpy(M, dv[9], sz[9] * sizeof(double), hipMemcpyDeviceToHost));
  NCHK(hipMemcpy(beta, dv[10], sz[10] * sizeof(double), hipMemcpyDeviceToHost));
  if (resp) NCHK(hipMemcpy(resp, dv[11], (size_t)ncol * 3 * n * nd * sizeof(double), hipMemcpyDeviceToHost));
done:
  for (int i = 0; i < 12; ++i)
    if (dv[i]) (void)hipFree(dv[i]);
  return rc;
}
