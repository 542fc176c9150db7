// ALI (approximate Lambda iteration) for one NLTE atom on gfx950.
//
// Restates atmosphere::nltepops (atmos/atmos_pop.cc:67-240):
//   - accumulation: for every ray and wavelength, the scalar opacity and
//     emissivity (the population-independent background given, plus the
//     atom's bound-free terms atom_rad.cc:312-437 and Voigt lines
//     atom_rad.cc:439-537 with the current populations), the scalar
//     short-characteristics formal solution with the diagonal Lambda*
//     (plpar/atmos_ppbez_rts.cc:21-219) and the sums of atom::add
//     (atmos/atmol/atom/atom.cc:1865-1927);
//   - new populations: atom::newpops (atom_pop.cc:27-135), rate matrix with
//     the local approximate operator (R_ij_local_ALO, atom_rad.cc:1732-1757)
//     or plain Lambda iteration (R_ij), bound-free R_i_cont / R_cont_i
//     (:1842-1854), conservation row for the most populated level, the
//     pivot-free Crout solve (cmlib/mathtools.cc:487-514), in double;
//   - iteration: Lambda-iteration fallback for negative populations, stop
//     at a relative change below tol (atmos_pop.cc:198-214).
// One workgroup per column, work-items over wavelengths; the depth
// recurrence of every wavelength advances in lock-step so that the sums of
// each depth are reduced across the workgroup (wavefront shuffles, then
// LDS) and added to the column's accumulators by the workgroup that owns the
// column.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "autotune.h"
#include "rt_common.h"
#include "synth_kernel.h"

namespace snapi {

constexpr int kAliThreads = 256;
constexpr int kAliMaxLines = 16;  // lines whose profiles a work-item carries between depths
constexpr int kAliMaxBf = 16;
constexpr int kAliSlots = 64;  // reduced sums per depth: 3 per line + 2 per bound-free <= 64
constexpr double kSaha = 2.0706615382E-16;  // cmlib/const.h:24

// Atom constants as kernel arguments.
struct AliAtom {
  int nlevel, ntrans, nline, nbf;
  double mass;
  int line_t[kAliMaxLines], line_lo[kAliMaxLines], line_up[kAliMaxLines];
  double lam0[kAliMaxLines], aul[kAliMaxLines], blu[kAliMaxLines], bul[kAliMaxLines], energy[kAliMaxLines];
  int bf_t[kAliMaxBf], bf_lo[kAliMaxBf], bf_up[kAliMaxBf];
  double gratio[kAliMaxBf], de[kAliMaxBf];
  // rate matrix structure
  int stage[SNAPI_ALI_MAX_LEVELS];
  int bf_of[SNAPI_ALI_MAX_LEVELS];                           // bound-free slot of a level or -1
  signed char pair_line[SNAPI_ALI_MAX_LEVELS][SNAPI_ALI_MAX_LEVELS];  // line slot between two levels or -1
};

struct AliGridArgs {
  int nlam, nray, bg_per_ray, tau_grid;
  double boundary;
  double ct[16], wray[16];
};

struct AliCol {  // device pointers of the per-column inputs
  const double *cs, *vproj, *gamma, *bgop, *bgem, *pops;
  const double *lam, *lamw, *fudge, *sigma;
  double *acc;
  const int *active;
};

template <typename R> struct AliTile {
  R *grid, *opref, *tinv, *vr;          // per depth (vr: velocity along the current ray / c)
  R *idld, *a, *lof, *emf, *accf;       // per line and depth
  R *nlo, *e1;                          // per bound-free and depth
  static __host__ __device__ size_t bytes(int nd, int nline, int nbf) {
    return sizeof(R) * (size_t)nd * (4 + 5 * nline + 2 * nbf);
  }
  static __device__ AliTile carve(R *p, int nd, int nline, int nbf) {
    AliTile t;
    t.grid = p; p += nd;
    t.opref = p; p += nd;
    t.tinv = p; p += nd;
    t.vr = p; p += nd;
    t.idld = p; p += (size_t)nd * nline;
    t.a = p; p += (size_t)nd * nline;
    t.lof = p; p += (size_t)nd * nline;
    t.emf = p; p += (size_t)nd * nline;
    t.accf = p; p += (size_t)nd * nline;
    t.nlo = p; p += (size_t)nd * nbf;
    t.e1 = p;
    return t;
  }
};

// opacity, emissivity and line profiles of one depth for one wavelength
template <typename R, int kL, int kB>
__device__ __forceinline__ void ali_opem(const AliAtom &A, const AliTile<R> &t, int nd, int z, double lam, R lamR,
                                         R hcl, R plc, R ple, R fud, const R sg[kB], R bo, R be, R &op, R &em,
                                         R prof[kL]) {
  R o = bo, e = be;
  const R ex = rexp<R>(-hcl * t.tinv[z]);  // exp(-h c / (lambda k T))
  const R B = plc * rrcp<R>(rexp<R>(ple * t.tinv[z]) - R(1));
#pragma unroll
  for (int b = 0; b < kB; ++b) {
    if (b >= A.nbf) continue;
    const R sig = sg[b];
    if (sig == R(0)) continue;
    const int i = b * nd + z;
    o += sig * (t.nlo[i] - t.e1[i] * ex) * fud;
    e += sig * t.e1[i] * (R(1) - ex) * B * fud;
  }
  const R hc4l = R(kH * kC / (4.0 * kPi)) * rrcp<R>(lamR);
#pragma unroll
  for (int j = 0; j < kL; ++j) {
    prof[j] = R(0);
    if (j >= A.nline) continue;
    const int i = j * nd + z;
    const R x = (R)(A.lam0[j] - lam) * t.idld[i] - lamR * t.vr[z] * t.idld[i];
    R H, F;
    voigt<R>(x, t.a[i], H, F);
    const R p = t.idld[i] * H;
    prof[j] = p;
    o += p * t.lof[i] * hc4l;
    e += p * t.emf[i] * hc4l;
  }
  op = o;
  em = e;
}

// Normalised opacity and source function of one depth.  In double exactly as
// the reference forms them (op / op_ref, then (em / op_ref) / (op / op_ref)).
// In float the wavelength grid's far-UV points (opacities up to 1e32 cm^-1
// against referent opacities down to 1e-14 cm^-1) would overflow: the
// normalised opacity saturates at 1e20, where every step is optically thick
// and the solution equals the local source function, and S = em / op is
// formed from the raw values.
template <typename R> struct AliNorm;
template <> struct AliNorm<double> {
  static __device__ __forceinline__ void apply(bool tg, double op, double em, double opref, double &oo, double &S) {
    oo = tg ? op / opref : op;
    S = tg ? (em / opref) / oo : em / oo;
  }
  static __device__ __forceinline__ double fb(double om, double o0, double op, double h1, double h2) {
    return fb_deriv<double>(om, o0, op, h1, h2);
  }
};
template <> struct AliNorm<float> {
  static __device__ __forceinline__ void apply(bool tg, float op, float em, float opref, float &oo, float &S) {
    oo = fminf(tg ? op * rrcp<float>(opref) : op, 1e20f);
    S = em * rrcp<float>(op);
  }
  // Fritsch-Butland derivative in harmonic form, 1 / (al / d2 + (1 - al) / d1),
  // whose products cannot overflow for the saturated opacities
  static __device__ __forceinline__ float fb(float om, float o0, float op, float h1, float h2) {
    const float d2 = (op - o0) * rrcp<float>(h2), d1 = (o0 - om) * rrcp<float>(h1);
    if ((d1 > 0.0f && d2 > 0.0f) || (d1 < 0.0f && d2 < 0.0f)) {
      const float al = (1.0f + h1 * rrcp<float>(h1 + h2)) * (1.0f / 3.0f);
      return rrcp<float>(al * rrcp<float>(d2) + (1.0f - al) * rrcp<float>(d1));
    }
    return 0.0f;
  }
};

// With kLds the column's sums stay in LDS, [depth][3 per line, 2 per
// bound-free], and reach the accumulators in global memory once, after the
// last ray; the additions happen in the same order either way.  kL and kB
// bound the atom's lines and bound-free transitions the work-items carry in
// registers, kS (32 or 64) the reduced slots per depth.
// Wavefronts per SIMD the float kernel is held to (128 VGPRs: 49 spilled in
// the 32-slot variant, 159 in the 64-slot one); the double ones keep their
// registers.  On 4096
// columns x 81 depths x 220 wavelengths x 10 rays: 2 waves (188 VGPRs)
// 5.99 s, 4 waves 5.12 s, 5 waves 7.17 s, 8 waves 8.98 s.
// The float variants of the small and mid atoms also come held to three and
// two wavefronts, among which the autotuner picks (autotune.h); on the Ca II
// workload five and six wavefronts measured 1.5x and 2x slower than four.
template <typename R> constexpr int kAliWaves = 1;
template <> constexpr int kAliWaves<float> = 4;
template <typename R, bool kLds, int kL, int kB, int kS, int W = kAliWaves<R>>
__global__ __launch_bounds__(kAliThreads, W) void ali_acc_kernel(const AliAtom A, const AliGridArgs G, int nd, AliCol P) {
  extern __shared__ unsigned char smem_raw[];
  const int col = blockIdx.x;
  if (P.active && !P.active[col]) return;  // uniform over the workgroup
  const int nline = A.nline, nbf = A.nbf, nlam = G.nlam;
  const int nwave = blockDim.x / 64, wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  double *red = reinterpret_cast<double *>(smem_raw);  // [2][nwave][kS]
  const int nsum = 3 * nline + 2 * nbf;
  double *sums = red + 2 * nwave * kS;  // [nd][nsum] with kLds
  R *sm = reinterpret_cast<R *>(sums + (kLds ? (size_t)nd * nsum : 0));
  AliTile<R> t = AliTile<R>::carve(sm, nd, nline, nbf);
  const double *cs = P.cs + (size_t)col * SNAPI_A_NF * nd;
  const double *pops = P.pops + (size_t)col * nd * A.nlevel;
  double *acc = P.acc + (size_t)col * nd * A.ntrans * 4;
  // per-depth and per-(transition, depth) quantities of this column
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    const double T = cs[SNAPI_A_T * nd + z];
    t.grid[z] = (R)cs[SNAPI_A_GRID * nd + z];
    t.opref[z] = (R)cs[SNAPI_A_OPREF * nd + z];
    t.tinv[z] = (R)(1.0 / T);
  }
  for (int i = threadIdx.x; i < nd * nline; i += blockDim.x) {
    const int j = i / nd, z = i - j * nd;
    const double T = cs[SNAPI_A_T * nd + z], vt = cs[SNAPI_A_VT * nd + z];
    const double l0 = A.lam0[j];
    const double dld = sqrt(2.0 * kK * T / A.mass + vt * vt) * l0 / kC;
    const double sf = l0 * l0 / (4.0 * kPi * kC);
    const double *n = pops + (size_t)z * A.nlevel;
    const double nlo = n[A.line_lo[j]], nup = n[A.line_up[j]];
    t.idld[i] = (R)(1.0 / dld);
    t.a[i] = (R)(sf * P.gamma[((size_t)col * A.ntrans + A.line_t[j]) * nd + z] / dld);
    t.lof[i] = (R)(nlo * A.blu[j] * (1.0 - nup / nlo * A.bul[j] / A.blu[j]));
    t.emf[i] = (R)(nup * A.aul[j]);
    t.accf[i] = (R)((nlo * A.blu[j] - nup * A.bul[j]) * A.energy[j] * 0.25 / kPi);
  }
  if (kLds)
    for (int i = threadIdx.x; i < nd * nsum; i += blockDim.x) sums[i] = 0.0;
  for (int i = threadIdx.x; i < nd * nbf; i += blockDim.x) {
    const int b = i / nd, z = i - b * nd;
    const double T = cs[SNAPI_A_T * nd + z], ne = cs[SNAPI_A_NE * nd + z];
    const double *n = pops + (size_t)z * A.nlevel;
    t.nlo[i] = (R)n[A.bf_lo[b]];
    t.e1[i] = (R)(n[A.bf_up[b]] * ne * kSaha * pow(T, -1.5) * A.gratio[b] * exp(A.de[b] / kK / T));
  }
  for (int r = 0; r < G.nray; ++r) {
    __syncthreads();  // the previous ray's readers of vr are done
    for (int z = threadIdx.x; z < nd; z += blockDim.x)
      t.vr[z] = (R)(P.vproj[((size_t)col * G.nray + r) * nd + z] / kC);
    __syncthreads();
    const double ct = G.ct[r];
    const R mct = (R)(-ct);
    const bool down = ct > 0.0;  // enters at the top
    auto zof = [&](int k) { return down ? k : nd - 1 - k; };
    const size_t bgo = ((size_t)col * (G.bg_per_ray ? G.nray : 1) + (G.bg_per_ray ? r : 0)) * nd * nlam;
    const double *bop = P.bgop + bgo, *bem = P.bgem + bgo;
    for (int l0 = 0; l0 < nlam; l0 += blockDim.x) {
      const int l = l0 + threadIdx.x;
      const bool valid = l < nlam;
      const int lc = valid ? l : nlam - 1;  // lanes past the grid compute a copy and add nothing
      const double lam = P.lam[lc];
      const R lamR = (R)lam;
      const R hcl = (R)(kH * kC / lam / kK);
      const R plc = (R)(2.0 * 6.62606957 * 8.9875517E-7 / lam / lam / lam / lam / lam);
      const R ple = (R)(1.4387769 / lam);
      const R fud = (R)P.fudge[lc];
      const double ww = valid ? P.lamw[lc] * G.wray[r] : 0.0;
      // per-wavelength constants: cross-sections, 2 h c^2 / lambda^5, h c / (lambda k)
      R sg[kB];
      double sw[kB];  // weight x cross-section x lambda / (h c) of the photo-/recombination integrals
#pragma unroll
      for (int b = 0; b < kB; ++b) {
        const double sgd = b < nbf ? P.sigma[(size_t)A.bf_t[b] * nlam + lc] : 0.0;
        sg[b] = (R)sgd;
        sw[b] = ww * sgd * lam / kH / kC;
      }
      const double c5 = 2.0 * kH * kC * kC / pow(lam, 5.0);
      const double hclk = kH * kC / lam / kK;
      R profc[kL], profn[kL];
      R opc, emc, opn, emn;
      // depth 0 and the look-ahead depth 1
      const int z0 = zof(0), z1 = zof(1);
      ali_opem<R, kL, kB>(A, t, nd, z0, lam, lamR, hcl, plc, ple, fud, sg, (R)bop[(size_t)z0 * nlam + lc],
                  (R)bem[(size_t)z0 * nlam + lc], opc, emc, profc);
      ali_opem<R, kL, kB>(A, t, nd, z1, lam, lamR, hcl, plc, ple, fud, sg, (R)bop[(size_t)z1 * nlam + lc],
                  (R)bem[(size_t)z1 * nlam + lc], opn, emn, profn);
      const bool tg = G.tau_grid != 0;
      R oom, oo0, Sm, S0;
      AliNorm<R>::apply(tg, opc, emc, t.opref[z0], oom, Sm);
      AliNorm<R>::apply(tg, opn, emn, t.opref[z1], oo0, S0);
      R h0 = (t.grid[z1] - t.grid[z0]) / mct;
      R Dm = (oo0 - oom) / h0;
      R I = down ? R(0) : Sm;
      if (!down && G.boundary >= 0.0) I = (R)G.boundary;
      R Ls = R(0);
      if (!down) Ls = (G.boundary == -1.0) ? R(1) : R(0);
      for (int k = 0; k < nd; ++k) {
        const int z = zof(k);
        R opd = opc;  // the total opacity of depth z (de-normalised as the reference does)
        if (k > 0) {
          // optical-depth step k-1 -> k (D of the last point one-sided)
          R D0, oop = R(0), hp = R(1), Sp = R(0), opp = R(0), emp = R(0);
          R profp[kL];
          if (k < nd - 1) {
            const int zn = zof(k + 1);
            ali_opem<R, kL, kB>(A, t, nd, zn, lam, lamR, hcl, plc, ple, fud, sg,
                        (R)bop[(size_t)zn * nlam + lc], (R)bem[(size_t)zn * nlam + lc], opp, emp, profp);
            AliNorm<R>::apply(tg, opp, emp, t.opref[zn], oop, Sp);
            hp = (t.grid[zn] - t.grid[z]) / mct;
            D0 = AliNorm<R>::fb(oom, oo0, oop, h0, hp);
          } else {
            D0 = (oo0 - oom) / h0;
          }
          const R C = R(0.5) * ((oo0 - h0 * R(0.5) * D0) + (oom + h0 * R(0.5) * Dm));
          const R tb = h0 * R(1.0 / 3.0) * (oom + oo0 + C);
          R etb, w0, w1;
          sc_weights<R>(tb, etb, w0, w1);
          I = I * etb + w0 * Sm + w1 * S0;
          Ls = w1;
          opd = (tg && sizeof(R) == 8) ? oo0 * t.opref[z] : opn;
#pragma unroll
          for (int j = 0; j < kL; ++j) profc[j] = profn[j];
          // shift the window: depth k+1 becomes the look-ahead of the next step
          oom = oo0;
          oo0 = oop;
          Sm = S0;
          S0 = Sp;
          Dm = D0;
          h0 = hp;
          opn = opp;
#pragma unroll
          for (int j = 0; j < kL; ++j) profn[j] = profp[j];
        }
        // contributions of this wavelength to the sums of depth z, in fixed
        // slots: line j at 3j (J), 3j+1 (Lambda*), 3j+2 (norm); bound-free b
        // at kS-2-2b (photoionisation) and kS-1-2b (recombination)
        const int buf = k & 1;
        double *rb = red + (size_t)buf * nwave * kS;
        const R iop = rrcp<R>(opd);
        R v[kS];
#pragma unroll
        for (int q = 0; q < kS; ++q) v[q] = R(0);
        const R wwr = (R)ww;
#pragma unroll
        for (int j = 0; j < kL; ++j) {
          if (j >= nline) continue;
          const R p = profc[j];
          // line_factor = (n_l B_lu - n_u B_ul) phi E / (4 pi op); weight phi line_factor / (4 pi)
          const R cf = wwr * p * (p * t.accf[j * nd + z] * iop) * R(0.25 / kPi);
          v[3 * j] = cf * I;
          v[3 * j + 1] = cf * Ls;
          v[3 * j + 2] = wwr * p * R(0.25 / kPi);
        }
        // exp(-h c / (lambda k T)): in double on the double path, in float
        // (whose result the float path rounds to anyway) on the mixed one
        double exd = 0.0;
        if (nbf) {
          if constexpr (sizeof(R) == 8)
            exd = exp(-hclk * (double)t.tinv[z]);
          else
            exd = (double)rexp<R>(-hcl * t.tinv[z]);
        }
#pragma unroll
        for (int b = 0; b < kB; ++b) {
          if (b >= nbf) continue;
          const double s = sw[b];
          const double Id = (double)I;
          v[kS - 2 - 2 * b] = (R)(s * Id);
          v[kS - 1 - 2 * b] = (R)(s * (Id + c5) * exd);
        }
        // butterfly reduction over the wavefront: at each step a lane keeps
        // half of its slots and adds its partner's copy of them; at the end
        // lane L holds the wavefront's sum of slot L mod kS (with 32 slots the
        // two half-wavefronts are added first)
        if (kS < 64) {
#pragma unroll
          for (int q = 0; q < kS; ++q) v[q] += __shfl_xor(v[q], 32, 64);
        }
#pragma unroll
        for (int o = kS / 2, half = kS / 2; o >= 1; o >>= 1, half >>= 1) {
          const bool upper = (lane & o) != 0;
#pragma unroll
          for (int q = 0; q < half; ++q) {
            const R lo = v[q], hi = v[q + half];
            const R recv = __shfl_xor(upper ? lo : hi, o, 64);
            v[q] = (upper ? hi : lo) + recv;
          }
        }
        if (lane < kS) rb[wave * kS + lane] = (double)v[0];
        __syncthreads();
        for (int qq = threadIdx.x; qq < kS; qq += blockDim.x) {
          int tr = -1, slot = 0;
          if (qq < 3 * nline) {
            tr = A.line_t[qq / 3];
            slot = qq % 3;
          } else if (qq >= kS - 2 * nbf) {
            const int r2 = kS - 1 - qq;
            tr = A.bf_t[r2 / 2];
            slot = (r2 & 1) ? 0 : 1;
          }
          if (tr < 0) continue;
          double s = 0.0;
          for (int w = 0; w < nwave; ++w) s += rb[w * kS + qq];
          if (kLds) {
            const int c = qq < 3 * nline ? qq : 3 * nline + 2 * ((kS - 1 - qq) / 2) + slot;
            sums[(size_t)z * nsum + c] += s;
            continue;
          }
          double *ap = acc + ((size_t)z * A.ntrans + tr) * 4;
          if (qq < 3 * nline) {
            if (slot == 0) {
              ap[0] += s;
              ap[1] += s;
            } else if (slot == 1) {
              ap[2] += s;
            } else {
              ap[3] += s;
            }
          } else {
            ap[slot] += s;
          }
        }
        // the other buffer is written next; this one is read until the
        // barrier of the following depth
      }
      __syncthreads();
    }
  }
  if (kLds) {
    for (int i = threadIdx.x; i < nd * nsum; i += blockDim.x) {
      const int z = i / nsum, c = i - z * nsum;
      const double v = sums[i];
      if (c < 3 * nline) {
        double *ap = acc + ((size_t)z * A.ntrans + A.line_t[c / 3]) * 4;
        const int slot = c % 3;
        if (slot == 0) {
          ap[0] += v;
          ap[1] += v;
        } else if (slot == 1) {
          ap[2] += v;
        } else {
          ap[3] += v;
        }
      } else {
        const int b = (c - 3 * nline) / 2, slot = (c - 3 * nline) & 1;
        acc[((size_t)z * A.ntrans + A.bf_t[b]) * 4 + slot] += v;
      }
    }
  }
}

// New populations of one (column, depth): thread per depth, matrix in LDS.
__global__ void ali_newpops_kernel(const AliAtom A, int ncol, int nd, const double *__restrict__ cs_all,
                                   const double *__restrict__ coll, const double *__restrict__ pops,
                                   const double *__restrict__ acc, int alo, double *__restrict__ newp,
                                   double *__restrict__ rc_col, int *__restrict__ neg_col,
                                   const int *__restrict__ mask) {
  extern __shared__ double lds[];
  const int n = A.nlevel, nt = A.ntrans;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)ncol * nd) return;
  const int col = (int)(gid / nd), z = (int)(gid % nd);
  if (mask && !mask[col]) return;
  double *M = lds + (size_t)threadIdx.x * (n * n + 2 * n);
  double *b = M + n * n, *y = b + n;
  const double *p = pops + ((size_t)col * nd + z) * n;
  const double *A4 = acc + ((size_t)col * nd + z) * nt * 4;
  const double *C = coll + ((size_t)col * nd + z) * n * n;
  const double *cs = cs_all + (size_t)col * SNAPI_A_NF * nd;
  const double T = cs[SNAPI_A_T * nd + z], ne = cs[SNAPI_A_NE * nd + z];
  for (int i = 0; i < n * n; ++i) M[i] = 0.0;
  for (int i = 0; i < n; ++i) b[i] = 0.0;
  for (int i = 0; i < n - 1; ++i) {
    for (int j = 0; j < n; ++j) {
      if (A.stage[j] != A.stage[i] || j == i) continue;
      const int s = A.pair_line[i][j];
      double Rout = 0.0, Rin = 0.0;
      if (s >= 0) {
        const int t = A.line_t[s];
        const double JJ = A4[t * 4 + 0] / A4[t * 4 + 3], LL = A4[t * 4 + 2] / A4[t * 4 + 3];
        const bool idown = (i == A.line_up[s]);  // i -> j is a downward transition
        const double nu = idown ? p[i] : p[j], nl = idown ? p[j] : p[i];
        const double Sl = nu * A.aul[s] / (nl * A.blu[s] - nu * A.bul[s]);
        const double Ao = idown ? A.aul[s] : 0.0, Bo = idown ? A.bul[s] : A.blu[s];
        const double Ai = idown ? 0.0 : A.aul[s], Bi = idown ? A.blu[s] : A.bul[s];
        if (alo) {
          Rout = Ao * (1.0 - LL) + Bo * (JJ - LL * Sl);
          Rin = Ai * (1.0 - LL) + Bi * (JJ - LL * Sl);
        } else {
          Rout = Ao + Bo * JJ;
          Rin = Ai + Bi * JJ;
        }
      }
      M[i * n + i] -= (Rout + C[i * n + j]);
      M[i * n + j] += (Rin + C[j * n + i]);
    }
    const int bs = A.bf_of[i];
    if (bs >= 0) {
      const int t = A.bf_t[bs], ic = A.bf_up[bs];
      double R = A4[t * 4 + 0];
      M[i * n + i] -= (R + C[i * n + ic]);
      M[ic * n + i] += (R + C[i * n + ic]);
      R = A4[t * 4 + 1] * A.gratio[bs] * exp(A.de[bs] / kK / T) * pow(T, -1.5) * ne * kSaha;
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
  // pivot-free Crout in place: L (with diagonal) below, unit U above
  for (int k = 0; k < n; ++k) {
    for (int i = k; i < n; ++i) {
      double sum = 0.0;
      for (int q = 0; q < k; ++q) sum += M[i * n + q] * M[q * n + k];
      M[i * n + k] -= sum;
    }
    for (int j = k + 1; j < n; ++j) {
      double sum = 0.0;
      for (int q = 0; q < k; ++q) sum += M[k * n + q] * M[q * n + j];
      M[k * n + j] = (M[k * n + j] - sum) / M[k * n + k];
    }
  }
  for (int i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int q = 0; q < i; ++q) sum += M[i * n + q] * y[q];
    y[i] = (b[i] - sum) / M[i * n + i];
  }
  double *x = newp + ((size_t)col * nd + z) * n;
  for (int i = n - 1; i >= 0; --i) {
    double sum = 0.0;
    for (int q = i + 1; q < n; ++q) sum += M[i * n + q] * x[q];
    x[i] = y[i] - sum;
  }
  double mx = 0.0;
  int neg = 0;
  for (int i = 0; i < n; ++i) {
    const double r = x[i] / p[i] - 1.0;
    if (r < -1.0) neg = 1;
    if (!(fabs(r) <= mx)) mx = fabs(r);  // a NaN ends up as the maximum
  }
  // non-negative doubles order like their bit patterns
  atomicMax(reinterpret_cast<unsigned long long *>(rc_col + col), (unsigned long long)__double_as_longlong(mx));
  if (neg) atomicOr(neg_col + col, 1);
}

// per-column bookkeeping of the iteration
__global__ void ali_update_kernel(int ncol, int nd, int n, int max_iter, double tol, int *__restrict__ active,
                                  int *__restrict__ retry, const double *__restrict__ rc, const int *__restrict__ neg,
                                  const double *__restrict__ newp, double *__restrict__ pops, int *__restrict__ iters,
                                  double *__restrict__ rc_out, int stage) {
  const int col = blockIdx.x;
  if (stage == 0) {  // after the approximate-operator solve: columns to redo with Lambda iteration
    if (threadIdx.x == 0) retry[col] = active[col] ? neg[col] : 0;
    return;
  }
  if (!active[col]) return;
  __syncthreads();
  const bool bad = neg[col] != 0;
  if (!bad)
    for (int i = threadIdx.x; i < nd * n; i += blockDim.x) pops[(size_t)col * nd * n + i] = newp[(size_t)col * nd * n + i];
  __syncthreads();
  if (threadIdx.x == 0) {
    iters[col] += 1;
    rc_out[col] = bad ? -1.0 : rc[col];
    if (bad || rc[col] < tol || iters[col] >= max_iter) active[col] = 0;
  }
}

__global__ void ali_reset_kernel(int ncol, const int *__restrict__ mask, double *__restrict__ rc,
                                 int *__restrict__ neg) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol || (mask && !mask[col])) return;
  rc[col] = 0.0;
  neg[col] = 0;
}

__global__ void ali_zero_kernel(size_t n_per_col, int ncol, const int *__restrict__ mask, double *__restrict__ p) {
  const size_t n = n_per_col * (size_t)ncol;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    if (!mask || mask[i / n_per_col]) p[i] = 0.0;
}

// Number of active columns, written (not added) by one workgroup, so the
// counter needs no clearing between iterations.
__global__ __launch_bounds__(256) void ali_count_kernel(int ncol, const int *__restrict__ active,
                                                        int *__restrict__ count) {
  __shared__ int part[256];
  int n = 0;
  for (int col = threadIdx.x; col < ncol; col += 256) n += active[col] ? 1 : 0;
  part[threadIdx.x] = n;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = part[0];
}

// Start of the iteration: every column active, nothing to redo, no
// iterations yet (kernels rather than memsets, so a captured stream records
// no memset nodes).
__global__ void ali_init_kernel(int ncol, int *__restrict__ active, int *__restrict__ retry, int *__restrict__ iters,
                                double *__restrict__ rc_out) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  active[col] = 1;
  retry[col] = 0;
  iters[col] = 0;
  rc_out[col] = 0.0;
}

// ---------------------------------------------------------------- host side
static int convert_atom(const snapi_ali_atom *a, AliAtom &A) {
  if (!a || a->nlevel < 2 || a->nlevel > SNAPI_ALI_MAX_LEVELS || a->ntrans < 1 || a->ntrans > SNAPI_ALI_MAX_TRANS)
    return SNAPI_ERR_ARG;
  memset(&A, 0, sizeof(A));
  A.nlevel = a->nlevel;
  A.ntrans = a->ntrans;
  A.mass = a->mass;
  for (int i = 0; i < SNAPI_ALI_MAX_LEVELS; ++i) {
    A.bf_of[i] = -1;
    for (int j = 0; j < SNAPI_ALI_MAX_LEVELS; ++j) A.pair_line[i][j] = -1;
  }
  for (int t = 0; t < a->ntrans; ++t) {
    const snapi_ali_transition &tr = a->tr[t];
    if (tr.lower < 0 || tr.lower >= a->nlevel || tr.upper < 0 || tr.upper >= a->nlevel || tr.lower == tr.upper)
      return SNAPI_ERR_ARG;
    if (tr.kind == SNAPI_TR_LINE) {
      if (A.nline >= kAliMaxLines) return SNAPI_ERR_UNSUPPORTED;
      const int j = A.nline++;
      A.line_t[j] = t;
      A.line_lo[j] = tr.lower;
      A.line_up[j] = tr.upper;
      A.lam0[j] = tr.lam0;
      A.aul[j] = tr.aul;
      A.blu[j] = tr.blu;
      A.bul[j] = tr.bul;
      A.energy[j] = tr.energy;
      A.pair_line[tr.lower][tr.upper] = A.pair_line[tr.upper][tr.lower] = (signed char)j;
      if (!(tr.blu > 0.0) || !(tr.lam0 > 0.0)) return SNAPI_ERR_ARG;
    } else if (tr.kind == SNAPI_TR_BF) {
      if (A.nbf >= kAliMaxBf || A.bf_of[tr.lower] >= 0) return SNAPI_ERR_UNSUPPORTED;
      const int b = A.nbf++;
      A.bf_t[b] = t;
      A.bf_lo[b] = tr.lower;
      A.bf_up[b] = tr.upper;
      A.gratio[b] = tr.gratio;
      A.de[b] = tr.de;
      A.bf_of[tr.lower] = b;
    } else {
      return SNAPI_ERR_ARG;
    }
  }
  // every level but the last ionises; the stage ends at a level whose
  // continuum is the next level
  for (int i = 0; i < a->nlevel - 1; ++i)
    if (A.bf_of[i] < 0) return SNAPI_ERR_UNSUPPORTED;
  if (A.bf_of[a->nlevel - 1] >= 0) return SNAPI_ERR_UNSUPPORTED;
  int s = 0;
  for (int i = 0; i < a->nlevel; ++i) {
    A.stage[i] = s;
    if (i + 1 < a->nlevel && A.bf_up[A.bf_of[i]] == i + 1) ++s;
  }
  return SNAPI_OK;
}

static int convert_grid(const snapi_ali_grid *g, AliGridArgs &G) {
  if (!g || g->nlam < 1 || g->nray < 1 || g->nray > 16 || !g->lam || !g->lamw || !g->fudge || !g->sigma ||
      !g->theta || !g->wray)
    return SNAPI_ERR_ARG;
  G.nlam = g->nlam;
  G.nray = g->nray;
  G.bg_per_ray = g->bg_per_ray ? 1 : 0;
  G.tau_grid = g->tau_grid ? 1 : 0;
  G.boundary = g->boundary;
  for (int r = 0; r < g->nray; ++r) {
    G.ct[r] = cos(g->theta[r]);
    G.wray[r] = g->wray[r];
    if (fabs(G.ct[r]) < 1e-6) return SNAPI_ERR_ARG;
  }
  return SNAPI_OK;
}

// Atoms of at most kSmallL lines and kSmallB bound-free transitions run the
// variant with 32 reduced slots and shorter register arrays.
constexpr int kSmallL = 6, kSmallB = 6;
// and atoms of at most kMidL lines and kMidB bound-free transitions (every
// level pair of a 5-level ion is a line, as the reference's opacity takes
// them: 10 lines and 5 bound-free for Ca II 5 + continuum) the 64-slot
// reduction with register arrays of their own size
constexpr int kMidL = 10, kMidB = 6;
// columns of the autotuner's sample
constexpr int kAliTuneColumns = 4096;

static size_t acc_smem(int precision, const AliAtom &A, int nd, int threads, bool lds_sums) {
  const size_t es = precision == SNAPI_FP64 ? 8 : 4;
  return sizeof(double) * 2 * (threads / 64) * kAliSlots + es * (size_t)nd * (4 + 5 * A.nline + 2 * A.nbf) +
         (lds_sums ? sizeof(double) * (size_t)nd * (3 * A.nline + 2 * A.nbf) : 0);
}

static int launch_acc(int precision, const AliAtom &A, const AliGridArgs &G, int ncol, int nd, const AliCol &P,
                      hipStream_t st) {
  const int threads = kAliThreads;
  if (3 * A.nline + 2 * A.nbf > kAliSlots) return SNAPI_ERR_UNSUPPORTED;
  // the column's sums in LDS when they fit beside the tile, else in global memory
  const bool lds = acc_smem(precision, A, nd, threads, true) <= 65536;
  const size_t smem = acc_smem(precision, A, nd, threads, lds);
  if (smem > 65536) return SNAPI_ERR_UNSUPPORTED;
  const bool small = A.nline <= kSmallL && A.nbf <= kSmallB && 3 * A.nline + 2 * A.nbf <= 32;
  const bool mid = A.nline <= kMidL && A.nbf <= kMidB;
  const dim3 g(ncol), b(threads);
  if (precision != SNAPI_FP64 && (small || mid) && lds) {
    // float, small or mid atom: occupancy bound chosen by the autotuner; a
    // timed run clears the sums first, as every caller does before the launch
    using AccFn = void (*)(const AliAtom, const AliGridArgs, int, AliCol);
    const AccFn fs[3] = {ali_acc_kernel<float, true, kSmallL, kSmallB, 32, 4>,
                         ali_acc_kernel<float, true, kSmallL, kSmallB, 32, 3>,
                         ali_acc_kernel<float, true, kSmallL, kSmallB, 32, 2>};
    const AccFn fm[3] = {ali_acc_kernel<float, true, kMidL, kMidB, 64, 4>, ali_acc_kernel<float, true, kMidL, kMidB, 64, 3>,
                         ali_acc_kernel<float, true, kMidL, kMidB, 64, 2>};
    const AccFn *fn = small ? fs : fm;
    const size_t per_col = (size_t)nd * A.ntrans * 4;
    auto run = [&](int c, int nc, bool clear) {
      if (clear)
        hipLaunchKernelGGL(ali_zero_kernel, dim3(1024), dim3(256), 0, st, per_col, nc, P.active, P.acc);
      hipLaunchKernelGGL(fn[c], dim3(nc), b, smem, st, A, G, nd, P);
      return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
    };
    const int ns = ncol < kAliTuneColumns ? ncol : kAliTuneColumns;
    const TuneKey key{TUNE_ALI, SNAPI_MIXED, nd, A.nline, A.nbf, G.nlam,
                      (tune_size_class(ncol, kAliTuneColumns) << 8) + G.nray};
    bool tuned = false;
    const int c = tune_choose(key, 3, 0, st, true, [&](int cand) {
      tuned = true;
      return run(cand, ns, true);
    });
    return run(c, ncol, tuned);
  }
  if (precision == SNAPI_FP64) {
    if (small && lds)
      hipLaunchKernelGGL((ali_acc_kernel<double, true, kSmallL, kSmallB, 32>), g, b, smem, st, A, G, nd, P);
    else if (mid && lds)
      hipLaunchKernelGGL((ali_acc_kernel<double, true, kMidL, kMidB, 64>), g, b, smem, st, A, G, nd, P);
    else if (lds)
      hipLaunchKernelGGL((ali_acc_kernel<double, true, kAliMaxLines, kAliMaxBf, 64>), g, b, smem, st, A, G, nd, P);
    else
      hipLaunchKernelGGL((ali_acc_kernel<double, false, kAliMaxLines, kAliMaxBf, 64>), g, b, smem, st, A, G, nd, P);
  } else {
    if (small && lds)
      hipLaunchKernelGGL((ali_acc_kernel<float, true, kSmallL, kSmallB, 32>), g, b, smem, st, A, G, nd, P);
    else if (mid && lds)
      hipLaunchKernelGGL((ali_acc_kernel<float, true, kMidL, kMidB, 64>), g, b, smem, st, A, G, nd, P);
    else if (lds)
      hipLaunchKernelGGL((ali_acc_kernel<float, true, kAliMaxLines, kAliMaxBf, 64>), g, b, smem, st, A, G, nd, P);
    else
      hipLaunchKernelGGL((ali_acc_kernel<float, false, kAliMaxLines, kAliMaxBf, 64>), g, b, smem, st, A, G, nd, P);
  }
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

static int newpops_threads(int n) {
  const size_t per = sizeof(double) * (size_t)(n * n + 2 * n);
  int th = (int)(65536 / per);
  if (th > 128) th = 128;
  th = (th / 32) * 32;
  return th < 32 ? 32 : th;
}

static int launch_newpops(const AliAtom &A, int ncol, int nd, const double *cs, const double *coll,
                          const double *pops, const double *acc, int alo, double *newp, double *rc, int *neg,
                          const int *mask, hipStream_t st) {
  const int th = newpops_threads(A.nlevel);
  const size_t smem = sizeof(double) * (size_t)th * (A.nlevel * A.nlevel + 2 * A.nlevel);
  const long long nthr = (long long)ncol * nd;
  hipLaunchKernelGGL(ali_reset_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, ncol, mask, rc, neg);
  hipLaunchKernelGGL(ali_newpops_kernel, dim3((unsigned)((nthr + th - 1) / th)), dim3(th), smem, st, A, ncol, nd, cs,
                     coll, pops, acc, alo, newp, rc, neg, mask);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

#define ACHK(x)                \
  do {                         \
    if ((x) != hipSuccess) {   \
      rc = SNAPI_ERR_DEVICE;   \
      goto done;               \
    }                          \
  } while (0)

struct AliBuffers {
  double *cs = nullptr, *vproj = nullptr, *gamma = nullptr, *bgop = nullptr, *bgem = nullptr, *pops = nullptr;
  double *lam = nullptr, *lamw = nullptr, *fudge = nullptr, *sigma = nullptr, *acc = nullptr;
  void release() {
    double *p[] = {cs, vproj, gamma, bgop, bgem, pops, lam, lamw, fudge, sigma, acc};
    for (double *q : p) hipFree(q);
  }
};

static int upload_inputs(const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol, int nd,
                         const double *colstate, const double *vproj, const double *gamma, const double *bgop,
                         const double *bgem, const double *pops, AliBuffers &B, AliCol &P) {
  const size_t nlam = grid->nlam, nray = grid->nray, nt = atom->ntrans;
  const size_t nbg = (size_t)ncol * (grid->bg_per_ray ? nray : 1) * nd * nlam;
  struct {
    double **dst;
    const void *src;
    size_t n;
  } up[] = {{&B.cs, colstate, (size_t)ncol * SNAPI_A_NF * nd},
            {&B.vproj, vproj, (size_t)ncol * nray * nd},
            {&B.gamma, gamma, (size_t)ncol * nt * nd},
            {&B.bgop, bgop, nbg},
            {&B.bgem, bgem, nbg},
            {&B.pops, pops, (size_t)ncol * nd * atom->nlevel},
            {&B.lam, grid->lam, nlam},
            {&B.lamw, grid->lamw, nlam},
            {&B.fudge, grid->fudge, nlam},
            {&B.sigma, grid->sigma, nt * nlam}};
  for (auto &u : up) {
    if (hipMalloc((void **)u.dst, u.n * sizeof(double)) != hipSuccess) return SNAPI_ERR_DEVICE;
    if (hipMemcpy(*u.dst, u.src, u.n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return SNAPI_ERR_DEVICE;
  }
  if (hipMalloc((void **)&B.acc, (size_t)ncol * nd * nt * 4 * sizeof(double)) != hipSuccess) return SNAPI_ERR_DEVICE;
  P.cs = B.cs;
  P.vproj = B.vproj;
  P.gamma = B.gamma;
  P.bgop = B.bgop;
  P.bgem = B.bgem;
  P.pops = B.pops;
  P.lam = B.lam;
  P.lamw = B.lamw;
  P.fudge = B.fudge;
  P.sigma = B.sigma;
  P.acc = B.acc;
  P.active = nullptr;
  return SNAPI_OK;
}

static bool bad_inputs(const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol, int nd, const void *a,
                       const void *b, const void *c, const void *d, const void *e) {
  return !atom || !grid || ncol <= 0 || nd < 3 || nd > SNAPI_MAX_ND || !a || !b || !c || !d || !e;
}

}  // namespace snapi

using namespace snapi;

extern "C" int snapi_ali_accumulate(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol,
                                    int nd, const double *colstate, const double *vproj, const double *gamma,
                                    const double *bgop, const double *bgem, const double *pops, double *acc) {
  if (bad_inputs(atom, grid, ncol, nd, colstate, vproj, gamma, bgop, bgem) || !pops || !acc) return SNAPI_ERR_ARG;
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  AliAtom A;
  AliGridArgs G;
  int rc = convert_atom(atom, A);
  if (rc) return rc;
  rc = convert_grid(grid, G);
  if (rc) return rc;
  AliBuffers B;
  AliCol P;
  rc = upload_inputs(atom, grid, ncol, nd, colstate, vproj, gamma, bgop, bgem, pops, B, P);
  if (rc) goto done;
  ACHK(hipMemset(B.acc, 0, (size_t)ncol * nd * atom->ntrans * 4 * sizeof(double)));
  rc = launch_acc(precision, A, G, ncol, nd, P, 0);
  if (rc) goto done;
  ACHK(hipDeviceSynchronize());
  ACHK(hipMemcpy(acc, B.acc, (size_t)ncol * nd * atom->ntrans * 4 * sizeof(double), hipMemcpyDeviceToHost));
done:
  B.release();
  return rc;
}

extern "C" int snapi_ali_newpops(const snapi_ali_atom *atom, int ncol, int nd, const double *colstate,
                                 const double *coll, const double *pops, const double *acc, int alo, double *newpops,
                                 double *relchange) {
  if (!atom || ncol <= 0 || nd < 1 || !colstate || !coll || !pops || !acc || !newpops || !relchange)
    return SNAPI_ERR_ARG;
  AliAtom A;
  int rc = convert_atom(atom, A);
  if (rc) return rc;
  const int n = atom->nlevel, nt = atom->ntrans;
  double *dcs = nullptr, *dcoll = nullptr, *dp = nullptr, *dacc = nullptr, *dnew = nullptr, *drc = nullptr;
  int *dneg = nullptr;
  std::vector<int> neg(ncol);
  ACHK(hipMalloc((void **)&dcs, (size_t)ncol * SNAPI_A_NF * nd * 8));
  ACHK(hipMalloc((void **)&dcoll, (size_t)ncol * nd * n * n * 8));
  ACHK(hipMalloc((void **)&dp, (size_t)ncol * nd * n * 8));
  ACHK(hipMalloc((void **)&dacc, (size_t)ncol * nd * nt * 4 * 8));
  ACHK(hipMalloc((void **)&dnew, (size_t)ncol * nd * n * 8));
  ACHK(hipMalloc((void **)&drc, (size_t)ncol * 8));
  ACHK(hipMalloc((void **)&dneg, (size_t)ncol * sizeof(int)));
  ACHK(hipMemcpy(dcs, colstate, (size_t)ncol * SNAPI_A_NF * nd * 8, hipMemcpyHostToDevice));
  ACHK(hipMemcpy(dcoll, coll, (size_t)ncol * nd * n * n * 8, hipMemcpyHostToDevice));
  ACHK(hipMemcpy(dp, pops, (size_t)ncol * nd * n * 8, hipMemcpyHostToDevice));
  ACHK(hipMemcpy(dacc, acc, (size_t)ncol * nd * nt * 4 * 8, hipMemcpyHostToDevice));
  rc = launch_newpops(A, ncol, nd, dcs, dcoll, dp, dacc, alo ? 1 : 0, dnew, drc, dneg, nullptr, 0);
  if (rc) goto done;
  ACHK(hipDeviceSynchronize());
  ACHK(hipMemcpy(newpops, dnew, (size_t)ncol * nd * n * 8, hipMemcpyDeviceToHost));
  ACHK(hipMemcpy(relchange, drc, (size_t)ncol * 8, hipMemcpyDeviceToHost));
  ACHK(hipMemcpy(neg.data(), dneg, (size_t)ncol * sizeof(int), hipMemcpyDeviceToHost));
  for (int c = 0; c < ncol; ++c)
    if (neg[c]) relchange[c] = -1.0;
done:
  hipFree(dcs); hipFree(dcoll); hipFree(dp); hipFree(dacc); hipFree(dnew); hipFree(drc); hipFree(dneg);
  return rc;
}

// Device work areas of the iteration beyond its inputs and outputs.
struct AliWork {
  double *newp = nullptr;  // [ncol][nd][nlevel] populations of the current solve
  double *rc = nullptr;    // [ncol] relative change of the current solve
  int *ints = nullptr;     // [3 ncol + 1]: active, retry, negative flags, active count
};

static size_t ali_work_doubles(int ncol, int nd, int nlevel) { return (size_t)ncol * nd * nlevel + ncol; }
static size_t ali_work_ints(int ncol) { return (size_t)ncol * 3 + 1; }

// The iteration on device-resident inputs (P: heavy arrays and the grid's
// arrays on the device, P.acc a scratch of ncol * nd * ntrans * 4 doubles;
// W: the work areas).  The only host traffic is one int per iteration.
static int ali_solve_core(int precision, const AliAtom &A, const AliGridArgs &G, int ncol, int nd, AliCol P,
                          const AliWork &W, const double *dcoll, double *dpops, int *diters, double *drco, double tol,
                          int max_iter, hipStream_t st) {
  const int n = A.nlevel, nt = A.ntrans;
  int rc = SNAPI_OK;
  double *dnew = W.newp, *drc = W.rc;
  int *active = W.ints, *retry = W.ints + ncol, *neg = W.ints + 2 * ncol, *count = W.ints + 3 * ncol;
  hipLaunchKernelGGL(ali_init_kernel, dim3((ncol + 255) / 256), dim3(256), 0, st, ncol, active, retry, diters, drco);
  ACHK(hipGetLastError());
  P.active = active;
  P.pops = dpops;
  for (int it = 0; it < max_iter; ++it) {
    int nact = 0;
    hipLaunchKernelGGL(ali_count_kernel, dim3(1), dim3(256), 0, st, ncol, active, count);
    ACHK(hipMemcpyAsync(&nact, count, sizeof(int), hipMemcpyDeviceToHost, st));
    ACHK(hipStreamSynchronize(st));
    if (nact == 0) break;
    hipLaunchKernelGGL(ali_zero_kernel, dim3(1024), dim3(256), 0, st, (size_t)nd * nt * 4, ncol, active, P.acc);
    rc = launch_acc(precision, A, G, ncol, nd, P, st);
    if (rc) goto done;
    rc = launch_newpops(A, ncol, nd, P.cs, dcoll, dpops, P.acc, 1, dnew, drc, neg, active, st);
    if (rc) goto done;
    hipLaunchKernelGGL(ali_update_kernel, dim3(ncol), dim3(64), 0, st, ncol, nd, n, max_iter, tol, active, retry, drc,
                       neg, dnew, dpops, diters, drco, 0);
    rc = launch_newpops(A, ncol, nd, P.cs, dcoll, dpops, P.acc, 0, dnew, drc, neg, retry, st);
    if (rc) goto done;
    hipLaunchKernelGGL(ali_update_kernel, dim3(ncol), dim3(256), 0, st, ncol, nd, n, max_iter, tol, active, retry,
                       drc, neg, dnew, dpops, diters, drco, 1);
    if (hipGetLastError() != hipSuccess) {
      rc = SNAPI_ERR_DEVICE;
      goto done;
    }
  }
  ACHK(hipStreamSynchronize(st));
done:
  return rc;
}

// The work areas from the device heap (freed by release_work).
static int alloc_work(int ncol, int nd, int nlevel, AliWork &W) {
  W = AliWork{};
  if (hipMalloc((void **)&W.newp, ali_work_doubles(ncol, nd, nlevel) * 8) != hipSuccess) return SNAPI_ERR_DEVICE;
  W.rc = W.newp + (size_t)ncol * nd * nlevel;
  if (hipMalloc((void **)&W.ints, ali_work_ints(ncol) * sizeof(int)) != hipSuccess) return SNAPI_ERR_DEVICE;
  return SNAPI_OK;
}
static void release_work(AliWork &W) {
  hipFree(W.newp);
  hipFree(W.ints);
  W = AliWork{};
}

extern "C" int snapi_ali_solve(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol, int nd,
                               const double *colstate, const double *vproj, const double *gamma, const double *bgop,
                               const double *bgem, const double *coll, double tol, int max_iter, double *pops,
                               int *iterations, double *relchange) {
  if (bad_inputs(atom, grid, ncol, nd, colstate, vproj, gamma, bgop, bgem) || !coll || !pops || !iterations ||
      !relchange || max_iter < 1)
    return SNAPI_ERR_ARG;
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  AliAtom A;
  AliGridArgs G;
  int rc = convert_atom(atom, A);
  if (rc) return rc;
  rc = convert_grid(grid, G);
  if (rc) return rc;
  const int n = atom->nlevel;
  AliBuffers B;
  AliCol P;
  AliWork W;
  double *dcoll = nullptr, *drco = nullptr;
  int *diters = nullptr;
  rc = upload_inputs(atom, grid, ncol, nd, colstate, vproj, gamma, bgop, bgem, pops, B, P);
  if (rc) goto done;
  ACHK(hipMalloc((void **)&dcoll, (size_t)ncol * nd * n * n * 8));
  ACHK(hipMemcpy(dcoll, coll, (size_t)ncol * nd * n * n * 8, hipMemcpyHostToDevice));
  ACHK(hipMalloc((void **)&drco, (size_t)ncol * 8));
  ACHK(hipMalloc((void **)&diters, (size_t)ncol * sizeof(int)));
  rc = alloc_work(ncol, nd, n, W);
  if (rc) goto done;
  rc = ali_solve_core(precision, A, G, ncol, nd, P, W, dcoll, B.pops, diters, drco, tol, max_iter, 0);
  if (rc) goto done;
  ACHK(hipMemcpy(pops, B.pops, (size_t)ncol * nd * n * 8, hipMemcpyDeviceToHost));
  ACHK(hipMemcpy(iterations, diters, (size_t)ncol * sizeof(int), hipMemcpyDeviceToHost));
  ACHK(hipMemcpy(relchange, drco, (size_t)ncol * 8, hipMemcpyDeviceToHost));
done:
  B.release();
  release_work(W);
  hipFree(dcoll); hipFree(drco); hipFree(diters);
  return rc;
}

extern "C" int snapi_ali_solve_dev(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid, int ncol,
                                   int nd, const double *colstate_dev, const double *vproj_dev,
                                   const double *gamma_dev, const double *bgop_dev, const double *bgem_dev,
                                   const double *coll_dev, double tol, int max_iter, double *pops_dev,
                                   int *iterations_dev, double *relchange_dev, void *stream) {
  if (bad_inputs(atom, grid, ncol, nd, colstate_dev, vproj_dev, gamma_dev, bgop_dev, bgem_dev) || !coll_dev ||
      !pops_dev || !iterations_dev || !relchange_dev || max_iter < 1)
    return SNAPI_ERR_ARG;
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  AliAtom A;
  AliGridArgs G;
  int rc = convert_atom(atom, A);
  if (rc) return rc;
  rc = convert_grid(grid, G);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const size_t nlam = grid->nlam, nt = atom->ntrans;
  double *g = nullptr, *acc = nullptr;
  AliCol P;
  AliWork W;
  ACHK(alloc_work(ncol, nd, atom->nlevel, W) == SNAPI_OK ? hipSuccess : hipErrorOutOfMemory);
  ACHK(hipMalloc((void **)&g, (3 * nlam + nt * nlam) * 8));
  ACHK(hipMalloc((void **)&acc, (size_t)ncol * nd * nt * 4 * 8));
  ACHK(hipMemcpyAsync(g, grid->lam, nlam * 8, hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(g + nlam, grid->lamw, nlam * 8, hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(g + 2 * nlam, grid->fudge, nlam * 8, hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(g + 3 * nlam, grid->sigma, nt * nlam * 8, hipMemcpyHostToDevice, st));
  P.cs = colstate_dev;
  P.vproj = vproj_dev;
  P.gamma = gamma_dev;
  P.bgop = bgop_dev;
  P.bgem = bgem_dev;
  P.pops = pops_dev;
  P.lam = g;
  P.lamw = g + nlam;
  P.fudge = g + 2 * nlam;
  P.sigma = g + 3 * nlam;
  P.acc = acc;
  P.active = nullptr;
  rc = ali_solve_core(precision, A, G, ncol, nd, P, W, coll_dev, pops_dev, iterations_dev, relchange_dev, tol,
                      max_iter, st);
done:
  release_work(W);
  hipFree(g);
  hipFree(acc);
  return rc;
}

// ------------------------------------------- allocation-free ALI iteration
// The work areas of snapi_ali_solve_dev_noalloc live in one fixed device
// array.  Each device that runs it holds its own copy (at its own address,
// looked up once per device) until the process ends.
namespace snapi {

static constexpr size_t kAliScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kAliScratchBytes];

static constexpr int kAliScratchDevices = 64;
static unsigned char *g_scratch_dev[kAliScratchDevices];
static std::mutex g_scratch_mu;

// Device address of this device's copy of `scratch` (null on failure).
static unsigned char *scratch_address() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kAliScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  }
  return g_scratch_dev[dev];
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the fixed area a problem needs: the grid's arrays, the sums, the
// work areas.
static size_t noalloc_bytes(int ncol, int nd, int nlevel, int ntrans, int nlam) {
  return round256((size_t)(3 + ntrans) * nlam * 8) + round256((size_t)ncol * nd * ntrans * 4 * 8) +
         round256(ali_work_doubles(ncol, nd, nlevel) * 8) + round256(ali_work_ints(ncol) * sizeof(int));
}

}  // namespace snapi

extern "C" size_t snapi_ali_noalloc_capacity(void) { return kAliScratchBytes; }

extern "C" int snapi_ali_solve_dev_noalloc(int precision, const snapi_ali_atom *atom, const snapi_ali_grid *grid,
                                           int ncol, int nd, const double *colstate_dev, const double *vproj_dev,
                                           const double *gamma_dev, const double *bgop_dev, const double *bgem_dev,
                                           const double *coll_dev, double tol, int max_iter, double *pops_dev,
                                           int *iterations_dev, double *relchange_dev, void *stream) {
  if (bad_inputs(atom, grid, ncol, nd, colstate_dev, vproj_dev, gamma_dev, bgop_dev, bgem_dev) || !coll_dev ||
      !pops_dev || !iterations_dev || !relchange_dev || max_iter < 1)
    return SNAPI_ERR_ARG;
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  AliAtom A;
  AliGridArgs G;
  int rc = convert_atom(atom, A);
  if (rc) return rc;
  rc = convert_grid(grid, G);
  if (rc) return rc;
  const size_t nlam = grid->nlam, nt = atom->ntrans;
  if (noalloc_bytes(ncol, nd, atom->nlevel, atom->ntrans, grid->nlam) > kAliScratchBytes) return SNAPI_ERR_CAPACITY;
  unsigned char *base = scratch_address();
  if (!base) return SNAPI_ERR_DEVICE;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // carve the fixed area: grid arrays, sums, work areas (256-byte aligned)
  unsigned char *q = base;
  double *g = reinterpret_cast<double *>(q);
  q += round256((3 + nt) * nlam * 8);
  double *acc = reinterpret_cast<double *>(q);
  q += round256((size_t)ncol * nd * nt * 4 * 8);
  AliWork W;
  W.newp = reinterpret_cast<double *>(q);
  W.rc = W.newp + (size_t)ncol * nd * atom->nlevel;
  q += round256(ali_work_doubles(ncol, nd, atom->nlevel) * 8);
  W.ints = reinterpret_cast<int *>(q);
  AliCol P;
  ACHK(hipMemcpyAsync(g, grid->lam, nlam * 8, hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(g + nlam, grid->lamw, nlam * 8, hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(g + 2 * nlam, grid->fudge, nlam * 8, hipMemcpyHostToDevice, st));
  ACHK(hipMemcpyAsync(g + 3 * nlam, grid->sigma, nt * nlam * 8, hipMemcpyHostToDevice, st));
  P.cs = colstate_dev;
  P.vproj = vproj_dev;
  P.gamma = gamma_dev;
  P.bgop = bgop_dev;
  P.bgem = bgem_dev;
  P.pops = pops_dev;
  P.lam = g;
  P.lamw = g + nlam;
  P.fudge = g + 2 * nlam;
  P.sigma = g + 3 * nlam;
  P.acc = acc;
  P.active = nullptr;
  // every value the iteration reads from the area is written first in this
  // call (init kernel, count kernel, per-iteration zeroing of the sums)
  rc = ali_solve_core(precision, A, G, ncol, nd, P, W, coll_dev, pops_dev, iterations_dev, relchange_dev, tol,
                      max_iter, st);
done:
  return rc;
}

