// Node Jacobian of the emergent Stokes vector by a forward-mode (tangent)
// sweep of the polarized formal solution on gfx950.
//
// The reference differentiates its synthesis along each node direction by
// perturbing the opacity matrices with the analytic local derivatives and
// re-running the formal solver twice per node (obs_stokes_responses_to_nodes,
// atmos/atmos_fit.cc:1205-1352, formal_pert_numerical, plpar/
// atmos_ppbez_rts.cc:1780-1857).  Here one work-item per (column,
// wavelength) carries the derivative of the intensity along every node
// direction through the same recurrence:
//   (1 + w1 K~_k) dI_k = dw0 (S_{k-1} - K~_{k-1} I_{k-1}) + w0 (dS_{k-1} - dK~_{k-1} I_{k-1})
//                      + dw1 (S_k - K~_k I_k) + w1 (dS_k - dK~_k I_k)
//                      + de I_{k-1} + (e - w0 K~_{k-1}) dI_{k-1},
// where dw0, dw1, de follow from the derivative of the optical-depth step
// (Bezier integral with the Fritsch-Butland derivative, whose partials carry
// the one-point look-ahead).  The local derivatives of (K, E) along vt,
// vLOS, |B|, inclination and azimuth restate the analytic part of
// atom::boundbound_op_em_vector_plus_pert (atmos/atmol/atom/
// atom_rad_synthesis.cc:771-1026) including its rho_V inclination term
// (:999); a node direction is the parameter derivative times the node map
// dq(z)/dnode.  Nothing is stored per depth: the recurrence is streamed and
// the one 4x4 factorisation of each depth serves the forward solve and all
// node directions.
#include <hip/hip_runtime.h>

#include <cmath>

#include "resp_kernel.h"
#include "autotune.h"
#include "deriv_kernel.h"
#include "rt_common.h"
#include "synth_kernel.h"

namespace snapi {

// u_q = dS_q - dK~_q I of one depth for every parameter in qmask: the
// derivative of the reduced source terms (K~ = K/eta_I - 1, S = E/eta_I).
template <typename R, unsigned ROWS>
__device__ __forceinline__ void reduced_u(const R G[nrows(ROWS)][11], R eta, const Kred<R> &k, const R S[4],
                                          const R I[4], unsigned qmask, R u[nrows(ROWS)][4], R ds0[nrows(ROWS)]) {
  const R ie = rrcp<R>(eta);
#pragma unroll
  for (int q = 0; q < nrows(ROWS); ++q) {  // q: position in the row set
    if (!(qmask & (1u << row_of(ROWS, q)))) {
      u[q][0] = u[q][1] = u[q][2] = u[q][3] = R(0);
      ds0[q] = R(0);
      continue;
    }
    const R ge = G[q][0];
    Kred<R> dk;
    dk.a = (G[q][1] - k.a * ge) * ie;
    dk.b = (G[q][2] - k.b * ge) * ie;
    dk.c = (G[q][3] - k.c * ge) * ie;
    dk.d = (G[q][6] - k.d * ge) * ie;
    dk.e = (G[q][5] - k.e * ge) * ie;
    dk.f = (G[q][4] - k.f * ge) * ie;
    R y[4];
    kmul<R>(dk, I, y);
#pragma unroll
    for (int s = 0; s < 4; ++s) u[q][s] = (G[q][7 + s] - S[s] * ge) * ie - y[s];
    ds0[q] = (G[q][7] - S[0] * ge) * ie;
  }
}

// A[q][i] for the node's parameter q as a dot product with the node's
// one-hot row oh[q] (uniform over the workgroup): a select chain on q would
// be folded into a dynamic index, which places the private array in scratch
// memory.
template <typename R, int NQR, int N>
__device__ __forceinline__ R pick(const R (&A)[NQR][N], const R oh[NQR], int i) {
  R r = oh[0] * A[0][i];
#pragma unroll
  for (int k = 1; k < NQR; ++k) r += oh[k] * A[k][i];
  return r;
}

template <typename R> __device__ __forceinline__ void sc_weights_dt(R tb, R &detb, R &dw0, R &dw1) {
  if (tb < R(1E-2)) {
    detb = R(-1) + tb - R(0.5) * tb * tb;
    dw0 = R(0.5) - R(2.0 / 3.0) * tb + R(3.0 / 8.0) * tb * tb;
    dw1 = R(0.5) - tb * R(1.0 / 3.0) + tb * tb * R(0.125);
  } else {
    const R e = rexp<R>(-tb), i = rrcp<R>(tb);
    detb = -e;
    dw0 = -i * i + e * (R(1) + i) + e * i * i;
    dw1 = i * i - e * i - e * i * i;
  }
}

template <typename R>
__device__ __forceinline__ R fb_deriv_p(R om, R o0, R op, R h1, R h2, R &gm, R &g0, R &gp) {
  const R ih1 = rrcp<R>(h1), ih2 = rrcp<R>(h2);
  const R d2 = (op - o0) * ih2, d1 = (o0 - om) * ih1;
  gm = g0 = gp = R(0);
  if (d1 * d2 > R(0)) {
    const R al = (R(1) + h1 * rrcp<R>(h1 + h2)) * R(1.0 / 3.0);
    const R den = al * d1 + (R(1) - al) * d2;
    const R iden = rrcp<R>(den);
    const R dD1 = d2 * d2 * (R(1) - al) * iden * iden;
    const R dD2 = d1 * d1 * al * iden * iden;
    gm = -dD1 * ih1;
    g0 = dD1 * ih1 - dD2 * ih2;
    gp = dD2 * ih2;
    return d1 * d2 * iden;
  }
  return R(0);
}

// One workgroup per column of one pass of up to kTanMaxNodes nodes,
// work-items over wavelengths.  Writes the emergent Stokes vector (pass 0)
// and J[col][l*4+s][n] = dI_s/dnode_n * nscale[n] * ws[s] * mask[col][l].
// ROWS: the parameter rows the kernel carries (a superset of the pass's own);
// with the temperature or density row the local derivatives read dstate and
// each node's direction is nmap[n] d/dq_pidx[n] + nmap2[n] d/dq_pidx2[n].
template <typename R, unsigned ROWS>
__device__ __forceinline__ void tangent_body(const KLines &lines, const SynthArgs &args, const RespArgs &ra,
                                             const R *__restrict__ state, const R *__restrict__ dstate,
                                             const double *__restrict__ lam, R *__restrict__ stokes,
                                             R *__restrict__ J, int *__restrict__ status, int pass) {
  extern __shared__ unsigned char smem_raw[];
  R *sm = reinterpret_cast<R *>(smem_raw);
  const int nd = args.nd, nl = args.nlambda, nf = args.nfield;
  const int col = ra.col0 + blockIdx.x;
  if (ra.active && !ra.active[col]) return;  // uniform over the workgroup
  const int n0 = pass * kTanMaxNodes;
  constexpr int NR = nrows(ROWS);
  constexpr bool kTh = has_row(ROWS, Q_T) || has_row(ROWS, Q_N);
  const int ng = min(ra.nnode - n0, kTanMaxNodes);
  const R *st = state + (size_t)col * nf * nd;
  ColumnTile<R> tile = ColumnTile<R>::carve(sm, nd, lines.nline);
  DerivTile<R> dt = DerivTile<R>::carve(sm + ColumnTile<R>::bytes(nd, lines.nline) / sizeof(R), nd, lines.nline, ng);
  ThermoTile<R> tht = ThermoTile<R>::carve(
      sm + (ColumnTile<R>::bytes(nd, lines.nline) + DerivTile<R>::bytes(nd, lines.nline, ng, kTh)) / sizeof(R), nd,
      lines.nline);
  stage_column<R, R>(lines, args, st, tile);
  stage_deriv<R, R>(lines, args, st, dt);
  if constexpr (kTh) stage_thermo<R, R>(lines, args, st, dstate + (size_t)col * 2 * nf * nd, tht);
  const R *nmg = static_cast<const R *>(ra.nmap) + ((size_t)col * ra.nnode + n0) * nd;
  for (int i = threadIdx.x; i < ng * nd; i += blockDim.x) dt.nm[i] = nmg[i];
  if constexpr (kTh) {
    const R *nmg2 = ra.nmap2 ? static_cast<const R *>(ra.nmap2) + ((size_t)col * ra.nnode + n0) * nd : nullptr;
    for (int i = threadIdx.x; i < ng * nd; i += blockDim.x) {
      const int j = i / nd;
      dt.nm2[i] = (nmg2 && ra.pidx2[n0 + j] >= 0) ? nmg2[i] : R(0);
    }
  }
  __syncthreads();
  // one-hot parameter rows of each node of the pass (the second: the
  // density row of a temperature node)
  R oh[kTanMaxNodes][NR], oh2[kTh ? kTanMaxNodes : 1][NR];
  unsigned qmask = 0;
#pragma unroll
  for (int j = 0; j < kTanMaxNodes; ++j) {
    const int q = j < ng ? qrow(ra.pidx[n0 + j]) : -1;
#pragma unroll
    for (int k = 0; k < NR; ++k) oh[j][k] = (q == row_of(ROWS, k)) ? R(1) : R(0);
    if (q >= 0) qmask |= 1u << q;
    if constexpr (kTh) {
      const int q2 = (j < ng && ra.nmap2 && ra.pidx2[n0 + j] >= 0) ? qrow(ra.pidx2[n0 + j]) : -1;
#pragma unroll
      for (int k = 0; k < NR; ++k) oh2[j][k] = (q2 == row_of(ROWS, k)) ? R(1) : R(0);
      if (q2 >= 0) qmask |= 1u << q2;
    }
  }
  // node j's derivative of entry i of the rows A at depth z
  auto ndir = [&](int j, int z, const auto &A, int i) {
    R r = dt.nm[j * nd + z] * pick<R>(A, oh[j], i);
    if constexpr (kTh) r += dt.nm2[j * nd + z] * pick<R>(A, oh2[j], i);
    return r;
  };
  const bool up = args.ct < 0.0;
  const R mct = (R)(-args.ct);
  auto zof = [&](int k) { return up ? nd - 1 - k : k; };
  const R *mk_all = static_cast<const R *>(ra.mask);
  // wavelengths in index order, as the synthesis kernel runs them
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    LamConst<R> c;
    lambda_consts<R>(lines, args, lam[l], c);
    R G[NR][11], u[NR][4], ds0[NR];
    R K0[7], E0[4], K1[7], E1[4];
    // bottom point: full derivatives, boundary I = (S_I, 0, 0, 0)
    const int z0 = zof(0), z1 = zof(1);
    deriv_at<R, true, ROWS>(lines, args, tile, dt, tht, c, z0, qmask, K0, E0, G);
    Kred<R> kp, kc;
    R Sp[4], Sc[4];
    reduce_ks<R>(K0, E0, kp, Sp);
    R Ip[4] = {up ? E0[0] / K0[0] : R(0), R(0), R(0), R(0)};
    reduced_u<R, ROWS>(G, K0[0], kp, Sp, Ip, qmask, u, ds0);
    R dI[kTanMaxNodes][4], v[kTanMaxNodes][4], dem[kTanMaxNodes], de0[kTanMaxNodes], dDm[kTanMaxNodes];
    R ds0r[NR][1];
#pragma unroll
    for (int k = 0; k < NR; ++k) ds0r[k][0] = ds0[k];
#pragma unroll
    for (int j = 0; j < kTanMaxNodes; ++j) {
      const bool on = j < ng;
      dem[j] = on ? ndir(j, z0, G, 0) : R(0);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        v[j][s] = on ? ndir(j, z0, u, s) : R(0);
        dI[j][s] = R(0);
      }
      // the bottom boundary I = S_I: its derivative is that of S_I
      R d0 = R(0);
      if (on) {
        d0 = dt.nm[j * nd + z0] * pick<R>(ds0r, oh[j], 0);
        if constexpr (kTh) d0 += dt.nm2[j * nd + z0] * pick<R>(ds0r, oh2[j], 0);
      }
      dI[j][0] = up ? d0 : R(0);
    }
    deriv_at<R, false, ROWS>(lines, args, tile, dt, tht, c, z1, qmask, K1, E1, G);
    reduce_ks<R>(K1, E1, kc, Sc);
    R oom = K0[0], oo0 = K1[0];
    R g0 = tile.grid[z1];
    R h0 = (g0 - tile.grid[z0]) / mct;
    R Dm = (oo0 - oom) / h0;
#pragma unroll
    for (int j = 0; j < kTanMaxNodes; ++j) {
      de0[j] = j < ng ? ndir(j, z1, G, 0) : R(0);
      dDm[j] = (de0[j] - dem[j]) / h0;
    }
    int rc = 0;
    for (int k = 1; k <= nd - 2; ++k) {
      const int zk = zof(k), zn = zof(k + 1);
      R Kn[7], En[4];
      deriv_at<R, false, ROWS>(lines, args, tile, dt, tht, c, zn, qmask, Kn, En, G);
      R dep[kTanMaxNodes];
#pragma unroll
      for (int j = 0; j < kTanMaxNodes; ++j) dep[j] = j < ng ? ndir(j, zn, G, 0) : R(0);
      const R oop = Kn[0];
      const R gp = tile.grid[zn];
      const R hp = (gp - g0) / mct;
      R fgm, fg0, fgp;
      const R D0 = fb_deriv_p<R>(oom, oo0, oop, h0, hp, fgm, fg0, fgp);
      const R C = R(0.5) * ((oo0 - h0 * R(0.5) * D0) + (oom + h0 * R(0.5) * Dm));
      const R tb = h0 * R(1.0 / 3.0) * (oom + oo0 + C);
      if (!(tb < R(1E10))) rc |= 1;
      R etb, w0, w1, detb, dw0, dw1;
      sc_weights<R>(tb, etb, w0, w1);
      sc_weights_dt<R>(tb, detb, dw0, dw1);
      R kv[4], B[4], In[4];
      kmul<R>(kp, Ip, kv);
#pragma unroll
      for (int s = 0; s < 4; ++s) B[s] = -w0 * kv[s] + w0 * Sp[s] + w1 * Sc[s] + Ip[s] * etb;
      LU4<R> f;
      factor_1pk<R>(kc, w1, f);
      solve_lu<R>(f, B, In);
      R Pp[4], Pc[4], kvc[4];
      kmul<R>(kc, In, kvc);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        Pp[s] = Sp[s] - kv[s];
        Pc[s] = Sc[s] - kvc[s];
      }
      // full derivatives of depth k, reduced with the new intensity
      {
        R Kd[7], Ed[4];
        deriv_at<R, true, ROWS>(lines, args, tile, dt, tht, c, zk, qmask, Kd, Ed, G);
        reduced_u<R, ROWS>(G, oo0, kc, Sc, In, qmask, u, ds0);
      }
      const R hh = h0 * h0 * R(1.0 / 12.0);
#pragma unroll
      for (int j = 0; j < kTanMaxNodes; ++j) {
        if (j >= ng) continue;
        const R dD0 = fgm * dem[j] + fg0 * de0[j] + fgp * dep[j];
        const R dtb = h0 * R(0.5) * (dem[j] + de0[j]) + hh * (dDm[j] - dD0);
        R kvd[4], rhs[4];
        kmul<R>(kp, dI[j], kvd);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const R us = ndir(j, zk, u, s);
          rhs[s] = dtb * (dw0 * Pp[s] + dw1 * Pc[s] + detb * Ip[s]) + w0 * v[j][s] + w1 * us + etb * dI[j][s] -
                   w0 * kvd[s];
          v[j][s] = us;
        }
        solve_lu<R>(f, rhs, dI[j]);
        dem[j] = de0[j];
        de0[j] = dep[j];
        dDm[j] = dD0;
      }
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        Ip[s] = In[s];
        Sp[s] = Sc[s];
      }
      kp = kc;
      reduce_ks<R>(Kn, En, kc, Sc);
      oom = oo0;
      oo0 = oop;
      Dm = D0;
      h0 = hp;
      g0 = gp;
    }
    const size_t o = (size_t)col * nl + l;
    if (pass == 0 && stokes)
      for (int s = 0; s < 4; ++s) stokes[o * 4 + s] = Ip[s];
    if (!(fabs((double)Ip[0]) < 1e300)) rc |= 2;
    if (rc) atomicOr(status, rc);
    const R mk = mk_all ? mk_all[o] : R(1);
#pragma unroll
    for (int j = 0; j < kTanMaxNodes; ++j) {
      if (j >= ng) continue;
      const int n = n0 + j;
      for (int s = 0; s < 4; ++s) J[(o * 4 + s) * ra.nnode + n] = dI[j][s] * (R)(ra.nscale[n] * ra.ws[s]) * mk;
    }
  }
}

// The float kernels are held to two wavefronts per SIMD (256 registers, some
// spilled): their registers would otherwise leave one wavefront per SIMD, and
// the sweep waits on memory and LDS; measured 1.34x faster on the inversion
// workload than at four wavefronts (128 VGPRs, heavy spilling).  The double
// kernels keep their registers.  Each kernel carries one of four row sets,
// the smallest that holds the pass's parameters is launched (fewer rows, fewer
// live registers): all five kinematic and magnetic rows, the field rows
// without the turbulent velocity, all seven rows, or turbulent velocity with
// temperature and density (a pass of temperature nodes).  Each comes in
// workgroups of 64, 128 and 256 work-items (one column, its wavelengths over
// the work-items), among which the autotuner picks per problem shape
// (autotune.h): larger workgroups share the column's staged tiles over more
// wavelengths at once, smaller ones keep more columns resident when the tiles
// fill the LDS.
constexpr unsigned kRowsField = (1u << Q_VL) | (1u << Q_B) | (1u << Q_TH) | (1u << Q_PH);
constexpr unsigned kRowsTemp = (1u << Q_VT) | (1u << Q_T) | (1u << Q_N);
constexpr int kTanSets = 4;
constexpr unsigned kTanRowSets[kTanSets] = {kRowsKin, kRowsField, kRowsAll, kRowsTemp};

#define SNAPI_TANGENT_KERNEL(NAME, R, ROWS, BOUNDS)                                                                \
  __global__ BOUNDS void NAME(const KLines lines, const SynthArgs args, const RespArgs ra,                        \
                              const R *__restrict__ state, const R *__restrict__ dstate,                          \
                              const double *__restrict__ lam, R *__restrict__ stokes, R *__restrict__ J,         \
                              int *__restrict__ status, int pass) {                                               \
    tangent_body<R, ROWS>(lines, args, ra, state, dstate, lam, stokes, J, status, pass);                          \
  }
#define SNAPI_TANGENT_SET(SUF, ROWS)                                                                              \
  SNAPI_TANGENT_KERNEL(tangent_kernel_##SUF##_f, float, ROWS, __launch_bounds__(64, 2))                          \
  SNAPI_TANGENT_KERNEL(tangent_kernel_##SUF##_f128, float, ROWS, __launch_bounds__(128, 2))                      \
  SNAPI_TANGENT_KERNEL(tangent_kernel_##SUF##_f256, float, ROWS, __launch_bounds__(256, 2))                      \
  SNAPI_TANGENT_KERNEL(tangent_kernel_##SUF##_d, double, ROWS, __launch_bounds__(64))                            \
  SNAPI_TANGENT_KERNEL(tangent_kernel_##SUF##_d128, double, ROWS, __launch_bounds__(128))                        \
  SNAPI_TANGENT_KERNEL(tangent_kernel_##SUF##_d256, double, ROWS, __launch_bounds__(256))
SNAPI_TANGENT_SET(kin, kRowsKin)
SNAPI_TANGENT_SET(field, kRowsField)
SNAPI_TANGENT_SET(all, kRowsAll)
SNAPI_TANGENT_SET(temp, kRowsTemp)
#undef SNAPI_TANGENT_SET
#undef SNAPI_TANGENT_KERNEL

template <typename R>
using TanFn = void (*)(const KLines, const SynthArgs, const RespArgs, const R *, const R *, const double *, R *, R *,
                       int *, int);
// Candidates of the autotuner (workgroup sizes), the default first.
constexpr int kTanCand = 3;
constexpr int kTanCandThreads[kTanCand] = {64, 128, 256};
static TanFn<float> tan_fn(float *, int set, int c) {
  static const TanFn<float> t[kTanSets][kTanCand] = {
      {tangent_kernel_kin_f, tangent_kernel_kin_f128, tangent_kernel_kin_f256},
      {tangent_kernel_field_f, tangent_kernel_field_f128, tangent_kernel_field_f256},
      {tangent_kernel_all_f, tangent_kernel_all_f128, tangent_kernel_all_f256},
      {tangent_kernel_temp_f, tangent_kernel_temp_f128, tangent_kernel_temp_f256}};
  return t[set][c];
}
static TanFn<double> tan_fn(double *, int set, int c) {
  static const TanFn<double> t[kTanSets][kTanCand] = {
      {tangent_kernel_kin_d, tangent_kernel_kin_d128, tangent_kernel_kin_d256},
      {tangent_kernel_field_d, tangent_kernel_field_d128, tangent_kernel_field_d256},
      {tangent_kernel_all_d, tangent_kernel_all_d128, tangent_kernel_all_d256},
      {tangent_kernel_temp_d, tangent_kernel_temp_d128, tangent_kernel_temp_d256}};
  return t[set][c];
}
// The row set to launch for a pass whose parameter rows are `rows`: the
// smallest that holds them (the sets' sizes: 5, 4, 7, 3).
static int tan_row_set(unsigned rows) {
  int best = -1;
  for (int s = 0; s < kTanSets; ++s)
    if ((rows & ~kTanRowSets[s]) == 0 &&
        (best < 0 || __builtin_popcount(kTanRowSets[s]) < __builtin_popcount(kTanRowSets[best])))
      best = s;
  return best;
}
// Columns of the autotuner's sample.
constexpr int kTanTuneColumns = 16384;

template <typename R>
int launch_tangent(const KLines &kl, const SynthArgs &a, RespArgs ra, int ncol, const void *ds_, const void *dd_,
                   const double *dl, void *dstk_, void *dJ_, hipStream_t stream) {
  if (ra.nnode <= 0 || ra.nnode > SNAPI_MAX_NODES || !ra.nmap) return SNAPI_ERR_ARG;
  bool thermo = false;
  for (int n = 0; n < ra.nnode; ++n) {
    if (ra.pidx[n] < 0 || ra.pidx[n] > 6) return SNAPI_ERR_ARG;
    thermo |= ra.pidx[n] < 2 || (ra.nmap2 && ra.pidx2[n] >= 0);
  }
  // temperature and density directions need the thermodynamic derivatives,
  // and H- at the middle wavelength (the response path's) to keep the
  // background's derivatives wavelength-independent
  if (thermo && (!dd_ || (a.hminus && !a.hminus_mid))) return SNAPI_ERR_UNSUPPORTED;
  const int wthreads = ((a.nlambda + 63) / 64) * 64;
  // candidates whose workgroup the wavelengths fill
  int ncand = 1;
  while (ncand < kTanCand && kTanCandThreads[ncand] <= wthreads) ++ncand;
  const int npass = (ra.nnode + kTanMaxNodes - 1) / kTanMaxNodes;
  int *status = device_status_slot();
  if (!status) return SNAPI_ERR_DEVICE;
  ra.col0 = 0;
  for (int pass = 0; pass < npass; ++pass) {
    unsigned rows = 0;
    for (int n = pass * kTanMaxNodes; n < ra.nnode && n < (pass + 1) * kTanMaxNodes; ++n) {
      rows |= 1u << qrow(ra.pidx[n]);
      if (ra.nmap2 && ra.pidx2[n] >= 0) rows |= 1u << qrow(ra.pidx2[n]);
    }
    const int set = tan_row_set(rows);
    if (set < 0) return SNAPI_ERR_ARG;
    const bool th = (kTanRowSets[set] & ((1u << Q_T) | (1u << Q_N))) != 0;
    const int ng = ra.nnode - pass * kTanMaxNodes < kTanMaxNodes ? ra.nnode - pass * kTanMaxNodes : kTanMaxNodes;
    const size_t smem = ColumnTile<R>::bytes(a.nd, kl.nline) + DerivTile<R>::bytes(a.nd, kl.nline, ng, th) +
                        (th ? ThermoTile<R>::bytes(a.nd, kl.nline) : 0);
    if (smem > 65536) return SNAPI_ERR_UNSUPPORTED;
    const R *st = static_cast<const R *>(ds_), *dd = static_cast<const R *>(dd_);
    R *so = static_cast<R *>(dstk_), *Jo = static_cast<R *>(dJ_);
    auto launch = [&](int c, int nc) {
      const int threads = kTanCandThreads[c] < wthreads ? kTanCandThreads[c] : wthreads;
      hipLaunchKernelGGL(tan_fn((R *)nullptr, set, c), dim3(nc), dim3(threads), smem, stream, kl, a, ra, st, dd, dl,
                         so, Jo, status, pass);
      return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
    };
    const TuneKey key{TUNE_TANGENT, sizeof(R) == 8 ? SNAPI_FP64 : SNAPI_MIXED, a.nd, kl.nline, kl.ncomp, a.nlambda,
                      (tune_size_class(ncol, kTanTuneColumns) << 8) + (set << 4) + ng};
    const int ns = ncol < kTanTuneColumns ? ncol : kTanTuneColumns;
    const int c = tune_choose(key, ncand, 0, stream, true, [&](int cand) { return launch(cand, ns); });
    const int rc = launch(c, ncol);
    if (rc != SNAPI_OK) return rc;
  }
  return SNAPI_OK;
}

template int launch_tangent<double>(const KLines &, const SynthArgs &, RespArgs, int, const void *, const void *,
                                    const double *, void *, void *, hipStream_t);
template int launch_tangent<float>(const KLines &, const SynthArgs &, RespArgs, int, const void *, const void *,
                                   const double *, void *, void *, hipStream_t);

}  // namespace snapi
