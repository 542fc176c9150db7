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
#include "rt_common.h"
#include "synth_kernel.h"

namespace snapi {

constexpr int kTanThreads = 64;
enum { Q_VT = 0, Q_VL = 1, Q_B = 2, Q_TH = 3, Q_PH = 4, NQ = 5 };

// LDS tile of the derivative factors of one column, after the ColumnTile.
template <typename R> struct DerivTile {
  R *sct, *dqt, *dut, *mhs;  // per depth: sin(th)cos(th), d qf/d th, d uf/d th, -sin(th)/2
  R *gv;                     // per line and depth: d ln(Doppler width) / d vt
  R *nm;                     // [node of the pass][z]: node map d q(z) / d node
  static __host__ __device__ size_t bytes(int nd, int nline, int nn) {
    return sizeof(R) * (size_t)nd * (4 + nline + nn);
  }
  static __device__ DerivTile carve(R *p, int nd, int nline) {
    DerivTile t;
    t.sct = p;
    t.dqt = p + nd;
    t.dut = p + 2 * nd;
    t.mhs = p + 3 * nd;
    t.gv = p + 4 * nd;
    t.nm = p + (size_t)(4 + nline) * nd;
    return t;
  }
};

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
template <typename R, bool FULL>
__device__ __forceinline__ void deriv_at(const KLines &L, const SynthArgs &a, const ColumnTile<R> &t,
                                         const DerivTile<R> &d, const LamConst<R> &c, int z, unsigned qmask,
                                         R K[7], R E[4], R G[NQ][11]) {
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
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int i = 0; i < 11; ++i) G[q][i] = R(0);
  const bool wvt = qmask & (1u << Q_VT), wvl = qmask & (1u << Q_VL), wb = qmask & (1u << Q_B);
  const bool wth = qmask & (1u << Q_TH), wph = qmask & (1u << Q_PH);
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
        G[Q_VT][0] += opl * gt;
        if (FULL) G[Q_VT][7] += eml * gt;
      }
      if (wvl) {
        G[Q_VL][0] += opl * gl;
        if (FULL) G[Q_VL][7] += eml * gl;
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
    auto add = [&](const R h[3], const R f[3], R *Kq, bool full) {
      const R hbr = h[2] + h[1], fbr = f[2] + f[1];
      const R eI = s2h * h[0] + c2q * hbr;
      Kq[0] += eI * opl;
      if (!full) return;
      const R hl = h[0] - R(0.5) * hbr, fl = f[0] - R(0.5) * fbr;
      const R eQ = qf * hl, eU = uf * hl, eV = vf * (h[2] - h[1]);
      Kq[1] += eQ * opl;
      Kq[2] += eU * opl;
      Kq[3] += eV * opl;
      Kq[4] += qf * fl * opl;
      Kq[5] += uf * fl * opl;
      Kq[6] += vf * (f[2] - f[1]) * opl;
      Kq[7] += eI * eml;
      Kq[8] += eQ * eml;
      Kq[9] += eU * eml;
      Kq[10] += eV * eml;
    };
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
      add(h, f, G[Q_VT], FULL);
    }
    if (wvl) {
      R h[3], f[3];
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        h[g] = cvl * HX[g];
        f[g] = cvl * FX[g];
      }
      add(h, f, G[Q_VL], FULL);
    }
    if (wb) {
      R h[3], f[3];
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        h[g] = cB * HXS[g];
        f[g] = cB * FXS[g];
      }
      add(h, f, G[Q_B], FULL);
    }
    const R hl = HS[0] - R(0.5) * (HS[2] + HS[1]), fl = FS[0] - R(0.5) * (FS[2] + FS[1]);
    if (wth) {
      // d/d inclination; rho_V as the reference writes it (cos, :999)
      const R sct = d.sct[z], dqt = d.dqt[z], dut = d.dut[z], mhs = d.mhs[z];
      const R eI = sct * hl;
      G[Q_TH][0] += eI * opl;
      if (FULL) {
        const R eQ = dqt * hl, eU = dut * hl, eV = mhs * (HS[2] - HS[1]);
        G[Q_TH][1] += eQ * opl;
        G[Q_TH][2] += eU * opl;
        G[Q_TH][3] += eV * opl;
        G[Q_TH][4] += dqt * fl * opl;
        G[Q_TH][5] += dut * fl * opl;
        G[Q_TH][6] += vf * (FS[2] - FS[1]) * opl;
        G[Q_TH][7] += eI * eml;
        G[Q_TH][8] += eQ * eml;
        G[Q_TH][9] += eU * eml;
        G[Q_TH][10] += eV * eml;
      }
    }
    if (wph && FULL) {
      // d/d azimuth: d qf = -2 uf, d uf = 2 qf
      G[Q_PH][1] += R(-2) * uf * hl * opl;
      G[Q_PH][2] += R(2) * qf * hl * opl;
      G[Q_PH][4] += R(-2) * uf * fl * opl;
      G[Q_PH][5] += R(2) * qf * fl * opl;
      G[Q_PH][8] += R(-2) * uf * hl * eml;
      G[Q_PH][9] += R(2) * qf * hl * eml;
    }
  }
  K[0] = KI; K[1] = KQ; K[2] = KU; K[3] = KV; K[4] = RQ; K[5] = RU; K[6] = RV;
  E[0] = EI; E[1] = EQ; E[2] = EU; E[3] = EV;
}

// u_q = dS_q - dK~_q I of one depth for every parameter in qmask: the
// derivative of the reduced source terms (K~ = K/eta_I - 1, S = E/eta_I).
template <typename R>
__device__ __forceinline__ void reduced_u(const R G[NQ][11], R eta, const Kred<R> &k, const R S[4], const R I[4],
                                          unsigned qmask, R u[NQ][4], R ds0[NQ]) {
  const R ie = rrcp<R>(eta);
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    if (!(qmask & (1u << q))) {
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
template <typename R, int N> __device__ __forceinline__ R pick(const R (&A)[NQ][N], const R oh[NQ], int i) {
  R r = oh[0] * A[0][i];
#pragma unroll
  for (int k = 1; k < NQ; ++k) r += oh[k] * A[k][i];
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

// One workgroup per (column, pass of up to kTanMaxNodes nodes), work-items
// over wavelengths.  Writes the emergent Stokes vector (pass 0) and
// J[col][l*4+s][n] = dI_s/dnode_n * nscale[n] * ws[s] * mask[col][l].
template <typename R>
__device__ __forceinline__ void tangent_body(const KLines &lines, const SynthArgs &args, const RespArgs &ra,
                                             const R *__restrict__ state, const double *__restrict__ lam,
                                             R *__restrict__ stokes, R *__restrict__ J, int *__restrict__ status) {
  extern __shared__ unsigned char smem_raw[];
  R *sm = reinterpret_cast<R *>(smem_raw);
  const int nd = args.nd, nl = args.nlambda, nf = args.nfield;
  const int col = ra.col0 + blockIdx.x;
  if (ra.active && !ra.active[col]) return;  // uniform over the workgroup
  const int n0 = blockIdx.y * kTanMaxNodes;
  const int ng = min(ra.nnode - n0, kTanMaxNodes);
  const R *st = state + (size_t)col * nf * nd;
  ColumnTile<R> tile = ColumnTile<R>::carve(sm, nd, lines.nline);
  DerivTile<R> dt = DerivTile<R>::carve(sm + ColumnTile<R>::bytes(nd, lines.nline) / sizeof(R), nd, lines.nline);
  // the synthesis kernel's grouped wavelength order (float path)
  int *order = nullptr;
  if (sizeof(R) == 4 && nl <= kOrderMax) {
    order = reinterpret_cast<int *>(smem_raw + ColumnTile<R>::bytes(nd, lines.nline) +
                                    DerivTile<R>::bytes(nd, lines.nline, ng));
    lambda_order(lines, lam, nl, reinterpret_cast<float *>(order + nl), order + 2 * nl, order);
  }
  stage_column<R, R>(lines, args, st, tile);
  stage_deriv<R, R>(lines, args, st, dt);
  const R *nmg = static_cast<const R *>(ra.nmap) + ((size_t)col * ra.nnode + n0) * nd;
  for (int i = threadIdx.x; i < ng * nd; i += blockDim.x) dt.nm[i] = nmg[i];
  __syncthreads();
  R oh[kTanMaxNodes][NQ];  // one-hot parameter row of each node of the pass
  unsigned qmask = 0;
#pragma unroll
  for (int j = 0; j < kTanMaxNodes; ++j) {
    const int q = j < ng ? ra.pidx[n0 + j] - 2 : -1;
#pragma unroll
    for (int k = 0; k < NQ; ++k) oh[j][k] = (q == k) ? R(1) : R(0);
    if (q >= 0) qmask |= 1u << q;
  }
  const bool up = args.ct < 0.0;
  const R mct = (R)(-args.ct);
  auto zof = [&](int k) { return up ? nd - 1 - k : k; };
  const R *mk_all = static_cast<const R *>(ra.mask);
  for (int i = threadIdx.x; i < nl; i += blockDim.x) {
    const int l = order ? order[i] : i;
    LamConst<R> c;
    lambda_consts<R>(lines, args, lam[l], c);
    R G[NQ][11], u[NQ][4], ds0[NQ];
    R K0[7], E0[4], K1[7], E1[4];
    // bottom point: full derivatives, boundary I = (S_I, 0, 0, 0)
    const int z0 = zof(0), z1 = zof(1);
    deriv_at<R, true>(lines, args, tile, dt, c, z0, qmask, K0, E0, G);
    Kred<R> kp, kc;
    R Sp[4], Sc[4];
    reduce_ks<R>(K0, E0, kp, Sp);
    R Ip[4] = {up ? E0[0] / K0[0] : R(0), R(0), R(0), R(0)};
    reduced_u<R>(G, K0[0], kp, Sp, Ip, qmask, u, ds0);
    R dI[kTanMaxNodes][4], v[kTanMaxNodes][4], dem[kTanMaxNodes], de0[kTanMaxNodes], dDm[kTanMaxNodes];
#pragma unroll
    for (int j = 0; j < kTanMaxNodes; ++j) {
      const R m = j < ng ? dt.nm[j * nd + z0] : R(0);
      dem[j] = m * pick<R, 11>(G, oh[j], 0);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        v[j][s] = m * pick<R, 4>(u, oh[j], s);
        dI[j][s] = R(0);
      }
      R d0 = R(0);
#pragma unroll
      for (int k = 0; k < NQ; ++k) d0 += oh[j][k] * ds0[k];
      dI[j][0] = up ? m * d0 : R(0);
    }
    deriv_at<R, false>(lines, args, tile, dt, c, z1, qmask, K1, E1, G);
    reduce_ks<R>(K1, E1, kc, Sc);
    R oom = K0[0], oo0 = K1[0];
    R g0 = tile.grid[z1];
    R h0 = (g0 - tile.grid[z0]) / mct;
    R Dm = (oo0 - oom) / h0;
#pragma unroll
    for (int j = 0; j < kTanMaxNodes; ++j) {
      const R m = j < ng ? dt.nm[j * nd + z1] : R(0);
      de0[j] = m * pick<R, 11>(G, oh[j], 0);
      dDm[j] = (de0[j] - dem[j]) / h0;
    }
    int rc = 0;
    for (int k = 1; k <= nd - 2; ++k) {
      const int zk = zof(k), zn = zof(k + 1);
      R Kn[7], En[4];
      deriv_at<R, false>(lines, args, tile, dt, c, zn, qmask, Kn, En, G);
      R dep[kTanMaxNodes];
#pragma unroll
      for (int j = 0; j < kTanMaxNodes; ++j) dep[j] = (j < ng ? dt.nm[j * nd + zn] : R(0)) * pick<R, 11>(G, oh[j], 0);
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
        deriv_at<R, true>(lines, args, tile, dt, c, zk, qmask, Kd, Ed, G);
        reduced_u<R>(G, oo0, kc, Sc, In, qmask, u, ds0);
      }
      const R hh = h0 * h0 * R(1.0 / 12.0);
#pragma unroll
      for (int j = 0; j < kTanMaxNodes; ++j) {
        if (j >= ng) continue;
        const R m = dt.nm[j * nd + zk];
          const R dD0 = fgm * dem[j] + fg0 * de0[j] + fgp * dep[j];
        const R dtb = h0 * R(0.5) * (dem[j] + de0[j]) + hh * (dDm[j] - dD0);
        R kvd[4], rhs[4];
        kmul<R>(kp, dI[j], kvd);
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const R us = m * pick<R, 4>(u, oh[j], s);
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
    if (blockIdx.y == 0 && stokes)
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

// The float kernel is held to two wavefronts per SIMD (256 registers, some
// spilled): its 392 registers would otherwise leave one wavefront per SIMD,
// and the sweep waits on memory and LDS; measured 1.34x faster on the
// inversion workload (four wavefronts, 128 VGPRs with 393 spilled, measured
// 0.7x).  The double kernel keeps its registers.
__global__ __launch_bounds__(kTanThreads, 2) void tangent_kernel_f(const KLines lines, const SynthArgs args,
                                                                  const RespArgs ra, const float *__restrict__ state,
                                                                  const double *__restrict__ lam,
                                                                  float *__restrict__ stokes, float *__restrict__ J,
                                                                  int *__restrict__ status) {
  tangent_body<float>(lines, args, ra, state, lam, stokes, J, status);
}
__global__ __launch_bounds__(kTanThreads) void tangent_kernel_d(const KLines lines, const SynthArgs args,
                                                               const RespArgs ra, const double *__restrict__ state,
                                                               const double *__restrict__ lam,
                                                               double *__restrict__ stokes, double *__restrict__ J,
                                                               int *__restrict__ status) {
  tangent_body<double>(lines, args, ra, state, lam, stokes, J, status);
}

template <typename R>
int launch_tangent(const KLines &kl, const SynthArgs &a, RespArgs ra, int ncol, const void *ds_, const double *dl,
                   void *dstk_, void *dJ_, hipStream_t stream) {
  if (ra.nnode <= 0 || ra.nnode > SNAPI_MAX_NODES || !ra.nmap) return SNAPI_ERR_ARG;
  for (int n = 0; n < ra.nnode; ++n)
    if (ra.pidx[n] < 2 || ra.pidx[n] > 6) return SNAPI_ERR_UNSUPPORTED;  // T and Nt need the equation of state
  int threads = ((a.nlambda + 63) / 64) * 64;
  if (threads > kTanThreads) threads = kTanThreads;
  const int npass = (ra.nnode + kTanMaxNodes - 1) / kTanMaxNodes;
  const int ng = ra.nnode < kTanMaxNodes ? ra.nnode : kTanMaxNodes;
  const size_t smem = ColumnTile<R>::bytes(a.nd, kl.nline) + DerivTile<R>::bytes(a.nd, kl.nline, ng) +
                      (sizeof(R) == 4 && a.nlambda <= kOrderMax ? (size_t)a.nlambda * 12 : 0);
  if (smem > 65536) return SNAPI_ERR_UNSUPPORTED;
  int *status = device_status_slot();
  if (!status) return SNAPI_ERR_DEVICE;
  ra.col0 = 0;
  if (sizeof(R) == 4)
    hipLaunchKernelGGL(tangent_kernel_f, dim3(ncol, npass), dim3(threads), smem, stream, kl, a, ra,
                       static_cast<const float *>(ds_), dl, static_cast<float *>(dstk_), static_cast<float *>(dJ_),
                       status);
  else
    hipLaunchKernelGGL(tangent_kernel_d, dim3(ncol, npass), dim3(threads), smem, stream, kl, a, ra,
                       static_cast<const double *>(ds_), dl, static_cast<double *>(dstk_), static_cast<double *>(dJ_),
                       status);
  if (hipGetLastError() != hipSuccess) return SNAPI_ERR_DEVICE;
  return SNAPI_OK;
}

template int launch_tangent<double>(const KLines &, const SynthArgs &, RespArgs, int, const void *, const double *,
                                    void *, void *, hipStream_t);
template int launch_tangent<float>(const KLines &, const SynthArgs &, RespArgs, int, const void *, const double *,
                                   void *, void *, hipStream_t);

}  // namespace snapi
