// Batched Levenberg-Marquardt inversion of many pixels on one GPU.
//
// Restates atmosphere::stokes_lm_fit (atmos/atmos_fit.cc:18-227) for the
// magnetic and kinematic node parameters (microturbulence, line-of-sight
// velocity, |B|, inclination, azimuth), with the column's thermodynamic state
// held fixed:
//   - node values -> depth stratification by the reference's monotone cubic
//     Bezier node interpolation (atmospheric_interpolation, cmlib/
//     mathtools.cc:1208-1315) on log tau; node responses by its central
//     difference of the interpolation (atmos_fit.cc:1250-1305);
//   - Jacobian by the forward-mode sweep over the node directions
//     (tangent.hip), by the adjoint sweep (responses.hip) contracted with the
//     node responses, or by central differences of the node values through
//     the synthesis (2 n + 1 syntheses);
//   - scaled Jacobian and residual (scale_rf, calc_residual :229-260), J^T J
//     on the matrix cores, multiplicative damping and the truncated-SVD solve
//     in FP64 (lm.hip), corrections scaled back and bracketed
//     (scale_corrections, model::bracket_parameter_values,
//     atmos_model.cc:957-977);
//   - accept if chi^2 decreases (lambda / sqrt(10), floor 1e-5), else
//     lambda * sqrt(10) and the Jacobian is reused; stop after three accepted
//     steps with relative chi^2 change below 1e-3 or at the iteration limit
//     (:163-181).
// Pixels live in a fixed set of slots; a pixel that finishes hands its slot
// to the next pixel of a device-side queue, so uneven iteration counts keep
// the launches full.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lm_kernel.h"
#include "resp_kernel.h"
#include "rt_common.h"
#include "synth_kernel.h"
#include "chem_device.h"

extern "C" int snapi_synth_stokes_dev(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol,
                                      int nd, const void *state_dev, int nlambda, const double *lam_dev,
                                      void *stokes_dev, void *stream);

namespace snapi {

// reference model parameter codes (atmos.h model, atmos_model.cc:630-710)
enum { NP_T = 1, NP_VT = 2, NP_VS = 3, NP_B = 4, NP_THETA = 5, NP_PHI = 6 };

struct NodeArgs {
  int nnode;
  int code[SNAPI_MAX_NODES];      // NP_*
  double ltau[SNAPI_MAX_NODES];   // node positions, log tau_500
  double scale[SNAPI_MAX_NODES];  // T 5000, vt 1e5, vs 1e5, B 1000, angles 1 (const.h:35-38)
  double step[SNAPI_MAX_NODES];   // node-response steps: vt/vs 20 cm/s, B 0.01 G, angles 1e-3 rad (const.h:27-32)
  int analytic;                   // node responses by exact derivatives instead of the reference's differences
};

// Forward-mode dual number: a value and its derivative along one node
// direction.  The node interpolation and the field projection are written
// once over V = double or Dual, so the analytic node responses follow the
// same arithmetic as the values.
struct Dual {
  double v, d;
};
__device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ inline Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ inline Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ inline Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ inline Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ inline Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ inline Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ inline Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
// the overloads below would hide the double ones inside this namespace
using ::acos;
using ::atan;
using ::cos;
using ::sin;
using ::sqrt;
__device__ inline double val(double a) { return a; }
__device__ inline double val(Dual a) { return a.v; }
__device__ inline Dual sin(Dual a) { return {sin(a.v), cos(a.v) * a.d}; }
__device__ inline Dual cos(Dual a) { return {cos(a.v), -sin(a.v) * a.d}; }
__device__ inline Dual sqrt(Dual a) {
  const double r = sqrt(a.v);
  return {r, 0.5 * a.d / r};
}
__device__ inline Dual acos(Dual a) { return {acos(a.v), -a.d / sqrt(1.0 - a.v * a.v)}; }
__device__ inline Dual atan(Dual a) { return {atan(a.v), a.d / (1.0 + a.v * a.v)}; }
template <typename V>
__device__ inline V constant(double c) {
  return V(c);
}
template <>
__device__ inline Dual constant<Dual>(double c) {
  return {c, 0.0};
}

// Monotone interior node derivatives (weighted harmonic mean of the
// neighbouring slopes, zero where they change sign) and the cubic Bezier
// segment through them, with linear extrapolation outside the nodes.
template <typename V>
__device__ V bezier_nodes(const double *t, const V *y, V *d, int n, double x) {
  for (int i = 1; i < n - 1; ++i) {
    const double h1 = t[i] - t[i - 1], h2 = t[i + 1] - t[i];
    const V d1 = (y[i] - y[i - 1]) / h1, d2 = (y[i + 1] - y[i]) / h2;
    d[i] = constant<V>(0.0);
    if (val(d1 * d2) > 0) {
      const double al = 0.3333333 * (1.0 + h2 / (h1 + h2));
      d[i] = d1 * d2 / (al * d2 + (1.0 - al) * d1);
    }
  }
  if (x < t[0]) return y[0] + d[0] * (x - t[0]);
  if (x > t[n - 1]) return y[n - 1] + d[n - 1] * (x - t[n - 1]);
  int ii = 0;
  for (ii = 0; ii < n - 1; ++ii)
    if (x <= t[ii + 1]) break;
  const double hh = t[ii + 1] - t[ii];
  const V E = y[ii] + hh / 3.0 * d[ii], F = y[ii + 1] - hh / 3.0 * d[ii + 1];
  const double u = (x - t[ii]) / hh;
  return (1.0 - u) * (1.0 - u) * (1.0 - u) * y[ii] + u * u * u * y[ii + 1] + 3.0 * u * (1.0 - u) * (1.0 - u) * E +
         3.0 * u * u * (1.0 - u) * F;
}

// atmospheric_interpolation for a non-temperature parameter at log tau x:
// cubic Bezier between nodes, zero end derivatives, interior derivatives by
// the weighted harmonic mean of the neighbouring slopes, linear (here:
// constant) extrapolation outside the node range.
template <typename V>
__device__ V interp_param(const NodeArgs &na, const V *v, int code, double x) {
  double t[SNAPI_MAX_NODES];
  V y[SNAPI_MAX_NODES], d[SNAPI_MAX_NODES];
  int n = 0;
  for (int i = 0; i < na.nnode; ++i)
    if (na.code[i] == code) {
      t[n] = na.ltau[i];
      y[n] = v[i];
      ++n;
    }
  if (n == 0) return constant<V>(0.0);
  if (n == 1) return y[0];
  d[0] = constant<V>(0.0);
  d[n - 1] = constant<V>(0.0);
  return bezier_nodes(t, y, d, n, x);
}

// Temperature gradient of the FAL-C model against log tau near the bottom
// (atmospheric_interpolation, cmlib/mathtools.cc:1227-1232): the bottom
// derivative of the temperature interpolation.
__constant__ double kFalcDT[11] = {1694.91525424, 2164.5021645, 2342.50815695, 2289.18827635, 2155.92183045,
                                   2094.90240415, 2080.69183003, 2066.44925898, 1960.27182436, 1826.11342193,
                                   1729.74468968};
__constant__ double kFalcTau[11] = {-0.01232, 0.08798, 0.2002, 0.32704, 0.46667, 0.61462,
                                    0.7674,   0.92221, 1.07711, 1.22829, 1.37282};

// interpol_1d (cmlib/mathtools.cc:1057-1117): quadratic Bezier between the
// bracketing points, Fritsch-Butland derivatives (one-sided at the ends).
// Called only strictly inside the table.
__device__ double interpol_1d(const double *f, const double *x, int N, double xi) {
  int i;
  for (i = 0; i <= N - 2; ++i)
    if (xi < x[i + 1]) break;
  double h1 = 0.0, h2 = 0.0, fd1, fd2;
  if (i > 0) {
    h1 = x[i] - x[i - 1];
    h2 = x[i + 1] - x[i];
    const double d1 = (f[i] - f[i - 1]) / h1, d2 = (f[i + 1] - f[i]) / h2;
    const double al = 1.0 / 3.0 * (1.0 + h2 / (h1 + h2));
    fd1 = (d1 * d2 > 0) ? d1 * d2 / (al * d2 + (1.0 - al) * d1) : 0.0;
  } else {
    fd1 = (f[1] - f[0]) / (x[1] - x[0]);
  }
  const double C0 = f[i] + h2 * 0.5 * fd1;
  if (i + 1 < N - 1) {
    h1 = x[i + 1] - x[i];
    h2 = x[i + 2] - x[i + 1];
    const double d1 = (f[i + 1] - f[i]) / h1, d2 = (f[i + 2] - f[i + 1]) / h2;
    const double al = 1.0 / 3.0 * (1.0 + h2 / (h1 + h2));
    fd2 = (d1 * d2 > 0) ? d1 * d2 / (al * d2 + (1.0 - al) * d1) : 0.0;
  } else {
    fd2 = (f[N - 1] - f[N - 2]) / (x[N - 1] - x[N - 2]);
  }
  const double C1 = f[i + 1] - h1 * 0.5 * fd2;
  const double C = (C0 + C1) * 0.5;
  const double u = (xi - x[i]) / (x[i + 1] - x[i]);
  return (1.0 - u) * (1.0 - u) * f[i] + u * u * f[i + 1] + u * 2.0 * (1.0 - u) * C;
}

// atmospheric_interpolation of the temperature: flat top, bottom derivative
// from the FAL-C table at the deepest node (its end values beyond the
// table), the Bezier segments of interp_param, clamped to 3400..20000 K; one
// node gives a constant.
template <typename V>
__device__ V interp_temp(const NodeArgs &na, const V *v, double x) {
  double t[SNAPI_MAX_NODES];
  V y[SNAPI_MAX_NODES], d[SNAPI_MAX_NODES];
  int n = 0;
  for (int i = 0; i < na.nnode; ++i)
    if (na.code[i] == NP_T) {
      t[n] = na.ltau[i];
      y[n] = v[i];
      ++n;
    }
  if (n == 0) return constant<V>(0.0);
  if (n == 1) return y[0];
  d[0] = constant<V>(0.0);
  const double tb = t[n - 1];
  d[n - 1] = constant<V>(tb < kFalcTau[0]    ? kFalcDT[0]
                         : tb > kFalcTau[10] ? kFalcDT[10]
                                             : interpol_1d(kFalcDT, kFalcTau, 11, tb));
  V q = bezier_nodes(t, y, d, n, x);
  if (val(q) < 3400.0) q = constant<V>(3400.0);
  if (val(q) > 20000.0) q = constant<V>(20000.0);
  return q;
}

// Field values of one depth from the node values, as the atmosphere stores
// and projects them for the mu = 1 ray: B = |B| (sin th cos ph, sin th sin
// ph, cos th), inclination acos(Bz/|B|), azimuth atan(By/Bx) (atmos/
// atmos_rad.cc:25-51).  out: vt, vlos, |B|, inclination [deg], azimuth [deg].
template <typename V>
struct DepthFieldsT {
  V vt, vlos, bm, inc, azi;
};
using DepthFields = DepthFieldsT<double>;
template <typename V>
__device__ DepthFieldsT<V> fields_at(const NodeArgs &na, const V *v, double x) {
  DepthFieldsT<V> f;
  f.vt = interp_param(na, v, NP_VT, x);
  f.vlos = interp_param(na, v, NP_VS, x);
  // build_from_nodes (atmos_pp.cc:121-167): a field without nodes is 1e-3 G
  // along 1e-3 rad, and the inclination stays at least 1e-3 rad
  bool hb = false, ht = false, hp = false;
  for (int i = 0; i < na.nnode; ++i) {
    hb |= na.code[i] == NP_B;
    ht |= na.code[i] == NP_THETA;
    hp |= na.code[i] == NP_PHI;
  }
  const V B = hb ? interp_param(na, v, NP_B, x) : constant<V>(1E-3);
  V th = ht ? interp_param(na, v, NP_THETA, x) : constant<V>(1E-3);
  if (fabs(val(th)) < 1E-3) th = constant<V>(1E-3);
  const V ph = hp ? interp_param(na, v, NP_PHI, x) : constant<V>(1E-3);
  const V bx = B * sin(th) * cos(ph), by = B * sin(th) * sin(ph), bz = B * cos(th);
  f.bm = sqrt(bx * bx + by * by + bz * bz);
  f.inc = acos(bz / f.bm) * 180.0 / kPi;
  f.azi = atan(by / bx) * 180.0 / kPi;
  return f;
}

// One workgroup per slot, threads over depth.  Writes the node-controlled
// fields of state (and, with nmap, the node responses d q_p(z) / d node).
template <typename R>
__global__ void nodes_to_state_kernel(NodeArgs na, int nslot, int nd, int nf, const double *__restrict__ nodes,
                                      size_t node_stride, R *__restrict__ state, size_t state_stride,
                                      R *__restrict__ nmap, const int *__restrict__ active) {
  const int slot = blockIdx.x;
  if (active && !active[slot]) return;
  const double *v = nodes + (size_t)slot * node_stride;
  R *st = state + (size_t)slot * state_stride;
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    const double x = log10(-(double)st[SNAPI_F_GRID * nd + z]);
    const DepthFields f = fields_at(na, v, x);
    st[SNAPI_F_VT * nd + z] = (R)f.vt;
    st[SNAPI_F_VLOS * nd + z] = (R)f.vlos;
    st[SNAPI_F_BMAG * nd + z] = (R)f.bm;
    st[SNAPI_F_BINC * nd + z] = (R)f.inc;
    st[SNAPI_F_BAZI * nd + z] = (R)f.azi;
    if (nmap && na.analytic) {
      // d(vt + vz + |B| + inc + azi [rad]) / dnode, one dual pass per node
      Dual vd[SNAPI_MAX_NODES];
      for (int i = 0; i < na.nnode; ++i) vd[i] = Dual{v[i], 0.0};
      for (int n = 0; n < na.nnode; ++n) {
        vd[n].d = 1.0;
        const DepthFieldsT<Dual> a = fields_at(na, (const Dual *)vd, x);
        vd[n].d = 0.0;
        const double dq = a.vt.d + a.vlos.d + a.bm.d + (a.inc.d + a.azi.d) * kPi / 180.0;
        nmap[((size_t)slot * na.nnode + n) * nd + z] = (R)dq;
      }
    } else if (nmap) {
      double vv[SNAPI_MAX_NODES];
      for (int i = 0; i < na.nnode; ++i) vv[i] = v[i];
      for (int n = 0; n < na.nnode; ++n) {
        const double h = na.step[n];
        vv[n] = v[n] + 0.5 * h;
        const DepthFields a = fields_at(na, vv, x);
        vv[n] = v[n] - 0.5 * h;
        const DepthFields b = fields_at(na, vv, x);
        vv[n] = v[n];
        // the reference differences vt + vz + |B| + inc + azi (radians); only
        // the node's own quantity changes
        const double qa = a.vt + a.vlos + a.bm + (a.inc + a.azi) * kPi / 180.0;
        const double qb = b.vt + b.vlos + b.bm + (b.inc + b.azi) * kPi / 180.0;
        nmap[((size_t)slot * na.nnode + n) * nd + z] = (R)((qa - qb) / h);
      }
    }
  }
}

// r = (S_obs - S) w_s mask (calc_residual), chi^2 = sum r^2 (calc_chisq)
template <typename R>
__global__ void residual_kernel(int nslot, int nl, const R *__restrict__ S, const R *__restrict__ obs,
                                const R *__restrict__ mask, double w0, double w1, double w2, double w3,
                                R *__restrict__ r, double *__restrict__ chi2, const int *__restrict__ active) {
  const int slot = blockIdx.x;
  if (active && !active[slot]) return;
  __shared__ double red[256];
  const double ws[4] = {w0, w1, w2, w3};
  double acc = 0.0;
  for (int i = threadIdx.x; i < nl * 4; i += blockDim.x) {
    const int l = i >> 2, s = i & 3;
    const double v = ((double)obs[(size_t)slot * nl * 4 + i] - (double)S[(size_t)slot * nl * 4 + i]) * ws[s] *
                     (double)mask[(size_t)slot * nl + l];
    if (r) r[(size_t)slot * nl * 4 + i] = (R)v;
    acc += v * v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) chi2[slot] = red[0];
}

// Central-difference Jacobian from the (2n+1) syntheses of each slot.
template <typename R>
__global__ void fdjac_kernel(NodeArgs na, int nslot, int nl, const R *__restrict__ Sall, const double *__restrict__ hv,
                             const R *__restrict__ mask, double w0, double w1, double w2, double w3,
                             R *__restrict__ S, R *__restrict__ J, const int *__restrict__ active) {
  const int slot = blockIdx.x;
  if (active && !active[slot]) return;
  const int nv = 2 * na.nnode + 1;
  const double ws[4] = {w0, w1, w2, w3};
  for (int i = threadIdx.x; i < nl * 4; i += blockDim.x) {
    const int l = i >> 2, s = i & 3;
    const size_t base = (size_t)slot * nv;
    S[(size_t)slot * nl * 4 + i] = Sall[(base + 0) * nl * 4 + i];
    const double mk = (double)mask[(size_t)slot * nl + l];
    for (int n = 0; n < na.nnode; ++n) {
      const double d = ((double)Sall[(base + 1 + 2 * n) * nl * 4 + i] - (double)Sall[(base + 2 + 2 * n) * nl * 4 + i]) /
                       (2.0 * hv[(size_t)slot * na.nnode + n]);
      J[((size_t)slot * nl * 4 + i) * na.nnode + n] = (R)(d * na.scale[n] * ws[s] * mk);
    }
  }
}

// node sets of the central differences: v = 0 base, 2n+1 / 2n+2 = node n +/- h
__global__ void fd_nodes_kernel(NodeArgs na, int nslot, const double *__restrict__ nodes, double *__restrict__ fdn,
                                double *__restrict__ hv, const int *__restrict__ active) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslot || (active && !active[slot])) return;
  const int nn = na.nnode, nv = 2 * nn + 1;
  const double *v = nodes + (size_t)slot * nn;
  for (int k = 0; k < nv; ++k)
    for (int i = 0; i < nn; ++i) fdn[((size_t)slot * nv + k) * nn + i] = v[i];
  for (int n = 0; n < nn; ++n) {
    const double h = fmax(1e-3 * fabs(v[n]), na.step[n]);
    hv[(size_t)slot * nn + n] = h;
    fdn[((size_t)slot * nv + 1 + 2 * n) * nn + n] += h;
    fdn[((size_t)slot * nv + 2 + 2 * n) * nn + n] -= h;
  }
}

// bracket_parameter_values (atmos_model.cc:957-977)
__device__ double bracket(int code, double v) {
  switch (code) {
    case NP_T: return fmin(fmax(v, 3400.0), 12000.0);
    case NP_VT: return fmin(fmax(v, 1E2), 15E5);
    case NP_VS: return fmin(fmax(v, -15E5), 15E5);
    case NP_B: return fmin(fmax(v, 1.0), 1E4);
  }
  return v;
}

__global__ void trial_kernel(NodeArgs na, int nslot, const double *__restrict__ nodes, const double *__restrict__ corr,
                             double *__restrict__ test, const int *__restrict__ active) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslot || !active[slot]) return;
  for (int n = 0; n < na.nnode; ++n) {
    const size_t i = (size_t)slot * na.nnode + n;
    test[i] = bracket(na.code[n], nodes[i] + corr[i] * na.scale[n]);
  }
}

struct SlotState {
  int *pix, *iter, *corrected, *ntrack, *active, *needj;
  // 1 where the Jacobian must march the hydrostatic equilibrium itself: a
  // slot just refilled; 0 after an accepted trial, whose temperature and
  // equilibrium profiles (at the now current nodes) are still in place
  int *march;
  double *lm, *chi2, *chi2t, *track;
};

// accept / reject (atmos_fit.cc:163-181)
__global__ void update_kernel(NodeArgs na, int nslot, int max_iter, SlotState ss, double *__restrict__ nodes,
                              const double *__restrict__ test) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= nslot || !ss.active[slot]) return;
  const double lmm = sqrt(10.0);
  bool stop = false;
  ss.iter[slot] += 1;
  if (ss.chi2t[slot] < ss.chi2[slot]) {
    for (int n = 0; n < na.nnode; ++n) nodes[(size_t)slot * na.nnode + n] = test[(size_t)slot * na.nnode + n];
    ss.lm[slot] = fmax(ss.lm[slot] / lmm, 1E-5);
    ss.corrected[slot] = 1;
    double *tr = ss.track + (size_t)slot * 3;
    int nt = ss.ntrack[slot];
    if (nt < 3) {
      tr[nt] = ss.chi2[slot];
      ++nt;
    } else {
      tr[0] = tr[1];
      tr[1] = tr[2];
      tr[2] = ss.chi2[slot];
    }
    ss.ntrack[slot] = nt;
    if (nt >= 3) {
      const double change = fabs((tr[1] - tr[2]) / tr[2]);
      if (change < 1E-3) stop = true;
    }
  } else {
    ss.lm[slot] *= lmm;
    ss.corrected[slot] = 0;
  }
  if (stop || ss.iter[slot] >= max_iter) ss.active[slot] = 0;
  ss.needj[slot] = ss.active[slot] && ss.corrected[slot];
  ss.march[slot] = 0;
}

// Finished slots write their pixel's result and take the next pixel of the
// queue (copying its state, observation, mask and start nodes).
template <typename R>
__global__ void refill_kernel(NodeArgs na, int nslot, int npix, int nf, int nd, int nl, int *__restrict__ queue,
                              SlotState ss, const R *__restrict__ pstate, const R *__restrict__ pobs,
                              const R *__restrict__ pmask, const double *__restrict__ pnodes0, double lm0,
                              R *__restrict__ sstate, R *__restrict__ tstate, R *__restrict__ sobs,
                              R *__restrict__ smask, double *__restrict__ snodes, double *__restrict__ out_nodes,
                              double *__restrict__ out_chi2, int *__restrict__ out_iter) {
  const int slot = blockIdx.x;
  __shared__ int next;
  if (ss.active[slot]) return;
  if (threadIdx.x == 0) {
    const int old = ss.pix[slot];
    if (old >= 0) {
      // the result with its angles polished (model::polish_angles,
      // atmos_model.cc:992-1001): inclination acos(cos), azimuth atan(tan)
      for (int n = 0; n < na.nnode; ++n) {
        double v = snodes[(size_t)slot * na.nnode + n];
        if (na.code[n] == NP_THETA) v = acos(cos(v));
        else if (na.code[n] == NP_PHI) v = atan(tan(v));
        out_nodes[(size_t)old * na.nnode + n] = v;
      }
      out_chi2[old] = ss.chi2[slot];
      out_iter[old] = ss.iter[slot];
    }
    next = atomicAdd(queue, 1);
    if (next >= npix) next = -1;
    ss.pix[slot] = next;
    if (next >= 0) {
      ss.iter[slot] = 0;
      ss.corrected[slot] = 1;
      ss.ntrack[slot] = 0;
      ss.lm[slot] = lm0;
      ss.active[slot] = 1;
      ss.needj[slot] = 1;
      ss.march[slot] = 1;
      // the start put in the parameter limits, as stokes_lm_fit does before
      // its first iteration (atmos_fit.cc:91)
      for (int n = 0; n < na.nnode; ++n)
        snodes[(size_t)slot * na.nnode + n] = bracket(na.code[n], pnodes0[(size_t)next * na.nnode + n]);
    } else {
      ss.needj[slot] = 0;
      ss.march[slot] = 0;
    }
  }
  __syncthreads();
  if (next < 0) return;
  const size_t ns = (size_t)nf * nd;
  for (size_t i = threadIdx.x; i < ns; i += blockDim.x) {
    const R v = pstate[(size_t)next * ns + i];
    sstate[(size_t)slot * ns + i] = v;
    tstate[(size_t)slot * ns + i] = v;
  }
  for (int i = threadIdx.x; i < nl * 4; i += blockDim.x) sobs[(size_t)slot * nl * 4 + i] = pobs[(size_t)next * nl * 4 + i];
  for (int i = threadIdx.x; i < nl; i += blockDim.x)
    smask[(size_t)slot * nl + i] = pmask ? pmask[(size_t)next * nl + i] : R(1);
}

template <typename R> __global__ void fill_kernel(size_t n, R v, R *__restrict__ p) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// per-(slot, node set) activity from the slots' flags
__global__ void expand_active_kernel(long n, int group, const int *__restrict__ active, int *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = active[i / group];
}

// One workgroup of 1024: count[0] = the active slots, total[0] += count[0].
__global__ __launch_bounds__(1024) void active_count_kernel(int nslot, const int *__restrict__ active,
                                                            int *__restrict__ count, long long *__restrict__ total) {
  __shared__ int part[1024];
  int n = 0;
  for (int s = threadIdx.x; s < nslot; s += blockDim.x) n += active[s] != 0;
  part[threadIdx.x] = n;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    count[0] = part[0];
    total[0] += part[0];
  }
}

// Runs iteration() until an iteration finds no active slot.  Iterations go
// in batches of k between host reads of the active count (k doubles while a
// batch takes under 2 ms, up to 32; the extra iterations of the last batch
// find nothing to do).  On a stream other than the null stream the first
// iteration runs eagerly and the second is captured into a HIP graph that
// every later iteration replays; SNAPI_LAUNCH=eager in the environment
// keeps every iteration eager.  count: the device word the iteration
// writes; pinned: one int of pinned host memory.
template <typename F>
static int run_iterations(F &iteration, hipStream_t st, const int *count, int *pinned) {
  const char *mode = getenv("SNAPI_LAUNCH");
  bool want_graph = st != nullptr && !(mode && !strcmp(mode, "eager"));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int rc = SNAPI_OK, done_iters = 0, k = 1;
  for (;;) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int b = 0; b < k && !rc; ++b, ++done_iters) {
      if (exec) {
        if (hipGraphLaunch(exec, st) != hipSuccess) rc = SNAPI_ERR_DEVICE;
        continue;
      }
      if (want_graph && done_iters >= 1) {
        // capture one iteration; on any failure stay eager
        want_graph = false;
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
          const int r = iteration();
          const hipError_t e = hipStreamEndCapture(st, &graph);
          if (r == SNAPI_OK && e == hipSuccess && graph &&
              hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
            if (hipGraphLaunch(exec, st) != hipSuccess) rc = SNAPI_ERR_DEVICE;
            continue;
          }
          (void)hipGetLastError();
          if (exec) (void)hipGraphExecDestroy(exec);
          if (graph) (void)hipGraphDestroy(graph);
          exec = nullptr;
          graph = nullptr;
        }
      }
      rc = iteration();
    }
    if (rc) break;
    if (hipMemcpyAsync(pinned, count, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = SNAPI_ERR_DEVICE;
      break;
    }
    if (*pinned == 0) break;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms < 2.0 && k < 32) k *= 2;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  return rc;
}

// ---- temperature nodes: the thermodynamic state rebuilt from the nodes ----

constexpr double kNodeStepT = 5.0;  // delta_T, the temperature-node step (cmlib/const.h:27)

// Temperature profiles of one pixel for the node responses
// (obs_stokes_responses_to_nodes, atmos_fit.cc:1225-1248): variant 0 the
// nodes as given, then per temperature node j the node raised by delta_T/2
// (variant 1+2j) and lowered from there by delta_T (2+2j).  Writes
// Tv[pix][nvar][nd] and the log tau of each depth (from the state's grid)
// for the hydrostatic march, ltv[pix][nvar][nd].
template <typename R>
__global__ void tprofiles_kernel(NodeArgs na, int npix, int nd, size_t state_stride, const R *__restrict__ state,
                                 const double *__restrict__ nodes, int nvar, double *__restrict__ Tv,
                                 double *__restrict__ ltv, const int *__restrict__ active) {
  const int pix = blockIdx.x;
  if (active && !active[pix]) return;
  const double *v0 = nodes + (size_t)pix * na.nnode;
  const R *st = state + (size_t)pix * state_stride;
  double v[SNAPI_MAX_NODES];
  for (int var = 0; var < nvar; ++var) {
    for (int i = 0; i < na.nnode; ++i) v[i] = v0[i];
    if (var > 0) {
      int j = (var - 1) / 2, n = 0;
      for (int i = 0; i < na.nnode; ++i)
        if (na.code[i] == NP_T && j-- == 0) n = i;
      v[n] += kNodeStepT * 0.5;
      if ((var - 1) % 2) v[n] -= kNodeStepT;
    }
    for (int z = threadIdx.x; z < nd; z += blockDim.x) {
      const double x = log10(-(double)st[SNAPI_F_GRID * nd + z]);
      Tv[((size_t)pix * nvar + var) * nd + z] = interp_temp(na, v, x);
      ltv[((size_t)pix * nvar + var) * nd + z] = x;
    }
  }
}

// Node responses of the temperature nodes, dT(z)/dnode and dNt(z)/dnode by
// the difference of the raised and lowered profiles over delta_T, into the
// node maps' rows of those nodes; the double state gets the given nodes' T
// and Nt and the other fields of state.
template <typename R>
__global__ void tmaps_kernel(NodeArgs na, int npix, int nd, int nf, int nvar, const double *__restrict__ Tv,
                             const double *__restrict__ Ntv, const R *__restrict__ state, R *__restrict__ nmap,
                             R *__restrict__ nmap2, double *__restrict__ state_d, const int *__restrict__ active) {
  const int pix = blockIdx.x;
  if (active && !active[pix]) return;
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    const double *T0 = Tv + (size_t)pix * nvar * nd, *N0 = Ntv + (size_t)pix * nvar * nd;
    for (int f = 0; f < nf; ++f)
      state_d[((size_t)pix * nf + f) * nd + z] = (double)state[((size_t)pix * nf + f) * nd + z];
    state_d[((size_t)pix * nf + SNAPI_F_T) * nd + z] = T0[z];
    state_d[((size_t)pix * nf + SNAPI_F_NT) * nd + z] = N0[z];
    if (!nmap) continue;  // the state alone
    int j = 0;
    for (int n = 0; n < na.nnode; ++n) {
      if (na.code[n] != NP_T) {
        nmap2[((size_t)pix * na.nnode + n) * nd + z] = R(0);
        continue;
      }
      const size_t a = (size_t)(1 + 2 * j) * nd + z, b = (size_t)(2 + 2 * j) * nd + z;
      double dT = T0[a];
      dT -= T0[b];
      dT /= kNodeStepT;
      double dN = N0[a];
      dN -= N0[b];
      dN /= kNodeStepT;
      nmap[((size_t)pix * na.nnode + n) * nd + z] = (R)dT;
      nmap2[((size_t)pix * na.nnode + n) * nd + z] = (R)dN;
      ++j;
    }
  }
}

// Analytic node responses of the temperature nodes: dT(z)/dnode, the
// interpolation's exact derivative at the given nodes, into dT
// [pix][ntn][nd] (for the hydrostatic tangent) and the node maps' rows.
// ltv [pix][nd]: log tau of each depth.
template <typename R>
__global__ void tderiv_kernel(NodeArgs na, int npix, int nd, int ntn, const double *__restrict__ nodes,
                              const double *__restrict__ ltv, double *__restrict__ dT, R *__restrict__ nmap,
                              const int *__restrict__ active) {
  const int pix = blockIdx.x;
  if (active && !active[pix]) return;
  Dual vd[SNAPI_MAX_NODES];
  for (int i = 0; i < na.nnode; ++i) vd[i] = Dual{nodes[(size_t)pix * na.nnode + i], 0.0};
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    const double x = ltv[(size_t)pix * nd + z];
    int j = 0;
    for (int n = 0; n < na.nnode; ++n) {
      if (na.code[n] != NP_T) continue;
      vd[n].d = 1.0;
      const Dual T = interp_temp(na, (const Dual *)vd, x);
      vd[n].d = 0.0;
      dT[((size_t)pix * ntn + j) * nd + z] = T.d;
      nmap[((size_t)pix * na.nnode + n) * nd + z] = (R)T.d;
      ++j;
    }
  }
}

// dNt/dnode rows of the node maps from the hydrostatic tangent dN
// [pix][ntn][nd]; zero for the other nodes.
template <typename R>
__global__ void tmap2_kernel(NodeArgs na, int npix, int nd, int ntn, const double *__restrict__ dN,
                             R *__restrict__ nmap2, const int *__restrict__ active) {
  const int pix = blockIdx.x;
  if (active && !active[pix]) return;
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    int j = 0;
    for (int n = 0; n < na.nnode; ++n)
      nmap2[((size_t)pix * na.nnode + n) * nd + z] =
          na.code[n] == NP_T ? (R)dN[((size_t)pix * ntn + j++) * nd + z] : R(0);
  }
}

// The temperature and density columns of the given nodes into a double state.
__global__ void set_tn_kernel(int nd, int nf, const double *__restrict__ Tv, const double *__restrict__ Ntv,
                              double *__restrict__ state) {
  const int pix = blockIdx.x;
  for (int z = threadIdx.x; z < nd; z += blockDim.x) {
    state[((size_t)pix * nf + SNAPI_F_T) * nd + z] = Tv[(size_t)pix * nd + z];
    state[((size_t)pix * nf + SNAPI_F_NT) * nd + z] = Ntv[(size_t)pix * nd + z];
  }
}

template <typename A, typename B>
__global__ void convert_kernel(size_t n, const A *__restrict__ a, B *__restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    b[i] = (B)a[i];
}

#define HIPCHK(x)                          \
  do {                                     \
    if ((x) != hipSuccess) {               \
      rc = SNAPI_ERR_DEVICE;               \
      goto done;                           \
    }                                      \
  } while (0)

// Species tables and state map of an inversion with temperature nodes.
struct ThermoArgs {
  const snapi_chem_table *chem, *chem_he;
  const snapi_state_map *map;
  double broadening;   // Gaussian spectral broadening width [cm], 0: none
  int npsf;            // spectral PSF samples, 0: none
  const double *psf;   // [npsf], host memory
};

// mask[slot][l] applied to J[slot][l*4+s][n] after the convolution.
template <typename R>
__global__ void mask_jacobian_kernel(int nl, int nc, const R *__restrict__ mask, R *__restrict__ J,
                                     const int *__restrict__ active) {
  const int slot = blockIdx.x;
  if (!active[slot]) return;
  for (int i = threadIdx.x; i < nl * nc; i += blockDim.x)
    J[(size_t)slot * nl * nc + i] *= mask[(size_t)slot * nl + i / nc];
}

template <typename R>
static int invert_impl(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                       const NodeArgs &na, int npix, int nd, const void *pstate_h, int nlambda, const double *lam_h,
                       const void *obs_h, const void *mask_h, const double *ws, double lm0, int max_iter, int nslot,
                       const double *nodes0, double *nodes_out, double *chi2_out, int *iter_out, long long *stats,
                       bool dev, hipStream_t st, const ThermoArgs *th = nullptr, const double *reg_h = nullptr) {
  KLines kl;
  SynthArgs a;
  int rc = prepare_synth(lines, ray, 1, nd, nlambda, kl, a);
  if (rc) return rc;
  const int nf = a.nfield, nn = na.nnode, nl = nlambda, nv = 2 * nn + 1;
  const size_t es = sizeof(R), ns = (size_t)nf * nd;
  // device buffers
  R *pstate = nullptr, *pobs = nullptr, *pmask = nullptr, *sstate = nullptr, *tstate = nullptr, *dzero = nullptr;
  R *sobs = nullptr, *smask = nullptr, *S = nullptr, *St = nullptr, *J = nullptr, *r = nullptr, *nmap = nullptr;
  R *fdstate = nullptr, *Sall = nullptr, *scr = nullptr;
  double *dlam = nullptr, *pnodes0 = nullptr, *snodes = nullptr, *tnodes = nullptr, *JTJ = nullptr, *rhs = nullptr;
  double *corr = nullptr, *onodes = nullptr, *ochi2 = nullptr, *fdn = nullptr, *hv = nullptr, *dbl = nullptr;
  int *ints = nullptr, *queue = nullptr, *oiter = nullptr, *count = nullptr, *pinned = nullptr;
  long long *total = nullptr;
  double *dreg = nullptr, *dpsf = nullptr;  // regularisation [nn][nn], spectral PSF
  // temperature nodes: profiles and densities of the rebuilds, double states
  // for the equation of state, the response path's state and derivatives
  ChemDevice cdev, cdev_he;
  double *Tv = nullptr, *ltv = nullptr, *Ntv = nullptr, *sd = nullptr, *rsd = nullptr, *dsd = nullptr,
         *tsd = nullptr, *escr = nullptr;
  R *rstate = nullptr, *rdst = nullptr, *nmap2 = nullptr;
  std::vector<double> lam_host;  // the convolution plans its grid on the host
  int ntn = 0;
  for (int n = 0; n < nn; ++n) ntn += na.code[n] == NP_T;
  // hydrostatic builds per pixel for the Jacobian: the nodes and each
  // temperature node +- delta_T/2, or the nodes alone with the tangent
  const int nvar = na.analytic ? 1 : 1 + 2 * ntn;
  double *dTd = nullptr, *dNd = nullptr, *hscr = nullptr;
  // difference Jacobian with temperature nodes: per (slot, node set)
  int *fd_active = nullptr;
  double *fdTv = nullptr, *fdltv = nullptr, *fdNtv = nullptr, *fdsd = nullptr, *fdescr = nullptr;
  // caller's buffers (device variant) or copies of the host arrays
  void *own[8] = {};
  SlotState ss;
  size_t per_col = 0;
  int chunk = 0;
  int *status = device_status_slot();
  if (!status) return SNAPI_ERR_DEVICE;
  if (th) {
    // the difference Jacobian differences the synthesis path's spectrum of
    // every node set; with instrument effects only the response-path
    // Jacobians are built
    const bool jac_ok = jacobian == SNAPI_JAC_ADJOINT || jacobian == SNAPI_JAC_TANGENT ||
                        (jacobian == SNAPI_JAC_DIFFERENCES && !(th->broadening > 0.0) && th->npsf == 0);
    if (!jac_ok || nf != SNAPI_NF_BASE + 2 * th->map->nline) {
      rc = !jac_ok ? SNAPI_ERR_UNSUPPORTED : SNAPI_ERR_ARG;
      goto done;
    }
    rc = chem_upload(th->chem, th->map, cdev, st);
    if (rc) goto done;
    rc = chem_upload(th->chem_he, nullptr, cdev_he, st);
    if (rc) goto done;
  }
  if (nslot <= 0) {
    // as many slots as the pixels, or as fit in half of the free device
    // memory: the bytes one slot holds in the buffers allocated below
    size_t b = es * ((size_t)ns * 4 + (size_t)nl * 17 + (size_t)nl * 4 * nn + (size_t)nn * nd) +
               8 * ((size_t)nn * 4 + (size_t)nn * nn + 6) + 4 * 7;
    if (th)
      b += 8 * ((size_t)3 * nvar * nd + (size_t)ns * 5) + chem_eos_scratch(cdev, nd, true) +
           es * ((size_t)ns * 3 + (size_t)nn * nd) +
           (na.analytic && ntn ? 8 * (size_t)2 * ntn * nd + chem_he_tangent_scratch(1, nd) : 0);
    if (jacobian == SNAPI_JAC_ADJOINT) b += resp_scratch_per_col(nl, nd, es);
    if (jacobian == SNAPI_JAC_DIFFERENCES) {
      b += es * ((size_t)nv * ns + (size_t)nv * nl * 4) + 8 * ((size_t)nv * nn + nn);
      if (th) b += (size_t)nv * (4 + 8 * (size_t)nd * 3 + 8 * (size_t)ns) + chem_eos_scratch(cdev, (long)nv * nd, false);
    }
    size_t fr = 0, tot = 0;
    nslot = npix;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr > 0) {
      const size_t fit = (fr / 2) / b;
      if (fit < (size_t)nslot) nslot = fit > 0 ? (int)fit : 1;
    }
  }
  chunk = nslot;
  if (dev) {
    pstate = (R *)pstate_h;
    pobs = (R *)obs_h;
    pmask = (R *)mask_h;
    dlam = (double *)lam_h;
    pnodes0 = (double *)nodes0;
    onodes = nodes_out;
    ochi2 = chi2_out;
    oiter = iter_out;
  } else {
    HIPCHK(hipMalloc((void **)&pstate, (size_t)npix * ns * es));
    HIPCHK(hipMalloc((void **)&pobs, (size_t)npix * nl * 4 * es));
    HIPCHK(hipMalloc((void **)&pmask, (size_t)npix * nl * es));
    HIPCHK(hipMalloc((void **)&dlam, (size_t)nl * 8));
    HIPCHK(hipMalloc((void **)&pnodes0, (size_t)npix * nn * 8));
    HIPCHK(hipMalloc((void **)&onodes, (size_t)npix * nn * 8));
    HIPCHK(hipMalloc((void **)&ochi2, (size_t)npix * 8));
    HIPCHK(hipMalloc((void **)&oiter, (size_t)npix * sizeof(int)));
    own[0] = pstate; own[1] = pobs; own[2] = pmask; own[3] = dlam;
    own[4] = pnodes0; own[5] = onodes; own[6] = ochi2; own[7] = oiter;
  }
  if (!pstate || !pobs || !dlam || !pnodes0 || !onodes || !ochi2 || !oiter) {
    rc = SNAPI_ERR_ARG;
    goto done;
  }
  HIPCHK(hipMalloc((void **)&sstate, (size_t)nslot * ns * es));
  HIPCHK(hipMalloc((void **)&tstate, (size_t)nslot * ns * es));
  HIPCHK(hipMalloc((void **)&dzero, (size_t)nslot * 2 * ns * es));
  HIPCHK(hipMalloc((void **)&sobs, (size_t)nslot * nl * 4 * es));
  HIPCHK(hipMalloc((void **)&smask, (size_t)nslot * nl * es));
  HIPCHK(hipMalloc((void **)&S, (size_t)nslot * nl * 4 * es));
  HIPCHK(hipMalloc((void **)&St, (size_t)nslot * nl * 4 * es));
  HIPCHK(hipMalloc((void **)&J, (size_t)nslot * nl * 4 * nn * es));
  HIPCHK(hipMalloc((void **)&r, (size_t)nslot * nl * 4 * es));
  HIPCHK(hipMalloc((void **)&nmap, (size_t)nslot * nn * nd * es));
  HIPCHK(hipMalloc((void **)&snodes, (size_t)nslot * nn * 8));
  HIPCHK(hipMalloc((void **)&tnodes, (size_t)nslot * nn * 8));
  HIPCHK(hipMalloc((void **)&JTJ, (size_t)nslot * nn * nn * 8));
  if (reg_h) {
    HIPCHK(hipMalloc((void **)&dreg, (size_t)nn * nn * 8));
    HIPCHK(hipMemcpy(dreg, reg_h, (size_t)nn * nn * 8, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMalloc((void **)&rhs, (size_t)nslot * nn * 8));
  HIPCHK(hipMalloc((void **)&corr, (size_t)nslot * nn * 8));
  HIPCHK(hipMalloc((void **)&dbl, (size_t)nslot * 6 * 8));
  HIPCHK(hipMalloc((void **)&ints, (size_t)nslot * 7 * sizeof(int)));
  HIPCHK(hipMalloc((void **)&queue, 4 * sizeof(int)));  // next pixel, active count, total (long long)
  count = queue + 1;
  total = reinterpret_cast<long long *>(queue + 2);
  HIPCHK(hipHostMalloc((void **)&pinned, sizeof(int), hipHostMallocDefault));
  ss.pix = ints;
  ss.iter = ints + nslot;
  ss.corrected = ints + 2 * nslot;
  ss.ntrack = ints + 3 * nslot;
  ss.active = ints + 4 * nslot;
  ss.needj = ints + 5 * nslot;
  ss.march = ints + 6 * nslot;
  ss.lm = dbl;
  ss.chi2 = dbl + nslot;
  ss.chi2t = dbl + 2 * nslot;
  ss.track = dbl + 3 * nslot;
  if (th) {
    HIPCHK(hipMalloc((void **)&Tv, (size_t)nslot * nvar * nd * 8));
    HIPCHK(hipMalloc((void **)&ltv, (size_t)nslot * nvar * nd * 8));
    HIPCHK(hipMalloc((void **)&Ntv, (size_t)nslot * nvar * nd * 8));
    HIPCHK(hipMalloc((void **)&sd, (size_t)nslot * ns * 8));
    HIPCHK(hipMalloc((void **)&rsd, (size_t)nslot * ns * 8));
    HIPCHK(hipMalloc((void **)&dsd, (size_t)nslot * 2 * ns * 8));
    HIPCHK(hipMalloc((void **)&tsd, (size_t)nslot * ns * 8));
    HIPCHK(hipMalloc((void **)&escr, chem_eos_scratch(cdev, (long)nslot * nd, true)));
    HIPCHK(hipMalloc((void **)&rstate, (size_t)nslot * ns * es));
    HIPCHK(hipMalloc((void **)&rdst, (size_t)nslot * 2 * ns * es));
    HIPCHK(hipMalloc((void **)&nmap2, (size_t)nslot * nn * nd * es));
    if (na.analytic && ntn) {
      HIPCHK(hipMalloc((void **)&dTd, (size_t)nslot * ntn * nd * 8));
      HIPCHK(hipMalloc((void **)&dNd, (size_t)nslot * ntn * nd * 8));
      HIPCHK(hipMalloc((void **)&hscr, chem_he_tangent_scratch(nslot, nd)));
    }
    if (th->npsf > 0) {
      HIPCHK(hipMalloc((void **)&dpsf, (size_t)th->npsf * 8));
      HIPCHK(hipMemcpy(dpsf, th->psf, (size_t)th->npsf * 8, hipMemcpyHostToDevice));
    }
    if (th->broadening > 0.0) {
      lam_host.resize(nl);
      HIPCHK(hipMemcpy(lam_host.data(), dev ? lam_h : (const double *)lam_h, (size_t)nl * 8,
                       dev ? hipMemcpyDeviceToHost : hipMemcpyHostToHost));
    }
  }
  if (jacobian == SNAPI_JAC_ADJOINT) {
    per_col = resp_scratch_per_col(nl, nd, es);
    size_t c = ((size_t)16 << 30) / (per_col ? per_col : 1);
    if (c < 1) c = 1;
    if (c > (size_t)nslot) c = nslot;
    chunk = (int)c;
    HIPCHK(hipMalloc((void **)&scr, per_col * chunk));
  } else if (jacobian == SNAPI_JAC_DIFFERENCES) {
    HIPCHK(hipMalloc((void **)&fdstate, (size_t)nslot * nv * ns * es));
    HIPCHK(hipMalloc((void **)&Sall, (size_t)nslot * nv * nl * 4 * es));
    HIPCHK(hipMalloc((void **)&fdn, (size_t)nslot * nv * nn * 8));
    HIPCHK(hipMalloc((void **)&hv, (size_t)nslot * nn * 8));
    if (th) {
      // every node set's own temperature, hydrostatic densities and
      // equation of state
      const size_t nsv = (size_t)nslot * nv;
      HIPCHK(hipMalloc((void **)&fd_active, nsv * sizeof(int)));
      HIPCHK(hipMalloc((void **)&fdTv, nsv * nd * 8));
      HIPCHK(hipMalloc((void **)&fdltv, nsv * nd * 8));
      HIPCHK(hipMalloc((void **)&fdNtv, nsv * nd * 8));
      HIPCHK(hipMalloc((void **)&fdsd, nsv * ns * 8));
      HIPCHK(hipMalloc((void **)&fdescr, chem_eos_scratch(cdev, (long)nsv * nd, false)));
    }
  }
  if (!dev) {
    HIPCHK(hipMemcpy(pstate, pstate_h, (size_t)npix * ns * es, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pobs, obs_h, (size_t)npix * nl * 4 * es, hipMemcpyHostToDevice));
    if (mask_h) HIPCHK(hipMemcpy(pmask, mask_h, (size_t)npix * nl * es, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dlam, lam_h, (size_t)nl * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pnodes0, nodes0, (size_t)npix * nn * 8, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemsetAsync(dzero, 0, (size_t)nslot * 2 * ns * es, st));
  HIPCHK(hipMemsetAsync(ints, 0, (size_t)nslot * 7 * sizeof(int), st));
  HIPCHK(hipMemsetAsync(dbl, 0, (size_t)nslot * 6 * 8, st));
  HIPCHK(hipMemsetAsync(queue, 0, 4 * sizeof(int), st));
  {
    std::vector<int> neg(nslot, -1);
    HIPCHK(hipMemcpyAsync(ss.pix, neg.data(), nslot * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  reset_status();
  {
    // With temperature nodes the fit is the reference's: its Jacobian and
    // current chi^2 come from obs_stokes_responses, whose H- opacity is taken
    // at the window's middle wavelength (mol.cc:414-469), its trials from
    // obs_stokes with H- at every wavelength (mol.cc:379-412).
    SynthArgs a_resp = a;
    snapi_ray ray_syn = *ray;
    if (th) {
      a_resp.hminus_mid = 1;
      ray_syn.hminus_mid = 0;
    }
    RespArgs ra = {};
    ra.nseed = 4;
    ra.sink = SINK_NODE;
    for (int s = 0; s < 4; ++s) ra.ws[s] = ws[s];
    ra.nnode = nn;
    for (int n = 0; n < nn; ++n) {
      // node parameter -> per-depth response index (T 0 and Nt 1, vt 2, vLOS 3, |B| 4, incl 5, azi 6)
      const int c = na.code[n];
      ra.pidx[n] = c == NP_T ? 0 : c == NP_VT ? 2 : c == NP_VS ? 3 : c == NP_B ? 4 : c == NP_THETA ? 5 : 6;
      ra.pidx2[n] = c == NP_T ? 1 : -1;
      ra.nscale[n] = na.scale[n];
    }
    ra.nmap = nmap;
    ra.nmap2 = th ? nmap2 : nullptr;
    // with instrument effects the Jacobian is convolved before the mask is
    // applied (convolve_response_with_gauss / _psf, then scale_rf)
    ra.mask = (th && (th->broadening > 0.0 || th->npsf > 0)) ? nullptr : smask;
    ra.active = ss.needj;
    // One LM iteration over all slots, enqueued on st: refill finished
    // slots, count the active ones (and add them to the device total), then
    // Jacobian, normal equations, trial and accept/reject.  Every kernel
    // skips inactive slots, so an iteration after the last pixel finished
    // changes nothing.
    auto iteration = [&]() -> int {
      hipLaunchKernelGGL(refill_kernel<R>, dim3(nslot), dim3(256), 0, st, na, nslot, npix, nf, nd, nl, queue, ss, pstate,
                         pobs, mask_h ? pmask : (R *)nullptr, pnodes0, lm0, sstate, tstate, sobs, smask, snodes, onodes,
                         ochi2, oiter);
      hipLaunchKernelGGL(active_count_kernel, dim3(1), dim3(1024), 0, st, nslot, ss.active, count, total);
        // model -> state and Jacobian where the model changed
        hipLaunchKernelGGL(nodes_to_state_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, nf, snodes,
                           (size_t)nn, sstate, ns, nmap, ss.needj);
        if (th && jacobian == SNAPI_JAC_DIFFERENCES) {
          // central differences of the synthesis path over the 2n + 1 node
          // sets, every set with its own temperature, hydrostatic equilibrium
          // and equation of state (forward_evaluate per set)
          const long nsv = (long)nslot * nv;
          hipLaunchKernelGGL(fd_nodes_kernel, dim3((nslot + 127) / 128), dim3(128), 0, st, na, nslot, snodes, fdn, hv,
                             ss.needj);
          for (int k = 0; k < nv; ++k) {
            if (hipMemcpy2DAsync(fdstate + (size_t)k * ns, nv * ns * es, sstate, ns * es, ns * es, nslot,
                                    hipMemcpyDeviceToDevice, st) != hipSuccess) return SNAPI_ERR_DEVICE;
            hipLaunchKernelGGL(nodes_to_state_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, nf,
                               fdn + (size_t)k * nn, (size_t)nv * nn, fdstate + (size_t)k * ns, (size_t)nv * ns,
                               (R *)nullptr, ss.needj);
          }
          hipLaunchKernelGGL(expand_active_kernel, dim3((unsigned)((nsv + 255) / 256)), dim3(256), 0, st, nsv, nv,
                             ss.needj, fd_active);
          hipLaunchKernelGGL(tprofiles_kernel<R>, dim3((unsigned)nsv), dim3(128), 0, st, na, (int)nsv, nd, ns, fdstate,
                             fdn, 1, fdTv, fdltv, fd_active);
          rc = chem_hequilibrium(cdev_he, nsv, nd, fdTv, fdltv, fdNtv, fd_active, 1, st);
          if (rc) return rc;
          hipLaunchKernelGGL(tmaps_kernel<R>, dim3((unsigned)nsv), dim3(128), 0, st, na, (int)nsv, nd, nf, 1, fdTv,
                             fdNtv, fdstate, (R *)nullptr, (R *)nullptr, fdsd, fd_active);
          rc = chem_eos(cdev, nsv, nd, fdsd, nullptr, nullptr, fdescr, fd_active, st);
          if (rc) return rc;
          hipLaunchKernelGGL((convert_kernel<double, R>), dim3(1024), dim3(256), 0, st, (size_t)nsv * ns, fdsd,
                             fdstate);
          rc = synth_dev_masked(precision, lines, &ray_syn, (int)nsv, nd, fdstate, nl, dlam, Sall, ss.needj, nv, st);
          if (rc) return rc;
          hipLaunchKernelGGL(fdjac_kernel<R>, dim3(nslot), dim3(256), 0, st, na, nslot, nl, Sall, hv, smask, ws[0], ws[1],
                             ws[2], ws[3], S, J, ss.needj);
        } else if (th) {
          // build_from_nodes at the nodes and at each temperature node +- 2.5 K,
          // then the response path's state and derivatives (the reference's
          // Jacobian and the current chi^2 come from that path)
          // with the analytic maps (one profile per slot) an accepted trial
          // leaves the temperature and equilibrium of the current nodes in
          // Tv, ltv and Ntv: only refilled slots march
          const int *hmask = na.analytic ? ss.march : ss.needj;
          hipLaunchKernelGGL(tprofiles_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, ns, sstate, snodes, nvar,
                             Tv, ltv, hmask);
          rc = chem_hequilibrium(cdev_he, (long)nslot * nvar, nd, Tv, ltv, Ntv, hmask, nvar, st);
          if (rc) return rc;
          if (na.analytic) {
            hipLaunchKernelGGL(tmaps_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, nf, 1, Tv, Ntv, sstate,
                               (R *)nullptr, (R *)nullptr, sd, ss.needj);
            if (ntn) {
              // the hydrostatic tangent's partials come from the first five
              // chemical equilibria of the equation-of-state pass at the same
              // points (T, T +- 2.5 K, Nt +- its half step)
              hipLaunchKernelGGL(tderiv_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, ntn, snodes, ltv, dTd,
                                 nmap, ss.needj);
              rc = chem_eos_he_tangent(cdev, cdev_he, nslot, nd, sd, rsd, dsd, escr, Tv, ltv, Ntv, ntn, dTd, dNd, hscr,
                                       ss.needj, st);
            } else {
              rc = chem_eos(cdev, nslot, nd, sd, rsd, dsd, escr, ss.needj, st);
            }
            if (rc) return rc;
            hipLaunchKernelGGL(tmap2_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, ntn, dNd, nmap2, ss.needj);
          } else {
            hipLaunchKernelGGL(tmaps_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, nf, nvar, Tv, Ntv, sstate,
                               nmap, nmap2, sd, ss.needj);
            rc = chem_eos(cdev, nslot, nd, sd, rsd, dsd, escr, ss.needj, st);
            if (rc) return rc;
          }
          hipLaunchKernelGGL((convert_kernel<double, R>), dim3(1024), dim3(256), 0, st, (size_t)nslot * ns, rsd, rstate);
          hipLaunchKernelGGL((convert_kernel<double, R>), dim3(1024), dim3(256), 0, st, (size_t)nslot * 2 * ns, dsd,
                             rdst);
          if (jacobian == SNAPI_JAC_TANGENT)
            rc = launch_tangent<R>(kl, a_resp, ra, nslot, rstate, rdst, dlam, S, J, st);
          else
            rc = launch_resp<R>(kl, a_resp, ra, nslot, rstate, rdst, dlam, S, J, scr, chunk, st);
          if (rc) return rc;
          // the Gaussian, or else the PSF, on the current spectrum and the
          // Jacobian (atmos_fit.cc:104-115)
          if (th->broadening > 0.0) {
            rc = spectral_convolve_launch(precision, nslot, nl, 4, lam_host.data(), dlam, th->broadening, S, ss.needj,
                                          st);
            if (!rc)
              rc = spectral_convolve_launch(precision, nslot, nl, 4 * nn, lam_host.data(), dlam, th->broadening, J,
                                            ss.needj, st);
            if (rc) return rc;
            hipLaunchKernelGGL(mask_jacobian_kernel<R>, dim3(nslot), dim3(256), 0, st, nl, 4 * nn, smask, J, ss.needj);
          } else if (th->npsf > 0) {
            rc = psf_convolve_launch(precision, nslot, nl, 4, th->npsf, dpsf, S, ss.needj, st);
            if (!rc) rc = psf_convolve_launch(precision, nslot, nl, 4 * nn, th->npsf, dpsf, J, ss.needj, st);
            if (rc) return rc;
            hipLaunchKernelGGL(mask_jacobian_kernel<R>, dim3(nslot), dim3(256), 0, st, nl, 4 * nn, smask, J, ss.needj);
          }
        } else if (jacobian == SNAPI_JAC_ADJOINT) {
          rc = launch_resp<R>(kl, a, ra, nslot, sstate, dzero, dlam, S, J, scr, chunk, st);
          if (rc) return rc;
        } else if (jacobian == SNAPI_JAC_TANGENT) {
          rc = launch_tangent<R>(kl, a, ra, nslot, sstate, nullptr, dlam, S, J, st);
          if (rc) return rc;
        } else {
          hipLaunchKernelGGL(fd_nodes_kernel, dim3((nslot + 127) / 128), dim3(128), 0, st, na, nslot, snodes, fdn, hv,
                             ss.needj);
          // (2n+1) copies of every slot's state with the node sets applied
          for (int k = 0; k < nv; ++k) {
            if (hipMemcpy2DAsync(fdstate + (size_t)k * ns, nv * ns * es, sstate, ns * es, ns * es, nslot,
                                    hipMemcpyDeviceToDevice, st) != hipSuccess) return SNAPI_ERR_DEVICE;
            hipLaunchKernelGGL(nodes_to_state_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, nf,
                               fdn + (size_t)k * nn, (size_t)nv * nn, fdstate + (size_t)k * ns, (size_t)nv * ns,
                               (R *)nullptr, ss.needj);
          }
          // fdstate is [slot][node set]: one synthesis over nslot * (2n+1) columns
          rc = synth_dev_masked(precision, lines, ray, nslot * nv, nd, fdstate, nl, dlam, Sall, ss.needj, nv, st);
          if (rc) return rc;
          hipLaunchKernelGGL(fdjac_kernel<R>, dim3(nslot), dim3(256), 0, st, na, nslot, nl, Sall, hv, smask, ws[0], ws[1],
                             ws[2], ws[3], S, J, ss.needj);
        }
        // chi^2 of the current model (where it changed)
        hipLaunchKernelGGL(residual_kernel<R>, dim3(nslot), dim3(256), 0, st, nslot, nl, S, sobs, smask, ws[0], ws[1],
                           ws[2], ws[3], r, ss.chi2, ss.needj);
        // normal equations and damped solve
        launch_lm_normal<R>(nslot, nl * 4, nn, J, r, JTJ, rhs, nullptr, ss.active, st);
        launch_lm_solve(nslot, nn, JTJ, rhs, ss.lm, corr, ss.active, st, dreg);
        // trial model and its synthesis
        hipLaunchKernelGGL(trial_kernel, dim3((nslot + 127) / 128), dim3(128), 0, st, na, nslot, snodes, corr, tnodes,
                           ss.active);
        hipLaunchKernelGGL(nodes_to_state_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, nf, tnodes,
                           (size_t)nn, tstate, ns, (R *)nullptr, ss.active);
        if (th) {
          // the trial's thermodynamics: temperature, hydrostatic equilibrium,
          // the synthesis path's equation of state (forward_evaluate)
          hipLaunchKernelGGL(tprofiles_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, ns, tstate, tnodes, 1,
                             Tv, ltv, ss.active);
          rc = chem_hequilibrium(cdev_he, nslot, nd, Tv, ltv, Ntv, ss.active, 1, st);
          if (rc) return rc;
          hipLaunchKernelGGL(tmaps_kernel<R>, dim3(nslot), dim3(128), 0, st, na, nslot, nd, nf, 1, Tv, Ntv, tstate,
                             (R *)nullptr, (R *)nullptr, tsd, ss.active);
          rc = chem_eos(cdev, nslot, nd, tsd, nullptr, nullptr, escr, ss.active, st);
          if (rc) return rc;
          hipLaunchKernelGGL((convert_kernel<double, R>), dim3(1024), dim3(256), 0, st, (size_t)nslot * ns, tsd, tstate);
        }
        rc = synth_dev_masked(precision, lines, th ? &ray_syn : ray, nslot, nd, tstate, nl, dlam, St, ss.active, 1, st);
        if (rc) return rc;
        if (th && th->broadening > 0.0) {  // forward_evaluate (atmos_fit.cc:717-737): Gaussian, then PSF
          rc = spectral_convolve_launch(precision, nslot, nl, 4, lam_host.data(), dlam, th->broadening, St, ss.active,
                                        st);
          if (rc) return rc;
        }
        if (th && th->npsf > 0) {
          rc = psf_convolve_launch(precision, nslot, nl, 4, th->npsf, dpsf, St, ss.active, st);
          if (rc) return rc;
        }
        hipLaunchKernelGGL(residual_kernel<R>, dim3(nslot), dim3(256), 0, st, nslot, nl, St, sobs, smask, ws[0], ws[1],
                           ws[2], ws[3], (R *)nullptr, ss.chi2t, ss.active);
      hipLaunchKernelGGL(update_kernel, dim3((nslot + 127) / 128), dim3(128), 0, st, na, nslot, max_iter, ss, snodes,
                         tnodes);
      return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
    };
    rc = run_iterations(iteration, st, count, pinned);
    if (rc) goto done;
    long long pixel_iters = 0;
    HIPCHK(hipMemcpyAsync(&pixel_iters, total, sizeof(long long), hipMemcpyDeviceToHost, st));
    if (hipStreamSynchronize(st) != hipSuccess) {
      rc = SNAPI_ERR_DEVICE;
      goto done;
    }
    if (!dev) {
      HIPCHK(hipMemcpy(nodes_out, onodes, (size_t)npix * nn * 8, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(chi2_out, ochi2, (size_t)npix * 8, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(iter_out, oiter, (size_t)npix * sizeof(int), hipMemcpyDeviceToHost));
    }
    if (stats) stats[0] = pixel_iters;
    rc = read_status();
  }
done:
  for (void *p : own) hipFree(p);
  hipFree(sstate); hipFree(tstate); hipFree(dzero);
  hipFree(sobs); hipFree(smask); hipFree(S); hipFree(St); hipFree(J); hipFree(r); hipFree(nmap);
  hipFree(fdstate); hipFree(Sall); hipFree(scr); hipFree(snodes);
  hipFree(fd_active); hipFree(fdTv); hipFree(fdltv); hipFree(fdNtv); hipFree(fdsd); hipFree(fdescr);
  hipFree(tnodes); hipFree(JTJ); hipFree(rhs); hipFree(corr); hipFree(fdn);
  hipFree(hv); hipFree(dbl); hipFree(ints); hipFree(queue);
  chem_free(cdev); chem_free(cdev_he);
  hipFree(Tv); hipFree(ltv); hipFree(Ntv); hipFree(sd); hipFree(rsd); hipFree(dsd); hipFree(tsd); hipFree(escr);
  hipFree(dTd); hipFree(dNd); hipFree(hscr);
  hipFree(rstate); hipFree(rdst); hipFree(nmap2);
  hipFree(dreg); hipFree(dpsf);
  if (pinned) (void)hipHostFree(pinned);
  return rc;
}

// Stokes vector and raw node Jacobian dS/dnode of npix pixels at given node
// values (obs_stokes_responses_to_nodes, atmos/atmos_fit.cc:1205-1352, for
// the kinematic and magnetic nodes).
template <typename R>
static int node_jacobian_impl(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                              NodeArgs na, int npix, int nd, const void *state_h, const double *nodes_h, int nlambda,
                              const double *lam_h, void *stokes_h, void *J_h) {
  KLines kl;
  SynthArgs a;
  int rc = prepare_synth(lines, ray, 1, nd, nlambda, kl, a);
  if (rc) return rc;
  const int nf = a.nfield, nn = na.nnode, nl = nlambda, nv = 2 * nn + 1;
  const size_t es = sizeof(R), ns = (size_t)nf * nd;
  for (int n = 0; n < nn; ++n) na.scale[n] = 1.0;  // unscaled derivatives
  R *dstate = nullptr, *S = nullptr, *J = nullptr, *nmap = nullptr, *dzero = nullptr, *scr = nullptr;
  R *fdstate = nullptr, *Sall = nullptr, *ones = nullptr;
  double *dlam = nullptr, *dnodes = nullptr, *fdn = nullptr, *hv = nullptr;
  RespArgs ra = {};
  HIPCHK(hipMalloc((void **)&dstate, (size_t)npix * ns * es));
  HIPCHK(hipMalloc((void **)&S, (size_t)npix * nl * 4 * es));
  HIPCHK(hipMalloc((void **)&J, (size_t)npix * nl * 4 * nn * es));
  HIPCHK(hipMalloc((void **)&nmap, (size_t)npix * nn * nd * es));
  HIPCHK(hipMalloc((void **)&dlam, (size_t)nl * 8));
  HIPCHK(hipMalloc((void **)&dnodes, (size_t)npix * nn * 8));
  HIPCHK(hipMemcpy(dstate, state_h, (size_t)npix * ns * es, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dlam, lam_h, (size_t)nl * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dnodes, nodes_h, (size_t)npix * nn * 8, hipMemcpyHostToDevice));
  reset_status();
  hipLaunchKernelGGL(nodes_to_state_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, nf, dnodes, (size_t)nn,
                     dstate, ns, nmap, (const int *)nullptr);
  ra.nseed = 4;
  ra.sink = SINK_NODE;
  for (int s = 0; s < 4; ++s) ra.ws[s] = 1.0;
  ra.nnode = nn;
  for (int n = 0; n < nn; ++n) {
    ra.pidx[n] = na.code[n] == NP_VT ? 2 : na.code[n] == NP_VS ? 3 : na.code[n] == NP_B ? 4 : na.code[n] == NP_THETA ? 5 : 6;
    ra.nscale[n] = 1.0;
  }
  ra.nmap = nmap;
  if (jacobian == SNAPI_JAC_TANGENT) {
    rc = launch_tangent<R>(kl, a, ra, npix, dstate, nullptr, dlam, S, J, 0);
    if (rc) goto done;
  } else if (jacobian == SNAPI_JAC_ADJOINT) {
    const size_t per_col = resp_scratch_per_col(nl, nd, es);
    size_t c = ((size_t)16 << 30) / (per_col ? per_col : 1);
    if (c < 1) c = 1;
    if (c > (size_t)npix) c = npix;
    HIPCHK(hipMalloc((void **)&scr, per_col * c));
    HIPCHK(hipMalloc((void **)&dzero, (size_t)npix * 2 * ns * es));
    HIPCHK(hipMemset(dzero, 0, (size_t)npix * 2 * ns * es));
    rc = launch_resp<R>(kl, a, ra, npix, dstate, dzero, dlam, S, J, scr, (int)c, 0);
    if (rc) goto done;
  } else {
    HIPCHK(hipMalloc((void **)&fdstate, (size_t)npix * nv * ns * es));
    HIPCHK(hipMalloc((void **)&Sall, (size_t)npix * nv * nl * 4 * es));
    HIPCHK(hipMalloc((void **)&fdn, (size_t)npix * nv * nn * 8));
    HIPCHK(hipMalloc((void **)&hv, (size_t)npix * nn * 8));
    HIPCHK(hipMalloc((void **)&ones, (size_t)npix * nl * es));
    hipLaunchKernelGGL(fill_kernel<R>, dim3(256), dim3(256), 0, 0, (size_t)npix * nl, R(1), ones);
    hipLaunchKernelGGL(fd_nodes_kernel, dim3((npix + 127) / 128), dim3(128), 0, 0, na, npix, dnodes, fdn, hv,
                       (const int *)nullptr);
    for (int k = 0; k < nv; ++k) {
      HIPCHK(hipMemcpy2DAsync(fdstate + (size_t)k * ns, nv * ns * es, dstate, ns * es, ns * es, npix,
                              hipMemcpyDeviceToDevice, 0));
      hipLaunchKernelGGL(nodes_to_state_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, nf,
                         fdn + (size_t)k * nn, (size_t)nv * nn, fdstate + (size_t)k * ns, (size_t)nv * ns,
                         (R *)nullptr, (const int *)nullptr);
    }
    rc = snapi_synth_stokes_dev(precision, lines, ray, npix * nv, nd, fdstate, nl, dlam, Sall, nullptr);
    if (rc) goto done;
    hipLaunchKernelGGL(fdjac_kernel<R>, dim3(npix), dim3(256), 0, 0, na, npix, nl, Sall, hv, ones, 1.0, 1.0, 1.0,
                       1.0, S, J, (const int *)nullptr);
  }
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    rc = SNAPI_ERR_DEVICE;
    goto done;
  }
  HIPCHK(hipMemcpy(stokes_h, S, (size_t)npix * nl * 4 * es, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(J_h, J, (size_t)npix * nl * 4 * nn * es, hipMemcpyDeviceToHost));
  rc = read_status();
done:
  hipFree(dstate); hipFree(S); hipFree(J); hipFree(nmap); hipFree(dzero); hipFree(scr); hipFree(fdstate);
  hipFree(Sall); hipFree(ones); hipFree(dlam); hipFree(dnodes); hipFree(fdn); hipFree(hv);
  return rc;
}

// obs_stokes_responses_to_nodes (atmos_fit.cc:1205-1352) with temperature
// nodes: build_from_nodes (temperature interpolation, hydrostatic
// equilibrium on the state's tau grid, atmos_pp.cc:58-180, 372-455), the
// response path's state and thermodynamic derivatives (obs_stokes_responses'
// preamble), the adjoint per-depth responses contracted with the node
// responses: T and Nt rows for the temperature nodes (2 x delta_T/2 rebuilds
// each), the interpolation's differences for the others.
template <typename R>
static int node_jacobian_eos_impl(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                                  NodeArgs na,
                                  const snapi_chem_table *chem, const snapi_chem_table *chem_he,
                                  const snapi_state_map *map, int npix, int nd, const void *state_h,
                                  const double *nodes_h, int nlambda, const double *lam_h, void *stokes_h, void *J_h) {
  KLines kl;
  SynthArgs a;
  int rc = prepare_synth(lines, ray, 1, nd, nlambda, kl, a);
  if (rc) return rc;
  const int nf = a.nfield, nn = na.nnode, nl = nlambda;
  if (nf != SNAPI_NF_BASE + 2 * map->nline) return SNAPI_ERR_ARG;
  int ntn = 0;
  for (int n = 0; n < nn; ++n) ntn += na.code[n] == NP_T;
  const int nvar = na.analytic ? 1 : 1 + 2 * ntn;
  const size_t es = sizeof(R), ns = (size_t)nf * nd;
  for (int n = 0; n < nn; ++n) na.scale[n] = 1.0;  // unscaled derivatives
  ChemDevice cdev_he;
  double *dTd = nullptr, *dNd = nullptr, *hscr = nullptr;
  R *dstate = nullptr, *rstate = nullptr, *rdst = nullptr, *S = nullptr, *J = nullptr, *nmap = nullptr,
    *nmap2 = nullptr, *scr = nullptr;
  double *dlam = nullptr, *dnodes = nullptr, *Tv = nullptr, *ltv = nullptr, *Ntv = nullptr, *sd = nullptr,
         *rsd = nullptr, *dsd = nullptr;
  RespArgs ra = {};
  HIPCHK(hipMalloc((void **)&dstate, (size_t)npix * ns * es));
  HIPCHK(hipMalloc((void **)&rstate, (size_t)npix * ns * es));
  HIPCHK(hipMalloc((void **)&rdst, (size_t)npix * 2 * ns * es));
  HIPCHK(hipMalloc((void **)&S, (size_t)npix * nl * 4 * es));
  HIPCHK(hipMalloc((void **)&J, (size_t)npix * nl * 4 * nn * es));
  HIPCHK(hipMalloc((void **)&nmap, (size_t)npix * nn * nd * es));
  HIPCHK(hipMalloc((void **)&nmap2, (size_t)npix * nn * nd * es));
  HIPCHK(hipMalloc((void **)&dlam, (size_t)nl * 8));
  HIPCHK(hipMalloc((void **)&dnodes, (size_t)npix * nn * 8));
  HIPCHK(hipMalloc((void **)&Tv, (size_t)npix * nvar * nd * 8));
  HIPCHK(hipMalloc((void **)&ltv, (size_t)npix * nvar * nd * 8));
  HIPCHK(hipMalloc((void **)&Ntv, (size_t)npix * nvar * nd * 8));
  HIPCHK(hipMalloc((void **)&sd, (size_t)npix * ns * 8));
  HIPCHK(hipMalloc((void **)&rsd, (size_t)npix * ns * 8));
  HIPCHK(hipMalloc((void **)&dsd, (size_t)npix * 2 * ns * 8));
  HIPCHK(hipMemcpy(dstate, state_h, (size_t)npix * ns * es, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dlam, lam_h, (size_t)nl * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(dnodes, nodes_h, (size_t)npix * nn * 8, hipMemcpyHostToDevice));
  reset_status();
  hipLaunchKernelGGL(nodes_to_state_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, nf, dnodes, (size_t)nn,
                     dstate, ns, nmap, (const int *)nullptr);
  hipLaunchKernelGGL(tprofiles_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, ns, dstate, dnodes, nvar, Tv,
                     ltv, (const int *)nullptr);
  if (hipGetLastError() != hipSuccess) {
    rc = SNAPI_ERR_DEVICE;
    goto done;
  }
  rc = snapi_hequilibrium_dev(chem_he, npix * nvar, nd, Tv, ltv, Ntv, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr, nullptr);
  if (rc) goto done;
  if (na.analytic) {
    if (ntn) {
      HIPCHK(hipMalloc((void **)&dTd, (size_t)npix * ntn * nd * 8));
      HIPCHK(hipMalloc((void **)&dNd, (size_t)npix * ntn * nd * 8));
      HIPCHK(hipMalloc((void **)&hscr, chem_he_tangent_scratch(npix, nd)));
      rc = chem_upload(chem_he, nullptr, cdev_he, 0);
      if (rc) goto done;
      hipLaunchKernelGGL(tderiv_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, ntn, dnodes, ltv, dTd, nmap,
                         (const int *)nullptr);
      rc = chem_he_tangent(cdev_he, npix, nd, Tv, ltv, Ntv, ntn, dTd, dNd, hscr, nullptr, 0);
      if (rc) goto done;
    }
    hipLaunchKernelGGL(tmap2_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, ntn, dNd, nmap2,
                       (const int *)nullptr);
    hipLaunchKernelGGL(tmaps_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, nf, 1, Tv, Ntv, dstate,
                       (R *)nullptr, (R *)nullptr, sd, (const int *)nullptr);
  } else {
    hipLaunchKernelGGL(tmaps_kernel<R>, dim3(npix), dim3(128), 0, 0, na, npix, nd, nf, nvar, Tv, Ntv, dstate, nmap,
                       nmap2, sd, (const int *)nullptr);
  }
  rc = snapi_eos_state_dev(chem, map, npix, nd, sd, rsd, dsd, nullptr);
  if (rc) goto done;
  hipLaunchKernelGGL((convert_kernel<double, R>), dim3(1024), dim3(256), 0, 0, (size_t)npix * ns, rsd, rstate);
  hipLaunchKernelGGL((convert_kernel<double, R>), dim3(1024), dim3(256), 0, 0, (size_t)npix * 2 * ns, dsd, rdst);
  ra.nseed = 4;
  ra.sink = SINK_NODE;
  for (int s = 0; s < 4; ++s) ra.ws[s] = 1.0;
  ra.nnode = nn;
  for (int n = 0; n < nn; ++n) {
    const int c = na.code[n];
    ra.pidx[n] = c == NP_T ? 0 : c == NP_VT ? 2 : c == NP_VS ? 3 : c == NP_B ? 4 : c == NP_THETA ? 5 : 6;
    ra.pidx2[n] = c == NP_T ? 1 : -1;
    ra.nscale[n] = 1.0;
  }
  ra.nmap = nmap;
  ra.nmap2 = nmap2;
  {
    // the response path's H- at the window's middle wavelength, as
    // obs_stokes_responses_to_nodes evaluates it (mol.cc:414-469)
    SynthArgs a_resp = a;
    a_resp.hminus_mid = 1;
    if (jacobian == SNAPI_JAC_TANGENT) {
      rc = launch_tangent<R>(kl, a_resp, ra, npix, rstate, rdst, dlam, S, J, 0);
    } else {
      const size_t per_col = resp_scratch_per_col(nl, nd, es);
      size_t c = ((size_t)16 << 30) / (per_col ? per_col : 1);
      if (c < 1) c = 1;
      if (c > (size_t)npix) c = npix;
      HIPCHK(hipMalloc((void **)&scr, per_col * c));
      rc = launch_resp<R>(kl, a_resp, ra, npix, rstate, rdst, dlam, S, J, scr, (int)c, 0);
    }
    if (rc) goto done;
  }
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    rc = SNAPI_ERR_DEVICE;
    goto done;
  }
  HIPCHK(hipMemcpy(stokes_h, S, (size_t)npix * nl * 4 * es, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(J_h, J, (size_t)npix * nl * 4 * nn * es, hipMemcpyDeviceToHost));
  rc = read_status();
done:
  hipFree(dstate); hipFree(rstate); hipFree(rdst); hipFree(S); hipFree(J); hipFree(nmap); hipFree(nmap2);
  hipFree(scr); hipFree(dlam); hipFree(dnodes); hipFree(Tv); hipFree(ltv); hipFree(Ntv); hipFree(sd);
  hipFree(rsd); hipFree(dsd); hipFree(dTd); hipFree(dNd); hipFree(hscr);
  chem_free(cdev_he);
  return rc;
}

}  // namespace snapi

using namespace snapi;

// Checks the node model and fills na; strips SNAPI_JAC_ANALYTIC_MAPS from
// jacobian into na.analytic.
static int setup_nodes(const snapi_line_table *lines, const snapi_ray *ray, const snapi_node_model *model, int npix,
                       int nd, int nlambda, int &jacobian, int max_iter, NodeArgs &na, bool temperature = false) {
  const int analytic = (jacobian & SNAPI_JAC_ANALYTIC_MAPS) != 0;
  jacobian &= ~SNAPI_JAC_ANALYTIC_MAPS;
  if (!lines || !ray || !model) return SNAPI_ERR_ARG;
  if (npix <= 0 || nd < 3 || nlambda <= 0 || max_iter <= 0) return SNAPI_ERR_ARG;
  if (model->nnode <= 0 || model->nnode > SNAPI_MAX_NODES - 1) return SNAPI_ERR_ARG;
  if (jacobian != SNAPI_JAC_ADJOINT && jacobian != SNAPI_JAC_DIFFERENCES && jacobian != SNAPI_JAC_TANGENT)
    return SNAPI_ERR_ARG;
  if (cos(ray->theta) > -0.999999) return SNAPI_ERR_UNSUPPORTED;  // node responses assume the mu = 1 ray
  na = NodeArgs{};
  na.nnode = model->nnode;
  na.analytic = analytic;
  for (int n = 0; n < na.nnode; ++n) {
    const int c = model->param[n];
    // temperature nodes need the equation of state (the _eos entry points)
    if (c < (temperature ? NP_T : NP_VT) || c > NP_PHI) return SNAPI_ERR_UNSUPPORTED;
    na.code[n] = c;
    na.ltau[n] = model->logtau[n];
    na.scale[n] = c == NP_T ? 5000.0 : c == NP_VT ? 1E5 : c == NP_VS ? 1E5 : c == NP_B ? 1000.0 : 1.0;
    na.step[n] = c == NP_T ? kNodeStepT : (c == NP_VT || c == NP_VS) ? 20.0 : c == NP_B ? 0.01 : 0.001;
  }
  for (int p = NP_T; p <= NP_PHI; ++p)
    if (model->reg_type[p] < 0 || model->reg_type[p] > 2 || !std::isfinite(model->reg_alpha[p])) return SNAPI_ERR_ARG;
  return SNAPI_OK;
}

// The regularisation of atmosphere::regularize_hessian / regularize_parameter
// (atmos/atmos_fit.cc:1371-1480): for each parameter code p with a
// regularisation type, R over the parameter's first run of consecutive
// nodes [from, to] (to is the node before the first node of another
// parameter after from; none when the run reaches the end of the list, as
// the reference's to stays 0 there), departure from zero R[i][i] = 1 on
// the run, Tikhonov R[i][i] = 1, R[i][i-1] = -1 for i in (from, to] and
// R[from][from] = -1, R[from][from+1] = 1; G += alpha^2 R^T R.  G [nn][nn]
// row-major; returns false when no parameter is regularised.
static bool reg_matrix(const snapi_node_model *model, const NodeArgs &na, double *G) {
  const int n = na.nnode;
  bool on = false;
  for (int i = 0; i < n * n; ++i) G[i] = 0.0;
  for (int p = NP_T; p <= NP_PHI; ++p) {
    const int type = model->reg_type[p];
    if (!type) continue;
    int from = -1;
    for (int i = 0; i < n; ++i)
      if (na.code[i] == p) {
        from = i;
        break;
      }
    if (from < 0) continue;  // no node of this parameter
    int to = -1;
    for (int i = from; i < n; ++i)
      if (na.code[i] != p) {
        to = i - 1;
        break;
      }
    double R[SNAPI_MAX_NODES][SNAPI_MAX_NODES] = {};
    if (type == 2) {
      for (int i = from + 1; i <= to; ++i) {
        R[i][i] = 1.0;
        R[i][i - 1] = -1.0;
      }
      R[from][from] = -1.0;
      if (from + 1 < n) R[from][from + 1] = 1.0;
    } else {
      for (int i = from; i <= to; ++i) R[i][i] = 1.0;
    }
    const double a2 = model->reg_alpha[p] * model->reg_alpha[p];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double rr = 0.0;
        for (int k = 0; k < n; ++k) rr += R[k][i] * R[k][j];
        G[i * n + j] += rr * a2;
      }
    on = true;
  }
  return on;
}

// Checks the fit's instrument effects (null: none) into th.
static int instrument_args(const snapi_instrument *inst, ThermoArgs &th) {
  if (!inst) return SNAPI_OK;
  if (!(inst->broadening >= 0.0) || inst->npsf < 0 || inst->npsf > 1024 || (inst->npsf > 0 && !inst->psf))
    return SNAPI_ERR_ARG;
  th.broadening = inst->broadening;
  th.npsf = inst->npsf;
  th.psf = inst->psf;
  return SNAPI_OK;
}

extern "C" int snapi_invert(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                            const snapi_node_model *model, int npix, int nd, const void *state, int nlambda,
                            const double *lam, const void *obs, const void *mask, const double *wstokes,
                            double lm_start, int max_iter, int nslot, double *nodes, double *chi2, int *iterations,
                            long long *pixel_iterations) {
  if (!state || !lam || !obs || !wstokes || !nodes || !chi2 || !iterations) return SNAPI_ERR_ARG;
  NodeArgs na;
  int rc = setup_nodes(lines, ray, model, npix, nd, nlambda, jacobian, max_iter, na);
  if (rc) return rc;
  if (nslot > npix) nslot = npix;  // nslot <= 0: chosen from the free device memory
  double G[SNAPI_MAX_NODES * SNAPI_MAX_NODES];
  const double *reg = reg_matrix(model, na, G) ? G : nullptr;
  if (precision == SNAPI_FP64)
    return invert_impl<double>(precision, jacobian, lines, ray, na, npix, nd, state, nlambda, lam, obs, mask, wstokes,
                               lm_start, max_iter, nslot, nodes, nodes, chi2, iterations, pixel_iterations, false, 0,
                               nullptr, reg);
  if (precision == SNAPI_MIXED)
    return invert_impl<float>(precision, jacobian, lines, ray, na, npix, nd, state, nlambda, lam, obs, mask, wstokes,
                              lm_start, max_iter, nslot, nodes, nodes, chi2, iterations, pixel_iterations, false, 0,
                              nullptr, reg);
  return SNAPI_ERR_ARG;
}

extern "C" int snapi_invert_dev(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                                const snapi_node_model *model, int npix, int nd, const void *state_dev, int nlambda,
                                const double *lam_dev, const void *obs_dev, const void *mask_dev,
                                const double *wstokes, double lm_start, int max_iter, int nslot,
                                const double *nodes0_dev, double *nodes_dev, double *chi2_dev, int *iterations_dev,
                                long long *pixel_iterations, void *stream) {
  if (!state_dev || !lam_dev || !obs_dev || !wstokes || !nodes0_dev || !nodes_dev || !chi2_dev || !iterations_dev)
    return SNAPI_ERR_ARG;
  NodeArgs na;
  int rc = setup_nodes(lines, ray, model, npix, nd, nlambda, jacobian, max_iter, na);
  if (rc) return rc;
  if (nslot > npix) nslot = npix;  // nslot <= 0: chosen from the free device memory
  hipStream_t st = (hipStream_t)stream;
  double G[SNAPI_MAX_NODES * SNAPI_MAX_NODES];
  const double *reg = reg_matrix(model, na, G) ? G : nullptr;
  if (precision == SNAPI_FP64)
    return invert_impl<double>(precision, jacobian, lines, ray, na, npix, nd, state_dev, nlambda, lam_dev, obs_dev,
                               mask_dev, wstokes, lm_start, max_iter, nslot, nodes0_dev, nodes_dev, chi2_dev,
                               iterations_dev, pixel_iterations, true, st, nullptr, reg);
  if (precision == SNAPI_MIXED)
    return invert_impl<float>(precision, jacobian, lines, ray, na, npix, nd, state_dev, nlambda, lam_dev, obs_dev,
                              mask_dev, wstokes, lm_start, max_iter, nslot, nodes0_dev, nodes_dev, chi2_dev,
                              iterations_dev, pixel_iterations, true, st, nullptr, reg);
  return SNAPI_ERR_ARG;
}

extern "C" int snapi_node_jacobian(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                                   const snapi_node_model *model, int npix, int nd, const void *state,
                                   const double *nodes, int nlambda, const double *lam, void *stokes, void *J) {
  if (!state || !nodes || !lam || !stokes || !J) return SNAPI_ERR_ARG;
  NodeArgs na;
  int rc = setup_nodes(lines, ray, model, npix, nd, nlambda, jacobian, 1, na);
  if (rc) return rc;
  if (precision == SNAPI_FP64)
    return node_jacobian_impl<double>(precision, jacobian, lines, ray, na, npix, nd, state, nodes, nlambda, lam,
                                      stokes, J);
  if (precision == SNAPI_MIXED)
    return node_jacobian_impl<float>(precision, jacobian, lines, ray, na, npix, nd, state, nodes, nlambda, lam,
                                     stokes, J);
  return SNAPI_ERR_ARG;
}

extern "C" int snapi_node_jacobian_eos(int precision, int jacobian, const snapi_line_table *lines,
                                       const snapi_ray *ray, const snapi_node_model *model,
                                       const snapi_chem_table *chem, const snapi_chem_table *chem_he,
                                       const snapi_state_map *map, int npix, int nd, const void *state,
                                       const double *nodes, int nlambda, const double *lam, void *stokes, void *J) {
  if (!state || !nodes || !lam || !stokes || !J || !chem || !chem_he || !map) return SNAPI_ERR_ARG;
  NodeArgs na;
  int rc = setup_nodes(lines, ray, model, npix, nd, nlambda, jacobian, 1, na, true);
  if (rc) return rc;
  int ntn = 0;
  for (int n = 0; n < na.nnode; ++n) ntn += na.code[n] == NP_T;
  if (ntn < 1) return SNAPI_ERR_ARG;  // build_from_nodes needs a temperature node
  if (jacobian != SNAPI_JAC_ADJOINT && jacobian != SNAPI_JAC_TANGENT) return SNAPI_ERR_UNSUPPORTED;
  if (precision == SNAPI_FP64)
    return node_jacobian_eos_impl<double>(precision, jacobian, lines, ray, na, chem, chem_he, map, npix, nd, state,
                                          nodes, nlambda, lam, stokes, J);
  if (precision == SNAPI_MIXED)
    return node_jacobian_eos_impl<float>(precision, jacobian, lines, ray, na, chem, chem_he, map, npix, nd, state,
                                         nodes, nlambda, lam, stokes, J);
  return SNAPI_ERR_ARG;
}

extern "C" int snapi_invert_eos_dev(int precision, int jacobian, const snapi_line_table *lines,
                                    const snapi_ray *ray, const snapi_node_model *model,
                                    const snapi_chem_table *chem, const snapi_chem_table *chem_he,
                                    const snapi_state_map *map, int npix, int nd, const void *state_dev, int nlambda,
                                    const double *lam_dev, const void *obs_dev, const void *mask_dev,
                                    const double *wstokes, const snapi_instrument *inst, double lm_start,
                                    int max_iter, int nslot, const double *nodes0_dev, double *nodes_dev,
                                    double *chi2_dev, int *iterations_dev, long long *pixel_iterations,
                                    void *stream) {
  if (!state_dev || !lam_dev || !obs_dev || !wstokes || !nodes0_dev || !nodes_dev || !chi2_dev || !iterations_dev ||
      !chem || !chem_he || !map)
    return SNAPI_ERR_ARG;
  NodeArgs na;
  int rc = setup_nodes(lines, ray, model, npix, nd, nlambda, jacobian, max_iter, na, true);
  if (rc) return rc;
  int ntn = 0;
  for (int n = 0; n < na.nnode; ++n) ntn += na.code[n] == NP_T;
  if (ntn < 1) return SNAPI_ERR_ARG;
  if (nslot > npix) nslot = npix;  // nslot <= 0: chosen from the free device memory
  ThermoArgs th{chem, chem_he, map, 0.0, 0, nullptr};
  double G[SNAPI_MAX_NODES * SNAPI_MAX_NODES];
  const double *reg = reg_matrix(model, na, G) ? G : nullptr;
  rc = instrument_args(inst, th);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (precision == SNAPI_FP64)
    return invert_impl<double>(precision, jacobian, lines, ray, na, npix, nd, state_dev, nlambda, lam_dev, obs_dev,
                               mask_dev, wstokes, lm_start, max_iter, nslot, nodes0_dev, nodes_dev, chi2_dev,
                               iterations_dev, pixel_iterations, true, st, &th, reg);
  if (precision == SNAPI_MIXED)
    return invert_impl<float>(precision, jacobian, lines, ray, na, npix, nd, state_dev, nlambda, lam_dev, obs_dev,
                              mask_dev, wstokes, lm_start, max_iter, nslot, nodes0_dev, nodes_dev, chi2_dev,
                              iterations_dev, pixel_iterations, true, st, &th, reg);
  return SNAPI_ERR_ARG;
}

extern "C" int snapi_invert_eos(int precision, int jacobian, const snapi_line_table *lines, const snapi_ray *ray,
                                const snapi_node_model *model, const snapi_chem_table *chem,
                                const snapi_chem_table *chem_he, const snapi_state_map *map, int npix, int nd,
                                const void *state, int nlambda, const double *lam, const void *obs, const void *mask,
                                const double *wstokes, const snapi_instrument *inst, double lm_start, int max_iter,
                                int nslot, double *nodes, double *chi2, int *iterations,
                                long long *pixel_iterations) {
  if (!state || !lam || !obs || !wstokes || !nodes || !chi2 || !iterations || !chem || !chem_he || !map)
    return SNAPI_ERR_ARG;
  NodeArgs na;
  int rc = setup_nodes(lines, ray, model, npix, nd, nlambda, jacobian, max_iter, na, true);
  if (rc) return rc;
  int ntn = 0;
  for (int n = 0; n < na.nnode; ++n) ntn += na.code[n] == NP_T;
  if (ntn < 1) return SNAPI_ERR_ARG;
  if (nslot > npix) nslot = npix;  // nslot <= 0: chosen from the free device memory
  ThermoArgs th{chem, chem_he, map, 0.0, 0, nullptr};
  double G[SNAPI_MAX_NODES * SNAPI_MAX_NODES];
  const double *reg = reg_matrix(model, na, G) ? G : nullptr;
  rc = instrument_args(inst, th);
  if (rc) return rc;
  if (precision == SNAPI_FP64)
    return invert_impl<double>(precision, jacobian, lines, ray, na, npix, nd, state, nlambda, lam, obs, mask, wstokes,
                               lm_start, max_iter, nslot, nodes, nodes, chi2, iterations, pixel_iterations, false, 0,
                               &th, reg);
  if (precision == SNAPI_MIXED)
    return invert_impl<float>(precision, jacobian, lines, ray, na, npix, nd, state, nlambda, lam, obs, mask, wstokes,
                              lm_start, max_iter, nslot, nodes, nodes, chi2, iterations, pixel_iterations, false, 0,
                              &th, reg);
  return SNAPI_ERR_ARG;
}

// build_from_nodes (atmos_pp.cc:58-180) with popsetup and compute_op_referent:
// the node-controlled fields, the temperature, hydrostatic equilibrium and
// the synthesis path's equation-of-state fields of each pixel's state.
extern "C" int snapi_build_from_nodes_dev(const snapi_node_model *model, const snapi_chem_table *chem,
                                          const snapi_chem_table *chem_he, const snapi_state_map *map, int npix,
                                          int nd, double *state_dev, const double *nodes_dev, void *stream) {
  if (!model || !chem || !chem_he || !map || !state_dev || !nodes_dev || npix <= 0 || nd < 3 ||
      nd > SNAPI_MAX_ND || model->nnode <= 0 || model->nnode > SNAPI_MAX_NODES - 1)
    return SNAPI_ERR_ARG;
  NodeArgs na{};
  na.nnode = model->nnode;
  int ntn = 0;
  for (int n = 0; n < na.nnode; ++n) {
    const int c = model->param[n];
    if (c < NP_T || c > NP_PHI) return SNAPI_ERR_UNSUPPORTED;
    na.code[n] = c;
    na.ltau[n] = model->logtau[n];
    ntn += c == NP_T;
  }
  if (ntn < 1) return SNAPI_ERR_ARG;
  const int nf = SNAPI_NF_BASE + 2 * map->nline;
  const size_t ns = (size_t)nf * nd;
  hipStream_t st = (hipStream_t)stream;
  ChemDevice cdev, cdev_he;
  double *Tv = nullptr, *ltv = nullptr, *Ntv = nullptr, *escr = nullptr;
  int rc = chem_upload(chem, map, cdev, st);
  if (!rc) rc = chem_upload(chem_he, nullptr, cdev_he, st);
  if (!rc && (hipMalloc((void **)&Tv, (size_t)npix * nd * 8) != hipSuccess ||
              hipMalloc((void **)&ltv, (size_t)npix * nd * 8) != hipSuccess ||
              hipMalloc((void **)&Ntv, (size_t)npix * nd * 8) != hipSuccess ||
              hipMalloc((void **)&escr, chem_eos_scratch(cdev, (long)npix * nd, false)) != hipSuccess))
    rc = SNAPI_ERR_DEVICE;
  if (!rc) {
    hipLaunchKernelGGL(nodes_to_state_kernel<double>, dim3(npix), dim3(128), 0, st, na, npix, nd, nf, nodes_dev,
                       (size_t)na.nnode, state_dev, ns, (double *)nullptr, (const int *)nullptr);
    hipLaunchKernelGGL(tprofiles_kernel<double>, dim3(npix), dim3(128), 0, st, na, npix, nd, ns, state_dev,
                       nodes_dev, 1, Tv, ltv, (const int *)nullptr);
    rc = hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
  }
  if (!rc) rc = chem_hequilibrium(cdev_he, npix, nd, Tv, ltv, Ntv, nullptr, 1, st);
  if (!rc) {
    hipLaunchKernelGGL(set_tn_kernel, dim3(npix), dim3(128), 0, st, nd, nf, Tv, Ntv, state_dev);
    rc = hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
  }
  if (!rc) rc = chem_eos(cdev, npix, nd, state_dev, nullptr, nullptr, escr, nullptr, st);
  if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = SNAPI_ERR_DEVICE;
  chem_free(cdev);
  chem_free(cdev_he);
  hipFree(Tv); hipFree(ltv); hipFree(Ntv); hipFree(escr);
  return rc;
}

extern "C" int snapi_build_from_nodes(const snapi_node_model *model, const snapi_chem_table *chem,
                                      const snapi_chem_table *chem_he, const snapi_state_map *map, int npix, int nd,
                                      double *state, const double *nodes) {
  if (!model || !map || !state || !nodes || npix <= 0 || nd <= 0) return SNAPI_ERR_ARG;
  const size_t n = (size_t)npix * (SNAPI_NF_BASE + 2 * map->nline) * nd;
  double *ds = nullptr, *dn = nullptr;
  int rc = SNAPI_OK;
  if (hipMalloc((void **)&ds, n * 8) != hipSuccess || hipMalloc((void **)&dn, (size_t)npix * model->nnode * 8) != hipSuccess ||
      hipMemcpy(ds, state, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dn, nodes, (size_t)npix * model->nnode * 8, hipMemcpyHostToDevice) != hipSuccess)
    rc = SNAPI_ERR_DEVICE;
  if (!rc) rc = snapi_build_from_nodes_dev(model, chem, chem_he, map, npix, nd, ds, dn, nullptr);
  if (!rc && hipMemcpy(state, ds, n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = SNAPI_ERR_DEVICE;
  hipFree(ds);
  hipFree(dn);
  return rc;
}
