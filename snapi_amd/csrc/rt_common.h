// Device building blocks shared by the synthesis, response and adjoint
// kernels: Humlicek W4 Voigt/Faraday-Voigt, short-characteristics weights,
// the Fritsch-Butland opacity derivative and the pivot-free 4x4 solve.
// Everything is templated on the arithmetic type R (double or float).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/snapi_gpu.h"

namespace snapi {

constexpr double kC = 2.99792458E10;        // cm/s      (cmlib/const.h:11)
constexpr double kH = 6.6260687652E-27;     // erg s     (cmlib/const.h:12)
constexpr double kK = 1.3806503E-16;        // erg/K     (cmlib/const.h:13)
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kSqrtPi = 1.77245386399759;  // as cmlib/profiles.cc:33
constexpr double kThomson = 6.65E-25;

template <typename R> __device__ __forceinline__ R rexp(R x);
template <> __device__ __forceinline__ double rexp<double>(double x) { return exp(x); }
template <> __device__ __forceinline__ float rexp<float>(float x) { return expf(x); }
template <typename R> __device__ __forceinline__ R rsin(R x);
template <> __device__ __forceinline__ double rsin<double>(double x) { return sin(x); }
template <> __device__ __forceinline__ float rsin<float>(float x) { return sinf(x); }
template <typename R> __device__ __forceinline__ R rsqrt_(R x);
template <> __device__ __forceinline__ double rsqrt_<double>(double x) { return sqrt(x); }
template <> __device__ __forceinline__ float rsqrt_<float>(float x) { return sqrtf(x); }
template <typename R> __device__ __forceinline__ R rcos(R x);
template <> __device__ __forceinline__ double rcos<double>(double x) { return cos(x); }
template <> __device__ __forceinline__ float rcos<float>(float x) { return cosf(x); }
template <typename R> __device__ __forceinline__ R rexp2(R x);
template <> __device__ __forceinline__ double rexp2<double>(double x) { return exp2(x); }
template <> __device__ __forceinline__ float rexp2<float>(float x) { return exp2f(x); }

// Reciprocal: IEEE division in double; the hardware v_rcp_f32 (1 ulp) in
// float, where the correctly rounded division costs about ten instructions.
// The float path's sine, cosine and exponential in the Voigt function's fourth
// region, where |Im u| = 2 a |v| < 12 and Re u = a^2 - v^2 < 1: the
// hardware instructions (v_sin_f32 on the argument in turns, v_exp_f32 on
// base 2), within a few 1e-7 of the library's results there; the double path
// keeps the library functions.
template <typename R> __device__ __forceinline__ R rsin_fast(R x) { return rsin<R>(x); }
template <> __device__ __forceinline__ float rsin_fast<float>(float x) {
  return __builtin_amdgcn_sinf(x * 0.15915494309189535f);
}
template <typename R> __device__ __forceinline__ R rcos_fast(R x) { return rcos<R>(x); }
template <> __device__ __forceinline__ float rcos_fast<float>(float x) {
  return __builtin_amdgcn_cosf(x * 0.15915494309189535f);
}
template <typename R> __device__ __forceinline__ R rexp_fast(R x) { return rexp<R>(x); }
template <> __device__ __forceinline__ float rexp_fast<float>(float x) {
  return __builtin_amdgcn_exp2f(x * 1.4426950408889634f);
}
template <typename R> __device__ __forceinline__ R rrcp(R x);
template <> __device__ __forceinline__ double rrcp<double>(double x) { return 1.0 / x; }
template <> __device__ __forceinline__ float rrcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

template <typename R> struct cplx {
  R re, im;
};
template <typename R> __device__ __forceinline__ cplx<R> cm(R a, R b) { return {a, b}; }
template <typename R> __device__ __forceinline__ cplx<R> mul(cplx<R> a, cplx<R> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __device__ __forceinline__ cplx<R> addr(cplx<R> a, R s) { return {a.re + s, a.im}; }
template <typename R> __device__ __forceinline__ cplx<R> rsub(R s, cplx<R> a) { return {s - a.re, -a.im}; }
template <typename R> __device__ __forceinline__ cplx<R> scl(cplx<R> a, R s) { return {a.re * s, a.im * s}; }
template <typename R> __device__ __forceinline__ cplx<R> cdiv(cplx<R> a, cplx<R> b) {
  R d = rrcp<R>(b.re * b.re + b.im * b.im);
  return {(a.re * b.re + a.im * b.im) * d, (a.im * b.re - a.re * b.im) * d};
}

// c - u p and t p + c: in float with the products fused (four
// multiply-adds each); in double as a product then a sum, which keeps the
// double kernels within their registers.
template <typename R> __device__ __forceinline__ cplx<R> c_minus_up(R c, cplx<R> u, cplx<R> p) {
  return rsub(c, mul(u, p));
}
template <> __device__ __forceinline__ cplx<float> c_minus_up<float>(float c, cplx<float> u, cplx<float> p) {
  return {fmaf(-u.re, p.re, fmaf(u.im, p.im, c)), fmaf(-u.re, p.im, -(u.im * p.re))};
}
template <typename R> __device__ __forceinline__ cplx<R> tp_plus_c(cplx<R> t, cplx<R> p, R c) {
  return addr(mul(t, p), c);
}
template <> __device__ __forceinline__ cplx<float> tp_plus_c<float>(cplx<float> t, cplx<float> p, float c) {
  return {fmaf(t.re, p.re, fmaf(-t.im, p.im, c)), fmaf(t.re, p.im, t.im * p.re)};
}
// sqrt(1 - s^2) of the fourth region: the float path takes the hardware
// square root (v_sqrt_f32, 1 ulp) on an argument in [0, 1].
template <typename R> __device__ __forceinline__ R rsqrt_unit(R x) { return rsqrt_<R>(x); }
template <> __device__ __forceinline__ float rsqrt_unit<float>(float x) { return __builtin_amdgcn_sqrtf(x); }

// Humlicek (1982) W4 rational approximation of w(z), z = v + i a, in its four
// regions; H = Re w / sqrt(pi), F = Im w / sqrt(pi).  Region 4 takes
// cos(Im u) as +sqrt(1 - sin^2(Im u)), the form the reference evaluates
// (cmlib/profiles.cc:7-37).  The region tests are uniform per lane.
template <typename R> __device__ __forceinline__ void voigt_raw(R v, R a, R &H, R &F);
template <typename R> __device__ __forceinline__ void voigt(R v, R a, R &H, R &F) {
  voigt_raw<R>(v, a, H, F);
  H *= R(1.0 / kSqrtPi);
  F *= R(1.0 / kSqrtPi);
}
// w(z) itself: H = Re w, F = Im w without the 1/sqrt(pi).
template <typename R> __device__ __forceinline__ void voigt_raw(R v, R a, R &H, R &F) {
  const R av = v < R(0) ? -v : v;
  const R s = av + a;
  const cplx<R> t = cm<R>(a, -v);
  const cplx<R> u = cm<R>(a * a - v * v, R(-2) * a * v);
  cplx<R> w;
  if (s >= R(15)) {
    w = cdiv(scl(t, R(0.5641896)), addr(u, R(0.5)));
  } else if (s > R(5.5)) {
    cplx<R> num = mul(t, addr(scl(u, R(0.5641896)), R(1.410474)));
    cplx<R> den = addr(mul(u, addr(u, R(3.0))), R(0.75));
    w = cdiv(num, den);
  } else if (a > (R(0.195) * av - R(0.176))) {
    cplx<R> num = addr(scl(t, R(0.5642236)), R(3.778987));
    num = tp_plus_c(t, num, R(11.96482));
    num = tp_plus_c(t, num, R(20.20933));
    num = tp_plus_c(t, num, R(16.4955));
    cplx<R> den = addr(t, R(6.699398));
    den = tp_plus_c(t, den, R(21.69274));
    den = tp_plus_c(t, den, R(39.27121));
    den = tp_plus_c(t, den, R(38.82363));
    den = tp_plus_c(t, den, R(16.4955));
    w = cdiv(num, den);
  } else {
    cplx<R> p = rsub(R(1.320522), scl(u, R(0.56419)));
    p = c_minus_up(R(35.76683), u, p);
    p = c_minus_up(R(219.0313), u, p);
    p = c_minus_up(R(1540.787), u, p);
    p = c_minus_up(R(3321.9905), u, p);
    p = c_minus_up(R(36183.31), u, p);
    cplx<R> w4 = mul(t, p);
    cplx<R> q = rsub(R(1.841439), u);
    q = c_minus_up(R(61.57037), u, q);
    q = c_minus_up(R(364.2191), u, q);
    q = c_minus_up(R(2186.181), u, q);
    q = c_minus_up(R(9022.228), u, q);
    q = c_minus_up(R(24322.84), u, q);
    q = c_minus_up(R(32066.6), u, q);
    const R su = rsin_fast<R>(u.im);
    const R e = rexp_fast<R>(u.re);
    cplx<R> r = cdiv(w4, q);
    w = cm<R>(e * rsqrt_unit<R>(R(1) - su * su) - r.re, e * su - r.im);
  }
  H = w.re;
  F = w.im;
}

// Linear short-characteristics weights with the Taylor branch below
// d tau = 1e-2 (plpar/atmos_ppbez_rts.cc:1067-1096).
template <typename R> __device__ __forceinline__ void sc_weights(R tb, R &etb, R &w0, R &w1) {
  if (tb < R(1E-2)) {
    etb = R(1) - tb + tb * tb * R(0.5) - tb * tb * tb * R(1.0 / 6.0);
    w0 = tb * (R(0.5) - tb * (R(1.0 / 3.0) - tb * R(0.125)));
    w1 = R(0.5) * tb * (R(1) - tb * R(1.0 / 3.0) * (R(1) - tb * R(0.25)));
  } else {
    etb = rexp<R>(-tb);
    const R itb = rrcp<R>(tb);
    w0 = itb - etb * (R(1) + itb);
    w1 = R(1) - itb + etb * itb;
  }
}

// Fritsch-Butland monotone derivative at an interior point
// (atmos_ppbez_rts.cc:966-979).
template <typename R> __device__ __forceinline__ R fb_deriv(R om, R o0, R op, R h1, R h2) {
  const R d2 = (op - o0) * rrcp<R>(h2);
  const R d1 = (o0 - om) * rrcp<R>(h1);
  if (d1 * d2 > R(0)) {
    const R al = (R(1) + h1 * rrcp<R>(h1 + h2)) * R(1.0 / 3.0);
    return d1 * d2 * rrcp<R>(al * d1 + (R(1) - al) * d2);
  }
  return R(0);
}

// Reduced propagation matrix K/eta_I - 1 of the Stokes vector equation:
//   [ 0   a   b   c ]   a,b,c = eta_Q,U,V / eta_I
//   [ a   0   d  -e ]   d = rho_V / eta_I  (reference K[2][3])
//   [ b  -d   0   f ]   e = rho_U / eta_I  (reference K[4][2])
//   [ c   e  -f   0 ]   f = rho_Q / eta_I  (reference K[3][4])
template <typename R> struct Kred {
  R a, b, c, d, e, f;
};

// y = Kred * x
template <typename R> __device__ __forceinline__ void kmul(const Kred<R> &k, const R x[4], R y[4]) {
  y[0] = k.a * x[1] + k.b * x[2] + k.c * x[3];
  y[1] = k.a * x[0] + k.d * x[2] - k.e * x[3];
  y[2] = k.b * x[0] - k.d * x[1] + k.f * x[3];
  y[3] = k.c * x[0] + k.e * x[1] - k.f * x[2];
}

// Pivot-free Crout factorisation of 1 + w1 Kred (cmlib/mathtools.cc:487-514:
// unit upper triangle, diagonal kept in L) and the solve with it; the factor
// is shared by the forward solve and the tangent solves of one depth.
template <typename R> struct LU4 {
  R D[4][4];
  R id[4];  // reciprocals of the diagonal of L
};
template <typename R> __device__ __forceinline__ void factor_1pk(const Kred<R> &k, R w1, LU4<R> &f) {
  R A[4][4];
  A[0][0] = R(1);       A[0][1] = w1 * k.a;   A[0][2] = w1 * k.b;   A[0][3] = w1 * k.c;
  A[1][0] = w1 * k.a;   A[1][1] = R(1);       A[1][2] = w1 * k.d;   A[1][3] = -w1 * k.e;
  A[2][0] = w1 * k.b;   A[2][1] = -w1 * k.d;  A[2][2] = R(1);       A[2][3] = w1 * k.f;
  A[3][0] = w1 * k.c;   A[3][1] = w1 * k.e;   A[3][2] = -w1 * k.f;  A[3][3] = R(1);
  R (&D)[4][4] = f.D;
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
    for (int i = kk; i < 4; ++i) {
      R sum = R(0);
#pragma unroll
      for (int p = 0; p < kk; ++p) sum += D[i][p] * D[p][kk];
      D[i][kk] = A[i][kk] - sum;
    }
    const R inv = rrcp<R>(D[kk][kk]);
    f.id[kk] = inv;
#pragma unroll
    for (int j = kk + 1; j < 4; ++j) {
      R sum = R(0);
#pragma unroll
      for (int p = 0; p < kk; ++p) sum += D[kk][p] * D[p][j];
      D[kk][j] = (A[kk][j] - sum) * inv;
    }
  }
}
template <typename R> __device__ __forceinline__ void solve_lu(const LU4<R> &f, const R b[4], R x[4]) {
  const R (&D)[4][4] = f.D;
  R y[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    R sum = R(0);
#pragma unroll
    for (int p = 0; p < i; ++p) sum += D[i][p] * y[p];
    y[i] = (b[i] - sum) * f.id[i];
  }
#pragma unroll
  for (int i = 3; i >= 0; --i) {
    R sum = R(0);
#pragma unroll
    for (int p = i + 1; p < 4; ++p) sum += D[i][p] * x[p];
    x[i] = y[i] - sum;
  }
}
// Solve (1 + w1 Kred) x = b.
template <typename R> __device__ __forceinline__ void solve_1pk(const Kred<R> &k, R w1, const R b[4], R x[4]) {
  LU4<R> f;
  factor_1pk<R>(k, w1, f);
  solve_lu<R>(f, b, x);
}

// Line constants passed by value as a kernel argument (a few KB).
struct KLine {
  double lam0;      // vacuum line centre [cm]
  double opf_l;     // B_lu h c / (4 pi lam0)
  double opf_u;     // B_ul h c / (4 pi lam0)
  double emf;       // A_ul h c / (4 pi lam0)
  double sf;        // lam0^2 / (4 pi c)
  double grad;      // radiative damping
  double sigma;     // van der Waals cross-section
  double alpha;     // its velocity exponent
  double twok_m;    // 2 k / mass
  double vmfac;     // 8 k / (pi mu), mu = 1.66e-24 / (1/1.008 + 1.66e-24/mass)
  double lbfac;     // 4.67e-5 lam0^2  (lambda_B per Gauss)
  int split;
  int nc[3];
  int c0;
};
struct KLines {
  int nline;
  int ncomp;
  KLine l[SNAPI_MAX_LINES];
  double shift[SNAPI_MAX_COMPONENTS];
  double str[SNAPI_MAX_COMPONENTS];
};

}  // namespace snapi

namespace snapi {

// W4 values and the derivative of each region's expression with respect to
// v (the analytic derivative of the approximant, as cmlib/profiles.cc:39-89
// evaluates it); dH/da = -dF/dv and dF/da = dH/dv since w is analytic.
template <typename R> __device__ __forceinline__ void voigt_d(R v, R a, R &H, R &F, R &dH, R &dF) {
  const R av = v < R(0) ? -v : v;
  const R s = av + a;
  const cplx<R> t = cm<R>(a, -v), dt = cm<R>(R(0), R(-1));
  const cplx<R> u = cm<R>(a * a - v * v, R(-2) * a * v), du = cm<R>(R(-2) * v, R(-2) * a);
  cplx<R> w, dw;
  if (s >= R(15)) {
    const cplx<R> den = addr(u, R(0.5));
    const cplx<R> num = scl(t, R(0.5641896));
    w = cdiv(num, den);
    const cplx<R> a1 = cdiv(scl(dt, R(0.5641896)), den);
    const cplx<R> a2 = mul(cdiv(w, den), du);
    dw = cm<R>(a1.re - a2.re, a1.im - a2.im);
  } else if (s > R(5.5)) {
    const cplx<R> num = addr(scl(u, R(0.5641896)), R(1.410474));
    const cplx<R> den = addr(mul(u, addr(u, R(3.0))), R(0.75));
    const cplx<R> tdu = mul(t, du);
    w = cdiv(mul(t, num), den);
    const cplx<R> a1 = cdiv(mul(dt, num), den);
    const cplx<R> a2 = cdiv(scl(tdu, R(0.5641896)), den);
    const cplx<R> a3 = mul(cdiv(cdiv(mul(tdu, num), den), den), addr(scl(u, R(2)), R(3.0)));
    dw = cm<R>(a1.re + a2.re - a3.re, a1.im + a2.im - a3.im);
  } else if (a > (R(0.195) * av - R(0.176))) {
    cplx<R> up = addr(scl(t, R(0.5642236)), R(3.778987));
    up = tp_plus_c(t, up, R(11.96482));
    up = tp_plus_c(t, up, R(20.20933));
    up = tp_plus_c(t, up, R(16.4955));
    cplx<R> lo = addr(t, R(6.699398));
    lo = tp_plus_c(t, lo, R(21.69274));
    lo = tp_plus_c(t, lo, R(39.27121));
    lo = tp_plus_c(t, lo, R(38.82363));
    lo = tp_plus_c(t, lo, R(16.4955));
    w = cdiv(up, lo);
    // d(up)/dt and d(lo)/dt in Horner form
    cplx<R> dup = addr(scl(t, R(0.5642236 * 4.0)), R(3.778987 * 3.0));
    dup = tp_plus_c(t, dup, R(11.96482 * 2.0));
    dup = tp_plus_c(t, dup, R(20.20933));
    cplx<R> dlo = addr(scl(t, R(5.0)), R(6.699398 * 4.0));
    dlo = tp_plus_c(t, dlo, R(21.69274 * 3.0));
    dlo = tp_plus_c(t, dlo, R(39.27121 * 2.0));
    dlo = tp_plus_c(t, dlo, R(38.82363));
    const cplx<R> nn = mul(dup, lo), mm = mul(dlo, up);
    dw = mul(dt, cdiv(cdiv(cm<R>(nn.re - mm.re, nn.im - mm.im), lo), lo));
  } else {
    cplx<R> P = rsub(R(1.320522), scl(u, R(0.56419)));
    P = c_minus_up(R(35.76683), u, P);
    P = c_minus_up(R(219.0313), u, P);
    P = c_minus_up(R(1540.787), u, P);
    P = c_minus_up(R(3321.9905), u, P);
    P = c_minus_up(R(36183.31), u, P);
    const cplx<R> w4 = mul(t, P);
    cplx<R> dP = rsub(R(1.320522 * 5.0), scl(u, R(0.56419 * 6.0)));
    dP = c_minus_up(R(35.76683 * 4.0), u, dP);
    dP = c_minus_up(R(219.0313 * 3.0), u, dP);
    dP = c_minus_up(R(1540.787 * 2.0), u, dP);
    dP = c_minus_up(R(3321.9905), u, dP);
    const cplx<R> g1 = mul(dt, P), g2 = mul(mul(t, du), dP);
    const cplx<R> dw4 = cm<R>(g1.re - g2.re, g1.im - g2.im);
    cplx<R> Q = rsub(R(1.841439), u);
    Q = c_minus_up(R(61.57037), u, Q);
    Q = c_minus_up(R(364.2191), u, Q);
    Q = c_minus_up(R(2186.181), u, Q);
    Q = c_minus_up(R(9022.228), u, Q);
    Q = c_minus_up(R(24322.84), u, Q);
    Q = c_minus_up(R(32066.6), u, Q);
    cplx<R> dQ = rsub(R(1.841439 * 6.0), scl(u, R(7.0)));
    dQ = c_minus_up(R(61.57037 * 5.0), u, dQ);
    dQ = c_minus_up(R(364.2191 * 4.0), u, dQ);
    dQ = c_minus_up(R(2186.181 * 3.0), u, dQ);
    dQ = c_minus_up(R(9022.228 * 2.0), u, dQ);
    dQ = c_minus_up(R(24322.84), u, dQ);
    dQ = scl(mul(du, dQ), R(-1));
    const R su = rsin_fast<R>(u.im);
    const R co = rsqrt_unit<R>(R(1) - su * su);
    const R dsu = rcos_fast<R>(u.im) * du.im;
    const R e = rexp_fast<R>(u.re);
    const cplx<R> r = cdiv(w4, Q);
    w = cm<R>(e * co - r.re, e * su - r.im);
    const cplx<R> n1 = mul(dw4, Q), n2 = mul(dQ, w4);
    // (n1 - n2) / Q^2 as two divisions: |Q|^4 overflows float for |u| ~ 30
    const cplx<R> q2 = cdiv(cdiv(cm<R>(n1.re - n2.re, n1.im - n2.im), Q), Q);
    dw = cm<R>(e * du.re * co + e * (-su * du.im) - q2.re, e * du.re * su + e * dsu - q2.im);
  }
  H = w.re * R(1.0 / kSqrtPi);
  F = w.im * R(1.0 / kSqrtPi);
  dH = dw.re * R(1.0 / kSqrtPi);
  dF = dw.im * R(1.0 / kSqrtPi);
}

}  // namespace snapi
