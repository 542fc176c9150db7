/*
 * Operation count of the oracle's polarized synthesis -- MEASUREMENT
 * INFRASTRUCTURE ONLY (the algorithmic flop count behind the synthesis
 * roofline).  Nothing in the product links or calls this file.
 *
 * snapi_oracle.c (the FP64 restatement of atmosphere::obs_stokes:
 * op_em_vector atmos/atmos_rad_synthesis.cc:20-72, the line and H-
 * opacities, atmos_ppbez::formal plpar/atmos_ppbez_rts.cc:918-1218) is
 * compiled here with every `double` replaced by `Flop`, a scalar that
 * performs the same FP64 arithmetic and counts it:
 *   add, sub, mul, div, sqrt, exp, log, log10, sin, cos = 1; pow = 3
 *   (log, mul, exp); fabs, negation, comparisons, conversions = 0.
 * The C code has no fused operations (it is compiled with
 * -ffp-contract=off), so a * b + c counts 2, as a fused multiply-add does.
 *
 * The restatement recomputes, for every wavelength, the quantities that do
 * not depend on it (Doppler widths, damping, level factors, the field
 * geometry), and for every depth the ones that depend on the wavelength
 * only (the H- polynomials in lambda).  An implementation evaluates each
 * of them once.  The count therefore numbers values: within one column an
 * operation is counted the first time it meets a given (operation,
 * operands) combination, and repeats of it are not (value numbering, as
 * common-subexpression elimination would leave them).
 *
 * Input (binary, little-endian, from oracle/opcount.py):
 *   int32 nline, ncomp, ncol, nd, nf, nlambda, tau_grid, hminus, hminus_mid
 *   f64 theta, lam_mid
 *   per line: f64 lam0 aul blu bul grad sigma alpha mass; int32 split
 *             ncomp[3] comp0
 *   per component: f64 shift strength
 *   f64 state[ncol][nf][nd], f64 lam[nlambda]
 * Output (text, one line per column): column, counted operations, raw
 * operations (every evaluation), then the emergent Stokes vectors of the
 * column as hexadecimal doubles (for the check against liboracle).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace opc {

// open-addressing set of (operation, operand bits) keys, cleared per column
struct Seen {
  std::vector<uint64_t> k0, k1;
  std::vector<uint32_t> k2;
  std::vector<uint8_t> used;
  size_t mask = 0, n = 0;
  void reset(size_t cap_log2) {
    const size_t cap = (size_t)1 << cap_log2;
    k0.assign(cap, 0);
    k1.assign(cap, 0);
    k2.assign(cap, 0);
    used.assign(cap, 0);
    mask = cap - 1;
    n = 0;
  }
  void grow() {
    std::vector<uint64_t> a0, a1;
    std::vector<uint32_t> a2;
    std::vector<uint8_t> u;
    a0.swap(k0);
    a1.swap(k1);
    a2.swap(k2);
    u.swap(used);
    size_t lg = 0;
    while (((size_t)1 << lg) < 2 * a0.size()) ++lg;
    reset(lg);
    for (size_t i = 0; i < a0.size(); ++i)
      if (u[i]) insert(a0[i], a1[i], a2[i]);
  }
  bool insert(uint64_t a, uint64_t b, uint32_t op) {
    if (2 * (n + 1) > mask + 1) grow();
    uint64_t h = a * 0x9E3779B97F4A7C15ull ^ (b + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full ^ op;
    h ^= h >> 29;
    for (size_t i = h & mask;; i = (i + 1) & mask) {
      if (!used[i]) {
        used[i] = 1;
        k0[i] = a;
        k1[i] = b;
        k2[i] = op;
        ++n;
        return true;
      }
      if (k0[i] == a && k1[i] == b && k2[i] == op) return false;
    }
  }
};
Seen seen;
uint64_t counted = 0, raw = 0;

inline uint64_t bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}
enum : uint32_t { ADD = 1, SUB, MUL, DIV, SQRT, EXP, LOG, LOG10, SIN, COS, POW };
inline double rec(uint32_t op, double r, double a, double b, int cost) {
  raw += cost;
  uint64_t x = bits(a), y = bits(b);
  if ((op == ADD || op == MUL) && x > y) {  // commutative: one key for both orders
    const uint64_t t = x;
    x = y;
    y = t;
  }
  if (seen.insert(x, y, op)) counted += cost;
  return r;
}

}  // namespace opc

struct Flop {
  double v;
  Flop() = default;
  constexpr Flop(double x) : v(x) {}
  explicit operator double() const { return v; }
  Flop &operator+=(Flop b);
  Flop &operator-=(Flop b);
  Flop &operator*=(Flop b);
  Flop &operator/=(Flop b);
};
inline Flop operator+(Flop a, Flop b) { return Flop(opc::rec(opc::ADD, a.v + b.v, a.v, b.v, 1)); }
inline Flop operator-(Flop a, Flop b) { return Flop(opc::rec(opc::SUB, a.v - b.v, a.v, b.v, 1)); }
inline Flop operator*(Flop a, Flop b) { return Flop(opc::rec(opc::MUL, a.v * b.v, a.v, b.v, 1)); }
inline Flop operator/(Flop a, Flop b) { return Flop(opc::rec(opc::DIV, a.v / b.v, a.v, b.v, 1)); }
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline Flop &Flop::operator+=(Flop b) { return *this = *this + b; }
inline Flop &Flop::operator-=(Flop b) { return *this = *this - b; }
inline Flop &Flop::operator*=(Flop b) { return *this = *this * b; }
inline Flop &Flop::operator/=(Flop b) { return *this = *this / b; }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
inline bool operator<=(Flop a, Flop b) { return a.v <= b.v; }
inline bool operator>=(Flop a, Flop b) { return a.v >= b.v; }
inline bool operator==(Flop a, Flop b) { return a.v == b.v; }
inline bool operator!=(Flop a, Flop b) { return a.v != b.v; }
inline Flop sqrt(Flop a) { return Flop(opc::rec(opc::SQRT, std::sqrt(a.v), a.v, 0.0, 1)); }
inline Flop exp(Flop a) { return Flop(opc::rec(opc::EXP, std::exp(a.v), a.v, 0.0, 1)); }
inline Flop log(Flop a) { return Flop(opc::rec(opc::LOG, std::log(a.v), a.v, 0.0, 1)); }
inline Flop log10(Flop a) { return Flop(opc::rec(opc::LOG10, std::log10(a.v), a.v, 0.0, 1)); }
inline Flop sin(Flop a) { return Flop(opc::rec(opc::SIN, std::sin(a.v), a.v, 0.0, 1)); }
inline Flop cos(Flop a) { return Flop(opc::rec(opc::COS, std::cos(a.v), a.v, 0.0, 1)); }
inline Flop pow(Flop a, Flop b) { return Flop(opc::rec(opc::POW, std::pow(a.v, b.v), a.v, b.v, 3)); }
inline Flop fabs(Flop a) { return Flop(std::fabs(a.v)); }

#define double Flop
#include "snapi_oracle.c"
#undef double

static void rd(FILE *f, void *p, size_t n) {
  if (fread(p, 1, n, f) != n) {
    fprintf(stderr, "opcount: short input\n");
    exit(2);
  }
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: opcount input.bin\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  int32_t h[9];
  rd(f, h, sizeof h);
  const int nline = h[0], ncomp = h[1], ncol = h[2], nd = h[3], nf = h[4], nl = h[5];
  if (nline < 0 || nline > SNAPI_MAX_LINES || ncomp < 0 || ncomp > SNAPI_MAX_COMPONENTS || ncol < 1 || nd < 2 ||
      nf != SNAPI_NF_BASE + 2 * nline || nl < 1) {
    fprintf(stderr, "opcount: bad header\n");
    return 2;
  }
  static snapi_line_table lt;
  std::memset(&lt, 0, sizeof lt);
  lt.nline = nline;
  lt.ncomp_total = ncomp;
  snapi_ray ray;
  std::memset(&ray, 0, sizeof ray);
  double d2[2];
  rd(f, d2, sizeof d2);
  ray.theta = d2[0];
  ray.lam_mid = d2[1];
  ray.tau_grid = h[6];
  ray.hminus = h[7];
  ray.hminus_mid = h[8];
  for (int i = 0; i < nline; ++i) {
    double d[8];
    int32_t n[5];
    rd(f, d, sizeof d);
    rd(f, n, sizeof n);
    snapi_line &L = lt.line[i];
    L.lam0 = d[0];
    L.aul = d[1];
    L.blu = d[2];
    L.bul = d[3];
    L.grad = d[4];
    L.sigma = d[5];
    L.alpha = d[6];
    L.mass = d[7];
    L.split = n[0];
    L.ncomp[0] = n[1];
    L.ncomp[1] = n[2];
    L.ncomp[2] = n[3];
    L.comp0 = n[4];
  }
  for (int c = 0; c < ncomp; ++c) {
    double d[2];
    rd(f, d, sizeof d);
    lt.comp[c].shift = d[0];
    lt.comp[c].strength = d[1];
  }
  std::vector<double> st((size_t)ncol * nf * nd), lam(nl);
  rd(f, st.data(), st.size() * 8);
  rd(f, lam.data(), lam.size() * 8);
  fclose(f);
  std::vector<Flop> stf(st.begin(), st.end()), lamf(lam.begin(), lam.end()), out((size_t)nl * 4);
  for (int c = 0; c < ncol; ++c) {
    opc::seen.reset(22);
    opc::counted = opc::raw = 0;
    orc_synth(&lt, &ray, 1, nd, stf.data() + (size_t)c * nf * nd, nl, lamf.data(), out.data(), 1);
    printf("%d %llu %llu", c, (unsigned long long)opc::counted, (unsigned long long)opc::raw);
    for (size_t i = 0; i < out.size(); ++i) printf(" %a", out[i].v);
    printf("\n");
  }
  return 0;
}
