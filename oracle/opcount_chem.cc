/*
 * Operation count of the oracle's chemical equilibrium and hydrostatic
 * march -- MEASUREMENT INFRASTRUCTURE ONLY (the algorithmic flop count
 * behind the rooflines of chemeq_kernel and hequilibrium_kernel).  Nothing in
 * the product links or calls this file.
 *
 * snapi_oracle_chem.c (the FP64 restatement of atmosphere::chemeq,
 * atmos/atmos_chm.cc:28-363, with the partition functions of
 * atmos/atmol/atom/partf.cc, atom::ionfrc, the dissociation constants, and
 * atmos_pp::enforce_hequilibrium, atmos/plpar/atmos_pp.cc:372-455) is
 * compiled here with every `double` replaced by `Flop`, the counting scalar
 * of oracle/opcount.cc: add, sub, mul, div, sqrt, exp, log, log10 = 1;
 * pow = 3; fabs, negation, comparisons, conversions = 0.  Within one column
 * an operation is counted the first time it meets a given (operation,
 * operands) combination (value numbering), so the temperature-only parts of
 * the equilibria at one depth count once, as the kernels evaluate them.
 *
 * Input (binary, from oracle/opcount.py): int32 mode (0: chemical
 * equilibrium of every point, 1: hydrostatic march of every column), ncol,
 * nd; the snapi_chem_table's bytes; f64 T[ncol][nd]; mode 0: f64
 * Nt[ncol][nd], mode 1: f64 logtau[ncol][nd].
 * Output (text, one line per column): column, counted operations, raw
 * operations, equilibria solved (mode 1: the march's iterations plus the
 * final pass), then Nt (mode 1) or Ne (mode 0) of every depth as
 * hexadecimal doubles, for the check against liboracle.
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


// the header's structures keep their double fields (the table is read as
// bytes); the restatement's own doubles count
#include "snapi_oracle.h"
inline bool isfinite(Flop a) { return std::isfinite(a.v); }
inline Flop pow(double a, Flop b) { return pow(Flop(a), b); }
inline Flop pow(Flop a, double b) { return pow(a, Flop(b)); }
inline Flop pow(int a, Flop b) { return pow(Flop((double)a), b); }
inline Flop pow(Flop a, int b) { return pow(a, Flop((double)b)); }
#define double Flop
#include "snapi_oracle_chem.c"
#undef double

static void rd(FILE *f, void *p, size_t n) {
  if (fread(p, 1, n, f) != n) {
    fprintf(stderr, "opcount_chem: short input\n");
    exit(2);
  }
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: opcount_chem input.bin\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  int32_t h[3];
  rd(f, h, sizeof h);
  const int mode = h[0], ncol = h[1], nd = h[2];
  if ((mode != 0 && mode != 1) || ncol < 1 || nd < 1) {
    fprintf(stderr, "opcount_chem: bad header\n");
    return 2;
  }
  static snapi_chem_table tb;
  rd(f, &tb, sizeof tb);
  std::vector<double> T((size_t)ncol * nd), X((size_t)ncol * nd);
  rd(f, T.data(), T.size() * 8);
  rd(f, X.data(), X.size() * 8);
  fclose(f);
  std::vector<Flop> Tf(nd), Xf(nd), o1(nd), o2(nd), o3(nd), o4(nd);
  std::vector<int> it(nd);
  for (int c = 0; c < ncol; ++c) {
    opc::seen.reset(22);
    opc::counted = opc::raw = 0;
    for (int z = 0; z < nd; ++z) {
      Tf[z] = T[(size_t)c * nd + z];
      Xf[z] = X[(size_t)c * nd + z];
    }
    long neq = 0;
    if (mode == 0) {
      for (int z = 0; z < nd; ++z) {
        Flop nh;
        chemeq_point(&tb, Tf[z], Xf[z], &o1[z], NULL, NULL, NULL, &o2[z], &nh);
      }
      neq = nd;
    } else {
      orc_hequilibrium(&tb, nd, Tf.data(), Xf.data(), o1.data(), o2.data(), o3.data(), o4.data(), it.data());
      for (int z = 0; z < nd; ++z) neq += it[z] + 1;
    }
    printf("%d %llu %llu %ld", c, (unsigned long long)opc::counted, (unsigned long long)opc::raw, neq);
    for (int z = 0; z < nd; ++z) printf(" %a", o1[z].v);
    printf("\n");
  }
  return 0;
}
