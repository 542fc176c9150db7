/*
 * Operation count of the oracle's ALI accumulation pass -- MEASUREMENT
 * INFRASTRUCTURE ONLY (the algorithmic flop count behind the roofline of
 * ali_acc_kernel).  Nothing in the product links or calls this file.
 *
 * snapi_oracle_ali.c's orc_ali_accumulate (the FP64 restatement of the
 * formal solutions and atom::add sums of atmosphere::nltepops,
 * atmos/atmos_pop.cc:133-195, atom.cc:1865-1927, atom_rad.cc:312-537) is
 * compiled here, with the synthesis restatement it calls (snapi_oracle.c:
 * the Voigt function, the scalar formal solution, the Planck function),
 * with every `double` replaced by `Flop`, the counting scalar of
 * oracle/opcount.cc (the atom's table is read as the bytes of its doubles): add, sub, mul, div, sqrt, exp, log, log10 = 1;
 * pow = 3; fabs, negation, comparisons, conversions = 0.  Within one column
 * an operation is counted the first time it meets a given (operation,
 * operands) combination (value numbering), so the quantities the restatement
 * recomputes per ray or wavelength but that do not depend on them count once.
 *
 * Input (binary, from oracle/opcount.py): int32 nd, nlam, nray, bg_per_ray,
 * tau_grid, nbg (1 or nray); f64 boundary; the snapi_ali_atom's bytes; f64
 * lam, lamw, fudge [nlam], sigma [ntrans][nlam], theta, wray [nray]; then
 * per column: colstate [SNAPI_A_NF][nd], vproj [nray][nd], gamma
 * [ntrans][nd], bgop, bgem [nbg][nd][nlam], pops [nd][nlevel]; int32 ncol
 * first of all.
 * Output (text, one line per column): column, counted operations, raw
 * operations, then the sums acc [nd][ntrans][4] as hexadecimal doubles.
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


inline bool isfinite(Flop a) { return std::isfinite(a.v); }
inline Flop pow(double a, Flop b) { return pow(Flop(a), b); }
inline Flop pow(Flop a, double b) { return pow(a, Flop(b)); }
inline Flop pow(int a, Flop b) { return pow(Flop((double)a), b); }
inline Flop pow(Flop a, int b) { return pow(a, Flop((double)b)); }
// the whole restatement in Flop: the synthesis file for the Voigt function,
// the scalar formal solution and the Planck function the ALI pass calls
#define double Flop
#include "snapi_oracle.c"
#undef PI_D
#include "snapi_oracle_ali.c"
#undef double

static void rd(FILE *f, void *p, size_t n) {
  if (fread(p, 1, n, f) != n) {
    fprintf(stderr, "opcount_ali: short input\n");
    exit(2);
  }
}
static std::vector<double> rdv(FILE *f, size_t n) {
  std::vector<double> v(n);
  rd(f, v.data(), n * 8);
  return v;
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: opcount_ali input.bin\n");
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  int32_t h[7];
  rd(f, h, sizeof h);
  const int ncol = h[0], nd = h[1], nlam = h[2], nray = h[3], nbg = h[6];
  double boundary;
  rd(f, &boundary, 8);
  static snapi_ali_atom at;
  rd(f, &at, sizeof at);
  if (ncol < 1 || nd < 2 || nlam < 1 || nray < 1 || (nbg != 1 && nbg != nray) || at.ntrans < 1 ||
      at.ntrans > SNAPI_ALI_MAX_TRANS || at.nlevel < 1 || at.nlevel > SNAPI_ALI_MAX_LEVELS) {
    fprintf(stderr, "opcount_ali: bad header\n");
    return 2;
  }
  const int nt = at.ntrans, nl = at.nlevel;
  std::vector<double> lam = rdv(f, nlam), lamw = rdv(f, nlam), fudge = rdv(f, nlam), sigma = rdv(f, (size_t)nt * nlam),
                      theta = rdv(f, nray), wray = rdv(f, nray);
  std::vector<Flop> lamf(lam.begin(), lam.end()), lamwf(lamw.begin(), lamw.end()), fudgef(fudge.begin(), fudge.end()),
      sigmaf(sigma.begin(), sigma.end()), thetaf(theta.begin(), theta.end()), wrayf(wray.begin(), wray.end());
  snapi_ali_grid gf;
  gf.nlam = nlam;
  gf.lam = lamf.data();
  gf.lamw = lamwf.data();
  gf.fudge = fudgef.data();
  gf.sigma = sigmaf.data();
  gf.nray = nray;
  gf.theta = thetaf.data();
  gf.wray = wrayf.data();
  gf.bg_per_ray = h[4];
  gf.boundary = Flop(boundary);
  gf.tau_grid = h[5];
  static_assert(sizeof(Flop) == sizeof(double), "Flop must be a plain double");
  for (int c = 0; c < ncol; ++c) {
    std::vector<double> cs = rdv(f, (size_t)SNAPI_A_NF * nd), vp = rdv(f, (size_t)nray * nd),
                        ga = rdv(f, (size_t)nt * nd), bo = rdv(f, (size_t)nbg * nd * nlam),
                        be = rdv(f, (size_t)nbg * nd * nlam), po = rdv(f, (size_t)nd * nl);
    std::vector<Flop> csf(cs.begin(), cs.end()), vpf(vp.begin(), vp.end()), gaf(ga.begin(), ga.end()),
        bof(bo.begin(), bo.end()), bef(be.begin(), be.end()), pof(po.begin(), po.end()), acc((size_t)nd * nt * 4);
    opc::seen.reset(22);
    opc::counted = opc::raw = 0;
    orc_ali_accumulate(&at, &gf, nd, csf.data(), vpf.data(), gaf.data(),
                       bof.data(), bef.data(), pof.data(), acc.data());
    printf("%d %llu %llu", c, (unsigned long long)opc::counted, (unsigned long long)opc::raw);
    for (size_t i = 0; i < acc.size(); ++i) printf(" %a", acc[i].v);
    printf("\n");
  }
  fclose(f);
  return 0;
}
