// Host-side runtime of the library: argument checks, conversion of the
// public line table into kernel constants, and the per-device status word.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "rt_common.h"
#include "synth_kernel.h"

namespace snapi {

static constexpr int kMaxDevices = 64;
static int *g_status[kMaxDevices];
static std::mutex g_mu;

int *device_status_slot() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_status[dev]) {
    if (hipMalloc((void **)&g_status[dev], sizeof(int)) != hipSuccess) return nullptr;
    if (hipMemset(g_status[dev], 0, sizeof(int)) != hipSuccess) return nullptr;
  }
  return g_status[dev];
}

void reset_status() {
  int *s = device_status_slot();
  // completed before returning, so work on any stream sees the cleared flag
  if (s && hipMemset(s, 0, sizeof(int)) == hipSuccess) (void)hipStreamSynchronize(nullptr);
}

int read_status() {
  int *s = device_status_slot();
  if (!s) return SNAPI_ERR_DEVICE;
  int h = 0;
  if (hipMemcpy(&h, s, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return SNAPI_ERR_DEVICE;
  if (h & 2) return SNAPI_ERR_NONFINITE;
  if (h & 1) return SNAPI_ERR_UNSUPPORTED;
  return SNAPI_OK;
}

// Line constants: the factors of atom_rad_synthesis.cc:403-433 and the
// Doppler/damping prefactors of atom_rad.cc:879-883 and atom.cc:939-960.
static void convert_lines(const snapi_line_table *t, KLines &kl) {
  kl.nline = t->nline;
  kl.ncomp = t->ncomp_total;
  for (int i = 0; i < t->nline; ++i) {
    const snapi_line &s = t->line[i];
    KLine &k = kl.l[i];
    const double hc4 = kH * kC / s.lam0 / 4.0 / kPi;
    k.lam0 = s.lam0;
    k.opf_l = s.blu * hc4;
    k.opf_u = s.bul * hc4;
    k.emf = s.aul * hc4;
    k.sf = s.lam0 * s.lam0 / (4.0 * kPi * kC);
    k.grad = s.grad;
    k.sigma = s.sigma;
    k.alpha = s.alpha;
    k.twok_m = 2.0 * kK / s.mass;
    const double red = 1.66E-24 / (1.0 / 1.008 + 1.66E-24 / s.mass);
    k.vmfac = 8.0 * kK / kPi / red;
    k.lbfac = 4.67E-5 * s.lam0 * s.lam0;
    k.split = s.split;
    for (int g = 0; g < 3; ++g) k.nc[g] = s.split ? s.ncomp[g] : 0;
    k.c0 = s.comp0;
  }
  for (int c = 0; c < t->ncomp_total; ++c) {
    kl.shift[c] = t->comp[c].shift;
    kl.str[c] = t->comp[c].strength;
  }
}

int prepare_synth(const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd, int nlambda, KLines &kl,
                  SynthArgs &a) {
  if (!lines || !ray) return SNAPI_ERR_ARG;
  if (ncol <= 0 || nd < 2 || nd > SNAPI_MAX_ND || nlambda <= 0) return SNAPI_ERR_ARG;
  if (lines->nline < 0 || lines->nline > SNAPI_MAX_LINES) return SNAPI_ERR_ARG;
  if (lines->ncomp_total < 0 || lines->ncomp_total > SNAPI_MAX_COMPONENTS) return SNAPI_ERR_ARG;
  for (int i = 0; i < lines->nline; ++i) {
    const snapi_line &s = lines->line[i];
    if (!(s.lam0 > 0) || !(s.mass > 0)) return SNAPI_ERR_ARG;
    if (s.split) {
      int n = s.ncomp[0] + s.ncomp[1] + s.ncomp[2];
      if (s.ncomp[0] < 0 || s.ncomp[1] < 0 || s.ncomp[2] < 0 || s.comp0 < 0 ||
          s.comp0 + n > lines->ncomp_total)
        return SNAPI_ERR_ARG;
    }
  }
  // the column tile must fit in LDS (64 KiB per workgroup)
  if (ColumnTile<double>::bytes(nd, lines->nline) > 65536) return SNAPI_ERR_UNSUPPORTED;
  const double ct = cos(ray->theta);
  if (ct == 0.0) return SNAPI_ERR_ARG;
  convert_lines(lines, kl);
  a.ncol = ncol;
  a.nd = nd;
  a.nlambda = nlambda;
  a.nfield = SNAPI_NF_BASE + 2 * lines->nline;
  a.ct = ct;
  a.tau_grid = ray->tau_grid ? 1 : 0;
  a.hminus = ray->hminus ? 1 : 0;
  a.hminus_mid = ray->hminus_mid ? 1 : 0;
  a.lam_mid = ray->lam_mid;
  return SNAPI_OK;
}

}  // namespace snapi

extern "C" int snapi_version(void) { return 100; }

extern "C" int snapi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
