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

}  // namespace snapi

// A caller-owned status word on one device (snapi_context_create).  The
// calling thread's bound context, when it lives on the current device,
// receives the status bits of every launch that thread makes.
struct snapi_context {
  int device;
  int *word;
};

namespace snapi {

static thread_local snapi_context *t_ctx = nullptr;

static int current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
  return dev;
}

int *device_status_slot() {
  const int dev = current_device();
  if (dev < 0) return nullptr;
  if (t_ctx && t_ctx->device == dev) return t_ctx->word;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_status[dev]) {
    int *w = nullptr;
    if (hipMalloc((void **)&w, sizeof(int)) != hipSuccess) return nullptr;
    if (hipMemset(w, 0, sizeof(int)) != hipSuccess) {
      (void)hipFree(w);
      return nullptr;
    }
    g_status[dev] = w;
  }
  return g_status[dev];
}

static int status_of(int h) {
  if (h & 2) return SNAPI_ERR_NONFINITE;
  if (h & 1) return SNAPI_ERR_UNSUPPORTED;
  return SNAPI_OK;
}

// Reads (and, with reset, clears) a status word after the work on `stream`.
static int take_status(int *w, hipStream_t stream, int reset) {
  if (!w) return SNAPI_ERR_DEVICE;
  if (hipStreamSynchronize(stream) != hipSuccess) return SNAPI_ERR_DEVICE;
  int h = 0;
  if (hipMemcpy(&h, w, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return SNAPI_ERR_DEVICE;
  if (reset && (hipMemset(w, 0, sizeof(int)) != hipSuccess || hipDeviceSynchronize() != hipSuccess))
    return SNAPI_ERR_DEVICE;
  return status_of(h);
}

int reset_status() {
  int *s = device_status_slot();
  if (!s) return SNAPI_ERR_DEVICE;
  // completed before returning, so work on any stream sees the cleared flag
  if (hipMemset(s, 0, sizeof(int)) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess)
    return SNAPI_ERR_DEVICE;
  return SNAPI_OK;
}

int read_status() {
  int *s = device_status_slot();
  if (!s) return SNAPI_ERR_DEVICE;
  int h = 0;
  if (hipMemcpy(&h, s, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return SNAPI_ERR_DEVICE;
  return status_of(h);
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

using namespace snapi;

extern "C" int snapi_version(void) { return 102; }

extern "C" int snapi_context_create(snapi_context **ctx) {
  if (!ctx) return SNAPI_ERR_ARG;
  *ctx = nullptr;
  const int dev = current_device();
  if (dev < 0) return SNAPI_ERR_DEVICE;
  int *w = nullptr;
  if (hipMalloc((void **)&w, sizeof(int)) != hipSuccess) return SNAPI_ERR_DEVICE;
  if (hipMemset(w, 0, sizeof(int)) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    (void)hipFree(w);
    return SNAPI_ERR_DEVICE;
  }
  *ctx = new snapi_context{dev, w};
  return SNAPI_OK;
}

extern "C" int snapi_context_destroy(snapi_context *ctx) {
  if (!ctx) return SNAPI_ERR_ARG;
  if (t_ctx == ctx) t_ctx = nullptr;
  const hipError_t e = hipFree(ctx->word);
  delete ctx;
  return e == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

extern "C" int snapi_context_bind(snapi_context *ctx) {
  t_ctx = ctx;
  return SNAPI_OK;
}

extern "C" int snapi_context_status(snapi_context *ctx, void *stream, int reset) {
  if (!ctx) return SNAPI_ERR_ARG;
  int prev = -1;
  if (hipGetDevice(&prev) != hipSuccess) return SNAPI_ERR_DEVICE;
  if (prev != ctx->device && hipSetDevice(ctx->device) != hipSuccess) return SNAPI_ERR_DEVICE;
  const int rc = take_status(ctx->word, static_cast<hipStream_t>(stream), reset);
  if (prev != ctx->device && hipSetDevice(prev) != hipSuccess) return SNAPI_ERR_DEVICE;
  return rc;
}

extern "C" int snapi_get_status(void *stream, int reset) {
  return take_status(device_status_slot(), static_cast<hipStream_t>(stream), reset);
}

extern "C" int snapi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
