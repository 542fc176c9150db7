// Polarized synthesis on gfx950: Zeeman opacity build fused with the
// Bezier (linear) short-characteristics formal solution, one work-item per
// (column, wavelength), the depth recurrence kept in registers and the
// column's per-depth quantities staged once in LDS.
//
// Restates atmosphere::obs_stokes (atmos/atmos_obs.cc:16-87):
//   op_em_vector          atmos/atmos_rad_synthesis.cc:20-72
//   boundbound_op_em_vector atmos/atmol/atom/atom_rad_synthesis.cc:396-505
//   h_minus_mol::op_em_vector atmos/atmol/mol/mol.cc:379-412
//   normalize_to_referent_opacity atmos/atmos_rad.cc:175-185
//   atmos_ppbez::formal (polarized) atmos/plpar/atmos_ppbez_rts.cc:918-1218
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "rt_common.h"
#include "synth_kernel.h"
#include "autotune.h"

namespace snapi {

template <typename R>
__device__ __forceinline__ void synth_body(const KLines &lines, const SynthArgs &args, const R *__restrict__ state,
                                           const double *__restrict__ lam, R *__restrict__ stokes,
                                           int *__restrict__ status, const int *__restrict__ active, int active_div) {
  extern __shared__ unsigned char smem_raw[];
  R *sm = reinterpret_cast<R *>(smem_raw);
  const int nd = args.nd;
  const int col = blockIdx.x;
  if (active && !active[col / active_div]) return;  // uniform over the workgroup
  const R *st = state + (size_t)col * args.nfield * nd;
  ColumnTile<R> tile = ColumnTile<R>::carve(sm, nd, lines.nline);
  RayGeo<R> geo = RayGeo<R>::carve(sm + ColumnTile<R>::bytes(nd, lines.nline) / sizeof(R), nd);
  stage_column<R, R, true>(lines, args, st, tile);
  stage_ray<R, R>(args, st, geo);
  __syncthreads();
  // wavelengths in index order: 64 neighbouring wavelengths of a wavefront
  // lie in the same W4 regions of most Zeeman components, so the wavefront
  // executes fewer of the four branches than any regrouping tried
  for (int l = threadIdx.x; l < args.nlambda; l += blockDim.x) {
    R out[4];
    int rc = solve_lambda<R>(lines, args, tile, geo, lam[l], out);
    R *o = stokes + ((size_t)col * args.nlambda + l) * 4;
    o[0] = out[0];
    o[1] = out[1];
    o[2] = out[2];
    o[3] = out[3];
    if (rc) atomicOr(status, rc);
  }
}

// The float kernel is held to eight wavefronts per SIMD (64 VGPRs, 228
// bytes per work-item spilled to scratch), which hides the latency of its
// LDS and kernel-argument reads and of its transcendentals far better than
// four at 113 VGPRs without spills; on the 512 x 512 x 71 cube: 4 waves
// 310 ms, 6 waves 214 ms, 7 waves 172 ms, 8 waves 160-166 ms.  The double
// kernel keeps its registers (239 VGPRs; five wavefronts would spill and
// measured 32 % slower).  Both come in a few occupancy bounds, among which
// the autotuner picks per problem shape (autotune.h); every bound computes
// the same results.
#define SNAPI_SYNTH_KERNEL(NAME, R, BOUNDS)                                                                      \
  __global__ BOUNDS void NAME(const KLines lines, const SynthArgs args, const R *__restrict__ state,             \
                              const double *__restrict__ lam, R *__restrict__ stokes, int *__restrict__ status,  \
                              const int *__restrict__ active, int active_div) {                                  \
    synth_body<R>(lines, args, state, lam, stokes, status, active, active_div);                                  \
  }
SNAPI_SYNTH_KERNEL(synth_kernel_f, float, __launch_bounds__(kSynthThreads, 8))
SNAPI_SYNTH_KERNEL(synth_kernel_f7, float, __launch_bounds__(kSynthThreads, 7))
SNAPI_SYNTH_KERNEL(synth_kernel_f6, float, __launch_bounds__(kSynthThreads, 6))
SNAPI_SYNTH_KERNEL(synth_kernel_d, double, __launch_bounds__(kSynthThreads))
SNAPI_SYNTH_KERNEL(synth_kernel_d3, double, __launch_bounds__(kSynthThreads, 3))
SNAPI_SYNTH_KERNEL(synth_kernel_d4, double, __launch_bounds__(kSynthThreads, 4))
#undef SNAPI_SYNTH_KERNEL

// Candidates of the autotuner, the default first.
using SynthFnF = void (*)(const KLines, const SynthArgs, const float *, const double *, float *, int *, const int *,
                          int);
using SynthFnD = void (*)(const KLines, const SynthArgs, const double *, const double *, double *, int *,
                          const int *, int);
template <typename R> struct SynthKernels;
template <> struct SynthKernels<float> {
  static constexpr int n = 3;
  static SynthFnF fn(int c) { return c == 1 ? synth_kernel_f7 : c == 2 ? synth_kernel_f6 : synth_kernel_f; }
};
template <> struct SynthKernels<double> {
  static constexpr int n = 3;
  static SynthFnD fn(int c) { return c == 1 ? synth_kernel_d3 : c == 2 ? synth_kernel_d4 : synth_kernel_d; }
};

// Columns of the autotuner's sample: enough workgroups to fill the chip.
constexpr int kSynthTuneColumns = 16384;

template <typename R>
static int launch_synth(const KLines &kl, const SynthArgs &a, const void *state, const double *lam, void *stokes,
                        int *status, hipStream_t stream, const int *active = nullptr, int active_div = 1) {
  int threads = ((a.nlambda + 63) / 64) * 64;
  if (threads > kSynthThreads) threads = kSynthThreads;
  const size_t smem = ColumnTile<R>::bytes(a.nd, kl.nline) + RayGeo<R>::bytes(a.nd);
  const int adiv = active_div > 0 ? active_div : 1;
  auto launch = [&](int c, int ncol) {
    hipLaunchKernelGGL(SynthKernels<R>::fn(c), dim3(ncol), dim3(threads), smem, stream, kl, a,
                       static_cast<const R *>(state), lam, static_cast<R *>(stokes), status, active, adiv);
    return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
  };
  const TuneKey key{TUNE_SYNTH, sizeof(R) == 8 ? SNAPI_FP64 : SNAPI_MIXED, a.nd, kl.nline, kl.ncomp, a.nlambda,
                    tune_size_class(a.ncol, kSynthTuneColumns)};
  const int ns = a.ncol < kSynthTuneColumns ? a.ncol : kSynthTuneColumns;
  const int c = tune_choose(key, SynthKernels<R>::n, 0, stream, true, [&](int cand) { return launch(cand, ns); });
  return launch(c, a.ncol);
}

// Synthesis of the columns whose flag active[col / active_div] is set (the
// others keep their output untouched): the trial and difference syntheses
// of the batched inversion, where finished slots are skipped.
int synth_dev_masked(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol, int nd,
                     const void *state_dev, int nlambda, const double *lam_dev, void *stokes_dev, const int *active,
                     int active_div, hipStream_t stream) {
  KLines kl;
  SynthArgs a;
  int rc = prepare_synth(lines, ray, ncol, nd, nlambda, kl, a);
  if (rc) return rc;
  if (!state_dev || !lam_dev || !stokes_dev) return SNAPI_ERR_ARG;
  int *status = device_status_slot();
  if (!status) return SNAPI_ERR_DEVICE;
  if (precision == SNAPI_FP64)
    return launch_synth<double>(kl, a, state_dev, lam_dev, stokes_dev, status, stream, active, active_div);
  if (precision == SNAPI_MIXED)
    return launch_synth<float>(kl, a, state_dev, lam_dev, stokes_dev, status, stream, active, active_div);
  return SNAPI_ERR_ARG;
}

}  // namespace snapi

using namespace snapi;

extern "C" int snapi_synth_stokes_dev(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol,
                                      int nd, const void *state_dev, int nlambda, const double *lam_dev,
                                      void *stokes_dev, void *stream) {
  KLines kl;
  SynthArgs a;
  int rc = prepare_synth(lines, ray, ncol, nd, nlambda, kl, a);
  if (rc) return rc;
  if (!state_dev || !lam_dev || !stokes_dev) return SNAPI_ERR_ARG;
  int *status = device_status_slot();
  if (!status) return SNAPI_ERR_DEVICE;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (precision == SNAPI_FP64)
    return launch_synth<double>(kl, a, state_dev, lam_dev, stokes_dev, status, s);
  if (precision == SNAPI_MIXED)
    return launch_synth<float>(kl, a, state_dev, lam_dev, stokes_dev, status, s);
  return SNAPI_ERR_ARG;
}

extern "C" int snapi_synth_stokes(int precision, const snapi_line_table *lines, const snapi_ray *ray, int ncol,
                                  int nd, const void *state, int nlambda, const double *lam, void *stokes) {
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  if (!state || !lam || !stokes || !lines || !ray || ncol <= 0 || nd < 2 || nlambda <= 0) return SNAPI_ERR_ARG;
  const size_t esz = precision == SNAPI_FP64 ? sizeof(double) : sizeof(float);
  const size_t nf = SNAPI_NF_BASE + 2 * (size_t)lines->nline;
  const size_t sbytes = (size_t)ncol * nf * nd * esz;
  const size_t obytes = (size_t)ncol * nlambda * 4 * esz;
  void *ds = nullptr, *dout = nullptr;
  double *dlam = nullptr;
  int rc = SNAPI_OK;
  if (hipMalloc(&ds, sbytes) != hipSuccess || hipMalloc(&dout, obytes) != hipSuccess ||
      hipMalloc((void **)&dlam, nlambda * sizeof(double)) != hipSuccess) {
    rc = SNAPI_ERR_DEVICE;
  } else {
    if (hipMemcpy(ds, state, sbytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dlam, lam, nlambda * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = SNAPI_ERR_DEVICE;
    else
      rc = reset_status();
    if (rc == SNAPI_OK)
      rc = snapi_synth_stokes_dev(precision, lines, ray, ncol, nd, ds, nlambda, dlam, dout, nullptr);
    if (rc == SNAPI_OK) {
      if (hipDeviceSynchronize() != hipSuccess) rc = SNAPI_ERR_DEVICE;
      else if (hipMemcpy(stokes, dout, obytes, hipMemcpyDeviceToHost) != hipSuccess) rc = SNAPI_ERR_DEVICE;
      else rc = read_status();
    }
  }
  if (ds) hipFree(ds);
  if (dout) hipFree(dout);
  if (dlam) hipFree(dlam);
  return rc;
}
