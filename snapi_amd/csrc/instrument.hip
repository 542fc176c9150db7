// Instrument effects on the emergent spectra and their Jacobians: the
// Gaussian spectral convolution the fit applies to both
// (convolve_spectra_with_gauss and convolve_response_with_gauss,
// cmlib/mathtools.cc:1350-1427, 1473-1498; atmosphere::stokes_lm_fit,
// atmos/atmos_fit.cc:104-110), the convolution with a sampled spectral PSF
// (convolve_spectra_with_psf and convolve_response_with_psf,
// mathtools.cc:1429-1465, 1500-1522; atmos_fit.cc:111-115, 730-734) and the
// scattered-light offset of observable::add_scattered_light
// (cmlib/obs.cc:412-428).
#include <hip/hip_runtime.h>

#include <cmath>

#include "rt_common.h"

namespace {

constexpr int kConvThreads = 256;
constexpr int kKernelHalf = 60;  // exp(-(d/2)^2) is exactly 0 in double beyond d = 60 fine steps

// interpol_1d (cmlib/mathtools.cc:1057-1117): quadratic Bezier between the
// bracketing points with Fritsch-Butland derivatives (one-sided at the
// ends), the end values outside.  f and x are strided by fs and 1.
__device__ double interpol_1d(const double *f, int fs, const double *x, int N, double xi) {
  if (!(xi < x[N - 1] && xi > x[0])) return xi >= x[N - 1] ? f[(N - 1) * fs] : f[0];
  // first i with xi < x[i + 1], by bisection (x ascending)
  int lo = 0, hi = N - 2;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xi < x[mid + 1])
      hi = mid;
    else
      lo = mid + 1;
  }
  const int i = lo;
  double h1 = 0.0, h2 = 0.0, fd1, fd2;
  const double fi = f[i * fs], fi1 = f[(i + 1) * fs];
  if (i > 0) {
    h1 = x[i] - x[i - 1];
    h2 = x[i + 1] - x[i];
    const double d1 = (fi - f[(i - 1) * fs]) / h1, d2 = (fi1 - fi) / h2;
    const double al = 1.0 / 3.0 * (1.0 + h2 / (h1 + h2));
    fd1 = (d1 * d2 > 0) ? d1 * d2 / (al * d2 + (1.0 - al) * d1) : 0.0;
  } else {
    fd1 = (f[fs] - f[0]) / (x[1] - x[0]);
  }
  const double C0 = fi + h2 * 0.5 * fd1;
  if (i + 1 < N - 1) {
    h1 = x[i + 1] - x[i];
    h2 = x[i + 2] - x[i + 1];
    const double d1 = (fi1 - fi) / h1, d2 = (f[(i + 2) * fs] - fi1) / h2;
    const double al = 1.0 / 3.0 * (1.0 + h2 / (h1 + h2));
    fd2 = (d1 * d2 > 0) ? d1 * d2 / (al * d2 + (1.0 - al) * d1) : 0.0;
  } else {
    fd2 = (f[(N - 1) * fs] - f[(N - 2) * fs]) / (x[N - 1] - x[N - 2]);
  }
  const double C1 = fi1 - h1 * 0.5 * fd2;
  const double C = (C0 + C1) * 0.5;
  const double u = (xi - x[i]) / (x[i + 1] - x[i]);
  return (1.0 - u) * (1.0 - u) * fi + u * u * fi1 + u * 2.0 * (1.0 - u) * C;
}

// One workgroup per (series, group of up to kc components): resample onto
// the even fine grid (LDS), convolve with the Gaussian by the trapezoidal
// rule over the fine points within kKernelHalf steps (the rest of the
// reference's sum is exact zeros), resample back.  data [nseries][nl][ncomp].
template <typename R>
__global__ __launch_bounds__(kConvThreads) void gauss_convolve_kernel(int nl, int ncomp, int kc,
                                                                      const double *__restrict__ lam, double width,
                                                                      int Nf, double l_min, double step,
                                                                      R *__restrict__ data,
                                                                      const int *__restrict__ active) {
  if (active && !active[blockIdx.x]) return;  // uniform over the workgroup
  extern __shared__ double sm[];
  double *lf = sm;                   // [Nf] fine wavelengths
  double *lm = lf + Nf;              // [nl] the series' wavelengths
  double *src = lm + nl;             // [nl][kc] input components
  double *Si = src + (size_t)nl * kc; // [Nf][kc] resampled
  double *Sc = Si + (size_t)Nf * kc;  // [Nf][kc] convolved
  const int c0 = blockIdx.y * kc, nc = min(kc, ncomp - c0);
  R *d = data + (size_t)blockIdx.x * nl * ncomp;
  for (int l = threadIdx.x; l < Nf; l += blockDim.x) lf[l] = l_min + l * step;
  for (int l = threadIdx.x; l < nl; l += blockDim.x) {
    lm[l] = lam[l];
    for (int c = 0; c < kc; ++c) src[l * kc + c] = c < nc ? (double)d[(size_t)l * ncomp + c0 + c] : 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < Nf * kc; i += blockDim.x) {
    const int l = i / kc, c = i % kc;
    Si[i] = interpol_1d(src + c, kc, lm, nl, lf[l]);
  }
  __syncthreads();
  const double rw = 0.56418958354 / width;
  for (int i = threadIdx.x; i < Nf * kc; i += blockDim.x) {
    const int l = i / kc, c = i % kc;
    const int a = max(0, l - kKernelHalf), b = min(Nf - 1, l + kKernelHalf);
    double acc = 0.0;
    for (int ll = a; ll <= b; ++ll) {
      const double dl = lf[ll] - lf[l];
      const double ker = rw * exp(-dl * dl / width / width);
      const double wl = (ll == 0 || ll == Nf - 1) ? step * 0.5 : step;
      acc += (Si[ll * kc + c] * ker) * wl;
    }
    Sc[i] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nl * kc; i += blockDim.x) {
    const int l = i / kc, c = i % kc;
    if (c < nc) d[(size_t)l * ncomp + c0 + c] = (R)interpol_1d(Sc + c, kc, lf, Nf, lm[l]);
  }
}

// The fine grid of convolve_spectra_with_gauss and a component group size
// whose LDS fits; false when even one component does not fit.
bool conv_plan(const double *lam_h, int nl, double width, int &Nf, double &l_min, double &step, int &kc,
               size_t &lds) {
  l_min = lam_h[0] - 10.0 * width;
  const double l_max = lam_h[nl - 1] + 10.0 * width;
  Nf = int((l_max - l_min) / (0.5 * width)) + 1;
  step = (l_max - l_min) / (Nf - 1);
  for (kc = 8; kc >= 1; kc /= 2) {
    lds = ((size_t)Nf + nl + (size_t)nl * kc + 2 * (size_t)Nf * kc) * sizeof(double);
    if (lds <= 60 * 1024) return true;
  }
  return false;
}

template <typename R>
int launch_conv(int nseries, int nl, int ncomp, const double *lam_h, const double *lam_d, double width, R *data,
                hipStream_t st, const int *active = nullptr) {
  int Nf, kc;
  double l_min, step;
  size_t lds;
  if (!conv_plan(lam_h, nl, width, Nf, l_min, step, kc, lds)) return SNAPI_ERR_UNSUPPORTED;
  const int ng = (ncomp + kc - 1) / kc;
  hipLaunchKernelGGL(gauss_convolve_kernel<R>, dim3(nseries, ng), dim3(kConvThreads), lds, st, nl, ncomp, kc, lam_d,
                     width, Nf, l_min, step, data, active);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

int check_args(int precision, int nseries, int nl, int ncomp, const double *lam, double width) {
  if ((precision != SNAPI_FP64 && precision != SNAPI_MIXED) || nseries <= 0 || nl < 3 || ncomp <= 0 || !lam ||
      !(width > 0.0) || nseries > 0x7fffffff)
    return SNAPI_ERR_ARG;
  for (int l = 1; l < nl; ++l)
    if (!(lam[l] > lam[l - 1])) return SNAPI_ERR_ARG;
  return SNAPI_OK;
}

// Sampled-PSF convolution.  One workgroup per (series, group of up to kc
// components): the series' components are staged in LDS in double, then
//   out[l] = sum_j in[l - h + j] psf[j] / norm,  h = (npsf - 1) / 2,
// over the samples that fall on the grid, norm = sum_j psf[j] (the full
// sum, also at the ends).  Terms are summed in j order as (in psf) / norm,
// the order of the reference's loop.  data [nseries][nl][ncomp].
constexpr int kPsfMax = 1024;
template <typename R>
__global__ __launch_bounds__(kConvThreads) void psf_convolve_kernel(int nl, int ncomp, int kc, int npsf,
                                                                    const double *__restrict__ psf,
                                                                    R *__restrict__ data,
                                                                    const int *__restrict__ active) {
  if (active && !active[blockIdx.x]) return;  // uniform over the workgroup
  extern __shared__ double sm[];
  double *p = sm;                // [npsf]
  double *src = p + npsf;        // [nl][kc]
  __shared__ double norm;
  const int c0 = blockIdx.y * kc, nc = min(kc, ncomp - c0);
  R *d = data + (size_t)blockIdx.x * nl * ncomp;
  for (int j = threadIdx.x; j < npsf; j += blockDim.x) p[j] = psf[j];
  for (int i = threadIdx.x; i < nl * kc; i += blockDim.x) {
    const int l = i / kc, c = i % kc;
    src[i] = c < nc ? (double)d[(size_t)l * ncomp + c0 + c] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int j = 0; j < npsf; ++j) s += p[j];
    norm = s;
  }
  __syncthreads();
  const int h = (npsf - 1) / 2;
  const double nrm = norm;
  for (int i = threadIdx.x; i < nl * kc; i += blockDim.x) {
    const int l = i / kc, c = i % kc;
    if (c >= nc) continue;
    const int j0 = max(0, h - l), j1 = min(npsf - 1, nl - 1 - l + h);
    double acc = 0.0;
    for (int j = j0; j <= j1; ++j) acc += src[(l - h + j) * kc + c] * p[j] / nrm;
    d[(size_t)l * ncomp + c0 + c] = (R)acc;
  }
}

template <typename R>
int launch_psf(int nseries, int nl, int ncomp, int npsf, const double *psf_d, R *data, hipStream_t st,
               const int *active = nullptr) {
  int kc = 16;
  while (kc > 1 && ((size_t)npsf + (size_t)nl * kc) * sizeof(double) > 48 * 1024) kc /= 2;
  const size_t lds = ((size_t)npsf + (size_t)nl * kc) * sizeof(double);
  if (lds > 48 * 1024) return SNAPI_ERR_UNSUPPORTED;
  const int ng = (ncomp + kc - 1) / kc;
  hipLaunchKernelGGL(psf_convolve_kernel<R>, dim3(nseries, ng), dim3(kConvThreads), lds, st, nl, ncomp, kc, npsf,
                     psf_d, data, active);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

int check_psf_args(int precision, int nseries, int nl, int ncomp, int npsf, const double *psf) {
  if ((precision != SNAPI_FP64 && precision != SNAPI_MIXED) || nseries <= 0 || nl <= 0 || ncomp <= 0 || npsf <= 0 ||
      npsf > kPsfMax || !psf)
    return SNAPI_ERR_ARG;
  return SNAPI_OK;
}

// S_I += fraction * continuum_level on every wavelength of every series.
template <typename R>
__global__ void scattered_light_kernel(size_t n, R add, R *__restrict__ stokes) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    stokes[i * 4] += add;
}

}  // namespace

namespace snapi {

// The PSF convolution for callers that hold the data and the PSF on the
// device already (the inversion); series with active[series] == 0 are left
// alone.  npsf in 1..1024.
int psf_convolve_launch(int precision, int nseries, int nl, int ncomp, int npsf, const double *psf_d, void *data,
                        const int *active, hipStream_t st) {
  return precision == SNAPI_FP64 ? launch_psf<double>(nseries, nl, ncomp, npsf, psf_d, (double *)data, st, active)
                                 : launch_psf<float>(nseries, nl, ncomp, npsf, psf_d, (float *)data, st, active);
}

// The convolution for callers that hold the data already (the inversion):
// series with active[series] == 0 are left alone.
int spectral_convolve_launch(int precision, int nseries, int nl, int ncomp, const double *lam_h,
                             const double *lam_d, double width, void *data, const int *active, hipStream_t st) {
  return precision == SNAPI_FP64
             ? launch_conv<double>(nseries, nl, ncomp, lam_h, lam_d, width, (double *)data, st, active)
             : launch_conv<float>(nseries, nl, ncomp, lam_h, lam_d, width, (float *)data, st, active);
}

}  // namespace snapi

extern "C" int snapi_spectral_convolve_dev(int precision, int nseries, int nlambda, int ncomp, const double *lam,
                                           const double *lam_dev, double width, void *data_dev, void *stream) {
  int rc = check_args(precision, nseries, nlambda, ncomp, lam, width);
  if (rc) return rc;
  if (!lam_dev || !data_dev) return SNAPI_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  rc = precision == SNAPI_FP64
           ? launch_conv<double>(nseries, nlambda, ncomp, lam, lam_dev, width, (double *)data_dev, st)
           : launch_conv<float>(nseries, nlambda, ncomp, lam, lam_dev, width, (float *)data_dev, st);
  return rc;
}

extern "C" int snapi_spectral_convolve(int precision, int nseries, int nlambda, int ncomp, const double *lam,
                                       double width, void *data) {
  int rc = check_args(precision, nseries, nlambda, ncomp, lam, width);
  if (rc) return rc;
  if (!data) return SNAPI_ERR_ARG;
  const size_t es = precision == SNAPI_FP64 ? 8 : 4, n = (size_t)nseries * nlambda * ncomp * es;
  double *dl = nullptr;
  void *dd = nullptr;
  if (hipMalloc((void **)&dl, (size_t)nlambda * 8) != hipSuccess || hipMalloc(&dd, n) != hipSuccess ||
      hipMemcpy(dl, lam, (size_t)nlambda * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dd, data, n, hipMemcpyHostToDevice) != hipSuccess)
    rc = SNAPI_ERR_DEVICE;
  if (!rc) rc = snapi_spectral_convolve_dev(precision, nseries, nlambda, ncomp, lam, dl, width, dd, nullptr);
  if (!rc && (hipDeviceSynchronize() != hipSuccess || hipMemcpy(data, dd, n, hipMemcpyDeviceToHost) != hipSuccess))
    rc = SNAPI_ERR_DEVICE;
  hipFree(dl);
  (void)hipFree(dd);
  return rc;
}

extern "C" int snapi_psf_convolve_dev(int precision, int nseries, int nlambda, int ncomp, int npsf,
                                      const double *psf_dev, void *data_dev, void *stream) {
  int rc = check_psf_args(precision, nseries, nlambda, ncomp, npsf, psf_dev);
  if (rc) return rc;
  if (!data_dev) return SNAPI_ERR_ARG;
  return snapi::psf_convolve_launch(precision, nseries, nlambda, ncomp, npsf, psf_dev, data_dev, nullptr,
                                    (hipStream_t)stream);
}

extern "C" int snapi_psf_convolve(int precision, int nseries, int nlambda, int ncomp, int npsf, const double *psf,
                                  void *data) {
  int rc = check_psf_args(precision, nseries, nlambda, ncomp, npsf, psf);
  if (rc) return rc;
  if (!data) return SNAPI_ERR_ARG;
  const size_t es = precision == SNAPI_FP64 ? 8 : 4, n = (size_t)nseries * nlambda * ncomp * es;
  double *dp = nullptr;
  void *dd = nullptr;
  if (hipMalloc((void **)&dp, (size_t)npsf * 8) != hipSuccess || hipMalloc(&dd, n) != hipSuccess ||
      hipMemcpy(dp, psf, (size_t)npsf * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dd, data, n, hipMemcpyHostToDevice) != hipSuccess)
    rc = SNAPI_ERR_DEVICE;
  if (!rc) rc = snapi_psf_convolve_dev(precision, nseries, nlambda, ncomp, npsf, dp, dd, nullptr);
  if (!rc && (hipDeviceSynchronize() != hipSuccess || hipMemcpy(data, dd, n, hipMemcpyDeviceToHost) != hipSuccess))
    rc = SNAPI_ERR_DEVICE;
  (void)hipFree(dp);
  (void)hipFree(dd);
  return rc;
}

extern "C" int snapi_add_scattered_light_dev(int precision, int nseries, int nlambda, double fraction,
                                             double continuum_level, void *stokes_dev, void *stream) {
  if ((precision != SNAPI_FP64 && precision != SNAPI_MIXED) || nseries <= 0 || nlambda <= 0 || !stokes_dev)
    return SNAPI_ERR_ARG;
  const size_t n = (size_t)nseries * nlambda;
  const double add = fraction * continuum_level;
  const unsigned nb = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  if (precision == SNAPI_FP64)
    hipLaunchKernelGGL(scattered_light_kernel<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, n, add,
                       (double *)stokes_dev);
  else
    hipLaunchKernelGGL(scattered_light_kernel<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, n, (float)add,
                       (float *)stokes_dev);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

extern "C" int snapi_add_scattered_light(int precision, int nseries, int nlambda, double fraction,
                                         double continuum_level, void *stokes) {
  if ((precision != SNAPI_FP64 && precision != SNAPI_MIXED) || nseries <= 0 || nlambda <= 0 || !stokes)
    return SNAPI_ERR_ARG;
  const size_t n = (size_t)nseries * nlambda * 4 * (precision == SNAPI_FP64 ? 8 : 4);
  void *dd = nullptr;
  int rc = SNAPI_OK;
  if (hipMalloc(&dd, n) != hipSuccess || hipMemcpy(dd, stokes, n, hipMemcpyHostToDevice) != hipSuccess)
    rc = SNAPI_ERR_DEVICE;
  if (!rc) rc = snapi_add_scattered_light_dev(precision, nseries, nlambda, fraction, continuum_level, dd, nullptr);
  if (!rc && (hipDeviceSynchronize() != hipSuccess || hipMemcpy(stokes, dd, n, hipMemcpyDeviceToHost) != hipSuccess))
    rc = SNAPI_ERR_DEVICE;
  (void)hipFree(dd);
  return rc;
}
