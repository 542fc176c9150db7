// Scalar (unpolarised) formal solution with the diagonal approximate Lambda
// operator, batched over independent (column, frequency, ray) problems: the
// inner solve of the ALI iteration.
// Restates atmos_ppbez::formal (scalar), atmos/plpar/atmos_ppbez_rts.cc:21-219:
// the same cubic-Hermite optical-depth steps as the polarized solver, linear
// source-function weights, Lambda*_z = w1, and at the boundary Lambda* = 1
// (outgoing ray, boundary = -1) or 0.
#include <hip/hip_runtime.h>

#include <cmath>

#include "rt_common.h"

namespace snapi {

__global__ __launch_bounds__(256) void formal_scalar_kernel(int nprob, int nd, const double *__restrict__ grid,
                                                           const double *__restrict__ op,
                                                           const double *__restrict__ em,
                                                           const double *__restrict__ ctv, double boundary,
                                                           double *__restrict__ I, double *__restrict__ L) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  const size_t base = (size_t)p * nd;
  const double *g = grid + base, *o = op + base, *e = em + base;
  double *Ip = I + base, *Lp = L ? L + base : nullptr;
  const double ct = ctv[p];
  const bool up = ct < 0.0;
  const double mct = -ct;
  auto zof = [&](int k) { return up ? nd - 1 - k : k; };
  const int z0 = zof(0);
  double Iprev = up ? e[z0] / o[z0] : 0.0;
  if (up && boundary >= 0) Iprev = boundary;
  Ip[z0] = Iprev;
  if (Lp) Lp[z0] = up ? (boundary == -1.0 ? 1.0 : 0.0) : 0.0;
  if (nd < 2) return;
  // opacity derivative at the first point: one-sided
  const int z1 = zof(1);
  double oom = o[z0], oo0 = o[z1];
  double gm = g[z0], g0 = g[z1];
  double h0 = (g0 - gm) / mct;
  double Dm = (oo0 - oom) / h0;
  double Sprev = e[z0] / o[z0];
  for (int k = 1; k < nd; ++k) {
    const int z = zof(k);
    double D0;
    double hp = 0.0, gp = 0.0, oop = 0.0;
    if (k < nd - 1) {
      const int zn = zof(k + 1);
      oop = o[zn];
      gp = g[zn];
      hp = (gp - g0) / mct;
      D0 = fb_deriv<double>(oom, oo0, oop, h0, hp);
    } else {
      D0 = (oo0 - oom) / h0;  // one-sided at the last point
    }
    const double C = 0.5 * ((oo0 - h0 * 0.5 * D0) + (oom + h0 * 0.5 * Dm));
    const double tb = h0 / 3.0 * (oom + oo0 + C);
    double etb, w0, w1;
    sc_weights<double>(tb, etb, w0, w1);
    const double Sz = e[z] / o[z];
    const double In = Iprev * etb + w0 * Sprev + w1 * Sz;
    Ip[z] = In;
    if (Lp) Lp[z] = w1;
    Iprev = In;
    Sprev = Sz;
    oom = oo0;
    oo0 = oop;
    Dm = D0;
    h0 = hp;
    g0 = gp;
  }
}

}  // namespace snapi

using namespace snapi;

extern "C" int snapi_formal_scalar(int nprob, int nd, const double *grid, const double *op, const double *em,
                                   const double *ct, double boundary, double *I, double *lstar) {
  if (nprob <= 0 || nd < 2 || !grid || !op || !em || !ct || !I) return SNAPI_ERR_ARG;
  const size_t n = (size_t)nprob * nd * sizeof(double);
  double *dg = nullptr, *dop = nullptr, *dem = nullptr, *dct = nullptr, *dI = nullptr, *dL = nullptr;
  int rc = SNAPI_OK;
  if (hipMalloc((void **)&dg, n) != hipSuccess || hipMalloc((void **)&dop, n) != hipSuccess ||
      hipMalloc((void **)&dem, n) != hipSuccess || hipMalloc((void **)&dct, nprob * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&dI, n) != hipSuccess || hipMalloc((void **)&dL, n) != hipSuccess) {
    rc = SNAPI_ERR_DEVICE;
  } else {
    if (hipMemcpy(dg, grid, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dop, op, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dem, em, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dct, ct, nprob * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = SNAPI_ERR_DEVICE;
  }
  if (rc == SNAPI_OK) {
    hipLaunchKernelGGL(formal_scalar_kernel, dim3((nprob + 255) / 256), dim3(256), 0, 0, nprob, nd, dg, dop, dem,
                       dct, boundary, dI, dL);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = SNAPI_ERR_DEVICE;
    else if (hipMemcpy(I, dI, n, hipMemcpyDeviceToHost) != hipSuccess ||
             (lstar && hipMemcpy(lstar, dL, n, hipMemcpyDeviceToHost) != hipSuccess))
      rc = SNAPI_ERR_DEVICE;
  }
  hipFree(dg);
  hipFree(dop);
  hipFree(dem);
  hipFree(dct);
  hipFree(dI);
  hipFree(dL);
  return rc;
}
