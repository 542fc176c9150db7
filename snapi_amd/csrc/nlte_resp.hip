// The two numerical kernels of the NLTE population responses on gfx950, in
// double with the reference's operation order (no fused multiply-adds):
//
//   formal_resp_kernel   atmos_ppbez::formal_with_responses
//                        (atmos/plpar/atmos_ppbez_rts.cc:349-778): scalar
//                        formal solution plus dI_z/dop_z' and dI_z/dem_z' for
//                        every pair of depths.  One wavefront per (column,
//                        problem); rows z advance along the ray, lanes own the
//                        columns z' of the two response matrices and keep the
//                        previous row in registers.  The reference's
//                        bezier_delimiter (1e10) keeps the weights linear, so
//                        the terms it multiplies by w2 = 0 are left out, and
//                        d(dtau)/d(op) has four non-zero entries per row,
//                        which are formed from the per-row Fritsch-Butland
//                        stencil derivatives staged in LDS.
//   crout_kernel,        the Crout factorisation without pivoting and the
//   crout_solve_kernel   solves of atom::compute_nlte_population_responses
//                        (atmos/atmol/atom/atom_responses.cc:171-200; Crout /
//                        solveCrout, cmlib/mathtools.cc:487-514).  One
//                        workgroup per system: the factorisation walks the
//                        columns as Crout does (dot products in the same
//                        order), the solves put one work-item on each
//                        right-hand side.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "nlte_device.h"
#include "rt_common.h"

#pragma clang fp contract(off)

namespace {

constexpr int kFrThreads = 64;   // one wavefront per formal problem
constexpr int kFrMaxPer = 4;     // depths per lane: nd <= 256
constexpr int kCroutThreads = 512;
constexpr int kSolveThreads = 256;

// Linear short-characteristics weights and their step derivatives
// (atmos_ppbez_rts.cc:615-636, 705-717).
__device__ __forceinline__ void lin_weights(double tb, double &etb, double &w0, double &w1, double &dw0,
                                            double &dw1) {
  etb = (tb < 1E-2) ? (1.0 - tb + tb * tb * 0.5 - tb * tb * tb / 6.0) : exp(-tb);
  if (tb < 1E-2) {
    w0 = tb * (0.5 - tb * (1.0 / 3.0 - tb / 8.0));
    dw0 = 1.0 / 2.0 - tb * (2.0 / 3.0 - 3.0 * tb / 8.0);
    w1 = 0.5 * tb * (1.0 - tb / 3.0 * (1.0 - tb / 4.0));
    dw1 = 0.5 - tb * (1.0 / 3.0 - tb / 8.0);
  } else {
    w0 = 1.0 / tb - etb * (1.0 + 1.0 / tb);
    dw0 = -w0 / tb + etb;
    w1 = 1.0 - 1.0 / tb + etb / tb;
    dw1 = (1.0 - w1) / tb - etb / tb;
  }
}

__global__ __launch_bounds__(kFrThreads) void formal_resp_kernel(int nprob, int nd, const double *__restrict__ grid,
                                                                 const double *__restrict__ op_all,
                                                                 const double *__restrict__ em_all,
                                                                 const double *__restrict__ theta, double boundary,
                                                                 double *__restrict__ I_all,
                                                                 double *__restrict__ L_all,
                                                                 double *__restrict__ Rop_all,
                                                                 double *__restrict__ Rem_all) {
  extern __shared__ double sm[];
  double *g = sm, *oo = sm + nd, *ee = sm + 2 * nd, *D = sm + 3 * nd, *dtau = sm + 4 * nd;
  double *dDb = sm + 5 * nd, *dDc = sm + 6 * nd, *dDf = sm + 7 * nd;  // row i: entries i-dz, i, i+dz
  const int pb = blockIdx.x;  // (column, problem)
  const int col = pb / nprob, prob = pb - col * nprob;
  const int lane = threadIdx.x;
  const size_t base = (size_t)pb * nd;
  for (int z = lane; z < nd; z += kFrThreads) {
    g[z] = grid[(size_t)col * nd + z];
    oo[z] = op_all[base + z];
    ee[z] = em_all[base + z];
  }
  __syncthreads();
  const double ct = cos(theta[prob]);
  const int dz = (ct > 0) ? 1 : -1, zl = (ct > 0) ? 0 : nd - 1, zh = (ct > 0) ? nd - 1 : 0;
  // Fritsch-Butland derivatives of the opacity and their derivatives with
  // respect to the opacities (atmos_ppbez_rts.cc:415-497)
  for (int z = lane; z < nd; z += kFrThreads) {
    double Dz = 0.0, b = 0.0, c = 0.0, f = 0.0;
    if (z == zl) {
      Dz = (oo[zl + dz] - oo[zl]) / (g[zl + dz] - g[zl]) * (-ct);
      c = -1.0 / (g[zl + dz] - g[zl]) * (-ct);
      f = 1.0 / (g[zl + dz] - g[zl]) * (-ct);
    } else if (z == zh) {
      Dz = (oo[zh] - oo[zh - dz]) / (g[zh] - g[zh - dz]) * (-ct);
      b = -1.0 / (g[zh] - g[zh - dz]) * (-ct);
      c = 1.0 / (g[zh] - g[zh - dz]) * (-ct);
    } else {
      const double h2 = (g[z + dz] - g[z]) / (-ct);
      const double h1 = (g[z] - g[z - dz]) / (-ct);
      const double d2 = (oo[z + dz] - oo[z]) / h2;
      const double d1 = (oo[z] - oo[z - dz]) / h1;
      const double alpha = (1.0 + h1 / (h1 + h2)) / 3.0;
      if (d1 * d2 > 0.0) {
        Dz = d1 * d2 / (alpha * d1 + (1.0 - alpha) * d2);
        const double den = (alpha * d1 + (1.0 - alpha) * d2);
        b = (-1.0 / h1 * d2 * den + alpha / h1 * d1 * d2) / den / den;
        c = (1.0 / h1 * d2 * den - 1.0 / h2 * d1 * den - alpha / h1 * d1 * d2 + (1.0 - alpha) / h2 * d1 * d2) / den /
            den;
        f = (1.0 / h2 * d1 * den - (1.0 - alpha) / h2 * d1 * d2) / den / den;
      }
    }
    D[z] = Dz;
    dDb[z] = b;
    dDc[z] = c;
    dDf[z] = f;
  }
  __syncthreads();
  for (int z = lane; z < nd; z += kFrThreads) {
    if (z == zl) {
      dtau[z] = 0.0;
    } else {
      const double C0 = oo[z] - (g[z] - g[z - dz]) / (-ct) / 2.0 * D[z];
      const double C1 = oo[z - dz] + (g[z] - g[z - dz]) / (-ct) / 2.0 * D[z - dz];
      const double C = 0.5 * (C0 + C1);
      dtau[z] = (g[z] - g[z - dz]) / (-ct) / 3.0 * (oo[z - dz] + oo[z] + C);
    }
  }
  __syncthreads();
  double *Iout = I_all + base, *Lout = L_all + base;
  double *Rop = Rop_all + base * nd, *Rem = Rem_all + base * nd;
  // boundary row (atmos_ppbez_rts.cc:382-413)
  double Iprev = (ct > 0) ? 0.0 : ee[zl] / oo[zl];
  if (ct < 0 && boundary >= 0) Iprev = boundary;
  double po[kFrMaxPer], pe[kFrMaxPer];
#pragma unroll
  for (int m = 0; m < kFrMaxPer; ++m) {
    const int zp = lane + m * kFrThreads;
    double vo = 0.0, ve = 0.0;
    if (ct <= 0 && boundary == -1 && zp == zl) {
      vo = -ee[zl] / oo[zl] / oo[zl];
      ve = 1.0 / oo[zl];
    }
    po[m] = vo;
    pe[m] = ve;
    if (zp < nd) {
      Rop[(size_t)zl * nd + zp] = vo;
      Rem[(size_t)zl * nd + zp] = ve;
    }
  }
  if (lane == 0) {
    Iout[zl] = Iprev;
    double Lz = 0.0;
    if (ct <= 0) Lz = (boundary == -1) ? 1.0 : 0.0;
    Lout[zl] = Lz;
  }
  // rows along the ray (atmos_ppbez_rts.cc:573-763 with w2 = 0)
  for (int z = zl + dz; (z - zh - dz) * dz < 0; z += dz) {
    const double tb = dtau[z];
    double etb, w0, w1, dw0, dw1;
    lin_weights(tb, etb, w0, w1, dw0, dw1);
    const double s0 = ee[z - dz] / oo[z - dz], s1 = ee[z] / oo[z];
    const double Inew = Iprev * etb + w0 * s0 + w1 * s1;
    // d(dtau_z)/d(op_j) for j = z-2dz, z-dz, z, z+dz (atmos_ppbez_rts.cc:500-512)
    const double hi = (g[z] - g[z - dz]) / (-ct);
    const double q = hi * hi / 12.0;
    const int zm = z - dz;
    const double x_m2 = q * dDb[zm] - q * 0.0;
    const double x_m1 = 1.0 / 2.0 * hi + (q * dDc[zm] - q * dDb[z]);
    const double x_0 = 1.0 / 2.0 * hi + (q * dDf[zm] - q * dDc[z]);
    const double x_p1 = 0.0 + (q * 0.0 - q * dDf[z]);
    const double a1 = -Iprev * etb, a2 = (s0 * dw0 + s1 * dw1);
    const double e_loc = w1 / oo[z], e_up = w0 / oo[zm];
    const double o_loc = -w1 * ee[z] / oo[z] / oo[z], o_up = -w0 * ee[zm] / oo[zm] / oo[zm];
#pragma unroll
    for (int m = 0; m < kFrMaxPer; ++m) {
      const int zp = lane + m * kFrThreads;
      double re = (zp == z) ? e_loc : (zp == zm) ? e_up : 0.0;
      double ro = (zp == z) ? o_loc : (zp == zm) ? o_up : 0.0;
      re += pe[m] * etb;
      ro += po[m] * etb;
      double x = 0.0;
      bool band = true;
      if (zp == z - 2 * dz)
        x = x_m2;
      else if (zp == zm)
        x = x_m1;
      else if (zp == z)
        x = x_0;
      else if (zp == z + dz)
        x = x_p1;
      else
        band = false;
      if (band) {
        ro += a1 * x;
        ro += a2 * x;
      }
      pe[m] = re;
      po[m] = ro;
      if (zp < nd) {
        Rop[(size_t)z * nd + zp] = ro;
        Rem[(size_t)z * nd + zp] = re;
      }
    }
    if (lane == 0) {
      Iout[z] = Inew;
      Lout[z] = w1;
    }
    Iprev = Inew;
  }
}

// Crout without pivoting (cmlib/mathtools.cc:487-500): column k of the lower
// triangle, then row k of the unit upper triangle, each entry a dot product
// over p < k in ascending order.  The lower triangle is also kept column by
// column in LT ([k][i] = L[i][k]) so that the column-k products read it
// coalesced; the step's shared operands (column k of U, row k of L) are
// staged in LDS, and each work-item loads its own operands kUnroll at a time
// before adding the products in order.  D holds the factors as the
// reference lays them out.
constexpr int kUnroll = 16;

__device__ __forceinline__ double ordered_dot(int k, const double *__restrict__ a, size_t astride,
                                              const double *__restrict__ b) {
  double sum = 0.;
  int p = 0;
  for (; p + kUnroll <= k; p += kUnroll) {
    double va[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) va[u] = a[(size_t)(p + u) * astride];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) sum += va[u] * b[p + u];
  }
  for (; p < k; ++p) sum += a[(size_t)p * astride] * b[p];
  return sum;
}

__global__ __launch_bounds__(kCroutThreads) void crout_kernel(int n, const double *__restrict__ S_all,
                                                              double *__restrict__ D_all,
                                                              double *__restrict__ LT_all) {
  extern __shared__ double stage[];  // column k of U [n], row k of L [n]
  double *ucol = stage, *lrow = stage + n;
  const size_t off = (size_t)blockIdx.x * n * n;
  const double *S = S_all + off;
  double *D = D_all + off;
  double *LT = LT_all + off;
  for (int k = 0; k < n; ++k) {
    for (int p = threadIdx.x; p < k; p += blockDim.x) {
      ucol[p] = D[(size_t)p * n + k];
      lrow[p] = D[(size_t)k * n + p];
    }
    __syncthreads();
    for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
      const double v = S[(size_t)i * n + k] - ordered_dot(k, LT + i, n, ucol);
      D[(size_t)i * n + k] = v;
      LT[(size_t)k * n + i] = v;
    }
    __syncthreads();
    const double dkk = D[(size_t)k * n + k];
    for (int j = k + 1 + threadIdx.x; j < n; j += blockDim.x)
      D[(size_t)k * n + j] = (S[(size_t)k * n + j] - ordered_dot(k, D + j, n, lrow)) / dkk;
    __syncthreads();
  }
}

// solveCrout (cmlib/mathtools.cc:502-514) for nrhs right-hand sides, one
// work-item each, blockDim.x of them per workgroup; each work-item's y, then
// x, lives in its own LDS column ([row][blockDim.x]) while the factors' rows
// are read uniformly.  B, X: [sys][nrhs][n].
__global__ __launch_bounds__(kSolveThreads) void crout_solve_kernel(int n, int nrhs, const double *__restrict__ LU_all,
                                                                    const double *__restrict__ B_all,
                                                                    double *__restrict__ X_all) {
  extern __shared__ double yl[];
  const int T = blockDim.x, tid = threadIdx.x;
  const int r = blockIdx.y * T + tid;
  if (r >= nrhs) return;
  const double *LU = LU_all + (size_t)blockIdx.x * n * n;
  const double *b = B_all + ((size_t)blockIdx.x * nrhs + r) * n;
  double *x = X_all + ((size_t)blockIdx.x * nrhs + r) * n;
  double *y = yl + tid;
  for (int i = 0; i < n; ++i) {
    double sum = 0.;
    const double *row = LU + (size_t)i * n;
    for (int k = 0; k < i; ++k) sum += row[k] * y[k * T];
    y[i * T] = (b[i] - sum) / row[i];
  }
  // back substitution: x overwrites y from the last row up
  for (int i = n - 1; i >= 0; --i) {
    double sum = 0.;
    const double *row = LU + (size_t)i * n;
    for (int k = i + 1; k < n; ++k) sum += row[k] * y[k * T];
    y[i * T] = (y[i * T] - sum);
  }
  for (int i = 0; i < n; ++i) x[i] = y[i * T];
}

// Right-hand sides per solve workgroup: as many LDS columns of n doubles as
// the device's per-workgroup LDS holds, at most kSolveThreads.
int solve_tile(int n) {
  static int lim = 0;
  if (!lim) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      v = 64 * 1024;
    lim = v;
  }
  const long t = (long)lim / ((long)n * (long)sizeof(double));
  return (int)std::max(0L, std::min<long>(kSolveThreads, t));
}

#define NCHK(x)              \
  do {                       \
    if ((x) != hipSuccess) { \
      rc = SNAPI_ERR_DEVICE; \
      goto done;             \
    }                        \
  } while (0)

}  // namespace

namespace snapi {

int launch_formal_resp(int ncol, int nprob, int nd, const double *grid, const double *op, const double *em,
                       const double *theta, double boundary, double *I, double *L, double *Rop, double *Rem,
                       hipStream_t st) {
  if (nd < 2 || nd > kFrThreads * kFrMaxPer) return SNAPI_ERR_UNSUPPORTED;
  const long nb = (long)ncol * nprob;
  if (nb > 0x7fffffffL) return SNAPI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(formal_resp_kernel, dim3((unsigned)nb), dim3(kFrThreads), 8 * nd * sizeof(double), st, nprob,
                     nd, grid, op, em, theta, boundary, I, L, Rop, Rem);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

int launch_response_solve(int nsys, int n, int nrhs, const double *M, const double *B, double *LU, double *LT,
                          double *X, hipStream_t st) {
  const int T = solve_tile(n);
  if (T < 1) return SNAPI_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(crout_kernel, dim3(nsys), dim3(kCroutThreads), 2 * (size_t)n * sizeof(double), st, n, M, LU,
                     LT);
  if (hipGetLastError() != hipSuccess) return SNAPI_ERR_DEVICE;
  const dim3 grid(nsys, (nrhs + T - 1) / T);
  hipLaunchKernelGGL(crout_solve_kernel, grid, dim3(T), (size_t)T * n * sizeof(double), st, n, nrhs, LU, B, X);
  return hipGetLastError() == hipSuccess ? SNAPI_OK : SNAPI_ERR_DEVICE;
}

}  // namespace snapi

using snapi::launch_formal_resp;
using snapi::launch_response_solve;

extern "C" int snapi_formal_responses_dev(int ncol, int nprob, int nd, const double *grid_dev, const double *op_dev,
                                          const double *em_dev, const double *theta_dev, double boundary,
                                          double *I_dev, double *L_dev, double *Rop_dev, double *Rem_dev,
                                          void *stream) {
  if (ncol <= 0 || nprob <= 0 || nd <= 0 || !grid_dev || !op_dev || !em_dev || !theta_dev || !I_dev || !L_dev ||
      !Rop_dev || !Rem_dev)
    return SNAPI_ERR_ARG;
  return launch_formal_resp(ncol, nprob, nd, grid_dev, op_dev, em_dev, theta_dev, boundary, I_dev, L_dev, Rop_dev,
                            Rem_dev, (hipStream_t)stream);
}

extern "C" int snapi_formal_responses(int ncol, int nprob, int nd, const double *grid, const double *op,
                                      const double *em, const double *theta, double boundary, double *I, double *L,
                                      double *Rop, double *Rem) {
  if (ncol <= 0 || nprob <= 0 || nd <= 0 || !grid || !op || !em || !theta || !I || !L || !Rop || !Rem)
    return SNAPI_ERR_ARG;
  if (nd < 2 || nd > kFrThreads * kFrMaxPer) return SNAPI_ERR_UNSUPPORTED;
  const size_t np = (size_t)ncol * nprob;
  int rc = SNAPI_OK;
  double *dg = nullptr, *dop = nullptr, *dem = nullptr, *dth = nullptr, *dI = nullptr, *dL = nullptr;
  double *dRo = nullptr, *dRe = nullptr;
  NCHK(hipMalloc((void **)&dg, (size_t)ncol * nd * 8));
  NCHK(hipMalloc((void **)&dop, np * nd * 8));
  NCHK(hipMalloc((void **)&dem, np * nd * 8));
  NCHK(hipMalloc((void **)&dth, (size_t)nprob * 8));
  NCHK(hipMalloc((void **)&dI, np * nd * 8));
  NCHK(hipMalloc((void **)&dL, np * nd * 8));
  NCHK(hipMalloc((void **)&dRo, np * nd * nd * 8));
  NCHK(hipMalloc((void **)&dRe, np * nd * nd * 8));
  NCHK(hipMemcpy(dg, grid, (size_t)ncol * nd * 8, hipMemcpyHostToDevice));
  NCHK(hipMemcpy(dop, op, np * nd * 8, hipMemcpyHostToDevice));
  NCHK(hipMemcpy(dem, em, np * nd * 8, hipMemcpyHostToDevice));
  NCHK(hipMemcpy(dth, theta, (size_t)nprob * 8, hipMemcpyHostToDevice));
  rc = launch_formal_resp(ncol, nprob, nd, dg, dop, dem, dth, boundary, dI, dL, dRo, dRe, nullptr);
  if (rc) goto done;
  NCHK(hipDeviceSynchronize());
  NCHK(hipMemcpy(I, dI, np * nd * 8, hipMemcpyDeviceToHost));
  NCHK(hipMemcpy(L, dL, np * nd * 8, hipMemcpyDeviceToHost));
  NCHK(hipMemcpy(Rop, dRo, np * nd * nd * 8, hipMemcpyDeviceToHost));
  NCHK(hipMemcpy(Rem, dRe, np * nd * nd * 8, hipMemcpyDeviceToHost));
done:
  hipFree(dg); hipFree(dop); hipFree(dem); hipFree(dth); hipFree(dI); hipFree(dL); hipFree(dRo); hipFree(dRe);
  return rc;
}

extern "C" int snapi_response_solve(int nsys, int n, int nrhs, const double *M, const double *B, double *X) {
  if (nsys <= 0 || n <= 0 || nrhs <= 0 || !M || !B || !X) return SNAPI_ERR_ARG;
  const size_t nn = (size_t)n * n;
  int rc = SNAPI_OK;
  double *dM = nullptr, *dB = nullptr, *dLU = nullptr, *dY = nullptr, *dX = nullptr;
  NCHK(hipMalloc((void **)&dM, nsys * nn * 8));
  NCHK(hipMalloc((void **)&dLU, nsys * nn * 8));
  NCHK(hipMalloc((void **)&dB, (size_t)nsys * nrhs * n * 8));
  NCHK(hipMalloc((void **)&dY, nsys * nn * 8));
  NCHK(hipMalloc((void **)&dX, (size_t)nsys * nrhs * n * 8));
  NCHK(hipMemcpy(dM, M, nsys * nn * 8, hipMemcpyHostToDevice));
  NCHK(hipMemcpy(dB, B, (size_t)nsys * nrhs * n * 8, hipMemcpyHostToDevice));
  rc = launch_response_solve(nsys, n, nrhs, dM, dB, dLU, dY, dX, nullptr);
  if (rc) goto done;
  NCHK(hipDeviceSynchronize());
  NCHK(hipMemcpy(X, dX, (size_t)nsys * nrhs * n * 8, hipMemcpyDeviceToHost));
done:
  hipFree(dM); hipFree(dB); hipFree(dLU); hipFree(dY); hipFree(dX);
  return rc;
}
