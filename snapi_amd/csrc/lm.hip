// Batched Levenberg-Marquardt normal-equation step on gfx950.
//
// Restates the inner step of atmosphere::stokes_lm_fit (atmos/atmos_fit.cc:
// 129-160): J^T J and J^T r (multiply_with_transpose / multiply_vector,
// cmlib/mathtools.cc:615-624, 573-582), the multiplicative damping
// JTJ_ii *= (1 + lambda) (:140), and solve_svd (cmlib/mathtools.cc:344-387):
// a pseudo-inverse that drops singular values below w_max / 1e6.
//
// Kernel 1 (one wavefront per pixel): the augmented Gram matrix
// [J | r]^T [J | r] on the matrix cores, 4 rows per MFMA step
// (v_mfma_f32_16x16x4_f32 for float J, v_mfma_f64_16x16x4_f64 for double J).
// Kernel 2 (one wavefront per pixel, FP64): cyclic Jacobi eigen-decomposition
// of the damped symmetric matrix in LDS; for a symmetric positive
// semi-definite matrix its singular values are |eigenvalues| and the
// truncated pseudo-inverse Q diag(1/lambda_i) Q^T equals solve_svd's.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/snapi_gpu.h"
#include "lm_kernel.h"

namespace snapi {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int kNP = 16;  // padded parameter count (J columns + r + zero)

// Lane l of a 16x16x4 MFMA holds A[l%16][l/16] and B[l/16][l%16]; for the
// Gram matrix both are Jaug[row0 + l/16][l%16].  Element i = 0..3 of the
// result in lane l is C[row(l/16, i)][l%16]: rows 4*(l/16)+i for the f32
// instruction, rows l/16 + 4*i for the f64 one.
template <typename T> struct Mfma;
template <> struct Mfma<float> {
  typedef floatx4 acc;
  static __device__ __forceinline__ acc step(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int sub, int i) { return 4 * sub + i; }
};
template <> struct Mfma<double> {
  typedef doublex4 acc;
  static __device__ __forceinline__ acc step(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int sub, int i) { return sub + 4 * i; }
};

template <typename T>
__global__ __launch_bounds__(256) void lm_normal_kernel(int npix, int nrow, int npar, const T *__restrict__ J,
                                                       const T *__restrict__ r, double *__restrict__ JTJ,
                                                       double *__restrict__ rhs, double *__restrict__ chi2,
                                                       const int *__restrict__ active) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= npix) return;  // whole wavefronts exit together
  if (active && !active[wave]) return;
  const T *Jp = J + (size_t)wave * nrow * npar;
  const T *rp = r + (size_t)wave * nrow;
  const int col = lane & 15, sub = lane >> 4;
  typename Mfma<T>::acc c = {0, 0, 0, 0};
  for (int row0 = 0; row0 < nrow; row0 += 4) {
    const int row = row0 + sub;
    T v = T(0);
    if (row < nrow) {
      if (col < npar) v = Jp[(size_t)row * npar + col];
      else if (col == npar) v = rp[row];
    }
    c = Mfma<T>::step(v, v, c);
  }
  double *G = JTJ + (size_t)wave * npar * npar;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int rr = Mfma<T>::row(sub, i);
    const double val = (double)c[i];
    if (rr < npar && col < npar) G[rr * npar + col] = val;
    if (rr == npar && col < npar) rhs[(size_t)wave * npar + col] = val;  // (J^T r)_col = r^T J
    if (rr == npar && col == npar && chi2) chi2[wave] = val;
  }
}

// One wavefront per pixel: damped matrix in LDS, cyclic Jacobi, truncated
// pseudo-inverse applied to rhs.
// reg (or null): [npar][npar] added to every pixel's J^T J before the
// damping, the regularisation of atmos_fit.cc:1371-1480.
__global__ __launch_bounds__(64) void lm_solve_kernel(int npix, int npar, const double *__restrict__ JTJ,
                                                     const double *__restrict__ rhs, const double *__restrict__ lm,
                                                     double *__restrict__ corr, const int *__restrict__ active,
                                                     const double *__restrict__ reg) {
  __shared__ double M[kNP][kNP + 1];
  __shared__ double Q[kNP][kNP + 1];
  __shared__ double rot[2];
  __shared__ double w[kNP];
  const int pix = blockIdx.x;
  const int j = threadIdx.x;
  const int n = npar;
  if (active && !active[pix]) return;  // uniform over the workgroup
  const double *G = JTJ + (size_t)pix * n * n;
  const double damp = 1.0 + lm[pix];
  for (int i = j; i < n * n; i += 64) {
    const int a = i / n, b = i - a * n;
    M[a][b] = (reg ? G[i] + reg[i] : G[i]) * (a == b ? damp : 1.0);
    Q[a][b] = (a == b) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    // convergence: off-diagonal mass relative to the diagonal
    if (j == 0) {
      double off = 0.0, dia = 0.0;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
          if (a == b) dia += M[a][b] * M[a][b];
          else off += M[a][b] * M[a][b];
        }
      rot[0] = (off <= 1e-30 * dia || off == 0.0) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (rot[0] != 0.0) break;
    __syncthreads();
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (j == 0) {
          const double apq = M[p][q];
          double cs = 1.0, sn = 0.0;
          if (apq != 0.0) {
            const double theta = (M[q][q] - M[p][p]) / (2.0 * apq);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            cs = 1.0 / sqrt(t * t + 1.0);
            sn = t * cs;
          }
          rot[0] = cs;
          rot[1] = sn;
        }
        __syncthreads();
        const double cs = rot[0], sn = rot[1];
        if (sn != 0.0) {
          // columns p, q of M and Q
          if (j < n) {
            const double mp = M[j][p], mq = M[j][q];
            M[j][p] = cs * mp - sn * mq;
            M[j][q] = sn * mp + cs * mq;
            const double qp = Q[j][p], qq = Q[j][q];
            Q[j][p] = cs * qp - sn * qq;
            Q[j][q] = sn * qp + cs * qq;
          }
          __syncthreads();
          // rows p, q of M
          if (j < n) {
            const double mp = M[p][j], mq = M[q][j];
            M[p][j] = cs * mp - sn * mq;
            M[q][j] = sn * mp + cs * mq;
          }
        }
        __syncthreads();
      }
  }
  if (j < n) w[j] = M[j][j];
  __syncthreads();
  // pseudo-inverse with the solve_svd cut: keep w_i if w_max / |w_i| <= 1e6
  if (j < n) {
    double wmax = 0.0;
    for (int a = 0; a < n; ++a) wmax = fmax(wmax, fabs(w[a]));
    double acc = 0.0;
    for (int a = 0; a < n; ++a) {
      const double wa = fabs(w[a]);
      if (wa == 0.0 || wmax / wa > 1E6) continue;
      double proj = 0.0;  // (Q^T rhs)_a
      for (int b = 0; b < n; ++b) proj += Q[b][a] * rhs[(size_t)pix * n + b];
      acc += Q[j][a] * proj / w[a];
    }
    corr[(size_t)pix * n + j] = acc;
  }
}

template <typename T>
void launch_lm_normal(int npix, int nrow, int npar, const T *J, const T *r, double *JTJ, double *rhs, double *chi2,
                      const int *active, hipStream_t s) {
  hipLaunchKernelGGL(lm_normal_kernel<T>, dim3((npix + 3) / 4), dim3(256), 0, s, npix, nrow, npar, J, r, JTJ, rhs,
                     chi2, active);
}
template void launch_lm_normal<float>(int, int, int, const float *, const float *, double *, double *, double *,
                                      const int *, hipStream_t);
template void launch_lm_normal<double>(int, int, int, const double *, const double *, double *, double *, double *,
                                       const int *, hipStream_t);

void launch_lm_solve(int npix, int npar, const double *JTJ, const double *rhs, const double *lm, double *corr,
                     const int *active, hipStream_t s, const double *reg) {
  hipLaunchKernelGGL(lm_solve_kernel, dim3(npix), dim3(64), 0, s, npix, npar, JTJ, rhs, lm, corr, active, reg);
}

}  // namespace snapi

using namespace snapi;

extern "C" int snapi_lm_step(int precision, int npix, int nrow, int npar, const void *J, const void *r,
                             const double *lm, double *JTJ, double *rhs, double *corr) {
  if (npix <= 0 || nrow <= 0 || npar <= 0 || npar > kNP - 1 || !J || !r || !lm || !JTJ || !rhs || !corr)
    return SNAPI_ERR_ARG;
  if (precision != SNAPI_FP64 && precision != SNAPI_MIXED) return SNAPI_ERR_ARG;
  const size_t es = precision == SNAPI_FP64 ? 8 : 4;
  void *dJ = nullptr, *dr = nullptr;
  double *dG = nullptr, *drhs = nullptr, *dlm = nullptr, *dcorr = nullptr;
  int rc = SNAPI_OK;
  const size_t nJ = (size_t)npix * nrow * npar, nr = (size_t)npix * nrow, nG = (size_t)npix * npar * npar;
  if (hipMalloc(&dJ, nJ * es) != hipSuccess || hipMalloc(&dr, nr * es) != hipSuccess ||
      hipMalloc((void **)&dG, nG * 8) != hipSuccess || hipMalloc((void **)&drhs, (size_t)npix * npar * 8) != hipSuccess ||
      hipMalloc((void **)&dlm, (size_t)npix * 8) != hipSuccess ||
      hipMalloc((void **)&dcorr, (size_t)npix * npar * 8) != hipSuccess) {
    rc = SNAPI_ERR_DEVICE;
  } else {
    if (hipMemcpy(dJ, J, nJ * es, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dr, r, nr * es, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dlm, lm, (size_t)npix * 8, hipMemcpyHostToDevice) != hipSuccess)
      rc = SNAPI_ERR_DEVICE;
  }
  if (rc == SNAPI_OK) {
    if (precision == SNAPI_FP64)
      launch_lm_normal<double>(npix, nrow, npar, (const double *)dJ, (const double *)dr, dG, drhs, nullptr, nullptr,
                               0);
    else
      launch_lm_normal<float>(npix, nrow, npar, (const float *)dJ, (const float *)dr, dG, drhs, nullptr, nullptr, 0);
    launch_lm_solve(npix, npar, dG, drhs, dlm, dcorr, nullptr, 0);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = SNAPI_ERR_DEVICE;
    else if (hipMemcpy(JTJ, dG, nG * 8, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(rhs, drhs, (size_t)npix * npar * 8, hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(corr, dcorr, (size_t)npix * npar * 8, hipMemcpyDeviceToHost) != hipSuccess)
      rc = SNAPI_ERR_DEVICE;
  }
  hipFree(dJ);
  hipFree(dr);
  hipFree(dG);
  hipFree(drhs);
  hipFree(dlm);
  hipFree(dcorr);
  return rc;
}
