// Launchers of the batched LM normal-equation kernels (lm.hip) for the
// inversion driver.
#pragma once
#include <hip/hip_runtime.h>

namespace snapi {

// [J | r]^T [J | r] on the matrix cores: JTJ [npix][npar][npar], rhs = J^T r
// [npix][npar], chi2 = r^T r [npix] (optional); pixels with active == 0 skipped.
template <typename T>
void launch_lm_normal(int npix, int nrow, int npar, const T *J, const T *r, double *JTJ, double *rhs, double *chi2,
                      const int *active, hipStream_t s);
// damped FP64 pseudo-inverse solve per pixel
void launch_lm_solve(int npix, int npar, const double *JTJ, const double *rhs, const double *lm, double *corr,
                     const int *active, hipStream_t s, const double *reg = nullptr);

}  // namespace snapi
