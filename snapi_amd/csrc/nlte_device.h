// Launchers shared by the NLTE response kernels (nlte_resp.hip) and the
// response-system assembly (nlte_system.hip).  All arrays are device
// pointers; the launches are enqueued on st.
#pragma once
#include <hip/hip_runtime.h>

namespace snapi {

// Scalar formal solution with the response of every intensity to every
// opacity and emissivity: grid [ncol][nd]; op, em [ncol][nprob][nd]; theta
// [nprob]; I, L [ncol][nprob][nd]; Rop, Rem [ncol][nprob][nd][nd].
int launch_formal_resp(int ncol, int nprob, int nd, const double *grid, const double *op, const double *em,
                       const double *theta, double boundary, double *I, double *L, double *Rop, double *Rem,
                       hipStream_t st);

// Pivot-free Crout factorisation and solves of nsys systems: M [nsys][n][n];
// B, X [nsys][nrhs][n]; LU and LT [nsys][n][n] scratch (the factors, and the
// lower factor column by column).
int launch_response_solve(int nsys, int n, int nrhs, const double *M, const double *B, double *LU, double *LT,
                          double *X, hipStream_t st);

}  // namespace snapi
