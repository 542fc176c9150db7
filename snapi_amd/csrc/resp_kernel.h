// Shared declarations of the adjoint response kernel (responses.hip), used
// by the batched inversion driver (invert.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "synth_kernel.h"

namespace snapi {

constexpr int kRespThreads = 128;
constexpr int kFwd = 15;  // per depth: K[7], E[4], I[4]
constexpr int kAdj = 11;  // per depth and seed: Kbar[7], Ebar[4]
enum { SINK_RESP = 0, SINK_GRAD = 1, SINK_NODE = 2 };

struct RespArgs {
  int nseed;          // 4 (responses, node Jacobian) or 1 (gradient)
  int sink;
  int col0;           // first column of this launch
  long long nwork;    // work-items of this launch (columns * nlambda)
  const void *obs;    // [ncol][nlambda][4] (gradient), state precision
  double ws[4];       // Stokes weights (gradient, node Jacobian)
  // node Jacobian: J[col][l*4+s][n] = sum_z dS_s/dq_{pidx[n]}(z) nmap[col][n][z] * nscale[n] * ws[s] * mask[col][l]
  int nnode;
  int pidx[SNAPI_MAX_NODES];
  double nscale[SNAPI_MAX_NODES];
  const void *nmap;   // [ncol][nnode][nd]
  // a second parameter row per node (temperature nodes move T and, through
  // hydrostatic equilibrium, Nt): with nmap2 set, node n with pidx2[n] >= 0
  // adds sum_z dS_s/dq_{pidx2[n]}(z) nmap2[col][n][z]
  int pidx2[SNAPI_MAX_NODES];
  const void *nmap2;  // [ncol][nnode][nd] or null
  const void *mask;   // [ncol][nlambda] or null
  const int *active;  // [ncol] or null: columns with 0 are skipped
  int staged_derivs;  // set by launch_resp: local derivatives from per-column LDS tiles
};


template <typename R>
int launch_resp(const KLines &kl, const SynthArgs &a, RespArgs ra, int ncol, const void *ds_, const void *dd_,
                const double *dl, void *dstk_, void *dout_, void *dscr_, int chunk, hipStream_t stream);
size_t resp_scratch_per_col(int nlambda, int nd, size_t esz);

// Node Jacobian by the forward-mode sweep (tangent.hip): uses ra.nnode,
// pidx (0..6), pidx2 and nmap2 (temperature nodes), nscale, ws, nmap, mask
// and active; writes stokes [ncol][nlambda][4] and J[ncol][nlambda*4][nnode].
// Temperature and density rows read dd_ (dstate [ncol][2][nfield][nd]).
constexpr int kTanMaxNodes = 8;  // node directions carried per workgroup
template <typename R>
int launch_tangent(const KLines &kl, const SynthArgs &a, RespArgs ra, int ncol, const void *ds_, const void *dd_,
                   const double *dl, void *dstk_, void *dJ_, hipStream_t stream);

}  // namespace snapi
