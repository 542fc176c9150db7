// Launch-shape autotuner: the kernels that come in several compiled shapes
// (occupancy bounds, workgroup sizes) time each shape on a sample of the
// caller's own problem the first time a problem shape is seen, and keep the
// fastest per device for the rest of the process.  Every shape of a kernel
// computes the same results bit for bit; only the launch differs.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

namespace snapi {

enum TuneKernel { TUNE_SYNTH = 0, TUNE_TANGENT = 1, TUNE_ALI = 2, TUNE_NKERNEL = 3 };

// Problem shape that a choice is kept for.
struct TuneKey {
  int kernel, precision, nd, nline, ncomp, nlambda, extra;
};

// The candidate to launch for `key`.  With tuning on, the first call for a
// key on a device, on a stream that is not being captured, runs
// run(c, sample) for every candidate c < ncand (sample: the caller's own
// reduced problem, e.g. its first columns), once untimed and twice timed
// with HIP events on `stream`, and keeps the fastest.  run returns a
// snapi_status; a candidate that fails is never chosen.  With tuning off,
// during stream capture, or when a key has no choice yet in a call that may
// not tune (tune = false), the default candidate dflt is used.  A candidate
// forced through snapi_autotune_force wins over both.
int tune_choose(const TuneKey &key, int ncand, int dflt, hipStream_t stream, bool tune,
                const std::function<int(int)> &run);

// Problem-size class of a launch over ncol columns, for TuneKey::extra: 0 at
// or above the sample size (the choice is made on a sample of that many
// columns), else the bit length of ncol (the choice is made on the whole
// call), so that a small call never decides for a large one.
inline int tune_size_class(int ncol, int sample) {
  if (ncol >= sample) return 0;
  int b = 0;
  while (ncol > 0) {
    ++b;
    ncol >>= 1;
  }
  return b;
}

}  // namespace snapi
