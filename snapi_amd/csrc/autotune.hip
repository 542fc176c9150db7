// Launch-shape autotuner (autotune.h) and its C entry points.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "autotune.h"
#include "rt_common.h"

namespace snapi {

namespace {

constexpr int kMaxCand = 8;

struct Entry {
  int choice;
  int ncand;
  float ms[kMaxCand];  // best timed run of each candidate, < 0 when it failed
};

using Key = std::tuple<int, int, int, int, int, int, int, int>;  // device + TuneKey

std::mutex g_mu;
std::map<Key, Entry> g_table;
int g_mode = -1;                 // -1: not read from the environment yet
int g_force[TUNE_NKERNEL] = {-1, -1, -1};

int mode_locked() {
  if (g_mode < 0) {
    const char *e = getenv("SNAPI_AUTOTUNE");
    g_mode = (e && e[0] == '0') ? 0 : 1;
  }
  return g_mode;
}

Key make_key(int dev, const TuneKey &k) {
  return Key(dev, k.kernel, k.precision, k.nd, k.nline, k.ncomp, k.nlambda, k.extra);
}

}  // namespace

int tune_choose(const TuneKey &key, int ncand, int dflt, hipStream_t stream, bool tune,
                const std::function<int(int)> &run) {
  if (ncand <= 1) return 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return dflt;
  const Key k = make_key(dev, key);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    const int f = (key.kernel >= 0 && key.kernel < TUNE_NKERNEL) ? g_force[key.kernel] : -1;
    if (f >= 0 && f < ncand) return f;
    auto it = g_table.find(k);
    if (it != g_table.end()) return it->second.choice;
    if (!mode_locked() || !tune) return dflt;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return dflt;
  Entry e;
  e.ncand = ncand < kMaxCand ? ncand : kMaxCand;
  e.choice = dflt;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return dflt;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return dflt;
  }
  float best = 0.0f;
  bool any = false;
  for (int c = 0; c < e.ncand; ++c) {
    e.ms[c] = -1.0f;
    if (run(c) != SNAPI_OK) continue;  // untimed: loads the code object
    float tmin = -1.0f;
    bool ok = true;
    for (int r = 0; r < 2 && ok; ++r) {
      ok = hipEventRecord(e0, stream) == hipSuccess && run(c) == SNAPI_OK &&
           hipEventRecord(e1, stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
      float ms = 0.0f;
      if (ok) ok = hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
      if (ok && (tmin < 0.0f || ms < tmin)) tmin = ms;
    }
    if (!ok) continue;
    e.ms[c] = tmin;
    if (!any || tmin < best) {
      best = tmin;
      e.choice = c;
      any = true;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::lock_guard<std::mutex> lk(g_mu);
  if (any) g_table[k] = e;
  return e.choice;
}

}  // namespace snapi

using namespace snapi;

extern "C" int snapi_autotune_mode(int mode) {
  std::lock_guard<std::mutex> lk(g_mu);
  const int old = mode_locked();
  if (mode == 0 || mode == 1) g_mode = mode;
  else if (mode != -1) return SNAPI_ERR_ARG;
  return old;
}

extern "C" int snapi_autotune_force(int kernel, int candidate) {
  if (kernel < 0 || kernel >= TUNE_NKERNEL || candidate < -1 || candidate >= kMaxCand) return SNAPI_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  g_force[kernel] = candidate;
  return SNAPI_OK;
}

extern "C" int snapi_autotune_clear(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table.clear();
  return SNAPI_OK;
}

extern "C" int snapi_autotune_report(char *buf, size_t n) {
  if (!buf || n == 0) return SNAPI_ERR_ARG;
  std::string s;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (const auto &kv : g_table) {
      char line[512];
      const Key &k = kv.first;
      const Entry &e = kv.second;
      int m = snprintf(line, sizeof line,
                       "{\"device\": %d, \"kernel\": %d, \"precision\": %d, \"nd\": %d, \"nline\": %d, "
                       "\"ncomp\": %d, \"nlambda\": %d, \"extra\": %d, \"choice\": %d, \"ms\": [",
                       std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k), std::get<4>(k),
                       std::get<5>(k), std::get<6>(k), std::get<7>(k), e.choice);
      for (int c = 0; c < e.ncand && m > 0 && (size_t)m < sizeof line; ++c)
        m += snprintf(line + m, sizeof line - m, c ? ", %.4f" : "%.4f", (double)e.ms[c]);
      if (m > 0 && (size_t)m < sizeof line) snprintf(line + m, sizeof line - m, "]}\n");
      s += line;
    }
  }
  if (s.size() + 1 > n) return SNAPI_ERR_CAPACITY;
  memcpy(buf, s.c_str(), s.size() + 1);
  return SNAPI_OK;
}
