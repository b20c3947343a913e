// tkfmi: per-kernel timing of the last timed call (TKFMI_TIME_KERNELS).
// HIP events with timing enabled are recorded around each kernel launch on
// the stream the kernel runs on; tkfmi_last_kernel_times() waits for them and
// reports the durations.  Event pairs are pooled per device and reused.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fb2d_internal.h"
#include "tkfmi.h"

namespace {

struct KRec {
  std::string name;
  int n_pairs;
  int64_t cells;
  int dev;
  int slot;
};

struct Pool {
  std::vector<hipEvent_t> a, b;
};

std::mutex g_pool_mu;
Pool g_pool[64];
thread_local std::vector<KRec> g_recs;
thread_local bool g_enabled = false;

}  // namespace

void tkfmi_ktime_begin_call(bool enabled) {
  g_enabled = enabled;
  if (enabled) g_recs.clear();
}

int tkfmi_ktime_start(const char* name, int n_pairs, int64_t cells, hipStream_t s, std::string& err) {
  if (!g_enabled) return -1;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) { err = "hipGetDevice failed"; return -2; }
  std::lock_guard<std::mutex> lk(g_pool_mu);
  Pool& p = g_pool[dev];
  const int slot = (int)g_recs.size();
  while ((int)p.a.size() <= slot) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      err = "hipEventCreate failed";
      return -2;
    }
    p.a.push_back(a);
    p.b.push_back(b);
  }
  if (hipEventRecord(p.a[slot], s) != hipSuccess) { err = "hipEventRecord failed"; return -2; }
  g_recs.push_back(KRec{name, n_pairs, cells, dev, slot});
  return slot;
}

int tkfmi_ktime_stop(int slot, hipStream_t s, std::string& err) {
  if (slot < 0) return TKFMI_OK;
  const KRec& r = g_recs[slot];
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (hipEventRecord(g_pool[r.dev].b[r.slot], s) != hipSuccess) { err = "hipEventRecord failed"; return TKFMI_EHIP; }
  return TKFMI_OK;
}

extern "C" int tkfmi_last_kernel_times(tkfmi_kernel_time* out, int32_t max) {
  if (!out && max > 0) return TKFMI_EINVAL;
  int n = 0;
  for (const KRec& r : g_recs) {
    if (n >= max) break;
    hipEvent_t a, b;
    {
      std::lock_guard<std::mutex> lk(g_pool_mu);
      a = g_pool[r.dev].a[r.slot];
      b = g_pool[r.dev].b[r.slot];
    }
    if (hipEventSynchronize(b) != hipSuccess) return TKFMI_EHIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return TKFMI_EHIP;
    tkfmi_kernel_time& o = out[n++];
    memset(&o, 0, sizeof(o));
    strncpy(o.name, r.name.c_str(), sizeof(o.name) - 1);
    o.n_pairs = r.n_pairs;
    o.cells = r.cells;
    o.ms = ms;
  }
  return n;
}
