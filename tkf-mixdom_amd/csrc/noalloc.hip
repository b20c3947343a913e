// tkfmi: tkfmi_fb2d_batched without any device allocation.
//
// The existing entry point takes its scratch memory (per-pair descriptors,
// the forward lattice, boundary rings between chained workgroups, ticket
// counters and flags) from a caller-provided workspace.  This variant takes
// it from `scratch`, a file-scope device array of TKFMI_NOALLOC_SCRATCH_BYTES
// that every GPU the library runs on holds until the process ends.  Its
// address is looked up once per device with hipGetSymbolAddress (in host
// code the array's name is not a device address) and that pointer is what
// the kernels get.  While the call runs, uploads copy from the caller's host
// memory and every launch goes on the caller's stream, so nothing is
// allocated: no device memory, pinned staging buffers, streams or events.
//
// Every call on a GPU shares that GPU's scratch, which keeps the last call's
// contents; the kernels clear what must start at zero (ticket counters and
// progress flags) at the start of each call.  Calls that overlap on one GPU
// (from several streams or threads) race on the scratch and must be
// serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "fb2d_internal.h"
#include "tkfmi.h"

namespace {

constexpr size_t kScratchBytes = TKFMI_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device
thread_local bool g_active = false;

// The current device's copy of `scratch` (looked up on the first call there).
int scratch_on_device(void** out, std::string& err) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    err = std::string("hipGetDevice: ") + hipGetErrorString(e);
    return TKFMI_EHIP;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return TKFMI_EHIP;
    }
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return TKFMI_OK;
}

// Sets the allocation-free mode for the duration of one call.
struct ActiveScope {
  ActiveScope() { g_active = true; }
  ~ActiveScope() { g_active = false; }
};

}  // namespace

bool tkfmi_noalloc_active() { return g_active; }

extern "C" int tkfmi_fb2d_batched_noalloc(const tkfmi_problem* prob, const tkfmi_outputs* out,
                                          void* stream, uint32_t flags) {
  tkfmi_clear_error();
  if (flags & (TKFMI_TIME_KERNELS | TKFMI_REUSE_DESCRIPTORS))
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_fb2d_batched_noalloc: TKFMI_TIME_KERNELS and "
                                    "TKFMI_REUSE_DESCRIPTORS are not available without a workspace");
  const size_t need = tkfmi_fb2d_workspace_bytes(prob, flags);
  if (need == 0)  // invalid problem: the regular entry point reports why, before any device work
    return tkfmi_fb2d_batched(prob, out, nullptr, 0, stream, flags);
  if (need > kScratchBytes)
    return tkfmi_fail(TKFMI_EINVAL, "tkfmi_fb2d_batched_noalloc: the batch needs " + std::to_string(need) +
                                        " bytes of scratch, more than the " + std::to_string(kScratchBytes) +
                                        " it holds; split the batch or call tkfmi_fb2d_batched with a workspace");
  void* ptr = nullptr;
  std::string err;
  const int rc = scratch_on_device(&ptr, err);
  if (rc != TKFMI_OK) return tkfmi_fail(rc, err);
  ActiveScope scope;
  return tkfmi_fb2d_batched(prob, out, ptr, kScratchBytes, stream, flags);
}
