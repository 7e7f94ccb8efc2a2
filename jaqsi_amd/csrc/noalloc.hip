// Fixed device scratch for the entry point that allocates nothing
// (jq_sv_forward_noalloc): callers under HIP graph capture or in serving
// loops that must not call hipMalloc.  The array lives in the code object,
// one copy per device, for the life of the process.
#include <hip/hip_runtime.h>

#include <mutex>

#include "jaqsi_amd.h"
#include "noalloc.h"

namespace jq {

namespace {
static __device__ __align__(256) unsigned char scratch[JQ_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(void** out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[dev] == nullptr) {
    // the host-side name of a __device__ array is only a stand-in: its
    // device address comes from the runtime, once per device (this device
    // is current)
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

}  // namespace jq
