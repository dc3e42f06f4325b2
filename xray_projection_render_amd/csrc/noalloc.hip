// noalloc.hip - the fixed device scratch region of xrp_render_device_noalloc.
//
// That entry point renders without allocating device or pinned host memory:
// its per-call data (stat counters, node table, deformations, views and the
// sample table) goes into this file-scope array instead of buffers the
// library allocates.  Each GPU holds its own copy of the array, at its own
// address, until the process ends; host code never uses the array's name as
// an address, only the pointer hipGetSymbolAddress returns for the device
// that was current when it was looked up.
#include <hip/hip_runtime.h>

#include <mutex>

#include "march.hpp"

namespace xrp {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
}  // namespace

int noalloc_scratch(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int current = -1;
    hipError_t e = hipGetDevice(&current);
    if (e != hipSuccess) return (int)e;
    if (current != device) return (int)hipErrorInvalidDevice;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return (int)hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}

}  // namespace xrp
