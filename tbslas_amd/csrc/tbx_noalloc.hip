// Device scratch of the allocation-free step entry point
// (tbx_semilag_insitu_noalloc): a file-scope device array of a fixed size,
// one copy per GPU, alive until the process ends.  Its host-side name is a
// stand-in and never a device address: the address each device holds it at
// is read once per device with hipGetSymbolAddress, with that device
// current, and only that address is handed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tbx_internal.h"

namespace tbx {
namespace {

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};

}  // namespace

hipError_t noalloc_scratch(void **ptr, size_t *bytes) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_ptr[dev] = p;
  }
  *ptr = g_ptr[dev];
  *bytes = kNoAllocScratchBytes;
  return hipSuccess;
}

}  // namespace tbx
