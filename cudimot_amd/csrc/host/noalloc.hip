// Scratch of the allocation-free entry point (cdm_mcmc_noalloc): a
// file-scope device array, one copy per GPU, held until the process ends.
// Host code never passes the array's name to a kernel or a copy (on the host
// it is only a stand-in for the symbol); it looks up each device's address
// once, with that device current, and uses that pointer.
#include <hip/hip_runtime.h>

#include <mutex>

#include "noalloc.h"

static __device__ __align__(256) unsigned char scratch[cdm::kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_ptr[kMaxDevices] = {};

}  // namespace

namespace cdm {

hipError_t noalloc_scratch(void** out) {
  *out = nullptr;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

}  // namespace cdm
