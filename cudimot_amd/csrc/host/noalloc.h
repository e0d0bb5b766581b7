// Library-owned scratch of the allocation-free entry points (noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace cdm {

// Bytes of the static device scratch on each GPU (64 MiB).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;

// Device address of the current device's scratch, looked up with
// hipGetSymbolAddress on the first call on that device and kept.
hipError_t noalloc_scratch(void** out);

}  // namespace cdm
