// Device fills the C-ABI issues as kernels, so that launch sequences recorded
// into a HIP graph and replayed reset their counters themselves (a memset
// node in a replayed graph is avoided).
#pragma once
#include <hip/hip_runtime.h>

namespace cdm {

// p[0..n) = 0 on stream s.
hipError_t zero_ints(int* p, int n, hipStream_t s);

}  // namespace cdm
