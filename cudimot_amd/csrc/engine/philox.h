// Counter-based random numbers for the MCMC sampler.
//
// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11): ten rounds of two
// 32x32->64-bit multiplies (0xD2511F53, 0xCD9E8D57) with a Weyl key schedule
// (0x9E3779B9, 0xBB67AE85).  Every draw of the sampler is addressed by a
// counter (voxel, jump, slot) under a 64-bit seed, so draws are independent
// of launch shape, voxel-group size, sub-volume partitioning or jump-block
// splitting, and a CPU implementation reproduces them exactly.
//   slot p < NPARAMS : proposal normal for parameter p, Metropolis uniform
//   slot NPARAMS     : the same pair for the Rician precision tau
// Words 0,1 feed a Box-Muller normal (computed in double), word 2 the
// uniform on (0,1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cdm {

struct U32x4 {
  uint32_t v[4];
};

__host__ __device__ __forceinline__ U32x4 philox4x32_10(U32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t a = (uint64_t)0xD2511F53u * c.v[0];
    const uint64_t b = (uint64_t)0xCD9E8D57u * c.v[2];
    U32x4 n;
    n.v[0] = (uint32_t)(b >> 32) ^ c.v[1] ^ k0;
    n.v[1] = (uint32_t)b;
    n.v[2] = (uint32_t)(a >> 32) ^ c.v[3] ^ k1;
    n.v[3] = (uint32_t)a;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

__host__ __device__ __forceinline__ double u01_open(uint32_t w) {
  return ((double)w + 0.5) * 2.3283064365386963e-10;  // (w + 1/2) / 2^32
}

struct McmcDraw {
  double normal, uniform;
};

__host__ __device__ __forceinline__ McmcDraw mcmc_draw(uint64_t seed, uint64_t voxel, uint32_t jump,
                                                       uint32_t slot) {
  U32x4 c;
  c.v[0] = (uint32_t)voxel;
  c.v[1] = (uint32_t)(voxel >> 32);
  c.v[2] = jump;
  c.v[3] = slot;
  const U32x4 w = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  McmcDraw d;
  d.normal = sqrt(-2.0 * log(u01_open(w.v[0]))) * cos(6.283185307179586 * u01_open(w.v[1]));
  d.uniform = u01_open(w.v[2]);
  return d;
}

}  // namespace cdm
