// ORACLE (test infrastructure only).  Model-independent C entry points.
#include <omp.h>
#include <stdint.h>
#include "philox_ref.h"

extern "C" {
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  orc_u32x4 c;
  for (int i = 0; i < 4; ++i) c.v[i] = ctr[i];
  const orc_u32x4 r = orc_philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = r.v[i];
}

void orc_draws(unsigned long long seed, unsigned long long voxel, uint32_t jump, uint32_t slot, double* normal,
               double* uniform) {
  orc_mcmc_draws(seed, voxel, jump, slot, normal, uniform);
}

// Threads an OpenMP parallel region of the oracle runs on (the team size
// actually formed, not the requested maximum).
int orc_omp_threads(void) {
  int n = 1;
#pragma omp parallel
  {
#pragma omp single
    n = omp_get_num_threads();
  }
  return n;
}
}
