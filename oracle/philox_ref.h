/* ORACLE (test infrastructure only).
 *
 * Philox4x32-10 counter-based generator, written from Salmon, Moraes, Dror &
 * Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11 (Sec. 4.2):
 * ten rounds of two 32x32->64 multiplies with multipliers 0xD2511F53 and
 * 0xCD9E8D57, and a Weyl key schedule with increments 0x9E3779B9 and
 * 0xBB67AE85.  Used by the CPU restatement of the MCMC sampler so that it
 * draws the same numbers as the GPU sampler (the reference draws from
 * cuRAND XORWOW, MCMC.cu:75,323,687, which cannot be reproduced here).
 */
#pragma once
#include <stdint.h>
#include <math.h>

typedef struct { uint32_t v[4]; } orc_u32x4;

static inline orc_u32x4 orc_philox4x32_10(orc_u32x4 ctr, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * ctr.v[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr.v[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    orc_u32x4 n;
    n.v[0] = hi1 ^ ctr.v[1] ^ k0;
    n.v[1] = lo1;
    n.v[2] = hi0 ^ ctr.v[3] ^ k1;
    n.v[3] = lo0;
    ctr = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return ctr;
}

/* Draw set for one (voxel, jump, slot) of the sampler: words 0,1 feed a
 * Box-Muller standard normal, word 2 the Metropolis uniform. */
static inline double orc_u01(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

static inline void orc_mcmc_draws(uint64_t seed, uint64_t voxel, uint32_t jump, uint32_t slot,
                                  double* normal, double* uniform) {
  orc_u32x4 c;
  c.v[0] = (uint32_t)voxel;
  c.v[1] = (uint32_t)(voxel >> 32);
  c.v[2] = jump;
  c.v[3] = slot;
  const orc_u32x4 w = orc_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u1 = orc_u01(w.v[0]), u2 = orc_u01(w.v[1]);
  *normal = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  *uniform = orc_u01(w.v[2]);
}
