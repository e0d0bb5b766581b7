// Bingham-NODDI with two fanning populations.
//   S = S0 [ (1-fiso) ((1-fFan2) F1 + fFan2 F2) + fiso exp(-b Diso) ]
//   Fk = (1-fintra_k) E_k + fintra_k I_k
// P: fiso, fFan2, fintra1, kappa1, beta1, th1, ph1, psi1,
//    fintra2, kappa2, beta2, th2, ph2, psi2.  CFP: g (3), b (1).  FixP: S0.
// Follows NODDI_Bingham_2Fans/modelfunctions.h:234-325: every derivative
// by central differences (the reference's NUMERICAL step rule), formed by
// the engine (CDM_NUMERICAL_MASK) from per-probe voxel caches.
//
// Engine cache hooks: per fan, the normaliser and hindered diffusivities
// (five 1F1 evaluations, the cache's jobs, depending on kappa and beta) and
// the local frame (depending on the three angles) are computed once per
// parameter vector instead of once per measurement; a step or probe in
// fintra or an angle recombines the kept 1F1 values.  MCMC keeps each fan's
// signal per measurement (Fk): a step in fiso or fFan2 recombines the kept
// signals, a step in one fan's parameters recomputes that fan's only.
#include "../NODDI_Bingham/bingham_math.h"

#define CDM_VOXEL_CACHE 26                // per fan: lnc4, dpar, dperp1, dperp2, frame (9)
#define CDM_VOXEL_CACHE_DEPENDS 0x3FFCu   // everything but fiso and fFan2
#define CDM_VOXEL_CACHE_JOBS 10           // per fan: 1F1(kappa, beta) and its four probes
#define CDM_VOXEL_CACHE_JOBS_DEPENDS 0x618u  // kappa1, beta1, kappa2, beta2
#define CDM_NUMERICAL_MASK 0x3FFFu        // all 14, as the reference
#define CDM_MEAS_CACHE 1                  // exp(-b Diso) per measurement
#define CDM_LM_JACOBIAN_COLUMNS 1         // 64-lane LM: the Jacobian column by column
#define CDM_MEAS_VOXEL_CACHE 2            // F1, F2 per measurement
#define CDM_MEAS_VOXEL_CACHE_DEPENDS 0x3FFCu
#define CDM_MEAS_VOXEL_CACHE_PARTIAL 1    // a step in one fan recomputes that fan
#define CDM_VOXEL_CACHE_PARTIAL 1         // ... and only that fan's cache entries and jobs

constexpr int b2f_fan = 13;  // cache values per fan

MACRO CDM_F64 Voxel_Cache_Job(int job, int npar, T* P, T* FixP) {
  return job < 5 ? bingham_fan_job(job, (CDM_F64)P[3], (CDM_F64)P[4])
                 : bingham_fan_job(job - 5, (CDM_F64)P[9], (CDM_F64)P[10]);
}

MACRO void Voxel_Cache_Combine(int npar, T* P, T* FixP, const CDM_F64* jobs, CDM_F64* vc) {
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int o = 6 * k;  // fan k's parameters start at P[2 + o]
    bingham_fan f;
    bingham_fan_combine((CDM_F64)P[2 + o], jobs + 5 * k, f);
    CDM_F64* c = vc + b2f_fan * k;
    c[0] = f.lnc4;
    c[1] = f.dpar;
    c[2] = f.dperp1;
    c[3] = f.dperp2;
    bingham_fan_frame((CDM_F64)P[5 + o], (CDM_F64)P[6 + o], (CDM_F64)P[7 + o], c + 4);
  }
}

// The jobs a step in `changed` moves: fan 1's five for kappa1, beta1, fan
// 2's for kappa2, beta2.
__device__ inline void Voxel_Cache_Job_Range(int changed, int& first, int& count) {
  first = changed >= 8 ? 5 : 0;
  count = (changed == 3 || changed == 4 || changed == 9 || changed == 10) ? 5 : 0;
}

// Voxel_Cache_Combine after a change of parameter `changed` alone: the
// moved fan's normaliser and diffusivities (fintra, kappa, beta) or its
// frame (an angle); every other entry copied from cur.
MACRO void Voxel_Cache_Update(int npar, T* P, T* FixP, const CDM_F64* jobs, const CDM_F64* cur, CDM_F64* vc,
                              int changed) {
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int o = 6 * k;
    CDM_F64* c = vc + b2f_fan * k;
    const CDM_F64* q = cur + b2f_fan * k;
    if (changed >= 2 + o && changed < 5 + o) {
      bingham_fan f;
      bingham_fan_combine((CDM_F64)P[2 + o], jobs + 5 * k, f);
      c[0] = f.lnc4;
      c[1] = f.dpar;
      c[2] = f.dperp1;
      c[3] = f.dperp2;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) c[j] = q[j];
    }
    if (changed >= 5 + o && changed < 8 + o) {
      bingham_fan_frame((CDM_F64)P[5 + o], (CDM_F64)P[6 + o], (CDM_F64)P[7 + o], c + 4);
    } else {
#pragma unroll
      for (int j = 4; j < b2f_fan; ++j) c[j] = q[j];
    }
  }
}

MACRO void Voxel_Cache(int npar, T* P, T* FixP, CDM_F64* vc) {
  CDM_F64 jobs[10];
  for (int j = 0; j < 10; ++j) jobs[j] = Voxel_Cache_Job(j, npar, P, FixP);
  Voxel_Cache_Combine(npar, P, FixP, jobs, vc);
}

MACRO void Meas_Cache(T* CFP, CDM_F64* mc) { mc[0] = exp(-(CDM_F64)CFP[3] * CDM_NODDI_DISO); }

// Signal of fan k (0, 1) from its cache values.
MACRO T b2f_fan_signal(int k, T* P, T* CFP, const CDM_F64* vc) {
  const CDM_F64* c = vc + b2f_fan * k;
  bingham_fan f;
  f.lnc4 = c[0];
  f.dpar = c[1];
  f.dperp1 = c[2];
  f.dperp2 = c[3];
  T gx, gy, gn;
  bingham_frame_local(c + 4, CFP, gx, gy, gn);
  const int o = 6 * k;
  const T e = bingham_fan_extra_local(gx, gy, gn, CFP[3], f);
  const T i = bingham_fan_intra_local(P[3 + o], P[4 + o], gx, gy, gn, CFP[3], f);
  return ((T)1 - P[2 + o]) * e + P[2 + o] * i;
}

MACRO T Predicted_Signal_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc) {
  const T fan1 = b2f_fan_signal(0, P, CFP, vc);
  const T fan2 = b2f_fan_signal(1, P, CFP, vc);
  const T aniso = ((T)1 - P[1]) * fan1 + P[1] * fan2;
  return FixP[0] * (((T)1 - P[0]) * aniso + P[0] * (T)mc[0]);
}

// The signal from the kept fan signals: the arithmetic of
// Predicted_Signal_Cached after its two fan evaluations.
MACRO T Predicted_Signal_MV(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc, const T* mvc) {
  const T fan1 = mvc[0];
  const T fan2 = mvc[1];
  const T aniso = ((T)1 - P[1]) * fan1 + P[1] * fan2;
  return FixP[0] * (((T)1 - P[0]) * aniso + P[0] * (T)mc[0]);
}

MACRO void Meas_Voxel_Cache(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc, T* mvc) {
  mvc[0] = b2f_fan_signal(0, P, CFP, vc);
  mvc[1] = b2f_fan_signal(1, P, CFP, vc);
}

// After a change of parameter `changed` alone: fan 1's parameters are
// P[2..7], fan 2's P[8..13]; fiso and fFan2 move neither fan's signal.
MACRO T Predicted_Signal_MV_Update(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc,
                                   const T* cur, T* mvc, int changed) {
  const bool all = changed < 0;
  if (all || (changed >= 2 && changed < 8))
    mvc[0] = b2f_fan_signal(0, P, CFP, vc);
  else
    mvc[0] = cur[0];
  if (all || changed >= 8)
    mvc[1] = b2f_fan_signal(1, P, CFP, vc);
  else
    mvc[1] = cur[1];
  return Predicted_Signal_MV(npar, P, CFP, FixP, vc, mc, mvc);
}

// No analytic entries: the engine forms all 14 by central differences.
MACRO void Partial_Derivatives_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc,
                                      T* derivatives) {}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  CDM_F64 vc[2 * b2f_fan], mc[1];
  Voxel_Cache(npar, P, FixP, vc);
  Meas_Cache(CFP, mc);
  return Predicted_Signal_Cached(npar, P, CFP, FixP, vc, mc);
}

MACRO bool ConstraintsMCMC(int npar, T* P) {
  if ((P[3] - (T)0.1) < P[4]) return false;
  if ((P[9] - (T)0.1) < P[10]) return false;
  return true;
}

// The reference's form, all 14 by central differences (used without the
// engine's central differences, e.g. by the CPU oracle).
MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  for (int i = 0; i < 14; ++i) derivatives[i] = NUMERICAL(i);
}

MACRO void FixConstraintsLM(int npar, T* P) {
  if (P[3] < P[4]) P[4] = P[3] - (T)0.5;
  if (P[9] < P[10]) P[10] = P[9] - (T)0.5;
}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
