// Bingham-NODDI: one fanning fibre population.
//   S = S0 [ (1-fiso) ((1-fintra) E + fintra I) + fiso exp(-b Diso) ]
// P: fiso, fintra, kappa, beta, th, ph, psi (kappa = kappa1,
// beta = kappa1 - kappa2).  CFP: g (3), b (1).  FixP: S0.
// Follows NODDI_Bingham/modelfunctions.h:37-246: analytic derivatives in
// fiso and fintra (the latter ignoring the tortuosity dependence of E, as
// the reference), central differences for the rest.
//
// Engine cache hooks: the normaliser 1F1(kappa, beta, 0) and the hindered
// diffusivities (five 1F1 evaluations) depend on fintra, kappa and beta
// only, so they are computed once per parameter vector (Voxel_Cache) instead
// of once per measurement; the intra-axonal signal of each measurement does
// not depend on fiso or fintra, so an MCMC step in those two reuses it
// (Meas_Voxel_Cache, Predicted_Signal_MV; Predicted_Signal_MV_Fill does both
// from one evaluation of the local gradient); the five 1F1 evaluations
// depend on kappa and beta only, so an MCMC step in fintra recombines them.
#include "bingham_math.h"

#define CDM_VOXEL_CACHE 4
#define CDM_VOXEL_CACHE_DEPENDS 0x0Eu       // fintra, kappa, beta
#define CDM_VOXEL_CACHE_JOBS 5              // its five 1F1 evaluations
#define CDM_VOXEL_CACHE_JOBS_DEPENDS 0x0Cu  // kappa, beta
#define CDM_NUMERICAL_MASK 0x7Cu            // kappa, beta, th, ph, psi
#define CDM_MEAS_VOXEL_CACHE 1              // intra-axonal signal per measurement
#define CDM_MEAS_VOXEL_CACHE_DEPENDS 0x7Cu  // kappa, beta, th, ph, psi
#define CDM_MEAS_VOXEL_CACHE_FILL 1         // cache row and signal together
#define CDM_SIGNAL_PARTIALS 1               // signal and analytic entries together
#define CDM_MEAS_CACHE 1                    // exp(-b Diso) per measurement
#define CDM_LM_ROW_PASSES 2                 // 64-lane LM: the Jacobian rows in two passes

MACRO void Voxel_Cache(int npar, T* P, T* FixP, CDM_F64* vc) {
  bingham_fan f;
  bingham_fan_setup((CDM_F64)P[1], (CDM_F64)P[2], (CDM_F64)P[3], f);
  vc[0] = f.lnc4;
  vc[1] = f.dpar;
  vc[2] = f.dperp1;
  vc[3] = f.dperp2;
}

// The isotropic compartment depends on the b-value only.
MACRO void Meas_Cache(T* CFP, CDM_F64* mc) { mc[0] = exp(-(CDM_F64)CFP[3] * CDM_NODDI_DISO); }

MACRO CDM_F64 Voxel_Cache_Job(int job, int npar, T* P, T* FixP) {
  return bingham_fan_job(job, (CDM_F64)P[2], (CDM_F64)P[3]);
}

MACRO void Voxel_Cache_Combine(int npar, T* P, T* FixP, const CDM_F64* jobs, CDM_F64* vc) {
  bingham_fan f;
  bingham_fan_combine((CDM_F64)P[1], jobs, f);
  vc[0] = f.lnc4;
  vc[1] = f.dpar;
  vc[2] = f.dperp1;
  vc[3] = f.dperp2;
}

MACRO void bingham1_parts(T* P, T* CFP, const CDM_F64* vc, T& fan, T& e, T& i) {
  bingham_fan f;
  f.lnc4 = vc[0];
  f.dpar = vc[1];
  f.dperp1 = vc[2];
  f.dperp2 = vc[3];
  bingham_fan_signal_parts(P[2], P[3], P[4], P[5], P[6], f, CFP, e, i);
  fan = ((T)1 - P[1]) * e + P[1] * i;
}

MACRO T Predicted_Signal_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc) {
  T fan, e, i;
  bingham1_parts(P, CFP, vc, fan, e, i);
  const T iso = (T)mc[0];
  return FixP[0] * (((T)1 - P[0]) * fan + P[0] * iso);
}

MACRO void Meas_Voxel_Cache(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc, T* mvc) {
  bingham_fan f;
  f.lnc4 = vc[0];
  f.dpar = vc[1];
  f.dperp1 = vc[2];
  f.dperp2 = vc[3];
  mvc[0] = bingham_fan_intra(P[2], P[3], P[4], P[5], P[6], f, CFP);
}

MACRO T Predicted_Signal_MV(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc, const T* mvc) {
  bingham_fan f;
  f.lnc4 = vc[0];
  f.dpar = vc[1];
  f.dperp1 = vc[2];
  f.dperp2 = vc[3];
  const T e = bingham_fan_extra(P[4], P[5], P[6], f, CFP);
  const T fan = ((T)1 - P[1]) * e + P[1] * mvc[0];
  const T iso = (T)mc[0];
  return FixP[0] * (((T)1 - P[0]) * fan + P[0] * iso);
}

// Meas_Voxel_Cache and Predicted_Signal_MV from one local gradient.
MACRO T Predicted_Signal_MV_Fill(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc, T* mvc) {
  T fan, e, i;
  bingham1_parts(P, CFP, vc, fan, e, i);
  mvc[0] = i;
  const T iso = (T)mc[0];
  return FixP[0] * (((T)1 - P[0]) * fan + P[0] * iso);
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  CDM_F64 vc[4], mc[1];
  Voxel_Cache(npar, P, FixP, vc);
  Meas_Cache(CFP, mc);
  return Predicted_Signal_Cached(npar, P, CFP, FixP, vc, mc);
}

// kappa must exceed beta by 0.05 (NODDI_Bingham/modelfunctions.h:169-178).
MACRO bool ConstraintsMCMC(int npar, T* P) { return !((P[2] - (T)0.05) < P[3]); }

// Analytic entries (fiso, fintra); the engine forms the others.
MACRO void Partial_Derivatives_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc,
                                      T* derivatives) {
  T fan, e, i;
  bingham1_parts(P, CFP, vc, fan, e, i);
  const T iso = (T)mc[0];
  derivatives[0] = FixP[0] * (iso - fan);
  derivatives[1] = FixP[0] * ((T)1 - P[0]) * (i - e);
}

// The signal of Predicted_Signal_Cached with the analytic entries of
// Partial_Derivatives_Cached, from one evaluation of the fan's parts.
MACRO T Predicted_Signal_Partials_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc,
                                         T* derivatives) {
  T fan, e, i;
  bingham1_parts(P, CFP, vc, fan, e, i);
  const T iso = (T)mc[0];
  derivatives[0] = FixP[0] * (iso - fan);
  derivatives[1] = FixP[0] * ((T)1 - P[0]) * (i - e);
  return FixP[0] * (((T)1 - P[0]) * fan + P[0] * iso);
}

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  CDM_F64 vc[4], mc[1];
  Voxel_Cache(npar, P, FixP, vc);
  Meas_Cache(CFP, mc);
  Partial_Derivatives_Cached(npar, P, CFP, FixP, vc, mc, derivatives);
  derivatives[2] = NUMERICAL(2);
  derivatives[3] = NUMERICAL(3);
  derivatives[4] = NUMERICAL(4);
  derivatives[5] = NUMERICAL(5);
  derivatives[6] = NUMERICAL(6);
}

// After LM: beta may not exceed kappa (NODDI_Bingham/modelfunctions.h:228-236).
MACRO void FixConstraintsLM(int npar, T* P) {
  if (P[2] < P[3]) P[3] = P[2] - (T)0.05;
}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
