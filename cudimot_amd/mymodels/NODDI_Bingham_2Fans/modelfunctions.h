// Bingham-NODDI with two fanning populations.
//   S = S0 [ (1-fiso) ((1-fFan2) F1 + fFan2 F2) + fiso exp(-b Diso) ]
//   Fk = (1-fintra_k) E_k + fintra_k I_k
// P: fiso, fFan2, fintra1, kappa1, beta1, th1, ph1, psi1,
//    fintra2, kappa2, beta2, th2, ph2, psi2.  CFP: g (3), b (1).  FixP: S0.
// Follows NODDI_Bingham_2Fans/modelfunctions.h:234-325 (all derivatives by
// central differences).  Engine cache hook: each fan's normaliser and
// hindered diffusivities are computed once per parameter vector.
#include "../NODDI_Bingham/bingham_math.h"

#define CDM_VOXEL_CACHE 8
#define CDM_VOXEL_CACHE_DEPENDS 0x71Cu  // fintra, kappa, beta of both fans

MACRO void Voxel_Cache(int npar, T* P, T* FixP, CDM_F64* vc) {
  bingham_fan f1, f2;
  bingham_fan_setup((CDM_F64)P[2], (CDM_F64)P[3], (CDM_F64)P[4], f1);
  bingham_fan_setup((CDM_F64)P[8], (CDM_F64)P[9], (CDM_F64)P[10], f2);
  vc[0] = f1.lnc4;
  vc[1] = f1.dpar;
  vc[2] = f1.dperp1;
  vc[3] = f1.dperp2;
  vc[4] = f2.lnc4;
  vc[5] = f2.dpar;
  vc[6] = f2.dperp1;
  vc[7] = f2.dperp2;
}

MACRO T Predicted_Signal_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc) {
  bingham_fan f1, f2;
  f1.lnc4 = vc[0];
  f1.dpar = vc[1];
  f1.dperp1 = vc[2];
  f1.dperp2 = vc[3];
  f2.lnc4 = vc[4];
  f2.dpar = vc[5];
  f2.dperp1 = vc[6];
  f2.dperp2 = vc[7];
  const T fan1 = bingham_fan_signal(P[2], P[3], P[4], P[5], P[6], P[7], f1, CFP);
  const T fan2 = bingham_fan_signal(P[8], P[9], P[10], P[11], P[12], P[13], f2, CFP);
  const T aniso = ((T)1 - P[1]) * fan1 + P[1] * fan2;
  const T iso = exp_gpu(-CFP[3] * (T)CDM_NODDI_DISO);
  return FixP[0] * (((T)1 - P[0]) * aniso + P[0] * iso);
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  CDM_F64 vc[8];
  Voxel_Cache(npar, P, FixP, vc);
  return Predicted_Signal_Cached(npar, P, CFP, FixP, vc, (const CDM_F64*)0);
}

MACRO bool ConstraintsMCMC(int npar, T* P) {
  if ((P[3] - (T)0.1) < P[4]) return false;
  if ((P[9] - (T)0.1) < P[10]) return false;
  return true;
}

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  for (int i = 0; i < 14; ++i) derivatives[i] = NUMERICAL(i);
}

MACRO void FixConstraintsLM(int npar, T* P) {
  if (P[3] < P[4]) P[4] = P[3] - (T)0.5;
  if (P[9] < P[10]) P[10] = P[9] - (T)0.5;
}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
