// Bingham-NODDI with two fanning populations.
//   S = S0 [ (1-fiso) ((1-fFan2) F1 + fFan2 F2) + fiso exp(-b Diso) ]
//   Fk = (1-fintra_k) E_k + fintra_k I_k
// P: fiso, fFan2, fintra1, kappa1, beta1, th1, ph1, psi1,
//    fintra2, kappa2, beta2, th2, ph2, psi2.  CFP: g (3), b (1).  FixP: S0.
// Follows NODDI_Bingham_2Fans/modelfunctions.h:234-325 (all derivatives by
// central differences).
#include "../NODDI_Bingham/bingham_math.h"

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  bingham_fan f1, f2;
  bingham_fan_setup((CDM_F64)P[2], (CDM_F64)P[3], (CDM_F64)P[4], f1);
  bingham_fan_setup((CDM_F64)P[8], (CDM_F64)P[9], (CDM_F64)P[10], f2);
  const T fan1 = bingham_fan_signal(P[2], P[3], P[4], P[5], P[6], P[7], f1, CFP);
  const T fan2 = bingham_fan_signal(P[8], P[9], P[10], P[11], P[12], P[13], f2, CFP);
  const T aniso = ((T)1 - P[1]) * fan1 + P[1] * fan2;
  const T iso = exp_gpu(-CFP[3] * (T)CDM_NODDI_DISO);
  return FixP[0] * (((T)1 - P[0]) * aniso + P[0] * iso);
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
