// Bingham-NODDI: one fanning fibre population.
//   S = S0 [ (1-fiso) ((1-fintra) E + fintra I) + fiso exp(-b Diso) ]
// P: fiso, fintra, kappa, beta, th, ph, psi (kappa = kappa1,
// beta = kappa1 - kappa2).  CFP: g (3), b (1).  FixP: S0.
// Follows NODDI_Bingham/modelfunctions.h:37-246: analytic derivatives in
// fiso and fintra (the latter ignoring the tortuosity dependence of E, as
// the reference), central differences for the rest.
#include "bingham_math.h"

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  bingham_fan f;
  bingham_fan_setup((CDM_F64)P[1], (CDM_F64)P[2], (CDM_F64)P[3], f);
  const T fan = bingham_fan_signal(P[1], P[2], P[3], P[4], P[5], P[6], f, CFP);
  const T iso = exp_gpu(-CFP[3] * (T)CDM_NODDI_DISO);
  return FixP[0] * (((T)1 - P[0]) * fan + P[0] * iso);
}

// kappa must exceed beta by 0.05 (NODDI_Bingham/modelfunctions.h:169-178).
MACRO bool ConstraintsMCMC(int npar, T* P) { return !((P[2] - (T)0.05) < P[3]); }

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  bingham_fan f;
  bingham_fan_setup((CDM_F64)P[1], (CDM_F64)P[2], (CDM_F64)P[3], f);
  const T fan = bingham_fan_signal(P[1], P[2], P[3], P[4], P[5], P[6], f, CFP);
  // compartments for the fintra derivative: fan = (1-fintra) E + fintra I
  const T e = bingham_fan_signal((T)0, P[2], P[3], P[4], P[5], P[6], f, CFP);
  const T i = bingham_fan_signal((T)1, P[2], P[3], P[4], P[5], P[6], f, CFP);
  const T iso = exp_gpu(-CFP[3] * (T)CDM_NODDI_DISO);
  derivatives[0] = FixP[0] * (iso - fan);
  derivatives[1] = FixP[0] * ((T)1 - P[0]) * (i - e);
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
