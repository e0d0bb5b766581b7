// NODDI with a Watson orientation distribution.
//   S = S0 [ (1-fiso) ((1-fintra) E + fintra I) + fiso exp(-b Diso) ]
// E: hindered compartment, exp(-b ((dpar_eq - dperp_eq) (g.n)^2 + dperp_eq))
//    with dperp = Dpar (1 - fintra) (tortuosity);
// I: intra-axonal sticks, 1/2 sum_l L_l(b Dpar) c_l(kappa) Y_2l(g.n),
//    floored at 1e-3 before the 1/2 when negative.
// P: fiso, fintra, kappa, th, ph.  CFP: g (3), b (1).  FixP: S0.
// Follows NODDI_Watson/modelfunctions.h:68-216.  The derivative in fintra
// is the reference's (it ignores the tortuosity dependence of E).
#include "watson_math.h"

MACRO void watson_compartments(T* P, T* CFP, T& extra, T& intra) {
  T st, ct, sp, cp;
  tsincos(P[3], &st, &ct);
  tsincos(P[4], &sp, &cp);
  T gn = CFP[0] * st * cp + CFP[1] * st * sp + CFP[2] * ct;
  // per-voxel quantities in double
  CDM_F64 dpar_eq, dperp_eq, coef[7], lgi[7];
  noddi_watson_hindered((CDM_F64)P[2], CDM_NODDI_DPAR * (1.0 - (CDM_F64)P[1]), dpar_eq, dperp_eq);
  noddi_watson_sh((CDM_F64)P[2], coef);
  noddi_lgi((CDM_F64)CFP[3] * CDM_NODDI_DPAR, lgi);
  extra = exp_gpu(-CFP[3] * ((T)(dpar_eq - dperp_eq) * gn * gn + (T)dperp_eq));
  if (gn > (T)1) gn = (T)1;
  if (gn < (T)-1) gn = (T)-1;
  T sh[7];
  noddi_even_sh(gn, sh);
  CDM_F64 s = 0.0;
  for (int l = 0; l < 7; ++l) s += lgi[l] * coef[l] * (CDM_F64)sh[l];
  if (s < 0) s = 1e-3;
  intra = (T)(0.5 * s);
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  T extra, intra;
  watson_compartments(P, CFP, extra, intra);
  const T iso = exp_gpu(-CFP[3] * (T)CDM_NODDI_DISO);
  const T aniso = ((T)1 - P[1]) * extra + P[1] * intra;
  return FixP[0] * (((T)1 - P[0]) * aniso + P[0] * iso);
}

MACRO bool ConstraintsMCMC(int npar, T* P) { return true; }

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  T extra, intra;
  watson_compartments(P, CFP, extra, intra);
  const T iso = exp_gpu(-CFP[3] * (T)CDM_NODDI_DISO);
  const T aniso = ((T)1 - P[1]) * extra + P[1] * intra;
  derivatives[0] = FixP[0] * (iso - aniso);
  derivatives[1] = FixP[0] * ((T)1 - P[0]) * (intra - extra);
  derivatives[2] = NUMERICAL(2);
  derivatives[3] = NUMERICAL(3);
  derivatives[4] = NUMERICAL(4);
}

MACRO void FixConstraintsLM(int npar, T* P) {}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
