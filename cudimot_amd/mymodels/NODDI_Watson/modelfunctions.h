// NODDI with a Watson orientation distribution.
//   S = S0 [ (1-fiso) ((1-fintra) E + fintra I) + fiso exp(-b Diso) ]
// E: hindered compartment, exp(-b ((dpar_eq - dperp_eq) (g.n)^2 + dperp_eq))
//    with dperp = Dpar (1 - fintra) (tortuosity);
// I: intra-axonal sticks, 1/2 sum_l L_l(b Dpar) c_l(kappa) Y_2l(g.n),
//    floored at 1e-3 before the 1/2 when negative.
// P: fiso, fintra, kappa, th, ph.  CFP: g (3), b (1).  FixP: S0.
// Follows NODDI_Watson/modelfunctions.h:68-216.  The derivative in fintra
// is the reference's (it ignores the tortuosity dependence of E).
//
// Engine cache hooks: the hindered diffusivities and the SH coefficients
// depend on (fintra, kappa) only (Voxel_Cache); the Legendre-Gaussian
// integrals and the isotropic term depend on the b-value only (Meas_Cache).
#include "watson_math.h"

#define CDM_VOXEL_CACHE 9              // dpar_eq, dperp_eq, c_0..c_6
#define CDM_VOXEL_CACHE_DEPENDS 0x06u  // fintra, kappa
#define CDM_MEAS_CACHE 8               // L_0..L_6 (b Dpar), exp(-b Diso)
#define CDM_NUMERICAL_MASK 0x1Cu       // kappa, th, ph
#define CDM_SIGNAL_PARTIALS 1          // signal and analytic entries together

MACRO void Voxel_Cache(int npar, T* P, T* FixP, CDM_F64* vc) {
  noddi_watson_hindered((CDM_F64)P[2], CDM_NODDI_DPAR * (1.0 - (CDM_F64)P[1]), vc[0], vc[1]);
  noddi_watson_sh((CDM_F64)P[2], vc + 2);
}

MACRO void Meas_Cache(T* CFP, CDM_F64* mc) {
  noddi_lgi((CDM_F64)CFP[3] * CDM_NODDI_DPAR, mc);
  mc[7] = exp(-(CDM_F64)CFP[3] * CDM_NODDI_DISO);
}

MACRO void watson_parts(T* P, T* CFP, const CDM_F64* vc, const CDM_F64* mc, T& extra, T& intra) {
  T st, ct, sp, cp;
  tsincos(P[3], &st, &ct);
  tsincos(P[4], &sp, &cp);
  T gn = CFP[0] * st * cp + CFP[1] * st * sp + CFP[2] * ct;
#ifdef CDM_WATSON_REFERENCE_FORMS
  extra = exp_gpu(-CFP[3] * ((T)(vc[0] - vc[1]) * gn * gn + (T)vc[1]));
#else
  // the hardware 2^x form in single precision: its error is below the float
  // rounding the argument already carries (exact exp in double)
  extra = texp_fast(-CFP[3] * ((T)(vc[0] - vc[1]) * gn * gn + (T)vc[1]));
#endif
  if (gn > (T)1) gn = (T)1;
  if (gn < (T)-1) gn = (T)-1;
  T sh[7];
  noddi_even_sh(gn, sh);
  CDM_F64 s = 0.0;
  for (int l = 0; l < 7; ++l) s += mc[l] * vc[2 + l] * (CDM_F64)sh[l];
  if (s < 0) s = 1e-3;
  intra = (T)(0.5 * s);
}

MACRO T Predicted_Signal_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc) {
  T extra, intra;
  watson_parts(P, CFP, vc, mc, extra, intra);
  const T aniso = ((T)1 - P[1]) * extra + P[1] * intra;
  return FixP[0] * (((T)1 - P[0]) * aniso + P[0] * (T)mc[7]);
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  CDM_F64 vc[9], mc[8];
  Voxel_Cache(npar, P, FixP, vc);
  Meas_Cache(CFP, mc);
  return Predicted_Signal_Cached(npar, P, CFP, FixP, vc, mc);
}

MACRO bool ConstraintsMCMC(int npar, T* P) { return true; }

// Analytic entries (fiso, fintra); the engine forms the others.
MACRO void Partial_Derivatives_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc,
                                      T* derivatives) {
  T extra, intra;
  watson_parts(P, CFP, vc, mc, extra, intra);
  const T aniso = ((T)1 - P[1]) * extra + P[1] * intra;
  derivatives[0] = FixP[0] * ((T)mc[7] - aniso);
  derivatives[1] = FixP[0] * ((T)1 - P[0]) * (intra - extra);
}

// The signal of Predicted_Signal_Cached with the analytic entries of
// Partial_Derivatives_Cached, from one evaluation of the parts.
MACRO T Predicted_Signal_Partials_Cached(int npar, T* P, T* CFP, T* FixP, const CDM_F64* vc, const CDM_F64* mc,
                                         T* derivatives) {
  T extra, intra;
  watson_parts(P, CFP, vc, mc, extra, intra);
  const T aniso = ((T)1 - P[1]) * extra + P[1] * intra;
  derivatives[0] = FixP[0] * ((T)mc[7] - aniso);
  derivatives[1] = FixP[0] * ((T)1 - P[0]) * (intra - extra);
  return FixP[0] * (((T)1 - P[0]) * aniso + P[0] * (T)mc[7]);
}

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  CDM_F64 vc[9], mc[8];
  Voxel_Cache(npar, P, FixP, vc);
  Meas_Cache(CFP, mc);
  Partial_Derivatives_Cached(npar, P, CFP, FixP, vc, mc, derivatives);
  derivatives[2] = NUMERICAL(2);
  derivatives[3] = NUMERICAL(3);
  derivatives[4] = NUMERICAL(4);
}

MACRO void FixConstraintsLM(int npar, T* P) {}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
