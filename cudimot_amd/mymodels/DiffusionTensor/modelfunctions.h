// Diffusion tensor.
//   S = S0 exp(-b g^T D g),  D = [[Dxx Dxy Dxz] [Dxy Dyy Dyz] [Dxz Dyz Dzz]]
// P: S0, Dxx, Dxy, Dxz, Dyy, Dyz, Dzz.  CFP: g (3), b (1).
// Follows DiffusionTensor/modelfunctions.h:20-114.
// The compartment exponentials use texp_fast (in single precision the
// hardware 2^y of y = x log2(e): relative error about |x| 2^-24 + 1 ulp,
// below what the float rounding of the argument b g^T D g itself
// brings); double-precision evaluations use the exact exp.

MACRO void dti_design(const T* CFP, T* q) {
  // q[k] = d(b g^T D g)/dD_k for the six tensor elements in parameter order.
  const T b = CFP[3];
  q[0] = b * CFP[0] * CFP[0];
  q[1] = (T)2 * b * CFP[0] * CFP[1];
  q[2] = (T)2 * b * CFP[0] * CFP[2];
  q[3] = b * CFP[1] * CFP[1];
  q[4] = (T)2 * b * CFP[1] * CFP[2];
  q[5] = b * CFP[2] * CFP[2];
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  T q[6];
  dti_design(CFP, q);
  T adc = (T)0;
#pragma unroll
  for (int k = 0; k < 6; ++k) adc += q[k] * P[1 + k];
  return P[0] * texp_fast(-adc);
}

MACRO bool ConstraintsMCMC(int npar, T* P) { return true; }

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  T q[6];
  dti_design(CFP, q);
  T adc = (T)0;
#pragma unroll
  for (int k = 0; k < 6; ++k) adc += q[k] * P[1 + k];
  const T e = texp_fast(-adc);
  derivatives[0] = e;
#pragma unroll
  for (int k = 0; k < 6; ++k) derivatives[1 + k] = -P[0] * q[k] * e;
}

MACRO void FixConstraintsLM(int npar, T* P) {}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
