// Two-parameter test model: S = a exp(-b x).  P: a, b.  CFP: x (1).
// Follows toymodel/modelfunctions.h:13-50.

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) { return P[0] * exp_gpu(-P[1] * CFP[0]); }

MACRO bool ConstraintsMCMC(int npar, T* P) { return true; }

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  const T e = exp_gpu(-P[1] * CFP[0]);
  derivatives[0] = e;
  derivatives[1] = -P[0] * CFP[0] * e;
}

MACRO void FixConstraintsLM(int npar, T* P) {}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
