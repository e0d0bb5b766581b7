// Ball and two sticks.
//   S = S0 * [ (1-f1-f2) exp(-b d) + f1 exp(-b d (g.n1)^2) + f2 exp(-b d (g.n2)^2) ]
// P: S0, d, f1, th1, ph1, f2, th2, ph2.  CFP: g (3), b (1).
// Follows Ball_2_Sticks/modelfunctions.h:7-123.
//
// Reference parity: the reference's dS/dph2 reuses the th2 chain factor
// (Ball_2_Sticks/modelfunctions.h:99-103), so its Levenberg-Marquardt moves
// ph2 along the th2 derivative.  That behaviour is kept by default so fits
// reproduce the reference; define CDM_B2S_EXACT_PHI2 to use the true
// derivative.
// The compartment exponentials use texp_fast (in single precision the
// hardware 2^y of y = x log2(e): relative error about |x| 2^-24 + 1 ulp,
// below what the float rounding of the argument b d (g.n)^2 itself
// brings); double-precision evaluations use the exact exp.

MACRO void b2s_stick(const T* CFP, T th, T ph, T& gn, T& dth, T& dph) {
  T st, ct, sp, cp;
  tsincos(th, &st, &ct);
  tsincos(ph, &sp, &cp);
  const T gxy = CFP[0] * cp + CFP[1] * sp;
  gn = st * gxy + CFP[2] * ct;
  dth = ct * gxy - st * CFP[2];
  dph = st * (CFP[1] * cp - CFP[0] * sp);
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  T g1, g2, u, v;
  b2s_stick(CFP, P[3], P[4], g1, u, v);
  b2s_stick(CFP, P[6], P[7], g2, u, v);
  const T bd = CFP[3] * P[1];
  const T ball = texp_fast(-bd);
  const T s1 = texp_fast(-bd * g1 * g1);
  const T s2 = texp_fast(-bd * g2 * g2);
  return P[0] * (((T)1 - P[2] - P[5]) * ball + P[2] * s1 + P[5] * s2);
}

MACRO bool ConstraintsMCMC(int npar, T* P) { return !(P[2] < P[5]); }

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  T g1, d1th, d1ph, g2, d2th, d2ph;
  b2s_stick(CFP, P[3], P[4], g1, d1th, d1ph);
  b2s_stick(CFP, P[6], P[7], g2, d2th, d2ph);
  const T bd = CFP[3] * P[1];
  const T ball = texp_fast(-bd);
  const T s1 = texp_fast(-bd * g1 * g1);
  const T s2 = texp_fast(-bd * g2 * g2);
  const T fb = (T)1 - P[2] - P[5];
  derivatives[0] = fb * ball + P[2] * s1 + P[5] * s2;
  derivatives[1] = -P[0] * CFP[3] * (fb * ball + P[2] * g1 * g1 * s1 + P[5] * g2 * g2 * s2);
  derivatives[2] = P[0] * (s1 - ball);
  const T c1 = -(T)2 * P[0] * bd * P[2] * g1 * s1;
  derivatives[3] = c1 * d1th;
  derivatives[4] = c1 * d1ph;
  derivatives[5] = P[0] * (s2 - ball);
  const T c2 = -(T)2 * P[0] * bd * P[5] * g2 * s2;
  derivatives[6] = c2 * d2th;
#ifdef CDM_B2S_EXACT_PHI2
  derivatives[7] = c2 * d2ph;
#else
  derivatives[7] = c2 * d2th;
#endif
}

// After LM: keep the larger stick first (swap f, th, ph of the two sticks).
MACRO void FixConstraintsLM(int npar, T* P) {
  if (P[2] < P[5]) {
    for (int k = 0; k < 3; ++k) {
      const T t = P[2 + k];
      P[2 + k] = P[5 + k];
      P[5 + k] = t;
    }
  }
}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
