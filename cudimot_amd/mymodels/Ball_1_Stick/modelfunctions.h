// Ball and one stick.
//   S = S0 * [ (1-f) exp(-b d) + f exp(-b d (g.n)^2) ]
//   n = (sin th cos ph, sin th sin ph, cos th)
// P: S0, d, f, th, ph.  CFP: g (3), b (1).
// Forward model and analytic Jacobian follow Ball_1_Stick/modelfunctions.h:11-77.
// The compartment exponentials use texp_fast (in single precision the
// hardware 2^y of y = x log2(e): relative error about |x| 2^-24 + 1 ulp,
// below what the float rounding of the argument b d (g.n)^2 itself
// brings); double-precision evaluations use the exact exp.

MACRO void b1s_geometry(const T* P, const T* CFP, T& gn, T& dgn_dth, T& dgn_dph) {
  T st, ct, sp, cp;
  tsincos(P[3], &st, &ct);
  tsincos(P[4], &sp, &cp);
  const T gxy = CFP[0] * cp + CFP[1] * sp;  // g projected on the stick azimuth
  gn = st * gxy + CFP[2] * ct;
  dgn_dth = ct * gxy - st * CFP[2];
  dgn_dph = st * (CFP[1] * cp - CFP[0] * sp);
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  T gn, a, b;
  b1s_geometry(P, CFP, gn, a, b);
  const T bd = CFP[3] * P[1];
  const T ball = texp_fast(-bd);
  const T stick = texp_fast(-bd * gn * gn);
  return P[0] * (((T)1 - P[2]) * ball + P[2] * stick);
}

MACRO bool ConstraintsMCMC(int npar, T* P) { return true; }

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  T gn, dth, dph;
  b1s_geometry(P, CFP, gn, dth, dph);
  const T bd = CFP[3] * P[1];
  const T ball = texp_fast(-bd);
  const T stick = texp_fast(-bd * gn * gn);
  const T f = P[2];
  derivatives[0] = ((T)1 - f) * ball + f * stick;
  derivatives[1] = -P[0] * CFP[3] * (((T)1 - f) * ball + f * gn * gn * stick);
  derivatives[2] = P[0] * (stick - ball);
  const T common = -(T)2 * P[0] * bd * f * gn * stick;  // dS/d(gn) / ... chain factor
  derivatives[3] = common * dth;
  derivatives[4] = common * dph;
}

MACRO void FixConstraintsLM(int npar, T* P) {}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
