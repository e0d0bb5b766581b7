// Ball and three sticks (no plain counterpart in the reference; written in
// the plug-in format by analogy with Ball_2_Sticks/modelfunctions.h:7-123).
//   S = S0 * [ (1-f1-f2-f3) exp(-b d) + sum_i fi exp(-b d (g.ni)^2) ]
// P: S0, d, f1, th1, ph1, f2, th2, ph2, f3, th3, ph3.  CFP: g (3), b (1).
// The Jacobian is the exact one (no ph2 quirk).
// The compartment exponentials use texp_fast (in single precision the
// hardware 2^y of y = x log2(e): relative error about |x| 2^-24 + 1 ulp,
// below what the float rounding of the argument b d (g.n)^2 itself
// brings); double-precision evaluations use the exact exp.

MACRO void b3s_stick(const T* CFP, T th, T ph, T& gn, T& dth, T& dph) {
  T st, ct, sp, cp;
  tsincos(th, &st, &ct);
  tsincos(ph, &sp, &cp);
  const T gxy = CFP[0] * cp + CFP[1] * sp;
  gn = st * gxy + CFP[2] * ct;
  dth = ct * gxy - st * CFP[2];
  dph = st * (CFP[1] * cp - CFP[0] * sp);
}

MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP) {
  const T bd = CFP[3] * P[1];
  T s = ((T)1 - P[2] - P[5] - P[8]) * texp_fast(-bd);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T g, u, v;
    b3s_stick(CFP, P[3 + 3 * k], P[4 + 3 * k], g, u, v);
    s += P[2 + 3 * k] * texp_fast(-bd * g * g);
  }
  return P[0] * s;
}

MACRO bool ConstraintsMCMC(int npar, T* P) { return !(P[2] < P[5]) && !(P[5] < P[8]); }

MACRO void Partial_Derivatives(int npar, T* P, T* CFP, T* FixP, T* derivatives) {
  const T bd = CFP[3] * P[1];
  const T ball = texp_fast(-bd);
  const T fb = (T)1 - P[2] - P[5] - P[8];
  T mix = fb * ball, dmix = fb * ball;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T g, dth, dph;
    b3s_stick(CFP, P[3 + 3 * k], P[4 + 3 * k], g, dth, dph);
    const T sk = texp_fast(-bd * g * g);
    const T fk = P[2 + 3 * k];
    mix += fk * sk;
    dmix += fk * g * g * sk;
    derivatives[2 + 3 * k] = P[0] * (sk - ball);
    const T c = -(T)2 * P[0] * bd * fk * g * sk;
    derivatives[3 + 3 * k] = c * dth;
    derivatives[4 + 3 * k] = c * dph;
  }
  derivatives[0] = mix;
  derivatives[1] = -P[0] * CFP[3] * dmix;
}

// After LM: order the sticks by decreasing fraction.
MACRO void FixConstraintsLM(int npar, T* P) {
  for (int pass = 0; pass < 2; ++pass) {
    for (int a = 0; a < 2 - pass; ++a) {
      const int i = 2 + 3 * a, j = 2 + 3 * (a + 1);
      if (P[i] < P[j]) {
        for (int k = 0; k < 3; ++k) {
          const T t = P[i + k];
          P[i + k] = P[j + k];
          P[j + k] = t;
        }
      }
    }
  }
}

MACRO T custom_priors(int id_p, T* P, int nmeas, T* CFP, T* FixP) { return (T)0; }
