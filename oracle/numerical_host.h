// ORACLE (test infrastructure only).  Central-difference derivative for
// plug-ins that write NUMERICAL(i); restates macro_numerical.h:14-31 of the
// reference in double precision.  Included inside a model's namespace after
// its modelparameters.h (needs NPARAMS).
MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP);

MACRO T numerical(int idpar, T* P, T* CFP, T* FixP) {
  T step = (T)TINY * nonzerosign(P[idpar]) * fabs_gpu(P[idpar]);
  step = min_gpu(max_gpu(fabs_gpu(step), (T)MINSTEP), (T)MAXSTEP);
  const T saved = P[idpar];
  P[idpar] += step * (T)0.5;
  const T a = Predicted_Signal(NPARAMS, P, CFP, FixP);
  P[idpar] -= 2 * (step * (T)0.5);
  const T b = Predicted_Signal(NPARAMS, P, CFP, FixP);
  P[idpar] = saved;
  return (a - b) / step;
}
