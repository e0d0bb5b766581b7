// Plug-in declaration ABI: MACRO, NUMERICAL(i) and the central-difference
// helper a plug-in's Partial_Derivatives may call.
//
// Names and semantics follow the reference (macro_numerical.h:1-32):
//   MACRO          template <typename T> __device__ inline
//   NUMERICAL(i)   numerical(i, P, CFP, FixP)
//   step           clamp(TINY*|P[i]|, MINSTEP, MAXSTEP), central difference
//                  (S(P[i]+h/2) - S(P[i]-h/2)) / h
// For a double instantiation the constants are the reference's
// (TINY 1e-5, MINSTEP 1e-6, MAXSTEP 0.2).  For a float instantiation the
// relative step is widened to 2e-3 and the floor to 1e-4: with a 24-bit
// mantissa a 1e-5 relative step leaves ~1e-2 relative error in the
// difference quotient, while a step near eps^(1/3) keeps it near 1e-4.
//
// This header expects NPARAMS to be defined (modelparameters.h) and the
// plug-in to define Predicted_Signal afterwards.  It is included once per
// model namespace, so it has no include guard; functions_gpu.h must already
// be included at global scope.
#include "functions_gpu.h"

#ifndef MACRO
#define MACRO template <typename T> __device__ inline
#endif
#define NUMERICAL(idpar) numerical(idpar, P, CFP, FixP);
#ifndef TINY
#define TINY 1e-5
#endif
#ifndef MAXSTEP
#define MAXSTEP 0.2
#endif
#ifndef MINSTEP
#define MINSTEP 1e-6
#endif
#ifndef nonzerosign
#define nonzerosign(a) ((a) < 0 ? -1 : 1)
#endif


MACRO T Predicted_Signal(int npar, T* P, T* CFP, T* FixP);

MACRO T numerical(int idpar, T* P, T* CFP, T* FixP) {
  T h = fabs_gpu((T)::cdm_fd_step<T>::rel * P[idpar]);
  h = min_gpu(max_gpu(h, (T)::cdm_fd_step<T>::lo), (T)::cdm_fd_step<T>::hi);
  const T keep = P[idpar];
  // Same increment sequence as the reference, so a double instantiation
  // probes bit-identical points.
  P[idpar] += h * (T)0.5;
  const T up = Predicted_Signal(NPARAMS, P, CFP, FixP);
  P[idpar] -= (T)2 * (h * (T)0.5);
  const T dn = Predicted_Signal(NPARAMS, P, CFP, FixP);
  P[idpar] = keep;
  return (up - dn) / h;
}
