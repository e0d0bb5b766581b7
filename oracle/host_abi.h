// ORACLE (test infrastructure only; never linked into the product).
//
// Host-side plug-in ABI used to compile model plug-ins (this project's and,
// for pinning, the reference's own mymodels/<name>/ headers) with g++ in
// double precision.  It restates the reference's plug-in environment:
//   functions_gpu.h:6-40      *_gpu wrappers (tan/sqrt/atan2/acos/pow -> double)
//   macro_numerical.h:1-32    MACRO, NUMERICAL(i), numerical(): central
//                             difference, step clamp(1e-5|p|, 1e-6, 0.2)
// plus the T-preserving helpers (tsqrt, tsincos, ...) that this project's
// plug-ins use.
#pragma once
#include <algorithm>
#include <cmath>
#include <math.h>

#ifndef __device__
#define __device__
#endif
#ifndef __host__
#define __host__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
using std::max;
using std::min;

inline float log_gpu(float x) { return logf(x); }
inline double log_gpu(double x) { return log(x); }
inline float exp_gpu(float x) { return expf(x); }
inline double exp_gpu(double x) { return exp(x); }
inline float fabs_gpu(float x) { return fabsf(x); }
inline double fabs_gpu(double x) { return fabs(x); }
inline float sin_gpu(float x) { return sinf(x); }
inline double sin_gpu(double x) { return sin(x); }
inline float cos_gpu(float x) { return cosf(x); }
inline double cos_gpu(double x) { return cos(x); }
inline float atan_gpu(float x) { return atanf(x); }
inline double atan_gpu(double x) { return atan(x); }
inline float min_gpu(float a, float b) { return fminf(a, b); }
inline double min_gpu(double a, double b) { return fmin(a, b); }
inline float max_gpu(float a, float b) { return fmaxf(a, b); }
inline double max_gpu(double a, double b) { return fmax(a, b); }
inline double tan_gpu(double x) { return tan(x); }
inline double sqrt_gpu(double x) { return sqrt(x); }
inline double atan2_gpu(double y, double x) { return atan2(y, x); }
inline double acos_gpu(double x) { return acos(x); }
inline double pow_gpu(double a, double b) { return pow(a, b); }

inline double tsqrt(double x) { return sqrt(x); }
inline double tatan2(double y, double x) { return atan2(y, x); }
inline double tacos(double x) { return acos(x); }
inline double ttan(double x) { return tan(x); }
inline double tlog(double x) { return log(x); }
inline double texp(double x) { return exp(x); }
inline void tsincos(double x, double* s, double* c) { *s = sin(x); *c = cos(x); }
inline double terf(double x) { return erf(x); }

// Starting-value and finishing helpers of this project's plug-ins: the
// starting values in single precision, as on the GPU (they only place the
// first iterate of a root finder); everything else exact in double.
inline float tcos_start(float x) { return cosf(x); }
inline float trsq_start(float x) { return 1.0f / sqrtf(x); }
inline float tsqrt_start(float x) { return sqrtf(x); }
inline double trcp_iter(double x) { return 1.0 / x; }
inline double texp_fast(double x) { return exp(x); }
inline double tdiv_fast(double a, double b) { return a / b; }
template <typename T> inline T texp_of(double y) { return (T)exp(y); }
template <typename T> inline T tsqrt_of(double c) { return (T)sqrt(c); }

#ifndef MACRO
#define MACRO template <typename T> __device__ inline
#endif
#define NUMERICAL(idpar) numerical(idpar, P, CFP, FixP);
#define TINY 1e-5
#define MAXSTEP 0.2
#define MINSTEP 1e-6
#define nonzerosign(a) ((a) < 0 ? -1 : 1)
