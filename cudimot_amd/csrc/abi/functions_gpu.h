// Plug-in math ABI for cudimot-style model plug-ins (mymodels/<name>/modelfunctions.h).
//
// A plug-in calls these names from its templated Predicted_Signal /
// Partial_Derivatives / constraints code.  The names and their precision
// contract follow the reference's device helper header (functions_gpu.h:6-40):
//   * log/exp/fabs/sin/cos/atan/min/max keep the argument's precision;
//   * tan/sqrt/atan2/acos/pow always return double, whatever the argument.
// The second group is kept for source compatibility with existing plug-ins.
// Plug-ins written for this engine use the T-preserving t* helpers below
// (tsqrt, tatan2, tacos, tpow, ttan) so that an FP32 instantiation stays in
// FP32 on the GPU.
//
// The same header compiles for the GPU (hipcc, device functions) and for the
// host (g++, plain inline functions), so one plug-in source serves the HIP
// kernels and the CPU code that checks them.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CDM_FUNC __device__ __forceinline__
#else
#include <cmath>
#include <math.h>
#include <algorithm>
#ifndef __device__
#define __device__
#endif
#ifndef __host__
#define __host__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#define CDM_FUNC inline
using std::max;
using std::min;
#endif

// Scalar type of a plug-in's double-precision parts (an operation-counting
// type in the FLOP-model build).
#ifndef CDM_F64
#define CDM_F64 double
#endif

CDM_FUNC float log_gpu(float x) { return logf(x); }
CDM_FUNC double log_gpu(double x) { return log(x); }
CDM_FUNC float exp_gpu(float x) { return expf(x); }
CDM_FUNC double exp_gpu(double x) { return exp(x); }
CDM_FUNC float fabs_gpu(float x) { return fabsf(x); }
CDM_FUNC double fabs_gpu(double x) { return fabs(x); }
CDM_FUNC float sin_gpu(float x) { return sinf(x); }
CDM_FUNC double sin_gpu(double x) { return sin(x); }
CDM_FUNC float cos_gpu(float x) { return cosf(x); }
CDM_FUNC double cos_gpu(double x) { return cos(x); }
CDM_FUNC float atan_gpu(float x) { return atanf(x); }
CDM_FUNC double atan_gpu(double x) { return atan(x); }
CDM_FUNC float min_gpu(float a, float b) { return fminf(a, b); }
CDM_FUNC double min_gpu(double a, double b) { return fmin(a, b); }
CDM_FUNC float max_gpu(float a, float b) { return fmaxf(a, b); }
CDM_FUNC double max_gpu(double a, double b) { return fmax(a, b); }

// Promoted to double for every argument type (reference contract).
CDM_FUNC double tan_gpu(float x) { return tan((double)x); }
CDM_FUNC double tan_gpu(double x) { return tan(x); }
CDM_FUNC double sqrt_gpu(float x) { return sqrt((double)x); }
CDM_FUNC double sqrt_gpu(double x) { return sqrt(x); }
CDM_FUNC double atan2_gpu(float y, float x) { return atan2((double)y, (double)x); }
CDM_FUNC double atan2_gpu(double y, double x) { return atan2(y, x); }
CDM_FUNC double acos_gpu(float x) { return acos((double)x); }
CDM_FUNC double acos_gpu(double x) { return acos(x); }
CDM_FUNC double pow_gpu(float a, float b) { return pow((double)a, (double)b); }
CDM_FUNC double pow_gpu(double a, double b) { return pow(a, b); }

// T-preserving helpers (engine extension, no counterpart in the reference).
CDM_FUNC float tsqrt(float x) { return sqrtf(x); }
CDM_FUNC double tsqrt(double x) { return sqrt(x); }
CDM_FUNC float tatan2(float y, float x) { return atan2f(y, x); }
CDM_FUNC double tatan2(double y, double x) { return atan2(y, x); }
CDM_FUNC float tacos(float x) { return acosf(x); }
CDM_FUNC double tacos(double x) { return acos(x); }
CDM_FUNC float ttan(float x) { return tanf(x); }
CDM_FUNC double ttan(double x) { return tan(x); }
CDM_FUNC float tlog(float x) { return logf(x); }
CDM_FUNC double tlog(double x) { return log(x); }
CDM_FUNC float texp(float x) { return expf(x); }
CDM_FUNC double texp(double x) { return exp(x); }
CDM_FUNC void tsincos(float x, float* s, float* c) { *s = sinf(x); *c = cosf(x); }
CDM_FUNC void tsincos(double x, double* s, double* c) { *s = sin(x); *c = cos(x); }
CDM_FUNC float terf(float x) { return erff(x); }
CDM_FUNC double terf(double x) { return erf(x); }

// Reciprocal for the iterates of a root finder (engine extension): on the
// GPU the hardware estimate refined by one Newton-Raphson step, accurate to
// a few ulp but not correctly rounded, so for values a following iteration
// corrects rather than for results; plain division on the host.
// Single-precision cosine for starting values (engine extension): the
// hardware approximation on the GPU (absolute error ~1e-6 for arguments of
// a few radians), cosf on the host.
#if defined(__HIP_DEVICE_COMPILE__)
CDM_FUNC float tcos_start(float x) { return __cosf(x); }
CDM_FUNC float trsq_start(float x) { return __builtin_amdgcn_rsqf(x); }
CDM_FUNC float tsqrt_start(float x) { return __builtin_amdgcn_sqrtf(x); }
#else
CDM_FUNC float tcos_start(float x) { return cosf(x); }
CDM_FUNC float trsq_start(float x) { return 1.0f / sqrtf(x); }
CDM_FUNC float tsqrt_start(float x) { return sqrtf(x); }
#endif

// Fast forms for single-precision terms whose argument already carries
// float rounding (engine extension): texp_fast(x) is the hardware 2^y of
// y = x log2(e), with relative error about |x| 2^-24 + 1 ulp, the size of
// the error the rounded argument itself brings; tdiv_fast(a, b) is a times
// the hardware reciprocal of b (about 2 ulp, b a normal number).  For
// double both are the exact functions.
#if defined(__HIP_DEVICE_COMPILE__)
CDM_FUNC float texp_fast(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504f); }
CDM_FUNC float tdiv_fast(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
#else
CDM_FUNC float texp_fast(float x) { return expf(x); }
CDM_FUNC float tdiv_fast(float a, float b) { return a / b; }
#endif
CDM_FUNC double texp_fast(double x) { return exp(x); }
CDM_FUNC double tdiv_fast(double a, double b) { return a / b; }

// A double-precision intermediate finished in the signal's precision T
// (engine extension):
//  * texp_of<T>(y) = exp(y).  For float the reduction y log2(e) = n + f,
//    |f| <= 1/2, is exact enough in double that only the hardware 2^f
//    (about 1 ulp on that interval) and an exact scaling by 2^n remain in
//    single precision: no float rounding of a large argument.
//  * tsqrt_of<T>(c) = sqrt(c) for c a normal positive number: the hardware
//    square root (1 ulp) for float, sqrt for double.
template <typename T> CDM_FUNC T texp_of(double y);
template <typename T> CDM_FUNC T tsqrt_of(double c);
template <> CDM_FUNC double texp_of<double>(double y) { return exp(y); }
template <> CDM_FUNC double tsqrt_of<double>(double c) { return sqrt(c); }
#if defined(__HIP_DEVICE_COMPILE__)
template <> CDM_FUNC float texp_of<float>(double y) {
  const double z = y * 1.4426950408889634;  // log2(e)
  const double n = rint(z);
  return __builtin_ldexpf(__builtin_amdgcn_exp2f((float)(z - n)), (int)n);
}
template <> CDM_FUNC float tsqrt_of<float>(double c) { return __builtin_amdgcn_sqrtf((float)c); }
#else
template <> CDM_FUNC float texp_of<float>(double y) { return (float)exp(y); }
template <> CDM_FUNC float tsqrt_of<float>(double c) { return sqrtf((float)c); }
#endif

#if defined(__HIP_DEVICE_COMPILE__)
CDM_FUNC double trcp_iter(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}
#else
CDM_FUNC double trcp_iter(double x) { return 1.0 / x; }
#endif

// Central-difference step for NUMERICAL(i) by precision (see
// macro_numerical.h): the reference's constants for double, a wider relative
// step and floor for float.
template <typename T> struct cdm_fd_step;
template <> struct cdm_fd_step<double> {
  static constexpr double rel = 1e-5, lo = 1e-6, hi = 0.2;
};
template <> struct cdm_fd_step<float> {
  static constexpr float rel = 2e-3f, lo = 1e-4f, hi = 0.2f;
};
