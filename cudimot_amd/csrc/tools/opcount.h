// Algorithmic operation counting for model plug-ins (host build only).
//
// Counted<B> wraps a float or double and tallies every arithmetic
// operation and math-library call made on it, by precision, with the rule:
// +, -, * = 1 flop (a fused multiply-add counts as its two operations),
// / and sqrt = 1, each transcendental (exp, log, sin, cos, tan, atan,
// atan2, acos, erf, pow) = 1.  Comparisons, copies and fabs/min/max are
// free.  A plug-in compiled with T = Counted<float> and CDM_F64 =
// Counted<double> gives the FP32 and FP64 work of one forward evaluation.
#pragma once
#include <cmath>
#include <type_traits>

namespace cdm_count {

struct Tally {
  double f32 = 0, f64 = 0, t32 = 0, t64 = 0;  // flops and transcendental calls by precision
};
inline Tally& tally() {
  static thread_local Tally t;
  return t;
}
template <class B>
inline void flop(double n = 1) {
  if (std::is_same<B, float>::value) tally().f32 += n;
  else tally().f64 += n;
}
template <class B>
inline void trans() {
  if (std::is_same<B, float>::value) {
    tally().f32 += 1;
    tally().t32 += 1;
  } else {
    tally().f64 += 1;
    tally().t64 += 1;
  }
}

template <class B>
struct Counted {
  B v;
  Counted() : v(0) {}
  template <class X, class = typename std::enable_if<std::is_arithmetic<X>::value>::type>
  Counted(X x) : v((B)x) {}
  template <class C>
  explicit Counted(const Counted<C>& o) : v((B)o.v) {}
  explicit operator int() const { return (int)v; }
  explicit operator double() const { return (double)v; }
  explicit operator float() const { return (float)v; }
  Counted operator-() const { return Counted(-v); }
  Counted& operator+=(const Counted& o) { flop<B>(); v += o.v; return *this; }
  Counted& operator-=(const Counted& o) { flop<B>(); v -= o.v; return *this; }
  Counted& operator*=(const Counted& o) { flop<B>(); v *= o.v; return *this; }
  Counted& operator/=(const Counted& o) { flop<B>(); v /= o.v; return *this; }
};

template <class B> Counted<B> mk(B x) { Counted<B> c; c.v = x; return c; }

#define CDM_BINOP(OP)                                                                                   \
  template <class B> Counted<B> operator OP(const Counted<B>& a, const Counted<B>& b) {                 \
    flop<B>(); return mk<B>(a.v OP b.v);                                                                 \
  }                                                                                                      \
  template <class B, class X, class = typename std::enable_if<std::is_arithmetic<X>::value>::type>      \
  Counted<B> operator OP(const Counted<B>& a, X b) { flop<B>(); return mk<B>(a.v OP (B)b); }              \
  template <class B, class X, class = typename std::enable_if<std::is_arithmetic<X>::value>::type>      \
  Counted<B> operator OP(X a, const Counted<B>& b) { flop<B>(); return mk<B>((B)a OP b.v); }
CDM_BINOP(+)
CDM_BINOP(-)
CDM_BINOP(*)
CDM_BINOP(/)
#undef CDM_BINOP

#define CDM_CMPOP(OP)                                                                                   \
  template <class B> bool operator OP(const Counted<B>& a, const Counted<B>& b) { return a.v OP b.v; }  \
  template <class B, class X, class = typename std::enable_if<std::is_arithmetic<X>::value>::type>      \
  bool operator OP(const Counted<B>& a, X b) { return a.v OP (B)b; }                                      \
  template <class B, class X, class = typename std::enable_if<std::is_arithmetic<X>::value>::type>      \
  bool operator OP(X a, const Counted<B>& b) { return (B)a OP b.v; }
CDM_CMPOP(<)
CDM_CMPOP(>)
CDM_CMPOP(<=)
CDM_CMPOP(>=)
CDM_CMPOP(==)
CDM_CMPOP(!=)
#undef CDM_CMPOP

#define CDM_FN1(NAME, EXPR, COST)                                  \
  template <class B> Counted<B> NAME(const Counted<B>& a) {      \
    COST; using std::NAME; return mk<B>((B)EXPR);                 \
  }
CDM_FN1(exp, exp(a.v), trans<B>())
CDM_FN1(log, log(a.v), trans<B>())
CDM_FN1(sin, sin(a.v), trans<B>())
CDM_FN1(cos, cos(a.v), trans<B>())
CDM_FN1(tan, tan(a.v), trans<B>())
CDM_FN1(atan, atan(a.v), trans<B>())
CDM_FN1(acos, acos(a.v), trans<B>())
CDM_FN1(erf, erf(a.v), trans<B>())
CDM_FN1(sqrt, sqrt(a.v), flop<B>())
CDM_FN1(fabs, fabs(a.v), (void)0)
#undef CDM_FN1
template <class B> Counted<B> atan2(const Counted<B>& y, const Counted<B>& x) {
  trans<B>(); return mk<B>((B)std::atan2(y.v, x.v));
}
template <class B> Counted<B> pow(const Counted<B>& a, const Counted<B>& b) {
  trans<B>(); return mk<B>((B)std::pow(a.v, b.v));
}
template <class B> Counted<B> fmin(const Counted<B>& a, const Counted<B>& b) { return a.v < b.v ? a : b; }
template <class B> Counted<B> fmax(const Counted<B>& a, const Counted<B>& b) { return a.v > b.v ? a : b; }
template <class B> Counted<B> fmin(const Counted<B>& a, double b) { return a.v < b ? a : mk<B>((B)b); }

}  // namespace cdm_count

typedef cdm_count::Counted<float> cdm_cf;
typedef cdm_count::Counted<double> cdm_cd;

// Plug-in ABI (functions_gpu.h names) for the counted types.
#define CDM_ABI_1(NAME, FN)                                         \
  inline cdm_cf NAME(cdm_cf x) { return cdm_count::FN(x); }        \
  inline cdm_cd NAME(cdm_cd x) { return cdm_count::FN(x); }
CDM_ABI_1(log_gpu, log)
CDM_ABI_1(exp_gpu, exp)
CDM_ABI_1(fabs_gpu, fabs)
CDM_ABI_1(sin_gpu, sin)
CDM_ABI_1(cos_gpu, cos)
CDM_ABI_1(atan_gpu, atan)
CDM_ABI_1(tsqrt, sqrt)
CDM_ABI_1(tacos, acos)
CDM_ABI_1(ttan, tan)
CDM_ABI_1(tlog, log)
CDM_ABI_1(texp, exp)
CDM_ABI_1(terf, erf)
#undef CDM_ABI_1
inline cdm_cd trcp_iter(cdm_cd x) { return 1.0 / x; }
template <typename T> inline T texp_of(cdm_cd y) { return cdm_count::exp(T(y)); }
inline cdm_cf texp_fast(cdm_cf x) { return cdm_count::exp(x); }
inline cdm_cd texp_fast(cdm_cd x) { return cdm_count::exp(x); }
inline cdm_cf tdiv_fast(cdm_cf a, cdm_cf b) { return a / b; }
inline cdm_cd tdiv_fast(cdm_cd a, cdm_cd b) { return a / b; }
template <typename T> inline T tsqrt_of(cdm_cd c) { return cdm_count::sqrt(T(c)); }
inline cdm_cd tan_gpu(cdm_cf x) { return cdm_count::tan(cdm_cd(x)); }
inline cdm_cd tan_gpu(cdm_cd x) { return cdm_count::tan(x); }
inline cdm_cd sqrt_gpu(cdm_cf x) { return cdm_count::sqrt(cdm_cd(x)); }
inline cdm_cd sqrt_gpu(cdm_cd x) { return cdm_count::sqrt(x); }
inline cdm_cd acos_gpu(cdm_cf x) { return cdm_count::acos(cdm_cd(x)); }
inline cdm_cd acos_gpu(cdm_cd x) { return cdm_count::acos(x); }
inline cdm_cd atan2_gpu(cdm_cf y, cdm_cf x) { return cdm_count::atan2(cdm_cd(y), cdm_cd(x)); }
inline cdm_cd atan2_gpu(cdm_cd y, cdm_cd x) { return cdm_count::atan2(y, x); }
inline cdm_cd pow_gpu(cdm_cf a, cdm_cf b) { return cdm_count::pow(cdm_cd(a), cdm_cd(b)); }
inline cdm_cd pow_gpu(cdm_cd a, cdm_cd b) { return cdm_count::pow(a, b); }
inline cdm_cf tatan2(cdm_cf y, cdm_cf x) { return cdm_count::atan2(y, x); }
inline cdm_cd tatan2(cdm_cd y, cdm_cd x) { return cdm_count::atan2(y, x); }
inline void tsincos(cdm_cf x, cdm_cf* s, cdm_cf* c) { *s = cdm_count::sin(x); *c = cdm_count::cos(x); }
inline void tsincos(cdm_cd x, cdm_cd* s, cdm_cd* c) { *s = cdm_count::sin(x); *c = cdm_count::cos(x); }
inline cdm_cf min_gpu(cdm_cf a, cdm_cf b) { return cdm_count::fmin(a, b); }
inline cdm_cd min_gpu(cdm_cd a, cdm_cd b) { return cdm_count::fmin(a, b); }
inline cdm_cf max_gpu(cdm_cf a, cdm_cf b) { return cdm_count::fmax(a, b); }
inline cdm_cd max_gpu(cdm_cd a, cdm_cd b) { return cdm_count::fmax(a, b); }
