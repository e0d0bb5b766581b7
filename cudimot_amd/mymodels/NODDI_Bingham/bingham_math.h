// Bingham-NODDI helpers (Tariq et al., NeuroImage 2016), as the reference's
// NODDI_Bingham plug-in uses them (NODDI_Bingham/modelfunctions.h:37-138,
// hyp_Sapprox_doublePrecision.h:8-131, eigenValues.h:12-51):
//  * the confluent hypergeometric function of a 3x3 matrix argument,
//    1F1(1/2; 3/2; X), by the second-order saddle-point approximation of
//    Kume & Wood (Biometrika 92, 2005), always in double, through the
//    characteristic polynomial of the argument instead of the reference's
//    eigenvalues by Smith's method (the same approximation;
//    eigenValues.h:12-51);
//  * hindered diffusivities from central differences (delta 5e-5) of
//    1F1(kappa, beta, 0), the parallel one in kappa and the first
//    perpendicular one in beta, the second from the trace.
// Scalar type of the double-precision parts (an operation-counting type
// when the plug-in is compiled for the FLOP model).
#ifndef CDM_F64
#define CDM_F64 double
#endif
#ifndef CDM_NODDI_DISO
#define CDM_NODDI_DISO 3.0e-3
#define CDM_NODDI_DPAR 1.7e-3
#endif

// 1F1(1/2; 3/2; X) for X = diag(0, beta, kappa) - r g g^T (r >= 0, g a unit
// vector): the intra-axonal argument of one measurement, and with r = 0 the
// normaliser's, without an eigen-decomposition.  With l_i the eigenvalues
// of L = -X = diag(0, -beta, -kappa) + r g g^T, everything the
// approximation needs is a symmetric function of the l_i, so it follows
// from the coefficients of
// E(s) = det(L - s I) = s3 - s2 s + s1 s^2 - s^3 (s1, s2, s3 the elementary
// symmetric functions, by the matrix determinant lemma):
//   * with y_i = 1 / (l_i - s), the elementary symmetric functions of the
//     y_i are e1 = E2/E, e2 = E1/E, e3 = 1/E (E1 = s1 - 3s,
//     E2 = s2 - 2 s1 s + 3 s^2).  Newton's identities give the power sums
//     p_j = sum_i y_i^j (p1 = e1, p2 = e1 p1 - 2 e2, p3 = e1 p2 - e2 p1 + 3 e3,
//     p4 = e1 p3 - e2 p2 + e3 p1), and K' = p1 / 2, K'' = p2 / 2, K''' = p3,
//     K4 = 3 p4; scaled by powers of E they are polynomials in E, E1, E2,
//     which leaves one reciprocal per Halley step and one for the result;
//   * the saddle point t < l_min solves K'(t) = e1 / 2 = 1, i.e. it is a
//     root of the cubic G(t) = E2 - 2E = 2t^3 + (3 - 2 s1) t^2
//     + 2 (s2 - s1) t + (s2 - 2 s3).  On t < l_min, G = 2E (K' - 1) changes
//     sign exactly once; every other root of G lies at or above l_min, and
//     the saddle point lies below l_min - 1/2, so it is the smallest root of
//     G and a simple one;
//   * that root comes from the trigonometric form of a cubic's roots,
//     evaluated in single precision (acos from Abramowitz & Stegun 4.4.45,
//     absolute error 7e-5, and the hardware cosine, square root and
//     reciprocal square root), and one Halley step on K'(t) - 1 in double
//     brings it to double precision: worst relative error of the result
//     7e-13 against an extended-precision bisection over 10^6 cases across
//     kappa, beta, b and gradient directions, double roots included, the
//     same as three steps from l_min - 1/2; for r = 0 (the normaliser)
//     1.3e-12 over 2 10^5 (kappa, beta) pairs, and 7e-8 in the normaliser's
//     central-difference derivatives.  The step count is the same on every
//     lane of a wave.
// The result is 1F1 = sqrt(c) exp(x) / 4 with c = 2 e3 / K2 and x = T2 - t;
// bingham_hyp_rank1_parts returns (c, x), bingham_hyp_rank1 the value.
__device__ inline void bingham_hyp_rank1_parts(CDM_F64 beta, CDM_F64 kappa, CDM_F64 r, CDM_F64 gx, CDM_F64 gy,
                                               CDM_F64 gn, CDM_F64& c_out, CDM_F64& x_out) {
  const CDM_F64 d2 = -beta, d3 = -kappa;
  const CDM_F64 w1 = gx * gx, w2 = gy * gy, w3 = gn * gn;
  const CDM_F64 s1 = d2 + d3 + r * (w1 + w2 + w3);
  const CDM_F64 s2 = d2 * d3 + r * (w1 * (d2 + d3) + w2 * d3 + w3 * d2);
  const CDM_F64 s3 = r * w1 * d2 * d3;
  // G(t) / 2 = t^3 + a t^2 + bq t + c0; its smallest root is
  // -a/3 + 2 sqrt(p) cos((acos(h / p^3/2) + 2 pi) / 3) with p = (a^2 - 3 bq) / 9
  // and h = (a (9 bq - 2 a^2) - 27 c0) / 54
  const CDM_F64 a = 1.5 - s1, bq = s2 - s1, c0 = 0.5 * s2 - s3;
  CDM_F64 pd = (a * a - 3.0 * bq) * (1.0 / 9.0);
  if (pd < 0.0) pd = 0.0;
  const CDM_F64 hd = (a * (9.0 * bq - 2.0 * a * a) - 27.0 * c0) * (1.0 / 54.0);
  const float p = (float)pd;
  const float rp = p > 1e-30f ? trsq_start(p) : 0.0f;  // 1 / sqrt(p), hardware estimate
  const float sp = p * rp;
  float c = (float)hd * (rp * rp * rp);
  c = fminf(fmaxf(c, -1.0f), 1.0f);
  const float ca = fabsf(c);
  float ac = tsqrt_start(1.0f - ca) * (1.5707288f + ca * (-0.2121144f + ca * (0.0742610f + ca * -0.0187293f)));
  if (c < 0.0f) ac = 3.14159265f - ac;
  const float t0 = (float)(a * (-1.0 / 3.0)) + 2.0f * sp * tcos_start((ac + 6.28318531f) * (1.0f / 3.0f));
  // In terms of E, A = E2 and B = E1 (so e1 = A/E, e2 = B/E, e3 = 1/E):
  // D = E^2 p2 = A^2 - 2 B E, P3 = E^3 p3 = A (A^2 - 3 B E) + 3 E^2 and
  // E^4 p4 = A P3 - B E D + A E^2.  With f = K' - 1 = (A - 2E) / (2E) the
  // Halley step 2 f K'' / (2 K''^2 - f K''') is (A - 2E) D E / (D^2 - (A - 2E) P3):
  // one reciprocal.
  CDM_F64 t = (CDM_F64)t0;
  {
    const CDM_F64 E = s3 - t * (s2 - t * (s1 - t));
    const CDM_F64 A = s2 - t * (2.0 * s1 - 3.0 * t);
    const CDM_F64 B = s1 - 3.0 * t;
    const CDM_F64 D = A * A - 2.0 * B * E;
    const CDM_F64 P3 = A * (A * A - 3.0 * B * E) + 3.0 * E * E;
    const CDM_F64 f2 = A - 2.0 * E;
    t -= f2 * D * E * trcp_iter(D * D - f2 * P3);
  }
  const CDM_F64 E = s3 - t * (s2 - t * (s1 - t));
  const CDM_F64 A = s2 - t * (2.0 * s1 - 3.0 * t);
  const CDM_F64 B = s1 - 3.0 * t;
  const CDM_F64 D = A * A - 2.0 * B * E;
  const CDM_F64 P3 = A * (A * A - 3.0 * B * E) + 3.0 * E * E;
  const CDM_F64 P4 = A * P3 - B * E * D + A * E * E;
  // K2 = D / (2 E^2), K3 = P3 / E^3, K4 = 3 P4 / E^4:
  // T2 = K4 / (8 K2^2) - 5 K3^2 / (24 K2^3) = (1.5 P4 - (5/3) P3^2 / D) / D^2, 2 e3 / K2 = 4 E / D
  const CDM_F64 iD = trcp_iter(D);
  const CDM_F64 T2 = iD * iD * (1.5 * P4 - (5.0 / 3.0) * P3 * P3 * iD);
  c_out = 4.0 * E * iD;
  x_out = T2 - t;
}

__device__ inline CDM_F64 bingham_hyp_rank1(CDM_F64 beta, CDM_F64 kappa, CDM_F64 r, CDM_F64 gx, CDM_F64 gy,
                                           CDM_F64 gn) {
  CDM_F64 c, x;
  bingham_hyp_rank1_parts(beta, kappa, r, gx, gy, gn, c, x);
  return 0.25 * sqrt(c) * exp(x);
}

// Per-fan quantities that depend only on (fintra, kappa, beta): the
// normaliser 1F1(kappa, beta, 0) and the hindered diffusivities.
struct bingham_fan {
  CDM_F64 lnc4;                    // log(4 1F1(kappa, beta, 0))
  CDM_F64 dpar, dperp1, dperp2;    // hindered diffusivities
};

// The five normaliser evaluations of a fan, one per job: 1F1(kappa, beta, 0)
// and its central-difference probes in kappa and in beta (delta 5e-5), each
// the r = 0 case of the rank-one form (1 at the zero matrix, as the
// reference's special case).
constexpr double bingham_fd_delta = 0.00005;
__device__ inline CDM_F64 bingham_fan_job(int job, CDM_F64 kappa, CDM_F64 beta) {
  const CDM_F64 delta = bingham_fd_delta;
  const CDM_F64 k = job == 1 ? kappa + delta : job == 2 ? kappa - delta : kappa;
  const CDM_F64 b = job == 3 ? beta + delta : job == 4 ? beta - delta : beta;
  if (k == 0.0 && b == 0.0) return 1.0;
  return bingham_hyp_rank1(b, k, 0.0, 1.0, 0.0, 0.0);
}

// Fan quantities from the five job values h[0..4].
__device__ inline void bingham_fan_combine(CDM_F64 fintra, const CDM_F64* h, bingham_fan& f) {
  const CDM_F64 dperp = CDM_NODDI_DPAR * (1.0 - fintra);
  const CDM_F64 dd = CDM_NODDI_DPAR - dperp;
  const CDM_F64 delta = bingham_fd_delta;
  f.lnc4 = log(4.0 * h[0]);
  f.dpar = dperp + dd * ((h[1] - h[2]) / (2.0 * delta)) / h[0];
  f.dperp1 = dperp + dd * ((h[3] - h[4]) / (2.0 * delta)) / h[0];
  f.dperp2 = CDM_NODDI_DPAR + 2 * dperp - f.dpar - f.dperp1;
}

__device__ inline void bingham_fan_setup(CDM_F64 fintra, CDM_F64 kappa, CDM_F64 beta, bingham_fan& f) {
  CDM_F64 h[5];
  for (int j = 0; j < 5; ++j) h[j] = bingham_fan_job(j, kappa, beta);
  bingham_fan_combine(fintra, h, f);
}

// Gradient in the fan's local frame for measurement CFP (g, b): components
// along the fibre n(th, ph), the fanning direction (from psi) and their cross
// product (NODDI_Bingham/modelfunctions.h:116-138).
MACRO void bingham_fan_local(T th, T ph, T psi, T* CFP, T& gx, T& gy, T& gn) {
  T st, ct, sp, cp, ss, cs;
  tsincos(th, &st, &ct);
  tsincos(ph, &sp, &cp);
  tsincos(psi, &ss, &cs);
  const T n0 = st * cp, n1 = st * sp, n2 = ct;
  const T f0 = -cs * sp - ct * cp * ss, f1 = cp * cs - ct * sp * ss, f2 = st * ss;
  const T gx_local0 = f1 * n2 - f2 * n1, gx_local1 = f2 * n0 - f0 * n2, gx_local2 = f0 * n1 - f1 * n0;
  gn = CFP[0] * n0 + CFP[1] * n1 + CFP[2] * n2;       // along the fibre
  gy = CFP[0] * f0 + CFP[1] * f1 + CFP[2] * f2;       // along the fanning direction
  gx = CFP[0] * gx_local0 + CFP[1] * gx_local1 + CFP[2] * gx_local2;
}

// The same frame once per parameter vector, in double: fr[0..2] = f x n,
// fr[3..5] = the fanning direction f, fr[6..8] = the fibre n; the local
// gradient of a measurement is then three dot products
// (bingham_frame_local).
__device__ inline void bingham_fan_frame(CDM_F64 th, CDM_F64 ph, CDM_F64 psi, CDM_F64* fr) {
  CDM_F64 st, ct, sp, cp, ss, cs;
  tsincos(th, &st, &ct);
  tsincos(ph, &sp, &cp);
  tsincos(psi, &ss, &cs);
  const CDM_F64 n0 = st * cp, n1 = st * sp, n2 = ct;
  const CDM_F64 f0 = -cs * sp - ct * cp * ss, f1 = cp * cs - ct * sp * ss, f2 = st * ss;
  fr[0] = f1 * n2 - f2 * n1;
  fr[1] = f2 * n0 - f0 * n2;
  fr[2] = f0 * n1 - f1 * n0;
  fr[3] = f0;
  fr[4] = f1;
  fr[5] = f2;
  fr[6] = n0;
  fr[7] = n1;
  fr[8] = n2;
}

MACRO void bingham_frame_local(const CDM_F64* fr, T* CFP, T& gx, T& gy, T& gn) {
  gx = CFP[0] * (T)fr[0] + CFP[1] * (T)fr[1] + CFP[2] * (T)fr[2];
  gy = CFP[0] * (T)fr[3] + CFP[1] * (T)fr[4] + CFP[2] * (T)fr[5];
  gn = CFP[0] * (T)fr[6] + CFP[1] * (T)fr[7] + CFP[2] * (T)fr[8];
}

// Hindered compartment of one fan from the local gradient (the azimuth
// ratio and the exponential in their fast forms: both arguments are
// products of single-precision values already).
MACRO T bingham_fan_extra_local(T gx, T gy, T gn, T b, const bingham_fan& f) {
  const T c2 = gn * gn, s2 = (T)1 - c2;
  // cos^2 of the azimuth atan2(gy, gx) in the plane normal to the fibre
  const T gxy = gx * gx + gy * gy;
  const T caz2 = gxy > (T)0 ? tdiv_fast(gx * gx, gxy) : (T)1, saz2 = (T)1 - caz2;
  return texp_fast(-b * ((T)f.dpar * c2 + s2 * ((T)f.dperp1 * saz2 + (T)f.dperp2 * caz2)));
}

// Intra-axonal compartment of one fan from the local gradient: 1F1 of
// diag(0, beta, kappa) - b Dpar g_local g_local^T over 1F1(kappa, beta, 0),
// i.e. sqrt(c) exp(x - log(4 nc)).  The saddle point, c and x are FP64; the
// square root and the exponential of the exponent difference are finished in
// the plug-in's precision T, the precision of the signal itself (FP64 in the
// LM Jacobian pass), the exponential from its double-precision argument.
MACRO T bingham_fan_intra_local(T kappa, T beta, T gx, T gy, T gn, T b, const bingham_fan& f) {
  const CDM_F64 r = (CDM_F64)b * CDM_NODDI_DPAR;
  // b = 0: the numerator is the normaliser itself, 1F1(diag(0, beta, kappa))
  if (!(r > 0.0)) return (T)1;
  CDM_F64 c, x;
  bingham_hyp_rank1_parts((CDM_F64)beta, (CDM_F64)kappa, r, (CDM_F64)gx, (CDM_F64)gy, (CDM_F64)gn, c, x);
  return tsqrt_of<T>(c) * texp_of<T>(x - f.lnc4);
}

MACRO T bingham_fan_extra(T th, T ph, T psi, const bingham_fan& f, T* CFP) {
  T gx, gy, gn;
  bingham_fan_local(th, ph, psi, CFP, gx, gy, gn);
  return bingham_fan_extra_local(gx, gy, gn, CFP[3], f);
}

MACRO T bingham_fan_intra(T kappa, T beta, T th, T ph, T psi, const bingham_fan& f, T* CFP) {
  T gx, gy, gn;
  bingham_fan_local(th, ph, psi, CFP, gx, gy, gn);
  return bingham_fan_intra_local(kappa, beta, gx, gy, gn, CFP[3], f);
}

// Hindered (e) and intra-axonal (i) signals of one fan.
MACRO void bingham_fan_signal_parts(T kappa, T beta, T th, T ph, T psi, const bingham_fan& f, T* CFP, T& extra,
                                    T& intra) {
  T gx, gy, gn;
  bingham_fan_local(th, ph, psi, CFP, gx, gy, gn);
  extra = bingham_fan_extra_local(gx, gy, gn, CFP[3], f);
  intra = bingham_fan_intra_local(kappa, beta, gx, gy, gn, CFP[3], f);
}

MACRO T bingham_fan_signal(T fintra, T kappa, T beta, T th, T ph, T psi, const bingham_fan& f, T* CFP) {
  T e, i;
  bingham_fan_signal_parts(kappa, beta, th, ph, psi, f, CFP, e, i);
  return ((T)1 - fintra) * e + fintra * i;
}
