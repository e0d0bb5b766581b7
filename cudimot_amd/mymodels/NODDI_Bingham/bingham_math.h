// Bingham-NODDI helpers (Tariq et al., NeuroImage 2016), as the reference's
// NODDI_Bingham plug-in uses them (NODDI_Bingham/modelfunctions.h:37-138,
// hyp_Sapprox_doublePrecision.h:8-131, eigenValues.h:12-51):
//  * the confluent hypergeometric function of a 3x3 matrix argument,
//    1F1(1/2; 3/2; X), by the second-order saddle-point approximation of
//    Kume & Wood (Biometrika 92, 2005), always in double;
//  * eigenvalues of a symmetric 3x3 matrix by Smith's trigonometric method
//    (Comm. ACM 4, 1961);
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

// Saddle point t of K'(t) = sum_i 1 / (2 (l_i - t)) = 1 with l_i = -x_i,
// t < min l_i.  K' is increasing and convex on (-inf, min l), K'(min l - 1/2)
// >= 1, so Newton's method started there decreases monotonically to the root.
__device__ inline CDM_F64 bingham_saddle(CDM_F64 l1, CDM_F64 l2, CDM_F64 l3) {
  const CDM_F64 lmin = fmin(l1, fmin(l2, l3));
  CDM_F64 t = lmin - 0.5;
  for (int it = 0; it < 60; ++it) {
    const CDM_F64 a = 1.0 / (l1 - t), b = 1.0 / (l2 - t), c = 1.0 / (l3 - t);
    const CDM_F64 f = 0.5 * (a + b + c) - 1.0;
    const CDM_F64 df = 0.5 * (a * a + b * b + c * c);
    const CDM_F64 tn = t - f / df;
    if (!(tn < t)) break;  // monotone from the right: stop when no progress
    t = tn;
  }
  return t;
}

// 1F1(1/2; 3/2; diag(x1, x2, x3)), normalised so that 1F1(0) = 1.
//   c1 = sqrt(2 pi / K2) * pi^0... : c1 = sqrt(2/K2) pi prod_i (l_i - t)^-1/2 e^-t
//   c3 = c1 exp(K4 / (8 K2^2) - 5 K3^2 / (24 K2^3)),   result = c3 / (4 pi)
// with K_j = (j-1)!/2 sum_i (l_i - t)^-j.
__device__ inline CDM_F64 bingham_hyp(CDM_F64 x1, CDM_F64 x2, CDM_F64 x3) {
  if (x1 == 0.0 && x2 == 0.0 && x3 == 0.0) return 1.0;
  const CDM_F64 pi = 3.14159265358979323846;
  const CDM_F64 l1 = -x1, l2 = -x2, l3 = -x3;
  const CDM_F64 t = bingham_saddle(l1, l2, l3);
  const CDM_F64 a = 1.0 / (l1 - t), b = 1.0 / (l2 - t), c = 1.0 / (l3 - t);
  const CDM_F64 K2 = 0.5 * (a * a + b * b + c * c);
  const CDM_F64 K3 = a * a * a + b * b * b + c * c * c;
  const CDM_F64 K4 = 3.0 * (a * a * a * a + b * b * b * b + c * c * c * c);
  const CDM_F64 R = sqrt(a * b * c);
  const CDM_F64 T2 = K4 / (8.0 * K2 * K2) - 5.0 * K3 * K3 / (24.0 * K2 * K2 * K2);
  const CDM_F64 c1 = sqrt(2.0 / K2) * pi * R * exp(-t);
  return c1 * exp(T2) / (4.0 * pi);
}

// Eigenvalues of a symmetric 3x3 matrix [a00 a01 a02; . a11 a12; . . a22],
// in descending order (Smith's method), in double: the intra-axonal
// normaliser feeds central differences in kappa and beta, whose steps are
// far below single-precision resolution of the eigenvalues.
__device__ inline void bingham_eig3(CDM_F64 a00, CDM_F64 a01, CDM_F64 a02, CDM_F64 a11, CDM_F64 a12, CDM_F64 a22,
                                    CDM_F64* e) {
  const CDM_F64 off = a01 * a01 + a02 * a02 + a12 * a12;
  if (off == 0.0) {
    e[0] = a00;
    e[1] = a11;
    e[2] = a22;
    return;
  }
  const CDM_F64 q = (a00 + a11 + a22) / 3.0;
  const CDM_F64 d0 = a00 - q, d1 = a11 - q, d2 = a22 - q;
  const CDM_F64 p = sqrt((d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * off) / 6.0);
  const CDM_F64 ip = 1.0 / p;
  const CDM_F64 b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip, b01 = a01 * ip, b02 = a02 * ip, b12 = a12 * ip;
  const CDM_F64 r = (b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b02 * b12) + b02 * (b01 * b12 - b02 * b11)) / 2.0;
  CDM_F64 phi;
  if (r <= -1.0) phi = 3.14159265358979323846 / 3.0;
  else if (r >= 1.0) phi = 0.0;
  else phi = acos(r) / 3.0;
  e[0] = q + 2.0 * p * cos(phi);
  e[2] = q + 2.0 * p * cos(phi + 2.0 * 3.14159265358979323846 / 3.0);
  e[1] = 3.0 * q - e[0] - e[2];
}

// Per-fan quantities that depend only on (fintra, kappa, beta): the
// normaliser 1F1(kappa, beta, 0) and the hindered diffusivities.
struct bingham_fan {
  CDM_F64 nc;                      // 1F1(kappa, beta, 0)
  CDM_F64 dpar, dperp1, dperp2;    // hindered diffusivities
};

__device__ inline void bingham_fan_setup(CDM_F64 fintra, CDM_F64 kappa, CDM_F64 beta, bingham_fan& f) {
  const CDM_F64 dperp = CDM_NODDI_DPAR * (1.0 - fintra);
  const CDM_F64 dd = CDM_NODDI_DPAR - dperp;
  const CDM_F64 delta = 0.00005;
  f.nc = bingham_hyp(kappa, beta, 0.0);
  CDM_F64 diff = (bingham_hyp(kappa + delta, beta, 0.0) - bingham_hyp(kappa - delta, beta, 0.0)) / (2.0 * delta);
  f.dpar = dperp + dd * diff / f.nc;
  diff = (bingham_hyp(kappa, beta + delta, 0.0) - bingham_hyp(kappa, beta - delta, 0.0)) / (2.0 * delta);
  f.dperp1 = dperp + dd * diff / f.nc;
  f.dperp2 = CDM_NODDI_DPAR + 2 * dperp - f.dpar - f.dperp1;
}

// Hindered (e) and intra-axonal (i) signals of one fan for gradient g and
// b-value b.  Frame: fibre n(th, ph), fanning direction from psi, and their
// cross product (NODDI_Bingham/modelfunctions.h:116-138).
MACRO void bingham_fan_signal_parts(T kappa, T beta, T th, T ph, T psi, const bingham_fan& f, T* CFP, T& extra,
                                    T& intra) {
  T st, ct, sp, cp, ss, cs;
  tsincos(th, &st, &ct);
  tsincos(ph, &sp, &cp);
  tsincos(psi, &ss, &cs);
  const T n0 = st * cp, n1 = st * sp, n2 = ct;
  const T f0 = -cs * sp - ct * cp * ss, f1 = cp * cs - ct * sp * ss, f2 = st * ss;
  const T gx_local0 = f1 * n2 - f2 * n1, gx_local1 = f2 * n0 - f0 * n2, gx_local2 = f0 * n1 - f1 * n0;
  const T gn = CFP[0] * n0 + CFP[1] * n1 + CFP[2] * n2;       // along the fibre
  const T gy = CFP[0] * f0 + CFP[1] * f1 + CFP[2] * f2;       // along the fanning direction
  const T gx = CFP[0] * gx_local0 + CFP[1] * gx_local1 + CFP[2] * gx_local2;
  const T b = CFP[3];
  // hindered compartment
  const T c2 = gn * gn, s2 = (T)1 - c2;
  const T az = tatan2(gy, gx);
  const T caz = cos_gpu(az);
  const T caz2 = caz * caz, saz2 = (T)1 - caz2;
  extra = exp_gpu(-b * ((T)f.dpar * c2 + s2 * ((T)f.dperp1 * saz2 + (T)f.dperp2 * caz2)));
  // intra-axonal compartment: 1F1 of diag(0, beta, kappa) - b Dpar g_local g_local^T
  const CDM_F64 pc = -(CDM_F64)b * CDM_NODDI_DPAR;
  const CDM_F64 dx = (CDM_F64)gx, dy = (CDM_F64)gy, dn = (CDM_F64)gn;
  CDM_F64 e[3];
  bingham_eig3(dx * dx * pc, dx * dy * pc, dx * dn * pc, (CDM_F64)beta + dy * dy * pc, dy * dn * pc,
               (CDM_F64)kappa + dn * dn * pc, e);
  intra = (T)(bingham_hyp(e[0], e[1], e[2]) / f.nc);
}

MACRO T bingham_fan_signal(T fintra, T kappa, T beta, T th, T ph, T psi, const bingham_fan& f, T* CFP) {
  T e, i;
  bingham_fan_signal_parts(kappa, beta, th, ph, psi, f, CFP, e, i);
  return ((T)1 - fintra) * e + fintra * i;
}
