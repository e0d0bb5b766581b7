// NODDI-Watson helpers: Dawson's integral, Watson hindered diffusivities,
// spherical-harmonic coefficients of the Watson density, Legendre-Gaussian
// integrals and even Legendre polynomials.  Formulas follow the NODDI
// toolbox as used by the reference (NODDI_Watson/WatsonFunctions.h:13-194,
// NODDI_Watson/LegendreFunctions.h:6-107); the per-voxel quantities are
// evaluated in double whatever the plug-in's precision.

// Scalar type of the double-precision parts (an operation-counting type
// when the plug-in is compiled for the FLOP model).
#ifndef CDM_F64
#define CDM_F64 double
#endif
#define CDM_NODDI_DISO 3.0e-3  // isotropic diffusivity (mm^2/s), diffusivities.h
#define CDM_NODDI_DPAR 1.7e-3  // intra-axonal parallel diffusivity (mm^2/s)

// Dawson's integral F(x) = exp(-x^2) int_0^x exp(t^2) dt.
// |x| < 0.2: x (1 - 2/3 x^2 (1 - 2/5 x^2 (1 - 2/7 x^2))).
// Otherwise Rybicki's sampling-theorem sum (Computers in Physics 3, 85,
// 1989) with h = 0.4 and the six nearest odd terms on each side of the
// closest even node n0 h:
//   F(x) ~ sign(x) / sqrt(pi) * exp(-xp^2) * sum_k w_k [e^{(2k+1) 2 xp h}/(n0+2k+1)
//                                                     + e^{-(2k+1) 2 xp h}/(n0-2k-1)]
// with xp = |x| - n0 h and w_k = exp(-((2k+1) h)^2).  This is the
// approximation (and its ~2e-7 accuracy) the reference uses.
__device__ inline CDM_F64 noddi_dawson(CDM_F64 x) {
  const CDM_F64 h = 0.4;
  const CDM_F64 w[6] = {0.8521437889662113, 0.23692775868212165, 0.01831563888873418,
                       0.0003936690406550776, 2.352575200009771e-06, 3.90893843426485e-09};
  if (fabs(x) < 0.2) {
    const CDM_F64 x2 = x * x;
    return x * (1.0 - (2.0 / 3.0) * x2 * (1.0 - 0.4 * x2 * (1.0 - (2.0 / 7.0) * x2)));
  }
  const CDM_F64 ax = fabs(x);
  const int n0 = 2 * (int)(0.5 * ax / h + 0.5);
  const CDM_F64 xp = ax - n0 * h;
  CDM_F64 e1 = exp(2.0 * xp * h);
  const CDM_F64 e2 = e1 * e1;
  CDM_F64 up = n0 + 1, dn = up - 2.0, acc = 0.0;
  for (int k = 0; k < 6; ++k, up += 2.0, dn -= 2.0, e1 *= e2) acc += w[k] * (e1 / up + 1.0 / (dn * e1));
  CDM_F64 r = fabs(exp(-xp * xp));
  if (x < 0) r = -r;
  return 0.5641895835 * r * acc;
}

// Hindered (extra-cellular) diffusivities of a Watson distribution of
// sticks with parallel diffusivity dpar and perpendicular dperp
// (WatsonFunctions.h:53-89): outputs the apparent parallel and
// perpendicular diffusivities of the orientation-averaged tensor.
__device__ inline void noddi_watson_hindered(CDM_F64 kappa, CDM_F64 dperp, CDM_F64& dpar_eq, CDM_F64& dperp_eq) {
  const CDM_F64 dpar = CDM_NODDI_DPAR;
  const CDM_F64 d = dpar - dperp;
  if (kappa < 1e-5) {
    const CDM_F64 s = dpar + 2.0 * dperp, k2 = kappa * kappa;
    dpar_eq = s / 3.0 + 4.0 * d * kappa / 45.0 + 8.0 * d * k2 / 945.0;
    dperp_eq = s / 3.0 - 2.0 * d * kappa / 45.0 - 4.0 * d * k2 / 945.0;
  } else {
    const CDM_F64 sk = sqrt(kappa);
    const CDM_F64 ratio = sk / noddi_dawson(sk);
    dpar_eq = (-d + 2.0 * dperp * kappa + d * ratio) / (2.0 * kappa);
    dperp_eq = (d + 2.0 * (dpar + dperp) * kappa - d * ratio) / (4.0 * kappa);
  }
}

// c_l by quadrature of the definition
//   c_l = sqrt(4 pi (4l+1)) * int_0^1 e^{k t^2} P_2l(t) dt / int_0^1 e^{k t^2} dt
// with the 24-point Gauss-Legendre rule (12 positive nodes, even integrand).
// Every term is O(1), so the absolute error stays near 1e-16.
__device__ inline void noddi_watson_sh_quad(CDM_F64 k, CDM_F64* c) {
  const CDM_F64 node[12][2] = {
      {0.06405689286260563, 0.12793819534675221}, {0.1911188674736163, 0.1258374563468283},
      {0.3150426796961634, 0.12167047292780342},  {0.4337935076260451, 0.11550566805372561},
      {0.5454214713888396, 0.1074442701159656},   {0.6480936519369755, 0.09761865210411406},
      {0.7401241915785544, 0.08619016153195329},  {0.820001985973903, 0.07334648141108041},
      {0.886415527004401, 0.05929858491543674},   {0.9382745520027328, 0.04427743881741955},
      {0.9747285559713095, 0.028531388628933743}, {0.9951872199970213, 0.012341229799987091}};
  CDM_F64 acc[7] = {0, 0, 0, 0, 0, 0, 0}, z = 0.0;
  for (int q = 0; q < 12; ++q) {
    const CDM_F64 t = node[q][0], wq = node[q][1] * exp(k * t * t);
    z += wq;
    CDM_F64 pm = 1.0, p = t;  // Legendre P_0, P_1
    acc[0] += wq;
    for (int l = 2; l <= 12; ++l) {
      const CDM_F64 pn = (t * (2 * l - 1) * p - (l - 1) * pm) / l;
      pm = p;
      p = pn;
      if ((l & 1) == 0) acc[l / 2] += wq * p;
    }
  }
  const CDM_F64 pi = 3.14159265358979323846;
  for (int l = 0; l < 7; ++l) c[l] = sqrt(4.0 * pi * (4 * l + 1)) * acc[l] / z;
}

// Coefficients c_l (l = 0, 2, ..., 12) of the Watson density in the
// normalised even spherical harmonics (WatsonFunctions.h:91-194):
// polynomial fits in log(kappa/30) above 30, closed forms in kappa,
// sqrt(kappa) and Dawson(sqrt(kappa)) for 3 < kappa <= 30, and a Taylor
// expansion at or below 0.4.  Deviation from the reference: in
// 0.4 < kappa <= 3 the closed forms cancel catastrophically (the reference's
// intra-cellular signal is off by up to ~1% at kappa = 0.41, and by more
// than 1e-6 up to kappa ~ 2), so that band uses noddi_watson_sh_quad.  With
// CDM_WATSON_REFERENCE_FORMS defined (the NODDI_Watson_refnum plug-in) the
// reference's closed forms are kept in that band too.
__device__ inline void noddi_watson_sh(CDM_F64 k, CDM_F64* c) {
  const CDM_F64 pi = 3.14159265358979323846;
  c[0] = 2.0 * sqrt(pi);
#ifndef CDM_WATSON_REFERENCE_FORMS
  if (k > 0.4 && k <= 3.0) {
    noddi_watson_sh_quad(k, c);
    return;
  }
#endif
  if (k > 30) {
    // fits in u = log(kappa) - log(30), coefficients of u^0..u^6
    const CDM_F64 fit[6][7] = {
        {7.52308, 0.411538, -0.214588, 0.0784091, -0.023981, 0.00731537, -0.0026467},
        {8.93718, 1.62147, -0.733421, 0.191568, -0.0202906, -0.00779095, 0.00574847},
        {8.87905, 3.35689, -1.15935, 0.0673053, 0.121857, -0.066642, 0.0180215},
        {7.84352, 5.03178, -1.0193, -0.426362, 0.328816, -0.0688176, -0.0229398},
        {6.30113, 6.09914, -0.16088, -1.05578, 0.338069, 0.0937157, -0.106935},
        {4.65678, 6.30069, 1.13754, -1.38393, -0.0134758, 0.331686, -0.105954}};
    const CDM_F64 u = log(k) - log(30.0);
    for (int l = 0; l < 6; ++l) {
      CDM_F64 v = fit[l][0], pw = u;
      for (int j = 1; j < 7; ++j) {
        v = v + fit[l][j] * pw;
        pw = pw * u;
      }
      c[l + 1] = v;
    }
  } else if (k > 0.4) {
    const CDM_F64 sk = sqrt(k);
    const CDM_F64 sk2 = sk * k, sk3 = sk2 * k, sk4 = sk3 * k, sk5 = sk4 * k, sk6 = sk5 * k;
    const CDM_F64 k2 = k * k, k3 = k2 * k, k4 = k3 * k, k5 = k4 * k, k6 = k5 * k;
    const CDM_F64 dw = noddi_dawson(sk);
    const CDM_F64 erfik = (2.0 / 1.772453850905516) * exp(sk * sk) * dw;  // erfi(sqrt k)
    const CDM_F64 ierfik = 1.0 / erfik;
    const CDM_F64 ek = exp(k);
    CDM_F64 v;
    v = 3.0 * sk - (3.0 + 2.0 * k) * dw;
    v = sqrt(5.0) * v * ek;
    c[1] = v * ierfik / k;
    v = (105.0 + 60.0 * k + 12.0 * k2) * dw;
    v = v - 105.0 * sk + 10.0 * sk2;
    v = 0.375 * v * ek / k2;
    c[2] = v * ierfik;
    v = -3465.0 - 1890.0 * k - 420.0 * k2 - 40.0 * k3;
    v = v * dw;
    v = v + 3465.0 * sk - 420.0 * sk2 + 84.0 * sk3;
    v = v * sqrt(13.0 * pi) / 64.0 / k3;
    c[3] = v / dw;
    v = 675675.0 + 360360.0 * k + 83160.0 * k2 + 10080.0 * k3 + 560.0 * k4;
    v = v * dw;
    v = v - 675675.0 * sk + 90090.0 * sk2 - 23100.0 * sk3 + 744.0 * sk4;
    v = sqrt(17.0) * v * ek;
    v = v / 512.0 / k4;
    c[4] = v * ierfik;
    v = -43648605.0 - 22972950.0 * k - 5405400.0 * k2 - 720720.0 * k3 - 55440.0 * k4 - 2016.0 * k5;
    v = v * dw;
    v = v + 43648605.0 * sk - 6126120.0 * sk2 + 1729728.0 * sk3 - 82368.0 * sk4 + 5104.0 * sk5;
    v = sqrt(21.0 * pi) * v / 4096.0 / k5;
    c[5] = v / dw;
    v = 7027425405.0 + 3666482820.0 * k + 872972100.0 * k2 + 122522400.0 * k3 + 10810800.0 * k4 + 576576.0 * k5 +
        14784.0 * k6;
    v = v * dw;
    v = v - 7027425405.0 * sk + 1018467450.0 * sk2 - 302630328.0 * sk3 + 17153136.0 * sk4 - 1553552.0 * sk5 +
        25376.0 * sk6;
    v = 5.0 * v * ek;
    v = v / 16384.0 / k6;
    c[6] = v * ierfik;
  } else {
    const CDM_F64 k2 = k * k, k3 = k2 * k, k4 = k3 * k, k5 = k4 * k, k6 = k5 * k;
    c[1] = (4.0 / 3.0 * k + 8.0 / 63.0 * k2) * sqrt(pi / 5.0);
    c[2] = (8.0 / 21.0 * k2 + 32.0 / 693.0 * k3) * (sqrt(pi) * 0.2);
    c[3] = (16.0 / 693.0 * k3 + 32.0 / 10395.0 * k4) * sqrt(pi / 13.0);
    c[4] = (32.0 / 19305.0 * k4) * sqrt(pi / 17.0);
    c[5] = 64.0 * sqrt(pi / 21.0) * k5 / 692835.0;
    c[6] = 128.0 * sqrt(pi) * k6 / 152108775.0;
  }
}

// Legendre-Gaussian integrals L_l(x) = int_0^1 exp(-x t^2) P_2l(t) dt,
// l = 0..6 (LegendreFunctions.h:6-57): for x > 0.05 the integrals
// I_j = int_0^1 t^2j exp(-x t^2) dt by upward recurrence from
// I_0 = sqrt(pi) erf(sqrt x) / (2 sqrt x) -- here doubled, as the reference
// -- then recombined with the Legendre coefficients; a Taylor expansion for
// small x.  Deviation from the reference: it takes erf in single precision
// (erff) inside its double path, and the upward recurrence multiplies that
// error ~60-fold by I_6 at b = 700 s/mm^2, enough to move the intra-cellular
// signal by up to ~1e-3 at large kappa; here erf is double, except with
// CDM_WATSON_REFERENCE_FORMS (the NODDI_Watson_refnum plug-in), which takes
// the reference's single-precision erf of the float-rounded argument.  That
// value is taken correctly rounded (the double erf rounded to float), one
// fixed choice: math libraries' erff differ from it by an ulp for a few
// percent of arguments (glibc 2.35: 3.6% of floats in [0.2, 4]; CUDA's
// erff allows 2 ulp), and one ulp of erff moves the reference's
// intra-cellular signal by up to 1.4e-2 (kappa 30) through the recurrence.
__device__ inline void noddi_lgi(CDM_F64 x, CDM_F64* L) {
  if (x > 0.05) {
    CDM_F64 I[7];
    const CDM_F64 sx = sqrt(x);
#ifdef CDM_WATSON_REFERENCE_FORMS
    I[0] = sqrt(3.14159265358979323846) * (CDM_F64)(float)erf((CDM_F64)(float)sx) / sx;
#else
    I[0] = sqrt(3.14159265358979323846) * erf(sx) / sx;
#endif
    const CDM_F64 dx = 1.0 / x, emx = -exp(-x);
    for (int i = 1; i < 7; ++i) {
      I[i] = emx + ((i + 1) - 1.5) * I[i - 1];
      I[i] = I[i] * dx;
    }
    L[0] = I[0];
    L[1] = -0.5 * I[0] + 1.5 * I[1];
    L[2] = 0.375 * I[0] - 3.75 * I[1] + 4.375 * I[2];
    L[3] = -0.3125 * I[0] + 6.5625 * I[1] - 19.6875 * I[2] + 14.4375 * I[3];
    L[4] = 0.2734375 * I[0] - 9.84375 * I[1] + 54.140625 * I[2] - 93.84375 * I[3] + 50.2734375 * I[4];
    L[5] = -(63.0 / 256.0) * I[0] + (3465.0 / 256.0) * I[1] - (30030.0 / 256.0) * I[2] + (90090.0 / 256.0) * I[3] -
           (109395.0 / 256.0) * I[4] + (46189.0 / 256.0) * I[5];
    L[6] = (231.0 / 1024.0) * I[0] - (18018.0 / 1024.0) * I[1] + (225225.0 / 1024.0) * I[2] -
           (1021020.0 / 1024.0) * I[3] + (2078505.0 / 1024.0) * I[4] - (1939938.0 / 1024.0) * I[5] +
           (676039.0 / 1024.0) * I[6];
  } else {
    const CDM_F64 x2 = x * x, x3 = x2 * x, x4 = x3 * x, x5 = x4 * x, x6 = x5 * x;
    L[0] = 2.0 - 2.0 * x / 3.0 + x2 / 5.0 - x3 / 21.0 + x4 / 108.0;
    L[1] = -4.0 * x / 15.0 + 4.0 * x2 / 35.0 - 2.0 * x3 / 63.0 + 2.0 * x4 / 297.0;
    L[2] = 8.0 * x2 / 315.0 - 8.0 * x3 / 693.0 + 4.0 * x4 / 1287.0;
    L[3] = -16.0 * x3 / 9009.0 + 16.0 * x4 / 19305.0;
    L[4] = 32.0 * x4 / 328185.0;
    L[5] = -64.0 * x5 / 14549535.0;
    L[6] = 128.0 * x6 / 760543875.0;
  }
}

// sqrt((4j+1)/(4 pi)) * P_2j(x) for j = 0..6, with the three-term Legendre
// recurrence run once up to degree 12.
MACRO void noddi_even_sh(T x, T* sh) {
  const T norm[7] = {(T)0.28209479177387814, (T)0.6307831305050401, (T)0.8462843753216345,
                     (T)1.0171072362820548,  (T)1.1631066229203197, (T)1.2927207364566028,
                     (T)1.4104739588693906};
  T pm = (T)1, p = x;  // P_0, P_1
  sh[0] = norm[0];
  for (int l = 2; l <= 12; ++l) {
    const T pn = (x * (T)(2 * l - 1) * p - (T)(l - 1) * pm) / (T)l;
    pm = p;
    p = pn;
    if ((l & 1) == 0) sh[l / 2] = norm[l / 2] * p;
  }
}
