// ORACLE (test infrastructure only; never linked into the product).
//
// FP64 CPU restatement of the reference's per-voxel fitting loops.  Included
// inside a model namespace after the plug-in (NPARAMS, Predicted_Signal,
// Partial_Derivatives, ConstraintsMCMC, FixConstraintsLM, custom_priors).
// Each function mirrors one reference kernel for one voxel, with the
// reference's control flow, constants and float-rounded bounds/priors:
//   orc_gridsearch   GridSearch.cu:25-47 (bounds), 49-95 (cost), 97-248 (kernel)
//   orc_levmar       Levenberg_Marquardt.cu:33-135 (transforms), 137-332
//                    (cost, gradient, Hessian), 336-407 (Gauss-Jordan),
//                    409-623 (kernel)
//   orc_mcmc_phase   MCMC.cu:34-68 (logIo), 71-324 (proposal, priors, bounds,
//                    tau, likelihood, Metropolis test), 327-683 (kernel),
//                    771-848 (burn-in then recording launch)
//   orc_predicted    getPredictedSignal.cu:22-92
//   orc_bic_aic      BIC_AIC.cu:61-201
// Sums over measurements follow the reference's partition: 32 strided
// partial sums combined by the same shfl_down tree.
// MCMC draws come from Philox4x32-10 counters (philox_ref.h) because the
// reference's cuRAND XORWOW streams cannot be reproduced.

enum { ORC_NOBOUNDS = 0, ORC_BMIN = 1, ORC_BMAX = 2, ORC_BMINMAX = 3 };
enum { ORC_NOPRIOR = 0, ORC_GAUSS = 1, ORC_GAMMA = 2, ORC_ARD = 3, ORC_SIN = 4, ORC_CUSTOM = 5 };

struct OrcSpec {
  int bound_types[NPARAMS];
  float bmin[NPARAMS], bmax[NPARAMS];
  int prior_types[NPARAMS];
  float pa[NPARAMS], pb[NPARAMS];
  int fixed[NPARAMS];
};

static const int ORC_WARP = 32;

// Sum over measurements with the reference's lane partition + tree.
template <class F>
static double orc_warp_sum(int nmeas, F term) {
  double part[ORC_WARP];
  for (int l = 0; l < ORC_WARP; ++l) {
    double acc = 0.0;
    for (int m = l; m < nmeas; m += ORC_WARP) acc += term(m);
    part[l] = acc;
  }
  for (int off = ORC_WARP / 2; off > 0; off >>= 1)
    for (int l = 0; l + off < ORC_WARP; ++l) part[l] += part[l + off];
  return part[0];
}

static double orc_sse(int nmeas, int cfp_ts, const double* meas, double* P, double* cfp, double* fixp) {
  return orc_warp_sum(nmeas, [&](int m) {
    double e = Predicted_Signal(NPARAMS, P, &cfp[m * cfp_ts], fixp) - meas[m];
    return e * e;
  });
}

// ---------------------------------------------------------------- GridSearch
static bool orc_gs_in_bounds(const OrcSpec& s, const double* p) {
  for (int i = 0; i < NPARAMS; ++i) {
    const int bt = s.bound_types[i];
    if (bt == ORC_BMIN) {
      if (p[i] < s.bmin[i]) return false;
    } else if (bt == ORC_BMAX) {
      if (p[i] > s.bmax[i]) return false;
    } else if (bt == ORC_BMINMAX) {
      if (p[i] < s.bmin[i]) return false;
      if (p[i] > s.bmax[i]) return false;
    }
  }
  return true;
}

// Returns the index of the kept combination (-1: none in bounds).
static int orc_gridsearch(const OrcSpec& s, int nmeas, int cfp_ts, const double* meas, double* params,
                          double* cfp, double* fixp, int ngp, const int* gparams, int ncombs,
                          const double* grid) {
  double best = 9e20;
  int kept = -1;
  double trial[NPARAMS];
  for (int c = 0; c < ncombs; ++c) {
    for (int i = 0; i < NPARAMS; ++i) trial[i] = params[i];
    for (int i = 0; i < ngp; ++i) trial[gparams[i]] = grid[c * ngp + i];
    if (!orc_gs_in_bounds(s, trial)) continue;
    const double cf = orc_sse(nmeas, cfp_ts, meas, trial, cfp, fixp);
    if (cf < best) {
      for (int i = 0; i < NPARAMS; ++i) params[i] = trial[i];
      best = cf;
      kept = c;
    }
  }
  return kept;
}

// ------------------------------------------------------ Levenberg-Marquardt
static const double ORC_CFTOL = 1.0e-8, ORC_LTOL = 1.0e20, ORC_EPS = 2.0e-16;
static const double ORC_TWODIVPI = 0.636619772367581, ORC_AVOID = 1e-4, ORC_SPEED = 10;

static void orc_minmax_inv(const OrcSpec& s, int i, double* p, double* t) {
  const double lo = s.bmin[i], hi = s.bmax[i];
  if (p[i] <= lo + ORC_AVOID) p[i] += ORC_AVOID;
  if (p[i] >= hi - ORC_AVOID) p[i] -= ORC_AVOID;
  t[i] = p[i] - (lo + hi) / 2.0;
  t[i] *= 2.0 / ((hi - lo) * ORC_TWODIVPI);
  t[i] = tan(t[i]) * ORC_SPEED;
}

static void orc_minmax_fwd(const OrcSpec& s, int i, double* p, double* t) {
  const double lo = s.bmin[i], hi = s.bmax[i];
  p[i] = (hi - lo) / 2.0;
  p[i] *= atan(t[i] / ORC_SPEED);
  p[i] *= ORC_TWODIVPI;
  p[i] += (lo + hi) / 2.0;
  if (p[i] <= lo + ORC_AVOID) {
    p[i] += ORC_AVOID;
    t[i] = p[i] - (lo + hi) / 2.0;
    t[i] *= 2.0 / ((hi - lo) * ORC_TWODIVPI);
    t[i] = tan(t[i]) * ORC_SPEED;
  }
  if (p[i] >= hi - ORC_AVOID) {
    p[i] -= ORC_AVOID;
    t[i] = p[i] - (lo + hi) / 2.0;
    t[i] *= 2.0 / ((hi - lo) * ORC_TWODIVPI);
    t[i] = tan(t[i]) * ORC_SPEED;
  }
}

static void orc_inv_all(const OrcSpec& s, double* p, double* t) {
  for (int i = 0; i < NPARAMS; ++i) {
    const int bt = s.bound_types[i];
    if (bt == ORC_BMIN) t[i] = log(p[i] - (double)s.bmin[i]);
    else if (bt == ORC_BMAX) t[i] = log((double)s.bmax[i] - p[i]);
    else if (bt == ORC_BMINMAX) orc_minmax_inv(s, i, p, t);
    else t[i] = p[i];
  }
}

static void orc_fwd_all(const OrcSpec& s, double* p, double* t) {
  for (int i = 0; i < NPARAMS; ++i) {
    const int bt = s.bound_types[i];
    if (bt == ORC_BMIN) p[i] = exp(t[i]) + (double)s.bmin[i];
    else if (bt == ORC_BMAX) p[i] = (double)s.bmax[i] - exp(t[i]);
    else if (bt == ORC_BMINMAX) orc_minmax_fwd(s, i, p, t);
    else p[i] = t[i];
  }
}

static void orc_dtransf(const OrcSpec& s, const double* t, double* d) {
  for (int i = 0; i < NPARAMS; ++i) {
    const int bt = s.bound_types[i];
    if (bt == ORC_BMIN) d[i] = d[i] * exp(t[i]);
    else if (bt == ORC_BMAX) d[i] = d[i] * (-exp(t[i]));
    else if (bt == ORC_BMINMAX) {
      // The reference subtracts the float bounds in float here (no cast).
      const float span = s.bmax[i] - s.bmin[i];
      d[i] = d[i] * (span / 2.0) * ORC_TWODIVPI;
      d[i] = d[i] * (ORC_SPEED / (ORC_SPEED * ORC_SPEED + t[i] * t[i]));
    }
  }
}

// Gradient 2 J^T r and Gauss-Newton Hessian 2 J^T J in the transformed space.
static void orc_grad_hess(const OrcSpec& s, int nmeas, int cfp_ts, const double* meas, double* p,
                          const double* t, double* cfp, double* fixp, double* g, double* hs) {
  // Per-lane partial sums per 32-measurement chunk, then the tree, as the
  // reference reduces each element chunk by chunk.
  for (int i = 0; i < NPARAMS; ++i) g[i] = 0.0;
  for (int i = 0; i < NPARAMS * NPARAMS; ++i) hs[i] = 0.0;
  const int chunks = (nmeas + ORC_WARP - 1) / ORC_WARP;
  static thread_local double J[ORC_WARP][NPARAMS];
  static thread_local double R[ORC_WARP];
  for (int c = 0; c < chunks; ++c) {
    for (int l = 0; l < ORC_WARP; ++l) {
      const int m = c * ORC_WARP + l;
      for (int i = 0; i < NPARAMS; ++i) J[l][i] = 0.0;
      R[l] = 0.0;
      if (m < nmeas) {
        R[l] = Predicted_Signal(NPARAMS, p, &cfp[m * cfp_ts], fixp) - meas[m];
        Partial_Derivatives(NPARAMS, p, &cfp[m * cfp_ts], fixp, J[l]);
        orc_dtransf(s, t, J[l]);
      }
    }
    double part[ORC_WARP];
    for (int i = 0; i < NPARAMS; ++i) {
      for (int l = 0; l < ORC_WARP; ++l) part[l] = 2.0 * R[l] * J[l][i];
      for (int off = ORC_WARP / 2; off > 0; off >>= 1)
        for (int l = 0; l + off < ORC_WARP; ++l) part[l] += part[l + off];
      g[i] += part[0];
      if (s.fixed[i]) g[i] = 0;
    }
    for (int i = 0; i < NPARAMS; ++i)
      for (int k = 0; k < NPARAMS; ++k) {
        for (int l = 0; l < ORC_WARP; ++l) part[l] = 2.0 * J[l][i] * J[l][k];
        for (int off = ORC_WARP / 2; off > 0; off >>= 1)
          for (int l = 0; l + off < ORC_WARP; ++l) part[l] += part[l + off];
        hs[i * NPARAMS + k] += part[0];
      }
  }
  for (int i = 0; i < NPARAMS; ++i)
    for (int k = 0; k < NPARAMS; ++k)
      if (s.fixed[i] || s.fixed[k]) hs[i * NPARAMS + k] = (i == k) ? 1.0 : 0.0;
}

// Gauss-Jordan elimination without pivoting on [hs | g]; x = hs^-1 g.
static void orc_gauss_jordan(const double* hs, const double* g, double* x) {
  double A[NPARAMS][NPARAMS + 1];
  for (int r = 0; r < NPARAMS; ++r) {
    for (int c = 0; c < NPARAMS; ++c) A[r][c] = hs[r * NPARAMS + c];
    A[r][NPARAMS] = g[r];
  }
  for (int col = 0; col < NPARAMS; ++col) {
    const double piv = A[col][col];
    for (int c = 0; c <= NPARAMS; ++c) A[col][c] = A[col][c] / piv;
    for (int r = col + 1; r < NPARAMS; ++r) {
      const double f = A[r][col];
      for (int c = 0; c <= NPARAMS; ++c) A[r][c] -= f * A[col][c];
    }
  }
  for (int col = NPARAMS - 1; col > 0; --col)
    for (int r = 0; r < col; ++r) {
      const double f = A[r][col];
      for (int c = 0; c <= NPARAMS; ++c) A[r][c] -= f * A[col][c];
    }
  for (int r = 0; r < NPARAMS; ++r) x[r] = A[r][NPARAMS];
}

// Returns the number of successful iterations; *final_cost receives the cost.
static int orc_levmar(const OrcSpec& s, int nmeas, int cfp_ts, const double* meas, double* p, double* cfp,
                      double* fixp, int max_iters, bool marquardt, double* final_cost) {
  double t[NPARAMS], g[NPARAMS], hs[NPARAMS * NPARAMS], step[NPARAMS];
  double lambda = 0.1, olambda = 0.0, pcf, ncf = 0.0;
  bool success = true, end = false;
  int iter = 0, nsucc = 0;
  orc_inv_all(s, p, t);
  pcf = orc_sse(nmeas, cfp_ts, meas, p, cfp, fixp);
  while (!(success && iter++ >= max_iters)) {
    if (success) orc_grad_hess(s, nmeas, cfp_ts, meas, p, t, cfp, fixp, g, hs);
    for (int i = 0; i < NPARAMS; ++i) {
      if (marquardt) hs[i * NPARAMS + i] = ((1 + lambda) / (1 + olambda)) * hs[i * NPARAMS + i];
      else hs[i * NPARAMS + i] += lambda - olambda;
    }
    orc_gauss_jordan(hs, g, step);
    for (int i = 0; i < NPARAMS; ++i) step[i] = t[i] - step[i];
    orc_fwd_all(s, p, step);
    ncf = orc_sse(nmeas, cfp_ts, meas, p, cfp, fixp);
    if ((success = (ncf < pcf))) {
      olambda = 0.0;
      for (int i = 0; i < NPARAMS; ++i) t[i] = step[i];
      lambda = lambda / 10.0;
      if (2.0 * fabs(pcf - ncf) <= ORC_CFTOL * (fabs(pcf) + fabs(ncf) + ORC_EPS)) end = true;
      pcf = ncf;
      ++nsucc;
    } else {
      olambda = lambda;
      lambda = lambda * 10.0;
      if (lambda > ORC_LTOL) end = true;
      orc_fwd_all(s, p, t);
    }
    if (end) break;
  }
  FixConstraintsLM(NPARAMS, p);
  if (final_cost) *final_cost = pcf;
  return nsucc;
}

// --------------------------------------------------------------------- MCMC
static double orc_logI0(double x) {
  double y, b = fabs(x);
  if (b < 3.75) {
    double a = x / 3.75;
    a *= a;
    y = 1.0 + a * (3.5156229 + a * (3.0899424 + a * (1.2067492 + a * (0.2659732 + a * (0.0360768 + a * 0.0045813)))));
    y = log(y);
  } else {
    const double a = 3.75 / b;
    y = b + log((0.39894228 +
                 a * (0.01328592 +
                      a * (0.00225319 +
                           a * (-0.00157565 +
                                a * (0.00916281 + a * (-0.02057706 + a * (0.02635537 + a * (-0.01647633 + a * 0.00392377)))))))) /
                sqrt(b));
  }
  return y;
}

static double orc_prior(const OrcSpec& s, int i, double* p, int nmeas, double* cfp, double* fixp) {
  switch (s.prior_types[i]) {
    case ORC_GAUSS: {
      const double sd2 = s.pb[i] * s.pb[i];  // float product, as the reference
      return (p[i] - s.pa[i]) * (p[i] - s.pa[i]) / (2 * sd2);
    }
    case ORC_GAMMA: return (1.0 - s.pa[i]) * log(p[i]) + s.pb[i] * p[i];
    case ORC_ARD: return s.pa[i] * log(p[i]);
    case ORC_SIN: return -log(fabs(sin(p[i]) / 2.0));
    case ORC_CUSTOM: return custom_priors(i, p, nmeas, cfp, fixp);
    default: return 0;
  }
}

static double orc_likelihood(bool rician, int nmeas, int cfp_ts, const double* meas, double* p, double tau,
                             double* cfp, double* fixp) {
  if (rician) {
    const double acc = orc_warp_sum(nmeas, [&](int m) {
      const double s = Predicted_Signal(NPARAMS, p, &cfp[m * cfp_ts], fixp);
      const double y = meas[m];
      return log(y) + (-0.5 * tau * (y * y + s * s) + orc_logI0(tau * s * y));
    });
    return -nmeas * log(tau) - acc;
  }
  const double sse = orc_sse(nmeas, cfp_ts, meas, p, cfp, fixp);
  return (nmeas / 2.0) * log(sse / 2.0);
}

struct OrcChain {
  double propSD[NPARAMS];
  double tau, tauSD;
};

// One launch of the reference sampler (burn-in when !recording).
// jump0 offsets the Philox jump counter so the recording launch draws
// fresh numbers.
static void orc_mcmc_phase(const OrcSpec& s, bool recording, bool update_prop, bool rician, int nmeas, int cfp_ts,
                           const double* meas, double* p, double* cfp, double* fixp, int niters, int nsamples,
                           int sampleevery, int upe, OrcChain& ch, double* samples /*[P][S]*/,
                           double* tau_samples /*[S]*/, uint64_t seed, uint64_t voxel, uint32_t jump0) {
  double priors[NPARAMS];
  int nacc[NPARAMS], nrej[NPARAMS];
  double tau = 0, tauSD = 0;
  int tacc = 0, trej = 0;
  for (int i = 0; i < NPARAMS; ++i) {
    nacc[i] = nrej[i] = 0;
    if (!recording) ch.propSD[i] = p[i] / 10.0;
  }
  if (rician && recording) {
    tau = ch.tau;
    tauSD = ch.tauSD;
  }
  // Clamp to bounds and initialise priors (MCMC.cu:97-139).
  for (int i = 0; i < NPARAMS; ++i) {
    const int bt = s.bound_types[i];
    if (bt == ORC_BMIN) {
      if (p[i] < s.bmin[i]) p[i] = s.bmin[i];
    } else if (bt == ORC_BMAX) {
      if (p[i] > s.bmax[i]) p[i] = s.bmax[i];
    } else if (bt == ORC_BMINMAX) {
      if (p[i] < s.bmin[i]) p[i] = s.bmin[i];
      else if (p[i] > s.bmax[i]) p[i] = s.bmax[i];
    }
    priors[i] = orc_prior(s, i, p, nmeas, cfp, fixp);
  }
  auto total_prior = [&]() {
    double tp = 0;
    for (int i = 0; i < NPARAMS; ++i) tp += priors[i];
    return tp;
  };
  if (rician && !recording) {
    // tau = 1 / variance of the residuals (MCMC.cu:193-251).
    const double mean = orc_warp_sum(nmeas, [&](int m) {
      return Predicted_Signal(NPARAMS, p, &cfp[m * cfp_ts], fixp) - meas[m];
    }) / nmeas;
    const double var = orc_warp_sum(nmeas, [&](int m) {
      const double e = mean - (Predicted_Signal(NPARAMS, p, &cfp[m * cfp_ts], fixp) - meas[m]);
      return e * e;
    }) / nmeas;
    tau = 1.0 / var;
    tauSD = tau / 2.0;
  }
  double like = orc_likelihood(rician, nmeas, cfp_ts, meas, p, tau, cfp, fixp);
  double energy = total_prior() + like, old_energy = 0;

  for (int it = 0; it < niters; ++it) {
    const uint32_t jump = jump0 + (uint32_t)it;
    if (rician) {
      double z, u;
      orc_mcmc_draws(seed, voxel, jump, NPARAMS, &z, &u);
      const double old = tau;
      tau = tau + z * tauSD;
      if (tau > 0) {
        like = orc_likelihood(true, nmeas, cfp_ts, meas, p, tau, cfp, fixp);
        old_energy = energy;
        energy = total_prior() + like;
        if (exp(old_energy - energy) > u) {
          ++tacc;
        } else {
          ++trej;
          tau = old;
          energy = old_energy;
        }
      } else {
        ++trej;
        tau = old;
      }
    }
    for (int i = 0; i < NPARAMS; ++i) {
      double z, u;
      orc_mcmc_draws(seed, voxel, jump, (uint32_t)i, &z, &u);
      const double old = p[i];
      if (!s.fixed[i]) p[i] = p[i] + z * ch.propSD[i];
      bool ok = ConstraintsMCMC(NPARAMS, p);
      if (ok) {
        const int bt = s.bound_types[i];
        if (bt == ORC_BMIN) ok = !(p[i] < s.bmin[i]);
        else if (bt == ORC_BMAX) ok = !(p[i] > s.bmax[i]);
        else if (bt == ORC_BMINMAX) ok = !(p[i] < s.bmin[i]) && !(p[i] > s.bmax[i]);
      }
      if (ok) {
        like = orc_likelihood(rician, nmeas, cfp_ts, meas, p, tau, cfp, fixp);
        const double old_prior = priors[i];
        priors[i] = orc_prior(s, i, p, nmeas, cfp, fixp);
        old_energy = energy;
        energy = total_prior() + like;
        if (exp(old_energy - energy) > u) {
          ++nacc[i];
        } else {
          ++nrej[i];
          p[i] = old;
          priors[i] = old_prior;
          energy = old_energy;
        }
      } else {
        ++nrej[i];
        p[i] = old;
      }
    }
    if (recording && (it % sampleevery) == 0) {
      const int k = it / sampleevery;
      for (int i = 0; i < NPARAMS; ++i) samples[i * nsamples + k] = p[i];
      if (rician) tau_samples[k] = tau;
    }
    if (!recording || update_prop) {
      if (it > 0 && (it % upe) == 0) {
        for (int i = 0; i < NPARAMS; ++i) {
          ch.propSD[i] *= sqrt((nacc[i] + 1.0) / (nrej[i] + 1.0));
          ch.propSD[i] = fmin(ch.propSD[i], 1e10);
          nacc[i] = nrej[i] = 0;
        }
        if (rician) {
          tauSD *= sqrt((tacc + 1.0) / (trej + 1.0));
          tauSD = fmin(tauSD, 1e10);
          tacc = trej = 0;
        }
      }
    }
  }
  if (rician && !recording) {
    ch.tau = tau;
    ch.tauSD = tauSD;
  }
}

// Burn-in launch then recording launch, as MCMC<T>::run (MCMC.cu:771-848).
static void orc_mcmc(const OrcSpec& s, bool rician, bool update_prop, int nmeas, int cfp_ts, const double* meas,
                     double* p, double* cfp, double* fixp, int nburn, int njumps, int sampleevery,
                     double* samples, double* tau_samples, uint64_t seed, uint64_t voxel) {
  const int nb = nburn < 1 ? 1 : nburn;
  const int nsamples = njumps / sampleevery;
  OrcChain ch;
  ch.tau = ch.tauSD = 0;
  orc_mcmc_phase(s, false, true, rician, nmeas, cfp_ts, meas, p, cfp, fixp, nb, nsamples, sampleevery, sampleevery,
                 ch, samples, tau_samples, seed, voxel, 0u);
  if (rician && tau_samples) tau_samples[0] = ch.tau;
  orc_mcmc_phase(s, true, update_prop, rician, nmeas, cfp_ts, meas, p, cfp, fixp, njumps, nsamples, sampleevery,
                 sampleevery, ch, samples, tau_samples, seed, voxel, (uint32_t)nb);
}

// ------------------------------------------------- predicted signal, BIC/AIC
static void orc_mean_params(const double* samples, int nsamples, double* mean) {
  for (int i = 0; i < NPARAMS; ++i) {
    if (nsamples > 1) {
      double v = 0;
      for (int k = 0; k < nsamples; ++k) v += samples[i * nsamples + k];
      mean[i] = v / nsamples;
    } else {
      mean[i] = samples[i];
    }
  }
}

static void orc_predicted(int nmeas, int cfp_ts, int nsamples, const double* samples, double* cfp, double* fixp,
                          double* out) {
  double mp[NPARAMS];
  orc_mean_params(samples, nsamples, mp);
  for (int m = 0; m < nmeas; ++m) out[m] = Predicted_Signal(NPARAMS, mp, &cfp[m * cfp_ts], fixp);
}

static void orc_bic_aic(bool rician, int nmeas, int cfp_ts, int nsamples, const double* meas, const double* samples,
                        const double* tau_samples, double* cfp, double* fixp, double* bic, double* aic) {
  double mp[NPARAMS];
  orc_mean_params(samples, nsamples, mp);
  double tau = 0;
  if (rician) {
    if (nsamples > 1) {
      double v = 0;
      for (int k = 0; k < nsamples; ++k) v += tau_samples[k];
      tau = v / nsamples;
    } else {
      tau = tau_samples[0];
    }
  }
  int np = NPARAMS;
  double ll;
  if (rician) {
    const double acc = orc_warp_sum(nmeas, [&](int m) {
      const double s = Predicted_Signal(NPARAMS, mp, &cfp[m * cfp_ts], fixp);
      const double y = meas[m];
      return log(y) + (-0.5 * tau * (y * y + s * s) + orc_logI0(tau * s * y));
    });
    np++;
    ll = nmeas * log(tau) + acc;
  } else {
    const double sse = orc_sse(nmeas, cfp_ts, meas, mp, cfp, fixp);
    ll = -(nmeas / 2.0) * log(sse / 2.0);
  }
  *aic = -2.0 * ll + np * 2.0;
  *bic = -2.0 * ll + np * log((double)nmeas);
}
