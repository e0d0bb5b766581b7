// Fitting kernels, instantiated once per model plug-in.
//
// This header is included inside a model namespace right after the
// plug-in's modelparameters.h and modelfunctions.h, so NPARAMS, MyType and
// the plug-in functions resolve unqualified.  Identifiers here are long on
// purpose: plug-in headers define short macros (H, MN, A1, ...).
//
// Work layout (all kernels): a block of 256 threads = 4 waves; a voxel group
// of L lanes owns one voxel (L = 1..64, chosen per model and measurement
// count); lane l of a group evaluates measurements l, l+L, ...  The per-
// voxel state (parameters, transforms, lambda, chain state) lives in
// registers, replicated in the L lanes of the group; sums over measurements
// use cdm::Group<L>::sum.  Common fixed parameters (gradient directions,
// b-values) are staged in LDS once per block.
//
// Precision: the plug-in's forward model and Jacobian run in MyType (FP32
// for the shipped plug-ins); costs and log-likelihoods accumulate in FP64;
// the LM normal equations are solved by an FP64 LDL^T (Cholesky) factorisation;
// chain and LM state are FP64.

typedef MyType cdm_tf;
// Precision of the Levenberg-Marquardt Jacobian pass.  A plug-in defines
// CDM_LM_JACOBIAN_DOUBLE when central differences of its single-precision
// forward model are noise-dominated along near-degenerate directions (the
// fanning angle of Bingham-NODDI as beta -> 0, crossing sticks started
// together); MCMC and the cost keep their own precision.
#ifdef CDM_LM_JACOBIAN_DOUBLE
typedef double cdm_jt;
typedef double cdm_lm_acc;  // J^T J and J^T r sums
#else
typedef cdm_tf cdm_jt;
typedef float cdm_lm_acc;
#endif
constexpr int cdm_np = NPARAMS;
constexpr int cdm_ntri = cdm::tri_size(NPARAMS);

// Optional invariant caches declared by the plug-in (engine extension, no
// counterpart in the reference's plug-in API):
//   CDM_VOXEL_CACHE n          n values that depend only on the parameters in
//   CDM_VOXEL_CACHE_DEPENDS b  bit mask b, computed by Voxel_Cache(npar, P,
//                              FixP, vc) once per parameter vector
//   CDM_MEAS_CACHE k           k values per measurement from Meas_Cache(CFP,
//                              mc), computed once per block into LDS
//   Predicted_Signal_Cached(npar, P, CFP, FixP, vc, mc)
//   CDM_MEAS_VOXEL_CACHE k     k values per (voxel, measurement) that depend
//   CDM_MEAS_VOXEL_CACHE_DEPENDS b  only on the parameters in bit mask b,
//                              from Meas_Voxel_Cache(npar, P, CFP, FixP, vc,
//                              mc, mvc); Predicted_Signal_MV(npar, P, CFP,
//                              FixP, vc, mc, mvc) uses them.  MCMC keeps them
//                              per voxel so that a step in another parameter
//                              reuses them
//   CDM_MEAS_VOXEL_CACHE_FILL  Predicted_Signal_MV_Fill(npar, P, CFP, FixP,
//                              vc, mc, mvc) fills mvc as Meas_Voxel_Cache and
//                              returns the signal of Predicted_Signal_MV,
//                              sharing the work the two have in common
//   CDM_MEAS_VOXEL_CACHE_PARTIAL  Predicted_Signal_MV_Update(npar, P, CFP,
//                              FixP, vc, mc, cur, mvc, changed) fills mvc
//                              after a change of parameter `changed` alone,
//                              recomputing only the entries that depend on
//                              it and copying the others from cur (all of
//                              them when changed < 0), and returns the
//                              signal of Predicted_Signal_MV
//   CDM_VOXEL_CACHE_JOBS_DEPENDS b  the jobs of CDM_VOXEL_CACHE_JOBS depend
//                              only on the parameters in b (a subset of
//                              CDM_VOXEL_CACHE_DEPENDS): an MCMC step in
//                              another cache parameter recombines the kept
//                              job values instead of recomputing them
//   CDM_VOXEL_CACHE_PARTIAL    with the cache's jobs: after an MCMC step in
//                              parameter `changed` alone, only the jobs
//                              [first, first + count) of
//                              Voxel_Cache_Job_Range(changed, first, count)
//                              are recomputed (the others kept), and
//                              Voxel_Cache_Update(npar, P, FixP, jobs, cur,
//                              vc, changed) fills vc recomputing only the
//                              entries that depend on `changed` and copying
//                              the others from cur, with the arithmetic of
//                              Voxel_Cache_Combine
//   CDM_NUMERICAL_MASK b       derivatives the engine forms by central
//                              differences (the reference's NUMERICAL step
//                              rule) with per-perturbation voxel caches;
//                              Partial_Derivatives_Cached fills the others
//   CDM_LM_ROW_PASSES          with CDM_NUMERICAL_MASK: 64-lane LM groups form
//                              the row-wise Jacobian in two passes over the
//                              measurements (the signal, the analytic entries
//                              and the first half of the central differences,
//                              then the rest), with half the probes live
//   CDM_LM_JACOBIAN_COLUMNS    with CDM_NUMERICAL_MASK and the cache's jobs:
//                              64-lane LM groups form the Jacobian column by
//                              column (cdm_lm_jacobian_columns), holding two
//                              probe caches at a time instead of all of them
//   CDM_SIGNAL_PARTIALS        Predicted_Signal_Partials_Cached(npar, P, CFP,
//                              FixP, vc, mc, d) returns the signal of
//                              Predicted_Signal_Cached and fills the analytic
//                              entries of Partial_Derivatives_Cached from one
//                              evaluation of the shared parts
// Without them the plug-in's Predicted_Signal / Partial_Derivatives are used.
#ifdef CDM_VOXEL_CACHE
constexpr int cdm_vc = CDM_VOXEL_CACHE;
constexpr unsigned cdm_vc_dep = CDM_VOXEL_CACHE_DEPENDS;
#else
constexpr int cdm_vc = 1;
constexpr unsigned cdm_vc_dep = 0;
#endif
#ifdef CDM_MEAS_CACHE
constexpr int cdm_mc = CDM_MEAS_CACHE;
#else
constexpr int cdm_mc = 0;
#endif
#ifdef CDM_MEAS_VOXEL_CACHE
constexpr int cdm_mvc = CDM_MEAS_VOXEL_CACHE;
constexpr unsigned cdm_mvc_dep = CDM_MEAS_VOXEL_CACHE_DEPENDS;
#else
constexpr int cdm_mvc = 0;
constexpr unsigned cdm_mvc_dep = 0;
#endif
// The column-wise LM Jacobian keeps the base point's per-measurement cache
// values in LDS, so that a probe recomputes only the entries its parameter
// moves (Predicted_Signal_MV_Update).
#if defined(CDM_LM_JACOBIAN_COLUMNS) && defined(CDM_MEAS_VOXEL_CACHE_PARTIAL)
constexpr int cdm_lm_base_mvc = CDM_MEAS_VOXEL_CACHE;
#else
constexpr int cdm_lm_base_mvc = 0;
#endif

template <class TF>
__device__ __forceinline__ void cdm_prepare(TF* p, TF* fx, CDM_F64* vc) {
#ifdef CDM_VOXEL_CACHE
  Voxel_Cache(cdm_np, p, fx, vc);
#endif
}

#ifdef CDM_VOXEL_CACHE_JOBS
constexpr int cdm_vc_jobs = CDM_VOXEL_CACHE_JOBS;
#ifdef CDM_VOXEL_CACHE_JOBS_DEPENDS
constexpr unsigned cdm_vcj_dep = CDM_VOXEL_CACHE_JOBS_DEPENDS;
#else
constexpr unsigned cdm_vcj_dep = cdm_vc_dep;
#endif
#else
constexpr int cdm_vc_jobs = 0;
constexpr unsigned cdm_vcj_dep = cdm_vc_dep;
#endif

#if defined(CDM_VOXEL_CACHE) && defined(CDM_VOXEL_CACHE_JOBS)
// The job values of the voxel cache at p, spread over the group's lanes
// (every lane ends with all of them).
template <int L, class TF>
__device__ __forceinline__ void cdm_cache_jobs_g(TF* p, TF* fx, CDM_F64* jr, int lane) {
  constexpr int rounds = (cdm_vc_jobs + L - 1) / L;
  CDM_F64 mine[rounds];
#pragma unroll
  for (int r = 0; r < rounds; ++r) {
    const int j = lane + r * L;
    mine[r] = Voxel_Cache_Job(j < cdm_vc_jobs ? j : 0, cdm_np, p, fx);
  }
#pragma unroll
  for (int j = 0; j < cdm_vc_jobs; ++j) jr[j] = cdm::Group<L>::bcast(mine[j / L], j % L);
}
#endif

#if defined(CDM_VOXEL_CACHE) && defined(CDM_VOXEL_CACHE_JOBS) && defined(CDM_VOXEL_CACHE_PARTIAL)
// After a step in parameter `changed`: the jobs that depend on it, spread
// over the group's lanes, the others copied from cur.
template <int L, class TF>
__device__ __forceinline__ void cdm_cache_jobs_part_g(TF* p, TF* fx, const CDM_F64* cur, CDM_F64* jr, int lane,
                                                      int changed) {
  int first = 0, count = 0;
  Voxel_Cache_Job_Range(changed, first, count);
  constexpr int rounds = (cdm_vc_jobs + L - 1) / L;
  CDM_F64 mine[rounds];
#pragma unroll
  for (int r = 0; r < rounds; ++r) {
    mine[r] = 0.0;
    if (r * L < count) {
      const int j = lane + r * L;
      mine[r] = Voxel_Cache_Job(first + (j < count ? j : 0), cdm_np, p, fx);
    }
  }
#pragma unroll
  for (int j = 0; j < cdm_vc_jobs; ++j) {
    const int k = j - first;
    if (k >= 0 && k < count) {
      CDM_F64 m = mine[0];
#pragma unroll
      for (int r = 1; r < rounds; ++r)
        if (k / L == r) m = mine[r];
      jr[j] = cdm::Group<L>::bcast(m, k % L);
    } else {
      jr[j] = cur[j];
    }
  }
}
#endif

// The voxel cache computed by a group of L lanes.  A plug-in that declares
// CDM_VOXEL_CACHE_JOBS n splits Voxel_Cache into n independent scalar jobs
// (Voxel_Cache_Job) and a combination (Voxel_Cache_Combine); lane l computes
// jobs l, l+L, ..., the group exchanges the results and every lane combines
// them, with the arithmetic of Voxel_Cache.
template <int L, class TF>
__device__ __forceinline__ void cdm_prepare_g(TF* p, TF* fx, CDM_F64* vc, int lane) {
#if defined(CDM_VOXEL_CACHE) && defined(CDM_VOXEL_CACHE_JOBS)
  if constexpr (L > 1) {
    CDM_F64 jr[cdm_vc_jobs];
    cdm_cache_jobs_g<L>(p, fx, jr, lane);
    Voxel_Cache_Combine(cdm_np, p, fx, jr, vc);
    return;
  }
#endif
  (void)lane;
  cdm_prepare(p, fx, vc);
}

template <class TF>
__device__ __forceinline__ TF cdm_signal(TF* p, TF* cfp, TF* fx, const CDM_F64* vc, const CDM_F64* mc) {
#ifdef CDM_VOXEL_CACHE
  return Predicted_Signal_Cached(cdm_np, p, cfp, fx, vc, mc);
#else
  return Predicted_Signal(cdm_np, p, cfp, fx);
#endif
}
constexpr int cdm_cfp_max = 16;
constexpr int cdm_fixp_max = 8;
constexpr int cdm_block = 256;

enum { cdm_NOBOUNDS = 0, cdm_BMIN = 1, cdm_BMAX = 2, cdm_BMINMAX = 3 };
enum { cdm_NOPRIOR = 0, cdm_GAUSS = 1, cdm_GAMMA = 2, cdm_ARD = 3, cdm_SIN = 4, cdm_CUSTOM = 5 };

// Bounds, priors and fixed flags, passed by value (the reference keeps them
// in __constant__ memory as float: GridSearch.cu:20-23,
// Levenberg_Marquardt.cu:28-31, MCMC.cu:24-30).
struct cdm_consts {
  int bound_types[cdm_np];
  float bmin[cdm_np], bmax[cdm_np];
  int prior_types[cdm_np];
  float pa[cdm_np], pb[cdm_np];
  int fixed[cdm_np];
  int dbg_voxel, dbg_cap;  // --debug=<voxel> records (dbg null: off)
  double* dbg;
};

// --debug=<voxel> (cudimotoptions.h:160-161; the reference prints the
// voxel's state from inside its kernels, e.g. Levenberg_Marquardt.cu:482-501,
// MCMC.cu:437-456): the methods write state records of that voxel to a
// device buffer instead, which the host prints.  Region r (0 GridSearch,
// 1 Levenberg-Marquardt, 2 MCMC) of c.dbg holds the number of records
// written, then c.dbg_cap records of cdm_dbg_width doubles: an index, three
// values and the parameters.  Record `slot` is written by every lane of the
// voxel's group (the same values to the same addresses: no branch on the
// lane index).
constexpr int cdm_dbg_width = 4 + cdm_np;
__host__ __device__ constexpr size_t cdm_dbg_region_doubles(int cap) { return 8 + (size_t)cap * cdm_dbg_width; }

template <class TP>
__device__ __forceinline__ void cdm_dbg_record(const cdm_consts& c, int region, int slot, double a0, double a1,
                                               double a2, double a3, const TP* p) {
  if (slot < 0 || slot >= c.dbg_cap) return;
  double* base = c.dbg + (size_t)region * cdm_dbg_region_doubles(c.dbg_cap);
  double* r = base + 8 + (size_t)slot * cdm_dbg_width;
  r[0] = a0;
  r[1] = a1;
  r[2] = a2;
  r[3] = a3;
#pragma unroll
  for (int k = 0; k < cdm_np; ++k) r[4 + k] = (double)p[k];
  base[0] = (double)(slot + 1);
}

struct cdm_geom {
  int nvox, nmeas, cfp_ts, fixp_ts;
  const float* meas;    // [nvox][nmeas]
  const float* cfp;     // [nmeas][cfp_ts]
  const float* fixp;    // [nvox][fixp_ts]
};

// ------------------------------------------------------------------ helpers

// LDS layout: CFP table (float, padded), then the measurement cache (double).
__host__ __device__ __forceinline__ int cdm_mc_offset(int nmeas, int cfp_ts) {
  return ((nmeas * cfp_ts + cdm_cfp_max + 1) / 2) * 2;  // in floats, 8-byte aligned
}

__device__ __forceinline__ const CDM_F64* cdm_stage_cfp(const cdm_geom& g, float* s_cfp) {
  const int n = g.nmeas * g.cfp_ts;
  for (int i = threadIdx.x; i < n + cdm_cfp_max; i += blockDim.x) s_cfp[i] = i < n ? g.cfp[i] : 0.f;
  CDM_F64* s_mc = reinterpret_cast<CDM_F64*>(s_cfp + cdm_mc_offset(g.nmeas, g.cfp_ts));
#ifdef CDM_MEAS_CACHE
  __syncthreads();
  for (int m = threadIdx.x; m < g.nmeas; m += blockDim.x)
    Meas_Cache(const_cast<float*>(&s_cfp[m * g.cfp_ts]), &s_mc[m * cdm_mc]);
#endif
  __syncthreads();
  return s_mc;
}

__device__ __forceinline__ const CDM_F64* cdm_mc_row(const CDM_F64* s_mc, int m) {
  return cdm_mc ? s_mc + m * cdm_mc : nullptr;
}

template <class TF>
__device__ __forceinline__ void cdm_load_fixp(const cdm_geom& g, int v, TF* fx) {
#pragma unroll
  for (int k = 0; k < cdm_fixp_max; ++k) fx[k] = k < g.fixp_ts ? (TF)g.fixp[(size_t)v * g.fixp_ts + k] : (TF)0;
}

// Pointer to one measurement's CFP row in the plug-in's precision.
template <class TF>
struct cdm_cfp_row {
  TF buf[cdm_cfp_max];
  __device__ __forceinline__ TF* at(const float* s_cfp, int m, int cts) {
#pragma unroll
    for (int k = 0; k < cdm_cfp_max; ++k) buf[k] = (TF)s_cfp[m * cts + k];
    return buf;
  }
};
template <>
struct cdm_cfp_row<float> {
  __device__ __forceinline__ float* at(const float* s_cfp, int m, int cts) {
    return const_cast<float*>(&s_cfp[m * cts]);
  }
};

// Sum of squared residuals for parameters p (in TF), FP64 sums.
template <int L, class TF>
__device__ __forceinline__ double cdm_sse(const cdm_geom& g, const float* s_cfp, const CDM_F64* s_mc,
                                          const float* mv, TF* p, TF* fx, int lane) {
  double acc = 0.0;
  cdm_cfp_row<TF> row;
  CDM_F64 vc[cdm_vc];
  cdm_prepare_g<L>(p, fx, vc, lane);
  for (int m = lane; m < g.nmeas; m += L) {
    const TF s = cdm_signal(p, row.at(s_cfp, m, g.cfp_ts), fx, vc, cdm_mc_row(s_mc, m));
    const TF r = s - (TF)mv[m];
    acc += (double)(r * r);
  }
  return cdm::Group<L>::sum(acc);
}

// ---------------------------------------------------------------- GridSearch
// GridSearch.cu:97-248.  Combinations whose grid values fall outside the
// float bounds are removed on the host (same test as checkBounds,
// GridSearch.cu:25-47); the kernel tests the non-grid parameters once per
// voxel.  An FP32 screen keeps the CDM_GS_TOPK best combinations in
// enumeration order (strict <, so ties keep the earliest); those are
// re-scored in FP64 and the FP64 minimum (earliest on ties) is kept.  When
// more than CDM_GS_TOPK combinations sit within the FP32 error band of the
// best, the voxel rescans every combination in FP64, so the chosen index is
// the FP64 argmin.

#ifndef CDM_GS_TOPK
#define CDM_GS_TOPK 4
#endif

struct cdm_gs_args {
  cdm_geom g;
  float* params;            // [nvox][NP] in/out
  int ngp;                  // grid parameters
  int gparams[cdm_np];      // their indices, ascending
  int ncombs;               // valid (in-bounds) combinations
  const float* grid_f;      // [ncombs][ngp]
  const double* grid_d;     // [ncombs][ngp]
  const int* comb_id;       // [ncombs] original enumeration index
  cdm_consts c;
  int* kept_out;            // [nvox] kept original index or -1 (nullable)
};

// One voxel's part of the search: the non-grid parameters' bounds test, the
// fixed parameters and the FP32 screen of combinations [c0, c1) into a
// sorted top-K list (strict <, so ties keep the earliest).
struct cdm_gs_voxel {
  float base[cdm_np];
  cdm_tf fx[cdm_fixp_max];
  float best_c[CDM_GS_TOPK];
  int best_i[CDM_GS_TOPK];
};

// Whether the voxel's non-grid parameters are inside the bounds (and the grid
// is not empty); loads its parameters and fixed parameters.
__device__ __forceinline__ bool cdm_gs_begin(const cdm_gs_args& a, int v, cdm_gs_voxel& w) {
  bool is_grid[cdm_np];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) {
    w.base[i] = a.params[(size_t)v * cdm_np + i];
    is_grid[i] = false;
  }
  for (int k = 0; k < a.ngp; ++k) {
#pragma unroll
    for (int i = 0; i < cdm_np; ++i)
      if (a.gparams[k] == i) is_grid[i] = true;
  }
  bool ok = a.ncombs > 0;
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) {
    if (is_grid[i]) continue;
    const double x = w.base[i];
    const int bt = a.c.bound_types[i];
    if ((bt == cdm_BMIN || bt == cdm_BMINMAX) && x < a.c.bmin[i]) ok = false;
    if ((bt == cdm_BMAX || bt == cdm_BMINMAX) && x > a.c.bmax[i]) ok = false;
  }
  if (ok) cdm_load_fixp(a.g, v, w.fx);
#pragma unroll
  for (int k = 0; k < CDM_GS_TOPK; ++k) {
    w.best_c[k] = INFINITY;
    w.best_i[k] = -1;
  }
  return ok;
}

// Inserts (cost, index) into the sorted list; equal costs keep the earlier
// entry first.
__device__ __forceinline__ void cdm_gs_insert(cdm_gs_voxel& w, float cost, int c) {
  if (cost < w.best_c[CDM_GS_TOPK - 1]) {
    float cc = cost;
    int ci = c;
#pragma unroll
    for (int k = 0; k < CDM_GS_TOPK; ++k) {
      if (cc < w.best_c[k]) {
        const float tc = w.best_c[k];
        const int ti = w.best_i[k];
        w.best_c[k] = cc;
        w.best_i[k] = ci;
        cc = tc;
        ci = ti;
      }
    }
  }
}

template <int L>
__device__ __forceinline__ void cdm_gs_screen(const cdm_gs_args& a, const float* s_cfp, const CDM_F64* s_mc,
                                              const float* mv, cdm_gs_voxel& w, int c0, int c1, int lane) {
  cdm_tf trial[cdm_np];
  for (int c = c0; c < c1; ++c) {
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) trial[i] = (cdm_tf)w.base[i];
    for (int k = 0; k < a.ngp; ++k) {
      const float gv = a.grid_f[(size_t)c * a.ngp + k];
#pragma unroll
      for (int i = 0; i < cdm_np; ++i)
        if (i == a.gparams[k]) trial[i] = (cdm_tf)gv;
    }
    const float cost = (float)cdm_sse<L>(a.g, s_cfp, s_mc, mv, trial, w.fx, lane);
    cdm_gs_insert(w, cost, c);
  }
}

// The sum of squares of the voxel's measurements (the FP32 error band).
template <int L>
__device__ __forceinline__ double cdm_gs_sumsq(const cdm_gs_args& a, const float* mv, int lane) {
  double sumsq = 0.0;
  for (int m = lane; m < a.g.nmeas; m += L) sumsq += (double)mv[m] * (double)mv[m];
  return cdm::Group<L>::sum(sumsq);
}

// Whether more than the K listed combinations may sit within the FP32 error
// band of the best (rounding of the forward model and of the residual sum,
// generous by a factor ~10): then every combination is rescored in FP64.
__device__ __forceinline__ bool cdm_gs_crowded(const cdm_gs_voxel& w, double sumsq) {
  const double b0 = w.best_c[0];
  const double band = 4e-5 * sqrt(fmax(b0, 0.0) * sumsq) + 1e-4 * b0 + 1e-8 * sumsq;
  return w.best_i[CDM_GS_TOPK - 1] >= 0 && (double)w.best_c[CDM_GS_TOPK - 1] <= b0 + band;
}

// FP64 cost of combination c.
template <int L>
__device__ __forceinline__ double cdm_gs_cost64(const cdm_gs_args& a, const float* s_cfp, const CDM_F64* s_mc,
                                                const float* mv, const cdm_gs_voxel& w, int c, int lane) {
  double trial_d[cdm_np];
  double fx_d[cdm_fixp_max];
#pragma unroll
  for (int k = 0; k < cdm_fixp_max; ++k) fx_d[k] = (double)w.fx[k];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) trial_d[i] = (double)w.base[i];
  for (int k = 0; k < a.ngp; ++k) {
    const double gv = a.grid_d[(size_t)c * a.ngp + k];
#pragma unroll
    for (int i = 0; i < cdm_np; ++i)
      if (i == a.gparams[k]) trial_d[i] = gv;
  }
  return cdm_sse<L>(a.g, s_cfp, s_mc, mv, trial_d, fx_d, lane);
}

// The FP64 minimum over combinations [c0, c1) (the earliest on ties).
template <int L>
__device__ __forceinline__ void cdm_gs_rescan(const cdm_gs_args& a, const float* s_cfp, const CDM_F64* s_mc,
                                              const float* mv, const cdm_gs_voxel& w, int c0, int c1, int lane,
                                              double& bestd, int& bestd_i) {
  for (int c = c0; c < c1; ++c) {
    const double cd = cdm_gs_cost64<L>(a, s_cfp, s_mc, mv, w, c, lane);
    if (cd < bestd) {
      bestd = cd;
      bestd_i = c;
    }
  }
}

// The FP64 rescore of the listed candidates, in ascending enumeration order
// (the earliest on ties).
template <int L>
__device__ __forceinline__ void cdm_gs_rescore_list(const cdm_gs_args& a, const float* s_cfp, const CDM_F64* s_mc,
                                                    const float* mv, const cdm_gs_voxel& w, int lane, double& bestd,
                                                    int& bestd_i) {
  int cand[CDM_GS_TOPK];
#pragma unroll
  for (int k = 0; k < CDM_GS_TOPK; ++k) cand[k] = w.best_i[k] < 0 ? INT_MAX : w.best_i[k];
#pragma unroll
  for (int i = 0; i < CDM_GS_TOPK; ++i)
#pragma unroll
    for (int k = 0; k + 1 < CDM_GS_TOPK - i; ++k)
      if (cand[k] > cand[k + 1]) {
        const int tmp = cand[k];
        cand[k] = cand[k + 1];
        cand[k + 1] = tmp;
      }
  for (int q = 0; q < CDM_GS_TOPK; ++q) {
    int c = cand[0];
#pragma unroll
    for (int k = 0; k < CDM_GS_TOPK; ++k)
      if (k == q) c = cand[k];
    if (c == INT_MAX) break;
    const double cd = cdm_gs_cost64<L>(a, s_cfp, s_mc, mv, w, c, lane);
    if (cd < bestd) {
      bestd = cd;
      bestd_i = c;
    }
  }
}

// The outputs for the chosen combination (-1: none): the grid parameters,
// the kept index, and the --debug record.
__device__ __forceinline__ void cdm_gs_write(const cdm_gs_args& a, const cdm_gs_voxel& w, int v, int lane, int chosen,
                                             double chosen_cost) {
  if (lane != 0) return;
  if (chosen < 0) {
    if (a.kept_out) a.kept_out[v] = -1;
    return;
  }
  for (int k = 0; k < a.ngp; ++k) a.params[(size_t)v * cdm_np + a.gparams[k]] = a.grid_f[(size_t)chosen * a.ngp + k];
  if (a.kept_out) a.kept_out[v] = a.comb_id[chosen];
  if (a.c.dbg && v == a.c.dbg_voxel) {
    // the kept combination, its cost (the FP64 rescore's, or the screen's
    // when the plug-in is double), the valid combinations, the parameters
    double pk[cdm_np];
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) pk[i] = (double)w.base[i];
    for (int k = 0; k < a.ngp; ++k) {
#pragma unroll
      for (int i = 0; i < cdm_np; ++i)
        if (i == a.gparams[k]) pk[i] = (double)a.grid_f[(size_t)chosen * a.ngp + k];
    }
    cdm_dbg_record(a.c, 0, 0, (double)a.comb_id[chosen], chosen_cost, (double)a.ncombs, 0.0, pk);
  }
}

// The rest of the search from the screened list: the FP64 rescore of the
// candidates (of every combination when crowded), then the outputs.
template <int L>
__device__ __forceinline__ void cdm_gs_finish(const cdm_gs_args& a, const float* s_cfp, const CDM_F64* s_mc,
                                              const float* mv, cdm_gs_voxel& w, int v, int lane) {
  const double sumsq = cdm_gs_sumsq<L>(a, mv, lane);
  int chosen = w.best_i[0];
  double chosen_cost = w.best_c[0];
  if constexpr (sizeof(cdm_tf) != sizeof(double)) {
    double bestd = INFINITY;
    int bestd_i = -1;
    if (cdm_gs_crowded(w, sumsq))
      cdm_gs_rescan<L>(a, s_cfp, s_mc, mv, w, 0, a.ncombs, lane, bestd, bestd_i);
    else
      cdm_gs_rescore_list<L>(a, s_cfp, s_mc, mv, w, lane, bestd, bestd_i);
    chosen = bestd_i;
    chosen_cost = bestd;
  }
  cdm_gs_write(a, w, v, lane, chosen, chosen_cost);
}

template <int L>
__global__ void __launch_bounds__(cdm_block) cdm_gridsearch_kernel(cdm_gs_args a) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = blockIdx.x * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  const float* mv = a.g.meas + (size_t)v * a.g.nmeas;
  cdm_gs_voxel w;
  if (!cdm_gs_begin(a, v, w)) {
    if (lane == 0 && a.kept_out) a.kept_out[v] = -1;
    return;
  }
  cdm_gs_screen<L>(a, s_cfp, s_mc, mv, w, 0, a.ncombs, lane);
  cdm_gs_finish<L>(a, s_cfp, s_mc, mv, w, v, lane);
}

// Small volumes (fewer voxel groups than the GPU holds): the screen split
// over S slices of the combinations, block (x, z) screening slice z of its
// voxels into split_c / split_i [nvox][S][K]; cdm_gridsearch_merge_kernel
// merges each voxel's S lists in slice order with the same insertion, which
// leaves the K smallest (cost, index) pairs, the sequential screen's list.
// A voxel that is not crowded is finished there; a crowded one is marked in
// crowd[v] and its FP64 rescan split the same way
// (cdm_gridsearch_rescan_kernel: slice z's minimum, the earliest on ties, into
// rescan_c / rescan_i [nvox][S]), then cdm_gridsearch_pick_kernel takes the
// minimum over the slices in order (strict <: the earliest on ties), the
// sequential rescan's choice.
struct cdm_gs_split {
  int S;
  float* list_c;     // [nvox][S][K]
  int* list_i;       // [nvox][S][K]
  int* crowd;        // [nvox]
  double* rescan_c;  // [nvox][S]
  int* rescan_i;     // [nvox][S]
};

__device__ __forceinline__ void cdm_gs_slice(const cdm_gs_args& a, int S, int z, int& c0, int& c1) {
  const int per = (a.ncombs + S - 1) / S;
  c0 = min(a.ncombs, z * per);
  c1 = min(a.ncombs, c0 + per);
}

template <int L>
__global__ void __launch_bounds__(cdm_block) cdm_gridsearch_split_kernel(cdm_gs_args a, cdm_gs_split sp) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = blockIdx.x * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  const int z = blockIdx.y;
  const float* mv = a.g.meas + (size_t)v * a.g.nmeas;
  cdm_gs_voxel w;
  if (cdm_gs_begin(a, v, w)) {
    int c0, c1;
    cdm_gs_slice(a, sp.S, z, c0, c1);
    cdm_gs_screen<L>(a, s_cfp, s_mc, mv, w, c0, c1, lane);
  }
  if (lane == 0) {
    const size_t o = ((size_t)v * sp.S + z) * CDM_GS_TOPK;
#pragma unroll
    for (int k = 0; k < CDM_GS_TOPK; ++k) {
      sp.list_c[o + k] = w.best_c[k];
      sp.list_i[o + k] = w.best_i[k];
    }
  }
}

// The voxel's merged screen list.
__device__ __forceinline__ void cdm_gs_merged(const cdm_gs_split& sp, int v, cdm_gs_voxel& w) {
  for (int z = 0; z < sp.S; ++z) {
    const size_t o = ((size_t)v * sp.S + z) * CDM_GS_TOPK;
#pragma unroll
    for (int k = 0; k < CDM_GS_TOPK; ++k) {
      const int ci = sp.list_i[o + k];
      if (ci >= 0) cdm_gs_insert(w, sp.list_c[o + k], ci);
    }
  }
}

template <int L>
__global__ void __launch_bounds__(cdm_block) cdm_gridsearch_merge_kernel(cdm_gs_args a, cdm_gs_split sp) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = blockIdx.x * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  const float* mv = a.g.meas + (size_t)v * a.g.nmeas;
  cdm_gs_voxel w;
  if (!cdm_gs_begin(a, v, w)) {
    if (lane == 0) sp.crowd[v] = 0;
    cdm_gs_write(a, w, v, lane, -1, 0.0);
    return;
  }
  cdm_gs_merged(sp, v, w);
  if constexpr (sizeof(cdm_tf) != sizeof(double)) {
    const double sumsq = cdm_gs_sumsq<L>(a, mv, lane);
    const bool crowded = cdm_gs_crowded(w, sumsq);
    if (lane == 0) sp.crowd[v] = crowded ? 1 : 0;
    if (crowded) return;
    double bestd = INFINITY;
    int bestd_i = -1;
    cdm_gs_rescore_list<L>(a, s_cfp, s_mc, mv, w, lane, bestd, bestd_i);
    cdm_gs_write(a, w, v, lane, bestd_i, bestd);
  } else {
    if (lane == 0) sp.crowd[v] = 0;
    cdm_gs_write(a, w, v, lane, w.best_i[0], w.best_c[0]);
  }
}

template <int L>
__global__ void __launch_bounds__(cdm_block) cdm_gridsearch_rescan_kernel(cdm_gs_args a, cdm_gs_split sp) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = blockIdx.x * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  if (!sp.crowd[v]) return;
  const int z = blockIdx.y;
  const float* mv = a.g.meas + (size_t)v * a.g.nmeas;
  cdm_gs_voxel w;
  (void)cdm_gs_begin(a, v, w);
  int c0, c1;
  cdm_gs_slice(a, sp.S, z, c0, c1);
  double bestd = INFINITY;
  int bestd_i = -1;
  cdm_gs_rescan<L>(a, s_cfp, s_mc, mv, w, c0, c1, lane, bestd, bestd_i);
  if (lane == 0) {
    sp.rescan_c[(size_t)v * sp.S + z] = bestd;
    sp.rescan_i[(size_t)v * sp.S + z] = bestd_i;
  }
}

template <int L>
__global__ void __launch_bounds__(cdm_block) cdm_gridsearch_pick_kernel(cdm_gs_args a, cdm_gs_split sp) {
  const int v = blockIdx.x * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  if (!sp.crowd[v]) return;
  cdm_gs_voxel w;
  (void)cdm_gs_begin(a, v, w);
  double bestd = INFINITY;
  int bestd_i = -1;
  for (int z = 0; z < sp.S; ++z) {
    const double cd = sp.rescan_c[(size_t)v * sp.S + z];
    const int ci = sp.rescan_i[(size_t)v * sp.S + z];
    if (ci >= 0 && cd < bestd) {
      bestd = cd;
      bestd_i = ci;
    }
  }
  cdm_gs_write(a, w, v, lane, bestd_i, bestd);
}

// ------------------------------------------------------ Levenberg-Marquardt
// Levenberg_Marquardt.cu:409-623, with the bound transforms of :37-135 in
// FP64 against float bounds.  The gradient and Hessian come from one fused
// pass in the Jacobian precision (the reference makes two passes over the
// same values); each lane accumulates its measurements (in FP32, or FP64
// with CDM_LM_JACOBIAN_DOUBLE) before the group sums, and the damped normal
// equations are solved by FP64 LDL^T.  The cost that decides acceptance is
// evaluated with the plug-in instantiated in double: an FP32 cost carries
// ~5e-7 relative rounding noise, which stops the iteration early in flat
// valleys (e.g. the kappa-beta ridge of Bingham-NODDI) where the reference
// keeps improving.  A zero pivot yields an inf/NaN step, which the cost test
// rejects, as in the reference.  Only successful iterations count toward
// max_iters; the loop also ends on a relative cost change <= 1e-8 or
// lambda > 1e20.

constexpr double cdm_lm_cftol = 1.0e-8, cdm_lm_ltol = 1.0e20, cdm_lm_eps = 2.0e-16;
constexpr double cdm_lm_twodivpi = 0.636619772367581, cdm_lm_avoid = 1e-4, cdm_lm_speed = 10;

struct cdm_lm_args {
  cdm_geom g;
  float* params;  // [nvox][NP] in/out
  int max_iters;
  cdm_consts c;
  int* iters_out;       // successful iterations per voxel (nullable)
  float* cost_out;      // final cost per voxel (nullable)
  int* queue;           // work-queue counter (queue kernel only; zero at launch)
};

__device__ __forceinline__ void cdm_minmax_t(double lo, double hi, double p, double& t) {
  t = p - (lo + hi) / 2.0;
  t *= 2.0 / ((hi - lo) * cdm_lm_twodivpi);
  t = tan(t) * cdm_lm_speed;
}

__device__ __forceinline__ void cdm_inv_transform(const cdm_consts& c, double* p, double* t) {
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) {
    const int bt = c.bound_types[i];
    const double lo = c.bmin[i], hi = c.bmax[i];
    if (bt == cdm_BMIN) {
      t[i] = log(p[i] - lo);
    } else if (bt == cdm_BMAX) {
      t[i] = log(hi - p[i]);
    } else if (bt == cdm_BMINMAX) {
      if (p[i] <= lo + cdm_lm_avoid) p[i] += cdm_lm_avoid;
      if (p[i] >= hi - cdm_lm_avoid) p[i] -= cdm_lm_avoid;
      cdm_minmax_t(lo, hi, p[i], t[i]);
    } else {
      t[i] = p[i];
    }
  }
}

__device__ __forceinline__ void cdm_transform(const cdm_consts& c, double* p, double* t) {
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) {
    const int bt = c.bound_types[i];
    const double lo = c.bmin[i], hi = c.bmax[i];
    if (bt == cdm_BMIN) {
      p[i] = exp(t[i]) + lo;
    } else if (bt == cdm_BMAX) {
      p[i] = hi - exp(t[i]);
    } else if (bt == cdm_BMINMAX) {
      double x = (hi - lo) / 2.0;
      x *= atan(t[i] / cdm_lm_speed);
      x *= cdm_lm_twodivpi;
      x += (lo + hi) / 2.0;
      if (x <= lo + cdm_lm_avoid) {
        x += cdm_lm_avoid;
        cdm_minmax_t(lo, hi, x, t[i]);
      }
      if (x >= hi - cdm_lm_avoid) {
        x -= cdm_lm_avoid;
        cdm_minmax_t(lo, hi, x, t[i]);
      }
      p[i] = x;
    } else {
      p[i] = t[i];
    }
  }
}

// d(param)/d(transformed param), Levenberg_Marquardt.cu:120-135.
__device__ __forceinline__ void cdm_dtransform(const cdm_consts& c, const double* t, double* d) {
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) {
    const int bt = c.bound_types[i];
    if (bt == cdm_BMIN) {
      d[i] = exp(t[i]);
    } else if (bt == cdm_BMAX) {
      d[i] = -exp(t[i]);
    } else if (bt == cdm_BMINMAX) {
      const float span = c.bmax[i] - c.bmin[i];  // float difference, as the reference
      d[i] = (span / 2.0) * cdm_lm_twodivpi * (cdm_lm_speed / (cdm_lm_speed * cdm_lm_speed + t[i] * t[i]));
    } else {
      d[i] = 1.0;
    }
  }
}

// Signal and Jacobian at one measurement.  With CDM_NUMERICAL_MASK the
// engine forms the listed derivatives by the reference's central difference
// (step clamp(rel |p|, lo, hi), probes p + h/2 and p + h/2 - 2 (h/2)), using
// voxel caches prepared once per perturbation; otherwise the plug-in's
// Partial_Derivatives is called.
#ifdef CDM_NUMERICAL_MASK
// Whether the row-wise form keeps every probe's voxel cache for the whole
// measurement loop; past this size (in doubles) it recomputes a probe's
// cache where the probe is evaluated instead (64-lane groups of plug-ins
// with CDM_LM_JACOBIAN_COLUMNS take the column-wise form, which keeps two).
constexpr bool cdm_jac_keep_probes =
    (1 + 2 * __builtin_popcount(CDM_NUMERICAL_MASK & cdm_vc_dep)) * cdm_vc <= 160;
#endif

template <class TJ>
struct cdm_jacobian_setup {
  CDM_F64 vc0[cdm_vc];
#ifdef CDM_NUMERICAL_MASK
  CDM_F64 vcp[cdm_np][cdm_vc], vcm[cdm_np][cdm_vc];  // unused (removed) without cdm_jac_keep_probes
  TJ h[cdm_np];
#endif
#ifdef CDM_NUMERICAL_MASK
  // Parameters of cache c: 0 = p, 2k+1 / 2k+2 = the k-th dependent probe
  // p_i + h/2, p_i + h/2 - 2 (h/2) (the probes of eval()).
  __device__ __forceinline__ void probe_params(const TJ* p, int c, TJ* q) const {
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) q[i] = p[i];
    int k = 0;
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      if (!((CDM_NUMERICAL_MASK >> i) & 1u) || !((cdm_vc_dep >> i) & 1u)) continue;
      const TJ up = p[i] + h[i] * (TJ)0.5;
      if (c == 2 * k + 1) q[i] = up;
      if (c == 2 * k + 2) q[i] = up - (TJ)2 * (h[i] * (TJ)0.5);
      ++k;
    }
  }
  // Storage of cache c (compile-time c after unrolling).
  __device__ __forceinline__ CDM_F64* cache_of(int c) {
    if (c == 0) return vc0;
    int k = 0;
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      if (!((CDM_NUMERICAL_MASK >> i) & 1u) || !((cdm_vc_dep >> i) & 1u)) continue;
      if (c == 2 * k + 1) return vcp[i];
      if (c == 2 * k + 2) return vcm[i];
      ++k;
    }
    return vc0;
  }
#endif
  // Voxel caches at p and at every dependent probe.  When the group has at
  // least as many lanes as caches, lane j computes cache j and the group
  // exchanges them, instead of every lane computing all of them.
  template <int L>
  __device__ __forceinline__ void init(TJ* p, TJ* fx, int lane) {
#ifdef CDM_NUMERICAL_MASK
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      if (!((CDM_NUMERICAL_MASK >> i) & 1u)) continue;
      TJ hi = fabs_gpu((TJ)::cdm_fd_step<TJ>::rel * p[i]);
      hi = min_gpu(max_gpu(hi, (TJ)::cdm_fd_step<TJ>::lo), (TJ)::cdm_fd_step<TJ>::hi);
      h[i] = hi;
    }
    constexpr int njobs = 1 + 2 * __builtin_popcount(CDM_NUMERICAL_MASK & cdm_vc_dep);
    if constexpr (!cdm_jac_keep_probes) {
      cdm_prepare_g<L>(p, fx, vc0, lane);
      return;
    }
#ifdef CDM_VOXEL_CACHE_JOBS
    if constexpr (cdm_vc > 0 && njobs > 1 && L >= 2) {
      // every (cache, job) pair on its own lane: pair t = c * jobs + j
      constexpr int T = njobs * cdm_vc_jobs, rounds = (T + L - 1) / L;
      CDM_F64 mine[rounds];
#pragma unroll
      for (int r = 0; r < rounds; ++r) {
        int t = lane + r * L;
        if (t >= T) t = 0;
        TJ q[cdm_np];
        probe_params(p, t / cdm_vc_jobs, q);
        mine[r] = Voxel_Cache_Job(t % cdm_vc_jobs, cdm_np, q, fx);
      }
#pragma unroll
      for (int c = 0; c < njobs; ++c) {
        CDM_F64 jr[cdm_vc_jobs];
#pragma unroll
        for (int j = 0; j < cdm_vc_jobs; ++j)
          jr[j] = cdm::Group<L>::bcast(mine[(c * cdm_vc_jobs + j) / L], (c * cdm_vc_jobs + j) % L);
        TJ q[cdm_np];
        probe_params(p, c, q);
        CDM_F64* dst = cache_of(c);
        Voxel_Cache_Combine(cdm_np, q, fx, jr, dst);
      }
      return;
    }
#endif
    if constexpr (cdm_vc > 0 && njobs > 1 && L >= njobs) {
      // job 0: p; job 2k+1 / 2k+2: the k-th dependent probe p_i + h/2, p_i + h/2 - 2 (h/2)
      const int job = lane < njobs ? lane : 0;
      TJ q[cdm_np];
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) q[i] = p[i];
      int k = 0;
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) {
        if (!((CDM_NUMERICAL_MASK >> i) & 1u) || !((cdm_vc_dep >> i) & 1u)) continue;
        const TJ up = p[i] + h[i] * (TJ)0.5;
        if (job == 2 * k + 1) q[i] = up;
        if (job == 2 * k + 2) q[i] = up - (TJ)2 * (h[i] * (TJ)0.5);
        ++k;
      }
      CDM_F64 mine[cdm_vc];
      cdm_prepare(q, fx, mine);
#pragma unroll
      for (int c = 0; c < cdm_vc; ++c) vc0[c] = cdm::Group<L>::bcast(mine[c], 0);
      k = 0;
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) {
        if (!((CDM_NUMERICAL_MASK >> i) & 1u) || !((cdm_vc_dep >> i) & 1u)) continue;
#pragma unroll
        for (int c = 0; c < cdm_vc; ++c) {
          vcp[i][c] = cdm::Group<L>::bcast(mine[c], 2 * k + 1);
          vcm[i][c] = cdm::Group<L>::bcast(mine[c], 2 * k + 2);
        }
        ++k;
      }
      return;
    }
    cdm_prepare(p, fx, vc0);
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      if (!((CDM_NUMERICAL_MASK >> i) & 1u) || !((cdm_vc_dep >> i) & 1u)) continue;
      const TJ keep = p[i];
      p[i] += h[i] * (TJ)0.5;
      cdm_prepare(p, fx, vcp[i]);
      p[i] -= (TJ)2 * (h[i] * (TJ)0.5);
      cdm_prepare(p, fx, vcm[i]);
      p[i] = keep;
    }
#else
    (void)lane;
    cdm_prepare(p, fx, vc0);
#endif
  }
  __device__ __forceinline__ TJ eval(TJ* p, TJ* cm, TJ* fx, const CDM_F64* mc, TJ* d) {
#ifdef CDM_NUMERICAL_MASK
    return eval_part<CDM_NUMERICAL_MASK, true>(p, cm, fx, mc, d);
#else
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) d[i] = (TJ)0;
    const TJ s = cdm_signal(p, cm, fx, vc0, mc);
    Partial_Derivatives(cdm_np, p, cm, fx, d);
    return s;
#endif
  }
#ifdef CDM_NUMERICAL_MASK
  // eval restricted to the central differences of the parameters in PROBES
  // (a subset of CDM_NUMERICAL_MASK), with the signal and the analytic
  // entries when SIGNAL (0 is returned otherwise); each entry is eval's.
  template <unsigned PROBES, bool SIGNAL>
  __device__ __forceinline__ TJ eval_part(TJ* p, TJ* cm, TJ* fx, const CDM_F64* mc, TJ* d) {
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) d[i] = (TJ)0;
    TJ s = (TJ)0;
    if constexpr (SIGNAL) {
#ifdef CDM_SIGNAL_PARTIALS
      s = Predicted_Signal_Partials_Cached(cdm_np, p, cm, fx, vc0, mc, d);
#else
      s = cdm_signal(p, cm, fx, vc0, mc);
      Partial_Derivatives_Cached(cdm_np, p, cm, fx, vc0, mc, d);
#endif
    }
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      if (!((CDM_NUMERICAL_MASK >> i) & 1u) || !((PROBES >> i) & 1u)) continue;
      const bool dep = (cdm_vc_dep >> i) & 1u;
      const TJ keep = p[i];
      TJ up, dn;
      if constexpr (cdm_jac_keep_probes) {
        p[i] += h[i] * (TJ)0.5;
        up = cdm_signal(p, cm, fx, dep ? vcp[i] : vc0, mc);
        p[i] -= (TJ)2 * (h[i] * (TJ)0.5);
        dn = cdm_signal(p, cm, fx, dep ? vcm[i] : vc0, mc);
      } else {
        CDM_F64 vq[cdm_vc];
        p[i] += h[i] * (TJ)0.5;
        if (dep) cdm_prepare(p, fx, vq);
        up = cdm_signal(p, cm, fx, dep ? vq : vc0, mc);
        p[i] -= (TJ)2 * (h[i] * (TJ)0.5);
        if (dep) cdm_prepare(p, fx, vq);
        dn = cdm_signal(p, cm, fx, dep ? vq : vc0, mc);
      }
      p[i] = keep;
      d[i] = (up - dn) / h[i];
    }
    return s;
  }
#endif
};

// The first ceil(n/2) parameters of a mask of n (the first pass of
// CDM_LM_ROW_PASSES).
constexpr unsigned cdm_lm_first_half(unsigned mask) {
  int keep = (__builtin_popcount(mask) + 1) / 2;
  unsigned out = 0;
  for (int b = 0; b < 32 && keep > 0; ++b)
    if ((mask >> b) & 1u) {
      out |= 1u << b;
      --keep;
    }
  return out;
}

// A 64-lane LM group (one whole wave) sums J^T J and J^T r through LDS: each
// lane stores the rows of its measurements, then lane e sums entry e (a
// packed lower-triangle element of J^T J, then the elements of J^T r) over
// all measurements in order, in FP64.  The per-lane accumulators of the
// other group sizes (P(P+1)/2 + P values live across every measurement's
// evaluation) are then never live with the Jacobian's evaluations, which
// with many parameters outgrow the register file.  The per-group scratch
// sits after the CFP table and the measurement cache in the dynamic LDS.
template <int L>
constexpr bool cdm_lm_lds_normal = (L == 64);

__host__ __device__ __forceinline__ size_t cdm_lm_scratch_offset(int nmeas, int cfp_ts) {
  const size_t b = (size_t)cdm_mc_offset(nmeas, cfp_ts) * sizeof(float) + (size_t)nmeas * cdm_mc * sizeof(CDM_F64);
  return (b + 15) / 16 * 16;
}

// The column-wise LM Jacobian builds the voxel caches of all its
// central-difference probes at once, one probe per lane, into LDS (after the
// base point's per-measurement cache values): the jobs of the probes whose
// parameter moves them, [2 x those parameters][jobs], then the caches of
// the probes whose parameter moves the cache, [2 x those parameters][cache].
#if defined(CDM_LM_JACOBIAN_COLUMNS) && defined(CDM_NUMERICAL_MASK) && defined(CDM_VOXEL_CACHE) && \
    defined(CDM_VOXEL_CACHE_JOBS)
constexpr unsigned cdm_lm_kdep = CDM_NUMERICAL_MASK & cdm_vc_dep;
constexpr unsigned cdm_lm_kjdep = CDM_NUMERICAL_MASK & cdm_vcj_dep;
constexpr size_t cdm_lm_probe_doubles =
    (size_t)2 * __builtin_popcount(cdm_lm_kjdep) * cdm_vc_jobs + (size_t)2 * __builtin_popcount(cdm_lm_kdep) * cdm_vc;
#else
constexpr size_t cdm_lm_probe_doubles = 0;
#endif

// Doubles of LDS scratch per LM voxel group: the [M][P+1] rows (J, r), then
// the P(P+1)/2 + P sums (J^T J packed, then J^T r; J^T J is kept there,
// doubled and with the fixed parameters' rows and columns replaced, for the
// damped solves of rejected steps), then the P(P+1)/2 entries of the
// factorised damped matrix, then (column-wise Jacobian with partial
// per-measurement caches) the base point's [M][cdm_lm_base_mvc] cache values,
// then the column-wise Jacobian's probe jobs and caches; none below 64 lanes.
__host__ __device__ __forceinline__ size_t cdm_lm_scratch_doubles(int L, int nmeas) {
  return L == 64 ? (size_t)nmeas * (cdm_np + 1) + 2 * cdm_ntri + cdm_np + (size_t)nmeas * cdm_lm_base_mvc +
                       cdm_lm_probe_doubles
                 : 0;
}

// LDS stores of one wave made visible to the wave's other lanes.
__device__ __forceinline__ void cdm_wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// sums[e] for e < P(P+1)/2: sum_m J_mi J_mj of packed element e = i(i+1)/2 + j;
// sums[P(P+1)/2 + i]: sum_m J_mi r_m.  rows: [nmeas][P+1].
static_assert(cdm::tri_ij_exact<cdm_np>(), "packed lower-triangle index");

template <int L>
__device__ __forceinline__ void cdm_lds_normal_sums(const CDM_F64* rows, CDM_F64* sums, int nmeas, int lane) {
  constexpr int T = cdm_ntri + cdm_np;
  for (int e = lane; e < T; e += L) {
    int i = 0, j = 0;
    if (e < cdm_ntri) {
      cdm::tri_ij<cdm_np>(e, i, j);
    } else {
      i = e - cdm_ntri;
      j = cdm_np;
    }
    double acc = 0.0;
    for (int m = 0; m < nmeas; ++m) acc += rows[m * (cdm_np + 1) + i] * rows[m * (cdm_np + 1) + j];
    sums[e] = acc;
  }
}

#if defined(CDM_LM_JACOBIAN_COLUMNS) && defined(CDM_NUMERICAL_MASK) && defined(CDM_VOXEL_CACHE) && \
    defined(CDM_VOXEL_CACHE_JOBS)
// Column-wise Jacobian of a 64-lane LM group into the LDS rows [M][P+1]
// (J columns, then the residual): one pass for the signal, the residual and
// the analytic entries, then one pass per central-difference parameter,
// whose two probes' voxel caches the group computes once (their 1F1 jobs
// spread over the lanes when the parameter moves them, a recombination of
// the base jobs otherwise, the base cache when the cache does not depend on
// it).  Only the base cache and the two probes' caches are live at a time;
// the row-wise form keeps every probe's cache for the whole measurement
// loop.  Each entry is the row-wise form's value: the same probes (the
// reference's step rule and increment sequence), caches and quotient.
constexpr bool cdm_lm_columns = true;
template <int L>
__device__ __forceinline__ void cdm_lm_jacobian_columns(const cdm_geom& g, const float* s_cfp, const CDM_F64* s_mc,
                                                        const float* mv, cdm_jt* pj, cdm_jt* fxj, const cdm_jt* dtf,
                                                        CDM_F64* rows, int lane) {
  static_assert(L == 64, "column-wise Jacobian: one voxel per wave");
  constexpr int np1 = cdm_np + 1;
  cdm_cfp_row<cdm_jt> rowj;
  CDM_F64 jr0[cdm_vc_jobs], vc0[cdm_vc];
  cdm_cache_jobs_g<L>(pj, fxj, jr0, lane);
  Voxel_Cache_Combine(cdm_np, pj, fxj, jr0, vc0);
  // the base point's per-measurement cache values, after the factor storage
  CDM_F64* base_mvc = rows + (size_t)g.nmeas * np1 + 2 * cdm_ntri + cdm_np;
  (void)base_mvc;
  // signal, residual and analytic entries
  for (int m = lane; m < g.nmeas; m += L) {
    cdm_jt* cm = rowj.at(s_cfp, m, g.cfp_ts);
    cdm_jt d[cdm_np];
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) d[i] = (cdm_jt)0;
#if defined(CDM_MEAS_VOXEL_CACHE_PARTIAL)
    cdm_jt mvc[cdm_mvc];
    const cdm_jt sig = Predicted_Signal_MV_Update(cdm_np, pj, cm, fxj, vc0, cdm_mc_row(s_mc, m), mvc, mvc, -1);
#pragma unroll
    for (int k = 0; k < cdm_mvc; ++k) base_mvc[m * cdm_mvc + k] = (CDM_F64)mvc[k];
    Partial_Derivatives_Cached(cdm_np, pj, cm, fxj, vc0, cdm_mc_row(s_mc, m), d);
#elif defined(CDM_SIGNAL_PARTIALS)
    const cdm_jt sig = Predicted_Signal_Partials_Cached(cdm_np, pj, cm, fxj, vc0, cdm_mc_row(s_mc, m), d);
#else
    const cdm_jt sig = cdm_signal(pj, cm, fxj, vc0, cdm_mc_row(s_mc, m));
    Partial_Derivatives_Cached(cdm_np, pj, cm, fxj, vc0, cdm_mc_row(s_mc, m), d);
#endif
#pragma unroll
    for (int i = 0; i < cdm_np; ++i)
      if (!((CDM_NUMERICAL_MASK >> i) & 1u)) rows[m * np1 + i] = (CDM_F64)(cdm_lm_acc)(d[i] * dtf[i]);
    rows[m * np1 + cdm_np] = (CDM_F64)(cdm_lm_acc)(sig - (cdm_jt)mv[m]);
  }
  // The probes' voxel caches, all at once: first the jobs of the probes
  // whose parameter moves them, pair (probe, job) on its own lane, then each
  // probe's cache on its own lane, both into LDS.  Every lane building every
  // probe's cache in turn cost more than the measurement loops.
  constexpr int nDep = __builtin_popcount(cdm_lm_kdep), nJdep = __builtin_popcount(cdm_lm_kjdep);
  static_assert(2 * nDep <= L, "one probe cache per lane");
  CDM_F64* probe_jobs = base_mvc + (size_t)g.nmeas * cdm_lm_base_mvc;  // [2 nJdep][jobs]
  CDM_F64* probe_vc = probe_jobs + (size_t)2 * nJdep * cdm_vc_jobs;    // [2 nDep][cache]
  // parameters of the probe on side `side` (0: p_i + h/2, 1: that - 2 (h/2))
  auto probe_point = [&](int i, int side, cdm_jt* q) {
    cdm_jt pi = (cdm_jt)0;
#pragma unroll
    for (int k = 0; k < cdm_np; ++k)
      if (k == i) pi = pj[k];
    cdm_jt h = fabs_gpu((cdm_jt)::cdm_fd_step<cdm_jt>::rel * pi);
    h = min_gpu(max_gpu(h, (cdm_jt)::cdm_fd_step<cdm_jt>::lo), (cdm_jt)::cdm_fd_step<cdm_jt>::hi);
    const cdm_jt up = pi + h * (cdm_jt)0.5;
    const cdm_jt pp = side == 0 ? up : up - (cdm_jt)2 * (h * (cdm_jt)0.5);
#pragma unroll
    for (int k = 0; k < cdm_np; ++k) q[k] = k == i ? pp : pj[k];
  };
  // index of the k-th parameter in a mask, and the parameters of a mask below i
  auto nth_param = [](unsigned mask, int k) {
    int idx = 0, c = 0;
#pragma unroll
    for (int b = 0; b < cdm_np; ++b)
      if ((mask >> b) & 1u) {
        if (c == k) idx = b;
        ++c;
      }
    return idx;
  };
  auto rank_of = [](unsigned mask, int i) {
    int c = 0;
#pragma unroll
    for (int b = 0; b < cdm_np; ++b)
      if (b < i && ((mask >> b) & 1u)) ++c;
    return c;
  };
  {
    constexpr int TJ = 2 * nJdep * cdm_vc_jobs;
#pragma unroll
    for (int r = 0; r < (TJ + L - 1) / L; ++r) {
      const int t = lane + r * L;
      if (t < TJ) {
        const int pr = t / cdm_vc_jobs;
        cdm_jt q[cdm_np];
        probe_point(nth_param(cdm_lm_kjdep, pr / 2), pr % 2, q);
        probe_jobs[t] = Voxel_Cache_Job(t % cdm_vc_jobs, cdm_np, q, fxj);
      }
    }
  }
  cdm_wave_lds_sync();
  if (lane < 2 * nDep) {
    const int i = nth_param(cdm_lm_kdep, lane / 2), side = lane % 2;
    cdm_jt q[cdm_np];
    probe_point(i, side, q);
    CDM_F64 jq[cdm_vc_jobs], vc[cdm_vc];
    const bool jdep = (cdm_lm_kjdep >> i) & 1u;
    const int pr = 2 * rank_of(cdm_lm_kjdep, i) + side;
#pragma unroll
    for (int j = 0; j < cdm_vc_jobs; ++j) jq[j] = jdep ? probe_jobs[pr * cdm_vc_jobs + j] : jr0[j];
    Voxel_Cache_Combine(cdm_np, q, fxj, jq, vc);
#pragma unroll
    for (int c = 0; c < cdm_vc; ++c) probe_vc[lane * cdm_vc + c] = vc[c];
  }
  cdm_wave_lds_sync();
  // one column per central-difference parameter (a loop, not unrolled: the
  // parameter index selects by compares, the body is one measurement loop)
#pragma unroll 1
  for (int i = 0; i < cdm_np; ++i) {
    if (!((CDM_NUMERICAL_MASK >> i) & 1u)) continue;
    cdm_jt pi = (cdm_jt)0, di = (cdm_jt)0;
#pragma unroll
    for (int k = 0; k < cdm_np; ++k)
      if (k == i) {
        pi = pj[k];
        di = dtf[k];
      }
    cdm_jt h = fabs_gpu((cdm_jt)::cdm_fd_step<cdm_jt>::rel * pi);
    h = min_gpu(max_gpu(h, (cdm_jt)::cdm_fd_step<cdm_jt>::lo), (cdm_jt)::cdm_fd_step<cdm_jt>::hi);
    const bool dep = (cdm_lm_kdep >> i) & 1u;
    const int t0 = 2 * rank_of(cdm_lm_kdep, i);
    // the probes one after the other: the upper probe's signals go to the
    // column's LDS entries (as doubles, exact for the signal's type), the
    // lower probe's pass forms the quotient from them
#pragma unroll 1
    for (int probe = 0; probe < 2; ++probe) {
      cdm_jt q[cdm_np];
      probe_point(i, probe, q);
      CDM_F64 vc[cdm_vc];
#pragma unroll
      for (int c = 0; c < cdm_vc; ++c) vc[c] = dep ? probe_vc[(t0 + probe) * cdm_vc + c] : vc0[c];
      for (int m = lane; m < g.nmeas; m += L) {
        cdm_jt* cm = rowj.at(s_cfp, m, g.cfp_ts);
#if defined(CDM_MEAS_VOXEL_CACHE_PARTIAL)
        // only the cache entries parameter i moves are recomputed; the rest
        // are the base point's (the same values a full evaluation gives)
        cdm_jt cur[cdm_mvc], tmp[cdm_mvc];
#pragma unroll
        for (int k = 0; k < cdm_mvc; ++k) cur[k] = (cdm_jt)base_mvc[m * cdm_mvc + k];
        const cdm_jt sq = Predicted_Signal_MV_Update(cdm_np, q, cm, fxj, vc, cdm_mc_row(s_mc, m), cur, tmp, i);
#else
        const cdm_jt sq = cdm_signal(q, cm, fxj, vc, cdm_mc_row(s_mc, m));
#endif
        if (probe == 0) {
          rows[m * np1 + i] = (CDM_F64)sq;
        } else {
          const cdm_jt su = (cdm_jt)rows[m * np1 + i];
          rows[m * np1 + i] = (CDM_F64)(cdm_lm_acc)(((su - sq) / h) * di);
        }
      }
    }
  }
}
#else
constexpr bool cdm_lm_columns = false;
template <int L>
__device__ __forceinline__ void cdm_lm_jacobian_columns(const cdm_geom&, const float*, const CDM_F64*, const float*,
                                                        cdm_jt*, cdm_jt*, const cdm_jt*, CDM_F64*, int) {}
#endif

// The damped solve of a 64-lane LM group with the matrices in LDS: hs the
// packed 2 J^T J (fixed parameters' rows and columns replaced), diag its
// damped diagonal, fac the factor's storage.  The LDL^T factorisation runs
// column by column, lane i computing row i's entry of each column (the
// register form's arithmetic for every entry, in the same order); every
// lane then substitutes for the whole step.  The per-voxel registers keep
// only vectors of P values, not the P(P+1)/2-entry matrices.
template <int L>
__device__ __forceinline__ void cdm_lds_damped_solve(const CDM_F64* hs, CDM_F64* fac, const double* diag,
                                                     const double* grad, double* step, double* dg, int lane) {
  static_assert(L == 64 && cdm_np <= 64, "one row per lane");
  for (int e = lane; e < cdm_ntri; e += L) {
    int i = 0, j = 0;
    cdm::tri_ij<cdm_np>(e, i, j);
    double di = 0.0;
#pragma unroll
    for (int k = 0; k < cdm_np; ++k)
      if (k == i) di = diag[k];
    fac[e] = (i == j) ? di : (double)hs[e];
  }
  cdm_wave_lds_sync();
#pragma unroll
  for (int j = 0; j < cdm_np; ++j) {
    double dj = fac[j * (j + 1) / 2 + j];
#pragma unroll
    for (int q = 0; q < j; ++q) dj -= fac[j * (j + 1) / 2 + q] * fac[j * (j + 1) / 2 + q] * dg[q];
    dg[j] = dj;
    const double inv = 1.0 / dj;
    const int i = lane;
    if (i > j && i < cdm_np) {
      const int ri = i * (i + 1) / 2;
      double x = fac[ri + j];
#pragma unroll
      for (int q = 0; q < j; ++q) x -= fac[ri + q] * fac[j * (j + 1) / 2 + q] * dg[q];
      fac[ri + j] = x * inv;
    }
    cdm_wave_lds_sync();
  }
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) {  // L y = g
    double x = grad[i];
#pragma unroll
    for (int q = 0; q < i; ++q) x -= fac[i * (i + 1) / 2 + q] * step[q];
    step[i] = x;
  }
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) step[i] /= dg[i];  // D z = y
#pragma unroll
  for (int i = cdm_np - 1; i >= 0; --i) {  // L^T x = z
    double x = step[i];
#pragma unroll
    for (int q = i + 1; q < cdm_np; ++q) x -= fac[q * (q + 1) / 2 + i] * step[q];
    step[i] = x;
  }
  // the next factorisation's stores must not pass these reads
  cdm_wave_lds_sync();
}

// The kernel bodies take the block index as an argument so that the batched
// persistent kernel (one launch over several models' voxels) can run them on
// tiles it takes from a work queue.  A body stages the CFP table (two block
// barriers, reached by every thread), then returns early for threads past
// the last voxel; it has no barrier after that point.
//
// One voxel's fit by one voxel group (lane = the thread's lane in it).
template <int L>
__device__ __forceinline__ void cdm_levmar_voxel(const cdm_lm_args& a, int marquardt, int v, int lane, float* s_cfp,
                                                 const CDM_F64* s_mc) {
  const float* mv = a.g.meas + (size_t)v * a.g.nmeas;
  cdm_tf fx[cdm_fixp_max];
  cdm_load_fixp(a.g, v, fx);

  double p[cdm_np], t[cdm_np];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) p[i] = a.params[(size_t)v * cdm_np + i];
  cdm_inv_transform(a.c, p, t);

  double fxd[cdm_fixp_max];
#pragma unroll
  for (int k = 0; k < cdm_fixp_max; ++k) fxd[k] = (double)fx[k];
  double pcf = cdm_sse<L>(a.g, s_cfp, s_mc, mv, p, fxd, lane);

  double lambda = 0.1, olambda = 0.0;
  bool success = true;
  int iter = 0, nsucc = 0;
  // --debug: record 0 the start, record k the state after attempt k
  // (attempt, cost, lambda after it, accepted, parameters)
  const bool dbg = a.c.dbg && v == a.c.dbg_voxel;
  int dbg_slot = 0;
  if (dbg) cdm_dbg_record(a.c, 1, dbg_slot++, 0.0, pcf, lambda, 1.0, p);
  cdm_lm_acc hs[cdm_ntri];  // 2 J^T J, packed lower triangle (undamped)
  double grad[cdm_np];  // 2 J^T r
  double diag[cdm_np];  // current damped diagonal

  while (!(success && iter++ >= a.max_iters)) {
    if (success) {
      double dt[cdm_np];
      cdm_dtransform(a.c, t, dt);
      cdm_jt dtf[cdm_np], pj[cdm_np], fxj[cdm_fixp_max];
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) {
        dtf[i] = (cdm_jt)dt[i];
        pj[i] = (cdm_jt)p[i];
      }
#pragma unroll
      for (int k = 0; k < cdm_fixp_max; ++k) fxj[k] = (cdm_jt)fxd[k];
      cdm_lm_acc gl[cdm_np];
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) gl[i] = 0;
#pragma unroll
      for (int k = 0; k < cdm_ntri; ++k) hs[k] = 0;
      cdm_cfp_row<cdm_jt> rowj;
      if constexpr (cdm_lm_lds_normal<L>) {
        CDM_F64* rows = reinterpret_cast<CDM_F64*>(reinterpret_cast<char*>(s_cfp) +
                                                   cdm_lm_scratch_offset(a.g.nmeas, a.g.cfp_ts)) +
                        (size_t)(threadIdx.x / L) * cdm_lm_scratch_doubles(L, a.g.nmeas);
        CDM_F64* sums = rows + (size_t)a.g.nmeas * (cdm_np + 1);
        if constexpr (cdm_lm_columns) {
          cdm_lm_jacobian_columns<L>(a.g, s_cfp, s_mc, mv, pj, fxj, dtf, rows, lane);
        } else {
          cdm_jacobian_setup<cdm_jt> jac;
          jac.template init<L>(pj, fxj, lane);
#if defined(CDM_LM_ROW_PASSES) && defined(CDM_NUMERICAL_MASK)
          // two passes over the measurements: the signal, the analytic
          // entries and the first half of the central differences, then the
          // other half (fewer probes live at a time); the same entries
          constexpr unsigned kFirst = cdm_lm_first_half(CDM_NUMERICAL_MASK);
          constexpr unsigned kSecond = CDM_NUMERICAL_MASK & ~kFirst;
          for (int m = lane; m < a.g.nmeas; m += L) {
            cdm_jt* cm = rowj.at(s_cfp, m, a.g.cfp_ts);
            cdm_jt d[cdm_np];
            const cdm_jt r =
                jac.template eval_part<kFirst, true>(pj, cm, fxj, cdm_mc_row(s_mc, m), d) - (cdm_jt)mv[m];
#pragma unroll
            for (int i = 0; i < cdm_np; ++i)
              if (!((kSecond >> i) & 1u)) rows[m * (cdm_np + 1) + i] = (CDM_F64)(cdm_lm_acc)(d[i] * dtf[i]);
            rows[m * (cdm_np + 1) + cdm_np] = (CDM_F64)(cdm_lm_acc)r;
          }
          for (int m = lane; m < a.g.nmeas; m += L) {
            cdm_jt* cm = rowj.at(s_cfp, m, a.g.cfp_ts);
            cdm_jt d[cdm_np];
            (void)jac.template eval_part<kSecond, false>(pj, cm, fxj, cdm_mc_row(s_mc, m), d);
#pragma unroll
            for (int i = 0; i < cdm_np; ++i)
              if ((kSecond >> i) & 1u) rows[m * (cdm_np + 1) + i] = (CDM_F64)(cdm_lm_acc)(d[i] * dtf[i]);
          }
#else
          for (int m = lane; m < a.g.nmeas; m += L) {
            cdm_jt* cm = rowj.at(s_cfp, m, a.g.cfp_ts);
            cdm_jt d[cdm_np];
            const cdm_jt r = jac.eval(pj, cm, fxj, cdm_mc_row(s_mc, m), d) - (cdm_jt)mv[m];
#pragma unroll
            for (int i = 0; i < cdm_np; ++i) rows[m * (cdm_np + 1) + i] = (CDM_F64)(cdm_lm_acc)(d[i] * dtf[i]);
            rows[m * (cdm_np + 1) + cdm_np] = (CDM_F64)(cdm_lm_acc)r;
          }
#endif
        }
        cdm_wave_lds_sync();
        cdm_lds_normal_sums<L>(rows, sums, a.g.nmeas, lane);
        cdm_wave_lds_sync();
#pragma unroll
        for (int i = 0; i < cdm_np; ++i) gl[i] = (cdm_lm_acc)sums[cdm_ntri + i];
        // 2 J^T J in place (fixed parameters: identity rows and columns),
        // kept in LDS for this iteration's damped solves
        for (int e = lane; e < cdm_ntri; e += L) {
          int i = 0, j = 0;
          cdm::tri_ij<cdm_np>(e, i, j);
          int fi = 0, fj = 0;
#pragma unroll
          for (int k = 0; k < cdm_np; ++k) {
            if (k == i) fi = a.c.fixed[k];
            if (k == j) fj = a.c.fixed[k];
          }
          cdm_lm_acc h = (cdm_lm_acc)2 * (cdm_lm_acc)sums[e];
          if (fi || fj) h = (i == j) ? (cdm_lm_acc)1 : (cdm_lm_acc)0;
          sums[e] = (CDM_F64)h;
        }
        // the next iteration's stores into `rows` must not pass this one's
        // reads, and every lane sees the doubled matrix
        cdm_wave_lds_sync();
#pragma unroll
        for (int i = 0; i < cdm_np; ++i) diag[i] = (double)(cdm_lm_acc)sums[i * (i + 1) / 2 + i];
      } else {
        cdm_jacobian_setup<cdm_jt> jac;
        jac.template init<L>(pj, fxj, lane);
        for (int m = lane; m < a.g.nmeas; m += L) {
          cdm_jt* cm = rowj.at(s_cfp, m, a.g.cfp_ts);
          cdm_jt d[cdm_np];
          const cdm_jt r = jac.eval(pj, cm, fxj, cdm_mc_row(s_mc, m), d) - (cdm_jt)mv[m];
          cdm_lm_acc df[cdm_np];
#pragma unroll
          for (int i = 0; i < cdm_np; ++i) df[i] = (cdm_lm_acc)(d[i] * dtf[i]);
          const cdm_lm_acc rf = (cdm_lm_acc)r;
#pragma unroll
          for (int i = 0; i < cdm_np; ++i) gl[i] += rf * df[i];
          int k = 0;
#pragma unroll
          for (int i = 0; i < cdm_np; ++i)
#pragma unroll
            for (int j = 0; j <= i; ++j) hs[k++] += df[i] * df[j];
        }
#pragma unroll
        for (int i = 0; i < cdm_np; ++i) gl[i] = cdm::Group<L>::sum(gl[i]);
#pragma unroll
        for (int k = 0; k < cdm_ntri; ++k) hs[k] = cdm::Group<L>::sum(hs[k]);
      }
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) grad[i] = a.c.fixed[i] ? 0.0 : 2.0 * (double)gl[i];
      if constexpr (!cdm_lm_lds_normal<L>) {
        int k = 0;
#pragma unroll
        for (int i = 0; i < cdm_np; ++i)
#pragma unroll
          for (int j = 0; j <= i; ++j, ++k) {
            cdm_lm_acc h = (cdm_lm_acc)2 * hs[k];
            if (a.c.fixed[i] || a.c.fixed[j]) h = (i == j) ? (cdm_lm_acc)1 : (cdm_lm_acc)0;
            hs[k] = h;
          }
#pragma unroll
        for (int i = 0; i < cdm_np; ++i) diag[i] = hs[i * (i + 1) / 2 + i];
      }
    }
    // Damping applied cumulatively to the kept diagonal, as the reference
    // damps its shared Hessian in place (Levenberg_Marquardt.cu:530-537).
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      if (marquardt) diag[i] = ((1 + lambda) / (1 + olambda)) * diag[i];
      else diag[i] += lambda - olambda;
    }
    // FP64 LDL^T factorisation of the damped matrix (square-root-free
    // Cholesky).  Its pivots are those of Gaussian elimination without
    // pivoting, so like the reference's Gauss-Jordan solve it gives a step
    // for any matrix with non-zero pivots, and inf/NaN (a rejected step)
    // for a zero pivot (Levenberg_Marquardt.cu:336-407).
    double step[cdm_np], dg[cdm_np];
    if constexpr (cdm_lm_lds_normal<L>) {
      CDM_F64* rows = reinterpret_cast<CDM_F64*>(reinterpret_cast<char*>(s_cfp) +
                                                 cdm_lm_scratch_offset(a.g.nmeas, a.g.cfp_ts)) +
                      (size_t)(threadIdx.x / L) * cdm_lm_scratch_doubles(L, a.g.nmeas);
      const CDM_F64* hsum = rows + (size_t)a.g.nmeas * (cdm_np + 1);
      cdm_lds_damped_solve<L>(hsum, const_cast<CDM_F64*>(hsum) + cdm_ntri + cdm_np, diag, grad, step, dg, lane);
    } else {
    double fac[cdm_ntri];
    {
      int k = 0;
#pragma unroll
      for (int i = 0; i < cdm_np; ++i)
#pragma unroll
        for (int j = 0; j <= i; ++j, ++k) fac[k] = (i == j) ? diag[i] : (double)hs[k];
#pragma unroll
      for (int j = 0; j < cdm_np; ++j) {
        double dj = fac[j * (j + 1) / 2 + j];
#pragma unroll
        for (int q = 0; q < j; ++q) dj -= fac[j * (j + 1) / 2 + q] * fac[j * (j + 1) / 2 + q] * dg[q];
        dg[j] = dj;
        const double inv = 1.0 / dj;
#pragma unroll
        for (int i = j + 1; i < cdm_np; ++i) {
          double x = fac[i * (i + 1) / 2 + j];
#pragma unroll
          for (int q = 0; q < j; ++q) x -= fac[i * (i + 1) / 2 + q] * fac[j * (j + 1) / 2 + q] * dg[q];
          fac[i * (i + 1) / 2 + j] = x * inv;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {  // L y = g
      double x = grad[i];
#pragma unroll
      for (int q = 0; q < i; ++q) x -= fac[i * (i + 1) / 2 + q] * step[q];
      step[i] = x;
    }
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) step[i] /= dg[i];  // D z = y
#pragma unroll
    for (int i = cdm_np - 1; i >= 0; --i) {  // L^T x = z
      double x = step[i];
#pragma unroll
      for (int q = i + 1; q < cdm_np; ++q) x -= fac[q * (q + 1) / 2 + i] * step[q];
      step[i] = x;
    }
    }
    double tn[cdm_np];
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) tn[i] = t[i] - step[i];
    cdm_transform(a.c, p, tn);
    const double ncf = cdm_sse<L>(a.g, s_cfp, s_mc, mv, p, fxd, lane);
    bool end = false;
    success = ncf < pcf;
    if (success) {
      olambda = 0.0;
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) t[i] = tn[i];
      lambda = lambda / 10.0;
      if (2.0 * fabs(pcf - ncf) <= cdm_lm_cftol * (fabs(pcf) + fabs(ncf) + cdm_lm_eps)) end = true;
      pcf = ncf;
      ++nsucc;
    } else {
      olambda = lambda;
      lambda = lambda * 10.0;
      if (lambda > cdm_lm_ltol) end = true;
      cdm_transform(a.c, p, t);
    }
    if (dbg) {
      cdm_dbg_record(a.c, 1, dbg_slot, (double)dbg_slot, success ? pcf : ncf, lambda, success ? 1.0 : 0.0, p);
      ++dbg_slot;
    }
    if (end) break;
  }
  FixConstraintsLM(cdm_np, p);
  // Every lane of the group holds the same results and stores them (the
  // same values to the same addresses): no branch on the lane index, which
  // in the work-queue loop would let the compiler split the group's lanes
  // across loop iterations.
  (void)lane;
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) a.params[(size_t)v * cdm_np + i] = (float)p[i];
  if (a.iters_out) a.iters_out[v] = nsucc;
  if (a.cost_out) a.cost_out[v] = (float)pcf;
}

// Static mapping: voxel group g of block `block` fits voxel block * groups + g.
template <int L>
__device__ __forceinline__ void cdm_levmar_body(const cdm_lm_args& a, int marquardt, int block) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = block * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  cdm_levmar_voxel<L>(a, marquardt, v, lane, s_cfp, s_mc);
}

// Work queue: LM iteration counts vary widely between voxels (from a few to
// the --iterLevMar cap), so a group that finishes takes the next voxel from
// a global counter instead of idling until the other groups of its wave
// finish theirs.  A voxel's arithmetic does not depend on which group fits
// it.  The grid holds only the blocks that can be resident at once; every
// group leaves when the counter passes the last voxel, and after at most
// nvox voxels in any case.
//
// The loop body has no branch on the lane index: every lane of the group
// executes the fetch (lane 0 adds 1, the others 0) and the broadcast takes
// lane 0's ticket.  A lane-0-only fetch at the top of the loop together
// with the lane-0-only stores at the bottom let the compiler form an inner
// loop in which the other lanes came back to the broadcast without lane 0
// having fetched, re-reading a stale index.
template <int L>
__device__ __forceinline__ void cdm_levmar_queue_body(const cdm_lm_args& a, int marquardt) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int lane = threadIdx.x % L;
  for (int k = 0; k < a.g.nvox; ++k) {
    const int ticket = atomicAdd(a.queue, lane == 0 ? 1 : 0);
    const int v = cdm::Group<L>::bcast(ticket, 0);
    if (v >= a.g.nvox) break;
    cdm_levmar_voxel<L>(a, marquardt, v, lane, s_cfp, s_mc);
  }
}

__global__ void cdm_queue_reset_kernel(int* q) { *q = 0; }

// W: resident 256-thread blocks per CU the register allocation must allow
// (1, 2 or 4; the compiler caps registers and spills beyond them).  Chosen
// per model and measurement count by the autotuner with L.
template <int L, int W>
__global__ void __launch_bounds__(cdm_block, W) cdm_levmar_queue_kernel(cdm_lm_args a, int marquardt) {
  cdm_levmar_queue_body<L>(a, marquardt);
}

template <int L, int W>
__global__ void __launch_bounds__(cdm_block, W) cdm_levmar_kernel(cdm_lm_args a, int marquardt) {
  cdm_levmar_body<L>(a, marquardt, blockIdx.x);
}

// --------------------------------------------------------------------- MCMC
// MCMC.cu:327-683: single-site Metropolis-within-Gibbs with adaptive
// proposal widths.  One launch runs `count` jumps of one phase (burn-in or
// recording) starting at jump `begin` of that phase; the chain state is kept
// in FP64 in `state` between launches, so a phase may be split into jump
// blocks (captured in a HIP graph by the host) without changing any draw.
// Phase initialisation (begin == 0) follows the reference's kernel prologue
// (propSD = param/10 in burn-in, clamp to bounds, priors, tau); the phase's
// last block writes params back.  Draws: Philox counters (philox.h) with
// jump = phase offset + iteration.

struct cdm_chain {  // FP64 chain state per voxel
  double p[cdm_np], prior[cdm_np], sd[cdm_np];
  double energy, tau, tau_sd;
  int nacc[cdm_np], nrej[cdm_np];
  int tau_acc, tau_rej;
};

// Per-parameter chain state of one voxel while a launch runs, kept in LDS
// (one slot per voxel group of the block) rather than registers: a step
// indexes it by the proposed parameter, and registers are the kernel's
// occupancy limit.  All lanes of the group write identical values.
// A large voxel cache (16 values or more) keeps the current state's copy,
// and its cache jobs, in the voxel's LDS slot rather than in registers: a
// step holds the proposal's cache in registers, and the current one is read
// to select or copy entries and written on acceptance.
constexpr bool cdm_vc_in_lds = cdm_vc >= 16;

struct cdm_vslot {
  double p[cdm_np], prior[cdm_np], sd[cdm_np];
  CDM_F64 vc[cdm_vc_in_lds ? cdm_vc : 1];
  CDM_F64 jr[cdm_vc_in_lds && cdm_vc_jobs > 0 ? cdm_vc_jobs : 1];
  int nacc[cdm_np], nrej[cdm_np];
};

// Byte offset of the slots in the dynamic LDS (after the CFP table and the
// measurement cache, 16-byte aligned).
__host__ __device__ __forceinline__ size_t cdm_vslot_offset(int nmeas, int cfp_ts) {
  const size_t b = (size_t)cdm_mc_offset(nmeas, cfp_ts) * sizeof(float) + (size_t)nmeas * cdm_mc * sizeof(CDM_F64);
  return (b + 15) / 16 * 16;
}

// The two per-measurement cache rows of each voxel slot in the dynamic LDS,
// after the slots: row_stride values apart (a row and its padding), 2 rows
// per slot.  Rows are rebuilt at the start of every launch, so they need
// not outlive it.
__host__ __device__ __forceinline__ size_t cdm_mcmc_rows_offset(int nmeas, int cfp_ts, int L) {
  const size_t b = cdm_vslot_offset(nmeas, cfp_ts) + (size_t)(cdm_block / L) * sizeof(cdm_vslot);
  return (b + 15) / 16 * 16;
}

// Values per LDS cache row: at least nmeas * cdm_mvc, and = 2 mod 16, so
// that the 16 four-lane groups of a wave (slots two rows apart) start on
// different banks.
__host__ __device__ __forceinline__ int cdm_mcmc_row_stride(int nmeas) {
  const int n = nmeas * (cdm_mvc > 0 ? cdm_mvc : 1);
  return n + ((2 - n % 16) + 16) % 16;
}

struct cdm_mcmc_args {
  cdm_geom g;
  float* params;         // [nvox][NP] in/out
  cdm_tf* mvc;           // [nvox][2][nmeas][cdm_mvc] per-measurement caches (cdm_mvc > 0, rows_lds == 0)
  int rows_lds;          // the cache rows live in LDS (cdm_mcmc_rows_offset) instead of mvc
  int row_stride;        // values between an LDS slot's two rows
  cdm_chain* state;      // [nvox]
  float* samples;        // [nvox][NP][nsamples] (recording)
  float* tau_samples;    // [nvox][nsamples] (Rician, recording)
  int nsamples, sampleevery, upe;
  int recording, update_prop;
  int begin, count, phase_len;
  unsigned jump_offset;  // Philox jump index of iteration 0 of this phase
  unsigned long long seed;
  long long voxel_offset;
  cdm_consts c;
};

__device__ __forceinline__ double cdm_logI0(double x) {
  // Numerical Recipes polynomial fits, as MCMC.cu:52-68.
  double y, b = fabs(x);
  if (b < 3.75) {
    double q = x / 3.75;
    q *= q;
    y = 1.0 + q * (3.5156229 + q * (3.0899424 + q * (1.2067492 + q * (0.2659732 + q * (0.0360768 + q * 0.0045813)))));
    y = log(y);
  } else {
    const double q = 3.75 / b;
    y = b + log((0.39894228 +
                 q * (0.01328592 +
                      q * (0.00225319 +
                           q * (-0.00157565 +
                                q * (0.00916281 + q * (-0.02057706 + q * (0.02635537 + q * (-0.01647633 + q * 0.00392377)))))))) /
                sqrt(b));
  }
  return y;
}

template <int L, bool RICIAN>
__device__ __forceinline__ double cdm_neg_loglik(const cdm_geom& g, const float* s_cfp, const CDM_F64* s_mc,
                                                 const float* mv, const double* p, double tau, cdm_tf* fx,
                                                 const CDM_F64* vc, int lane) {
  cdm_tf pf[cdm_np];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) pf[i] = (cdm_tf)p[i];
  cdm_cfp_row<cdm_tf> row;
  double acc = 0.0;
  for (int m = lane; m < g.nmeas; m += L) {
    const cdm_tf s = cdm_signal(pf, row.at(s_cfp, m, g.cfp_ts), fx, vc, cdm_mc_row(s_mc, m));
    if constexpr (RICIAN) {
      const double y = mv[m], sd = (double)s;
      acc += log(y) + (-0.5 * tau * (y * y + sd * sd) + cdm_logI0(tau * sd * y));
    } else {
      const cdm_tf r = s - (cdm_tf)mv[m];
      acc += (double)(r * r);
    }
  }
  acc = cdm::Group<L>::sum(acc);
  if constexpr (RICIAN) return -g.nmeas * log(tau) - acc;
  else return (g.nmeas / 2.0) * log(acc / 2.0);
}

// The same with the per-measurement caches: MODE 1 reads them from `row`
// ([nmeas][cdm_mvc] of this voxel), MODE 2 computes them, stores them in
// `row` and uses them.  Each lane reads only what it wrote itself.
template <int L, bool RICIAN, int MODE>
__device__ __forceinline__ double cdm_neg_loglik_mv(const cdm_geom& g, const float* s_cfp, const CDM_F64* s_mc,
                                                    const float* mv, const double* p, double tau, cdm_tf* fx,
                                                    const CDM_F64* vc, cdm_tf* row, int lane,
                                                    const cdm_tf* row_cur = nullptr, int changed = -1) {
#ifdef CDM_MEAS_VOXEL_CACHE
  cdm_tf pf[cdm_np];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) pf[i] = (cdm_tf)p[i];
  cdm_cfp_row<cdm_tf> crow;
  double acc = 0.0;
  for (int m = lane; m < g.nmeas; m += L) {
    cdm_tf* cfpm = crow.at(s_cfp, m, g.cfp_ts);
    const CDM_F64* mcm = cdm_mc_row(s_mc, m);
    cdm_tf mvc[cdm_mvc];
    cdm_tf s;
    if constexpr (MODE == 2) {
#if defined(CDM_MEAS_VOXEL_CACHE_PARTIAL)
      cdm_tf cur[cdm_mvc];
#pragma unroll
      for (int k = 0; k < cdm_mvc; ++k) cur[k] = changed >= 0 ? row_cur[m * cdm_mvc + k] : (cdm_tf)0;
      s = Predicted_Signal_MV_Update(cdm_np, pf, cfpm, fx, vc, mcm, cur, mvc, changed);
#elif defined(CDM_MEAS_VOXEL_CACHE_FILL)
      s = Predicted_Signal_MV_Fill(cdm_np, pf, cfpm, fx, vc, mcm, mvc);
#else
      Meas_Voxel_Cache(cdm_np, pf, cfpm, fx, vc, mcm, mvc);
      s = Predicted_Signal_MV(cdm_np, pf, cfpm, fx, vc, mcm, mvc);
#endif
#pragma unroll
      for (int k = 0; k < cdm_mvc; ++k) row[m * cdm_mvc + k] = mvc[k];
    } else {
#pragma unroll
      for (int k = 0; k < cdm_mvc; ++k) mvc[k] = row[m * cdm_mvc + k];
      s = Predicted_Signal_MV(cdm_np, pf, cfpm, fx, vc, mcm, mvc);
    }
    if constexpr (RICIAN) {
      const double y = mv[m], sd = (double)s;
      acc += log(y) + (-0.5 * tau * (y * y + sd * sd) + cdm_logI0(tau * sd * y));
    } else {
      const cdm_tf r = s - (cdm_tf)mv[m];
      acc += (double)(r * r);
    }
  }
  acc = cdm::Group<L>::sum(acc);
  if constexpr (RICIAN) return -g.nmeas * log(tau) - acc;
  else return (g.nmeas / 2.0) * log(acc / 2.0);
#else
  return cdm_neg_loglik<L, RICIAN>(g, s_cfp, s_mc, mv, p, tau, fx, vc, lane);
#endif
}

// Voxel cache of a double parameter vector, computed in the plug-in's precision.
__device__ __forceinline__ void cdm_prepare_d(const double* p, cdm_tf* fx, CDM_F64* vc) {
  cdm_tf pf[cdm_np];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) pf[i] = (cdm_tf)p[i];
  cdm_prepare(pf, fx, vc);
}
template <int L>
__device__ __forceinline__ void cdm_prepare_dg(const double* p, cdm_tf* fx, CDM_F64* vc, int lane) {
  cdm_tf pf[cdm_np];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) pf[i] = (cdm_tf)p[i];
  cdm_prepare_g<L>(pf, fx, vc, lane);
}

template <class TP>
__device__ __forceinline__ double cdm_prior(const cdm_consts& c, int i, TP* p, int nmeas, float* cfp, cdm_tf* fx) {
  const int pt = c.prior_types[i];
  if (pt == cdm_GAUSS) {
    const double sd2 = c.pb[i] * c.pb[i];
    return (p[i] - c.pa[i]) * (p[i] - c.pa[i]) / (2 * sd2);
  } else if (pt == cdm_GAMMA) {
    return (1.0 - c.pa[i]) * log(p[i]) + c.pb[i] * p[i];
  } else if (pt == cdm_ARD) {
    return c.pa[i] * log(p[i]);
  } else if (pt == cdm_SIN) {
    return -log(fabs(sin(p[i]) / 2.0));
  } else if (pt == cdm_CUSTOM) {
    cdm_tf pf[cdm_np];
#pragma unroll
    for (int k = 0; k < cdm_np; ++k) pf[k] = (cdm_tf)p[k];
    cdm_cfp_row<cdm_tf> row;
    return (double)custom_priors(i, pf, nmeas, row.at(cfp, 0, 0), fx);
  }
  return 0.0;
}

template <int L, bool RICIAN>
__device__ __forceinline__ void cdm_mcmc_body(const cdm_mcmc_args& a, int block) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = block * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  const float* mv = a.g.meas + (size_t)v * a.g.nmeas;
  cdm_tf fx[cdm_fixp_max];
  cdm_load_fixp(a.g, v, fx);
  const uint64_t gvox = (uint64_t)(a.voxel_offset + v);
  cdm_chain* st = &a.state[v];

  cdm_vslot* vs = reinterpret_cast<cdm_vslot*>(reinterpret_cast<char*>(s_cfp) +
                                               cdm_vslot_offset(a.g.nmeas, a.g.cfp_ts)) +
                  threadIdx.x / L;
  double* p = vs->p;
  double* prior = vs->prior;
  double* sd = vs->sd;
  int* nacc = vs->nacc;
  int* nrej = vs->nrej;
  double energy, tau = 0.0, tau_sd = 0.0;
  int tau_acc = 0, tau_rej = 0;

  if (a.begin == 0) {
    // Phase prologue (MCMC.cu:411-494).
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      p[i] = a.recording ? st->p[i] : (double)a.params[(size_t)v * cdm_np + i];
      nacc[i] = nrej[i] = 0;
      sd[i] = a.recording ? st->sd[i] : p[i] / 10.0;
    }
    if (RICIAN && a.recording) {
      tau = st->tau;
      tau_sd = st->tau_sd;
    }
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      const int bt = a.c.bound_types[i];
      if (bt == cdm_BMIN) {
        if (p[i] < a.c.bmin[i]) p[i] = a.c.bmin[i];
      } else if (bt == cdm_BMAX) {
        if (p[i] > a.c.bmax[i]) p[i] = a.c.bmax[i];
      } else if (bt == cdm_BMINMAX) {
        if (p[i] < a.c.bmin[i]) p[i] = a.c.bmin[i];
        else if (p[i] > a.c.bmax[i]) p[i] = a.c.bmax[i];
      }
    }
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) prior[i] = cdm_prior(a.c, i, p, a.g.nmeas, s_cfp, fx);
    if (RICIAN && !a.recording) {
      // tau = 1 / variance of the residuals (MCMC.cu:193-251)
      cdm_tf pf[cdm_np];
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) pf[i] = (cdm_tf)p[i];
      CDM_F64 vc0[cdm_vc];
      cdm_prepare(pf, fx, vc0);
      cdm_cfp_row<cdm_tf> row;
      double s1 = 0.0;
      for (int m = lane; m < a.g.nmeas; m += L)
        s1 += (double)(cdm_signal(pf, row.at(s_cfp, m, a.g.cfp_ts), fx, vc0, cdm_mc_row(s_mc, m)) - (cdm_tf)mv[m]);
      const double mean = cdm::Group<L>::sum(s1) / a.g.nmeas;
      double s2 = 0.0;
      for (int m = lane; m < a.g.nmeas; m += L) {
        const double e =
            mean - (double)(cdm_signal(pf, row.at(s_cfp, m, a.g.cfp_ts), fx, vc0, cdm_mc_row(s_mc, m)) -
                            (cdm_tf)mv[m]);
        s2 += e * e;
      }
      tau = 1.0 / (cdm::Group<L>::sum(s2) / a.g.nmeas);
      tau_sd = tau / 2.0;
    }
    double tp = 0.0;
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) tp += prior[i];
    CDM_F64 vc0[cdm_vc];
    cdm_prepare_d(p, fx, vc0);
    energy = tp + cdm_neg_loglik<L, RICIAN>(a.g, s_cfp, s_mc, mv, p, tau, fx, vc0, lane);
  } else {
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      p[i] = st->p[i];
      prior[i] = st->prior[i];
      sd[i] = st->sd[i];
      nacc[i] = st->nacc[i];
      nrej[i] = st->nrej[i];
    }
    energy = st->energy;
    tau = st->tau;
    tau_sd = st->tau_sd;
    tau_acc = st->tau_acc;
    tau_rej = st->tau_rej;
  }

  // voxel cache of the current state (a deterministic function of p); with
  // a job-dependency mask narrower than the cache's, the current job values
  // are kept too, so that a step in a parameter the jobs do not depend on
  // only recombines them
#if defined(CDM_VOXEL_CACHE) && defined(CDM_VOXEL_CACHE_JOBS)
  constexpr bool job_split = cdm_vcj_dep != cdm_vc_dep;
#else
  constexpr bool job_split = false;
#endif
  CDM_F64 vc_reg[cdm_vc_in_lds ? 1 : cdm_vc];
  CDM_F64 jr_reg[!cdm_vc_in_lds && job_split ? cdm_vc_jobs : 1];
  CDM_F64* vc_cur = cdm_vc_in_lds ? vs->vc : vc_reg;
  CDM_F64* jr_cur = cdm_vc_in_lds ? vs->jr : jr_reg;
  if constexpr (job_split) {
#if defined(CDM_VOXEL_CACHE) && defined(CDM_VOXEL_CACHE_JOBS)
    cdm_tf pf[cdm_np];
#pragma unroll
    for (int k = 0; k < cdm_np; ++k) pf[k] = (cdm_tf)p[k];
    cdm_cache_jobs_g<L>(pf, fx, jr_cur, lane);
    Voxel_Cache_Combine(cdm_np, pf, fx, jr_cur, vc_cur);
#endif
  } else {
    cdm_prepare_dg<L>(p, fx, vc_cur, lane);
  }
  // per-measurement caches: two rows per voxel, `cur` holds the current
  // state's; in the block's LDS when they fit there, else in global memory.
  // Each lane reads only the entries it wrote itself, so neither needs a
  // barrier; row 0 is filled at the start of every launch.
  cdm_tf* mvc_rows = nullptr;
  size_t rs = 0;
  if constexpr (cdm_mvc > 0) {
    if (a.rows_lds) {
      rs = (size_t)a.row_stride;
      mvc_rows = reinterpret_cast<cdm_tf*>(reinterpret_cast<char*>(s_cfp) +
                                           cdm_mcmc_rows_offset(a.g.nmeas, a.g.cfp_ts, L)) +
                 (size_t)(threadIdx.x / L) * 2 * rs;
    } else {
      rs = (size_t)a.g.nmeas * cdm_mvc;
      mvc_rows = a.mvc + (size_t)v * 2 * rs;
    }
  }
  int cur = 0;
  if constexpr (cdm_mvc > 0)
    (void)cdm_neg_loglik_mv<L, RICIAN, 2>(a.g, s_cfp, s_mc, mv, p, tau, fx, vc_cur, mvc_rows, lane);
  const int end = a.begin + a.count;
  for (int it = a.begin; it < end; ++it) {
    const uint32_t jump = a.jump_offset + (uint32_t)it;
    // The jump's draws do not depend on the chain, so the group computes them
    // up front, slot s on lane s % L (round s / L), and a step takes its slot's
    // pair from that lane: each lane runs ceil(slots / L) Philox blocks and
    // Box-Muller transforms instead of all of them.  Same values as drawing
    // slot by slot.
    // (One lane per voxel draws each slot where it is used.)
    constexpr int nslots = cdm_np + (RICIAN ? 1 : 0);
    constexpr int rounds = L > 1 ? (nslots + L - 1) / L : 1;
    double dz[rounds], du[rounds];
    if constexpr (L > 1) {
#pragma unroll
      for (int r = 0; r < rounds; ++r) {
        int slot = lane + r * L;
        slot = slot < nslots ? slot : nslots - 1;
        const cdm::McmcDraw d = cdm::mcmc_draw(a.seed, gvox, jump, (uint32_t)slot);
        dz[r] = d.normal;
        du[r] = d.uniform;
      }
    }
    auto draw_of = [&](int slot) {
      if constexpr (L == 1) return cdm::mcmc_draw(a.seed, gvox, jump, (uint32_t)slot);
      double z = dz[0], u = du[0];
#pragma unroll
      for (int r = 1; r < rounds; ++r)
        if (slot / L == r) {
          z = dz[r];
          u = du[r];
        }
      cdm::McmcDraw d;
      d.normal = cdm::Group<L>::bcast(z, slot % L);
      d.uniform = cdm::Group<L>::bcast(u, slot % L);
      return d;
    };
    if constexpr (RICIAN) {
      const cdm::McmcDraw d = draw_of(cdm_np);
      const double old = tau;
      tau = tau + d.normal * tau_sd;
      if (tau > 0.0) {
        double tp = 0.0;
#pragma unroll
        for (int i = 0; i < cdm_np; ++i) tp += prior[i];
        const double e_new =
            tp + (cdm_mvc > 0 ? cdm_neg_loglik_mv<L, true, 1>(a.g, s_cfp, s_mc, mv, p, tau, fx, vc_cur,
                                                              mvc_rows + cur * rs, lane)
                              : cdm_neg_loglik<L, true>(a.g, s_cfp, s_mc, mv, p, tau, fx, vc_cur, lane));
        if (exp(energy - e_new) > d.uniform) {
          ++tau_acc;
          energy = e_new;
        } else {
          ++tau_rej;
          tau = old;
        }
      } else {
        ++tau_rej;
        tau = old;
      }
    }
#pragma unroll 1
    for (int i = 0; i < cdm_np; ++i) {
      const cdm::McmcDraw d = draw_of(i);
      const double old = p[i];
      if (!a.c.fixed[i]) p[i] = p[i] + d.normal * sd[i];
      bool ok = ConstraintsMCMC(cdm_np, p);
      const int bt = a.c.bound_types[i];
      if ((bt == cdm_BMIN || bt == cdm_BMINMAX) && p[i] < a.c.bmin[i]) ok = false;
      if ((bt == cdm_BMAX || bt == cdm_BMINMAX) && p[i] > a.c.bmax[i]) ok = false;
      if (ok) {
        // the voxel cache changes only with the parameters it depends on
        const bool dep = (cdm_vc_dep >> i) & 1u;
        const bool mdep = (cdm_mvc_dep >> i) & 1u;
        const bool jdep = (cdm_vcj_dep >> i) & 1u;
        CDM_F64 vc_new[cdm_vc];
        CDM_F64 jr_new[job_split ? cdm_vc_jobs : 1];
        if (dep) {
          if constexpr (job_split) {
#if defined(CDM_VOXEL_CACHE) && defined(CDM_VOXEL_CACHE_JOBS)
            cdm_tf pf[cdm_np];
#pragma unroll
            for (int k = 0; k < cdm_np; ++k) pf[k] = (cdm_tf)p[k];
#ifdef CDM_VOXEL_CACHE_PARTIAL
            // only what parameter i moves: its jobs, its cache entries
            if (jdep) {
              cdm_cache_jobs_part_g<L>(pf, fx, jr_cur, jr_new, lane, i);
            } else {
#pragma unroll
              for (int j = 0; j < cdm_vc_jobs; ++j) jr_new[j] = jr_cur[j];
            }
            Voxel_Cache_Update(cdm_np, pf, fx, jr_new, vc_cur, vc_new, i);
#else
            if (jdep) {
              cdm_cache_jobs_g<L>(pf, fx, jr_new, lane);
            } else {
#pragma unroll
              for (int j = 0; j < cdm_vc_jobs; ++j) jr_new[j] = jr_cur[j];
            }
            Voxel_Cache_Combine(cdm_np, pf, fx, jr_new, vc_new);
#endif
#endif
          } else {
            cdm_prepare_dg<L>(p, fx, vc_new, lane);
          }
        }
        // element-wise select: a pointer select between two register arrays
        // would place both in scratch memory
        CDM_F64 vcu[cdm_vc];
#pragma unroll
        for (int k = 0; k < cdm_vc; ++k) vcu[k] = dep ? vc_new[k] : vc_cur[k];
        double like;
        if constexpr (cdm_mvc > 0) {
          like = mdep ? cdm_neg_loglik_mv<L, RICIAN, 2>(a.g, s_cfp, s_mc, mv, p, tau, fx, vcu,
                                                        mvc_rows + (1 - cur) * rs, lane, mvc_rows + cur * rs, i)
                      : cdm_neg_loglik_mv<L, RICIAN, 1>(a.g, s_cfp, s_mc, mv, p, tau, fx, vcu, mvc_rows + cur * rs,
                                                        lane);
        } else {
          like = cdm_neg_loglik<L, RICIAN>(a.g, s_cfp, s_mc, mv, p, tau, fx, vcu, lane);
        }
        const double old_prior = prior[i];
        prior[i] = cdm_prior(a.c, i, p, a.g.nmeas, s_cfp, fx);
        double tp = 0.0;
#pragma unroll
        for (int k = 0; k < cdm_np; ++k) tp += prior[k];
        const double e_new = tp + like;
        if (exp(energy - e_new) > d.uniform) {
          ++nacc[i];
          energy = e_new;
#pragma unroll
          for (int k = 0; k < cdm_vc; ++k) vc_cur[k] = vcu[k];
          if constexpr (job_split) {
#pragma unroll
            for (int j = 0; j < cdm_vc_jobs; ++j) jr_cur[j] = (dep && jdep) ? jr_new[j] : jr_cur[j];
          }
          if (mdep) cur = 1 - cur;
        } else {
          ++nrej[i];
          p[i] = old;
          prior[i] = old_prior;
        }
      } else {
        ++nrej[i];
        p[i] = old;
      }
    }
    // --debug: one record per jump (jump, energy, tau, acceptances since the
    // last proposal update, parameters)
    if (a.c.dbg && v == a.c.dbg_voxel) {
      int acc = 0;
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) acc += nacc[i];
      cdm_dbg_record(a.c, 2, (int)jump, (double)jump, energy, tau, (double)acc, p);
    }
    // every lane of the group holds the same state and stores it (no branch
    // on the lane index inside the jump loop)
    if (a.recording && (it % a.sampleevery) == 0) {
      const int k = it / a.sampleevery;
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) a.samples[((size_t)v * cdm_np + i) * a.nsamples + k] = (float)p[i];
      if (RICIAN) a.tau_samples[(size_t)v * a.nsamples + k] = (float)tau;
    }
    if ((!a.recording || a.update_prop) && it > 0 && (it % a.upe) == 0) {
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) {
        sd[i] *= sqrt((nacc[i] + 1.0) / (nrej[i] + 1.0));
        sd[i] = fmin(sd[i], 1e10);
        nacc[i] = nrej[i] = 0;
      }
      if (RICIAN) {
        tau_sd *= sqrt((tau_acc + 1.0) / (tau_rej + 1.0));
        tau_sd = fmin(tau_sd, 1e10);
        tau_acc = tau_rej = 0;
      }
    }
  }

  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < cdm_np; ++i) {
      st->p[i] = p[i];
      st->prior[i] = prior[i];
      st->sd[i] = sd[i];
      st->nacc[i] = nacc[i];
      st->nrej[i] = nrej[i];
    }
    st->energy = energy;
    st->tau = tau;
    st->tau_sd = tau_sd;
    st->tau_acc = tau_acc;
    st->tau_rej = tau_rej;
    if (end == a.phase_len) {
#pragma unroll
      for (int i = 0; i < cdm_np; ++i) a.params[(size_t)v * cdm_np + i] = (float)p[i];
      if (RICIAN && !a.recording && a.tau_samples) a.tau_samples[(size_t)v * a.nsamples] = (float)tau;
    }
  }
}

template <int L, bool RICIAN, int W>
__global__ void __launch_bounds__(cdm_block, W) cdm_mcmc_kernel(cdm_mcmc_args a) {
  cdm_mcmc_body<L, RICIAN>(a, blockIdx.x);
}

// ------------------------------------------------- predicted signal, BIC/AIC
// getPredictedSignal.cu:22-92 and BIC_AIC.cu:61-201: the arithmetic mean of
// the samples (angles included) drives one forward evaluation per
// measurement.

struct cdm_post_args {
  cdm_geom g;
  int nsamples;
  const float* samples;      // [nvox][NP][nsamples]
  const float* tau_samples;  // [nvox][nsamples] (Rician BIC)
  float* out;                // predicted [nvox][nmeas]
  float* bic;                // [nvox]
  float* aic;                // [nvox]
};

__device__ __forceinline__ void cdm_mean_sample(const cdm_post_args& a, int v, double* mp) {
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) {
    const float* s = a.samples + ((size_t)v * cdm_np + i) * a.nsamples;
    if (a.nsamples > 1) {
      double acc = 0.0;
      for (int k = 0; k < a.nsamples; ++k) acc += s[k];
      mp[i] = acc / a.nsamples;
    } else {
      mp[i] = s[0];
    }
  }
}

template <int L>
__global__ void __launch_bounds__(cdm_block) cdm_predicted_kernel(cdm_post_args a) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = blockIdx.x * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  cdm_tf fx[cdm_fixp_max];
  cdm_load_fixp(a.g, v, fx);
  double mp[cdm_np];
  cdm_mean_sample(a, v, mp);
  cdm_tf pf[cdm_np];
#pragma unroll
  for (int i = 0; i < cdm_np; ++i) pf[i] = (cdm_tf)mp[i];
  cdm_cfp_row<cdm_tf> row;
  CDM_F64 vc[cdm_vc];
  cdm_prepare(pf, fx, vc);
  for (int m = lane; m < a.g.nmeas; m += L)
    a.out[(size_t)v * a.g.nmeas + m] = (float)cdm_signal(pf, row.at(s_cfp, m, a.g.cfp_ts), fx, vc, cdm_mc_row(s_mc, m));
}

template <int L, bool RICIAN>
__global__ void __launch_bounds__(cdm_block) cdm_bic_aic_kernel(cdm_post_args a) {
  extern __shared__ float s_cfp[];
  const CDM_F64* s_mc = cdm_stage_cfp(a.g, s_cfp);
  const int v = blockIdx.x * (cdm_block / L) + threadIdx.x / L;
  const int lane = threadIdx.x % L;
  if (v >= a.g.nvox) return;
  const float* mv = a.g.meas + (size_t)v * a.g.nmeas;
  cdm_tf fx[cdm_fixp_max];
  cdm_load_fixp(a.g, v, fx);
  double mp[cdm_np];
  cdm_mean_sample(a, v, mp);
  double tau = 0.0;
  if (RICIAN) {
    const float* ts = a.tau_samples + (size_t)v * a.nsamples;
    if (a.nsamples > 1) {
      double acc = 0.0;
      for (int k = 0; k < a.nsamples; ++k) acc += ts[k];
      tau = acc / a.nsamples;
    } else {
      tau = ts[0];
    }
  }
  CDM_F64 vc[cdm_vc];
  cdm_prepare_d(mp, fx, vc);
  const double nll = cdm_neg_loglik<L, RICIAN>(a.g, s_cfp, s_mc, mv, mp, tau, fx, vc, lane);
  // nll is -loglik (Rician) or (M/2) log(SSE/2) (Gaussian): loglik = -nll.
  const int np = cdm_np + (RICIAN ? 1 : 0);
  if (lane == 0) {
    a.aic[v] = (float)(2.0 * nll + np * 2.0);
    a.bic[v] = (float)(2.0 * nll + np * log((double)a.g.nmeas));
  }
}
