// Host-side launchers of one model's kernels (included in the model
// namespace after engine.h and args.inl).  Grid: ceil(nvox / (256/L)) blocks
// of 256 threads; dynamic LDS holds the staged CFP table and the measurement
// cache.

#define CDM_DISPATCH_L(L, CALL)            \
  switch (L) {                             \
    case 1: { constexpr int LG = 1; CALL; } break;   \
    case 4: { constexpr int LG = 4; CALL; } break;   \
    case 16: { constexpr int LG = 16; CALL; } break; \
    case 64: { constexpr int LG = 64; CALL; } break; \
    default: return hipErrorInvalidValue;  \
  }

#define CDM_DISPATCH_W(W, CALL)            \
  switch (W) {                             \
    case 1: { constexpr int WG = 1; CALL; } break; \
    case 2: { constexpr int WG = 2; CALL; } break; \
    case 4: { constexpr int WG = 4; CALL; } break; \
    default: return hipErrorInvalidValue;  \
  }

template <int L>
static void op_gridsearch_split(const cdm::Geom& g, const cdm_gs_args& a, const cdm_gs_split& sp, hipStream_t s) {
  const dim3 grid = op_grid(g.nvox, L), grid_s(grid.x, (unsigned)sp.S);
  hipLaunchKernelGGL(cdm_gridsearch_split_kernel<L>, grid_s, dim3(cdm_block), op_lds(g), s, a, sp);
  hipLaunchKernelGGL(cdm_gridsearch_merge_kernel<L>, grid, dim3(cdm_block), op_lds(g), s, a, sp);
  if constexpr (sizeof(cdm_tf) != sizeof(double)) {
    hipLaunchKernelGGL(cdm_gridsearch_rescan_kernel<L>, grid_s, dim3(cdm_block), op_lds(g), s, a, sp);
    hipLaunchKernelGGL(cdm_gridsearch_pick_kernel<L>, grid, dim3(cdm_block), 0, s, a, sp);
  }
}

// S > 1: the search split over S slices of the combinations, in the scratch
// `split` (cdm::gs_split_bytes(nvox, S) bytes): the screen
// (cdm_gridsearch_split_kernel), the merge that finishes the voxels that are
// not crowded (cdm_gridsearch_merge_kernel), the crowded voxels' FP64 rescan
// (cdm_gridsearch_rescan_kernel) and their choice (cdm_gridsearch_pick_kernel).
static hipError_t op_gridsearch(int L, const cdm::Geom& g, float* params, int ngp, const int* gparams, int ncombs,
                                const float* grid_f, const double* grid_d, const int* comb_id, const cdm::Consts& c,
                                int* kept, int S, void* split, hipStream_t s) {
  if (g.nvox <= 0) return hipSuccess;
  if (ngp > cdm_np || g.cfp_ts > cdm_cfp_max || g.fixp_ts > cdm_fixp_max) return hipErrorInvalidValue;
  cdm_gs_args a;
  a.g = op_geom(g);
  a.params = params;
  a.ngp = ngp;
  for (int i = 0; i < cdm_np; ++i) a.gparams[i] = i < ngp ? gparams[i] : -1;
  a.ncombs = ncombs;
  a.grid_f = grid_f;
  a.grid_d = grid_d;
  a.comb_id = comb_id;
  a.c = op_consts(c);
  a.kept_out = kept;
  if (S > 1 && split) {
    const size_t n = (size_t)g.nvox * S;
    cdm_gs_split sp;
    sp.S = S;
    sp.rescan_c = (double*)split;
    sp.list_c = (float*)(sp.rescan_c + n);
    sp.list_i = (int*)(sp.list_c + n * CDM_GS_TOPK);
    sp.rescan_i = sp.list_i + n * CDM_GS_TOPK;
    sp.crowd = sp.rescan_i + n;
    CDM_DISPATCH_L(L, op_gridsearch_split<LG>(g, a, sp, s));
    return hipGetLastError();
  }
  CDM_DISPATCH_L(L, hipLaunchKernelGGL(cdm_gridsearch_kernel<LG>, op_grid(g.nvox, LG), dim3(cdm_block), op_lds(g),
                                       s, a));
  return hipGetLastError();
}

// Resident blocks per CU of the work-queue LM kernel <L, W> at `lds` bytes,
// queried once per shape (host-side occupancy calculation).
template <int L, int W>
static int op_levmar_queue_residency(size_t lds) {
  static int cached = 0;
  static size_t cached_lds = 0;
  if (cached <= 0 || cached_lds != lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (cdm_levmar_queue_kernel<L, W>), cdm_block, lds) !=
            hipSuccess || n <= 0)
      n = W;
    cached = n;
    cached_lds = lds;
  }
  return cached;
}

template <int L, int W>
static void op_levmar_queue_launch(const cdm_lm_args& a, int marquardt, int nvox, size_t lds, int ncu, hipStream_t s) {
  const unsigned need = op_grid(nvox, L).x;
  const unsigned cap = (unsigned)op_levmar_queue_residency<L, W>(lds) * (unsigned)(ncu > 0 ? ncu : 1);
  hipLaunchKernelGGL((cdm_levmar_queue_kernel<L, W>), dim3(need < cap ? need : cap), dim3(cdm_block), lds, s, a,
                     marquardt);
}

// queue: a device int of the model handle; non-null selects the work-queue
// kernel (grid: the resident blocks, at most one per voxel group), null the
// static mapping.  ncu: compute units of the device.
static hipError_t op_levmar(int L, int W, const cdm::Geom& g, float* params, int max_iters, int marquardt,
                            const cdm::Consts& c, int* iters, float* cost, int* queue, int ncu, hipStream_t s) {
  if (g.nvox <= 0) return hipSuccess;
  if (g.cfp_ts > cdm_cfp_max || g.fixp_ts > cdm_fixp_max) return hipErrorInvalidValue;
  L = op_lm_lanes(g, L);
  if (op_lds_lm(g, L) > cdm_lds_max) return hipErrorInvalidValue;
  cdm_lm_args a = op_lm_args(g, params, max_iters, c, iters, cost);
  if (!queue) {
    CDM_DISPATCH_L(L, CDM_DISPATCH_W(W, hipLaunchKernelGGL((cdm_levmar_kernel<LG, WG>), op_grid(g.nvox, LG),
                                                           dim3(cdm_block), op_lds_lm(g, LG), s, a, marquardt)));
    return hipGetLastError();
  }
  a.queue = queue;
  hipLaunchKernelGGL(cdm_queue_reset_kernel, dim3(1), dim3(1), 0, s, queue);
  const size_t lds = op_lds_lm(g, L);
  CDM_DISPATCH_L(L, CDM_DISPATCH_W(W, (op_levmar_queue_launch<LG, WG>(a, marquardt, g.nvox, lds, ncu, s))));
  return hipGetLastError();
}

static hipError_t op_mcmc(int L, int W, int rician, const cdm::Geom& g, const cdm::McmcLaunch& m,
                          const cdm::Consts& c, hipStream_t s) {
  if (g.nvox <= 0) return hipSuccess;
  if (g.cfp_ts > cdm_cfp_max || g.fixp_ts > cdm_fixp_max) return hipErrorInvalidValue;
  // the Rician likelihood is compiled at the default register budget only
  const int We = rician ? 1 : W;
  const cdm_mcmc_args a = op_mcmc_args(g, L, We, m, c);
  if (rician) {
    CDM_DISPATCH_L(L, hipLaunchKernelGGL((cdm_mcmc_kernel<LG, true, 1>), op_grid(g.nvox, LG), dim3(cdm_block),
                                         op_lds_mcmc(g, LG, We), s, a));
  } else {
    CDM_DISPATCH_L(L, CDM_DISPATCH_W(W, hipLaunchKernelGGL((cdm_mcmc_kernel<LG, false, WG>), op_grid(g.nvox, LG),
                                                           dim3(cdm_block), op_lds_mcmc(g, LG, We), s, a)));
  }
  return hipGetLastError();
}

// W: the launch's effective register budget (1 for the Rician likelihood).
static int op_mcmc_rows_lds(int L, int W, const cdm::Geom& g) {
  return op_mcmc_rows_lds_bytes(g, L, W) > 0 ? 1 : 0;
}

static hipError_t op_predicted(int L, const cdm::Geom& g, int nsamples, const float* samples, float* out,
                               hipStream_t s) {
  if (g.nvox <= 0) return hipSuccess;
  if (g.cfp_ts > cdm_cfp_max || g.fixp_ts > cdm_fixp_max) return hipErrorInvalidValue;
  cdm_post_args a;
  a.g = op_geom(g);
  a.nsamples = nsamples;
  a.samples = samples;
  a.tau_samples = nullptr;
  a.out = out;
  a.bic = a.aic = nullptr;
  CDM_DISPATCH_L(L, hipLaunchKernelGGL(cdm_predicted_kernel<LG>, op_grid(g.nvox, LG), dim3(cdm_block), op_lds(g), s,
                                       a));
  return hipGetLastError();
}

static hipError_t op_bic_aic(int L, int rician, const cdm::Geom& g, int nsamples, const float* samples,
                             const float* tau_samples, float* bic, float* aic, hipStream_t s) {
  if (g.nvox <= 0) return hipSuccess;
  if (g.cfp_ts > cdm_cfp_max || g.fixp_ts > cdm_fixp_max) return hipErrorInvalidValue;
  cdm_post_args a;
  a.g = op_geom(g);
  a.nsamples = nsamples;
  a.samples = samples;
  a.tau_samples = tau_samples;
  a.out = nullptr;
  a.bic = bic;
  a.aic = aic;
  if (rician) {
    CDM_DISPATCH_L(L, hipLaunchKernelGGL((cdm_bic_aic_kernel<LG, true>), op_grid(g.nvox, LG), dim3(cdm_block),
                                         op_lds(g), s, a));
  } else {
    CDM_DISPATCH_L(L, hipLaunchKernelGGL((cdm_bic_aic_kernel<LG, false>), op_grid(g.nvox, LG), dim3(cdm_block),
                                         op_lds(g), s, a));
  }
  return hipGetLastError();
}

#undef CDM_DISPATCH_L
#undef CDM_DISPATCH_W
