// Kernel-argument builders of one model (included in the model namespace
// after engine.h): host-side geometry, bounds/priors and launch options into
// the kernels' argument structs, the dynamic LDS size and the grid.  Used by
// the per-model launchers (launch.inl) and by the batched launch.

static cdm_consts op_consts(const cdm::Consts& in) {
  cdm_consts c;
  for (int i = 0; i < cdm_np; ++i) {
    c.bound_types[i] = in.bound_types[i];
    c.bmin[i] = in.bmin[i];
    c.bmax[i] = in.bmax[i];
    c.prior_types[i] = in.prior_types[i];
    c.pa[i] = in.pa[i];
    c.pb[i] = in.pb[i];
    c.fixed[i] = in.fixed[i];
  }
  c.dbg_voxel = in.dbg_voxel;
  c.dbg_cap = in.dbg_cap;
  c.dbg = in.dbg;
  return c;
}

static cdm_geom op_geom(const cdm::Geom& g) {
  cdm_geom k;
  k.nvox = g.nvox;
  k.nmeas = g.nmeas;
  k.cfp_ts = g.cfp_ts;
  k.fixp_ts = g.fixp_ts;
  k.meas = g.meas;
  k.cfp = g.cfp;
  k.fixp = g.fixp;
  return k;
}

// LDS of a CU (gfx950), the most one block can have.
constexpr size_t cdm_lds_max = 160 * 1024;

// MCMC keeps the per-measurement cache rows in LDS when the block's LDS then
// leaves room for the W blocks per CU of the launch's register budget (W = 1:
// the whole LDS of the CU; 2: 80 KB), else in global memory.
static size_t op_mcmc_rows_lds_bytes(const cdm::Geom& g, int L, int W) {
  if (cdm_mvc <= 0) return 0;
  const size_t rows = (size_t)(cdm_block / L) * 2 * (size_t)cdm_mcmc_row_stride(g.nmeas) * sizeof(cdm_tf);
  const size_t total = cdm_mcmc_rows_offset(g.nmeas, g.cfp_ts, L) + rows;
  return total <= cdm_lds_max / (size_t)(W > 1 ? W : 1) ? rows : 0;
}

static size_t op_lds_mcmc(const cdm::Geom& g, int L, int W) {
  const size_t rows = op_mcmc_rows_lds_bytes(g, L, W);
  if (rows) return cdm_mcmc_rows_offset(g.nmeas, g.cfp_ts, L) + rows;
  return cdm_vslot_offset(g.nmeas, g.cfp_ts) + (size_t)(cdm_block / L) * sizeof(cdm_vslot);
}

static size_t op_lds(const cdm::Geom& g) {
  return (size_t)cdm_mc_offset(g.nmeas, g.cfp_ts) * sizeof(float) + (size_t)g.nmeas * cdm_mc * sizeof(CDM_F64);
}

// LM launches: the 64-lane groups' J^T J scratch after the shared tables.
static size_t op_lds_lm(const cdm::Geom& g, int L) {
  const size_t extra = cdm_lm_scratch_doubles(L, g.nmeas);
  if (!extra) return op_lds(g);
  return cdm_lm_scratch_offset(g.nmeas, g.cfp_ts) + (size_t)(cdm_block / L) * extra * sizeof(CDM_F64);
}

// LM voxel-group size actually launched: 64-lane groups keep the rows of J
// in LDS, and past the LDS of a CU they fall back to 16 lanes (per-lane
// accumulators).
static int op_lm_lanes(const cdm::Geom& g, int L) {
  return (L == 64 && op_lds_lm(g, 64) > cdm_lds_max) ? 16 : L;
}

static dim3 op_grid(int nvox, int L) {
  const int per_block = cdm_block / L;
  return dim3((unsigned)((nvox + per_block - 1) / per_block));
}

static cdm_lm_args op_lm_args(const cdm::Geom& g, float* params, int max_iters, const cdm::Consts& c, int* iters,
                              float* cost) {
  cdm_lm_args a;
  a.g = op_geom(g);
  a.params = params;
  a.max_iters = max_iters;
  a.c = op_consts(c);
  a.iters_out = iters;
  a.cost_out = cost;
  a.queue = nullptr;
  return a;
}

// L, W: the launch's voxel-group size and register budget (they decide where
// the cache rows live and their LDS layout).
static cdm_mcmc_args op_mcmc_args(const cdm::Geom& g, int L, int W, const cdm::McmcLaunch& m,
                                  const cdm::Consts& c) {
  cdm_mcmc_args a;
  a.g = op_geom(g);
  a.params = m.params;
  a.mvc = (cdm_tf*)m.mvc;
  a.rows_lds = op_mcmc_rows_lds_bytes(g, L, W) > 0 ? 1 : 0;
  a.row_stride = cdm_mcmc_row_stride(g.nmeas);
  a.state = (cdm_chain*)m.state;
  a.samples = m.samples;
  a.tau_samples = m.tau_samples;
  a.nsamples = m.nsamples;
  a.sampleevery = m.sampleevery;
  a.upe = m.upe;
  a.recording = m.recording;
  a.update_prop = m.update_prop;
  a.begin = m.begin;
  a.count = m.count;
  a.phase_len = m.phase_len;
  a.jump_offset = m.jump_offset;
  a.seed = m.seed;
  a.voxel_offset = m.voxel_offset;
  a.c = op_consts(c);
  return a;
}
