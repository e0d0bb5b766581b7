// Model-independent view of one compiled plug-in: sizes and kernel launchers.
// Each model translation unit (generated from model_unit.hip.in) fills one
// ModelOps table; the C-ABI layer (capi.cpp) dispatches through it.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <string>

namespace cdm {

constexpr int kMaxParams = 32;

struct Consts {
  int bound_types[kMaxParams];
  float bmin[kMaxParams], bmax[kMaxParams];
  int prior_types[kMaxParams];
  float pa[kMaxParams], pb[kMaxParams];
  int fixed[kMaxParams];
  // --debug=<voxel>: the methods' state records of that voxel (dbg null: off)
  int dbg_voxel, dbg_cap;
  double* dbg;
};

struct Geom {
  int nvox, nmeas, cfp_ts, fixp_ts;
  const float* meas;
  const float* cfp;
  const float* fixp;
};

struct McmcLaunch {
  float* params;
  void* state;
  void* mvc;  // per-measurement cache rows (mvc_bytes_per_meas x 2 x nmeas per voxel)
  float* samples;
  float* tau_samples;
  int nsamples, sampleevery, upe;
  int recording, update_prop;
  int begin, count, phase_len;
  unsigned jump_offset;
  unsigned long long seed;
  long long voxel_offset;
};

enum Method { kGridSearch = 0, kLevMar = 1, kMcmc = 2, kPredicted = 3, kBicAic = 4 };

// Scratch of a GridSearch split over S slices (per voxel and slice: an FP64
// rescan minimum, 4 screened costs and indices, the rescan index; per voxel
// a crowded flag).
inline size_t gs_split_bytes(int nvox, int S) {
  const size_t n = (size_t)nvox * S;
  return n * (sizeof(double) + 4 * (sizeof(float) + sizeof(int)) + sizeof(int)) + (size_t)nvox * sizeof(int);
}

struct ModelOps {
  const char* name;
  int nparams;
  int ncfp;
  const int* cfp_sizes;
  int nfixp;
  const int* fixp_sizes;
  int forward_is_double;
  size_t chain_bytes;
  size_t mvc_bytes_per_meas;  // per-measurement cache of one state (0: none)
  int cfp_max, fixp_max;
  // S > 1: the search in S slices of the combinations, with gs_split_bytes
  // of scratch at `split`
  hipError_t (*gridsearch)(int L, const Geom& g, float* params, int ngp, const int* gparams, int ncombs,
                           const float* grid_f, const double* grid_d, const int* comb_id, const Consts& c,
                           int* kept, int S, void* split, hipStream_t s);
  // W: register budget, in resident 256-thread blocks per CU (1, 2 or 4)
  // queue: device int for the work-queue kernel (null: static voxel mapping);
  // ncu: compute units of the device
  hipError_t (*levmar)(int L, int W, const Geom& g, float* params, int max_iters, int marquardt, const Consts& c,
                       int* iters, float* cost, int* queue, int ncu, hipStream_t s);
  hipError_t (*mcmc)(int L, int W, int rician, const Geom& g, const McmcLaunch& m, const Consts& c, hipStream_t s);
  hipError_t (*predicted)(int L, const Geom& g, int nsamples, const float* samples, float* out, hipStream_t s);
  hipError_t (*bic_aic)(int L, int rician, const Geom& g, int nsamples, const float* samples,
                        const float* tau_samples, float* bic, float* aic, hipStream_t s);
  // 1 when an MCMC launch of group size L keeps the per-measurement cache
  // rows in LDS (the state buffer then needs no rows after the chains)
  int (*mcmc_rows_lds)(int L, int W, const Geom& g);
};

// Voxel-group sizes compiled for every model and method.
constexpr int kGroupSizes[] = {1, 4, 16, 64};

const ModelOps* find_model(const char* name);
int model_count();
const ModelOps* model_at(int i);

// ---------------------------------------------------------------- batches
// A batch unit compiles several plug-ins into one persistent kernel per
// method: workgroups take voxel tiles (one block's worth of one item) from a
// work queue until every item's tiles are done.
enum BatchMode { kBatchLevMar = 0, kBatchMcmc = 1, kBatchMcmcRician = 2 };

struct BatchLaunchItem {
  int model;  // index in the unit's model list
  int L;      // voxel-group size
  Geom g;
  Consts c;
  // Levenberg-Marquardt
  float* params;
  int max_iters;
  int* iters;
  float* cost;
  // MCMC
  McmcLaunch mc;
};

struct BatchOps {
  const char* name;
  int nmodels;
  const char* const* models;
  size_t item_bytes;  // device bytes of one prepared item
  // Host-side preparation of nitems device items (into host_out, item_bytes
  // each, in the given order); returns the total number of tiles and the
  // dynamic LDS bytes a workgroup needs.  Non-zero on an invalid item.
  int (*prepare)(int mode, int nitems, const BatchLaunchItem* items, void* host_out, int* total_tiles,
                 size_t* lds_bytes);
  // Workgroups of the persistent launch: resident blocks per CU x CUs,
  // capped by the tiles (queried before any stream capture).
  hipError_t (*grid)(int mode, int total_tiles, size_t lds_bytes, int* grid);
  // One persistent launch over prepared items (dev_items) with a zeroed
  // work-queue counter.
  hipError_t (*launch)(int mode, const void* dev_items, int nitems, int total_tiles, int* queue, size_t lds_bytes,
                       int marquardt, int grid, hipStream_t s);
};

int batch_count();
const BatchOps* batch_at(int i);

// Error reporting shared by every C-ABI source file: stores the message for
// cdm_last_error() on this thread and returns the status (-1 argument error,
// -2 HIP error).
int set_error(const std::string& msg, int status = -1);

}  // namespace cdm
