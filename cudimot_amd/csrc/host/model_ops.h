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

struct ModelOps {
  const char* name;
  int nparams;
  int ncfp;
  const int* cfp_sizes;
  int nfixp;
  const int* fixp_sizes;
  int forward_is_double;
  size_t chain_bytes;
  int cfp_max, fixp_max;
  hipError_t (*gridsearch)(int L, const Geom& g, float* params, int ngp, const int* gparams, int ncombs,
                           const float* grid_f, const double* grid_d, const int* comb_id, const Consts& c,
                           int* kept, hipStream_t s);
  hipError_t (*levmar)(int L, const Geom& g, float* params, int max_iters, int marquardt, const Consts& c,
                       int* iters, float* cost, hipStream_t s);
  hipError_t (*mcmc)(int L, int rician, const Geom& g, const McmcLaunch& m, const Consts& c, hipStream_t s);
  hipError_t (*predicted)(int L, const Geom& g, int nsamples, const float* samples, float* out, hipStream_t s);
  hipError_t (*bic_aic)(int L, int rician, const Geom& g, int nsamples, const float* samples,
                        const float* tau_samples, float* bic, float* aic, hipStream_t s);
};

// Voxel-group sizes compiled for every model and method.
constexpr int kGroupSizes[] = {1, 4, 16, 64};

const ModelOps* find_model(const char* name);
int model_count();
const ModelOps* model_at(int i);

// Error reporting shared by every C-ABI source file: stores the message for
// cdm_last_error() on this thread and returns the status (-1 argument error,
// -2 HIP error).
int set_error(const std::string& msg, int status = -1);

}  // namespace cdm
