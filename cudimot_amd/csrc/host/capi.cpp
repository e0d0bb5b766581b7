// C-ABI of the fitting engine (declarations and reference citations in
// include/cudimot_amd.h).  Host-side state per model handle: the parsed
// specification, device copies of the search grid, the MCMC chain-state
// buffer and the launch choices (voxel-group size per method, MCMC blocks).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/cudimot_amd.h"
#include "fill.h"
#include "model_ops.h"
#include "noalloc.h"
#include "spec.h"

namespace {

// A captured launch sequence, kept instantiated and replayed while its key
// (every value the launches bake into the graph's nodes) is unchanged.
struct GraphCache {
  std::vector<unsigned char> key;
  hipGraphExec_t exec = nullptr;
  void clear() {
    if (exec) (void)hipGraphExecDestroy(exec);
    exec = nullptr;
    key.clear();
  }
};

// Counters of the graph caches (cdm_graph_stats).
struct GraphStats {
  long long captures = 0, replays = 0;
};

}  // namespace

struct cdm_model {
  const cdm::ModelOps* ops = nullptr;
  cdm::ModelSpec spec;
  int group[5] = {0, 0, 0, 0, 0};
  int occ[5] = {0, 0, 0, 0, 0};  // register budget (resident blocks per CU); 0: automatic
  // autotuned choices for this model: (measurement count, method, lanes, budget)
  struct Tuned {
    int nmeas, method, lanes;
    double ms;  // measured time of the tuning run (0 if not given)
    int occ;    // register budget (1 if not given)
  };
  std::vector<Tuned> tuned;
  int mcmc_block = 0;
  int mcmc_graph = 0;
  // device grid (valid combinations only)
  bool grid_dirty = true;
  int ncombs_valid = 0;
  float* d_grid_f = nullptr;
  double* d_grid_d = nullptr;
  int* d_comb_id = nullptr;
  // MCMC chain state
  void* d_state = nullptr;
  size_t state_bytes = 0;
  // LM work-queue counter (one int; calls on one handle are serialised)
  int* d_queue = nullptr;
  // scratch of a GridSearch split over slices of the combinations
  // (cdm::gs_split_bytes), grown as needed
  void* d_gs_split = nullptr;
  size_t gs_split_bytes = 0;
  // MCMC phases recorded into HIP graphs (jump blocks with use_graph), one
  // per phase, and the private stream they are captured on
  GraphCache mcmc_graphs[2];
  hipStream_t capture_stream = nullptr;
  GraphStats graph_stats;
  // the stream of the handle's last launch and an event after it: a call on
  // another stream waits for it, since the queue counter and the chain
  // state belong to the handle
  hipStream_t last_stream = nullptr;
  hipEvent_t last_done = nullptr;
  bool has_last = false;
  // --debug=<voxel>: device records of that voxel, three regions (GridSearch,
  // LM, MCMC) of debug_cap records each (d_debug null: off)
  int debug_voxel = -1;
  int debug_cap = 0;
  double* d_debug = nullptr;
};

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return -1;
}

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return -2;
}

std::string lib_dir() {
  Dl_info info;
  if (dladdr((void*)&cdm_version, &info) && info.dli_fname) {
    std::string p(info.dli_fname);
    const size_t s = p.rfind('/');
    return s == std::string::npos ? std::string(".") : p.substr(0, s);
  }
  return ".";
}

cdm::Consts make_consts(const cdm_model* m) {
  cdm::Consts c;
  std::memset(&c, 0, sizeof(c));
  const cdm::ModelSpec& s = m->spec;
  for (int i = 0; i < s.nparams && i < cdm::kMaxParams; ++i) {
    c.bound_types[i] = s.bound_types[i];
    c.bmin[i] = (float)s.bmin[i];
    c.bmax[i] = (float)s.bmax[i];
    c.prior_types[i] = s.prior_types[i];
    c.pa[i] = (float)s.pa[i];
    c.pb[i] = (float)s.pb[i];
    c.fixed[i] = s.fixed[i];
  }
  c.dbg_voxel = m->debug_voxel;
  c.dbg_cap = m->debug_cap;
  c.dbg = m->d_debug;
  return c;
}

// Doubles of one --debug region: the record count, then debug_cap records of
// (index, three values, the parameters).
size_t debug_region_doubles(const cdm_model* m) { return 8 + (size_t)m->debug_cap * (4 + m->ops->nparams); }

// Clears a method's --debug records before its launches (a fill kernel: the
// count is the region's first double, zero bits).
hipError_t debug_reset(const cdm_model* m, int region, hipStream_t s) {
  if (!m->d_debug) return hipSuccess;
  return cdm::zero_ints(reinterpret_cast<int*>(m->d_debug + (size_t)region * debug_region_doubles(m)), 2, s);
}

int tsize(int n, const int* sizes) {
  int t = 0;
  for (int i = 0; i < n; ++i) t += sizes[i];
  return t;
}

bool valid_group(int L) { return L == 1 || L == 4 || L == 16 || L == 64; }

// Built-in voxel-group choice when no tuned entry exists: more lanes per
// voxel when a voxel has many measurements, so each lane keeps a handful of
// them; LM keeps its normal matrix per voxel, so it favours wide groups.
int default_group(int method, int nmeas) {
  if (method == cdm::kLevMar) return nmeas >= 48 ? 64 : 16;
  if (nmeas >= 192) return 16;
  return 4;
}

// Explicit choice, else the tuned entry of the nearest measurement count
// (log ratio), else the built-in rule.  Tuned and built-in choices get an
// occupancy floor: a launch over few voxels widens the group (x4 steps, up to
// 64 lanes) until the grid covers about two waves per SIMD of the GPU, since
// a narrow group over a small volume leaves most CUs idle and each voxel's
// work serial.
constexpr long long kLaneSlots = 256LL * 4 * 64 * 2;

// Tuned entry of the nearest measurement count (log ratio), or null.
const cdm_model::Tuned* tuned_for(const cdm_model* m, int method, int nmeas) {
  const cdm_model::Tuned* best = nullptr;
  double dist = 0;
  for (const auto& t : m->tuned) {
    if (t.method != method) continue;
    const double d = std::fabs(std::log((double)t.nmeas / (double)std::max(nmeas, 1)));
    if (!best || d < dist) {
      best = &t;
      dist = d;
    }
  }
  return best;
}

int group_for(const cdm_model* m, int method, int nmeas, int nvox) {
  if (m->group[method]) return m->group[method];
  const cdm_model::Tuned* t = tuned_for(m, method, nmeas);
  int L = t ? t->lanes : default_group(method, nmeas);
  while (nvox > 0 && L < 64 && (long long)nvox * L < kLaneSlots) L *= 4;
  return L;
}

// Register budget of the LM and MCMC kernels: explicit, else tuned, else 1
// (the compiler's own allocation).
int occ_for(const cdm_model* m, int method, int nmeas) {
  if (m->occ[method]) return m->occ[method];
  const cdm_model::Tuned* t = tuned_for(m, method, nmeas);
  return t ? t->occ : 1;
}

bool valid_occ(int w) { return w == 1 || w == 2 || w == 4; }

const char* const kMethodNames[5] = {"gridsearch", "levmar", "mcmc", "predicted", "bic_aic"};

// Tuning table lines: "<model> <nmeas> <method> <lanes> [ms]"; '#' starts a
// comment.  Returns the number of entries kept for this model, -1 if the
// file cannot be read.
int read_tuning(cdm_model* m, const std::string& path, std::string* err) {
  std::ifstream in(path);
  if (!in) {
    *err = "cannot read tuning table " + path;
    return -1;
  }
  std::vector<cdm_model::Tuned> got;
  std::string line;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream ss(line);
    std::string model, method;
    int nmeas = 0, lanes = 0;
    double ms = 0;
    if (!(ss >> model)) continue;
    if (!(ss >> nmeas >> method >> lanes) || nmeas <= 0 || !valid_group(lanes)) {
      *err = path + ":" + std::to_string(lineno) + ": expected '<model> <nmeas> <method> <lanes>' with lanes 1/4/16/64";
      return -1;
    }
    if (model != m->ops->name) continue;
    int k = -1;
    for (int i = 0; i < 5; ++i)
      if (method == kMethodNames[i]) k = i;
    if (k < 0) {
      *err = path + ":" + std::to_string(lineno) + ": unknown method '" + method + "'";
      return -1;
    }
    int occ = 1;
    if (!(ss >> ms)) ms = 0;
    else if (!(ss >> occ)) occ = 1;
    if (!valid_occ(occ)) {
      *err = path + ":" + std::to_string(lineno) + ": register budget must be 1, 2 or 4";
      return -1;
    }
    got.push_back({nmeas, k, lanes, ms, occ});
  }
  m->tuned = got;
  return (int)got.size();
}

cdm::Geom make_geom(const cdm_model* m, int nvox, int nmeas, const float* meas, const float* cfp, const float* fixp) {
  cdm::Geom g;
  g.nvox = nvox;
  g.nmeas = nmeas;
  g.cfp_ts = tsize(m->ops->ncfp, m->ops->cfp_sizes);
  g.fixp_ts = tsize(m->ops->nfixp, m->ops->fixp_sizes);
  g.meas = meas;
  g.cfp = cfp;
  g.fixp = fixp;
  return g;
}

int check_geom(const cdm_model* m, int nvox, int nmeas, const void* meas, const void* cfp, const void* fixp,
               bool need_meas) {
  if (!m || !m->ops) return fail("null model handle");
  if (nvox < 0 || nmeas <= 0) return fail("the number of voxels must be >= 0 and of measurements > 0");
  if (need_meas && nvox > 0 && !meas) return fail("measurements pointer is null");
  const int cts = tsize(m->ops->ncfp, m->ops->cfp_sizes), fts = tsize(m->ops->nfixp, m->ops->fixp_sizes);
  if (cts > 0 && !cfp) return fail("this model needs common fixed parameters (--CFP) but cfp is null");
  if (fts > 0 && nvox > 0 && !fixp) return fail("this model needs per-voxel fixed parameters (--FixP) but fixp is null");
  if (cts > m->ops->cfp_max) return fail("CFP rows per measurement exceed the engine limit");
  if (fts > m->ops->fixp_max) return fail("FixP values per voxel exceed the engine limit");
  return 0;
}

void free_grid(cdm_model* m) {
  if (m->d_grid_f) (void)hipFree(m->d_grid_f);
  if (m->d_grid_d) (void)hipFree(m->d_grid_d);
  if (m->d_comb_id) (void)hipFree(m->d_comb_id);
  m->d_grid_f = nullptr;
  m->d_grid_d = nullptr;
  m->d_comb_id = nullptr;
}

// Combinations whose grid values pass the float bounds test of
// GridSearch.cu:25-47; the non-grid parameters are tested per voxel on the
// device.
int upload_grid(cdm_model* m) {
  if (!m->grid_dirty) return 0;
  free_grid(m);
  const cdm::ModelSpec& s = m->spec;
  const int ngp = (int)s.grid_params.size();
  std::vector<float> gf;
  std::vector<double> gd;
  std::vector<int> id;
  for (int c = 0; c < s.grid_combs; ++c) {
    bool ok = true;
    for (int k = 0; k < ngp; ++k) {
      const int p = s.grid_params[k];
      const double v = s.grid[(size_t)c * ngp + k];
      const float lo = (float)s.bmin[p], hi = (float)s.bmax[p];
      const int bt = s.bound_types[p];
      if ((bt == cdm::BMIN || bt == cdm::BMINMAX) && v < lo) ok = false;
      if ((bt == cdm::BMAX || bt == cdm::BMINMAX) && v > hi) ok = false;
    }
    if (!ok) continue;
    for (int k = 0; k < ngp; ++k) {
      gd.push_back(s.grid[(size_t)c * ngp + k]);
      gf.push_back((float)s.grid[(size_t)c * ngp + k]);
    }
    id.push_back(c);
  }
  m->ncombs_valid = (int)id.size();
  if (m->ncombs_valid > 0) {
    hipError_t e;
    if ((e = hipMalloc((void**)&m->d_grid_f, gf.size() * sizeof(float))) != hipSuccess) return hip_fail(e, "grid");
    if ((e = hipMalloc((void**)&m->d_grid_d, gd.size() * sizeof(double))) != hipSuccess) return hip_fail(e, "grid");
    if ((e = hipMalloc((void**)&m->d_comb_id, id.size() * sizeof(int))) != hipSuccess) return hip_fail(e, "grid");
    if ((e = hipMemcpy(m->d_grid_f, gf.data(), gf.size() * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(e, "grid upload");
    if ((e = hipMemcpy(m->d_grid_d, gd.data(), gd.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(e, "grid upload");
    if ((e = hipMemcpy(m->d_comb_id, id.data(), id.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
      return hip_fail(e, "grid upload");
  }
  m->grid_dirty = false;
  return 0;
}

// MCMC chain state of nvox voxels, followed (256-byte aligned) by the two
// per-measurement cache rows of each voxel when the plug-in declares them.
size_t mvc_offset(const cdm_model* m, int nvox) {
  return (m->ops->chain_bytes * (size_t)nvox + 255) / 256 * 256;
}

// The handle's chain-state buffer, grown to at least `bytes`.
int ensure_state(cdm_model* m, size_t bytes) {
  if (m->state_bytes >= bytes) return 0;
  if (m->d_state) (void)hipFree(m->d_state);
  m->d_state = nullptr;
  m->state_bytes = 0;
  hipError_t e = hipMalloc(&m->d_state, bytes);
  if (e != hipSuccess) return hip_fail(e, "MCMC chain state");
  m->state_bytes = bytes;
  return 0;
}

void* mvc_rows(cdm_model* m, int nvox) {
  return m->ops->mvc_bytes_per_meas ? (void*)((char*)m->d_state + mvc_offset(m, nvox)) : nullptr;
}

// Replays the graph of `gc` on `s` when its key equals `key`; otherwise
// records the launches that `issue` makes on the private capture stream `*cs`
// (created on first use: a capture needs a stream other than the null
// stream) into a new graph, instantiates it, keeps it with `key` and launches
// it on `s`, which keeps the caller's stream order.
template <class F>
hipError_t run_graph_cached(GraphCache& gc, hipStream_t* cs, GraphStats* st, const std::vector<unsigned char>& key,
                            hipStream_t s, F issue) {
  if (gc.exec && gc.key == key) {
    ++st->replays;
    return hipGraphLaunch(gc.exec, s);
  }
  gc.clear();
  hipError_t e = hipSuccess;
  if (!*cs) e = hipStreamCreateWithFlags(cs, hipStreamNonBlocking);
  if (e != hipSuccess) {
    *cs = nullptr;
    return e;
  }
  e = hipStreamBeginCapture(*cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return e;
  const hipError_t ei = issue(*cs);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(*cs, &g);
  e = ei != hipSuccess ? ei : ec;
  if (e == hipSuccess) e = hipGraphInstantiate(&gc.exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    gc.exec = nullptr;
    return e;
  }
  gc.key = key;
  ++st->captures;
  return hipGraphLaunch(gc.exec, s);
}

// Appends the bytes of a value to a graph key.
template <class V>
void key_add(std::vector<unsigned char>* key, const V& v) {
  const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
  key->insert(key->end(), p, p + sizeof(V));
}

// Orders a call of handle m on stream s after the handle's last launch when
// that was on another stream (the work-queue counter and the chain state
// belong to the handle).  Skipped while s is being captured: a graph cannot
// wait on work outside it, and the capturing caller orders its streams.
hipError_t order_after_last(cdm_model* m, hipStream_t s) {
  if (!m->has_last || m->last_stream == s) return hipSuccess;
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(s, &cst);
  if (e != hipSuccess || cst != hipStreamCaptureStatusNone) return e;
  return hipStreamWaitEvent(s, m->last_done, 0);
}

// Records the end of a call of handle m on stream s (not while s is being
// captured).
hipError_t note_last(cdm_model* m, hipStream_t s) {
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(s, &cst);
  if (e != hipSuccess || cst != hipStreamCaptureStatusNone) return e;
  if (!m->last_done && (e = hipEventCreateWithFlags(&m->last_done, hipEventDisableTiming)) != hipSuccess) {
    m->last_done = nullptr;
    return e;
  }
  if ((e = hipEventRecord(m->last_done, s)) != hipSuccess) return e;
  m->last_stream = s;
  m->has_last = true;
  return hipSuccess;
}

}  // namespace

int cdm::set_error(const std::string& msg, int status) {
  g_err = msg;
  return status;
}

extern "C" {

int cdm_version(void) { return 100; }

const char* cdm_last_error(void) { return g_err.c_str(); }

int cdm_list_models(char* buf, int buflen) {
  std::string s;
  for (int i = 0; i < cdm::model_count(); ++i) {
    if (i) s += ",";
    s += cdm::model_at(i)->name;
  }
  if (buf && buflen > 0) {
    std::strncpy(buf, s.c_str(), (size_t)buflen - 1);
    buf[buflen - 1] = 0;
  }
  return (int)s.size();
}

int cdm_model_open(const char* name, const char* priors_file, cdm_model** out) {
  if (!out) return fail("null output handle");
  *out = nullptr;
  if (!name) return fail("null model name");
  const cdm::ModelOps* ops = cdm::find_model(name);
  if (!ops) return fail(std::string("model not compiled into this library: ") + name);
  cdm_model* m = new cdm_model();
  m->ops = ops;
  const std::string path =
      priors_file ? std::string(priors_file) : lib_dir() + "/../mymodels/" + std::string(name) + "/modelpriors";
  const std::string err = cdm::parse_priors_file(path, ops->nparams, m->spec);
  if (!err.empty()) {
    delete m;
    return fail(err);
  }
  // the library's own tuning table, when one was written next to it
  std::string terr;
  const std::string tpath = lib_dir() + "/../tuning.txt";
  if (std::ifstream(tpath).good() && read_tuning(m, tpath, &terr) < 0) {
    delete m;
    return fail(terr);
  }
  *out = m;
  return 0;
}

void cdm_model_close(cdm_model* m) {
  if (!m) return;
  if (m->has_last && m->last_done) (void)hipEventSynchronize(m->last_done);
  free_grid(m);
  if (m->d_state) (void)hipFree(m->d_state);
  if (m->d_queue) (void)hipFree(m->d_queue);
  if (m->d_gs_split) (void)hipFree(m->d_gs_split);
  for (GraphCache& g : m->mcmc_graphs) g.clear();
  if (m->capture_stream) (void)hipStreamDestroy(m->capture_stream);
  if (m->last_done) (void)hipEventDestroy(m->last_done);
  if (m->d_debug) (void)hipFree(m->d_debug);
  delete m;
}

int cdm_model_set_debug(cdm_model* m, int voxel, int max_records) {
  if (!m) return fail("null model handle");
  if (voxel >= 0 && max_records <= 0) return fail("--debug needs room for at least one record");
  if (m->has_last && m->last_done) (void)hipEventSynchronize(m->last_done);
  if (m->d_debug) (void)hipFree(m->d_debug);
  m->d_debug = nullptr;
  m->debug_voxel = -1;
  m->debug_cap = 0;
  if (voxel < 0) return 0;
  m->debug_cap = max_records;
  const size_t bytes = 3 * debug_region_doubles(m) * sizeof(double);
  hipError_t e = hipMalloc((void**)&m->d_debug, bytes);
  if (e == hipSuccess) e = hipMemset(m->d_debug, 0, bytes);
  if (e != hipSuccess) {
    if (m->d_debug) (void)hipFree(m->d_debug);
    m->d_debug = nullptr;
    m->debug_cap = 0;
    return hip_fail(e, "--debug records");
  }
  m->debug_voxel = voxel;
  return 0;
}

int cdm_model_debug_trace(cdm_model* m, int method, double* out, int max_records, int* nrecords, void* stream) {
  if (!m || !nrecords) return fail("bad arguments");
  *nrecords = 0;
  const int region = method == cdm::kGridSearch ? 0 : method == cdm::kLevMar ? 1 : method == cdm::kMcmc ? 2 : -1;
  if (region < 0) return fail("--debug records exist for GridSearch, Levenberg-Marquardt and MCMC only");
  if (!m->d_debug) return 0;
  const size_t n = debug_region_doubles(m);
  std::vector<double> host(n);
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = hipMemcpyAsync(host.data(), m->d_debug + (size_t)region * n, n * sizeof(double),
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_fail(e, "--debug records");
  const int have = std::min((int)host[0], m->debug_cap);
  const int w = 4 + m->ops->nparams;
  const int k = std::min(have, std::max(max_records, 0));
  if (out && k > 0) std::memcpy(out, host.data() + 8, (size_t)k * w * sizeof(double));
  *nrecords = have;
  return 0;
}

int cdm_model_graph_stats(const cdm_model* m, long long* captures, long long* replays) {
  if (!m) return fail("null model handle");
  if (captures) *captures = m->graph_stats.captures;
  if (replays) *replays = m->graph_stats.replays;
  return 0;
}

int cdm_model_info(const cdm_model* m, int* nparams, int* cfp_tsize, int* fixp_tsize) {
  if (!m) return fail("null model handle");
  if (nparams) *nparams = m->ops->nparams;
  if (cfp_tsize) *cfp_tsize = tsize(m->ops->ncfp, m->ops->cfp_sizes);
  if (fixp_tsize) *fixp_tsize = tsize(m->ops->nfixp, m->ops->fixp_sizes);
  return 0;
}

int cdm_model_sizes(const cdm_model* m, int* ncfp, int* cfp_sizes, int* nfixp, int* fixp_sizes) {
  if (!m) return fail("null model handle");
  if (ncfp) *ncfp = m->ops->ncfp;
  if (nfixp) *nfixp = m->ops->nfixp;
  for (int i = 0; cfp_sizes && i < m->ops->ncfp; ++i) cfp_sizes[i] = m->ops->cfp_sizes[i];
  for (int i = 0; fixp_sizes && i < m->ops->nfixp; ++i) fixp_sizes[i] = m->ops->fixp_sizes[i];
  return 0;
}

int cdm_model_spec(const cdm_model* m, int* bound_types, double* bmin, double* bmax, int* prior_types, double* pa,
                   double* pb, int* fixed, double* p_init, int* init_provided) {
  if (!m) return fail("null model handle");
  const cdm::ModelSpec& s = m->spec;
  for (int i = 0; i < s.nparams; ++i) {
    if (bound_types) bound_types[i] = s.bound_types[i];
    if (bmin) bmin[i] = s.bmin[i];
    if (bmax) bmax[i] = s.bmax[i];
    if (prior_types) prior_types[i] = s.prior_types[i];
    if (pa) pa[i] = s.pa[i];
    if (pb) pb[i] = s.pb[i];
    if (fixed) fixed[i] = s.fixed[i];
    if (p_init) p_init[i] = s.init_provided ? s.p_init[i] : 0.0;
  }
  if (init_provided) *init_provided = s.init_provided ? 1 : 0;
  return 0;
}

int cdm_model_set_fixed(cdm_model* m, const char* list) {
  if (!m) return fail("null model handle");
  const std::string err = cdm::parse_fixed(list ? list : "", m->spec);
  return err.empty() ? 0 : fail(err);
}

int cdm_model_load_grid(cdm_model* m, const char* src, int is_text, int* ncombs, int* ngrid_params) {
  if (!m) return fail("null model handle");
  if (!src) return fail("null grid source");
  const std::string err = is_text ? cdm::parse_grid_text(src, m->spec) : cdm::parse_grid_file(src, m->spec);
  if (!err.empty()) return fail(err);
  m->grid_dirty = true;
  if (ncombs) *ncombs = m->spec.grid_combs;
  if (ngrid_params) *ngrid_params = (int)m->spec.grid_params.size();
  return 0;
}

int cdm_model_grid(const cdm_model* m, int* grid_params, double* grid) {
  if (!m) return fail("null model handle");
  if (grid_params) std::copy(m->spec.grid_params.begin(), m->spec.grid_params.end(), grid_params);
  if (grid) std::copy(m->spec.grid.begin(), m->spec.grid.end(), grid);
  return 0;
}

int cdm_model_set_group(cdm_model* m, int method, int lanes) {
  if (!m) return fail("null model handle");
  if (method < 0 || method > 4) return fail("unknown method");
  if (lanes != 0 && !valid_group(lanes)) return fail("voxel-group size must be 1, 4, 16 or 64 lanes");
  m->group[method] = lanes;
  return 0;
}

int cdm_model_set_occupancy(cdm_model* m, int method, int blocks) {
  if (!m) return fail("null model handle");
  if (method != cdm::kLevMar && method != cdm::kMcmc) return fail("register budgets apply to LM and MCMC");
  if (blocks != 0 && !valid_occ(blocks)) return fail("register budget must be 1, 2 or 4 blocks per CU (0: automatic)");
  m->occ[method] = blocks;
  return 0;
}

int cdm_model_get_occupancy(const cdm_model* m, int method, int nmeas) {
  if (!m || method < 0 || method > 4) return fail("bad arguments");
  return occ_for(m, method, nmeas);
}

int cdm_model_load_tuning(cdm_model* m, const char* path) {
  if (!m) return fail("null model handle");
  if (!path) {
    m->tuned.clear();
    return 0;
  }
  std::string err;
  const int n = read_tuning(m, path, &err);
  return n < 0 ? fail(err) : n;
}

int cdm_model_get_group(const cdm_model* m, int method, int nmeas, int nvox) {
  if (!m || method < 0 || method > 4) return fail("bad arguments");
  return group_for(m, method, nmeas, nvox);
}

int cdm_model_set_mcmc_blocks(cdm_model* m, int jumps_per_launch, int use_graph) {
  if (!m) return fail("null model handle");
  if (jumps_per_launch < 0) return fail("jumps per launch must be >= 0");
  m->mcmc_block = jumps_per_launch;
  m->mcmc_graph = use_graph ? 1 : 0;
  return 0;
}

// Slices of the combinations for a GridSearch whose voxel groups alone
// leave the GPU mostly idle: enough (at most 16, at least 64 combinations
// each) for about 8 blocks per CU; 1 for volumes that fill it.
int gs_splits(int L, int nvox, int ncombs) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) {
    (void)hipGetLastError();
    return 1;
  }
  const long long per_block = 256 / L;
  const long long blocks = ((long long)nvox + per_block - 1) / per_block;
  const long long target = 8LL * ncu;
  if (blocks >= target) return 1;
  long long S = std::min<long long>(16, (target + blocks - 1) / blocks);
  S = std::min<long long>(S, std::max(1, ncombs / 64));
  return (int)std::max<long long>(1, S);
}

int cdm_gridsearch(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
                   const float* fixp, int* kept, void* stream) {
  int r = check_geom(m, nvox, nmeas, meas, cfp, fixp, true);
  if (r) return r;
  if (m->spec.grid_combs <= 0) return fail("GridSearch needs a grid (--gridSearch file with search[i]=(...) lines)");
  if ((r = upload_grid(m))) return r;
  if (nvox == 0) return 0;
  const cdm::Geom g = make_geom(m, nvox, nmeas, meas, cfp, fixp);
  const cdm::Consts c = make_consts(m);
  hipStream_t s = (hipStream_t)stream;
  const int L = group_for(m, cdm::kGridSearch, nmeas, nvox);
  const int S = gs_splits(L, nvox, m->ncombs_valid);
  hipError_t e = hipSuccess;
  if (S > 1) {
    // the lists belong to the handle: order after its last launch
    const size_t bytes = cdm::gs_split_bytes(nvox, S);
    if (m->gs_split_bytes < bytes) {
      if ((e = order_after_last(m, s)) != hipSuccess) return hip_fail(e, "GridSearch stream order");
      if (m->has_last && (e = hipStreamSynchronize(m->last_stream)) != hipSuccess)
        return hip_fail(e, "GridSearch scratch");
      if (m->d_gs_split) (void)hipFree(m->d_gs_split);
      m->d_gs_split = nullptr;
      m->gs_split_bytes = 0;
      if ((e = hipMalloc(&m->d_gs_split, bytes)) != hipSuccess) {
        m->d_gs_split = nullptr;
        return hip_fail(e, "GridSearch scratch");
      }
      m->gs_split_bytes = bytes;
    }
    if ((e = order_after_last(m, s)) != hipSuccess) return hip_fail(e, "GridSearch stream order");
  }
  if ((e = debug_reset(m, 0, s)) != hipSuccess) return hip_fail(e, "GridSearch --debug records");
  e = m->ops->gridsearch(L, g, params, (int)m->spec.grid_params.size(), m->spec.grid_params.data(), m->ncombs_valid,
                         m->d_grid_f, m->d_grid_d, m->d_comb_id, c, kept, S, S > 1 ? m->d_gs_split : nullptr, s);
  if (e != hipSuccess) return hip_fail(e, "GridSearch launch");
  if (S > 1) (void)note_last(m, s);
  return 0;
}

int cdm_levmar(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
               const float* fixp, int max_iters, int marquardt, int* iters, float* cost, void* stream) {
  int r = check_geom(m, nvox, nmeas, meas, cfp, fixp, true);
  if (r) return r;
  if (max_iters < 0) return fail("--iterLevMar must be >= 0");
  if (nvox == 0) return 0;
  const cdm::Geom g = make_geom(m, nvox, nmeas, meas, cfp, fixp);
  if (!m->d_queue) {
    const hipError_t ea = hipMalloc((void**)&m->d_queue, sizeof(int));
    if (ea != hipSuccess) {
      m->d_queue = nullptr;
      return hip_fail(ea, "Levenberg-Marquardt work queue");
    }
  }
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess)
    ncu = 0;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e = order_after_last(m, s);
  if (e != hipSuccess) return hip_fail(e, "Levenberg-Marquardt stream order");
  if ((e = debug_reset(m, 1, s)) != hipSuccess) return hip_fail(e, "Levenberg-Marquardt --debug records");
  e = m->ops->levmar(group_for(m, cdm::kLevMar, nmeas, nvox), occ_for(m, cdm::kLevMar, nmeas), g, params, max_iters,
                     marquardt ? 1 : 0, make_consts(m), iters, cost, m->d_queue, ncu, s);
  if (e != hipSuccess) return hip_fail(e, "Levenberg-Marquardt launch");
  e = note_last(m, s);
  return e == hipSuccess ? 0 : hip_fail(e, "Levenberg-Marquardt stream order");
}

}  // extern "C"

namespace {

// Argument checks shared by cdm_mcmc and cdm_mcmc_noalloc (no device work).
int check_mcmc(cdm_model* m, int nvox, int nmeas, const float* meas, const float* cfp, const float* fixp, int njumps,
               int sampleevery, int rician, const float* samples, const float* tau_samples) {
  int r = check_geom(m, nvox, nmeas, meas, cfp, fixp, true);
  if (r) return r;
  if (sampleevery <= 0) return fail("--sampleevery must be > 0");
  if (njumps < sampleevery)
    return fail("If MCMC is run, the method must record at least one sample. The number of jumps (--nj) must be "
                "greater than or equal to space between samples (--se)");
  if (!samples && nvox > 0) return fail("samples pointer is null");
  if (rician && !tau_samples && nvox > 0) return fail("--rician needs tau_samples");
  if (nvox > 0 && njumps % sampleevery != 0)
    return fail("--nj must be a multiple of --se (the reference writes past the sample buffer otherwise)");
  return 0;
}

// Register budget an MCMC launch runs at (the Rician likelihood is compiled
// at the default budget only); with the voxel group it decides where the
// per-measurement cache rows live.
int mcmc_budget(const cdm_model* m, int nmeas, int rician) { return rician ? 1 : occ_for(m, cdm::kMcmc, nmeas); }

// Bytes of chain state an MCMC call over nvox voxels needs: the chains, then
// the per-measurement cache rows unless the launch keeps them in LDS.
size_t mcmc_state_bytes(const cdm_model* m, int nvox, int nmeas, const float* cfp, const float* fixp, int rician) {
  const int L = group_for(m, cdm::kMcmc, nmeas, nvox);
  const cdm::Geom g = make_geom(m, nvox, nmeas, nullptr, cfp, fixp);
  if (!m->ops->mvc_bytes_per_meas || m->ops->mcmc_rows_lds(L, mcmc_budget(m, nmeas, rician), g))
    return m->ops->chain_bytes * (size_t)nvox;
  return mvc_offset(m, nvox) + m->ops->mvc_bytes_per_meas * 2 * (size_t)nmeas * (size_t)nvox;
}

// Burn-in then recording, each phase as one launch or as blocks of
// jumps_per_launch jumps (0: whole phases), over the chain state at `state`
// (nvox voxels, their cache rows after it).  use_graph: record a blocked
// phase into a HIP graph kept by the handle and replayed by later calls.
int run_mcmc(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp, const float* fixp,
             int nburn, int njumps, int sampleevery, int update_proposal, int rician, unsigned long long seed,
             long long voxel_offset, float* samples, float* tau_samples, void* state, int jumps_per_launch,
             bool use_graph, hipStream_t s) {
  const int nb = nburn < 1 ? 1 : nburn;  // MCMC.cu:702-703
  const int nsamples = njumps / sampleevery;
  const cdm::Geom g = make_geom(m, nvox, nmeas, meas, cfp, fixp);
  const cdm::Consts c = make_consts(m);
  const int L = group_for(m, cdm::kMcmc, nmeas, nvox);
  const int W = occ_for(m, cdm::kMcmc, nmeas);

  cdm::McmcLaunch ml;
  ml.params = params;
  ml.state = state;
  ml.mvc = (m->ops->mvc_bytes_per_meas && !m->ops->mcmc_rows_lds(L, mcmc_budget(m, nmeas, rician), g))
               ? (void*)((char*)state + mvc_offset(m, nvox))
               : nullptr;
  ml.samples = samples;
  ml.tau_samples = tau_samples;
  ml.nsamples = nsamples;
  ml.sampleevery = sampleevery;
  ml.upe = sampleevery;
  ml.seed = seed;
  ml.voxel_offset = voxel_offset;
  for (int phase = 0; phase < 2; ++phase) {
    const int len = phase == 0 ? nb : njumps;
    const int blk = jumps_per_launch > 0 ? jumps_per_launch : len;
    ml.recording = phase;
    ml.update_prop = phase == 0 ? 1 : (update_proposal ? 1 : 0);
    ml.phase_len = len;
    ml.jump_offset = phase == 0 ? 0u : (unsigned)nb;
    auto issue = [&](hipStream_t q) -> hipError_t {
      for (int b = 0; b < len; b += blk) {
        ml.begin = b;
        ml.count = std::min(blk, len - b);
        const hipError_t e = m->ops->mcmc(L, W, rician ? 1 : 0, g, ml, c, q);
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    };
    hipError_t e;
    if (use_graph && len > blk) {
      // the phase's graph is replayed while every argument its launches
      // bake in is the same (the jump-block offsets follow from len, blk)
      std::vector<unsigned char> key;
      cdm::McmcLaunch mk = ml;
      mk.begin = mk.count = 0;
      key_add(&key, g);
      key_add(&key, c);
      key_add(&key, mk);
      const int shape[5] = {L, W, rician, blk, len};
      key_add(&key, shape);
      e = run_graph_cached(m->mcmc_graphs[phase], &m->capture_stream, &m->graph_stats, key, s, issue);
    } else {
      e = issue(s);
    }
    if (e != hipSuccess) return hip_fail(e, "MCMC launch");
  }
  return 0;
}

// MCMC over the handle's own chain state, ordered after the handle's
// previous call (arguments already checked).
int mcmc_call(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp, const float* fixp,
              int nburn, int njumps, int sampleevery, int update_proposal, int rician, unsigned long long seed,
              long long voxel_offset, float* samples, float* tau_samples, int jumps_per_launch, bool use_graph,
              hipStream_t s) {
  int r = ensure_state(m, mcmc_state_bytes(m, nvox, nmeas, cfp, fixp, rician));
  if (r) return r;
  hipError_t e = order_after_last(m, s);
  if (e != hipSuccess) return hip_fail(e, "MCMC stream order");
  if ((e = debug_reset(m, 2, s)) != hipSuccess) return hip_fail(e, "MCMC --debug records");
  r = run_mcmc(m, nvox, nmeas, meas, params, cfp, fixp, nburn, njumps, sampleevery, update_proposal, rician, seed,
               voxel_offset, samples, tau_samples, m->d_state, jumps_per_launch, use_graph, s);
  if (r) return r;
  e = note_last(m, s);
  return e == hipSuccess ? 0 : hip_fail(e, "MCMC stream order");
}

}  // namespace

extern "C" {

int cdm_mcmc(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
             const float* fixp, int nburn, int njumps, int sampleevery, int update_proposal, int rician,
             unsigned long long seed, long long voxel_offset, float* samples, float* tau_samples, void* stream) {
  int r = check_mcmc(m, nvox, nmeas, meas, cfp, fixp, njumps, sampleevery, rician, samples, tau_samples);
  if (r || nvox == 0) return r;
  return mcmc_call(m, nvox, nmeas, meas, params, cfp, fixp, nburn, njumps, sampleevery, update_proposal, rician, seed,
                   voxel_offset, samples, tau_samples, m->mcmc_block, m->mcmc_graph != 0, (hipStream_t)stream);
}

int cdm_mcmc_noalloc(cdm_model* m, int nvox, int nmeas, const float* meas, float* params, const float* cfp,
                     const float* fixp, int nburn, int njumps, int sampleevery, int update_proposal, int rician,
                     unsigned long long seed, long long voxel_offset, float* samples, float* tau_samples,
                     void* stream) {
  int r = check_mcmc(m, nvox, nmeas, meas, cfp, fixp, njumps, sampleevery, rician, samples, tau_samples);
  if (r || nvox == 0) return r;
  const size_t need = mcmc_state_bytes(m, nvox, nmeas, cfp, fixp, rician);
  if (need > cdm::kNoAllocScratchBytes) {
    const size_t per = (need + (size_t)nvox - 1) / (size_t)nvox;
    return fail("cdm_mcmc_noalloc: " + std::to_string(nvox) + " voxels x " + std::to_string(nmeas) +
                " measurements need " + std::to_string(need) + " bytes of chain state, more than the " +
                std::to_string(cdm::kNoAllocScratchBytes) + "-byte allocation-free scratch (at most about " +
                std::to_string(cdm::kNoAllocScratchBytes / per) +
                " voxels per call for this model; split the volume and pass each part's voxel_offset)");
  }
  void* scratch = nullptr;
  hipError_t e = cdm::noalloc_scratch(&scratch);
  if (e != hipSuccess) return hip_fail(e, "allocation-free scratch lookup");
  if ((e = debug_reset(m, 2, (hipStream_t)stream)) != hipSuccess) return hip_fail(e, "MCMC --debug records");
  return run_mcmc(m, nvox, nmeas, meas, params, cfp, fixp, nburn, njumps, sampleevery, update_proposal, rician, seed,
                  voxel_offset, samples, tau_samples, scratch, m->mcmc_block, false, (hipStream_t)stream);
}

int cdm_predicted_signal(cdm_model* m, int nvox, int nmeas, int nsamples, const float* samples, const float* cfp,
                         const float* fixp, float* out, void* stream) {
  int r = check_geom(m, nvox, nmeas, nullptr, cfp, fixp, false);
  if (r) return r;
  if (nsamples <= 0) return fail("nsamples must be > 0");
  if (nvox == 0) return 0;
  const cdm::Geom g = make_geom(m, nvox, nmeas, nullptr, cfp, fixp);
  const hipError_t e =
      m->ops->predicted(group_for(m, cdm::kPredicted, nmeas, nvox), g, nsamples, samples, out, (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "predicted-signal launch");
}

int cdm_bic_aic(cdm_model* m, int nvox, int nmeas, int nsamples, int rician, const float* meas, const float* samples,
                const float* tau_samples, const float* cfp, const float* fixp, float* bic, float* aic, void* stream) {
  int r = check_geom(m, nvox, nmeas, meas, cfp, fixp, true);
  if (r) return r;
  if (nsamples <= 0) return fail("nsamples must be > 0");
  if (rician && !tau_samples && nvox > 0) return fail("Rician BIC/AIC needs tau samples");
  if (nvox == 0) return 0;
  const cdm::Geom g = make_geom(m, nvox, nmeas, meas, cfp, fixp);
  const hipError_t e = m->ops->bic_aic(group_for(m, cdm::kBicAic, nmeas, nvox), rician ? 1 : 0, g, nsamples, samples,
                                       tau_samples, bic, aic, (hipStream_t)stream);
  return e == hipSuccess ? 0 : hip_fail(e, "BIC/AIC launch");
}

int cdm_rician_clean(int nvox, int nmeas, float* meas) {
  if (nvox < 0 || nmeas <= 0 || (!meas && nvox > 0)) return fail("bad arguments");
  for (int v = 0; v < nvox; ++v) {
    float* x = meas + (size_t)v * nmeas;
    float mn = x[0], mx = x[0];
    for (int k = 1; k < nmeas; ++k) {
      mn = std::min(mn, x[k]);
      mx = std::max(mx, x[k]);
    }
    if (mn <= 0 && mx > 0) {
      float minpos = mx;
      for (int k = 0; k < nmeas; ++k)
        if (x[k] > 0 && x[k] < minpos) minpos = x[k];
      for (int k = 0; k < nmeas; ++k)
        if (x[k] <= 0) x[k] = minpos;
    }
  }
  return 0;
}

int cdm_batch_units(char* buf, int buflen) {
  std::string s;
  for (int i = 0; i < cdm::batch_count(); ++i) {
    const cdm::BatchOps* b = cdm::batch_at(i);
    if (i) s += ";";
    s += b->name;
    s += ":";
    for (int k = 0; k < b->nmodels; ++k) {
      if (k) s += ",";
      s += b->models[k];
    }
  }
  if (buf && buflen > 0) {
    std::strncpy(buf, s.c_str(), (size_t)buflen - 1);
    buf[buflen - 1] = 0;
  }
  return (int)s.size();
}

}  // extern "C"

namespace {

// Batch unit holding every item's model, and each item's model index in it.
const cdm::BatchOps* find_batch(int nitems, const cdm_batch_item* items, std::vector<int>* index) {
  for (int u = 0; u < cdm::batch_count(); ++u) {
    const cdm::BatchOps* b = cdm::batch_at(u);
    std::vector<int> idx;
    for (int i = 0; i < nitems; ++i) {
      int k = -1;
      for (int j = 0; j < b->nmodels; ++j)
        if (items[i].model && std::strcmp(b->models[j], items[i].model->ops->name) == 0) k = j;
      if (k < 0) break;
      idx.push_back(k);
    }
    if ((int)idx.size() == nitems) {
      *index = idx;
      return b;
    }
  }
  return nullptr;
}

int check_batch(int nitems, const cdm_batch_item* items, bool mcmc, const cdm::BatchOps** unit,
                std::vector<int>* index, long long* total_vox) {
  if (nitems <= 0 || !items) return fail("a batch needs at least one item");
  *total_vox = 0;
  for (int i = 0; i < nitems; ++i) {
    const cdm_batch_item& it = items[i];
    int r = check_geom(it.model, it.nvox, it.nmeas, it.meas, it.cfp, it.fixp, true);
    if (r) return fail("batch item " + std::to_string(i) + ": " + g_err);
    if (it.nvox <= 0) return fail("batch item " + std::to_string(i) + ": no voxels");
    if (!it.params) return fail("batch item " + std::to_string(i) + ": params pointer is null");
    if (mcmc && !it.samples) return fail("batch item " + std::to_string(i) + ": samples pointer is null");
    for (int j = 0; j < i; ++j)
      if (items[j].model == it.model)
        return fail("batch items " + std::to_string(j) + " and " + std::to_string(i) +
                    " share a model handle (each item needs its own handle: it owns the chain state)");
    *total_vox += it.nvox;
  }
  *unit = find_batch(nitems, items, index);
  if (!*unit) {
    std::string names;
    for (int i = 0; i < nitems; ++i) names += (i ? "," : "") + std::string(items[i].model->ops->name);
    return fail("no batch unit compiled for the models " + names +
                " (build with python -m cudimot_amd.build --batch NAME=" + names + ")");
  }
  return 0;
}

// An item whose voxels alone fill the GPU several times over (blocks of its
// tuned voxel group >= 8 per CU by default, cdm_batch_set_fill) runs through
// the per-model kernels at its
// own tuned launch shape, voxel group and register budget, with the same
// draws and results; the persistent kernel, whose register allocation is the
// largest of its models' and sets one occupancy for all of them, takes the
// smaller items, which it packs together to fill the GPU.
std::atomic<int> g_fill_blocks_per_cu{8};

bool fills_gpu(const cdm_model* m, int method, int nmeas, int nvox) {
  const int per_cu = g_fill_blocks_per_cu.load();
  if (per_cu <= 0) return false;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) {
    (void)hipGetLastError();
    return false;
  }
  const long long per_block = 256 / group_for(m, method, nmeas, nvox);
  return ((long long)nvox + per_block - 1) / per_block >= (long long)per_cu * ncu;
}

// Items ordered by estimated work per voxel (heaviest first), so the queue
// hands out the long tiles early and the short ones fill the tail.  The
// estimate is the tuning table's time for the model and method when present,
// else measurements x parameters.
std::vector<int> batch_order(int nitems, const cdm_batch_item* items, int method) {
  std::vector<double> w(nitems);
  for (int i = 0; i < nitems; ++i) {
    const cdm_model* m = items[i].model;
    double t = 0;
    for (const auto& e : m->tuned)
      if (e.method == method && e.nmeas > 0) t = std::max(t, e.ms * items[i].nmeas / e.nmeas);
    w[i] = t > 0 ? t : (double)items[i].nmeas * m->ops->nparams;
  }
  std::vector<int> order(nitems);
  for (int i = 0; i < nitems; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return w[a] > w[b]; });
  return order;
}

// Device item table, work-queue counters and pinned host staging of the
// batched launches, one set per (method, device), grown as needed and kept
// for the process.  A call first waits (on the host) for the previous call on
// the same set, then rewrites it; its own launches stay asynchronous on the
// caller's stream.  Calls are serialised by g_batch_mutex.
struct BatchWorkspace {
  void* d_items = nullptr;
  size_t items_cap = 0;
  int* d_queues = nullptr;
  size_t queues_cap = 0;
  unsigned char* h_stage = nullptr;
  size_t stage_cap = 0;
  hipEvent_t done = nullptr;
  bool pending = false;
  GraphCache graphs[2];  // per MCMC phase
  hipStream_t capture_stream = nullptr;
};

constexpr int kBatchDevices = 64;
std::mutex g_batch_mutex;
BatchWorkspace g_batch[2][kBatchDevices];
GraphStats g_batch_stats;

hipError_t batch_acquire(int mode, size_t item_bytes, size_t nqueues, BatchWorkspace** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kBatchDevices) return hipErrorInvalidDevice;
  BatchWorkspace& w = g_batch[mode][dev];
  if (w.pending && (e = hipEventSynchronize(w.done)) != hipSuccess) return e;
  w.pending = false;
  if (!w.done && (e = hipEventCreateWithFlags(&w.done, hipEventDisableTiming)) != hipSuccess) {
    w.done = nullptr;
    return e;
  }
  if (w.items_cap < item_bytes) {
    if (w.d_items) (void)hipFree(w.d_items);
    w.d_items = nullptr;
    w.items_cap = 0;
    if ((e = hipMalloc(&w.d_items, item_bytes)) != hipSuccess) return e;
    w.items_cap = item_bytes;
  }
  if (w.stage_cap < item_bytes) {
    if (w.h_stage) (void)hipHostFree(w.h_stage);
    w.h_stage = nullptr;
    w.stage_cap = 0;
    if ((e = hipHostMalloc((void**)&w.h_stage, item_bytes, hipHostMallocDefault)) != hipSuccess) return e;
    w.stage_cap = item_bytes;
  }
  if (w.queues_cap < nqueues) {
    if (w.d_queues) (void)hipFree(w.d_queues);
    w.d_queues = nullptr;
    w.queues_cap = 0;
    if ((e = hipMalloc((void**)&w.d_queues, nqueues * sizeof(int))) != hipSuccess) return e;
    w.queues_cap = nqueues;
  }
  *out = &w;
  return hipSuccess;
}

// Stages the prepared items into the workspace and copies them to the device.
hipError_t batch_upload(BatchWorkspace* w, const std::vector<unsigned char>& host, hipStream_t s) {
  std::memcpy(w->h_stage, host.data(), host.size());
  return hipMemcpyAsync(w->d_items, w->h_stage, host.size(), hipMemcpyHostToDevice, s);
}

// Orders the items' handles' work on s and records its end on them.
hipError_t batch_order(int nitems, const cdm_batch_item* items, hipStream_t s) {
  for (int i = 0; i < nitems; ++i) {
    const hipError_t e = order_after_last(items[i].model, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
hipError_t batch_note(int nitems, const cdm_batch_item* items, BatchWorkspace* w, hipStream_t s) {
  for (int i = 0; i < nitems; ++i) {
    const hipError_t e = note_last(items[i].model, s);
    if (e != hipSuccess) return e;
  }
  hipStreamCaptureStatus cst = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(s, &cst);
  if (e != hipSuccess || cst != hipStreamCaptureStatusNone) return e;
  if ((e = hipEventRecord(w->done, s)) != hipSuccess) return e;
  w->pending = true;
  return hipSuccess;
}

}  // namespace

extern "C" {

// One persistent launch per method over all the items (the work queue hands
// out tiles of every model).
static int batch_levmar_persistent(int nitems, const cdm_batch_item* items, int max_iters, int marquardt,
                                   void* stream) {
  const cdm::BatchOps* unit = nullptr;
  std::vector<int> index;
  long long total = 0;
  int r = check_batch(nitems, items, false, &unit, &index, &total);
  if (r) return r;
  hipStream_t s = (hipStream_t)stream;
  const std::vector<int> order = batch_order(nitems, items, cdm::kLevMar);
  std::vector<cdm::BatchLaunchItem> li(nitems);
  for (int k = 0; k < nitems; ++k) {
    const cdm_batch_item& it = items[order[k]];
    cdm::BatchLaunchItem& b = li[k];
    b.model = index[order[k]];
    b.L = group_for(it.model, cdm::kLevMar, it.nmeas, (int)std::min<long long>(total, INT_MAX));
    b.g = make_geom(it.model, it.nvox, it.nmeas, it.meas, it.cfp, it.fixp);
    b.c = make_consts(it.model);
    b.params = it.params;
    b.max_iters = max_iters;
    b.iters = it.iters;
    b.cost = it.cost;
  }
  std::vector<unsigned char> host(unit->item_bytes * nitems);
  int tiles = 0;
  size_t lds = 0;
  if (unit->prepare(cdm::kBatchLevMar, nitems, li.data(), host.data(), &tiles, &lds))
    return fail("batch LM: an item does not fit the engine limits");
  std::lock_guard<std::mutex> lock(g_batch_mutex);
  BatchWorkspace* w = nullptr;
  hipError_t e = batch_acquire(cdm::kBatchLevMar, host.size(), 1, &w);
  if (e == hipSuccess) e = batch_order(nitems, items, s);
  for (int i = 0; i < nitems && e == hipSuccess; ++i) e = debug_reset(items[i].model, 1, s);
  if (e == hipSuccess) e = batch_upload(w, host, s);
  if (e == hipSuccess) e = cdm::zero_ints(w->d_queues, 1, s);
  int grid = 0;
  if (e == hipSuccess) e = unit->grid(cdm::kBatchLevMar, tiles, lds, &grid);
  if (e == hipSuccess)
    e = unit->launch(cdm::kBatchLevMar, w->d_items, nitems, tiles, w->d_queues, lds, marquardt ? 1 : 0, grid, s);
  if (e == hipSuccess) e = batch_note(nitems, items, w, s);
  return e == hipSuccess ? 0 : hip_fail(e, "batch LM");
}

static int batch_mcmc_persistent(int nitems, const cdm_batch_item* items, int nburn, int njumps, int sampleevery,
                                 int update_proposal, int rician, unsigned long long seed, int jumps_per_launch,
                                 int use_graph, void* stream) {
  const cdm::BatchOps* unit = nullptr;
  std::vector<int> index;
  long long total = 0;
  int r = check_batch(nitems, items, true, &unit, &index, &total);
  if (r) return r;
  if (sampleevery <= 0) return fail("--sampleevery must be > 0");
  if (njumps < sampleevery || njumps % sampleevery != 0)
    return fail("--nj must be a positive multiple of --se");
  if (jumps_per_launch < 0) return fail("jumps per launch must be >= 0");
  const int nb = nburn < 1 ? 1 : nburn;  // MCMC.cu:702-703
  const int nsamples = njumps / sampleevery;
  const int mode = rician ? cdm::kBatchMcmcRician : cdm::kBatchMcmc;
  hipStream_t s = (hipStream_t)stream;
  for (int i = 0; i < nitems; ++i)
    if (rician && !items[i].tau_samples) return fail("batch item " + std::to_string(i) + ": --rician needs tau_samples");
  for (int i = 0; i < nitems; ++i) {
    const cdm_model* im = items[i].model;
    const size_t bytes = mvc_offset(im, items[i].nvox) +
                         im->ops->mvc_bytes_per_meas * 2 * (size_t)items[i].nmeas * (size_t)items[i].nvox;
    if ((r = ensure_state(items[i].model, bytes))) return r;
  }
  const std::vector<int> order = batch_order(nitems, items, cdm::kMcmc);
  // one prepared item set per launch: phases x jump blocks
  struct Launch {
    int phase, begin, count;
  };
  std::vector<Launch> launches;
  for (int phase = 0; phase < 2; ++phase) {
    const int len = phase == 0 ? nb : njumps;
    const int blk = jumps_per_launch > 0 ? jumps_per_launch : len;
    for (int b = 0; b < len; b += blk) launches.push_back({phase, b, std::min(blk, len - b)});
  }
  const size_t per = unit->item_bytes * nitems;
  std::vector<unsigned char> host(per * launches.size());
  std::vector<cdm::BatchLaunchItem> li(nitems);
  int tiles = 0;
  size_t lds = 0;
  for (size_t q = 0; q < launches.size(); ++q) {
    const Launch& L = launches[q];
    for (int k = 0; k < nitems; ++k) {
      const cdm_batch_item& it = items[order[k]];
      cdm::BatchLaunchItem& b = li[k];
      b.model = index[order[k]];
      b.L = group_for(it.model, cdm::kMcmc, it.nmeas, (int)std::min<long long>(total, INT_MAX));
      b.g = make_geom(it.model, it.nvox, it.nmeas, it.meas, it.cfp, it.fixp);
      b.c = make_consts(it.model);
      b.params = it.params;
      cdm::McmcLaunch& ml = b.mc;
      ml.params = it.params;
      ml.state = it.model->d_state;
      ml.mvc = mvc_rows(it.model, it.nvox);
      ml.samples = it.samples;
      ml.tau_samples = it.tau_samples;
      ml.nsamples = nsamples;
      ml.sampleevery = sampleevery;
      ml.upe = sampleevery;
      ml.seed = seed;
      ml.voxel_offset = it.voxel_offset;
      ml.recording = L.phase;
      ml.update_prop = L.phase == 0 ? 1 : (update_proposal ? 1 : 0);
      ml.phase_len = L.phase == 0 ? nb : njumps;
      ml.jump_offset = L.phase == 0 ? 0u : (unsigned)nb;
      ml.begin = L.begin;
      ml.count = L.count;
    }
    if (unit->prepare(mode, nitems, li.data(), host.data() + q * per, &tiles, &lds))
      return fail("batch MCMC: an item does not fit the engine limits");
  }
  std::lock_guard<std::mutex> lock(g_batch_mutex);
  BatchWorkspace* w = nullptr;
  hipError_t e = batch_acquire(cdm::kBatchMcmc, host.size(), launches.size(), &w);
  if (e == hipSuccess) e = batch_order(nitems, items, s);
  for (int i = 0; i < nitems && e == hipSuccess; ++i) e = debug_reset(items[i].model, 2, s);
  if (e == hipSuccess) e = batch_upload(w, host, s);
  if (e != hipSuccess) return hip_fail(e, "batch MCMC set-up");
  int grid = 0;
  if ((e = unit->grid(mode, tiles, lds, &grid)) != hipSuccess) return hip_fail(e, "batch MCMC set-up");
  size_t q = 0;
  for (int phase = 0; phase < 2 && e == hipSuccess; ++phase) {
    size_t q_end = q;
    while (q_end < launches.size() && launches[q_end].phase == phase) ++q_end;
    const size_t q0 = q;
    // the phase's counters are zeroed by a kernel inside the sequence, so a
    // replayed graph resets them itself
    auto issue = [&](hipStream_t st) -> hipError_t {
      hipError_t el = cdm::zero_ints(w->d_queues + q0, (int)(q_end - q0), st);
      for (size_t k = q0; k < q_end && el == hipSuccess; ++k)
        el = unit->launch(mode, (unsigned char*)w->d_items + k * per, nitems, tiles, w->d_queues + k, lds,
                          0, grid, st);
      return el;
    };
    if (use_graph && q_end - q0 > 1) {
      std::vector<unsigned char> key;
      const long long shape[9] = {(long long)(size_t)w->d_items, (long long)(size_t)w->d_queues, (long long)per,
                                  nitems, tiles, (long long)lds, grid, (long long)((q0 << 32) | q_end), mode};
      key_add(&key, shape);
      e = run_graph_cached(w->graphs[phase], &w->capture_stream, &g_batch_stats, key, s, issue);
    } else {
      e = issue(s);
    }
    q = q_end;
  }
  if (e == hipSuccess) e = batch_note(nitems, items, w, s);
  return e == hipSuccess ? 0 : hip_fail(e, "batch MCMC");
}

int cdm_batch_levmar(int nitems, const cdm_batch_item* items, int max_iters, int marquardt, void* stream) {
  const cdm::BatchOps* unit = nullptr;
  std::vector<int> index;
  long long total = 0;
  int r = check_batch(nitems, items, false, &unit, &index, &total);
  if (r) return r;
  if (max_iters < 0) return fail("--iterLevMar must be >= 0");
  std::vector<cdm_batch_item> rest;
  for (int i = 0; i < nitems; ++i) {
    const cdm_batch_item& it = items[i];
    if (!fills_gpu(it.model, cdm::kLevMar, it.nmeas, it.nvox)) {
      rest.push_back(it);
      continue;
    }
    r = cdm_levmar(it.model, it.nvox, it.nmeas, it.meas, it.params, it.cfp, it.fixp, max_iters, marquardt, it.iters,
                   it.cost, stream);
    if (r) return fail("batch item " + std::to_string(i) + ": " + g_err);
  }
  if (rest.empty()) return 0;
  return batch_levmar_persistent((int)rest.size(), rest.data(), max_iters, marquardt, stream);
}

int cdm_batch_mcmc(int nitems, const cdm_batch_item* items, int nburn, int njumps, int sampleevery,
                   int update_proposal, int rician, unsigned long long seed, int jumps_per_launch, int use_graph,
                   void* stream) {
  const cdm::BatchOps* unit = nullptr;
  std::vector<int> index;
  long long total = 0;
  int r = check_batch(nitems, items, true, &unit, &index, &total);
  if (r) return r;
  if (sampleevery <= 0) return fail("--sampleevery must be > 0");
  if (njumps < sampleevery || njumps % sampleevery != 0)
    return fail("--nj must be a positive multiple of --se");
  if (jumps_per_launch < 0) return fail("jumps per launch must be >= 0");
  for (int i = 0; i < nitems; ++i)
    if (rician && !items[i].tau_samples) return fail("batch item " + std::to_string(i) + ": --rician needs tau_samples");
  std::vector<cdm_batch_item> rest;
  for (int i = 0; i < nitems; ++i) {
    const cdm_batch_item& it = items[i];
    if (!fills_gpu(it.model, cdm::kMcmc, it.nmeas, it.nvox)) {
      rest.push_back(it);
      continue;
    }
    r = mcmc_call(it.model, it.nvox, it.nmeas, it.meas, it.params, it.cfp, it.fixp, nburn, njumps, sampleevery,
                  update_proposal, rician, seed, it.voxel_offset, it.samples, it.tau_samples, jumps_per_launch,
                  use_graph != 0, (hipStream_t)stream);
    if (r) return fail("batch item " + std::to_string(i) + ": " + g_err);
  }
  if (rest.empty()) return 0;
  return batch_mcmc_persistent((int)rest.size(), rest.data(), nburn, njumps, sampleevery, update_proposal, rician,
                               seed, jumps_per_launch, use_graph, stream);
}

int cdm_batch_set_fill(int blocks_per_cu) {
  if (blocks_per_cu < 0) return fail("blocks per CU must be >= 0");
  g_fill_blocks_per_cu.store(blocks_per_cu);
  return 0;
}

int cdm_batch_graph_stats(long long* captures, long long* replays) {
  std::lock_guard<std::mutex> lock(g_batch_mutex);
  if (captures) *captures = g_batch_stats.captures;
  if (replays) *replays = g_batch_stats.replays;
  return 0;
}

int cdm_sync(void* stream) {
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "device synchronise");
  e = hipGetLastError();
  return e == hipSuccess ? 0 : hip_fail(e, "device error");
}

}  // extern "C"
