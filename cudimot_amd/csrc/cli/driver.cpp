// Host driver: the reference's split_parts_X, X (the fit) and merge_parts_X
// programs over the engine's C-ABI, with their options, part-file formats,
// checks and messages (split_parts.cc:37-281, cudimot.cc:25-132 with
// dMRI_Data.cu:33-152 and Parameters.cu:23-512, merge_parts.cc:21-150).
//
// The program's name selects the step the way the reference's per-model
// binaries do: split_parts_<Model>, merge_parts_<Model>, or <Model> for the
// fit; under any other name the first argument is the step (split, fit,
// merge) and --model names the plug-in.  One part is fitted in one pass:
// the reference cuts a part into sub-parts of at most 12,800 voxels to fit
// older GPUs (dMRI_Data.cu:65-79); results per voxel do not depend on it.
#include <hip/hip_runtime_api.h>
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cli/nifti_io.h"
#include "cli/options.h"
#include "cli/parts.h"
#include "cudimot_amd.h"

namespace fs = std::filesystem;
using namespace cdm_cli;

namespace {

[[noreturn]] void die(const std::string& msg) {
  std::cerr << "CUDIMOT Error: " << msg << std::endl;
  std::exit(-1);
}

void check(int rc, const char* what) {
  if (rc != 0) die(std::string(what) + ": " + cdm_last_error());
}

void hcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

std::string part_dir(const Options& o, int id) { return o.partsdir + "/part_" + std::to_string(id); }

// FSL's Log::makeDir / setthenmakeDir: a new directory (name+ if taken,
// unless --forcedir) holding the command line (cudimotoptions.cc:37-48).
void make_logdir(const Options& o, int argc, char** argv) {
  std::string d = o.logdir;
  if (!o.forcedir)
    while (fs::exists(d)) d += "+";
  std::error_code ec;
  fs::create_directories(d, ec);
  if (ec) die("cannot create log directory " + d);
  std::ofstream log(d + "/logfile");
  for (int a = 0; a < argc; ++a) log << argv[a] << " ";
  log << "\n---------------------------------------------\n\n";
  std::cout << "Log directory is: " << d << std::endl;
}

struct Model {
  cdm_model* h = nullptr;
  int nparams = 0, cfp_tsize = 0, fixp_tsize = 0, ncfp = 0, nfixp = 0;
  std::vector<int> cfp_sizes, fixp_sizes;
  std::vector<double> p_init;
  bool init_provided = false;
  ~Model() { cdm_model_close(h); }
};

void open_model(const Options& o, Model& m) {
  check(cdm_model_open(o.model.c_str(), o.priorsfile.empty() ? nullptr : o.priorsfile.c_str(), &m.h),
        "opening the model");
  check(cdm_model_info(m.h, &m.nparams, &m.cfp_tsize, &m.fixp_tsize), "model info");
  check(cdm_model_sizes(m.h, &m.ncfp, nullptr, &m.nfixp, nullptr), "model sizes");
  m.cfp_sizes.resize(m.ncfp);
  m.fixp_sizes.resize(m.nfixp);
  check(cdm_model_sizes(m.h, nullptr, m.cfp_sizes.data(), nullptr, m.fixp_sizes.data()), "model sizes");
  m.p_init.resize(m.nparams);
  int ip = 0;
  check(cdm_model_spec(m.h, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, m.p_init.data(), &ip),
        "model spec");
  m.init_provided = ip != 0;
}

// MISCMATHS::read_ascii_matrix: one matrix row per non-empty line,
// whitespace- or comma-separated numbers.
std::vector<std::vector<double>> read_ascii_matrix(const std::string& path) {
  std::ifstream in(path);
  if (!in) die("Unable to open ASCII matrix file: " + path);
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(in, line)) {
    std::replace(line.begin(), line.end(), ',', ' ');
    std::istringstream ss(line);
    std::vector<double> r;
    std::string tok;
    while (ss >> tok) {
      char* end = nullptr;
      const double v = std::strtod(tok.c_str(), &end);
      if (*end) die("non-numeric entry '" + tok + "' in " + path);
      r.push_back(v);
    }
    if (!r.empty()) rows.push_back(r);
  }
  for (const auto& r : rows)
    if (r.size() != rows[0].size()) die("rows of different lengths in " + path);
  return rows;
}

std::vector<std::string> read_lines(const std::string& path, const char* what) {
  std::ifstream in(path);
  if (!in) die(std::string("Unable to open ") + what + " file: " + path);
  std::vector<std::string> lines;
  std::string l;
  while (std::getline(in, l)) {
    if (!l.empty() && l.back() == '\r') l.pop_back();
    lines.push_back(l);
  }
  return lines;
}

// ---------------------------------------------------------------- split
void split_parts(const Options& o) {
  if (o.gridSearch.empty() && o.no_LevMar && !o.runMCMC)
    die("You must select at least one method to fit the model: GridSearch, Levenberg_Marquardt or MCMC");
  Volume data, mask;
  std::string e = read_volume(o.datafile, data);
  if (!e.empty()) die(e);
  e = read_volume(o.maskfile, mask);
  if (!e.empty()) die(e);
  std::vector<double> dataM;
  int nvox = 0;
  e = masked_matrix(data, mask, dataM, nvox);
  if (!e.empty()) die(e);
  const int nmeas = data.nt;
  if (nmeas <= 0) die("The number of diffusion-weighted measurements must be greater than 0 in the input volume");
  if (nvox <= 0) die("The number of voxels must be greater than 0");
  if (o.nParts <= 0) die("The number of parts must be greater than 0");
  if (nvox < o.nParts) die("The number of parts/jobs must be lower than number of voxels");
  for (int i = 0; i < o.nParts; ++i) {
    std::error_code ec;
    fs::create_directory(part_dir(o, i), ec);
    if (ec) die("cannot create " + part_dir(o, i) + ": " + ec.message());
  }
  const int size_part = nvox / o.nParts;
  auto first = [&](int i) { return i * size_part; };
  auto count = [&](int i) { return i == o.nParts - 1 ? nvox - first(i) : size_part; };
  auto save = [&](const std::vector<double>& M, int rows, const std::string& name) {
    for (int i = 0; i < o.nParts; ++i) {
      Part p;
      p.nvox = count(i);
      p.nrows = rows;
      p.v.assign(M.begin() + (size_t)first(i) * rows, M.begin() + (size_t)(first(i) + count(i)) * rows);
      const std::string err = write_part(part_dir(o, i) + "/" + name, p, o.part_rowmajor);
      if (!err.empty()) die(err);
    }
  };
  save(dataM, nmeas, "data");

  Model m;
  open_model(o, m);
  if (o.set("--init_params")) {
    const std::vector<std::string> lines = read_lines(o.init_params, "Initialization Parameter");
    if ((int)lines.size() > m.nparams)
      die("Too many lines for specifying the initialization of the parameters in the file: " + o.init_params +
          ". The number of lines must match the number of parameters in this model: " + std::to_string(m.nparams));
    if ((int)lines.size() != m.nparams)
      die("The number of volumes (lines) provided for initializing the parameters in: " + o.init_params +
          " does not match the number of parameters of this model: " + std::to_string(m.nparams) +
          ". If a parameter does not need initialization, its line can be empty.");
    for (int k = 0; k < m.nparams; ++k) {
      std::vector<double> vals;
      if (!lines[k].empty()) {
        Volume pv;
        e = read_volume(lines[k], pv);
        if (!e.empty()) die(e);
        if (pv.nx != mask.nx || pv.ny != mask.ny || pv.nz != mask.nz)
          die("The size of the mask and the volume used for initilizing the parameters: " + lines[k] +
              " does not match");
        if (pv.nt != 1) die("The volume used for initilize the parameters: " + lines[k] + " must be a 3D volume");
        int nv = 0;
        e = masked_matrix(pv, mask, vals, nv);
        if (!e.empty()) die(e);
      } else {
        vals.assign(nvox, m.init_provided ? m.p_init[k] : 0.0);
      }
      save(vals, 1, "ParamInit_" + std::to_string(k));
    }
  }
  if (m.nfixp > 0 && !o.set("--FixP"))
    die("Expected a list with " + std::to_string(m.nfixp) +
        " NIfTI files for specifying the Fixed Parameters of the model: Use option --FixP");
  if (m.nfixp > 0) {
    const std::vector<std::string> lines = read_lines(o.FixP, "Fixed Parameter");
    int id = 0;
    for (; id < (int)lines.size() && id < m.nfixp; ++id) {
      if (lines[id].empty())
        die("Please remove the empty lines in the file used for specifying the Fixed Parameters: " + o.FixP);
      Volume fv;
      e = read_volume(lines[id], fv);
      if (!e.empty()) die(e);
      if (fv.nx != mask.nx || fv.ny != mask.ny || fv.nz != mask.nz)
        die("The size of the mask and the volume used for specifying the Fixed Parameters: " + lines[id] +
            " does not match");
      if (fv.nt != m.fixp_sizes[id])
        die("The NIfTI file used for specifying the Fixed Parameters: " + lines[id] + " must have " +
            std::to_string(m.fixp_sizes[id]) + " volumes");
      std::vector<double> vals;
      int nv = 0;
      e = masked_matrix(fv, mask, vals, nv);
      if (!e.empty()) die(e);
      save(vals, fv.nt, "FixParam_" + std::to_string(id));
    }
    if (id != m.nfixp)
      die("The number of NIfTI files (lines) provided in the file for specifying the Fixed Parameters: " + o.FixP +
          " does not match the number of parameters of this model: " + std::to_string(m.nfixp));
  } else if (o.set("--FixP")) {
    die("This model does not need any Fixed Parameter, however the user has provided a Fixed Parameter file.");
  }
}

// ------------------------------------------------------------------ fit
template <typename T>
struct Dev {
  T* p = nullptr;
  explicit Dev(size_t n) {
    if (n) hcheck(hipMalloc((void**)&p, n * sizeof(T)), "hipMalloc");
  }
  ~Dev() {
    if (p) (void)hipFree(p);
  }
  void up(const std::vector<T>& h) { hcheck(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "copy"); }
  void down(std::vector<T>& h) { hcheck(hipMemcpy(h.data(), p, h.size() * sizeof(T), hipMemcpyDeviceToHost), "copy"); }
};

void write_rows(const Options& o, const std::string& name, int nvox, int nrows, const std::vector<float>& src,
                size_t stride, size_t offset) {
  Part p;
  p.nvox = nvox;
  p.nrows = nrows;
  p.v.resize((size_t)nvox * nrows);
  for (int v = 0; v < nvox; ++v)
    for (int r = 0; r < nrows; ++r) p.v[(size_t)v * nrows + r] = src[(size_t)v * stride + offset + r];
  const std::string e = write_part(part_dir(o, o.idPart) + "/" + name, p, o.part_rowmajor);
  if (!e.empty()) die(e);
}

void fit(const Options& o) {
  timeval t1, t2;
  gettimeofday(&t1, nullptr);
  if (o.gridSearch.empty() && o.no_LevMar && !o.runMCMC)
    die("You must select at least one method to fit the model: GridSearch, Levenberg_Marquardt or MCMC");
  Part data;
  std::string e = read_part(part_dir(o, o.idPart) + "/data", data, o.part_rowmajor);
  if (!e.empty()) die(e);
  const int nvox = data.nvox, nmeas = data.nrows;
  std::cout << "Number of Voxels to compute: " << nvox << std::endl;
  std::cout << "Number of Measurements: " << nmeas << std::endl;

  Model m;
  open_model(o, m);
  if (!o.fixed.empty()) check(cdm_model_set_fixed(m.h, o.fixed.c_str()), "--fixed");
  if (!o.gridSearch.empty()) check(cdm_model_load_grid(m.h, o.gridSearch.c_str(), 0, nullptr, nullptr), "--gridSearch");
  const int P = m.nparams;

  // Starting values: --init_params parts > P_init > 0 (Parameters.cu:42-87).
  std::vector<float> params((size_t)nvox * P);
  for (int k = 0; k < P; ++k) {
    if (o.set("--init_params")) {
      Part pi;
      const std::string f = part_dir(o, o.idPart) + "/ParamInit_" + std::to_string(k);
      e = read_part(f, pi, o.part_rowmajor);
      if (!e.empty()) die(e);
      if (pi.nvox != nvox || pi.nrows != 1)
        die("The amount of data in the input file " + f + " for initializing the parameters is not correct");
      for (int v = 0; v < nvox; ++v) params[(size_t)v * P + k] = (float)pi.v[v];
    } else {
      for (int v = 0; v < nvox; ++v) params[(size_t)v * P + k] = m.init_provided ? (float)m.p_init[k] : 0.0f;
    }
  }
  // Common fixed parameters: one ASCII matrix (size x nmeas) per CFP
  // (Parameters.cu:93-142), stored [nmeas][CFP_Tsize].
  std::vector<float> cfp((size_t)nmeas * m.cfp_tsize);
  if (o.set("--CFP")) {
    int nset = 0, cum = 0;
    for (const std::string& line : read_lines(o.CFP, "Common Fixed Parameters")) {
      if (line.empty()) continue;
      if (nset >= m.ncfp)
        die("The number of Common Fixed Parameters provided in file: " + o.CFP +
            " is not correct for this model. The number of Common Fixed Parameters must be " + std::to_string(m.ncfp));
      const auto vals = read_ascii_matrix(line);
      const int size = m.cfp_sizes[nset];
      if ((int)vals.size() != size)
        die("Common Fixed Parameters number " + std::to_string(nset) + " and file " + line +
            " do not match the dimensions specified for this model");
      if ((int)vals[0].size() != nmeas)
        die("Common Fixed Parameters number " + std::to_string(nset) + " and file " + line +
            " do not match the number of measurements: " + std::to_string(nmeas));
      for (int i = 0; i < size; ++i)
        for (int j = 0; j < nmeas; ++j) cfp[(size_t)cum + i + (size_t)j * m.cfp_tsize] = (float)vals[i][j];
      ++nset;
      cum += size;
    }
    if (nset != m.ncfp)
      die("The number of Common Fixed Parameters provided in file: " + o.CFP +
          " is not correct for this model. The number of Common Fixed Parameters must be " + std::to_string(m.ncfp));
  } else if (m.cfp_tsize) {
    die("The user must provide a file with a list of common fixed parameters for this model: Use option --CFP");
  }
  // Per-voxel fixed parameters from the FixParam_k parts (Parameters.cu:151-200).
  std::vector<float> fixp((size_t)nvox * m.fixp_tsize);
  if (o.set("--FixP")) {
    int cum = 0;
    for (int k = 0; k < m.nfixp; ++k) {
      Part fp;
      const std::string f = part_dir(o, o.idPart) + "/FixParam_" + std::to_string(k);
      e = read_part(f, fp, o.part_rowmajor);
      if (!e.empty()) die(e);
      if (fp.nvox != nvox || fp.nrows != m.fixp_sizes[k])
        die("The amount of data in the intermediate file " + f + " with Fixed Parameters is not correct");
      for (int v = 0; v < nvox; ++v)
        for (int r = 0; r < fp.nrows; ++r) fixp[(size_t)v * m.fixp_tsize + cum + r] = (float)fp.v[(size_t)v * fp.nrows + r];
      cum += fp.nrows;
    }
  } else if (m.fixp_tsize) {
    die("The user must provide a file with a list of Fixed Parameters for this model: Use option --FixP");
  }

  std::vector<float> meas(data.v.begin(), data.v.end());
  if (o.rician) check(cdm_rician_clean(nvox, nmeas, meas.data()), "Rician data preparation");
  if (o.runMCMC && (o.sampleevery <= 0 || o.njumps % o.sampleevery != 0))
    die("--nj must be a positive multiple of --se (the reference writes past its sample buffer otherwise)");
  const int nsamples = o.runMCMC ? o.njumps / o.sampleevery : 1;

  // MCMC random streams are addressed by the voxel's index in the whole
  // volume: parts 0..N-2 hold part 0's voxel count (split_parts.cc:80-94).
  long long voxel_offset = 0;
  if (o.idPart > 0) {
    int n0 = 0, r0 = 0;
    if (read_part_header(part_dir(o, 0) + "/data", n0, r0).empty())
      voxel_offset = (long long)o.idPart * n0;
    else
      voxel_offset = (long long)o.idPart * nvox;
  }

  std::cout << "\n\n\nPart 1 of 1: processing " << nvox << " voxels" << std::endl;
  Dev<float> d_meas((size_t)nvox * nmeas), d_params((size_t)nvox * P), d_cfp(cfp.size()), d_fixp(fixp.size());
  d_meas.up(meas);
  d_params.up(params);
  if (!cfp.empty()) d_cfp.up(cfp);
  if (!fixp.empty()) d_fixp.up(fixp);
  const float* cfp_p = cfp.empty() ? nullptr : d_cfp.p;
  const float* fixp_p = fixp.empty() ? nullptr : d_fixp.p;
  // --debug=<voxel>: the methods' records of that voxel, printed after each
  // method (GridSearch.cu:284-286 reads the option the same way)
  const int debug_voxel = o.debug.empty() ? -1 : std::atoi(o.debug.c_str());
  if (debug_voxel >= 0) {
    const int cap = std::max({1, (o.runMCMC ? std::max(o.nburn, 1) + o.njumps : 0), 32 * (o.iterLevMar + 1)});
    check(cdm_model_set_debug(m.h, debug_voxel, cap), "--debug");
  }
  auto print_debug = [&](int method, const char* title) {
    if (debug_voxel < 0 || debug_voxel >= nvox) return;
    int n = 0;
    check(cdm_model_debug_trace(m.h, method, nullptr, 0, &n, nullptr), "--debug");
    std::vector<double> rec((size_t)std::max(n, 1) * (4 + P));
    check(cdm_model_debug_trace(m.h, method, rec.data(), n, &n, nullptr), "--debug");
    std::printf("\n ----- %s GPU algorithm: voxel %i -----\n", title, debug_voxel);
    for (int k = 0; k < n; ++k) {
      const double* r = &rec[(size_t)k * (4 + P)];
      if (method == CDM_GRIDSEARCH)
        std::printf("Kept combination %d of %d, cost %f\n", (int)r[0], (int)r[2], r[1]);
      else if (method == CDM_LEVMAR && k == 0)
        std::printf("Initial cost %f\n", r[1]);
      else if (method == CDM_LEVMAR)
        std::printf("Iteration %d: cost %f, lambda %g, %s\n", (int)r[0], r[1], r[2],
                    r[3] != 0.0 ? "accepted" : "rejected");
      else
        std::printf("Jump %d: energy %f, accepted since the last proposal update %d%s\n", (int)r[0], r[1],
                    (int)r[3], o.rician ? (", tau " + std::to_string(r[2])).c_str() : "");
      for (int i = 0; i < P; ++i)
        std::printf("%s Parameter[%i]: %f\n", method == CDM_LEVMAR && k == 0 ? "Initial" : "", i, r[4 + i]);
    }
    std::printf("--------------------------------------------------------\n");
  };
  if (!o.gridSearch.empty()) {
    check(cdm_gridsearch(m.h, nvox, nmeas, d_meas.p, d_params.p, cfp_p, fixp_p, nullptr, nullptr), "GridSearch");
    print_debug(CDM_GRIDSEARCH, "GridSearch");
  }
  if (!o.no_LevMar) {
    check(cdm_levmar(m.h, nvox, nmeas, d_meas.p, d_params.p, cfp_p, fixp_p, o.iterLevMar, !o.no_Marquardt, nullptr,
                     nullptr, nullptr),
          "Levenberg-Marquardt");
    print_debug(CDM_LEVMAR, "Levenberg_Marquardt");
  }
  Dev<float> d_samples(o.runMCMC ? (size_t)nvox * P * nsamples : 0);
  Dev<float> d_tau(o.runMCMC && o.rician ? (size_t)nvox * nsamples : 0);
  if (o.runMCMC) {
    check(cdm_mcmc(m.h, nvox, nmeas, d_meas.p, d_params.p, cfp_p, fixp_p, o.nburn, o.njumps, o.sampleevery,
                   !o.no_updateproposal, o.rician, (unsigned long long)o.seed, voxel_offset, d_samples.p, d_tau.p,
                   nullptr),
          "MCMC");
    print_debug(CDM_MCMC, "MCMC");
  }
  // Predicted signal and BIC/AIC from the parameters (no MCMC) or the
  // samples (Parameters.cu:320-358).
  const float* src = o.runMCMC ? d_samples.p : d_params.p;
  Dev<float> d_pred(o.getPredictedSignal ? (size_t)nvox * nmeas : 0);
  Dev<float> d_bic(o.BIC_AIC ? nvox : 0), d_aic(o.BIC_AIC ? nvox : 0);
  if (o.getPredictedSignal)
    check(cdm_predicted_signal(m.h, nvox, nmeas, nsamples, src, cfp_p, fixp_p, d_pred.p, nullptr), "predicted signal");
  if (o.BIC_AIC)
    check(cdm_bic_aic(m.h, nvox, nmeas, nsamples, o.rician && o.runMCMC, d_meas.p, src, d_tau.p, cfp_p, fixp_p,
                      d_bic.p, d_aic.p, nullptr),
          "BIC/AIC");
  check(cdm_sync(nullptr), "fit");

  // Outputs (Parameters.cu:385-512): Param_k_samples [nsamples x nvox]
  // (the final parameters when MCMC did not run), Tau_samples,
  // PredictedSignal, BIC, AIC.
  std::vector<float> samples((size_t)nvox * P * nsamples);
  if (o.runMCMC)
    d_samples.down(samples);
  else
    d_params.down(samples);
  for (int k = 0; k < P; ++k)
    write_rows(o, "Param_" + std::to_string(k) + "_samples", nvox, nsamples, samples, (size_t)P * nsamples,
               (size_t)k * nsamples);
  if (o.runMCMC && o.rician) {
    std::vector<float> tau((size_t)nvox * nsamples);
    d_tau.down(tau);
    write_rows(o, "Tau_samples", nvox, nsamples, tau, nsamples, 0);
  }
  if (o.getPredictedSignal) {
    std::vector<float> pred((size_t)nvox * nmeas);
    d_pred.down(pred);
    write_rows(o, "PredictedSignal", nvox, nmeas, pred, nmeas, 0);
  }
  if (o.BIC_AIC) {
    std::vector<float> b(nvox), a(nvox);
    d_bic.down(b);
    d_aic.down(a);
    write_rows(o, "BIC", nvox, 1, b, 1, 0);
    write_rows(o, "AIC", nvox, 1, a, 1, 0);
  }
  gettimeofday(&t2, nullptr);
  std::cout << "\nPart processed in: " << (t2.tv_sec - t1.tv_sec) + 1e-6 * (t2.tv_usec - t1.tv_usec) << " seconds"
            << std::endl;
}

// ---------------------------------------------------------------- merge
void join_parts(const Options& o, const Volume& mask, const std::string& name, int nrows) {
  std::vector<double> all;
  int total = 0;
  for (int i = 0; i < o.nParts; ++i) {
    Part p;
    const std::string f = part_dir(o, i) + "/" + name;
    std::string e = read_part(f, p, o.part_rowmajor);
    if (!e.empty()) die(e);
    if (nrows == -1) nrows = p.nrows;  // PredictedSignal: taken from the first part
    if (p.nrows != nrows) die("The amount of data in the intermediate output file: " + f + " is not correct.");
    all.insert(all.end(), p.v.begin(), p.v.end());
    total += p.nvox;
  }
  const std::string e = write_masked(o.outputdir + "/" + name, mask, all, total, nrows);
  if (!e.empty()) die(e);
}

void merge_parts(const Options& o) {
  Volume mask;
  std::string e = read_volume(o.maskfile, mask);
  if (!e.empty()) die(e);
  const int nsamples = o.runMCMC ? (o.sampleevery > 0 ? o.njumps / o.sampleevery : 0) : 1;
  if (nsamples <= 0) die("The number of samples must be greater than 0");
  if (o.nParts <= 0) die("The number of parts must be greater than 0");
  Model m;
  open_model(o, m);
  for (int k = 0; k < m.nparams; ++k) join_parts(o, mask, "Param_" + std::to_string(k) + "_samples", nsamples);
  if (o.getPredictedSignal) join_parts(o, mask, "PredictedSignal", -1);
  if (o.BIC_AIC) {
    join_parts(o, mask, "BIC", 1);
    join_parts(o, mask, "AIC", 1);
  }
  if (o.rician && o.runMCMC) join_parts(o, mask, "Tau_samples", nsamples);
  if (!o.keepTmp) {
    std::error_code ec;
    fs::remove_all(o.partsdir, ec);
  }
}

// --------------------------------------------------------------- finish
// The <Model>_finish.sh steps (mymodels/<Model>/<Model>_finish.sh): rename
// the merged Param_k_samples volumes, then write sample means, mean fibre
// and fanning orientations and the NODDI dispersion maps, computed on the
// device by cdm_sample_mean / cdm_dyads / cdm_noddi_maps.
struct FinishSpec {
  const char* model;
  std::vector<const char*> names;  // Param_k_samples -> <name>_samples (nullptr: keep)
  std::vector<int> means;          // mean_<name>
  struct Dir {
    int th, ph, psi;
    const char* out;
  };
  std::vector<Dir> dirs;  // psi < 0: make_dyadic_vectors (+ _dispersion); else getFanningOrientation
  struct Noddi {
    int kappa, beta;
    const char* sfx;
  };
  std::vector<Noddi> noddi;
  std::vector<int> fsum;     // mean_fsum: sum of these means (Ball_2_Sticks_finish.sh:45-47)
  bool tensor = false;       // DiffusionTensor: Param_1..6 -> their means, and Tensor (DiffusionTensor_finish.sh:22-31)
};

const std::vector<FinishSpec>& finish_specs() {
  static const std::vector<FinishSpec> specs = {
      // Ball_1_Stick_finish.sh:21-31
      {"Ball_1_Stick", {"S0", "d", "f1", "th1", "ph1"}, {0, 1, 2}, {{3, 4, -1, "dyads1"}}, {}, {}},
      // Ball_2_Sticks_finish.sh:22-47 (maskdyads thresholds not reproduced)
      {"Ball_2_Sticks", {"S0", "d", "f1", "th1", "ph1", "f2", "th2", "ph2"}, {0, 1, 2, 5},
       {{3, 4, -1, "dyads1"}, {6, 7, -1, "dyads2"}}, {}, {2, 5}},
      // no reference script: by analogy with Ball_2_Sticks
      {"Ball_3_Sticks", {"S0", "d", "f1", "th1", "ph1", "f2", "th2", "ph2", "f3", "th3", "ph3"}, {0, 1, 2, 5, 8},
       {{3, 4, -1, "dyads1"}, {6, 7, -1, "dyads2"}, {9, 10, -1, "dyads3"}}, {}, {2, 5, 8}},
      // NODDI_Watson_finish.sh:22-35
      {"NODDI_Watson", {"fiso", "fintra", "kappa", "th", "ph"}, {0, 1, 2}, {{3, 4, -1, "dyads1"}}, {{2, -1, ""}}, {}},
      // NODDI_Bingham_finish.sh:24-56
      {"NODDI_Bingham", {"fiso", "fintra", "kappa", "beta", "th", "ph", "psi"}, {0, 1, 2, 3},
       {{4, 5, -1, "dyads1"}, {4, 5, 6, "fanningDir"}}, {{2, 3, ""}}, {}},
      // NODDI_Bingham_2Fans_finish.sh:23-80
      {"NODDI_Bingham_2Fans",
       {"fiso", "fFan2", "fintra_f1", "kappa_f1", "beta_f1", "th_f1", "ph_f1", "psi_f1", "fintra_f2", "kappa_f2",
        "beta_f2", "th_f2", "ph_f2", "psi_f2"},
       {0, 1, 2, 3, 4, 8, 9, 10},
       {{5, 6, -1, "dyads1"}, {11, 12, -1, "dyads2"}, {5, 6, 7, "fanningDir_f1"}, {11, 12, 13, "fanningDir_f2"}},
       {{3, 4, "_f1"}, {9, 10, "_f2"}}, {}},
      // DiffusionTensor_finish.sh:20-32 (-tensor_decomp not reproduced)
      {"DiffusionTensor", {"S0", nullptr, nullptr, nullptr, nullptr, nullptr, nullptr}, {0}, {}, {}, {}, true},
  };
  return specs;
}

std::string nifti_name(const std::string& dir, const std::string& base) {
  const std::string f = find_volume(dir + "/" + base);
  if (f.empty()) die("cannot find " + dir + "/" + base + " (.nii or .nii.gz)");
  return f;
}

void finish(const std::string& model, const std::string& dir) {
  const FinishSpec* spec = nullptr;
  for (const FinishSpec& s : finish_specs())
    if (model == s.model) spec = &s;
  if (!spec) die("no finish step for model " + model);
  Volume mask;
  std::string e = read_volume(dir + "/nodif_brain_mask", mask);
  if (!e.empty()) die(e);
  const int P = (int)spec->names.size();
  int nvox = 0, ns = 0;
  std::vector<float> samples;  // [nvox][P][ns]
  for (int k = 0; k < P; ++k) {
    Volume v;
    e = read_volume(dir + "/Param_" + std::to_string(k) + "_samples", v);
    if (!e.empty()) die(e);
    std::vector<double> m;
    int nv = 0;
    e = masked_matrix(v, mask, m, nv);
    if (!e.empty()) die(e);
    if (k == 0) {
      nvox = nv;
      ns = v.nt;
      samples.assign((size_t)nvox * P * ns, 0.0f);
    } else if (v.nt != ns) {
      die("Param_" + std::to_string(k) + "_samples has " + std::to_string(v.nt) + " samples, expected " +
          std::to_string(ns));
    }
    for (int i = 0; i < nvox; ++i)
      for (int s = 0; s < ns; ++s) samples[((size_t)i * P + k) * ns + s] = (float)m[(size_t)i * ns + s];
  }
  for (int k = 0; k < P; ++k)
    if (spec->names[k]) {
      const std::string from = nifti_name(dir, "Param_" + std::to_string(k) + "_samples");
      const std::string ext = from.substr(from.find_last_of('/') + 1 + std::string("Param_" + std::to_string(k) + "_samples").size());
      fs::rename(from, dir + "/" + spec->names[k] + "_samples" + ext);
    }
  Dev<float> d_s(samples.size()), d_mean((size_t)nvox * P);
  d_s.up(samples);
  check(cdm_sample_mean(nvox, P, ns, d_s.p, d_mean.p, nullptr), "sample means");
  std::vector<float> mean((size_t)nvox * P);
  check(cdm_sync(nullptr), "sample means");
  d_mean.down(mean);
  auto column = [&](const std::vector<float>& src, int stride, int col, int width) {
    std::vector<double> out((size_t)nvox * width);
    for (int i = 0; i < nvox; ++i)
      for (int w = 0; w < width; ++w) out[(size_t)i * width + w] = src[(size_t)i * stride + col + w];
    return out;
  };
  auto save = [&](const std::string& name, const std::vector<double>& vals, int nt) {
    const std::string err = write_masked(dir + "/" + name, mask, vals, nvox, nt);
    if (!err.empty()) die(err);
  };
  for (int k : spec->means) save(std::string("mean_") + spec->names[k], column(mean, P, k, 1), 1);
  if (!spec->fsum.empty()) {
    std::vector<double> f(nvox, 0.0);
    for (int k : spec->fsum)
      for (int i = 0; i < nvox; ++i) f[i] += mean[(size_t)i * P + k];
    save("mean_fsum", f, 1);
  }
  if (spec->tensor) {
    std::vector<double> t((size_t)nvox * 6);
    for (int k = 1; k <= 6; ++k) {
      save("Param_" + std::to_string(k) + "_samples", column(mean, P, k, 1), 1);
      for (int i = 0; i < nvox; ++i) t[(size_t)i * 6 + k - 1] = mean[(size_t)i * P + k];
    }
    save("Tensor", t, 6);
  }
  for (const FinishSpec::Dir& d : spec->dirs) {
    Dev<float> d_dy((size_t)nvox * 3), d_disp(nvox);
    check(cdm_dyads(nvox, P, ns, d_s.p, d.th, d.ph, d.psi, d_dy.p, d.psi < 0 ? d_disp.p : nullptr, nullptr), "dyads");
    std::vector<float> dy((size_t)nvox * 3), disp(nvox);
    check(cdm_sync(nullptr), "dyads");
    d_dy.down(dy);
    save(d.out, column(dy, 3, 0, 3), 3);
    if (d.psi < 0) {
      d_disp.down(disp);
      save(std::string(d.out) + "_dispersion", column(disp, 1, 0, 1), 1);
    }
  }
  for (const FinishSpec::Noddi& nd : spec->noddi) {
    const int w = nd.beta < 0 ? 1 : 5;
    Dev<float> d_maps((size_t)nvox * w);
    check(cdm_noddi_maps(nvox, P, ns, d_s.p, nd.kappa, nd.beta, d_maps.p, nullptr), "NODDI maps");
    std::vector<float> maps((size_t)nvox * w);
    check(cdm_sync(nullptr), "NODDI maps");
    d_maps.down(maps);
    const std::string sfx = nd.sfx;
    if (nd.beta < 0) {
      save("OD" + sfx, column(maps, 1, 0, 1), 1);
      continue;
    }
    std::vector<double> k2((size_t)nvox * ns);  // the k2 = kappa - beta sample volume
    for (int i = 0; i < nvox; ++i)
      for (int s = 0; s < ns; ++s)
        k2[(size_t)i * ns + s] = (float)(samples[((size_t)i * P + nd.kappa) * ns + s] -
                                         samples[((size_t)i * P + nd.beta) * ns + s]);
    save("k2" + sfx + "_samples", k2, ns);
    save("mean_k2" + sfx, column(maps, 5, 0, 1), 1);
    save("ODIp" + sfx, column(maps, 5, 1, 1), 1);
    save("ODIs" + sfx, column(maps, 5, 2, 1), 1);
    save("ODItot" + sfx, column(maps, 5, 3, 1), 1);
    save("DA" + sfx, column(maps, 5, 4, 1), 1);
  }
}

bool starts_with(const std::string& s, const std::string& p) { return s.compare(0, p.size(), p) == 0; }

bool is_model(const std::string& name) {
  char buf[4096];
  const int n = cdm_list_models(buf, sizeof(buf));
  std::stringstream ss(std::string(buf, std::min<int>(n, sizeof(buf) - 1)));
  std::string t;
  while (std::getline(ss, t, ','))
    if (t == name) return true;
  return false;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string prog = fs::path(argv[0]).filename().string();
  std::string step, model;
  int first = 1;
  if (starts_with(prog, "split_parts_")) {
    step = "split";
    model = prog.substr(12);
  } else if (starts_with(prog, "merge_parts_")) {
    step = "merge";
    model = prog.substr(12);
  } else if (is_model(prog)) {
    step = "fit";
    model = prog;
  } else if (prog.size() > 7 && prog.compare(prog.size() - 7, 7, "_finish") == 0) {
    step = "finish";
    model = prog.substr(0, prog.size() - 7);
  } else {
    if (argc < 2 || argv[1][0] == '-') {
      std::cerr << "Usage: " << prog << " split|fit|merge --model=<plug-in> [options]\n"
                << "       " << prog << " finish --model=<plug-in> <output directory>\n\n"
                << usage(prog);
      return 2;
    }
    step = argv[1];
    first = 2;
    if (step != "split" && step != "fit" && step != "merge" && step != "finish") {
      std::cerr << "unknown step '" << step << "' (split, fit, merge or finish)\n";
      return 2;
    }
  }
  if (step == "finish") {  // <Model>_finish.sh <directory>
    std::string dir;
    for (int a = first; a < argc; ++a) {
      const std::string s = argv[a];
      if (starts_with(s, "--model="))
        model = s.substr(8);
      else if (dir.empty() && s[0] != '-')
        dir = s;
      else {
        std::cerr << "Usage: " << prog << " finish --model=<plug-in> <output directory>\n";
        return 2;
      }
    }
    if (dir.empty() || model.empty()) {
      std::cerr << "Usage: " << prog << " finish --model=<plug-in> <output directory>\n"
                << "expects the merged Param_k_samples volumes and nodif_brain_mask in the directory\n";
      return 2;
    }
    finish(model, dir);
    return 0;
  }
  Options o;
  const std::string err = parse_options(argc, argv, first, o);
  if (!err.empty()) {
    std::cerr << usage(prog) << "\n" << err << std::endl;
    return 2;
  }
  if (o.model.empty()) o.model = model;
  if (o.help || !missing_compulsory(o).empty() || o.model.empty()) {
    std::cerr << usage(prog);
    for (const std::string& n : missing_compulsory(o)) std::cerr << "missing compulsory option " << n << "\n";
    if (o.model.empty()) std::cerr << "missing --model\n";
    return 2;
  }
  make_logdir(o, argc, argv);
  if (step == "split")
    split_parts(o);
  else if (step == "fit")
    fit(o);
  else
    merge_parts(o);
  return 0;
}
