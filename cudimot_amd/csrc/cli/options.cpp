#include "cli/options.h"

#include <algorithm>
#include <cstdlib>
#include <sstream>

namespace cdm_cli {
namespace {

enum Kind { kBool, kInt, kStr };

struct Spec {
  const char* names;  // comma-separated aliases, long name last
  Kind kind;
  bool compulsory;
  const char* help;
};

// Order and wording follow cudimotoptions.h:78-168.
const Spec kSpecs[] = {
    {"-h,--help", kBool, false, "Display this message"},
    {"--ld,--logdir", kStr, false, "Log directory (default is logdir)"},
    {"--forcedir", kBool, false, "Use the actual directory name given - i.e. don't add + to make a new directory"},
    {"--data", kStr, true, "Data file"},
    {"--maskfile", kStr, true, "Mask file"},
    {"--partsdir", kStr, true, "Directory where different parts of the data/results will be stored"},
    {"--outputdir", kStr, true, "Directory where to write the output files"},
    {"--idPart", kInt, true, "Number of the part of the dataset to process [0..N-1]"},
    {"--nParts", kInt, true, "Total number of parts of the dataset [1..N]"},
    {"--nj,--njumps", kInt, false, "Num of jumps to be made by MCMC (default is 1250)"},
    {"--bi,--burnin", kInt, false, "Total num of jumps at start of MCMC to be discarded (default is 1000)"},
    {"--se,--sampleevery", kInt, false, "Num of jumps for each sample (MCMC) (default is 25)"},
    {"--upe,--updateproposalevery", kInt, false,
     "Num of jumps for each update to the proposal density std (MCMC) (default is 40; proposals adapt every --se "
     "jumps, as in the reference)"},
    {"--no_updateproposal", kBool, false, "Do not update the proposal density std during the recording step of MCMC"},
    {"--seed", kInt, false, "Seed for pseudo random number generator"},
    {"--gridSearch", kStr, false, "Run gridSearch algorithm using the values specified in a file"},
    {"--runMCMC", kBool, false, "Run MCMC algorithm and get distribution of estimates"},
    {"--no_LevMar", kBool, false, "Do not run Levenberg-Marquardt algorithm"},
    {"--no_Marquardt", kBool, false, "Do not use Marquardt contribution in Levenberg algorithm"},
    {"--iterLevMar", kInt, false, "Maximum number of iterations in Levenberg(-Marquardt) algorithm (default is 200)"},
    {"--rician", kBool, false, "Use Rician noise modelling in MCMC"},
    {"--keepTmp", kBool, false, "Do not remove the temporal directory created for storing the data/results parts"},
    {"--getPredictedSignal", kBool, false, "Save the predicted signal by the model at the end"},
    {"--CFP", kStr, false,
     "File with a list of ASCII files for specifying the common (to all voxels) fixed parameters of the model"},
    {"--FixP", kStr, false, "File with a list of NIfTI files for specifying the fixed parameters of the model"},
    {"--fixed", kStr, false, "List of the Ids of the fixed parameters separated by commas. For example: --fixed=2,4"},
    {"--init_params", kStr, false, "File with a list of NIfTI files for the initialization of the model parameters"},
    {"--debug", kStr, false,
     "Specify a voxel for debugging. Some variables at certain steps of the algorithms will be printed (use few "
     "iterations)"},
    {"--BIC_AIC", kBool, false, "Calculate Bayesian and Akaike Information Criteria at the end"},
    {"--priors", kStr, false, "File with parameters information (initialization, bounds and priors)"},
    {"--model", kStr, false, "Model plug-in (when the program name does not carry it)"},
    {"--part_rowmajor", kBool, false, "Part files in FSL <= 6.0.5 (NEWMAT, row-major) order"},
};

std::vector<std::string> split(const std::string& s, char c) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string t;
  while (std::getline(ss, t, c)) out.push_back(t);
  return out;
}

std::string long_name(const Spec& s) { return split(s.names, ',').back(); }

void* target(Options& o, const std::string& n) {
  if (n == "--help") return &o.help;
  if (n == "--logdir") return &o.logdir;
  if (n == "--forcedir") return &o.forcedir;
  if (n == "--data") return &o.datafile;
  if (n == "--maskfile") return &o.maskfile;
  if (n == "--partsdir") return &o.partsdir;
  if (n == "--outputdir") return &o.outputdir;
  if (n == "--idPart") return &o.idPart;
  if (n == "--nParts") return &o.nParts;
  if (n == "--njumps") return &o.njumps;
  if (n == "--burnin") return &o.nburn;
  if (n == "--sampleevery") return &o.sampleevery;
  if (n == "--updateproposalevery") return &o.updateproposalevery;
  if (n == "--no_updateproposal") return &o.no_updateproposal;
  if (n == "--seed") return &o.seed;
  if (n == "--gridSearch") return &o.gridSearch;
  if (n == "--runMCMC") return &o.runMCMC;
  if (n == "--no_LevMar") return &o.no_LevMar;
  if (n == "--no_Marquardt") return &o.no_Marquardt;
  if (n == "--iterLevMar") return &o.iterLevMar;
  if (n == "--rician") return &o.rician;
  if (n == "--keepTmp") return &o.keepTmp;
  if (n == "--getPredictedSignal") return &o.getPredictedSignal;
  if (n == "--CFP") return &o.CFP;
  if (n == "--FixP") return &o.FixP;
  if (n == "--fixed") return &o.fixed;
  if (n == "--init_params") return &o.init_params;
  if (n == "--debug") return &o.debug;
  if (n == "--BIC_AIC") return &o.BIC_AIC;
  if (n == "--priors") return &o.priorsfile;
  if (n == "--model") return &o.model;
  if (n == "--part_rowmajor") return &o.part_rowmajor;
  return nullptr;
}

const Spec* find(const std::string& name) {
  for (const Spec& s : kSpecs)
    for (const std::string& a : split(s.names, ','))
      if (a == name) return &s;
  return nullptr;
}

}  // namespace

bool Options::set(const std::string& n) const { return std::find(given.begin(), given.end(), n) != given.end(); }

std::string parse_options(int argc, char** argv, int first, Options& o) {
  for (int i = first; i < argc; ++i) {
    std::string a = argv[i];
    if (a.size() < 2 || a[0] != '-') return "unexpected argument: " + a;
    std::string name = a, value;
    bool has_value = false;
    const size_t eq = a.find('=');
    if (eq != std::string::npos) {
      name = a.substr(0, eq);
      value = a.substr(eq + 1);
      has_value = true;
    }
    const Spec* s = find(name);
    if (!s) return "unknown option: " + name;
    const std::string ln = long_name(*s);
    void* t = target(o, ln);
    if (s->kind == kBool) {
      if (has_value) return "option " + name + " takes no value";
      *(bool*)t = true;
    } else {
      if (!has_value) {
        if (i + 1 >= argc) return "option " + name + " requires a value";
        value = argv[++i];
      }
      if (s->kind == kInt) {
        char* end = nullptr;
        const long v = std::strtol(value.c_str(), &end, 10);
        if (value.empty() || *end) return "option " + name + " needs an integer, got '" + value + "'";
        *(int*)t = (int)v;
      } else {
        *(std::string*)t = value;
      }
    }
    o.given.push_back(ln);
  }
  return "";
}

std::vector<std::string> missing_compulsory(const Options& o) {
  std::vector<std::string> out;
  for (const Spec& s : kSpecs)
    if (s.compulsory && !o.set(long_name(s))) out.push_back(long_name(s));
  return out;
}

std::string usage(const std::string& program) {
  std::ostringstream u;
  u << "Usage: " << program << " [options]\n\nCompulsory arguments (You MUST set one or more of):\n";
  for (int pass = 0; pass < 2; ++pass) {
    if (pass) u << "\nOptional arguments (You may optionally specify one or more of):\n";
    for (const Spec& s : kSpecs)
      if (s.compulsory == (pass == 0)) u << "\t" << s.names << "\t" << s.help << "\n";
  }
  return u.str();
}

}  // namespace cdm_cli
