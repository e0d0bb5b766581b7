// Command-line options of the reference's three programs (split_parts_X,
// X, merge_parts_X), with the reference's names, aliases and defaults
// (cudimotoptions.h:78-168).  All three parse the same set; --data,
// --maskfile, --partsdir, --outputdir, --idPart and --nParts are compulsory
// (cudimotoptions.cc:31-35 prints the usage and exits with status 2 when
// one is missing).
#pragma once
#include <string>
#include <vector>

namespace cdm_cli {

struct Options {
  bool help = false;
  std::string logdir = "logdir";
  bool forcedir = false;
  std::string datafile = "data";
  std::string maskfile = "nodif_brain_mask";
  std::string partsdir, outputdir;
  int idPart = 0, nParts = 0;
  int njumps = 1250, nburn = 1000, sampleevery = 25, updateproposalevery = 40;
  bool no_updateproposal = false;
  int seed = 8219;
  bool no_LevMar = false, no_Marquardt = false;
  int iterLevMar = 200;
  std::string gridSearch;
  bool runMCMC = false, rician = false, keepTmp = false, getPredictedSignal = false;
  std::string CFP, FixP, fixed, init_params, debug;
  bool BIC_AIC = false;
  std::string priorsfile;
  // Engine additions (no reference counterpart).
  std::string model;         // --model: plug-in name when the program name does not carry it
  bool part_rowmajor = false;  // --part_rowmajor: part files in FSL <= 6.0.5 (NEWMAT) order
  std::vector<std::string> given;  // long names of the options on the command line
  bool set(const std::string& long_name) const;
};

// Parse argv[first..]; returns "" or an error message.
std::string parse_options(int argc, char** argv, int first, Options& o);
// Names of compulsory options missing from the command line.
std::vector<std::string> missing_compulsory(const Options& o);
std::string usage(const std::string& program);

}  // namespace cdm_cli
