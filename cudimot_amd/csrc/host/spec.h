// Model specification parsed from text files, with the reference's rules.
//
//   modelpriors   Model.cu:16-299   P_init[...], bounds[i]=(min,max|min,|,max),
//                                   prior[i]=Gaussian(a,b)|Gamma(a,b)|ARD(a)|sin()|custom()
//   --fixed       Model.cu:279-298  comma-separated parameter ids
//   grid file     Model.cu:302-393  search[i]=(v0,v1,...); combinations are the
//                                   cartesian product with the lowest grid
//                                   parameter outermost
// Reference behaviours kept on purpose: a line containing "//" anywhere is
// skipped; only "prior[" is recognised ("priors[" lines are ignored); numbers
// are read with strtod; errors are reported (here as a status and message,
// where the reference prints and exits).
#pragma once
#include <string>
#include <vector>

namespace cdm {

enum BoundType { NOBOUNDS = 0, BMIN = 1, BMAX = 2, BMINMAX = 3 };
enum PriorType { NOPRIOR = 0, GAUSSPRIOR = 1, GAMMAPRIOR = 2, ARDPRIOR = 3, SINPRIOR = 4, CUSTOM = 5 };

struct ModelSpec {
  int nparams = 0;
  bool init_provided = false;
  std::vector<double> p_init;
  std::vector<int> bound_types;
  std::vector<double> bmin, bmax;
  std::vector<int> prior_types;
  std::vector<double> pa, pb;
  std::vector<int> fixed;

  // Grid search
  std::vector<int> grid_params;  // ascending parameter ids
  int grid_combs = 0;
  std::vector<double> grid;      // [grid_combs][grid_params.size()]
};

// Each returns an empty string on success, else an error message.
std::string parse_priors_text(const std::string& text, int nparams, ModelSpec& out);
std::string parse_priors_file(const std::string& path, int nparams, ModelSpec& out);
std::string parse_fixed(const std::string& list, ModelSpec& spec);
std::string parse_grid_text(const std::string& text, ModelSpec& spec);
std::string parse_grid_file(const std::string& path, ModelSpec& spec);

}  // namespace cdm
