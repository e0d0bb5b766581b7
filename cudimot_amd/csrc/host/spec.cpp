#include "spec.h"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace cdm {

namespace {

// string::find / rfind as the reference stores them: in an int, npos -> -1.
int ifind(const std::string& s, const std::string& t) { return (int)s.find(t); }
int irfind(const std::string& s, const std::string& t) { return (int)s.rfind(t); }

// substr with an int length, where a negative length means "to the end"
// (the size_t conversion the reference relies on).
std::string isub(const std::string& s, int pos, int len) {
  if (pos < 0 || pos > (int)s.size()) return std::string();
  if (len < 0) return s.substr((size_t)pos);
  return s.substr((size_t)pos, (size_t)len);
}

double num(const std::string& s) { return strtod(s.c_str(), nullptr); }

bool read_file(const std::string& path, std::string& out) {
  std::ifstream f(path.c_str());
  if (!f.is_open()) return false;
  std::stringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

}  // namespace

std::string parse_priors_text(const std::string& text, int nparams, ModelSpec& out) {
  out = ModelSpec();
  out.nparams = nparams;
  out.bound_types.assign(nparams, NOBOUNDS);
  out.bmin.assign(nparams, 0.0);
  out.bmax.assign(nparams, 0.0);
  out.prior_types.assign(nparams, NOPRIOR);
  out.pa.assign(nparams, 0.0);
  out.pb.assign(nparams, 0.0);
  out.fixed.assign(nparams, 0);
  if (nparams <= 0) return "Number of parameters must be greater than 0";

  std::istringstream in(text);
  std::string line;
  const std::string mark_init = "P_init[", mark_bound = "bounds[", mark_prior = "prior[";
  while (std::getline(in, line)) {
    if (ifind(line, "//") >= 0) continue;  // Model.cu:71-74
    int pos0, pos1;
    if ((pos0 = ifind(line, mark_init)) >= 0) {
      out.init_provided = true;
      pos0 += (int)mark_init.size();
      pos1 = irfind(line, "]");
      std::string vals = isub(line, pos0, pos1 - pos0);
      int c;
      while ((c = ifind(vals, ",")) >= 0) {
        out.p_init.push_back(num(vals.substr(0, (size_t)c)));
        vals.erase(0, (size_t)c + 1);
      }
      out.p_init.push_back(num(vals));
      if ((int)out.p_init.size() != nparams) {
        std::ostringstream e;
        e << "The number of values for initializating the parameters " << out.p_init.size()
          << " do not match the number of parameters of this model " << nparams;
        return e.str();
      }
    }
    if ((pos0 = ifind(line, mark_bound)) >= 0) {
      pos0 += (int)mark_bound.size();
      pos1 = irfind(line, "]");
      const int id = atoi(isub(line, pos0, pos1 - pos0).c_str());
      if (id >= nparams || id < 0) {
        std::ostringstream e;
        e << "Wrong number of parameter [" << id << "] when specifying the bounds";
        return e.str();
      }
      out.bound_types[id] = BMINMAX;
      pos1 = irfind(line, "(");
      pos0 = pos1 + 1;
      pos1 = irfind(line, ",");
      int len = pos1 - pos0;
      if (len == 0) out.bound_types[id] = BMAX;
      else out.bmin[id] = num(isub(line, pos0, len));
      pos0 = pos1 + 1;
      pos1 = irfind(line, ")");
      len = pos1 - pos0;
      if (len == 0 && out.bound_types[id] == BMINMAX) {
        out.bound_types[id] = BMIN;
      } else if (len == 0) {
        return "For bounded priors, at least a min or a max value must be specified: bounds[x]=(min,), "
               "bounds[x]=(,max) or bounds[x]=(min,max)";
      } else {
        out.bmax[id] = num(isub(line, pos0, len));
      }
    }
    if ((pos0 = ifind(line, mark_prior)) >= 0) {
      pos0 += (int)mark_prior.size();
      pos1 = irfind(line, "]");
      const int id = atoi(isub(line, pos0, pos1 - pos0).c_str());
      if (id >= nparams || id < 0) {
        std::ostringstream e;
        e << "Wrong number of parameter [" << id << "] when specifying the priors";
        return e.str();
      }
      pos0 = irfind(line, "=") + 1;
      pos1 = irfind(line, "(");
      const std::string type = isub(line, pos0, pos1 - pos0);
      int pt;
      if (type == "Gaussian") pt = GAUSSPRIOR;
      else if (type == "Gamma") pt = GAMMAPRIOR;
      else if (type == "ARD") pt = ARDPRIOR;
      else if (type == "sin") pt = SINPRIOR;
      else if (type == "custom") pt = CUSTOM;
      else return "Prior " + type + " is a not a recognised prior type";
      out.prior_types[id] = pt;
      if (pt <= ARDPRIOR) {
        pos0 = pos1 + 1;
        pos1 = irfind(line, ",");
        int len = pos1 - pos0;
        if (len == 0) return "A prior of type " + type + " needs a first argument";
        out.pa[id] = num(isub(line, pos0, len));
        if (pt <= GAMMAPRIOR) {
          pos0 = pos1 + 1;
          pos1 = irfind(line, ")");
          len = pos1 - pos0;
          if (len == 0) return "A prior of type " + type + " needs a second argument";
          out.pb[id] = num(isub(line, pos0, len));
        }
      }
    }
  }
  return std::string();
}

std::string parse_priors_file(const std::string& path, int nparams, ModelSpec& out) {
  std::string text;
  if (!read_file(path, text)) return "Unable to open file with Priors information: " + path;
  return parse_priors_text(text, nparams, out);
}

std::string parse_fixed(const std::string& list, ModelSpec& spec) {
  spec.fixed.assign(spec.nparams, 0);
  std::stringstream ss(list);
  std::vector<int> ids;
  int i;
  while (ss >> i) {
    if (i >= spec.nparams || i < 0) {
      std::ostringstream e;
      e << "Wrong number of parameter " << i << " when specifying the fixed parameters";
      return e.str();
    }
    ids.push_back(i);
    if (ss.peek() == ',') ss.ignore();
  }
  for (int id : ids) spec.fixed[id] = 1;
  return std::string();
}

namespace {
void set_grid(int level, const std::vector<std::vector<double>>& vals, const std::vector<int>& gp,
              std::vector<double>& comb, std::vector<double>& grid) {
  const std::vector<double>& v = vals[gp[level]];
  for (size_t i = 0; i < v.size(); ++i) {
    comb[level] = v[i];
    if (level == (int)gp.size() - 1) grid.insert(grid.end(), comb.begin(), comb.end());
    else set_grid(level + 1, vals, gp, comb, grid);
  }
}
}  // namespace

std::string parse_grid_text(const std::string& text, ModelSpec& spec) {
  std::vector<std::vector<double>> vals(spec.nparams);
  std::istringstream in(text);
  std::string line;
  const std::string mark = "search[";
  while (std::getline(in, line)) {
    if (ifind(line, "//") >= 0) continue;
    int pos0, pos1;
    if ((pos0 = ifind(line, mark)) >= 0) {
      pos0 += (int)mark.size();
      pos1 = irfind(line, "]");
      const int id = atoi(isub(line, pos0, pos1 - pos0).c_str());
      if (id >= spec.nparams || id < 0) {
        std::ostringstream e;
        e << "Wrong number of parameter [" << id << "] when specifying the GridSearch values";
        return e.str();
      }
      pos1 = irfind(line, "(");
      pos0 = pos1 + 1;
      pos1 = irfind(line, ")");
      std::string v = isub(line, pos0, pos1 - pos0);
      int c;
      while ((c = ifind(v, ",")) >= 0) {
        vals[id].push_back(num(v.substr(0, (size_t)c)));
        v.erase(0, (size_t)c + 1);
      }
      vals[id].push_back(num(v));
    }
  }
  spec.grid_params.clear();
  spec.grid.clear();
  spec.grid_combs = 1;
  for (int i = 0; i < spec.nparams; ++i)
    if (!vals[i].empty()) {
      spec.grid_params.push_back(i);
      spec.grid_combs *= (int)vals[i].size();
    }
  if (spec.grid_params.empty()) {
    spec.grid_combs = 0;
    return std::string();
  }
  std::vector<double> comb(spec.grid_params.size());
  spec.grid.reserve((size_t)spec.grid_combs * spec.grid_params.size());
  set_grid(0, vals, spec.grid_params, comb, spec.grid);
  return std::string();
}

std::string parse_grid_file(const std::string& path, ModelSpec& spec) {
  std::string text;
  if (!read_file(path, text)) return "Unable to open GridSearch file configuration: " + path;
  return parse_grid_text(text, spec);
}

}  // namespace cdm
