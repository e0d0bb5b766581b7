#include "cli/parts.h"

#include <cstdint>
#include <fstream>

namespace cdm_cli {

std::string read_part_header(const std::string& path, int& nvox, int& nrows) {
  std::ifstream in(path, std::ios::binary);
  if (!in) return "cannot open part file " + path;
  int32_t h[2];
  int64_t nbytes;
  in.read((char*)h, 8);
  in.read((char*)&nbytes, 8);
  if (!in) return path + ": truncated part header";
  nvox = h[0];
  nrows = h[1];
  if (nvox <= 0 || nrows <= 0) return path + ": the numbers of voxels and rows must be greater than 0";
  if (nbytes != (int64_t)nvox * nrows * 8)
    return path + ": byte count " + std::to_string(nbytes) + " does not match " + std::to_string(nvox) + " x " +
           std::to_string(nrows) + " doubles";
  return "";
}

std::string read_part(const std::string& path, Part& p, bool row_major) {
  std::string e = read_part_header(path, p.nvox, p.nrows);
  if (!e.empty()) return e;
  std::ifstream in(path, std::ios::binary);
  in.seekg(16);
  const size_t n = (size_t)p.nvox * p.nrows;
  std::vector<double> raw(n);
  in.read((char*)raw.data(), (std::streamsize)(n * 8));
  if (!in) return path + ": data shorter than its header says";
  if (!row_major) {
    p.v.swap(raw);
  } else {
    p.v.resize(n);
    for (int r = 0; r < p.nrows; ++r)
      for (int c = 0; c < p.nvox; ++c) p.v[(size_t)c * p.nrows + r] = raw[(size_t)r * p.nvox + c];
  }
  return "";
}

std::string write_part(const std::string& path, const Part& p, bool row_major) {
  std::ofstream out(path, std::ios::binary);
  if (!out) return "cannot write part file " + path;
  const int32_t h[2] = {p.nvox, p.nrows};
  const int64_t nbytes = (int64_t)p.nvox * p.nrows * 8;
  out.write((const char*)h, 8);
  out.write((const char*)&nbytes, 8);
  if (!row_major) {
    out.write((const char*)p.v.data(), (std::streamsize)nbytes);
  } else {
    std::vector<double> raw(p.v.size());
    for (int r = 0; r < p.nrows; ++r)
      for (int c = 0; c < p.nvox; ++c) raw[(size_t)r * p.nvox + c] = p.v[(size_t)c * p.nrows + r];
    out.write((const char*)raw.data(), (std::streamsize)nbytes);
  }
  return out ? "" : "error writing " + path;
}

}  // namespace cdm_cli
