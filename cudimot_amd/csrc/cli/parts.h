// Intermediate part files of the reference's split / fit / merge steps
// (split_parts.cc:19-34 writes them, dMRI_Data.cu:39-63 and
// Parameters.cu:52-69,166-181 read them, Parameters.cu:385-506 writes the
// results, merge_parts.cc:21-61 joins them): int32 nvox, int32 nrows,
// int64 byte count, then the nrows x nvox FP64 matrix.  FSL >= 6.0.6
// (armawrap, the reference Makefile's FSL_GE_606 path) stores it
// column-major: each voxel's nrows values are contiguous.  FSL <= 6.0.5
// (NEWMAT) stores it row-major; `row_major` selects that order.
#pragma once
#include <string>
#include <vector>

namespace cdm_cli {

struct Part {
  int nvox = 0, nrows = 0;
  std::vector<double> v;  // [nvox][nrows]: voxel-major whatever the file order
};

std::string read_part(const std::string& path, Part& p, bool row_major = false);
std::string write_part(const std::string& path, const Part& p, bool row_major = false);
// Header only (nvox, nrows); "" or an error.
std::string read_part_header(const std::string& path, int& nvox, int& nrows);

}  // namespace cdm_cli
