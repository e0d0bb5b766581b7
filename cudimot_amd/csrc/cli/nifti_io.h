// NIfTI-1 single-file volumes (.nii, .nii.gz) for the host driver: the
// subset of FSL's newimage that the reference's split/merge steps use
// (read_volume4D, volume4D::matrix(mask), setmatrix(mask), save_volume4D;
// split_parts.cc:48-54, merge_parts.cc:55-60).
#pragma once
#include <string>
#include <vector>

namespace cdm_cli {

struct Volume {
  unsigned char hdr[348];  // native-endian NIfTI-1 header (geometry kept for outputs)
  int nx = 0, ny = 0, nz = 0, nt = 0;
  std::vector<float> data;  // x fastest, then y, z, t; scaled by scl_slope/inter
  size_t nvox3() const { return (size_t)nx * ny * nz; }
};

// Read `name` (with or without .nii / .nii.gz).  Returns "" or an error.
std::string read_volume(const std::string& name, Volume& v);

// Resolve the file a read of `name` would open ("" if none).
std::string find_volume(const std::string& name);

// Rows = volumes (t), columns = in-mask voxels in x-fastest order
// (newimage's volume4D::matrix(mask)); out is column-major [nvox][nt].
// A voxel is in the mask when its value is > 0.
std::string masked_matrix(const Volume& v, const Volume& mask, std::vector<double>& out, int& nvox);

// Write a FLOAT32 4-D volume with mask's geometry: values [nvox][nt]
// (column-major, one column per in-mask voxel) scattered into the mask,
// zero outside; cal_min/cal_max = the volume's min/max (merge_parts.cc:57-59).
// The file type follows $FSLOUTPUTTYPE (NIFTI -> .nii, else .nii.gz).
std::string write_masked(const std::string& name, const Volume& mask, const std::vector<double>& values, int nvox,
                         int nt, std::string* path_out = nullptr);

// Write a FLOAT32 volume of the given size with geometry from `like`.
std::string write_volume(const std::string& name, const Volume& like, int nx, int ny, int nz, int nt,
                         const std::vector<float>& data, std::string* path_out = nullptr);

}  // namespace cdm_cli
