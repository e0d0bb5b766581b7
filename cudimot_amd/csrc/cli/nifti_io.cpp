// NIfTI-1 reader/writer on zlib (gzread reads plain and gzip files alike).
// Header offsets are those of the NIfTI-1 standard (nifti1.h, 348-byte
// header, single-file "n+1" magic).
#include "cli/nifti_io.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>

namespace cdm_cli {
namespace {

// Byte offsets of the NIfTI-1 header fields used here.
enum : int {
  kSizeofHdr = 0,
  kDim = 40,  // short[8]
  kDatatype = 70,
  kBitpix = 72,
  kPixdim = 76,  // float[8]
  kVoxOffset = 108,
  kSclSlope = 112,
  kSclInter = 116,
  kCalMax = 124,
  kCalMin = 128,
  kMagic = 344,
};

template <typename T>
T get(const unsigned char* h, int off) {
  T v;
  std::memcpy(&v, h + off, sizeof(T));
  return v;
}
template <typename T>
void put(unsigned char* h, int off, T v) {
  std::memcpy(h + off, &v, sizeof(T));
}

void swap_bytes(unsigned char* p, int n) { std::reverse(p, p + n); }

// Every multi-byte field of the header as (offset, width, count), for
// reading files written on the other endianness.
void swap_header(unsigned char* h) {
  static const int fields[][3] = {
      {0, 4, 1},    {32, 4, 1},   {36, 2, 1},  {40, 2, 8},  {56, 4, 3},  {68, 2, 4},  {76, 4, 8},
      {108, 4, 3},  {120, 2, 1},  {124, 4, 4}, {140, 4, 2}, {252, 2, 2}, {256, 4, 6}, {280, 4, 12},
  };
  for (const auto& f : fields)
    for (int k = 0; k < f[2]; ++k) swap_bytes(h + f[0] + k * f[1], f[1]);
}

bool ends_with(const std::string& s, const std::string& t) {
  return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

bool exists(const std::string& p) { return std::ifstream(p).good(); }

std::string output_path(const std::string& name) {
  if (ends_with(name, ".nii") || ends_with(name, ".nii.gz")) return name;
  const char* t = std::getenv("FSLOUTPUTTYPE");
  if (t && std::string(t) == "NIFTI") return name + ".nii";
  return name + ".nii.gz";
}

std::string write_file(const std::string& path, const unsigned char* hdr, const std::vector<float>& data) {
  unsigned char ext[4] = {0, 0, 0, 0};
  if (ends_with(path, ".gz")) {
    gzFile f = gzopen(path.c_str(), "wb6");
    if (!f) return "cannot write " + path;
    bool ok = gzwrite(f, hdr, 348) == 348 && gzwrite(f, ext, 4) == 4;
    const char* p = (const char*)data.data();
    size_t left = data.size() * sizeof(float);
    while (ok && left) {
      const unsigned chunk = (unsigned)std::min<size_t>(left, 1u << 30);
      ok = gzwrite(f, p, chunk) == (int)chunk;
      p += chunk;
      left -= chunk;
    }
    if (gzclose(f) != Z_OK || !ok) return "error writing " + path;
    return "";
  }
  std::ofstream o(path, std::ios::binary);
  if (!o) return "cannot write " + path;
  o.write((const char*)hdr, 348);
  o.write((const char*)ext, 4);
  o.write((const char*)data.data(), (std::streamsize)(data.size() * sizeof(float)));
  return o ? "" : "error writing " + path;
}

}  // namespace

std::string find_volume(const std::string& name) {
  if (ends_with(name, ".nii") || ends_with(name, ".nii.gz")) return exists(name) ? name : "";
  for (const char* e : {".nii.gz", ".nii"})
    if (exists(name + e)) return name + e;
  return "";
}

std::string read_volume(const std::string& name, Volume& v) {
  const std::string path = find_volume(name);
  if (path.empty()) return "cannot find NIfTI volume " + name + " (.nii or .nii.gz)";
  gzFile f = gzopen(path.c_str(), "rb");
  if (!f) return "cannot open " + path;
  std::vector<unsigned char> buf;
  {
    unsigned char tmp[1 << 16];
    int n;
    while ((n = gzread(f, tmp, sizeof(tmp))) > 0) buf.insert(buf.end(), tmp, tmp + n);
    const bool bad = n < 0;
    gzclose(f);
    if (bad) return "error reading " + path;
  }
  if (buf.size() < 352) return path + ": too short for a NIfTI-1 file";
  std::memcpy(v.hdr, buf.data(), 348);
  bool swapped = false;
  if (get<int32_t>(v.hdr, kSizeofHdr) != 348) {
    swap_header(v.hdr);
    swapped = true;
    if (get<int32_t>(v.hdr, kSizeofHdr) != 348) return path + ": not a NIfTI-1 header";
  }
  if (std::memcmp(v.hdr + kMagic, "n+1", 4) != 0) return path + ": not a single-file NIfTI-1 volume (magic n+1)";
  const int ndim = get<int16_t>(v.hdr, kDim);
  if (ndim < 1 || ndim > 7) return path + ": bad dim[0]";
  int d[8];
  for (int i = 0; i < 8; ++i) d[i] = get<int16_t>(v.hdr, kDim + 2 * i);
  for (int i = 1; i <= ndim; ++i)
    if (d[i] < 1) return path + ": bad dimension";
  for (int i = ndim + 1; i < 8; ++i) d[i] = 1;
  for (int i = 5; i <= 7; ++i)
    if (d[i] != 1) return path + ": volumes with more than 4 dimensions are not supported";
  v.nx = d[1];
  v.ny = d[2];
  v.nz = d[3];
  v.nt = d[4];
  const int dtype = get<int16_t>(v.hdr, kDatatype);
  int width = 0;
  switch (dtype) {
    case 2: case 256: width = 1; break;     // uint8, int8
    case 4: case 512: width = 2; break;     // int16, uint16
    case 8: case 768: case 16: width = 4; break;  // int32, uint32, float32
    case 64: case 1024: case 1280: width = 8; break;  // float64, int64, uint64
    default: return path + ": unsupported NIfTI datatype " + std::to_string(dtype);
  }
  const size_t n = v.nvox3() * (size_t)v.nt;
  const size_t off = (size_t)get<float>(v.hdr, kVoxOffset);
  if (off < 348 || off + n * width > buf.size()) return path + ": data shorter than its header says";
  float slope = get<float>(v.hdr, kSclSlope), inter = get<float>(v.hdr, kSclInter);
  const bool scale = slope != 0.0f && std::isfinite(slope) && !(slope == 1.0f && inter == 0.0f);
  if (!std::isfinite(inter)) inter = 0.0f;
  v.data.resize(n);
  unsigned char* p = buf.data() + off;
  for (size_t i = 0; i < n; ++i, p += width) {
    if (swapped) swap_bytes(p, width);
    double x;
    switch (dtype) {
      case 2: x = *(uint8_t*)p; break;
      case 256: x = *(int8_t*)p; break;
      case 4: x = get<int16_t>(p, 0); break;
      case 512: x = get<uint16_t>(p, 0); break;
      case 8: x = get<int32_t>(p, 0); break;
      case 768: x = get<uint32_t>(p, 0); break;
      case 16: x = get<float>(p, 0); break;
      case 64: x = get<double>(p, 0); break;
      case 1024: x = (double)get<int64_t>(p, 0); break;
      default: x = (double)get<uint64_t>(p, 0); break;
    }
    v.data[i] = scale ? (float)(x * slope + inter) : (float)x;
  }
  return "";
}

std::string masked_matrix(const Volume& v, const Volume& mask, std::vector<double>& out, int& nvox) {
  if (v.nx != mask.nx || v.ny != mask.ny || v.nz != mask.nz)
    return "volume and mask sizes differ (" + std::to_string(v.nx) + "x" + std::to_string(v.ny) + "x" +
           std::to_string(v.nz) + " vs " + std::to_string(mask.nx) + "x" + std::to_string(mask.ny) + "x" +
           std::to_string(mask.nz) + ")";
  const size_t n3 = v.nvox3();
  nvox = 0;
  for (size_t i = 0; i < n3; ++i) nvox += mask.data[i] > 0;
  out.assign((size_t)nvox * v.nt, 0.0);
  size_t c = 0;
  for (size_t i = 0; i < n3; ++i) {
    if (!(mask.data[i] > 0)) continue;
    for (int t = 0; t < v.nt; ++t) out[c * v.nt + t] = (double)v.data[(size_t)t * n3 + i];
    ++c;
  }
  return "";
}

std::string write_volume(const std::string& name, const Volume& like, int nx, int ny, int nz, int nt,
                         const std::vector<float>& data, std::string* path_out) {
  unsigned char h[348];
  std::memcpy(h, like.hdr, 348);
  const int16_t dims[8] = {(int16_t)(nt > 1 ? 4 : 3), (int16_t)nx, (int16_t)ny, (int16_t)nz, (int16_t)nt, 1, 1, 1};
  for (int i = 0; i < 8; ++i) put<int16_t>(h, kDim + 2 * i, dims[i]);
  put<int16_t>(h, kDatatype, 16);
  put<int16_t>(h, kBitpix, 32);
  if (nt > 1 && !(get<float>(h, kPixdim + 16) > 0)) put<float>(h, kPixdim + 16, 1.0f);
  put<float>(h, kVoxOffset, 352.0f);
  put<float>(h, kSclSlope, 1.0f);
  put<float>(h, kSclInter, 0.0f);
  float mn = 0, mx = 0;
  if (!data.empty()) {
    mn = mx = data[0];
    for (float x : data) {
      mn = std::min(mn, x);
      mx = std::max(mx, x);
    }
  }
  put<float>(h, kCalMax, mx);
  put<float>(h, kCalMin, mn);
  std::memcpy(h + kMagic, "n+1", 4);
  const std::string path = output_path(name);
  if (path_out) *path_out = path;
  return write_file(path, h, data);
}

std::string write_masked(const std::string& name, const Volume& mask, const std::vector<double>& values, int nvox,
                         int nt, std::string* path_out) {
  const size_t n3 = mask.nvox3();
  size_t inmask = 0;
  for (size_t i = 0; i < n3; ++i) inmask += mask.data[i] > 0;
  if ((size_t)nvox != inmask)
    return "mask has " + std::to_string(inmask) + " voxels, results have " + std::to_string(nvox);
  std::vector<float> vol(n3 * (size_t)nt, 0.0f);
  size_t c = 0;
  for (size_t i = 0; i < n3; ++i) {
    if (!(mask.data[i] > 0)) continue;
    for (int t = 0; t < nt; ++t) vol[(size_t)t * n3 + i] = (float)values[c * nt + t];
    ++c;
  }
  return write_volume(name, mask, mask.nx, mask.ny, mask.nz, nt, vol, path_out);
}

}  // namespace cdm_cli
