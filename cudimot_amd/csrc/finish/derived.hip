// Derived maps of the reference's <Model>_finish.sh steps, computed from
// the MCMC samples [nvox][nparams][nsamples] where they lie in device
// memory: per-parameter sample means (fslmaths -Tmean), mean fibre
// orientations from the mean dyadic tensor (FSL make_dyadic_vectors; the
// reference's utils/getFanningOrientation.cc:39-76 for Bingham fanning
// directions) and the NODDI dispersion indices
// (NODDI_Watson_finish.sh:35, NODDI_Bingham_finish.sh:46-56).
//
// All three read each sample once and write a few values per voxel: HBM
// bound, one thread per output row.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "../../../include/cudimot_amd.h"
#include "../host/model_ops.h"

namespace {

constexpr int kBlock = 256;
constexpr float kTwoDivPi = 0.636619772367581f;  // the finish scripts' Two_div_pi

__global__ __launch_bounds__(kBlock) void sample_mean_kernel(long long rows, int ns, const float* __restrict__ x,
                                                             float* __restrict__ out) {
  const long long r = (long long)blockIdx.x * kBlock + threadIdx.x;
  if (r >= rows) return;
  const float* p = x + r * ns;
  double s = 0.0;
  for (int k = 0; k < ns; ++k) s += p[k];
  out[r] = (float)(s / ns);
}

// Cyclic Jacobi rotations of a symmetric 3x3 matrix a (upper triangle
// a00 a01 a02 a11 a12 a22); v receives the eigenvectors as columns.
__device__ void jacobi3(double a[3][3], double v[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i][j] = i == j;
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = fabs(a[0][1]) + fabs(a[0][2]) + fabs(a[1][2]);
    const double diag = fabs(a[0][0]) + fabs(a[1][1]) + fabs(a[2][2]);
    if (off <= 1e-15 * diag || off == 0.0) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0.0) continue;
        // rotation angle that zeroes a[p][q]: tan(2t) = 2 a_pq / (a_qq - a_pp)
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // columns p, q of A
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // rows p, q
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void dyads_kernel(int n, int np, int ns, const float* __restrict__ x, int th,
                                                       int ph, int psi, float* __restrict__ dyads,
                                                       float* __restrict__ disp) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n) return;
  const float* base = x + (size_t)v * np * ns;
  double m[3][3] = {};
  for (int k = 0; k < ns; ++k) {
    const float t = base[th * ns + k], f = base[ph * ns + k];
    float st, ct, sf, cf;
    sincosf(t, &st, &ct);
    sincosf(f, &sf, &cf);
    float d[3];
    if (psi < 0) {  // fibre direction
      d[0] = st * cf;
      d[1] = st * sf;
      d[2] = ct;
    } else {  // fanning direction (getFanningOrientation.cc:49-51)
      float sp, cp;
      sincosf(base[psi * ns + k], &sp, &cp);
      d[0] = -cp * sf - ct * cf * sp;
      d[1] = cf * cp - ct * sf * sp;
      d[2] = st * sp;
    }
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) m[i][j] += (double)d[i] * d[j];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) m[j][i] = (m[i][j] /= ns);
  double vec[3][3];
  jacobi3(m, vec);
  int b = 0;
  for (int i = 1; i < 3; ++i)
    if (m[i][i] > m[b][b]) b = i;
  double e[3] = {vec[0][b], vec[1][b], vec[2][b]};
  const double nrm = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  // An eigenvector's sign is arbitrary: report the one with z >= 0 (then
  // y, then x, for vectors in the xy-plane).
  const double key = e[2] != 0.0 ? e[2] : (e[1] != 0.0 ? e[1] : e[0]);
  const double sg = (key < 0 ? -1.0 : 1.0) / (nrm > 0 ? nrm : 1.0);
  for (int i = 0; i < 3; ++i) dyads[(size_t)v * 3 + i] = (float)(e[i] * sg);
  if (disp) {
    double lmax = 0.0;
    for (int i = 0; i < 3; ++i) lmax = fmax(lmax, fabs(m[i][i]));
    disp[v] = (float)(1.0 - lmax);
  }
}

__global__ __launch_bounds__(kBlock) void noddi_maps_kernel(int n, int np, int ns, const float* __restrict__ x,
                                                            int kappa, int beta, float* __restrict__ maps) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= n) return;
  const float* base = x + (size_t)v * np * ns;
  double sk = 0.0, sb = 0.0, sk2 = 0.0;
  for (int k = 0; k < ns; ++k) {
    const float kk = base[kappa * ns + k];
    sk += kk;
    if (beta >= 0) {
      const float bb = base[beta * ns + k];
      sb += bb;
      sk2 += (float)(kk - bb);  // the k2 = kappa - beta sample volume
    }
  }
  const float mk = (float)(sk / ns);
  if (beta < 0) {  // Watson: OD = 2/pi atan(1/kappa)
    maps[v] = atanf(1.0f / mk) * kTwoDivPi;
    return;
  }
  const float mb = (float)(sb / ns), mk2 = (float)(sk2 / ns);
  float* o = maps + (size_t)v * 5;
  o[0] = mk2;
  o[1] = atanf(1.0f / mk2) * kTwoDivPi;                                // ODIp
  o[2] = atanf(1.0f / mk) * kTwoDivPi;                                 // ODIs
  o[3] = atanf(fabsf(sqrtf((1.0f / mk) * (1.0f / mk2)))) * kTwoDivPi;  // ODItot
  o[4] = atanf(mb / mk2) * kTwoDivPi;                                  // DA
}

int launch_check(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return cdm::set_error(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

}  // namespace

extern "C" {

int cdm_sample_mean(int nvox, int nparams, int nsamples, const float* samples, float* mean, void* stream) {
  if (nvox < 0 || nparams <= 0 || nsamples <= 0) return cdm::set_error("cdm_sample_mean: bad sizes");
  if (nvox == 0) return 0;
  if (!samples || !mean) return cdm::set_error("cdm_sample_mean: null pointer");
  const long long rows = (long long)nvox * nparams;
  hipLaunchKernelGGL(sample_mean_kernel, dim3((unsigned)((rows + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                     (hipStream_t)stream, rows, nsamples, samples, mean);
  return launch_check("cdm_sample_mean");
}

int cdm_dyads(int nvox, int nparams, int nsamples, const float* samples, int theta, int phi, int psi, float* dyads,
              float* dispersion, void* stream) {
  if (nvox < 0 || nparams <= 0 || nsamples <= 0) return cdm::set_error("cdm_dyads: bad sizes");
  if (theta < 0 || theta >= nparams || phi < 0 || phi >= nparams || psi >= nparams)
    return cdm::set_error("cdm_dyads: parameter index out of range");
  if (nvox == 0) return 0;
  if (!samples || !dyads) return cdm::set_error("cdm_dyads: null pointer");
  hipLaunchKernelGGL(dyads_kernel, dim3((nvox + kBlock - 1) / kBlock), dim3(kBlock), 0, (hipStream_t)stream, nvox,
                     nparams, nsamples, samples, theta, phi, psi, dyads, dispersion);
  return launch_check("cdm_dyads");
}

int cdm_noddi_maps(int nvox, int nparams, int nsamples, const float* samples, int kappa, int beta, float* maps,
                   void* stream) {
  if (nvox < 0 || nparams <= 0 || nsamples <= 0) return cdm::set_error("cdm_noddi_maps: bad sizes");
  if (kappa < 0 || kappa >= nparams || beta >= nparams) return cdm::set_error("cdm_noddi_maps: parameter index out of range");
  if (nvox == 0) return 0;
  if (!samples || !maps) return cdm::set_error("cdm_noddi_maps: null pointer");
  hipLaunchKernelGGL(noddi_maps_kernel, dim3((nvox + kBlock - 1) / kBlock), dim3(kBlock), 0, (hipStream_t)stream, nvox,
                     nparams, nsamples, samples, kappa, beta, maps);
  return launch_check("cdm_noddi_maps");
}

}  // extern "C"
