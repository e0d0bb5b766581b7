// Orientation initialisation from a tensor fit: the principal and second
// eigenvectors of the diffusion tensor give the fibre direction (theta, phi)
// and the fanning angle psi of the Bingham/Watson plug-ins' frame.
// Restates the reference's pipeline utilities: dtifit V1/V2
// (utils/Run_dtifit.sh:36-46), cart2spherical (utils/cart2spherical.cc:17-69,
// FSL miscmaths cart2sph: theta = acos(z/|v|), phi = atan2(y, x)) and
// initialise_Psi (utils/initialise_Psi.cc:18-106), which solves for psi in the
// frame fan(psi) = cos(psi) a + sin(psi) b with a = (-sin phi, cos phi, 0),
// b = (-cos theta cos phi, -cos theta sin phi, sin theta)
// (NODDI_Bingham/modelfunctions.h:99-121); a and b are orthonormal and normal
// to the fibre, so psi = atan2(V2.b, V2.a).  The eigenvectors come from cyclic
// Jacobi rotations in FP64 (one thread per voxel).
#include <hip/hip_runtime.h>

#include <string>

#include "../../../include/cudimot_amd.h"
#include "../host/model_ops.h"

namespace {

// Eigen-decomposition of a symmetric 3x3 matrix: eigenvalues descending in
// w, eigenvectors in the columns of V.
__device__ void jacobi3(double A[3][3], double w[3], double V[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; ++sweep) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    const double scale = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
    if (off <= 1e-30 * scale || off == 0.0) break;
    for (int p = 0; p < 2; ++p) {
      for (int q = p + 1; q < 3; ++q) {
        if (A[p][q] == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- A J
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^T A
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
  int idx[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (A[idx[j]][idx[j]] > A[idx[i]][idx[i]]) {
        const int t = idx[i];
        idx[i] = idx[j];
        idx[j] = t;
      }
  double Vs[3][3];
  for (int i = 0; i < 3; ++i) {
    w[i] = A[idx[i]][idx[i]];
    for (int k = 0; k < 3; ++k) Vs[k][i] = V[k][idx[i]];
  }
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) V[k][i] = Vs[k][i];
}

__global__ void tensor_frame_kernel(int n, const float* __restrict__ dti, int stride, int first, float* __restrict__ out,
                                    int out_stride) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const float* d = dti + (size_t)v * stride + first;  // Dxx Dxy Dxz Dyy Dyz Dzz
  double A[3][3] = {{d[0], d[1], d[2]}, {d[1], d[3], d[4]}, {d[2], d[4], d[5]}};
  double w[3], V[3][3];
  jacobi3(A, w, V);
  const double x = V[0][0], y = V[1][0], z = V[2][0];
  const double r = sqrt(x * x + y * y + z * z);
  const double th = r > 0 ? acos(fmin(fmax(z / r, -1.0), 1.0)) : 1.5707963267948966;
  const double ph = r > 0 ? atan2(y, x) : 1.5707963267948966;
  const double sp = sin(ph), cp = cos(ph), ct = cos(th), st = sin(th);
  const double f0 = V[0][1], f1 = V[1][1], f2 = V[2][1];
  const double fa = -sp * f0 + cp * f1;                     // V2 . a
  const double fb = -ct * cp * f0 - ct * sp * f1 + st * f2;  // V2 . b
  float* o = out + (size_t)v * out_stride;
  o[0] = (float)th;
  o[1] = (float)ph;
  o[2] = (float)atan2(fb, fa);
  o[3] = (float)w[0];
  o[4] = (float)w[1];
  o[5] = (float)w[2];
}

}  // namespace

extern "C" int cdm_tensor_frame(int n, const float* dti, int stride, int first, float* out, void* stream) {
  if (n < 0 || stride < first + 6 || first < 0) return cdm::set_error("tensor frame: need stride >= first + 6");
  if (n == 0) return 0;
  if (!dti || !out) return cdm::set_error("null device pointer");
  hipLaunchKernelGGL(tensor_frame_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, dti, stride,
                     first, out, 6);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : cdm::set_error(std::string("tensor frame launch: ") + hipGetErrorString(e), -2);
}
