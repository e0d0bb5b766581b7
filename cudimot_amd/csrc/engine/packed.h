// Packed lower triangle of a symmetric NP x NP matrix: element (i, j) with
// 0 <= j <= i < NP is stored at e = i (i + 1) / 2 + j.  The LM kernels keep
// J^T J in this order; the 64-lane LDS sums hand entry e to lane e % 64.
#pragma once
#include <hip/hip_runtime.h>

namespace cdm {

constexpr int tri_size(int np) { return np * (np + 1) / 2; }

// (i, j) of packed element e (0 <= e < tri_size(NP)); unrolled over the NP
// rows into selects, with no data-dependent loop.
template <int NP>
__host__ __device__ constexpr void tri_ij(int e, int& i, int& j) {
  i = 0;
  j = 0;
  int k = 0;  // first element of row ii
  for (int ii = 0; ii < NP; ++ii) {
    if (e >= k && e < k + ii + 1) {
      i = ii;
      j = e - k;
    }
    k += ii + 1;
  }
}

// Every packed element maps back to itself, inside the triangle.
template <int NP>
constexpr bool tri_ij_exact() {
  for (int e = 0; e < tri_size(NP); ++e) {
    int i = -1, j = -1;
    tri_ij<NP>(e, i, j);
    if (i < 0 || i >= NP || j < 0 || j > i || i * (i + 1) / 2 + j != e) return false;
  }
  return true;
}

}  // namespace cdm
