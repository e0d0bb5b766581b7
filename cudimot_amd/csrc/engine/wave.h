// Cross-lane sums for voxel groups on a 64-lane CDNA wavefront.
//
// A voxel group is L consecutive lanes (L a power of two, 1..64) that share
// one voxel: lane l of the group handles measurements l, l+L, l+2L, ...
// Group::sum returns the group total in every lane of the group, so all
// lanes keep identical per-voxel state and run the voxel's serial steps in
// lockstep, with no leader lane, LDS state or barrier.
//
// Butterfly stages: xor 1 and xor 2 with DPP quad_perm, 4 and 8 with DPP
// row_half_mirror / row_mirror (valid once the smaller groups already hold
// equal sums), 16 with ds_swizzle (xor mode), 32 with ds_bpermute.  Each
// stage adds two values that are equal on both partner lanes, so every lane
// ends with a bit-identical result.
#pragma once
#include <hip/hip_runtime.h>

namespace cdm {

__device__ __forceinline__ int lane_dpp_quad_xor1(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ int lane_dpp_quad_xor2(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
}
__device__ __forceinline__ int lane_dpp_half_mirror(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
}
__device__ __forceinline__ int lane_dpp_mirror(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
}
__device__ __forceinline__ int lane_swz_xor16(int v) { return __builtin_amdgcn_ds_swizzle(v, 0x401F); }
__device__ __forceinline__ int lane_xor32(int v) {
  const int src = (int)(__lane_id() ^ 32u);
  return __builtin_amdgcn_ds_bpermute(src << 2, v);
}

template <int STAGE>
__device__ __forceinline__ int lane_exchange(int v) {
  if constexpr (STAGE == 1) return lane_dpp_quad_xor1(v);
  else if constexpr (STAGE == 2) return lane_dpp_quad_xor2(v);
  else if constexpr (STAGE == 4) return lane_dpp_half_mirror(v);
  else if constexpr (STAGE == 8) return lane_dpp_mirror(v);
  else if constexpr (STAGE == 16) return lane_swz_xor16(v);
  else return lane_xor32(v);
}

template <int STAGE>
__device__ __forceinline__ float exchange(float v) {
  return __int_as_float(lane_exchange<STAGE>(__float_as_int(v)));
}
template <int STAGE>
__device__ __forceinline__ double exchange(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_exchange<STAGE>((int)(b & 0xffffffffll));
  const int hi = lane_exchange<STAGE>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int STAGE>
__device__ __forceinline__ int exchange(int v) {
  return lane_exchange<STAGE>(v);
}

template <int L>
struct Group {
  static_assert(L == 1 || L == 2 || L == 4 || L == 8 || L == 16 || L == 32 || L == 64, "group size");
  template <class V>
  __device__ __forceinline__ static V sum(V v) {
    if constexpr (L >= 2) v = v + exchange<1>(v);
    if constexpr (L >= 4) v = v + exchange<2>(v);
    if constexpr (L >= 8) v = v + exchange<4>(v);
    if constexpr (L >= 16) v = v + exchange<8>(v);
    if constexpr (L >= 32) v = v + exchange<16>(v);
    if constexpr (L >= 64) v = v + exchange<32>(v);
    return v;
  }
  // Logical OR over the group (flags kept as 0/1 ints).
  __device__ __forceinline__ static int any(int f) { return sum(f) != 0; }
  // The value held by group lane `src` (0 <= src < L), in every lane of the
  // group (ds_bpermute from lane (first lane of the group) + src).
  __device__ __forceinline__ static int bcast(int v, int src) {
    if constexpr (L == 1) return v;
    const int from = (int)(__lane_id() & ~(unsigned)(L - 1)) + src;
    return __builtin_amdgcn_ds_bpermute(from << 2, v);
  }
  __device__ __forceinline__ static float bcast(float v, int src) { return __int_as_float(bcast(__float_as_int(v), src)); }
  __device__ __forceinline__ static double bcast(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = bcast((int)(b & 0xffffffffll), src);
    const int hi = bcast((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
};

}  // namespace cdm
