// Wave64 reductions shared by the strip kernels: fixed-order sums (the same
// value in every lane, bit-reproducible) and the compensated (two-sum) form.
#pragma once

#include <hip/hip_runtime.h>

namespace petstrip {

// Sum over the 64 lanes of a wave; every lane returns the same value and the
// order of additions is fixed.  Rows of 16 lanes are reduced with DPP
// (quad_perm 1032, quad_perm 2301, row_ror 4, row_ror 8), then the four row
// sums are combined as (r0 + r1) + (r2 + r3).
__device__ __forceinline__ float wave_sum(float v) {
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xb1, 0xf, 0xf, false));
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4e, 0xf, 0xf, false));
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xf, 0xf, false));
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, false));
  float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return (r0 + r1) + (r2 + r3);
}

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// s + e = a + b exactly (Knuth's two-sum).
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Sum over each aligned segment of 64 / PACK lanes (PACK = 1, 2, 4), every
// lane of a segment returning its sum: the rows of 16 as in wave_sum, then
// for two segments (r0 + r1) and (r2 + r3), the partial sums wave_sum adds.
template <int PACK>
__device__ __forceinline__ float seg_sum(float v) {
  static_assert(PACK == 1 || PACK == 2 || PACK == 4, "segments of 64, 32 or 16 lanes");
  if constexpr (PACK == 1) {
    return wave_sum(v);
  } else {
    v += __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xb1, 0xf, 0xf, false));
    v += __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4e, 0xf, 0xf, false));
    v += __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xf, 0xf, false));
    v += __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, false));
    if constexpr (PACK == 4) return v;
    const float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    const float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    const float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    const float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return __lane_id() < 32 ? r0 + r1 : r2 + r3;
  }
}

// Compensated sum of (hi, lo) pairs over each aligned segment of SEG lanes;
// partners combine in lane order so both compute the same value.
template <int SEG = 64>
__device__ __forceinline__ float seg_sum_comp(float hi, float lo) {
  for (int off = 1; off < SEG; off <<= 1) {
    float oh = __shfl_xor(hi, off, 64);
    float ol = __shfl_xor(lo, off, 64);
    bool upper = (__lane_id() & off) != 0;
    float a = upper ? oh : hi, b = upper ? hi : oh;
    float la = upper ? ol : lo, lb = upper ? lo : ol;
    float s, e;
    two_sum(a, b, s, e);
    hi = s;
    lo = e + (la + lb);
  }
  return hi + lo;
}

// Compensated wave sum of (hi, lo) pairs.
__device__ __forceinline__ float wave_sum_comp(float hi, float lo) {
  return seg_sum_comp<64>(hi, lo);
}

}  // namespace petstrip
