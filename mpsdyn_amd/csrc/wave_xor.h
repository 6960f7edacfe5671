// Butterfly exchange across a wave64 without LDS: lane i receives the value that lane i ^ O
// holds (O = 32, 16, 8, 4, 2, 1).  __shfl_xor compiles to ds_bpermute, which goes through the
// LDS crossbar that every wave of the CU shares; a workgroup that reduces many values per step
// queues behind it.  Here the exchange is VALU lane movement:
//   O = 32, 16: v_permlane32_swap / v_permlane16_swap of the value with itself, which leaves
//               the lower and the upper half (row pair) in two registers; a lane takes the one
//               that holds its partner;
//   O = 8:      DPP row_ror:8 (within a row of 16, a rotation by 8 is the xor);
//   O = 4:      DPP row_ror:4 on lanes with bit 2 set (they read lane i - 4), row_ror:12 on the
//               others (they read lane i + 4);
//   O = 2, 1:   DPP quad_perm [2,3,0,1] / [1,0,3,2].
// The values moved are exactly __shfl_xor's, so a reduction built on these returns the same
// bits.  All 64 lanes must be active.
#pragma once
#include <hip/hip_runtime.h>

namespace mpsd {

template <int O>
__device__ __forceinline__ unsigned xor_lane_u32(unsigned v) {
  const unsigned lane = __lane_id();
  if constexpr (O == 32) {
    const auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? p[0] : p[1];
  } else if constexpr (O == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? p[0] : p[1];
  } else if constexpr (O == 8) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xF, 0xF, false);
  } else if constexpr (O == 4) {
    const unsigned down = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xF, 0xF, false);
    const unsigned up = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x12C, 0xF, 0xF, false);
    return (lane & 4) ? down : up;
  } else if constexpr (O == 2) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);
  } else {
    static_assert(O == 1, "xor_lane: O is a power of two below 64");
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);
  }
}

template <int O>
__device__ __forceinline__ double xor_lane(double x) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = xor_lane_u32<O>((unsigned)b), hi = xor_lane_u32<O>((unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <int O>
__device__ __forceinline__ double2 xor_lane(double2 v) {
  return double2{xor_lane<O>(v.x), xor_lane<O>(v.y)};
}

// Sum, minimum and maximum over the wave with the butterfly order of a __shfl_xor loop from
// O = 32 down to 1; every lane returns the same value.
__device__ __forceinline__ double wave_sum_x(double v) {
  v += xor_lane<32>(v);
  v += xor_lane<16>(v);
  v += xor_lane<8>(v);
  v += xor_lane<4>(v);
  v += xor_lane<2>(v);
  v += xor_lane<1>(v);
  return v;
}
__device__ __forceinline__ double wave_min_x(double v) {
  v = fmin(v, xor_lane<32>(v));
  v = fmin(v, xor_lane<16>(v));
  v = fmin(v, xor_lane<8>(v));
  v = fmin(v, xor_lane<4>(v));
  v = fmin(v, xor_lane<2>(v));
  v = fmin(v, xor_lane<1>(v));
  return v;
}
__device__ __forceinline__ double wave_max_x(double v) {
  v = fmax(v, xor_lane<32>(v));
  v = fmax(v, xor_lane<16>(v));
  v = fmax(v, xor_lane<8>(v));
  v = fmax(v, xor_lane<4>(v));
  v = fmax(v, xor_lane<2>(v));
  v = fmax(v, xor_lane<1>(v));
  return v;
}

}  // namespace mpsd
