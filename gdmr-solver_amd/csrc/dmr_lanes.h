// Lane exchanges for the fixed-order wavefront butterflies, without the LDS crossbar that
// __shfl_xor goes through (ds_bpermute, one LDS round trip per 32 bits and step).  Partner
// lane ^ m comes from DPP within a 16-lane row (quad_perm for m = 1, 2; row_ror for m = 4, 8)
// and from the gfx950 row swaps v_permlane16_swap (m = 16) and v_permlane32_swap (m = 32).  The
// pairs are those of __shfl_xor, so a butterfly gives the same bits; every lane of the group
// that exchanges (the 32-lane half for width 32, the wavefront for 64) must be active.
// They suit a single latency-bound butterfly (a Jacobi rotation's dot product, a reflector's norm):
// each step is a short dependent chain of VALU operations instead of an LDS round trip.  Many
// independent butterflies at once (a workgroup's column sums) are better served by ds_bpermute,
// whose LDS operations overlap, so the multi-column reductions keep __shfl_xor.
// Device code only: included by the .hip translation units.
#pragma once

#include <hip/hip_runtime.h>

namespace dmr {

template <int M> __device__ inline int lane_xor(int v) {
  static_assert(M == 1 || M == 2 || M == 4 || M == 8 || M == 16 || M == 32, "xor distance");
  const int lane = (int)(threadIdx.x & 63);
  if constexpr (M == 1) {
    return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (M == 2) {
    return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (M == 4) {
    // row_ror:12 brings lane + 4, row_ror:4 lane - 4 (within the row).
    const int up = __builtin_amdgcn_update_dpp(0, v, 0x12C, 0xF, 0xF, false);
    const int dn = __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xF, false);
    return (lane & 4) ? dn : up;
  } else if constexpr (M == 8) {
    return __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (M == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (lane & 16) ? (int)r[0] : (int)r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (lane & 32) ? (int)r[0] : (int)r[1];
  }
}
template <int M> __device__ inline float lane_xor(float v) {
  return __int_as_float(lane_xor<M>(__float_as_int(v)));
}
template <int M> __device__ inline double lane_xor(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = lane_xor<M>((int)(b & 0xffffffffLL)), hi = lane_xor<M>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// v + partner over W = 32 or 64 lanes, distances W/2 down to 1 (the order of the __shfl_xor
// butterflies it replaces); every lane gets the sum.
template <int W, typename T> __device__ inline T xor_sum(T v) {
  static_assert(W == 32 || W == 64, "width");
  if constexpr (W == 64) v = v + lane_xor<32>(v);
  v = v + lane_xor<16>(v);
  v = v + lane_xor<8>(v);
  v = v + lane_xor<4>(v);
  v = v + lane_xor<2>(v);
  v = v + lane_xor<1>(v);
  return v;
}
// fmax over the 64 lanes of the wavefront, distances 32 down to 1.
__device__ inline double xor_max64(double v) {
  v = fmax(v, lane_xor<32>(v));
  v = fmax(v, lane_xor<16>(v));
  v = fmax(v, lane_xor<8>(v));
  v = fmax(v, lane_xor<4>(v));
  v = fmax(v, lane_xor<2>(v));
  v = fmax(v, lane_xor<1>(v));
  return v;
}

}  // namespace dmr
