// wave.h -- cross-lane sums within a wavefront by DPP (no LDS traffic).
#pragma once

#include <hip/hip_runtime.h>

namespace myqc {

// Sum over the 64 lanes of a wavefront (all active), returned wave-uniform:
// within each row of 16 lanes by four DPP exchanges (lanes xor 1, xor 2,
// the mirrored half-row, the mirrored row), then the four row sums read from
// lanes 0, 16, 32, 48.  No LDS traffic.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true);
}
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i<CTRL>((int)b), hi = dpp_i<CTRL>((int)(b >> 32));
    return v + __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
__device__ __forceinline__ double lane_value(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
template <typename T>
__device__ __forceinline__ T wave_total(T v) {
    v = dpp_add<0xB1>(v);    // quad_perm [1,0,3,2]
    v = dpp_add<0x4E>(v);    // quad_perm [2,3,0,1]
    v = dpp_add<0x141>(v);   // row_half_mirror
    v = dpp_add<0x140>(v);   // row_mirror
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}

// Sum over aligned groups of G lanes (G = 1, 2, 4 or 8; every lane of a
// group active): lanes xor 1, xor 2 within each quad, then the mirrored
// half-row; every lane of the group ends with the group's sum.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8, "group of 1, 2, 4 or 8 lanes");
    if constexpr (G >= 2) v = dpp_add<0xB1>(v);
    if constexpr (G >= 4) v = dpp_add<0x4E>(v);
    if constexpr (G >= 8) v = dpp_add<0x141>(v);
    return v;
}

}  // namespace myqc
