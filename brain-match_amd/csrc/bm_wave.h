// Cross-lane reductions on a 64-lane wavefront with DPP and ds_swizzle.
//
// Inside a 16-lane row, DPP reads a neighbour's register as an operand
// modifier of the VALU instruction (no LDS traffic): quad_perm swaps lanes
// 1 and 2 apart, row_half_mirror and row_mirror pair the 4- and 8-lane halves.
// Across the 16-lane halves of a 32-lane group ds_swizzle exchanges
// registers through the LDS crossbar without touching LDS memory.  Across
// rows the four row results are read into scalar registers (v_readlane).
// All lanes of every row taking part must be active.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bm_wave {

// DPP controls (GFX9 encoding)
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;  // lane i <- 7 - i within 8 lanes
constexpr int kRowMirror = 0x140;      // lane i <- 15 - i within 16 lanes
constexpr int kSwizzleXor16 = 0x401F;  // ds_swizzle bitmask mode: lane ^ 16 within 32 lanes

template <int kCtrl>
__device__ __forceinline__ int dpp(int v) {
    return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, false);
}

template <int kCtrl>
__device__ __forceinline__ unsigned long long dpp64(unsigned long long v) {
    const int lo = dpp<kCtrl>((int)(uint32_t)v), hi = dpp<kCtrl>((int)(uint32_t)(v >> 32));
    return ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo;
}

// sum over each aligned group of kWidth lanes (8, 16, 32 or 64), in every lane of the group
template <int kWidth>
__device__ __forceinline__ int group_sum(int v) {
    static_assert(kWidth == 8 || kWidth == 16 || kWidth == 32 || kWidth == 64, "group width");
    v += dpp<kQuadXor1>(v);
    v += dpp<kQuadXor2>(v);
    v += dpp<kRowHalfMirror>(v);
    if constexpr (kWidth >= 16) v += dpp<kRowMirror>(v);
    if constexpr (kWidth == 32) v += __builtin_amdgcn_ds_swizzle(v, kSwizzleXor16);
    if constexpr (kWidth == 64)
        v = __builtin_amdgcn_readlane(v, 0) + __builtin_amdgcn_readlane(v, 16) + __builtin_amdgcn_readlane(v, 32) +
            __builtin_amdgcn_readlane(v, 48);
    return v;
}

__device__ __forceinline__ int wave_sum(int v) { return group_sum<64>(v); }

__device__ __forceinline__ unsigned long long readlane64(unsigned long long v, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
}

// 64-bit sum over the wave, in every lane
__device__ __forceinline__ long long wave_sum64(long long v) {
    unsigned long long u = (unsigned long long)v;
    u += dpp64<kQuadXor1>(u);
    u += dpp64<kQuadXor2>(u);
    u += dpp64<kRowHalfMirror>(u);
    u += dpp64<kRowMirror>(u);
    return (long long)(readlane64(u, 0) + readlane64(u, 16) + readlane64(u, 32) + readlane64(u, 48));
}

// max / min of a 64-bit key over the wave, in every lane
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k) {
    unsigned long long t;
    t = dpp64<kQuadXor1>(k);
    k = t > k ? t : k;
    t = dpp64<kQuadXor2>(k);
    k = t > k ? t : k;
    t = dpp64<kRowHalfMirror>(k);
    k = t > k ? t : k;
    t = dpp64<kRowMirror>(k);
    k = t > k ? t : k;
    const unsigned long long a = readlane64(k, 0), b = readlane64(k, 16), c = readlane64(k, 32),
                             d = readlane64(k, 48);
    const unsigned long long ab = a > b ? a : b, cd = c > d ? c : d;
    return ab > cd ? ab : cd;
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long k) {
    unsigned long long t;
    t = dpp64<kQuadXor1>(k);
    k = t < k ? t : k;
    t = dpp64<kQuadXor2>(k);
    k = t < k ? t : k;
    t = dpp64<kRowHalfMirror>(k);
    k = t < k ? t : k;
    t = dpp64<kRowMirror>(k);
    k = t < k ? t : k;
    const unsigned long long a = readlane64(k, 0), b = readlane64(k, 16), c = readlane64(k, 32),
                             d = readlane64(k, 48);
    const unsigned long long ab = a < b ? a : b, cd = c < d ? c : d;
    return ab < cd ? ab : cd;
}

}  // namespace bm_wave
