// Wavefront (64-lane) helpers for gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fl {

__device__ inline unsigned lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// Lanes of the wave (restricted to `active`) holding the same 8-bit value.
__device__ inline uint64_t match_u8(unsigned v, uint64_t active) {
    uint64_t peers = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint64_t bal = __ballot((v >> b) & 1u);
        peers &= ((v >> b) & 1u) ? bal : ~bal;
    }
    return peers;
}

// DPP lane swaps (VALU, no LDS): CTRL 0xB1 = quad_perm(1,0,3,2) (lane xor 1),
// 0x4E = quad_perm(2,3,0,1) (lane xor 2), 0x141 = row_half_mirror (lane i of
// each 8-lane group takes lane 7 - i).  Every lane of the wave must be active.
template <int CTRL>
__device__ inline float dpp_swap(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ inline double dpp_swap(double v) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over each aligned group of 8 lanes, every lane receiving it: pairs,
// then pairs of pairs, then the two quads (after two rounds lane 7 - i holds
// the other quad's sum).  The same fixed tree as xor shuffles by 1, 2, 4, so
// the bits do not depend on the launch; all lanes of the wave must be active.
template <typename R>
__device__ inline R group8_sum(R v) {
    v += dpp_swap<0xB1>(v);
    v += dpp_swap<0x4E>(v);
    v += dpp_swap<0x141>(v);
    return v;
}

// Inclusive wave scan of u32 (Kogge-Stone over shuffles).
__device__ inline uint32_t wave_inclusive_scan(uint32_t v) {
    const unsigned lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= (unsigned)d) v += t;
    }
    return v;
}

__device__ inline uint64_t wave_inclusive_scan64(uint64_t v) {
    const unsigned lane = lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t t = __shfl_up(v, d, 64);
        if (lane >= (unsigned)d) v += t;
    }
    return v;
}

template <typename T>
__device__ inline T atomic_load_agent(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ inline void atomic_store_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace fl
