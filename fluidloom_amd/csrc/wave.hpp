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
