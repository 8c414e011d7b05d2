// Device-side open-addressing hash table (u64 key -> u32 value).
//
// Semantics follow kernels/hashmap/hash_build.cl:19-67 and
// kernels/hashmap/hash_query.cl:22-56: Fibonacci hashing
// slot = (key * 0x9e3779b97f4a7c15) & (capacity - 1), linear probing, at most
// 32 probes, EMPTY = ~0 keys, INVALID = ~0u values.  Insertion claims a slot
// with a 64-bit CAS; the value is combined with atomicMin so that a key
// inserted more than once deterministically keeps its smallest value.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fl {

constexpr uint32_t kMaxProbes = 32;

__device__ inline uint64_t hash_slot(uint64_t key, uint64_t cap) {
    return (key * 0x9e3779b97f4a7c15ull) & (cap - 1);
}

// Returns the number of probes used (kMaxProbes when the key was dropped).
__device__ inline uint32_t hash_insert(uint64_t* tkeys, uint32_t* tvals, uint64_t cap, uint64_t key,
                                       uint32_t val) {
    uint64_t s = hash_slot(key, cap);
    for (uint32_t p = 0; p < kMaxProbes; ++p) {
        unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&tkeys[s]),
                                            0xFFFFFFFFFFFFFFFFull, (unsigned long long)key);
        if (prev == 0xFFFFFFFFFFFFFFFFull || prev == key) {
            atomicMin(&tvals[s], val);
            return p;
        }
        s = (s + 1) & (cap - 1);
    }
    return kMaxProbes;
}

__device__ inline uint32_t hash_lookup(const uint64_t* __restrict__ tkeys, const uint32_t* __restrict__ tvals,
                                       uint64_t cap, uint64_t key, uint32_t* probes = nullptr) {
    if (key == 0xFFFFFFFFFFFFFFFFull) {
        if (probes) *probes = 0;
        return 0xFFFFFFFFu;
    }
    uint64_t s = hash_slot(key, cap);
    for (uint32_t p = 0; p < kMaxProbes; ++p) {
        uint64_t k = tkeys[s];
        if (k == key) {
            if (probes) *probes = p;
            return tvals[s];
        }
        if (k == 0xFFFFFFFFFFFFFFFFull) {
            if (probes) *probes = p;
            return 0xFFFFFFFFu;
        }
        s = (s + 1) & (cap - 1);
    }
    if (probes) *probes = kMaxProbes;
    return 0xFFFFFFFFu;
}

}  // namespace fl
