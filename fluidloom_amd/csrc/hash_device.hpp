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

// Whether slot s holds `key` after the call (claimed now or before); false
// when another key holds it.  A slot's key never changes once set, so a
// relaxed load decides every case but an empty slot, and only that one needs
// the CAS: a key inserted millions of times (a hot duplicate) then costs
// loads, not serialised atomics on one address.
__device__ inline bool claim_slot(uint64_t* tkeys, uint64_t s, uint64_t key) {
    const uint64_t cur = __hip_atomic_load(&tkeys[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == key) return true;
    if (cur != 0xFFFFFFFFFFFFFFFFull) return false;
    const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long*>(&tkeys[s]),
                                              0xFFFFFFFFFFFFFFFFull, (unsigned long long)key);
    return prev == 0xFFFFFFFFFFFFFFFFull || prev == key;
}

// Keeps the smallest value of a slot.  Stored values only decrease, so a
// loaded value at or below `val` means the atomic can be skipped.
__device__ inline void keep_min(uint32_t* tvals, uint64_t s, uint32_t val) {
    if (__hip_atomic_load(&tvals[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > val) atomicMin(&tvals[s], val);
}

// Returns the number of probes used (kMaxProbes when the key was dropped).
__device__ inline uint32_t hash_insert(uint64_t* tkeys, uint32_t* tvals, uint64_t cap, uint64_t key,
                                       uint32_t val) {
    uint64_t s = hash_slot(key, cap);
    for (uint32_t p = 0; p < kMaxProbes; ++p) {
        if (claim_slot(tkeys, s, key)) {
            keep_min(tvals, s, val);
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

// Exact map for the AMR cell lookups (not the reference's table): the
// packed cell keys (L<<57 | x<<38 | y<<19 | z) carry z in their low bits, so
// the slot is the top log2(cap) bits of the Fibonacci product instead of the
// low ones, and probing runs until a hit or an empty slot.  The capacity
// rule keeps the load <= 0.6, so an empty slot always exists and no key is
// ever dropped.
__device__ inline uint64_t cell_slot(uint64_t key, uint64_t cap) {
    const int bits = 63 - __clzll((long long)cap);
    return (key * 0x9e3779b97f4a7c15ull) >> (64 - bits);
}

__device__ inline void cell_insert(uint64_t* tkeys, uint32_t* tvals, uint64_t cap, uint64_t key, uint32_t val) {
    uint64_t s = cell_slot(key, cap);
    for (uint64_t p = 0; p < cap; ++p) {
        if (claim_slot(tkeys, s, key)) {
            keep_min(tvals, s, val);
            return;
        }
        s = (s + 1) & (cap - 1);
    }
}

__device__ inline uint32_t cell_lookup(const uint64_t* __restrict__ tkeys, const uint32_t* __restrict__ tvals,
                                       uint64_t cap, uint64_t key) {
    if (key == 0xFFFFFFFFFFFFFFFFull) return 0xFFFFFFFFu;
    uint64_t s = cell_slot(key, cap);
    for (uint64_t p = 0; p < cap; ++p) {
        const uint64_t k = tkeys[s];
        if (k == key) return tvals[s];
        if (k == 0xFFFFFFFFFFFFFFFFull) return 0xFFFFFFFFu;
        s = (s + 1) & (cap - 1);
    }
    return 0xFFFFFFFFu;
}

// N lookups advanced together, one probe step each per round: their
// dependent loads overlap instead of one chain waiting on another.  The
// value is loaded beside the key so a hit costs no extra round.  out[j]:
// the stored value, or 0xFFFFFFFF when key[j] is absent (or is EMPTY).
template <int N>
__device__ inline void cell_lookup_n(const uint64_t* __restrict__ tkeys, const uint32_t* __restrict__ tvals,
                                     uint64_t cap, const uint64_t (&key)[N], uint32_t (&out)[N]) {
    uint64_t s[N];
    unsigned open = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        out[j] = 0xFFFFFFFFu;
        s[j] = cell_slot(key[j], cap);
        if (key[j] != 0xFFFFFFFFFFFFFFFFull) open |= 1u << j;
    }
    for (uint64_t p = 0; p < cap && open; ++p) {
        uint64_t k[N];
        uint32_t v[N];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            k[j] = 0xFFFFFFFFFFFFFFFFull;
            v[j] = 0xFFFFFFFFu;
            if ((open >> j) & 1u) {
                k[j] = tkeys[s[j]];
                v[j] = tvals[s[j]];
            }
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (!((open >> j) & 1u)) continue;
            if (k[j] == key[j]) {
                out[j] = v[j];
                open &= ~(1u << j);
            } else if (k[j] == 0xFFFFFFFFFFFFFFFFull) {
                open &= ~(1u << j);
            } else {
                s[j] = (s[j] + 1) & (cap - 1);
            }
        }
    }
}

// Cell-key map for the AMR kernels (packed keys L<<57 | x<<38 | y<<19 | z),
// exact like the map above but laid out for spatial locality: the eight
// cells of one 2x2x2 block (same level, coordinates equal but for their
// lowest bits) hash to one aligned group of eight slots, each at the offset
// of its octant.  A cell's D3Q19 neighbours then lie in at most 8 groups (6
// of the 18 in its own), each 64 B of keys and 32 B of values, and the cells
// of one refined region share those lines, where a per-cell hash sends every
// lookup to two random lines.  Full blocks collide as units, so a displaced
// key sits several slots past its home (about 5 on average at load 0.57):
// lookups therefore scan whole groups, the eight keys of a group in one
// round (four 16 B loads), keeping the linear-probing order -- slots from
// the home offset to the end of its group, then whole groups -- so a round
// count is the group displacement (1.7 on average at load 0.57).  Keys whose
// low bits are all zero (coarse anchors) fill a group from its first slot:
// bucketed linear probing at the same load.  The key arrays must be 16 B
// aligned.
__device__ inline uint64_t cellkey_slot(uint64_t key, uint64_t cap) {
    constexpr uint64_t kLow = (1ull << 38) | (1ull << 19) | 1ull;
    const int bits = 63 - __clzll((long long)cap);  // cap >= 8
    const uint64_t group = ((key & ~kLow) * 0x9e3779b97f4a7c15ull) >> (64 - bits);
    const uint64_t oct = (key & 1ull) | ((key >> 18) & 2ull) | ((key >> 36) & 4ull);
    return (group & ~7ull) | oct;
}

// Bit t set where group slot t (of the eight at g) holds `key`, in `hit`, or
// is empty, in `emp`.
__device__ inline void cellkey_scan(const uint64_t* __restrict__ tkeys, uint64_t g, uint64_t key, unsigned& hit,
                                    unsigned& emp) {
    const ulonglong2* p = reinterpret_cast<const ulonglong2*>(tkeys + g);
    hit = emp = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const ulonglong2 v = p[q];
        hit |= (v.x == key ? 1u : 0u) << (2 * q) | (v.y == key ? 1u : 0u) << (2 * q + 1);
        emp |= (v.x == 0xFFFFFFFFFFFFFFFFull ? 1u : 0u) << (2 * q) |
               (v.y == 0xFFFFFFFFFFFFFFFFull ? 1u : 0u) << (2 * q + 1);
    }
}

// Insertion probes slot by slot with claim_slot (coherent loads, a CAS only
// on an empty slot): the eight cells of a displaced block usually come from
// one wave and would race for the same free slots under a group scan.  The
// table is the same linear-probing table either way.
__device__ inline void cellkey_insert(uint64_t* tkeys, uint32_t* tvals, uint64_t cap, uint64_t key, uint32_t val) {
    uint64_t s = cellkey_slot(key, cap);
    for (uint64_t p = 0; p < cap; ++p) {
        if (claim_slot(tkeys, s, key)) {
            keep_min(tvals, s, val);
            return;
        }
        s = (s + 1) & (cap - 1);
    }
}

// N lookups advanced together, one group per lookup per round.  out[j]: the
// stored value (kValues) or 0 (presence only), 0xFFFFFFFF when key[j] is
// absent or EMPTY.  Values are loaded after the probe rounds, all at once.
template <int N, bool kValues = true>
__device__ inline void cellkey_lookup_n(const uint64_t* __restrict__ tkeys, const uint32_t* __restrict__ tvals,
                                        uint64_t cap, const uint64_t (&key)[N], uint32_t (&out)[N]) {
    uint64_t s[N];
    unsigned open = 0, found = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        out[j] = 0xFFFFFFFFu;
        s[j] = cellkey_slot(key[j], cap);
        if (key[j] != 0xFFFFFFFFFFFFFFFFull) open |= 1u << j;
    }
    for (uint64_t r = 0; r < cap && open; r += 8) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (!((open >> j) & 1u)) continue;
            const uint64_t g = s[j] & ~7ull;
            unsigned hit, emp;
            cellkey_scan(tkeys, g, key[j], hit, emp);
            const unsigned m = (hit | emp) & (0xFFu << (s[j] & 7ull));
            if (!m) {
                s[j] = (g + 8) & (cap - 1);
                continue;
            }
            const unsigned t = (unsigned)__builtin_ctz(m);
            open &= ~(1u << j);
            if ((hit >> t) & 1u) {
                found |= 1u << j;
                s[j] = g + t;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
        if ((found >> j) & 1u) out[j] = kValues ? tvals[s[j]] : 0u;
}

__device__ inline uint32_t cellkey_lookup(const uint64_t* __restrict__ tkeys, const uint32_t* __restrict__ tvals,
                                          uint64_t cap, uint64_t key) {
    const uint64_t k[1] = {key};
    uint32_t o[1];
    cellkey_lookup_n<1>(tkeys, tvals, cap, k, o);
    return o[0];
}

}  // namespace fl
