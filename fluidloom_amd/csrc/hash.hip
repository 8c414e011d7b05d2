// a7: GPU hash table kernels (clear / build / query) behind HashTableManager.
#include <cmath>

#include "fl_internal.hpp"
#include "hash_device.hpp"

namespace fl {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_clear(uint64_t* __restrict__ tk, uint32_t* __restrict__ tv,
                                                  uint64_t cap) {
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * kBlock) {
        tk[i] = kEmptyKey;
        tv[i] = kInvalid;
    }
}

__global__ __launch_bounds__(kBlock) void k_build(uint64_t* __restrict__ tk, uint32_t* __restrict__ tv,
                                                  uint64_t cap, const uint64_t* __restrict__ keys,
                                                  const uint32_t* __restrict__ vals, size_t n,
                                                  uint32_t* __restrict__ probes) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        uint64_t k = keys[i];
        uint32_t p = 0;
        if (k != kEmptyKey) p = hash_insert(tk, tv, cap, k, vals ? vals[i] : (uint32_t)i);
        if (probes) probes[i] = p;
    }
}

__global__ __launch_bounds__(kBlock) void k_query(const uint64_t* __restrict__ tk,
                                                  const uint32_t* __restrict__ tv, uint64_t cap,
                                                  const uint64_t* __restrict__ q, uint32_t* __restrict__ out,
                                                  size_t n, uint32_t* __restrict__ probes) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        uint32_t p;
        out[i] = hash_lookup(tk, tv, cap, q[i], &p);
        if (probes) probes[i] = p;
    }
}

template <bool kCellKeys>
__global__ __launch_bounds__(kBlock) void k_cell_build(uint64_t* __restrict__ tk, uint32_t* __restrict__ tv,
                                                       uint64_t cap, const uint64_t* __restrict__ keys, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const uint64_t k = keys[i];
        if (k == kEmptyKey) continue;
        if (kCellKeys)
            cellkey_insert(tk, tv, cap, k, (uint32_t)i);
        else
            cell_insert(tk, tv, cap, k, (uint32_t)i);
    }
}

unsigned cap_grid(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    return (unsigned)(g ? g : 1);
}

}  // namespace

// used by the AMR kernels
fl_status hash_clear_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, hipStream_t s) {
    hipLaunchKernelGGL(k_clear, dim3(cap_grid(cap)), dim3(kBlock), 0, s, tk, tv, cap);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status hash_build_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, const uint64_t* keys,
                          const uint32_t* vals, size_t n, uint32_t* probes, hipStream_t s) {
    if (n == 0) return FL_OK;
    hipLaunchKernelGGL(k_build, dim3(cap_grid(n)), dim3(kBlock), 0, s, tk, tv, cap, keys, vals, n, probes);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

// Exact maps (value = smallest index holding the key).  Their lookups stop at
// the first empty slot, so the load must stay <= 0.6: with the capacity
// rule's 2^30 ceiling that bounds n by kMaxMapCells (which also keeps
// indices inside u32 below the kFiner / kInvalid markers).
template <bool kCellKeys>
fl_status map_build(uint64_t* tk, uint32_t* tv, uint64_t cap, const uint64_t* keys, size_t n, hipStream_t s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(n <= kMaxMapCells, "too many cells for the cell map (at most 644,245,094)");
    FL_REQUIRE((double)n <= 0.6 * (double)cap, "cell map capacity below n / 0.6");
    FL_REQUIRE(cap >= 8 && (reinterpret_cast<uintptr_t>(tk) & 15u) == 0, "cell map keys must be 16 B aligned");
    hipLaunchKernelGGL(k_cell_build<kCellKeys>, dim3(cap_grid(n)), dim3(kBlock), 0, s, tk, tv, cap, keys, n);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

// keys of any packing, per-key slots (the lattice layout's brick map)
fl_status cellmap_build_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, const uint64_t* keys, size_t n,
                             hipStream_t s) {
    return map_build<false>(tk, tv, cap, keys, n, s);
}

// packed AMR cell keys, 2x2x2-block groups (the AMR kernels' cell map)
fl_status cellkey_map_build_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, const uint64_t* keys, size_t n,
                                 hipStream_t s) {
    return map_build<true>(tk, tv, cap, keys, n, s);
}

}  // namespace fl

using namespace fl;

extern "C" {

// HashTableManager::computeCapacity (HashTableManager.cpp:47-57):
// 2^clamp(ceil(log2(n / 0.6)), 20, 30)
uint64_t fl_hash_capacity(size_t n) {
    double m = (double)n / (double)0.6f;
    int l = 0;
    while (l < 63 && std::ldexp(1.0, l) < m) ++l;
    if (l < 20) l = 20;
    if (l > 30) l = 30;
    return 1ull << l;
}

fl_status fl_hash_clear(uint64_t* d_tkeys, uint32_t* d_tvals, uint64_t capacity, fl_stream s) {
    FL_REQUIRE(d_tkeys && d_tvals, "null pointer");
    FL_REQUIRE(capacity && (capacity & (capacity - 1)) == 0, "capacity must be a power of two");
    return hash_clear_impl(d_tkeys, d_tvals, capacity, FL_STREAM(s));
}

fl_status fl_hash_build(uint64_t* d_tkeys, uint32_t* d_tvals, uint64_t capacity, const uint64_t* d_keys,
                        const uint32_t* d_vals, size_t n, uint32_t* d_probe_counts, fl_stream s) {
    FL_REQUIRE(d_tkeys && d_tvals, "null pointer");
    FL_REQUIRE(capacity && (capacity & (capacity - 1)) == 0, "capacity must be a power of two");
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_keys != nullptr, "null keys");
    return hash_build_impl(d_tkeys, d_tvals, capacity, d_keys, d_vals, n, d_probe_counts, FL_STREAM(s));
}

fl_status fl_hash_query(const uint64_t* d_tkeys, const uint32_t* d_tvals, uint64_t capacity,
                        const uint64_t* d_query, uint32_t* d_out, size_t n, uint32_t* d_probe_counts,
                        fl_stream s) {
    FL_REQUIRE(d_tkeys && d_tvals, "null pointer");
    FL_REQUIRE(capacity && (capacity & (capacity - 1)) == 0, "capacity must be a power of two");
    if (n == 0) return FL_OK;
    FL_REQUIRE(d_query && d_out, "null pointer");
    hipLaunchKernelGGL(k_query, dim3(cap_grid(n)), dim3(kBlock), 0, FL_STREAM(s), d_tkeys, d_tvals, capacity,
                       d_query, d_out, n, d_probe_counts);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

}  // extern "C"
