// Device-side tile scan with decoupled look-back, shared by the scan,
// compaction, split/merge and layout kernels so that predicate, scan and
// scatter fuse into one pass over the data.
//
// Workspace (u64 words): ws[0] = dynamic tile counter, ws[1 + t] = status of
// tile t: bits 62-63 flag (1 aggregate, 2 inclusive prefix), bits 0-61 value.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "wave.hpp"

namespace fl {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr uint64_t kScanAgg = 1ull << 62;
constexpr uint64_t kScanInc = 2ull << 62;
constexpr uint64_t kScanVal = (1ull << 62) - 1ull;

inline size_t scan_tiles(size_t n) { return (n + kScanTile - 1) / kScanTile; }
inline size_t scan_ws_words(size_t n) { return 2 + scan_tiles(n); }

struct ScanShared {
    uint64_t wsum[kScanThreads / 64];
    uint64_t prefix;
    uint32_t tile;
};

__device__ inline uint32_t scan_tile_id(ScanShared& sh, uint64_t* ws) {
    if (threadIdx.x == 0)
        sh.tile = (uint32_t)atomicAdd(reinterpret_cast<unsigned long long*>(ws), 1ull);
    __syncthreads();
    return sh.tile;
}

__device__ inline uint64_t wave_sum64(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Wave 0 publishes this tile's total and walks back over the earlier tiles'
// status words (decoupled look-back) to the tile's exclusive prefix, which
// it leaves in sh.prefix.  When `total` is non-null the last tile writes the
// grand total there.  Called by wave 0 only.
__device__ inline void tile_lookback(ScanShared& sh, uint64_t* ws, uint32_t tile, uint64_t tile_total,
                                     uint64_t* total, size_t n, size_t tile_elems = kScanTile) {
    const unsigned lane = lane_id();
    uint64_t* st = ws + 1;
    uint64_t excl = 0;
    if (tile == 0) {
        if (lane == 0) atomic_store_agent(&st[0], kScanInc | tile_total);
    } else {
        if (lane == 0) atomic_store_agent(&st[tile], kScanAgg | tile_total);
        int64_t j = (int64_t)tile - 1;
        for (;;) {
            int64_t idx = j - (int64_t)lane;
            uint64_t s = idx >= 0 ? atomic_load_agent(&st[idx]) : kScanInc;
            uint64_t f = s >> 62;
            uint64_t incm = __ballot(f == 2);
            uint64_t zerom = __ballot(f == 0);
            int first = incm ? (__ffsll((long long)incm) - 1) : 64;
            uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
            if (zerom & need) continue;
            uint64_t val = ((int)lane <= first) ? (s & kScanVal) : 0ull;
            excl += wave_sum64(val);
            if (first < 64) break;
            j -= 64;
        }
        if (lane == 0) atomic_store_agent(&st[tile], kScanInc | (excl + tile_total));
    }
    if (lane == 0) {
        sh.prefix = excl;
        if (total && (size_t)(tile + 1) * tile_elems >= n) *total = excl + tile_total;
    }
}

// Returns the exclusive prefix (over the whole array) of this thread's first
// element; `mine` is the thread's total.  When `total` is non-null the last
// tile writes the grand total there.  Contains __syncthreads().
__device__ inline uint64_t scan_tile_prefix(ScanShared& sh, uint64_t* ws, uint32_t tile, uint64_t mine,
                                            uint64_t* total, size_t n, size_t tile_elems = kScanTile) {
    const int w = threadIdx.x >> 6;
    const unsigned lane = lane_id();
    uint64_t inc = wave_inclusive_scan64(mine);
    if (lane == 63) sh.wsum[w] = inc;
    __syncthreads();
    uint64_t woff = 0, tile_total = 0;
#pragma unroll
    for (int i = 0; i < kScanThreads / 64; ++i) {
        uint64_t s = sh.wsum[i];
        if (i < w) woff += s;
        tile_total += s;
    }
    if (w == 0) tile_lookback(sh, ws, tile, tile_total, total, n, tile_elems);
    __syncthreads();
    return sh.prefix + woff + inc - mine;
}

// Striped tiles: element j of thread t is tile * kScanTile + j * kScanThreads
// + t, so every load and store of a round j is coalesced.  Given a thread's
// predicate bits (bit j = element j kept), writes rel[j] = position of each
// kept element among the tile's kept elements (in element order) and returns
// the tile's exclusive prefix over the whole array.  When `total` is
// non-null the last tile writes the grand total there.  Contains
// __syncthreads().
struct StripedShared {
    ScanShared sc;
    uint32_t cnt[kScanItems][kScanThreads / 64];
};

__device__ inline uint64_t striped_positions(StripedShared& sh, uint64_t* ws, uint32_t tile, uint32_t pmask,
                                             uint64_t* total, size_t n, uint32_t* rel) {
    const int w = threadIdx.x >> 6;
    const unsigned lane = lane_id();
    const uint64_t below = lanemask_lt();
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t b = __ballot((pmask >> j) & 1u);
        rel[j] = (uint32_t)__popcll(b & below);
        if (lane == 0) sh.cnt[j][w] = (uint32_t)__popcll(b);
    }
    __syncthreads();
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        uint32_t before = 0, all = 0;
#pragma unroll
        for (int v = 0; v < kScanThreads / 64; ++v) {
            const uint32_t c = sh.cnt[j][v];
            if (v < w) before += c;
            all += c;
        }
        rel[j] += run + before;
        run += all;
    }
    if (w == 0) tile_lookback(sh.sc, ws, tile, run, total, n);
    __syncthreads();
    return sh.sc.prefix;
}

}  // namespace fl
