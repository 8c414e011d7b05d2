// a4: stable LSD radix sort of (u64 key, u32 value) pairs for gfx950.
//
// One upfront histogram pass computes the digit histograms of every pass
// (LDS-privatised, 8-bit digits), one tiny kernel scans them and flags the
// passes whose digit is constant (skipped).  Each remaining pass is a single
// "onesweep" kernel: a workgroup takes the next 4096-key tile (dynamic tile
// id, so look-back always waits on running workgroups), ranks keys stably
// with 64-lane ballot matching (wave-striped layout: wave w owns 1024
// consecutive keys, item j of lane l is key w*1024 + 64*j + l, so ballot
// order == index order), publishes its per-digit counts, resolves the global
// digit offsets by decoupled look-back over the predecessors' counts, sorts
// the tile by digit in LDS and writes runs of equal digits contiguously.
// The reference (kernels/hashmap/radix_sort.cl:77-105) scatters with
// atomic_inc and is unstable; this sort is stable.
#include "fl_internal.hpp"
#include "wave.hpp"

namespace fl {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;

struct SortLayout {
    size_t hist, base, triv, ctr, status, total;
};

SortLayout sort_layout(size_t n) {
    SortLayout L;
    size_t ntiles = (n + kTile - 1) / kTile;
    L.hist = 0;
    L.base = L.hist + 8 * 256 * 4;
    L.triv = L.base + 8 * 256 * 4;
    L.ctr = L.triv + 8 * 4;
    L.status = (L.ctr + 8 * 4 + 255) & ~size_t(255);
    L.total = L.status + 8 * ntiles * 256 * 4;
    return L;
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// Upfront digit histograms of all passes in one read of the keys.  Each
// wave counts into its own LDS copy (fewer conflicting LDS atomics), each
// thread takes kHistItems keys of a block tile with all loads issued first.
// Lanes holding the same digit as the wave's first active lane are counted
// with one ballot and added by one lane: a digit shared by most of the wave
// (the zero low digits of coarse cells' keys, a constant high digit) would
// otherwise serialise 64 atomics on one LDS word.
constexpr int kHistItems = 8;

__global__ __launch_bounds__(kThreads) void k_hist(const uint64_t* __restrict__ keys, size_t n,
                                                   int begin_bit, int end_bit, int npass,
                                                   uint32_t* __restrict__ ghist) {
    __shared__ uint32_t h[kWaves][8 * 256];
    for (int i = threadIdx.x; i < kWaves * 8 * 256; i += kThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    uint32_t* hw = h[threadIdx.x >> 6];
    const size_t tile = (size_t)kThreads * kHistItems;
    for (size_t t0 = (size_t)blockIdx.x * tile; t0 < n; t0 += (size_t)gridDim.x * tile) {
        uint64_t k[kHistItems];
#pragma unroll
        for (int j = 0; j < kHistItems; ++j) {
            const size_t i = t0 + (size_t)j * kThreads + threadIdx.x;
            k[j] = i < n ? keys[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < kHistItems; ++j) {
            if (t0 + (size_t)j * kThreads + threadIdx.x >= n) continue;
            for (int p = 0; p < npass; ++p) {
                const int b = begin_bit + 8 * p;
                const int nb = end_bit - b < 8 ? end_bit - b : 8;
                const unsigned d = (unsigned)((k[j] >> b) & ((1ull << nb) - 1ull));
                const unsigned d0 = __builtin_amdgcn_readfirstlane(d);
                const uint64_t same = __ballot(d == d0);
                if (d != d0)
                    atomicAdd(&hw[p * 256 + d], 1u);
                else if (lane_id() == (unsigned)__builtin_ctzll(same))
                    atomicAdd(&hw[p * 256 + d0], (uint32_t)__builtin_popcountll(same));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npass * 256; i += kThreads) {
        uint32_t v = 0;
        for (int w = 0; w < kWaves; ++w) v += h[w][i];
        if (v) atomicAdd(&ghist[i], v);
    }
}

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ uint32_t block_excl_scan256(uint32_t v, uint32_t* wsum, uint32_t* total) {
    uint32_t inc = wave_inclusive_scan(v);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int i = 0; i < kWaves; ++i) {
        uint32_t s = wsum[i];
        if (i < w) off += s;
        tot += s;
    }
    __syncthreads();
    if (total) *total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kThreads) void k_scan_hist(const uint32_t* __restrict__ hist,
                                                        uint32_t* __restrict__ base,
                                                        uint32_t* __restrict__ triv, int npass,
                                                        uint32_t n) {
    __shared__ uint32_t wsum[kWaves];
    for (int p = 0; p < npass; ++p) {
        uint32_t c = hist[p * 256 + threadIdx.x];
        uint32_t e = block_excl_scan256(c, wsum, nullptr);
        base[p * 256 + threadIdx.x] = e;
        if (threadIdx.x == 0) triv[p] = 0;
        __syncthreads();
        if (c == n) triv[p] = 1;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kThreads) void k_onesweep(const uint64_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       uint64_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout, size_t n,
                                                       int shift, uint32_t mask,
                                                       const uint32_t* __restrict__ gbase,
                                                       uint32_t* __restrict__ counter,
                                                       uint32_t* __restrict__ status) {
    __shared__ uint64_t skey[kTile];
    __shared__ uint32_t sval[kTile];
    __shared__ uint32_t whist[kWaves][256];
    __shared__ uint32_t tstart[256];
    __shared__ uint32_t gofs[256];
    __shared__ uint32_t wsum[kWaves];
    __shared__ uint32_t s_tile;

    const int t = threadIdx.x;
    const int w = t >> 6;
    const unsigned lane = lane_id();
    if (t == 0) s_tile = atomicAdd(counter, 1u);
    for (int i = t; i < kWaves * 256; i += kThreads) (&whist[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const size_t tbase = (size_t)tile * kTile;
    const size_t wbase = tbase + (size_t)w * 64 * kItems;

    uint64_t k[kItems];
    uint32_t v[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        size_t idx = wbase + (size_t)j * 64 + lane;
        bool ok = idx < n;
        k[j] = ok ? kin[idx] : 0ull;
        v[j] = ok ? vin[idx] : 0u;
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        size_t idx = wbase + (size_t)j * 64 + lane;
        bool ok = idx < n;
        uint64_t act = __ballot(ok);
        unsigned d = (unsigned)(k[j] >> shift) & mask;
        uint64_t peers = match_u8(d, act);
        uint32_t before = ok ? whist[w][d] : 0u;
        rank[j] = before + (uint32_t)__popcll(peers & lt);
        if (ok && (peers & lt) == 0) whist[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();

    // per-digit counts of this tile and per-wave offsets
    uint32_t run = 0;
    for (int i = 0; i < kWaves; ++i) {
        uint32_t c = whist[i][t];
        whist[i][t] = run;
        run += c;
    }
    uint32_t start = block_excl_scan256(run, wsum, nullptr);
    tstart[t] = start;

    // decoupled look-back on this digit
    uint32_t* st = status + (size_t)tile * 256;
    uint32_t excl = 0;
    if (tile == 0) {
        atomic_store_agent(&st[t], kFlagInc | run);
    } else {
        atomic_store_agent(&st[t], kFlagAgg | run);
        int j = (int)tile - 1;
        while (j >= 0) {
            uint32_t s = atomic_load_agent(&status[(size_t)j * 256 + t]);
            uint32_t f = s & ~kCountMask;
            if (f == 0) continue;
            excl += s & kCountMask;
            if (f == kFlagInc) break;
            --j;
        }
        atomic_store_agent(&st[t], kFlagInc | (excl + run));
    }
    gofs[t] = gbase[t] + excl;
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        size_t idx = wbase + (size_t)j * 64 + lane;
        if (idx < n) {
            unsigned d = (unsigned)(k[j] >> shift) & mask;
            uint32_t pos = tstart[d] + whist[w][d] + rank[j];
            skey[pos] = k[j];
            sval[pos] = v[j];
        }
    }
    __syncthreads();
    const uint32_t nvalid = (uint32_t)((n - tbase) < (size_t)kTile ? (n - tbase) : (size_t)kTile);
    for (uint32_t p = t; p < nvalid; p += kThreads) {
        uint64_t key = skey[p];
        unsigned d = (unsigned)(key >> shift) & mask;
        uint32_t o = gofs[d] + (p - tstart[d]);
        kout[o] = key;
        vout[o] = sval[p];
    }
}

__global__ void k_iota(uint32_t* __restrict__ v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = (uint32_t)i;
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

size_t fl_sort_temp_bytes(size_t n) { return sort_layout(n).total; }

fl_status fl_iota_u32(uint32_t* d_vals, size_t n, fl_stream s) {
    if (n == 0) return FL_OK;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 256) > 4096 ? 4096 : grid_for(n, 256)), dim3(256), 0,
                       FL_STREAM(s), d_vals, n);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_sort_pairs_u64u32(uint64_t* d_keys, uint32_t* d_vals, uint64_t* d_keys_alt,
                               uint32_t* d_vals_alt, size_t n, int begin_bit, int end_bit,
                               void* d_temp, size_t temp_bytes, fl_stream s) {
    if (n < 2) return FL_OK;
    FL_REQUIRE(d_keys && d_vals && d_keys_alt && d_vals_alt && d_temp, "null pointer");
    FL_REQUIRE(n <= kCountMask, "n must be below 2^30");
    FL_REQUIRE(begin_bit >= 0 && begin_bit < end_bit && end_bit <= 64, "bad bit range");
    SortLayout L = sort_layout(n);
    FL_REQUIRE(temp_bytes >= L.total, "temp buffer too small (see fl_sort_temp_bytes)");
    hipStream_t st = FL_STREAM(s);
    char* tb = static_cast<char*>(d_temp);
    uint32_t* hist = reinterpret_cast<uint32_t*>(tb + L.hist);
    uint32_t* base = reinterpret_cast<uint32_t*>(tb + L.base);
    uint32_t* triv = reinterpret_cast<uint32_t*>(tb + L.triv);
    uint32_t* ctr = reinterpret_cast<uint32_t*>(tb + L.ctr);
    uint32_t* status = reinterpret_cast<uint32_t*>(tb + L.status);
    const int npass = (end_bit - begin_bit + 7) / 8;
    const size_t ntiles = (n + kTile - 1) / kTile;

    size_t zero_words = L.total / 4;
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(zero_words, 256) > 4096 ? 4096 : grid_for(zero_words, 256)),
                       dim3(256), 0, st, reinterpret_cast<uint32_t*>(tb), zero_words);
    FL_CHECK_LAUNCH();
    unsigned hg = grid_for(n, kThreads * kHistItems);
    if (hg > 8192) hg = 8192;
    hipLaunchKernelGGL(k_hist, dim3(hg), dim3(kThreads), 0, st, d_keys, n, begin_bit, end_bit, npass, hist);
    FL_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan_hist, dim3(1), dim3(kThreads), 0, st, hist, base, triv, npass, (uint32_t)n);
    FL_CHECK_LAUNCH();
    uint32_t htriv[8] = {0};
    FL_HIP(hipMemcpyAsync(htriv, triv, sizeof(uint32_t) * npass, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));

    uint64_t* ck = d_keys; uint32_t* cv = d_vals;
    uint64_t* ak = d_keys_alt; uint32_t* av = d_vals_alt;
    for (int p = 0; p < npass; ++p) {
        if (htriv[p]) continue;
        int b = begin_bit + 8 * p;
        int nb = end_bit - b < 8 ? end_bit - b : 8;
        uint32_t mask = (uint32_t)((1u << nb) - 1u);
        hipLaunchKernelGGL(k_onesweep, dim3((unsigned)ntiles), dim3(kThreads), 0, st, ck, cv, ak, av, n, b,
                           mask, base + p * 256, ctr + p, status + (size_t)p * ntiles * 256);
        FL_CHECK_LAUNCH();
        std::swap(ck, ak);
        std::swap(cv, av);
    }
    if (ck != d_keys) {
        FL_HIP(hipMemcpyAsync(d_keys, ck, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
        FL_HIP(hipMemcpyAsync(d_vals, cv, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
    return FL_OK;
}

}  // extern "C"
