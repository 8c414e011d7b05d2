// a5: single-pass exclusive scan of u32 with decoupled look-back.
//
// Replaces the reference's three-launch Blelloch scan
// (src/adaptation/kernels/compact_cells.cl:61-125), which has no carry above
// its second level and is wrong for n > 65,536.  A workgroup takes the next
// 4096-element tile (dynamic id), scans 16 consecutive elements per thread,
// scans the 256 thread totals with wave shuffles, publishes its tile total
// and resolves its prefix by look-back over 64-bit (flag, value) words.
#include <mutex>

#include "fl_internal.hpp"
#include "scan_device.hpp"
#include "wave.hpp"

namespace fl {
namespace {

__global__ void k_zero64(uint64_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// 64 elements per thread, 16K per tile: the look-back's inclusive prefix
// advances about one 64-tile window per round trip, so with small tiles the
// tile count, not the bytes, bounds the scan (64M elements: 0.27 ms with
// 4K-element tiles, 0.20 ms with these).
constexpr int kScanRun = 64;
constexpr size_t kScanRunTile = (size_t)kScanThreads * kScanRun;
inline size_t scan_run_tiles(size_t n) { return (n + kScanRunTile - 1) / kScanRunTile; }

__global__ __launch_bounds__(kScanThreads) void k_scan(const uint32_t* __restrict__ in,
                                                       uint32_t* __restrict__ out,
                                                       uint64_t* __restrict__ total, size_t n,
                                                       uint64_t* __restrict__ ws) {
    __shared__ ScanShared sh;
    const uint32_t tile = scan_tile_id(sh, ws);
    const size_t base = (size_t)tile * kScanRunTile + (size_t)threadIdx.x * kScanRun;
    // a thread's run is 256 B: with both arrays 16 B aligned (the caller may
    // pass any pointer) it moves as sixteen 16 B accesses
    const bool vec = base + kScanRun <= n && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    uint32_t v[kScanRun];
    if (vec) {
        const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int q = 0; q < kScanRun / 4; ++q) {
            const uint4 u = p[q];
            v[4 * q] = u.x;
            v[4 * q + 1] = u.y;
            v[4 * q + 2] = u.z;
            v[4 * q + 3] = u.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < kScanRun; ++j) {
            size_t i = base + j;
            v[j] = i < n ? in[i] : 0u;
        }
    }
    uint64_t acc = 0;
#pragma unroll
    for (int j = 0; j < kScanRun; ++j) acc += v[j];
    uint64_t excl = scan_tile_prefix(sh, ws, tile, acc, total, n, kScanRunTile);
#pragma unroll
    for (int j = 0; j < kScanRun; ++j) {
        const uint32_t x = v[j];
        v[j] = (uint32_t)excl;
        excl += x;
    }
    if (vec) {
        uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int q = 0; q < kScanRun / 4; ++q) p[q] = make_uint4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    } else {
#pragma unroll
        for (int j = 0; j < kScanRun; ++j) {
            size_t i = base + j;
            if (i < n) out[i] = v[j];
        }
    }
}

// Clears `words` workspace words (the tile counter and the tile status
// words must start at zero), then runs the scan: two kernel launches, no
// memset, so a captured stream records only kernel nodes.
fl_status scan_launch(const uint32_t* d_in, uint32_t* d_out, uint64_t* d_total, size_t n, uint64_t* ws,
                      size_t words, hipStream_t st) {
    hipLaunchKernelGGL(k_zero64, dim3(grid_for(words, 256) > 2048 ? 2048 : grid_for(words, 256)), dim3(256), 0,
                       st, ws, words);
    FL_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan, dim3(scan_run_tiles(n)), dim3(kScanThreads), 0, st, d_in, d_out, d_total, n, ws);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

// ---- allocation-free variant ----------------------------------------------
// Its look-back workspace (tile counter + one status word per 16K-element
// tile) lives in the library's own device array below, so the call needs no
// allocation and no caller workspace.  Bounded at n <= 2^31 elements:
// 2 + 2^31 / 16384 = 131,074 words (1 MiB) per device, held until the
// process ends.  Every call on a device shares that device's array, so
// calls that overlap on one device (two streams, two threads) race.
constexpr size_t kNoAllocMaxN = size_t(1) << 31;
constexpr size_t kNoAllocWords = 2 + kNoAllocMaxN / kScanRunTile;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch_ptr[kMaxDevices] = {};

}  // namespace
}  // namespace fl

static __device__ __align__(256) uint64_t scratch[fl::kNoAllocWords];

namespace fl {
namespace {

// A zero-element scan's total, written by a kernel (no memset node).
__global__ void k_store_zero(uint64_t* __restrict__ p) {
    if (threadIdx.x == 0) *p = 0ull;
}

}  // namespace

// Device address of this device's copy of `scratch`, looked up once per
// device with that device current (host code never passes the array's name
// itself: on the host it is only a stand-in for the symbol).
fl_status noalloc_scratch_setup(int device, uint64_t** out) {
    FL_REQUIRE(device >= 0 && device < kMaxDevices, "device id out of range for the scan scratch");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        FL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FL_REQUIRE(p != nullptr, "hipGetSymbolAddress returned no address for the scan scratch");
        g_scratch_ptr[device] = static_cast<uint64_t*>(p);
    }
    if (out) *out = g_scratch_ptr[device];
    return FL_OK;
}

}  // namespace fl

using namespace fl;

extern "C" {

size_t fl_scan_temp_bytes(size_t n) { return scan_ws_words(n) * sizeof(uint64_t); }

fl_status fl_scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t* d_total, size_t n,
                                void* d_temp, size_t temp_bytes, fl_stream s) {
    hipStream_t st = FL_STREAM(s);
    if (n == 0) {
        if (d_total) FL_HIP(hipMemsetAsync(d_total, 0, sizeof(uint64_t), st));
        return FL_OK;
    }
    FL_REQUIRE(d_in && d_out && d_temp, "null pointer");
    size_t words = scan_ws_words(n);
    FL_REQUIRE(temp_bytes >= words * sizeof(uint64_t), "temp buffer too small (see fl_scan_temp_bytes)");
    return scan_launch(d_in, d_out, d_total, n, static_cast<uint64_t*>(d_temp), words, st);
}

size_t fl_scan_noalloc_max_elements(void) { return kNoAllocMaxN; }

fl_status fl_scan_exclusive_u32_noalloc(const uint32_t* d_in, uint32_t* d_out, uint64_t* d_total, size_t n,
                                        fl_stream s) {
    FL_REQUIRE(n <= kNoAllocMaxN, "n exceeds the allocation-free scan's limit of 2^31 elements");
    int dev = 0;
    FL_HIP(hipGetDevice(&dev));
    uint64_t* ws = nullptr;
    fl_status r = noalloc_scratch_setup(dev, &ws);
    if (r) return r;
    hipStream_t st = FL_STREAM(s);
    if (n == 0) {
        if (d_total) {
            hipLaunchKernelGGL(k_store_zero, dim3(1), dim3(64), 0, st, d_total);
            FL_CHECK_LAUNCH();
        }
        return FL_OK;
    }
    FL_REQUIRE(d_in && d_out, "null pointer");
    const size_t words = 2 + scan_run_tiles(n);
    return scan_launch(d_in, d_out, d_total, n, ws, words, st);
}

}  // extern "C"
