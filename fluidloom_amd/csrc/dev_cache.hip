// Cache of device blocks for the lattice layout builds and the solvers'
// tables and population buffers.  A layout build takes about a hundred
// temporaries and a regrid frees and allocates every table of a level, so
// hipMalloc / hipFree (the latter waits for the whole device) would cost
// tens of milliseconds per regrid.  Freed blocks are kept per device, after
// their stream has drained, and handed to the next request that fits them;
// block sizes above 1 MiB are rounded up to eight classes per octave so
// that slightly different sizes from one regrid to the next reuse blocks.
// The cache holds at most an eighth of the device's memory (32 GiB at most);
// fl_release_cached_memory() returns every cached block to HIP, and an
// allocation that fails releases the cache once and tries again.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <unordered_map>

#include "fl_internal.hpp"

namespace fl {
namespace {

struct Block {
    size_t bytes;
    int device;
};

struct Cache {
    std::mutex mu;
    std::unordered_map<void*, Block> live;             // handed out
    std::map<int, std::multimap<size_t, void*>> free;  // per device: size -> block
    std::map<int, size_t> cached, cap;                 // per device
};

Cache& cache() {
    static Cache* c = new Cache();  // never destroyed: blocks may be freed at exit
    return *c;
}

size_t round_bytes(size_t b) {
    if (b <= 256) return 256;
    if (b <= (1u << 20)) return (b + 255) & ~size_t(255);
    int e = 63 - __builtin_clzll(b);  // 2^e <= b < 2^(e+1)
    const size_t step = size_t(1) << (e - 3);
    return (b + step - 1) & ~(step - 1);
}

// the cap of the current device
size_t device_cap() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || tot == 0) return size_t(4) << 30;
    const size_t c = tot / 8;
    return c < (size_t(32) << 30) ? c : (size_t(32) << 30);
}

// hipFree of every cached block of `dev` (the caller holds the lock and has
// made `dev` current).
void release_locked(Cache& c, int dev) {
    auto it = c.free.find(dev);
    if (it == c.free.end()) return;
    for (auto& kv : it->second) (void)hipFree(kv.second);
    it->second.clear();
    c.cached[dev] = 0;
}

}  // namespace

fl_status cache_alloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return FL_OK;
    int dev = 0;
    FL_HIP(hipGetDevice(&dev));
    const size_t rb = round_bytes(bytes);
    Cache& c = cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto& fl = c.free[dev];
    auto it = fl.lower_bound(rb);
    if (it != fl.end() && it->first <= rb + rb / 8) {
        *p = it->second;
        c.live[*p] = Block{it->first, dev};
        c.cached[dev] -= it->first;
        fl.erase(it);
        return FL_OK;
    }
    hipError_t e = hipMalloc(p, rb);
    if (e == hipErrorOutOfMemory && c.cached[dev]) {
        (void)hipGetLastError();
        release_locked(c, dev);
        e = hipMalloc(p, rb);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    c.live[*p] = Block{rb, dev};
    return FL_OK;
}

void cache_free(void* p, hipStream_t s) {
    if (!p) return;
    Cache& c = cache();
    std::lock_guard<std::mutex> g(c.mu);
    auto it = c.live.find(p);
    if (it == c.live.end()) return;
    const Block b = it->second;
    c.live.erase(it);
    int cur = 0;
    (void)hipGetDevice(&cur);
    // the block's last user may still run on s; the next user can be on
    // another stream, so the block is cached only once s has drained
    if (cur != b.device || hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    if (!c.cap.count(b.device)) c.cap[b.device] = device_cap();
    if (c.cached[b.device] + b.bytes > c.cap[b.device]) {
        (void)hipFree(p);
        return;
    }
    c.free[b.device].emplace(b.bytes, p);
    c.cached[b.device] += b.bytes;
}

}  // namespace fl

extern "C" fl_status fl_release_cached_memory(void) {
    fl::Cache& c = fl::cache();
    std::lock_guard<std::mutex> g(c.mu);
    int cur = 0;
    FL_HIP(hipGetDevice(&cur));
    for (auto& kv : c.free) {
        if (kv.second.empty()) continue;
        FL_HIP(hipSetDevice(kv.first));
        fl::release_locked(c, kv.first);
    }
    FL_HIP(hipSetDevice(cur));
    return FL_OK;
}

extern "C" fl_status fl_cached_memory_bytes(uint64_t* bytes) {
    FL_REQUIRE(bytes, "null output");
    fl::Cache& c = fl::cache();
    std::lock_guard<std::mutex> g(c.mu);
    uint64_t t = 0;
    for (auto& kv : c.cached) t += kv.second;
    *bytes = t;
    return FL_OK;
}
