// Cache of device blocks for the lattice layout builds and the solvers'
// tables and population buffers.  A layout build takes about a hundred
// temporaries and a regrid frees and allocates every table of a level, so
// hipMalloc / hipFree (the latter waits for the whole device) would cost
// tens of milliseconds per regrid.  A freed block is kept per device with
// the stream of its last user and an event recorded there: a request on the
// same stream reuses it at once (stream order keeps the new work behind the
// old), a request from another stream first waits for that event.  Waiting
// for the stream at every free instead cost a regrid of the 76M-leaf sphere
// about a thousand synchronisations (10-16 ms).  Block sizes above 1 MiB
// are rounded up to eight classes per octave so that slightly different
// sizes from one regrid to the next reuse blocks.  The cache holds at most
// an eighth of the device's memory (32 GiB at most);
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

// A cached block: the stream of its last user and an event recorded there
// at the free (null when the free waited for the stream instead).
struct FreeBlock {
    void* p;
    hipStream_t s;
    hipEvent_t ev;
};

struct Cache {
    std::mutex mu;
    std::unordered_map<void*, Block> live;                 // handed out
    std::map<int, std::multimap<size_t, FreeBlock>> free;  // per device: size -> block
    std::map<int, std::vector<hipEvent_t>> events;         // per device: spare events
    std::map<int, size_t> cached, cap;                     // per device
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
// made `dev` current; hipFree waits for the device, so pending users finish
// first).
void release_locked(Cache& c, int dev) {
    auto it = c.free.find(dev);
    if (it == c.free.end()) return;
    for (auto& kv : it->second) {
        (void)hipFree(kv.second.p);
        if (kv.second.ev) c.events[dev].push_back(kv.second.ev);
    }
    it->second.clear();
    c.cached[dev] = 0;
}

}  // namespace

fl_status cache_alloc(void** p, size_t bytes, hipStream_t s) {
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
        // among blocks of this size, one last used on s needs no wait
        auto pick = it;
        for (auto j = it; j != fl.end() && j->first == it->first; ++j)
            if (j->second.s == s) {
                pick = j;
                break;
            }
        const FreeBlock fb = pick->second;
        const size_t sz = pick->first;
        if (fb.ev) {
            if (fb.s != s) {
                const hipError_t e = s ? hipStreamWaitEvent(s, fb.ev, 0) : hipEventSynchronize(fb.ev);
                if (e != hipSuccess) return hip_fail(e, "cache_alloc wait", __FILE__, __LINE__);
            }
            c.events[dev].push_back(fb.ev);
        }
        *p = fb.p;
        c.live[*p] = Block{sz, dev};
        c.cached[dev] -= sz;
        fl.erase(pick);
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
    if (cur != b.device) {
        (void)hipFree(p);
        return;
    }
    if (!c.cap.count(b.device)) c.cap[b.device] = device_cap();
    if (c.cached[b.device] + b.bytes > c.cap[b.device]) {
        (void)hipFree(p);
        return;
    }
    // the block's last user may still run on s: an event marks the point
    // after which it is free
    hipEvent_t ev = nullptr;
    auto& spare = c.events[b.device];
    if (!spare.empty()) {
        ev = spare.back();
        spare.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        ev = nullptr;
    }
    if (ev && hipEventRecord(ev, s) != hipSuccess) {
        spare.push_back(ev);
        ev = nullptr;
    }
    if (!ev && hipStreamSynchronize(s) != hipSuccess) {
        (void)hipFree(p);
        return;
    }
    c.free[b.device].emplace(b.bytes, FreeBlock{p, s, ev});
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
