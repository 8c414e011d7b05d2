// Internal helpers shared by every translation unit of libfluidloom_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "fluidloom_hip.h"

namespace fl {

// Records the message returned by fl_last_error_string() for this thread.
void set_error(const std::string& msg);
// Records a HIP failure; returns FL_E_HIP (or FL_E_OOM for out-of-memory).
int hip_fail(hipError_t e, const char* what, const char* file, int line);

// Device blocks from the library's block cache (dev_cache.hip) on the
// current device.  cache_free keeps the block with s, the stream of its last
// user; cache_alloc for stream s hands it out at once when it was last used
// on s, else after that stream has passed the free (a null s waits on the
// host).
fl_status cache_alloc(void** p, size_t bytes, hipStream_t s);
void cache_free(void* p, hipStream_t s);

constexpr uint64_t kEmptyKey = 0xFFFFFFFFFFFFFFFFull;
constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint32_t kFiner = 0xFFFFFFFEu;
// Largest cell count an exact cell map holds at load <= 0.6 within the
// capacity rule's 2^30-slot ceiling: floor(0.6 * 2^30).
constexpr size_t kMaxMapCells = 644245094;

inline unsigned grid_for(size_t n, unsigned block) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g == 0 ? 1 : g);
}

}  // namespace fl

#define FL_HIP(call)                                                        \
    do {                                                                    \
        hipError_t fl_e_ = (call);                                          \
        if (fl_e_ != hipSuccess) return fl::hip_fail(fl_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define FL_CHECK_LAUNCH() FL_HIP(hipGetLastError())

#define FL_TRY(x)                \
    do {                         \
        fl_status fl_r_ = (x);   \
        if (fl_r_) return fl_r_; \
    } while (0)

#define FL_REQUIRE(cond, msg)                                               \
    do {                                                                    \
        if (!(cond)) {                                                      \
            fl::set_error(std::string(__func__) + ": " + (msg));            \
            return FL_E_INVALID;                                            \
        }                                                                   \
    } while (0)

#define FL_STREAM(s) (reinterpret_cast<hipStream_t>(s))
