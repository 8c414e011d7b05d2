// Stable 64-bit key / 32-bit value sort on the device (the reference's
// RadixSort, include/fluidloom/core/hashmap/RadixSort.h:14-60, whose driver
// is a no-op in the reference).  sortByKey sorts in place; the buffers keep
// their identity.
#pragma once
#include <cstdint>

#include "fluidloom/core/backend/IBackend.h"

namespace fluidloom {
namespace hashmap {

class RadixSort {
public:
    explicit RadixSort(IBackend* backend);
    ~RadixSort();
    RadixSort(const RadixSort&) = delete;
    RadixSort& operator=(const RadixSort&) = delete;

    // Returns the sort time in milliseconds (device events).
    double sortByKey(DeviceBufferPtr& keys_buffer, DeviceBufferPtr& values_buffer, size_t count);

    struct Stats {
        double total_time_ms;
        double histogram_time_ms;
        double prefix_sum_time_ms;
        double scatter_time_ms;
    };
    const Stats& getLastStats() const { return last_stats_; }

private:
    IBackend* backend_;
    DeviceBufferPtr temp_keys_;
    DeviceBufferPtr temp_values_;
    DeviceBufferPtr histogram_buffer_;
    Stats last_stats_{};
};

}  // namespace hashmap
}  // namespace fluidloom
