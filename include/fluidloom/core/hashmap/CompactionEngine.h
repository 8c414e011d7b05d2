// Compaction + Hilbert permutation (include/fluidloom/core/hashmap/CompactionEngine.h
// of the reference): filter (state != 0 && key != EMPTY), stable sort by key
// on the device, permutation[old] = new (HASH_INVALID_VALUE for dropped).
#pragma once
#include <memory>
#include <vector>

#include "fluidloom/core/backend/IBackend.h"
#include "fluidloom/core/hashmap/HashTable.h"
#include "fluidloom/core/hashmap/RadixSort.h"

namespace fluidloom {
namespace hashmap {

struct CompactionResult {
    size_t num_valid_cells;
    std::vector<uint32_t> permutation;
    double compaction_time_ms;
};

class CompactionEngine {
public:
    explicit CompactionEngine(IBackend* backend);
    ~CompactionEngine();
    CompactionEngine(const CompactionEngine&) = delete;
    CompactionEngine& operator=(const CompactionEngine&) = delete;

    CompactionResult compactAndPermute(const std::vector<uint64_t>& hilbert_indices,
                                       const std::vector<uint32_t>& cell_state);

    // Gathers each field (element_size bytes per cell) into permuted order in
    // place: new[perm[old]] = old for every kept cell.
    void applyPermutationToFields(const std::vector<void*>& field_device_ptrs, size_t element_size,
                                  size_t old_count, const std::vector<uint32_t>& permutation);

    size_t getPeakMemoryUsage() const { return peak_memory_usage_; }

private:
    IBackend* backend_;
    std::unique_ptr<RadixSort> sorter_;
    DeviceBufferPtr scratch_keys_;
    DeviceBufferPtr scratch_values_;
    size_t peak_memory_usage_ = 0;
    void ensureScratchAllocated(size_t count);
};

}  // namespace hashmap
}  // namespace fluidloom
