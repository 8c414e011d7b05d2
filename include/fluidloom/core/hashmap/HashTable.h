// Hash table constants and descriptor (include/fluidloom/core/hashmap/HashTable.h:10-61
// of the reference; values unchanged).
#pragma once
#include <cstddef>
#include <cstdint>

namespace fluidloom {
namespace hashmap {

static constexpr uint64_t HASH_EMPTY_KEY = 0xFFFFFFFFFFFFFFFFULL;
static constexpr uint32_t HASH_INVALID_VALUE = 0xFFFFFFFFU;
static constexpr float MAX_LOAD_FACTOR = 0.6f;
static constexpr uint32_t INITIAL_CAPACITY_POW2 = 20;
static constexpr size_t HASH_TABLE_ALIGNMENT = 64;

struct alignas(HASH_TABLE_ALIGNMENT) HashTable {
    uint64_t capacity;
    uint64_t size;
    uint32_t log2_capacity;
    void* keys_device_ptr;
    void* values_device_ptr;

    static inline uint64_t hashKey(uint64_t key) { return key * 0x9e3779b97f4a7c15ULL; }
    inline uint64_t getSlot(uint64_t key) const { return hashKey(key) & (capacity - 1); }
};

struct HashTableMetadata {
    size_t num_cells;
    size_t active_capacity;
    size_t bytes_allocated;
    uint32_t build_time_ms;
    float average_probe_count;
    uint32_t max_probe_count;
};

struct HashBuildConfig {
    static constexpr uint32_t MAX_PROBE_LIMIT = 32;
    static constexpr bool ENABLE_STATS = true;
};

}  // namespace hashmap
}  // namespace fluidloom
