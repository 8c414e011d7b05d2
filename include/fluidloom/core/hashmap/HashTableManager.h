// GPU hash table lifecycle (include/fluidloom/core/hashmap/HashTableManager.h
// of the reference, whose rebuild/query are stubs there): capacity rule
// 2^clamp(ceil(log2(n/0.6)),20,30), device clear/build/query.
#pragma once
#include <memory>
#include <vector>

#include "fluidloom/core/backend/IBackend.h"
#include "fluidloom/core/hashmap/CompactionEngine.h"
#include "fluidloom/core/hashmap/HashTable.h"

namespace fluidloom {
namespace hashmap {

class HashTableManager {
public:
    explicit HashTableManager(IBackend* backend);
    ~HashTableManager();
    HashTableManager(const HashTableManager&) = delete;
    HashTableManager& operator=(const HashTableManager&) = delete;

    double rebuild(const std::vector<uint64_t>& hilbert_indices, const std::vector<uint32_t>& array_indices);
    std::vector<uint32_t> query(const std::vector<uint64_t>& query_keys);

    void* getKeysDevicePtr() const { return table_keys_ ? table_keys_->getDevicePointer() : nullptr; }
    void* getValuesDevicePtr() const { return table_values_ ? table_values_->getDevicePointer() : nullptr; }
    uint64_t getCapacity() const { return current_table_.capacity; }
    const HashTable& getHashTable() const { return current_table_; }
    const HashTableMetadata& getMetadata() const { return metadata_; }

    static uint64_t computeCapacity(size_t num_cells);

private:
    IBackend* backend_;
    std::unique_ptr<CompactionEngine> compactor_;
    HashTable current_table_{};
    HashTableMetadata metadata_{};
    DeviceBufferPtr table_keys_;
    DeviceBufferPtr table_values_;
    void allocateTable(size_t num_cells);
    void clearTable();
};

}  // namespace hashmap
}  // namespace fluidloom
