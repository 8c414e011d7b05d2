// C++ API: RadixSort, CompactionEngine, HashTableManager on the device.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>

#include "fluidloom/common/FluidLoomError.h"
#include "fluidloom/core/backend/HIPBackend.h"
#include "fluidloom/core/hashmap/HashTableManager.h"
#include "fluidloom_hip.h"

namespace fluidloom {
namespace hashmap {
namespace {

fl_stream stream_of(IBackend* b) {
    if (b->getType() != BackendType::HIP) FL_THROW(BackendError, "device hash map requires the HIP backend");
    return static_cast<HIPBackend*>(b)->stream();
}

double elapsed_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// ------------------------------------------------------------------ sort
RadixSort::RadixSort(IBackend* backend) : backend_(backend) {
    if (!backend) throw std::invalid_argument("Backend must not be null");
}
RadixSort::~RadixSort() = default;

double RadixSort::sortByKey(DeviceBufferPtr& keys, DeviceBufferPtr& values, size_t count) {
    if (!keys || !values) FL_THROW(BackendError, "null buffer");
    if (keys->getSize() < count * 8 || values->getSize() < count * 4) FL_THROW(BackendError, "buffers too small");
    fl_stream s = stream_of(backend_);
    auto t0 = std::chrono::steady_clock::now();
    if (count >= 2) {
        if (!temp_keys_ || temp_keys_->getSize() < count * 8) {
            temp_keys_ = backend_->allocateBuffer(count * 8);
            temp_values_ = backend_->allocateBuffer(count * 4);
        }
        size_t tb = fl_sort_temp_bytes(count);
        if (!histogram_buffer_ || histogram_buffer_->getSize() < tb) histogram_buffer_ = backend_->allocateBuffer(tb);
        FL_CHECK_STATUS(fl_sort_pairs_u64u32(static_cast<uint64_t*>(keys->getDevicePointer()),
                                             static_cast<uint32_t*>(values->getDevicePointer()),
                                             static_cast<uint64_t*>(temp_keys_->getDevicePointer()),
                                             static_cast<uint32_t*>(temp_values_->getDevicePointer()), count, 0, 64,
                                             histogram_buffer_->getDevicePointer(), tb, s));
        FL_CHECK_STATUS(fl_stream_sync(s));
    }
    double ms = elapsed_ms(t0);
    last_stats_ = Stats{ms, 0.0, 0.0, ms};
    return ms;
}

// ------------------------------------------------------------------ compaction
CompactionEngine::CompactionEngine(IBackend* backend)
    : backend_(backend), sorter_(std::make_unique<RadixSort>(backend)) {
    if (!backend) throw std::invalid_argument("Backend must not be null");
}
CompactionEngine::~CompactionEngine() = default;

void CompactionEngine::ensureScratchAllocated(size_t count) {
    if (scratch_keys_ && scratch_keys_->getSize() >= count * 8) return;
    scratch_keys_ = backend_->allocateBuffer(count * 8);
    scratch_values_ = backend_->allocateBuffer(count * 4);
    peak_memory_usage_ = count * 12;
}

CompactionResult CompactionEngine::compactAndPermute(const std::vector<uint64_t>& keys,
                                                     const std::vector<uint32_t>& state) {
    if (keys.size() != state.size()) throw std::invalid_argument("Mismatched input sizes");
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = keys.size();
    // active cells (state != 0 && key != EMPTY), CompactionEngine.cpp:46-55
    std::vector<uint64_t> vk;
    std::vector<uint32_t> vi;
    vk.reserve(n);
    vi.reserve(n);
    for (size_t i = 0; i < n; ++i)
        if (state[i] != 0 && keys[i] != HASH_EMPTY_KEY) {
            vk.push_back(keys[i]);
            vi.push_back((uint32_t)i);
        }
    const size_t m = vk.size();
    std::vector<uint32_t> perm(n, HASH_INVALID_VALUE);
    if (m == 0) return CompactionResult{0, std::move(perm), elapsed_ms(t0)};
    ensureScratchAllocated(m);
    backend_->copyHostToDevice(vk.data(), *scratch_keys_, m * 8);
    backend_->copyHostToDevice(vi.data(), *scratch_values_, m * 4);
    sorter_->sortByKey(scratch_keys_, scratch_values_, m);
    std::vector<uint32_t> sorted(m);
    backend_->copyDeviceToHost(*scratch_values_, sorted.data(), m * 4);
    for (uint32_t k = 0; k < m; ++k) perm[sorted[k]] = k;
    return CompactionResult{m, std::move(perm), elapsed_ms(t0)};
}

void CompactionEngine::applyPermutationToFields(const std::vector<void*>& ptrs, size_t elem, size_t old_count,
                                                const std::vector<uint32_t>& perm) {
    if (perm.size() != old_count) throw std::invalid_argument("permutation size mismatch");
    if (elem != 4) FL_THROW(BackendError, "applyPermutationToFields supports 4-byte elements");
    fl_stream s = stream_of(backend_);
    std::vector<uint32_t> src, dst;
    for (uint32_t i = 0; i < old_count; ++i)
        if (perm[i] != HASH_INVALID_VALUE) {
            src.push_back(i);
            dst.push_back(perm[i]);
        }
    if (src.empty()) return;
    auto dsrc = backend_->allocateBuffer(src.size() * 4, src.data());
    auto ddst = backend_->allocateBuffer(dst.size() * 4, dst.data());
    auto pack = backend_->allocateBuffer(src.size() * 4);
    for (void* p : ptrs) {
        FL_CHECK_STATUS(fl_pack_fields(static_cast<const float*>(p), static_cast<const uint32_t*>(dsrc->getDevicePointer()),
                                       src.size(), static_cast<uint8_t*>(pack->getDevicePointer()), 0, s));
        FL_CHECK_STATUS(fl_unpack_fields(static_cast<const uint8_t*>(pack->getDevicePointer()), 0,
                                         static_cast<const uint32_t*>(ddst->getDevicePointer()), dst.size(),
                                         static_cast<float*>(p), s));
    }
    FL_CHECK_STATUS(fl_stream_sync(s));
}

// ------------------------------------------------------------------ hash table
HashTableManager::HashTableManager(IBackend* backend)
    : backend_(backend), compactor_(std::make_unique<CompactionEngine>(backend)) {
    if (!backend) throw std::invalid_argument("Backend must not be null");
}
HashTableManager::~HashTableManager() = default;

uint64_t HashTableManager::computeCapacity(size_t num_cells) { return fl_hash_capacity(num_cells); }

void HashTableManager::allocateTable(size_t n) {
    uint64_t cap = computeCapacity(n);
    if (cap == current_table_.capacity && table_keys_) return;
    table_keys_ = backend_->allocateBuffer(cap * 8);
    table_values_ = backend_->allocateBuffer(cap * 4);
    current_table_.capacity = cap;
    uint32_t l = 0;
    while ((1ull << l) < cap) ++l;
    current_table_.log2_capacity = l;
    current_table_.keys_device_ptr = table_keys_->getDevicePointer();
    current_table_.values_device_ptr = table_values_->getDevicePointer();
    metadata_.bytes_allocated = cap * 12;
    metadata_.active_capacity = (size_t)(cap * MAX_LOAD_FACTOR);
}

void HashTableManager::clearTable() {
    if (!table_keys_) return;
    FL_CHECK_STATUS(fl_hash_clear(static_cast<uint64_t*>(table_keys_->getDevicePointer()),
                                  static_cast<uint32_t*>(table_values_->getDevicePointer()), current_table_.capacity,
                                  stream_of(backend_)));
    current_table_.size = 0;
}

double HashTableManager::rebuild(const std::vector<uint64_t>& keys, const std::vector<uint32_t>& vals) {
    if (keys.size() != vals.size()) throw std::invalid_argument("Mismatched input sizes");
    auto t0 = std::chrono::steady_clock::now();
    const size_t n = keys.size();
    allocateTable(n);
    clearTable();
    fl_stream s = stream_of(backend_);
    uint32_t max_probe = 0;
    double avg_probe = 0.0;
    if (n) {
        auto dk = backend_->allocateBuffer(n * 8, keys.data());
        auto dv = backend_->allocateBuffer(n * 4, vals.data());
        auto dp = backend_->allocateBuffer(n * 4);
        FL_CHECK_STATUS(fl_hash_build(static_cast<uint64_t*>(table_keys_->getDevicePointer()),
                                      static_cast<uint32_t*>(table_values_->getDevicePointer()),
                                      current_table_.capacity, static_cast<const uint64_t*>(dk->getDevicePointer()),
                                      static_cast<const uint32_t*>(dv->getDevicePointer()), n,
                                      static_cast<uint32_t*>(dp->getDevicePointer()), s));
        std::vector<uint32_t> probes(n);
        backend_->copyDeviceToHost(*dp, probes.data(), n * 4);
        uint64_t sum = 0;
        for (uint32_t p : probes) {
            sum += p;
            max_probe = std::max(max_probe, p);
        }
        avg_probe = (double)sum / (double)n;
    }
    FL_CHECK_STATUS(fl_stream_sync(s));
    double ms = elapsed_ms(t0);
    current_table_.size = n;
    metadata_.num_cells = n;
    metadata_.build_time_ms = (uint32_t)ms;
    metadata_.average_probe_count = (float)avg_probe;
    metadata_.max_probe_count = max_probe;
    return ms;
}

std::vector<uint32_t> HashTableManager::query(const std::vector<uint64_t>& q) {
    std::vector<uint32_t> out(q.size(), HASH_INVALID_VALUE);
    if (q.empty() || !table_keys_) return out;
    fl_stream s = stream_of(backend_);
    auto dq = backend_->allocateBuffer(q.size() * 8, q.data());
    auto dr = backend_->allocateBuffer(q.size() * 4);
    FL_CHECK_STATUS(fl_hash_query(static_cast<const uint64_t*>(table_keys_->getDevicePointer()),
                                  static_cast<const uint32_t*>(table_values_->getDevicePointer()),
                                  current_table_.capacity, static_cast<const uint64_t*>(dq->getDevicePointer()),
                                  static_cast<uint32_t*>(dr->getDevicePointer()), q.size(), nullptr, s));
    backend_->copyDeviceToHost(*dr, out.data(), q.size() * 4);
    return out;
}

}  // namespace hashmap
}  // namespace fluidloom
