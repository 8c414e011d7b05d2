// Per-field device storage (include/fluidloom/core/fields/SOAFieldManager.h of
// the reference), true structure-of-arrays: component c of a field with n
// cells starts at base + c * pitch where pitch = n * type size rounded up to
// 256 bytes (the reference hands out base + c * bytesPerCell, which is not a
// component array).  getAllocationSize() returns the cell count as the
// reference's tests expect; versions/dirty flags behave as there.
#pragma once
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "fluidloom/core/backend/IBackend.h"
#include "fluidloom/core/fields/FieldDescriptor.h"

namespace fluidloom {
namespace fields {

class SOAFieldManager {
public:
    explicit SOAFieldManager(IBackend* backend);
    SOAFieldManager(const SOAFieldManager&) = delete;
    SOAFieldManager& operator=(const SOAFieldManager&) = delete;

    FieldHandle allocate(const FieldDescriptor& desc, size_t num_cells);
    void resize(FieldHandle handle, size_t new_num_cells);
    void deallocate(FieldHandle handle);
    void* getDevicePtr(FieldHandle handle, uint16_t component = 0) const;
    const FieldDescriptor& getDescriptor(FieldHandle handle) const;
    size_t getAllocationSize(FieldHandle handle) const;
    size_t getComponentPitch(FieldHandle handle) const;
    void markDirty(FieldHandle handle);
    void markClean(FieldHandle handle);
    bool isDirty(FieldHandle handle) const;
    uint64_t getVersion(FieldHandle handle) const;
    size_t getTotalMemoryUsage() const;
    size_t getMemoryUsage(FieldHandle handle) const;

private:
    struct FieldState {
        FieldDescriptor descriptor;
        size_t num_cells{0};
        std::vector<void*> device_ptrs;
        DeviceBufferPtr device_buffer;
        uint64_t version{1};
        size_t pitch{0};
    };
    IBackend* backend_;
    std::unordered_map<uint64_t, FieldState> fields_;
    mutable std::mutex mutex_;
    FieldState& state(FieldHandle h);
    const FieldState& state(FieldHandle h) const;
    static size_t pitchFor(const FieldDescriptor& d, size_t n);
};

}  // namespace fields
}  // namespace fluidloom
