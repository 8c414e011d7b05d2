// Process-wide field registry (include/fluidloom/core/registry/FieldRegistry.h
// of the reference): by name and by id under a shared mutex.
#pragma once
#include <optional>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "fluidloom/core/fields/FieldDescriptor.h"

namespace fluidloom {
namespace registry {

class FieldRegistry {
public:
    static FieldRegistry& instance();
    FieldRegistry(const FieldRegistry&) = delete;
    FieldRegistry& operator=(const FieldRegistry&) = delete;
    bool registerField(const fields::FieldDescriptor& desc);
    std::optional<fields::FieldDescriptor> lookupByName(const std::string& name) const;
    std::optional<fields::FieldDescriptor> lookupById(fields::FieldHandle handle) const;
    bool exists(const std::string& name) const;
    std::vector<std::string> getAllNames() const;
    void clear();

private:
    FieldRegistry() = default;
    mutable std::shared_mutex mutex_;
    std::unordered_map<std::string, fields::FieldDescriptor> by_name_;
    std::unordered_map<uint64_t, fields::FieldDescriptor> by_id_;
};

}  // namespace registry
}  // namespace fluidloom
