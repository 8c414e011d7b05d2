// Field descriptors (include/fluidloom/core/fields/FieldDescriptor.h:10-91 of
// the reference): same enums, validation (name 1-64 chars, 1-32 components,
// halo depth <= 2) and FNV-1a 64-bit ids of the name.
#pragma once
#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace fluidloom {
namespace fields {

enum class FieldType : uint8_t { FLOAT32 = 0, FLOAT64 = 1, INT32 = 2, INT64 = 3, UINT8 = 4, FLOAT16 = 5 };
enum class AveragingRule : uint8_t { ARITHMETIC = 0, VOLUME_WEIGHTED = 1, MINIMUM = 2, MAXIMUM = 3, SKIP = 4 };
enum class SolidScheme : uint8_t { ZERO = 0, BOUNCE_BACK = 1, NEUMANN = 2, SKIP = 3, CUSTOM = 4 };

// 64-bit FNV-1a of a string (offset basis 0xcbf29ce484222325, prime 0x100000001b3).
uint64_t fnv1a_64(const std::string& s);
size_t fieldTypeSize(FieldType t);

struct alignas(64) FieldDescriptor {
    uint64_t id{};
    std::string name;
    FieldType type{FieldType::FLOAT32};
    uint16_t num_components{1};
    uint8_t halo_depth{0};
    AveragingRule avg_rule{AveragingRule::ARITHMETIC};
    SolidScheme solid_scheme{SolidScheme::ZERO};
    std::vector<double> default_value;
    bool is_visualization_field{false};
    bool is_mask{false};
    bool is_material{false};
    struct {
        uint64_t version{0};
        size_t bytes_allocated{0};
        void* device_ptr{nullptr};
        bool is_dirty{false};
    } gpu_state;

    FieldDescriptor() = default;
    FieldDescriptor(const std::string& field_name, FieldType field_type, uint16_t components, uint8_t halo = 0,
                    AveragingRule avg = AveragingRule::ARITHMETIC, SolidScheme solid = SolidScheme::ZERO);
    size_t bytesPerCell() const;
    bool isValid() const;
};

struct FieldHandle {
    uint64_t id;
    explicit FieldHandle(uint64_t field_id) : id(field_id) {}
    bool operator==(const FieldHandle& o) const { return id == o.id; }
    bool operator<(const FieldHandle& o) const { return id < o.id; }
};

}  // namespace fields
}  // namespace fluidloom

namespace std {
template <>
struct hash<fluidloom::fields::FieldHandle> {
    size_t operator()(const fluidloom::fields::FieldHandle& h) const noexcept { return std::hash<uint64_t>{}(h.id); }
};
}  // namespace std
