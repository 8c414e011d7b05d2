// SoA pack-buffer layout (include/fluidloom/halo/PackBufferLayout.h:18-61 of
// the reference): field-major then component-major blocks, each block
// max_cells = capacity_bytes / cell_size_bytes long.
#pragma once
#include <cstddef>
#include <string>
#include <vector>

namespace fluidloom {
namespace halo {

struct PackBufferLayout {
    size_t capacity_bytes{0};
    size_t used_bytes{0};
    struct FieldLayout {
        std::string field_name;
        size_t num_components{0};
        size_t bytes_per_component{0};
        size_t offset_in_cell{0};  // bytes of the fields before this one in a cell
    };
    std::vector<FieldLayout> fields;
    size_t cell_size_bytes{0};

    size_t maxCells() const { return cell_size_bytes ? capacity_bytes / cell_size_bytes : 0; }

    // Byte offset of (field, component, cell) in the pack buffer.
    size_t getOffset(size_t field_idx, size_t component_idx, size_t cell_idx) const {
        if (field_idx >= fields.size()) return 0;
        const size_t m = maxCells();
        if (m == 0) return 0;
        const FieldLayout& f = fields[field_idx];
        return f.offset_in_cell * m + component_idx * f.bytes_per_component * m + cell_idx * f.bytes_per_component;
    }

    bool validate() const {
        if (capacity_bytes == 0 || used_bytes > capacity_bytes || cell_size_bytes == 0) return false;
        for (const auto& f : fields)
            if (f.bytes_per_component % 4 != 0) return false;
        return true;
    }

    void addField(const std::string& name, size_t ncomp, size_t bytes_per_comp) {
        fields.push_back(FieldLayout{name, ncomp, bytes_per_comp, cell_size_bytes});
        cell_size_bytes += ncomp * bytes_per_comp;
    }
};

}  // namespace halo
}  // namespace fluidloom
