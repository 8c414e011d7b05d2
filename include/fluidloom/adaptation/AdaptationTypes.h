// Results and configuration of the adaptation engines
// (include/fluidloom/adaptation/AdaptationTypes.h:27-157 of the reference).
// OpenCL handles are replaced by backend-neutral ones: fl_context (unused,
// kept for signature parity), fl_stream (hipStream_t), fl_mem (device
// pointer), fl_event (hipEvent_t, owned by the result).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "fluidloom/adaptation/CellDescriptor.h"

namespace fluidloom {
using fl_context = void*;
using fl_queue = void*;
using fl_mem = void*;
using fl_device_event = void*;

namespace adaptation {

void releaseEvent(fl_device_event e);

struct SplitResult {
    bool success = false;
    size_t num_children = 0;
    size_t num_parents_split = 0;
    std::vector<CellDescriptor> children;
    std::vector<uint32_t> parent_to_child_map;
    std::vector<uint32_t> split_parent_indices;
    std::vector<float> child_fields;  // [child][ncomp] copies of the parent values
    fl_device_event event = nullptr;
    size_t device_memory_used = 0;
    SplitResult() = default;
    SplitResult(const SplitResult&) = delete;
    SplitResult& operator=(const SplitResult&) = delete;
    SplitResult(SplitResult&& o) noexcept { *this = std::move(o); }
    SplitResult& operator=(SplitResult&& o) noexcept;
    ~SplitResult() { releaseEvent(event); }
};

struct MergeResult {
    bool success = false;
    size_t num_parents_created = 0;
    size_t num_children_merged = 0;
    std::vector<CellDescriptor> parents;
    std::vector<float> averaged_fields;
    std::unordered_set<size_t> merged_child_indices;
    std::vector<uint32_t> group_to_parent_map;
    fl_device_event event = nullptr;
    double mass_before_merge = 0.0;
    double mass_after_merge = 0.0;
    double conservation_error = 0.0;
    MergeResult() = default;
    MergeResult(const MergeResult&) = delete;
    MergeResult& operator=(const MergeResult&) = delete;
    MergeResult(MergeResult&& o) noexcept { *this = std::move(o); }
    MergeResult& operator=(MergeResult&& o) noexcept;
    ~MergeResult() { releaseEvent(event); }
};

struct BalanceResult {
    bool converged = false;
    uint32_t iterations = 0;
    uint32_t total_violations_detected = 0;
    uint32_t total_cells_marked_for_balance = 0;
    size_t final_cell_count = 0;
    double max_level_difference = 0.0;
    struct IterationStats {
        uint32_t violations_this_iter = 0;
        uint32_t cells_marked_this_iter = 0;
        double time_ms = 0.0;
    };
    std::vector<IterationStats> per_iteration_stats;
};

enum class BalanceMode : int { REFERENCE_TEST_POINTS = 0, D3Q19_NEIGHBOUR_CUBES = 1 };

struct AdaptationConfig {
    uint8_t max_refinement_level = 8;
    uint8_t min_refinement_level = 0;
    size_t max_cells_per_split_batch = 65536;
    size_t max_cells_per_merge_batch = 32768;
    uint32_t max_balance_iterations = 10;
    bool enforce_2_1_balance = true;
    bool cascade_refinement = true;
    uint8_t cascade_depth = 2;
    bool allow_modifying_geometry_moving = false;
    std::string default_averaging_rule = "arithmetic";
    bool validate_conservation = true;
    double conservation_tolerance = 0.001;
    double buffer_growth_factor = 1.5;
    size_t initial_buffer_capacity = 1048576;
    // build additions: balance rule, finest-unit level of the reference
    // rule, level-0 extents, merge balance check
    BalanceMode balance_mode = BalanceMode::REFERENCE_TEST_POINTS;
    uint8_t anchor_level = 8;
    int32_t root_extent[3] = {1, 1, 1};
    bool merge_balance_check = false;

    void validate() const {
        if (max_refinement_level > 19) throw std::invalid_argument("max_refinement_level cannot exceed 19");
        if (min_refinement_level > max_refinement_level)
            throw std::invalid_argument("min_refinement_level must be <= max_refinement_level");
        if (cascade_depth > max_refinement_level)
            throw std::invalid_argument("cascade_depth cannot exceed max_refinement_level");
    }
};

}  // namespace adaptation
}  // namespace fluidloom
