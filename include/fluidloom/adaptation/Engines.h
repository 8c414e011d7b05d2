// Adaptation engines over device SoA cell arrays (SplitEngine.h,
// MergeEngine.h, BalanceEnforcer.h and AdaptationEngine.h of the reference,
// one header here).  Method names, argument order and results follow the
// reference; every stage runs as HIP kernels with no host-side hash tables
// or prefix sums.
#pragma once
#include "fluidloom/adaptation/AdaptationTypes.h"

namespace fluidloom {
namespace adaptation {

class SplitEngine {
public:
    SplitEngine(fl_context context, fl_queue queue, const AdaptationConfig& config);
    ~SplitEngine() = default;
    SplitResult split(fl_mem parent_x, fl_mem parent_y, fl_mem parent_z, fl_mem parent_level,
                      fl_mem parent_states, fl_mem refine_flags, fl_mem parent_material_id, size_t num_parents,
                      fl_mem parent_fields = nullptr, uint32_t num_field_components = 0);

private:
    fl_queue m_queue;
    AdaptationConfig m_config;
};

class MergeEngine {
public:
    MergeEngine(fl_context context, fl_queue queue, const AdaptationConfig& config);
    ~MergeEngine() = default;
    MergeResult merge(fl_mem child_x, fl_mem child_y, fl_mem child_z, fl_mem child_level, fl_mem child_states,
                      fl_mem refine_flags, fl_mem child_material_id, size_t num_children,
                      fl_mem child_fields = nullptr, uint32_t num_field_components = 0);

private:
    fl_queue m_queue;
    AdaptationConfig m_config;
};

class BalanceEnforcer {
public:
    BalanceEnforcer(fl_context context, fl_queue queue, const AdaptationConfig& config);
    ~BalanceEnforcer() = default;
    BalanceResult enforce(fl_mem coord_x, fl_mem coord_y, fl_mem coord_z, fl_mem levels, fl_mem cell_states,
                          fl_mem refine_flags, size_t num_cells);

private:
    fl_queue m_queue;
    AdaptationConfig m_config;
};

// balance -> merge -> split -> compaction in one device pass (fl_regrid).
// Mutates the caller's handle slots: on return they point to new arrays
// (the old ones are freed), *num_cells and *capacity are updated.  Output
// order: survivors, children (parent order), merged parents (leader order).
class AdaptationEngine {
public:
    AdaptationEngine(fl_context context, fl_queue queue, const AdaptationConfig& config);
    ~AdaptationEngine() = default;
    fl_device_event adapt(fl_mem* coord_x, fl_mem* coord_y, fl_mem* coord_z, fl_mem* levels, fl_mem* cell_states,
                          fl_mem* refine_flags, fl_mem* material_id, size_t* num_cells, size_t* capacity,
                          fl_mem* fields = nullptr, uint32_t num_field_components = 0);
    const BalanceResult& lastBalance() const { return m_last_balance; }

private:
    fl_queue m_queue;
    AdaptationConfig m_config;
    BalanceResult m_last_balance;
};

}  // namespace adaptation
}  // namespace fluidloom
