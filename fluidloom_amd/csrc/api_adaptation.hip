// C++ API: SplitEngine, MergeEngine, BalanceEnforcer, AdaptationEngine over
// fl_split / fl_merge / fl_balance / fl_regrid.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "fluidloom/adaptation/Engines.h"
#include "fluidloom/common/FluidLoomError.h"
#include "fluidloom_hip.h"

namespace fluidloom {
namespace adaptation {
namespace {

struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    explicit DevBuf(size_t n) {
        if (n) FL_CHECK_STATUS(fl_malloc(&p, n));
    }
    ~DevBuf() {
        if (p) (void)fl_free(p);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void* release() {
        void* q = p;
        p = nullptr;
        return q;
    }
};

struct CellArrays {
    DevBuf x, y, z, level, state, flag, material;
    explicit CellArrays(size_t n)
        : x(n * 4), y(n * 4), z(n * 4), level(n), state(n), flag(n * 4), material(n * 4) {}
    fl_cells view() {
        return fl_cells{(int32_t*)x.p, (int32_t*)y.p, (int32_t*)z.p, (uint8_t*)level.p, (uint8_t*)state.p,
                        (int32_t*)flag.p, (uint32_t*)material.p};
    }
};

template <typename T>
std::vector<T> d2h(const void* d, size_t n, fl_stream s) {
    std::vector<T> v(n);
    if (n) FL_CHECK_STATUS(fl_memcpy_d2h(v.data(), d, n * sizeof(T), s));
    return v;
}

std::vector<CellDescriptor> cells_to_host(fl_cells c, size_t n, fl_stream s) {
    auto x = d2h<int32_t>(c.x, n, s), y = d2h<int32_t>(c.y, n, s), z = d2h<int32_t>(c.z, n, s);
    auto l = d2h<uint8_t>(c.level, n, s), st = d2h<uint8_t>(c.state, n, s);
    auto m = d2h<uint32_t>(c.material, n, s);
    std::vector<CellDescriptor> out(n);
    for (size_t i = 0; i < n; ++i) {
        out[i].x = x[i]; out[i].y = y[i]; out[i].z = z[i];
        out[i].level = l[i]; out[i].state = st[i]; out[i].material_id = m[i];
    }
    return out;
}

fl_device_event record(fl_stream s) {
    fl_event e = nullptr;
    FL_CHECK_STATUS(fl_event_create(&e));
    FL_CHECK_STATUS(fl_event_record(e, s));
    return e;
}

fl_regrid_config make_cfg(const AdaptationConfig& c) {
    fl_regrid_config r{};
    r.balance_mode = c.enforce_2_1_balance ? (int)c.balance_mode : -1;
    r.anchor_level = c.anchor_level;
    r.max_split_level = c.max_refinement_level;
    r.max_balance_iters = (int)c.max_balance_iterations;
    for (int a = 0; a < 3; ++a) r.root[a] = c.root_extent[a];
    r.merge_balance_check = c.merge_balance_check ? 1 : 0;
    r.ncomp = 0;
    r.avg_rule = c.default_averaging_rule == "arithmetic" ? 0 : 1;
    return r;
}

fl_cells cells_of(fl_mem x, fl_mem y, fl_mem z, fl_mem l, fl_mem s, fl_mem f, fl_mem m) {
    return fl_cells{(int32_t*)x, (int32_t*)y, (int32_t*)z, (uint8_t*)l, (uint8_t*)s, (int32_t*)f, (uint32_t*)m};
}

}  // namespace

void releaseEvent(fl_device_event e) {
    if (e) (void)fl_event_destroy(e);
}

SplitResult& SplitResult::operator=(SplitResult&& o) noexcept {
    if (this != &o) {
        releaseEvent(event);
        success = o.success; num_children = o.num_children; num_parents_split = o.num_parents_split;
        children = std::move(o.children); parent_to_child_map = std::move(o.parent_to_child_map);
        split_parent_indices = std::move(o.split_parent_indices); child_fields = std::move(o.child_fields);
        event = o.event; o.event = nullptr; device_memory_used = o.device_memory_used;
    }
    return *this;
}

MergeResult& MergeResult::operator=(MergeResult&& o) noexcept {
    if (this != &o) {
        releaseEvent(event);
        success = o.success; num_parents_created = o.num_parents_created; num_children_merged = o.num_children_merged;
        parents = std::move(o.parents); averaged_fields = std::move(o.averaged_fields);
        merged_child_indices = std::move(o.merged_child_indices); group_to_parent_map = std::move(o.group_to_parent_map);
        event = o.event; o.event = nullptr; mass_before_merge = o.mass_before_merge;
        mass_after_merge = o.mass_after_merge; conservation_error = o.conservation_error;
    }
    return *this;
}

// ------------------------------------------------------------------ split
SplitEngine::SplitEngine(fl_context, fl_queue queue, const AdaptationConfig& config) : m_queue(queue), m_config(config) {
    m_config.validate();
}

SplitResult SplitEngine::split(fl_mem px, fl_mem py, fl_mem pz, fl_mem pl, fl_mem ps, fl_mem rf, fl_mem pm, size_t n,
                               fl_mem pfields, uint32_t ncomp) {
    SplitResult res;
    fl_cells c = cells_of(px, py, pz, pl, ps, rf, pm);
    DevBuf cs(std::max<size_t>(n, 1) * 4);
    size_t tb = fl_regrid_temp_bytes(n);
    DevBuf tmp(tb);
    uint64_t ns = 0;
    fl_cells none{};
    int st = fl_split(&c, n, m_config.max_refinement_level, (uint32_t*)cs.p, &none, 0, tmp.p, tb, &ns, m_queue);
    if (st != FL_OK && st != FL_E_CAPACITY) FL_CHECK_STATUS(st);
    CellArrays ch(std::max<uint64_t>(8 * ns, 1));
    fl_cells cv = ch.view();
    if (ns) FL_CHECK_STATUS(fl_split(&c, n, m_config.max_refinement_level, (uint32_t*)cs.p, &cv, 8 * ns, tmp.p, tb, &ns,
                                     m_queue));
    res.num_parents_split = ns;
    res.num_children = 8 * ns;
    res.children = cells_to_host(cv, 8 * ns, m_queue);
    res.parent_to_child_map = d2h<uint32_t>(cs.p, n, m_queue);
    for (size_t i = 0; i < n; ++i)
        if (res.parent_to_child_map[i] != INVALID_INDEX) res.split_parent_indices.push_back((uint32_t)i);
    if (pfields && ncomp && ns) {
        std::vector<uint32_t> parent(8 * ns);
        for (size_t k = 0; k < res.split_parent_indices.size(); ++k)
            for (int q = 0; q < 8; ++q) parent[8 * k + q] = res.split_parent_indices[k];
        DevBuf dpar(parent.size() * 4), dout(parent.size() * ncomp * 4);
        FL_CHECK_STATUS(fl_memcpy_h2d(dpar.p, parent.data(), parent.size() * 4, m_queue));
        FL_CHECK_STATUS(fl_ghost_fill_c2f((const float*)pfields, (const uint32_t*)dpar.p, parent.size(), (int)ncomp,
                                          (float*)dout.p, m_queue));
        res.child_fields = d2h<float>(dout.p, parent.size() * ncomp, m_queue);
    }
    res.device_memory_used = 8 * ns * 22 + n * 4;
    res.success = true;
    res.event = record(m_queue);
    return res;
}

// ------------------------------------------------------------------ merge
MergeEngine::MergeEngine(fl_context, fl_queue queue, const AdaptationConfig& config) : m_queue(queue), m_config(config) {
    m_config.validate();
}

MergeResult MergeEngine::merge(fl_mem cx, fl_mem cy, fl_mem cz, fl_mem cl, fl_mem cst, fl_mem rf, fl_mem cm, size_t n,
                               fl_mem cfields, uint32_t ncomp) {
    MergeResult res;
    fl_cells c = cells_of(cx, cy, cz, cl, cst, rf, cm);
    DevBuf gid(std::max<size_t>(n, 1) * 4), lead(std::max<size_t>(n, 1) * 4);
    size_t cap = n / 8 + 1;
    CellArrays par(cap);
    fl_cells pv = par.view();
    size_t tb = fl_regrid_temp_bytes(n);
    DevBuf tmp(tb);
    uint64_t ng = 0;
    FL_CHECK_STATUS(fl_merge(&c, n, (uint32_t*)gid.p, (uint32_t*)lead.p, &pv, cap, tmp.p, tb, &ng, m_queue));
    res.num_parents_created = ng;
    res.num_children_merged = 8 * ng;
    res.parents = cells_to_host(pv, ng, m_queue);
    auto g = d2h<uint32_t>(gid.p, n, m_queue);
    for (size_t i = 0; i < n; ++i)
        if (g[i] != INVALID_INDEX) res.merged_child_indices.insert(i);
    res.group_to_parent_map.resize(ng);
    for (size_t k = 0; k < ng; ++k) res.group_to_parent_map[k] = (uint32_t)k;
    if (cfields && ncomp && ng) {
        auto x = d2h<int32_t>(cx, n, m_queue), y = d2h<int32_t>(cy, n, m_queue), z = d2h<int32_t>(cz, n, m_queue);
        std::vector<uint32_t> child8(8 * ng);
        for (size_t i = 0; i < n; ++i)
            if (g[i] != INVALID_INDEX) child8[8 * g[i] + ((x[i] & 1) | ((y[i] & 1) << 1) | ((z[i] & 1) << 2))] = (uint32_t)i;
        DevBuf d8(child8.size() * 4), dout(ng * ncomp * 4);
        FL_CHECK_STATUS(fl_memcpy_h2d(d8.p, child8.data(), child8.size() * 4, m_queue));
        int rule = m_config.default_averaging_rule == "arithmetic" ? 0 : 1;
        FL_CHECK_STATUS(fl_ghost_fill_f2c((const float*)cfields, (const uint32_t*)d8.p, ng, (int)ncomp, rule,
                                          (float*)dout.p, m_queue));
        res.averaged_fields = d2h<float>(dout.p, ng * ncomp, m_queue);
        auto f = d2h<float>(cfields, n * ncomp, m_queue);
        double before = 0, after = 0;
        for (size_t i = 0; i < n; ++i)
            if (g[i] != INVALID_INDEX) before += f[i * ncomp] * (rule == 0 ? 0.125 : 1.0);
        for (size_t k = 0; k < ng; ++k) after += res.averaged_fields[k * ncomp];
        res.mass_before_merge = before;
        res.mass_after_merge = after;
        res.conservation_error = std::abs(after - before);
    }
    res.success = true;
    res.event = record(m_queue);
    return res;
}

// ------------------------------------------------------------------ balance
BalanceEnforcer::BalanceEnforcer(fl_context, fl_queue queue, const AdaptationConfig& config)
    : m_queue(queue), m_config(config) {
    m_config.validate();
}

BalanceResult BalanceEnforcer::enforce(fl_mem x, fl_mem y, fl_mem z, fl_mem levels, fl_mem states, fl_mem flags,
                                       size_t n) {
    BalanceResult res;
    if (n == 0) return res;
    fl_cells c = cells_of(x, y, z, levels, states, flags, nullptr);
    fl_regrid_config cfg = make_cfg(m_config);
    if (cfg.balance_mode < 0) cfg.balance_mode = (int)m_config.balance_mode;
    size_t tb = fl_regrid_temp_bytes(n);
    DevBuf tmp(tb);
    fl_regrid_result r{};
    FL_CHECK_STATUS(fl_balance(&cfg, &c, n, tmp.p, tb, &r, m_queue));
    res.converged = r.balance_converged != 0;
    res.iterations = (uint32_t)r.balance_iterations;
    res.total_violations_detected = (uint32_t)r.balance_violations;
    res.total_cells_marked_for_balance = (uint32_t)r.balance_marked;
    res.final_cell_count = n;
    return res;
}

// ------------------------------------------------------------------ engine
AdaptationEngine::AdaptationEngine(fl_context, fl_queue queue, const AdaptationConfig& config)
    : m_queue(queue), m_config(config) {
    m_config.validate();
}

fl_device_event AdaptationEngine::adapt(fl_mem* cx, fl_mem* cy, fl_mem* cz, fl_mem* lv, fl_mem* st, fl_mem* rf,
                                        fl_mem* mat, size_t* num_cells, size_t* capacity, fl_mem* fields,
                                        uint32_t ncomp) {
    const size_t n = *num_cells;
    fl_cells c = cells_of(*cx, *cy, *cz, *lv, *st, *rf, *mat);
    fl_regrid_config cfg = make_cfg(m_config);
    size_t tb = fl_regrid_temp_bytes(n);
    DevBuf tmp(tb);
    m_last_balance = BalanceResult{};
    if (cfg.balance_mode >= 0 && n) {
        fl_regrid_result r{};
        FL_CHECK_STATUS(fl_balance(&cfg, &c, n, tmp.p, tb, &r, m_queue));
        m_last_balance.converged = r.balance_converged != 0;
        m_last_balance.iterations = (uint32_t)r.balance_iterations;
        m_last_balance.total_violations_detected = (uint32_t)r.balance_violations;
        m_last_balance.total_cells_marked_for_balance = (uint32_t)r.balance_marked;
    }
    // count splits (dry run) to size the output
    uint64_t ns = 0;
    fl_cells none{};
    int s0 = fl_split(&c, n, cfg.max_split_level, nullptr, &none, 0, tmp.p, tb, &ns, m_queue);
    if (s0 != FL_OK && s0 != FL_E_CAPACITY) FL_CHECK_STATUS(s0);
    size_t need = n + 7 * ns;
    size_t newcap = std::max(*capacity, need);
    if (need > *capacity) newcap = std::max(need, (size_t)(need * m_config.buffer_growth_factor));
    CellArrays out(std::max<size_t>(newcap, 1));
    fl_cells ov = out.view();
    DevBuf fout(fields && ncomp ? newcap * ncomp * 4 : 0);
    cfg.balance_mode = -1;
    cfg.ncomp = fields && ncomp ? (int)ncomp : 0;
    fl_regrid_result r{};
    FL_CHECK_STATUS(fl_regrid(&cfg, &c, n, &ov, newcap, cfg.ncomp ? (const float*)*fields : nullptr,
                              cfg.ncomp ? (float*)fout.p : nullptr, tmp.p, tb, &r, m_queue));
    FL_CHECK_STATUS(fl_stream_sync(m_queue));
    // release the caller's old arrays and hand over the new ones
    fl_mem* slots[7] = {cx, cy, cz, lv, st, rf, mat};
    DevBuf* bufs[7] = {&out.x, &out.y, &out.z, &out.level, &out.state, &out.flag, &out.material};
    for (int k = 0; k < 7; ++k) {
        (void)fl_free(*slots[k]);
        *slots[k] = bufs[k]->release();
    }
    if (cfg.ncomp) {
        (void)fl_free(*fields);
        *fields = fout.release();
    }
    *num_cells = (size_t)r.num_cells_out;
    *capacity = newcap;
    return record(m_queue);
}

}  // namespace adaptation
}  // namespace fluidloom
