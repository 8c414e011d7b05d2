// Device-side lattice layout: the brick tables, link words and ghost lists
// of every level built by kernels from leaf arrays that stay on the device,
// plus the per-level brick maps (brick key -> brick) that slot lookups use
// afterwards (queries, regrid population transfer).
//
// The tables are the ones the host builder in lbm_solver.hip produces, in
// the same order, so both layouts are interchangeable bit for bit:
//   bricks      a level's leaf bricks plus the 26-neighbour dilation of its
//               active (fluid-holding) bricks, ordered active first, then
//               by instance, then along the Hilbert curve of the brick
//               coordinates (16 levels);
//   nbr         27 entries per brick (13 = itself), -1 where absent;
//   cflags      per slot: 2 bits per direction (regular, solid, wall,
//               outflow) and the active bit, for fluid leaves;
//   bflags      bit 0: simple brick (64 fluid leaves, all links regular);
//   c2f groups  one per covering coarse leaf (ascending coarse slot), 8
//               child slots and 8 direction masks per group, 7 capture
//               indices into the coarse level's capture list (ascending,
//               unique slots);
//   f2c lists   ghosts with their 8 children, stably ordered by the first
//               child's slot;
//   force links slot * 32 + direction of every solid link of an active
//               brick, in slot order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "fl_internal.hpp"

namespace fl {
namespace lbm {

// Per-instance lattice data the layout reads (one record per instance).
struct LayoutInst {
    int32_t ext0[3];    // level-0 extents in cells
    int32_t max_level;  // finest level of the instance
    int32_t bc[6];      // faces -x,+x,-y,+y,-z,+z: FL_BC_*
};

// One level's layout on the device (arrays allocated through the caller's
// allocator, which owns them) and its sizes.
struct LayoutLevel {
    int32_t* nbr = nullptr;
    int32_t* bcoord = nullptr;
    uint16_t* binst = nullptr;
    uint32_t* bflags = nullptr;
    uint64_t* cflags = nullptr;
    uint32_t* c2f_gchild = nullptr;
    uint32_t* c2f_gcap = nullptr;
    uint32_t* c2f_gmask = nullptr;
    uint32_t* cap_slots = nullptr;
    uint32_t* f2c_ghost = nullptr;
    uint32_t* f2c_child = nullptr;
    uint32_t* force_links = nullptr;
    uint32_t* abeg = nullptr;  // [ninst + 1] active-brick range of each instance
    uint32_t* lbeg = nullptr;  // [ninst + 1] force-link range of each instance
    uint64_t* map_keys = nullptr;  // brick map: packed brick key -> brick index
    uint32_t* map_vals = nullptr;
    uint64_t map_cap = 0;
    uint32_t nb = 0, nactive = 0;
    uint64_t nfluid = 0;
    std::vector<uint64_t> nfluid_inst;
    size_t n_c2f = 0, n_groups = 0, n_cap = 0, n_f2c = 0, n_links = 0;
    std::vector<uint32_t> abeg_h, lbeg_h;  // host copies of abeg / lbeg
};

// Slots of a level's brick map: a power of two holding nb bricks at load
// <= 0.5 (at least 64).
inline uint64_t brick_map_capacity(uint64_t nb) {
    uint64_t c = 64;
    while (c < 2 * nb) c <<= 1;
    return c;
}
// Device bytes of a level's brick map.
inline uint64_t brick_map_bytes(uint64_t nb) { return brick_map_capacity(nb) * (sizeof(uint64_t) + sizeof(uint32_t)); }

// Persistent device allocation owned by the caller (the solver).
using DevAlloc = std::function<fl_status(void**, size_t)>;

// Builds every level's layout from device leaf arrays (native coordinates,
// level, state 0 fluid / 1 solid); leaves [offsets[k], offsets[k+1]) belong
// to instance k (offsets on the host).  Mesh errors return FL_E_MESH with
// the message in *err.
fl_status build_layout_device(const std::vector<LayoutInst>& inst, int nlevels, const size_t* h_offsets,
                              const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                              const uint8_t* st, size_t n, hipStream_t s, const DevAlloc& alloc, LayoutLevel* out,
                              std::string* err);

// The brick map of a layout whose bricks were built on the host.
fl_status build_brick_map(const int32_t* bcoord, const uint16_t* binst, uint32_t nb, hipStream_t s,
                          const DevAlloc& alloc, LayoutLevel& L);

// Device view of one level for slot lookups.
struct LevelMap {
    const uint64_t* keys;
    const uint32_t* vals;
    uint64_t cap;
    const uint64_t* cflags;  // active bit = fluid leaf
};
constexpr int kMaxLevels = 24;
struct LevelMaps {
    LevelMap m[kMaxLevels];
    int nlevels;
};

// Slots of fluid leaves (x, y, z, level) of instance `inst`, -1 where the
// position is not a fluid leaf; inputs and output on the device.
fl_status launch_query_slots(const LevelMaps& maps, const LayoutInst* d_inst, int inst, const int32_t* x,
                             const int32_t* y, const int32_t* z, const uint8_t* lv, size_t n, int64_t* out,
                             hipStream_t s);

// Population transfer lists of a regrid for new level l: every fluid slot
// of the new layout takes its populations from the old layout's leaf at the
// same position (copy), from its old parent (split), from its 8 old
// children (merge), or restarts from equilibrium (rest).  Lists in slot
// order; their device arrays are allocated with `alloc`.
struct RemapLists {
    uint32_t *copy_dst = nullptr, *copy_src = nullptr, *split_dst = nullptr, *split_src = nullptr;
    uint32_t *merge_dst = nullptr, *merge_src8 = nullptr, *rest_dst = nullptr;
    uint32_t ncopy = 0, nsplit = 0, nmerge = 0, nrest = 0;
};
fl_status build_remap_lists(const LevelMaps& old_maps, const LayoutInst* d_inst, int l, const LayoutLevel& nl,
                            hipStream_t s, const DevAlloc& alloc, RemapLists& out);

}  // namespace lbm
}  // namespace fl
