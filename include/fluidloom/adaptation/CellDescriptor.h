// Cell descriptors and states (include/fluidloom/adaptation/CellDescriptor.h:32-121
// of the reference), native per-level coordinates: child q of (x,y,z,L) is
// ((x<<1)|(q&1), (y<<1)|(q>>1&1), (z<<1)|(q>>2&1), L+1); parent is >>1.
// Plain integer members instead of bitfields (levels up to 19 here).
#pragma once
#include <cstdint>

namespace fluidloom {
namespace adaptation {

static constexpr uint8_t MAX_REFINEMENT_LEVEL = 8;
static constexpr uint32_t INVALID_INDEX = 0xFFFFFFFFu;

enum class CellState : uint8_t {
    FLUID = 0,
    SOLID = 1,
    GEOMETRY_STATIC = 2,
    GEOMETRY_MOVING = 3,
    UNALLOCATED = 4,
    GHOST = 5
};

struct CellDescriptor {
    int32_t x = 0, y = 0, z = 0;
    uint8_t level = 0;
    uint8_t state = 0;
    uint32_t material_id = 0;

    void getParent(CellDescriptor& p) const {
        p.x = x >> 1; p.y = y >> 1; p.z = z >> 1;
        p.level = (uint8_t)(level - 1);
        p.state = state;
        p.material_id = material_id;
    }
    void getChild(unsigned q, CellDescriptor& c) const {
        c.x = (x << 1) | (int32_t)(q & 1);
        c.y = (y << 1) | (int32_t)((q >> 1) & 1);
        c.z = (z << 1) | (int32_t)((q >> 2) & 1);
        c.level = (uint8_t)(level + 1);
        c.state = state;
        c.material_id = material_id;
    }
    bool isSiblingOf(const CellDescriptor& o) const {
        return level == o.level && level != 0 && (x >> 1) == (o.x >> 1) && (y >> 1) == (o.y >> 1) &&
               (z >> 1) == (o.z >> 1);
    }
    bool isFirstSibling() const { return (x & 1) == 0 && (y & 1) == 0 && (z & 1) == 0; }
    bool operator==(const CellDescriptor& o) const {
        return x == o.x && y == o.y && z == o.z && level == o.level;
    }
};

}  // namespace adaptation
}  // namespace fluidloom
