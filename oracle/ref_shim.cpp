// extern "C" entry points over the reference's own Hilbert codec sources
// (compiled in place from the reference tree by oracle/Makefile).  Used only
// to cross-check the oracle's restatement and to write golden vectors.
#include <cstdint>
#include "fluidloom/core/hilbert/HilbertCodec.h"

namespace fluidloom { namespace adaptation {
uint64_t hilbert_encode_3d(int32_t x, int32_t y, int32_t z, uint8_t max_level);
} }

extern "C" uint64_t ref_hilbert_encode(int32_t x, int32_t y, int32_t z, int level) {
    return fluidloom::hilbert::encode(x, y, z, static_cast<uint8_t>(level));
}
extern "C" void ref_hilbert_decode(uint64_t h, int level, int32_t* x, int32_t* y, int32_t* z) {
    fluidloom::hilbert::decode(h, static_cast<uint8_t>(level), *x, *y, *z);
}
extern "C" uint64_t ref_rotdir_encode(int32_t x, int32_t y, int32_t z, int max_level) {
    return fluidloom::adaptation::hilbert_encode_3d(x, y, z, static_cast<uint8_t>(max_level));
}
extern "C" int ref_is_valid(uint64_t h, int level) {
    return fluidloom::hilbert::isValid(h, static_cast<uint8_t>(level)) ? 1 : 0;
}
