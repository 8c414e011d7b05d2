// Host-side launchers of the lattice kernels (lbm_kernels.hip).  Every
// launcher covers all instances of a batch at one level.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_device.hpp"

namespace fl {
namespace lbm {

// Collide-stream kernel variants (bricks per workgroup x XCD-aware remap).
int collide_variants();
const char* collide_variant_name(int v);

template <typename T>
hipError_t launch_collide(const LevelDev& L, const T* A, T* B, int variant, hipStream_t s);
// advances a level's stochastic-rounding clock (one kernel node, graph-safe)
// Bit 1 of bflags on every brick (of the first nslots / 64) with a
// pressure-outlet link, and *flag |= 1 when there is one (the caller zeroes
// *flag)
hipError_t launch_any_outlet(const uint64_t* cflags, uint32_t* bflags, size_t nslots, uint32_t* flag, hipStream_t s);
hipError_t launch_tick(uint32_t* tick, hipStream_t s);
template <typename T>
hipError_t launch_capture(const LevelDev& L, const T* A, const uint32_t* slots, uint32_t ncap, void* cap,
                          const uint8_t* skip, hipStream_t s);
// capture staging map of a level (nslots slots) for its ncap captures (k_cap_map)
hipError_t launch_cap_map(const uint32_t* slots, uint32_t ncap, const uint64_t* cflags, uint32_t* bflags,
                          uint32_t* cap_of, size_t nslots, uint8_t* staged, hipStream_t s);
template <typename T>
hipError_t launch_fill_c2f(const LevelDev& Lf, const uint32_t* gchild, const uint32_t* gcap, const uint32_t* gmask,
                           uint32_t ngroups, const void* cap0, const void* cap1, uint32_t ncap, double alpha, T* A,
                           hipStream_t s);
// stage, staged: the child sums the fine collide staged (LevelDev::stage) and
// which ghosts have them; stage == nullptr gathers every ghost's children
template <typename T>
hipError_t launch_fill_f2c(const LevelDev& Lc, const uint32_t* ghost, const uint32_t* child8, uint32_t n,
                           const LevelDev& Lf, const T* Af, T* Ac, const void* stage, const uint8_t* staged,
                           hipStream_t s);
// f2c staging map of a fine level (nbricks bricks) for the coarse level's n
// f2c ghosts (k_stage_map)
hipError_t launch_stage_map(const uint32_t* child8, uint32_t n, const uint64_t* cflags, uint32_t* bflags,
                            uint32_t* f2c_blk, size_t nbricks, uint8_t* staged, hipStream_t s);
template <typename T>
hipError_t launch_init(const LevelDev& L, T* A, T* B, hipStream_t s, bool skip_simple);
// FP64 reductions: per-brick (or per-link) values, then one workgroup per
// instance over its contiguous range (fixed order: results do not depend on
// the other instances of the batch).
template <typename T>
hipError_t launch_brick_sums(const LevelDev& L, const T* A, double vol, double* out, hipStream_t s);
template <typename T>
hipError_t launch_brick_du(const LevelDev& L, const T* A, float* uprev, float* uacc, float inv_m, double* out,
                           hipStream_t s);
template <typename T>
hipError_t launch_velocity_sample(const LevelDev& L, const T* A, float* uacc, hipStream_t s);
hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
hipError_t launch_inst_reduce(const double* vals, int ncomp, const uint32_t* beg, int ninst, int op, int first,
                              double* acc, hipStream_t s);
template <typename T>
hipError_t launch_link_force(const uint32_t* links, uint32_t n, size_t S, const T* A, double scale, double* out,
                             hipStream_t s);
template <typename T>
hipError_t launch_extract(const T* A, size_t S, const int64_t* slots, size_t n, double* out, hipStream_t s);
template <typename T>
hipError_t launch_inject(T* A, T* B, size_t S, const int64_t* slots, size_t n, const double* in, hipStream_t s);
template <typename T>
hipError_t launch_vorticity(const LevelDev& L, const T* A, const int64_t* slots, size_t n, double scale, double* out,
                            hipStream_t s);

// One level's population transfer after a regrid.
template <typename T>
struct RemapArgs {
    LevelDev Ln{};                     // new level l
    T* A = nullptr;
    T* B = nullptr;
    const T* copy_src_buf = nullptr;  // old level, current buffer
    size_t So = 0;
    const uint32_t* copy_dst = nullptr;
    const uint32_t* copy_src = nullptr;
    uint32_t ncopy = 0;
    LevelDev Lc{};                     // old level l-1
    const T* split_prev = nullptr;     // its previous buffer
    const uint32_t* split_dst = nullptr;
    const uint32_t* split_src = nullptr;
    uint32_t nsplit = 0;
    LevelDev Lf{};                     // old level l+1
    const T* merge_prev = nullptr;
    const uint32_t* merge_dst = nullptr;
    const uint32_t* merge_src8 = nullptr;
    uint32_t nmerge = 0;
    const uint32_t* rest_dst = nullptr;
    uint32_t nrest = 0;
};
template <typename T>
hipError_t launch_remap(const RemapArgs<T>& r, hipStream_t s);
template <typename T>
hipError_t launch_ghost_eq(const uint32_t* c2f_gchild, const uint32_t* c2f_gcap, const uint32_t* cap_slots, uint32_t nc2f,
                           const T* Ac, size_t Sc, const uint32_t* f2c_ghost, const uint32_t* f2c_child, uint32_t nf2c,
                           const T* Af, size_t Sf, T* A, T* B, size_t S, hipStream_t s);

}  // namespace lbm
}  // namespace fl
