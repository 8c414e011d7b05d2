// Host-side launchers of the lattice kernels (lbm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_device.hpp"

namespace fl {
namespace lbm {

template <typename T>
hipError_t launch_collide(const LevelDev& L, const BCDev& bc, const T* A, T* B, double inv_tau, hipStream_t s);
template <typename T>
hipError_t launch_capture(const LevelDev& L, const BCDev& bc, const T* A, const uint32_t* slots, uint32_t ncap,
                          void* cap, hipStream_t s);
template <typename T>
hipError_t launch_fill_c2f(const uint32_t* ghost, const uint32_t* capidx, uint32_t n, const void* cap0,
                           const void* cap1, uint32_t ncap, double alpha, double fac, T* A, size_t S,
                           hipStream_t s);
template <typename T>
hipError_t launch_fill_f2c(const uint32_t* ghost, const uint32_t* child8, uint32_t n, const LevelDev& Lf,
                           const BCDev& bc, const T* Af, double fac, T* Ac, size_t Sc, hipStream_t s);
template <typename T>
hipError_t launch_init(size_t S, double rho, const double u[3], T* A, T* B, hipStream_t s);
unsigned totals_blocks(const LevelDev& L);
template <typename T>
hipError_t launch_totals(const LevelDev& L, const T* A, double vol, double* partial, unsigned nblk, hipStream_t s);
unsigned force_blocks(uint32_t n);
template <typename T>
hipError_t launch_force(const uint32_t* links, uint32_t n, size_t S, const T* A, double scale, double* partial,
                        unsigned nblk, hipStream_t s);
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
    T* A = nullptr;
    T* B = nullptr;
    size_t S = 0;
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
    double split_fac = 0;
    LevelDev Lf{};                     // old level l+1
    const T* merge_prev = nullptr;
    const uint32_t* merge_dst = nullptr;
    const uint32_t* merge_src8 = nullptr;
    uint32_t nmerge = 0;
    double merge_fac = 0;
    const uint32_t* rest_dst = nullptr;
    uint32_t nrest = 0;
    BCDev bc{};
};
template <typename T>
hipError_t launch_remap(const RemapArgs<T>& r, hipStream_t s);
template <typename T>
hipError_t launch_ghost_eq(const uint32_t* c2f_ghost, const uint32_t* c2f_cap, const uint32_t* cap_slots, uint32_t nc2f,
                           const T* Ac, size_t Sc, const uint32_t* f2c_ghost, const uint32_t* f2c_child, uint32_t nf2c,
                           const T* Af, size_t Sf, T* A, T* B, size_t S, hipStream_t s);

}  // namespace lbm
}  // namespace fl
