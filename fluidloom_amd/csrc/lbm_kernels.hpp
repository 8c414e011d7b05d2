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

}  // namespace lbm
}  // namespace fl
