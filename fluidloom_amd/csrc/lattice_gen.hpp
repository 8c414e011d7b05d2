// Generated by fluidloom_amd/lattice.py from the lattice declarations there;
// do not edit.  The lattice kernels take their velocities, weights and
// opposites from these tables.
#pragma once

namespace fl {
namespace lattice {

// Lattice D3Q19: Q = 19, cs^2 = 1/3
constexpr int D3Q19_Q = 19;
constexpr double D3Q19_CS2 = 1.0 / 3.0;
constexpr int D3Q19_C[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}, {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
constexpr double D3Q19_W[19] = {1.0 / 3.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr int D3Q19_OPP[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

}  // namespace lattice
}  // namespace fl
