// D3Q19 lattice constants and per-cell device arithmetic.
//
// Velocity order: rest, +x, -x, +y, -y, +z, -z (the order of the lattice
// fixture in tests/unit/parsing/test_parsing.cpp:23-29), then the twelve
// edge directions in opposite pairs, so opp(i) = i + 1 for odd i and i - 1
// for even i > 0 -- the table LatticeDescriptor::computeOpposites
// (src/parsing/registry/LatticeDescriptor.cpp:30-49) derives.  Weights 1/3,
// 1/18, 1/36; cs^2 = 1/3.
//
// Storage: populations are SoA f[q][slot] with slot = brick * 64 + local,
// local = lx + 4 ly + 16 lz inside a 4x4x4 brick.  Three storage types:
// float, double, and _Float16 holding f_q - w_q.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fl {
namespace lbm {

constexpr int Q = 19;

__host__ __device__ constexpr int cx(int i) {
    return i == 1 || i == 7 || i == 9 || i == 11 || i == 13 ? 1
         : (i == 2 || i == 8 || i == 10 || i == 12 || i == 14 ? -1 : 0);
}
__host__ __device__ constexpr int cy(int i) {
    return i == 3 || i == 7 || i == 10 || i == 15 || i == 17 ? 1
         : (i == 4 || i == 8 || i == 9 || i == 16 || i == 18 ? -1 : 0);
}
__host__ __device__ constexpr int cz(int i) {
    return i == 5 || i == 11 || i == 14 || i == 15 || i == 18 ? 1
         : (i == 6 || i == 12 || i == 13 || i == 16 || i == 17 ? -1 : 0);
}
__host__ __device__ constexpr int opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }
__host__ __device__ constexpr double wq(int i) {
    return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
}

// Storage codecs.
template <typename T>
struct Store;
template <>
struct Store<float> {
    using R = float;
    __device__ static inline float load(const float* p, size_t i, int) { return p[i]; }
    __device__ static inline void store(float* p, size_t i, int, float v) { p[i] = v; }
};
template <>
struct Store<double> {
    using R = double;
    __device__ static inline double load(const double* p, size_t i, int) { return p[i]; }
    __device__ static inline void store(double* p, size_t i, int, double v) { p[i] = v; }
};
template <>
struct Store<_Float16> {
    using R = float;
    __device__ static inline float load(const _Float16* p, size_t i, int q) {
        return (float)wq(q) + (float)p[i];
    }
    __device__ static inline void store(_Float16* p, size_t i, int q, float v) {
        p[i] = (_Float16)(v - (float)wq(q));
    }
};

template <typename R>
__device__ inline R feq(int i, R rho, R ux, R uy, R uz, R usq) {
    const R cu = (R)cx(i) * ux + (R)cy(i) * uy + (R)cz(i) * uz;
    return (R)wq(i) * rho * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * usq);
}

template <typename R>
__device__ inline void moments(const R* f, R& rho, R& ux, R& uy, R& uz) {
    R r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        r += f[i];
        if (cx(i)) jx += cx(i) > 0 ? f[i] : -f[i];
        if (cy(i)) jy += cy(i) > 0 ? f[i] : -f[i];
        if (cz(i)) jz += cz(i) > 0 ? f[i] : -f[i];
    }
    rho = r;
    ux = jx / r;
    uy = jy / r;
    uz = jz / r;
}

// BGK: f* = f - (f - feq) / tau (in place)
template <typename R>
__device__ inline void bgk(R* f, R inv_tau) {
    R rho, ux, uy, uz;
    moments(f, rho, ux, uy, uz);
    const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = f[i] - (f[i] - feq(i, rho, ux, uy, uz, usq)) * inv_tau;
}

// Per-cell link codes (2 bits per direction in a u64; bit 63 = active fluid
// leaf): 0 regular pull, 1 bounce-back from a solid, 2 domain wall (moving
// wall term of the face), 3 outflow copy from the position clamped into the
// domain.
constexpr uint64_t kActiveBit = 1ull << 63;
enum : unsigned { kLinkRegular = 0, kLinkSolid = 1, kLinkWall = 2, kLinkOutflow = 3 };

struct LevelDev {
    const int32_t* nbr;      // [brick][27], -1 = none
    const int32_t* bcoord;   // [brick][3]
    const uint32_t* bflags;  // bit 0: simple brick (all 64 active, all links regular)
    const uint64_t* cflags;  // [slot]
    size_t S;                // slot stride = nbricks * 64
    int32_t ext[3];          // domain extent in cells at this level
    int32_t nactive;         // active bricks (they come first)
};

struct BCDev {
    int32_t type[6];
    float u[6][3];
    double ud[6][3];
    double rho0;
};

// Wall face for an outside source (axes z, y, x in that order, WALL faces only).
__device__ inline int wall_face(const BCDev& bc, const int32_t* ext, int sx, int sy, int sz) {
    int p[3] = {sx, sy, sz};
    for (int a = 2; a >= 0; --a) {
        int f = p[a] < 0 ? 2 * a : (p[a] >= ext[a] ? 2 * a + 1 : -1);
        if (f >= 0 && bc.type[f] == 0) return f;
    }
    return -1;
}

// Slot of the same-level position (lx+dx, ...) relative to a brick, via the
// brick neighbour table (global loads; used by list kernels).
__device__ inline int64_t slot_of(const LevelDev& L, int b, int px, int py, int pz) {
    int nb = L.nbr[b * 27 + 13 + (px >> 2) + 3 * (py >> 2) + 9 * (pz >> 2)];
    if (nb < 0) return -1;
    return (int64_t)nb * 64 + (px & 3) + 4 * (py & 3) + 16 * (pz & 3);
}

// Generic pull of the 19 incoming populations of one slot (list kernels).
template <typename T, typename R>
__device__ inline void gather_slot(const LevelDev& L, const BCDev& bc, const T* __restrict__ A, int64_t slot,
                                   R* f) {
    const int b = (int)(slot >> 6);
    const int l = (int)(slot & 63);
    const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
    const uint64_t w = L.cflags[slot];
    const size_t S = L.S;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const unsigned code = (unsigned)(w >> (2 * i)) & 3u;
        const int px = lx - cx(i), py = ly - cy(i), pz = lz - cz(i);
        if (code == kLinkRegular) {
            f[i] = Store<T>::load(A, (size_t)i * S + (size_t)slot_of(L, b, px, py, pz), i);
        } else if (code == kLinkSolid) {
            f[i] = Store<T>::load(A, (size_t)opp(i) * S + (size_t)slot, opp(i));
        } else if (code == kLinkWall) {
            const int32_t* bc3 = L.bcoord + 3 * b;
            int face = wall_face(bc, L.ext, bc3[0] * 4 + px, bc3[1] * 4 + py, bc3[2] * 4 + pz);
            R uw = 0;
            if (face >= 0) {
                if (sizeof(R) == 8)
                    uw = (R)(cx(i) * bc.ud[face][0] + cy(i) * bc.ud[face][1] + cz(i) * bc.ud[face][2]);
                else
                    uw = (R)((float)cx(i) * bc.u[face][0] + (float)cy(i) * bc.u[face][1] +
                             (float)cz(i) * bc.u[face][2]);
            }
            f[i] = Store<T>::load(A, (size_t)opp(i) * S + (size_t)slot, opp(i)) +
                   (R)6 * (R)wq(i) * (R)bc.rho0 * uw;
        } else {
            const int32_t* bc3 = L.bcoord + 3 * b;
            int gx = bc3[0] * 4 + px, gy = bc3[1] * 4 + py, gz = bc3[2] * 4 + pz;
            int qx = gx < 0 ? 0 : (gx >= L.ext[0] ? L.ext[0] - 1 : gx);
            int qy = gy < 0 ? 0 : (gy >= L.ext[1] ? L.ext[1] - 1 : gy);
            int qz = gz < 0 ? 0 : (gz >= L.ext[2] ? L.ext[2] - 1 : gz);
            int64_t s2 = slot_of(L, b, qx - bc3[0] * 4, qy - bc3[1] * 4, qz - bc3[2] * 4);
            f[i] = Store<T>::load(A, (size_t)i * S + (size_t)s2, i);
        }
    }
}

}  // namespace lbm
}  // namespace fl
