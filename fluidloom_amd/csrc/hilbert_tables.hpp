// Hilbert codec tables and device helpers shared by the key, AMR and lattice
// kernels.
//
// Core codec: the 12-state table of src/core/hilbert/HilbertCodec.cpp:13-42
// (entry = (next_state << 3) | curve_digit, indexed [state][(z<<2)|(y<<1)|x]),
// and its inverse (:27-42).  On the device the table is expanded per
// workgroup into LDS as a three-level table T3[state][q2 q1 q0] =
// (next_state << 9) | (d2 d1 d0) so that a key of level L costs ceil(L/3)
// dependent LDS lookups instead of L.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fl {

__device__ __constant__ static const uint8_t kHT[12][8] = {
    {8, 17, 27, 18, 47, 38, 28, 37},  {16, 71, 1, 62, 51, 52, 2, 61},
    {0, 75, 95, 76, 9, 10, 86, 85},   {4, 77, 55, 78, 3, 66, 80, 65},
    {50, 49, 45, 46, 67, 88, 68, 7},  {6, 57, 39, 72, 5, 58, 84, 83},
    {12, 11, 29, 90, 79, 32, 30, 89}, {74, 91, 73, 40, 69, 92, 70, 15},
    {14, 13, 81, 82, 63, 36, 24, 35}, {20, 31, 19, 56, 53, 54, 42, 41},
    {26, 93, 43, 44, 25, 94, 64, 23}, {22, 87, 21, 60, 33, 48, 34, 59},
};
__device__ __constant__ static const uint8_t kIHT[12][8] = {
    {8, 17, 19, 26, 30, 39, 37, 44},  {16, 2, 6, 52, 53, 63, 59, 65},
    {0, 12, 13, 73, 75, 87, 86, 90},  {86, 71, 69, 4, 0, 73, 75, 50},
    {93, 49, 48, 68, 70, 42, 43, 7},  {75, 57, 61, 87, 86, 4, 0, 34},
    {37, 95, 91, 9, 8, 26, 30, 76},   {43, 74, 72, 89, 93, 68, 70, 15},
    {30, 82, 83, 39, 37, 9, 8, 60},   {59, 47, 46, 18, 16, 52, 53, 25},
    {70, 28, 24, 42, 43, 89, 93, 23}, {53, 36, 38, 63, 59, 18, 16, 81},
};
// Rotation/direction codec tables of src/adaptation/utils/HilbertCodec3D.cpp:7-27.
__device__ __constant__ static const uint8_t kROT[8][8] = {
    {0, 1, 2, 3, 4, 5, 6, 7}, {1, 0, 7, 6, 5, 4, 3, 2}, {2, 3, 0, 1, 6, 7, 4, 5},
    {3, 2, 5, 4, 7, 6, 1, 0}, {4, 5, 6, 7, 0, 1, 2, 3}, {5, 4, 3, 2, 1, 0, 7, 6},
    {6, 7, 4, 5, 2, 3, 0, 1}, {7, 6, 1, 0, 3, 2, 5, 4}};
__device__ __constant__ static const uint8_t kDIR[8][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0}, {3, 0, 0, 5, 5, 0, 0, 3}, {4, 4, 0, 3, 3, 0, 4, 4},
    {3, 5, 5, 0, 0, 5, 5, 3}, {2, 2, 2, 0, 0, 2, 2, 2}, {3, 0, 0, 5, 5, 0, 0, 3},
    {4, 4, 0, 3, 3, 0, 4, 4}, {3, 5, 5, 0, 0, 5, 5, 3}};

// LDS-resident three-level tables (12 KB each).
struct HilbertLds {
    uint16_t t3[12 * 512];
    uint8_t t1[12 * 8];
};

// Fill the encode (inverse == false) or decode (inverse == true) tables.
// Must be followed by __syncthreads() before use.
__device__ inline void hilbert_lds_init(HilbertLds& L, bool inverse) {
    const uint8_t(*T)[8] = inverse ? kIHT : kHT;
    for (int e = threadIdx.x; e < 12 * 512; e += blockDim.x) {
        int s = e >> 9, q = e & 511;
        int v0 = T[s][(q >> 6) & 7];
        int s1 = v0 >> 3;
        int v1 = T[s1][(q >> 3) & 7];
        int s2 = v1 >> 3;
        int v2 = T[s2][q & 7];
        L.t3[e] = (uint16_t)(((v2 >> 3) << 9) | ((v0 & 7) << 6) | ((v1 & 7) << 3) | (v2 & 7));
    }
    for (int e = threadIdx.x; e < 96; e += blockDim.x) L.t1[e] = T[e >> 3][e & 7];
}

// hilbert::encode semantics: 21-bit coordinate mask, MSB first, 3*level bits.
__device__ inline uint64_t hilbert_encode_lds(const HilbertLds& L, uint32_t x, uint32_t y, uint32_t z,
                                              int level) {
    x &= 0x1FFFFFu; y &= 0x1FFFFFu; z &= 0x1FFFFFu;
    uint64_t h = 0;
    int s = 0;
    int i = level - 1;
    // leading (level % 3) single steps
    for (int r = level % 3; r > 0; --r, --i) {
        int q = (int)((((z >> i) & 1u) << 2) | (((y >> i) & 1u) << 1) | ((x >> i) & 1u));
        int v = L.t1[s * 8 + q];
        h = (h << 3) | (uint64_t)(v & 7);
        s = v >> 3;
    }
    for (; i >= 2; i -= 3) {
        uint32_t bx = (x >> (i - 2)) & 7u, by = (y >> (i - 2)) & 7u, bz = (z >> (i - 2)) & 7u;
        // interleave three quadrants (MSB first): q = q2 q1 q0, q_k = (z_k y_k x_k)
        uint32_t q = 0;
        q |= (((bz >> 2) & 1u) << 8) | (((by >> 2) & 1u) << 7) | (((bx >> 2) & 1u) << 6);
        q |= (((bz >> 1) & 1u) << 5) | (((by >> 1) & 1u) << 4) | (((bx >> 1) & 1u) << 3);
        q |= ((bz & 1u) << 2) | ((by & 1u) << 1) | (bx & 1u);
        int v = L.t3[s * 512 + q];
        h = (h << 9) | (uint64_t)(v & 511);
        s = v >> 9;
    }
    return h;  // already < 2^(3*level); the reference's canonical mask is a no-op
}

__device__ inline void hilbert_decode_lds(const HilbertLds& L, uint64_t h, int level, int32_t* ox,
                                          int32_t* oy, int32_t* oz) {
    uint32_t x = 0, y = 0, z = 0;
    int s = 0;
    int rem = level;  // digits left, MSB first
    for (int r = level % 3; r > 0; --r) {
        --rem;
        int d = (int)((h >> (3 * rem)) & 7u);
        int v = L.t1[s * 8 + d];
        int q = v & 7;
        s = v >> 3;
        x = (x << 1) | (uint32_t)(q & 1);
        y = (y << 1) | (uint32_t)((q >> 1) & 1);
        z = (z << 1) | (uint32_t)((q >> 2) & 1);
    }
    while (rem >= 3) {
        rem -= 3;
        int d = (int)((h >> (3 * rem)) & 511u);
        int v = L.t3[s * 512 + d];
        s = v >> 9;
        for (int k = 2; k >= 0; --k) {
            int q = (v >> (3 * k)) & 7;
            x = (x << 1) | (uint32_t)(q & 1);
            y = (y << 1) | (uint32_t)((q >> 1) & 1);
            z = (z << 1) | (uint32_t)((q >> 2) & 1);
        }
    }
    *ox = (int32_t)x; *oy = (int32_t)y; *oz = (int32_t)z;
}

// Direct (non-LDS) single-level encode used by kernels whose LDS is busy.
__device__ inline uint64_t hilbert_encode_direct(uint32_t x, uint32_t y, uint32_t z, int level) {
    x &= 0x1FFFFFu; y &= 0x1FFFFFu; z &= 0x1FFFFFu;
    uint64_t h = 0;
    int s = 0;
    for (int i = level - 1; i >= 0; --i) {
        int q = (int)((((z >> i) & 1u) << 2) | (((y >> i) & 1u) << 1) | ((x >> i) & 1u));
        int v = kHT[s][q];
        h = (h << 3) | (uint64_t)(v & 7);
        s = v >> 3;
    }
    return h;
}

// Unique cell key used by every cell hash table:
// (level << 57) | (x << 38) | (y << 19) | z for coordinates below 2^19 and
// level <= 19 (never equal to the EMPTY key).
__host__ __device__ inline uint64_t cell_hkey(uint32_t x, uint32_t y, uint32_t z, int level) {
    return ((uint64_t)level << 57) | ((uint64_t)x << 38) | ((uint64_t)y << 19) | (uint64_t)z;
}
constexpr int32_t kKeyCoordLimit = 1 << 19;

}  // namespace fl
