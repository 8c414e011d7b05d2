// D3Q19 lattice constants and per-cell device arithmetic.
//
// The velocities, weights and opposites come from lattice_gen.hpp, which
// fluidloom_amd/lattice.py generates from its lattice declaration (the
// HIP counterpart of LatticeDescriptor::generateOpenCLCode,
// src/parsing/registry/LatticeDescriptor.cpp:51-80).  Velocity order: rest,
// +x, -x, +y, -y, +z, -z (the lattice fixture in
// tests/unit/parsing/test_parsing.cpp:23-29), then the twelve edge
// directions in opposite pairs.  Weights 1/3, 1/18, 1/36; cs^2 = 1/3.
//
// Storage: slot = brick * 64 + local, local = lx + 4 ly + 16 lz inside a
// 4x4x4 brick; populations are 64-cell brick planes in plane groups (pidx).  Four storage types:
// float, double, _Float16 holding f_q - w_q, and fp16c holding (f_q - w_q) / w_q
// (fp16c_codec.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fp16c_codec.hpp"
#include "lattice_gen.hpp"

namespace fl {
namespace lbm {

constexpr int Q = lattice::D3Q19_Q;
static_assert(Q == 19, "the brick kernels are written for D3Q19");

__host__ __device__ constexpr int cx(int i) { return lattice::D3Q19_C[i][0]; }
__host__ __device__ constexpr int cy(int i) { return lattice::D3Q19_C[i][1]; }
__host__ __device__ constexpr int cz(int i) { return lattice::D3Q19_C[i][2]; }
__host__ __device__ constexpr int opp(int i) { return lattice::D3Q19_OPP[i]; }
__host__ __device__ constexpr double wq(int i) { return lattice::D3Q19_W[i]; }

// The kernels rely on opposite pairs (2k-1, 2k) after the rest velocity.
constexpr bool pairs_adjacent() {
    for (int i = 1; i < Q; ++i)
        if (opp(i) != ((i & 1) ? i + 1 : i - 1)) return false;
    return opp(0) == 0;
}
static_assert(pairs_adjacent(), "lattice declaration must list opposite velocities in adjacent pairs");

// Plane groups.  Planes 16g .. 16g+15 form group g (the last group holds the
// remaining Q - 16 = 3), and inside a group each brick's 64-cell planes sit
// next to each other: group g starts at element 16g * S, and brick b's run
// in it at 64 * (group size) * b.  One brick's reads and writes are then two
// contiguous runs per buffer (4 KiB and 768 B in FP32) instead of 19 runs of
// 256 B, which HBM serves with fewer row activations.  (All 19 planes of a
// brick in one run was far slower: neighbouring waves then crowd into a
// narrow address window.  Groups of 8 measured 1-3 % slower.)
constexpr int kPlaneGroup = 16;
__host__ __device__ constexpr int plane_group_size(int i) {
    return (i / kPlaneGroup + 1) * kPlaneGroup <= Q ? kPlaneGroup : Q % kPlaneGroup;
}
// Element offset of plane i's first brick, and the distance between bricks.
__host__ __device__ constexpr size_t plane_base(int i, size_t S) {
    return (size_t)(i - i % kPlaneGroup) * S + (size_t)(i % kPlaneGroup) * 64;
}
__host__ __device__ constexpr size_t brick_stride(int i) { return (size_t)64 * plane_group_size(i); }
// Element index of population i of a slot (brick * 64 + local) in a level's
// population array of Q * S elements.
__host__ __device__ constexpr size_t pidx(int i, size_t S, size_t slot) {
    return plane_base(i, S) + (slot >> 6) * brick_stride(i) + (slot & 63);
}

namespace detail {
// every (plane, brick) pair maps to its own 64-element run of [0, Q * S)
constexpr bool plane_groups_tile(size_t nb) {
    const size_t S = nb * 64, n = Q * nb;
    unsigned long long seen[4] = {0, 0, 0, 0};
    if (n > 256) return false;
    for (int i = 0; i < Q; ++i)
        for (size_t b = 0; b < nb; ++b) {
            const size_t e = pidx(i, S, b * 64);
            if (e % 64 || e / 64 >= n) return false;
            const size_t r = e / 64;
            if (seen[r / 64] >> (r % 64) & 1ull) return false;
            seen[r / 64] |= 1ull << (r % 64);
        }
    return true;
}
}  // namespace detail
static_assert(detail::plane_groups_tile(1) && detail::plane_groups_tile(3) && detail::plane_groups_tile(13),
              "plane groups must tile the population array");

// Storage codecs.
template <typename T>
struct Store;
// store_nt: streaming store of the collide output (not re-read by the same
// launch), kept out of the caches that serve neighbour re-reads.
//
// Besides the population itself (load / store), every codec reads and writes
// its deviation h = f - w_q (dev / store_dev*), which the stepping kernels
// compute with: density and momentum are then sums of small numbers
// (rho = 1 + sum h), so neither the rounding of the weights to the storage
// type nor the fixed low bits that 16-bit storage leaves in w + h bias them.
template <>
struct Store<float> {
    using R = float;
    __device__ static inline float load(const float* p, size_t i, int) { return p[i]; }
    __device__ static inline void store(float* p, size_t i, int, float v) { p[i] = v; }
    __device__ static inline void store_nt(float* p, size_t i, int, float v) { __builtin_nontemporal_store(v, p + i); }
    __device__ static inline float dev(const float* p, size_t i, int q) { return p[i] - (float)wq(q); }
    __device__ static inline void store_dev(float* p, size_t i, int q, float h) { p[i] = (float)wq(q) + h; }
    __device__ static inline void store_dev_nt(float* p, size_t i, int q, float h) {
        __builtin_nontemporal_store((float)wq(q) + h, p + i);
    }
};
template <>
struct Store<double> {
    using R = double;
    __device__ static inline double load(const double* p, size_t i, int) { return p[i]; }
    __device__ static inline void store(double* p, size_t i, int, double v) { p[i] = v; }
    __device__ static inline void store_nt(double* p, size_t i, int, double v) {
        __builtin_nontemporal_store(v, p + i);
    }
    __device__ static inline double dev(const double* p, size_t i, int q) { return p[i] - wq(q); }
    __device__ static inline void store_dev(double* p, size_t i, int q, double h) { p[i] = wq(q) + h; }
    __device__ static inline void store_dev_nt(double* p, size_t i, int q, double h) {
        __builtin_nontemporal_store(wq(q) + h, p + i);
    }
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
    __device__ static inline void store_nt(_Float16* p, size_t i, int q, float v) {
        __builtin_nontemporal_store((_Float16)(v - (float)wq(q)), p + i);
    }
    // Stochastic rounding of f - w to FP16 (r: 32 random bits) with the
    // gfx950 conversion instruction (v_cvt_sr_f16_f32): the value rounds up
    // with probability equal to the dropped fraction, also for FP16
    // subnormals.  Its mean over random r equals the input to within the
    // sampling error (measured at 4M draws for normal, subnormal and
    // below-subnormal inputs).  Keeps slow updates (below half an FP16 ulp per
    // step) from being lost, which round-to-nearest does every step.
    __device__ static inline void store_sr(_Float16* p, size_t i, int q, float v, uint32_t r) {
        store_dev_sr(p, i, q, v - (float)wq(q), r);
    }
    __device__ static inline float dev(const _Float16* p, size_t i, int) { return (float)p[i]; }
    // dev() at a 32-bit byte offset from a wave-uniform base: the load then
    // takes the base in scalar registers and only the offset in a vector
    // register (no 64-bit address per lane).  A plane group spans below 2^32
    // bytes in 16-bit storage (bricks per level < 2^21, 1024 cells per brick).
    __device__ static inline float dev_at(const _Float16* base, unsigned boff, int) {
        return (float)*reinterpret_cast<const _Float16*>(reinterpret_cast<const char*>(base) + boff);
    }
    __device__ static inline void store_dev(_Float16* p, size_t i, int, float h) { p[i] = (_Float16)h; }
    __device__ static inline void store_dev_sr(_Float16* p, size_t i, int, float h, uint32_t r) {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        const h2 keep = {(_Float16)0, (_Float16)0};
        __builtin_nontemporal_store(__builtin_amdgcn_cvt_sr_f16_f32(keep, h, r, false)[0], p + i);
    }
};

template <>
struct Store<fp16c> {
    using R = float;
    // f = w + w x: the code float times 2^113 w in one fma (exact scaling)
    __device__ static inline float load(const fp16c* p, size_t i, int q) {
        const float w = (float)wq(q);
        return fmaf(fp16c_code_float(p[i].bits), w * kFp16cUp, w);
    }
    __device__ static inline uint16_t enc(float v, int q) {
        return fp16c_encode_rne((v - (float)wq(q)) * (float)(1.0 / wq(q)));
    }
    __device__ static inline void store(fp16c* p, size_t i, int q, float v) { p[i].bits = enc(v, q); }
    __device__ static inline void store_nt(fp16c* p, size_t i, int q, float v) {
        __builtin_nontemporal_store(enc(v, q), &p[i].bits);
    }
    // (f - w) (2^-113 / w): for |x| >= 2^-13 the same rounding as (f - w)(1 / w)
    __device__ static inline void store_sr(fp16c* p, size_t i, int q, float v, uint32_t r) {
        store_dev_sr(p, i, q, v - (float)wq(q), r);
    }
    // h = w x: the code float times 2^113 w
    __device__ static inline float dev(const fp16c* p, size_t i, int q) {
        return fp16c_code_float(p[i].bits) * ((float)wq(q) * kFp16cUp);
    }
    // dev() at a 32-bit byte offset from a wave-uniform base (as for _Float16)
    __device__ static inline float dev_at(const fp16c* base, unsigned boff, int q) {
        const uint16_t bits = *reinterpret_cast<const uint16_t*>(reinterpret_cast<const char*>(base) + boff);
        return fp16c_code_float(bits) * ((float)wq(q) * kFp16cUp);
    }
    __device__ static inline void store_dev(fp16c* p, size_t i, int q, float h) {
        p[i].bits = fp16c_encode_rne(h * (float)(1.0 / wq(q)));
    }
    __device__ static inline void store_dev_sr(fp16c* p, size_t i, int q, float h, uint32_t r) {
        const float xs = h * ((float)(1.0 / wq(q)) * kFp16cDown);
        __builtin_nontemporal_store(fp16c_encode_sr_scaled(xs, r), &p[i].bits);
    }
};

// 32-bit integer finaliser ("lowbias32", C. Wellons' hash-prospector) used
// to draw the stochastic-rounding bits from (cell, level, step, direction).
__device__ inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// c_i . u from the nonzero components of c_i only (i a compile-time constant
// after unrolling).  The same value as cx*ux + cy*uy + cz*uz for finite u:
// the dropped terms are products with 0, which IEEE arithmetic does not let
// the compiler remove by itself (0 * inf is NaN), and +-1 products are exact.
template <typename R>
__device__ __attribute__((always_inline)) inline R cdot(int i, R ux, R uy, R uz) {
    R s = 0;
    bool any = false;
    if (cx(i)) {
        s = cx(i) > 0 ? ux : -ux;
        any = true;
    }
    if (cy(i)) {
        const R t = cy(i) > 0 ? uy : -uy;
        s = any ? s + t : t;
        any = true;
    }
    if (cz(i)) {
        const R t = cz(i) > 0 ? uz : -uz;
        s = any ? s + t : t;
    }
    return s;
}

template <typename R>
__device__ __attribute__((always_inline)) inline R fmadd(R a, R b, R c) {
    if constexpr (sizeof(R) == 8)
        return __builtin_fma(a, b, c);
    else
        return __builtin_fmaf(a, b, c);
}

template <typename R>
__device__ inline R feq(int i, R rho, R ux, R uy, R uz, R usq) {
    const R cu = cdot(i, ux, uy, uz);
    return (R)wq(i) * rho * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * usq);
}

template <typename R>
__device__ inline void density_momentum(const R* f, R& rho, R& jx, R& jy, R& jz) {
    R r = 0, ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        r += f[i];
        if (cx(i)) ax += cx(i) > 0 ? f[i] : -f[i];
        if (cy(i)) ay += cy(i) > 0 ? f[i] : -f[i];
        if (cz(i)) az += cz(i) > 0 ? f[i] : -f[i];
    }
    rho = r;
    jx = ax;
    jy = ay;
    jz = az;
}

template <typename R>
__device__ inline void moments(const R* f, R& rho, R& ux, R& uy, R& uz) {
    R jx, jy, jz;
    density_momentum(f, rho, jx, jy, jz);
    const R ir = (R)1 / rho;
    ux = jx * ir;
    uy = jy * ir;
    uz = jz * ir;
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

// ---- deviation form: h_i = f_i - w_i --------------------------------------
// Equilibrium deviation feq_i - w_i = w_i (d + rho (3 cu + 4.5 cu^2 - 1.5 u^2))
// with d = rho - 1 (from sum_i w_i = 1 and sum_i w_i c_i = 0).
template <typename R>
__device__ inline R feq_dev(int i, R d, R rho, R ux, R uy, R uz, R usq) {
    const R cu = cdot(i, ux, uy, uz);
    return (R)wq(i) * (d + rho * ((R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * usq));
}

// d = sum h = rho - 1, j = sum c h (= sum c f)
template <typename R>
__device__ inline void dev_density_momentum(const R* h, R& d, R& jx, R& jy, R& jz) {
    density_momentum(h, d, jx, jy, jz);
}

template <typename R>
__device__ inline void dev_moments(const R* h, R& d, R& rho, R& ux, R& uy, R& uz) {
    R jx, jy, jz;
    density_momentum(h, d, jx, jy, jz);
    rho = (R)1 + d;
    const R ir = (R)1 / rho;
    ux = jx * ir;
    uy = jy * ir;
    uz = jz * ir;
}

// BGK on deviations, h* = h - (h - (feq - w)) / tau (in place): the same
// update as bgk() on f = w + h, arranged over the opposite pairs (i, i+1)
// of the lattice, which share |c_i . u| and the weight.  With omega = 1 / tau
// and c0 = d - 1.5 rho u^2,
//   omega (feq - w)_i = omega w (c0 + 4.5 rho cu^2) +- 3 omega w rho cu = S +- A,
//   h*_i = (1 - omega) h_i + S + A,  h*_{i+1} = (1 - omega) h_{i+1} + S - A,
// and density and momentum come from the pair sums and differences.  About
// 130 operations per cell against 240 for the direction-by-direction form
// (the collide kernels in 16-bit storage are bound by instruction issue as
// much as by memory).  1 / rho in FP32: the hardware reciprocal and one
// Newton step.
template <typename R>
__device__ __attribute__((always_inline)) inline void bgk_dev(R* h, R inv_tau) {
    constexpr int P = (Q - 1) / 2;
    R sum[P], dif[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        sum[p] = h[2 * p + 1] + h[2 * p + 2];
        dif[p] = h[2 * p + 1] - h[2 * p + 2];
    }
    R d = h[0];
#pragma unroll
    for (int p = 0; p < P; ++p) d += sum[p];
    // j = sum over pairs of c_i (h_i - h_{i+1}), nonzero components only
    R j[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R s = 0;
        bool any = false;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int c = lattice::D3Q19_C[2 * p + 1][a];
            if (!c) continue;
            const R t = c > 0 ? dif[p] : -dif[p];
            s = any ? s + t : t;
            any = true;
        }
        j[a] = s;
    }
    const R rho = (R)1 + d;
    R ir;
    if constexpr (sizeof(R) == 8) {
        ir = (R)1 / rho;
    } else {
        // the hardware reciprocal (about 1 ulp, not correctly rounded) and one
        // Newton step: without the step, the batched cavities' FP32 velocity
        // error against the FP64 run doubled (median 4.0e-6 -> 7.6e-6)
        const R r0 = __builtin_amdgcn_rcpf(rho);
        ir = fmadd(r0, fmadd(-rho, r0, (R)1), r0);
    }
    const R ux = j[0] * ir, uy = j[1] * ir, uz = j[2] * ir;
    const R usq = fmadd(ux, ux, fmadd(uy, uy, uz * uz));
    const R c0 = fmadd((R)-1.5 * rho, usq, d);
    const R r45 = (R)4.5 * rho;
    const R om1 = (R)1 - inv_tau;
    const R r3w = (R)3 * rho * inv_tau;
    h[0] = fmadd(om1, h[0], (inv_tau * (R)wq(0)) * c0);
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const int i = 2 * p + 1;
        const R cu = cdot(i, ux, uy, uz);
        const R S = (inv_tau * (R)wq(i)) * fmadd(r45 * cu, cu, c0);
        const R A = (r3w * (R)wq(i)) * cu;
        h[i] = fmadd(om1, h[i], S + A);
        h[i + 1] = fmadd(om1, h[i + 1], S - A);
    }
}

// Per-cell link codes (2 bits per direction in a u64; bit 63 = active fluid
// leaf): 0 regular pull, 1 bounce-back from a solid, 2 domain wall (moving
// wall term of the face), 3 pressure outlet: the population of the position
// clamped into the domain, its density set to the instance's rho0
// (outflow_link).
constexpr uint64_t kActiveBit = 1ull << 63;

// c2f capture indices per ghost: the covering coarse cell, then its coarse
// fluid neighbours along -x, +x, -y, +y, -z, +z (kNoCap when absent)
constexpr int kC2FCaps = 7;
// A capture buffer of ncap cells: Q population planes, then one record of
// (density, momentum x, y, z) per cell starting at cap_moments(ncap), which
// rounds up to a multiple of 4 elements so a record is one aligned 16 B (FP32)
// or 32 B (FP64) access.  Size: cap_elems(ncap).
__host__ __device__ constexpr size_t cap_moments(size_t ncap) { return ((size_t)Q * ncap + 3) & ~(size_t)3; }
__host__ __device__ constexpr size_t cap_elems(size_t ncap) { return cap_moments(ncap) + 4 * ncap; }
constexpr uint32_t kNoCap = 0xFFFFFFFFu;
enum : unsigned { kLinkRegular = 0, kLinkSolid = 1, kLinkWall = 2, kLinkOutflow = 3 };

// Per-instance constants (one record per instance of a batch).
struct InstDev {
    int32_t bc_type[6];   // faces -x,+x,-y,+y,-z,+z
    float bc_u[6][3];
    double bc_ud[6][3];
    double rho0;          // density in the moving-wall term
    double tau0;          // level-0 relaxation time
    double rho_init;      // initial equilibrium
    double u_init[3];
    int32_t ext0[3];      // level-0 extents in cells
    int32_t live;         // 0: the instance is not stepped any more
};

struct LevelDev {
    const int32_t* nbr;      // [brick][27], -1 = none
    const int32_t* bcoord;   // [brick][3]
    const uint32_t* bflags;  // bit 0: simple brick (all 64 active, all links regular);
                             // bit 1: a cell of the brick has a pressure-outlet link;
                             // bit 2: a 2x2x2 block of the brick is the children of a
                             // staged f2c ghost of the coarser level (f2c_blk);
                             // bit 3: a cell of the brick is a staged capture (cap_of)
    const uint64_t* cflags;  // [slot]
    const uint16_t* binst;   // [brick] instance id
    const float* itau;       // [instance] 1 / tau at this level
    const double* itau_d;
    const InstDev* inst;     // [instance]
    const uint32_t* tick;    // collide launches of this level so far (stochastic-rounding clock)
    size_t S;                // slot stride = nbricks * 64
    int32_t nactive;         // active bricks (they come first)
    int32_t level;
    int32_t outlet;          // 1 when a fluid leaf of the level has a pressure-outlet link
    // f2c staging (this level's collide feeding the coarser level's f2c fill):
    // f2c_blk[brick * 8 + block] = index of the coarse f2c ghost whose children
    // are that 2x2x2 block (kNoCap: none); when stage_on, the collide writes
    // those ghosts' child sums to stage ([23][n_stage], the coarse level's
    // Store<T>::R), block = (lx >> 1) + 2 (ly >> 1) + 4 (lz >> 1)
    const uint32_t* f2c_blk;
    void* stage;
    uint32_t n_stage;
    int32_t stage_on;
    // capture staging (this level's collide writing the capture that follows
    // it): cap_of[slot] = the slot's capture index (kNoCap: none, or not
    // staged); when cap_on, the collide writes those cells' pulled deviations
    // and (rho - 1, momentum) into cap_out (ncap captures, the k_capture layout)
    const uint32_t* cap_of;
    void* cap_out;
    uint32_t ncap;
    int32_t cap_on;
};

__host__ __device__ inline double tau_at(double tau0, int level) {
    return (tau0 - 0.5) * (double)(1u << level) + 0.5;
}

// Wall face for an outside source (axes z, y, x in that order, WALL faces only).
__device__ inline int wall_face(const InstDev& I, int level, int sx, int sy, int sz) {
    int p[3] = {sx, sy, sz};
    for (int a = 2; a >= 0; --a) {
        const int e = I.ext0[a] << level;
        int f = p[a] < 0 ? 2 * a : (p[a] >= e ? 2 * a + 1 : -1);
        if (f >= 0 && I.bc_type[f] == 0) return f;
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

// Pressure outlet (non-equilibrium extrapolation at a fixed density): the
// population of the source cell s2 -- the pulled position clamped into the
// domain -- with the source's density replaced by rho_out at the source's
// velocity,
//   f_i = f_i(s2) + (rho_out - rho_s) w_i (1 + 3 c_i.u_s + 4.5 (c_i.u_s)^2 - 1.5 u_s^2),
// i.e. feq_i(rho_out, u_s) + f_i(s2) - feq_i(rho_s, u_s) (feq is linear in
// rho).  A plain copy (zero gradient) leaves the density level free, and
// under a velocity inlet the mass then grows without bound; this anchors the
// outlet pressure.  kDev: A holds deviations h = f - w (rho_s = 1 + sum h,
// j_s = sum c h) and the result is a deviation.
template <typename T, typename R, bool kDev>
__device__ inline R outflow_link(const LevelDev& L, const T* __restrict__ A, size_t s2, int i, R rho_out) {
    R f[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j)
        f[j] = kDev ? Store<T>::dev(A, pidx(j, L.S, s2), j) : Store<T>::load(A, pidx(j, L.S, s2), j);
    R rho, jx, jy, jz;
    density_momentum(f, rho, jx, jy, jz);
    if (kDev) rho += (R)1;
    const R ir = (R)1 / rho;
    const R ux = jx * ir, uy = jy * ir, uz = jz * ir;
    const R cu = (R)cx(i) * ux + (R)cy(i) * uy + (R)cz(i) * uz;
    const R usq = ux * ux + uy * uy + uz * uz;
    return f[i] + (rho_out - rho) * (R)wq(i) * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - (R)1.5 * usq);
}

// A non-regular link of direction i of `slot` (brick b, local offsets px..):
// bounce-back from a solid, moving-wall bounce-back, or the pressure outlet.
// kOutlet = false leaves the outlet out of the code (levels whose LevelDev
// says outlet == 0 have no such link): its 19 extra loads, inlined into
// every direction of the collide's unrolled pull loop, slowed the whole
// kernel by a quarter even where no link takes them.
template <typename T, typename R, bool kOutlet = true>
__device__ inline R special_link(const LevelDev& L, const T* __restrict__ A, int b, size_t slot, int i, unsigned code,
                                 int px, int py, int pz) {
    if (code == kLinkSolid) return Store<T>::load(A, pidx(opp(i), L.S, slot), opp(i));
    const InstDev& I = L.inst[L.binst[b]];
    const int32_t* c3 = L.bcoord + 3 * b;
    if (code == kLinkWall) {
        const int face = wall_face(I, L.level, c3[0] * 4 + px, c3[1] * 4 + py, c3[2] * 4 + pz);
        R uw = 0;
        if (face >= 0) {
            if (sizeof(R) == 8)
                uw = (R)(cx(i) * I.bc_ud[face][0] + cy(i) * I.bc_ud[face][1] + cz(i) * I.bc_ud[face][2]);
            else
                uw = (R)((float)cx(i) * I.bc_u[face][0] + (float)cy(i) * I.bc_u[face][1] +
                         (float)cz(i) * I.bc_u[face][2]);
        }
        return Store<T>::load(A, pidx(opp(i), L.S, slot), opp(i)) + (R)6 * (R)wq(i) * (R)I.rho0 * uw;
    }
    if constexpr (!kOutlet) return (R)0;  // unreachable on a level without outlet links
    // outlet: the source clamped into the domain along the axes it leaves
    // through a non-periodic face; along a periodic axis the offset stays
    // as it is and the brick neighbour table (which wraps periodic faces)
    // resolves the image, as the layout's source_of does
    const int p[3] = {px, py, pz};
    int r[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int e = I.ext0[a] << L.level;
        const int g = c3[a] * 4 + p[a];
        r[a] = p[a];
        if (g < 0 && I.bc_type[2 * a] != FL_BC_PERIODIC) r[a] = -c3[a] * 4;
        else if (g >= e && I.bc_type[2 * a + 1] != FL_BC_PERIODIC) r[a] = e - 1 - c3[a] * 4;
    }
    const int64_t s2 = slot_of(L, b, r[0], r[1], r[2]);
    return outflow_link<T, R, false>(L, A, (size_t)s2, i, (R)I.rho0);
}

// special_link on deviations: w_opp(i) = w_i, so bounce-back and the wall
// term carry over unchanged.
template <typename T, typename R, bool kOutlet = true>
__device__ inline R special_link_dev(const LevelDev& L, const T* __restrict__ A, int b, size_t slot, int i,
                                     unsigned code, int px, int py, int pz) {
    if (code == kLinkSolid) return Store<T>::dev(A, pidx(opp(i), L.S, slot), opp(i));
    const InstDev& I = L.inst[L.binst[b]];
    const int32_t* c3 = L.bcoord + 3 * b;
    if (code == kLinkWall) {
        const int face = wall_face(I, L.level, c3[0] * 4 + px, c3[1] * 4 + py, c3[2] * 4 + pz);
        R uw = 0;
        if (face >= 0) {
            if (sizeof(R) == 8)
                uw = (R)(cx(i) * I.bc_ud[face][0] + cy(i) * I.bc_ud[face][1] + cz(i) * I.bc_ud[face][2]);
            else
                uw = (R)((float)cx(i) * I.bc_u[face][0] + (float)cy(i) * I.bc_u[face][1] +
                         (float)cz(i) * I.bc_u[face][2]);
        }
        return Store<T>::dev(A, pidx(opp(i), L.S, slot), opp(i)) + (R)6 * (R)wq(i) * (R)I.rho0 * uw;
    }
    if constexpr (!kOutlet) return (R)0;  // unreachable on a level without outlet links
    const int p[3] = {px, py, pz};
    int r[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int e = I.ext0[a] << L.level;
        const int g = c3[a] * 4 + p[a];
        r[a] = p[a];
        if (g < 0 && I.bc_type[2 * a] != FL_BC_PERIODIC) r[a] = -c3[a] * 4;
        else if (g >= e && I.bc_type[2 * a + 1] != FL_BC_PERIODIC) r[a] = e - 1 - c3[a] * 4;
    }
    const int64_t s2 = slot_of(L, b, r[0], r[1], r[2]);
    return outflow_link<T, R, true>(L, A, (size_t)s2, i, (R)I.rho0);
}

// Offset of the source that cell (lx, ly, lz) of a brick pulls in direction
// i, from plane i's base (A + plane_base(i, S)), in units of 1 / kScale
// elements (kScale = 2 gives byte offsets into 16-bit storage, below 2^32).
// NB: the brick's 27 neighbour entries (13 = the brick itself); the source
// brick is one of at most four of them, chosen by whether the pulled
// position leaves the brick along each axis (static indices only, so NB
// stays in registers).  Bricks per level are below 2^21 (the layout's
// limit): a 24-bit multiply (full rate; the 3-plane group's stride 192 is
// not a power of two).
template <unsigned kScale = 1>
__device__ __attribute__((always_inline)) inline unsigned pull_offset(const int (&NB)[27], int lx, int ly, int lz,
                                                                      int i) {
    const int px = lx - cx(i), py = ly - cy(i), pz = lz - cz(i);
    const bool ex = (unsigned)px > 3u, ey = (unsigned)py > 3u, ez = (unsigned)pz > 3u;
    const int kx = -cx(i), ky = -3 * cy(i), kz = -9 * cz(i);
    const int n000 = NB[13], n100 = NB[13 + kx], n010 = NB[13 + ky], n110 = NB[13 + kx + ky];
    const int n001 = NB[13 + kz], n101 = NB[13 + kx + kz], n011 = NB[13 + ky + kz], n111 = NB[13 + kx + ky + kz];
    const int nb = ez ? (ey ? (ex ? n111 : n011) : (ex ? n101 : n001)) : (ey ? (ex ? n110 : n010) : (ex ? n100 : n000));
    return __umul24((unsigned)nb, (unsigned)brick_stride(i) * kScale) +
           kScale * (unsigned)((px & 3) + 4 * (py & 3) + 16 * (pz & 3));
}

// The neighbour-row entries one cell can pull from: its own brick and the
// seven bricks on its octant's side.  A cell with lx < 2 leaves the brick
// only towards -x (lx = 0 pulling along +x), one with lx >= 2 only towards
// +x; likewise along y and z.  o[dz][dy][dx] = row[13 + dx sx + 3 dy sy +
// 9 dz sz] with s = -1 / +1 by octant: 8 loads in place of the row's 27.
struct Octant {
    int o000, o001, o010, o011, o100, o101, o110, o111;
};
__device__ __attribute__((always_inline)) inline Octant octant_row(const int* __restrict__ row, int lx, int ly,
                                                                   int lz) {
    const int sx = lx < 2 ? -1 : 1, sy = ly < 2 ? -3 : 3, sz = lz < 2 ? -9 : 9;
    return Octant{row[13],      row[13 + sx],      row[13 + sy],      row[13 + sy + sx],
                  row[13 + sz], row[13 + sz + sx], row[13 + sz + sy], row[13 + sz + sy + sx]};
}

// pull_offset from the octant's row entries (octant_row): a pulled position
// that leaves the brick along an axis leaves it on the octant's side.
__device__ __attribute__((always_inline)) inline unsigned pull_offset_oct(const Octant& O, int lx, int ly, int lz,
                                                                          int i) {
    const int px = lx - cx(i), py = ly - cy(i), pz = lz - cz(i);
    const bool ex = (unsigned)px > 3u, ey = (unsigned)py > 3u, ez = (unsigned)pz > 3u;
    // select between values, not between the fields themselves: a choice of
    // lvalues becomes a load through a chosen address, which keeps the
    // struct in scratch memory
    const int a000 = O.o000, a001 = O.o001, a010 = O.o010, a011 = O.o011;
    const int a100 = O.o100, a101 = O.o101, a110 = O.o110, a111 = O.o111;
    const int nb = ez ? (ey ? (ex ? a111 : a110) : (ex ? a101 : a100)) : (ey ? (ex ? a011 : a010) : (ex ? a001 : a000));
    return __umul24((unsigned)nb, (unsigned)brick_stride(i)) + (unsigned)((px & 3) + 4 * (py & 3) + 16 * (pz & 3));
}

// Generic pull of the 19 incoming populations of one slot (list kernels).
// w: the slot's link word (L.cflags[slot]), which callers may load early.
template <typename T, typename R>
__device__ inline void gather_slot(const LevelDev& L, const T* __restrict__ A, int64_t slot, uint64_t w, R* f) {
    const int b = (int)(slot >> 6);
    const int l = (int)(slot & 63);
    const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
    if ((w & ((1ull << (2 * Q)) - 1)) == 0) {
        // every link regular (the common case, e.g. the children at a level
        // interface): the cell's octant of the neighbour row first, then 19
        // independent population loads (a regular link's source brick exists)
        const Octant O = octant_row(L.nbr + (size_t)b * 27, lx, ly, lz);
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = Store<T>::load(A + plane_base(i, L.S), pull_offset_oct(O, lx, ly, lz, i), i);
        return;
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const unsigned code = (unsigned)(w >> (2 * i)) & 3u;
        const int px = lx - cx(i), py = ly - cy(i), pz = lz - cz(i);
        if (code == kLinkRegular)
            f[i] = Store<T>::load(A, pidx(i, L.S, (size_t)slot_of(L, b, px, py, pz)), i);
        else
            f[i] = special_link<T, R>(L, A, b, (size_t)slot, i, code, px, py, pz);
    }
}

template <typename T, typename R>
__device__ inline void gather_slot(const LevelDev& L, const T* __restrict__ A, int64_t slot, R* f) {
    gather_slot<T, R>(L, A, slot, L.cflags[slot], f);
}

// gather_slot on deviations h = f - w (the fills' and captures' input).
template <typename T, typename R>
__device__ inline void gather_slot_dev(const LevelDev& L, const T* __restrict__ A, int64_t slot, uint64_t w, R* h) {
    const int b = (int)(slot >> 6);
    const int l = (int)(slot & 63);
    const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
    if ((w & ((1ull << (2 * Q)) - 1)) == 0) {
        const Octant O = octant_row(L.nbr + (size_t)b * 27, lx, ly, lz);
#pragma unroll
        for (int i = 0; i < Q; ++i) h[i] = Store<T>::dev(A + plane_base(i, L.S), pull_offset_oct(O, lx, ly, lz, i), i);
        return;
    }
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const unsigned code = (unsigned)(w >> (2 * i)) & 3u;
        const int px = lx - cx(i), py = ly - cy(i), pz = lz - cz(i);
        if (code == kLinkRegular)
            h[i] = Store<T>::dev(A, pidx(i, L.S, (size_t)slot_of(L, b, px, py, pz)), i);
        else
            h[i] = special_link_dev<T, R>(L, A, b, (size_t)slot, i, code, px, py, pz);
    }
}

template <typename T, typename R>
__device__ inline void gather_slot_dev(const LevelDev& L, const T* __restrict__ A, int64_t slot, R* h) {
    gather_slot_dev<T, R>(L, A, slot, L.cflags[slot], h);
}

}  // namespace lbm
}  // namespace fl
