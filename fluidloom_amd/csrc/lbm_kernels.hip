// a14/a15: D3Q19 lattice kernels for one refinement level of a batch of
// instances (a single solve is a batch of one).
//
//   k_collide   fused pull-stream + BGK over the active bricks: one 64-lane
//               wave per 4x4x4 brick, lane = cell.  The brick's 27-entry
//               neighbour row is held in scalar registers, so a regular
//               pull costs one population load plus a few lane selects;
//               cells of "simple" bricks (all 64 active, no wall links)
//               never read their link word.  Variants: BPW = 1..16 bricks
//               (waves) per workgroup, with or without an XCD-aware remap
//               (each XCD takes one contiguous Hilbert range of bricks), or
//               persistent; the solver's autotuner picks one per level.
//               16-bit storages (FP16 f - w, FP16C) round the output
//               stochastically, with bits drawn from the cell's coordinates,
//               level and the level's step clock.
//   k_capture   pre-collision state of coarse cells that feed c2f ghosts
//   k_fill_c2f  coarse -> fine ghost populations: density and velocity
//               interpolated linearly from the covering coarse cell and its
//               coarse fluid neighbours, f_neq rescaled
//   k_fill_f2c  fine -> coarse ghost populations (8-child average)
//   k_init, k_brick_sums, k_brick_du, k_inst_reduce, k_link_force,
//   k_extract, k_inject, k_vorticity, k_remap_*, k_ghost_eq_*: set-up,
//   deterministic FP64 reductions per instance, queries, regrid transfer.
#include <type_traits>

#include "fl_internal.hpp"
#include "lbm_device.hpp"
#include "lbm_kernels.hpp"
#include "wave.hpp"

namespace fl {
namespace lbm {
namespace {

constexpr int kWG = 256;

__device__ inline unsigned xcd_remap(unsigned bid, unsigned nblk) {
    const unsigned q = nblk / 8, r = nblk % 8;
    const unsigned x = bid % 8, k = bid / 8;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// XCD-aware in chunks of C workgroups: the k-th workgroup dispatched to XCD x
// (workgroup b goes to XCD b % 8) takes item k % C of chunk (k / C) * 8 + x.
// Each XCD then works through runs of C consecutive groups (a brick's
// neighbours mostly in its own L2) while the eight XCDs stay inside one
// window of 8C groups.  Groups past the last whole window keep their index.
__device__ inline unsigned xcd_chunk_remap(unsigned bid, unsigned nblk, unsigned C) {
    const unsigned win = 8u * C;
    if (bid >= (nblk / win) * win) return bid;
    const unsigned x = bid % 8u, k = bid / 8u;
    return ((k / C) * 8u + x) * C + k % C;
}

// Rounding bits of slot s (brick b = s >> 6, cell s & 63) of a level: the
// brick's coordinates inside its instance, the cell, the level and the
// level's step clock, so a ghost rounds the same way alone and in a batch.
__device__ inline uint32_t slot_bits(const LevelDev& L, size_t s) {
    const int32_t* c3 = L.bcoord + 3 * (s >> 6);
    return mix32(mix32((uint32_t)c3[0] * 0x9E3779B1u ^ mix32((uint32_t)c3[1] * 0x85EBCA77u ^
                                                              mix32((uint32_t)c3[2] * 0xC2B2AE3Du)) ^
                       ((uint32_t)L.level << 24) ^ (*L.tick * 0x27D4EB2Fu)) +
                 (uint32_t)(s & 63));
}

// Ghost population store of the fills, from the deviation h = f - w:
// stochastic rounding for the 16-bit storages (a ghost is interpolated afresh
// every sub-step from slowly varying values, so round-to-nearest would feed
// the fine level the same error at every step), plain stores otherwise.
// r: the slot's bits; direction i draws r + i * 0x9E3779B9, as the collide
// does.
template <typename T>
__device__ inline void store_ghost(T* p, size_t idx, int i, typename Store<T>::R h, uint32_t r) {
    if constexpr (std::is_same<T, _Float16>::value || std::is_same<T, fp16c>::value)
        Store<T>::store_dev_sr(p, idx, i, h, r + (uint32_t)i * 0x9E3779B9u);
    else
        Store<T>::store_dev(p, idx, i, h);
}

template <typename R>
__device__ inline R inv_tau_of(const LevelDev& L, int inst) {
    if (sizeof(R) == 8) return (R)L.itau_d[inst];
    return (R)L.itau[inst];
}

// v of the lane whose index differs in the bits of m (m < 32): ds_swizzle in
// its bit mode (and 0x1F, xor m), no lane-index arithmetic
template <int M, typename R>
__device__ __attribute__((always_inline)) inline R swz_xor(R v) {
    constexpr int kPat = (M << 10) | 0x1F;
    if constexpr (sizeof(R) == 8) {
        const long long x = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_ds_swizzle((int)(x & 0xFFFFFFFFll), kPat);
        const int hi = __builtin_amdgcn_ds_swizzle((int)(x >> 32), kPat);
        return __builtin_bit_cast(R, ((long long)hi << 32) | (long long)(unsigned)lo);
    } else {
        return __builtin_bit_cast(R, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), kPat));
    }
}

// f2c staging, run by the collide of a fine level's second sub-step on the
// bricks that hold children of staged coarse f2c ghosts (bflags bit 2).  h:
// the cell's pulled deviations, the values the coarse level's f2c fill would
// gather from this collide's input at its next step (nothing writes that
// buffer in between).  Each child contributes (rho - 1, momentum, f_neq) as
// in k_fill_f2c; the 2x2x2 block's sum is formed across lanes 1, 4 and 16
// apart (x, then y, then z: group8_sum's tree, so the same bits), and the
// block's eight lanes store the 23 sums of their ghost.  Staged children
// have regular links only (k_stage_map), so their pulled values are the
// fill's gather bit for bit.  The exchanges are
// ds_swizzle: __shfl_xor's lane-index arithmetic here made the 16-bit
// collide instances allocate all 256 registers.  Every lane of a
// staged block is an active fluid leaf (the map admits no other), so the
// shuffles within the block read live lanes.
template <typename R>
__device__ __attribute__((always_inline)) inline void stage_children(const LevelDev& L, const R* h, int b, int lane) {
    const int blk = ((lane >> 1) & 1) | ((lane >> 2) & 2) | ((lane >> 3) & 4);
    const int q = (lane & 1) | ((lane >> 1) & 2) | ((lane >> 2) & 4);
    const uint32_t k = L.f2c_blk[(size_t)b * 8 + blk];
    R d, jx, jy, jz;
    dev_density_momentum(h, d, jx, jy, jz);
    const R rho = (R)1 + d;
    const R ir = (R)1 / rho;
    const R vx = jx * ir, vy = jy * ir, vz = jz * ir;
    const R usq = vx * vx + vy * vy + vz * vz;
    R* __restrict__ st = static_cast<R*>(L.stage);
#pragma unroll
    for (int j = 0; j < 4 + Q; ++j) {
        R v = j == 0 ? d : j == 1 ? jx : j == 2 ? jy : j == 3 ? jz
                     : h[j - 4] - feq_dev(j - 4, d, rho, vx, vy, vz, usq);
        v += swz_xor<1>(v);
        v += swz_xor<4>(v);
        v += swz_xor<16>(v);
        if (k != kNoCap && (j & 7) == q) st[(size_t)j * L.n_stage + k] = v;
    }
}

// One wave per brick, lane = cell.  The brick's 27 neighbour entries are
// wave-uniform loads (scalar registers); for direction i a lane's source
// brick is one of at most four of them, chosen by whether the pulled
// position leaves the brick along each axis.  Offsets from a plane's base
// are 32-bit (below 8 * slots, and the layout keeps slots below 2^27), on a
// wave-uniform plane base.
// kStaging: the f2c and capture staging code (LevelDev::stage_on, cap_on)
// is compiled in.  The outlet instances leave it out, also on the path they
// delegate plain bricks to: a second copy of it in one kernel body made the
// 16-bit outlet instances allocate all 256 registers.  The solver never
// asks an outlet level to stage.
template <typename T, bool kOutlet, bool kStaging = !kOutlet>
__device__ __attribute__((always_inline)) inline void collide_brick(const LevelDev& L, const T* __restrict__ A,
                                                                    T* __restrict__ B, int b_in) {
    using R = typename Store<T>::R;
    const int b = __builtin_amdgcn_readfirstlane(b_in);
    if (b >= L.nactive) return;
    // What the population loads need comes from b alone: the neighbour row and
    // the brick flags are one round of scalar loads.  The instance (its live
    // flag and relaxation rate) only matters to the collision and the stores,
    // so its two dependent loads overlap the population loads, and the live
    // flag is tested after them; a stopped instance keeps its layout, so
    // reading its populations stays in bounds.
    // binst is 16-bit: the wave-uniform dword holding it keeps the load on
    // the scalar unit (device allocations are whole multiples of 256 B, so
    // the dword's other half is inside the array).
    const int* __restrict__ row = L.nbr + (size_t)b * 27;
    int NB[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) NB[k] = row[k];
    const uint32_t bf = L.bflags[b];
    const uint32_t ipair = reinterpret_cast<const uint32_t*>(L.binst)[b >> 1];
    if constexpr (kOutlet) {
        // only the bricks with an outlet link take the outlet code
        if (!(bf & 2u)) {
            collide_brick<T, false, false>(L, A, B, b);
            return;
        }
    }
    const int inst = (int)((b & 1) ? ipair >> 16 : ipair & 0xFFFFu);
    // The instance's live flag and relaxation rate, and in 16-bit storage the
    // brick's rounding hash, are read here, ahead of every store, so they are
    // scalar loads issued before the population loads.  Read where they are
    // used, after a possible staging store, they were vector loads issued
    // only once the populations had arrived.
    const int live = L.inst[inst].live;
    const R itau = inv_tau_of<R>(L, inst);
    constexpr bool k16 = std::is_same<T, _Float16>::value || std::is_same<T, fp16c>::value;
    uint32_t hb = 0;
    if constexpr (k16) {
        // rounding bits from the cell's own coordinates, level and step, so an
        // instance rounds the same way alone and inside any batch: the brick's
        // coordinates, level and step hashed once per wave (wave-uniform:
        // scalar-unit multiplies)
        const int32_t* c3 = L.bcoord + 3 * b;
        hb = mix32((uint32_t)c3[0] * 0x9E3779B1u ^
                   mix32((uint32_t)c3[1] * 0x85EBCA77u ^
                         mix32((uint32_t)c3[2] * 0xC2B2AE3Du ^ ((uint32_t)L.level << 24) ^ (*L.tick * 0x27D4EB2Fu))));
    }
    const int lane = (int)(threadIdx.x & 63);
    const bool simple = (bf & 1u) != 0;
    const unsigned slot = (unsigned)b * 64u + (unsigned)lane;
    const uint64_t w = simple ? 0ull : L.cflags[slot];
    const bool active = simple || (w & kActiveBit);
    const int lx = lane & 3, ly = (lane >> 2) & 3, lz = lane >> 4;
    const size_t S = L.S;
    // pulled source of direction i: offset from plane i's base (A + plane_base)
    auto source = [&](int i) -> unsigned { return pull_offset(NB, lx, ly, lz, i); };
    // pulled deviation of direction i in 16-bit storage: byte offsets from a
    // scalar plane base
    auto dev16 = [&](int i) -> R {
        if constexpr (sizeof(T) == 2)
            return Store<T>::dev_at(A + plane_base(i, S), pull_offset<2>(NB, lx, ly, lz, i), i);
        else
            return (R)0;
    };
    // The cell's populations as deviations h = f - w: density and momentum
    // then come from sums of small numbers, so neither the weights' rounding
    // to the storage type nor the low bits that 16-bit storage leaves fixed
    // in every w + h bias them.  16-bit storage reads h directly; FP32 and
    // FP64 read f and subtract w after the last load has been issued (a
    // subtraction next to each load made the compiler wait on the loads one
    // by one: 2.21 ms against 1.74 ms for the sphere's level-3 launch).
    constexpr bool kDevLoads = sizeof(T) == 2;
    R f[Q];
    if (sizeof(T) == 2 && simple) {
        // 16-bit storage, every cell fluid, every link regular: 19 independent
        // loads back to back.  FP32 and FP64 keep the interleaved form below,
        // which measured faster for FP32.
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = dev16(i);
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            f[i] = 0;
            if (active) {
                const unsigned code = (unsigned)(w >> (2 * i)) & 3u;
                if (code == kLinkRegular)
                    f[i] = kDevLoads ? dev16(i) : Store<T>::load(A + plane_base(i, S), source(i), i);
                else
                    f[i] = kDevLoads ? special_link_dev<T, R, kOutlet>(L, A, b, slot, i, code, lx - cx(i), ly - cy(i),
                                                                       lz - cz(i))
                                     : special_link<T, R, kOutlet>(L, A, b, slot, i, code, lx - cx(i), ly - cy(i),
                                                                   lz - cz(i));
            }
        }
    }
    if (!active || !live) return;
    if (!kDevLoads) {
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] -= (R)wq(i);
    }
    if constexpr (kStaging) {
    if (L.stage_on && (bf & 4u)) stage_children<R>(L, f, b, lane);
    if (L.cap_on && (bf & 8u)) {
        // the capture after this collide gathers these very values (k_capture
        // on the collide's input); staged captures have regular links only,
        // so the pulled values are the gather's bit for bit
        const uint32_t kc = L.cap_of[slot];
        if (kc != kNoCap) {
            R m[4];
            dev_density_momentum(f, m[0], m[1], m[2], m[3]);
            R* __restrict__ cap = static_cast<R*>(L.cap_out);
#pragma unroll
            for (int i = 0; i < Q; ++i) cap[(size_t)i * L.ncap + kc] = f[i];
            typedef R R4 __attribute__((ext_vector_type(4)));
            *reinterpret_cast<R4*>(cap + cap_moments(L.ncap) + 4 * (size_t)kc) = R4{m[0], m[1], m[2], m[3]};
        }
    }
    }
    bgk_dev(f, itau);
    if constexpr (k16) {
        // One hash of the brick's hash hb plus the lane gives the cell's h.
        // Direction i draws h + i * 0x9E3779B9 (a Weyl sequence: one add per
        // direction, where the kernel is bound by VALU issue in 16-bit
        // storage).  Each draw is as uniform over cells and steps as h, so
        // each direction's rounding stays unbiased; the 19 draws of one cell
        // are spread evenly over the 32-bit range.
        const uint32_t h = mix32(hb + (uint32_t)lane);
        if constexpr (std::is_same<T, _Float16>::value) {
            // v_cvt_sr_f16_f32 writes one half of its register and keeps the
            // other: converting every direction into the same register keeps
            // a fresh zeroed register (one move) per direction out
            typedef _Float16 h2 __attribute__((ext_vector_type(2)));
            h2 acc = {(_Float16)0, (_Float16)0};
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                acc = __builtin_amdgcn_cvt_sr_f16_f32(acc, f[i], h + (uint32_t)i * 0x9E3779B9u, false);
                __builtin_nontemporal_store(acc[0], B + plane_base(i, S) + (size_t)b * brick_stride(i) + lane);
            }
        } else {
#pragma unroll
            for (int i = 0; i < Q; ++i)
                Store<T>::store_dev_sr(B + plane_base(i, S), (unsigned)b * (unsigned)brick_stride(i) + (unsigned)lane,
                                       i, f[i], h + (uint32_t)i * 0x9E3779B9u);
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i)
            Store<T>::store_dev_nt(B + plane_base(i, S), (unsigned)b * (unsigned)brick_stride(i) + (unsigned)lane, i,
                                   f[i]);
    }
}

// MODE 0: workgroup g takes brick group g (consecutive groups land on
// different XCDs); 1: XCD-aware remap, each XCD one contiguous eighth of the
// groups; 2: persistent -- a grid sized to the resident capacity, each
// workgroup walking a contiguous chunk of groups (chunks of one XCD
// adjacent); 3 / 4 / 5 / 6: XCD-aware in chunks of 64 / 128 / 256 / 512
// bricks (xcd_chunk_remap).
// On the 3-level sphere's finest level (980k bricks) the chunked forms are
// the fastest: mode 1 keeps neighbours in one L2 but streams eight distant
// windows, mode 0 streams one window but splits neighbours over all L2s.
// kOutlet: the level has pressure-outlet links (LevelDev::outlet).
template <typename T, int BPW, int MODE, bool kOutlet>
__global__ __launch_bounds__(64 * BPW) void k_collide(LevelDev L, const T* __restrict__ A, T* __restrict__ B,
                                                      unsigned ngroups) {
    const int wave = (int)(threadIdx.x >> 6);
    if (MODE == 2) {
        const unsigned wg = xcd_remap(blockIdx.x, gridDim.x);
        const unsigned per = (ngroups + gridDim.x - 1) / gridDim.x;
        const unsigned g0 = wg * per;
        const unsigned g1 = g0 + per < ngroups ? g0 + per : ngroups;
        for (unsigned g = g0; g < g1; ++g) collide_brick<T, kOutlet>(L, A, B, (int)(g * BPW) + wave);
    } else {
        constexpr unsigned kChunk = MODE == 3 ? 64u : MODE == 4 ? 128u : MODE == 5 ? 256u : 512u;
        const unsigned g = MODE == 1   ? xcd_remap(blockIdx.x, gridDim.x)
                           : MODE >= 3 ? xcd_chunk_remap(blockIdx.x, gridDim.x, kChunk / BPW)
                                       : blockIdx.x;
        collide_brick<T, kOutlet>(L, A, B, (int)(g * BPW) + wave);
    }
}

// Bricks with a pressure-outlet link (a 2-bit link code of 3 among a cell's
// 19 directions) get bit 1 of their flags, and *flag |= 1 when the level has
// any.  A wave covers one brick per round (the grid stride is a multiple of
// 64).
__global__ __launch_bounds__(kWG) void k_any_outlet(const uint64_t* __restrict__ cflags, uint32_t* __restrict__ bflags,
                                                    size_t n, uint32_t* __restrict__ flag) {
    constexpr uint64_t kEven = 0x1555555555ull;  // bit 2i of each direction i = 0..18
    bool any = false;
    for (size_t s = (size_t)blockIdx.x * kWG + threadIdx.x; s < n; s += (size_t)gridDim.x * kWG) {
        const uint64_t w = cflags[s];
        const bool o = (w & kActiveBit) && (w & (w >> 1) & kEven);
        if (__ballot(o) && (threadIdx.x & 63) == 0) atomicOr(&bflags[s >> 6], 2u);
        any = any || o;
    }
    if (__ballot(any) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

__global__ __launch_bounds__(kWG) void k_fill_u32_kernel(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * kWG + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_tick(uint32_t* tick) {
    if (threadIdx.x == 0) *tick += 1u;
}

// skip (optional): captures the collide already wrote (LevelDev::cap_on)
template <typename T>
__global__ __launch_bounds__(kWG) void k_capture(LevelDev L, const T* __restrict__ A,
                                                 const uint32_t* __restrict__ cap_slots, uint32_t ncap,
                                                 typename Store<T>::R* __restrict__ cap,
                                                 const uint8_t* __restrict__ skip) {
    using R = typename Store<T>::R;
    // deviations h = f - w and the record (rho - 1, momentum)
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < ncap; k += gridDim.x * kWG) {
        if (skip && skip[k]) continue;
        R f[Q], m[4];
        gather_slot_dev<T, R>(L, A, cap_slots[k], f);
        dev_density_momentum(f, m[0], m[1], m[2], m[3]);
#pragma unroll
        for (int i = 0; i < Q; ++i) cap[(size_t)i * ncap + k] = f[i];
        typedef R R4 __attribute__((ext_vector_type(4)));
        *reinterpret_cast<R4*>(cap + cap_moments(ncap) + 4 * (size_t)k) = R4{m[0], m[1], m[2], m[3]};
    }
}

// factor of the c2f fill for a ghost at level l of an instance with tau0
__device__ inline double c2f_factor(double tau0, int l) {
    const double tf = tau_at(tau0, l), tc = tau_at(tau0, l - 1);
    return (1.0 - 1.0 / tf) * (tf / (2.0 * tc));
}
__device__ inline double f2c_factor(double tau0, int l) {
    const double tc = tau_at(tau0, l), tf = tau_at(tau0, l + 1);
    return (1.0 - 1.0 / tc) * (2.0 * tc / tf);
}

// Coarse -> fine, eight lanes per group of ghosts sharing a covering coarse
// leaf P (lane q = the child at fine-coordinate parity x + 2y + 4z).  The
// group loads P's time-interpolated populations and its neighbours'
// time-interpolated density and momentum once, three values per lane, and
// shares them with shuffles.  Each ghost then takes P's moments plus a
// linear correction along each axis to its own centre (offset 1/4 of a
// coarse cell; central slope when both coarse neighbours along the axis are
// fluid leaves, one-sided when one is, none otherwise) and P's rescaled
// non-equilibrium part.
template <typename T>
__global__ __launch_bounds__(kWG) void k_fill_c2f(LevelDev Lf, const uint32_t* __restrict__ gchild,
                                                  const uint32_t* __restrict__ gcap,
                                                  const uint32_t* __restrict__ gmask, uint32_t ngroups,
                                                  const typename Store<T>::R* __restrict__ cap0,
                                                  const typename Store<T>::R* __restrict__ cap1, uint32_t ncap,
                                                  typename Store<T>::R alpha, T* __restrict__ A) {
    using R = typename Store<T>::R;
    const uint32_t t = blockIdx.x * kWG + threadIdx.x;
    const uint32_t k = t >> 3;
    const int q = (int)(t & 7u);
    // every lane of a group shares k, so a group is wholly in or out
    const bool on = k < ngroups;
    uint32_t ci[kC2FCaps];
    uint32_t g = kNoCap, want = 0;
    bool live = false;
    double tau0 = 0;
    R pv[3] = {0, 0, 0}, mv[4] = {0, 0, 0, 0};
    if (on) {
#pragma unroll
        for (int j = 0; j < kC2FCaps; ++j) ci[j] = gcap[(size_t)kC2FCaps * k + j];
        g = gchild[(size_t)8 * k + q];
        want = gmask[(size_t)8 * k + q];
        // the ghost's instance, looked up while the capture loads are in flight
        if (g != kNoCap) {
            const InstDev& I = Lf.inst[Lf.binst[g >> 6]];
            live = I.live != 0;
            tau0 = I.tau0;
        }
        // alpha == 0 (the first fine sub-step of a coarse step) needs only the
        // earlier capture: the same values, half the loads
        const bool two = alpha != (R)0;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int i = q + 8 * r;  // population i of P
            if (i < Q) {
                const R a0 = cap0[(size_t)i * ncap + ci[0]];
                pv[r] = two ? ((R)1 - alpha) * a0 + alpha * cap1[(size_t)i * ncap + ci[0]] : a0;
            }
        }
        // lane a < 6: the moment record of neighbour a, one access per capture
        if (q < 6) {
            const uint32_t c = ci[1 + q];
            if (c != kNoCap) {
                typedef R R4 __attribute__((ext_vector_type(4)));
                const R4 a0 = *reinterpret_cast<const R4*>(cap0 + cap_moments(ncap) + 4 * (size_t)c);
                if (two) {
                    const R4 a1 = *reinterpret_cast<const R4*>(cap1 + cap_moments(ncap) + 4 * (size_t)c);
                    mv[0] = ((R)1 - alpha) * a0.x + alpha * a1.x;
                    mv[1] = ((R)1 - alpha) * a0.y + alpha * a1.y;
                    mv[2] = ((R)1 - alpha) * a0.z + alpha * a1.z;
                    mv[3] = ((R)1 - alpha) * a0.w + alpha * a1.w;
                } else {
                    mv[0] = a0.x;
                    mv[1] = a0.y;
                    mv[2] = a0.z;
                    mv[3] = a0.w;
                }
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < kC2FCaps; ++j) ci[j] = kNoCap;
    }
    R f[Q], nm[24];
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = __shfl(pv[i >> 3], i & 7, 8);
    // nm[4a + c]: moment c of neighbour a, held by lane a
#pragma unroll
    for (int j = 0; j < 24; ++j) nm[j] = __shfl(mv[j & 3], j >> 2, 8);
    if (!live) return;
    const R fac = (R)c2f_factor(tau0, Lf.level);
    // f[] holds P's deviations; moments as (rho - 1, u), the neighbours'
    // records as (rho - 1, momentum)
    R mp[4], rp;
    dev_moments(f, mp[0], rp, mp[1], mp[2], mp[3]);
    R phi[4] = {mp[0], mp[1], mp[2], mp[3]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const bool hm = ci[1 + 2 * a] != kNoCap, hq = ci[2 + 2 * a] != kNoCap;
        R mm[4] = {0, 0, 0, 0}, mq[4] = {0, 0, 0, 0};
        if (hm) {
            const R* v = nm + 4 * (2 * a);
            const R ir = (R)1 / ((R)1 + v[0]);
            mm[0] = v[0];
            mm[1] = v[1] * ir;
            mm[2] = v[2] * ir;
            mm[3] = v[3] * ir;
        }
        if (hq) {
            const R* v = nm + 4 * (2 * a + 1);
            const R ir = (R)1 / ((R)1 + v[0]);
            mq[0] = v[0];
            mq[1] = v[1] * ir;
            mq[2] = v[2] * ir;
            mq[3] = v[3] * ir;
        }
        const R dl = ((q >> a) & 1) ? (R)0.25 : (R)-0.25;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            R gr = 0;
            if (hm && hq) gr = (R)0.5 * (mq[j] - mm[j]);
            else if (hq) gr = mq[j] - mp[j];
            else if (hm) gr = mp[j] - mm[j];
            phi[j] += dl * gr;
        }
    }
    const R usq = phi[1] * phi[1] + phi[2] * phi[2] + phi[3] * phi[3];
    const R usp = mp[1] * mp[1] + mp[2] * mp[2] + mp[3] * mp[3];
    const R rphi = (R)1 + phi[0];
    // only the populations some fluid leaf pulls from this ghost
    const uint32_t rb = slot_bits(Lf, g);
    auto put = [&](int i, R e, R em) {
        if ((want >> i) & 1u) store_ghost<T>(A, pidx(i, Lf.S, g), i, e + fac * (f[i] - em), rb);
    };
    // Equilibrium deviations (feq_dev) of the interpolated state (phi) and of
    // P's own over the opposite pairs (i, i+1), which share |c_i . u| and the
    // weight: feq_dev = w (c0 + 4.5 rho cu^2) +- 3 w rho cu, c0 = d - 1.5 rho u^2.
    const R c0p = fmadd((R)-1.5 * rphi, usq, phi[0]), r45p = (R)4.5 * rphi, r3p = (R)3 * rphi;
    const R c0m = fmadd((R)-1.5 * rp, usp, mp[0]), r45m = (R)4.5 * rp, r3m = (R)3 * rp;
    put(0, (R)wq(0) * c0p, (R)wq(0) * c0m);
#pragma unroll
    for (int k = 0; k < (Q - 1) / 2; ++k) {
        const int i = 2 * k + 1;
        const R cp = cdot(i, phi[1], phi[2], phi[3]), cm = cdot(i, mp[1], mp[2], mp[3]);
        const R sp = (R)wq(i) * fmadd(r45p * cp, cp, c0p), ap = ((R)wq(i) * r3p) * cp;
        const R sm = (R)wq(i) * fmadd(r45m * cm, cm, c0m), am = ((R)wq(i) * r3m) * cm;
        put(i, sp + ap, sm + am);
        put(i + 1, sp - ap, sm - am);
    }
}

template <typename T>
__device__ inline void child_average(const LevelDev& Lf, const T* __restrict__ Af, const uint32_t* c8,
                                     typename Store<T>::R& r, typename Store<T>::R& ux, typename Store<T>::R& uy,
                                     typename Store<T>::R& uz, typename Store<T>::R* aneq) {
    using R = typename Store<T>::R;
    R sr = 0, sjx = 0, sjy = 0, sjz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) aneq[i] = 0;
    for (int q = 0; q < 8; ++q) {
        R f[Q];
        gather_slot<T, R>(Lf, Af, c8[q], f);
        R rho, vx, vy, vz;
        moments(f, rho, vx, vy, vz);
        const R usq = vx * vx + vy * vy + vz * vz;
        sr += rho;
        sjx += rho * vx;
        sjy += rho * vy;
        sjz += rho * vz;
#pragma unroll
        for (int i = 0; i < Q; ++i) aneq[i] += f[i] - feq(i, rho, vx, vy, vz, usq);
    }
    r = sr * (R)0.125;
    ux = (sjx * (R)0.125) / r;
    uy = (sjy * (R)0.125) / r;
    uz = (sjz * (R)0.125) / r;
#pragma unroll
    for (int i = 0; i < Q; ++i) aneq[i] *= (R)0.125;
}

// Fine -> coarse: eight lanes per ghost, lane q pulling child q; the group
// sums density, momentum and non-equilibrium parts with DPP lane swaps (a
// fixed tree, so the result does not depend on the launch) and each lane
// then stores up to three of the ghost's 19 populations.  The child's slot
// and link word are loaded alongside the ghost's instance check, so the
// dependent loads before the population reads are two deep, not four.
template <typename T>
__global__ __launch_bounds__(kWG) void k_fill_f2c(LevelDev Lc, const uint32_t* __restrict__ ghost,
                                                  const uint32_t* __restrict__ child8, uint32_t n, LevelDev Lf,
                                                  const T* __restrict__ Af, T* __restrict__ Ac,
                                                  const typename Store<T>::R* __restrict__ stage,
                                                  const uint8_t* __restrict__ staged) {
    using R = typename Store<T>::R;
    const uint32_t t = blockIdx.x * kWG + threadIdx.x;
    const uint32_t k = t >> 3;
    const int q = (int)(t & 7u);
    // every lane of a group shares k, so a group is wholly in or out, and
    // wholly staged or not
    bool on = k < n;
    bool stg = false;
    size_t g = 0;
    uint32_t child = 0;
    uint64_t w = 0;
    if (on) {
        g = ghost[k];
        stg = stage && staged[k];
        if (!stg) {
            child = child8[(size_t)8 * k + q];
            w = Lf.cflags[child];
        }
        on = Lc.inst[Lc.binst[g >> 6]].live != 0;
    }
    R s[4 + Q];
#pragma unroll
    for (int j = 0; j < 4 + Q; ++j) s[j] = 0;
    if (stg) {
        // the group sums the fine collide formed (stage_children), the same bits
#pragma unroll
        for (int j = 0; j < 4 + Q; ++j) s[j] = stage[(size_t)j * n + k];
    } else {
        if (on) {
            // the child's deviations h = f - w: s = (rho - 1, momentum, f_neq)
            R f[Q];
            gather_slot_dev<T, R>(Lf, Af, child, w, f);
            R d, jx, jy, jz;
            dev_density_momentum(f, d, jx, jy, jz);
            const R rho = (R)1 + d;
            const R ir = (R)1 / rho;
            const R vx = jx * ir, vy = jy * ir, vz = jz * ir;
            const R usq = vx * vx + vy * vy + vz * vz;
            s[0] = d;
            s[1] = jx;
            s[2] = jy;
            s[3] = jz;
#pragma unroll
            for (int i = 0; i < Q; ++i) s[4 + i] = f[i] - feq_dev(i, d, rho, vx, vy, vz, usq);
        }
        // every lane of the group is active here (whole waves, no lane has
        // left, the branch is group-uniform); the sum stays within the group
#pragma unroll
        for (int j = 0; j < 4 + Q; ++j) s[j] = group8_sum(s[j]);
    }
    if (!on) return;
    const R fac = (R)f2c_factor(Lc.inst[Lc.binst[g >> 6]].tau0, Lc.level);
    const R dm = s[0] * (R)0.125;
    const R r = (R)1 + dm;
    const R ux = (s[1] * (R)0.125) / r, uy = (s[2] * (R)0.125) / r, uz = (s[3] * (R)0.125) / r;
    const R usq = ux * ux + uy * uy + uz * uz;
    const uint32_t rb = slot_bits(Lc, g);
#pragma unroll
    for (int i = 0; i < Q; ++i)
        if ((i & 7) == q)
            store_ghost<T>(Ac, pidx(i, Lc.S, g), i, feq_dev(i, dm, r, ux, uy, uz, usq) + fac * (s[4 + i] * (R)0.125),
                           rb);
}

// skip_simple: a regrid's population transfer writes every fluid leaf of the
// new layout in both buffers next, so the bricks whose 64 cells are all
// fluid leaves ("simple") need no initial state.
template <typename T>
__global__ __launch_bounds__(kWG) void k_init(LevelDev L, T* __restrict__ A, T* __restrict__ B, int skip_simple) {
    using R = typename Store<T>::R;
    for (size_t s = (size_t)blockIdx.x * kWG + threadIdx.x; s < L.S; s += (size_t)gridDim.x * kWG) {
        if (skip_simple && (s >> 6) < (size_t)L.nactive && (L.bflags[s >> 6] & 1u)) continue;
        const InstDev& I = L.inst[L.binst[s >> 6]];
        const R ux = (R)I.u_init[0], uy = (R)I.u_init[1], uz = (R)I.u_init[2], rho = (R)I.rho_init;
        const R usq = (R)(I.u_init[0] * I.u_init[0] + I.u_init[1] * I.u_init[1] + I.u_init[2] * I.u_init[2]);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            R e = feq(i, rho, ux, uy, uz, usq);
            Store<T>::store(A, pidx(i, L.S, s), i, e);
            Store<T>::store(B, pidx(i, L.S, s), i, e);
        }
    }
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// One wave per active brick: FP64 (mass, jx, jy, jz) * vol of its active
// cells, summed in a fixed butterfly order.
template <typename T>
__global__ __launch_bounds__(kWG) void k_brick_sums(LevelDev L, const T* __restrict__ A, double vol,
                                                    double* __restrict__ out) {
    const int b = (int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6));
    if (b >= L.nactive) return;
    const size_t s = (size_t)b * 64 + (threadIdx.x & 63);
    const bool act = (L.bflags[b] & 1u) != 0 || (L.cflags[s] & kActiveBit);
    double r = 0, jx = 0, jy = 0, jz = 0;
    if (act) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const double f = (double)Store<T>::load(A, pidx(i, L.S, s), i);
            r += f;
            jx += cx(i) * f;
            jy += cy(i) * f;
            jz += cz(i) * f;
        }
    }
    r = wave_sum(r * vol);
    jx = wave_sum(jx * vol);
    jy = wave_sum(jy * vol);
    jz = wave_sum(jz * vol);
    if ((threadIdx.x & 63) == 0) {
        out[(size_t)b * 4 + 0] = r;
        out[(size_t)b * 4 + 1] = jx;
        out[(size_t)b * 4 + 2] = jy;
        out[(size_t)b * 4 + 3] = jz;
    }
}

// Velocity of an active cell (FP64 sums over the stored populations).
template <typename T>
__device__ inline void cell_velocity(const LevelDev& L, const T* __restrict__ A, size_t s, float* u) {
    double r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const double f = (double)Store<T>::load(A, pidx(i, L.S, s), i);
        r += f;
        jx += cx(i) * f;
        jy += cy(i) * f;
        jz += cz(i) * f;
    }
    u[0] = (float)(jx / r);
    u[1] = (float)(jy / r);
    u[2] = (float)(jz / r);
}

// One wave per active brick: max over its active cells of |u - u_prev|
// (infinity norm), u_prev replaced by u.  With inv_m > 0, u is the window
// mean uacc * inv_m of the velocities sampled since the last call (and the
// accumulator is cleared for the next window); with inv_m = 0, u is the
// current velocity.
template <typename T>
__global__ __launch_bounds__(kWG) void k_brick_du(LevelDev L, const T* __restrict__ A, float* __restrict__ uprev,
                                                  float* __restrict__ uacc, float inv_m, double* __restrict__ out) {
    const int b = (int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6));
    if (b >= L.nactive) return;
    const size_t s = (size_t)b * 64 + (threadIdx.x & 63);
    const bool act = (L.bflags[b] & 1u) != 0 || (L.cflags[s] & kActiveBit);
    double d = 0;
    if (act) {
        float u[3];
        if (inv_m > 0.0f) {
            for (int c = 0; c < 3; ++c) {
                u[c] = uacc[s * 3 + c] * inv_m;
                uacc[s * 3 + c] = 0.0f;
            }
        } else {
            cell_velocity<T>(L, A, s, u);
        }
        for (int c = 0; c < 3; ++c) {
            d = fmax(d, fabs((double)u[c] - (double)uprev[s * 3 + c]));
            uprev[s * 3 + c] = u[c];
        }
    }
    d = wave_max(d);
    if ((threadIdx.x & 63) == 0) out[b] = d;
}

// uacc += u at every active cell (one velocity sample of the current window).
template <typename T>
__global__ __launch_bounds__(kWG) void k_velocity_sample(LevelDev L, const T* __restrict__ A,
                                                         float* __restrict__ uacc) {
    const int b = (int)(blockIdx.x * (kWG / 64) + (threadIdx.x >> 6));
    if (b >= L.nactive) return;
    const size_t s = (size_t)b * 64 + (threadIdx.x & 63);
    if (!((L.bflags[b] & 1u) != 0 || (L.cflags[s] & kActiveBit))) return;
    float u[3];
    cell_velocity<T>(L, A, s, u);
    for (int c = 0; c < 3; ++c) uacc[s * 3 + c] += u[c];
}

// One workgroup per instance: reduce the per-brick (or per-link) values of
// the instance's range [beg[i], beg[i+1]) in a fixed order; op 0 = sum,
// 1 = max.  first != 0 stores, else accumulates into acc[i * ncomp + c].
__global__ __launch_bounds__(kWG) void k_inst_reduce(const double* __restrict__ vals, int ncomp,
                                                     const uint32_t* __restrict__ beg, int op, int first,
                                                     double* __restrict__ acc) {
    const int inst = blockIdx.x;
    const uint32_t b0 = beg[inst], b1 = beg[inst + 1];
    __shared__ double red[kWG];
    for (int c = 0; c < ncomp; ++c) {
        double v = 0;
        for (uint32_t k = b0 + threadIdx.x; k < b1; k += kWG) {
            const double x = vals[(size_t)k * ncomp + c];
            v = op ? fmax(v, x) : v + x;
        }
        red[threadIdx.x] = v;
        __syncthreads();
        for (int st = kWG / 2; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st)
                red[threadIdx.x] = op ? fmax(red[threadIdx.x], red[threadIdx.x + st])
                                      : red[threadIdx.x] + red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double& a = acc[(size_t)inst * ncomp + c];
            a = first ? red[0] : (op ? fmax(a, red[0]) : a + red[0]);
        }
        __syncthreads();
    }
}

// Momentum exchange per solid link (slot * 32 + q): -2 c_q f*_opp(q) * scale.
template <typename T>
__global__ __launch_bounds__(kWG) void k_link_force(const uint32_t* __restrict__ links, uint32_t n, size_t S,
                                                    const T* __restrict__ A, double scale, double* __restrict__ out) {
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        const uint32_t e = links[k];
        const size_t s = e >> 5;
        const int q = (int)(e & 31u);
        double fo = 0;
        int c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
        for (int i = 1; i < Q; ++i)
            if (i == q) {
                fo = (double)Store<T>::load(A, pidx(opp(i), S, s), opp(i));
                c0 = cx(i);
                c1 = cy(i);
                c2 = cz(i);
            }
        out[(size_t)k * 3 + 0] = -2.0 * c0 * fo * scale;
        out[(size_t)k * 3 + 1] = -2.0 * c1 * fo * scale;
        out[(size_t)k * 3 + 2] = -2.0 * c2 * fo * scale;
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_extract(const T* __restrict__ A, size_t S, const int64_t* __restrict__ slots,
                                                 size_t n, double* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * kWG + threadIdx.x; k < n; k += (size_t)gridDim.x * kWG) {
        const int64_t s = slots[k];
#pragma unroll
        for (int i = 0; i < Q; ++i) out[k * Q + i] = s < 0 ? 0.0 : (double)Store<T>::load(A, pidx(i, S, s), i);
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_inject(T* __restrict__ A, T* __restrict__ B, size_t S,
                                                const int64_t* __restrict__ slots, size_t n,
                                                const double* __restrict__ in) {
    using R = typename Store<T>::R;
    for (size_t k = (size_t)blockIdx.x * kWG + threadIdx.x; k < n; k += (size_t)gridDim.x * kWG) {
        const int64_t s = slots[k];
        if (s < 0) continue;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            Store<T>::store(A, pidx(i, S, s), i, (R)in[k * Q + i]);
            Store<T>::store(B, pidx(i, S, s), i, (R)in[k * Q + i]);
        }
    }
}

// |vorticity| (level-0 units) at listed fluid-leaf slots.
template <typename T>
__device__ inline bool leaf_velocity(const LevelDev& L, const T* __restrict__ A, int64_t s, double* u) {
    if (s < 0) return false;
    const bool simple = (L.bflags[s >> 6] & 1u) != 0;
    if (!simple && !(L.cflags[s] & kActiveBit)) return false;
    double r = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        double f = (double)Store<T>::load(A, pidx(i, L.S, (size_t)s), i);
        r += f;
        jx += cx(i) * f;
        jy += cy(i) * f;
        jz += cz(i) * f;
    }
    u[0] = jx / r;
    u[1] = jy / r;
    u[2] = jz / r;
    return true;
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_vorticity(LevelDev L, const T* __restrict__ A, const int64_t* __restrict__ slots,
                                                   size_t n, double scale, double* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * kWG + threadIdx.x; k < n; k += (size_t)gridDim.x * kWG) {
        const int64_t s = slots[k];
        double u0[3];
        if (!leaf_velocity<T>(L, A, s, u0)) {
            out[k] = 0.0;
            continue;
        }
        const int b = (int)(s >> 6), l = (int)(s & 63);
        const int lx = l & 3, ly = (l >> 2) & 3, lz = l >> 4;
        double g[3][3];
        for (int a = 0; a < 3; ++a) {
            const int dx = a == 0, dy = a == 1, dz = a == 2;
            double up[3], um[3];
            const bool hp = leaf_velocity<T>(L, A, slot_of(L, b, lx + dx, ly + dy, lz + dz), up);
            const bool hm = leaf_velocity<T>(L, A, slot_of(L, b, lx - dx, ly - dy, lz - dz), um);
            for (int c = 0; c < 3; ++c)
                g[a][c] = hp && hm ? 0.5 * (up[c] - um[c]) : (hp ? up[c] - u0[c] : (hm ? u0[c] - um[c] : 0.0));
        }
        const double wx = g[1][2] - g[2][1], wy = g[2][0] - g[0][2], wz = g[0][1] - g[1][0];
        out[k] = sqrt(wx * wx + wy * wy + wz * wz) * scale;
    }
}

// Population transfer of a regrid: copy / split child / merged parent / rest.
// The copy moves the stored bits (an unsigned integer of T's size); the 19
// loads of a cell are issued before its 38 non-temporal stores.
template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_copy(const T* __restrict__ Ao_, size_t So, const uint32_t* __restrict__ dst,
                                                    const uint32_t* __restrict__ src, uint32_t n, T* __restrict__ A_,
                                                    T* __restrict__ B_, size_t S) {
    using U = typename std::conditional<sizeof(T) == 2, uint16_t,
                                        typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type>::type;
    static_assert(sizeof(U) == sizeof(T), "storage element size");
    const U* __restrict__ Ao = reinterpret_cast<const U*>(Ao_);
    U* __restrict__ A = reinterpret_cast<U*>(A_);
    U* __restrict__ B = reinterpret_cast<U*>(B_);
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        const uint32_t s0 = src[k], d0 = dst[k];
        U v[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) v[i] = Ao[pidx(i, So, s0)];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            __builtin_nontemporal_store(v[i], A + pidx(i, S, d0));
            __builtin_nontemporal_store(v[i], B + pidx(i, S, d0));
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_split(LevelDev Ln, LevelDev Lc, const T* __restrict__ Acprev,
                                                     const uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                                     uint32_t n, T* __restrict__ A, T* __restrict__ B) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        const R fac = (R)c2f_factor(Ln.inst[Ln.binst[dst[k] >> 6]].tau0, Ln.level);
        R f[Q];
        gather_slot<T, R>(Lc, Acprev, src[k], f);
        R rho, ux, uy, uz;
        moments(f, rho, ux, uy, uz);
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R e = feq(i, rho, ux, uy, uz, usq);
            const R v = e + fac * (f[i] - e);
            Store<T>::store(A, pidx(i, Ln.S, dst[k]), i, v);
            Store<T>::store(B, pidx(i, Ln.S, dst[k]), i, v);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_merge(LevelDev Ln, LevelDev Lf, const T* __restrict__ Afprev,
                                                     const uint32_t* __restrict__ dst, const uint32_t* __restrict__ src8,
                                                     uint32_t n, T* __restrict__ A, T* __restrict__ B) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        const R fac = (R)f2c_factor(Ln.inst[Ln.binst[dst[k] >> 6]].tau0, Ln.level);
        R r, ux, uy, uz, aneq[Q];
        child_average<T>(Lf, Afprev, src8 + (size_t)8 * k, r, ux, uy, uz, aneq);
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R v = feq(i, r, ux, uy, uz, usq) + fac * aneq[i];
            Store<T>::store(A, pidx(i, Ln.S, dst[k]), i, v);
            Store<T>::store(B, pidx(i, Ln.S, dst[k]), i, v);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_remap_rest(const uint32_t* __restrict__ dst, uint32_t n, T* __restrict__ A,
                                                    T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R v = (R)wq(i);
            Store<T>::store(A, pidx(i, S, dst[k]), i, v);
            Store<T>::store(B, pidx(i, S, dst[k]), i, v);
        }
    }
}

// Ghost slots after a regrid: equilibrium of the covering coarse leaf (c2f)
// or of the 8 children's average density and momentum (f2c), both buffers.
template <typename T>
__global__ __launch_bounds__(kWG) void k_ghost_eq_c2f(const uint32_t* __restrict__ gchild, const uint32_t* __restrict__ gcap,
                                                      const uint32_t* __restrict__ cap_slots, uint32_t ngroups,
                                                      const T* __restrict__ Ac, size_t Sc, T* __restrict__ A,
                                                      T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    const size_t n = (size_t)ngroups * 8;
    for (size_t t = (size_t)blockIdx.x * kWG + threadIdx.x; t < n; t += (size_t)gridDim.x * kWG) {
        const uint32_t g = gchild[t];
        if (g == kNoCap) continue;
        const size_t cs = cap_slots[gcap[(t >> 3) * kC2FCaps]];
        R f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = Store<T>::load(Ac, pidx(i, Sc, cs), i);
        R rho, ux, uy, uz;
        moments(f, rho, ux, uy, uz);
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R e = feq(i, rho, ux, uy, uz, usq);
            Store<T>::store(A, pidx(i, S, g), i, e);
            Store<T>::store(B, pidx(i, S, g), i, e);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kWG) void k_ghost_eq_f2c(const uint32_t* __restrict__ ghost, const uint32_t* __restrict__ child8,
                                                      uint32_t n, const T* __restrict__ Af, size_t Sf, T* __restrict__ A,
                                                      T* __restrict__ B, size_t S) {
    using R = typename Store<T>::R;
    for (uint32_t k = blockIdx.x * kWG + threadIdx.x; k < n; k += gridDim.x * kWG) {
        R sr = 0, sjx = 0, sjy = 0, sjz = 0;
        for (int q = 0; q < 8; ++q) {
            const size_t cs = child8[(size_t)8 * k + q];
            R f[Q];
#pragma unroll
            for (int i = 0; i < Q; ++i) f[i] = Store<T>::load(Af, pidx(i, Sf, cs), i);
            R rho, ux, uy, uz;
            moments(f, rho, ux, uy, uz);
            sr += rho;
            sjx += rho * ux;
            sjy += rho * uy;
            sjz += rho * uz;
        }
        const R r = sr * (R)0.125;
        const R ux = (sjx * (R)0.125) / r, uy = (sjy * (R)0.125) / r, uz = (sjz * (R)0.125) / r;
        const R usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const R e = feq(i, r, ux, uy, uz, usq);
            Store<T>::store(A, pidx(i, S, ghost[k]), i, e);
            Store<T>::store(B, pidx(i, S, ghost[k]), i, e);
        }
    }
}

unsigned lgrid(size_t n) {
    size_t g = (n + kWG - 1) / kWG;
    if (g > 4096) g = 4096;
    return (unsigned)(g ? g : 1);
}

int device_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

template <typename T, int BPW, int MODE, bool kOutlet = false>
hipError_t collide_variant(const LevelDev& L, const T* A, T* B, hipStream_t s) {
    const unsigned ngroups = (unsigned)((L.nactive + BPW - 1) / BPW);
    unsigned grid = ngroups;
    if (MODE == 2) {
        static int per_cu = 0;
        if (per_cu == 0) {
            int nb = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_collide<T, BPW, MODE, kOutlet>, 64 * BPW, 0) !=
                    hipSuccess ||
                nb <= 0)
                nb = 8 / BPW > 0 ? 8 / BPW : 1;
            per_cu = nb;
        }
        const unsigned cap = (unsigned)(per_cu * device_cus());
        if (grid > cap) grid = cap;
    }
    hipLaunchKernelGGL((k_collide<T, BPW, MODE, kOutlet>), dim3(grid), dim3(64 * BPW), 0, s, L, A, B, ngroups);
    return hipGetLastError();
}

}  // namespace

int collide_variants() { return 21; }

const char* collide_variant_name(int v) {
    static const char* names[] = {"bpw1",     "bpw2",         "bpw4",         "bpw1-xcd",     "bpw2-xcd", "bpw4-xcd",
                                  "bpw1-persist", "bpw2-persist", "bpw4-persist", "bpw8",     "bpw8-xcd", "bpw16",
                                  "bpw16-xcd", "bpw2-xc128", "bpw4-xc64", "bpw4-xc128", "bpw8-xc128",
                                  "bpw4-xc256", "bpw8-xc256", "bpw4-xc512", "bpw8-xc512"};
    return v >= 0 && v < 21 ? names[v] : "?";
}

template <typename T>
hipError_t launch_collide(const LevelDev& L, const T* A, T* B, int variant, hipStream_t s) {
    if (L.nactive == 0) return hipSuccess;
    if (L.outlet) {
        // levels with pressure-outlet links (at most the few that touch an
        // outflow face) run one of three outlet instances, by the variant's
        // workgroup-to-brick mapping; every variant gives the same bits
        switch (variant) {
            case 3: case 4: case 5: case 10: case 12: return collide_variant<T, 4, 1, true>(L, A, B, s);
            case 0: case 1: case 2: case 6: case 7: case 8: case 9: case 11: return collide_variant<T, 4, 0, true>(L, A, B, s);
            default: return collide_variant<T, 4, 4, true>(L, A, B, s);
        }
    }
    switch (variant) {
        case 0: return collide_variant<T, 1, 0>(L, A, B, s);
        case 1: return collide_variant<T, 2, 0>(L, A, B, s);
        case 2: return collide_variant<T, 4, 0>(L, A, B, s);
        case 3: return collide_variant<T, 1, 1>(L, A, B, s);
        case 4: return collide_variant<T, 2, 1>(L, A, B, s);
        case 5: return collide_variant<T, 4, 1>(L, A, B, s);
        case 6: return collide_variant<T, 1, 2>(L, A, B, s);
        case 7: return collide_variant<T, 2, 2>(L, A, B, s);
        case 8: return collide_variant<T, 4, 2>(L, A, B, s);
        case 9: return collide_variant<T, 8, 0>(L, A, B, s);
        case 10: return collide_variant<T, 8, 1>(L, A, B, s);
        case 11: return collide_variant<T, 16, 0>(L, A, B, s);
        case 12: return collide_variant<T, 16, 1>(L, A, B, s);
        case 13: return collide_variant<T, 2, 4>(L, A, B, s);
        case 14: return collide_variant<T, 4, 3>(L, A, B, s);
        case 15: return collide_variant<T, 4, 4>(L, A, B, s);
        case 16: return collide_variant<T, 8, 4>(L, A, B, s);
        case 17: return collide_variant<T, 4, 5>(L, A, B, s);
        case 18: return collide_variant<T, 8, 5>(L, A, B, s);
        case 19: return collide_variant<T, 4, 6>(L, A, B, s);
        default: return collide_variant<T, 8, 6>(L, A, B, s);
    }
}

hipError_t launch_any_outlet(const uint64_t* cflags, uint32_t* bflags, size_t nslots, uint32_t* flag, hipStream_t s) {
    if (nslots) hipLaunchKernelGGL(k_any_outlet, dim3(lgrid(nslots)), dim3(kWG), 0, s, cflags, bflags, nslots, flag);
    return hipGetLastError();
}

hipError_t launch_tick(uint32_t* tick, hipStream_t s) {
    hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, tick);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_capture(const LevelDev& L, const T* A, const uint32_t* slots, uint32_t ncap, void* cap,
                          const uint8_t* skip, hipStream_t s) {
    if (ncap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_capture<T>, dim3(lgrid(ncap)), dim3(kWG), 0, s, L, A, slots, ncap,
                       static_cast<typename Store<T>::R*>(cap), skip);
    return hipGetLastError();
}

// Capture staging map of a level: capture k is staged when its slot is an
// active fluid leaf whose links are all regular pulls (the collide's pulled
// values are then the capture's gather bit for bit); cap_of[slot] = k and
// the brick's bflags gain bit 3.  cap_of starts all kNoCap.
__global__ __launch_bounds__(kWG) void k_cap_map(const uint32_t* __restrict__ slots, uint32_t ncap,
                                                 const uint64_t* __restrict__ cflags, uint32_t* __restrict__ bflags,
                                                 uint32_t* __restrict__ cap_of, uint8_t* __restrict__ staged) {
    const uint32_t k = blockIdx.x * kWG + threadIdx.x;
    if (k >= ncap) return;
    const uint32_t sl = slots[k];
    const uint64_t w = cflags[sl];
    const bool ok = (w & kActiveBit) != 0 && (w & ((1ull << (2 * Q)) - 1)) == 0;
    staged[k] = ok ? 1 : 0;
    if (!ok) return;
    cap_of[sl] = k;
    atomicOr(bflags + (sl >> 6), 8u);
}

hipError_t launch_cap_map(const uint32_t* slots, uint32_t ncap, const uint64_t* cflags, uint32_t* bflags,
                          uint32_t* cap_of, size_t nslots, uint8_t* staged, hipStream_t s) {
    if (nslots)
        hipLaunchKernelGGL(k_fill_u32_kernel, dim3((unsigned)((nslots + kWG - 1) / kWG)), dim3(kWG), 0, s, cap_of,
                           nslots, kNoCap);
    if (ncap) hipLaunchKernelGGL(k_cap_map, dim3((ncap + kWG - 1) / kWG), dim3(kWG), 0, s, slots, ncap, cflags, bflags,
                                 cap_of, staged);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_c2f(const LevelDev& Lf, const uint32_t* gchild, const uint32_t* gcap, const uint32_t* gmask,
                           uint32_t ngroups, const void* cap0, const void* cap1, uint32_t ncap, double alpha, T* A,
                           hipStream_t s) {
    if (ngroups == 0) return hipSuccess;
    using R = typename Store<T>::R;
    // eight lanes per group, one pass (no grid stride: a group must stay whole)
    const size_t blocks = ((size_t)ngroups * 8 + kWG - 1) / kWG;
    hipLaunchKernelGGL(k_fill_c2f<T>, dim3((unsigned)blocks), dim3(kWG), 0, s, Lf, gchild, gcap, gmask, ngroups,
                       static_cast<const R*>(cap0), static_cast<const R*>(cap1), ncap, (R)alpha, A);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_fill_f2c(const LevelDev& Lc, const uint32_t* ghost, const uint32_t* child8, uint32_t n,
                           const LevelDev& Lf, const T* Af, T* Ac, const void* stage, const uint8_t* staged,
                           hipStream_t s) {
    if (n == 0) return hipSuccess;
    using R = typename Store<T>::R;
    // eight lanes per ghost, one pass (no grid stride: a group must stay whole)
    const size_t blocks = ((size_t)n * 8 + kWG - 1) / kWG;
    hipLaunchKernelGGL(k_fill_f2c<T>, dim3((unsigned)blocks), dim3(kWG), 0, s, Lc, ghost, child8, n, Lf, Af, Ac,
                       static_cast<const R*>(stage), staged);
    return hipGetLastError();
}

// The staging map of fine level Lf for the coarse level's f2c ghosts: ghost
// k is staged when its eight children are active fluid leaves of one brick
// (they form a 2x2x2 block of it) whose links are all regular pulls; then
// f2c_blk[brick * 8 + block] = k and the brick's bflags gain bit 2.  (A wall
// or outlet link's value is formed from the population in the FP32 collide
// and from the deviation in the fill's gather, which round differently; a
// regular pull is the same load either way.)  f2c_blk starts all kNoCap.
__global__ __launch_bounds__(kWG) void k_stage_map(const uint32_t* __restrict__ child8, uint32_t n,
                                                   const uint64_t* __restrict__ cflags, uint32_t* __restrict__ bflags,
                                                   uint32_t* __restrict__ f2c_blk, uint8_t* __restrict__ staged) {
    const uint32_t k = blockIdx.x * kWG + threadIdx.x;
    if (k >= n) return;
    const uint32_t c0 = child8[(size_t)8 * k];
    bool ok = (c0 & 0x15u) == 0;  // the block's even corner
    for (int q = 0; q < 8 && ok; ++q) {
        const uint32_t c = child8[(size_t)8 * k + q];
        const uint64_t w = cflags[c];
        ok = (c >> 6) == (c0 >> 6) && (w & kActiveBit) != 0 && (w & ((1ull << (2 * Q)) - 1)) == 0;
    }
    staged[k] = ok ? 1 : 0;
    if (!ok) return;
    const uint32_t l = c0 & 63u;
    const uint32_t blk = ((l >> 1) & 1u) | ((l >> 2) & 2u) | ((l >> 3) & 4u);
    f2c_blk[(size_t)(c0 >> 6) * 8 + blk] = k;
    atomicOr(bflags + (c0 >> 6), 4u);
}

hipError_t launch_stage_map(const uint32_t* child8, uint32_t n, const uint64_t* cflags, uint32_t* bflags,
                            uint32_t* f2c_blk, size_t nbricks, uint8_t* staged, hipStream_t s) {
    if (nbricks) {
        const size_t m = nbricks * 8;
        hipLaunchKernelGGL(k_fill_u32_kernel, dim3((unsigned)((m + kWG - 1) / kWG)), dim3(kWG), 0, s, f2c_blk, m,
                           kNoCap);
    }
    if (n) hipLaunchKernelGGL(k_stage_map, dim3((n + kWG - 1) / kWG), dim3(kWG), 0, s, child8, n, cflags, bflags,
                              f2c_blk, staged);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_init(const LevelDev& L, T* A, T* B, hipStream_t s, bool skip_simple) {
    if (L.S == 0) return hipSuccess;
    hipLaunchKernelGGL(k_init<T>, dim3(lgrid(L.S)), dim3(kWG), 0, s, L, A, B, skip_simple ? 1 : 0);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_brick_sums(const LevelDev& L, const T* A, double vol, double* out, hipStream_t s) {
    if (L.nactive == 0) return hipSuccess;
    hipLaunchKernelGGL(k_brick_sums<T>, dim3((L.nactive + 3) / 4), dim3(kWG), 0, s, L, A, vol, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_brick_du(const LevelDev& L, const T* A, float* uprev, float* uacc, float inv_m, double* out,
                           hipStream_t s) {
    if (L.nactive == 0) return hipSuccess;
    hipLaunchKernelGGL(k_brick_du<T>, dim3((L.nactive + 3) / 4), dim3(kWG), 0, s, L, A, uprev, uacc, inv_m, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_velocity_sample(const LevelDev& L, const T* A, float* uacc, hipStream_t s) {
    if (L.nactive == 0) return hipSuccess;
    hipLaunchKernelGGL(k_velocity_sample<T>, dim3((L.nactive + 3) / 4), dim3(kWG), 0, s, L, A, uacc);
    return hipGetLastError();
}

__global__ __launch_bounds__(kWG) void k_fill_f32(float* __restrict__ p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * kWG + threadIdx.x; i < n; i += (size_t)gridDim.x * kWG) p[i] = v;
}

hipError_t launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t g = (n + kWG - 1) / kWG;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)(g > 65536 ? 65536 : g)), dim3(kWG), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_inst_reduce(const double* vals, int ncomp, const uint32_t* beg, int ninst, int op, int first,
                              double* acc, hipStream_t s) {
    if (ninst == 0) return hipSuccess;
    hipLaunchKernelGGL(k_inst_reduce, dim3(ninst), dim3(kWG), 0, s, vals, ncomp, beg, op, first, acc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_link_force(const uint32_t* links, uint32_t n, size_t S, const T* A, double scale, double* out,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_link_force<T>, dim3(lgrid(n)), dim3(kWG), 0, s, links, n, S, A, scale, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_extract(const T* A, size_t S, const int64_t* slots, size_t n, double* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_extract<T>, dim3(lgrid(n)), dim3(kWG), 0, s, A, S, slots, n, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_inject(T* A, T* B, size_t S, const int64_t* slots, size_t n, const double* in, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_inject<T>, dim3(lgrid(n)), dim3(kWG), 0, s, A, B, S, slots, n, in);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_vorticity(const LevelDev& L, const T* A, const int64_t* slots, size_t n, double scale, double* out,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vorticity<T>, dim3(lgrid(n)), dim3(kWG), 0, s, L, A, slots, n, scale, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_remap(const RemapArgs<T>& r, hipStream_t s) {
    if (r.ncopy)
        hipLaunchKernelGGL(k_remap_copy<T>, dim3(lgrid(r.ncopy)), dim3(kWG), 0, s, r.copy_src_buf, r.So, r.copy_dst,
                           r.copy_src, r.ncopy, r.A, r.B, r.Ln.S);
    if (r.nsplit)
        hipLaunchKernelGGL(k_remap_split<T>, dim3(lgrid(r.nsplit)), dim3(kWG), 0, s, r.Ln, r.Lc, r.split_prev,
                           r.split_dst, r.split_src, r.nsplit, r.A, r.B);
    if (r.nmerge)
        hipLaunchKernelGGL(k_remap_merge<T>, dim3(lgrid(r.nmerge)), dim3(kWG), 0, s, r.Ln, r.Lf, r.merge_prev,
                           r.merge_dst, r.merge_src8, r.nmerge, r.A, r.B);
    if (r.nrest)
        hipLaunchKernelGGL(k_remap_rest<T>, dim3(lgrid(r.nrest)), dim3(kWG), 0, s, r.rest_dst, r.nrest, r.A, r.B,
                           r.Ln.S);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ghost_eq(const uint32_t* c2f_gchild, const uint32_t* c2f_gcap, const uint32_t* cap_slots, uint32_t nc2f,
                           const T* Ac, size_t Sc, const uint32_t* f2c_ghost, const uint32_t* f2c_child, uint32_t nf2c,
                           const T* Af, size_t Sf, T* A, T* B, size_t S, hipStream_t s) {
    if (nc2f)
        hipLaunchKernelGGL(k_ghost_eq_c2f<T>, dim3(lgrid((size_t)nc2f * 8)), dim3(kWG), 0, s, c2f_gchild, c2f_gcap, cap_slots, nc2f,
                           Ac, Sc, A, B, S);
    if (nf2c)
        hipLaunchKernelGGL(k_ghost_eq_f2c<T>, dim3(lgrid(nf2c)), dim3(kWG), 0, s, f2c_ghost, f2c_child, nf2c, Af, Sf,
                           A, B, S);
    return hipGetLastError();
}

#define FL_INST(T)                                                                                            \
    template hipError_t launch_collide<T>(const LevelDev&, const T*, T*, int, hipStream_t);                    \
    template hipError_t launch_capture<T>(const LevelDev&, const T*, const uint32_t*, uint32_t, void*,            \
                                          const uint8_t*, hipStream_t);                                          \
    template hipError_t launch_fill_c2f<T>(const LevelDev&, const uint32_t*, const uint32_t*, const uint32_t*,  \
                                           uint32_t, const void*, const void*, uint32_t, double, T*, hipStream_t); \
    template hipError_t launch_fill_f2c<T>(const LevelDev&, const uint32_t*, const uint32_t*, uint32_t,         \
                                           const LevelDev&, const T*, T*, const void*, const uint8_t*,          \
                                           hipStream_t);                                                        \
    template hipError_t launch_init<T>(const LevelDev&, T*, T*, hipStream_t, bool);                             \
    template hipError_t launch_brick_sums<T>(const LevelDev&, const T*, double, double*, hipStream_t);          \
    template hipError_t launch_brick_du<T>(const LevelDev&, const T*, float*, float*, float, double*, hipStream_t); \
    template hipError_t launch_velocity_sample<T>(const LevelDev&, const T*, float*, hipStream_t);               \
    template hipError_t launch_link_force<T>(const uint32_t*, uint32_t, size_t, const T*, double, double*,      \
                                             hipStream_t);                                                      \
    template hipError_t launch_extract<T>(const T*, size_t, const int64_t*, size_t, double*, hipStream_t);     \
    template hipError_t launch_inject<T>(T*, T*, size_t, const int64_t*, size_t, const double*, hipStream_t);  \
    template hipError_t launch_vorticity<T>(const LevelDev&, const T*, const int64_t*, size_t, double, double*, \
                                            hipStream_t);                                                      \
    template hipError_t launch_remap<T>(const RemapArgs<T>&, hipStream_t);                                     \
    template hipError_t launch_ghost_eq<T>(const uint32_t*, const uint32_t*, const uint32_t*, uint32_t, const T*, \
                                           size_t, const uint32_t*, const uint32_t*, uint32_t, const T*, size_t, \
                                           T*, T*, size_t, hipStream_t);

FL_INST(float)
FL_INST(double)
FL_INST(_Float16)
FL_INST(fp16c)

}  // namespace lbm
}  // namespace fl
