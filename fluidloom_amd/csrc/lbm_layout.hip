// Device-side lattice layout builder (see lbm_layout.hpp for the tables).
//
// Per level:
//   1. the brick keys of the level's leaves (instance, brick coordinates)
//      and whether each brick holds a fluid leaf: sort + unique;
//   2. the active bricks' 26-neighbour dilation (periodic faces wrap):
//      sort + unique again -> the level's bricks;
//   3. order: stable sort by (instance, Hilbert key of the brick
//      coordinates), then a stable partition putting active bricks first;
//   4. the brick map (exact cell map: packed brick key -> brick index),
//      neighbour rows, leaf slot kinds (duplicates and missing bricks are
//      mesh errors).
// Then, with every level's leaf kinds known: ghost kinds from the leaves one
// level up and down, the fluid leaves' link words with the direction masks
// of the ghosts they pull from, force links, and the ghost lists (scan
// compaction in slot order; c2f groups by a stable sort on the covering
// coarse slot; capture lists sorted and unique; f2c ghosts stably sorted by
// their first child).  Only sizes and per-instance counts come back to the
// host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fl_internal.hpp"
#include "hash_device.hpp"
#include "hilbert_tables.hpp"
#include "lbm_device.hpp"
#include "lbm_layout.hpp"
#include "wave.hpp"

namespace fl {
fl_status hash_clear_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, hipStream_t s);
fl_status cellmap_build_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, const uint64_t* keys, size_t n,
                             hipStream_t s);
}  // namespace fl

namespace fl {
namespace lbm {
namespace {

constexpr unsigned kB = 256;
enum : uint8_t { T_NONE = 0, T_FLUID = 1, T_SOLID = 2, T_C2F = 3, T_F2C = 4, T_WALL = 5 };
enum : uint32_t {
    kErrLevel = 1,
    kErrDup = 2,
    kErrMissing = 4,
    kErrLink = 8,
    kErrParent = 16,
    kErrDomain = 32,
    kErrOrder = 64,
};

__host__ __device__ inline bool regular_t(uint8_t t) { return t == T_FLUID || t == T_C2F || t == T_F2C; }

__device__ inline uint64_t pack_bkey(uint32_t inst, int32_t bx, int32_t by, int32_t bz) {
    return ((uint64_t)inst << 48) | ((uint64_t)(uint16_t)bx << 32) | ((uint64_t)(uint16_t)by << 16) |
           (uint64_t)(uint16_t)bz;
}
__device__ inline int32_t bkey_coord(uint64_t k, int a) { return (int32_t)((k >> (32 - 16 * a)) & 0xFFFF); }

inline unsigned blocks_for(size_t n) {
    size_t g = (n + kB - 1) / kB;
    if (g == 0) g = 1;
    return (unsigned)(g > 65536 * 4 ? 65536 * 4 : g);
}

// Brick coordinates of a neighbour after periodic wrap; false when outside.
__device__ inline bool wrap_brick_d(const LayoutInst& I, int l, int32_t* c) {
    for (int a = 0; a < 3; ++a) {
        const int32_t nb = (I.ext0[a] << l) >> 2;
        if (c[a] >= 0 && c[a] < nb) continue;
        const int f = c[a] < 0 ? 2 * a : 2 * a + 1;
        if (I.bc[f] != FL_BC_PERIODIC) return false;
        c[a] = (c[a] + nb) % nb;
    }
    return true;
}

__device__ inline int64_t map_slot(const LevelMap& M, const LayoutInst& I, int inst, int l, int32_t x, int32_t y,
                                   int32_t z) {
    if (x < 0 || y < 0 || z < 0 || x >= (I.ext0[0] << l) || y >= (I.ext0[1] << l) || z >= (I.ext0[2] << l))
        return -1;
    const uint32_t b = cell_lookup(M.keys, M.vals, M.cap, pack_bkey((uint32_t)inst, x >> 2, y >> 2, z >> 2));
    if (b == kInvalid) return -1;
    return (int64_t)b * 64 + (x & 3) + 4 * (y & 3) + 16 * (z & 3);
}

// source_of of the host builder: the source of direction i from (x, y, z)
// after periodic wrap; 0 when inside, else 1 with *face the wall face (axes
// z, y, x) or -1 for outflow, whose source is clamped into the domain.
__device__ inline int source_of_d(const LayoutInst& I, int l, int x, int y, int z, int i, int* s3, int* face) {
    int q[3] = {x - cx(i), y - cy(i), z - cz(i)};
    int out[3] = {0, 0, 0};
    bool any = false;
    for (int a = 0; a < 3; ++a) {
        const int e = I.ext0[a] << l;
        const int f = q[a] < 0 ? 2 * a : (q[a] >= e ? 2 * a + 1 : -1);
        if (f < 0) continue;
        if (I.bc[f] == FL_BC_PERIODIC) {
            q[a] = (q[a] + e) % e;
        } else {
            out[a] = f + 1;
            any = true;
        }
    }
    s3[0] = q[0];
    s3[1] = q[1];
    s3[2] = q[2];
    if (!any) return 0;
    *face = -1;
    for (int a = 2; a >= 0; --a)
        if (out[a] && I.bc[out[a] - 1] == FL_BC_WALL) {
            *face = out[a] - 1;
            break;
        }
    if (*face < 0)
        for (int a = 0; a < 3; ++a)
            if (out[a]) s3[a] = q[a] < 0 ? 0 : (I.ext0[a] << l) - 1;
    return 1;
}

// Per-key counters (counters[key] += v for every lane with `on`) summed in
// a wave-uniform running total while consecutive contributions share one
// key, so a wave issues an atomic only when the key changes and at flush().
// Every lane of the wave calls add() and flush().  Kernels using it run on
// kCountGrid workgroups, so each wave covers a long run of elements.
constexpr unsigned kCountGrid = 2048;
struct WaveCounter {
    uint32_t key = kInvalid;
    unsigned long long acc = 0;
    __device__ void flush(unsigned long long* counters) {
        if (key != kInvalid && acc && lane_id() == 0) atomicAdd(&counters[key], acc);
        acc = 0;
    }
    __device__ void add(unsigned long long* counters, uint32_t k, bool on, unsigned long long v) {
        const uint64_t m = __ballot(on);
        if (!m) return;
        const int leader = __ffsll((long long)m) - 1;
        const uint32_t k0 = __shfl(k, leader, 64);
        if (__ballot(on && k != k0) == 0ull) {
            unsigned long long sum = on ? v : 0ull;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) sum += __shfl_xor(sum, d, 64);
            if (k0 != key) {
                flush(counters);
                key = k0;
            }
            acc += sum;
        } else if (on) {
            atomicAdd(&counters[k], v);
        }
    }
};

// --------------------------------------------------------------- kernels
__global__ void k_fill32(uint32_t* p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) p[i] = v;
}
__global__ void k_fill8(uint8_t* p, size_t n, uint8_t v) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) p[i] = v;
}
__global__ void k_fill64(uint64_t* p, size_t n, uint64_t v) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) p[i] = v;
}

// instance of every leaf, the per-level leaf counts and the domain checks
__global__ void k_leaf_scan(const int32_t* __restrict__ X, const int32_t* __restrict__ Y,
                            const int32_t* __restrict__ Z, const uint8_t* __restrict__ LV, size_t n,
                            const uint64_t* __restrict__ offs, int ninst, const LayoutInst* __restrict__ inst,
                            uint16_t* __restrict__ linst, unsigned long long* __restrict__ lcount,
                            uint32_t* __restrict__ err) {
    // per-level counts in wave registers (whole waves iterate together)
    unsigned long long cnt[kMaxLevels];
#pragma unroll
    for (int l = 0; l < kMaxLevels; ++l) cnt[l] = 0;
    uint32_t e = 0;
    for (size_t base = (size_t)blockIdx.x * kB; base < n; base += (size_t)gridDim.x * kB) {
        const size_t j = base + threadIdx.x;
        int lok = -1;
        if (j < n) {
            int lo = 0, hi = ninst;  // last k with offs[k] <= j
            while (hi - lo > 1) {
                const int mid = (lo + hi) / 2;
                if (offs[mid] <= j) lo = mid; else hi = mid;
            }
            linst[j] = (uint16_t)lo;
            const LayoutInst& I = inst[lo];
            const int l = LV[j];
            if (l > I.max_level) {
                e |= kErrLevel;
            } else if (X[j] < 0 || Y[j] < 0 || Z[j] < 0 || X[j] >= (I.ext0[0] << l) || Y[j] >= (I.ext0[1] << l) ||
                       Z[j] >= (I.ext0[2] << l)) {
                e |= kErrDomain;
            } else {
                lok = l;
            }
        }
#pragma unroll
        for (int l = 0; l < kMaxLevels; ++l) cnt[l] += (unsigned long long)__popcll(__ballot(lok == l));
    }
    if (e) atomicOr(err, e);
    if (lane_id() == 0)
#pragma unroll
        for (int l = 0; l < kMaxLevels; ++l)
            if (cnt[l]) atomicAdd(&lcount[l], cnt[l]);
}

// leaf brick keys into a set: value 0 when the brick holds a fluid leaf, 1
// otherwise (the set keeps the smallest value of a key)
__global__ void k_insert_leaf_bricks(const int32_t* __restrict__ X, const int32_t* __restrict__ Y,
                                     const int32_t* __restrict__ Z, const uint8_t* __restrict__ LV,
                                     const uint8_t* __restrict__ ST, const uint16_t* __restrict__ linst, size_t n, int l,
                                     uint64_t* __restrict__ tk, uint32_t* __restrict__ tv, uint64_t cap) {
    // whole waves iterate together: a lane skips its insert when the lane
    // before it holds the same (brick, value), as most neighbours in a
    // space-filling-curve ordered leaf array do
    const int lane = (int)lane_id();
    for (size_t base = (size_t)blockIdx.x * kB; base < n; base += (size_t)gridDim.x * kB) {
        const size_t j = base + threadIdx.x;
        uint64_t key = kEmptyKey;
        uint32_t val = 0;
        if (j < n && LV[j] == l) {
            key = pack_bkey(linst[j], X[j] >> 2, Y[j] >> 2, Z[j] >> 2);
            val = ST[j] == 0 ? 0u : 1u;
        }
        const uint64_t pkey = __shfl(key, lane - 1, 64);
        const uint32_t pval = __shfl(val, lane - 1, 64);
        if (key != kEmptyKey && (lane == 0 || pkey != key || pval != val)) cell_insert(tk, tv, cap, key, val);
    }
}

// occupied slots of a set
__global__ void k_occupied(const uint64_t* __restrict__ tk, uint64_t cap, uint32_t* __restrict__ flag) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < cap; i += (size_t)gridDim.x * kB)
        flag[i] = tk[i] != kEmptyKey ? 1u : 0u;
}

// the occupied slots' keys and their active flag (value 0) in slot order
__global__ void k_extract(const uint64_t* __restrict__ tk, const uint32_t* __restrict__ tv, uint64_t cap,
                          const uint32_t* __restrict__ flag, const uint32_t* __restrict__ excl,
                          uint64_t* __restrict__ keys, uint32_t* __restrict__ act) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < cap; i += (size_t)gridDim.x * kB)
        if (flag[i]) {
            keys[excl[i]] = tk[i];
            act[excl[i]] = tv[i] == 0u ? 1u : 0u;
        }
}

// the leaf bricks (value 0 active / 1 not) and the 26-neighbour dilation of
// the active ones (value 2, which never lowers a leaf brick's value)
__global__ void k_insert_dilated(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ act, size_t m, int l,
                                 const LayoutInst* __restrict__ inst, uint64_t* __restrict__ tk,
                                 uint32_t* __restrict__ tv, uint64_t cap) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < m; i += (size_t)gridDim.x * kB) {
        const uint64_t k = keys[i];
        cell_insert(tk, tv, cap, k, act[i] ? 0u : 1u);
        if (!act[i]) continue;
        const uint32_t in = (uint32_t)(k >> 48);
        const LayoutInst& I = inst[in];
        for (int t = 0; t < 27; ++t) {
            if (t == 13) continue;
            int32_t c[3] = {bkey_coord(k, 0) + t % 3 - 1, bkey_coord(k, 1) + (t / 3) % 3 - 1, bkey_coord(k, 2) + t / 9 - 1};
            if (wrap_brick_d(I, l, c)) cell_insert(tk, tv, cap, pack_bkey(in, c[0], c[1], c[2]), 2u);
        }
    }
}

// occupied leaf slots of a level (a duplicate leaf leaves fewer than leaves)
__global__ void k_count_placed(const uint8_t* __restrict__ stype, size_t n, unsigned long long* __restrict__ count) {
    unsigned long long c = 0;
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB)
        c += stype[i] == T_FLUID || stype[i] == T_SOLID;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane_id() == 0 && c) atomicAdd(count, c);
}

// segment heads of a sorted key array (EMPTY keys, sorted last, are dropped)
__global__ void k_heads(const uint64_t* __restrict__ keys, size_t n, uint32_t* __restrict__ head) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB)
        head[i] = keys[i] != kEmptyKey && (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// unique keys with the OR of their values
__global__ void k_unique_or(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                            const uint32_t* __restrict__ head, const uint32_t* __restrict__ excl, size_t n,
                            uint64_t* __restrict__ okeys, uint32_t* __restrict__ ovals) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) {
        if (keys[i] == kEmptyKey) continue;
        const uint32_t seg = excl[i] + head[i] - 1u;
        if (head[i]) okeys[seg] = keys[i];
        if (vals[i]) atomicOr(&ovals[seg], vals[i]);
    }
}

// order key: (instance, Hilbert key of the brick coordinates on 16 levels)
// (instance, Hilbert key of 16 levels) packed as inst << hbits | key.  The
// curve's start cell is the origin at every scale, so brick coordinates
// below 2^k give keys below 2^(3k): with hbits = 3k the packing keeps the
// order of the pair and the sort only needs its low bits.
__global__ void k_order_keys(const uint64_t* __restrict__ bkeys, size_t nb, int hbits, uint64_t* __restrict__ okey,
                             uint32_t* __restrict__ oidx, uint32_t* __restrict__ err) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < nb; i += (size_t)gridDim.x * kB) {
        const uint64_t k = bkeys[i];
        const uint64_t h = hilbert_encode_direct((uint32_t)bkey_coord(k, 0), (uint32_t)bkey_coord(k, 1),
                                                 (uint32_t)bkey_coord(k, 2), 16);
        if (h >> hbits) atomicOr(err, kErrOrder);
        okey[i] = ((k >> 48) << hbits) | h;
        oidx[i] = (uint32_t)i;
    }
}

__global__ void k_gather_flag(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ flag, size_t n,
                              uint32_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) out[i] = flag[idx[i]];
}

// stable partition: flagged first (in order), then the rest (in order)
__global__ void k_partition(const uint32_t* __restrict__ idx, const uint32_t* __restrict__ flag,
                            const uint32_t* __restrict__ excl, size_t n, size_t nflag, uint32_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) {
        const size_t p = flag[i] ? excl[i] : nflag + (i - excl[i]);
        out[p] = idx[i];
    }
}

__global__ void k_bricks_final(const uint32_t* __restrict__ order, const uint64_t* __restrict__ ukeys, size_t nb,
                               uint64_t* __restrict__ bkey, int32_t* __restrict__ bcoord,
                               uint16_t* __restrict__ binst) {
    for (size_t b = (size_t)blockIdx.x * kB + threadIdx.x; b < nb; b += (size_t)gridDim.x * kB) {
        const uint64_t k = ukeys[order[b]];
        bkey[b] = k;
        for (int a = 0; a < 3; ++a) bcoord[3 * b + a] = bkey_coord(k, a);
        binst[b] = (uint16_t)(k >> 48);
    }
}

__global__ void k_bkeys_of(const int32_t* __restrict__ bcoord, const uint16_t* __restrict__ binst, size_t nb,
                           uint64_t* __restrict__ bkey) {
    for (size_t b = (size_t)blockIdx.x * kB + threadIdx.x; b < nb; b += (size_t)gridDim.x * kB)
        bkey[b] = pack_bkey(binst[b], bcoord[3 * b], bcoord[3 * b + 1], bcoord[3 * b + 2]);
}

__global__ void k_nbr(const int32_t* __restrict__ bcoord, const uint16_t* __restrict__ binst, size_t nb, int l,
                      const LayoutInst* __restrict__ inst, LevelMap M, int32_t* __restrict__ nbr) {
    for (size_t t = (size_t)blockIdx.x * kB + threadIdx.x; t < nb * 27; t += (size_t)gridDim.x * kB) {
        const size_t b = t / 27;
        const int k = (int)(t % 27);
        const uint32_t in = binst[b];
        int32_t c[3] = {bcoord[3 * b] + k % 3 - 1, bcoord[3 * b + 1] + (k / 3) % 3 - 1, bcoord[3 * b + 2] + k / 9 - 1};
        int32_t v = -1;
        if (wrap_brick_d(inst[in], l, c)) {
            const uint32_t r = cell_lookup(M.keys, M.vals, M.cap, pack_bkey(in, c[0], c[1], c[2]));
            v = r == kInvalid ? -1 : (int32_t)r;
        }
        nbr[t] = v;
    }
}

// leaf slot kinds; fluid leaves counted per instance
__global__ void k_place(const int32_t* __restrict__ X, const int32_t* __restrict__ Y, const int32_t* __restrict__ Z,
                        const uint8_t* __restrict__ LV, const uint8_t* __restrict__ ST,
                        const uint16_t* __restrict__ linst, size_t n, int l, const LayoutInst* __restrict__ inst,
                        LevelMap M, uint8_t* __restrict__ stype, unsigned long long* __restrict__ nfl,
                        uint32_t* __restrict__ err) {
    uint32_t e = 0;
    WaveCounter wc;
    // whole waves iterate together (the count below is a wave operation)
    for (size_t base = (size_t)blockIdx.x * kB; base < n; base += (size_t)gridDim.x * kB) {
        const size_t j = base + threadIdx.x;
        bool fluid = false;
        int in = 0;
        if (j < n && LV[j] == l) {
            in = linst[j];
            const int64_t s = map_slot(M, inst[in], in, l, X[j], Y[j], Z[j]);
            if (s < 0) {
                e |= kErrMissing;
            } else {
                fluid = ST[j] == 0;
                stype[s] = fluid ? T_FLUID : T_SOLID;
            }
        }
        wc.add(nfl, (uint32_t)in, fluid, 1ull);
    }
    wc.flush(nfl);
    if (e) atomicOr(err, e);
}

// the kinds of the other slots, from the leaves one level up and down
__global__ void k_classify(size_t nb, int l, const int32_t* __restrict__ bcoord, const uint16_t* __restrict__ binst,
                           const LayoutInst* __restrict__ inst, uint8_t* __restrict__ stype, int has_coarse,
                           LevelMap C, const uint8_t* __restrict__ ctype, int has_fine, LevelMap F,
                           const uint8_t* __restrict__ ftype) {
    for (size_t s = (size_t)blockIdx.x * kB + threadIdx.x; s < nb * 64; s += (size_t)gridDim.x * kB) {
        if (stype[s] != T_NONE) continue;
        const size_t b = s >> 6;
        const int loc = (int)(s & 63);
        const int in = binst[b];
        const LayoutInst& I = inst[in];
        const int32_t x = bcoord[3 * b] * 4 + (loc & 3), y = bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                      z = bcoord[3 * b + 2] * 4 + (loc >> 4);
        if (has_coarse) {
            const int64_t cs = map_slot(C, I, in, l - 1, x >> 1, y >> 1, z >> 1);
            const uint8_t ct = cs < 0 ? T_NONE : ctype[cs];
            if (ct == T_FLUID) { stype[s] = T_C2F; continue; }
            if (ct == T_SOLID) { stype[s] = T_WALL; continue; }
        }
        if (has_fine) {
            int nf = 0, ns = 0;
            for (int q = 0; q < 8; ++q) {
                const int64_t fs = map_slot(F, I, in, l + 1, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
                const uint8_t t = fs < 0 ? T_NONE : ftype[fs];
                nf += t == T_FLUID;
                ns += t == T_SOLID;
            }
            if (nf == 8) stype[s] = T_F2C;
            else if (nf + ns == 8) stype[s] = T_WALL;
        }
    }
}

// link words of the fluid leaves, one wave per brick (lane = cell)
__global__ __launch_bounds__(kB) void k_links(size_t nb, int l, const int32_t* __restrict__ bcoord,
                                              const uint16_t* __restrict__ binst,
                                              const LayoutInst* __restrict__ inst, const int32_t* __restrict__ nbr,
                                              const uint8_t* __restrict__ stype, LevelMap M,
                                              uint64_t* __restrict__ cflags, uint32_t* __restrict__ needed,
                                              uint32_t* __restrict__ fmask, uint32_t* __restrict__ bflags,
                                              uint32_t* __restrict__ bcnt, uint32_t* __restrict__ err) {
    const size_t b = (size_t)blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
    if (b >= nb) return;
    const int loc = (int)(threadIdx.x & 63);
    const size_t s = b * 64 + loc;
    const int in = binst[b];
    const LayoutInst& I = inst[in];
    const bool fluid = stype[s] == T_FLUID;
    bool simple = fluid;
    uint64_t w = 0;
    uint32_t fm = 0;
    if (fluid) {
        const int32_t x = bcoord[3 * b] * 4 + (loc & 3), y = bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                      z = bcoord[3 * b + 2] * 4 + (loc >> 4);
        w = kActiveBit;
        for (int i = 1; i < Q; ++i) {
            int s3[3], face;
            unsigned code;
            if (source_of_d(I, l, x, y, z, i, s3, &face)) {
                if (face >= 0) {
                    code = kLinkWall;
                } else {
                    const int64_t s2 = map_slot(M, I, in, l, s3[0], s3[1], s3[2]);
                    const uint8_t t2 = s2 < 0 ? T_NONE : stype[s2];
                    if (regular_t(t2)) {
                        code = kLinkOutflow;
                        // the outlet reads every population of its source
                        if (t2 != T_FLUID) atomicOr(&needed[s2], (1u << Q) - 1u);
                    } else if (t2 == T_SOLID || t2 == T_WALL) {
                        code = kLinkSolid;
                    } else {
                        atomicOr(err, kErrLink);
                        code = kLinkSolid;
                    }
                }
            } else {
                const int px = (loc & 3) - cx(i), py = ((loc >> 2) & 3) - cy(i), pz = (loc >> 4) - cz(i);
                const int nb2 = nbr[b * 27 + 13 + (px >> 2) + 3 * (py >> 2) + 9 * (pz >> 2)];
                const int64_t s2 = nb2 < 0 ? -1 : (int64_t)nb2 * 64 + (px & 3) + 4 * (py & 3) + 16 * (pz & 3);
                const uint8_t t2 = s2 < 0 ? T_NONE : stype[s2];
                if (regular_t(t2)) {
                    code = kLinkRegular;
                    if (t2 != T_FLUID) atomicOr(&needed[s2], 1u << i);
                } else if (t2 == T_SOLID || t2 == T_WALL) {
                    code = kLinkSolid;
                    fm |= 1u << i;
                } else {
                    atomicOr(err, kErrLink);
                    code = kLinkSolid;
                }
            }
            w |= (uint64_t)code << (2 * i);
            if (code != kLinkRegular) simple = false;
        }
    }
    cflags[s] = w;
    fmask[s] = fm;
    const bool all = __ballot(!simple) == 0ull;
    uint32_t c = (uint32_t)__popc(fm);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_xor(c, d, 64);
    if (loc == 0) {
        bflags[b] = all ? 1u : 0u;
        bcnt[b] = c;
    }
}

// force links of the active bricks in slot order (slot * 32 + direction)
__global__ __launch_bounds__(kB) void k_force_write(size_t nact, const uint32_t* __restrict__ fmask,
                                                    const uint32_t* __restrict__ boff,
                                                    uint32_t* __restrict__ out) {
    const size_t b = (size_t)blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
    if (b >= nact) return;
    const int loc = (int)(threadIdx.x & 63);
    const size_t s = b * 64 + loc;
    const uint32_t fm = fmask[s];
    const uint32_t c = (uint32_t)__popc(fm);
    uint32_t inc = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(inc, d, 64);
        if (loc >= d) inc += v;
    }
    uint32_t p = boff[b] + inc - c;
    for (int i = 1; i < Q; ++i)
        if ((fm >> i) & 1u) out[p++] = (uint32_t)(s * 32 + i);
}

// per-instance active bricks and force links
__global__ void k_inst_counts(size_t nact, const uint16_t* __restrict__ binst, const uint32_t* __restrict__ bcnt,
                              unsigned long long* __restrict__ acnt, unsigned long long* __restrict__ lcnt) {
    WaveCounter wa, wl;
    for (size_t base = (size_t)blockIdx.x * kB; base < nact; base += (size_t)gridDim.x * kB) {
        const size_t b = base + threadIdx.x;
        const bool on = b < nact;
        const uint32_t in = on ? binst[b] : 0u;
        wa.add(acnt, in, on, 1ull);
        wl.add(lcnt, in, on && bcnt[b] != 0u, on ? (unsigned long long)bcnt[b] : 0ull);
    }
    wa.flush(acnt);
    wl.flush(lcnt);
}

__global__ void k_ghost_flags(const uint8_t* __restrict__ stype, const uint32_t* __restrict__ needed, size_t n,
                              uint8_t kind, uint32_t* __restrict__ flag) {
    for (size_t s = (size_t)blockIdx.x * kB + threadIdx.x; s < n; s += (size_t)gridDim.x * kB)
        flag[s] = stype[s] == kind && needed[s] ? 1u : 0u;
}

__global__ void k_scatter_idx(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ excl, size_t n,
                              uint32_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB)
        if (flag[i]) out[excl[i]] = (uint32_t)i;
}

__global__ void k_scatter_val(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ excl, size_t n,
                              uint32_t skip, uint64_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB)
        if (vals[i] != skip) out[excl[i]] = vals[i];
}

__global__ void k_ne_flag(const uint32_t* __restrict__ vals, size_t n, uint32_t skip, uint32_t* __restrict__ flag) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB)
        flag[i] = vals[i] != skip ? 1u : 0u;
}

// covering coarse leaf and child position of every c2f ghost
__global__ void k_c2f_parent(const uint32_t* __restrict__ ghost, size_t ng, const int32_t* __restrict__ bcoord,
                             const uint16_t* __restrict__ binst, const LayoutInst* __restrict__ inst, int l,
                             LevelMap C, uint64_t* __restrict__ par, uint32_t* __restrict__ idx,
                             uint8_t* __restrict__ chq, uint32_t* __restrict__ err) {
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < ng; k += (size_t)gridDim.x * kB) {
        const uint32_t g = ghost[k];
        const size_t b = g >> 6;
        const int loc = (int)(g & 63);
        const int in = binst[b];
        const int32_t fx = bcoord[3 * b] * 4 + (loc & 3), fy = bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                      fz = bcoord[3 * b + 2] * 4 + (loc >> 4);
        const int64_t P = map_slot(C, inst[in], in, l - 1, fx >> 1, fy >> 1, fz >> 1);
        if (P < 0) atomicOr(err, kErrParent);
        par[k] = P < 0 ? 0ull : (uint64_t)P;
        idx[k] = (uint32_t)k;
        chq[k] = (uint8_t)((fx & 1) | ((fy & 1) << 1) | ((fz & 1) << 2));
    }
}

__global__ void k_group_fill(const uint64_t* __restrict__ par, const uint32_t* __restrict__ idx,
                             const uint32_t* __restrict__ head, const uint32_t* __restrict__ excl, size_t ng,
                             const uint32_t* __restrict__ ghost, const uint8_t* __restrict__ chq,
                             const uint32_t* __restrict__ needed, uint32_t* __restrict__ gchild,
                             uint32_t* __restrict__ gmask, uint32_t* __restrict__ gpar) {
    for (size_t e = (size_t)blockIdx.x * kB + threadIdx.x; e < ng; e += (size_t)gridDim.x * kB) {
        const uint32_t g = excl[e] + head[e] - 1u;
        const uint32_t k = idx[e];
        const uint32_t gs = ghost[k];
        const int q = chq[k];
        gchild[8 * (size_t)g + q] = gs;
        gmask[8 * (size_t)g + q] = needed[gs];
        if (head[e]) gpar[g] = (uint32_t)par[e];
    }
}

// capture sources of each group: the covering leaf P, then P's coarse fluid
// neighbours -x, +x, -y, +y, -z, +z (across periodic faces)
__global__ void k_group_src(const uint32_t* __restrict__ gpar, size_t G, const int32_t* __restrict__ cbcoord,
                            const uint16_t* __restrict__ cbinst, const LayoutInst* __restrict__ inst, int lc,
                            LevelMap C, const uint8_t* __restrict__ ctype, uint32_t* __restrict__ src) {
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < G; k += (size_t)gridDim.x * kB) {
        const uint32_t P = gpar[k];
        const size_t cb = P >> 6;
        const int loc = (int)(P & 63);
        const int in = cbinst[cb];
        const LayoutInst& I = inst[in];
        const int32_t pc[3] = {cbcoord[3 * cb] * 4 + (loc & 3), cbcoord[3 * cb + 1] * 4 + ((loc >> 2) & 3),
                               cbcoord[3 * cb + 2] * 4 + (loc >> 4)};
        src[k * kC2FCaps] = P;
        for (int a = 0; a < 3; ++a)
            for (int d = 0; d < 2; ++d) {
                uint32_t v = kNoCap;
                int32_t q[3] = {pc[0], pc[1], pc[2]};
                q[a] += d ? 1 : -1;
                const int32_t e = I.ext0[a] << lc;
                bool ok = true;
                if (q[a] < 0 || q[a] >= e) {
                    if (I.bc[2 * a + d] != FL_BC_PERIODIC) ok = false;
                    else q[a] = (q[a] + e) % e;
                }
                if (ok) {
                    const int64_t s2 = map_slot(C, I, in, lc, q[0], q[1], q[2]);
                    if (s2 >= 0 && ctype[s2] == T_FLUID) v = (uint32_t)s2;
                }
                src[k * kC2FCaps + 1 + 2 * a + d] = v;
            }
    }
}

__global__ void k_lower_bound(const uint32_t* __restrict__ src, size_t n, const uint32_t* __restrict__ cap,
                              size_t ncap, uint32_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) {
        const uint32_t v = src[i];
        if (v == kNoCap) {
            out[i] = kNoCap;
            continue;
        }
        size_t lo = 0, hi = ncap;
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (cap[mid] < v) lo = mid + 1; else hi = mid;
        }
        out[i] = (uint32_t)lo;
    }
}

__global__ void k_u64_to_u32(const uint64_t* __restrict__ a, size_t n, uint32_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) out[i] = (uint32_t)a[i];
}

// the 8 children of every f2c ghost (fine-coordinate parity order) and the
// ordering key (first child's slot)
__global__ void k_f2c_children(const uint32_t* __restrict__ ghost, size_t ng, const int32_t* __restrict__ bcoord,
                               const uint16_t* __restrict__ binst, const LayoutInst* __restrict__ inst, int l,
                               LevelMap F, uint32_t* __restrict__ child8, uint64_t* __restrict__ key,
                               uint32_t* __restrict__ idx) {
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < ng; k += (size_t)gridDim.x * kB) {
        const uint32_t g = ghost[k];
        const size_t b = g >> 6;
        const int loc = (int)(g & 63);
        const int in = binst[b];
        const int32_t x = bcoord[3 * b] * 4 + (loc & 3), y = bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                      z = bcoord[3 * b + 2] * 4 + (loc >> 4);
        for (int q = 0; q < 8; ++q) {
            const int64_t c = map_slot(F, inst[in], in, l + 1, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
            child8[8 * k + q] = (uint32_t)c;
        }
        key[k] = child8[8 * k];
        idx[k] = (uint32_t)k;
    }
}

__global__ void k_f2c_perm(const uint32_t* __restrict__ idx, size_t ng, const uint32_t* __restrict__ ghost,
                           const uint32_t* __restrict__ child8, uint32_t* __restrict__ og, uint32_t* __restrict__ oc) {
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < ng; k += (size_t)gridDim.x * kB) {
        const uint32_t j = idx[k];
        og[k] = ghost[j];
        for (int q = 0; q < 8; ++q) oc[8 * k + q] = child8[8 * (size_t)j + q];
    }
}

// ------------------------------------------------------- queries, remap
__global__ void k_query(LevelMaps maps, const LayoutInst* __restrict__ inst, int in, const int32_t* __restrict__ X,
                        const int32_t* __restrict__ Y, const int32_t* __restrict__ Z, const uint8_t* __restrict__ LV,
                        size_t n, int64_t* __restrict__ out) {
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < n; k += (size_t)gridDim.x * kB) {
        const int l = LV[k];
        int64_t s = -1;
        if (l < maps.nlevels) {
            const LevelMap& M = maps.m[l];
            s = map_slot(M, inst[in], in, l, X[k], Y[k], Z[k]);
            if (s >= 0 && !(M.cflags[s] & kActiveBit)) s = -1;
        }
        out[k] = s;
    }
}

// 1 copy, 2 split, 3 merge, 4 rest, 0 not a fluid leaf
__global__ void k_remap_class(LevelMaps om, const LayoutInst* __restrict__ inst, int l, size_t nact,
                              const int32_t* __restrict__ bcoord, const uint16_t* __restrict__ binst,
                              const uint64_t* __restrict__ cflags, uint8_t* __restrict__ cls,
                              uint32_t* __restrict__ src) {
    for (size_t s = (size_t)blockIdx.x * kB + threadIdx.x; s < nact * 64; s += (size_t)gridDim.x * kB) {
        uint8_t c = 0;
        uint32_t v = 0;
        if (cflags[s] & kActiveBit) {
            const size_t b = s >> 6;
            const int loc = (int)(s & 63);
            const int in = binst[b];
            const LayoutInst& I = inst[in];
            const int32_t x = bcoord[3 * b] * 4 + (loc & 3), y = bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                          z = bcoord[3 * b + 2] * 4 + (loc >> 4);
            c = 4;
            if (l < om.nlevels) {
                const LevelMap& M = om.m[l];
                const int64_t os = map_slot(M, I, in, l, x, y, z);
                if (os >= 0 && (M.cflags[os] & kActiveBit)) {
                    c = 1;
                    v = (uint32_t)os;
                }
            }
            if (c == 4 && l > 0 && l - 1 < om.nlevels) {
                const LevelMap& M = om.m[l - 1];
                const int64_t ps = map_slot(M, I, in, l - 1, x >> 1, y >> 1, z >> 1);
                if (ps >= 0 && (M.cflags[ps] & kActiveBit)) {
                    c = 2;
                    v = (uint32_t)ps;
                }
            }
            if (c == 4 && l + 1 < om.nlevels) {
                const LevelMap& M = om.m[l + 1];
                bool all = true;
                for (int q = 0; q < 8 && all; ++q) {
                    const int64_t ch = map_slot(M, I, in, l + 1, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
                    all = ch >= 0 && (M.cflags[ch] & kActiveBit);
                }
                if (all) c = 3;
            }
        }
        cls[s] = c;
        src[s] = v;
    }
}

__global__ void k_class_flag(const uint8_t* __restrict__ cls, size_t n, uint8_t want, uint32_t* __restrict__ flag) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB)
        flag[i] = cls[i] == want ? 1u : 0u;
}

__global__ void k_gather_u32(const uint32_t* __restrict__ idx, size_t n, const uint32_t* __restrict__ v,
                             uint32_t* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * kB + threadIdx.x; i < n; i += (size_t)gridDim.x * kB) out[i] = v[idx[i]];
}

__global__ void k_merge_src(LevelMaps om, const LayoutInst* __restrict__ inst, int l, const uint32_t* __restrict__ dst,
                            size_t n, const int32_t* __restrict__ bcoord, const uint16_t* __restrict__ binst,
                            uint32_t* __restrict__ src8) {
    for (size_t k = (size_t)blockIdx.x * kB + threadIdx.x; k < n; k += (size_t)gridDim.x * kB) {
        const uint32_t s = dst[k];
        const size_t b = s >> 6;
        const int loc = (int)(s & 63);
        const int in = binst[b];
        const int32_t x = bcoord[3 * b] * 4 + (loc & 3), y = bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                      z = bcoord[3 * b + 2] * 4 + (loc >> 4);
        for (int q = 0; q < 8; ++q)
            src8[8 * k + q] = (uint32_t)map_slot(om.m[l + 1], inst[in], in, l + 1, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1),
                                                 2 * z + ((q >> 2) & 1));
    }
}

// ------------------------------------------------------- host helpers
// Temporaries of one scope on stream s, from the block cache; returned to it
// when the scope ends.
struct Tmp {
    hipStream_t s;
    std::vector<void*> p;
    explicit Tmp(hipStream_t st) : s(st) {}
    Tmp(const Tmp&) = delete;
    Tmp& operator=(const Tmp&) = delete;
    ~Tmp() {
        for (void* q : p) cache_free(q, s);
    }
    template <typename T>
    fl_status get(T** out, size_t n) {
        void* q = nullptr;
        const fl_status st = cache_alloc(&q, (n ? n : 1) * sizeof(T), s);
        if (st) return st;
        p.push_back(q);
        *out = static_cast<T*>(q);
        return FL_OK;
    }
};


template <typename T>
fl_status pget(const DevAlloc& alloc, T** out, size_t n) {
    *out = nullptr;
    if (n == 0) return FL_OK;
    void* q = nullptr;
    FL_TRY(alloc(&q, n * sizeof(T)));
    *out = static_cast<T*>(q);
    return FL_OK;
}

fl_status sync_read(void* h, const void* d, size_t bytes, hipStream_t s) {
    FL_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s));
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
}

// exclusive scan of n u32 flags/counts into out; *total = their sum
fl_status scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint64_t* total, hipStream_t s) {
    *total = 0;
    if (n == 0) return FL_OK;
    Tmp t(s);
    const size_t tb = fl_scan_temp_bytes(n);
    uint8_t* ws;
    uint64_t* dt;
    FL_TRY(t.get(&ws, tb));
    FL_TRY(t.get(&dt, 1));
    FL_TRY(fl_scan_exclusive_u32(in, out, dt, n, ws, tb, reinterpret_cast<fl_stream>(s)));
    return sync_read(total, dt, 8, s);
}

// stable sort of (keys, vals) over bits [0, end_bit)
fl_status sort_pairs(uint64_t* k, uint32_t* v, size_t n, int end_bit, hipStream_t s) {
    if (n <= 1) return FL_OK;
    Tmp t(s);
    uint64_t* ka;
    uint32_t* va;
    uint8_t* ws;
    const size_t tb = fl_sort_temp_bytes(n);
    FL_TRY(t.get(&ka, n));
    FL_TRY(t.get(&va, n));
    FL_TRY(t.get(&ws, tb));
    FL_TRY(fl_sort_pairs_u64u32(k, v, ka, va, n, 0, end_bit, ws, tb, reinterpret_cast<fl_stream>(s)));
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
}

// unique keys (with the OR of their values) of a sorted array, EMPTY keys dropped
fl_status unique_or(Tmp& t, const uint64_t* keys, const uint32_t* vals, size_t n, uint64_t** okeys, uint32_t** ovals,
                    size_t* m, hipStream_t s) {
    uint32_t *head, *excl;
    FL_TRY(t.get(&head, n));
    FL_TRY(t.get(&excl, n));
    hipLaunchKernelGGL(k_heads, dim3(blocks_for(n)), dim3(kB), 0, s, keys, n, head);
    FL_CHECK_LAUNCH();
    uint64_t tot = 0;
    FL_TRY(scan_u32(head, excl, n, &tot, s));
    *m = tot;
    FL_TRY(t.get(okeys, tot));
    FL_TRY(t.get(ovals, tot));
    hipLaunchKernelGGL(k_fill32, dim3(blocks_for(tot)), dim3(kB), 0, s, *ovals, tot, 0u);
    hipLaunchKernelGGL(k_unique_or, dim3(blocks_for(n)), dim3(kB), 0, s, keys, vals, head, excl, n, *okeys, *ovals);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

// indices i < n with flag[i] != 0, in order (into `out`, allocated with `get`)
template <typename Get>
fl_status compact(const uint32_t* flag, size_t n, uint32_t** out, size_t* m, hipStream_t s, Get&& get) {
    Tmp t(s);
    uint32_t* excl;
    FL_TRY(t.get(&excl, n));
    uint64_t tot = 0;
    FL_TRY(scan_u32(flag, excl, n, &tot, s));
    *m = tot;
    FL_TRY(get(out, tot));
    if (tot) {
        hipLaunchKernelGGL(k_scatter_idx, dim3(blocks_for(n)), dim3(kB), 0, s, flag, excl, n, *out);
        FL_CHECK_LAUNCH();
    }
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
}

LevelMap level_map(const LayoutLevel& L) { return LevelMap{L.map_keys, L.map_vals, L.map_cap, L.cflags}; }

fl_status map_build(const uint64_t* bkey, uint32_t nb, hipStream_t s, const DevAlloc& alloc, LayoutLevel& L) {
    L.map_cap = brick_map_capacity(nb);
    FL_TRY(pget(alloc, &L.map_keys, L.map_cap));
    FL_TRY(pget(alloc, &L.map_vals, L.map_cap));
    FL_TRY(hash_clear_impl(L.map_keys, L.map_vals, L.map_cap, s));
    if (nb) FL_TRY(cellmap_build_impl(L.map_keys, L.map_vals, L.map_cap, bkey, nb, s));
    return FL_OK;
}

std::string mesh_message(uint32_t e) {
    if (e & kErrLevel) return "leaf level above the instance's max_level";
    if (e & kErrDomain) return "leaf outside the domain";
    if (e & kErrMissing) return "internal: leaf brick missing";
    if (e & kErrDup) return "duplicate leaf";
    if (e & kErrLink)
        return "a fluid leaf reads a position that is neither a leaf nor a ghost "
               "(mesh not 2:1 balanced over D3Q19 neighbour cubes)";
    if (e & kErrParent) return "c2f ghost with no covering coarse leaf";
    if (e & kErrOrder) return "internal: brick order key out of range";
    return "mesh error";
}

}  // namespace

fl_status build_brick_map(const int32_t* bcoord, const uint16_t* binst, uint32_t nb, hipStream_t s,
                          const DevAlloc& alloc, LayoutLevel& L) {
    Tmp t(s);
    uint64_t* bk;
    FL_TRY(t.get(&bk, nb));
    if (nb) {
        hipLaunchKernelGGL(k_bkeys_of, dim3(blocks_for(nb)), dim3(kB), 0, s, bcoord, binst, (size_t)nb, bk);
        FL_CHECK_LAUNCH();
    }
    FL_TRY(map_build(bk, nb, s, alloc, L));
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
}

fl_status build_layout_device(const std::vector<LayoutInst>& inst, int NL, const size_t* h_offsets, const int32_t* X,
                              const int32_t* Y, const int32_t* Z, const uint8_t* LV, const uint8_t* ST, size_t n,
                              hipStream_t s, const DevAlloc& alloc, LayoutLevel* out, std::string* err) {
    const int NI = (int)inst.size();
    FL_REQUIRE(NL >= 1 && NL <= kMaxLevels, "level count out of range");
    Tmp g(s);  // lives for the whole build
    LayoutInst* dinst;
    uint64_t* doffs;
    uint32_t* derr;
    uint16_t* linst;
    unsigned long long* lcount;
    FL_TRY(g.get(&dinst, NI));
    FL_TRY(g.get(&doffs, NI + 1));
    FL_TRY(g.get(&derr, 1));
    FL_TRY(g.get(&linst, n));
    FL_TRY(g.get(&lcount, kMaxLevels));
    std::vector<uint64_t> hoffs(h_offsets, h_offsets + NI + 1);
    FL_HIP(hipMemcpyAsync(dinst, inst.data(), NI * sizeof(LayoutInst), hipMemcpyHostToDevice, s));
    FL_HIP(hipMemcpyAsync(doffs, hoffs.data(), (NI + 1) * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kB), 0, s, derr, (size_t)1, 0u);
    hipLaunchKernelGGL(k_fill64, dim3(1), dim3(kB), 0, s, reinterpret_cast<uint64_t*>(lcount), (size_t)kMaxLevels,
                       (uint64_t)0);
    if (n) {
        hipLaunchKernelGGL(k_leaf_scan, dim3(kCountGrid), dim3(kB), 0, s, X, Y, Z, LV, n, doffs, NI, dinst, linst,
                           lcount, derr);
        FL_CHECK_LAUNCH();
    }
    uint32_t herr = 0;
    unsigned long long hcount[kMaxLevels];
    FL_TRY(sync_read(&herr, derr, 4, s));
    FL_TRY(sync_read(hcount, lcount, sizeof(hcount), s));
    if (herr) {
        *err = mesh_message(herr);
        return FL_E_MESH;
    }
    auto palloc = [&](auto** p, size_t k) { return pget(alloc, p, k); };
    std::vector<uint8_t*> stype(NL, nullptr);
    std::vector<uint32_t*> needed(NL, nullptr);
    // ---- bricks, brick maps, neighbour rows, leaf kinds (per level)
    for (int l = 0; l < NL; ++l) {
        LayoutLevel& L = out[l];
        L = LayoutLevel();
        const size_t nl = hcount[l];
        Tmp t(s);
        uint64_t* bkeys = nullptr;
        uint32_t* bact = nullptr;
        size_t nb = 0;
        if (nl) {
            // leaf bricks: a set keyed by brick (value 0 when active)
            const uint64_t capa = brick_map_capacity(nl);
            uint64_t *ta, *lkeys;
            uint32_t *va, *fl, *ex, *lact;
            FL_TRY(t.get(&ta, capa));
            FL_TRY(t.get(&va, capa));
            FL_TRY(t.get(&fl, capa));
            FL_TRY(t.get(&ex, capa));
            FL_TRY(hash_clear_impl(ta, va, capa, s));
            hipLaunchKernelGGL(k_insert_leaf_bricks, dim3(blocks_for(n)), dim3(kB), 0, s, X, Y, Z, LV, ST, linst, n, l,
                               ta, va, capa);
            hipLaunchKernelGGL(k_occupied, dim3(blocks_for(capa)), dim3(kB), 0, s, ta, capa, fl);
            FL_CHECK_LAUNCH();
            uint64_t m = 0;
            FL_TRY(scan_u32(fl, ex, capa, &m, s));
            FL_TRY(t.get(&lkeys, m));
            FL_TRY(t.get(&lact, m));
            hipLaunchKernelGGL(k_extract, dim3(blocks_for(capa)), dim3(kB), 0, s, ta, va, capa, fl, ex, lkeys, lact);
            FL_CHECK_LAUNCH();
            uint64_t na = 0;
            {
                uint32_t* ex2;
                FL_TRY(t.get(&ex2, m));
                FL_TRY(scan_u32(lact, ex2, m, &na, s));
            }
            // the level's bricks: leaf bricks plus the active ones' dilation
            const uint64_t capb = brick_map_capacity(m + 26 * na);
            uint64_t* tb;
            uint32_t *vb, *fb, *eb;
            FL_TRY(t.get(&tb, capb));
            FL_TRY(t.get(&vb, capb));
            FL_TRY(t.get(&fb, capb));
            FL_TRY(t.get(&eb, capb));
            FL_TRY(hash_clear_impl(tb, vb, capb, s));
            hipLaunchKernelGGL(k_insert_dilated, dim3(blocks_for(m)), dim3(kB), 0, s, lkeys, lact, (size_t)m, l, dinst,
                               tb, vb, capb);
            hipLaunchKernelGGL(k_occupied, dim3(blocks_for(capb)), dim3(kB), 0, s, tb, capb, fb);
            FL_CHECK_LAUNCH();
            uint64_t nbt = 0;
            FL_TRY(scan_u32(fb, eb, capb, &nbt, s));
            nb = nbt;
            FL_TRY(t.get(&bkeys, nb));
            FL_TRY(t.get(&bact, nb));
            hipLaunchKernelGGL(k_extract, dim3(blocks_for(capb)), dim3(kB), 0, s, tb, vb, capb, fb, eb, bkeys, bact);
            FL_CHECK_LAUNCH();
        }
        if ((uint64_t)nb * 64 >= (1ull << 32) / 32) {
            *err = "level too large for 32-bit slot lists";
            return FL_E_MESH;
        }
        L.nb = (uint32_t)nb;
        // order: stable by (instance, Hilbert key), then active bricks first
        uint64_t* okey;
        uint32_t *oidx, *flag, *excl, *order;
        FL_TRY(t.get(&okey, nb));
        FL_TRY(t.get(&oidx, nb));
        FL_TRY(t.get(&flag, nb));
        FL_TRY(t.get(&excl, nb));
        FL_TRY(t.get(&order, nb));
        uint64_t nact = 0;
        if (nb) {
            // bits of the largest brick coordinate and of the instance index
            int kb = 1, ib = 0;
            for (const LayoutInst& I : inst)
                for (int a = 0; a < 3; ++a)
                    while ((((int64_t)I.ext0[a] << l) + 3) / 4 > (int64_t(1) << kb)) ++kb;
            while ((NI - 1) >> ib) ++ib;
            if (kb > 16) {
                *err = "brick coordinates beyond the 16-level Hilbert order";
                return FL_E_MESH;
            }
            hipLaunchKernelGGL(k_order_keys, dim3(blocks_for(nb)), dim3(kB), 0, s, bkeys, nb, 3 * kb, okey, oidx,
                               derr);
            FL_CHECK_LAUNCH();
            FL_TRY(sort_pairs(okey, oidx, nb, 3 * kb + ib, s));
            hipLaunchKernelGGL(k_gather_flag, dim3(blocks_for(nb)), dim3(kB), 0, s, oidx, bact, nb, flag);
            FL_TRY(scan_u32(flag, excl, nb, &nact, s));
            hipLaunchKernelGGL(k_partition, dim3(blocks_for(nb)), dim3(kB), 0, s, oidx, flag, excl, nb, (size_t)nact,
                               order);
            FL_CHECK_LAUNCH();
        }
        L.nactive = (uint32_t)nact;
        uint64_t* bkey;
        FL_TRY(g.get(&bkey, nb));
        FL_TRY(palloc(&L.bcoord, nb * 3));
        FL_TRY(palloc(&L.binst, nb));
        if (nb) {
            hipLaunchKernelGGL(k_bricks_final, dim3(blocks_for(nb)), dim3(kB), 0, s, order, bkeys, nb, bkey, L.bcoord,
                               L.binst);
            FL_CHECK_LAUNCH();
        }
        FL_TRY(map_build(bkey, (uint32_t)nb, s, alloc, L));
        FL_TRY(palloc(&L.nbr, nb * 27));
        if (nb) {
            hipLaunchKernelGGL(k_nbr, dim3(blocks_for(nb * 27)), dim3(kB), 0, s, L.bcoord, L.binst, nb, l, dinst,
                               level_map(L), L.nbr);
            FL_CHECK_LAUNCH();
        }
        // leaf kinds
        const size_t slots = nb * 64;
        unsigned long long *nfl, *placed;
        FL_TRY(g.get(&stype[l], slots));
        FL_TRY(g.get(&needed[l], slots));
        FL_TRY(t.get(&nfl, NI + 1));
        placed = nfl + NI;
        hipLaunchKernelGGL(k_fill8, dim3(blocks_for(slots)), dim3(kB), 0, s, stype[l], slots, (uint8_t)T_NONE);
        hipLaunchKernelGGL(k_fill32, dim3(blocks_for(slots)), dim3(kB), 0, s, needed[l], slots, 0u);
        hipLaunchKernelGGL(k_fill64, dim3(blocks_for(NI + 1)), dim3(kB), 0, s, reinterpret_cast<uint64_t*>(nfl),
                           (size_t)(NI + 1), (uint64_t)0);
        if (nl) {
            hipLaunchKernelGGL(k_place, dim3(kCountGrid), dim3(kB), 0, s, X, Y, Z, LV, ST, linst, n, l, dinst,
                               level_map(L), stype[l], nfl, derr);
            hipLaunchKernelGGL(k_count_placed, dim3(kCountGrid), dim3(kB), 0, s, stype[l], slots, placed);
            FL_CHECK_LAUNCH();
        }
        L.nfluid_inst.assign(NI + 1, 0);
        FL_TRY(sync_read(L.nfluid_inst.data(), nfl, (NI + 1) * 8, s));
        const uint32_t dup = L.nfluid_inst[NI] != nl ? kErrDup : 0u;
        L.nfluid_inst.resize(NI);
        L.nfluid = 0;
        for (uint64_t v : L.nfluid_inst) L.nfluid += v;
        FL_TRY(sync_read(&herr, derr, 4, s));
        herr |= dup;
        if (herr) {
            *err = mesh_message(herr);
            return FL_E_MESH;
        }
    }
    // ---- ghost kinds (leaf kinds of every level are final)
    for (int l = 0; l < NL; ++l) {
        const LayoutLevel& L = out[l];
        if (!L.nb) continue;
        const bool hc = l > 0, hf = l + 1 < NL;
        hipLaunchKernelGGL(k_classify, dim3(blocks_for((size_t)L.nb * 64)), dim3(kB), 0, s, (size_t)L.nb, l, L.bcoord,
                           L.binst, dinst, stype[l], hc ? 1 : 0, hc ? level_map(out[l - 1]) : LevelMap{},
                           hc ? stype[l - 1] : nullptr, hf ? 1 : 0, hf ? level_map(out[l + 1]) : LevelMap{},
                           hf ? stype[l + 1] : nullptr);
        FL_CHECK_LAUNCH();
    }
    // ---- link words, force links, per-instance ranges
    for (int l = 0; l < NL; ++l) {
        LayoutLevel& L = out[l];
        const size_t nb = L.nb, slots = nb * 64;
        Tmp t(s);
        uint32_t *fmask, *bcnt, *boff;
        FL_TRY(palloc(&L.cflags, slots));
        FL_TRY(palloc(&L.bflags, nb));
        FL_TRY(t.get(&fmask, slots));
        FL_TRY(t.get(&bcnt, nb));
        FL_TRY(t.get(&boff, nb));
        if (nb) {
            hipLaunchKernelGGL(k_links, dim3((unsigned)((nb + kB / 64 - 1) / (kB / 64))), dim3(kB), 0, s, nb, l,
                               L.bcoord, L.binst, dinst, L.nbr, stype[l], level_map(L), L.cflags, needed[l], fmask,
                               L.bflags, bcnt, derr);
            FL_CHECK_LAUNCH();
        }
        uint64_t nlinks = 0;
        FL_TRY(scan_u32(bcnt, boff, L.nactive, &nlinks, s));
        L.n_links = nlinks;
        FL_TRY(palloc(&L.force_links, nlinks));
        if (nlinks) {
            hipLaunchKernelGGL(k_force_write, dim3((unsigned)((L.nactive + kB / 64 - 1) / (kB / 64))), dim3(kB), 0, s,
                               (size_t)L.nactive, fmask, boff, L.force_links);
            FL_CHECK_LAUNCH();
        }
        unsigned long long *acnt, *lcnt;
        FL_TRY(t.get(&acnt, NI));
        FL_TRY(t.get(&lcnt, NI));
        hipLaunchKernelGGL(k_fill64, dim3(blocks_for(NI)), dim3(kB), 0, s, reinterpret_cast<uint64_t*>(acnt),
                           (size_t)NI, (uint64_t)0);
        hipLaunchKernelGGL(k_fill64, dim3(blocks_for(NI)), dim3(kB), 0, s, reinterpret_cast<uint64_t*>(lcnt),
                           (size_t)NI, (uint64_t)0);
        if (L.nactive) {
            hipLaunchKernelGGL(k_inst_counts, dim3(kCountGrid), dim3(kB), 0, s, (size_t)L.nactive, L.binst,
                               bcnt, acnt, lcnt);
            FL_CHECK_LAUNCH();
        }
        std::vector<uint64_t> ha(NI), hl(NI);
        FL_TRY(sync_read(ha.data(), acnt, NI * 8, s));
        FL_TRY(sync_read(hl.data(), lcnt, NI * 8, s));
        std::vector<uint32_t> abeg(NI + 1, 0), lbeg(NI + 1, 0);
        for (int k = 0; k < NI; ++k) {
            abeg[k + 1] = abeg[k] + (uint32_t)ha[k];
            lbeg[k + 1] = lbeg[k] + (uint32_t)hl[k];
        }
        L.abeg_h = abeg;
        L.lbeg_h = lbeg;
        FL_TRY(palloc(&L.abeg, NI + 1));
        FL_TRY(palloc(&L.lbeg, NI + 1));
        FL_HIP(hipMemcpyAsync(L.abeg, abeg.data(), (NI + 1) * 4, hipMemcpyHostToDevice, s));
        FL_HIP(hipMemcpyAsync(L.lbeg, lbeg.data(), (NI + 1) * 4, hipMemcpyHostToDevice, s));
        FL_HIP(hipStreamSynchronize(s));
    }
    FL_TRY(sync_read(&herr, derr, 4, s));
    if (herr) {
        *err = mesh_message(herr);
        return FL_E_MESH;
    }
    // ---- ghost lists
    std::vector<uint32_t*> c2f(NL, nullptr), f2c(NL, nullptr);
    std::vector<size_t> nc2f(NL, 0), nf2c(NL, 0);
    auto gget = [&](uint32_t** p, size_t k) { return g.get(p, k); };
    for (int l = 0; l < NL; ++l) {
        const size_t slots = (size_t)out[l].nb * 64;
        if (!slots) continue;
        Tmp t(s);
        uint32_t* flag;
        FL_TRY(t.get(&flag, slots));
        hipLaunchKernelGGL(k_ghost_flags, dim3(blocks_for(slots)), dim3(kB), 0, s, stype[l], needed[l], slots,
                           (uint8_t)T_C2F, flag);
        FL_TRY(compact(flag, slots, &c2f[l], &nc2f[l], s, gget));
        hipLaunchKernelGGL(k_ghost_flags, dim3(blocks_for(slots)), dim3(kB), 0, s, stype[l], needed[l], slots,
                           (uint8_t)T_F2C, flag);
        FL_TRY(compact(flag, slots, &f2c[l], &nf2c[l], s, gget));
    }
    for (int l = 0; l < NL; ++l) out[l].n_c2f = nc2f[l];
    // c2f groups of level l and the capture list of level l - 1
    for (int l = 1; l < NL; ++l) {
        LayoutLevel& L = out[l];
        LayoutLevel& Cl = out[l - 1];
        const size_t ng = nc2f[l];
        if (!ng) continue;
        Tmp t(s);
        uint64_t* par;
        uint32_t *idx, *head, *excl;
        uint8_t* chq;
        FL_TRY(t.get(&par, ng));
        FL_TRY(t.get(&idx, ng));
        FL_TRY(t.get(&chq, ng));
        FL_TRY(t.get(&head, ng));
        FL_TRY(t.get(&excl, ng));
        hipLaunchKernelGGL(k_c2f_parent, dim3(blocks_for(ng)), dim3(kB), 0, s, c2f[l], ng, L.bcoord, L.binst, dinst, l,
                           level_map(Cl), par, idx, chq, derr);
        FL_CHECK_LAUNCH();
        FL_TRY(sync_read(&herr, derr, 4, s));
        if (herr) {
            *err = mesh_message(herr);
            return FL_E_MESH;
        }
        FL_TRY(sort_pairs(par, idx, ng, 32, s));
        hipLaunchKernelGGL(k_heads, dim3(blocks_for(ng)), dim3(kB), 0, s, par, ng, head);
        uint64_t G = 0;
        FL_TRY(scan_u32(head, excl, ng, &G, s));
        L.n_groups = G;
        uint32_t* gpar;
        FL_TRY(t.get(&gpar, G));
        FL_TRY(palloc(&L.c2f_gchild, 8 * G));
        FL_TRY(palloc(&L.c2f_gmask, 8 * G));
        hipLaunchKernelGGL(k_fill32, dim3(blocks_for(8 * G)), dim3(kB), 0, s, L.c2f_gchild, (size_t)(8 * G), kNoCap);
        hipLaunchKernelGGL(k_fill32, dim3(blocks_for(8 * G)), dim3(kB), 0, s, L.c2f_gmask, (size_t)(8 * G), 0u);
        hipLaunchKernelGGL(k_group_fill, dim3(blocks_for(ng)), dim3(kB), 0, s, par, idx, head, excl, ng, c2f[l], chq,
                           needed[l], L.c2f_gchild, L.c2f_gmask, gpar);
        FL_CHECK_LAUNCH();
        const size_t ns = G * kC2FCaps;
        uint32_t *src, *sflag, *sexcl, *dummy;
        uint64_t* skeys;
        FL_TRY(t.get(&src, ns));
        FL_TRY(t.get(&sflag, ns));
        FL_TRY(t.get(&sexcl, ns));
        hipLaunchKernelGGL(k_group_src, dim3(blocks_for(G)), dim3(kB), 0, s, gpar, (size_t)G, Cl.bcoord, Cl.binst,
                           dinst, l - 1, level_map(Cl), stype[l - 1], src);
        hipLaunchKernelGGL(k_ne_flag, dim3(blocks_for(ns)), dim3(kB), 0, s, src, ns, kNoCap, sflag);
        FL_CHECK_LAUNCH();
        uint64_t nv = 0;
        FL_TRY(scan_u32(sflag, sexcl, ns, &nv, s));
        FL_TRY(t.get(&skeys, nv));
        FL_TRY(t.get(&dummy, nv));
        hipLaunchKernelGGL(k_scatter_val, dim3(blocks_for(ns)), dim3(kB), 0, s, src, sexcl, ns, kNoCap, skeys);
        hipLaunchKernelGGL(k_fill32, dim3(blocks_for(nv)), dim3(kB), 0, s, dummy, (size_t)nv, 0u);
        FL_CHECK_LAUNCH();
        FL_TRY(sort_pairs(skeys, dummy, nv, 32, s));
        uint64_t* ukeys;
        uint32_t* uvals;
        size_t ncap = 0;
        FL_TRY(unique_or(t, skeys, dummy, nv, &ukeys, &uvals, &ncap, s));
        Cl.n_cap = ncap;
        FL_TRY(palloc(&Cl.cap_slots, ncap));
        if (ncap) hipLaunchKernelGGL(k_u64_to_u32, dim3(blocks_for(ncap)), dim3(kB), 0, s, ukeys, ncap, Cl.cap_slots);
        FL_TRY(palloc(&L.c2f_gcap, ns));
        hipLaunchKernelGGL(k_lower_bound, dim3(blocks_for(ns)), dim3(kB), 0, s, src, ns, Cl.cap_slots, ncap, L.c2f_gcap);
        FL_CHECK_LAUNCH();
        FL_HIP(hipStreamSynchronize(s));
    }
    // f2c ghosts of level l with their children on level l + 1
    for (int l = 0; l + 1 < NL; ++l) {
        LayoutLevel& L = out[l];
        const size_t ng = nf2c[l];
        L.n_f2c = ng;
        if (!ng) continue;
        Tmp t(s);
        uint32_t *ch8, *idx;
        uint64_t* key;
        FL_TRY(t.get(&ch8, 8 * ng));
        FL_TRY(t.get(&idx, ng));
        FL_TRY(t.get(&key, ng));
        hipLaunchKernelGGL(k_f2c_children, dim3(blocks_for(ng)), dim3(kB), 0, s, f2c[l], ng, L.bcoord, L.binst, dinst,
                           l, level_map(out[l + 1]), ch8, key, idx);
        FL_CHECK_LAUNCH();
        FL_TRY(sort_pairs(key, idx, ng, 32, s));
        FL_TRY(palloc(&L.f2c_ghost, ng));
        FL_TRY(palloc(&L.f2c_child, 8 * ng));
        hipLaunchKernelGGL(k_f2c_perm, dim3(blocks_for(ng)), dim3(kB), 0, s, idx, ng, f2c[l], ch8, L.f2c_ghost,
                           L.f2c_child);
        FL_CHECK_LAUNCH();
        FL_HIP(hipStreamSynchronize(s));
    }
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
}

fl_status launch_query_slots(const LevelMaps& maps, const LayoutInst* d_inst, int inst, const int32_t* x,
                             const int32_t* y, const int32_t* z, const uint8_t* lv, size_t n, int64_t* out,
                             hipStream_t s) {
    if (!n) return FL_OK;
    hipLaunchKernelGGL(k_query, dim3(blocks_for(n)), dim3(kB), 0, s, maps, d_inst, inst, x, y, z, lv, n, out);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status build_remap_lists(const LevelMaps& om, const LayoutInst* d_inst, int l, const LayoutLevel& nl, hipStream_t s,
                            const DevAlloc& alloc, RemapLists& R) {
    R = RemapLists();
    const size_t slots = (size_t)nl.nactive * 64;
    if (!slots) return FL_OK;
    Tmp t(s);
    uint8_t* cls;
    uint32_t *src, *flag;
    FL_TRY(t.get(&cls, slots));
    FL_TRY(t.get(&src, slots));
    FL_TRY(t.get(&flag, slots));
    hipLaunchKernelGGL(k_remap_class, dim3(blocks_for(slots)), dim3(kB), 0, s, om, d_inst, l, (size_t)nl.nactive,
                       nl.bcoord, nl.binst, nl.cflags, cls, src);
    FL_CHECK_LAUNCH();
    auto aget = [&](uint32_t** p, size_t k) { return pget(alloc, p, k); };
    size_t m = 0;
    const uint8_t kinds[4] = {1, 2, 3, 4};
    uint32_t** dst[4] = {&R.copy_dst, &R.split_dst, &R.merge_dst, &R.rest_dst};
    uint32_t* cnt[4] = {&R.ncopy, &R.nsplit, &R.nmerge, &R.nrest};
    for (int c = 0; c < 4; ++c) {
        hipLaunchKernelGGL(k_class_flag, dim3(blocks_for(slots)), dim3(kB), 0, s, cls, slots, kinds[c], flag);
        FL_TRY(compact(flag, slots, dst[c], &m, s, aget));
        *cnt[c] = (uint32_t)m;
    }
    FL_TRY(pget(alloc, &R.copy_src, R.ncopy));
    FL_TRY(pget(alloc, &R.split_src, R.nsplit));
    FL_TRY(pget(alloc, &R.merge_src8, 8 * (size_t)R.nmerge));
    if (R.ncopy) hipLaunchKernelGGL(k_gather_u32, dim3(blocks_for(R.ncopy)), dim3(kB), 0, s, R.copy_dst, (size_t)R.ncopy, src, R.copy_src);
    if (R.nsplit)
        hipLaunchKernelGGL(k_gather_u32, dim3(blocks_for(R.nsplit)), dim3(kB), 0, s, R.split_dst, (size_t)R.nsplit, src,
                           R.split_src);
    if (R.nmerge)
        hipLaunchKernelGGL(k_merge_src, dim3(blocks_for(R.nmerge)), dim3(kB), 0, s, om, d_inst, l, R.merge_dst,
                           (size_t)R.nmerge, nl.bcoord, nl.binst, R.merge_src8);
    FL_CHECK_LAUNCH();
    FL_HIP(hipStreamSynchronize(s));
    return FL_OK;
}

}  // namespace lbm
}  // namespace fl
