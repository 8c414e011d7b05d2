// a6, a9-a13: adaptation kernels -- face neighbours, 2:1 balance, split,
// merge, compaction/append and the fused device regrid pass.
//
// Every kernel works on SoA cell arrays in native per-level coordinates
// (child = (p<<1)|bit, parent = c>>1: split_cells.cl:90-92,
// merge_cells.cl:43-45).  Cell lookups go through a GPU hash keyed by the
// unique cell key (L << 57) | (x << 38) | (y << 19) | z whose value is the smallest
// index holding that cell, which makes every query result deterministic.
// Orderings that the reference leaves to atomic_inc (merge group ids,
// parent order) are fixed by scans: groups follow leader index order,
// children follow parent order, compaction is stable.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fl_internal.hpp"
#include "hash_device.hpp"
#include "hilbert_tables.hpp"
#include "scan_device.hpp"

namespace fl {
fl_status hash_clear_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, hipStream_t s);
fl_status cellkey_map_build_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, const uint64_t* keys, size_t n,
                                 hipStream_t s);
fl_status hash_build_impl(uint64_t* tk, uint32_t* tv, uint64_t cap, const uint64_t* keys,
                          const uint32_t* vals, size_t n, uint32_t* probes, hipStream_t s);

namespace {

constexpr int kBlock = 256;

struct CellsC {
    const int32_t* x;
    const int32_t* y;
    const int32_t* z;
    const uint8_t* level;
    const uint8_t* state;
    const int32_t* flag;
    const uint32_t* material;
};

struct CellsW {
    int32_t* x;
    int32_t* y;
    int32_t* z;
    uint8_t* level;
    uint8_t* state;
    int32_t* flag;
    uint32_t* material;
};

CellsC cview(const fl_cells* c) {
    return CellsC{c->x, c->y, c->z, c->level, c->state, c->flag, c->material};
}
CellsW wview(const fl_cells* c) {
    return CellsW{c->x, c->y, c->z, c->level, c->state, c->flag, c->material};
}

struct CellHash {
    const uint64_t* tk;
    const uint32_t* tv;
    uint64_t cap;
};

__device__ inline uint32_t find_cell(const CellHash& h, int l, int32_t x, int32_t y, int32_t z) {
    if (l < 0 || l > 19 || x < 0 || y < 0 || z < 0 || x >= kKeyCoordLimit || y >= kKeyCoordLimit ||
        z >= kKeyCoordLimit)
        return kInvalid;
    return cellkey_lookup(h.tk, h.tv, h.cap, cell_hkey((uint32_t)x, (uint32_t)y, (uint32_t)z, l));
}

// The key find_cell looks up, or EMPTY (absent) outside the key's range.
__device__ inline uint64_t cell_key_or_empty(int l, int32_t x, int32_t y, int32_t z) {
    if (l < 0 || l > 19 || x < 0 || y < 0 || z < 0 || x >= kKeyCoordLimit || y >= kKeyCoordLimit ||
        z >= kKeyCoordLimit)
        return kEmptyKey;
    return cell_hkey((uint32_t)x, (uint32_t)y, (uint32_t)z, l);
}

struct Root {
    int32_t n[3];
};

__device__ inline bool in_domain(int l, const Root& r, int32_t x, int32_t y, int32_t z) {
    return x >= 0 && y >= 0 && z >= 0 && (int64_t)x < ((int64_t)r.n[0] << l) && (int64_t)y < ((int64_t)r.n[1] << l) &&
           (int64_t)z < ((int64_t)r.n[2] << l);
}

__device__ __constant__ static const int8_t kD18[18][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},   {0, 0, -1},
    {1, 1, 0},  {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1},
    {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},  {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};
__device__ __constant__ static const int8_t kF6[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0},
                                                         {0, 1, 0},  {0, 0, -1}, {0, 0, 1}};

// ---------------------------------------------------------------- keys/hash
__global__ __launch_bounds__(kBlock) void k_hkeys(CellsC c, size_t n, uint64_t* __restrict__ keys) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int32_t a = c.x[i], b = c.y[i], d = c.z[i];
        keys[i] = (c.level[i] <= 19 && a >= 0 && b >= 0 && d >= 0 && a < kKeyCoordLimit && b < kKeyCoordLimit &&
                   d < kKeyCoordLimit)
                      ? cell_hkey((uint32_t)a, (uint32_t)b, (uint32_t)d, c.level[i])
                      : kEmptyKey;
    }
}

unsigned sgrid(size_t n) {
    size_t g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    return (unsigned)(g ? g : 1);
}

fl_status build_cell_hash(CellsC c, size_t n, uint64_t* keys, uint64_t* tk, uint32_t* tv, uint64_t cap,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_hkeys, dim3(sgrid(n)), dim3(kBlock), 0, s, c, n, keys);
    FL_CHECK_LAUNCH();
    fl_status st = hash_clear_impl(tk, tv, cap, s);
    if (st) return st;
    return cellkey_map_build_impl(tk, tv, cap, keys, n, s);
}

// ---------------------------------------------------------------- a9
// Per face: the first cell holding the face point at levels L, L-1, .., 0;
// else kFiner when a level L+1 cell touches it; else kInvalid.  The six
// faces are searched in lockstep, level by level (cellkey_lookup_n), so a
// cell's lookups overlap; each face's result is the same first hit.
__global__ __launch_bounds__(kBlock) void k_face_nbr(CellsC c, size_t n, const Root root, CellHash h,
                                                     uint32_t* __restrict__ nbr) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int L = c.level[i];
        const int32_t x = c.x[i], y = c.y[i], z = c.z[i];
        int32_t px[6], py[6], pz[6];
        uint32_t r[6];
        unsigned pend = 0;
#pragma unroll
        for (int f = 0; f < 6; ++f) {
            px[f] = x + kF6[f][0];
            py[f] = y + kF6[f][1];
            pz[f] = z + kF6[f][2];
            r[f] = kInvalid;
            if (in_domain(L, root, px[f], py[f], pz[f])) pend |= 1u << f;
        }
        for (int l = L; l >= 0 && pend; --l) {
            const int sh = L - l;
            uint64_t key[6];
            uint32_t got[6];
#pragma unroll
            for (int f = 0; f < 6; ++f)
                key[f] = ((pend >> f) & 1u) ? cell_key_or_empty(l, px[f] >> sh, py[f] >> sh, pz[f] >> sh) : kEmptyKey;
            cellkey_lookup_n<6>(h.tk, h.tv, h.cap, key, got);
#pragma unroll
            for (int f = 0; f < 6; ++f)
                if (((pend >> f) & 1u) && got[f] != kInvalid) {
                    r[f] = got[f];
                    pend &= ~(1u << f);
                }
        }
        for (int f = 0; f < 6; ++f) {
            if (!((pend >> f) & 1u)) continue;
            uint64_t key[8];
            uint32_t got[8];
#pragma unroll
            for (int q = 0; q < 8; ++q)
                key[q] = cell_key_or_empty(L + 1, 2 * px[f] + (q & 1), 2 * py[f] + ((q >> 1) & 1),
                                           2 * pz[f] + ((q >> 2) & 1));
            cellkey_lookup_n<8, false>(h.tk, h.tv, h.cap, key, got);
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if (got[q] != kInvalid) r[f] = kFiner;
        }
#pragma unroll
        for (int f = 0; f < 6; ++f) nbr[6 * i + f] = r[f];
    }
}

// ---------------------------------------------------------------- a12
// shadow = level + (flag > 0) (update_shadow_levels, BalanceEnforcer.cpp:53-62);
// with_flags = 0 gives the plain copy of the stored levels that the reference
// loop starts from (BalanceEnforcer.cpp:165-167).
__global__ __launch_bounds__(kBlock) void k_shadow(CellsC c, size_t n, int with_flags,
                                                   uint8_t* __restrict__ shadow) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        shadow[i] = (uint8_t)(c.level[i] + (with_flags && c.flag[i] > 0 ? 1 : 0));
}

// mode-0 key: anchor (finest units of level M <= 19) with the holder's shadow level
__global__ __launch_bounds__(kBlock) void k_anchor_keys(CellsC c, size_t n, int M,
                                                        const uint8_t* __restrict__ shadow,
                                                        uint64_t* __restrict__ keys) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        int L = c.level[i];
        if (L > M) {
            keys[i] = kEmptyKey;
            continue;
        }
        uint32_t ax = (uint32_t)c.x[i] << (M - L), ay = (uint32_t)c.y[i] << (M - L),
                 az = (uint32_t)c.z[i] << (M - L);
        keys[i] = cell_hkey(ax, ay, az, shadow[i]);
    }
}

// balance_enforce.cl:57-147 with the shadow levels of BalanceEnforcer.cpp:182-187:
// six test points (anchor-1 / anchor+size per axis, size from the shadow
// level), each located by probing levels 0..M; violation when the first
// holder found is more than one level finer than this cell's shadow level.
__global__ __launch_bounds__(kBlock) void k_detect0(CellsC c, size_t n, int M,
                                                    const uint8_t* __restrict__ shadow, CellHash h,
                                                    uint8_t* __restrict__ viol, uint32_t* __restrict__ count) {
    uint32_t local = 0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        uint8_t v = 0;
        if (c.state[i] == 0 && c.level[i] <= M) {
            const int s = shadow[i];
            const int L = c.level[i];
            const int32_t ax = c.x[i] << (M - L), ay = c.y[i] << (M - L), az = c.z[i] << (M - L);
            const int32_t size = s <= M ? (1 << (M - s)) : 0;
            // per face point: the first level l = 0..M holding it decides; the
            // six points are probed in lockstep (cellkey_lookup_n), which gives
            // the same first hits, and the cell violates if any face does
            int32_t px[6], py[6], pz[6];
#pragma unroll
            for (int t = 0; t < 6; ++t) {
                px[t] = ax;
                py[t] = ay;
                pz[t] = az;
                const int32_t d = (t & 1) ? size : -1;
                if ((t >> 1) == 0) px[t] += d; else if ((t >> 1) == 1) py[t] += d; else pz[t] += d;
            }
            unsigned pend = 0x3Fu;
            for (int l = 0; l <= M && pend && !v; ++l) {
                const int32_t msk = ~((1 << (M - l)) - 1);
                uint64_t key[6];
                uint32_t got[6];
#pragma unroll
                for (int t = 0; t < 6; ++t) {
                    const int32_t qx = px[t] & msk, qy = py[t] & msk, qz = pz[t] & msk;
                    const bool in = qx >= 0 && qy >= 0 && qz >= 0 && qx < (1 << M) && qy < (1 << M) && qz < (1 << M);
                    key[t] = ((pend >> t) & 1u) && in ? cell_hkey((uint32_t)qx, (uint32_t)qy, (uint32_t)qz, l)
                                                      : kEmptyKey;
                }
                cellkey_lookup_n<6, false>(h.tk, h.tv, h.cap, key, got);
#pragma unroll
                for (int t = 0; t < 6; ++t)
                    if (((pend >> t) & 1u) && got[t] != kInvalid) {
                        if (l > s + 1) v = 1;
                        pend &= ~(1u << t);
                    }
            }
        }
        viol[i] = v;
        local += v;
    }
    if (local) atomicAdd(count, local);
}

// Leaves overlapping the level-L cube p: is any at shadow > s + 1?
__device__ bool cube_violates(const CellHash& h, const uint8_t* __restrict__ shadow, int L, int32_t px,
                              int32_t py, int32_t pz, int s) {
    for (int l = L; l >= 0; --l) {
        int sh = L - l;
        if (find_cell(h, l, px >> sh, py >> sh, pz >> sh) != kInvalid) return false;
    }
    for (int q = 0; q < 8; ++q) {
        int32_t cx = 2 * px + (q & 1), cy = 2 * py + ((q >> 1) & 1), cz = 2 * pz + ((q >> 2) & 1);
        uint32_t j = find_cell(h, L + 1, cx, cy, cz);
        if (j != kInvalid) {
            if (shadow[j] > s + 1) return true;
            continue;
        }
        if (s <= L) return true;
        for (int r = 0; r < 8; ++r) {
            uint32_t g = find_cell(h, L + 2, 2 * cx + (r & 1), 2 * cy + ((r >> 1) & 1), 2 * cz + ((r >> 2) & 1));
            if (g == kInvalid) return true;
            if (shadow[g] > s + 1) return true;
        }
    }
    return false;
}

// The cell violates if any of its 18 in-domain neighbour cubes does.  The
// common case, a cube that is itself a leaf at level L (no violation), is
// probed first for all directions, six lookups in lockstep at a time; only
// the directions it leaves open take the full test, whose first probe is
// that same lookup, so the answer is unchanged.
__device__ bool detect_mode1(const CellHash& h, const uint8_t* __restrict__ shadow, int L, int32_t x,
                             int32_t y, int32_t z, int s, const Root root) {
    uint32_t open = 0;
    for (int d0 = 0; d0 < 18; d0 += 6) {
        uint64_t key[6];
        uint32_t got[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int d = d0 + k;
            const int32_t px = x + kD18[d][0], py = y + kD18[d][1], pz = z + kD18[d][2];
            key[k] = kEmptyKey;
            if (in_domain(L, root, px, py, pz)) {
                open |= 1u << d;
                key[k] = cell_key_or_empty(L, px, py, pz);
            }
        }
        cellkey_lookup_n<6, false>(h.tk, h.tv, h.cap, key, got);
#pragma unroll
        for (int k = 0; k < 6; ++k)
            if (got[k] != kInvalid) open &= ~(1u << (d0 + k));
    }
    for (int d = 0; d < 18; ++d) {
        if (!((open >> d) & 1u)) continue;
        const int32_t px = x + kD18[d][0], py = y + kD18[d][1], pz = z + kD18[d][2];
        if (cube_violates(h, shadow, L, px, py, pz, s)) return true;
    }
    return false;
}

__global__ __launch_bounds__(kBlock) void k_detect1(CellsC c, size_t n, const Root root,
                                                    const uint8_t* __restrict__ shadow, CellHash h,
                                                    uint8_t* __restrict__ viol, uint32_t* __restrict__ count) {
    uint32_t local = 0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        uint8_t v = detect_mode1(h, shadow, c.level[i], c.x[i], c.y[i], c.z[i], shadow[i], root) ? 1 : 0;
        viol[i] = v;
        local += v;
    }
    if (local) atomicAdd(count, local);
}

// mark_cascading_refinement (balance_enforce.cl:150-172) + update_shadow_levels
// (BalanceEnforcer.cpp:53-62)
__global__ __launch_bounds__(kBlock) void k_mark(CellsW c, size_t n, int max_split_level,
                                                 const uint8_t* __restrict__ viol, uint8_t* __restrict__ shadow,
                                                 uint32_t* __restrict__ count) {
    uint32_t local = 0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        if (viol[i] && c.state[i] == 0 && c.level[i] < max_split_level && c.flag[i] <= 0) {
            c.flag[i] = 1;
            ++local;
        }
    }
    if (local) atomicAdd(count, local);
}

// ---------------------------------------------------------------- a11 merge
// Leader test of merge_cells.cl:22-95 against the smallest-index holders;
// with check_balance the group is also dropped when the parent (level L-1)
// would have a D3Q19 neighbour cube holding cells of shadow level > L.
__device__ bool merge_leader(CellsC c, size_t i, const CellHash& h, const uint8_t* __restrict__ shadow,
                             int check_balance, const Root root, uint32_t sib[8]) {
    if (c.flag[i] != -1 || c.state[i] != 0) return false;
    const int32_t x = c.x[i], y = c.y[i], z = c.z[i];
    const int L = c.level[i];
    if ((x & 1) || (y & 1) || (z & 1) || L == 0) return false;
    if (find_cell(h, L, x, y, z) != (uint32_t)i) return false;
    for (int q = 0; q < 8; ++q) {
        uint32_t j = find_cell(h, L, x | (q & 1), y | ((q >> 1) & 1), z | ((q >> 2) & 1));
        if (j == kInvalid || c.flag[j] != -1 || c.state[j] != 0) return false;
        sib[q] = j;
    }
    if (check_balance && detect_mode1(h, shadow, L - 1, x >> 1, y >> 1, z >> 1, L - 1, root))
        return false;
    return true;
}

// Fused: leader predicate -> scan (group id) -> group ids on siblings,
// leader list and parent cells (create_parent_cells, merge_cells.cl:162-199).
__global__ __launch_bounds__(kScanThreads) void k_merge(CellsC c, size_t n, CellHash h,
                                                        const uint8_t* __restrict__ shadow, int check_balance,
                                                        const Root root, uint32_t* __restrict__ group_id,
                                                        uint32_t* __restrict__ leaders, CellsW par,
                                                        size_t par_off_unused, uint64_t* __restrict__ ws,
                                                        uint64_t* __restrict__ ngroups) {
    __shared__ ScanShared sh;
    const uint32_t tile = scan_tile_id(sh, ws);
    const size_t base = (size_t)tile * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t lmask = 0;
    for (int j = 0; j < kScanItems; ++j) {
        size_t i = base + j;
        uint32_t sib[8];
        if (i < n && merge_leader(c, i, h, shadow, check_balance, root, sib)) lmask |= 1u << j;
    }
    uint64_t g = scan_tile_prefix(sh, ws, tile, (uint64_t)__popc(lmask), ngroups, n);
    for (int j = 0; j < kScanItems; ++j) {
        if (!(lmask & (1u << j))) continue;
        size_t i = base + j;
        uint32_t sib[8];
        merge_leader(c, i, h, shadow, check_balance, root, sib);
        for (int q = 0; q < 8; ++q) group_id[sib[q]] = (uint32_t)g;
        leaders[g] = (uint32_t)i;
        ++g;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) p[i] = v;
}

__global__ void k_zero64(uint64_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Parents and merged fields at out[off + g], off read from the device.
__global__ __launch_bounds__(kBlock) void k_parents(CellsC c, const uint32_t* __restrict__ leaders,
                                                    const uint64_t* __restrict__ ngroups,
                                                    const uint64_t* __restrict__ off0,
                                                    const uint64_t* __restrict__ off1, int off1_scale,
                                                    CellsW out, size_t cap, CellHash h,
                                                    const float* __restrict__ fin, float* __restrict__ fout,
                                                    int ncomp, int rule) {
    const uint64_t G = *ngroups;
    const uint64_t off = (off0 ? *off0 : 0) + (off1 ? *off1 * (uint64_t)off1_scale : 0);
    for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; g < G; g += (uint64_t)gridDim.x * kBlock) {
        const uint32_t i = leaders[g];
        const uint64_t o = off + g;
        if (o >= cap) continue;
        out.x[o] = c.x[i] >> 1;
        out.y[o] = c.y[i] >> 1;
        out.z[o] = c.z[i] >> 1;
        out.level[o] = (uint8_t)(c.level[i] - 1);
        out.state[o] = c.state[i];
        out.material[o] = c.material[i];
        out.flag[o] = 0;
        if (ncomp > 0) {
            uint32_t sib[8];
            const int L = c.level[i];
            for (int q = 0; q < 8; ++q)
                sib[q] = find_cell(h, L, c.x[i] | (q & 1), c.y[i] | ((q >> 1) & 1), c.z[i] | ((q >> 2) & 1));
            for (int k = 0; k < ncomp; ++k) {
                float s = 0.0f;
                for (int q = 0; q < 8; ++q) {
                    float v = fin[(size_t)sib[q] * ncomp + k];
                    s += rule == 0 ? v * 0.125f : v;
                }
                fout[o * ncomp + k] = s;
            }
        }
    }
}

// ---------------------------------------------------------------- a6 compact
// mark_valid_cells + compact_cells (compact_cells.cl:9-172) in one pass.
__global__ __launch_bounds__(kScanThreads) void k_compact(CellsC c, size_t n, const uint32_t* __restrict__ group_id,
                                                          CellsW out, size_t cap, const float* __restrict__ fin,
                                                          float* __restrict__ fout, int ncomp,
                                                          uint64_t* __restrict__ ws, uint64_t* __restrict__ nkeep) {
    // striped tile: element j of this thread is t0 + j * kScanThreads + tid,
    // so each round's loads and the kept elements' stores are coalesced
    __shared__ StripedShared sh;
    const uint32_t tile = scan_tile_id(sh.sc, ws);
    const size_t t0 = (size_t)tile * kScanTile + threadIdx.x;
    uint32_t kmask = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = t0 + (size_t)j * kScanThreads;
        if (i < n && !(c.flag[i] > 0) && (group_id ? group_id[i] == kInvalid : true)) kmask |= 1u << j;
    }
    uint32_t rel[kScanItems];
    const uint64_t base = striped_positions(sh, ws, tile, kmask, nkeep, n, rel);
    // positions recomputed round by round (ballot rank + per-wave counts), so
    // the rounds run as a loop rather than unrolled with all loads in flight
    const int w = threadIdx.x >> 6;
    const uint64_t below = lanemask_lt();
    uint64_t run = base;
#pragma unroll 1
    for (int j = 0; j < kScanItems; ++j) {
        const bool mine = (kmask >> j) & 1u;
        const uint64_t bal = __ballot(mine);
        uint32_t cnt = 0, before = 0;
#pragma unroll
        for (int v = 0; v < kScanThreads / 64; ++v) {
            const uint32_t cv = sh.cnt[j][v];
            if (v < w) before += cv;
            cnt += cv;
        }
        const uint64_t o = run + before + (uint32_t)__popcll(bal & below);
        run += cnt;
        if (!mine) continue;
        const size_t i = t0 + (size_t)j * kScanThreads;
        if (o < cap) {
            out.x[o] = c.x[i]; out.y[o] = c.y[i]; out.z[o] = c.z[i];
            out.level[o] = c.level[i]; out.state[o] = c.state[i];
            out.material[o] = c.material[i]; out.flag[o] = 0;
            for (int k = 0; k < ncomp; ++k) fout[o * ncomp + k] = fin[i * ncomp + k];
        }
    }
}

// ---------------------------------------------------------------- a10 split
// split_count_and_allocate + split_generate_children + interpolate_split_fields
// (split_cells.cl:15-132) fused: predicate -> scan -> 8 children at
// out_off + 8*rank (child q = ((p<<1)|(q&1), (p<<1)|(q>>1&1), (p<<1)|(q>>2&1))).
__global__ __launch_bounds__(kScanThreads) void k_split(CellsC c, size_t n, int max_level,
                                                        uint32_t* __restrict__ child_start, CellsW out,
                                                        const uint64_t* __restrict__ out_off, size_t cap,
                                                        const float* __restrict__ fin, float* __restrict__ fout,
                                                        int ncomp, uint64_t* __restrict__ ws,
                                                        uint64_t* __restrict__ nsplit) {
    // striped tile (element j of this thread is t0 + j * kScanThreads + tid):
    // coalesced predicate loads; split parents of one round write their
    // children as consecutive 8-cell blocks
    __shared__ StripedShared sh;
    const uint32_t tile = scan_tile_id(sh.sc, ws);
    const size_t t0 = (size_t)tile * kScanTile + threadIdx.x;
    uint32_t smask = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = t0 + (size_t)j * kScanThreads;
        if (i < n && c.flag[i] > 0 && c.state[i] == 0 && c.level[i] < max_level) smask |= 1u << j;
    }
    uint32_t rel[kScanItems];
    const uint64_t rbase = striped_positions(sh, ws, tile, smask, nsplit, n, rel);
    const uint64_t off = out_off ? *out_off : 0;
    // round by round: the round's split parents go to LDS in rank order, then
    // the whole block writes their 8 * count children with consecutive lanes
    // on consecutive cells
    __shared__ int32_t px[kScanThreads], py[kScanThreads], pz[kScanThreads];
    __shared__ uint32_t pmat[kScanThreads], pidx_[kScanThreads];
    __shared__ uint8_t plev[kScanThreads], pst[kScanThreads];
    const int w = threadIdx.x >> 6;
    const uint64_t below = lanemask_lt();
    uint32_t run = 0;  // kept elements of the earlier rounds
#pragma unroll 1
    for (int j = 0; j < kScanItems; ++j) {
        const size_t i = t0 + (size_t)j * kScanThreads;
        const bool mine = (smask >> j) & 1u;
        const uint64_t bal = __ballot(mine);
        uint32_t cnt = 0, before = 0;
#pragma unroll
        for (int v = 0; v < kScanThreads / 64; ++v) {
            const uint32_t cv = sh.cnt[j][v];
            if (v < w) before += cv;
            cnt += cv;
        }
        if (i < n) {
            if (mine) {
                const uint32_t k = before + (uint32_t)__popcll(bal & below);
                px[k] = c.x[i];
                py[k] = c.y[i];
                pz[k] = c.z[i];
                plev[k] = c.level[i];
                pst[k] = c.state[i];
                pmat[k] = c.material[i];
                pidx_[k] = (uint32_t)i;
                if (child_start) child_start[i] = (uint32_t)(8 * (rbase + run + k));
            } else if (child_start) {
                child_start[i] = kInvalid;
            }
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < 8 * cnt; k += kScanThreads) {
            const uint32_t p = k >> 3;
            const int q = (int)(k & 7u);
            const uint64_t o = off + 8 * (rbase + run) + k;
            if (o >= cap) continue;
            out.x[o] = (px[p] << 1) | (q & 1);
            out.y[o] = (py[p] << 1) | ((q >> 1) & 1);
            out.z[o] = (pz[p] << 1) | ((q >> 2) & 1);
            out.level[o] = (uint8_t)(plev[p] + 1);
            out.state[o] = pst[p];
            out.material[o] = pmat[p];
            out.flag[o] = 0;
            for (int m = 0; m < ncomp; ++m) fout[o * ncomp + m] = fin[(size_t)pidx_[p] * ncomp + m];
        }
        run += cnt;
        __syncthreads();
    }
}

// ---------------------------------------------------------------- criteria
// Distance (level-0 units) from c to the cube of cell (x,y,z,L).
__device__ inline double cube_dist2(int32_t x, int32_t y, int32_t z, int L, double cx, double cy, double cz) {
    const double h = 1.0 / (double)(1 << L);
    const double lo[3] = {x * h, y * h, z * h};
    const double c[3] = {cx, cy, cz};
    double d2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        double d = c[a] < lo[a] ? lo[a] - c[a] : (c[a] > lo[a] + h ? c[a] - lo[a] - h : 0.0);
        d2 += d * d;
    }
    return d2;
}

__global__ __launch_bounds__(kBlock) void k_mark_ball(CellsW c, size_t n, double cx, double cy, double cz,
                                                      double r, int target, uint32_t* __restrict__ count) {
    uint32_t local = 0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        if (c.state[i] == 0 && c.level[i] < target && cube_dist2(c.x[i], c.y[i], c.z[i], c.level[i], cx, cy, cz) <= r * r) {
            c.flag[i] = 1;
            ++local;
        }
    }
    if (local) atomicAdd(count, local);
}

// state = 1 (SOLID) where the cell centre lies inside the ball.
__global__ __launch_bounds__(kBlock) void k_mark_solid(CellsW c, size_t n, double cx, double cy, double cz, double r,
                                                       uint32_t* __restrict__ count) {
    uint32_t local = 0;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const double h = 1.0 / (double)(1 << c.level[i]);
        const double dx = (c.x[i] + 0.5) * h - cx, dy = (c.y[i] + 0.5) * h - cy, dz = (c.z[i] + 0.5) * h - cz;
        if (dx * dx + dy * dy + dz * dz <= r * r) {
            c.state[i] = 1;
            ++local;
        }
    }
    if (local) atomicAdd(count, local);
}

// ---------------------------------------------------------------- synthetic inputs
// splitmix64 (Steele, Lea & Flood 2014 finaliser constants)
__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// level ~ U{0..max_level}, coords ~ U[0, 2^level): four independent
// splitmix64 streams keyed by (seed, index).
__global__ __launch_bounds__(kBlock) void k_gen_cells(uint64_t seed, size_t n, int max_level, CellsW c) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const uint64_t b = seed * 0x100000001B3ull + (uint64_t)i * 4;
        const uint64_t r0 = splitmix64(b), r1 = splitmix64(b + 1), r2 = splitmix64(b + 2), r3 = splitmix64(b + 3);
        const int L = (int)((r0 >> 32) % (uint64_t)(max_level + 1));
        const uint32_t m = (1u << L) - 1u;
        c.level[i] = (uint8_t)L;
        c.x[i] = (int32_t)((uint32_t)(r1 >> 32) & m);
        c.y[i] = (int32_t)((uint32_t)(r2 >> 32) & m);
        c.z[i] = (int32_t)((uint32_t)(r3 >> 32) & m);
        c.state[i] = 0;
        c.flag[i] = 0;
        c.material[i] = 0;
    }
}

// split flag: Bernoulli(p_split) from stream (seed ^ 0x5A17, i)
__global__ __launch_bounds__(kBlock) void k_flag_split(uint64_t seed, size_t n, uint64_t thresh, CellsW c) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        c.flag[i] = splitmix64((seed ^ 0x5A17ull) * 0x100000001B3ull + i) < thresh ? 1 : 0;
}

// merge flags: a leader (smallest-index holder of an all-even position,
// level > 0) whose 8 siblings all exist is chosen with Bernoulli(p_merge) from
// stream (seed ^ 0x3E26, leader index); its siblings still at flag 0 get -1.
__global__ __launch_bounds__(kBlock) void k_flag_merge(uint64_t seed, size_t n, uint64_t thresh, CellHash h, CellsW c) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const int32_t x = c.x[i], y = c.y[i], z = c.z[i];
        const int L = c.level[i];
        if (L == 0 || (x & 1) || (y & 1) || (z & 1)) continue;
        if (find_cell(h, L, x, y, z) != (uint32_t)i) continue;
        if (splitmix64((seed ^ 0x3E26ull) * 0x100000001B3ull + i) >= thresh) continue;
        uint32_t sib[8];
        bool ok = true;
        for (int q = 0; q < 8 && ok; ++q) {
            sib[q] = find_cell(h, L, x | (q & 1), y | ((q >> 1) & 1), z | ((q >> 2) & 1));
            ok = sib[q] != kInvalid;
        }
        if (!ok) continue;
        for (int q = 0; q < 8; ++q)
            if (c.flag[sib[q]] == 0) c.flag[sib[q]] = -1;
    }
}

// ---------------------------------------------------------------- workspace
struct RegridLayout {
    size_t keys, tk, tv, shadow, viol, gid, leaders, counters, ws0, ws1, ws2, total;
    uint64_t cap;
};

inline size_t al(size_t v) { return (v + 255) & ~size_t(255); }

RegridLayout regrid_layout(size_t n) {
    RegridLayout L;
    // twice the public table's capacity (load <= 0.3): the map's 2x2x2 block
    // groups are displaced as units, and at half the load a wave's lookups
    // finish in about two group rounds instead of four
    L.cap = std::min<uint64_t>(2 * fl_hash_capacity(n), 1ull << 31);
    size_t o = 0;
    L.keys = o; o = al(o + n * 8);
    L.tk = o; o = al(o + L.cap * 8);
    L.tv = o; o = al(o + L.cap * 4);
    L.shadow = o; o = al(o + n);
    L.viol = o; o = al(o + n);
    L.gid = o; o = al(o + n * 4);
    L.leaders = o; o = al(o + n * 4);
    L.counters = o; o = al(o + 16 * 8);
    size_t w = scan_ws_words(n) * 8;
    L.ws0 = o; o = al(o + w);
    L.ws1 = o; o = al(o + w);
    L.ws2 = o; o = al(o + w);
    L.total = o;
    return L;
}

struct Ctx {
    char* base;
    RegridLayout L;
    hipStream_t s;
    template <typename T>
    T* at(size_t off) const { return reinterpret_cast<T*>(base + off); }
};

fl_status run_balance(const fl_regrid_config* cfg, fl_cells* cells, size_t n, const Ctx& x, fl_regrid_result* res) {
    CellsC c = cview(cells);
    CellsW cw = wview(cells);
    uint8_t* shadow = x.at<uint8_t>(x.L.shadow);
    uint8_t* viol = x.at<uint8_t>(x.L.viol);
    uint64_t* keys = x.at<uint64_t>(x.L.keys);
    uint64_t* tk = x.at<uint64_t>(x.L.tk);
    uint32_t* tv = x.at<uint32_t>(x.L.tv);
    uint32_t* cnt = x.at<uint32_t>(x.L.counters);
    CellHash h{tk, tv, x.L.cap};
    res->balance_iterations = 0;
    res->balance_violations = 0;
    res->balance_marked = 0;
    res->balance_converged = 0;
    if (cfg->balance_mode < 0 || n == 0) {
        res->balance_converged = 1;
        return FL_OK;
    }
    if (cfg->balance_mode == 0) {
        FL_REQUIRE(cfg->anchor_level >= 0 && cfg->anchor_level <= 19, "anchor_level must be in 0..19");
        FL_REQUIRE(cfg->max_split_level <= cfg->anchor_level, "max_split_level must not exceed anchor_level");
    } else {
        fl_status st = build_cell_hash(c, n, keys, tk, tv, x.L.cap, x.s);
        if (st) return st;
    }
    // mode 0 starts from the stored levels; flags set before the call enter
    // the shadow only after the first marking pass, as in the reference
    hipLaunchKernelGGL(k_shadow, dim3(sgrid(n)), dim3(kBlock), 0, x.s, c, n, cfg->balance_mode != 0 ? 1 : 0,
                       shadow);
    FL_CHECK_LAUNCH();
    for (int it = 0; it < cfg->max_balance_iters; ++it) {
        FL_HIP(hipMemsetAsync(cnt, 0, 8, x.s));
        if (cfg->balance_mode == 0) {
            hipLaunchKernelGGL(k_anchor_keys, dim3(sgrid(n)), dim3(kBlock), 0, x.s, c, n, cfg->anchor_level,
                               shadow, keys);
            FL_CHECK_LAUNCH();
            fl_status st = hash_clear_impl(tk, tv, x.L.cap, x.s);
            if (st) return st;
            st = cellkey_map_build_impl(tk, tv, x.L.cap, keys, n, x.s);
            if (st) return st;
            hipLaunchKernelGGL(k_detect0, dim3(sgrid(n)), dim3(kBlock), 0, x.s, c, n, cfg->anchor_level, shadow,
                               h, viol, cnt);
        } else {
            hipLaunchKernelGGL(k_detect1, dim3(sgrid(n)), dim3(kBlock), 0, x.s, c, n, Root{{cfg->root[0], cfg->root[1], cfg->root[2]}}, shadow, h,
                               viol, cnt);
        }
        FL_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_mark, dim3(sgrid(n)), dim3(kBlock), 0, x.s, cw, n, cfg->max_split_level, viol,
                           shadow, cnt + 1);
        FL_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_shadow, dim3(sgrid(n)), dim3(kBlock), 0, x.s, c, n, 1, shadow);
        FL_CHECK_LAUNCH();
        uint32_t hc[2];
        FL_HIP(hipMemcpyAsync(hc, cnt, 8, hipMemcpyDeviceToHost, x.s));
        FL_HIP(hipStreamSynchronize(x.s));
        res->balance_iterations = it + 1;
        res->balance_violations += hc[0];
        if (hc[0] == 0) {
            res->balance_converged = 1;
            break;
        }
        res->balance_marked += hc[1];
        if (hc[1] == 0) break;
    }
    return FL_OK;
}

}  // namespace
}  // namespace fl

using namespace fl;

extern "C" {

fl_status fl_face_neighbors(const fl_cells* c, size_t n, const int32_t* root, uint64_t* d_keys_scratch,
                            uint64_t* d_tkeys, uint32_t* d_tvals, uint64_t capacity, uint32_t* d_nbr, fl_stream s) {
    if (n == 0) return FL_OK;
    FL_REQUIRE(c && c->x && c->y && c->z && c->level && d_keys_scratch && d_tkeys && d_tvals && d_nbr && root,
               "null pointer");
    FL_REQUIRE(n <= kMaxMapCells, "too many cells for the cell map (at most 644,245,094)");
    FL_REQUIRE(capacity >= fl_hash_capacity(n) && (capacity & (capacity - 1)) == 0,
               "capacity must be a power of two >= fl_hash_capacity(n)");
    hipStream_t st = FL_STREAM(s);
    fl_status r = build_cell_hash(cview(c), n, d_keys_scratch, d_tkeys, d_tvals, capacity, st);
    if (r) return r;
    CellHash h{d_tkeys, d_tvals, capacity};
    hipLaunchKernelGGL(k_face_nbr, dim3(sgrid(n)), dim3(kBlock), 0, st, cview(c), n, Root{{root[0], root[1], root[2]}},
                       h, d_nbr);
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_mark_ball(fl_cells* c, size_t n, double cx, double cy, double cz, double radius, int target_level,
                       uint64_t* h_marked, fl_stream s) {
    FL_REQUIRE(c && h_marked, "null pointer");
    hipStream_t st = FL_STREAM(s);
    uint32_t* cnt = nullptr;
    FL_HIP(hipMallocAsync(reinterpret_cast<void**>(&cnt), 4, st));
    FL_HIP(hipMemsetAsync(cnt, 0, 4, st));
    if (n) hipLaunchKernelGGL(k_mark_ball, dim3(sgrid(n)), dim3(kBlock), 0, st, wview(c), n, cx, cy, cz, radius,
                              target_level, cnt);
    uint32_t h = 0;
    FL_HIP(hipMemcpyAsync(&h, cnt, 4, hipMemcpyDeviceToHost, st));
    FL_HIP(hipFreeAsync(cnt, st));
    FL_HIP(hipStreamSynchronize(st));
    *h_marked = h;
    return FL_OK;
}

fl_status fl_mark_solid_ball(fl_cells* c, size_t n, double cx, double cy, double cz, double radius,
                             uint64_t* h_marked, fl_stream s) {
    FL_REQUIRE(c && h_marked, "null pointer");
    hipStream_t st = FL_STREAM(s);
    uint32_t* cnt = nullptr;
    FL_HIP(hipMallocAsync(reinterpret_cast<void**>(&cnt), 4, st));
    FL_HIP(hipMemsetAsync(cnt, 0, 4, st));
    if (n) hipLaunchKernelGGL(k_mark_solid, dim3(sgrid(n)), dim3(kBlock), 0, st, wview(c), n, cx, cy, cz, radius,
                              cnt);
    uint32_t h = 0;
    FL_HIP(hipMemcpyAsync(&h, cnt, 4, hipMemcpyDeviceToHost, st));
    FL_HIP(hipFreeAsync(cnt, st));
    FL_HIP(hipStreamSynchronize(st));
    *h_marked = h;
    return FL_OK;
}

fl_status fl_gen_cells(fl_cells* c, size_t n, uint64_t seed, int max_level, fl_stream s) {
    FL_REQUIRE(c && c->x && c->y && c->z && c->level && c->state && c->flag && c->material, "null pointer");
    FL_REQUIRE(max_level >= 0 && max_level <= 19, "max_level must be in 0..19");
    if (n) hipLaunchKernelGGL(k_gen_cells, dim3(sgrid(n)), dim3(kBlock), 0, FL_STREAM(s), seed, n, max_level, wview(c));
    FL_CHECK_LAUNCH();
    return FL_OK;
}

fl_status fl_gen_flags(fl_cells* c, size_t n, uint64_t seed, double p_split, double p_merge, void* d_temp,
                       size_t temp_bytes, fl_stream s) {
    FL_REQUIRE(c && c->x && c->flag, "null pointer");
    FL_REQUIRE(p_split >= 0 && p_split <= 1 && p_merge >= 0 && p_merge <= 1, "probabilities must be in [0,1]");
    RegridLayout L = regrid_layout(n > 0 ? n : 1);
    FL_REQUIRE(temp_bytes >= L.total, "temp buffer too small (see fl_regrid_temp_bytes)");
    if (n == 0) return FL_OK;
    hipStream_t st = FL_STREAM(s);
    auto th = [](double p) { return p >= 1.0 ? ~0ull : (uint64_t)std::ldexp(p, 64); };
    hipLaunchKernelGGL(k_flag_split, dim3(sgrid(n)), dim3(kBlock), 0, st, seed, n, th(p_split), wview(c));
    FL_CHECK_LAUNCH();
    char* b = static_cast<char*>(d_temp);
    fl_status r = build_cell_hash(cview(c), n, reinterpret_cast<uint64_t*>(b + L.keys), reinterpret_cast<uint64_t*>(b + L.tk),
                                  reinterpret_cast<uint32_t*>(b + L.tv), L.cap, st);
    if (r) return r;
    CellHash h{reinterpret_cast<uint64_t*>(b + L.tk), reinterpret_cast<uint32_t*>(b + L.tv), L.cap};
    hipLaunchKernelGGL(k_flag_merge, dim3(sgrid(n)), dim3(kBlock), 0, st, seed, n, th(p_merge), h, wview(c));
    FL_CHECK_LAUNCH();
    return FL_OK;
}

size_t fl_regrid_temp_bytes(size_t n) { return regrid_layout(n > 0 ? n : 1).total; }

fl_status fl_balance(const fl_regrid_config* cfg, fl_cells* c, size_t n, void* d_temp, size_t temp_bytes,
                     fl_regrid_result* res, fl_stream s) {
    FL_REQUIRE(cfg && c && res, "null pointer");
    FL_REQUIRE(n <= kMaxMapCells, "too many cells for the cell map (at most 644,245,094)");
    RegridLayout L = regrid_layout(n > 0 ? n : 1);
    FL_REQUIRE(n == 0 || temp_bytes >= L.total, "temp buffer too small (see fl_regrid_temp_bytes)");
    Ctx x{static_cast<char*>(d_temp), L, FL_STREAM(s)};
    return run_balance(cfg, c, n, x, res);
}

fl_status fl_split(const fl_cells* c, size_t n, int max_split_level, uint32_t* d_child_start, fl_cells* children,
                   size_t child_capacity, void* d_temp, size_t temp_bytes, uint64_t* h_num_split, fl_stream s) {
    FL_REQUIRE(c && children && h_num_split, "null pointer");
    hipStream_t st = FL_STREAM(s);
    size_t w = scan_ws_words(n > 0 ? n : 1);
    FL_REQUIRE(temp_bytes >= (w + 1) * 8, "temp buffer too small (see fl_regrid_temp_bytes)");
    uint64_t* ws = static_cast<uint64_t*>(d_temp);
    uint64_t* cnt = ws + w;
    hipLaunchKernelGGL(k_zero64, dim3(grid_for(w + 1, 256)), dim3(256), 0, st, ws, w + 1);
    FL_CHECK_LAUNCH();
    if (n > 0) {
        hipLaunchKernelGGL(k_split, dim3((unsigned)scan_tiles(n)), dim3(kScanThreads), 0, st, cview(c), n,
                           max_split_level, d_child_start, wview(children), (const uint64_t*)nullptr,
                           child_capacity, (const float*)nullptr, (float*)nullptr, 0, ws, cnt);
        FL_CHECK_LAUNCH();
    }
    FL_HIP(hipMemcpyAsync(h_num_split, cnt, 8, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    if (*h_num_split * 8 > child_capacity) {
        fl::set_error("fl_split: child capacity too small");
        return FL_E_CAPACITY;
    }
    return FL_OK;
}

fl_status fl_merge(const fl_cells* c, size_t n, uint32_t* d_group_id, uint32_t* d_group_leader, fl_cells* parents,
                   size_t parent_capacity, void* d_temp, size_t temp_bytes, uint64_t* h_num_groups, fl_stream s) {
    FL_REQUIRE(c && parents && h_num_groups && d_group_id && d_group_leader, "null pointer");
    FL_REQUIRE(n <= kMaxMapCells, "too many cells for the cell map (at most 644,245,094)");
    RegridLayout L = regrid_layout(n > 0 ? n : 1);
    FL_REQUIRE(temp_bytes >= L.total, "temp buffer too small (see fl_regrid_temp_bytes)");
    Ctx x{static_cast<char*>(d_temp), L, FL_STREAM(s)};
    hipStream_t st = x.s;
    uint64_t* cnt = x.at<uint64_t>(L.counters);
    uint64_t* ws = x.at<uint64_t>(L.ws0);
    size_t w = scan_ws_words(n > 0 ? n : 1);
    hipLaunchKernelGGL(k_zero64, dim3(grid_for(w, 256)), dim3(256), 0, st, ws, w);
    hipLaunchKernelGGL(k_zero64, dim3(1), dim3(64), 0, st, cnt, (size_t)16);
    FL_CHECK_LAUNCH();
    if (n > 0) {
        fl_status r = build_cell_hash(cview(c), n, x.at<uint64_t>(L.keys), x.at<uint64_t>(L.tk),
                                      x.at<uint32_t>(L.tv), L.cap, st);
        if (r) return r;
        hipLaunchKernelGGL(k_fill_u32, dim3(sgrid(n)), dim3(kBlock), 0, st, d_group_id, n, kInvalid);
        CellHash h{x.at<uint64_t>(L.tk), x.at<uint32_t>(L.tv), L.cap};
        hipLaunchKernelGGL(k_merge, dim3((unsigned)scan_tiles(n)), dim3(kScanThreads), 0, st, cview(c), n, h,
                           (const uint8_t*)nullptr, 0, Root{{0, 0, 0}}, d_group_id, d_group_leader, wview(parents), (size_t)0,
                           ws, cnt);
        FL_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_parents, dim3(sgrid(n / 8 + 1)), dim3(kBlock), 0, st, cview(c), d_group_leader, cnt,
                           (const uint64_t*)nullptr, (const uint64_t*)nullptr, 0, wview(parents), parent_capacity, h,
                           (const float*)nullptr, (float*)nullptr, 0, 0);
        FL_CHECK_LAUNCH();
    }
    FL_HIP(hipMemcpyAsync(h_num_groups, cnt, 8, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    if (*h_num_groups > parent_capacity) {
        fl::set_error("fl_merge: parent capacity too small");
        return FL_E_CAPACITY;
    }
    return FL_OK;
}

fl_status fl_regrid(const fl_regrid_config* cfg, fl_cells* in, size_t n, fl_cells* out, size_t out_capacity,
                    const float* d_fields_in, float* d_fields_out, void* d_temp, size_t temp_bytes,
                    fl_regrid_result* res, fl_stream s) {
    FL_REQUIRE(cfg && in && out && res, "null pointer");
    FL_REQUIRE(cfg->ncomp == 0 || (d_fields_in && d_fields_out), "fields required when ncomp > 0");
    FL_REQUIRE(n <= kMaxMapCells, "too many cells for the cell map (at most 644,245,094)");
    RegridLayout L = regrid_layout(n > 0 ? n : 1);
    FL_REQUIRE(temp_bytes >= L.total, "temp buffer too small (see fl_regrid_temp_bytes)");
    std::memset(res, 0, sizeof(*res));
    Ctx x{static_cast<char*>(d_temp), L, FL_STREAM(s)};
    hipStream_t st = x.s;
    fl_status r = run_balance(cfg, in, n, x, res);
    if (r) return r;
    uint64_t* cnt = x.at<uint64_t>(L.counters);  // [0] groups [1] kept [2] split
    size_t w = scan_ws_words(n > 0 ? n : 1);
    hipLaunchKernelGGL(k_zero64, dim3(grid_for(3 * w, 256)), dim3(256), 0, st, x.at<uint64_t>(L.ws0),
                       (L.ws2 - L.ws0) / 8 + w);
    hipLaunchKernelGGL(k_zero64, dim3(1), dim3(64), 0, st, cnt, (size_t)16);
    FL_CHECK_LAUNCH();
    if (n > 0) {
        CellsC c = cview(in);
        // a D3Q19-rule balance (mode 1) has built this map over the same
        // cells and changed only their flags; the reference rule's map holds
        // anchor keys and is replaced
        if (cfg->balance_mode <= 0) {
            r = build_cell_hash(c, n, x.at<uint64_t>(L.keys), x.at<uint64_t>(L.tk), x.at<uint32_t>(L.tv), L.cap, st);
            if (r) return r;
        }
        CellHash h{x.at<uint64_t>(L.tk), x.at<uint32_t>(L.tv), L.cap};
        uint32_t* gid = x.at<uint32_t>(L.gid);
        uint32_t* leaders = x.at<uint32_t>(L.leaders);
        uint8_t* shadow = x.at<uint8_t>(L.shadow);
        hipLaunchKernelGGL(k_shadow, dim3(sgrid(n)), dim3(kBlock), 0, st, c, n, 1, shadow);
        hipLaunchKernelGGL(k_fill_u32, dim3(sgrid(n)), dim3(kBlock), 0, st, gid, n, kInvalid);
        hipLaunchKernelGGL(k_merge, dim3((unsigned)scan_tiles(n)), dim3(kScanThreads), 0, st, c, n, h, shadow,
                           cfg->merge_balance_check, Root{{cfg->root[0], cfg->root[1], cfg->root[2]}}, gid, leaders, wview(out), (size_t)0,
                           x.at<uint64_t>(L.ws0), cnt + 0);
        FL_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_compact, dim3((unsigned)scan_tiles(n)), dim3(kScanThreads), 0, st, c, n,
                           (const uint32_t*)gid, wview(out), out_capacity, d_fields_in, d_fields_out, cfg->ncomp,
                           x.at<uint64_t>(L.ws1), cnt + 1);
        FL_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_split, dim3((unsigned)scan_tiles(n)), dim3(kScanThreads), 0, st, c, n,
                           cfg->max_split_level, (uint32_t*)nullptr, wview(out), (const uint64_t*)(cnt + 1),
                           out_capacity, d_fields_in, d_fields_out, cfg->ncomp, x.at<uint64_t>(L.ws2), cnt + 2);
        FL_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_parents, dim3(sgrid(n / 8 + 1)), dim3(kBlock), 0, st, c, (const uint32_t*)leaders,
                           (const uint64_t*)cnt, (const uint64_t*)(cnt + 1), (const uint64_t*)(cnt + 2), 8,
                           wview(out), out_capacity, h, d_fields_in, d_fields_out, cfg->ncomp, cfg->avg_rule);
        FL_CHECK_LAUNCH();
    }
    uint64_t hc[3];
    FL_HIP(hipMemcpyAsync(hc, cnt, 24, hipMemcpyDeviceToHost, st));
    FL_HIP(hipStreamSynchronize(st));
    res->num_groups = hc[0];
    res->num_split = hc[2];
    res->num_children = 8 * hc[2];
    res->num_cells_out = hc[1] + 8 * hc[2] + hc[0];
    if (res->num_cells_out > out_capacity) {
        fl::set_error("fl_regrid: output capacity too small");
        return FL_E_CAPACITY;
    }
    return FL_OK;
}

}  // extern "C"
