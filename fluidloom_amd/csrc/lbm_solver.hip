// a15/a16: AMR D3Q19 solver -- brick layout builder, coarse/fine step driver
// and the fl_lbm_* C ABI.
//
// The reference's LBM step is an empty DSL stub
// (src/parsing/SimulationBuilder.cpp:394-400) launched through
// KernelNode::execute (src/runtime/KernelNode.cpp:37-106) over fields held by
// SOAFieldManager (src/core/fields/SOAFieldManager.cpp:17-131).  This solver
// is the build's own specification of that step:
//
//  - Leaves (native coords + level, state 0 fluid / 1 solid) are stored per
//    level in 4x4x4 bricks; a level's bricks are the bricks holding its
//    leaves plus their 26 neighbours (so every ghost a leaf reads has a
//    slot), ordered active-first, each group along the Hilbert curve of the
//    brick coordinates.  Populations are SoA [q][brick*64 + local].
//  - Slot kinds: fluid leaf, solid leaf, c2f ghost (covered by a fluid leaf
//    one level coarser), f2c ghost (all 8 children are fluid leaves one
//    level finer), wall (covered by solid).  A fluid leaf may only read
//    regular slots (fluid, c2f, f2c) or walls; anything else is a mesh error
//    (the D3Q19 neighbour-cube balance of the regrid guarantees it).
//  - Time stepping: acoustic scaling tau_l = 2^l (tau0 - 1/2) + 1/2 with
//    recursive sub-cycling; the coarse/fine exchange rescales f_neq as
//    described in lbm_kernels.hip.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fl_internal.hpp"
#include "lbm_device.hpp"
#include "lbm_kernels.hpp"
#include "lbm_layout.hpp"

namespace fl {
namespace lbm {
namespace {

enum : uint8_t { T_NONE = 0, T_FLUID = 1, T_SOLID = 2, T_C2F = 3, T_F2C = 4, T_WALL = 5 };

inline bool regular(uint8_t t) { return t == T_FLUID || t == T_C2F || t == T_F2C; }

// Host copy of the 12-state table (src/core/hilbert/HilbertCodec.cpp:13-26),
// used only to order bricks along the curve.
const uint8_t kHostHT[12][8] = {
    {8, 17, 27, 18, 47, 38, 28, 37},  {16, 71, 1, 62, 51, 52, 2, 61},
    {0, 75, 95, 76, 9, 10, 86, 85},   {4, 77, 55, 78, 3, 66, 80, 65},
    {50, 49, 45, 46, 67, 88, 68, 7},  {6, 57, 39, 72, 5, 58, 84, 83},
    {12, 11, 29, 90, 79, 32, 30, 89}, {74, 91, 73, 40, 69, 92, 70, 15},
    {14, 13, 81, 82, 63, 36, 24, 35}, {20, 31, 19, 56, 53, 54, 42, 41},
    {26, 93, 43, 44, 25, 94, 64, 23}, {22, 87, 21, 60, 33, 48, 34, 59},
};

uint64_t host_encode(uint32_t x, uint32_t y, uint32_t z, int level) {
    uint64_t h = 0;
    int s = 0;
    for (int i = level - 1; i >= 0; --i) {
        int q = (int)((((z >> i) & 1u) << 2) | (((y >> i) & 1u) << 1) | ((x >> i) & 1u));
        int v = kHostHT[s][q];
        h = (h << 3) | (uint64_t)(v & 7);
        s = v >> 3;
    }
    return h;
}

template <class F>
void parallel_for(size_t n, F&& f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    if (nt > 16) nt = 16;
    if (n < 4096 || nt == 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    const size_t chunk = 1024;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (;;) {
                size_t b = next.fetch_add(chunk);
                if (b >= n) break;
                size_t e = std::min(n, b + chunk);
                for (size_t i = b; i < e; ++i) f(i);
            }
        });
    for (auto& t : th) t.join();
}

// Lattice parameters and leaf range of one instance of a batch.
struct InstHost {
    fl_lbm_params p{};
    size_t first = 0, count = 0;
};

inline uint64_t bkey(uint32_t inst, int32_t bx, int32_t by, int32_t bz) {
    return ((uint64_t)inst << 48) | ((uint64_t)(uint16_t)bx << 32) | ((uint64_t)(uint16_t)by << 16) |
           (uint64_t)(uint16_t)bz;
}
inline uint32_t bkey_inst(uint64_t k) { return (uint32_t)(k >> 48); }
inline int32_t bkey_c(uint64_t k, int a) { return (int32_t)((k >> (32 - 16 * a)) & 0xFFFF); }

// parallel_chunks(n, f): f(begin, end) over [0, n) in chunks, on up to 16
// threads; chunk results are combined by the caller.
template <class F>
void parallel_chunks(size_t n, size_t chunk, F&& f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 4;
    if (nt > 16) nt = 16;
    const size_t nchunks = (n + chunk - 1) / chunk;
    if (nchunks <= 1 || nt == 1) {
        for (size_t c = 0; c < nchunks; ++c) f(c * chunk, std::min(n, (c + 1) * chunk));
        return;
    }
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (unsigned t = 0; t < nt && t < nchunks; ++t)
        th.emplace_back([&] {
            for (;;) {
                const size_t c = next.fetch_add(1);
                if (c >= nchunks) break;
                f(c * chunk, std::min(n, (c + 1) * chunk));
            }
        });
    for (auto& t : th) t.join();
}

// Brick index of one instance at one level: a dense grid of brick numbers
// (-1 = none) when the level's brick grid is small enough, else the level's
// hash map.
constexpr size_t kDenseBricksMax = size_t(1) << 27;

struct BrickGrid {
    int32_t n[3] = {0, 0, 0};
    bool dense = false;
    std::vector<int32_t> idx;
    size_t at(int32_t bx, int32_t by, int32_t bz) const { return ((size_t)bz * n[1] + by) * n[0] + bx; }
    size_t cells() const { return (size_t)n[0] * n[1] * n[2]; }
};

struct HostLevel {
    int level = 0;
    std::vector<std::array<int32_t, 3>> iext;  // [instance] extents in cells at this level
    uint32_t nb = 0, nactive = 0;
    uint64_t nfluid = 0;
    std::vector<uint64_t> nfluid_inst;
    std::vector<BrickGrid> grid;                // [instance]
    std::unordered_map<uint64_t, int32_t> bmap;  // bricks of instances without a dense grid
    std::vector<int32_t> bcoord, nbr;
    std::vector<uint16_t> binst;
    std::vector<uint32_t> bflags;
    std::vector<uint64_t> cflags;
    std::vector<uint8_t> stype;
    std::vector<uint32_t> needed;  // per ghost slot: bit i set when a fluid leaf pulls population i from it
    std::vector<uint32_t> c2f_ghost, cap_slots, f2c_ghost, f2c_child, force_links;
    // c2f ghosts grouped by covering coarse leaf: 8 child slots per group
    // (kNoCap where that child is not a ghost) and 7 capture indices
    std::vector<uint32_t> c2f_gchild, c2f_gcap, c2f_gmask;
    std::vector<uint32_t> abeg;  // [ninst+1] active-brick range of each instance
    std::vector<uint32_t> lbeg;  // [ninst+1] solid-link range of each instance
    // list sizes (set by either builder; the host vectors above exist only
    // when the layout was built on the host)
    size_t n_c2f = 0, n_groups = 0, n_cap = 0, n_f2c = 0, n_links = 0;

    int32_t brick_at(int inst, int32_t bx, int32_t by, int32_t bz) const {
        const BrickGrid& g = grid[(size_t)inst];
        if (bx < 0 || by < 0 || bz < 0 || bx >= g.n[0] || by >= g.n[1] || bz >= g.n[2]) return -1;
        if (g.dense) return g.idx[g.at(bx, by, bz)];
        auto it = bmap.find(bkey((uint32_t)inst, bx, by, bz));
        return it == bmap.end() ? -1 : it->second;
    }
    int64_t slot_at(int inst, int32_t x, int32_t y, int32_t z) const {
        const auto& e = iext[(size_t)inst];
        if (x < 0 || y < 0 || z < 0 || x >= e[0] || y >= e[1] || z >= e[2]) return -1;
        const int32_t b = brick_at(inst, x >> 2, y >> 2, z >> 2);
        if (b < 0) return -1;
        return (int64_t)b * 64 + (x & 3) + 4 * (y & 3) + 16 * (z & 3);
    }
    uint8_t type_at(int inst, int32_t x, int32_t y, int32_t z) const {
        int64_t s = slot_at(inst, x, y, z);
        return s < 0 ? T_NONE : stype[(size_t)s];
    }
};

struct DevLevel {
    int32_t* nbr = nullptr;
    int32_t* bcoord = nullptr;
    uint16_t* binst = nullptr;
    uint32_t* bflags = nullptr;
    uint64_t* cflags = nullptr;
    uint32_t* c2f_gchild = nullptr;
    uint32_t* c2f_gcap = nullptr;
    uint32_t* c2f_gmask = nullptr;
    uint32_t* cap_slots = nullptr;
    uint32_t* f2c_ghost = nullptr;
    uint32_t* f2c_child = nullptr;
    uint32_t* force_links = nullptr;
    uint32_t* abeg = nullptr;
    uint32_t* lbeg = nullptr;
    uint64_t* map_keys = nullptr;  // brick map: packed brick key -> brick (slot lookups)
    uint32_t* map_vals = nullptr;
    uint64_t map_cap = 0;
    float* itau = nullptr;
    double* itau_d = nullptr;
    double* bvals = nullptr;  // [max(nactive * 4, nlinks * 3)] reduction staging
    float* uprev = nullptr;   // [nactive * 64 * 3] velocity snapshot (allocated on first use)
    float* uacc = nullptr;    // [nactive * 64 * 3] sum of the window's velocity samples (first use)
    void* buf[2] = {nullptr, nullptr};
    void* cap[2] = {nullptr, nullptr};
    // cap[0] holds this level's capture of buf[1 - cur] from the end of the
    // previous step of the current stepping pass (see advance)
    bool cap_fresh = false;
    // f2c staging: this (coarse) level's staged child sums, [23][n_f2c] in
    // the compute type, and which ghosts have them (LevelDev::stage); as a
    // fine level, its bricks' block -> coarse ghost map (LevelDev::f2c_blk).
    // stage_fresh: the finer level's last collide staged this level's
    // ghosts in the current pass.
    void* f2c_stage = nullptr;
    uint8_t* f2c_staged = nullptr;
    uint32_t* f2c_blk = nullptr;
    bool stage_fresh = false;
    // capture staging: slot -> capture index of the captures this level's
    // collide writes itself (LevelDev::cap_of), and which captures those are
    uint32_t* cap_of = nullptr;
    uint8_t* cap_staged = nullptr;
    int cur = 0;
    LevelDev ld{};
};

size_t store_bytes(int storage) {
    return storage == FL_STORE_F64 ? 8 : ((storage == FL_STORE_F16 || storage == FL_STORE_F16C) ? 2 : 4);
}
size_t compute_bytes(int storage) { return storage == FL_STORE_F64 ? 8 : 4; }

}  // namespace
}  // namespace lbm
}  // namespace fl

struct fl_lbm_solver {
    std::vector<fl::lbm::InstHost> inst;
    std::vector<fl::lbm::InstDev> hinst;
    std::vector<fl::lbm::LayoutInst> linst;
    fl::lbm::LayoutInst* dlinst = nullptr;
    int storage = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int L = 0;
    fl::lbm::HostLevel h[24];
    fl::lbm::DevLevel d[24];
    fl::lbm::InstDev* dinst = nullptr;
    std::vector<void*> allocs;
    double* acc = nullptr;  // [ninst * 4] per-instance reduction results
    uint32_t* ticks = nullptr;  // [24] collide launches per level (FP16 stochastic-rounding clock)
    bool own_ticks = false;
    int64_t* qslots = nullptr;
    double* qbuf = nullptr;
    size_t qcap = 0;
    uint64_t device_bytes = 0;
    int variant[24];
    bool use_graph = false;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    uint64_t graph_launches = 0;
    int nsamples = 0;  // velocity samples in the current steady-state window

    fl_lbm_solver() {
        for (int l = 0; l < 24; ++l) variant[l] = 4;  // bpw2-xcd until tuned
    }
    void drop_graphs() {
        for (int k = 0; k < 2; ++k) {
            if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            gexec[k] = nullptr;
            graph[k] = nullptr;
        }
    }
    ~fl_lbm_solver() {
        drop_graphs();
        for (void* a : allocs) fl::cache_free(a, stream);
        if (qslots) (void)hipFree(qslots);
        if (qbuf) (void)hipFree(qbuf);
        if (own_ticks && ticks) (void)hipFree(ticks);
        // the blocks cached above are tagged with this stream: its work ends
        // before the handle can be reused by a new stream
        if (own_stream && stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
    int ninst() const { return (int)inst.size(); }
};

namespace fl {
namespace lbm {
namespace {

template <typename T>
fl_status dev_upload(fl_lbm_solver* S, T** dst, const std::vector<T>& v) {
    *dst = nullptr;
    if (v.empty()) return FL_OK;
    FL_TRY(fl::cache_alloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T), S->stream));
    S->allocs.push_back(*dst);
    S->device_bytes += v.size() * sizeof(T);
    FL_HIP(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, S->stream));
    return FL_OK;
}

fl_status dev_alloc(fl_lbm_solver* S, void** dst, size_t bytes) {
    *dst = nullptr;
    if (bytes == 0) return FL_OK;
    FL_TRY(fl::cache_alloc(dst, bytes, S->stream));
    S->allocs.push_back(*dst);
    S->device_bytes += bytes;
    return FL_OK;
}

fl_status mesh_error(const std::string& m) {
    fl::set_error("fl_lbm_create: " + m);
    return FL_E_MESH;
}

// Source of direction i from (x,y,z) after periodic wrap.  Returns 0 when
// inside; else 1 with *face = wall face (axes z,y,x) or -1 for outflow, in
// which case (sx,sy,sz) is clamped into the domain.
int source_of(const fl_lbm_params& p, const std::array<int32_t, 3>& ext, int x, int y, int z, int i, int* sx, int* sy,
              int* sz, int* face) {
    int q[3] = {x - cx(i), y - cy(i), z - cz(i)};
    int out[3] = {0, 0, 0};
    bool any = false;
    for (int a = 0; a < 3; ++a) {
        int f = q[a] < 0 ? 2 * a : (q[a] >= ext[a] ? 2 * a + 1 : -1);
        if (f < 0) continue;
        if (p.bc[f] == FL_BC_PERIODIC) {
            q[a] = (q[a] + ext[a]) % ext[a];
        } else {
            out[a] = f + 1;
            any = true;
        }
    }
    *sx = q[0]; *sy = q[1]; *sz = q[2];
    if (!any) return 0;
    *face = -1;
    for (int a = 2; a >= 0; --a)
        if (out[a] && p.bc[out[a] - 1] == FL_BC_WALL) {
            *face = out[a] - 1;
            break;
        }
    if (*face < 0) {
        for (int a = 0; a < 3; ++a)
            if (out[a]) q[a] = q[a] < 0 ? 0 : ext[a] - 1;
        *sx = q[0]; *sy = q[1]; *sz = q[2];
    }
    return 1;
}

// Brick coordinates of a neighbour after periodic wrap; false when outside.
inline bool wrap_brick(const fl_lbm_params& p, const std::array<int32_t, 3>& nbx, int32_t* c) {
    for (int a = 0; a < 3; ++a) {
        if (c[a] >= 0 && c[a] < nbx[a]) continue;
        int f = c[a] < 0 ? 2 * a : 2 * a + 1;
        if (p.bc[f] != FL_BC_PERIODIC) return false;
        c[a] = (c[a] + nbx[a]) % nbx[a];
    }
    return true;
}

fl_status build_layout(fl_lbm_solver* S, const int32_t* X, const int32_t* Y, const int32_t* Z, const uint8_t* LV,
                       const uint8_t* ST, size_t n) {
    const int NL = S->L;
    const int NI = S->ninst();
    // leaf -> instance
    std::vector<uint16_t> linst(n);
    for (int k = 0; k < NI; ++k)
        for (size_t j = S->inst[k].first; j < S->inst[k].first + S->inst[k].count; ++j) linst[j] = (uint16_t)k;
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        H.level = l;
        H.iext.resize(NI);
        for (int k = 0; k < NI; ++k) {
            const fl_lbm_params& p = S->inst[k].p;
            H.iext[k] = {p.nx0 << l, p.ny0 << l, p.nz0 << l};
        }
    }
    // 1-2. bricks per level: the leaf bricks plus the 26-neighbour dilation of
    //      the active (fluid-holding) bricks.  Instances whose brick grid is
    //      small enough mark a dense byte grid (bit 0 leaf brick, bit 1
    //      active, bit 2 dilation); the others collect and sort keys.
    for (size_t j = 0; j < n; ++j) {
        const int in = linst[j];
        if (LV[j] > S->inst[in].p.max_level) return mesh_error("leaf level above the instance's max_level");
        const auto& e = S->h[LV[j]].iext[in];
        if (X[j] < 0 || Y[j] < 0 || Z[j] < 0 || X[j] >= e[0] || Y[j] >= e[1] || Z[j] >= e[2])
            return mesh_error("leaf outside the domain");
    }
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        H.grid.assign(NI, BrickGrid());
        struct E { int act; uint32_t inst; uint64_t h; uint64_t k; };
        std::vector<E> es;
        for (int in = 0; in < NI; ++in) {
            const fl_lbm_params& p = S->inst[in].p;
            BrickGrid& g = H.grid[in];
            for (int a = 0; a < 3; ++a) g.n[a] = H.iext[in][a] / 4;
            const std::array<int32_t, 3> nbx = {g.n[0], g.n[1], g.n[2]};
            const size_t j0 = S->inst[in].first, j1 = j0 + S->inst[in].count;
            if (l > p.max_level) continue;
            g.dense = g.cells() <= kDenseBricksMax;
            if (g.dense) {
                std::vector<uint8_t> fl(g.cells(), 0);
                parallel_chunks(j1 - j0, 1 << 16, [&](size_t b0, size_t b1) {
                    for (size_t j = j0 + b0; j < j0 + b1; ++j) {
                        if (LV[j] != l) continue;
                        const size_t c = g.at(X[j] >> 2, Y[j] >> 2, Z[j] >> 2);
                        const uint8_t bits = ST[j] == 0 ? 3 : 1;
                        if ((__atomic_load_n(&fl[c], __ATOMIC_RELAXED) & bits) != bits)
                            __atomic_fetch_or(&fl[c], bits, __ATOMIC_RELAXED);
                    }
                });
                parallel_chunks(g.cells(), 1 << 16, [&](size_t c0, size_t c1) {
                    for (size_t c = c0; c < c1; ++c) {
                        if (!(__atomic_load_n(&fl[c], __ATOMIC_RELAXED) & 2)) continue;
                        const int32_t bx = (int32_t)(c % g.n[0]), by = (int32_t)((c / g.n[0]) % g.n[1]),
                                      bz = (int32_t)(c / ((size_t)g.n[0] * g.n[1]));
                        for (int dz = -1; dz <= 1; ++dz)
                            for (int dy = -1; dy <= 1; ++dy)
                                for (int dx = -1; dx <= 1; ++dx) {
                                    int32_t q[3] = {bx + dx, by + dy, bz + dz};
                                    if (!wrap_brick(p, nbx, q)) continue;
                                    uint8_t& f = fl[g.at(q[0], q[1], q[2])];
                                    if (!(__atomic_load_n(&f, __ATOMIC_RELAXED) & 4))
                                        __atomic_fetch_or(&f, (uint8_t)4, __ATOMIC_RELAXED);
                                }
                    }
                });
                // collect in grid order, chunk by chunk
                const size_t chunk = 1 << 18, nch = (g.cells() + chunk - 1) / chunk;
                std::vector<std::vector<E>> part(nch);
                parallel_chunks(g.cells(), chunk, [&](size_t c0, size_t c1) {
                    auto& v = part[c0 / chunk];
                    for (size_t c = c0; c < c1; ++c) {
                        if (!fl[c]) continue;
                        const int32_t bx = (int32_t)(c % g.n[0]), by = (int32_t)((c / g.n[0]) % g.n[1]),
                                      bz = (int32_t)(c / ((size_t)g.n[0] * g.n[1]));
                        v.push_back(E{(fl[c] & 2) ? 0 : 1, (uint32_t)in,
                                      host_encode((uint32_t)bx, (uint32_t)by, (uint32_t)bz, 16),
                                      bkey((uint32_t)in, bx, by, bz)});
                    }
                });
                for (auto& v : part) es.insert(es.end(), v.begin(), v.end());
            } else {
                std::vector<uint64_t> leaf, act;
                for (size_t j = j0; j < j1; ++j) {
                    if (LV[j] != l) continue;
                    const uint64_t k = bkey((uint32_t)in, X[j] >> 2, Y[j] >> 2, Z[j] >> 2);
                    leaf.push_back(k);
                    if (ST[j] == 0) act.push_back(k);
                }
                std::sort(act.begin(), act.end());
                act.erase(std::unique(act.begin(), act.end()), act.end());
                std::vector<uint64_t> all = leaf;
                for (uint64_t k : act)
                    for (int dz = -1; dz <= 1; ++dz)
                        for (int dy = -1; dy <= 1; ++dy)
                            for (int dx = -1; dx <= 1; ++dx) {
                                int32_t q[3] = {bkey_c(k, 0) + dx, bkey_c(k, 1) + dy, bkey_c(k, 2) + dz};
                                if (wrap_brick(p, nbx, q)) all.push_back(bkey((uint32_t)in, q[0], q[1], q[2]));
                            }
                std::sort(all.begin(), all.end());
                all.erase(std::unique(all.begin(), all.end()), all.end());
                for (uint64_t k : all)
                    es.push_back(E{std::binary_search(act.begin(), act.end(), k) ? 0 : 1, (uint32_t)in,
                                   host_encode((uint32_t)bkey_c(k, 0), (uint32_t)bkey_c(k, 1),
                                               (uint32_t)bkey_c(k, 2), 16),
                                   k});
            }
        }
        // order: active first, then by instance, then along the Hilbert curve
        std::sort(es.begin(), es.end(), [](const E& a, const E& b) {
            if (a.act != b.act) return a.act < b.act;
            if (a.inst != b.inst) return a.inst < b.inst;
            return a.h < b.h;
        });
        H.nb = (uint32_t)es.size();
        H.nactive = 0;
        while (H.nactive < H.nb && es[H.nactive].act == 0) ++H.nactive;
        H.bcoord.resize((size_t)H.nb * 3);
        H.binst.resize(H.nb);
        H.abeg.assign(NI + 1, 0);
        for (int in = 0; in < NI; ++in)
            if (H.grid[in].dense) H.grid[in].idx.assign(H.grid[in].cells(), -1);
        for (uint32_t b = 0; b < H.nb; ++b) {
            const uint64_t k = es[b].k;
            const uint32_t in = bkey_inst(k);
            for (int a = 0; a < 3; ++a) H.bcoord[3 * b + a] = bkey_c(k, a);
            H.binst[b] = (uint16_t)in;
            BrickGrid& g = H.grid[in];
            if (g.dense)
                g.idx[g.at(bkey_c(k, 0), bkey_c(k, 1), bkey_c(k, 2))] = (int32_t)b;
            else
                H.bmap.emplace(k, (int32_t)b);
            if (b < H.nactive) H.abeg[in + 1]++;
        }
        for (int k = 0; k < NI; ++k) H.abeg[k + 1] += H.abeg[k];
        H.nbr.assign((size_t)H.nb * 27, -1);
        parallel_chunks(H.nb, 4096, [&](size_t b0, size_t b1) {
            for (size_t b = b0; b < b1; ++b) {
                const uint32_t in = H.binst[b];
                const fl_lbm_params& p = S->inst[in].p;
                const BrickGrid& g = H.grid[in];
                const std::array<int32_t, 3> nbx = {g.n[0], g.n[1], g.n[2]};
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            int32_t c[3] = {H.bcoord[3 * b] + dx, H.bcoord[3 * b + 1] + dy, H.bcoord[3 * b + 2] + dz};
                            if (!wrap_brick(p, nbx, c)) continue;
                            H.nbr[b * 27 + 13 + dx + 3 * dy + 9 * dz] = H.brick_at((int)in, c[0], c[1], c[2]);
                        }
            }
        });
        H.stype.assign((size_t)H.nb * 64, T_NONE);
        H.needed.assign((size_t)H.nb * 64, 0);
        H.cflags.assign((size_t)H.nb * 64, 0);
        H.bflags.assign(H.nb, 0);
    }
    // 3. leaf slots (chunk-local counters)
    {
        std::atomic<uint64_t> placed{0}, missing{0};
        std::vector<std::atomic<uint64_t>> nfl((size_t)NL * NI);
        for (auto& a : nfl) a.store(0);
        parallel_chunks(n, 1 << 16, [&](size_t j0, size_t j1) {
            uint64_t pl = 0, mi = 0;
            std::vector<uint64_t> loc((size_t)NL * NI, 0);
            for (size_t k = j0; k < j1; ++k) {
                HostLevel& H = S->h[LV[k]];
                int64_t s = H.slot_at(linst[k], X[k], Y[k], Z[k]);
                if (s < 0) { ++mi; continue; }
                uint8_t t = ST[k] == 0 ? T_FLUID : T_SOLID;
                uint8_t prev = __atomic_exchange_n(&H.stype[(size_t)s], t, __ATOMIC_RELAXED);
                if (prev == T_NONE) ++pl;
                if (t == T_FLUID) ++loc[(size_t)LV[k] * NI + linst[k]];
            }
            placed.fetch_add(pl);
            missing.fetch_add(mi);
            for (size_t q = 0; q < loc.size(); ++q)
                if (loc[q]) nfl[q].fetch_add(loc[q]);
        });
        if (missing.load()) return mesh_error("internal: leaf brick missing");
        if (placed.load() != n) return mesh_error("duplicate leaf");
        for (int l = 0; l < NL; ++l) {
            HostLevel& H = S->h[l];
            H.nfluid_inst.assign(NI, 0);
            H.nfluid = 0;
            for (int k = 0; k < NI; ++k) {
                H.nfluid_inst[k] = nfl[(size_t)l * NI + k].load();
                H.nfluid += H.nfluid_inst[k];
            }
        }
    }
    // 4. classify the non-leaf slots from the leaves one level up / down
    //    (reads only leaf kinds, which are final after step 3)
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        parallel_for(H.nb, [&](size_t b) {
            const int in = H.binst[b];
            for (int loc = 0; loc < 64; ++loc) {
                size_t s = b * 64 + loc;
                if (H.stype[s] != T_NONE) continue;
                int32_t x = H.bcoord[3 * b] * 4 + (loc & 3), y = H.bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                        z = H.bcoord[3 * b + 2] * 4 + (loc >> 4);
                if (l > 0) {
                    const HostLevel& C = S->h[l - 1];
                    int64_t cs = C.slot_at(in, x >> 1, y >> 1, z >> 1);
                    uint8_t ct = cs < 0 ? T_NONE : C.stype[(size_t)cs];
                    if (ct == T_FLUID) { H.stype[s] = T_C2F; continue; }
                    if (ct == T_SOLID) { H.stype[s] = T_WALL; continue; }
                }
                if (l + 1 < NL) {
                    const HostLevel& F = S->h[l + 1];
                    int nf = 0, ns = 0;
                    for (int q = 0; q < 8; ++q) {
                        uint8_t t = F.type_at(in, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
                        nf += t == T_FLUID;
                        ns += t == T_SOLID;
                    }
                    if (nf == 8) H.stype[s] = T_F2C;
                    else if (nf + ns == 8) H.stype[s] = T_WALL;
                }
            }
        });
    }
    // 5. link codes of every fluid leaf
    std::atomic<int> bad{0};
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        std::vector<std::vector<uint32_t>> blinks(H.nactive);
        parallel_for(H.nb, [&](size_t b) {
            const int in = H.binst[b];
            const fl_lbm_params& p = S->inst[in].p;
            bool simple = true;
            for (int loc = 0; loc < 64; ++loc) {
                size_t s = b * 64 + loc;
                if (H.stype[s] != T_FLUID) { simple = false; continue; }
                int32_t x = H.bcoord[3 * b] * 4 + (loc & 3), y = H.bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                        z = H.bcoord[3 * b + 2] * 4 + (loc >> 4);
                uint64_t w = kActiveBit;
                for (int i = 1; i < Q; ++i) {
                    int sx, sy, sz, face;
                    unsigned code;
                    if (source_of(p, H.iext[in], x, y, z, i, &sx, &sy, &sz, &face)) {
                        if (face >= 0) {
                            code = kLinkWall;
                        } else {
                            int64_t s2 = H.slot_at(in, sx, sy, sz);
                            uint8_t t2 = s2 < 0 ? T_NONE : H.stype[(size_t)s2];
                            if (regular(t2)) {
                                code = kLinkOutflow;
                                // the outlet reads every population of its source
                                if (t2 != T_FLUID)
                                    __atomic_fetch_or(&H.needed[(size_t)s2], (1u << Q) - 1u, __ATOMIC_RELAXED);
                            } else if (t2 == T_SOLID || t2 == T_WALL) {
                                code = kLinkSolid;
                            } else {
                                bad.store(1);
                                code = kLinkSolid;
                            }
                        }
                    } else {
                        const int px = (loc & 3) - cx(i), py = ((loc >> 2) & 3) - cy(i), pz = (loc >> 4) - cz(i);
                        const int nb = H.nbr[b * 27 + 13 + (px >> 2) + 3 * (py >> 2) + 9 * (pz >> 2)];
                        int64_t s2 = nb < 0 ? -1 : (int64_t)nb * 64 + (px & 3) + 4 * (py & 3) + 16 * (pz & 3);
                        uint8_t t2 = s2 < 0 ? T_NONE : H.stype[(size_t)s2];
                        if (regular(t2)) {
                            code = kLinkRegular;
                            if (t2 != T_FLUID) __atomic_fetch_or(&H.needed[(size_t)s2], 1u << i, __ATOMIC_RELAXED);
                        } else if (t2 == T_SOLID || t2 == T_WALL) {
                            code = kLinkSolid;
                            blinks[b].push_back((uint32_t)(s * 32 + i));
                        } else {
                            bad.store(1);
                            code = kLinkSolid;
                        }
                    }
                    w |= (uint64_t)code << (2 * i);
                    if (code != kLinkRegular) simple = false;
                }
                H.cflags[s] = w;
            }
            H.bflags[b] = simple ? 1u : 0u;
        });
        // solid links in slot order: contiguous per instance (active bricks
        // are grouped by instance)
        H.lbeg.assign(NI + 1, 0);
        for (uint32_t b = 0; b < H.nactive; ++b) {
            H.force_links.insert(H.force_links.end(), blinks[b].begin(), blinks[b].end());
            H.lbeg[H.binst[b] + 1] += (uint32_t)blinks[b].size();
        }
        for (int k = 0; k < NI; ++k) H.lbeg[k + 1] += H.lbeg[k];
    }
    if (bad.load()) return mesh_error("a fluid leaf reads a position that is neither a leaf nor a ghost "
                                      "(mesh not 2:1 balanced over D3Q19 neighbour cubes)");
    // 6. ghost lists
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        for (size_t s = 0; s < H.stype.size(); ++s) {
            if (!H.needed[s]) continue;
            if (H.stype[s] == T_C2F) H.c2f_ghost.push_back((uint32_t)s);
            if (H.stype[s] == T_F2C) H.f2c_ghost.push_back((uint32_t)s);
        }
    }
    // c2f sources.  Ghosts are grouped by their covering coarse leaf P (one
    // group per P, child q = fine-coordinate parity x + 2y + 4z); each group
    // captures P and P's coarse fluid neighbours -x, +x, -y, +y, -z, +z
    // (across periodic faces), once per coarse step.
    for (int l = 1; l < NL; ++l) {
        HostLevel& H = S->h[l];
        HostLevel& C = S->h[l - 1];
        const size_t ng = H.c2f_ghost.size();
        std::vector<int64_t> par(ng);
        std::vector<uint8_t> chq(ng);
        parallel_chunks(ng, 4096, [&](size_t k0, size_t k1) {
            for (size_t k = k0; k < k1; ++k) {
                const uint32_t g = H.c2f_ghost[k];
                const size_t b = g / 64;
                const int loc = (int)(g % 64);
                const int32_t fx = H.bcoord[3 * b] * 4 + (loc & 3), fy = H.bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                              fz = H.bcoord[3 * b + 2] * 4 + (loc >> 4);
                par[k] = C.slot_at(H.binst[b], fx >> 1, fy >> 1, fz >> 1);
                chq[k] = (uint8_t)((fx & 1) | ((fy & 1) << 1) | ((fz & 1) << 2));
            }
        });
        for (size_t k = 0; k < ng; ++k)
            if (par[k] < 0) return mesh_error("c2f ghost with no covering coarse leaf");
        std::vector<uint32_t> order(ng);
        for (size_t k = 0; k < ng; ++k) order[k] = (uint32_t)k;
        std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return par[x] < par[y]; });
        std::vector<int64_t> gpar;
        H.c2f_gchild.clear();
        H.c2f_gmask.clear();
        for (uint32_t k : order) {
            if (gpar.empty() || gpar.back() != par[k]) {
                gpar.push_back(par[k]);
                H.c2f_gchild.insert(H.c2f_gchild.end(), 8, kNoCap);
                H.c2f_gmask.insert(H.c2f_gmask.end(), 8, 0u);
            }
            H.c2f_gchild[8 * (gpar.size() - 1) + chq[k]] = H.c2f_ghost[k];
            H.c2f_gmask[8 * (gpar.size() - 1) + chq[k]] = H.needed[H.c2f_ghost[k]];
        }
        const size_t G = gpar.size();
        std::vector<int64_t> src(G * kC2FCaps, -1);
        parallel_chunks(G, 4096, [&](size_t k0, size_t k1) {
            for (size_t k = k0; k < k1; ++k) {
                const int64_t P = gpar[k];
                const size_t cb = (size_t)P / 64;
                const int loc = (int)(P % 64);
                const int in = C.binst[cb];
                const fl_lbm_params& p = S->inst[in].p;
                const int32_t pc[3] = {C.bcoord[3 * cb] * 4 + (loc & 3), C.bcoord[3 * cb + 1] * 4 + ((loc >> 2) & 3),
                                       C.bcoord[3 * cb + 2] * 4 + (loc >> 4)};
                src[k * kC2FCaps] = P;
                for (int a = 0; a < 3; ++a)
                    for (int d = 0; d < 2; ++d) {
                        int32_t q[3] = {pc[0], pc[1], pc[2]};
                        q[a] += d ? 1 : -1;
                        const int32_t e = C.iext[in][a];
                        if (q[a] < 0 || q[a] >= e) {
                            if (p.bc[2 * a + d] != FL_BC_PERIODIC) continue;
                            q[a] = (q[a] + e) % e;
                        }
                        const int64_t s2 = C.slot_at(in, q[0], q[1], q[2]);
                        if (s2 >= 0 && C.stype[(size_t)s2] == T_FLUID) src[k * kC2FCaps + 1 + 2 * a + d] = s2;
                    }
            }
        });
        std::vector<uint32_t> uniq;
        uniq.reserve(src.size());
        for (int64_t v : src)
            if (v >= 0) uniq.push_back((uint32_t)v);
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        C.cap_slots = uniq;
        H.c2f_gcap.assign(src.size(), kNoCap);
        for (size_t k = 0; k < src.size(); ++k)
            if (src[k] >= 0)
                H.c2f_gcap[k] = (uint32_t)(std::lower_bound(uniq.begin(), uniq.end(), (uint32_t)src[k]) - uniq.begin());
    }
    for (int l = 0; l + 1 < NL; ++l) {
        HostLevel& H = S->h[l];
        const HostLevel& F = S->h[l + 1];
        H.f2c_child.resize(H.f2c_ghost.size() * 8);
        for (size_t k = 0; k < H.f2c_ghost.size(); ++k) {
            size_t s = H.f2c_ghost[k];
            size_t b = s / 64;
            int loc = (int)(s % 64);
            int32_t x = H.bcoord[3 * b] * 4 + (loc & 3), y = H.bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                    z = H.bcoord[3 * b + 2] * 4 + (loc >> 4);
            for (int q = 0; q < 8; ++q)
                H.f2c_child[8 * k + q] =
                    (uint32_t)F.slot_at(H.binst[b], 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
        }
        // ghosts in the order of their first child's fine slot: the eight
        // ghosts over one fine brick become adjacent, so a wave's gathers read
        // one brick's planes (and their neighbours') instead of eight bricks'
        const size_t n = H.f2c_ghost.size();
        std::vector<uint32_t> ord(n);
        for (size_t k = 0; k < n; ++k) ord[k] = (uint32_t)k;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](uint32_t a, uint32_t b2) { return H.f2c_child[8 * (size_t)a] < H.f2c_child[8 * (size_t)b2]; });
        std::vector<uint32_t> g2(n), c2(8 * n);
        for (size_t k = 0; k < n; ++k) {
            g2[k] = H.f2c_ghost[ord[k]];
            for (int q = 0; q < 8; ++q) c2[8 * k + q] = H.f2c_child[8 * (size_t)ord[k] + q];
        }
        H.f2c_ghost.swap(g2);
        H.f2c_child.swap(c2);
    }
    for (int l = 0; l < NL; ++l)
        if ((uint64_t)S->h[l].nb * 64 >= (1ull << 32) / 32)
            return mesh_error("level too large for 32-bit slot lists");
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        H.n_c2f = H.c2f_ghost.size();
        H.n_groups = H.c2f_gcap.size() / kC2FCaps;
        H.n_cap = H.cap_slots.size();
        H.n_f2c = H.f2c_ghost.size();
        H.n_links = H.force_links.size();
    }
    return FL_OK;
}

// Device bytes the layout needs (buffers, tables, lists, staging).
uint64_t layout_bytes(const fl_lbm_solver* S) {
    const size_t sb = store_bytes(S->storage), cb = compute_bytes(S->storage);
    const int NI = S->ninst();
    uint64_t t = (uint64_t)NI * (sizeof(InstDev) + 4 * sizeof(double)) + (uint64_t)S->L * sizeof(uint32_t);
    for (int l = 0; l < S->L; ++l) {
        const HostLevel& H = S->h[l];
        t += H.nbr.size() * 4 + H.bcoord.size() * 4 + H.binst.size() * 2 + H.bflags.size() * 4 + H.cflags.size() * 8;
        t += (H.c2f_gchild.size() + H.c2f_gcap.size() + H.c2f_gmask.size() + H.cap_slots.size() + H.f2c_ghost.size() +
              H.f2c_child.size() +
              H.force_links.size() + H.abeg.size() + H.lbeg.size()) * 4;
        t += (uint64_t)NI * (sizeof(float) + sizeof(double));
        t += std::max<uint64_t>((uint64_t)H.nactive * 4, (uint64_t)H.force_links.size() * 3) * 8;
        t += 2ull * H.nb * 64 * Q * sb + 2ull * cap_elems(H.cap_slots.size()) * cb;
        t += brick_map_bytes(H.nb);
        // f2c staging: sums and flags per f2c ghost, the finer level's block map
        if (!H.f2c_ghost.empty() && l + 1 < S->L)
            t += H.f2c_ghost.size() * ((4 + Q) * cb + 1) + (uint64_t)S->h[l + 1].nb * 8 * sizeof(uint32_t);
        // capture staging: slot -> capture map and flags
        if (!H.cap_slots.empty()) t += (uint64_t)H.nb * 64 * sizeof(uint32_t) + H.cap_slots.size();
    }
    return t;
}

fl_status solver_alloc(fl_lbm_solver* S, void** p, size_t bytes) { return dev_alloc(S, p, bytes); }

// Device tables of a layout built on the host, and the brick map that slot
// lookups use with either builder.
fl_status upload_host_tables(fl_lbm_solver* S, int l) {
    HostLevel& H = S->h[l];
    DevLevel& D = S->d[l];
    fl_status r;
    if ((r = dev_upload(S, &D.nbr, H.nbr))) return r;
    if ((r = dev_upload(S, &D.bcoord, H.bcoord))) return r;
    if ((r = dev_upload(S, &D.binst, H.binst))) return r;
    if ((r = dev_upload(S, &D.bflags, H.bflags))) return r;
    if ((r = dev_upload(S, &D.cflags, H.cflags))) return r;
    if ((r = dev_upload(S, &D.c2f_gchild, H.c2f_gchild))) return r;
    if ((r = dev_upload(S, &D.c2f_gcap, H.c2f_gcap))) return r;
    if ((r = dev_upload(S, &D.c2f_gmask, H.c2f_gmask))) return r;
    if ((r = dev_upload(S, &D.cap_slots, H.cap_slots))) return r;
    if ((r = dev_upload(S, &D.f2c_ghost, H.f2c_ghost))) return r;
    if ((r = dev_upload(S, &D.f2c_child, H.f2c_child))) return r;
    if ((r = dev_upload(S, &D.force_links, H.force_links))) return r;
    if ((r = dev_upload(S, &D.abeg, H.abeg))) return r;
    if ((r = dev_upload(S, &D.lbeg, H.lbeg))) return r;
    LayoutLevel M;
    const DevAlloc alloc = [S](void** p, size_t b) { return solver_alloc(S, p, b); };
    if ((r = build_brick_map(D.bcoord, D.binst, H.nb, S->stream, alloc, M))) return r;
    D.map_keys = M.map_keys;
    D.map_vals = M.map_vals;
    D.map_cap = M.map_cap;
    return FL_OK;
}

// Buffers, per-level constants and the initial state; the layout tables are
// uploaded first unless the device builder made them.  remap_follows: a
// regrid's population transfer writes every fluid leaf next (bricks of fluid
// leaves only are then not initialised).
template <typename T>
fl_status upload_and_init(fl_lbm_solver* S, bool tables_on_device, bool remap_follows) {
    const size_t sb = sizeof(T);
    const size_t cb = compute_bytes(S->storage);
    const int NI = S->ninst();
    fl_status r;
    if ((r = dev_upload(S, &S->dinst, S->hinst))) return r;
    if ((r = dev_upload(S, &S->dlinst, S->linst))) return r;
    {
        void* a;
        if ((r = dev_alloc(S, &a, (size_t)NI * 4 * sizeof(double)))) return r;
        S->acc = static_cast<double*>(a);
    }
    for (int l = 0; l < S->L; ++l) {
        HostLevel& H = S->h[l];
        DevLevel& D = S->d[l];
        if (!tables_on_device && (r = upload_host_tables(S, l))) return r;
        std::vector<float> itf(NI);
        std::vector<double> itd(NI);
        for (int k = 0; k < NI; ++k) {
            itd[k] = 1.0 / tau_at(S->inst[k].p.tau0, l);
            itf[k] = (float)itd[k];
        }
        if ((r = dev_upload(S, &D.itau, itf))) return r;
        if ((r = dev_upload(S, &D.itau_d, itd))) return r;
        void* bv;
        const size_t nbv = std::max<size_t>((size_t)H.nactive * 4, H.n_links * 3);
        if ((r = dev_alloc(S, &bv, nbv * sizeof(double)))) return r;
        D.bvals = static_cast<double*>(bv);
        const size_t slots = (size_t)H.nb * 64;
        for (int k = 0; k < 2; ++k) {
            if ((r = dev_alloc(S, &D.buf[k], slots * Q * sb))) return r;
            if ((r = dev_alloc(S, &D.cap[k], cap_elems(H.n_cap) * cb))) return r;
        }
        D.cur = 0;
        D.ld.nbr = D.nbr;
        D.ld.bcoord = D.bcoord;
        D.ld.bflags = D.bflags;
        D.ld.cflags = D.cflags;
        D.ld.binst = D.binst;
        D.ld.itau = D.itau;
        D.ld.itau_d = D.itau_d;
        D.ld.inst = S->dinst;
        D.ld.S = slots;
        D.ld.nactive = (int32_t)H.nactive;
        D.ld.level = l;
        D.ld.tick = S->ticks + l;
        FL_HIP(launch_init<T>(D.ld, static_cast<T*>(D.buf[0]), static_cast<T*>(D.buf[1]), S->stream, remap_follows));
    }
    // which levels have pressure-outlet links: the collide runs its outlet
    // instances only there (the link words decide, whichever builder made them)
    uint32_t* flags = nullptr;
    if ((r = dev_alloc(S, reinterpret_cast<void**>(&flags), (size_t)S->L * sizeof(uint32_t)))) return r;
    FL_HIP(hipMemsetAsync(flags, 0, (size_t)S->L * sizeof(uint32_t), S->stream));
    for (int l = 0; l < S->L; ++l)
        FL_HIP(launch_any_outlet(S->d[l].cflags, S->d[l].bflags, (size_t)S->h[l].nactive * 64, flags + l, S->stream));
    std::vector<uint32_t> hf((size_t)S->L);
    FL_HIP(hipMemcpyAsync(hf.data(), flags, (size_t)S->L * sizeof(uint32_t), hipMemcpyDeviceToHost, S->stream));
    FL_HIP(hipStreamSynchronize(S->stream));
    for (int l = 0; l < S->L; ++l) S->d[l].ld.outlet = hf[(size_t)l] ? 1 : 0;
    // f2c staging: which coarse f2c ghosts the finer level's collide can
    // stage (children all active fluid leaves of one brick), per fine brick
    // and block
    for (int l = 0; l + 1 < S->L; ++l) {
        const HostLevel& H = S->h[l];
        DevLevel& D = S->d[l];
        DevLevel& F = S->d[l + 1];
        if (H.n_f2c == 0) continue;
        const size_t nbf = (size_t)S->h[l + 1].nb;
        if ((r = dev_alloc(S, &D.f2c_stage, (size_t)(4 + Q) * H.n_f2c * compute_bytes(S->storage)))) return r;
        if ((r = dev_alloc(S, reinterpret_cast<void**>(&D.f2c_staged), H.n_f2c))) return r;
        if ((r = dev_alloc(S, reinterpret_cast<void**>(&F.f2c_blk), nbf * 8 * sizeof(uint32_t)))) return r;
        FL_HIP(launch_stage_map(D.f2c_child, (uint32_t)H.n_f2c, F.cflags, F.bflags, F.f2c_blk, nbf, D.f2c_staged,
                                S->stream));
        F.ld.f2c_blk = F.f2c_blk;
        F.ld.stage = D.f2c_stage;
        F.ld.n_stage = (uint32_t)H.n_f2c;
        F.ld.stage_on = 0;
    }
    // capture staging: which of a level's captures its own collide can write
    for (int l = 0; l < S->L; ++l) {
        const HostLevel& H = S->h[l];
        DevLevel& D = S->d[l];
        if (H.n_cap == 0) continue;
        const size_t ns = (size_t)H.nb * 64;
        if ((r = dev_alloc(S, reinterpret_cast<void**>(&D.cap_of), ns * sizeof(uint32_t)))) return r;
        if ((r = dev_alloc(S, reinterpret_cast<void**>(&D.cap_staged), H.n_cap))) return r;
        FL_HIP(launch_cap_map(D.cap_slots, (uint32_t)H.n_cap, D.cflags, D.bflags, D.cap_of, ns, D.cap_staged,
                              S->stream));
        D.ld.cap_of = D.cap_of;
        D.ld.ncap = (uint32_t)H.n_cap;
        D.ld.cap_on = 0;
    }
    return FL_OK;
}

// stage_sub: this step is the second sub-step of the coarser level's step,
// whose collide stages the coarser level's f2c ghosts (LevelDev::stage).
template <typename T>
fl_status advance(fl_lbm_solver* S, int l, bool stage_sub) {
    DevLevel& D = S->d[l];
    const HostLevel& H = S->h[l];
    // levels are advanced while any deeper level holds fluid leaves (a mesh
    // may leave coarse levels empty, e.g. a uniformly refined one)
    bool fine = false;
    for (int k = l + 1; k < S->L; ++k) fine = fine || S->h[k].nfluid > 0;
    hipStream_t st = S->stream;
    const uint32_t ncap = (uint32_t)H.n_cap;
    if (fine) {
        DevLevel& F = S->d[l + 1];
        // The f2c fill gathers the children's pulled populations from the
        // input of the finer level's last collide, which (from the second
        // step of a pass on) staged their sums for the ghosts it could
        // (nothing writes that input in between): those ghosts read the sums.
        FL_HIP(launch_fill_f2c<T>(D.ld, D.f2c_ghost, D.f2c_child, (uint32_t)H.n_f2c, F.ld,
                                  static_cast<const T*>(F.buf[1 - F.cur]), static_cast<T*>(D.buf[D.cur]),
                                  D.stage_fresh ? D.f2c_stage : nullptr, D.f2c_staged, st));
        // The capture before the collide gathers from buf[1 - cur]: the
        // buffer the previous step's capture after its collide gathered from,
        // unchanged since (the f2c fill above writes buf[cur], the finer
        // levels write only their own buffers).  So within a pass only the
        // first step captures it; later steps take the previous capture.
        if (!D.cap_fresh)
            FL_HIP(launch_capture<T>(D.ld, static_cast<const T*>(D.buf[1 - D.cur]), D.cap_slots, ncap, D.cap[0],
                                     nullptr, st));
    }
    // staging: the collide writes the staged f2c sums of the coarser level
    // (second sub-step) and the staged captures that follow it; levels with
    // outlet links run the outlet instances, which have no staging code
    const bool cap_stage = fine && D.cap_of && !D.ld.outlet;
    LevelDev ld = D.ld;
    ld.stage_on = stage_sub && ld.stage && !ld.outlet ? 1 : 0;
    ld.cap_on = cap_stage ? 1 : 0;
    ld.cap_out = D.cap[1];
    FL_HIP(launch_collide<T>(ld, static_cast<const T*>(D.buf[D.cur]), static_cast<T*>(D.buf[1 - D.cur]),
                             S->variant[l], st));
    if (S->storage == FL_STORE_F16 || S->storage == FL_STORE_F16C) FL_HIP(launch_tick(S->ticks + l, st));
    D.cur ^= 1;
    if (fine) {
        FL_HIP(launch_capture<T>(D.ld, static_cast<const T*>(D.buf[1 - D.cur]), D.cap_slots, ncap, D.cap[1],
                                 cap_stage ? D.cap_staged : nullptr, st));
        DevLevel& F = S->d[l + 1];
        const HostLevel& HF = S->h[l + 1];
        for (int sub = 0; sub < 2; ++sub) {
            FL_HIP(launch_fill_c2f<T>(F.ld, F.c2f_gchild, F.c2f_gcap, F.c2f_gmask, (uint32_t)HF.n_groups, D.cap[0], D.cap[1],
                                      ncap, 0.5 * sub, static_cast<T*>(F.buf[F.cur]), st));
            fl_status r = advance<T>(S, l + 1, sub == 1);
            if (r) return r;
        }
        // this step's capture after the collide is the next step's before it
        std::swap(D.cap[0], D.cap[1]);
        D.cap_fresh = true;
        D.stage_fresh = D.f2c_stage != nullptr && !S->d[l + 1].ld.outlet;
    }
    return FL_OK;
}

// Start of a stepping pass (an eager run of steps or a graph capture): every
// level captures afresh at its first step, so a pass never depends on
// captures made before it (a graph replays the same way whatever ran
// between its launches).
void begin_pass(fl_lbm_solver* S) {
    for (int l = 0; l < S->L; ++l) {
        S->d[l].cap_fresh = false;
        S->d[l].stage_fresh = false;
    }
}

fl_status advance_any(fl_lbm_solver* S) {
    switch (S->storage) {
        case FL_STORE_F32: return advance<float>(S, 0, false);
        case FL_STORE_F16: return advance<_Float16>(S, 0, false);
        case FL_STORE_F16C: return advance<fp16c>(S, 0, false);
        default: return advance<double>(S, 0, false);
    }
}

// Two steps captured into a graph.  After two steps every level's buffer
// parity is back where it started; levels >= 1 flip an even number of times
// per step, so only level 0's parity selects the graph.
fl_status capture_graph(fl_lbm_solver* S, int k) {
    begin_pass(S);
    FL_HIP(hipStreamBeginCapture(S->stream, hipStreamCaptureModeRelaxed));
    fl_status r = advance_any(S);
    if (!r) r = advance_any(S);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(S->stream, &g);
    if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    FL_HIP(e);
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        FL_HIP(e);
    }
    S->graph[k] = g;
    S->gexec[k] = x;
    return FL_OK;
}

fl_status dispatch_step(fl_lbm_solver* S, int nsteps) {
    int t = 0;
    if (S->use_graph) {
        for (; t + 2 <= nsteps; t += 2) {
            const int k = S->d[0].cur;
            if (!S->gexec[k]) {
                fl_status r = capture_graph(S, k);
                if (r) return r;
            }
            FL_HIP(hipGraphLaunch(S->gexec[k], S->stream));
            ++S->graph_launches;
        }
    }
    if (t < nsteps) begin_pass(S);
    for (; t < nsteps; ++t) {
        fl_status r = advance_any(S);
        if (r) return r;
    }
    return FL_OK;
}

LevelMaps level_maps(const fl_lbm_solver* S) {
    LevelMaps m{};
    m.nlevels = S->L;
    for (int l = 0; l < S->L; ++l) {
        const DevLevel& D = S->d[l];
        m.m[l] = LevelMap{D.map_keys, D.map_vals, D.map_cap, D.cflags};
    }
    return m;
}

// slots of the query leaves of one instance (looked up on the device in the
// brick maps); -1 when not a fluid leaf
fl_status query_slots(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z,
                      const uint8_t* lv, size_t n, std::vector<int64_t>& slots) {
    slots.assign(n, -1);
    if (n == 0) return FL_OK;
    struct Buf {
        void* p = nullptr;
        ~Buf() { if (p) (void)hipFree(p); }
    } bx, by, bz, bl, bo;
    FL_HIP(hipMalloc(&bx.p, n * 4));
    FL_HIP(hipMalloc(&by.p, n * 4));
    FL_HIP(hipMalloc(&bz.p, n * 4));
    FL_HIP(hipMalloc(&bl.p, n));
    FL_HIP(hipMalloc(&bo.p, n * 8));
    FL_HIP(hipMemcpyAsync(bx.p, x, n * 4, hipMemcpyHostToDevice, S->stream));
    FL_HIP(hipMemcpyAsync(by.p, y, n * 4, hipMemcpyHostToDevice, S->stream));
    FL_HIP(hipMemcpyAsync(bz.p, z, n * 4, hipMemcpyHostToDevice, S->stream));
    FL_HIP(hipMemcpyAsync(bl.p, lv, n, hipMemcpyHostToDevice, S->stream));
    fl_status r = launch_query_slots(level_maps(S), S->dlinst, inst, static_cast<const int32_t*>(bx.p),
                                     static_cast<const int32_t*>(by.p), static_cast<const int32_t*>(bz.p),
                                     static_cast<const uint8_t*>(bl.p), n, static_cast<int64_t*>(bo.p), S->stream);
    if (r) return r;
    FL_HIP(hipMemcpyAsync(slots.data(), bo.p, n * 8, hipMemcpyDeviceToHost, S->stream));
    FL_HIP(hipStreamSynchronize(S->stream));
    return FL_OK;
}

fl_status ensure_query(fl_lbm_solver* S, size_t n) {
    if (n <= S->qcap) return FL_OK;
    if (S->qslots) (void)hipFree(S->qslots);
    if (S->qbuf) (void)hipFree(S->qbuf);
    S->qslots = nullptr;
    S->qbuf = nullptr;
    S->qcap = 0;
    FL_HIP(hipMalloc(reinterpret_cast<void**>(&S->qslots), n * sizeof(int64_t)));
    FL_HIP(hipMalloc(reinterpret_cast<void**>(&S->qbuf), n * Q * sizeof(double)));
    S->qcap = n;
    return FL_OK;
}

template <typename T>
fl_status get_f(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                size_t n, double* out) {
    std::vector<int64_t> slots;
    {
        fl_status qr = query_slots(S, inst, x, y, z, lv, n, slots);
        if (qr) return qr;
    }
    std::vector<double> tmp;
    for (size_t k = 0; k < n; ++k)
        if (lv[k] >= S->L)
            for (int i = 0; i < Q; ++i) out[k * Q + i] = std::nan("");
    for (int l = 0; l < S->L; ++l) {
        std::vector<int64_t> ls;
        std::vector<size_t> idx;
        for (size_t k = 0; k < n; ++k)
            if (lv[k] == l) { ls.push_back(slots[k]); idx.push_back(k); }
        if (ls.empty()) continue;
        fl_status r = ensure_query(S, ls.size());
        if (r) return r;
        FL_HIP(hipMemcpyAsync(S->qslots, ls.data(), ls.size() * sizeof(int64_t), hipMemcpyHostToDevice, S->stream));
        DevLevel& D = S->d[l];
        FL_HIP(launch_extract<T>(static_cast<const T*>(D.buf[D.cur]), D.ld.S, S->qslots, ls.size(), S->qbuf,
                                 S->stream));
        tmp.resize(ls.size() * Q);
        FL_HIP(hipMemcpyAsync(tmp.data(), S->qbuf, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, S->stream));
        FL_HIP(hipStreamSynchronize(S->stream));
        for (size_t j = 0; j < ls.size(); ++j)
            for (int i = 0; i < Q; ++i)
                out[idx[j] * Q + i] = ls[j] < 0 ? std::nan("") : tmp[j * Q + i];
    }
    return FL_OK;
}

template <typename T>
fl_status set_f(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                size_t n, const double* in) {
    std::vector<int64_t> slots;
    {
        fl_status qr = query_slots(S, inst, x, y, z, lv, n, slots);
        if (qr) return qr;
    }
    for (int l = 0; l < S->L; ++l) {
        std::vector<int64_t> ls;
        std::vector<double> vals;
        for (size_t k = 0; k < n; ++k)
            if (lv[k] == l) {
                ls.push_back(slots[k]);
                vals.insert(vals.end(), in + k * Q, in + (k + 1) * Q);
            }
        if (ls.empty()) continue;
        fl_status r = ensure_query(S, ls.size());
        if (r) return r;
        FL_HIP(hipMemcpyAsync(S->qslots, ls.data(), ls.size() * sizeof(int64_t), hipMemcpyHostToDevice, S->stream));
        FL_HIP(hipMemcpyAsync(S->qbuf, vals.data(), vals.size() * sizeof(double), hipMemcpyHostToDevice, S->stream));
        DevLevel& D = S->d[l];
        FL_HIP(launch_inject<T>(static_cast<T*>(D.buf[0]), static_cast<T*>(D.buf[1]), D.ld.S, S->qslots, ls.size(),
                                S->qbuf, S->stream));
        FL_HIP(hipStreamSynchronize(S->stream));
    }
    return FL_OK;
}

// Per-instance FP64 (mass, jx, jy, jz): per-brick wave sums, then one
// workgroup per instance over its brick range, accumulated over levels.
template <typename T>
fl_status totals(fl_lbm_solver* S, double* out4) {
    const int NI = S->ninst();
    bool first = true;
    for (int l = 0; l < S->L; ++l) {
        DevLevel& D = S->d[l];
        if (D.ld.nactive == 0) continue;
        FL_HIP(launch_brick_sums<T>(D.ld, static_cast<const T*>(D.buf[D.cur]), std::ldexp(1.0, -3 * l), D.bvals,
                                    S->stream));
        FL_HIP(launch_inst_reduce(D.bvals, 4, D.abeg, NI, 0, first ? 1 : 0, S->acc, S->stream));
        first = false;
    }
    if (first) {
        for (int k = 0; k < 4 * NI; ++k) out4[k] = 0.0;
        return FL_OK;
    }
    FL_HIP(hipMemcpyAsync(out4, S->acc, (size_t)NI * 4 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    FL_HIP(hipStreamSynchronize(S->stream));
    return FL_OK;
}

template <typename T>
fl_status force(fl_lbm_solver* S, double* out3) {
    const int NI = S->ninst();
    bool first = true;
    for (int l = 0; l < S->L; ++l) {
        DevLevel& D = S->d[l];
        const uint32_t nl = (uint32_t)S->h[l].n_links;
        if (nl == 0) continue;
        FL_HIP(launch_link_force<T>(D.force_links, nl, D.ld.S, static_cast<const T*>(D.buf[D.cur]),
                                    std::ldexp(1.0, -2 * l), D.bvals, S->stream));
        FL_HIP(launch_inst_reduce(D.bvals, 3, D.lbeg, NI, 0, first ? 1 : 0, S->acc, S->stream));
        first = false;
    }
    if (first) {
        for (int k = 0; k < 3 * NI; ++k) out3[k] = 0.0;
        return FL_OK;
    }
    FL_HIP(hipMemcpyAsync(out3, S->acc, (size_t)NI * 3 * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    FL_HIP(hipStreamSynchronize(S->stream));
    return FL_OK;
}

// The window's velocity accumulator of a level, allocated and cleared on
// first use (a fill kernel, not a memset).
fl_status level_uacc(fl_lbm_solver* S, DevLevel& D) {
    if (D.uacc) return FL_OK;
    void* a;
    const size_t n = (size_t)D.ld.nactive * 64 * 3;
    fl_status r = dev_alloc(S, &a, n * sizeof(float));
    if (r) return r;
    D.uacc = static_cast<float*>(a);
    FL_HIP(launch_fill_f32(D.uacc, n, 0.0f, S->stream));
    return FL_OK;
}

template <typename T>
fl_status velocity_sample(fl_lbm_solver* S) {
    for (int l = 0; l < S->L; ++l) {
        DevLevel& D = S->d[l];
        if (D.ld.nactive == 0) continue;
        fl_status r = level_uacc(S, D);
        if (r) return r;
        FL_HIP(launch_velocity_sample<T>(D.ld, static_cast<const T*>(D.buf[D.cur]), D.uacc, S->stream));
    }
    ++S->nsamples;
    return FL_OK;
}

template <typename T>
fl_status velocity_change(fl_lbm_solver* S, double* du) {
    const int NI = S->ninst();
    bool fresh = false, first = true;
    const float inv_m = S->nsamples > 0 ? 1.0f / (float)S->nsamples : 0.0f;
    for (int l = 0; l < S->L; ++l) {
        DevLevel& D = S->d[l];
        if (D.ld.nactive == 0) continue;
        if (!D.uprev) {
            void* a;
            fl_status r = dev_alloc(S, &a, (size_t)D.ld.nactive * 64 * 3 * sizeof(float));
            if (r) return r;
            D.uprev = static_cast<float*>(a);
            fresh = true;
        }
        if (inv_m > 0.0f) {
            fl_status r = level_uacc(S, D);
            if (r) return r;
        }
        FL_HIP(launch_brick_du<T>(D.ld, static_cast<const T*>(D.buf[D.cur]), D.uprev, D.uacc, inv_m, D.bvals,
                                  S->stream));
        FL_HIP(launch_inst_reduce(D.bvals, 1, D.abeg, NI, 1, first ? 1 : 0, S->acc, S->stream));
        first = false;
    }
    S->nsamples = 0;
    if (first) {
        for (int k = 0; k < NI; ++k) du[k] = 0.0;
        return FL_OK;
    }
    FL_HIP(hipMemcpyAsync(du, S->acc, (size_t)NI * sizeof(double), hipMemcpyDeviceToHost, S->stream));
    FL_HIP(hipStreamSynchronize(S->stream));
    if (fresh)
        for (int k = 0; k < NI; ++k) du[k] = HUGE_VAL;
    return FL_OK;
}

fl_status check_params(const fl_lbm_params* p) {
    FL_REQUIRE(p->max_level >= 0 && p->max_level < 20, "max_level must be in 0..19");
    FL_REQUIRE(p->nx0 > 0 && p->ny0 > 0 && p->nz0 > 0 && p->nx0 % 4 == 0 && p->ny0 % 4 == 0 && p->nz0 % 4 == 0,
               "level-0 extents must be positive multiples of 4");
    FL_REQUIRE(((int64_t)std::max(p->nx0, std::max(p->ny0, p->nz0)) << p->max_level) <= (int64_t)1 << 18,
               "finest-level extent must not exceed 2^18 cells");
    FL_REQUIRE(p->tau0 > 0.5, "tau0 must exceed 1/2");
    FL_REQUIRE(p->storage >= 0 && p->storage <= 3, "unknown storage type");
    for (int f = 0; f < 6; ++f) FL_REQUIRE(p->bc[f] >= 0 && p->bc[f] <= 2, "unknown boundary type");
    for (int a = 0; a < 3; ++a)
        FL_REQUIRE((p->bc[2 * a] == FL_BC_PERIODIC) == (p->bc[2 * a + 1] == FL_BC_PERIODIC),
                   "periodic faces must come in pairs");
    return FL_OK;
}

InstDev make_inst_dev(const fl_lbm_params& p) {
    InstDev I{};
    for (int f = 0; f < 6; ++f) {
        I.bc_type[f] = p.bc[f];
        for (int a = 0; a < 3; ++a) {
            I.bc_u[f][a] = (float)p.bc_u[f][a];
            I.bc_ud[f][a] = p.bc_u[f][a];
        }
    }
    I.rho0 = p.rho0;
    I.tau0 = p.tau0;
    I.rho_init = p.rho_init;
    for (int a = 0; a < 3; ++a) I.u_init[a] = p.u_init[a];
    I.ext0[0] = p.nx0;
    I.ext0[1] = p.ny0;
    I.ext0[2] = p.nz0;
    I.live = 1;
    return I;
}

// Instances of a batch: parameter checks and the per-instance records.
fl_status init_instances(const fl_lbm_params* params, int ninst, const size_t* offsets,
                         std::unique_ptr<fl_lbm_solver>& out) {
    FL_REQUIRE(params && offsets, "null pointer");
    FL_REQUIRE(ninst >= 1 && ninst <= 65535, "instance count must be in 1..65535");
    FL_REQUIRE(offsets[0] == 0, "offsets[0] must be 0");
    std::unique_ptr<fl_lbm_solver> S(new fl_lbm_solver());
    S->storage = params[0].storage;
    int maxl = 0;
    for (int k = 0; k < ninst; ++k) {
        fl_status r = check_params(&params[k]);
        if (r) return r;
        FL_REQUIRE(params[k].storage == S->storage, "every instance of a batch must use the same storage");
        FL_REQUIRE(offsets[k + 1] >= offsets[k], "offsets must be non-decreasing");
        InstHost ih;
        ih.p = params[k];
        ih.first = offsets[k];
        ih.count = offsets[k + 1] - offsets[k];
        S->inst.push_back(ih);
        S->hinst.push_back(make_inst_dev(params[k]));
        LayoutInst li{};
        li.ext0[0] = params[k].nx0;
        li.ext0[1] = params[k].ny0;
        li.ext0[2] = params[k].nz0;
        li.max_level = params[k].max_level;
        for (int f = 0; f < 6; ++f) li.bc[f] = params[k].bc[f];
        S->linst.push_back(li);
        maxl = std::max(maxl, params[k].max_level);
    }
    S->L = maxl + 1;
    out = std::move(S);
    return FL_OK;
}

// Host half of creation with the host builder: validation and the layout
// (no device work).  fl_lbm_plan_bytes uses it; so does creation when the
// host builder is selected (fl_lbm_layout_on_host).
fl_status plan_solver(const fl_lbm_params* params, int ninst, const size_t* offsets, const int32_t* h_x,
                      const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level, const uint8_t* h_state,
                      std::unique_ptr<fl_lbm_solver>& out) {
    std::unique_ptr<fl_lbm_solver> S;
    fl_status r = init_instances(params, ninst, offsets, S);
    if (r) return r;
    const size_t n = offsets[ninst];
    FL_REQUIRE(n == 0 || (h_x && h_y && h_z && h_level && h_state), "null leaf arrays");
    r = build_layout(S.get(), h_x, h_y, h_z, h_level, h_state, n);
    if (r) return r;
    out = std::move(S);
    return FL_OK;
}

// The device builder: every table made from device leaf arrays; only sizes
// and per-instance counts come back to the host.
fl_status device_layout(fl_lbm_solver* S, const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                        const uint8_t* d_level, const uint8_t* d_state, size_t n) {
    std::vector<size_t> offs(S->ninst() + 1, 0);
    for (int k = 0; k < S->ninst(); ++k) offs[k + 1] = S->inst[k].first + S->inst[k].count;
    std::vector<LayoutLevel> lv(S->L);
    std::string err;
    const DevAlloc alloc = [S](void** p, size_t b) { return solver_alloc(S, p, b); };
    fl_status r = build_layout_device(S->linst, S->L, offs.data(), d_x, d_y, d_z, d_level, d_state, n, S->stream,
                                      alloc, lv.data(), &err);
    if (r == FL_E_MESH) return mesh_error(err);
    if (r) return r;
    for (int l = 0; l < S->L; ++l) {
        const LayoutLevel& Lv = lv[l];
        HostLevel& H = S->h[l];
        DevLevel& D = S->d[l];
        H.level = l;
        H.nb = Lv.nb;
        H.nactive = Lv.nactive;
        H.nfluid = Lv.nfluid;
        H.nfluid_inst = Lv.nfluid_inst;
        H.n_c2f = Lv.n_c2f;
        H.n_groups = Lv.n_groups;
        H.n_cap = Lv.n_cap;
        H.n_f2c = Lv.n_f2c;
        H.n_links = Lv.n_links;
        H.abeg = Lv.abeg_h;
        H.lbeg = Lv.lbeg_h;
        D.nbr = Lv.nbr;
        D.bcoord = Lv.bcoord;
        D.binst = Lv.binst;
        D.bflags = Lv.bflags;
        D.cflags = Lv.cflags;
        D.c2f_gchild = Lv.c2f_gchild;
        D.c2f_gcap = Lv.c2f_gcap;
        D.c2f_gmask = Lv.c2f_gmask;
        D.cap_slots = Lv.cap_slots;
        D.f2c_ghost = Lv.f2c_ghost;
        D.f2c_child = Lv.f2c_child;
        D.force_links = Lv.force_links;
        D.abeg = Lv.abeg;
        D.lbeg = Lv.lbeg;
        D.map_keys = Lv.map_keys;
        D.map_vals = Lv.map_vals;
        D.map_cap = Lv.map_cap;
    }
    return FL_OK;
}

bool g_host_layout = false;  // fl_lbm_layout_on_host

fl_status attach_stream(fl_lbm_solver* S, hipStream_t stream) {
    if (stream) {
        S->stream = stream;
    } else {
        // an own stream so that the step can be graph-captured (the null
        // stream cannot be); it synchronises with the null stream
        FL_HIP(hipStreamCreate(&S->stream));
        S->own_stream = true;
    }
    return FL_OK;
}

// Stream and rounding clock of a new solver: its own, or for a regrid the
// previous layout's (with its instances' live flags).
fl_status attach(fl_lbm_solver* S, hipStream_t stream, const fl_lbm_solver* prev) {
    if (prev) {
        S->stream = prev->stream;
        S->ticks = prev->ticks;
        S->hinst = prev->hinst;
        return FL_OK;
    }
    fl_status r = attach_stream(S, stream);
    if (r) return r;
    FL_HIP(hipMalloc(reinterpret_cast<void**>(&S->ticks), 24 * sizeof(uint32_t)));
    S->own_ticks = true;
    FL_HIP(hipMemsetAsync(S->ticks, 0, 24 * sizeof(uint32_t), S->stream));
    return FL_OK;
}

fl_status init_any(fl_lbm_solver* S, bool tables_on_device, bool remap_follows) {
    switch (S->storage) {
        case FL_STORE_F32: return upload_and_init<float>(S, tables_on_device, remap_follows);
        case FL_STORE_F16: return upload_and_init<_Float16>(S, tables_on_device, remap_follows);
        case FL_STORE_F16C: return upload_and_init<fp16c>(S, tables_on_device, remap_follows);
        default: return upload_and_init<double>(S, tables_on_device, remap_follows);
    }
}

// A solver (layout, buffers, initial state) for leaves on the device.
fl_status build_on_device(const fl_lbm_params* params, int ninst, const size_t* offsets, const int32_t* d_x,
                          const int32_t* d_y, const int32_t* d_z, const uint8_t* d_level, const uint8_t* d_state,
                          hipStream_t stream, const fl_lbm_solver* prev, std::unique_ptr<fl_lbm_solver>& out) {
    std::unique_ptr<fl_lbm_solver> S;
    fl_status r = init_instances(params, ninst, offsets, S);
    if (r) return r;
    const size_t n = offsets[ninst];
    FL_REQUIRE(n == 0 || (d_x && d_y && d_z && d_level && d_state), "null leaf arrays");
    if ((r = attach(S.get(), stream, prev))) return r;
    if ((r = device_layout(S.get(), d_x, d_y, d_z, d_level, d_state, n))) return r;
    if ((r = init_any(S.get(), true, prev != nullptr))) return r;
    out = std::move(S);
    return FL_OK;
}

// The same with the host builder (the reference the device layout is tested
// against; leaves on the host).
fl_status build_on_host(const fl_lbm_params* params, int ninst, const size_t* offsets, const int32_t* h_x,
                        const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level, const uint8_t* h_state,
                        hipStream_t stream, const fl_lbm_solver* prev, std::unique_ptr<fl_lbm_solver>& out) {
    std::unique_ptr<fl_lbm_solver> S;
    fl_status r = plan_solver(params, ninst, offsets, h_x, h_y, h_z, h_level, h_state, S);
    if (r) return r;
    if ((r = attach(S.get(), stream, prev))) return r;
    if ((r = init_any(S.get(), false, prev != nullptr))) return r;
    out = std::move(S);
    return FL_OK;
}

// Leaves given on the host: the host builder when selected, otherwise one
// upload and the device builder.
fl_status build_from_host(const fl_lbm_params* params, int ninst, const size_t* offsets, const int32_t* h_x,
                          const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level, const uint8_t* h_state,
                          hipStream_t stream, const fl_lbm_solver* prev, std::unique_ptr<fl_lbm_solver>& out) {
    FL_REQUIRE(params && offsets && ninst >= 1, "null pointer");
    if (g_host_layout)
        return build_on_host(params, ninst, offsets, h_x, h_y, h_z, h_level, h_state, stream, prev, out);
    const size_t n = offsets[ninst];
    FL_REQUIRE(n == 0 || (h_x && h_y && h_z && h_level && h_state), "null leaf arrays");
    struct Buf {
        void* p = nullptr;
        ~Buf() { if (p) (void)hipFree(p); }
    } bx, by, bz, bl, bs;
    const size_t m = n ? n : 1;
    FL_HIP(hipMalloc(&bx.p, m * 4));
    FL_HIP(hipMalloc(&by.p, m * 4));
    FL_HIP(hipMalloc(&bz.p, m * 4));
    FL_HIP(hipMalloc(&bl.p, m));
    FL_HIP(hipMalloc(&bs.p, m));
    if (n) {
        FL_HIP(hipMemcpy(bx.p, h_x, n * 4, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(by.p, h_y, n * 4, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(bz.p, h_z, n * 4, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(bl.p, h_level, n, hipMemcpyHostToDevice));
        FL_HIP(hipMemcpy(bs.p, h_state, n, hipMemcpyHostToDevice));
    }
    fl_status r = build_on_device(params, ninst, offsets, static_cast<const int32_t*>(bx.p),
                                  static_cast<const int32_t*>(by.p), static_cast<const int32_t*>(bz.p),
                                  static_cast<const uint8_t*>(bl.p), static_cast<const uint8_t*>(bs.p), stream, prev,
                                  out);
    if (out && out->stream) (void)hipStreamSynchronize(out->stream);
    return r;
}

template <typename T>
fl_status vorticity(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z,
                    const uint8_t* lv, size_t n, double* out) {
    std::vector<int64_t> slots;
    {
        fl_status qr = query_slots(S, inst, x, y, z, lv, n, slots);
        if (qr) return qr;
    }
    for (size_t k = 0; k < n; ++k) out[k] = 0.0;
    std::vector<double> tmp;
    for (int l = 0; l < S->L; ++l) {
        std::vector<int64_t> ls;
        std::vector<size_t> idx;
        for (size_t k = 0; k < n; ++k)
            if (lv[k] == l) { ls.push_back(slots[k]); idx.push_back(k); }
        if (ls.empty()) continue;
        fl_status r = ensure_query(S, ls.size());
        if (r) return r;
        FL_HIP(hipMemcpyAsync(S->qslots, ls.data(), ls.size() * 8, hipMemcpyHostToDevice, S->stream));
        DevLevel& D = S->d[l];
        FL_HIP(launch_vorticity<T>(D.ld, static_cast<const T*>(D.buf[D.cur]), S->qslots, ls.size(),
                                   std::ldexp(1.0, l), S->qbuf, S->stream));
        tmp.resize(ls.size());
        FL_HIP(hipMemcpyAsync(tmp.data(), S->qbuf, tmp.size() * 8, hipMemcpyDeviceToHost, S->stream));
        FL_HIP(hipStreamSynchronize(S->stream));
        for (size_t j = 0; j < ls.size(); ++j) out[idx[j]] = tmp[j];
    }
    return FL_OK;
}

// Population transfer from the old solver O into the freshly built N (same
// instances, new leaves).  The transfer lists come from the device
// (build_remap_lists: lookups in O's brick maps).
template <typename T>
fl_status remap(fl_lbm_solver* O, fl_lbm_solver* N) {
    std::vector<void*> tmp_allocs;
    fl_status st = FL_OK;
    const DevAlloc talloc = [&tmp_allocs, N](void** p, size_t b) -> fl_status {
        *p = nullptr;
        if (b == 0) return FL_OK;
        FL_TRY(fl::cache_alloc(p, b, N->stream));
        tmp_allocs.push_back(*p);
        return FL_OK;
    };
    const LevelMaps om = level_maps(O);
    for (int l = 0; l < N->L && st == FL_OK; ++l) {
        const HostLevel& H = N->h[l];
        DevLevel& D = N->d[l];
        LayoutLevel nl;
        nl.nactive = H.nactive;
        nl.bcoord = D.bcoord;
        nl.binst = D.binst;
        nl.cflags = D.cflags;
        RemapLists R;
        st = build_remap_lists(om, N->dlinst, l, nl, N->stream, talloc, R);
        if (st) break;
        RemapArgs<T> a;
        a.Ln = D.ld;
        a.A = static_cast<T*>(D.buf[0]);
        a.B = static_cast<T*>(D.buf[1]);
        if (l < O->L) {
            const DevLevel& OD = O->d[l];
            a.copy_src_buf = static_cast<const T*>(OD.buf[OD.cur]);
            a.So = OD.ld.S;
        }
        a.copy_dst = R.copy_dst;
        a.copy_src = R.copy_src;
        a.ncopy = R.ncopy;
        if (l > 0 && l - 1 < O->L) {
            const DevLevel& OC = O->d[l - 1];
            a.Lc = OC.ld;
            a.split_prev = static_cast<const T*>(OC.buf[1 - OC.cur]);
        }
        a.split_dst = R.split_dst;
        a.split_src = R.split_src;
        a.nsplit = R.nsplit;
        if (l + 1 < O->L) {
            const DevLevel& OF = O->d[l + 1];
            a.Lf = OF.ld;
            a.merge_prev = static_cast<const T*>(OF.buf[1 - OF.cur]);
        }
        a.merge_dst = R.merge_dst;
        a.merge_src8 = R.merge_src8;
        a.nmerge = R.nmerge;
        a.rest_dst = R.rest_dst;
        a.nrest = R.nrest;
        hipError_t e = launch_remap<T>(a, N->stream);
        if (e != hipSuccess) st = fl::hip_fail(e, "launch_remap", __FILE__, __LINE__);
        D.cur = 0;
    }
    for (int l = 0; l < N->L && st == FL_OK; ++l) {
        const HostLevel& H = N->h[l];
        DevLevel& D = N->d[l];
        const T* Ac = l > 0 ? static_cast<const T*>(N->d[l - 1].buf[0]) : nullptr;
        const size_t Sc = l > 0 ? N->d[l - 1].ld.S : 0;
        const uint32_t* capsl = l > 0 ? N->d[l - 1].cap_slots : nullptr;
        const T* Af = l + 1 < N->L ? static_cast<const T*>(N->d[l + 1].buf[0]) : nullptr;
        const size_t Sf = l + 1 < N->L ? N->d[l + 1].ld.S : 0;
        hipError_t e = launch_ghost_eq<T>(D.c2f_gchild, D.c2f_gcap, capsl, (uint32_t)H.n_groups, Ac, Sc,
                                          D.f2c_ghost, D.f2c_child, (uint32_t)H.n_f2c, Af, Sf,
                                          static_cast<T*>(D.buf[0]), static_cast<T*>(D.buf[1]), D.ld.S, N->stream);
        if (e != hipSuccess) st = fl::hip_fail(e, "launch_ghost_eq", __FILE__, __LINE__);
    }
    hipError_t e = hipStreamSynchronize(N->stream);
    for (void* p : tmp_allocs) fl::cache_free(p, N->stream);
    if (st) return st;
    if (e != hipSuccess) return fl::hip_fail(e, "remap", __FILE__, __LINE__);
    return FL_OK;
}

// Back-to-back launches of one level's collide-stream, timed with HIP
// events on the solver's stream.  Every launch reads the current buffer and
// writes the other one; the other buffer (which the next step's captures and
// fills read) is saved first and restored after, so the solver state is
// unchanged.
template <typename T>
fl_status time_collide(fl_lbm_solver* S, int level, int variant, int iters, float* ms) {
    DevLevel& D = S->d[level];
    const size_t bytes = D.ld.S * Q * sizeof(T);
    void* save = nullptr;
    FL_HIP(hipMalloc(&save, bytes));
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipMemcpyAsync(save, D.buf[1 - D.cur], bytes, hipMemcpyDeviceToDevice, S->stream);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventRecord(a, S->stream);
    for (int it = 0; it < iters && e == hipSuccess; ++it)
        e = launch_collide<T>(D.ld, static_cast<const T*>(D.buf[D.cur]), static_cast<T*>(D.buf[1 - D.cur]), variant,
                              S->stream);
    if (e == hipSuccess) e = hipEventRecord(b, S->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(D.buf[1 - D.cur], save, bytes, hipMemcpyDeviceToDevice, S->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(S->stream);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(save);
    FL_HIP(e);
    *ms = t / iters;
    return FL_OK;
}

fl_status time_any(fl_lbm_solver* S, int level, int variant, int iters, float* ms) {
    switch (S->storage) {
        case FL_STORE_F32: return time_collide<float>(S, level, variant, iters, ms);
        case FL_STORE_F16: return time_collide<_Float16>(S, level, variant, iters, ms);
        case FL_STORE_F16C: return time_collide<fp16c>(S, level, variant, iters, ms);
        default: return time_collide<double>(S, level, variant, iters, ms);
    }
}

}  // namespace
}  // namespace lbm
}  // namespace fl

using namespace fl;
using namespace fl::lbm;

#define FL_SOLVER_INST(S, inst) \
    FL_REQUIRE(S, "null solver"); \
    FL_REQUIRE((inst) >= 0 && (inst) < (S)->ninst(), "instance index out of range")

namespace fl {
namespace lbm {
// Populations of S carried over to the new layout N, which then replaces S's.
fl_status adopt_layout(fl_lbm_solver* S, std::unique_ptr<fl_lbm_solver>& N) {
    fl_status r;
    switch (S->storage) {
        case FL_STORE_F32: r = remap<float>(S, N.get()); break;
        case FL_STORE_F16: r = remap<_Float16>(S, N.get()); break;
        case FL_STORE_F16C: r = remap<fp16c>(S, N.get()); break;
        default: r = remap<double>(S, N.get()); break;
    }
    if (r) return r;
    // the handle keeps its identity, stream, variants and graph setting; its
    // layout becomes the new mesh (the old one is freed with N)
    S->drop_graphs();
    std::swap(S->inst, N->inst);
    std::swap(S->L, N->L);
    for (int l = 0; l < 24; ++l) {
        std::swap(S->h[l], N->h[l]);
        std::swap(S->d[l], N->d[l]);
    }
    std::swap(S->dinst, N->dinst);
    std::swap(S->dlinst, N->dlinst);
    std::swap(S->acc, N->acc);
    std::swap(S->allocs, N->allocs);
    std::swap(S->device_bytes, N->device_bytes);
    // the velocity snapshots and window sums belonged to the old layout
    S->nsamples = 0;
    return FL_OK;
}
}  // namespace lbm
}  // namespace fl

extern "C" {

fl_status fl_lbm_create(const fl_lbm_params* p, const int32_t* h_x, const int32_t* h_y, const int32_t* h_z,
                        const uint8_t* h_level, const uint8_t* h_state, size_t n, fl_stream s,
                        fl_lbm_solver** out) {
    FL_REQUIRE(p && out, "null pointer");
    const size_t off[2] = {0, n};
    return fl_lbm_create_batch(p, 1, off, h_x, h_y, h_z, h_level, h_state, s, out);
}

fl_status fl_lbm_create_batch(const fl_lbm_params* params, int ninst, const size_t* offsets, const int32_t* h_x,
                              const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level,
                              const uint8_t* h_state, fl_stream s, fl_lbm_solver** out) {
    FL_REQUIRE(params && offsets && out, "null pointer");
    *out = nullptr;
    std::unique_ptr<fl_lbm_solver> S;
    fl_status r = build_from_host(params, ninst, offsets, h_x, h_y, h_z, h_level, h_state, FL_STREAM(s), nullptr, S);
    if (r) return r;
    *out = S.release();
    return FL_OK;
}

fl_status fl_lbm_create_batch_device(const fl_lbm_params* params, int ninst, const size_t* offsets,
                                     const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                                     const uint8_t* d_level, const uint8_t* d_state, fl_stream s,
                                     fl_lbm_solver** out) {
    FL_REQUIRE(params && offsets && out, "null pointer");
    *out = nullptr;
    std::unique_ptr<fl_lbm_solver> S;
    fl_status r = build_on_device(params, ninst, offsets, d_x, d_y, d_z, d_level, d_state, FL_STREAM(s), nullptr, S);
    if (r) return r;
    *out = S.release();
    return FL_OK;
}

fl_status fl_lbm_layout_on_host(int on) {
    g_host_layout = on != 0;
    return FL_OK;
}

fl_status fl_lbm_layout_get(const fl_lbm_solver* S, int level, fl_lbm_layout* out) {
    FL_REQUIRE(S && out, "null pointer");
    FL_REQUIRE(level >= 0 && level < S->L, "level out of range");
    const DevLevel& D = S->d[level];
    const HostLevel& H = S->h[level];
    out->nbr = D.nbr;
    out->bcoord = D.bcoord;
    out->binst = D.binst;
    out->bflags = D.bflags;
    out->cflags = D.cflags;
    out->c2f_gchild = D.c2f_gchild;
    out->c2f_gcap = D.c2f_gcap;
    out->c2f_gmask = D.c2f_gmask;
    out->cap_slots = D.cap_slots;
    out->f2c_ghost = D.f2c_ghost;
    out->f2c_child = D.f2c_child;
    out->force_links = D.force_links;
    out->abeg = D.abeg;
    out->lbeg = D.lbeg;
    out->nb = H.nb;
    out->nactive = H.nactive;
    out->n_groups = H.n_groups;
    out->n_cap = H.n_cap;
    out->n_f2c = H.n_f2c;
    out->n_links = H.n_links;
    return FL_OK;
}

fl_status fl_lbm_plan_bytes(const fl_lbm_params* params, int ninst, const size_t* offsets, const int32_t* h_x,
                            const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level, const uint8_t* h_state,
                            uint64_t* device_bytes) {
    FL_REQUIRE(device_bytes, "null pointer");
    std::unique_ptr<fl_lbm_solver> S;
    fl_status r = plan_solver(params, ninst, offsets, h_x, h_y, h_z, h_level, h_state, S);
    if (r) return r;
    *device_bytes = layout_bytes(S.get());
    return FL_OK;
}

fl_status fl_lbm_destroy(fl_lbm_solver* solver) {
    if (!solver) return FL_OK;
    (void)hipStreamSynchronize(solver->stream);
    delete solver;
    return FL_OK;
}

fl_status fl_lbm_step(fl_lbm_solver* solver, int nsteps) {
    FL_REQUIRE(solver, "null solver");
    FL_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    return dispatch_step(solver, nsteps);
}

fl_status fl_lbm_num_instances(const fl_lbm_solver* S, int* n) {
    FL_REQUIRE(S && n, "null pointer");
    *n = S->ninst();
    return FL_OK;
}

fl_status fl_lbm_stats_get(const fl_lbm_solver* S, fl_lbm_stats* out) {
    FL_REQUIRE(S && out, "null pointer");
    std::memset(out, 0, sizeof(*out));
    out->nlevels = S->L;
    for (int l = 0; l < S->L; ++l) {
        const HostLevel& H = S->h[l];
        out->leaves[l] = H.nfluid;
        out->bricks[l] = H.nb;
        out->active_bricks[l] = H.nactive;
        out->c2f[l] = H.n_c2f;
        out->f2c[l] = H.n_f2c;
        out->capture[l] = H.n_cap;
        out->updates_per_step += H.nfluid << l;
        out->variant[l] = S->variant[l];
    }
    out->device_bytes = S->device_bytes;
    out->ninst = S->ninst();
    out->graph = S->use_graph ? 1 : 0;
    out->graph_launches = S->graph_launches;
    return FL_OK;
}

fl_status fl_lbm_inst_leaves(const fl_lbm_solver* S, int inst, uint64_t* leaves, uint64_t* updates_per_step) {
    FL_SOLVER_INST(S, inst);
    FL_REQUIRE(leaves && updates_per_step, "null pointer");
    *updates_per_step = 0;
    for (int l = 0; l < 24; ++l) leaves[l] = 0;
    for (int l = 0; l < S->L; ++l) {
        leaves[l] = S->h[l].nfluid_inst[inst];
        *updates_per_step += leaves[l] << l;
    }
    return FL_OK;
}

fl_status fl_lbm_get_f_inst(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z,
                            const uint8_t* lv, size_t n, double* f) {
    FL_SOLVER_INST(S, inst);
    FL_REQUIRE(n == 0 || (x && y && z && lv && f), "null pointer");
    switch (S->storage) {
        case FL_STORE_F32: return get_f<float>(S, inst, x, y, z, lv, n, f);
        case FL_STORE_F16: return get_f<_Float16>(S, inst, x, y, z, lv, n, f);
        case FL_STORE_F16C: return get_f<fp16c>(S, inst, x, y, z, lv, n, f);
        default: return get_f<double>(S, inst, x, y, z, lv, n, f);
    }
}

fl_status fl_lbm_get_f(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                       size_t n, double* f) {
    return fl_lbm_get_f_inst(S, 0, x, y, z, lv, n, f);
}

fl_status fl_lbm_set_f_inst(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z,
                            const uint8_t* lv, size_t n, const double* f) {
    FL_SOLVER_INST(S, inst);
    FL_REQUIRE(n == 0 || (x && y && z && lv && f), "null pointer");
    switch (S->storage) {
        case FL_STORE_F32: return set_f<float>(S, inst, x, y, z, lv, n, f);
        case FL_STORE_F16: return set_f<_Float16>(S, inst, x, y, z, lv, n, f);
        case FL_STORE_F16C: return set_f<fp16c>(S, inst, x, y, z, lv, n, f);
        default: return set_f<double>(S, inst, x, y, z, lv, n, f);
    }
}

fl_status fl_lbm_set_f(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                       size_t n, const double* f) {
    return fl_lbm_set_f_inst(S, 0, x, y, z, lv, n, f);
}

fl_status fl_lbm_moments_inst(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z,
                              const uint8_t* lv, size_t n, double* rho, double* u) {
    FL_REQUIRE(S && (n == 0 || (rho && u)), "null pointer");
    std::vector<double> f(n * Q);
    fl_status r = fl_lbm_get_f_inst(S, inst, x, y, z, lv, n, f.data());
    if (r) return r;
    for (size_t k = 0; k < n; ++k) {
        double rr = 0, jx = 0, jy = 0, jz = 0;
        for (int i = 0; i < Q; ++i) {
            double v = f[k * Q + i];
            rr += v;
            jx += cx(i) * v;
            jy += cy(i) * v;
            jz += cz(i) * v;
        }
        rho[k] = rr;
        u[3 * k] = jx / rr;
        u[3 * k + 1] = jy / rr;
        u[3 * k + 2] = jz / rr;
    }
    return FL_OK;
}

fl_status fl_lbm_moments(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                         size_t n, double* rho, double* u) {
    return fl_lbm_moments_inst(S, 0, x, y, z, lv, n, rho, u);
}

fl_status fl_lbm_totals_batch(fl_lbm_solver* S, double* mass, double* mom) {
    FL_REQUIRE(S && mass && mom, "null pointer");
    std::vector<double> a((size_t)S->ninst() * 4);
    fl_status r;
    switch (S->storage) {
        case FL_STORE_F32: r = totals<float>(S, a.data()); break;
        case FL_STORE_F16: r = totals<_Float16>(S, a.data()); break;
        case FL_STORE_F16C: r = totals<fp16c>(S, a.data()); break;
        default: r = totals<double>(S, a.data()); break;
    }
    if (r) return r;
    for (int k = 0; k < S->ninst(); ++k) {
        mass[k] = a[4 * k];
        for (int c = 0; c < 3; ++c) mom[3 * k + c] = a[4 * k + 1 + c];
    }
    return FL_OK;
}

fl_status fl_lbm_totals(fl_lbm_solver* S, double* mass, double* mom) {
    FL_REQUIRE(S && mass && mom, "null pointer");
    std::vector<double> m(S->ninst()), j((size_t)S->ninst() * 3);
    fl_status r = fl_lbm_totals_batch(S, m.data(), j.data());
    if (r) return r;
    *mass = 0;
    mom[0] = mom[1] = mom[2] = 0;
    for (int k = 0; k < S->ninst(); ++k) {
        *mass += m[k];
        for (int c = 0; c < 3; ++c) mom[c] += j[3 * k + c];
    }
    return FL_OK;
}

fl_status fl_lbm_force_batch(fl_lbm_solver* S, double* F) {
    FL_REQUIRE(S && F, "null pointer");
    switch (S->storage) {
        case FL_STORE_F32: return force<float>(S, F);
        case FL_STORE_F16: return force<_Float16>(S, F);
        case FL_STORE_F16C: return force<fp16c>(S, F);
        default: return force<double>(S, F);
    }
}

fl_status fl_lbm_force(fl_lbm_solver* S, double* F) {
    FL_REQUIRE(S && F, "null pointer");
    std::vector<double> a((size_t)S->ninst() * 3);
    fl_status r = fl_lbm_force_batch(S, a.data());
    if (r) return r;
    F[0] = F[1] = F[2] = 0;
    for (int k = 0; k < S->ninst(); ++k)
        for (int c = 0; c < 3; ++c) F[c] += a[3 * k + c];
    return FL_OK;
}

fl_status fl_lbm_velocity_sample(fl_lbm_solver* S) {
    FL_REQUIRE(S, "null solver");
    switch (S->storage) {
        case FL_STORE_F32: return velocity_sample<float>(S);
        case FL_STORE_F16: return velocity_sample<_Float16>(S);
        case FL_STORE_F16C: return velocity_sample<fp16c>(S);
        default: return velocity_sample<double>(S);
    }
}

fl_status fl_lbm_velocity_change(fl_lbm_solver* S, double* du) {
    FL_REQUIRE(S && du, "null pointer");
    switch (S->storage) {
        case FL_STORE_F32: return velocity_change<float>(S, du);
        case FL_STORE_F16: return velocity_change<_Float16>(S, du);
        case FL_STORE_F16C: return velocity_change<fp16c>(S, du);
        default: return velocity_change<double>(S, du);
    }
}

fl_status fl_lbm_set_live(fl_lbm_solver* S, int inst, int live) {
    FL_SOLVER_INST(S, inst);
    if (S->hinst[inst].live && !live) {
        // the instance's last state sits in the current buffer of each level;
        // copy it into the other one so that it reads the same whatever the
        // buffer parity after later steps
        const size_t sb = store_bytes(S->storage);
        for (int l = 0; l < S->L; ++l) {
            const HostLevel& H = S->h[l];
            DevLevel& D = S->d[l];
            const size_t b0 = H.abeg[inst], b1 = H.abeg[inst + 1];
            if (b1 == b0) continue;
            // the instance's active bricks are one run in every plane group
            for (int q = 0; q < Q; q += kPlaneGroup) {
                const size_t off = (plane_base(q, D.ld.S) + b0 * brick_stride(q)) * sb;
                FL_HIP(hipMemcpyAsync(static_cast<char*>(D.buf[1 - D.cur]) + off,
                                      static_cast<const char*>(D.buf[D.cur]) + off, (b1 - b0) * brick_stride(q) * sb,
                                      hipMemcpyDeviceToDevice, S->stream));
            }
        }
    }
    S->hinst[inst].live = live ? 1 : 0;
    FL_HIP(hipMemcpyAsync(&S->dinst[inst].live, &S->hinst[inst].live, sizeof(int32_t), hipMemcpyHostToDevice,
                          S->stream));
    FL_HIP(hipStreamSynchronize(S->stream));
    return FL_OK;
}

fl_status fl_lbm_set_graph(fl_lbm_solver* S, int enable) {
    FL_REQUIRE(S, "null solver");
    S->use_graph = enable != 0;
    if (!S->use_graph) S->drop_graphs();
    return FL_OK;
}

int fl_lbm_num_variants(void) { return collide_variants(); }

const char* fl_lbm_variant_name(int v) { return collide_variant_name(v); }

fl_status fl_lbm_set_variant(fl_lbm_solver* S, int level, int variant) {
    FL_REQUIRE(S, "null solver");
    FL_REQUIRE(level >= 0 && level < S->L, "level out of range");
    FL_REQUIRE(variant >= 0 && variant < collide_variants(), "unknown collide variant");
    S->variant[level] = variant;
    S->drop_graphs();
    return FL_OK;
}

fl_status fl_lbm_time_collide_variant(fl_lbm_solver* S, int level, int variant, int iters, float* ms_per_launch) {
    FL_REQUIRE(S && ms_per_launch, "null pointer");
    FL_REQUIRE(level >= 0 && level < S->L && iters > 0, "bad level or iteration count");
    FL_REQUIRE(variant >= 0 && variant < collide_variants(), "unknown collide variant");
    return time_any(S, level, variant, iters, ms_per_launch);
}

fl_status fl_lbm_time_collide(fl_lbm_solver* S, int level, int iters, float* ms_per_launch) {
    FL_REQUIRE(S && ms_per_launch, "null pointer");
    FL_REQUIRE(level >= 0 && level < S->L && iters > 0, "bad level or iteration count");
    return time_any(S, level, S->variant[level], iters, ms_per_launch);
}

fl_status fl_lbm_autotune(fl_lbm_solver* S, int iters, int* chosen) {
    FL_REQUIRE(S, "null solver");
    FL_REQUIRE(iters > 0, "iters must be positive");
    for (int l = 0; l < S->L; ++l) {
        if (S->h[l].nactive == 0) {
            if (chosen) chosen[l] = S->variant[l];
            continue;
        }
        // timing leaves the solver state unchanged.  Every variant once (after
        // one warm-up launch), then the three fastest three more times each,
        // alternating, keeping each one's best: clocks and cache state drift
        // during a sweep, and the leading variants are often within 1-3 %.
        const int nv = collide_variants();
        std::vector<float> t((size_t)nv, 1e30f);
        for (int v = 0; v < nv; ++v) {
            fl_status r = time_any(S, l, v, 1, &t[(size_t)v]);
            if (!r) r = time_any(S, l, v, iters, &t[(size_t)v]);
            if (r) return r;
        }
        std::vector<int> order((size_t)nv);
        for (int v = 0; v < nv; ++v) order[(size_t)v] = v;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return t[(size_t)a] < t[(size_t)b]; });
        const int top = nv < 3 ? nv : 3;
        for (int rep = 0; rep < 3; ++rep)
            for (int k = 0; k < top; ++k) {
                const int v = order[(size_t)k];
                float ms;
                if (fl_status r = time_any(S, l, v, iters, &ms)) return r;
                t[(size_t)v] = std::min(t[(size_t)v], ms);
            }
        int best = order[0];
        for (int k = 1; k < top; ++k)
            if (t[(size_t)order[(size_t)k]] < t[(size_t)best]) best = order[(size_t)k];
        S->variant[l] = best;
        if (chosen) chosen[l] = best;
    }
    S->drop_graphs();
    return FL_OK;
}

fl_status fl_lbm_vorticity_inst(fl_lbm_solver* S, int inst, const int32_t* x, const int32_t* y, const int32_t* z,
                                const uint8_t* lv, size_t n, double* out) {
    FL_SOLVER_INST(S, inst);
    FL_REQUIRE(n == 0 || (x && y && z && lv && out), "null pointer");
    switch (S->storage) {
        case FL_STORE_F32: return vorticity<float>(S, inst, x, y, z, lv, n, out);
        case FL_STORE_F16: return vorticity<_Float16>(S, inst, x, y, z, lv, n, out);
        case FL_STORE_F16C: return vorticity<fp16c>(S, inst, x, y, z, lv, n, out);
        default: return vorticity<double>(S, inst, x, y, z, lv, n, out);
    }
}

fl_status fl_lbm_vorticity(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                           size_t n, double* out) {
    return fl_lbm_vorticity_inst(S, 0, x, y, z, lv, n, out);
}

fl_status fl_lbm_regrid_batch(fl_lbm_solver* S, const size_t* offsets, const int32_t* x, const int32_t* y,
                              const int32_t* z, const uint8_t* lv, const uint8_t* state) {
    FL_REQUIRE(S && offsets, "null pointer");
    std::vector<fl_lbm_params> ps;
    for (const auto& ih : S->inst) ps.push_back(ih.p);
    std::unique_ptr<fl_lbm_solver> N;
    fl_status r = build_from_host(ps.data(), S->ninst(), offsets, x, y, z, lv, state, S->stream, S, N);
    if (r) return r;
    return adopt_layout(S, N);
}

fl_status fl_lbm_regrid_batch_device(fl_lbm_solver* S, const size_t* offsets, const int32_t* d_x, const int32_t* d_y,
                                     const int32_t* d_z, const uint8_t* d_level, const uint8_t* d_state) {
    FL_REQUIRE(S && offsets, "null pointer");
    std::vector<fl_lbm_params> ps;
    for (const auto& ih : S->inst) ps.push_back(ih.p);
    std::unique_ptr<fl_lbm_solver> N;
    fl_status r = build_on_device(ps.data(), S->ninst(), offsets, d_x, d_y, d_z, d_level, d_state, S->stream, S, N);
    if (r) return r;
    return adopt_layout(S, N);
}

fl_status fl_lbm_regrid(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                        const uint8_t* state, size_t n) {
    FL_REQUIRE(S, "null solver");
    FL_REQUIRE(S->ninst() == 1, "fl_lbm_regrid is for single-instance solvers; use fl_lbm_regrid_batch");
    const size_t off[2] = {0, n};
    return fl_lbm_regrid_batch(S, off, x, y, z, lv, state);
}

fl_status fl_lbm_stream(fl_lbm_solver* S, fl_stream* s) {
    FL_REQUIRE(S && s, "null pointer");
    *s = reinterpret_cast<fl_stream>(S->stream);
    return FL_OK;
}

}  // extern "C"
