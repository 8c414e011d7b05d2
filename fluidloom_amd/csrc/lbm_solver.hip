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

namespace fl {
namespace lbm {
namespace {

enum : uint8_t { T_NONE = 0, T_FLUID = 1, T_SOLID = 2, T_C2F = 3, T_F2C = 4, T_WALL = 5 };

inline bool regular(uint8_t t) { return t == T_FLUID || t == T_C2F || t == T_F2C; }

inline uint64_t bkey(int32_t bx, int32_t by, int32_t bz) {
    return ((uint64_t)(uint32_t)bx << 42) | ((uint64_t)(uint32_t)by << 21) | (uint64_t)(uint32_t)bz;
}

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

struct HostLevel {
    int32_t ext[3] = {0, 0, 0};
    int32_t nbx[3] = {0, 0, 0};
    uint32_t nb = 0, nactive = 0;
    uint64_t nfluid = 0;
    std::unordered_map<uint64_t, int32_t> bmap;
    std::vector<int32_t> bcoord, nbr;
    std::vector<uint32_t> bflags;
    std::vector<uint64_t> cflags;
    std::vector<uint8_t> stype, needed;
    std::vector<int64_t> c2fsrc;
    std::vector<uint32_t> c2f_ghost, c2f_cap, cap_slots, f2c_ghost, f2c_child, force_links;

    int64_t slot_at(int32_t x, int32_t y, int32_t z) const {
        if (x < 0 || y < 0 || z < 0 || x >= ext[0] || y >= ext[1] || z >= ext[2]) return -1;
        auto it = bmap.find(bkey(x >> 2, y >> 2, z >> 2));
        if (it == bmap.end()) return -1;
        return (int64_t)it->second * 64 + (x & 3) + 4 * (y & 3) + 16 * (z & 3);
    }
    uint8_t type_at(int32_t x, int32_t y, int32_t z) const {
        int64_t s = slot_at(x, y, z);
        return s < 0 ? T_NONE : stype[(size_t)s];
    }
};

struct DevLevel {
    int32_t* nbr = nullptr;
    int32_t* bcoord = nullptr;
    uint32_t* bflags = nullptr;
    uint64_t* cflags = nullptr;
    uint32_t* c2f_ghost = nullptr;
    uint32_t* c2f_cap = nullptr;
    uint32_t* cap_slots = nullptr;
    uint32_t* f2c_ghost = nullptr;
    uint32_t* f2c_child = nullptr;
    uint32_t* force_links = nullptr;
    void* buf[2] = {nullptr, nullptr};
    void* cap[2] = {nullptr, nullptr};
    int cur = 0;
    double tau = 1.0;
    LevelDev ld{};
};

size_t store_bytes(int storage) { return storage == FL_STORE_F64 ? 8 : (storage == FL_STORE_F16 ? 2 : 4); }
size_t compute_bytes(int storage) { return storage == FL_STORE_F64 ? 8 : 4; }

}  // namespace
}  // namespace lbm
}  // namespace fl

struct fl_lbm_solver {
    fl_lbm_params p{};
    hipStream_t stream = nullptr;
    int L = 0;
    fl::lbm::HostLevel h[24];
    fl::lbm::DevLevel d[24];
    fl::lbm::BCDev bc{};
    std::vector<void*> allocs;
    double* partial = nullptr;
    int64_t* qslots = nullptr;
    double* qbuf = nullptr;
    size_t qcap = 0;
    uint64_t device_bytes = 0;

    ~fl_lbm_solver() {
        for (void* a : allocs) (void)hipFree(a);
    }
};

namespace fl {
namespace lbm {
namespace {

template <typename T>
fl_status dev_upload(fl_lbm_solver* S, T** dst, const std::vector<T>& v) {
    *dst = nullptr;
    if (v.empty()) return FL_OK;
    FL_HIP(hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T)));
    S->allocs.push_back(*dst);
    S->device_bytes += v.size() * sizeof(T);
    FL_HIP(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, S->stream));
    return FL_OK;
}

fl_status dev_alloc(fl_lbm_solver* S, void** dst, size_t bytes) {
    *dst = nullptr;
    if (bytes == 0) return FL_OK;
    FL_HIP(hipMalloc(dst, bytes));
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
int source_of(const fl_lbm_params& p, const int32_t* ext, int x, int y, int z, int i, int* sx, int* sy, int* sz,
              int* face) {
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

fl_status build_layout(fl_lbm_solver* S, const int32_t* X, const int32_t* Y, const int32_t* Z, const uint8_t* LV,
                       const uint8_t* ST, size_t n) {
    const fl_lbm_params& p = S->p;
    const int NL = S->L;
    // 1. leaf bricks per level
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        H.ext[0] = p.nx0 << l; H.ext[1] = p.ny0 << l; H.ext[2] = p.nz0 << l;
        for (int a = 0; a < 3; ++a) H.nbx[a] = H.ext[a] / 4;
    }
    std::vector<std::vector<uint64_t>> active_keys(NL), leaf_keys(NL);
    {
        std::vector<uint64_t> last(NL, ~0ull), lastf(NL, ~0ull);
        for (size_t k = 0; k < n; ++k) {
            int l = LV[k];
            if (l >= NL) return mesh_error("leaf level above max_level");
            const HostLevel& H = S->h[l];
            if (X[k] < 0 || Y[k] < 0 || Z[k] < 0 || X[k] >= H.ext[0] || Y[k] >= H.ext[1] || Z[k] >= H.ext[2])
                return mesh_error("leaf outside the domain");
            uint64_t bk = bkey(X[k] >> 2, Y[k] >> 2, Z[k] >> 2);
            if (bk != last[l]) { leaf_keys[l].push_back(bk); last[l] = bk; }
            if (ST[k] == 0 && bk != lastf[l]) { active_keys[l].push_back(bk); lastf[l] = bk; }
        }
    }
    // 2. bricks = leaf bricks + 26-neighbour dilation of active bricks
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        auto& ak = active_keys[l];
        std::sort(ak.begin(), ak.end());
        ak.erase(std::unique(ak.begin(), ak.end()), ak.end());
        std::vector<uint64_t> all = leaf_keys[l];
        all.reserve(all.size() + ak.size() * 27);
        for (uint64_t k : ak) {
            int32_t bx = (int32_t)(k >> 42), by = (int32_t)((k >> 21) & 0x1FFFFF), bz = (int32_t)(k & 0x1FFFFF);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int32_t c[3] = {bx + dx, by + dy, bz + dz};
                        bool ok = true;
                        for (int a = 0; a < 3; ++a) {
                            if (c[a] >= 0 && c[a] < H.nbx[a]) continue;
                            int f = c[a] < 0 ? 2 * a : 2 * a + 1;
                            if (p.bc[f] == FL_BC_PERIODIC) c[a] = (c[a] + H.nbx[a]) % H.nbx[a];
                            else ok = false;
                        }
                        if (ok) all.push_back(bkey(c[0], c[1], c[2]));
                    }
        }
        std::sort(all.begin(), all.end());
        all.erase(std::unique(all.begin(), all.end()), all.end());
        // order: active first, each group by Hilbert key of the brick coords
        struct E { int act; uint64_t h; uint64_t k; };
        std::vector<E> es(all.size());
        parallel_for(all.size(), [&](size_t i) {
            uint64_t k = all[i];
            int32_t bx = (int32_t)(k >> 42), by = (int32_t)((k >> 21) & 0x1FFFFF), bz = (int32_t)(k & 0x1FFFFF);
            es[i] = E{std::binary_search(ak.begin(), ak.end(), k) ? 0 : 1,
                      host_encode((uint32_t)bx, (uint32_t)by, (uint32_t)bz, 21), k};
        });
        std::sort(es.begin(), es.end(), [](const E& a, const E& b) {
            return a.act != b.act ? a.act < b.act : a.h < b.h;
        });
        H.nb = (uint32_t)es.size();
        H.nactive = (uint32_t)ak.size();
        H.bcoord.resize((size_t)H.nb * 3);
        H.bmap.reserve(H.nb * 2);
        for (uint32_t b = 0; b < H.nb; ++b) {
            uint64_t k = es[b].k;
            H.bcoord[3 * b] = (int32_t)(k >> 42);
            H.bcoord[3 * b + 1] = (int32_t)((k >> 21) & 0x1FFFFF);
            H.bcoord[3 * b + 2] = (int32_t)(k & 0x1FFFFF);
            H.bmap.emplace(k, (int32_t)b);
        }
        H.nbr.assign((size_t)H.nb * 27, -1);
        parallel_for(H.nb, [&](size_t b) {
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int32_t c[3] = {H.bcoord[3 * b] + dx, H.bcoord[3 * b + 1] + dy, H.bcoord[3 * b + 2] + dz};
                        bool ok = true;
                        for (int a = 0; a < 3; ++a) {
                            if (c[a] >= 0 && c[a] < H.nbx[a]) continue;
                            int f = c[a] < 0 ? 2 * a : 2 * a + 1;
                            if (p.bc[f] == FL_BC_PERIODIC) c[a] = (c[a] + H.nbx[a]) % H.nbx[a];
                            else ok = false;
                        }
                        if (!ok) continue;
                        auto it = H.bmap.find(bkey(c[0], c[1], c[2]));
                        if (it != H.bmap.end()) H.nbr[b * 27 + 13 + dx + 3 * dy + 9 * dz] = it->second;
                    }
        });
        H.stype.assign((size_t)H.nb * 64, T_NONE);
        H.needed.assign((size_t)H.nb * 64, 0);
        H.cflags.assign((size_t)H.nb * 64, 0);
        H.c2fsrc.assign((size_t)H.nb * 64, -1);
        H.bflags.assign(H.nb, 0);
    }
    // 3. leaf slots
    {
        std::atomic<uint64_t> placed{0}, missing{0};
        std::vector<std::atomic<uint64_t>> nfl(NL);
        for (auto& a : nfl) a.store(0);
        parallel_for(n, [&](size_t k) {
            HostLevel& H = S->h[LV[k]];
            int64_t s = H.slot_at(X[k], Y[k], Z[k]);
            if (s < 0) { missing.fetch_add(1); return; }
            uint8_t t = ST[k] == 0 ? T_FLUID : T_SOLID;
            uint8_t prev = __atomic_exchange_n(&H.stype[(size_t)s], t, __ATOMIC_RELAXED);
            if (prev == T_NONE) placed.fetch_add(1);
            if (t == T_FLUID) nfl[LV[k]].fetch_add(1);
        });
        if (missing.load()) return mesh_error("internal: leaf brick missing");
        if (placed.load() != n) return mesh_error("duplicate leaf");
        for (int l = 0; l < NL; ++l) S->h[l].nfluid = nfl[l].load();
    }
    // 4. classify the non-leaf slots from the leaves one level up / down
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        parallel_for(H.nb, [&](size_t b) {
            for (int loc = 0; loc < 64; ++loc) {
                size_t s = b * 64 + loc;
                if (H.stype[s] != T_NONE) continue;
                int32_t x = H.bcoord[3 * b] * 4 + (loc & 3), y = H.bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                        z = H.bcoord[3 * b + 2] * 4 + (loc >> 4);
                if (l > 0) {
                    const HostLevel& C = S->h[l - 1];
                    int64_t cs = C.slot_at(x >> 1, y >> 1, z >> 1);
                    uint8_t ct = cs < 0 ? T_NONE : C.stype[(size_t)cs];
                    if (ct == T_FLUID) { H.stype[s] = T_C2F; H.c2fsrc[s] = cs; continue; }
                    if (ct == T_SOLID) { H.stype[s] = T_WALL; continue; }
                }
                if (l + 1 < NL) {
                    const HostLevel& F = S->h[l + 1];
                    int nf = 0, ns = 0;
                    for (int q = 0; q < 8; ++q) {
                        uint8_t t = F.type_at(2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
                        nf += t == T_FLUID;
                        ns += t == T_SOLID;
                    }
                    if (nf == 8) H.stype[s] = T_F2C;
                    else if (nf + ns == 8) H.stype[s] = T_WALL;
                }
            }
        });
    }
    // The f2c/c2f classification above read only leaf kinds of neighbouring
    // levels (T_FLUID/T_SOLID are final before step 4 starts, and step 4
    // only writes non-leaf kinds).
    // 5. link codes of every fluid leaf
    std::atomic<int> bad{0};
    for (int l = 0; l < NL; ++l) {
        HostLevel& H = S->h[l];
        parallel_for(H.nb, [&](size_t b) {
            bool simple = true;
            std::vector<uint32_t> links;
            for (int loc = 0; loc < 64; ++loc) {
                size_t s = b * 64 + loc;
                if (H.stype[s] != T_FLUID) { simple = false; continue; }
                int32_t x = H.bcoord[3 * b] * 4 + (loc & 3), y = H.bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                        z = H.bcoord[3 * b + 2] * 4 + (loc >> 4);
                uint64_t w = kActiveBit;
                for (int i = 1; i < Q; ++i) {
                    int sx, sy, sz, face;
                    unsigned code;
                    if (source_of(p, H.ext, x, y, z, i, &sx, &sy, &sz, &face)) {
                        if (face >= 0) {
                            code = kLinkWall;
                        } else {
                            int64_t s2 = H.slot_at(sx, sy, sz);
                            uint8_t t2 = s2 < 0 ? T_NONE : H.stype[(size_t)s2];
                            if (regular(t2)) {
                                code = kLinkOutflow;
                                __atomic_store_n(&H.needed[(size_t)s2], (uint8_t)1, __ATOMIC_RELAXED);
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
                            __atomic_store_n(&H.needed[(size_t)s2], (uint8_t)1, __ATOMIC_RELAXED);
                        } else if (t2 == T_SOLID || t2 == T_WALL) {
                            code = kLinkSolid;
                            links.push_back((uint32_t)(s * 32 + i));
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
            if (!links.empty()) {
                static std::mutex mu;
                std::lock_guard<std::mutex> g(mu);
                H.force_links.insert(H.force_links.end(), links.begin(), links.end());
            }
        });
        std::sort(H.force_links.begin(), H.force_links.end());
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
    for (int l = 1; l < NL; ++l) {
        HostLevel& H = S->h[l];
        HostLevel& C = S->h[l - 1];
        std::vector<uint32_t> srcs;
        srcs.reserve(H.c2f_ghost.size());
        for (uint32_t g : H.c2f_ghost) srcs.push_back((uint32_t)H.c2fsrc[g]);
        std::vector<uint32_t> uniq = srcs;
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        C.cap_slots = uniq;
        H.c2f_cap.resize(srcs.size());
        for (size_t k = 0; k < srcs.size(); ++k)
            H.c2f_cap[k] = (uint32_t)(std::lower_bound(uniq.begin(), uniq.end(), srcs[k]) - uniq.begin());
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
                    (uint32_t)F.slot_at(2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
        }
    }
    for (int l = 0; l < NL; ++l)
        if ((uint64_t)S->h[l].nb * 64 >= (1ull << 32) / 32)
            return mesh_error("level too large for 32-bit slot lists");
    return FL_OK;
}

template <typename T>
fl_status upload_and_init(fl_lbm_solver* S) {
    const size_t sb = sizeof(T);
    const size_t cb = compute_bytes(S->p.storage);
    size_t maxpart = 1024 * 4;
    for (int l = 0; l < S->L; ++l) {
        HostLevel& H = S->h[l];
        DevLevel& D = S->d[l];
        fl_status r;
        if ((r = dev_upload(S, &D.nbr, H.nbr))) return r;
        if ((r = dev_upload(S, &D.bcoord, H.bcoord))) return r;
        if ((r = dev_upload(S, &D.bflags, H.bflags))) return r;
        if ((r = dev_upload(S, &D.cflags, H.cflags))) return r;
        if ((r = dev_upload(S, &D.c2f_ghost, H.c2f_ghost))) return r;
        if ((r = dev_upload(S, &D.c2f_cap, H.c2f_cap))) return r;
        if ((r = dev_upload(S, &D.cap_slots, H.cap_slots))) return r;
        if ((r = dev_upload(S, &D.f2c_ghost, H.f2c_ghost))) return r;
        if ((r = dev_upload(S, &D.f2c_child, H.f2c_child))) return r;
        if ((r = dev_upload(S, &D.force_links, H.force_links))) return r;
        const size_t slots = (size_t)H.nb * 64;
        for (int k = 0; k < 2; ++k) {
            if ((r = dev_alloc(S, &D.buf[k], slots * Q * sb))) return r;
            if ((r = dev_alloc(S, &D.cap[k], H.cap_slots.size() * Q * cb))) return r;
        }
        D.tau = std::ldexp(S->p.tau0 - 0.5, l) + 0.5;
        D.cur = 0;
        D.ld.nbr = D.nbr;
        D.ld.bcoord = D.bcoord;
        D.ld.bflags = D.bflags;
        D.ld.cflags = D.cflags;
        D.ld.S = slots;
        for (int a = 0; a < 3; ++a) D.ld.ext[a] = H.ext[a];
        D.ld.nactive = (int32_t)H.nactive;
        FL_HIP(launch_init<T>(slots, S->p.rho_init, S->p.u_init, static_cast<T*>(D.buf[0]),
                              static_cast<T*>(D.buf[1]), S->stream));
        maxpart = std::max(maxpart, (size_t)4 * 4096);
    }
    void* part;
    fl_status r = dev_alloc(S, &part, maxpart * sizeof(double));
    if (r) return r;
    S->partial = static_cast<double*>(part);
    FL_HIP(hipStreamSynchronize(S->stream));
    return FL_OK;
}

template <typename T>
fl_status advance(fl_lbm_solver* S, int l) {
    DevLevel& D = S->d[l];
    const HostLevel& H = S->h[l];
    const bool fine = l + 1 < S->L && S->h[l + 1].nfluid > 0;
    hipStream_t st = S->stream;
    const uint32_t ncap = (uint32_t)H.cap_slots.size();
    if (fine) {
        DevLevel& F = S->d[l + 1];
        const HostLevel& HF = S->h[l + 1];
        double fac = (1.0 - 1.0 / D.tau) * (2.0 * D.tau / F.tau);
        FL_HIP(launch_fill_f2c<T>(D.f2c_ghost, D.f2c_child, (uint32_t)H.f2c_ghost.size(), F.ld, S->bc,
                                  static_cast<const T*>(F.buf[1 - F.cur]), fac, static_cast<T*>(D.buf[D.cur]),
                                  D.ld.S, st));
        (void)HF;
        FL_HIP(launch_capture<T>(D.ld, S->bc, static_cast<const T*>(D.buf[1 - D.cur]), D.cap_slots, ncap, D.cap[0],
                                 st));
    }
    FL_HIP(launch_collide<T>(D.ld, S->bc, static_cast<const T*>(D.buf[D.cur]), static_cast<T*>(D.buf[1 - D.cur]),
                             1.0 / D.tau, st));
    D.cur ^= 1;
    if (fine) {
        FL_HIP(launch_capture<T>(D.ld, S->bc, static_cast<const T*>(D.buf[1 - D.cur]), D.cap_slots, ncap, D.cap[1],
                                 st));
        DevLevel& F = S->d[l + 1];
        const HostLevel& HF = S->h[l + 1];
        double fac = (1.0 - 1.0 / F.tau) * (F.tau / (2.0 * D.tau));
        for (int sub = 0; sub < 2; ++sub) {
            FL_HIP(launch_fill_c2f<T>(F.c2f_ghost, F.c2f_cap, (uint32_t)HF.c2f_ghost.size(), D.cap[0], D.cap[1], ncap,
                                      0.5 * sub, fac, static_cast<T*>(F.buf[F.cur]), F.ld.S, st));
            fl_status r = advance<T>(S, l + 1);
            if (r) return r;
        }
    }
    return FL_OK;
}

fl_status dispatch_step(fl_lbm_solver* S, int nsteps) {
    for (int t = 0; t < nsteps; ++t) {
        fl_status r;
        switch (S->p.storage) {
            case FL_STORE_F32: r = advance<float>(S, 0); break;
            case FL_STORE_F16: r = advance<_Float16>(S, 0); break;
            default: r = advance<double>(S, 0); break;
        }
        if (r) return r;
    }
    return FL_OK;
}

// host slots of the query leaves; -1 when not a fluid leaf of this mesh
fl_status query_slots(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                      size_t n, std::vector<int64_t>& slots) {
    slots.resize(n);
    for (size_t k = 0; k < n; ++k) {
        int l = lv[k];
        slots[k] = l < S->L ? S->h[l].slot_at(x[k], y[k], z[k]) : -1;
        if (slots[k] >= 0 && S->h[l].stype[(size_t)slots[k]] != T_FLUID) slots[k] = -1;
    }
    return FL_OK;
}

fl_status ensure_query(fl_lbm_solver* S, size_t n) {
    if (n <= S->qcap) return FL_OK;
    if (S->qslots) (void)hipFree(S->qslots);
    if (S->qbuf) (void)hipFree(S->qbuf);
    S->qslots = nullptr;
    S->qbuf = nullptr;
    FL_HIP(hipMalloc(reinterpret_cast<void**>(&S->qslots), n * sizeof(int64_t)));
    FL_HIP(hipMalloc(reinterpret_cast<void**>(&S->qbuf), n * Q * sizeof(double)));
    S->qcap = n;
    return FL_OK;
}

template <typename T>
fl_status get_f(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv, size_t n,
                double* out) {
    std::vector<int64_t> slots;
    query_slots(S, x, y, z, lv, n, slots);
    std::vector<double> tmp;
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
fl_status set_f(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv, size_t n,
                const double* in) {
    std::vector<int64_t> slots;
    query_slots(S, x, y, z, lv, n, slots);
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

template <typename T>
fl_status totals(fl_lbm_solver* S, double* mass, double* mom) {
    double acc[4] = {0, 0, 0, 0};
    std::vector<double> part;
    for (int l = 0; l < S->L; ++l) {
        DevLevel& D = S->d[l];
        if (D.ld.nactive == 0) continue;
        unsigned nb = totals_blocks(D.ld);
        FL_HIP(launch_totals<T>(D.ld, static_cast<const T*>(D.buf[D.cur]), std::ldexp(1.0, -3 * l), S->partial, nb,
                                S->stream));
        part.resize((size_t)nb * 4);
        FL_HIP(hipMemcpyAsync(part.data(), S->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                              S->stream));
        FL_HIP(hipStreamSynchronize(S->stream));
        for (unsigned b = 0; b < nb; ++b)
            for (int c = 0; c < 4; ++c) acc[c] += part[(size_t)b * 4 + c];
    }
    *mass = acc[0];
    mom[0] = acc[1]; mom[1] = acc[2]; mom[2] = acc[3];
    return FL_OK;
}

template <typename T>
fl_status force(fl_lbm_solver* S, double* F) {
    double acc[3] = {0, 0, 0};
    std::vector<double> part;
    for (int l = 0; l < S->L; ++l) {
        DevLevel& D = S->d[l];
        uint32_t nl = (uint32_t)S->h[l].force_links.size();
        if (nl == 0) continue;
        unsigned nb = force_blocks(nl);
        FL_HIP(launch_force<T>(D.force_links, nl, D.ld.S, static_cast<const T*>(D.buf[D.cur]),
                               std::ldexp(1.0, -2 * l), S->partial, nb, S->stream));
        part.resize((size_t)nb * 3);
        FL_HIP(hipMemcpyAsync(part.data(), S->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                              S->stream));
        FL_HIP(hipStreamSynchronize(S->stream));
        for (unsigned b = 0; b < nb; ++b)
            for (int c = 0; c < 3; ++c) acc[c] += part[(size_t)b * 3 + c];
    }
    F[0] = acc[0]; F[1] = acc[1]; F[2] = acc[2];
    return FL_OK;
}

fl_status create_solver(const fl_lbm_params* p, const int32_t* h_x, const int32_t* h_y, const int32_t* h_z,
                        const uint8_t* h_level, const uint8_t* h_state, size_t n, hipStream_t stream,
                        std::unique_ptr<fl_lbm_solver>& out) {
    FL_REQUIRE(n == 0 || (h_x && h_y && h_z && h_level && h_state), "null leaf arrays");
    FL_REQUIRE(p->max_level >= 0 && p->max_level < 20, "max_level must be in 0..19");
    FL_REQUIRE(p->nx0 > 0 && p->ny0 > 0 && p->nz0 > 0 && p->nx0 % 4 == 0 && p->ny0 % 4 == 0 && p->nz0 % 4 == 0,
               "level-0 extents must be positive multiples of 4");
    FL_REQUIRE(p->tau0 > 0.5, "tau0 must exceed 1/2");
    FL_REQUIRE(p->storage >= 0 && p->storage <= 2, "unknown storage type");
    for (int f = 0; f < 6; ++f) FL_REQUIRE(p->bc[f] >= 0 && p->bc[f] <= 2, "unknown boundary type");
    for (int a = 0; a < 3; ++a)
        FL_REQUIRE((p->bc[2 * a] == FL_BC_PERIODIC) == (p->bc[2 * a + 1] == FL_BC_PERIODIC),
                   "periodic faces must come in pairs");
    std::unique_ptr<fl_lbm_solver> S(new fl_lbm_solver());
    S->p = *p;
    S->L = p->max_level + 1;
    S->stream = stream;
    for (int f = 0; f < 6; ++f) {
        S->bc.type[f] = p->bc[f];
        for (int a = 0; a < 3; ++a) {
            S->bc.u[f][a] = (float)p->bc_u[f][a];
            S->bc.ud[f][a] = p->bc_u[f][a];
        }
    }
    S->bc.rho0 = p->rho0;
    fl_status r = build_layout(S.get(), h_x, h_y, h_z, h_level, h_state, n);
    if (r) return r;
    switch (p->storage) {
        case FL_STORE_F32: r = upload_and_init<float>(S.get()); break;
        case FL_STORE_F16: r = upload_and_init<_Float16>(S.get()); break;
        default: r = upload_and_init<double>(S.get()); break;
    }
    if (r) return r;
    out = std::move(S);
    return FL_OK;
}

template <typename T>
fl_status vorticity(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                    size_t n, double* out) {
    std::vector<int64_t> slots;
    query_slots(S, x, y, z, lv, n, slots);
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

// Population transfer from the old solver O into the freshly built N.
template <typename T>
fl_status remap(fl_lbm_solver* O, fl_lbm_solver* N) {
    std::vector<void*> tmp_allocs;
    auto up = [&](const std::vector<uint32_t>& v) -> const uint32_t* {
        if (v.empty()) return nullptr;
        void* d = nullptr;
        if (hipMalloc(&d, v.size() * 4) != hipSuccess) return nullptr;
        tmp_allocs.push_back(d);
        (void)hipMemcpyAsync(d, v.data(), v.size() * 4, hipMemcpyHostToDevice, N->stream);
        return static_cast<const uint32_t*>(d);
    };
    fl_status st = FL_OK;
    for (int l = 0; l < N->L && st == FL_OK; ++l) {
        const HostLevel& H = N->h[l];
        std::vector<uint32_t> cdst, csrc, sdst, ssrc, mdst, msrc, rdst;
        for (uint32_t b = 0; b < H.nb; ++b)
            for (int loc = 0; loc < 64; ++loc) {
                const size_t s = (size_t)b * 64 + loc;
                if (H.stype[s] != T_FLUID) continue;
                const int32_t x = H.bcoord[3 * b] * 4 + (loc & 3), y = H.bcoord[3 * b + 1] * 4 + ((loc >> 2) & 3),
                              z = H.bcoord[3 * b + 2] * 4 + (loc >> 4);
                const HostLevel& OH = O->h[l];
                int64_t os = OH.slot_at(x, y, z);
                if (os >= 0 && OH.stype[(size_t)os] == T_FLUID) {
                    cdst.push_back((uint32_t)s);
                    csrc.push_back((uint32_t)os);
                    continue;
                }
                if (l > 0) {
                    const HostLevel& OC = O->h[l - 1];
                    int64_t ps = OC.slot_at(x >> 1, y >> 1, z >> 1);
                    if (ps >= 0 && OC.stype[(size_t)ps] == T_FLUID) {
                        sdst.push_back((uint32_t)s);
                        ssrc.push_back((uint32_t)ps);
                        continue;
                    }
                }
                if (l + 1 < O->L) {
                    const HostLevel& OF = O->h[l + 1];
                    uint32_t ch[8];
                    bool all = true;
                    for (int q = 0; q < 8 && all; ++q) {
                        int64_t c = OF.slot_at(2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + ((q >> 2) & 1));
                        all = c >= 0 && OF.stype[(size_t)c] == T_FLUID;
                        if (all) ch[q] = (uint32_t)c;
                    }
                    if (all) {
                        mdst.push_back((uint32_t)s);
                        msrc.insert(msrc.end(), ch, ch + 8);
                        continue;
                    }
                }
                rdst.push_back((uint32_t)s);
            }
        RemapArgs<T> a;
        DevLevel& D = N->d[l];
        a.A = static_cast<T*>(D.buf[0]);
        a.B = static_cast<T*>(D.buf[1]);
        a.S = D.ld.S;
        a.bc = N->bc;
        const DevLevel& OD = O->d[l];
        a.copy_src_buf = static_cast<const T*>(OD.buf[OD.cur]);
        a.So = OD.ld.S;
        a.copy_dst = up(cdst);
        a.copy_src = up(csrc);
        a.ncopy = (uint32_t)cdst.size();
        if (l > 0) {
            const DevLevel& OC = O->d[l - 1];
            a.Lc = OC.ld;
            a.split_prev = static_cast<const T*>(OC.buf[1 - OC.cur]);
            a.split_fac = (1.0 - 1.0 / D.tau) * (D.tau / (2.0 * OC.tau));
        }
        a.split_dst = up(sdst);
        a.split_src = up(ssrc);
        a.nsplit = (uint32_t)sdst.size();
        if (l + 1 < O->L) {
            const DevLevel& OF = O->d[l + 1];
            a.Lf = OF.ld;
            a.merge_prev = static_cast<const T*>(OF.buf[1 - OF.cur]);
            a.merge_fac = (1.0 - 1.0 / D.tau) * (2.0 * D.tau / OF.tau);
        }
        a.merge_dst = up(mdst);
        a.merge_src8 = up(msrc);
        a.nmerge = (uint32_t)mdst.size();
        a.rest_dst = up(rdst);
        a.nrest = (uint32_t)rdst.size();
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
        hipError_t e = launch_ghost_eq<T>(D.c2f_ghost, D.c2f_cap, capsl, (uint32_t)H.c2f_ghost.size(), Ac, Sc,
                                          D.f2c_ghost, D.f2c_child, (uint32_t)H.f2c_ghost.size(), Af, Sf,
                                          static_cast<T*>(D.buf[0]), static_cast<T*>(D.buf[1]), D.ld.S, N->stream);
        if (e != hipSuccess) st = fl::hip_fail(e, "launch_ghost_eq", __FILE__, __LINE__);
    }
    hipError_t e = hipStreamSynchronize(N->stream);
    for (void* p : tmp_allocs) (void)hipFree(p);
    if (st) return st;
    if (e != hipSuccess) return fl::hip_fail(e, "remap", __FILE__, __LINE__);
    return FL_OK;
}

}  // namespace
}  // namespace lbm
}  // namespace fl

using namespace fl;
using namespace fl::lbm;

extern "C" {

fl_status fl_lbm_create(const fl_lbm_params* p, const int32_t* h_x, const int32_t* h_y, const int32_t* h_z,
                        const uint8_t* h_level, const uint8_t* h_state, size_t n, fl_stream s,
                        fl_lbm_solver** out) {
    FL_REQUIRE(p && out, "null pointer");
    *out = nullptr;
    std::unique_ptr<fl_lbm_solver> S;
    fl_status r = create_solver(p, h_x, h_y, h_z, h_level, h_state, n, FL_STREAM(s), S);
    if (r) return r;
    *out = S.release();
    return FL_OK;
}

fl_status fl_lbm_destroy(fl_lbm_solver* solver) {
    if (!solver) return FL_OK;
    (void)hipStreamSynchronize(solver->stream);
    if (solver->qslots) (void)hipFree(solver->qslots);
    if (solver->qbuf) (void)hipFree(solver->qbuf);
    delete solver;
    return FL_OK;
}

fl_status fl_lbm_step(fl_lbm_solver* solver, int nsteps) {
    FL_REQUIRE(solver, "null solver");
    FL_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    return dispatch_step(solver, nsteps);
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
        out->c2f[l] = H.c2f_ghost.size();
        out->f2c[l] = H.f2c_ghost.size();
        out->capture[l] = H.cap_slots.size();
        out->updates_per_step += H.nfluid << l;
    }
    out->device_bytes = S->device_bytes;
    return FL_OK;
}

fl_status fl_lbm_get_f(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                       size_t n, double* f) {
    FL_REQUIRE(S && (n == 0 || (x && y && z && lv && f)), "null pointer");
    switch (S->p.storage) {
        case FL_STORE_F32: return get_f<float>(S, x, y, z, lv, n, f);
        case FL_STORE_F16: return get_f<_Float16>(S, x, y, z, lv, n, f);
        default: return get_f<double>(S, x, y, z, lv, n, f);
    }
}

fl_status fl_lbm_set_f(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                       size_t n, const double* f) {
    FL_REQUIRE(S && (n == 0 || (x && y && z && lv && f)), "null pointer");
    switch (S->p.storage) {
        case FL_STORE_F32: return set_f<float>(S, x, y, z, lv, n, f);
        case FL_STORE_F16: return set_f<_Float16>(S, x, y, z, lv, n, f);
        default: return set_f<double>(S, x, y, z, lv, n, f);
    }
}

fl_status fl_lbm_moments(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                         size_t n, double* rho, double* u) {
    FL_REQUIRE(S && (n == 0 || (rho && u)), "null pointer");
    std::vector<double> f(n * Q);
    fl_status r = fl_lbm_get_f(S, x, y, z, lv, n, f.data());
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

fl_status fl_lbm_totals(fl_lbm_solver* S, double* mass, double* mom) {
    FL_REQUIRE(S && mass && mom, "null pointer");
    switch (S->p.storage) {
        case FL_STORE_F32: return totals<float>(S, mass, mom);
        case FL_STORE_F16: return totals<_Float16>(S, mass, mom);
        default: return totals<double>(S, mass, mom);
    }
}

fl_status fl_lbm_force(fl_lbm_solver* S, double* F) {
    FL_REQUIRE(S && F, "null pointer");
    switch (S->p.storage) {
        case FL_STORE_F32: return force<float>(S, F);
        case FL_STORE_F16: return force<_Float16>(S, F);
        default: return force<double>(S, F);
    }
}

fl_status fl_lbm_time_collide(fl_lbm_solver* S, int level, int iters, float* ms_per_launch) {
    FL_REQUIRE(S && ms_per_launch, "null pointer");
    FL_REQUIRE(level >= 0 && level < S->L && iters > 0, "bad level or iteration count");
    DevLevel& D = S->d[level];
    hipEvent_t a, b;
    FL_HIP(hipEventCreate(&a));
    FL_HIP(hipEventCreate(&b));
    FL_HIP(hipEventRecord(a, S->stream));
    for (int it = 0; it < iters; ++it) {
        hipError_t e;
        switch (S->p.storage) {
            case FL_STORE_F32:
                e = launch_collide<float>(D.ld, S->bc, static_cast<const float*>(D.buf[D.cur]),
                                          static_cast<float*>(D.buf[1 - D.cur]), 1.0 / D.tau, S->stream);
                break;
            case FL_STORE_F16:
                e = launch_collide<_Float16>(D.ld, S->bc, static_cast<const _Float16*>(D.buf[D.cur]),
                                             static_cast<_Float16*>(D.buf[1 - D.cur]), 1.0 / D.tau, S->stream);
                break;
            default:
                e = launch_collide<double>(D.ld, S->bc, static_cast<const double*>(D.buf[D.cur]),
                                           static_cast<double*>(D.buf[1 - D.cur]), 1.0 / D.tau, S->stream);
                break;
        }
        FL_HIP(e);
        D.cur ^= 1;
    }
    FL_HIP(hipEventRecord(b, S->stream));
    FL_HIP(hipEventSynchronize(b));
    float ms = 0;
    FL_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_launch = ms / iters;
    return FL_OK;
}

fl_status fl_lbm_vorticity(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                           size_t n, double* out) {
    FL_REQUIRE(S && (n == 0 || (x && y && z && lv && out)), "null pointer");
    switch (S->p.storage) {
        case FL_STORE_F32: return vorticity<float>(S, x, y, z, lv, n, out);
        case FL_STORE_F16: return vorticity<_Float16>(S, x, y, z, lv, n, out);
        default: return vorticity<double>(S, x, y, z, lv, n, out);
    }
}

fl_status fl_lbm_regrid(fl_lbm_solver* S, const int32_t* x, const int32_t* y, const int32_t* z, const uint8_t* lv,
                        const uint8_t* state, size_t n) {
    FL_REQUIRE(S, "null solver");
    std::unique_ptr<fl_lbm_solver> N;
    fl_lbm_params p = S->p;
    fl_status r = create_solver(&p, x, y, z, lv, state, n, S->stream, N);
    if (r) return r;
    switch (S->p.storage) {
        case FL_STORE_F32: r = remap<float>(S, N.get()); break;
        case FL_STORE_F16: r = remap<_Float16>(S, N.get()); break;
        default: r = remap<double>(S, N.get()); break;
    }
    if (r) return r;
    // the solver handle keeps its identity; its contents become the new mesh
    std::swap(S->h, N->h);
    std::swap(S->d, N->d);
    std::swap(S->allocs, N->allocs);
    std::swap(S->partial, N->partial);
    std::swap(S->device_bytes, N->device_bytes);
    return FL_OK;
}

fl_status fl_lbm_stream(fl_lbm_solver* S, fl_stream* s) {
    FL_REQUIRE(S && s, "null pointer");
    *s = reinterpret_cast<fl_stream>(S->stream);
    return FL_OK;
}

}  // extern "C"
