/*
 * fl_oracle.h -- CPU oracle for the AMR lattice-Boltzmann hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker.
 *
 * Each function restates the reference algorithm it follows (file:line in
 * fedres/fluidloom) in plain, single-threaded C so that the GPU kernels can be
 * compared against it bit for bit (integer work) or within a stated tolerance
 * (floating point).  Pinning of each function is listed in oracle.py.
 */
#ifndef FL_ORACLE_H
#define FL_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1: 12-state table codec (src/core/hilbert/HilbertCodec.cpp:44-115) */
uint64_t or_hilbert_encode(int32_t x, int32_t y, int32_t z, int level);
void or_hilbert_decode(uint64_t h, int level, int32_t* x, int32_t* y, int32_t* z);
/* prefix-aligned 63-bit key: encode(x,y,z,L) << 3*(21-L) */
uint64_t or_key63(int32_t x, int32_t y, int32_t z, int level);
/* unique hash key of a cell: (L << 57) | (x << 38) | (y << 19) | z for L <= 19
 * and coordinates in [0, 2^19); EMPTY (~0) otherwise */
uint64_t or_cell_hkey(int32_t x, int32_t y, int32_t z, int level);
void or_hilbert_encode_n(const int32_t* x, const int32_t* y, const int32_t* z,
                         const uint8_t* level, int level_scalar, uint64_t* out, size_t n);
void or_hilbert_decode_n(const uint64_t* h, const uint8_t* level, int level_scalar,
                         int32_t* x, int32_t* y, int32_t* z, size_t n);
void or_key63_n(const int32_t* x, const int32_t* y, const int32_t* z,
                const uint8_t* level, uint64_t* out, size_t n);

/* ---- a2: rotation/direction codec (src/adaptation/utils/HilbertCodec3D.cpp:29-53) */
uint64_t or_rotdir_encode(int32_t x, int32_t y, int32_t z, int max_level);
void or_rotdir_encode_n(const int32_t* x, const int32_t* y, const int32_t* z, int max_level,
                        uint64_t* out, size_t n);

/* ---- a4: stable LSD radix sort of (u64 key, u32 value) pairs
 * (kernels/hashmap/radix_sort.cl:14-105, made stable) over bits [begin_bit,end_bit) */
void or_sort_pairs_u64u32(uint64_t* keys, uint32_t* vals, size_t n, int begin_bit, int end_bit);

/* ---- a5: exclusive scan (src/adaptation/kernels/compact_cells.cl:61-125) */
uint64_t or_scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n);

/* ---- a7: Fibonacci-hash open-addressing table (kernels/hashmap/hash_build.cl:19-82,
 * kernels/hashmap/hash_query.cl:22-56).  Sequential insertion in index order. */
uint64_t or_hash_capacity(size_t n);
void or_hash_build(uint64_t* tkeys, uint32_t* tvals, uint64_t cap, const uint64_t* keys,
                   const uint32_t* vals, size_t n, uint32_t* probe_counts);
void or_hash_query(const uint64_t* tkeys, const uint32_t* tvals, uint64_t cap,
                   const uint64_t* q, uint32_t* out, size_t n);

/* ---- AMR cell arrays (SoA, native per-level coordinates) ---------------- */
typedef struct {
    int32_t* x;
    int32_t* y;
    int32_t* z;
    uint8_t* level;
    uint8_t* state;
    int32_t* flag;     /* refine flag: >0 split, -1 merge candidate */
    uint32_t* material;
    size_t n;
} or_cells;

/* a9: face-neighbour table: nbr[6*i+f], f = -x,+x,-y,+y,-z,+z; root = level-0
 * extents in cells (level L extents root << L) */
void or_face_neighbors(const or_cells* c, const int32_t* root, uint32_t* nbr);

/* a10: split (src/adaptation/kernels/split_cells.cl:15-106) */
size_t or_split_count(const or_cells* c, int max_level, uint32_t* child_start);
void or_split_generate(const or_cells* c, const uint32_t* child_start, or_cells* children);

/* a11: merge (src/adaptation/kernels/merge_cells.cl:11-199); group ids by leader index
 * order.  With check_balance a group is dropped when its parent would break
 * the D3Q19 neighbour-cube balance (mode 1 below).  Returns number of groups. */
size_t or_merge_mark(const or_cells* c, int check_balance, const int32_t* root, uint32_t* group_id,
                     uint32_t* group_leader);
void or_merge_parents(const or_cells* c, const uint32_t* group_leader, size_t ngroups,
                      or_cells* parents);
void or_merge_fields(const uint32_t* group_leader, size_t ngroups, const or_cells* c,
                     const float* fields, int ncomp, int rule, float* parent_fields);

/* a12: 2:1 balance.  mode 0 = reference test-point semantics
 * (src/adaptation/kernels/balance_enforce.cl:57-172 with shadow levels,
 * BalanceEnforcer.cpp:145-263) in anchor units of 2^anchor_level;
 * mode 1 = D3Q19 neighbour-cube balance used by the lattice solver.
 * Returns iterations; writes converged flag and totals. */
int or_balance(const or_cells* c, int mode, int anchor_level, int max_split_level,
               int max_iters, const int32_t* root, int* converged, uint32_t* total_violations,
               uint32_t* total_marked);

/* a6: compaction + append (compact_cells.cl:9-214, AdaptationEngine.cpp:417-656)
 * survivors (flag<=0 && group==INVALID) in order, then children, then parents.
 * Returns new count. */
size_t or_compact_append(const or_cells* c, const uint32_t* group_id, const or_cells* children,
                         const or_cells* parents, or_cells* out);

/* splitmix64-based seeded generator used for synthetic workloads */
uint64_t or_splitmix64(uint64_t x);
void or_gen_cells(uint64_t seed, size_t n, int max_level, int32_t* x, int32_t* y, int32_t* z,
                  uint8_t* level);

void or_gen_flags(const or_cells* c, uint64_t seed, uint64_t split_thresh, uint64_t merge_thresh);

#ifdef __cplusplus
}
#endif
#endif
