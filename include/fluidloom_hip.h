/*
 * fluidloom_hip.h -- C ABI of libfluidloom_hip.so, the MI355X (gfx950)
 * implementation of FluidLoom's AMR lattice-Boltzmann hot path.
 *
 * Conventions
 *  - Every entry point returns fl_status (FL_OK == 0).  On failure the
 *    message is available from fl_last_error_string() (per thread).  This is
 *    the C form of the reference's FL_THROW/BackendError convention
 *    (include/fluidloom/common/FluidLoomError.h:77-81); the C++ layer turns a
 *    non-zero status back into an exception.
 *  - Pointers named d_* are device pointers, h_* are host pointers.  All
 *    device work is ordered on the given stream (NULL = default stream);
 *    calls return after enqueueing unless stated otherwise.
 *  - No entry point allocates device memory except fl_malloc, the *_create
 *    functions and fl_lbm_regrid (which rebuilds the solver's layout).
 *    Scratch space is passed in by the caller; its size comes from the
 *    matching *_temp_bytes query.
 */
#ifndef FLUIDLOOM_HIP_H
#define FLUIDLOOM_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int fl_status;
enum {
    FL_OK = 0,
    FL_E_INVALID = 1,  /* bad argument (size, null pointer, range) */
    FL_E_OOM = 2,      /* device allocation failed */
    FL_E_HIP = 3,      /* HIP runtime error */
    FL_E_MESH = 4,     /* mesh violates the coarse/fine coupling rules */
    FL_E_CAPACITY = 5  /* an output capacity is too small */
};
typedef void* fl_stream; /* hipStream_t */
typedef void* fl_event;  /* hipEvent_t */

const char* fl_last_error_string(void);
const char* fl_version(void);

/* ---- device, memory, streams (IBackend.h:31-48 of the reference) ------- */
fl_status fl_init(int device);
fl_status fl_device_count(int* count);
fl_status fl_device_info(int device, char* name, size_t name_len, size_t* total_mem,
                         int* cu_count, int* gcn_arch);
fl_status fl_malloc(void** d_ptr, size_t bytes);
fl_status fl_free(void* d_ptr);
fl_status fl_memcpy_h2d(void* d_dst, const void* h_src, size_t bytes, fl_stream s);
fl_status fl_memcpy_d2h(void* h_dst, const void* d_src, size_t bytes, fl_stream s);
fl_status fl_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, fl_stream s);
fl_status fl_memset(void* d_dst, int value, size_t bytes, fl_stream s);
/* Plain vectorised device-to-device copy kernel (16 B per lane): the
 * bandwidth reference for the HBM-bound kernels and the calibration of the
 * profiler's HBM byte counters.  No counterpart in the reference. */
fl_status fl_copy_kernel(void* d_dst, const void* d_src, size_t bytes, fl_stream s);
fl_status fl_stream_create(fl_stream* s);
fl_status fl_stream_destroy(fl_stream s);
fl_status fl_stream_sync(fl_stream s);
fl_status fl_device_sync(void);
/* Device block cache.  Solver tables, population buffers and the layout
 * builders' temporaries come from a per-device cache that keeps freed
 * blocks (at most an eighth of the device's memory, 32 GiB at most) for
 * later requests, so regrids do not pay hipMalloc / hipFree per table.
 * fl_release_cached_memory returns every cached block to HIP (for callers
 * sharing the device with other allocators); fl_cached_memory_bytes reports
 * the bytes held.  No counterpart in the reference. */
fl_status fl_release_cached_memory(void);
fl_status fl_cached_memory_bytes(uint64_t* bytes);
fl_status fl_event_create(fl_event* e);
fl_status fl_event_destroy(fl_event e);
fl_status fl_event_record(fl_event e, fl_stream s);
fl_status fl_event_elapsed_ms(fl_event start, fl_event stop, float* ms);
fl_status fl_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ---- a1/a2: Hilbert keys ------------------------------------------------
 * fl_hilbert_encode replaces hilbert::encode (src/core/hilbert/HilbertCodec.cpp:44-80)
 * and the test_hilbert_encode kernel (kernels/hilbert.cl:104-115).  d_level
 * may be NULL, then level_scalar applies to every element. */
fl_status fl_hilbert_encode(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                            const uint8_t* d_level, int level_scalar, uint64_t* d_key,
                            size_t n, fl_stream s);
/* replaces hilbert::decode (HilbertCodec.cpp:82-115), kernels/hilbert.cl:118-130 */
fl_status fl_hilbert_decode(const uint64_t* d_key, const uint8_t* d_level, int level_scalar,
                            int32_t* d_x, int32_t* d_y, int32_t* d_z, size_t n, fl_stream s);
/* The same codec over int4-strided coordinates (the cl_int3 / 16-byte layout
 * of test_hilbert_encode / test_hilbert_decode, kernels/hilbert.cl:104-130). */
fl_status fl_hilbert_encode_aos4(const int32_t* d_xyzw, int level, uint64_t* d_key, size_t n,
                                 fl_stream s);
fl_status fl_hilbert_decode_aos4(const uint64_t* d_key, int level, int32_t* d_xyzw, size_t n,
                                 fl_stream s);
/* prefix-aligned 63-bit key encode(x,y,z,L) << 3*(21-L); with pack_level != 0
 * the level is or-ed into the low 5 bits (requires every level <= 19) so that
 * one sort orders by (key, level). */
fl_status fl_hilbert_key63(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                           const uint8_t* d_level, int pack_level, uint64_t* d_key, size_t n,
                           fl_stream s);
/* unique cell key (L << 57) | (x << 38) | (y << 19) | z (hash-table key of
 * every cell lookup; level <= 19, coordinates in [0, 2^19), else EMPTY) */
fl_status fl_cell_hkey(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                       const uint8_t* d_level, uint64_t* d_key, size_t n, fl_stream s);
/* replaces adaptation::hilbert_encode_3d (src/adaptation/utils/HilbertCodec3D.cpp:29-53) */
fl_status fl_rotdir_encode(const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                           int max_level, uint64_t* d_key, size_t n, fl_stream s);

/* ---- a4: stable LSD radix sort of (u64 key, u32 value) pairs -------------
 * Replaces RadixSort::sortByKey (src/core/hashmap/RadixSort.cpp:55-104) and
 * kernels/hashmap/radix_sort.cl:14-122.  Sorts bits [begin_bit, end_bit) in
 * 8-bit digits; passes whose digit is constant are skipped.  Result is in
 * d_keys/d_vals; d_keys_alt/d_vals_alt are ping-pong buffers of n elements. */
size_t fl_sort_temp_bytes(size_t n);
fl_status fl_sort_pairs_u64u32(uint64_t* d_keys, uint32_t* d_vals, uint64_t* d_keys_alt,
                               uint32_t* d_vals_alt, size_t n, int begin_bit, int end_bit,
                               void* d_temp, size_t temp_bytes, fl_stream s);
/* fills d_vals[i] = i */
fl_status fl_iota_u32(uint32_t* d_vals, size_t n, fl_stream s);

/* ---- a5: exclusive scan (single-pass, decoupled look-back) ---------------
 * Replaces AdaptationEngine::exclusiveScan (AdaptationEngine.cpp:360-415) and
 * scan_local/scan_add_sums (compact_cells.cl:61-125); correct for any n.
 * d_total (may be NULL) receives the sum as u64. */
size_t fl_scan_temp_bytes(size_t n);
fl_status fl_scan_exclusive_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t* d_total,
                                size_t n, void* d_temp, size_t temp_bytes, fl_stream s);

/* Allocation-free form of fl_scan_exclusive_u32: no counterpart in the
 * reference.  For callers that cannot allocate device memory or pass a
 * workspace (code under HIP graph capture, tight serving loops).  Same
 * kernels and results; the look-back workspace is the library's own device
 * array (1 MiB per device, held until the process ends), so n is limited to
 * fl_scan_noalloc_max_elements() = 2^31 (larger n: FL_E_INVALID).  Runs on
 * the current device; call fl_init(device) first so that the array's
 * address is looked up outside any capture.  Calls on one device share the
 * array: calls that overlap on one device (other streams or threads) race.
 * Issues kernels only (no allocation, no memset). */
size_t fl_scan_noalloc_max_elements(void);
fl_status fl_scan_exclusive_u32_noalloc(const uint32_t* d_in, uint32_t* d_out, uint64_t* d_total,
                                        size_t n, fl_stream s);

/* ---- a7: GPU hash table ---------------------------------------------------
 * Replaces hash_clear/hash_build/hash_query (kernels/hashmap/hash_build.cl:19-82,
 * hash_query.cl:22-98) behind HashTableManager (HashTableManager.cpp:47-159).
 * Fibonacci hash, linear probing, at most 32 probes.  A duplicate key keeps
 * the smallest value.  d_probe_counts may be NULL. */
uint64_t fl_hash_capacity(size_t n);
fl_status fl_hash_clear(uint64_t* d_tkeys, uint32_t* d_tvals, uint64_t capacity, fl_stream s);
fl_status fl_hash_build(uint64_t* d_tkeys, uint32_t* d_tvals, uint64_t capacity,
                        const uint64_t* d_keys, const uint32_t* d_vals, size_t n,
                        uint32_t* d_probe_counts, fl_stream s);
fl_status fl_hash_query(const uint64_t* d_tkeys, const uint32_t* d_tvals, uint64_t capacity,
                        const uint64_t* d_query, uint32_t* d_out, size_t n,
                        uint32_t* d_probe_counts, fl_stream s);

/* ---- AMR cell arrays (SoA, native per-level coordinates) ----------------- */
typedef struct {
    int32_t* x;
    int32_t* y;
    int32_t* z;
    uint8_t* level;
    uint8_t* state;    /* CellState: 0 FLUID, 1 SOLID, ... (CellDescriptor.h:107-114) */
    int32_t* flag;     /* refine flag: > 0 split, -1 merge candidate */
    uint32_t* material;
} fl_cells;

/* a9: face-neighbour table nbr[6*i+f] (f = -x,+x,-y,+y,-z,+z): index of the
 * finest cell at level <= L covering the same-level face neighbour, 0xFFFFFFFE
 * when that position is covered only by finer cells, 0xFFFFFFFF outside the
 * domain (level-0 extents root[3], level L extents root << L).  Uses a cell
 * hash (capacity >= fl_hash_capacity(n), power of two; d_tkeys 16 B aligned,
 * as any hipMalloc pointer is) and n u64 of key scratch that the caller
 * provides; the table is rebuilt here. */
fl_status fl_face_neighbors(const fl_cells* c, size_t n, const int32_t* root, uint64_t* d_keys_scratch,
                            uint64_t* d_tkeys, uint32_t* d_tvals, uint64_t capacity, uint32_t* d_nbr,
                            fl_stream s);

/* ---- a10-a13: adaptation -------------------------------------------------
 * One regrid pass on device, no host round trips except the final counts:
 * balance (mode 0 = reference test points, BalanceEnforcer.cpp:145-263 with
 * balance_enforce.cl:57-172; mode 1 = D3Q19 neighbour-cube balance), split
 * (split_cells.cl:15-106), merge (merge_cells.cl:11-199, group ids in leader
 * order, fixed-order field sums), compaction and append
 * (compact_cells.cl:9-214, AdaptationEngine.cpp:417-656).  Output order:
 * survivors, children (parent order, child 0..7), parents (leader order). */
typedef struct {
    int balance_mode;       /* 0 reference test points, 1 D3Q19 cubes, -1 off */
    int anchor_level;       /* finest-unit level for mode 0 */
    int max_split_level;    /* no split at or above this level */
    int max_balance_iters;
    int32_t root[3];        /* level-0 extents in cells (level L: root << L) */
    int merge_balance_check;/* 1: drop merges that would break mode-1 balance */
    int ncomp;              /* AoS float field components to remap (0 = none) */
    int avg_rule;           /* merge averaging: 0 arithmetic mean, 1 sum */
} fl_regrid_config;

typedef struct {
    uint64_t num_cells_out;
    uint64_t num_split;
    uint64_t num_children;
    uint64_t num_groups;
    uint64_t balance_iterations;
    uint64_t balance_violations;
    uint64_t balance_marked;
    int32_t balance_converged;
    int32_t pad;
} fl_regrid_result;

/* Workspace of fl_regrid / fl_balance / fl_merge / fl_gen_flags for n cells
 * (16 B aligned, as any hipMalloc pointer is): keys, the cell map at twice
 * fl_hash_capacity(n) slots, shadow levels, group ids and scan words. */
size_t fl_regrid_temp_bytes(size_t n);
/* in/out cells may not alias; out must hold n + 7*n cells (worst case).
 * d_fields_in/out: AoS [cell][ncomp] float (may be NULL when ncomp == 0). */
fl_status fl_regrid(const fl_regrid_config* cfg, fl_cells* in, size_t n, fl_cells* out,
                    size_t out_capacity, const float* d_fields_in, float* d_fields_out,
                    void* d_temp, size_t temp_bytes, fl_regrid_result* res, fl_stream s);

/* Stage entry points (each a single device pass) for the C++ engines. */
fl_status fl_balance(const fl_regrid_config* cfg, fl_cells* c, size_t n, void* d_temp,
                     size_t temp_bytes, fl_regrid_result* res, fl_stream s);
/* split: d_child_start[i] = 8*rank or 0xFFFFFFFF; returns parent count */
fl_status fl_split(const fl_cells* c, size_t n, int max_split_level, uint32_t* d_child_start,
                   fl_cells* children, size_t child_capacity, void* d_temp, size_t temp_bytes,
                   uint64_t* h_num_split, fl_stream s);
fl_status fl_merge(const fl_cells* c, size_t n, uint32_t* d_group_id, uint32_t* d_group_leader,
                   fl_cells* parents, size_t parent_capacity, void* d_temp, size_t temp_bytes,
                   uint64_t* h_num_groups, fl_stream s);

/* Seeded synthetic inputs of the AMR integer micro-benchmark (splitmix64
 * streams keyed by (seed, index); identical to the oracle's generator):
 * fl_gen_cells: level ~ U{0..max_level}, coords ~ U[0, 2^level), state 0,
 * flag 0.  fl_gen_flags: split flag 1 with probability p_split; each complete
 * sibling group whose leader draws below p_merge gets -1 on its members that
 * are still at 0 (temp: fl_regrid_temp_bytes(n)). */
fl_status fl_gen_cells(fl_cells* c, size_t n, uint64_t seed, int max_level, fl_stream s);
fl_status fl_gen_flags(fl_cells* c, size_t n, uint64_t seed, double p_split, double p_merge,
                       void* d_temp, size_t temp_bytes, fl_stream s);

/* Refinement criteria used by the benchmark meshes (synchronous):
 * fl_mark_ball sets flag = 1 on FLUID cells with level < target_level whose
 * cube (level-0 units) is within `radius` of the centre; fl_mark_solid_ball
 * sets state = SOLID where the cell centre lies inside the ball. */
fl_status fl_mark_ball(fl_cells* c, size_t n, double cx, double cy, double cz, double radius,
                       int target_level, uint64_t* h_marked, fl_stream s);
fl_status fl_mark_solid_ball(fl_cells* c, size_t n, double cx, double cy, double cz,
                             double radius, uint64_t* h_marked, fl_stream s);

/* ---- a14: coarse/fine ghost fill and halo pack ----------------------------
 * Index-list gather/scatter replacing pack_fields/unpack_fields
 * (kernels/halo_pack.cl:3-45): packed[off/4 + k] = field[idx[k]] and the
 * inverse.  d_idx may be NULL (identity, as in the reference). */
fl_status fl_pack_fields(const float* d_field, const uint32_t* d_idx, size_t count,
                         uint8_t* d_packed, size_t offset_bytes, fl_stream s);
fl_status fl_unpack_fields(const uint8_t* d_packed, size_t offset_bytes, const uint32_t* d_idx,
                           size_t count, float* d_field, fl_stream s);
/* coarse->fine: fine[k][c] = coarse[parent[k]][c] (split_cells.cl:110-132);
 * fine->coarse: coarse[g][c] = rule0 ? sum_q fine[child[8g+q]][c]*0.125 : sum */
fl_status fl_ghost_fill_c2f(const float* d_coarse, const uint32_t* d_parent, size_t nfine,
                            int ncomp, float* d_fine, fl_stream s);
fl_status fl_ghost_fill_f2c(const float* d_fine, const uint32_t* d_child8, size_t ncoarse,
                            int ncomp, int rule, float* d_coarse, fl_stream s);

/* ---- a15/a16: AMR D3Q19 lattice-Boltzmann solver ------------------------ */
enum { FL_BC_WALL = 0, FL_BC_OUTFLOW = 1, FL_BC_PERIODIC = 2 };
/* F32: float; F16: IEEE binary16 holding f - w_q; F64: double (the GPU FP64
 * oracle); F16C: (f - w_q) / w_q in a 16-bit format with a 4-bit exponent
 * and an 11-bit mantissa (|x| < 4).  Both 16-bit forms round the
 * collide-stream output stochastically. */
enum { FL_STORE_F32 = 0, FL_STORE_F16 = 1, FL_STORE_F64 = 2, FL_STORE_F16C = 3 };

typedef struct {
    int32_t nx0, ny0, nz0;  /* level-0 grid */
    int32_t max_level;
    double tau0;
    int32_t bc[6];          /* faces -x,+x,-y,+y,-z,+z */
    double bc_u[6][3];      /* wall velocities */
    double rho0;            /* density in the moving-wall term */
    int32_t storage;        /* FL_STORE_* */
    int32_t pad;
    double rho_init;
    double u_init[3];
} fl_lbm_params;

typedef struct {
    uint64_t leaves[24];    /* fluid leaves per level */
    uint64_t bricks[24];    /* 4x4x4 bricks per level (active + ghost-only) */
    uint64_t active_bricks[24];
    uint64_t c2f[24], f2c[24], capture[24];
    int32_t nlevels;
    int32_t pad;
    uint64_t updates_per_step;   /* sum over levels of leaves * 2^level */
    uint64_t device_bytes;
    int32_t variant[24];    /* collide-stream variant per level (fl_lbm_variant_name) */
    int32_t ninst;          /* instances in the batch */
    int32_t graph;          /* 1 when steps replay a captured hipGraph */
    uint64_t graph_launches;
} fl_lbm_stats;

typedef struct fl_lbm_solver fl_lbm_solver;

/* Leaves are host arrays (native coords, state 0 fluid / 1 solid).  A NULL
 * stream makes the solver create (and own) a stream of its own, so that its
 * step can be graph-captured.  Single-instance form of fl_lbm_create_batch. */
fl_status fl_lbm_create(const fl_lbm_params* p, const int32_t* h_x, const int32_t* h_y,
                        const int32_t* h_z, const uint8_t* h_level, const uint8_t* h_state,
                        size_t n, fl_stream s, fl_lbm_solver** out);
/* (No counterpart in the reference: batches, graph replay, autotuning and the
 * serving queries below are this library's own.)
 * A batch of independent instances in one layout: instance k owns leaves
 * [offsets[k], offsets[k+1]) of the concatenated arrays and its own
 * params[k] (grid, tau, boundaries, levels); every instance uses the same
 * storage.  Bricks of all instances are concatenated per level (grouped by
 * instance, Hilbert order inside), so one launch per level steps the whole
 * batch.  Results of an instance do not depend on the rest of the batch. */
fl_status fl_lbm_create_batch(const fl_lbm_params* params, int ninst, const size_t* offsets,
                              const int32_t* h_x, const int32_t* h_y, const int32_t* h_z,
                              const uint8_t* h_level, const uint8_t* h_state, fl_stream s,
                              fl_lbm_solver** out);
/* Leaves already on the device (for instance the output of fl_regrid): the
 * layout -- bricks, neighbour rows, link words, ghost lists, force links --
 * is built by kernels and no leaf travels through the host.  fl_lbm_create*
 * upload host leaves once and take this path.  (The reference builds its
 * index tables on the host: MergeEngine.cpp:90-138, BalanceEnforcer.cpp:
 * 100-143.) */
fl_status fl_lbm_create_batch_device(const fl_lbm_params* params, int ninst, const size_t* offsets,
                                     const int32_t* d_x, const int32_t* d_y, const int32_t* d_z,
                                     const uint8_t* d_level, const uint8_t* d_state, fl_stream s,
                                     fl_lbm_solver** out);
/* The layout of one level as device pointers and sizes (read-only views
 * owned by the solver): the tables both builders produce.  No counterpart
 * in the reference. */
typedef struct {
    const void* nbr;          /* int32 [nb][27] */
    const void* bcoord;       /* int32 [nb][3] */
    const void* binst;        /* uint16 [nb] */
    const void* bflags;       /* uint32 [nb] */
    const void* cflags;       /* uint64 [nb * 64] */
    const void* c2f_gchild;   /* uint32 [n_groups][8] */
    const void* c2f_gcap;     /* uint32 [n_groups][7] */
    const void* c2f_gmask;    /* uint32 [n_groups][8] */
    const void* cap_slots;    /* uint32 [n_cap] */
    const void* f2c_ghost;    /* uint32 [n_f2c] */
    const void* f2c_child;    /* uint32 [n_f2c][8] */
    const void* force_links;  /* uint32 [n_links] */
    const void* abeg;         /* uint32 [ninst + 1] */
    const void* lbeg;         /* uint32 [ninst + 1] */
    uint64_t nb, nactive, n_groups, n_cap, n_f2c, n_links;
} fl_lbm_layout;
fl_status fl_lbm_layout_get(const fl_lbm_solver* solver, int level, fl_lbm_layout* out);
/* Process-wide: on != 0 makes fl_lbm_create* and fl_lbm_regrid* build the
 * layout on the host instead (the reference implementation the device
 * builder is tested against; fl_lbm_plan_bytes always uses it). */
fl_status fl_lbm_layout_on_host(int on);
/* Device bytes fl_lbm_create_batch would allocate for these leaves (host
 * layout only; no device work) -- the admission footprint of a request. */
fl_status fl_lbm_plan_bytes(const fl_lbm_params* params, int ninst, const size_t* offsets,
                            const int32_t* h_x, const int32_t* h_y, const int32_t* h_z,
                            const uint8_t* h_level, const uint8_t* h_state, uint64_t* device_bytes);
fl_status fl_lbm_num_instances(const fl_lbm_solver* solver, int* n);
/* fluid leaves per level of one instance and its updates per coarse step */
fl_status fl_lbm_inst_leaves(const fl_lbm_solver* solver, int inst, uint64_t* h_leaves24,
                             uint64_t* h_updates_per_step);
fl_status fl_lbm_destroy(fl_lbm_solver* solver);
fl_status fl_lbm_step(fl_lbm_solver* solver, int nsteps);
fl_status fl_lbm_stats_get(const fl_lbm_solver* solver, fl_lbm_stats* out);
/* host-side queries at given leaves (synchronous); the forms without _inst
 * address instance 0 */
fl_status fl_lbm_get_f_inst(fl_lbm_solver* solver, int inst, const int32_t* h_x,
                            const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level,
                            size_t n, double* h_f);
fl_status fl_lbm_set_f_inst(fl_lbm_solver* solver, int inst, const int32_t* h_x,
                            const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level,
                            size_t n, const double* h_f);
fl_status fl_lbm_moments_inst(fl_lbm_solver* solver, int inst, const int32_t* h_x,
                              const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level,
                              size_t n, double* h_rho, double* h_u);
fl_status fl_lbm_vorticity_inst(fl_lbm_solver* solver, int inst, const int32_t* h_x,
                                const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level,
                                size_t n, double* h_out);
fl_status fl_lbm_get_f(fl_lbm_solver* solver, const int32_t* h_x, const int32_t* h_y,
                       const int32_t* h_z, const uint8_t* h_level, size_t n, double* h_f);
fl_status fl_lbm_set_f(fl_lbm_solver* solver, const int32_t* h_x, const int32_t* h_y,
                       const int32_t* h_z, const uint8_t* h_level, size_t n, const double* h_f);
fl_status fl_lbm_moments(fl_lbm_solver* solver, const int32_t* h_x, const int32_t* h_y,
                         const int32_t* h_z, const uint8_t* h_level, size_t n, double* h_rho,
                         double* h_u);
/* |vorticity| in level-0 units at the given leaves (regrid criterion):
 * central differences over same-level fluid leaves, one-sided at level
 * edges (synchronous). */
fl_status fl_lbm_vorticity(fl_lbm_solver* solver, const int32_t* h_x, const int32_t* h_y,
                           const int32_t* h_z, const uint8_t* h_level, size_t n, double* h_out);
/* Replace the mesh by a new leaf list (from fl_regrid) and carry the
 * populations over: unchanged leaves are copied, children of split cells
 * get the parent's state with f_neq rescaled, merged parents the 8-child
 * average -- the AdaptMeshNode step of the reference
 * (src/runtime/AdaptMeshNode.cpp:9-64), whose GPU path drops the fields
 * (AdaptationEngine.cpp:540-549).  Rebuilds the solver's device layout. */
fl_status fl_lbm_regrid(fl_lbm_solver* solver, const int32_t* h_x, const int32_t* h_y,
                        const int32_t* h_z, const uint8_t* h_level, const uint8_t* h_state, size_t n);
/* Batch form: instance k's leaves are [offsets[k], offsets[k+1]).  Live
 * flags, collide variants and the graph setting are kept. */
fl_status fl_lbm_regrid_batch(fl_lbm_solver* solver, const size_t* offsets, const int32_t* h_x,
                              const int32_t* h_y, const int32_t* h_z, const uint8_t* h_level,
                              const uint8_t* h_state);
/* The same with the new leaves on the device (no leaf travels through the
 * host; the population transfer lists are built on the device too). */
fl_status fl_lbm_regrid_batch_device(fl_lbm_solver* solver, const size_t* offsets, const int32_t* d_x,
                                     const int32_t* d_y, const int32_t* d_z, const uint8_t* d_level,
                                     const uint8_t* d_state);
/* FP64 device reductions (synchronous).  Per instance: per-brick wave sums,
 * then one workgroup per instance over its brick range in a fixed order.
 * fl_lbm_totals / fl_lbm_force sum the instances. */
fl_status fl_lbm_totals(fl_lbm_solver* solver, double* h_mass, double* h_mom3);
fl_status fl_lbm_force(fl_lbm_solver* solver, double* h_force3);
fl_status fl_lbm_totals_batch(fl_lbm_solver* solver, double* h_mass, double* h_mom3n);
fl_status fl_lbm_force_batch(fl_lbm_solver* solver, double* h_force3n);
/* Per instance, max over its fluid leaves of |u - u_snapshot| (infinity
 * norm over components), then snapshot = u.  The first call takes the
 * snapshot and reports HUGE_VAL.  Steady-state test of the serving loop.
 * When fl_lbm_velocity_sample was called since the last call, u is the mean
 * of those samples (a window mean, which averages out the per-step noise of
 * 16-bit storage's stochastic rounding) and the sum restarts; otherwise u
 * is the current velocity.  No counterpart in the reference (it has no
 * lattice solver). */
fl_status fl_lbm_velocity_change(fl_lbm_solver* solver, double* h_du);
/* Adds the current velocity of every fluid leaf (FP32) to the window sum
 * that the next fl_lbm_velocity_change averages.  Stream-ordered, no sync. */
fl_status fl_lbm_velocity_sample(fl_lbm_solver* solver);
/* A non-live instance is no longer stepped (its slots keep their values);
 * the serving loop retires finished requests this way. */
fl_status fl_lbm_set_live(fl_lbm_solver* solver, int inst, int live);
/* enable != 0: fl_lbm_step replays a hipGraph of two captured steps
 * (captured on first use, re-captured after a regrid or variant change). */
fl_status fl_lbm_set_graph(fl_lbm_solver* solver, int enable);
fl_status fl_lbm_stream(fl_lbm_solver* solver, fl_stream* s);
/* Collide-stream variants (bricks per workgroup x workgroup-to-brick
 * mapping: in order, XCD-contiguous, XCD chunks of 64 / 128 bricks,
 * persistent) and the autotuner: it times every variant on the live layout
 * (`iters` launches each), re-times the three fastest three more times each
 * and keeps the fastest per level (h_chosen: one entry per level, may be
 * NULL).  All variants give the same bits. */
int fl_lbm_num_variants(void);
const char* fl_lbm_variant_name(int variant);
fl_status fl_lbm_set_variant(fl_lbm_solver* solver, int level, int variant);
fl_status fl_lbm_autotune(fl_lbm_solver* solver, int iters, int* h_chosen);
fl_status fl_lbm_time_collide_variant(fl_lbm_solver* solver, int level, int variant, int iters,
                                      float* ms_per_launch);
/* Measurement: `iters` back-to-back collide-stream launches of one level on
 * the solver's stream, timed with HIP events.  The solver state is left
 * unchanged (the output buffer is saved and restored around the launches,
 * which needs one level buffer of temporary device memory). */
fl_status fl_lbm_time_collide(fl_lbm_solver* solver, int level, int iters, float* ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif
