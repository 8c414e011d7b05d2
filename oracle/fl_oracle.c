/*
 * fl_oracle.c -- CPU oracle (test infrastructure only; never linked into the
 * product library).  Plain single-threaded C restating the reference's
 * hot-path algorithms; see fl_oracle.h for the function list and oracle.py
 * for how each function is pinned.
 */
#include "fl_oracle.h"

#include <stdlib.h>
#include <string.h>

#define OR_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

static int in_domain(int l, const int32_t* root, int32_t x, int32_t y, int32_t z) {
    return x >= 0 && y >= 0 && z >= 0 && (int64_t)x < ((int64_t)root[0] << l) &&
           (int64_t)y < ((int64_t)root[1] << l) && (int64_t)z < ((int64_t)root[2] << l);
}

/* ------------------------------------------------------------------------ */
/* a1: 12-state table codec.
 * Table restated from src/core/hilbert/HilbertCodec.cpp:13-42 (format:
 * (next_state << 3) | curve_digit, indexed [state][(z<<2)|(y<<1)|x]).       */
static const uint8_t HT[12][8] = {
    {8, 17, 27, 18, 47, 38, 28, 37},  {16, 71, 1, 62, 51, 52, 2, 61},
    {0, 75, 95, 76, 9, 10, 86, 85},   {4, 77, 55, 78, 3, 66, 80, 65},
    {50, 49, 45, 46, 67, 88, 68, 7},  {6, 57, 39, 72, 5, 58, 84, 83},
    {12, 11, 29, 90, 79, 32, 30, 89}, {74, 91, 73, 40, 69, 92, 70, 15},
    {14, 13, 81, 82, 63, 36, 24, 35}, {20, 31, 19, 56, 53, 54, 42, 41},
    {26, 93, 43, 44, 25, 94, 64, 23}, {22, 87, 21, 60, 33, 48, 34, 59},
};
static const uint8_t IHT[12][8] = {
    {8, 17, 19, 26, 30, 39, 37, 44},  {16, 2, 6, 52, 53, 63, 59, 65},
    {0, 12, 13, 73, 75, 87, 86, 90},  {86, 71, 69, 4, 0, 73, 75, 50},
    {93, 49, 48, 68, 70, 42, 43, 7},  {75, 57, 61, 87, 86, 4, 0, 34},
    {37, 95, 91, 9, 8, 26, 30, 76},   {43, 74, 72, 89, 93, 68, 70, 15},
    {30, 82, 83, 39, 37, 9, 8, 60},   {59, 47, 46, 18, 16, 52, 53, 25},
    {70, 28, 24, 42, 43, 89, 93, 23}, {53, 36, 38, 63, 59, 18, 16, 81},
};

/* HilbertCodec.cpp:44-80: coordinates masked to 21 bits, MSB first, result
 * masked to 3*level bits (the mask is a no-op for level in 1..21). */
uint64_t or_hilbert_encode(int32_t x, int32_t y, int32_t z, int level) {
    uint32_t ux = (uint32_t)x & 0x1FFFFFu, uy = (uint32_t)y & 0x1FFFFFu, uz = (uint32_t)z & 0x1FFFFFu;
    uint64_t h = 0;
    int s = 0;
    for (int i = level - 1; i >= 0; --i) {
        int q = (int)((((uz >> i) & 1u) << 2) | (((uy >> i) & 1u) << 1) | ((ux >> i) & 1u));
        int v = HT[s][q];
        h = (h << 3) | (uint64_t)(v & 7);
        s = v >> 3;
    }
    unsigned shift = (unsigned)(uint8_t)(64 - level * 3);
    if (shift < 64) { h <<= shift; h >>= shift; }
    return h;
}

/* HilbertCodec.cpp:82-115 */
void or_hilbert_decode(uint64_t h, int level, int32_t* x, int32_t* y, int32_t* z) {
    if (level == 0) { *x = *y = *z = 0; return; }
    uint64_t t = h << (64 - level * 3);
    uint32_t ux = 0, uy = 0, uz = 0;
    int s = 0;
    for (int i = 0; i < level; ++i) {
        int d = (int)((t >> 61) & 7u);
        int v = IHT[s][d];
        int q = v & 7;
        s = v >> 3;
        ux = (ux << 1) | (uint32_t)(q & 1);
        uy = (uy << 1) | (uint32_t)((q >> 1) & 1);
        uz = (uz << 1) | (uint32_t)((q >> 2) & 1);
        t <<= 3;
    }
    *x = (int32_t)ux; *y = (int32_t)uy; *z = (int32_t)uz;
}

uint64_t or_key63(int32_t x, int32_t y, int32_t z, int level) {
    return or_hilbert_encode(x, y, z, level) << (3 * (21 - level));
}

uint64_t or_cell_hkey(int32_t x, int32_t y, int32_t z, int level) {
    if (level < 0 || level > 19 || x < 0 || y < 0 || z < 0 || x >= (1 << 19) || y >= (1 << 19) ||
        z >= (1 << 19))
        return OR_EMPTY_KEY;
    return ((uint64_t)level << 57) | ((uint64_t)x << 38) | ((uint64_t)y << 19) | (uint64_t)z;
}

void or_hilbert_encode_n(const int32_t* x, const int32_t* y, const int32_t* z,
                         const uint8_t* level, int level_scalar, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i)
        out[i] = or_hilbert_encode(x[i], y[i], z[i], level ? level[i] : level_scalar);
}

void or_hilbert_decode_n(const uint64_t* h, const uint8_t* level, int level_scalar,
                         int32_t* x, int32_t* y, int32_t* z, size_t n) {
    for (size_t i = 0; i < n; ++i)
        or_hilbert_decode(h[i], level ? level[i] : level_scalar, &x[i], &y[i], &z[i]);
}

void or_key63_n(const int32_t* x, const int32_t* y, const int32_t* z,
                const uint8_t* level, uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = or_key63(x[i], y[i], z[i], level[i]);
}

/* ------------------------------------------------------------------------ */
/* a2: rotation/direction codec, tables restated from
 * src/adaptation/utils/HilbertCodec3D.cpp:7-27; loop :29-53 (no coordinate
 * masking; bits above max_level are ignored). */
static const uint8_t ROT[8][8] = {
    {0, 1, 2, 3, 4, 5, 6, 7}, {1, 0, 7, 6, 5, 4, 3, 2}, {2, 3, 0, 1, 6, 7, 4, 5},
    {3, 2, 5, 4, 7, 6, 1, 0}, {4, 5, 6, 7, 0, 1, 2, 3}, {5, 4, 3, 2, 1, 0, 7, 6},
    {6, 7, 4, 5, 2, 3, 0, 1}, {7, 6, 1, 0, 3, 2, 5, 4}};
static const uint8_t DIR[8][8] = {
    {0, 0, 0, 0, 0, 0, 0, 0}, {3, 0, 0, 5, 5, 0, 0, 3}, {4, 4, 0, 3, 3, 0, 4, 4},
    {3, 5, 5, 0, 0, 5, 5, 3}, {2, 2, 2, 0, 0, 2, 2, 2}, {3, 0, 0, 5, 5, 0, 0, 3},
    {4, 4, 0, 3, 3, 0, 4, 4}, {3, 5, 5, 0, 0, 5, 5, 3}};

uint64_t or_rotdir_encode(int32_t x, int32_t y, int32_t z, int max_level) {
    uint64_t h = 0;
    int r = 0;
    for (int l = max_level - 1; l >= 0; --l) {
        int b = ((x >> l) & 1) | (((y >> l) & 1) << 1) | (((z >> l) & 1) << 2);
        h = (h << 3) | ROT[r][b];
        r = DIR[r][b];
    }
    return h;
}

void or_rotdir_encode_n(const int32_t* x, const int32_t* y, const int32_t* z, int max_level,
                        uint64_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = or_rotdir_encode(x[i], y[i], z[i], max_level);
}

/* ------------------------------------------------------------------------ */
/* a4: LSD radix sort, 8-bit digits (radix_sort.cl:14-105).  The reference
 * scatter claims positions with atomic_inc and is therefore unstable; the
 * oracle (and the build) use the stable counting-sort scatter. */
void or_sort_pairs_u64u32(uint64_t* keys, uint32_t* vals, size_t n, int begin_bit, int end_bit) {
    if (n < 2) return;
    uint64_t* k2 = (uint64_t*)malloc(n * sizeof(uint64_t));
    uint32_t* v2 = (uint32_t*)malloc(n * sizeof(uint32_t));
    uint64_t* ks = keys; uint32_t* vs = vals; uint64_t* kd = k2; uint32_t* vd = v2;
    for (int b = begin_bit; b < end_bit; b += 8) {
        int nb = end_bit - b < 8 ? end_bit - b : 8;
        uint64_t mask = (1ull << nb) - 1;
        size_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (size_t i = 0; i < n; ++i) cnt[((ks[i] >> b) & mask) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; ++i) {
            size_t p = cnt[(ks[i] >> b) & mask]++;
            kd[p] = ks[i]; vd[p] = vs[i];
        }
        uint64_t* tk = ks; ks = kd; kd = tk;
        uint32_t* tv = vs; vs = vd; vd = tv;
    }
    if (ks != keys) {
        memcpy(keys, ks, n * sizeof(uint64_t));
        memcpy(vals, vs, n * sizeof(uint32_t));
    }
    free(k2); free(v2);
}

/* ------------------------------------------------------------------------ */
/* a5: exclusive scan; total returned (AdaptationEngine.cpp:407-412 computes
 * total = out[n-1] + in[n-1]).  Unlike compact_cells.cl:61-125 this is
 * correct for any n. */
uint64_t or_scan_exclusive_u32(const uint32_t* in, uint32_t* out, size_t n) {
    uint64_t acc = 0;
    for (size_t i = 0; i < n; ++i) { uint32_t v = in[i]; out[i] = (uint32_t)acc; acc += v; }
    return acc;
}

/* ------------------------------------------------------------------------ */
/* a7: HashTableManager::computeCapacity (HashTableManager.cpp:47-57):
 * 2^clamp(ceil(log2(n / 0.6)), 20, 30). */
uint64_t or_hash_capacity(size_t n) {
    double m = (double)n / (double)0.6f;
    int l = 0;
    while (l < 63 && (double)(1ull << l) < m) ++l;
    if (l < 20) l = 20;
    if (l > 30) l = 30;
    return 1ull << l;
}

#define OR_EMPTY 0xFFFFFFFFFFFFFFFFull
#define OR_INVALID 0xFFFFFFFFu
#define OR_PROBES 32

/* hash_build.cl:19-67: slot = (key * 0x9e3779b97f4a7c15) & (cap-1), linear
 * probing, at most 32 probes, EMPTY keys skipped.  A duplicate key keeps the
 * smallest value (the reference overwrites in arbitrary order; the build
 * fixes the winner so that query results are deterministic). */
void or_hash_build(uint64_t* tkeys, uint32_t* tvals, uint64_t cap, const uint64_t* keys,
                   const uint32_t* vals, size_t n, uint32_t* probe_counts) {
    for (uint64_t s = 0; s < cap; ++s) { tkeys[s] = OR_EMPTY; tvals[s] = OR_INVALID; }
    for (size_t i = 0; i < n; ++i) {
        uint64_t k = keys[i];
        if (k == OR_EMPTY) { if (probe_counts) probe_counts[i] = 0; continue; }
        uint64_t s = (k * 0x9e3779b97f4a7c15ull) & (cap - 1);
        uint32_t p = 0;
        for (; p < OR_PROBES; ++p) {
            if (tkeys[s] == OR_EMPTY) { tkeys[s] = k; tvals[s] = vals[i]; break; }
            if (tkeys[s] == k) { if (vals[i] < tvals[s]) tvals[s] = vals[i]; break; }
            s = (s + 1) & (cap - 1);
        }
        if (probe_counts) probe_counts[i] = p;
    }
}

/* hash_query.cl:22-56 */
void or_hash_query(const uint64_t* tkeys, const uint32_t* tvals, uint64_t cap,
                   const uint64_t* q, uint32_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        uint64_t k = q[i];
        out[i] = OR_INVALID;
        if (k == OR_EMPTY) continue;
        uint64_t s = (k * 0x9e3779b97f4a7c15ull) & (cap - 1);
        for (uint32_t p = 0; p < OR_PROBES; ++p) {
            if (tkeys[s] == k) { out[i] = tvals[s]; break; }
            if (tkeys[s] == OR_EMPTY) break;
            s = (s + 1) & (cap - 1);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Cell lookup shared by neighbour, merge and balance restatements: a map
 * from the unique cell key (level, coords) to the smallest index holding it. */
typedef struct { uint64_t* k; uint32_t* v; uint64_t cap; int bits; } cellmap;

/* the slot is the top bits of the Fibonacci product: the packed key carries z
 * in its low bits, so low product bits would cluster cells that share y, z */
static uint64_t cm_slot(const cellmap* m, uint64_t k) { return (k * 0x9e3779b97f4a7c15ull) >> (64 - m->bits); }

static void cm_build(cellmap* m, const or_cells* c, const uint64_t* keys) {
    uint64_t cap = 1024;
    int bits = 10;
    while (cap < 2 * (uint64_t)c->n + 16) { cap <<= 1; ++bits; }
    m->cap = cap;
    m->bits = bits;
    m->k = (uint64_t*)malloc(cap * sizeof(uint64_t));
    m->v = (uint32_t*)malloc(cap * sizeof(uint32_t));
    for (uint64_t s = 0; s < cap; ++s) { m->k[s] = OR_EMPTY; m->v[s] = OR_INVALID; }
    for (size_t i = 0; i < c->n; ++i) {
        uint64_t k = keys[i];
        if (k == OR_EMPTY) continue;
        uint64_t s = cm_slot(m, k);
        for (;;) {
            if (m->k[s] == OR_EMPTY) { m->k[s] = k; m->v[s] = (uint32_t)i; break; }
            if (m->k[s] == k) break; /* keep smallest index */
            s = (s + 1) & (cap - 1);
        }
    }
}
static uint32_t cm_get(const cellmap* m, uint64_t k) {
    uint64_t s = cm_slot(m, k);
    for (;;) {
        if (m->k[s] == k) return m->v[s];
        if (m->k[s] == OR_EMPTY) return OR_INVALID;
        s = (s + 1) & (m->cap - 1);
    }
}
static void cm_free(cellmap* m) { free(m->k); free(m->v); }

static uint32_t cm_cell(const cellmap* m, int l, int32_t x, int32_t y, int32_t z) {
    uint64_t k = or_cell_hkey(x, y, z, l);
    if (k == OR_EMPTY_KEY) return OR_INVALID;
    return cm_get(m, k);
}

static void cell_keys(const or_cells* c, uint64_t* keys) {
    for (size_t i = 0; i < c->n; ++i) keys[i] = or_cell_hkey(c->x[i], c->y[i], c->z[i], c->level[i]);
}

/* ------------------------------------------------------------------------ */
/* a9: face neighbours.  For face f of cell (x,y,z,L) the same-level position
 * p = (x,y,z) + e_f is looked up at levels L, L-1, ..., 0 (p >> (L-l)); the
 * first (finest) covering cell wins.  Positions outside the domain
 * [0, root << L) give INVALID; positions not covered at level <= L but with
 * a child leaf at L+1 give FINER (0xFFFFFFFE).  This is the point-location
 * probe of balance_enforce.cl:91-146 in native coordinates. */
void or_face_neighbors(const or_cells* c, const int32_t* root, uint32_t* nbr) {
    cellmap m;
    uint64_t* keys = (uint64_t*)malloc((c->n ? c->n : 1) * sizeof(uint64_t));
    cell_keys(c, keys);
    cm_build(&m, c, keys);
    static const int F[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    for (size_t i = 0; i < c->n; ++i) {
        int L = c->level[i];
        for (int f = 0; f < 6; ++f) {
            int32_t px = c->x[i] + F[f][0], py = c->y[i] + F[f][1], pz = c->z[i] + F[f][2];
            uint32_t r = OR_INVALID;
            if (in_domain(L, root, px, py, pz)) {
                for (int l = L; l >= 0; --l) {
                    int sh = L - l;
                    uint32_t j = cm_cell(&m, l, px >> sh, py >> sh, pz >> sh);
                    if (j != OR_INVALID) { r = j; break; }
                }
                if (r == OR_INVALID) {
                    for (int q = 0; q < 8; ++q) {
                        if (cm_cell(&m, L + 1, 2 * px + (q & 1), 2 * py + ((q >> 1) & 1),
                                    2 * pz + ((q >> 2) & 1)) != OR_INVALID) {
                            r = 0xFFFFFFFEu;
                            break;
                        }
                    }
                }
            }
            nbr[6 * i + f] = r;
        }
    }
    cm_free(&m);
    free(keys);
}

/* ------------------------------------------------------------------------ */
/* a10: split.  split_cells.cl:15-50 predicate (flag>0 && FLUID && level <
 * max_level) with the exclusive prefix of SplitEngine.cpp:150-161 giving the
 * child block start 8*rank; children per split_cells.cl:53-106. */
size_t or_split_count(const or_cells* c, int max_level, uint32_t* child_start) {
    size_t r = 0;
    for (size_t i = 0; i < c->n; ++i) {
        int ok = c->flag[i] > 0 && c->state[i] == 0 && c->level[i] < max_level;
        child_start[i] = ok ? (uint32_t)(8 * r) : OR_INVALID;
        r += (size_t)ok;
    }
    return r;
}

void or_split_generate(const or_cells* c, const uint32_t* child_start, or_cells* ch) {
    for (size_t i = 0; i < c->n; ++i) {
        uint32_t s = child_start[i];
        if (s == OR_INVALID) continue;
        for (int q = 0; q < 8; ++q) {
            size_t k = (size_t)s + (size_t)q;
            ch->x[k] = (c->x[i] << 1) | (q & 1);
            ch->y[k] = (c->y[i] << 1) | ((q >> 1) & 1);
            ch->z[k] = (c->z[i] << 1) | ((q >> 2) & 1);
            ch->level[k] = (uint8_t)(c->level[i] + 1);
            ch->state[k] = c->state[i];
            ch->material[k] = c->material[i];
            ch->flag[k] = 0;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* a11: merge.  A leader is a FLUID cell with flag == -1 and all-even
 * coordinates that is the smallest index holding its (level, coords); its
 * group forms when the smallest-index holders of all 8 sibling positions at
 * the same level are FLUID with flag == -1 (merge_cells.cl:11-106).  Group
 * ids follow leader index order (the reference uses atomic_inc). */
static int cube_violates(const or_cells* c, const cellmap* m, const uint8_t* shadow, int L,
                         int32_t px, int32_t py, int32_t pz, int s, const int32_t* root);
static const int D18[18][3];

size_t or_merge_mark(const or_cells* c, int check_balance, const int32_t* root, uint32_t* group_id,
                     uint32_t* group_leader) {
    cellmap m;
    uint64_t* keys = (uint64_t*)malloc((c->n ? c->n : 1) * sizeof(uint64_t));
    uint8_t* shadow = (uint8_t*)malloc(c->n ? c->n : 1);
    for (size_t i = 0; i < c->n; ++i) shadow[i] = (uint8_t)(c->level[i] + (c->flag[i] > 0 ? 1 : 0));
    cell_keys(c, keys);
    cm_build(&m, c, keys);
    for (size_t i = 0; i < c->n; ++i) group_id[i] = OR_INVALID;
    size_t g = 0;
    for (size_t i = 0; i < c->n; ++i) {
        if (c->flag[i] != -1 || c->state[i] != 0) continue;
        if ((c->x[i] & 1) || (c->y[i] & 1) || (c->z[i] & 1)) continue;
        if (cm_get(&m, keys[i]) != (uint32_t)i) continue;
        uint32_t sib[8];
        int ok = 1;
        for (int q = 0; q < 8 && ok; ++q) {
            uint32_t j = cm_cell(&m, c->level[i], c->x[i] | (q & 1), c->y[i] | ((q >> 1) & 1),
                                 c->z[i] | ((q >> 2) & 1));
            if (j == OR_INVALID || c->flag[j] != -1 || c->state[j] != 0) ok = 0;
            sib[q] = j;
        }
        if (!ok || c->level[i] == 0) continue;
        if (check_balance) {
            /* the parent (level L-1, shadow L-1) must not see a D3Q19
             * neighbour cube holding cells of shadow level > L */
            int L = c->level[i] - 1;
            int32_t px = c->x[i] >> 1, py = c->y[i] >> 1, pz = c->z[i] >> 1;
            int bad = 0;
            for (int d = 0; d < 18 && !bad; ++d) {
                int32_t qx = px + D18[d][0], qy = py + D18[d][1], qz = pz + D18[d][2];
                if (!in_domain(L, root, qx, qy, qz)) continue;
                bad = cube_violates(c, &m, shadow, L, qx, qy, qz, L, root);
            }
            if (bad) continue;
        }
        for (int q = 0; q < 8; ++q) group_id[sib[q]] = (uint32_t)g;
        group_leader[g] = (uint32_t)i;
        ++g;
    }
    cm_free(&m);
    free(keys);
    free(shadow);
    return g;
}

/* create_parent_cells (merge_cells.cl:162-199) */
void or_merge_parents(const or_cells* c, const uint32_t* group_leader, size_t ng, or_cells* p) {
    for (size_t g = 0; g < ng; ++g) {
        uint32_t i = group_leader[g];
        p->x[g] = c->x[i] >> 1;
        p->y[g] = c->y[i] >> 1;
        p->z[g] = c->z[i] >> 1;
        p->level[g] = (uint8_t)(c->level[i] - 1);
        p->state[g] = c->state[i];
        p->material[g] = c->material[i];
        p->flag[g] = 0;
    }
}

/* merge_fields (merge_cells.cl:123-159): children summed in sibling order
 * q = 0..7 (fixed order instead of CAS float adds), times 0.125 for rule 0. */
void or_merge_fields(const uint32_t* group_leader, size_t ng, const or_cells* c,
                     const float* fields, int ncomp, int rule, float* pf) {
    cellmap m;
    uint64_t* keys = (uint64_t*)malloc((c->n ? c->n : 1) * sizeof(uint64_t));
    cell_keys(c, keys);
    cm_build(&m, c, keys);
    for (size_t g = 0; g < ng; ++g) {
        uint32_t i = group_leader[g];
        for (int k = 0; k < ncomp; ++k) {
            float s = 0.0f;
            for (int q = 0; q < 8; ++q) {
                uint32_t j = cm_cell(&m, c->level[i], c->x[i] | (q & 1), c->y[i] | ((q >> 1) & 1),
                                     c->z[i] | ((q >> 2) & 1));
                float v = fields[(size_t)j * ncomp + k];
                s += rule == 0 ? v * 0.125f : v;
            }
            pf[g * ncomp + k] = s;
        }
    }
    cm_free(&m);
    free(keys);
}

/* ------------------------------------------------------------------------ */
/* a12: balance. */
static const int D18[18][3] = {
    {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 0, 1}, {-1, 0, -1},
    {1, 0, -1}, {-1, 0, 1}, {0, 1, 1}, {0, -1, -1}, {0, 1, -1}, {0, -1, 1}};

/* anchor-unit map for mode 0: key = (anchor coords interleaved into hkey at
 * anchor_level) ; value = smallest index. */
static int detect_mode0(const or_cells* c, const cellmap* am, const uint8_t* shadow, size_t i,
                        int M) {
    int s = shadow[i];
    int L = c->level[i];
    if (L > M) return 0;
    int32_t ax = c->x[i] << (M - L), ay = c->y[i] << (M - L), az = c->z[i] << (M - L);
    int32_t size = (s <= M) ? (1 << (M - s)) : 0;
    int32_t tp[6][3] = {{ax - 1, ay, az}, {ax + size, ay, az}, {ax, ay - 1, az},
                        {ax, ay + size, az}, {ax, ay, az - 1}, {ax, ay, az + size}};
    for (int n = 0; n < 6; ++n) {
        for (int l = 0; l <= M; ++l) {
            int32_t msk = ~((1 << (M - l)) - 1);
            int32_t px = tp[n][0] & msk, py = tp[n][1] & msk, pz = tp[n][2] & msk;
            if (px < 0 || py < 0 || pz < 0 || px >= (1 << M) || py >= (1 << M) || pz >= (1 << M))
                continue;
            /* anchors are unique per (anchor, shadow level) lookup: the map
             * holds, for each anchor, the list head; we scan matches. */
            uint64_t k = or_hilbert_encode(px, py, pz, 21);
            uint64_t sidx = (k * 0x9e3779b97f4a7c15ull) & (am->cap - 1);
            int found = 0;
            for (;;) {
                if (am->k[sidx] == OR_EMPTY) break;
                if (am->k[sidx] == k) {
                    uint32_t j = am->v[sidx];
                    if (shadow[j] == l) { found = 1; break; }
                }
                sidx = (sidx + 1) & (am->cap - 1);
            }
            if (found) {
                if (l > s + 1) return 1;
                break;
            }
        }
    }
    return 0;
}


/* Leaves overlapping the level-L cube P (native coords at level L): is any
 * of them at shadow level > s + 1?  Covered-by-coarser-or-equal: no.  Else
 * descend at most two levels (s <= L+1). */
static int cube_violates(const or_cells* c, const cellmap* m, const uint8_t* shadow, int L,
                         int32_t px, int32_t py, int32_t pz, int s, const int32_t* root) {
    for (int l = L; l >= 0; --l) {
        int sh = L - l;
        if (cm_cell(m, l, px >> sh, py >> sh, pz >> sh) != OR_INVALID) return 0;
    }
    for (int q = 0; q < 8; ++q) {
        int32_t cx = 2 * px + (q & 1), cy = 2 * py + ((q >> 1) & 1), cz = 2 * pz + ((q >> 2) & 1);
        uint32_t j = cm_cell(m, L + 1, cx, cy, cz);
        if (j != OR_INVALID) {
            if (shadow[j] > s + 1) return 1;
            continue;
        }
        if (s <= L) return 1; /* descendants at level >= L+2 */
        for (int r = 0; r < 8; ++r) {
            uint32_t g = cm_cell(m, L + 2, 2 * cx + (r & 1), 2 * cy + ((r >> 1) & 1),
                                 2 * cz + ((r >> 2) & 1));
            if (g == OR_INVALID) return 1; /* descendants at level >= L+3 */
            if (shadow[g] > s + 1) return 1;
        }
    }
    (void)c; (void)root;
    return 0;
}

static int detect_mode1(const or_cells* c, const cellmap* m, const uint8_t* shadow, size_t i,
                        const int32_t* root) {
    int L = c->level[i];
    for (int d = 0; d < 18; ++d) {
        int32_t px = c->x[i] + D18[d][0], py = c->y[i] + D18[d][1], pz = c->z[i] + D18[d][2];
        if (!in_domain(L, root, px, py, pz)) continue;
        if (cube_violates(c, m, shadow, L, px, py, pz, shadow[i], root)) return 1;
    }
    return 0;
}

int or_balance(const or_cells* c, int mode, int anchor_level, int max_split_level,
               int max_iters, const int32_t* root, int* converged, uint32_t* total_violations,
               uint32_t* total_marked) {
    size_t n = c->n;
    uint8_t* shadow = (uint8_t*)malloc(n ? n : 1);
    uint8_t* viol = (uint8_t*)malloc(n ? n : 1);
    uint64_t* keys = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    cellmap m;
    if (mode == 0) {
        /* map anchor -> all indices: keep every index in its own slot so the
         * probe chain visits holders in index order (insertion order). */
        uint64_t cap = 1024;
        while (cap < 2 * (uint64_t)n + 16) cap <<= 1;
        m.cap = cap;
        m.k = (uint64_t*)malloc(cap * sizeof(uint64_t));
        m.v = (uint32_t*)malloc(cap * sizeof(uint32_t));
        for (uint64_t s = 0; s < cap; ++s) { m.k[s] = OR_EMPTY; m.v[s] = OR_INVALID; }
        for (size_t i = 0; i < n; ++i) {
            int L = c->level[i];
            if (L > anchor_level) continue;
            uint64_t k = or_hilbert_encode(c->x[i] << (anchor_level - L), c->y[i] << (anchor_level - L),
                                           c->z[i] << (anchor_level - L), 21);
            uint64_t s = (k * 0x9e3779b97f4a7c15ull) & (cap - 1);
            while (m.k[s] != OR_EMPTY) s = (s + 1) & (cap - 1);
            m.k[s] = k; m.v[s] = (uint32_t)i;
        }
    } else {
        cell_keys(c, keys);
        cm_build(&m, c, keys);
    }
    /* Mode 0 starts from a copy of the stored levels (BalanceEnforcer.cpp:
     * 165-167): split flags set before the call count only after the first
     * marking pass re-derives the shadow (:248-254). Mode 1, the lattice rule,
     * counts the requested splits from the start. */
    for (size_t i = 0; i < n; ++i)
        shadow[i] = (uint8_t)(c->level[i] + (mode != 0 && c->flag[i] > 0 ? 1 : 0));
    int it = 0;
    *converged = 0;
    *total_violations = 0;
    *total_marked = 0;
    for (it = 0; it < max_iters;) {
        uint32_t nv = 0, nm = 0;
        for (size_t i = 0; i < n; ++i) {
            viol[i] = 0;
            if (mode == 0 && c->state[i] != 0) continue;
            int v = mode == 0 ? detect_mode0(c, &m, shadow, i, anchor_level)
                              : detect_mode1(c, &m, shadow, i, root);
            viol[i] = (uint8_t)v;
            nv += (uint32_t)v;
        }
        ++it;
        *total_violations += nv;
        if (nv == 0) { *converged = 1; break; }
        for (size_t i = 0; i < n; ++i) {
            if (!viol[i] || c->state[i] != 0 || c->level[i] >= max_split_level) continue;
            if (c->flag[i] <= 0) { c->flag[i] = 1; ++nm; }
        }
        *total_marked += nm;
        if (nm == 0) break;
        for (size_t i = 0; i < n; ++i) shadow[i] = (uint8_t)(c->level[i] + (c->flag[i] > 0 ? 1 : 0));
    }
    cm_free(&m);
    free(shadow); free(viol); free(keys);
    return it;
}

/* ------------------------------------------------------------------------ */
/* a6: compaction + append. */
static void copy_cell(or_cells* d, size_t k, const or_cells* s, size_t i) {
    d->x[k] = s->x[i]; d->y[k] = s->y[i]; d->z[k] = s->z[i];
    d->level[k] = s->level[i]; d->state[k] = s->state[i];
    d->material[k] = s->material[i]; d->flag[k] = 0;
}

size_t or_compact_append(const or_cells* c, const uint32_t* group_id, const or_cells* ch,
                         const or_cells* par, or_cells* out) {
    size_t k = 0;
    for (size_t i = 0; i < c->n; ++i) {
        int keep = !(c->flag[i] > 0) && group_id[i] == OR_INVALID;
        if (keep) copy_cell(out, k++, c, i);
    }
    for (size_t i = 0; ch && i < ch->n; ++i) copy_cell(out, k++, ch, i);
    for (size_t i = 0; par && i < par->n; ++i) copy_cell(out, k++, par, i);
    return k;
}

/* ------------------------------------------------------------------------ */
uint64_t or_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* Synthetic cells: level ~ U{0..max_level}, coords ~ U[0, 2^level) from
 * independent splitmix64 streams keyed by (seed, index). */
void or_gen_cells(uint64_t seed, size_t n, int max_level, int32_t* x, int32_t* y, int32_t* z,
                  uint8_t* level) {
    for (size_t i = 0; i < n; ++i) {
        uint64_t b = seed * 0x100000001B3ull + (uint64_t)i * 4;
        uint64_t r0 = or_splitmix64(b), r1 = or_splitmix64(b + 1), r2 = or_splitmix64(b + 2),
                 r3 = or_splitmix64(b + 3);
        int L = (int)((r0 >> 32) % (uint64_t)(max_level + 1));
        uint32_t m = (1u << L) - 1u;
        level[i] = (uint8_t)L;
        x[i] = (int32_t)((uint32_t)(r1 >> 32) & m);
        y[i] = (int32_t)((uint32_t)(r2 >> 32) & m);
        z[i] = (int32_t)((uint32_t)(r3 >> 32) & m);
    }
}

/* Flags of the AMR micro-benchmark, identical to fl_gen_flags: split with
 * probability p_split (stream seed ^ 0x5A17), complete sibling groups whose
 * leader draws below p_merge (stream seed ^ 0x3E26) get -1 on members still
 * at 0.  Thresholds are p * 2^64 compared with the raw 64-bit draw. */
void or_gen_flags(const or_cells* c, uint64_t seed, uint64_t split_thresh, uint64_t merge_thresh) {
    size_t n = c->n;
    for (size_t i = 0; i < n; ++i)
        c->flag[i] = or_splitmix64((seed ^ 0x5A17ull) * 0x100000001B3ull + i) < split_thresh ? 1 : 0;
    cellmap m;
    uint64_t* keys = (uint64_t*)malloc((n ? n : 1) * sizeof(uint64_t));
    cell_keys(c, keys);
    cm_build(&m, c, keys);
    for (size_t i = 0; i < n; ++i) {
        int32_t x = c->x[i], y = c->y[i], z = c->z[i];
        int L = c->level[i];
        if (L == 0 || (x & 1) || (y & 1) || (z & 1)) continue;
        if (cm_cell(&m, L, x, y, z) != (uint32_t)i) continue;
        if (or_splitmix64((seed ^ 0x3E26ull) * 0x100000001B3ull + i) >= merge_thresh) continue;
        uint32_t sib[8];
        int ok = 1;
        for (int q = 0; q < 8 && ok; ++q) {
            sib[q] = cm_cell(&m, L, x | (q & 1), y | ((q >> 1) & 1), z | ((q >> 2) & 1));
            ok = sib[q] != OR_INVALID;
        }
        if (!ok) continue;
        for (int q = 0; q < 8; ++q)
            if (c->flag[sib[q]] == 0) c->flag[sib[q]] = -1;
    }
    cm_free(&m);
    free(keys);
}
