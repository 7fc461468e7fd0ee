/*
 * stc_pipeline.hip — per-timestep co-occurrence pipeline on gfx950 and the
 * C ABI of include/stcdb.h.
 *
 * One timestep (the work of process_with_gpu, src/cuda/geometry.cu:1573-2218):
 *
 *  1. prep      per point: QTree cell (16384 x 16384 grid over the map, as
 *               cuda_build_qtree geometry.cu:1270-1279), its Morton code, and
 *               the point's longitude scale dpm_lon(y) (stc_geo.h).
 *  2. sort      stable radix sort of the Morton codes.
 *  3. tree      top-down QTree build over the sorted codes: a block splits when
 *               it holds more than grid_capacity points; blocks at the finest
 *               level split once more into four halves. This is the QTree that
 *               the reference's bottom-up merge builds (geometry.cu:1295-1393),
 *               with the same node box expressions. Counts come from binary
 *               searches on the sorted codes, not a 2 GiB dense histogram.
 *  4. partition per point: descend by mid_x/mid_y to its leaf and record the
 *               deepest ancestor whose top/right edge is at least the border
 *               buffer away (cuda_partition, geometry.cu:221-271).
 *  5. leaves    stable sort of points by leaf: each leaf's members in id order,
 *               stored in leaf order as SoA x, y, dpm_lon, id. A leaf keeps its
 *               lowest grid_amplify * grid_capacity ids as targets.
 *  6. lookup    border points search the subtree under their last valid node
 *               for leaves within reach on their top or right (cuda_filtering,
 *               geometry.cu:295-392): a counting pass, a prefix sum, and an
 *               emitting pass, so no global atomics and a deterministic order.
 *  7. refine    one wavefront per check unit; lanes stride over the target
 *               leaf's members with coalesced SoA loads; hits are compacted with
 *               a ballot and one atomic per wavefront (cuda_refinement,
 *               geometry.cu:424-497).
 *  8. meetings  open-addressing table keyed by (oid1 << 32 | oid2) with 64-bit
 *               CAS: insert or extend, MBR by ordered-integer atomic min/max.
 *  9. identify  full-table scan: meetings not seen this step expire; those that
 *               lasted at least min_meet_time and stay inside the map emit two
 *               records each (cuda_identify_meetings, geometry.cu:595-737).
 *               Expired entries keep their key (so probe chains stay intact)
 *               until the table is rehashed.
 * 10. flush     when the memtable exceeds kv_restriction (stc_flush.hip).
 */
#include "stc_internal.h"
#include "stc_geo.h"
#include "stc_gen.h"
#include "stc_flush.h"
#include "stc_table.h"
#include "stc_query.h"
#include "../../include/stcdb.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#define ONE_DIM 16384u
#define NLEVELS 15
#define T_BRANCH 1u
#define T_LEAF 2u
#define STACK_DEPTH 96

static thread_local std::string g_last_fail;

int stc_hip_fail(hipError_t e, const char *expr, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", expr, hipGetErrorString(e), file, line);
    g_last_fail = buf;
    return STC_E_HIP;
}

/* ------------------------------------------------------------------ */
/* device structures                                                   */
/* ------------------------------------------------------------------ */

struct Node {                 /* QTSchema (src/index/QTree.h:49-62) without counters */
    double lo0, lo1, hi0, hi1;
    double mid_x, mid_y;
    uint32_t child[4];
    uint32_t type, leaf;
    uint32_t code, level;     /* Morton code and level of the block (branches) */
};

struct FrontierItem { uint32_t node, code; };

struct PairHit {              /* 24 B: key and the float MBR of the two positions */
    unsigned long long key;
    float mbr[4];
};

struct Counters {             /* device-side counters, copied back once per sync point */
    unsigned long long units;
    unsigned long long candidates;
    unsigned long long pairs;
    unsigned long long ambiguous;
    unsigned long long pairs_listed;
    unsigned long long new_keys;
    unsigned long long continuing;
    unsigned long long ambiguous_nodes;
    unsigned long long live, expired, retired, freed, tombs;
    unsigned long long kv_count;
    uint32_t nodes, leaves;
    uint32_t frontier[NLEVELS + 1];
    uint32_t bad_points;
    uint32_t border;
    uint32_t overflow;
    uint32_t sid_count, oversize_oid_count;
    uint32_t table_fail;
};


__device__ inline uint32_t part1by1(uint32_t x) {
    x &= 0x0000ffff;
    x = (x | (x << 8)) & 0x00FF00FF;
    x = (x | (x << 4)) & 0x0F0F0F0F;
    x = (x | (x << 2)) & 0x33333333;
    x = (x | (x << 1)) & 0x55555555;
    return x;
}
__device__ inline uint32_t compact1by1(uint32_t x) {
    x &= 0x55555555;
    x = (x | (x >> 1)) & 0x33333333;
    x = (x | (x >> 2)) & 0x0F0F0F0F;
    x = (x | (x >> 4)) & 0x00FF00FF;
    x = (x | (x >> 8)) & 0x0000FFFF;
    return x;
}

/* home slot of a key in a table of nb buckets: a 64-bit mix of the key
 * mapped onto [0, nb) by a multiply-shift (nb need not be a power of two) */
__device__ inline uint64_t hash_slot(unsigned long long key, uint64_t nb) {
    uint64_t z = key + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint64_t)(((unsigned __int128)z * nb) >> 64);
}

/* ------------------------------------------------------------------ */
/* kernels: prep, tree, partition                                      */
/* ------------------------------------------------------------------ */

struct MapBox { double lo0, lo1, hi0, hi1; };

__global__ void k_prep(const double2 *__restrict__ xy, uint32_t n, MapBox m, uint32_t *__restrict__ codes,
                       uint32_t *__restrict__ idx, double *__restrict__ dlon, double *__restrict__ idlon,
                       Counters *cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double2 p = xy[i];
    double fx = (p.x - m.lo0) / (m.hi0 - m.lo0) * ONE_DIM;
    double fy = (p.y - m.lo1) / (m.hi1 - m.lo1) * ONE_DIM;
    uint32_t code = 0;
    if (fx > -1.0 && fx < (double)ONE_DIM && fy > -1.0 && fy < (double)ONE_DIM) {
        code = part1by1((uint32_t)fx) | (part1by1((uint32_t)fy) << 1);
    } else {
        atomicAdd(&cnt->bad_points, 1u);
    }
    codes[i] = code;
    idx[i] = i;
    double d = stc_dpm_lon(p.y);
    dlon[i] = d;
    idlon[i] = 1.0 / d;
}

/* box of block (x, y) with xdim blocks per side (geometry.cu:1306-1311) */
__device__ inline void block_box(Node &nd, MapBox m, uint32_t x, uint32_t y, uint32_t xdim) {
    nd.lo0 = m.lo0 + (double)x * (m.hi0 - m.lo0) / (double)xdim;
    nd.lo1 = m.lo1 + (double)y * (m.hi1 - m.lo1) / (double)xdim;
    nd.hi0 = m.lo0 + (double)(x + 1) * (m.hi0 - m.lo0) / (double)xdim;
    nd.hi1 = m.lo1 + (double)(y + 1) * (m.hi1 - m.lo1) / (double)xdim;
    nd.mid_x = (nd.lo0 + nd.hi0) / 2;
    nd.mid_y = (nd.lo1 + nd.hi1) / 2;
}

/* leaf child i of parent, by halving (geometry.cu:1318-1334) */
__device__ inline void halving_leaf(Node &nd, const Node &p, int i) {
    double xhalf = p.mid_x - p.lo0;
    double yhalf = p.mid_y - p.lo1;
    nd.lo0 = p.lo0 + (double)(i % 2 == 1) * xhalf;
    nd.lo1 = p.lo1 + (double)(i / 2 == 1) * yhalf;
    nd.hi0 = p.mid_x + (double)(i % 2 == 1) * xhalf;
    nd.hi1 = p.mid_y + (double)(i / 2 == 1) * yhalf;
    nd.mid_x = (nd.lo0 + nd.hi0) / 2;
    nd.mid_y = (nd.lo1 + nd.hi1) / 2;
}

__device__ inline uint32_t lower_bound_u32(const uint32_t *a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__global__ void k_tree_root(Node *nodes, MapBox m, FrontierItem *frontier, Counters *cnt, uint32_t n, uint32_t cap) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Node nd;
    block_box(nd, m, 0, 0, 1);
    nd.type = T_BRANCH;
    nd.leaf = 0;
    nd.code = 0;
    nd.level = NLEVELS - 1;
    for (int c = 0; c < 4; c++) nd.child[c] = 0;
    nodes[0] = nd;
    cnt->nodes = 1;
    cnt->leaves = 0;
    for (int l = 0; l <= NLEVELS; l++) cnt->frontier[l] = 0;
    if (n > cap) {
        frontier[0].node = 0;
        frontier[0].code = 0;
        cnt->frontier[NLEVELS - 1] = 1;
    }
}

/* expand the branch blocks of level L (L >= 1): each of the four sub-blocks
 * becomes a branch when it holds more than cap points, else a halving leaf */
__global__ void k_tree_level(Node *nodes, MapBox m, const uint32_t *sorted_codes, uint32_t n, uint32_t cap, int L,
                             const FrontierItem *fin, FrontierItem *fout, Counters *cnt, uint32_t node_cap,
                             uint32_t front_cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt->frontier[L]) return;
    FrontierItem f = fin[i];
    Node parent = nodes[f.node];
    int Lc = L - 1;
    uint32_t xdim = ONE_DIM >> Lc;
    uint32_t child_ids[4];
    uint32_t nb = 0;
    bool is_branch[4];
    for (int c = 0; c < 4; c++) {
        uint32_t sub = (f.code << 2) | (uint32_t)c;
        uint32_t lo = sub << (2 * Lc);
        uint32_t a = lower_bound_u32(sorted_codes, n, lo);
        uint32_t b = lower_bound_u32(sorted_codes, n, lo + (1u << (2 * Lc)));
        is_branch[c] = (b - a) > cap;
        nb += is_branch[c];
    }
    uint32_t base = atomicAdd(&cnt->nodes, 4u);
    if (base + 4 > node_cap) { atomicAdd(&cnt->table_fail, 1u); return; }
    uint32_t nleaf = 4 - nb;
    uint32_t lbase = nleaf ? atomicAdd(&cnt->leaves, nleaf) : 0;
    uint32_t fbase = nb ? atomicAdd(&cnt->frontier[Lc], nb) : 0;
    if (fbase + nb > front_cap) { atomicAdd(&cnt->table_fail, 1u); return; }
    for (int c = 0; c < 4; c++) {
        uint32_t id = base + c;
        child_ids[c] = id;
        Node nd;
        uint32_t sub = (f.code << 2) | (uint32_t)c;
        for (int k = 0; k < 4; k++) nd.child[k] = 0;
        if (is_branch[c]) {
            block_box(nd, m, compact1by1(sub), compact1by1(sub >> 1), xdim);
            nd.type = T_BRANCH;
            nd.leaf = 0;
            nd.code = sub;
            nd.level = Lc;
            fout[fbase].node = id;
            fout[fbase].code = sub;
            fbase++;
        } else {
            halving_leaf(nd, parent, c);
            nd.type = T_LEAF;
            nd.leaf = lbase++;
            nd.code = sub;
            nd.level = Lc;
        }
        nodes[id] = nd;
    }
    for (int c = 0; c < 4; c++) nodes[f.node].child[c] = child_ids[c];
}

/* finest level: each overfull cell becomes a branch of four halving leaves */
__global__ void k_tree_finest(Node *nodes, const FrontierItem *fin, Counters *cnt, uint32_t node_cap) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt->frontier[0]) return;
    FrontierItem f = fin[i];
    Node parent = nodes[f.node];
    uint32_t base = atomicAdd(&cnt->nodes, 4u);
    if (base + 4 > node_cap) { atomicAdd(&cnt->table_fail, 1u); return; }
    uint32_t lbase = atomicAdd(&cnt->leaves, 4u);
    for (int c = 0; c < 4; c++) {
        Node nd;
        halving_leaf(nd, parent, c);
        for (int k = 0; k < 4; k++) nd.child[k] = 0;
        nd.type = T_LEAF;
        nd.leaf = lbase + c;
        nd.code = f.code;
        nd.level = 0;
        nodes[base + c] = nd;
        nodes[f.node].child[c] = base + c;
    }
}

__global__ void k_leaf_nodes(const Node *nodes, uint32_t n_nodes, uint32_t *leaf_node) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    if (nodes[i].type == T_LEAF) leaf_node[nodes[i].leaf] = i;
}

__global__ void k_partition(const double2 *__restrict__ xy, uint32_t n, const Node *__restrict__ nodes, double xbuf,
                            double ybuf, uint32_t *__restrict__ pt_leaf, uint32_t *__restrict__ pt_lv,
                            uint32_t *__restrict__ leaf_count, Counters *cnt) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool border = false;
    if (p < n) {
        double2 q = xy[p];
        uint32_t cur = 0, lv = 0;
        for (int depth = 0; depth < 64; depth++) {
            const Node &nd = nodes[cur];
            int loc = (q.y > nd.mid_y) * 2 + (q.x > nd.mid_x);
            cur = nd.child[loc];
            const Node &c = nodes[cur];
            if (q.x + xbuf < c.hi0 && q.y + ybuf < c.hi1) lv = cur;
            if (c.type == T_LEAF) break;
        }
        uint32_t leaf = nodes[cur].leaf;
        pt_leaf[p] = leaf;
        border = (lv != cur);
        pt_lv[p] = border ? lv : 0xFFFFFFFFu;
        atomicAdd(&leaf_count[leaf], 1u);
    }
    uint64_t b = __ballot(border);
    if ((threadIdx.x & (STC_WAVE - 1)) == 0 && b) atomicAdd(&cnt->border, (uint32_t)__popcll(b));
}

__global__ void k_iota(uint32_t *a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

/* leaf-ordered SoA and each point's rank within its leaf */
__global__ void k_gather(const double2 *__restrict__ xy, const double *__restrict__ dlon,
                         const uint32_t *__restrict__ lp, const uint32_t *__restrict__ pt_leaf,
                         const uint32_t *__restrict__ leaf_start, uint32_t n, double *__restrict__ lx,
                         double *__restrict__ ly, double *__restrict__ ld, uint32_t *__restrict__ pt_rank) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t p = lp[j];
    double2 q = xy[p];
    lx[j] = q.x;
    ly[j] = q.y;
    ld[j] = dlon[p];
    pt_rank[p] = j - leaf_start[pt_leaf[p]];
}

__global__ void k_overflow(const uint32_t *leaf_count, uint32_t nl, uint32_t buf, Counters *cnt) {
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t over = 0;
    if (l < nl && leaf_count[l] > buf) over = leaf_count[l] - buf;
    if (over) atomicAdd(&cnt->overflow, over);
}

/* ------------------------------------------------------------------ */
/* border lookup (two passes)                                          */
/* ------------------------------------------------------------------ */

template <bool EMIT>
__global__ void k_lookup(const double2 *__restrict__ xy, const double *__restrict__ dlon, uint32_t n,
                         const Node *__restrict__ nodes, const uint32_t *__restrict__ pt_lv, double r, double band,
                         uint32_t *__restrict__ count, const uint32_t *__restrict__ offset,
                         uint2 *__restrict__ units, Counters *cnt) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint32_t lv = pt_lv[p];
    if (lv == 0xFFFFFFFFu) {
        if (!EMIT) count[p] = 0;
        return;
    }
    double2 q = xy[p];
    double dl = dlon[p];
    uint32_t stk[STACK_DEPTH];
    int sp = 0;
    stk[sp++] = lv;
    uint32_t found = 0, amb = 0;
    uint32_t out = EMIT ? offset[p] : 0;
    while (sp > 0) {
        uint32_t cur = stk[--sp];
        const Node &nd = nodes[cur];
        for (int i = 0; i < 4; i++) {
            uint32_t c = nd.child[i];
            const Node &ch = nodes[c];
            double d = stc_box_dist(ch.lo0, ch.lo1, ch.hi0, ch.hi1, q.x, q.y, dl);
            if (fabs(d - r) <= band) amb++;
            if (d <= r) {
                if (ch.type == T_LEAF) {
                    if (q.y < ch.lo1 || (q.y < ch.hi1 && q.x < ch.lo0)) {
                        if (EMIT) units[out + found] = make_uint2(p, ch.leaf);
                        found++;
                    }
                } else if (sp < STACK_DEPTH) {
                    stk[sp++] = c;
                }
            }
        }
    }
    if (!EMIT) {
        count[p] = found;
        if (amb) atomicAdd(&cnt->ambiguous_nodes, (unsigned long long)amb);
    }
}

/* ------------------------------------------------------------------ */
/* meeting table                                                       */
/* ------------------------------------------------------------------ */

#define MAX_PROBES 1024u

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = STC_WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, STC_WAVE);
    return v;
}

__global__ void k_table_clear(Bucket *t, uint64_t nb) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    Bucket b;
    b.key = EMPTY_KEY;
    b.start = 0;
    b.end = 0xFFFFFFFFu;
    b.mbr[0] = b.mbr[1] = 0xFFFFFFFFu;
    b.mbr[2] = b.mbr[3] = 0u;
    t[i] = b;
}

/* linear probing from the key's hash slot; claims an empty bucket with a
 * 64-bit CAS. Returns null after MAX_PROBES probes (table too full). */
__device__ inline Bucket *table_find_or_claim(Bucket *t, uint64_t nb, unsigned long long key, bool *claimed) {
    uint64_t h = hash_slot(key, nb);
    for (uint32_t probe = 0; probe < MAX_PROBES; probe++) {
        Bucket *b = &t[h];
        unsigned long long k = __atomic_load_n(&b->key, __ATOMIC_RELAXED);
        if (k == key) { *claimed = false; return b; }
        if (key_free(k)) {
            unsigned long long prev = atomicCAS(&b->key, k, key);
            if (prev == k) { *claimed = true; return b; }
            if (prev == key) { *claimed = false; return b; }
        }
        h = (h + 1 == nb) ? 0 : h + 1;
    }
    return nullptr;
}

/* Insert or extend one pair seen at `now` (geometry.cu:469-494, without the
 * 5-probe drop). Within a step a key is updated only by the threads that
 * found this very pair, and all of them carry the same two positions, so the
 * bucket is read once and rewritten with plain stores: concurrent writers
 * store identical values and the result is the same in any order. Only
 * claiming an empty bucket needs a CAS. Returns -1 when the table is full,
 * 1 for a new key, 2 when it continues a meeting seen at now - 1, else 0. */
__device__ inline int table_upsert(Bucket *t, uint64_t nb, unsigned long long key, const float *mbr, uint32_t now) {
    uint64_t h = hash_slot(key, nb);
    for (uint32_t probe = 0; probe < MAX_PROBES; probe++) {
        Bucket *b = &t[h];
        Bucket cur; /* the key half first; the MBR half only for the matching bucket */
        {
            const uint4 lo4 = *(const uint4 *)b; /* key, start, end */
            cur.key = ((unsigned long long)lo4.y << 32) | lo4.x;
            cur.start = lo4.z;
            cur.end = lo4.w;
        }
        if (cur.key == key) {
            const uint4 hi4 = *((const uint4 *)b + 1);
            cur.mbr[0] = hi4.x; cur.mbr[1] = hi4.y; cur.mbr[2] = hi4.z; cur.mbr[3] = hi4.w;
        }
        bool claimed = false, raced = false;
        if (key_free(cur.key)) {
            unsigned long long prev = atomicCAS(&b->key, cur.key, key);
            if (prev == cur.key) {
                claimed = true;
            } else if (prev == key) { /* another thread claimed this very pair this step */
                cur = *b;
                raced = true;
            } else {
                h = (h + 1 == nb) ? 0 : h + 1;
                continue;
            }
        } else if (cur.key != key) {
            h = (h + 1 == nb) ? 0 : h + 1;
            continue;
        }
        if (claimed) {
            b->start = now;
            b->end = now;
            b->mbr[0] = f2o(mbr[0]);
            b->mbr[1] = f2o(mbr[1]);
            b->mbr[2] = f2o(mbr[2]);
            b->mbr[3] = f2o(mbr[3]);
            return 1;
        }
        uint32_t old_end = cur.end;
        if (old_end != now && old_end != now - 1) b->start = now; /* restarted (retired key) or racing a claim */
        b->end = now;
        {
            const uint32_t n0 = min(cur.mbr[0], f2o(mbr[0])), n1 = min(cur.mbr[1], f2o(mbr[1]));
            const uint32_t n2 = max(cur.mbr[2], f2o(mbr[2])), n3 = max(cur.mbr[3], f2o(mbr[3]));
            if (n0 != cur.mbr[0] || n1 != cur.mbr[1] || n2 != cur.mbr[2] || n3 != cur.mbr[3])
                *((uint4 *)b + 1) = make_uint4(n0, n1, n2, n3); /* only a grown MBR is written */
        }
        /* a free bucket's end field reads 0xFFFFFFFF until its claimer writes it: at
         * now == 0 that equals now - 1, so nothing continues into step 0 */
        return (!raced && now != 0 && old_end == now - 1) ? 2 : 0;
    }
    return -1;
}

/* insert a list of pairs (host-resolved ambiguous pairs) */
__global__ void k_table_insert(Bucket *t, uint64_t nb, const PairHit *__restrict__ pairs, uint64_t np, uint32_t now,
                               Counters *cnt) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    PairHit h = pairs[i];
    int rc = table_upsert(t, nb, h.key, h.mbr, now);
    if (rc < 0) atomicAdd(&cnt->table_fail, 1u);
    if (rc == 1) atomicAdd(&cnt->new_keys, 1ull);
    if (rc == 2) atomicAdd(&cnt->continuing, 1ull);
}

/* Persisted meeting units back to pair keys: the Cantor key of
 * (min pid, max pid) (geometry.cu:453-455) inverted exactly, by integer search
 * around the double estimate of getpid1/getpid2 (cuda_util.cuh:55-70). A key
 * that does not name two distinct objects below n sets bad. */
__global__ void k_units_decode(const stc_meeting_unit *__restrict__ u, uint32_t nu, uint32_t n,
                               uint64_t *__restrict__ keys, uint32_t *__restrict__ idx, Counters *cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    const uint64_t z = u[i].key;
    uint64_t w = (uint64_t)floor((sqrt(8.0 * (double)z + 1.0) - 1.0) / 2.0);
    while (w > 0 && (unsigned __int128)w * (w + 1) / 2 > z) w--;
    while ((unsigned __int128)(w + 1) * (w + 2) / 2 <= z) w++;
    const uint64_t y = z - (uint64_t)((unsigned __int128)w * (w + 1) / 2), x = w - y;
    if (!(x < y && y < n)) {
        atomicAdd(&cnt->bad_points, 1u);
        keys[i] = 0;
    } else {
        keys[i] = (x << 32) | y;
    }
    idx[i] = i;
}

/* the units in pair order as PairHits; a pair named twice sets table_fail's
 * neighbour flag (bad_points) */
__global__ void k_units_gather(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                               const stc_meeting_unit *__restrict__ u, uint32_t nu, PairHit *__restrict__ out,
                               Counters *cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    if (i > 0 && keys[i] == keys[i - 1]) atomicAdd(&cnt->bad_points, 1u);
    PairHit h;
    h.key = keys[i];
    const stc_meeting_unit m = u[idx[i]];
    h.mbr[0] = m.mbr[0]; h.mbr[1] = m.mbr[1]; h.mbr[2] = m.mbr[2]; h.mbr[3] = m.mbr[3];
    out[i] = h;
}

/* ------------------------------------------------------------------ */
/* refinement fused with the table insert                               */
/* ------------------------------------------------------------------ */

__global__ __launch_bounds__(256) void k_refine(
    const double2 *__restrict__ xy, const double *__restrict__ dlon, const double *__restrict__ idlon, uint32_t n,
    const uint32_t *__restrict__ pt_leaf, const uint32_t *__restrict__ pt_rank, const uint32_t *__restrict__ leaf_start,
    const uint32_t *__restrict__ leaf_count, const uint32_t *__restrict__ lp, const double *__restrict__ lx,
    const double *__restrict__ ly, const uint2 *__restrict__ units, uint64_t n_units_total, uint32_t buf, double r,
    double band, double r2_lo, double r2_hi, Bucket *table, uint64_t nb, uint32_t now,
    PairHit *__restrict__ pairs, uint64_t pair_cap,
    uint2 *__restrict__ amb, uint32_t amb_cap, Counters *cnt) {
    const int lane = threadIdx.x & (STC_WAVE - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / STC_WAVE;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / STC_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (STC_WAVE - lane));
    unsigned long long cand = 0, hits = 0, fresh = 0, fails = 0, cont = 0;
    for (uint64_t u = wid; u < n_units_total; u += nw) {
        uint32_t p, leaf, begin;
        if (u < n) {
            p = (uint32_t)u;
            leaf = pt_leaf[p];
            begin = pt_rank[p] + 1;
        } else {
            uint2 un = units[u - n];
            p = un.x;
            leaf = un.y;
            begin = 0;
        }
        uint32_t s = leaf_count[leaf];
        if (s > buf) s = buf;
        if (begin >= s) continue;
        uint32_t st = leaf_start[leaf];
        double2 pp = xy[p];
        double pd = dlon[p];
        double pinv = idlon[p];
        float pfx = (float)pp.x, pfy = (float)pp.y;
        for (uint32_t m0 = begin; m0 < s; m0 += STC_WAVE) {
            uint32_t m = m0 + lane;
            bool hit = false, ambg = false;
            uint32_t q = 0;
            double qx = 0, qy = 0;
            if (m < s) {
                qx = lx[st + m];
                qy = ly[st + m];
                /* squared distance with reciprocal multiplies; it is within
                 * 1.2e-15 (relative) of the formula's, so only a band of 1e-12
                 * around r^2 needs the exact divisions and square root */
                double ax = (pp.x - qx) * pinv;
                double ay = (pp.y - qy) * STC_INV_DPM_LAT;
                double s2 = ax * ax + ay * ay;
                if (s2 < r2_lo) {
                    hit = true;
                } else if (s2 <= r2_hi) {
                    double d = stc_point_dist(pp.x, pp.y, qx, qy, pd);
                    if (fabs(d - r) <= band) ambg = true;
                    else hit = d <= r;
                }
                if (hit || ambg) q = lp[st + m];
                cand++;
            }
            if (hit) {
                uint32_t a = p < q ? p : q, b = p < q ? q : p;
                unsigned long long key = ((unsigned long long)a << 32) | b;
                float qfx = (float)qx, qfy = (float)qy;
                float mbr[4] = {fminf(pfx, qfx), fminf(pfy, qfy), fmaxf(pfx, qfx), fmaxf(pfy, qfy)};
                /* once the table has overflowed this pass only counts; the host
                 * grows the table and repeats the pass */
                int rc = __atomic_load_n(&cnt->table_fail, __ATOMIC_RELAXED) ? -1
                                                                             : table_upsert(table, nb, key, mbr, now);
                hits++;
                if (rc == 1) fresh++;
                if (rc == 2) cont++;
                if (rc < 0) fails++;
                if (pairs) { /* inspection copy of the step's pairs */
                    unsigned long long k = atomicAdd(&cnt->pairs_listed, 1ull);
                    if (k < pair_cap) {
                        PairHit h;
                        h.key = key;
                        h.mbr[0] = mbr[0]; h.mbr[1] = mbr[1]; h.mbr[2] = mbr[2]; h.mbr[3] = mbr[3];
                        pairs[k] = h;
                    }
                }
            }
            if (ambg) {
                unsigned long long k = atomicAdd(&cnt->ambiguous, 1ull);
                if (k < amb_cap) amb[k] = make_uint2(p, q);
            }
        }
    }
    (void)lt;
    cand = wave_sum_u64(cand);
    hits = wave_sum_u64(hits);
    fresh = wave_sum_u64(fresh);
    fails = wave_sum_u64(fails);
    cont = wave_sum_u64(cont);
    if (lane == 0) {
        if (cont) atomicAdd(&cnt->continuing, cont);
        if (cand) atomicAdd(&cnt->candidates, cand);
        if (hits) atomicAdd(&cnt->pairs, hits);
        if (fresh) atomicAdd(&cnt->new_keys, fresh);
        if (fails) atomicAdd(&cnt->table_fail, (uint32_t)fails);
    }
}

/* float_to_uint (cuda_util.cuh:115-119), saturating as the reference GPU converts */


struct MemtableDev {
    uint64_t *keys;        /* the reference's 128-bit key as (lo, hi) words:
                              hi = pid << 26 | target, lo = duration << 32 | end offset */
    float4 *boxes;
    uint64_t cap;
    uint32_t *same_pid_count;
    uint4 *o_boxs;         /* quantized per-object MBR: low0, low1, high0, high1 */
    uint16_t *sids;
};

#define ID_STAGE 128 /* per-wavefront staging of records to emit */

/* write the two memtable records of each staged meeting (geometry.cu:641-705) */
__device__ inline void emit_staged(const Bucket *st, uint32_t staged, int lane, uint32_t max_meet,
                                   uint32_t end_time_min, MemtableDev mt, Counters *cnt) {
    unsigned long long kbase = 0;
    if (lane == 0) kbase = atomicAdd(&cnt->kv_count, 2ull * staged);
    kbase = __shfl(kbase, 0, STC_WAVE);
    for (uint32_t e = lane; e < staged; e += STC_WAVE) {
        Bucket b = st[e];
        float lo0 = o2f(b.mbr[0]), lo1 = o2f(b.mbr[1]), hi0 = o2f(b.mbr[2]), hi1 = o2f(b.mbr[3]);
        uint32_t duration = b.end - b.start;
        uint32_t pid = (uint32_t)(b.key >> 32), target = (uint32_t)b.key;
        if (duration > max_meet) {
            atomicAdd(&cnt->oversize_oid_count, 1u);
            mt.sids[pid] = 1;
            mt.sids[target] = 1;
        }
        unsigned long long k = kbase + 2ull * e;
        uint32_t q0 = f2q(lo0), q1 = f2q(lo1), q2 = f2q(hi0), q3 = f2q(hi1);
        for (int sdir = 0; sdir < 2; sdir++) {
            uint32_t a = sdir ? target : pid, c = sdir ? pid : target;
            if (k + sdir < mt.cap) {
                mt.keys[2 * (k + sdir)] = ((uint64_t)duration << 32) | (uint32_t)(b.end - end_time_min);
                mt.keys[2 * (k + sdir) + 1] = ((uint64_t)a << 26) | (uint64_t)c;
                mt.boxes[k + sdir] = make_float4(lo0, lo1, hi0, hi1);
            }
            if (atomicAdd(&mt.same_pid_count[a], 1u) == 0) atomicAdd(&cnt->sid_count, 1u);
            atomicMin(&mt.o_boxs[a].x, q0);
            atomicMax(&mt.o_boxs[a].z, q2);
            atomicMin(&mt.o_boxs[a].y, q1);
            atomicMax(&mt.o_boxs[a].w, q3);
        }
    }
}

__device__ inline uint64_t slot_next(uint64_t j, uint64_t nb) { return (j + 1 == nb) ? 0 : j + 1; }

__device__ inline void write_marker(Bucket *b, unsigned long long key) {
    Bucket x;
    x.key = key;
    x.start = 0;
    x.end = 0xFFFFFFFFu;
    x.mbr[0] = x.mbr[1] = 0xFFFFFFFFu;
    x.mbr[2] = x.mbr[3] = 0u;
    *b = x;
}

/* Stage the buckets of the lanes with `emit` set, writing the staged records
 * out first when the stage could overflow. Wave-uniform. */
__device__ inline void stage_push(Bucket *st, uint32_t &staged, bool emit, const Bucket &b, int lane, uint64_t lt,
                                  uint32_t max_meet, uint32_t end_time_min, MemtableDev mt, Counters *cnt) {
    uint64_t em = __ballot(emit);
    if (!em) return;
    if (staged + (uint32_t)__popcll(em) > ID_STAGE) {
        __builtin_amdgcn_wave_barrier();
        emit_staged(st, staged, lane, max_meet, end_time_min, mt, cnt);
        staged = 0;
        __builtin_amdgcn_wave_barrier();
    }
    if (emit) st[staged + __popcll(em & lt)] = b;
    staged += (uint32_t)__popcll(em);
}

struct IdentifyArgs {
    uint32_t now, min_meet, max_meet, end_time_min;
    MapBox m;
};

__device__ inline bool meeting_emits(const Bucket &b, const IdentifyArgs &a) {
    float lo0 = o2f(b.mbr[0]), lo1 = o2f(b.mbr[1]), hi0 = o2f(b.mbr[2]), hi1 = o2f(b.mbr[3]);
    return a.now - b.start >= a.min_meet + 1 && (double)lo0 >= a.m.lo0 && (double)hi0 <= a.m.hi0 &&
           (double)lo1 >= a.m.lo1 && (double)hi1 <= a.m.hi1;
}

struct IdentifyTally {
    unsigned long long live = 0, exp = 0, ret = 0, freed = 0, tomb = 0;
};

/* Backward-shift deletion over one probe cluster that starts after the free
 * bucket `owner`, walked bucket by bucket from global memory by the lanes with
 * `walking` set (wave-uniform call). Holes are kept in a 64-slot window behind
 * the walker; a hole that leaves the window becomes a tombstone, an occupied
 * bucket that matches no key. Used for clusters longer than the identify
 * window, which are rare. */
__device__ void walk_cluster(Bucket *t, uint64_t nb, uint64_t owner, bool walking, const IdentifyArgs &a,
                             Bucket *st, uint32_t &staged, int lane, uint64_t lt, MemtableDev mt, Counters *cnt,
                             IdentifyTally &tl) {
    const unsigned long long free_now = FREE_KEY(a.now), tomb = TOMB_KEY(a.now);
    uint64_t j = slot_next(owner, nb), wb = j, holes = 0;
    uint32_t rel = 0;
    while (__ballot(walking)) {
        bool emit = false;
        Bucket b;
        if (walking) {
            b = t[j];
            if (key_free(b.key)) {
                while (holes) {
                    int q = __builtin_ctzll(holes);
                    holes &= holes - 1;
                    uint64_t dst = wb + q;
                    if (dst >= nb) dst -= nb;
                    write_marker(&t[dst], free_now);
                    tl.freed++;
                }
                walking = false;
            } else {
                if (rel == STC_WAVE) {
                    if (holes & 1ull) {
                        write_marker(&t[wb], tomb);
                        tl.tomb++;
                    }
                    holes >>= 1;
                    wb = slot_next(wb, nb);
                    rel--;
                }
                if (b.end == a.now && !key_tomb(b.key)) {
                    tl.live++;
                    if (holes) {
                        uint64_t h = hash_slot(b.key, nb);
                        uint64_t back = (j >= h) ? j - h : j + nb - h;
                        uint64_t mask = holes;
                        if (back < rel) mask &= ~0ull << (rel - back);
                        if (mask) {
                            int q = __builtin_ctzll(mask);
                            uint64_t dst = wb + q;
                            if (dst >= nb) dst -= nb;
                            t[dst] = b;
                            holes ^= 1ull << q;
                            holes |= 1ull << rel;
                        }
                    }
                } else {
                    if (!key_tomb(b.key) && b.end == a.now - 1) {
                        tl.exp++;
                        emit = meeting_emits(b, a);
                    } else {
                        tl.ret++;
                    }
                    holes |= 1ull << rel;
                }
                j = slot_next(j, nb);
                rel++;
            }
        }
        stage_push(st, staged, emit, b, lane, lt, a.max_meet, a.end_time_min, mt, cnt);
    }
}

/* identify completed meetings (geometry.cu:595-737) and delete every bucket
 * that is not live at `now` from the table.
 *
 * Deletion is linear probing's backward shift, applied per probe cluster (a
 * maximal run of occupied buckets): each live bucket moves to the first hole
 * at or after its home slot, holes being dead buckets and the slots that moved
 * buckets vacated, so every remaining key keeps an unbroken probe chain; the
 * holes left at the end become free buckets.
 *
 * Each wavefront owns a contiguous range of 64-bucket chunks and the clusters
 * that start in them (after a bucket that was free when the pass began; slots
 * freed by the pass carry the tag `now` and start nothing). A chunk is handled
 * with the next chunk as lookahead, one bucket per lane and half: ballots give
 * the free, live and dead masks, each lane finds its cluster from them, and
 * the hole-filling runs as a scalar loop over 128-bit masks that visits only
 * live buckets with a hole before them in their cluster. The rare cluster
 * that leaves the 128-bucket window is walked from memory (walk_cluster).
 * Clusters are disjoint, so no two lanes ever write the same bucket, and a
 * cluster's buckets are read before its owner writes any of them.
 *
 * Expiring buckets that emit records are staged per wavefront and written in
 * batches, so the memtable counter sees one atomic per batch. */
__global__ __launch_bounds__(256) void k_identify(Bucket *t, uint64_t nb, IdentifyArgs a, MemtableDev mt,
                                                  Counters *cnt) {
    __shared__ Bucket stage[256 / STC_WAVE][ID_STAGE];
    const int lane = threadIdx.x & (STC_WAVE - 1), w = threadIdx.x / STC_WAVE;
    const uint64_t lt = (1ull << lane) - 1;           /* lanes below */
    const uint64_t gt = ~((2ull << lane) - 1);        /* lanes above */
    const unsigned long long free_now = FREE_KEY(a.now);
    const uint64_t nchunks = nb / STC_WAVE;
    const uint64_t nwaves = (uint64_t)gridDim.x * (256 / STC_WAVE);
    const uint64_t wid = (uint64_t)blockIdx.x * (256 / STC_WAVE) + w;
    const uint64_t c_begin = wid * nchunks / nwaves, c_end = (wid + 1) * nchunks / nwaves;
    Bucket *st = stage[w];
    IdentifyTally tl;
    uint32_t staged = 0;
    if (c_begin < c_end) {
        auto at = [&](uint64_t chunk) -> uint64_t {
            uint64_t p = chunk * STC_WAVE + lane;
            return p >= nb ? p - nb : p;
        };
        Bucket A = t[at(c_begin)], B = t[at(c_begin + 1)];
        for (uint64_t c = c_begin; c < c_end; c++) {
            Bucket N;
            if (c + 1 < c_end) N = t[at(c + 2)]; /* prefetch */
            const uint64_t pa = at(c), pb = at(c + 1);
            const bool fA = key_free(A.key), fB = key_free(B.key);
            const bool lA = !fA && !key_tomb(A.key) && A.end == a.now;
            const bool lB = !fB && !key_tomb(B.key) && B.end == a.now;
            const uint64_t FA = __ballot(fA), FB = __ballot(fB);
            const uint64_t OA = __ballot(fA && A.key != free_now);
            /* A-half lane: owned when the nearest free slot below it is an original free slot of this chunk */
            bool ownA = false, ownB = false, longA = false, longB = false, pastHoleA = false, pastHoleB = false;
            uint64_t below = FA & lt;
            int pf = 63 - __builtin_clzll(below | 1ull); /* valid when below != 0 */
            if (!fA && below && ((OA >> pf) & 1ull)) {
                ownA = true;
                longA = (FA & gt) == 0 && FB == 0;
            }
            /* B-half lane: owned when no free slot lies between the last free slot of the chunk and it */
            int pl = 63 - __builtin_clzll(FA | 1ull);
            if (!fB && (FB & lt) == 0 && FA && ((OA >> pl) & 1ull)) {
                ownB = true;
                longB = (FB & gt) == 0;
            }
            const bool partA = ownA && !longA, partB = ownB && !longB;
            const bool dA = partA && !lA, dB = partB && !lB;
            uint64_t HA = __ballot(dA), HB = __ballot(dB); /* holes: the owned dead buckets */
            if (partA && lA) pastHoleA = (HA & lt & ~((2ull << pf) - 1)) != 0;
            if (partB && lB) pastHoleB = (HB & lt) != 0 || (HA & ~((2ull << pl) - 1)) != 0;
            /* tallies and emission */
            bool emitA = false, emitB = false;
            if (partA) {
                if (lA) tl.live++;
                else if (!key_tomb(A.key) && A.end == a.now - 1) { tl.exp++; emitA = meeting_emits(A, a); }
                else tl.ret++;
            }
            if (partB) {
                if (lB) tl.live++;
                else if (!key_tomb(B.key) && B.end == a.now - 1) { tl.exp++; emitB = meeting_emits(B, a); }
                else tl.ret++;
            }
            stage_push(st, staged, emitA, A, lane, lt, a.max_meet, a.end_time_min, mt, cnt);
            stage_push(st, staged, emitB, B, lane, lt, a.max_meet, a.end_time_min, mt, cnt);
            /* home slots, relative to the window, of the live buckets that may move;
             * a bucket in its home slot never moves */
            int hA = 0, hB = 0;
            if (pastHoleA) {
                uint64_t h = hash_slot(A.key, nb);
                uint64_t back = (pa >= h) ? pa - h : pa + nb - h;
                hA = back > (uint64_t)lane ? 0 : lane - (int)back;
                pastHoleA = back != 0;
            }
            if (pastHoleB) {
                uint64_t h = hash_slot(B.key, nb);
                uint64_t back = (pb >= h) ? pb - h : pb + nb - h;
                hB = back > (uint64_t)(64 + lane) ? 0 : 64 + lane - (int)back;
                pastHoleB = back != 0;
            }
            uint64_t MA = __ballot(pastHoleA), MB = __ballot(pastHoleB);
            int dstA = -1, dstB = -1;
            while (MA) { /* scalar: first hole at or after the home slot */
                int x = __builtin_ctzll(MA);
                MA &= MA - 1;
                int hx = __builtin_amdgcn_readlane(hA, x);
                uint64_t e = HA & (~0ull << hx) & ((1ull << x) - 1);
                if (e) {
                    int q = __builtin_ctzll(e);
                    HA ^= 1ull << q;
                    HA |= 1ull << x;
                    if (lane == x) dstA = q;
                }
            }
            while (MB) {
                int x = __builtin_ctzll(MB);
                MB &= MB - 1;
                int hx = __builtin_amdgcn_readlane(hB, x);
                uint64_t eA = hx < 64 ? (HA & (~0ull << hx)) : 0ull;
                uint64_t eB = HB & (~0ull << (hx < 64 ? 0 : hx - 64)) & ((1ull << x) - 1);
                if (eA) {
                    int q = __builtin_ctzll(eA);
                    HA ^= 1ull << q;
                    HB |= 1ull << x;
                    if (lane == x) dstB = q;
                } else if (eB) {
                    int q = __builtin_ctzll(eB);
                    HB ^= 1ull << q;
                    HB |= 1ull << x;
                    if (lane == x) dstB = 64 + q;
                }
            }
            const uint64_t c0 = c * STC_WAVE;
            auto slot = [&](int r) -> uint64_t {
                uint64_t p = c0 + (uint64_t)r;
                return p >= nb ? p - nb : p;
            };
            if (dstA >= 0) t[slot(dstA)] = A;
            if (dstB >= 0) t[slot(dstB)] = B;
            if ((HA >> lane) & 1ull) {
                write_marker(&t[pa], free_now);
                tl.freed++;
            }
            if ((HB >> lane) & 1ull) {
                write_marker(&t[pb], free_now);
                tl.freed++;
            }
            /* the cluster after the chunk's last free slot, when it leaves the window */
            const uint64_t LG = __ballot(longA || longB);
            const bool long_owner = LG != 0 && lane == pl;
            walk_cluster(t, nb, pa, long_owner, a, st, staged, lane, lt, mt, cnt, tl);
            A = B;
            B = N;
        }
    }
    if (staged) {
        __builtin_amdgcn_wave_barrier();
        emit_staged(st, staged, lane, a.max_meet, a.end_time_min, mt, cnt);
    }
    unsigned long long v0 = wave_sum_u64(tl.live), v1 = wave_sum_u64(tl.exp), v2 = wave_sum_u64(tl.ret),
                       v3 = wave_sum_u64(tl.freed), v4 = wave_sum_u64(tl.tomb);
    if (lane == 0) {
        if (v0) atomicAdd(&cnt->live, v0);
        if (v1) atomicAdd(&cnt->expired, v1);
        if (v2) atomicAdd(&cnt->retired, v2);
        if (v3) atomicAdd(&cnt->freed, v3);
        if (v4) atomicAdd(&cnt->tombs, v4);
    }
}

/* Rehash in three passes over the old table, chunked by workgroup:
 * count the kept buckets per chunk, prefix-sum the counts, and compact the
 * kept buckets in order into a staging array; the old table is then freed
 * and the staged buckets are inserted into the new one. */
#define RH_CHUNK 8192u

__device__ inline bool keep_bucket(const Bucket &b, uint32_t keep_from) {
    return !key_free(b.key) && !key_tomb(b.key) && b.end != 0xFFFFFFFFu && b.end >= keep_from;
}

__global__ __launch_bounds__(256) void k_stage_count(const Bucket *src, uint64_t nsrc, uint32_t keep_from,
                                                     uint32_t *chunk_count) {
    __shared__ uint32_t wsum[4];
    uint64_t lo = (uint64_t)blockIdx.x * RH_CHUNK, hi = min(lo + RH_CHUNK, nsrc);
    uint32_t c = 0;
    for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) c += keep_bucket(src[i], keep_from);
    for (int o = STC_WAVE / 2; o > 0; o >>= 1) c += __shfl_down(c, o, STC_WAVE);
    if ((threadIdx.x & (STC_WAVE - 1)) == 0) wsum[threadIdx.x / STC_WAVE] = c;
    __syncthreads();
    if (threadIdx.x == 0) chunk_count[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(256) void k_stage_write(const Bucket *src, uint64_t nsrc, uint32_t keep_from,
                                                     const uint32_t *chunk_offset, Bucket *stage) {
    __shared__ uint32_t wbase[5];
    const int lane = threadIdx.x & (STC_WAVE - 1), w = threadIdx.x / STC_WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (STC_WAVE - lane));
    uint64_t lo = (uint64_t)blockIdx.x * RH_CHUNK, hi = min(lo + RH_CHUNK, nsrc);
    uint64_t out = chunk_offset[blockIdx.x];
    for (uint64_t base = lo; base < hi; base += blockDim.x) {
        uint64_t i = base + threadIdx.x;
        Bucket b;
        bool k = false;
        if (i < hi) {
            b = src[i];
            k = keep_bucket(b, keep_from);
        }
        uint64_t m = __ballot(k);
        if (lane == 0) wbase[w] = (uint32_t)__popcll(m);
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t acc = 0;
            for (int q = 0; q < 4; q++) { uint32_t v = wbase[q]; wbase[q] = acc; acc += v; }
            wbase[4] = acc;
        }
        __syncthreads();
        if (k) stage[out + wbase[w] + __popcll(m & lt)] = b;
        out += wbase[4];
        __syncthreads();
    }
}

__global__ void k_table_unstage(const Bucket *stage, uint64_t n, Bucket *dst, uint64_t ndst, Counters *cnt) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Bucket b = stage[i];
    bool claimed;
    Bucket *d = table_find_or_claim(dst, ndst, b.key, &claimed);
    if (!d) { atomicAdd(&cnt->table_fail, 1u); return; }
    d->start = b.start;
    d->end = b.end;
    d->mbr[0] = b.mbr[0];
    d->mbr[1] = b.mbr[1];
    d->mbr[2] = b.mbr[2];
    d->mbr[3] = b.mbr[3];
}

__global__ void k_fill_u32(uint32_t *a, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_obox(uint4 *a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_uint4(100000000u, 100000000u, 0u, 0u); /* cuda_init_o_boxs */
}

/* ------------------------------------------------------------------ */
/* host context                                                        */
/* ------------------------------------------------------------------ */

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    int ensure(size_t n) {
        if (n <= cap && p) return 0;
        size_t nc = std::max<size_t>(n + n / 2, 1024);
        if (p) { hipFree(p); p = nullptr; cap = 0; }
        if (hipMalloc((void **)&p, nc * sizeof(T)) != hipSuccess) { p = nullptr; return -1; }
        cap = nc;
        return 0;
    }
    void release() { if (p) hipFree(p); p = nullptr; cap = 0; }
};

/* A flushed run: metadata on the host, payload in a pinned block that the
 * copy stream fills while later steps run. */
struct RunHost {
    stc_run_info info;
    std::vector<stc_ctf_meta> ctfs;
    RunLayout lay;
    bool has_sorted = false;
    bool finalized = false;
    PinnedBlock blk;
    uint8_t *dev = nullptr;           /* HBM copy for queries (device_runs) */
    QueryRunDev qdev;
    std::vector<CtfStats> stats;      /* per CTF, for planning range queries */
};

/* Pinned host blocks for run payloads, recycled once their copies are done. */
struct PinnedPool {
    std::vector<PinnedBlock> idle, retiring;
    void reclaim() {
        for (size_t i = 0; i < retiring.size();) {
            if (hipEventQuery(retiring[i].ev) == hipSuccess) {
                idle.push_back(retiring[i]);
                retiring[i] = retiring.back();
                retiring.pop_back();
            } else {
                i++;
            }
        }
    }
    PinnedBlock acquire(size_t bytes) {
        reclaim();
        int best = -1;
        for (size_t i = 0; i < idle.size(); i++)
            if (idle[i].cap >= bytes && (best < 0 || idle[i].cap < idle[best].cap)) best = (int)i;
        if (best >= 0) {
            PinnedBlock b = idle[best];
            idle.erase(idle.begin() + best);
            return b;
        }
        if (!idle.empty()) { /* too small: replace the largest idle block */
            PinnedBlock b = idle.back();
            idle.pop_back();
            (void)hipHostFree(b.p);
            (void)hipEventDestroy(b.ev);
        }
        PinnedBlock b;
        size_t cap = bytes + bytes / 4 + (1u << 20);
        if (hipHostMalloc((void **)&b.p, cap, hipHostMallocDefault) != hipSuccess) return PinnedBlock();
        if (hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) {
            (void)hipHostFree(b.p);
            return PinnedBlock();
        }
        b.cap = cap;
        return b;
    }
    void release(PinnedBlock b) {
        if (b.p) retiring.push_back(b);
    }
    void clear() {
        for (auto *v : {&idle, &retiring})
            for (auto &b : *v) {
                (void)hipEventSynchronize(b.ev);
                (void)hipHostFree(b.p);
                (void)hipEventDestroy(b.ev);
            }
        idle.clear();
        retiring.clear();
    }
};

static PinnedBlock pool_acquire(void *pool, size_t bytes) { return ((PinnedPool *)pool)->acquire(bytes); }

struct StcCtx {
    stc_config cfg;
    MapBox map;
    std::string err;
    hipStream_t s = nullptr;
    uint32_t n = 0, cap = 0, buf = 0;
    double x_buffer = 0, y_buffer = 0, band = 0;
    int device = 0;

    DBuf<double2> xy_in;
    DBuf<uint32_t> codes, idx, codes_alt, idx_alt, temp;
    DBuf<double> dlon, idlon, lx, ly, ld;
    DBuf<Node> nodes;
    DBuf<FrontierItem> frontier_a, frontier_b;
    DBuf<uint32_t> leaf_count, leaf_start, leaf_node;
    DBuf<uint32_t> pt_leaf, pt_lv, pt_rank, lp, lkeys;
    DBuf<uint32_t> ucount, uoffset;
    DBuf<uint2> units;
    DBuf<PairHit> pairs;
    DBuf<uint2> amb;
    Bucket *table = nullptr;
    uint64_t table_buckets = 0;
    int log2b = 0;
    /* memtable */
    uint64_t *rec_keys = nullptr;
    float4 *rec_boxes = nullptr;
    uint64_t rec_cap = 0;
    uint32_t *same_pid_count = nullptr;
    uint4 *o_boxs = nullptr;
    uint16_t *sids = nullptr;
    Counters *dcnt = nullptr;
    Counters *hcnt = nullptr;  /* pinned */
    Counters *hinit = nullptr; /* pinned: initial counter values of a step */
    uint32_t sid_count = 0, oversize_oid_count = 0;
    uint32_t end_time_min = 0;
    uint64_t kv_count = 0;
    uint64_t last_live = 0;
    uint64_t occupied = 0;
    uint64_t last_new = 0;
    uint64_t last_listed = 0;
    uint32_t last_t = 0;
    bool stepped = false;
    uint32_t last_nodes = 0, last_leaves = 0;
    uint64_t last_pairs = 0;
    std::vector<PairHit> host_pairs_extra;
    DBuf<PairHit> amb_pairs;
    DBuf<stc_meeting_unit> units_in;  /* replayed meeting units and their pair keys */
    DBuf<uint64_t> ukeys, ukeys_alt;
    DBuf<uint32_t> uidx, uidx_alt, utemp;
    DBuf<uint32_t> chunk_count, chunk_offset;
    DBuf<Bucket> stage;
    std::vector<RunHost> runs;
    PinnedPool pinned;
    QueryBufs *qb = nullptr;
    FlushWork flushw;
    hipEvent_t ev[10];
};

static int fail(StcCtx *c, int code, const std::string &msg) {
    c->err = msg;
    return code;
}
#define CK(expr)                                                                   \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            stc_hip_fail(e_, #expr, __FILE__, __LINE__);                           \
            return fail(c, STC_E_HIP, g_last_fail);                                \
        }                                                                          \
    } while (0)
#define CKRC(expr)                                                                 \
    do {                                                                           \
        int rc_ = (expr);                                                          \
        if (rc_) return fail(c, rc_ < 0 ? rc_ : STC_E_HIP, g_last_fail.empty() ? #expr : g_last_fail); \
    } while (0)

static inline uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

static int sync_counters(StcCtx *c) {
    CK(hipMemcpyAsync(c->hcnt, c->dcnt, sizeof(Counters), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return 0;
}

/* the identify pass works on 64-bucket chunks: table sizes are multiples of 64 */
static uint64_t round_buckets(uint64_t nb) { return (nb + 63) & ~63ull; }

static int alloc_table(StcCtx *c, uint64_t nb, Bucket **out) {
    Bucket *t = nullptr;
    CK(hipMalloc((void **)&t, nb * sizeof(Bucket)));
    hipLaunchKernelGGL(k_table_clear, dim3(blocks_for(nb, 256)), dim3(256), 0, c->s, t, nb);
    CK(hipGetLastError());
    *out = t;
    return 0;
}

static int ensure_records(StcCtx *c, uint64_t need) {
    if (need <= c->rec_cap) return 0;
    uint64_t nc = std::max<uint64_t>(need + need / 2, 1 << 16);
    uint64_t *k = nullptr;
    float4 *b = nullptr;
    CK(hipMalloc((void **)&k, nc * 16));
    CK(hipMalloc((void **)&b, nc * 16));
    if (c->kv_count) {
        CK(hipMemcpyAsync(k, c->rec_keys, c->kv_count * 16, hipMemcpyDeviceToDevice, c->s));
        CK(hipMemcpyAsync(b, c->rec_boxes, c->kv_count * 16, hipMemcpyDeviceToDevice, c->s));
    }
    CK(hipStreamSynchronize(c->s));
    if (c->rec_keys) hipFree(c->rec_keys);
    if (c->rec_boxes) hipFree(c->rec_boxes);
    c->rec_keys = k;
    c->rec_boxes = b;
    c->rec_cap = nc;
    return 0;
}

static int ilog2_ceil(uint64_t v) {
    int l = 0;
    while ((1ull << l) < v) l++;
    return l;
}

/* Move the entries last seen at or after keep_from into a table sized for
 * expected_keys at a load of at most 1/3 (see the kernels above): fewer
 * probes per insert make the refine pass faster than the larger table makes
 * the identify scan slower. The kept entries go to a staging buffer that
 * stays allocated between rehashes; the table is refilled in place unless it
 * must grow. */
/* Copy a just-flushed run out of the flush buffers into its own HBM block
 * (packed keys, CTF metadata and bases, sids, oversize buffer) for queries. */
static int keep_run_on_device(StcCtx *c, RunHost &run) {
    const FlushWork &w = c->flushw;
    const uint32_t C = run.info.ctf_count, n = run.info.num_objects, nt = run.info.oversize_count;
    const uint64_t nf = run.info.kv_count;
    size_t off = 0;
    auto place = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    const size_t o_keys = place(run.info.key_bytes), o_ctfs = place(C * sizeof(stc_ctf_meta)),
                 o_bb = place((C + 1) * 8ull), o_rb = place((C + 1) * 4ull), o_sids = place(n * 2ull),
                 o_ok = place(nt * 16ull), o_ob = place(nt * 16ull);
    CK(hipMalloc((void **)&run.dev, std::max<size_t>(off, 256)));
    hipStream_t s = c->s;
    uint8_t *d = run.dev;
    if (run.info.key_bytes) CK(hipMemcpyAsync(d + o_keys, w.key_bytes, run.info.key_bytes, hipMemcpyDeviceToDevice, s));
    /* the host metadata is final but for flag 8, which queries do not read */
    CK(hipMemcpyAsync(d + o_ctfs, run.ctfs.data(), C * sizeof(stc_ctf_meta), hipMemcpyHostToDevice, s));
    CK(hipMemcpyAsync(d + o_bb, w.ctf_byte_base, (C + 1) * 8ull, hipMemcpyDeviceToDevice, s));
    CK(hipMemcpyAsync(d + o_rb, w.ctf_base, (C + 1) * 4ull, hipMemcpyDeviceToDevice, s));
    CK(hipMemcpyAsync(d + o_sids, w.sids_out, n * 2ull, hipMemcpyDeviceToDevice, s));
    if (nt) {
        CK(hipMemcpyAsync(d + o_ok, w.keys_out + 2 * nf, nt * 16ull, hipMemcpyDeviceToDevice, s));
        CK(hipMemcpyAsync(d + o_ob, w.boxes_out + nf, nt * 16ull, hipMemcpyDeviceToDevice, s));
    }
    QueryRunDev &q = run.qdev;
    q.key_bytes = d + o_keys;
    q.ctfs = (const stc_ctf_meta *)(d + o_ctfs);
    q.byte_base = (const uint64_t *)(d + o_bb);
    q.rec_base = (const uint32_t *)(d + o_rb);
    q.sids = (const uint16_t *)(d + o_sids);
    q.o_keys = (const uint64_t *)(d + o_ok);
    q.o_boxes = (const float4 *)(d + o_ob);
    q.ctf_count = C;
    q.n_tail = nt;
    q.end_time_min = run.info.end_time_min;
    q.source = 0; /* set per query from the run's position */
    CK(hipStreamSynchronize(s)); /* the host metadata vector is read by the copy above */
    return stc_query_ctf_stats(q, run.stats, s) ? fail(c, STC_E_HIP, std::string("run statistics: ") + g_last_fail) : 0;
}

static int table_rehash(StcCtx *c, uint32_t keep_from, uint64_t expected_keys) {
    hipStream_t s = c->s;
    uint64_t nold = c->table_buckets;
    uint32_t chunks = (uint32_t)((nold + RH_CHUNK - 1) / RH_CHUNK);
    if (c->chunk_count.ensure(chunks + 1) || c->chunk_offset.ensure(chunks + 1))
        return fail(c, STC_E_CAPACITY, "rehash chunk counts");
    hipLaunchKernelGGL(k_stage_count, dim3(chunks), dim3(256), 0, s, c->table, nold, keep_from, c->chunk_count.p);
    CK(hipGetLastError());
    CKRC(stc_exclusive_scan_u32(c->chunk_count.p, c->chunk_offset.p, chunks, c->chunk_offset.p + chunks, c->temp.p,
                                s));
    uint32_t kept = 0;
    CK(hipMemcpyAsync(&kept, c->chunk_offset.p + chunks, 4, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    if (kept) {
        if (c->stage.ensure(kept)) return fail(c, STC_E_CAPACITY, "rehash staging buffer");
        hipLaunchKernelGGL(k_stage_write, dim3(chunks), dim3(256), 0, s, c->table, nold, keep_from,
                           c->chunk_offset.p, c->stage.p);
        CK(hipGetLastError());
    }
    uint64_t want = round_buckets(std::max<uint64_t>(expected_keys, kept) * 3 + 1024);
    uint64_t nb = nold;
    if (want > nold) {
        CK(hipStreamSynchronize(s));
        CK(hipFree(c->table));
        c->table = nullptr;
        Bucket *nt = nullptr;
        CKRC(alloc_table(c, want, &nt));
        c->table = nt;
        nb = want;
    } else {
        hipLaunchKernelGGL(k_table_clear, dim3(blocks_for(nb, 256)), dim3(256), 0, s, c->table, nb);
        CK(hipGetLastError());
    }
    c->table_buckets = nb;
    if (kept) {
        hipLaunchKernelGGL(k_table_unstage, dim3(blocks_for(kept, 256)), dim3(256), 0, s, c->stage.p,
                           (uint64_t)kept, c->table, nb, c->dcnt);
        CK(hipGetLastError());
    }
    CK(hipStreamSynchronize(s));
    c->occupied = kept;
    return 0;
}

extern "C" {

int stc_device_count(int32_t *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) { *count = 0; return STC_E_HIP; }
    *count = n;
    return STC_OK;
}

int stc_device_info_get(int32_t device, stc_device_info *info) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return STC_E_HIP;
    memset(info, 0, sizeof(*info));
    info->device_id = device;
    snprintf(info->name, sizeof(info->name), "%s", prop.gcnArchName);
    info->mem_size = prop.totalGlobalMem;
    info->num_cus = prop.multiProcessorCount;
    info->clock_khz = prop.clockRate;
    return STC_OK;
}

int stc_create(const stc_config *cfg, const stc_box *mbr, void **out) {
    if (!cfg || !mbr || !out) return STC_E_ARG;
    *out = nullptr;
    StcCtx *c = new StcCtx();
    c->cfg = *cfg;
    c->map = MapBox{mbr->low[0], mbr->low[1], mbr->high[0], mbr->high[1]};
    c->n = cfg->num_objects;
    c->cap = cfg->grid_capacity;
    c->buf = (uint32_t)(cfg->grid_amplify * cfg->grid_capacity);
    uint32_t C = cfg->ctf_count, split = (uint32_t)std::sqrt((double)C);
    if (c->n == 0 || c->cap == 0 || c->buf == 0 || !(cfg->reach_distance > 0) || C == 0 || split * split != C ||
        !(mbr->high[0] > mbr->low[0]) || !(mbr->high[1] > mbr->low[1]) || c->n >= (1u << 26)) {
        delete c;
        return STC_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        delete c;
        return STC_E_HIP;
    }
    hipGetDevice(&c->device);
    /* border buffers from the map's lowest latitude, host formula (filter.cpp:30-31) */
    double absla = std::fabs(mbr->low[1]);
    if (absla == 90) absla = 89;
    c->x_buffer = cfg->reach_distance * (360.0 / (std::sin((90 - absla) * STC_PI_REF / 180) * 40076.0 * 1000.0));
    c->y_buffer = cfg->reach_distance * STC_DPM_LAT;
    c->band = 16.0 * (std::nextafter(cfg->reach_distance, INFINITY) - cfg->reach_distance);
    c->end_time_min = cfg->start_time + cfg->min_meet_time; /* process.cpp:404 */

    auto bad = [&](const char *what) {
        c->err = what;
        stc_destroy(c);
        return STC_E_HIP;
    };
    if (hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking) != hipSuccess) return bad("stream");
    for (int i = 0; i < 10; i++) hipEventCreate(&c->ev[i]);
    uint32_t n = c->n;
    size_t node_cap = 1 + 4ull * NLEVELS * (n / (c->cap + 1) + 1);
    size_t front_cap = n / (c->cap + 1) + 2;
    if (c->xy_in.ensure(n) || c->codes.ensure(n) || c->idx.ensure(n) || c->codes_alt.ensure(n) ||
        c->idx_alt.ensure(n) || c->dlon.ensure(n) || c->idlon.ensure(n) || c->lx.ensure(n) || c->ly.ensure(n) || c->ld.ensure(n) ||
        c->nodes.ensure(node_cap) || c->frontier_a.ensure(front_cap) || c->frontier_b.ensure(front_cap) ||
        c->pt_leaf.ensure(n) || c->pt_lv.ensure(n) || c->pt_rank.ensure(n) || c->lp.ensure(n) ||
        c->lkeys.ensure(n) || c->ucount.ensure(n) || c->uoffset.ensure(n) || c->units.ensure(4ull * n) ||
        c->pairs.ensure(std::max<size_t>(4ull * n, 1 << 16)) || c->amb.ensure(1 << 16) ||
        c->leaf_count.ensure(node_cap) || c->leaf_start.ensure(node_cap) || c->leaf_node.ensure(node_cap) ||
        c->temp.ensure(stc_radix_temp_words((uint32_t)std::max<size_t>(n, node_cap)) + 64))
        return bad("device allocation failed");
    if (hipMalloc((void **)&c->dcnt, sizeof(Counters)) != hipSuccess) return bad("counters");
    if (hipHostMalloc((void **)&c->hcnt, sizeof(Counters), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->hinit, sizeof(Counters), hipHostMallocDefault) != hipSuccess)
        return bad("pinned");
    uint64_t nb = cfg->num_meeting_buckets ? cfg->num_meeting_buckets : 4ull * n;
    c->table_buckets = round_buckets(std::max<uint64_t>(nb, 1024));
    if (alloc_table(c, c->table_buckets, &c->table)) return bad("table");
    if (hipMalloc((void **)&c->same_pid_count, 4ull * n) != hipSuccess ||
        hipMalloc((void **)&c->o_boxs, 16ull * n) != hipSuccess || hipMalloc((void **)&c->sids, 2ull * n) != hipSuccess)
        return bad("per-object state");
    hipMemsetAsync(c->same_pid_count, 0, 4ull * n, c->s);
    hipMemsetAsync(c->sids, 0, 2ull * n, c->s);
    hipLaunchKernelGGL(k_fill_obox, dim3(blocks_for(n, 256)), dim3(256), 0, c->s, c->o_boxs, n);
    if (ensure_records(c, std::max<uint64_t>(1 << 16, (uint64_t)cfg->kv_restriction + 2))) return bad("records");
    if (stc_flush_init(&c->flushw, n, C)) return bad("flush buffers");
    if (hipStreamSynchronize(c->s) != hipSuccess) return bad("init sync");
    *out = c;
    return STC_OK;
}

int stc_destroy(void *vc) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return STC_OK;
    if (c->s) hipStreamSynchronize(c->s);
    c->xy_in.release(); c->codes.release(); c->idx.release(); c->codes_alt.release(); c->idx_alt.release();
    c->temp.release(); c->dlon.release(); c->idlon.release(); c->lx.release(); c->ly.release(); c->ld.release(); c->nodes.release();
    c->frontier_a.release(); c->frontier_b.release(); c->leaf_count.release(); c->leaf_start.release();
    c->leaf_node.release(); c->pt_leaf.release(); c->pt_lv.release(); c->pt_rank.release(); c->lp.release();
    c->lkeys.release(); c->ucount.release(); c->uoffset.release(); c->units.release(); c->pairs.release();
    c->amb.release();
    c->amb_pairs.release();
    c->units_in.release(); c->ukeys.release(); c->ukeys_alt.release();
    c->uidx.release(); c->uidx_alt.release(); c->utemp.release();
    c->chunk_count.release();
    c->chunk_offset.release();
    c->stage.release();
    if (c->table) hipFree(c->table);
    if (c->rec_keys) hipFree(c->rec_keys);
    if (c->rec_boxes) hipFree(c->rec_boxes);
    if (c->same_pid_count) hipFree(c->same_pid_count);
    if (c->o_boxs) hipFree(c->o_boxs);
    if (c->sids) hipFree(c->sids);
    if (c->dcnt) hipFree(c->dcnt);
    if (c->hcnt) hipHostFree(c->hcnt);
    if (c->hinit) hipHostFree(c->hinit);
    stc_flush_free(&c->flushw);
    for (auto &r : c->runs) {
        c->pinned.release(r.blk);
        if (r.dev) (void)hipFree(r.dev);
    }
    c->runs.clear();
    c->pinned.clear();
    stc_query_bufs_free(c->qb);
    c->qb = nullptr;
    for (int i = 0; i < 10; i++) if (c->ev[i]) hipEventDestroy(c->ev[i]);
    if (c->s) hipStreamDestroy(c->s);
    delete c;
    return STC_OK;
}

const char *stc_last_error(void *vc) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return g_last_fail.c_str();
    return c->err.c_str();
}

/* host re-evaluation of ambiguous pair decisions with the C library's sin,
 * the reference CPU formula (geometry.h:40-48, util.h:59-66) */
static double host_dpm_lon(double lat) {
    return 360.0 / (std::sin((90 - std::fabs(lat)) * STC_PI_REF / 180) * 40076.0 * 1000.0);
}

struct StepGeom {
    uint32_t nodes, leaves;
    uint64_t units, pairs, ambiguous;
    uint32_t rebuilt;
};
static int step_tail(StcCtx *c, uint32_t t, StepGeom &g, stc_step_out *out);

static int step_impl(StcCtx *c, const double2 *xy, uint32_t t, stc_step_out *out) {
    const uint32_t n = c->n;
    hipStream_t s = c->s;
    Counters *cnt = c->dcnt;
    memset(c->hinit, 0, sizeof(Counters));
    c->hinit->kv_count = c->kv_count;
    c->hinit->sid_count = c->sid_count;
    c->hinit->oversize_oid_count = c->oversize_oid_count;
    CK(hipMemcpyAsync(cnt, c->hinit, sizeof(Counters), hipMemcpyHostToDevice, s));
    CK(hipEventRecord(c->ev[0], s));

    /* 1-2: cells, Morton codes, dpm_lon; sort codes */
    hipLaunchKernelGGL(k_prep, dim3(blocks_for(n, 256)), dim3(256), 0, s, xy, n, c->map, c->codes.p, c->idx.p,
                       c->dlon.p, c->idlon.p, cnt);
    CK(hipGetLastError());
    CKRC(stc_radix_sort_u32(c->codes.p, c->idx.p, c->codes_alt.p, c->idx_alt.p, n, 28, c->temp.p, s));

    /* 3: tree */
    uint32_t node_cap = (uint32_t)c->nodes.cap;
    hipLaunchKernelGGL(k_tree_root, dim3(1), dim3(64), 0, s, c->nodes.p, c->map, c->frontier_a.p, cnt, n, c->cap);
    FrontierItem *fin = c->frontier_a.p, *fout = c->frontier_b.p;
    for (int L = NLEVELS - 1; L >= 1; L--) {
        uint64_t bound = std::min<uint64_t>(1ull << (2 * (NLEVELS - 1 - L)), c->frontier_a.cap);
        hipLaunchKernelGGL(k_tree_level, dim3(blocks_for(bound, 128)), dim3(128), 0, s, c->nodes.p, c->map,
                           c->codes.p, n, c->cap, L, fin, fout, cnt, node_cap, (uint32_t)c->frontier_a.cap);
        std::swap(fin, fout);
    }
    {
        uint64_t bound = c->frontier_a.cap;
        hipLaunchKernelGGL(k_tree_finest, dim3(blocks_for(bound, 128)), dim3(128), 0, s, c->nodes.p, fin, cnt,
                           node_cap);
    }
    CK(hipGetLastError());
    CKRC(sync_counters(c));
    Counters h = *c->hcnt;
    if (h.bad_points) return fail(c, STC_E_INPUT, "point outside the map box");
    if (h.table_fail) return fail(c, STC_E_CAPACITY, "QTree node capacity exceeded");
    if (n <= c->cap) return fail(c, STC_E_ARG, "root does not split: too few objects for grid_capacity");
    uint32_t n_nodes = h.nodes, n_leaves = h.leaves;
    c->last_nodes = n_nodes;
    c->last_leaves = n_leaves;

    /* 4: partition */
    CK(hipMemsetAsync(c->leaf_count.p, 0, 4ull * n_leaves, s));
    hipLaunchKernelGGL(k_leaf_nodes, dim3(blocks_for(n_nodes, 256)), dim3(256), 0, s, c->nodes.p, n_nodes,
                       c->leaf_node.p);
    hipLaunchKernelGGL(k_partition, dim3(blocks_for(n, 256)), dim3(256), 0, s, xy, n, c->nodes.p, c->x_buffer,
                       c->y_buffer, c->pt_leaf.p, c->pt_lv.p, c->leaf_count.p, cnt);
    CK(hipGetLastError());
    /* 5: leaves in id order */
    CKRC(stc_exclusive_scan_u32(c->leaf_count.p, c->leaf_start.p, n_leaves, nullptr, c->temp.p, s));
    CK(hipMemcpyAsync(c->lkeys.p, c->pt_leaf.p, 4ull * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, s, c->lp.p, n);
    CKRC(stc_radix_sort_u32(c->lkeys.p, c->lp.p, c->codes_alt.p, c->idx_alt.p, n,
                            std::max(1, ilog2_ceil(n_leaves)), c->temp.p, s));
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, xy, c->dlon.p, c->lp.p, c->pt_leaf.p,
                       c->leaf_start.p, n, c->lx.p, c->ly.p, c->ld.p, c->pt_rank.p);
    hipLaunchKernelGGL(k_overflow, dim3(blocks_for(n_leaves, 256)), dim3(256), 0, s, c->leaf_count.p, n_leaves,
                       c->buf, cnt);
    CK(hipGetLastError());
    CK(hipEventRecord(c->ev[1], s));

    /* 6: border lookup (count, scan, emit) */
    double r = c->cfg.reach_distance;
    hipLaunchKernelGGL(k_lookup<false>, dim3(blocks_for(n, 128)), dim3(128), 0, s, xy, c->dlon.p, n, c->nodes.p,
                       c->pt_lv.p, r, c->band, c->ucount.p, (const uint32_t *)nullptr, (uint2 *)nullptr, cnt);
    CK(hipGetLastError());
    uint32_t *units_total = (uint32_t *)((char *)cnt + offsetof(Counters, units));
    CKRC(stc_exclusive_scan_u32(c->ucount.p, c->uoffset.p, n, units_total, c->temp.p, s));
    CKRC(sync_counters(c));
    h = *c->hcnt;
    uint64_t n_units = (uint32_t)h.units;
    if (c->units.ensure(n_units + 1)) return fail(c, STC_E_CAPACITY, "unit buffer allocation failed");
    hipLaunchKernelGGL(k_lookup<true>, dim3(blocks_for(n, 128)), dim3(128), 0, s, xy, c->dlon.p, n, c->nodes.p,
                       c->pt_lv.p, r, c->band, c->ucount.p, (const uint32_t *)c->uoffset.p, c->units.p, cnt);
    CK(hipGetLastError());
    CK(hipEventRecord(c->ev[2], s));

    /* 7-8: refinement fused with the meeting-table insert. First make room:
     * after this step's new keys at most half of the buckets hold keys. */
    uint32_t rebuilt = 0;
    uint64_t est_new = c->stepped ? c->last_new / 2 + 1024 : (uint64_t)n;
    if ((c->occupied + est_new) * 10 > c->table_buckets * 6) {
        CKRC(table_rehash(c, t - 1, c->last_live + est_new));
        rebuilt = 1;
    }
    const bool keep = (c->cfg.keep_pairs & 1) != 0;
    uint64_t total_units = (uint64_t)n + n_units;
    uint32_t grid = (uint32_t)std::min<uint64_t>(blocks_for(total_units * STC_WAVE, 256), 256u * 64u);
    for (int attempt = 0;; attempt++) {
        CK(hipMemsetAsync(&cnt->candidates, 0, 6 * sizeof(unsigned long long), s)); /* candidates .. continuing */
        CK(hipMemsetAsync(&cnt->table_fail, 0, sizeof(uint32_t), s));
        CK(hipEventRecord(c->ev[6], s));
        hipLaunchKernelGGL(k_refine, dim3(grid), dim3(256), 0, s, xy, c->dlon.p, c->idlon.p, n, c->pt_leaf.p,
                           c->pt_rank.p, c->leaf_start.p, c->leaf_count.p, c->lp.p, c->lx.p, c->ly.p, c->units.p,
                           total_units, c->buf, r, c->band, r * r * (1.0 - 1e-12), r * r * (1.0 + 1e-12), c->table,
                           c->table_buckets, t, keep ? c->pairs.p : (PairHit *)nullptr,
                           (uint64_t)c->pairs.cap, c->amb.p, (uint32_t)c->amb.cap, cnt);
        CK(hipGetLastError());
        CK(hipEventRecord(c->ev[7], s));
        CKRC(sync_counters(c));
        h = *c->hcnt;
        bool redo = false;
        if (h.table_fail) {
            /* grow, keeping the keys seen last step and this step; reinserting is idempotent */
            CKRC(table_rehash(c, t - 1, c->last_live + h.pairs + 1024));
            rebuilt = 1;
            redo = true;
            CK(hipMemsetAsync(&cnt->table_fail, 0, sizeof(uint32_t), s));
        }
        if (keep && h.pairs_listed > c->pairs.cap) {
            if (c->pairs.ensure(h.pairs_listed + 1024)) return fail(c, STC_E_CAPACITY, "pair list allocation failed");
            redo = true;
        }
        if (h.ambiguous > c->amb.cap) {
            if (c->amb.ensure(h.ambiguous + 1024)) return fail(c, STC_E_CAPACITY, "ambiguous list allocation failed");
            redo = true;
        }
        if (!redo) break;
        if (attempt >= 8) return fail(c, STC_E_CAPACITY, "meeting table cannot hold the step's pairs");
    }
    uint64_t n_pairs = h.pairs;
    uint64_t n_amb = h.ambiguous;
    c->last_new = h.new_keys;
    c->occupied += h.new_keys;
    c->last_listed = keep ? h.pairs_listed : 0;
    c->host_pairs_extra.clear();
    if (n_amb) {
        /* decide the band with the C library's sin and insert the accepted pairs */
        std::vector<uint2> ab(n_amb);
        CK(hipMemcpy(ab.data(), c->amb.p, n_amb * sizeof(uint2), hipMemcpyDeviceToHost));
        std::vector<PairHit> extra;
        for (const uint2 &e : ab) {
            double2 pp, qq;
            CK(hipMemcpy(&pp, xy + e.x, sizeof(double2), hipMemcpyDeviceToHost));
            CK(hipMemcpy(&qq, xy + e.y, sizeof(double2), hipMemcpyDeviceToHost));
            double dx = (pp.x - qq.x) / host_dpm_lon(pp.y);
            double dy = (pp.y - qq.y) / STC_DPM_LAT;
            if (std::sqrt(dx * dx + dy * dy) <= r) {
                PairHit ph;
                uint32_t a = std::min(e.x, e.y), b = std::max(e.x, e.y);
                ph.key = ((unsigned long long)a << 32) | b;
                float pfx = (float)pp.x, pfy = (float)pp.y, qfx = (float)qq.x, qfy = (float)qq.y;
                ph.mbr[0] = std::min(pfx, qfx); ph.mbr[1] = std::min(pfy, qfy);
                ph.mbr[2] = std::max(pfx, qfx); ph.mbr[3] = std::max(pfy, qfy);
                extra.push_back(ph);
            }
        }
        if (!extra.empty()) {
            if (c->amb_pairs.ensure(extra.size())) return fail(c, STC_E_CAPACITY, "ambiguous pair buffer");
            CK(hipMemcpy(c->amb_pairs.p, extra.data(), extra.size() * sizeof(PairHit), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_table_insert, dim3(blocks_for(extra.size(), 256)), dim3(256), 0, s, c->table,
                               c->table_buckets, c->amb_pairs.p, (uint64_t)extra.size(), t, cnt);
            CK(hipGetLastError());
            n_pairs += extra.size();
            c->occupied += extra.size();
            c->host_pairs_extra = extra;
        }
    }
    c->last_pairs = n_pairs;
    CK(hipEventRecord(c->ev[3], s));
    StepGeom g{n_nodes, n_leaves, n_units, n_pairs, n_amb, rebuilt};
    return step_tail(c, t, g, out);
}

/* Steps 9-10, shared by the position step and the replay of persisted
 * meeting units: identify, rebuild when needed, flush, report. */
static int step_tail(StcCtx *c, uint32_t t, StepGeom &g, stc_step_out *out) {
    const uint32_t n = c->n;
    hipStream_t s = c->s;
    Counters *cnt = c->dcnt;
    uint32_t rebuilt = g.rebuilt;
    /* 9: identify: meetings live at t - 1 and not continued now expire */
    CKRC(sync_counters(c));
    Counters h = *c->hcnt;
    uint64_t h_cont = c->hcnt->continuing;
    uint64_t expiring = c->last_live > h_cont ? c->last_live - h_cont : 0;
    uint64_t need_rec = c->kv_count + 2 * expiring + 2;
    if (ensure_records(c, need_rec)) return fail(c, STC_E_CAPACITY, "record buffer allocation failed");
    MemtableDev mt{c->rec_keys, c->rec_boxes, c->rec_cap, c->same_pid_count, c->o_boxs, c->sids};
    /* one wavefront per range of at least 16 chunks */
    uint32_t id_grid = (uint32_t)std::max<uint64_t>(std::min<uint64_t>(c->table_buckets / (64 * 16 * 4), 4096), 1);
    IdentifyArgs ia{t, c->cfg.min_meet_time, c->cfg.max_meet_time, c->end_time_min, c->map};
    CK(hipEventRecord(c->ev[8], s));
    hipLaunchKernelGGL(k_identify, dim3(id_grid), dim3(256), 0, s, c->table, c->table_buckets, ia, mt, cnt);
    CK(hipEventRecord(c->ev[9], s));
    CK(hipGetLastError());
    CKRC(sync_counters(c));
    h = *c->hcnt;
    if (h.table_fail) return fail(c, STC_E_CAPACITY, "meeting table full");
    c->kv_count = h.kv_count;
    c->sid_count = h.sid_count;
    c->oversize_oid_count = h.oversize_oid_count;
    c->last_live = h.live;
    uint64_t occupied = h.live + h.tombs;
    c->occupied = occupied;
    /* rebuild once live keys and tombstones fill 60% of the table */
    if (occupied * 10 > c->table_buckets * 6) {
        CKRC(table_rehash(c, t, h.live));
        rebuilt = 1;
        occupied = c->occupied;
    }
    CK(hipEventRecord(c->ev[4], s));

    /* 10: flush */
    uint32_t flushed = 0;
    if (c->kv_count > c->cfg.kv_restriction) {
        RunHost run;
        FlushArgs fa{};
        fa.n = n;
        fa.ctf_count = c->cfg.ctf_count;
        fa.kv_count = c->kv_count;
        fa.kv_restriction = c->cfg.kv_restriction;
        fa.end_time_min = c->end_time_min;
        fa.cur_time = t;
        fa.max_meet_time = c->cfg.max_meet_time;
        fa.area_threshold = c->cfg.area_threshold;
        fa.compat = c->cfg.compat;
        fa.keep_sorted = (c->cfg.keep_pairs & 2) ? 1 : 0;
        fa.map[0] = c->map.lo0; fa.map[1] = c->map.lo1; fa.map[2] = c->map.hi0; fa.map[3] = c->map.hi1;
        fa.sid_count = c->sid_count;
        fa.oversize_oid_count = c->oversize_oid_count;
        fa.rec_keys = c->rec_keys;
        fa.rec_boxes = c->rec_boxes;
        fa.same_pid_count = c->same_pid_count;
        fa.o_boxs = c->o_boxs;
        fa.sids = c->sids;
        FlushOut fo;
        fo.acquire = pool_acquire;
        fo.acquire_ctx = &c->pinned;
        int rc = stc_flush_run(&c->flushw, &fa, &fo, s);
        if (rc) {
            c->pinned.release(fo.blk);
            return fail(c, rc, std::string("flush: ") + g_last_fail);
        }
        run.info = fo.info;
        run.ctfs.swap(fo.ctfs);
        run.lay = fo.lay;
        run.has_sorted = fo.has_sorted;
        run.blk = fo.blk;
        if (c->cfg.device_runs) CKRC(keep_run_on_device(c, run));
        c->runs.push_back(std::move(run));
        /* reset the memtable (geometry.cu:2111-2131) */
        CK(hipMemsetAsync(c->same_pid_count, 0, 4ull * n, s));
        CK(hipMemsetAsync(c->sids, 0, 2ull * n, s));
        hipLaunchKernelGGL(k_fill_obox, dim3(blocks_for(n, 256)), dim3(256), 0, s, c->o_boxs, n);
        CK(hipGetLastError());
        c->kv_count = 0;
        c->sid_count = 0;
        c->oversize_oid_count = 0;
        c->end_time_min = t; /* process.cpp:271 with end_time_max = cur_time */
        flushed = 1;
    }
    CK(hipEventRecord(c->ev[5], s));
    CK(hipEventSynchronize(c->ev[5]));

    c->stepped = true;
    c->last_t = t;
    if (out) {
        memset(out, 0, sizeof(*out));
        out->nodes = g.nodes;
        out->leaves = g.leaves;
        out->border_points = h.border;
        out->cross_units = g.units;
        out->candidates = h.candidates;
        out->pairs = g.pairs;
        out->active_meetings = h.live;
        out->expired = h.expired;
        out->kv_count = h.kv_count;
        out->overflow_points = h.overflow;
        out->ambiguous_pairs = g.ambiguous;
        out->ambiguous_nodes = h.ambiguous_nodes;
        out->table_buckets = c->table_buckets;
        out->table_occupied = occupied;
        out->flushed = flushed;
        out->rebuilt = rebuilt;
        hipEventElapsedTime(&out->ms_partition, c->ev[0], c->ev[1]);
        hipEventElapsedTime(&out->ms_filter, c->ev[1], c->ev[2]);
        hipEventElapsedTime(&out->ms_refine, c->ev[2], c->ev[3]);
        hipEventElapsedTime(&out->ms_meetings, c->ev[3], c->ev[4]);
        hipEventElapsedTime(&out->ms_flush, c->ev[4], c->ev[5]);
        hipEventElapsedTime(&out->ms_total, c->ev[0], c->ev[5]);
        hipEventElapsedTime(&out->ms_refine_kernel, c->ev[6], c->ev[7]);
        hipEventElapsedTime(&out->ms_identify_kernel, c->ev[8], c->ev[9]);
        out->continuing = h_cont;
        out->new_keys = h.new_keys;
    }
    return STC_OK;
}

int stc_step_device(void *vc, const double *xy_dev, uint32_t t, stc_step_out *out) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !xy_dev) return STC_E_ARG;
    if (t >= 0xFFFFFFFEu) return fail(c, STC_E_ARG, "timestep out of range");
    if (c->stepped && t <= c->last_t) return fail(c, STC_E_STATE, "timesteps must increase");
    uint64_t kv_before = c->kv_count;
    int rc = step_impl(c, (const double2 *)xy_dev, t, out);
    if (rc == STC_OK && out) out->new_records = out->kv_count - kv_before;
    return rc;
}

int stc_step(void *vc, const double *xy_host, uint32_t t, stc_step_out *out) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !xy_host) return STC_E_ARG;
    CK(hipMemcpyAsync(c->xy_in.p, xy_host, 16ull * c->n, hipMemcpyHostToDevice, c->s));
    return stc_step_device(c, (const double *)c->xy_in.p, t, out);
}

/* A step from persisted meeting units (the reference's load_meetings_pers
 * mode, geometry.cu:1609-1622): the units' pairs are inserted or extended in
 * the meeting table with their MBRs, then steps 9-10 run as after refinement. */
int stc_step_meetings(void *vc, const stc_meeting_unit *units, uint64_t nu64, uint32_t t, stc_step_out *out) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || (!units && nu64)) return STC_E_ARG;
    if (t >= 0xFFFFFFFEu) return fail(c, STC_E_ARG, "timestep out of range");
    if (c->stepped && t <= c->last_t) return fail(c, STC_E_STATE, "timesteps must increase");
    if (nu64 >= 0xFFFFFFFFull) return fail(c, STC_E_ARG, "too many meeting units for one step");
    const uint32_t nu = (uint32_t)nu64;
    const uint64_t kv_before = c->kv_count;
    hipStream_t s = c->s;
    Counters *cnt = c->dcnt;
    memset(c->hinit, 0, sizeof(Counters));
    c->hinit->kv_count = c->kv_count;
    c->hinit->sid_count = c->sid_count;
    c->hinit->oversize_oid_count = c->oversize_oid_count;
    CK(hipMemcpyAsync(cnt, c->hinit, sizeof(Counters), hipMemcpyHostToDevice, s));
    CK(hipEventRecord(c->ev[0], s));
    CK(hipEventRecord(c->ev[1], s));
    CK(hipEventRecord(c->ev[2], s));
    if (c->units_in.ensure(nu + 1ull) || c->ukeys.ensure(nu + 1ull) || c->ukeys_alt.ensure(nu + 1ull) ||
        c->uidx.ensure(nu + 1ull) || c->uidx_alt.ensure(nu + 1ull) ||
        c->utemp.ensure(stc_radix_temp_words(nu + 1) + 64) || c->amb_pairs.ensure(nu + 1ull))
        return fail(c, STC_E_CAPACITY, "meeting unit buffers");
    if (nu) {
        CK(hipMemcpyAsync(c->units_in.p, units, (size_t)nu * sizeof(stc_meeting_unit), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_units_decode, dim3(blocks_for(nu, 256)), dim3(256), 0, s, c->units_in.p, nu, c->n,
                           c->ukeys.p, c->uidx.p, cnt);
        CK(hipGetLastError());
        CKRC(stc_radix_sort_u64(c->ukeys.p, c->uidx.p, c->ukeys_alt.p, c->uidx_alt.p, nu, 0,
                                32 + std::max(1, ilog2_ceil(c->n)), c->utemp.p, s));
        hipLaunchKernelGGL(k_units_gather, dim3(blocks_for(nu, 256)), dim3(256), 0, s, c->ukeys.p, c->uidx.p,
                           c->units_in.p, nu, c->amb_pairs.p, cnt);
        CK(hipGetLastError());
    }
    CKRC(sync_counters(c));
    if (c->hcnt->bad_points)
        return fail(c, STC_E_INPUT, "meeting units must name distinct pairs of objects below num_objects "
                                    "by the Cantor key of (min, max)");
    uint32_t rebuilt = 0;
    if ((c->occupied + nu) * 10 > c->table_buckets * 6) {
        CKRC(table_rehash(c, t - 1, c->last_live + nu));
        rebuilt = 1;
    }
    Counters h{};
    for (int attempt = 0;; attempt++) {
        CK(hipMemsetAsync(&cnt->new_keys, 0, 2 * sizeof(unsigned long long), s)); /* new_keys, continuing */
        CK(hipMemsetAsync(&cnt->table_fail, 0, sizeof(uint32_t), s));
        CK(hipEventRecord(c->ev[6], s));
        if (nu)
            hipLaunchKernelGGL(k_table_insert, dim3(blocks_for(nu, 256)), dim3(256), 0, s, c->table,
                               c->table_buckets, c->amb_pairs.p, (uint64_t)nu, t, cnt);
        CK(hipGetLastError());
        CK(hipEventRecord(c->ev[7], s));
        CKRC(sync_counters(c));
        h = *c->hcnt;
        if (!h.table_fail) break;
        if (attempt >= 8) return fail(c, STC_E_CAPACITY, "meeting table cannot hold the step's pairs");
        CKRC(table_rehash(c, t - 1, c->last_live + nu + 1024)); /* reinserting is idempotent */
        rebuilt = 1;
    }
    c->last_new = h.new_keys;
    c->occupied += h.new_keys;
    c->host_pairs_extra.clear();
    c->last_listed = 0;
    if ((c->cfg.keep_pairs & 1) && nu) {
        if (c->pairs.ensure(nu)) return fail(c, STC_E_CAPACITY, "pair list allocation failed");
        CK(hipMemcpyAsync(c->pairs.p, c->amb_pairs.p, (size_t)nu * sizeof(PairHit), hipMemcpyDeviceToDevice, s));
        c->last_listed = nu;
    }
    c->last_pairs = nu;
    CK(hipEventRecord(c->ev[3], s));
    StepGeom g{0, 0, 0, nu, 0, rebuilt};
    int rc = step_tail(c, t, g, out);
    if (rc == STC_OK && out) out->new_records = out->kv_count - kv_before;
    return rc;
}

/* The last step's pairs as the reference persists them per second
 * (geometry.cu:456-466, 1718-1723), in pair order. */
int stc_get_step_meetings(void *vc, stc_meeting_unit *units, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    if (!(c->cfg.keep_pairs & 1)) return fail(c, STC_E_STATE, "pairs are kept only with keep_pairs bit 0 set");
    std::vector<PairHit> ph(c->last_listed);
    if (c->last_listed) CK(hipMemcpy(ph.data(), c->pairs.p, c->last_listed * sizeof(PairHit), hipMemcpyDeviceToHost));
    for (const PairHit &e : c->host_pairs_extra) ph.push_back(e);
    std::sort(ph.begin(), ph.end(), [](const PairHit &a, const PairHit &b) { return a.key < b.key; });
    ph.erase(std::unique(ph.begin(), ph.end(), [](const PairHit &a, const PairHit &b) { return a.key == b.key; }),
             ph.end());
    *n = ph.size();
    if (units)
        for (uint64_t i = 0; i < std::min<uint64_t>(cap, ph.size()); i++) {
            const uint64_t a = ph[i].key >> 32, b = ph[i].key & 0xFFFFFFFFull;
            units[i].key = (a + b) * (a + b + 1) / 2 + b;
            units[i].start = units[i].end = c->last_t;
            memcpy(units[i].mbr, ph[i].mbr, sizeof(units[i].mbr));
        }
    return STC_OK;
}

int stc_get_pairs(void *vc, uint64_t *keys, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    if (!(c->cfg.keep_pairs & 1)) return fail(c, STC_E_STATE, "pairs are kept only with keep_pairs bit 0 set");
    std::vector<PairHit> ph(c->last_listed);
    if (c->last_listed) CK(hipMemcpy(ph.data(), c->pairs.p, c->last_listed * sizeof(PairHit), hipMemcpyDeviceToHost));
    std::vector<uint64_t> k(ph.size());
    for (size_t i = 0; i < ph.size(); i++) k[i] = ph[i].key;
    for (const PairHit &e : c->host_pairs_extra) k.push_back(e.key);
    std::sort(k.begin(), k.end());
    k.erase(std::unique(k.begin(), k.end()), k.end());
    *n = k.size();
    if (keys) memcpy(keys, k.data(), std::min<uint64_t>(cap, k.size()) * 8);
    return STC_OK;
}

int stc_get_active(void *vc, uint64_t *keys, uint32_t *start, uint32_t *end, float *mbr, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    std::vector<Bucket> tb(c->table_buckets);
    CK(hipMemcpy(tb.data(), c->table, c->table_buckets * sizeof(Bucket), hipMemcpyDeviceToHost));
    std::vector<Bucket> live;
    for (const Bucket &b : tb)
        if (!key_free(b.key) && !key_tomb(b.key) && b.end == c->last_t) live.push_back(b);
    std::sort(live.begin(), live.end(), [](const Bucket &a, const Bucket &b) { return a.key < b.key; });
    *n = live.size();
    for (uint64_t i = 0; i < std::min<uint64_t>(cap, live.size()); i++) {
        if (keys) keys[i] = live[i].key;
        if (start) start[i] = live[i].start;
        if (end) end[i] = live[i].end;
        if (mbr) for (int q = 0; q < 4; q++) mbr[4 * i + q] = o2f(live[i].mbr[q]);
    }
    return STC_OK;
}

int stc_get_records(void *vc, uint64_t *keys, float *boxes, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    *n = c->kv_count;
    uint64_t m = std::min<uint64_t>(cap, c->kv_count);
    if (m && keys) CK(hipMemcpy(keys, c->rec_keys, m * 16, hipMemcpyDeviceToHost));
    if (m && boxes) CK(hipMemcpy(boxes, c->rec_boxes, m * 16, hipMemcpyDeviceToHost));
    return STC_OK;
}

int stc_get_leaves(void *vc, double *boxes, uint32_t *counts, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    std::vector<Node> nd(c->last_nodes);
    std::vector<uint32_t> lc(c->last_leaves);
    CK(hipMemcpy(nd.data(), c->nodes.p, nd.size() * sizeof(Node), hipMemcpyDeviceToHost));
    if (!lc.empty()) CK(hipMemcpy(lc.data(), c->leaf_count.p, lc.size() * 4, hipMemcpyDeviceToHost));
    struct L { double lo1, lo0, hi0, hi1; uint32_t cnt; };
    std::vector<L> ls;
    for (const Node &x : nd)
        if (x.type == T_LEAF) ls.push_back(L{x.lo1, x.lo0, x.hi0, x.hi1, lc[x.leaf]});
    std::sort(ls.begin(), ls.end(), [](const L &a, const L &b) {
        if (a.lo1 != b.lo1) return a.lo1 < b.lo1;
        return a.lo0 < b.lo0;
    });
    *n = ls.size();
    for (uint64_t i = 0; i < std::min<uint64_t>(cap, ls.size()); i++) {
        if (boxes) {
            boxes[4 * i] = ls[i].lo0; boxes[4 * i + 1] = ls[i].lo1;
            boxes[4 * i + 2] = ls[i].hi0; boxes[4 * i + 3] = ls[i].hi1;
        }
        if (counts) counts[i] = ls[i].cnt;
    }
    return STC_OK;
}

int stc_num_runs(void *vc, uint32_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    *n = (uint32_t)c->runs.size();
    return STC_OK;
}

int stc_get_run_info(void *vc, uint32_t run, stc_run_info *info) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !info || run >= c->runs.size()) return STC_E_ARG;
    *info = c->runs[run].info;
    return STC_OK;
}

int stc_get_run(void *vc, uint32_t run, stc_ctf_meta *ctfs, uint8_t *key_bytes, uint8_t *bitmaps, uint16_t *sids,
                uint64_t *o_keys, float *o_boxes, uint8_t *o_bitmap, uint64_t *sorted_keys, float *sorted_boxes) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || run >= c->runs.size()) return STC_E_ARG;
    RunHost &R = c->runs[run];
    CK(hipEventSynchronize(R.blk.ev));
    const uint8_t *src = R.blk.p;
    const RunLayout &L = R.lay;
    if (!R.finalized) {
        if (stc_flush_raster_capped(src + L.counters))
            for (auto &m : R.ctfs) m.flags |= 8;
        R.finalized = true;
    }
    const stc_run_info &I = R.info;
    if (ctfs) memcpy(ctfs, R.ctfs.data(), R.ctfs.size() * sizeof(stc_ctf_meta));
    if (key_bytes) memcpy(key_bytes, src + L.key_bytes, I.key_bytes);
    if (bitmaps) memcpy(bitmaps, src + L.bitmaps, (size_t)I.ctf_count * 8192);
    if (sids) memcpy(sids, src + L.sids, (size_t)I.num_objects * 2);
    if (o_keys) memcpy(o_keys, src + L.o_keys, (size_t)I.oversize_count * 16);
    if (o_boxes) memcpy(o_boxes, src + L.o_boxes, (size_t)I.oversize_count * 16);
    if (o_bitmap) memcpy(o_bitmap, src + L.o_bitmap, 8192);
    if (R.has_sorted && sorted_keys) memcpy(sorted_keys, src + L.sorted_keys, (size_t)I.kv_count * 16);
    if (R.has_sorted && sorted_boxes) memcpy(sorted_boxes, src + L.sorted_boxes, (size_t)I.kv_count * 16);
    return STC_OK;
}

int stc_sync(void *vc) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return STC_E_ARG;
    CK(hipStreamSynchronize(c->s));
    if (c->flushw.cs) CK(hipStreamSynchronize(c->flushw.cs));
    return STC_OK;
}

int stc_live_search(void *vc, const uint32_t *pids, uint32_t n, stc_hit *hits, uint64_t cap, uint64_t *n_hits) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || (!pids && n) || !n_hits) return STC_E_ARG;
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    /* live at t and established (t - start >= min_meet + 1): start <= t - min_meet - 1 and end >= t,
     * the overlap predicate of an object query with t0 = t, t1 = t - min_meet - 1 */
    const uint32_t t = c->last_t, m = c->cfg.min_meet_time + 1;
    if (t < m) {
        *n_hits = 0;
        return STC_OK;
    }
    std::vector<stc_query> q(n);
    for (uint32_t i = 0; i < n; i++) {
        memset(&q[i], 0, sizeof(stc_query));
        q[i].kind = STC_Q_OID;
        q[i].oid = pids[i];
        q[i].t0 = t;
        q[i].t1 = t - m;
    }
    return stc_query_batch(vc, q.data(), n, hits, cap, n_hits, nullptr);
}

/* x86-64 images of the reference's CTF and CTB classes (step_merge.h:64-131) */
struct RefCtfImage {
    float mbr[4];
    uint64_t keys, bitmap;
    uint32_t kv_capacity, start_time_min, start_time_max, end_time_min, end_time_max, bitmap_size;
    uint16_t key_bit, id_bit, duration_bit, end_bit, low_x_bit, low_y_bit, edge_bit, mbr_bit, x_grid, y_grid;
    uint32_t pad;
};
struct RefCtbImage {
    uint64_t ctfs;
    uint32_t ctf_count, start_time_min, start_time_max, end_time_min, end_time_max, pad0;
    uint64_t first_widpid, sids, bitmaps, bitmap_mbrs, ctf_capacity;
    uint32_t o_count, pad1;
    uint64_t o_keys, o_boxes, o_bitmaps;
    uint32_t o_start_time_min, o_start_time_max, o_end_time_min, o_end_time_max;
    uint64_t box_rtree;
};
static_assert(sizeof(RefCtfImage) == 80 && sizeof(RefCtbImage) == 128, "reference struct images");

static bool write_file(const std::string &path, const void *a, size_t na, const void *b = nullptr, size_t nb = 0,
                       const void *c = nullptr, size_t nc = 0, const void *d = nullptr, size_t nd = 0,
                       const void *e = nullptr, size_t ne = 0) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = true;
    const void *p[] = {a, b, c, d, e};
    const size_t n[] = {na, nb, nc, nd, ne};
    for (int i = 0; i < 5 && ok; i++)
        if (n[i]) ok = fwrite(p[i], 1, n[i], f) == n[i];
    ok = (fclose(f) == 0) && ok;
    return ok;
}

int stc_dump_run(void *vc, uint32_t run, const char *dir, const char *prefix, uint32_t ctb) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !dir || !prefix || run >= c->runs.size()) return STC_E_ARG;
    RunHost &R = c->runs[run];
    CK(hipEventSynchronize(R.blk.ev));
    const uint8_t *src = R.blk.p;
    const RunLayout &L = R.lay;
    const stc_run_info &I = R.info;
    const std::string base = std::string(dir) + "/" + prefix;
    uint64_t off = 0;
    uint32_t smin = 0xFFFFFFFFu, smax = 0;
    for (uint32_t k = 0; k < I.ctf_count; k++) {
        const stc_ctf_meta &m = R.ctfs[k];
        const uint64_t nbytes = (uint64_t)m.kv_capacity * (m.key_bit / 8);
        const std::string tag = std::to_string(ctb) + "-" + std::to_string(k);
        if (!write_file(base + "_SSTable_" + tag, src + L.key_bytes + off, nbytes))
            return fail(c, STC_E_ARG, "cannot write " + base + "_SSTable_" + tag);
        off += nbytes;
        RefCtfImage img;
        memset(&img, 0, sizeof(img));
        memcpy(img.mbr, m.mbr, sizeof(img.mbr));
        img.kv_capacity = m.kv_capacity;
        img.start_time_min = m.start_time_min;
        img.start_time_max = m.start_time_max;
        img.end_time_min = m.end_time_min;
        img.end_time_max = m.end_time_max;
        img.bitmap_size = m.bitmap_size;
        img.key_bit = m.key_bit;
        img.id_bit = m.id_bit;
        img.duration_bit = m.duration_bit;
        img.end_bit = m.end_bit;
        img.low_x_bit = m.low_x_bit;
        img.low_y_bit = m.low_y_bit;
        img.edge_bit = m.edge_bit;
        img.mbr_bit = m.mbr_bit;
        img.x_grid = m.x_grid;
        img.y_grid = m.y_grid;
        if (!write_file(base + "_STcL" + tag, &img, sizeof(img), src + L.bitmaps + (uint64_t)k * 8192,
                        std::min<uint32_t>(m.bitmap_size, 8192)))
            return fail(c, STC_E_ARG, "cannot write " + base + "_STcL" + tag);
        smin = std::min(smin, m.start_time_min);
        smax = std::max(smax, m.start_time_max);
    }
    RefCtbImage img;
    memset(&img, 0, sizeof(img));
    img.ctf_count = I.ctf_count;
    img.start_time_min = I.ctf_count ? smin : 0;
    img.start_time_max = smax;
    img.end_time_min = I.end_time_min;
    img.end_time_max = I.cur_time;
    img.o_count = I.oversize_count;
    img.o_start_time_min = I.o_start_min;
    img.o_start_time_max = I.o_start_max;
    img.o_end_time_min = I.o_end_min;
    img.o_end_time_max = I.o_end_max;
    if (!write_file(base + "_CTB" + std::to_string(ctb), &img, sizeof(img), src + L.sids, (size_t)I.num_objects * 2,
                    src + L.o_keys, (size_t)I.oversize_count * 16, src + L.o_boxes, (size_t)I.oversize_count * 16,
                    src + L.o_bitmap, 8192))
        return fail(c, STC_E_ARG, "cannot write " + base + "_CTB" + std::to_string(ctb));
    return STC_OK;
}

/* Releases the host copies of all flushed runs. Copies still in flight finish
 * before their memory is reused. */
int stc_drop_runs(void *vc) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return STC_E_ARG;
    for (auto &r : c->runs) {
        c->pinned.release(r.blk);
        if (r.dev) {
            (void)hipStreamSynchronize(c->s);
            (void)hipFree(r.dev);
            r.dev = nullptr;
        }
    }
    c->runs.clear();
    return STC_OK;
}

/* Merges runs [first, first + count) into one run in their place. Every record
 * of the inputs (CTF records decoded as compaction.cpp:320-336 does, oversize
 * records as stored) forms a memtable whose end offsets are relative to the
 * earliest input base (compaction.cpp:483-487); it is flushed by the step's
 * own flush with cur_time = the latest input's. Objects oversize in any input
 * stay oversize (compaction.cpp:259-268). */
int stc_compact_runs(void *vc, uint32_t first, uint32_t count) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || count == 0) return STC_E_ARG;
    if ((uint64_t)first + count > c->runs.size()) return fail(c, STC_E_ARG, "run range out of bounds");
    if (count == 1) return STC_OK;
    const uint32_t n = c->n;
    uint64_t total = 0;
    uint32_t base = 0xFFFFFFFFu, tmax = 0;
    std::vector<uint64_t> in_ctf(count, 0); /* records held in each input's CTFs */
    for (uint32_t r = first; r < first + count; r++) {
        const RunHost &R = c->runs[r];
        if (!R.dev) return fail(c, STC_E_STATE, "compaction reads runs kept in HBM (stc_config.device_runs)");
        for (const auto &m : R.ctfs) in_ctf[r - first] += m.kv_capacity;
        total += in_ctf[r - first] + R.info.oversize_count;
        base = std::min(base, R.info.end_time_min);
        tmax = std::max(tmax, R.info.cur_time);
    }
    hipStream_t s = c->s;
    uint64_t *keys = nullptr;
    float4 *boxes = nullptr;
    uint32_t *spc = nullptr, *dcount = nullptr;
    uint4 *obox = nullptr;
    uint16_t *sids = nullptr;
    auto release = [&]() {
        (void)hipStreamSynchronize(s);
        for (void *p : {(void *)keys, (void *)boxes, (void *)spc, (void *)dcount, (void *)obox, (void *)sids})
            if (p) (void)hipFree(p);
    };
    if (hipMalloc((void **)&keys, std::max<uint64_t>(total, 1) * 16) ||
        hipMalloc((void **)&boxes, std::max<uint64_t>(total, 1) * 16) || hipMalloc((void **)&spc, 4ull * n) ||
        hipMalloc((void **)&dcount, 8) || hipMalloc((void **)&obox, 16ull * n) || hipMalloc((void **)&sids, 2ull * n)) {
        release();
        return fail(c, STC_E_CAPACITY, "compaction buffers");
    }
    int rc = STC_OK;
    uint32_t h_count[2] = {0, 0}; /* objects with records, bad ids */
    if (hipMemsetAsync(spc, 0, 4ull * n, s) || hipMemsetAsync(sids, 0, 2ull * n, s) ||
        hipMemsetAsync(dcount, 0, 8, s))
        rc = STC_E_HIP;
    if (!rc) {
        hipLaunchKernelGGL(k_fill_obox, dim3(blocks_for(n, 256)), dim3(256), 0, s, obox, n);
        if (hipGetLastError() != hipSuccess) rc = STC_E_HIP;
    }
    uint64_t out0 = 0;
    for (uint32_t r = first; r < first + count && !rc; r++) {
        const RunHost &R = c->runs[r];
        rc = stc_compact_decode(R.qdev, in_ctf[r - first], n, base, out0, keys, boxes, spc, obox, sids, dcount + 1,
                                s);
        out0 += in_ctf[r - first] + R.info.oversize_count;
    }
    if (!rc) rc = stc_count_objects(spc, n, dcount, s);
    if (!rc && (hipMemcpyAsync(h_count, dcount, 8, hipMemcpyDeviceToHost, s) || hipStreamSynchronize(s)))
        rc = STC_E_HIP;
    if (rc) {
        release();
        return fail(c, rc, std::string("compaction decode: ") + g_last_fail);
    }
    if (h_count[1]) {
        release();
        return fail(c, STC_E_STATE, "an input run holds a record whose object id is not below num_objects");
    }
    FlushArgs fa{};
    fa.n = n;
    fa.ctf_count = c->cfg.ctf_count;
    fa.kv_count = total;
    fa.kv_restriction = (uint32_t)std::min<uint64_t>((uint64_t)c->cfg.kv_restriction * count, 0xFFFFFFFFu);
    fa.end_time_min = base;
    fa.cur_time = tmax;
    fa.max_meet_time = c->cfg.max_meet_time;
    fa.area_threshold = c->cfg.area_threshold;
    fa.compat = c->cfg.compat;
    fa.keep_sorted = (c->cfg.keep_pairs & 2) ? 1 : 0;
    fa.sticky_oversize = 1;
    fa.map[0] = c->map.lo0; fa.map[1] = c->map.lo1; fa.map[2] = c->map.hi0; fa.map[3] = c->map.hi1;
    fa.sid_count = h_count[0];
    fa.oversize_oid_count = 0; /* counted by the flush: objects with records that are oversize */
    fa.rec_keys = keys;
    fa.rec_boxes = boxes;
    fa.same_pid_count = spc;
    fa.o_boxs = obox;
    fa.sids = sids;
    FlushOut fo;
    fo.acquire = pool_acquire;
    fo.acquire_ctx = &c->pinned;
    rc = stc_flush_run(&c->flushw, &fa, &fo, s);
    if (rc) {
        c->pinned.release(fo.blk);
        release();
        return fail(c, rc, std::string("compaction flush: ") + g_last_fail);
    }
    RunHost run;
    run.info = fo.info;
    run.ctfs.swap(fo.ctfs);
    run.lay = fo.lay;
    run.has_sorted = fo.has_sorted;
    run.blk = fo.blk;
    rc = keep_run_on_device(c, run);
    release();
    if (rc) {
        c->pinned.release(run.blk);
        if (run.dev) (void)hipFree(run.dev);
        return rc;
    }
    CK(hipStreamSynchronize(s));
    for (uint32_t r = first; r < first + count; r++) {
        c->pinned.release(c->runs[r].blk);
        (void)hipFree(c->runs[r].dev);
    }
    c->runs.erase(c->runs.begin() + first, c->runs.begin() + first + count);
    c->runs.insert(c->runs.begin() + first, std::move(run));
    return STC_OK;
}

int stc_query_batch(void *vc, const stc_query *q, uint32_t nq, stc_hit *hits, uint64_t cap, uint64_t *n_hits,
              uint64_t *counts) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || (!q && nq) || !n_hits) return STC_E_ARG;
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    for (uint32_t i = 0; i < nq; i++)
        if (q[i].kind < STC_Q_OID || q[i].kind > STC_Q_PAIR) return fail(c, STC_E_ARG, "unknown query kind");
    if (!c->qb) c->qb = stc_query_bufs_new();
    QueryState st;
    st.table = c->table;
    st.nb = c->table_buckets;
    st.last_t = c->last_t;
    st.stepped = 1;
    st.rec_keys = c->rec_keys;
    st.rec_boxes = c->rec_boxes;
    st.n_rec = c->kv_count;
    st.end_time_min = c->end_time_min;
    st.num_objects = c->n;
    st.map[0] = c->map.lo0;
    st.map[1] = c->map.lo1;
    st.map[2] = c->map.hi0;
    st.map[3] = c->map.hi1;
    for (uint32_t r = 0; r < c->runs.size(); r++) {
        if (!c->runs[r].dev) continue; /* kept on the host only */
        QueryRunDev d = c->runs[r].qdev;
        d.source = 2 + r;
        st.runs.push_back(d);
        st.run_ctfs.push_back(c->runs[r].ctfs.data());
        st.run_stats.push_back(c->runs[r].stats.data());
    }
    int rc = stc_query_run(c->qb, st, q, nq, hits, cap, n_hits, counts, c->s);
    if (rc) return fail(c, rc, std::string("query: ") + g_last_fail);
    return STC_OK;
}

} /* extern "C" */
