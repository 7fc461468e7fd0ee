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
 *  7. refine    the combinations of the check units (cuda_refinement,
 *               geometry.cu:424-497) regrouped by the owner of their pair (its
 *               smaller id): one wavefront per owner counts its pairs, a prefix
 *               sum gives each owner a segment, and a second pass writes them.
 *  8. meetings  the live meetings are key-sorted segments, one per owner; a
 *  9. identify  merge per owner joins the step's pairs with the previous
 *               step's segment (insert/extend, geometry.cu:469-494) and turns
 *               meetings without a pair into records when they lasted at least
 *               min_meet_time and stay inside the map (cuda_identify_meetings,
 *               geometry.cu:595-737).
 * 10. flush     when the memtable exceeds kv_restriction (stc_flush.hip).
 */
#include "stc_internal.h"
#include "stc_geo.h"
#include "stc_gen.h"
#include "stc_flush.h"
#include "stc_table.h"
#include "stc_schema.h"
#include "stc_query.h"
#include "../../include/stcdb.h"

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <atomic>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <type_traits>

/* 4-byte aligned vectors for the streaming stores of the 12-byte slots and 28-byte buckets */
typedef uint32_t u32x4u __attribute__((ext_vector_type(4), aligned(4)));
typedef uint32_t u32x3u __attribute__((ext_vector_type(3), aligned(4)));

#define ONE_DIM 16384u
#define NLEVELS 15
#define STACK_DEPTH 96
#define STC_QUERY_RESERVE (1ull << 30) /* device memory a query batch leaves free */

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

struct FrontierItem { uint32_t node, code; };

struct PairHit {              /* 24 B: key and the float MBR of the two positions (replayed meeting units) */
    unsigned long long key;
    float mbr[4];
};

/* 12 B (packed, 4-byte aligned): the partner and its position as floats.
 * The slots of a step are laid out owner by owner (slot i belongs to the
 * owner a with poff[a] <= i < poff[a + 1]), so the owner is not stored; its
 * position is posf[owner], and the pair's MBR is the box of the two (the box
 * of their float conversions, as the reference's f_box of the two points). */
struct __attribute__((packed, aligned(4))) PairSlot {
    uint32_t partner;
    float px, py;
};
static_assert(sizeof(PairSlot) == 12, "a pair slot is 12 bytes");

/* a slot's key (owner << 32 | partner) and ordered-int MBR (f2o of low0, low1,
 * high0, high1); a replayed PairHit carries both itself */
__device__ inline unsigned long long slot_key(const PairHit &e, uint32_t) { return e.key; }
__device__ inline unsigned long long slot_key(const PairSlot &e, uint32_t owner) {
    return ((unsigned long long)owner << 32) | e.partner;
}
__device__ inline uint4 slot_mbr(const PairHit &e, const float2 *, uint32_t) {
    return make_uint4(f2o(e.mbr[0]), f2o(e.mbr[1]), f2o(e.mbr[2]), f2o(e.mbr[3]));
}
__device__ inline uint4 slot_mbr(const PairSlot &e, const float2 *posf, uint32_t owner) {
    const float2 o = posf[owner];
    return make_uint4(f2o(fminf(o.x, e.px)), f2o(fminf(o.y, e.py)), f2o(fmaxf(o.x, e.px)), f2o(fmaxf(o.y, e.py)));
}

struct Counters {             /* device-side counters, copied back once per sync point */
    unsigned long long units;
    unsigned long long new_keys;
    unsigned long long continuing;
    unsigned long long ambiguous_nodes;
    unsigned long long live, expired;
    unsigned long long kv_count;
    uint32_t nodes, leaves;
    uint32_t frontier[NLEVELS + 1];
    uint32_t bad_points;
    uint32_t overflow;
    uint32_t oversize_oid_count;
    uint32_t table_fail;
    uint32_t lookup_overflow;          /* border lookups whose node stack was full */
    uint32_t border_part[64];          /* border points, by partition workgroup mod 64 */
    /* the collect pass's counters, from candidates to pk_over: cleared together before each pass */
    unsigned long long candidates;
    unsigned long long pairs;          /* slots reserved by the collect pass (pairs and band combinations) */
    unsigned long long ambiguous;
    unsigned long long pk_used[64];    /* psrc slots reserved in each slice by the collecting waves */
    unsigned long long cand_part[64], pairs_part[64]; /* the collect pass's totals, by workgroup mod 64 */
    uint32_t big, huge, giant;         /* segments left to the wave sort, the block sort, the global-scratch sort */
    uint32_t pk_over;                  /* a collecting wave's reservation exceeded its slice of psrc */
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

/* ------------------------------------------------------------------ */
/* kernels: prep, tree, partition                                      */
/* ------------------------------------------------------------------ */

__global__ void k_prep(const double2 *__restrict__ xy, uint32_t n, MapBox m, uint32_t *__restrict__ codes,
                       uint32_t *__restrict__ idx, double *__restrict__ dlon, float2 *__restrict__ posf,
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
    dlon[i] = stc_dpm_lon(p.y);
    if (posf) posf[i] = make_float2((float)p.x, (float)p.y);
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

/* Workgroups go to the 8 XCDs in turn (workgroup b to XCD b % 8), each XCD
 * with its own L2. This id gives each XCD a contiguous eighth of the grid, so
 * the owners that the waves of one XCD handle at a time are neighbours in leaf
 * order and share their candidates' cache lines (with the plain id, each
 * XCD's L2 would hold the neighbourhoods of the whole grid's owners). A grid
 * that is not a multiple of 8 keeps the plain id. */
__device__ inline uint32_t xcd_block_id() {
    const uint32_t nb = gridDim.x, b = blockIdx.x;
    return (nb & 7u) ? b : (b & 7u) * (nb >> 3) + (b >> 3);
}

/* One thread per point, points in Morton order (midx): a wave's points descend
 * the same subtrees, and a leaf's points are consecutive, so each run of
 * equal leaves in a wave adds its length to leaf_count once. */
__global__ void k_partition(const double2 *__restrict__ xy, const uint32_t *__restrict__ midx, uint32_t n,
                            const Node *__restrict__ nodes, double xbuf, double ybuf, uint32_t *__restrict__ pt_leaf,
                            uint32_t *__restrict__ pt_lv, uint32_t *__restrict__ leaf_count, Counters *cnt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (STC_WAVE - 1);
    bool border = false;
    uint32_t leaf = 0xFFFFFFFFu;
    if (j < n) {
        const uint32_t p = midx[j];
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
        leaf = nodes[cur].leaf;
        pt_leaf[p] = leaf;
        border = (lv != cur);
        pt_lv[p] = border ? lv : 0xFFFFFFFFu;
    }
    /* runs of equal leaves: the head of each run adds the run's length */
    const uint32_t prev = __shfl_up(leaf, 1, STC_WAVE);
    const bool head = j < n && (lane == 0 || prev != leaf);
    const uint64_t heads = __ballot(head), valid = __ballot(j < n);
    if (head) {
        const uint64_t later = lane == STC_WAVE - 1 ? 0ull : (heads >> (lane + 1));
        const uint32_t end = later ? (uint32_t)(lane + __ffsll((long long)later)) : (uint32_t)__popcll(valid);
        atomicAdd(&leaf_count[leaf], end - (uint32_t)lane);
    }
    /* border points: summed per workgroup, then one update of a counter of 64 (the host adds them) */
    __shared__ uint32_t bsum[1024 / STC_WAVE];
    const uint64_t b = __ballot(border);
    if (lane == 0) bsum[threadIdx.x / STC_WAVE] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tb = 0;
        for (uint32_t k = 0; k < blockDim.x / STC_WAVE; k++) tb += bsum[k];
        if (tb) atomicAdd(&cnt->border_part[blockIdx.x % 64], tb);
    }
}

/* positions of the two points of each band combination, for the host's decision */
__global__ void k_amb_gather(const double2 *__restrict__ xy, const uint2 *__restrict__ amb, uint32_t n_amb,
                             double4 *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_amb) return;
    const uint2 e = amb[i];
    const double2 p = xy[e.x], q = xy[e.y];
    out[i] = make_double4(p.x, p.y, q.x, q.y);
}

__global__ void k_iota(uint32_t *a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

/* leaf-ordered SoA and each point's rank within its leaf */
__global__ void k_gather(const double2 *__restrict__ xy, const double *__restrict__ dlon,
                         const uint32_t *__restrict__ lp, const uint32_t *__restrict__ pt_leaf,
                         const uint32_t *__restrict__ leaf_start, uint32_t n, double2 *__restrict__ lxy,
                         double *__restrict__ ld, uint32_t *__restrict__ pt_rank,
                         uint32_t *__restrict__ lleaf) {
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    uint32_t p = lp[j];
    double2 q = xy[p];
    lxy[j] = q;
    ld[j] = dlon[p];
    const uint32_t L = pt_leaf[p];
    pt_rank[p] = j - leaf_start[L];
    lleaf[j] = L;
}

/* each point's cross units in leaf order (first unit, count) */
__global__ void k_units_lorder(const uint32_t *__restrict__ lp, const uint32_t *__restrict__ uoffset,
                               const uint32_t *__restrict__ ucount, uint32_t n, uint2 *__restrict__ lunits) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t p = lp[j];
    lunits[j] = make_uint2(uoffset[p], ucount[p]);
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

/* One thread per point, points in leaf order: a wave's points share a leaf
 * and mostly the same subtrees, so their descents stay together. Outputs are
 * per point id (count[p], units from offset[p]). */
template <bool EMIT>
__global__ void k_lookup(const uint32_t *__restrict__ lp, const double2 *__restrict__ lxy,
                         const double *__restrict__ ld, uint32_t n,
                         const Node *__restrict__ nodes, const uint32_t *__restrict__ pt_lv, double r, double band,
                         uint32_t *__restrict__ count, const uint32_t *__restrict__ offset,
                         uint2 *__restrict__ units, Counters *cnt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint32_t p = lp[j];
    uint32_t lv = pt_lv[p];
    if (lv == 0xFFFFFFFFu) {
        if (!EMIT) count[p] = 0;
        return;
    }
    const double2 q = lxy[j];
    const double dl = ld[j];
    uint32_t stk[STACK_DEPTH];
    int sp = 0;
    stk[sp++] = lv;
    uint32_t found = 0, amb = 0;
    bool full = false;
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
                } else {
                    full = true;
                }
            }
        }
    }
    if (!EMIT) {
        count[p] = found;
        if (amb) atomicAdd(&cnt->ambiguous_nodes, (unsigned long long)amb);
        if (full) atomicAdd(&cnt->lookup_overflow, 1u);
    }
}

/* ------------------------------------------------------------------ */
/* meetings: pairs grouped by owner, merged with the live meetings      */
/* ------------------------------------------------------------------ */
/*
 * The live meetings after a step are an array of 28 B buckets sorted by key
 * (oid1 << 32 | oid2): one segment per owner oid1 (the pair's smaller id), in
 * owner order, each segment's live buckets followed by tombstones for the
 * step's duplicate or rejected pairs. A step finds its pairs owner by owner,
 * once (collect each owner's pair sources, then place them sorted into the
 * owner's segment), and one merge per owner
 * joins the new segment with the previous step's: a pair continues a meeting
 * of the previous step or opens one (insert/extend, geometry.cu:469-494), a
 * meeting of the previous step without a pair now expires and, when it lasted
 * long enough, becomes two memtable records (identify, geometry.cu:595-737).
 * Every pass streams whole segments instead of touching one random bucket per
 * pair.
 */

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = STC_WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, STC_WAVE);
    return v;
}

/* geometry of the step that the pair passes read */
/* a border point of the reverse lists in one 32-byte read: position, 1/dpm_lon
 * and id (its dpm_lon, needed only inside the guard band, is kept apart) */
struct __align__(32) RevPt {
    double x, y, il;
    uint32_t pid, pad;
};

struct PairGeo {
    const double2 *xy;
    const double *dlon;
    const double2 *lxy;                  /* positions in leaf order (one 16-byte read per candidate) */
    const uint32_t *pt_leaf, *pt_rank, *leaf_start, *leaf_count, *lp;
    const double *ld;                    /* dpm_lon in leaf order */
    const uint32_t *lleaf;               /* leaf of each point, in leaf order */
    const uint2 *lunits;                 /* (first cross unit, count) of each point, in leaf order */
    const uint2 *urange;                 /* per cross unit: its leaf's kept members above the unit's point (first, count) */
    const uint4 *prange;                 /* per point in leaf order: its later leaf members and the border
                                            points above it with a unit to its leaf (first, count, first, count) */
    const uint2 *units;                  /* cross units (pid, leaf), point after point */
    const uint32_t *uoffset, *ucount;    /* a point's cross units: units[uoffset[p] ..] */
    const uint32_t *rev_start, *rev_pid; /* per leaf, by pid: the border points with a cross unit to it */
    const RevPt *rev;                    /* their positions, 1/dpm_lon and ids, in rev_pid order */
    const double *rev_dl;                /* their dpm_lon */
    uint32_t n, buf;
    double r, band, r2_lo, r2_hi;
};

/* The T1 decision for a probe point and another point: 1 within r, 0 beyond,
 * 2 inside the guard band around r (decided on the host with the C library's
 * sin, geometry.h:40-48). The squared distance with reciprocal multiplies is
 * within 1.2e-15 (relative) of the formula's, so only a band of 1e-12 around
 * r^2 needs the exact divisions and square root (cuda_util.cuh:41-53). */
__device__ inline int pair_decide(double2 pp, const double *pd, double pinv, double qx, double qy, const PairGeo &g) {
    const double ax = (pp.x - qx) * pinv;
    const double ay = (pp.y - qy) * STC_INV_DPM_LAT;
    const double s2 = ax * ax + ay * ay;
    if (s2 < g.r2_lo) return 1;
    if (s2 <= g.r2_hi) {
        const double d = stc_point_dist(pp.x, pp.y, qx, qy, *pd); /* dpm_lon of the probe: read only here */
        if (fabs(d - g.r) <= g.band) return 2;
        return d <= g.r ? 1 : 0;
    }
    return 0;
}

__device__ inline float4 pair_mbr(float ax, float ay, float bx, float by) {
    return make_float4(fminf(ax, bx), fminf(ay, by), fmaxf(ax, bx), fmaxf(ay, by));
}

#define SRC_REV 0x80000000u /* source index tag: an index into the reverse lists, not the leaf-ordered arrays */

/* Every (unit, member) combination of the step whose pair is owned by a,
 * visited by the 64 lanes of a wave, at most one per lane per call of
 * f(valid, other, decision, probe, mbr, source) (source: the other point's
 * index in the leaf-ordered arrays, or SRC_REV | its index in the reverse
 * lists):
 *   (i)   a against the later (larger-id) kept members of its own leaf (the
 *         inside unit of geometry.cu:258; probe a);
 *   (ii)  a against the kept members above a of the leaves of a's cross units
 *         (geometry.cu:295-392; probe a);
 *   (iii) the border points above a that have a cross unit to a's leaf, when a
 *         is a kept member of it (probe: the border point).
 * Each combination of the reference's refinement (geometry.cu:424-497) is
 * visited exactly once, by the owner of its pair. Wave-uniform. */
/* first index i in [0, n) with arr[i] > v (n when none) for ascending arr: the
 * wave's lanes probe 64 evenly spaced entries per round, so a search takes one
 * or two dependent loads where a binary search takes log2(n). Wave-uniform. */
__device__ inline uint32_t wave_upper_bound(const uint32_t *arr, uint32_t n, uint32_t v, int lane) {
    uint32_t lo = 0, hi = n; /* the answer lies in [lo, hi] */
    while (lo < hi) {
        const uint32_t span = hi - lo;
        if (span <= STC_WAVE) {
            const bool gt = (uint32_t)lane < span ? arr[lo + lane] > v : true;
            const uint64_t m = __ballot(gt);
            return m ? lo + (uint32_t)__ffsll((long long)m) - 1 : hi;
        }
        const uint32_t step = (span + STC_WAVE - 1) / STC_WAVE;
        const uint32_t idx = lo + (uint32_t)lane * step;
        const bool gt = idx < hi ? arr[idx] > v : true;
        const uint64_t m = __ballot(gt);
        const uint32_t j = m ? (uint32_t)__ffsll((long long)m) - 1 : STC_WAVE;
        if (j == 0) return lo;
        const uint32_t nlo = lo + (j - 1) * step + 1, nhi = min(lo + j * step, hi);
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

template <bool OTHER, class F>
__device__ inline void owner_combos_by_source(uint32_t w, uint32_t a, const PairGeo &g, int lane, F &f) {
    const uint32_t L = g.lleaf[w];
    uint32_t s = g.leaf_count[L];
    if (s > g.buf) s = g.buf;
    const uint32_t st = g.leaf_start[L], rank = w - st;
    const double2 pa = g.lxy[w];
    const double pda = g.ld[w], pia = 1.0 / pda; /* dlon[a] (k_prep) and its reciprocal */
    const float afx = (float)pa.x, afy = (float)pa.y;
    /* (i) */
    for (uint32_t m0 = rank + 1; m0 < s; m0 += STC_WAVE) {
        const uint32_t m = m0 + lane;
        const bool v = m < s;
        int dec = 0;
        uint32_t q = 0;
        float4 mb = make_float4(0.f, 0.f, 0.f, 0.f);
        if (v) {
            const double2 qq = g.lxy[st + m];
            const double qx = qq.x, qy = qq.y;
            dec = pair_decide(pa, &g.ld[w], pia, qx, qy, g);
            if (OTHER) q = g.lp[st + m];
            mb = pair_mbr(afx, afy, (float)qx, (float)qy);
        }
        f(v, q, dec, a, mb, st + m);
    }
    /* (ii) */
    const uint2 un = g.lunits[w];
    const uint32_t u0 = un.x, nu = un.y;
    for (uint32_t k = 0; k < nu; k++) {
        const uint32_t L2 = g.units[u0 + k].y;
        uint32_t s2 = g.leaf_count[L2];
        if (s2 > g.buf) s2 = g.buf;
        const uint32_t st2 = g.leaf_start[L2];
        for (uint32_t m0 = wave_upper_bound(g.lp + st2, s2, a, lane); m0 < s2; m0 += STC_WAVE) {
            const uint32_t m = m0 + lane;
            const bool v = m < s2;
            int dec = 0;
            uint32_t q = 0;
            float4 mb = make_float4(0.f, 0.f, 0.f, 0.f);
            if (v) {
                const double2 qq = g.lxy[st2 + m];
                const double qx = qq.x, qy = qq.y;
                dec = pair_decide(pa, &g.ld[w], pia, qx, qy, g);
                if (OTHER) q = g.lp[st2 + m];
                mb = pair_mbr(afx, afy, (float)qx, (float)qy);
            }
            f(v, q, dec, a, mb, st2 + m);
        }
    }
    /* (iii) */
    if (rank < s) {
        const uint32_t r0 = g.rev_start[L], r1 = g.rev_start[L + 1];
        const double qx = g.lxy[st + rank].x, qy = g.lxy[st + rank].y;
        const float qfx = (float)qx, qfy = (float)qy;
        for (uint32_t i0 = r0 + wave_upper_bound(g.rev_pid + r0, r1 - r0, a, lane); i0 < r1; i0 += STC_WAVE) {
            const uint32_t i = i0 + lane;
            const bool v = i < r1;
            int dec = 0;
            uint32_t p = 0;
            float4 mb = make_float4(0.f, 0.f, 0.f, 0.f);
            if (v) {
                if (OTHER) p = g.rev[i].pid;
                const RevPt rp = g.rev[i];
                const double2 pp = make_double2(rp.x, rp.y);
                dec = pair_decide(pp, &g.rev_dl[i], rp.il, qx, qy, g);
                mb = pair_mbr((float)pp.x, (float)pp.y, qfx, qfy);
            }
            f(v, p, dec, p, mb, SRC_REV | i);
        }
    }
}

/* the number of combinations owner_combos visits for the owner at leaf-order
 * place w: its later leaf members, the members above it of its cross units'
 * leaves, the border points above it with a unit to its leaf. Wave-uniform. */
__device__ inline uint32_t owner_candidates(uint32_t w, const PairGeo &g, int lane) {
    const uint4 pr = g.prange[w];
    const uint2 un = g.lunits[w];
    unsigned long long t = 0;
    for (uint32_t k = (uint32_t)lane; k < un.y; k += STC_WAVE) t += g.urange[un.x + k].y;
    t = wave_sum_u64(t);
    return pr.y + pr.w + (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)t);
}

/* The same combinations as owner_combos_by_source, with the owner's sources
 * laid end to end so that every call of f but the last fills all 64 lanes
 * (OTHER false: the other point's id and the probe are left out of the call,
 * for a caller that rebuilds them from the source index when it needs them):
 * lane j describes source j (j = 0: (i); 1 .. nu: the cross units; nu + 1:
 * (iii)) by its first index and length, and each lane finds the source of its
 * combination from the prefix of the lengths. Owners with more than 62 cross
 * units take the per-source loops. on_total receives the number of
 * combinations before the first call of f. Wave-uniform. */
template <bool OTHER, class F, class T>
__device__ inline void owner_combos(uint32_t w, uint32_t a, const PairGeo &g, int lane, F &f, T &&on_total) {
    const uint2 un = g.lunits[w];
    const uint32_t u0 = un.x, nu = un.y;
    if (nu + 2 > STC_WAVE) {
        on_total(owner_candidates(w, g, lane));
        owner_combos_by_source<OTHER>(w, a, g, lane, f);
        return;
    }
    const uint4 pr = g.prange[w];
    const double2 pa = g.lxy[w];
    const double pda = g.ld[w], pia = 1.0 / pda; /* dlon[a] (k_prep) and its reciprocal */
    const float afx = (float)pa.x, afy = (float)pa.y;
    /* source table: first index (into the leaf-ordered arrays, or rev_pid for
     * the last source) and length, held by lane j for source j */
    uint32_t sbase = 0, slen = 0;
    if (lane == 0) {
        sbase = pr.x;
        slen = pr.y;
    } else if ((uint32_t)lane <= nu) {
        const uint2 ur = g.urange[u0 + lane - 1];
        sbase = ur.x;
        slen = ur.y;
    } else if ((uint32_t)lane == nu + 1) {
        sbase = pr.z;
        slen = pr.w;
    }
    /* exclusive prefix of the lengths */
    uint32_t scum = slen;
    for (int o = 1; o < STC_WAVE; o <<= 1) {
        const uint32_t y = __shfl_up(scum, o, STC_WAVE);
        if (lane >= o) scum += y;
    }
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane((int)scum, STC_WAVE - 1);
    on_total(total);
    scum -= slen;
    const uint32_t nseg = nu + 2;
    const double qax = pa.x, qay = pa.y; /* a's position as a member of its leaf */
    for (uint32_t c0 = 0; c0 < total; c0 += STC_WAVE) {
        const uint32_t c = c0 + lane;
        const bool v = c < total;
        uint32_t j = 0;
        for (uint32_t k = 1; k < nseg; k++) /* k is wave-uniform: a lane read, no exchange */
            if ((uint32_t)__builtin_amdgcn_readlane((int)scum, (int)k) <= c) j = k;
        const uint32_t jb = __shfl(sbase, (int)j, STC_WAVE), jc = __shfl(scum, (int)j, STC_WAVE);
        int dec = 0;
        uint32_t q = 0, probe = a, src = 0;
        float4 mb = make_float4(0.f, 0.f, 0.f, 0.f);
        if (v) {
            const uint32_t idx = jb + (c - jc);
            src = idx;
            if (j + 1 < nseg) { /* a member of the owner's leaf or of a cross unit's leaf; probe a */
                if (OTHER) q = g.lp[idx]; /* read with the position, not after the decision */
                const double2 qq = g.lxy[idx];
                const double qx = qq.x, qy = qq.y;
                dec = pair_decide(pa, &g.ld[w], pia, qx, qy, g);
                mb = pair_mbr(afx, afy, (float)qx, (float)qy);
            } else { /* a border point with a cross unit to a's leaf; probe the border point */
                const RevPt rp = g.rev[idx];
                const uint32_t p = rp.pid;
                const double2 pp = make_double2(rp.x, rp.y);
                dec = pair_decide(pp, &g.rev_dl[idx], rp.il, qax, qay, g);
                q = p;
                probe = p;
                src = SRC_REV | idx;
                mb = pair_mbr((float)pp.x, (float)pp.y, (float)qax, (float)qay);
            }
        }
        f(v, q, dec, probe, mb, src);
    }
}

/* split the cross units into (leaf, pid) key/value arrays for the reverse lists */
__global__ void k_units_split(const uint2 *__restrict__ units, uint32_t nu, uint32_t *__restrict__ leaf,
                              uint32_t *__restrict__ pid) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    const uint2 u = units[i];
    leaf[i] = u.y;
    pid[i] = u.x;
}

/* each leaf's first entry in the leaf-sorted reverse lists (rev_start[n_leaves] = nu) */
__global__ void k_rev_starts(const uint32_t *__restrict__ sorted_leaf, uint32_t nu, uint32_t n_leaves,
                             uint32_t *__restrict__ rev_start) {
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L <= n_leaves) rev_start[L] = lower_bound_u32(sorted_leaf, nu, L);
}

/* the border points' data in reverse-list order, so that owners read it coalesced */
__global__ void k_rev_gather(const uint32_t *__restrict__ rev_pid, uint32_t nr, const double2 *__restrict__ xy,
                             const double *__restrict__ dlon, RevPt *__restrict__ rev, double *__restrict__ rev_dl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    const uint32_t p = rev_pid[i];
    const double2 q = xy[p];
    const double d = dlon[p];
    RevPt r;
    r.x = q.x;
    r.y = q.y;
    r.il = 1.0 / d; /* the reciprocal of dlon[p], as the owners' loops compute theirs */
    r.pid = p;
    r.pad = 0;
    rev[i] = r;
    rev_dl[i] = d;
}

/* per cross unit (p, leaf): the leaf's kept members above p, as an index range
 * into the leaf-ordered arrays */
__global__ void k_unit_ranges(const uint2 *__restrict__ units, uint32_t nu, const uint32_t *__restrict__ leaf_start,
                              const uint32_t *__restrict__ leaf_count, const uint32_t *__restrict__ lp, uint32_t buf,
                              uint2 *__restrict__ urange) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const uint2 un = units[u];
    uint32_t s2 = leaf_count[un.y];
    if (s2 > buf) s2 = buf;
    const uint32_t st2 = leaf_start[un.y];
    const uint32_t m0 = lower_bound_u32(lp + st2, s2, un.x + 1);
    urange[u] = make_uint2(st2 + m0, s2 - m0);
}

/* per point in leaf order: its later kept members in its leaf, and (when it is
 * kept) the border points above it with a cross unit to its leaf */
__global__ void k_point_ranges(const uint32_t *__restrict__ lp, const uint32_t *__restrict__ lleaf, uint32_t n,
                               const uint32_t *__restrict__ leaf_start, const uint32_t *__restrict__ leaf_count,
                               const uint32_t *__restrict__ rev_start, const uint32_t *__restrict__ rev_pid,
                               uint32_t buf, uint4 *__restrict__ prange) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= n) return;
    const uint32_t a = lp[w], L = lleaf[w];
    uint32_t s = leaf_count[L];
    if (s > buf) s = buf;
    const uint32_t st = leaf_start[L], rank = w - st;
    uint4 r = make_uint4(0, 0, 0, 0);
    if (rank + 1 < s) {
        r.x = w + 1;
        r.y = s - rank - 1;
    }
    if (rank < s) {
        const uint32_t r0 = rev_start[L], r1 = rev_start[L + 1];
        const uint32_t i0 = r0 + lower_bound_u32(rev_pid + r0, r1 - r0, a + 1);
        r.z = i0;
        r.w = r1 - i0;
    }
    prange[w] = r;
}

#define WCAP 512u            /* slots of a segment the place pass sorts in registers */
#define WSORT_CAP 1024u      /* slots of the segments k_seg_sort_wave sorts in registers */

__device__ inline uint64_t lanemask_lt(int lane) { return lane == 0 ? 0ull : (~0ull >> (STC_WAVE - lane)); }

#define SRC_AMB 0x40000000u  /* key flag: a band combination, a pair only if the host accepted it */
#define PK_REGION 4096u      /* source slots a collecting wave reserves at a time */
#define PK_SLICES 64u        /* psrc slices, each with its own reservation counter */
static_assert(sizeof(((Counters *)nullptr)->pk_used) == PK_SLICES * sizeof(unsigned long long), "one counter per slice");

/* The pair pass: per owner (one wave, owners in leaf order), every
 * combination is decided once; those that are pairs or lie in the guard band
 * take a slot of the owner's segment and their sources (the other point's
 * index in the leaf-ordered arrays or SRC_REV | its reverse-list index, with
 * SRC_AMB on band combinations) are written, unsorted, to a run of psrc that
 * the wave reserved (PK_REGION slots at a time, fewer for a wave with few
 * owners left, or the owner's candidate count when larger, so an owner's
 * sources stay contiguous). psrc is split into PK_SLICES equal slices, wave w
 * reserving in slice w % PK_SLICES from that slice's counter, so that no one
 * address takes every reservation (with one owner per wave, as in a small
 * step, a single counter serialises the whole pass). Per owner: the slot count
 * (pcount, by id) and the run's start and length (pkrun, leaf order); band
 * combinations are listed as (probe, other) for the host; segments longer
 * than the place pass sorts go to `big` or `huge`. A reservation beyond its
 * slice writes nothing and sets pk_over (the host grows psrc and repeats the
 * pass; pk_used tells by how much). */
__global__ __launch_bounds__(256, 8) void k_pair_collect(PairGeo g, uint32_t *__restrict__ pcount,
                                                      uint2 *__restrict__ pkrun, uint32_t *__restrict__ psrc,
                                                      uint64_t pk_cap, uint2 *__restrict__ amb, uint32_t amb_cap,
                                                      uint32_t *__restrict__ big, uint32_t *__restrict__ huge,
                                                      Counters *cnt) {
    const int lane = threadIdx.x & (STC_WAVE - 1);
    /* wave-uniform values go through readfirstlane: scalar registers, scalar loads, scalar loop control */
    const uint64_t wid = (uint32_t)__builtin_amdgcn_readfirstlane((int)((xcd_block_id() * blockDim.x + threadIdx.x) / STC_WAVE));
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / STC_WAVE;
    const uint64_t lt = lanemask_lt(lane);
    unsigned long long cand = 0, total = 0;
    uint64_t rcur = 0, rend = 0; /* the wave's reserved run of psrc */
    bool over = false;
    const uint32_t slice_id = (uint32_t)(wid % PK_SLICES);
    const uint64_t slice = pk_cap / PK_SLICES, slice_base = slice_id * slice;
    for (uint64_t w = wid; w < g.n; w += nw) {
        const uint32_t a = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.lp[w]);
        bool room = true;
        uint32_t *out = psrc;
        auto reserve = [&](uint32_t tot) {
            if (rend - rcur < tot) {
                /* PK_REGION slots, or fewer when the wave's remaining owners need less at this
                 * owner's count each (a wave of a small step has one owner) */
                const uint64_t left = (g.n - 1 - w) / nw + 1;
                const uint64_t want = left * tot < PK_REGION ? left * tot : PK_REGION;
                const uint64_t sz = tot > want ? tot : want;
                unsigned long long b0 = 0;
                if (lane == 0) b0 = atomicAdd(&cnt->pk_used[slice_id], (unsigned long long)sz);
                b0 = __shfl(b0, 0, STC_WAVE);
                rcur = slice_base + b0;
                rend = rcur + sz;
            }
            room = rend <= slice_base + slice;
            over |= !room;
            out = psrc + rcur;
        };
        uint32_t fill = 0;
        auto f = [&](bool v, uint32_t, int dec, uint32_t, float4, uint32_t src) {
            cand += v;
            const bool take = v && dec != 0;
            const uint64_t m = __ballot(take);
            if (take) {
                if (room) __builtin_nontemporal_store(dec == 2 ? (src | SRC_AMB) : src, out + fill + (uint32_t)__popcll(m & lt));
                if (dec == 2) { /* (probe, other): the border point probes in a reverse-list combination */
                    const bool rev = (src & SRC_REV) != 0;
                    const uint32_t o = rev ? g.rev_pid[src & ~SRC_REV] : g.lp[src];
                    const unsigned long long k = atomicAdd(&cnt->ambiguous, 1ull);
                    if (k < amb_cap) amb[k] = rev ? make_uint2(o, a) : make_uint2(a, o);
                }
            }
            fill += (uint32_t)__popcll(m);
        };
        owner_combos<false>((uint32_t)w, a, g, lane, f, reserve);
        if (lane == 0) {
            pcount[a] = fill;
            pkrun[w] = make_uint2((uint32_t)rcur, fill);
            total += fill;
            if (fill > WCAP) {
                if (fill <= WSORT_CAP) big[atomicAdd(&cnt->big, 1u)] = a;
                else huge[atomicAdd(&cnt->huge, 1u)] = a;
            }
        }
        rcur += fill;
    }
    /* totals: summed per workgroup, then one update of a counter of 64 (the host adds them), so
     * that no address takes an update per wave */
    __shared__ unsigned long long bsum[256 / STC_WAVE][2];
    cand = wave_sum_u64(cand);
    if (lane == 0) {
        bsum[threadIdx.x / STC_WAVE][0] = cand;
        bsum[threadIdx.x / STC_WAVE][1] = total;
        if (over) atomicOr(&cnt->pk_over, 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c0 = 0, t0 = 0;
        for (uint32_t k = 0; k < blockDim.x / STC_WAVE; k++) {
            c0 += bsum[k][0];
            t0 += bsum[k][1];
        }
        const uint32_t part = blockIdx.x % PK_SLICES;
        if (c0) atomicAdd(&cnt->cand_part[part], c0);
        if (t0) atomicAdd(&cnt->pairs_part[part], t0);
    }
}

__device__ inline bool sorted_contains(const unsigned long long *a, uint32_t n, unsigned long long v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && a[lo] == v;
}



/* the value that lane (lane ^ J) holds, J a power of two below 64: quad
 * permutes (DPP) for J = 1, 2, the bit-mask swizzle within 32 lanes for
 * J = 4 .. 16, a lane permute for J = 32 */
template <int J>
__device__ inline uint32_t xor_lane32(uint32_t v, int lane) {
    static_assert(J >= 1 && J < STC_WAVE && (J & (J - 1)) == 0, "lane distance");
    if constexpr (J == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false); /* [1,0,3,2] */
    else if constexpr (J == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false); /* [2,3,0,1] */
    else if constexpr (J < 32) return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1F | (J << 10));
    else return (uint32_t)__builtin_amdgcn_ds_bpermute((lane ^ J) << 2, (int)v);
}

template <int J>
__device__ inline unsigned long long xor_lane(unsigned long long v, int lane) {
    const uint32_t lo = xor_lane32<J>((uint32_t)v, lane), hi = xor_lane32<J>((uint32_t)(v >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
}

template <int J>
__device__ inline uint32_t xor_lane(uint32_t v, int lane) {
    return xor_lane32<J>(v, lane);
}

/* one compare-exchange step of a bitonic stage k at lane distance J: each
 * lane keeps the smaller key when its block of k ascends and it is the lower
 * lane of its pair (or neither), the larger otherwise (keys of 32 or 64 bits) */
template <int K, int k, int J, class T>
__device__ inline void bitonic_lane_step(T (&x)[K], int lane) {
#pragma unroll
    for (int r = 0; r < K; r++) {
        const T y = xor_lane<J>(x[r], lane);
        const bool keep_min = (((r * STC_WAVE + lane) & k) == 0) == ((lane & J) == 0);
        const bool le = x[r] <= y;
        x[r] = le == keep_min ? x[r] : y;
    }
}

template <int K, int k, int J, class T>
__device__ inline void bitonic_stage_steps(T (&x)[K], int lane) {
    if constexpr (J >= STC_WAVE) { /* across registers of the lane */
#pragma unroll
        for (int r = 0; r < K; r++) {
            const int r2 = r ^ (J / STC_WAVE);
            if (r2 > r) {
                const bool asc = ((r * STC_WAVE) & k) == 0;
                const T u = x[r], v = x[r2];
                const bool sw = asc ? (u > v) : (u < v);
                x[r] = sw ? v : u;
                x[r2] = sw ? u : v;
            }
        }
    } else {
        bitonic_lane_step<K, k, J>(x, lane);
    }
    if constexpr (J > 1) bitonic_stage_steps<K, k, J / 2>(x, lane);
}

template <int K, int k, class T>
__device__ inline void bitonic_stages(T (&x)[K], int lane) {
    bitonic_stage_steps<K, k, k / 2>(x, lane);
    if constexpr (k < STC_WAVE * K) bitonic_stages<K, 2 * k>(x, lane);
}

/* ascending bitonic sort of 64 * K keys held by a wave, element r * 64 + lane
 * in x[r] of the lane: stages across lanes exchange through DPP, swizzles and
 * permutes, stages across registers swap within the lane */
template <int K, class T>
__device__ inline void reg_bitonic(T (&x)[K], int lane) {
    bitonic_stages<K, 2>(x, lane);
}

/* A collected source as the sort takes it, (partner << 32 | source): the
 * partner is the source's id; a band combination the host did not accept
 * becomes a dead slot (partner DEAD_B), its band flag cleared. */
__device__ inline unsigned long long settle_key(uint32_t src, uint32_t a, const PairGeo &g,
                                                const unsigned long long *acc, uint32_t n_acc) {
    const uint32_t idx = src & ~(SRC_REV | SRC_AMB);
    const uint32_t q = (src & SRC_REV) ? g.rev[idx].pid : g.lp[idx];
    uint32_t b = q;
    if ((src & SRC_AMB) && !sorted_contains(acc, n_acc, ((unsigned long long)a << 32) | q)) b = DEAD_B;
    return ((unsigned long long)b << 32) | (src & ~SRC_AMB);
}

/* Sort an owner's collected sources (srcs[0, slots), slots <= 64 * K) by
 * partner in registers and write its segment: live pairs in partner order,
 * then the dead slots (band combinations the host rejected, repeats of a
 * pair). The sort key is (partner, source); each pair's position is read from
 * its source after the sort. */
template <int K>
__device__ inline void write_sorted_segment(const uint32_t *srcs, uint32_t slots, uint32_t a, uint32_t w,
                                            const PairGeo &g, const unsigned long long *acc, uint32_t n_acc,
                                            PairSlot *__restrict__ seg, int lane) {
    const uint64_t lt = lanemask_lt(lane);
    unsigned long long x[K];
    uint32_t sv[K];
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t i = r * STC_WAVE + lane;
        sv[r] = i < slots ? __builtin_nontemporal_load(srcs + i) : 0u;
    }
#pragma unroll
    for (int r = 0; r < K; r++) x[r] = r * STC_WAVE + lane < slots ? settle_key(sv[r], a, g, acc, n_acc) : ~0ull;
    reg_bitonic<K>(x, lane);
    bool live[K];
    uint32_t nlive = 0;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const unsigned long long up = __shfl_up(x[r], 1, STC_WAVE);
        const unsigned long long last = r > 0 ? __shfl(x[r > 0 ? r - 1 : 0], STC_WAVE - 1, STC_WAVE) : ~0ull;
        const unsigned long long prev = lane > 0 ? up : last;
        const uint32_t i = r * STC_WAVE + lane;
        const uint32_t b = (uint32_t)(x[r] >> 32);
        live[r] = i < slots && b != DEAD_B && (i == 0 || (uint32_t)(prev >> 32) != b);
        nlive += (uint32_t)__popcll(__ballot(live[r]));
    }
    uint32_t lr = 0, dr = 0;
    /* the sources' positions are loaded PB slots at a time, all of a group's loads issued before
     * the first wait (a lane past the end reads the owner's own position) */
    constexpr int PB = K < 4 ? K : (K > 8 ? 1 : 4);
#pragma unroll
    for (int r0 = 0; r0 < K; r0 += PB) {
        double2 pos[PB];
#pragma unroll
        for (int b = 0; b < PB; b++) {
            const uint32_t i = (r0 + b) * STC_WAVE + lane;
            const uint32_t src = (uint32_t)x[r0 + b];
            const double2 *pp = i >= slots ? g.lxy + w
                                           : ((src & SRC_REV) ? (const double2 *)(g.rev + (src & ~SRC_REV)) : g.lxy + src);
            pos[b] = *pp;
        }
#pragma unroll
        for (int b = 0; b < PB; b++) {
            const int r = r0 + b;
            const uint32_t i = r * STC_WAVE + lane;
            const bool valid = i < slots;
            const uint64_t ml = __ballot(live[r]), md = __ballot(valid && !live[r]);
            if (valid) {
                /* the partner and its position (the source's), streamed out past the caches: the
                 * merge reads the slots once, a step's worth of traffic later */
                const uint32_t partner = live[r] ? (uint32_t)(x[r] >> 32) : DEAD_B;
                const float px = (float)pos[b].x, py = (float)pos[b].y;
                const u32x3u v = {partner, __float_as_uint(px), __float_as_uint(py)};
                __builtin_nontemporal_store(v, (u32x3u *)(seg + (live[r] ? lr + (uint32_t)__popcll(ml & lt)
                                                                          : nlive + dr + (uint32_t)__popcll(md & lt))));
            }
            lr += (uint32_t)__popcll(ml);
            dr += (uint32_t)__popcll(md);
        }
    }
}

/* the wave's LDS accesses so far complete before its next ones */
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* The place pass's form of the above, with the wave's LDS buffer lds (64 * K
 * 8-byte entries): each source's partner id and position are read together,
 * straight after the source (one dependent read instead of two), the
 * positions parked in lds by slot as the floats the slot stores, and the sort
 * key is (partner, slot), so that the sorted slots find their positions in
 * lds. With T 32 bits the key is partner << log2(64 K) | slot, the partner
 * field pbits wide and all ones for a dead slot (ids are below 2^pbits - 1),
 * which halves the sort's lane exchanges and selects; for K = 8 it is partner
 * << 8 | slot / 2 (at most 510 slots, so that the padding stays above every
 * key), the partners parked in lds beside the positions telling the two slots
 * of a pair apart (two slots with one key hold one partner: either position
 * serves); with T 64 bits it is partner << 32 | slot. The buffer lds holds
 * 96 K 8-byte entries. The sources arrive as ascending runs (each source
 * leaf's members and the reverse list are in id order), so a segment that is
 * one run is not sorted and one of two runs is merged (the second run
 * reversed behind the padding through lds makes the keys bitonic, and the last
 * stage of the network merges them); more runs take the full sort. */
template <int K, class T>
__device__ inline void place_segment(const uint32_t *srcs, uint32_t slots, uint32_t a, uint32_t w, const PairGeo &g,
                                     const unsigned long long *acc, uint32_t n_acc, PairSlot *__restrict__ seg,
                                     int lane, unsigned long long *lds, uint32_t pbits) {
    const uint64_t lt = lanemask_lt(lane);
    constexpr uint32_t P = K * STC_WAVE;
    constexpr bool PAIRED = sizeof(T) == 4 && K == 8; /* the key holds slot / 2 */
    constexpr int SB = sizeof(T) == 8 ? 32 : (K == 1 ? 6 : K == 2 ? 7 : 8);
    static_assert(K == 1 || K == 2 || K == 4 || K == 8, "segment width");
    const uint32_t pdead = sizeof(T) == 8 ? DEAD_B : (1u << pbits) - 1u;
    T x[K];
    float2 pf[K];
    uint32_t sv[K];
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t i = r * STC_WAVE + lane;
        sv[r] = i < slots ? __builtin_nontemporal_load(srcs + i) : 0u;
    }
    /* the partner's id and position of every slot, GB slots' loads issued before their first wait
     * (a lane past the end reads the owner's own entries): four at a time for the widest segments,
     * whose eight double positions in flight at once would spill */
    constexpr int GB = K > 4 ? 4 : K;
#pragma unroll
    for (int r0 = 0; r0 < K; r0 += GB) {
        uint32_t q[GB];
        double2 pd[GB];
#pragma unroll
        for (int u = 0; u < GB; u++) {
            const int r = r0 + u;
            const uint32_t i = r * STC_WAVE + lane;
            const uint32_t src = sv[r], idx = src & ~(SRC_REV | SRC_AMB);
            const bool rev = i < slots && (src & SRC_REV) != 0;
            const uint32_t li = i < slots ? idx : w;
            q[u] = rev ? g.rev[idx].pid : g.lp[li];
            pd[u] = rev ? *(const double2 *)(g.rev + idx) : g.lxy[li];
        }
#pragma unroll
        for (int u = 0; u < GB; u++) {
            const int r = r0 + u;
            const uint32_t i = r * STC_WAVE + lane;
            uint32_t b = q[u];
            if ((sv[r] & SRC_AMB) && !sorted_contains(acc, n_acc, ((unsigned long long)a << 32) | b)) b = pdead;
            x[r] = i < slots ? (((T)b << SB) | (T)(PAIRED ? i >> 1 : i)) : ~(T)0;
            pf[r] = make_float2((float)pd[u].x, (float)pd[u].y);
        }
    }
    /* PAIRED: each slot's key partner, parked past the first 64 K entries before the two-run
     * reversal below permutes the keys (the reversal and the positions use the first 64 K) */
    uint32_t *lpart = (uint32_t *)(lds + P);
    if constexpr (PAIRED) {
#pragma unroll
        for (int r = 0; r < K; r++)
            if (r * STC_WAVE + lane < slots) lpart[r * STC_WAVE + lane] = (uint32_t)(x[r] >> SB);
    }
    /* descents: places i (0 < i < slots) whose key is below its predecessor's */
    uint32_t d = 0, p = 0;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const T up = __shfl_up(x[r], 1, STC_WAVE);
        const T last = r > 0 ? __shfl(x[r > 0 ? r - 1 : 0], STC_WAVE - 1, STC_WAVE) : (T)0;
        const T prev = lane > 0 ? up : last;
        const uint32_t i = r * STC_WAVE + lane;
        const uint64_t m = __ballot(i < slots && x[r] < prev);
        d += (uint32_t)__popcll(m);
        if (m) p = r * STC_WAVE + 63 - (uint32_t)__clzll((long long)m);
    }
    if (d == 1) {
        /* runs [0, p) and [p, slots): the second reversed at the end, the padding between */
        T *lk = (T *)lds;
#pragma unroll
        for (int r = 0; r < K; r++) {
            const uint32_t i = r * STC_WAVE + lane;
            const uint32_t t = i < p ? i : (i < slots ? P - 1 - (i - p) : p + (i - slots));
            lk[t] = x[r];
        }
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < K; r++) x[r] = lk[r * STC_WAVE + lane];
        wave_lds_sync(); /* the buffer takes the positions next */
    }
    float2 *lpos = (float2 *)lds;
#pragma unroll
    for (int r = 0; r < K; r++)
        if (r * STC_WAVE + lane < slots) lpos[r * STC_WAVE + lane] = pf[r];
    wave_lds_sync();
    if (d == 1) bitonic_stage_steps<K, P, P / 2>(x, lane);
    else if (d > 1) reg_bitonic<K>(x, lane);
    bool live[K];
    uint32_t nlive = 0;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const T up = __shfl_up(x[r], 1, STC_WAVE);
        const T last = r > 0 ? __shfl(x[r > 0 ? r - 1 : 0], STC_WAVE - 1, STC_WAVE) : ~(T)0;
        const T prev = lane > 0 ? up : last;
        const uint32_t i = r * STC_WAVE + lane;
        const uint32_t b = (uint32_t)(x[r] >> SB);
        live[r] = i < slots && b != pdead && (i == 0 || (uint32_t)(prev >> SB) != b);
        nlive += (uint32_t)__popcll(__ballot(live[r]));
    }
    uint32_t lr = 0, dr = 0;
#pragma unroll
    for (int r = 0; r < K; r++) {
        const uint32_t i = r * STC_WAVE + lane;
        const bool valid = i < slots;
        const uint64_t ml = __ballot(live[r]), md = __ballot(valid && !live[r]);
        if (valid) {
            /* the partner and its position, streamed out past the caches: the merge reads the
             * slots once, a step's worth of traffic later */
            uint32_t si = (uint32_t)x[r] & (P - 1);
            if constexpr (PAIRED) {
                si = 2 * ((uint32_t)x[r] & (P / 2 - 1));
                si += lpart[si] != (uint32_t)(x[r] >> SB);
            }
            const float2 pv = lpos[si];
            const uint32_t partner = live[r] ? (uint32_t)(x[r] >> SB) : DEAD_B;
            const u32x3u v = {partner, __float_as_uint(pv.x), __float_as_uint(pv.y)};
            __builtin_nontemporal_store(v, (u32x3u *)(seg + (live[r] ? lr + (uint32_t)__popcll(ml & lt)
                                                                      : nlive + dr + (uint32_t)__popcll(md & lt))));
        }
        lr += (uint32_t)__popcll(ml);
        dr += (uint32_t)__popcll(md);
    }
    wave_lds_sync(); /* the buffer is reused by the wave's next owner */
}


/* Segments of WCAP + 1 .. WSORT_CAP slots: one wave each, sorted in
 * registers and written like the place pass's own segments. */
__global__ __launch_bounds__(256, 5) void k_seg_sort_wave(PairGeo g, const uint32_t *__restrict__ owners,
                                                       uint32_t n_owners, const uint32_t *__restrict__ poff,
                                                       const uint2 *__restrict__ pkrun,
                                                       const uint32_t *__restrict__ psrc,
                                                       const unsigned long long *__restrict__ acc, uint32_t n_acc,
                                                       PairSlot *__restrict__ pn) {
    const int lane = threadIdx.x & (STC_WAVE - 1);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / STC_WAVE;
    const uint32_t nw = (gridDim.x * blockDim.x) / STC_WAVE;
    for (uint32_t oi = wid; oi < n_owners; oi += nw) {
        const uint32_t a = owners[oi];
        const uint32_t base = poff[a], slots = poff[a + 1] - base;
        const uint32_t w = g.leaf_start[g.pt_leaf[a]] + g.pt_rank[a]; /* a's place in leaf order */
        write_sorted_segment<WSORT_CAP / STC_WAVE>(psrc + pkrun[w].x, slots, a, w, g, acc, n_acc, pn + base, lane);
    }
}

/* Segments longer than WSORT_CAP slots: their collected sources as pair
 * slots (key, the partner's position), unsorted, at the segment's place in
 * ptmp for the block sorts (k_seg_sort). One wave per owner. */
__global__ void k_huge_expand(PairGeo g, const uint32_t *__restrict__ owners, uint32_t n_owners,
                              const uint32_t *__restrict__ poff, const uint2 *__restrict__ pkrun,
                              const uint32_t *__restrict__ psrc, const unsigned long long *__restrict__ acc,
                              uint32_t n_acc, PairSlot *__restrict__ ptmp) {
    const int lane = threadIdx.x & (STC_WAVE - 1);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / STC_WAVE;
    const uint32_t nw = (gridDim.x * blockDim.x) / STC_WAVE;
    for (uint32_t oi = wid; oi < n_owners; oi += nw) {
        const uint32_t a = owners[oi];
        const uint32_t base = poff[a], slots = poff[a + 1] - base;
        const uint32_t w = g.leaf_start[g.pt_leaf[a]] + g.pt_rank[a];
        const uint32_t *srcs = psrc + pkrun[w].x;
        for (uint32_t i = lane; i < slots; i += STC_WAVE) {
            const unsigned long long k = settle_key(srcs[i], a, g, acc, n_acc);
            const uint32_t src = (uint32_t)k;
            const double2 pos = (src & SRC_REV) ? *(const double2 *)(g.rev + (src & ~SRC_REV)) : g.lxy[src];
            PairSlot h;
            h.partner = (uint32_t)(k >> 32);
            h.px = (float)pos.x;
            h.py = (float)pos.y;
            ptmp[base + i] = h;
        }
    }
}

/* The place pass: per owner (one wave, owners in leaf order) with at most
 * WCAP slots, its collected keys sorted in registers and written to pn as the
 * owner's segment sorted by partner, its live pairs first and its dead slots
 * behind them, so that pn is sorted by key. pbits: the bits of the object
 * count; a segment sorts 32-bit keys when pbits and its slot index (its slot
 * pair index, for 257 to 510 slots) fit in 32. Five waves per SIMD: at six
 * the 32-bit sort of the widest segments spills. */
__global__ __launch_bounds__(256, 5) void k_pair_place(PairGeo g, const uint32_t *__restrict__ poff,
                                                    const uint2 *__restrict__ pkrun,
                                                    const uint32_t *__restrict__ psrc,
                                                    const unsigned long long *__restrict__ acc, uint32_t n_acc,
                                                    PairSlot *__restrict__ pn, uint32_t pbits) {
    __shared__ unsigned long long wrun[256 / STC_WAVE][WCAP + WCAP / 2]; /* keys or positions, then partners */
    const int lane = threadIdx.x & (STC_WAVE - 1);
    unsigned long long *runbuf = wrun[threadIdx.x / STC_WAVE];
    const uint64_t wid = (uint32_t)__builtin_amdgcn_readfirstlane((int)((xcd_block_id() * blockDim.x + threadIdx.x) / STC_WAVE));
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / STC_WAVE;
    for (uint64_t w = wid; w < g.n; w += nw) {
        const uint2 run = pkrun[w];
        const uint32_t slots = (uint32_t)__builtin_amdgcn_readfirstlane((int)run.y);
        if (slots == 0 || slots > WCAP) continue;
        const uint32_t a = (uint32_t)__builtin_amdgcn_readfirstlane((int)g.lp[w]);
        const uint32_t *keys = psrc + (uint32_t)__builtin_amdgcn_readfirstlane((int)run.x);
        PairSlot *seg = pn + poff[a];
#define PLACE(KK, BITS)                                                                                          \
    do {                                                                                                         \
        if (pbits + (BITS) <= 32u) place_segment<KK, uint32_t>(keys, slots, a, (uint32_t)w, g, acc, n_acc, seg, lane, runbuf, pbits); \
        else place_segment<KK, unsigned long long>(keys, slots, a, (uint32_t)w, g, acc, n_acc, seg, lane, runbuf, pbits); \
    } while (0)
        if (slots <= STC_WAVE) PLACE(1, 6);
        else if (slots <= 2 * STC_WAVE) PLACE(2, 7);
        else if (slots <= 4 * STC_WAVE) PLACE(4, 8);
        else if (slots <= 2 * 255) PLACE(8, 8); /* slot pairs below 255 keep the padding above every key */
        else place_segment<8, unsigned long long>(keys, slots, a, (uint32_t)w, g, acc, n_acc, seg, lane, runbuf, pbits);
#undef PLACE
    }
}

/* replayed pairs (sorted by key): the segment length of each owner */
__global__ void k_owner_hist(const PairHit *__restrict__ pn, uint32_t np, uint32_t *__restrict__ pcount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    atomicAdd(&pcount[(uint32_t)(pn[i].key >> 32)], 1u);
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

/* the units in pair order as PairHits; a pair named twice sets bad_points */
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

struct MemtableDev {
    uint64_t *keys;        /* the reference's 128-bit key as (lo, hi) words:
                              hi = pid << 26 | target, lo = duration << 32 | end offset */
    float4 *boxes;
    uint64_t cap;
    uint32_t *same_pid_count;
    uint4 *o_boxs;         /* quantized per-object MBR: low0, low1, high0, high1 */
    uint16_t *sids;
};

struct MergeArgs {
    uint32_t now, min_meet, max_meet, end_time_min;
    MapBox m;
    uint32_t prev;         /* the step that built the previous array: its live meetings' end time */
};

/* a meeting that expires at `now` is recorded when it lasted long enough and
 * its MBR lies in the map (geometry.cu:665-670) */
__device__ inline bool meeting_emits(const Bucket &b, const MergeArgs &a) {
    float lo0 = o2f(b.mbr[0]), lo1 = o2f(b.mbr[1]), hi0 = o2f(b.mbr[2]), hi1 = o2f(b.mbr[3]);
    return a.now - b.start >= a.min_meet + 1 && (double)lo0 >= a.m.lo0 && (double)hi0 <= a.m.hi0 &&
           (double)lo1 >= a.m.lo1 && (double)hi1 <= a.m.hi1;
}

/* the two records of an expired meeting at k, k + 1 (geometry.cu:717-734);
 * end: the step that last saw it */
__device__ inline void write_meeting_records(const Bucket &b, uint32_t end, unsigned long long k, uint32_t end_time_min,
                                             const MemtableDev &mt) {
    const float4 box = make_float4(o2f(b.mbr[0]), o2f(b.mbr[1]), o2f(b.mbr[2]), o2f(b.mbr[3]));
    const uint32_t duration = end - b.start;
    const uint32_t pid = (uint32_t)(b.key >> 32), target = (uint32_t)b.key;
    for (int sdir = 0; sdir < 2; sdir++) {
        const uint32_t a = sdir ? target : pid, c = sdir ? pid : target;
        if (k + sdir < mt.cap) {
            mt.keys[2 * (k + sdir)] = ((uint64_t)duration << 32) | (uint32_t)(end - end_time_min);
            mt.keys[2 * (k + sdir) + 1] = ((uint64_t)a << 26) | (uint64_t)c;
            mt.boxes[k + sdir] = box;
        }
    }
}

/* per-object memtable state of the records [from, to) (geometry.cu:705-733):
 * record counts, quantised MBRs, and oversize meetings (one per record pair)
 * marking both objects. The objects with records are counted by the flush.
 * A box bound is updated only when the record moves it past the value read
 * first: within a memtable the bounds only shrink (low) or grow (high), so a
 * value read before another thread's update still proves an update redundant. */
__global__ void k_records_account(const uint64_t *__restrict__ keys, const float4 *__restrict__ boxes, uint64_t from,
                                  uint64_t to, uint32_t max_meet, MemtableDev mt, Counters *cnt) {
    const uint64_t i = from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (STC_WAVE - 1);
    bool oversize = false;
    if (i < to) {
        const uint64_t lo = keys[2 * i], hi = keys[2 * i + 1];
        const uint32_t a = (uint32_t)(hi >> 26) & 0x3FFFFFFu, c = (uint32_t)hi & 0x3FFFFFFu;
        const uint32_t duration = (uint32_t)(lo >> 32);
        if (((i - from) & 1) == 0 && duration > max_meet) {
            oversize = true;
            mt.sids[a] = 1;
            mt.sids[c] = 1;
        }
        const float4 b = boxes[i];
        atomicAdd(&mt.same_pid_count[a], 1u);
        const uint4 cur = mt.o_boxs[a];
        const uint32_t q0 = f2q(b.x), q1 = f2q(b.y), q2 = f2q(b.z), q3 = f2q(b.w);
        if (q0 < cur.x) atomicMin(&mt.o_boxs[a].x, q0);
        if (q2 > cur.z) atomicMax(&mt.o_boxs[a].z, q2);
        if (q1 < cur.y) atomicMin(&mt.o_boxs[a].y, q1);
        if (q3 > cur.w) atomicMax(&mt.o_boxs[a].w, q3);
    }
    /* one counter update per wave: a counter updated by every thread serialises on its address */
    const uint64_t om = __ballot(oversize);
    if (lane == 0 && om) atomicAdd(&cnt->oversize_oid_count, (uint32_t)__popcll(om));
}

/* ascending bitonic sort of s[0, P) (P a power of two) by a block */
template <int THREADS>
__device__ inline void block_bitonic(unsigned long long *s, uint32_t P) {
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += THREADS) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long x = s[i], y = s[ixj];
                    if ((x > y) == ((i & k) == 0)) {
                        s[i] = y;
                        s[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
}

/* rank of the set flags across the block, in thread order; *total = all set */
template <int THREADS>
__device__ inline uint32_t block_rank(bool f, uint32_t *wsum, uint32_t *total) {
    const int lane = threadIdx.x & (STC_WAVE - 1), wv = threadIdx.x / STC_WAVE;
    const uint64_t m = __ballot(f);
    const uint32_t r = (uint32_t)__popcll(m & lanemask_lt(lane));
    if (lane == 0) wsum[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int w = 0; w < THREADS / STC_WAVE; w++) {
        const uint32_t x = wsum[w];
        if (w < wv) off += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + r;
}

/* Segments longer than WSORT_CAP, expanded unsorted into ptmp: the same
 * ordering by a block, with the sort keys in LDS (LCAP entries) or, with
 * GLOBAL, in a per-block slice of gcap entries of global scratch. Longer
 * segments go to `spill` (or set table_fail when spill is null). */
template <int THREADS, int LCAP, bool GLOBAL>
__global__ __launch_bounds__(THREADS) void k_seg_sort(const uint32_t *__restrict__ owners, uint32_t n_owners,
                                                      const uint32_t *__restrict__ poff, const PairSlot *__restrict__ ptmp,
                                                      PairSlot *__restrict__ pn, uint32_t *__restrict__ spill,
                                                      uint32_t *spill_count, unsigned long long *gsrt, uint32_t gcap,
                                                      Counters *cnt) {
    __shared__ unsigned long long lsrt[GLOBAL ? 1 : LCAP];
    __shared__ uint32_t wsum[THREADS / STC_WAVE];
    unsigned long long *keys = GLOBAL ? gsrt + (size_t)blockIdx.x * gcap : lsrt;
    const uint32_t cap = GLOBAL ? gcap : (uint32_t)LCAP;
    const int tid = threadIdx.x;
    for (uint32_t oi = blockIdx.x; oi < n_owners; oi += gridDim.x) {
        const uint32_t a = owners[oi];
        const uint32_t base = poff[a], slots = poff[a + 1] - base;
        if (slots > cap) {
            if (tid == 0) {
                if (spill) spill[atomicAdd(spill_count, 1u)] = a;
                else atomicOr(&cnt->table_fail, 1u);
            }
            continue;
        }
        uint32_t P = 1;
        while (P < slots) P <<= 1;
        for (uint32_t i = tid; i < P; i += THREADS)
            keys[i] = i < slots ? (((unsigned long long)ptmp[base + i].partner << 32) | i) : ~0ull;
        __syncthreads();
        block_bitonic<THREADS>(keys, P);
        uint32_t nlive = 0;
        for (uint32_t k0 = 0; k0 < slots; k0 += THREADS) {
            const uint32_t k = k0 + tid;
            bool live = false;
            if (k < slots) {
                const uint32_t b = (uint32_t)(keys[k] >> 32);
                live = b != DEAD_B && (k == 0 || (uint32_t)(keys[k - 1] >> 32) != b);
            }
            uint32_t total;
            (void)block_rank<THREADS>(live, wsum, &total);
            nlive += total;
        }
        uint32_t lr = 0, dr = 0;
        for (uint32_t k0 = 0; k0 < slots; k0 += THREADS) {
            const uint32_t k = k0 + tid;
            bool live = false;
            if (k < slots) {
                const uint32_t b = (uint32_t)(keys[k] >> 32);
                live = b != DEAD_B && (k == 0 || (uint32_t)(keys[k - 1] >> 32) != b);
            }
            uint32_t tl, td;
            const uint32_t rl = block_rank<THREADS>(live, wsum, &tl);
            const uint32_t rd = block_rank<THREADS>(k < slots && !live, wsum, &td);
            if (k < slots) {
                PairSlot e = ptmp[base + (uint32_t)keys[k]];
                uint32_t dest;
                if (live) {
                    dest = lr + rl;
                } else {
                    dest = nlive + dr + rd;
                    e.partner = DEAD_B;
                }
                pn[base + dest] = e;
            }
            lr += tl;
            dr += td;
        }
        __syncthreads();
    }
}

#define MTILE 128u   /* new pairs per merge tile (one wave) */
#define OCAP 256u    /* previous meetings of a tile held whole in the wave's LDS */
#define MSTAGE 32u   /* expired meetings a merge wave stages before reserving their records */

/* write the wave's staged meetings' records: one counter reservation per batch */
__device__ inline void flush_stage(const Bucket *stage, uint32_t staged, const MergeArgs &ma, const MemtableDev &mt,
                                   Counters *cnt, int lane) {
    unsigned long long kb = 0;
    if (lane == 0) kb = atomicAdd(&cnt->kv_count, 2ull * staged);
    kb = __shfl(kb, 0, STC_WAVE);
    for (uint32_t e = lane; e < staged; e += STC_WAVE)
        write_meeting_records(stage[e], ma.prev, kb + 2ull * e, ma.end_time_min, mt);
    __builtin_amdgcn_wave_barrier();
}

/* the owner of each merge tile's first slot (slot i belongs to the owner a
 * with poff[a] <= i < poff[a + 1]; each tile's first slot lies in exactly one
 * non-empty segment) */
__global__ void k_tile_owner(const uint32_t *__restrict__ poff, uint32_t n, uint32_t *__restrict__ towner) {
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const uint64_t p0 = poff[a], p1 = poff[a + 1];
    for (uint64_t t = (p0 + MTILE - 1) / MTILE; t * MTILE < p1; t++) towner[t] = a;
}

/* Start of each merge tile's range in the previous meeting array: the number
 * of previous keys below the tile's first new key (found in that key's owner
 * segment). tb[n_tiles] = n_o. The search starts where the key's place among
 * the owner's new slots, scaled to the owner's previous segment, puts it (the
 * two segments mostly hold the same partners) and widens by doubling steps
 * until it brackets the key, so a few dependent reads usually suffice. */
template <class S>
__global__ void k_tile_bounds(const S *__restrict__ pn, uint64_t n_p, const Bucket *__restrict__ mold,
                              const uint32_t *__restrict__ ooff, const uint32_t *__restrict__ poff, uint64_t n_o,
                              uint32_t n_tiles, const uint32_t *__restrict__ towner, uint32_t *__restrict__ tb) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t > n_tiles) return;
    if (t == 0 || t == n_tiles || (uint64_t)t * MTILE >= n_p) {
        tb[t] = t == 0 ? 0u : (uint32_t)n_o;
        return;
    }
    const uint64_t first = (uint64_t)t * MTILE;
    uint32_t owner = 0;
    if constexpr (std::is_same<S, PairSlot>::value) owner = towner[t];
    const unsigned long long key = slot_key(pn[first], owner);
    const uint32_t a = (uint32_t)(key >> 32);
    uint32_t lo = ooff[a], hi = ooff[a + 1];
    const uint32_t p0 = poff[a], np_a = poff[a + 1] - p0;
    if (hi > lo && np_a) {
        /* bracket [lo, hi) around the scaled guess g: keys below lo are < key, keys from hi on are >= key */
        const uint32_t g = lo + (uint32_t)(((uint64_t)(first - p0) * (hi - lo)) / np_a);
        if (g < hi && mold[g].key < key) {
            uint32_t step = 1, l = g + 1;
            while (l < hi) {
                const uint32_t probe = min(l + step - 1, hi - 1);
                if (mold[probe].key < key) {
                    l = probe + 1;
                    step <<= 1;
                } else {
                    hi = probe;
                    break;
                }
            }
            lo = l;
        } else {
            uint32_t step = 1, r = g; /* mold[r] >= key or r == hi */
            while (r > lo) {
                const uint32_t probe = r - min(step, r - lo);
                if (mold[probe].key < key) {
                    lo = probe + 1;
                    break;
                }
                r = probe;
                step <<= 1;
            }
            hi = r;
        }
    }
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (mold[mid].key < key) lo = mid + 1;
        else hi = mid;
    }
    tb[t] = lo;
}

__device__ inline uint32_t lower_bound_u64(const unsigned long long *s, uint32_t n, unsigned long long v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ inline uint32_t lower_bound_bucket(const Bucket *s, uint32_t n, unsigned long long v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid].key < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

/* Merge of the step's pairs (pn, sorted by key, dead slots behind each owner's
 * live pairs) with the previous step's meetings (mold, sorted the same way),
 * tile by tile, one wave per tile: the tile's range of previous meetings (tb)
 * is staged whole in the wave's LDS, then each new slot is read once and
 * joined with it. A live pair continues the previous meeting of its key
 * (start kept, MBR widened) or opens one (insert/extend, geometry.cu:469-494);
 * a dead slot becomes a tombstone; a previous meeting whose key has no pair
 * expires and, when it lasted long enough inside the map, becomes two records
 * (identify, geometry.cu:595-737). mnew[i] is built from pn[i]. A range
 * longer than OCAP is searched in global memory. */
#define OPF (OCAP / STC_WAVE)   /* previous meetings per lane loaded at a tile's start */
#define SPF (MTILE / STC_WAVE)  /* pair slots per lane loaded at a tile's start */

template <class S>
__global__ __launch_bounds__(256) void k_meet_merge(const S *__restrict__ pn, uint64_t n_p,
                                                    const Bucket *__restrict__ mold, const uint32_t *__restrict__ tb,
                                                    uint32_t n_tiles, Bucket *__restrict__ mnew, MergeArgs ma,
                                                    MemtableDev mt, const float2 *__restrict__ posf,
                                                    const uint32_t *__restrict__ poff, const uint32_t *__restrict__ towner,
                                                    uint32_t n_obj, Counters *cnt) {
    constexpr bool SLOTS = std::is_same<S, PairSlot>::value; /* owners from poff, not from the slots */
    __shared__ unsigned long long wpk[256 / STC_WAVE][MTILE];
    __shared__ Bucket wob[256 / STC_WAVE][OCAP];
    __shared__ Bucket wstage[256 / STC_WAVE][MSTAGE];
    const int lane = threadIdx.x & (STC_WAVE - 1), wv = threadIdx.x / STC_WAVE;
    unsigned long long *pk = wpk[wv];
    Bucket *ob_l = wob[wv];
    Bucket *stage = wstage[wv];
    const uint64_t lt = lanemask_lt(lane);
    const uint32_t wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) / STC_WAVE);
    const uint32_t nw = (gridDim.x * blockDim.x) / STC_WAVE;
    const uint32_t n_o = tb[n_tiles];
    const uint64_t plast = n_p ? n_p - 1 : 0;
    uint32_t staged = 0;
    unsigned long long n_live = 0, n_exp = 0, n_fresh = 0, n_cont = 0;
    /* a tile's first owner a0 and the next tile's a1 (n_obj - 1 after the last tile): the owners
     * of the tile's slots lie in [a0, a1]; a step without pairs leaves tile 0 unset, hence the clamps */
    auto tile_owner = [&](uint32_t t) { return t < n_tiles ? min(towner[t], n_obj - 1) : n_obj - 1; };
    uint32_t o0 = 0, o1 = 0, on0 = 0, on1 = 0, a0 = 0, a1 = 0, an0 = 0, an1 = 0;
    if (wid < n_tiles) {
        o0 = tb[wid];
        o1 = tb[wid + 1];
        if constexpr (SLOTS) {
            a0 = tile_owner(wid);
            a1 = max(a0, tile_owner(wid + 1));
        }
    }
    for (uint32_t t = wid; t < n_tiles; t += nw, o0 = on0, o1 = on1, a0 = an0, a1 = an1) {
        { /* the next tile's range (and owners), read ahead */
            const uint32_t tn = t + nw < n_tiles ? t + nw : t;
            on0 = tb[tn];
            on1 = tb[tn + 1];
            if constexpr (SLOTS) {
                an0 = tile_owner(tn);
                an1 = max(an0, tile_owner(tn + 1));
            }
        }
        const uint64_t i0 = (uint64_t)t * MTILE;
        const uint32_t np = i0 < n_p ? (uint32_t)min<uint64_t>(MTILE, n_p - i0) : 0u;
        /* all loads of the tile are issued before the first wait; indices are clamped into the
         * arrays (each holds at least one entry) and lanes past the tile's ends ignore theirs */
        S ps[SPF];
        Bucket po[OPF];
#pragma unroll
        for (int k = 0; k < SPF; k++) ps[k] = pn[min<uint64_t>(i0 + lane + k * STC_WAVE, plast)];
        const uint32_t olast = o1 > o0 ? o1 - 1 : (n_o ? min(o0, n_o - 1) : 0u);
#pragma unroll
        for (int k = 0; k < OPF; k++) po[k] = mold[min(o0 + lane + k * STC_WAVE, olast)];
        /* the starts of the 64 owners after a0 (a0 owns the tile's first slot), read with the rest */
        uint32_t bst = 0;
        if constexpr (SLOTS) {
            const uint64_t ai = (uint64_t)a0 + 1 + (uint32_t)lane;
            bst = poff[ai < n_obj ? ai : n_obj];
        }
        if (o1 < o0) { /* the pairs were not sorted by key: report it, use nothing */
            if (lane == 0) atomicOr(&cnt->table_fail, 2u);
            continue;
        }
        const uint32_t no = o1 - o0;
        const bool oin = no <= OCAP;
        /* staged unconditionally (entries past the range are never searched), so the loads stay
         * ahead of every branch */
#pragma unroll
        for (int k = 0; k < OPF; k++) ob_l[lane + k * STC_WAVE] = po[k];
        __builtin_amdgcn_wave_barrier();
        /* each slot's owner: a0 plus the number of the owners a0 + 1 .. a1 that start at or before
         * the slot (their starts ascend; an empty segment's start equals the next one's, so both
         * count and the later owner stands). The starts read with the tile's loads cover the next
         * 64 owners; nb of them lie inside the tile (a prefix of the lanes). A dense tile has one to
         * three, counted directly; up to 63 are searched by binary lifting over the lanes; a tile
         * with more (a sparse stream's owners, most without pairs) lifts over poff itself. Lanes
         * past the tile's end take its last owner. */
        uint32_t own[SPF];
#pragma unroll
        for (int k = 0; k < SPF; k++) own[k] = a0;
        if constexpr (SLOTS) {
            if (np) {
                const uint32_t range = a1 - a0;
                uint32_t slot[SPF], c[SPF];
#pragma unroll
                for (int k = 0; k < SPF; k++) {
                    slot[k] = (uint32_t)min<uint64_t>(i0 + lane + k * STC_WAVE, i0 + np - 1);
                    c[k] = 0;
                }
                const uint32_t nb =
                    (uint32_t)__popcll(__ballot((uint32_t)lane < range && (uint64_t)bst < i0 + np));
                if (nb <= 8) {
                    for (uint32_t j = 0; j < nb; j++) {
                        const uint32_t bj = (uint32_t)__builtin_amdgcn_readlane((int)bst, (int)j);
#pragma unroll
                        for (int k = 0; k < SPF; k++) c[k] += slot[k] >= bj;
                    }
                } else if (nb < STC_WAVE) {
                    for (uint32_t st = STC_WAVE / 2; st; st >>= 1) {
                        if (st > nb) continue; /* wave-uniform */
#pragma unroll
                        for (int k = 0; k < SPF; k++) {
                            const uint32_t p = c[k] + st;
                            const uint32_t v = (uint32_t)__shfl((int)bst, (int)min(p - 1, (uint32_t)STC_WAVE - 1), STC_WAVE);
                            if (p <= nb && v <= slot[k]) c[k] = p;
                        }
                    }
                } else {
                    uint32_t st = 1u << (31 - __builtin_clz(range));
                    for (; st; st >>= 1) {
#pragma unroll
                        for (int k = 0; k < SPF; k++) {
                            const uint32_t p = c[k] + st;
                            if (p <= range && poff[a0 + p] <= slot[k]) c[k] = p;
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < SPF; k++) own[k] = a0 + c[k];
            }
        }
        /* the join, instantiated for previous meetings in LDS and, for a range longer than
         * OCAP, in global memory (so the LDS searches wait on nothing but LDS) */
        auto join = [&](const Bucket *obk) {
            /* every lane joins its slots, in range or not (lanes past np store nothing), so that no
             * load of the tile is moved behind a branch */
#pragma unroll
            for (int k = 0; k < SPF; k++) {
                const uint32_t i = lane + k * STC_WAVE;
                const bool valid = i < np;
                const S e = ps[k];
                const unsigned long long key = slot_key(e, own[k]);
                pk[i] = key;
                const bool dead = (uint32_t)key == DEAD_B;
                const uint32_t j = lower_bound_bucket(obk, no, key);
                const uint4 q = slot_mbr(e, posf, own[k]);
                const uint32_t q0 = q.x, q1 = q.y, q2 = q.z, q3 = q.w;
                const Bucket ob = obk[j < no ? j : (no ? no - 1 : 0u)];
                const bool cont = j < no && ob.key == key;
                Bucket nb;
                nb.key = key;
                nb.start = dead ? 0u : (cont ? ob.start : ma.now);
                nb.mbr[0] = dead ? 0u : (cont ? min(ob.mbr[0], q0) : q0);
                nb.mbr[1] = dead ? 0u : (cont ? min(ob.mbr[1], q1) : q1);
                nb.mbr[2] = dead ? 0u : (cont ? max(ob.mbr[2], q2) : q2);
                nb.mbr[3] = dead ? 0u : (cont ? max(ob.mbr[3], q3) : q3);
                if (valid) {
                    if (!dead) {
                        n_live++;
                        if (cont) n_cont++;
                        else n_fresh++;
                    }
                    /* streamed out past the caches: the array is read again only by the next step */
                    uint32_t *d = (uint32_t *)(mnew + i0 + i);
                    const u32x4u lo = {(uint32_t)nb.key, (uint32_t)(nb.key >> 32), nb.start, nb.mbr[0]};
                    const u32x3u hi = {nb.mbr[1], nb.mbr[2], nb.mbr[3]};
                    __builtin_nontemporal_store(lo, (u32x4u *)d);
                    __builtin_nontemporal_store(hi, (u32x3u *)(d + 4));
                }
            }
            __builtin_amdgcn_wave_barrier();
            for (uint32_t j0 = 0; j0 < no; j0 += STC_WAVE) {
                const uint32_t j = j0 + lane;
                bool emit = false;
                Bucket ob;
                if (j < no) {
                    const unsigned long long key = obk[j].key;
                    if ((uint32_t)key != DEAD_B) {
                        const uint32_t k = lower_bound_u64(pk, np, key);
                        if (!(k < np && pk[k] == key)) {
                            n_exp++;
                            ob = obk[j];
                            emit = meeting_emits(ob, ma);
                        }
                    }
                }
                const uint64_t em = __ballot(emit);
                if (em) {
                    const uint32_t total = (uint32_t)__popcll(em);
                    if (staged + total > MSTAGE) {
                        __builtin_amdgcn_wave_barrier();
                        flush_stage(stage, staged, ma, mt, cnt, lane);
                        staged = 0;
                    }
                    const uint32_t put = staged + (uint32_t)__popcll(em & lt);
                    if (emit && put < MSTAGE) stage[put] = ob;
                    /* a batch beyond the stage (more than MSTAGE at once) is written directly */
                    if (total > MSTAGE) {
                        unsigned long long kb = 0;
                        if (lane == 0) kb = atomicAdd(&cnt->kv_count, 2ull * (total - MSTAGE));
                        kb = __shfl(kb, 0, STC_WAVE);
                        if (emit && put >= MSTAGE) write_meeting_records(ob, ma.prev, kb + 2ull * (put - MSTAGE), ma.end_time_min, mt);
                        staged = MSTAGE;
                    } else {
                        staged += total;
                    }
                }
            }
        };
        if (oin) join(ob_l);
        else join(mold + o0);
        __builtin_amdgcn_wave_barrier(); /* the wave's LDS is reused by its next tile */
    }
    if (staged) {
        __builtin_amdgcn_wave_barrier();
        flush_stage(stage, staged, ma, mt, cnt, lane);
    }
    /* one update of each counter per workgroup: updates of one address serialise */
    __shared__ unsigned long long bsum[256 / STC_WAVE][4];
    n_live = wave_sum_u64(n_live);
    n_exp = wave_sum_u64(n_exp);
    n_fresh = wave_sum_u64(n_fresh);
    n_cont = wave_sum_u64(n_cont);
    if (lane == 0) {
        bsum[wv][0] = n_live;
        bsum[wv][1] = n_exp;
        bsum[wv][2] = n_fresh;
        bsum[wv][3] = n_cont;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long v = 0;
        for (int k = 0; k < 256 / STC_WAVE; k++) v += bsum[k][threadIdx.x];
        unsigned long long *dst[4] = {&cnt->live, &cnt->expired, &cnt->new_keys, &cnt->continuing};
        if (v) atomicAdd(dst[threadIdx.x], v);
    }
}

#define SEG_CAP 4096u  /* LDS sort keys of the four-wave segment sort */

__global__ void k_fill_obox(uint4 *a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_uint4(100000000u, 100000000u, 0u, 0u); /* cuda_init_o_boxs */
}

/* ------------------------------------------------------------------ */
/* host context                                                        */
/* ------------------------------------------------------------------ */

/* ------------------------------------------------------------------ */
/* the allocation-free live search: one scratch array per device         */
/* ------------------------------------------------------------------ */
/* Layout: [0, 256) hit counter; then NA_MAX_PIDS search keys
 * (pid << 32 | index in the call's list); then NA_MAX_HITS staged hits. */
#define NA_MAX_PIDS 16384u
#define NA_MAX_HITS 262144u
#define NA_KEYS_AT 256u
#define NA_HITS_AT (NA_KEYS_AT + NA_MAX_PIDS * 8u)
#define NA_SCRATCH_BYTES (NA_HITS_AT + NA_MAX_HITS * (uint32_t)sizeof(stc_hit))
static __device__ __align__(256) uint8_t scratch[NA_SCRATCH_BYTES];

/* scratch's device address on each device, resolved once per device */
static std::mutex g_na_mu;
static uint8_t *g_na_ptr[64];

static int noalloc_setup(int device) {
    if (device < 0 || device >= 64) return STC_E_ARG;
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (g_na_ptr[device]) return STC_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return stc_hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_na_ptr[device] = (uint8_t *)p;
    return STC_OK;
}

/* the search ids, sorted as (pid, index) in LDS by one workgroup; zeroes the
 * hit counter (the array keeps the previous call's contents) */
__global__ __launch_bounds__(1024) void k_na_prep(uint8_t *sc, uint32_t n) {
    __shared__ unsigned long long k[NA_MAX_PIDS];
    const uint32_t *pids = (const uint32_t *)(sc + NA_HITS_AT); /* staged by the host copy */
    uint32_t P = 1;
    while (P < n) P <<= 1;
    for (uint32_t i = threadIdx.x; i < P; i += blockDim.x)
        k[i] = i < n ? (((unsigned long long)pids[i] << 32) | i) : ~0ull;
    if (threadIdx.x == 0) *(unsigned long long *)sc = 0ull;
    __syncthreads();
    for (uint32_t size = 2; size <= P; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const unsigned long long a = k[i], b = k[j];
                    if ((a > b) == up) { k[i] = b; k[j] = a; }
                }
            }
            __syncthreads();
        }
    unsigned long long *out = (unsigned long long *)(sc + NA_KEYS_AT);
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) out[i] = k[i];
}

__device__ inline void na_emit(uint8_t *sc, uint32_t qi, uint32_t pid, uint32_t target, const Bucket &b, uint32_t end) {
    const unsigned long long at = atomicAdd((unsigned long long *)sc, 1ull);
    if (at >= NA_MAX_HITS) return;
    stc_hit h;
    h.query = qi;
    h.pid = pid;
    h.target = target;
    h.start = b.start;
    h.end = end;
    h.source = 0;
    for (int i = 0; i < 4; i++) h.mbr[i] = o2f(b.mbr[i]);
    ((stc_hit *)(sc + NA_HITS_AT))[at] = h;
}

/* every search id equal to `pid`: the hits seen from pid */
__device__ inline void na_match(uint8_t *sc, uint32_t n, uint32_t pid, uint32_t target, const Bucket &b, uint32_t end) {
    const unsigned long long *keys = (const unsigned long long *)(sc + NA_KEYS_AT);
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint32_t)(keys[mid] >> 32) < pid) lo = mid + 1; else hi = mid;
    }
    for (uint32_t i = lo; i < n && (uint32_t)(keys[i] >> 32) == pid; i++) na_emit(sc, (uint32_t)keys[i], pid, target, b, end);
}

/* live meetings of the last step established by t_est (start <= t_est) */
__global__ void k_na_live(const Bucket *__restrict__ t, uint64_t nb, uint32_t last_t, uint32_t t_est, uint8_t *sc,
                          uint32_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
        const Bucket b = t[i];
        if (!bucket_live(b) || b.start > t_est) continue;
        const uint32_t a = (uint32_t)(b.key >> 32), c = (uint32_t)b.key;
        na_match(sc, n, a, c, b, last_t);
        na_match(sc, n, c, a, b, last_t);
    }
}

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
    std::mutex mu; /* an asynchronous flush acquires from its own thread */
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
        std::lock_guard<std::mutex> lk(mu);
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
        std::lock_guard<std::mutex> lk(mu);
        if (b.p) retiring.push_back(b);
    }
    void clear() {
        std::lock_guard<std::mutex> lk(mu);
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
    DBuf<double> dlon, ld;
    DBuf<float2> posf;                   /* positions by id as floats: the owners' side of the pair slots */
    DBuf<double2> lxy;
    DBuf<Node> nodes;
    DBuf<FrontierItem> frontier_a, frontier_b;
    DBuf<uint32_t> leaf_count, leaf_start, leaf_node;
    DBuf<uint32_t> pt_leaf, pt_lv, pt_rank, lp, lkeys;
    DBuf<uint32_t> lleaf;                /* leaf of each point, in leaf order */
    DBuf<uint2> lunits;                  /* first cross unit and count of each point, in leaf order */
    DBuf<uint2> urange;                  /* per cross unit: member range above its point */
    DBuf<uint4> prange;                  /* per point in leaf order: own-leaf and reverse-list ranges */
    DBuf<uint32_t> ucount, uoffset;
    DBuf<uint2> units;
    DBuf<uint2> amb;                     /* band combinations (probe, other) for the host */
    DBuf<double4> amb_xy;                /* their positions (probe x, y, other x, y) */
    /* per leaf, by pid: the border points with a cross unit to it */
    DBuf<uint32_t> rev_leaf, rev_pid, rev_leaf_alt, rev_pid_alt, rev_start;
    DBuf<RevPt> rev;
    DBuf<double> rev_dl;
    /* the step's pairs by owner segment, and the live meetings: two arrays of
     * segments (offsets [n + 1], live buckets per segment [n]) used in turn */
    DBuf<uint32_t> pcount;
    DBuf<PairSlot> ptmp, pnew;           /* the long segments' pair slots as found, and all slots sorted by key */
    DBuf<PairHit> pbox;                  /* a replayed step's pairs (key and MBR), sorted by key */
    bool pairs_boxed = false;            /* the last step's pairs are in pbox (replayed units), not pnew */
    DBuf<uint32_t> psrc;                 /* collected pair sources, a run per owner */
    DBuf<uint2> pkrun;                   /* per owner in leaf order: its run of psrc (start, length) */
    DBuf<Bucket> meet[2];
    DBuf<uint32_t> moff[2];
    DBuf<uint32_t> tbounds;              /* merge tiles' ranges in the previous meetings */
    DBuf<uint32_t> towner;               /* merge tiles' first owners (the slots carry no owner) */
    int cur = 0;                         /* meet[cur] holds the meetings of the last step */
    uint64_t mlen = 0;                   /* buckets of meet[cur] (live and tombstones) */
    uint64_t plen = 0;                   /* entries of pnew (the last step's pair slots) */
    DBuf<unsigned long long> acc;        /* band pairs the host accepted, sorted */
    DBuf<uint32_t> spill_big, spill_huge, spill_giant;
    DBuf<unsigned long long> gsrt;       /* global sort keys of very large segments */
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
    uint32_t oversize_oid_count = 0;
    uint32_t end_time_min = 0;
    uint64_t kv_count = 0;
    uint64_t last_live = 0;
    uint64_t last_expired = 0;
    uint32_t last_t = 0;
    bool stepped = false;
    uint32_t last_nodes = 0, last_leaves = 0;
    DBuf<stc_meeting_unit> units_in;  /* replayed meeting units and their pair keys */
    DBuf<uint64_t> ukeys, ukeys_alt;
    DBuf<uint32_t> uidx, uidx_alt, utemp;
    std::vector<RunHost> runs;
    /* asynchronous flush (cfg.async_flush): a step's flush runs on fs from a host thread until the
     * next step's merge, or any call that reads runs, joins it */
    hipStream_t fs = nullptr;
    hipEvent_t fev[3] = {};              /* go (recorded on s), start and end (on fs) */
    std::thread fthr;
    bool frun = false;                   /* a flush thread started and not yet joined */
    int frc = 0;                         /* its status and message */
    std::string ferr;
    FlushOut fout;
    bool flush_fresh = false;            /* the joined run's copies wait for the next collect pass */
    bool fms_pending = false;            /* a joined flush whose device time is not reported yet */
    float flush_done_ms = 0;
    bool drop_pending = false;           /* stc_drop_runs came while a flush thread ran */
    bool lp_valid = false;               /* lp holds a position step's leaf order (a permutation of the points) */
    PinnedPool pinned;
    std::vector<PinnedBlock> deferred;   /* blocks of dropped runs whose host copy is not queued yet */
    QueryBufs *qb = nullptr;
    FlushWork flushw;
    hipEvent_t ev[14];
    bool dyn = false;                    /* dynamic schema: the QTree persists across steps */
    SchemaState sch;
    /* Serving: query batches run on their own stream against the state the
     * last completed step published, while the next step runs. A batch holds
     * a lease on the meeting array and the memtable epoch it reads; a step
     * waits for the leases on what it overwrites (the other meeting array,
     * and the memtable from index 0 once a flush emptied it). */
    hipStream_t qs = nullptr;
    std::mutex q_mu;                     /* one query batch at a time: its buffers and stream */
    std::mutex pub_mu;                   /* pub_*, readers_* */
    std::mutex err_mu;                   /* err: a query thread and the stepping thread can both fail */
    std::condition_variable pub_cv;
    QueryState pub;                      /* the state queries read */
    bool pub_valid = false;
    int pub_meet = 0;                    /* meeting array of pub */
    uint32_t epoch = 0;                  /* memtable epoch: +1 at every flush */
    uint32_t pub_epoch = 0;
    uint32_t readers_meet[2] = {0, 0};
    uint32_t readers_epoch[2] = {0, 0};
    uint32_t readers_total = 0;
};

/* publish the state after a step (or a change of the runs) to queries */
static void publish(StcCtx *c) {
    QueryState st;
    st.table = c->meet[c->cur].p;
    st.nb = c->mlen;
    st.last_t = c->last_t;
    st.stepped = c->stepped ? 1 : 0;
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
    std::lock_guard<std::mutex> lk(c->pub_mu);
    c->pub = std::move(st);
    c->pub_meet = c->cur;
    c->pub_epoch = c->epoch;
    c->pub_valid = c->stepped;
}

/* wait until no query batch reads meeting array mi (and, with old_memtable,
 * the memtable of an earlier epoch) */
static void wait_readers(StcCtx *c, int mi, bool old_memtable) {
    std::unique_lock<std::mutex> lk(c->pub_mu);
    c->pub_cv.wait(lk, [&] {
        return c->readers_meet[mi] == 0 && (!old_memtable || c->readers_epoch[(c->epoch + 1) & 1] == 0);
    });
}
/* wait until no query batch is in flight */
static void wait_all_readers(StcCtx *c) {
    std::unique_lock<std::mutex> lk(c->pub_mu);
    c->pub_cv.wait(lk, [&] { return c->readers_total == 0; });
}

/* the QTree the step's kernels descend */
static inline Node *tree_nodes(StcCtx *c) { return c->dyn ? c->sch.nodes : c->nodes.p; }

static int fail(StcCtx *c, int code, const std::string &msg) {
    std::lock_guard<std::mutex> lk(c->err_mu);
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

/* return the status of joining a pending asynchronous flush (its message is already set) */
#define JOIN_FLUSH(c)                  \
    do {                               \
        const int jrc_ = join_flush(c); \
        if (jrc_) return jrc_;         \
    } while (0)

static int sync_counters(StcCtx *c) {
    CK(hipMemcpyAsync(c->hcnt, c->dcnt, sizeof(Counters), hipMemcpyDeviceToHost, c->s));
    CK(hipStreamSynchronize(c->s));
    return 0;
}

static void wait_all_readers(StcCtx *c);
static int ensure_records(StcCtx *c, uint64_t need) {
    if (need <= c->rec_cap) return 0;
    wait_all_readers(c); /* the old buffers are freed below */
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

/* the last step's pairs (its pair slots without the dead ones), sorted by key */
static int step_pairs(StcCtx *c, std::vector<PairHit> &ph) {
    ph.resize(c->plen);
    if (c->pairs_boxed) {
        if (c->plen) CK(hipMemcpy(ph.data(), c->pbox.p, c->plen * sizeof(PairHit), hipMemcpyDeviceToHost));
    } else if (c->plen) { /* slots, owner by owner: the MBR is the box of the owner's and the partner's float positions */
        std::vector<PairSlot> sl(c->plen);
        std::vector<float2> pos(c->n);
        std::vector<uint32_t> off(c->n + 1ull);
        CK(hipMemcpy(sl.data(), c->pnew.p, c->plen * sizeof(PairSlot), hipMemcpyDeviceToHost));
        CK(hipMemcpy(pos.data(), c->posf.p, (size_t)c->n * sizeof(float2), hipMemcpyDeviceToHost));
        CK(hipMemcpy(off.data(), c->moff[c->cur].p, (c->n + 1ull) * 4, hipMemcpyDeviceToHost));
        for (uint32_t a = 0; a < c->n; a++)
            for (uint64_t i = off[a]; i < off[a + 1] && i < c->plen; i++) {
                const float2 o = pos[a];
                ph[i].key = ((uint64_t)a << 32) | sl[i].partner;
                ph[i].mbr[0] = std::fmin(o.x, sl[i].px);
                ph[i].mbr[1] = std::fmin(o.y, sl[i].py);
                ph[i].mbr[2] = std::fmax(o.x, sl[i].px);
                ph[i].mbr[3] = std::fmax(o.y, sl[i].py);
            }
    }
    ph.erase(std::remove_if(ph.begin(), ph.end(), [](const PairHit &e) { return (uint32_t)e.key == DEAD_B; }),
             ph.end());
    std::sort(ph.begin(), ph.end(), [](const PairHit &a, const PairHit &b) { return a.key < b.key; });
    return 0;
}

/* A run's pinned block back to the pool. With finish_copy, a host copy owed to
 * the block still happens: the block waits in `deferred` until the copy is
 * queued (the pool then reuses it once the copy's event has fired). Without
 * it, a copy that was never queued is dropped. */
static void release_block(StcCtx *c, PinnedBlock b, bool finish_copy = false) {
    if (finish_copy && c->flushw.job_set && c->flushw.job.dst == b.p) {
        c->deferred.push_back(b);
        return;
    }
    stc_flush_cancel(&c->flushw, b.p);
    c->pinned.release(b);
}

/* deferred blocks whose copy has been queued go back to the pool */
static void settle_deferred(StcCtx *c) {
    for (size_t i = 0; i < c->deferred.size();) {
        if (c->flushw.job_set && c->flushw.job.dst == c->deferred[i].p) {
            i++;
            continue;
        }
        c->pinned.release(c->deferred[i]);
        c->deferred[i] = c->deferred.back();
        c->deferred.pop_back();
    }
}

/* queue the host copies of the last flushed run on the copy stream. None
 * while a flush thread owns the flush buffers; a run just taken from one waits
 * for the next collect pass (copies to the host slow concurrent kernels that
 * write memory, and the collect pass mostly computes) unless `now`. */
static int queue_run_copies(StcCtx *c, bool now = true) {
    if (c->frun) return 0;
    if (c->flush_fresh && !now) return 0;
    c->flush_fresh = false;
    if (stc_flush_release(&c->flushw, c->s)) return STC_E_HIP;
    settle_deferred(c);
    return 0;
}

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
    return stc_query_ctf_stats(q, run.ctfs.data(), run.stats, s) ? fail(c, STC_E_HIP, std::string("run statistics: ") + g_last_fail) : 0;
}

/* release every run's host block (a queued copy still completes) and HBM block */
static void drop_all_runs(StcCtx *c) {
    wait_all_readers(c); /* the runs' device blocks are freed below */
    for (auto &r : c->runs) {
        release_block(c, r.blk, true);
        if (r.dev) {
            (void)hipStreamSynchronize(c->s);
            (void)hipFree(r.dev);
            r.dev = nullptr;
        }
    }
    c->runs.clear();
    publish(c);
}

/* Wait for an asynchronous flush and take its run: the main stream waits for
 * the flush's device work (the next merge reuses the memtable that the flush
 * reads and resets), the run joins the list (and HBM, with device_runs), and
 * the state with it is published. */
static int join_flush(StcCtx *c) {
    if (!c->frun) return STC_OK;
    c->fthr.join();
    c->frun = false;
    settle_deferred(c);
    if (c->frc) {
        release_block(c, c->fout.blk);
        c->fout = FlushOut();
        return fail(c, c->frc, "flush: " + c->ferr);
    }
    CK(hipStreamWaitEvent(c->s, c->fev[2], 0));
    CK(hipEventSynchronize(c->fev[2]));
    CK(hipEventElapsedTime(&c->flush_done_ms, c->fev[1], c->fev[2]));
    c->fms_pending = true;
    RunHost run;
    run.info = c->fout.info;
    run.ctfs.swap(c->fout.ctfs);
    run.lay = c->fout.lay;
    run.has_sorted = c->fout.has_sorted;
    run.blk = c->fout.blk;
    c->fout = FlushOut();
    if (c->cfg.device_runs) CKRC(keep_run_on_device(c, run));
    c->runs.push_back(std::move(run));
    c->flush_fresh = true;
    if (c->drop_pending) { /* stc_drop_runs came while the flush ran */
        c->drop_pending = false;
        drop_all_runs(c);
    } else {
        publish(c);
    }
    return STC_OK;
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
    /* pair sources carry two tag bits above a 30-bit index */
    if (cfg->num_objects >= (1u << 30)) {
        g_last_fail = "num_objects must be below 2^30";
        return STC_E_ARG;
    }
    StcCtx *c = new StcCtx();
    c->err.reserve(1024); /* a failure message then needs no allocation */
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
    if (noalloc_setup(c->device)) {
        c->err = g_last_fail;
        stc_destroy(c);
        return STC_E_HIP;
    }
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
    if (hipStreamCreateWithFlags(&c->qs, hipStreamNonBlocking) != hipSuccess) return bad("query stream");
    /* the flush stream only when it is used: HIP spreads a process's streams over its hardware queues in
     * creation order, and with four streams per context every context's main stream would share one
     * queue (measured: 64 batched streams of 5M objects step at 163M instead of 310M point-updates/s) */
    if (cfg->async_flush) {
        if (hipStreamCreateWithFlags(&c->fs, hipStreamNonBlocking) != hipSuccess) return bad("flush stream");
        for (int i = 0; i < 3; i++)
            if (hipEventCreate(&c->fev[i]) != hipSuccess) return bad("flush events");
    }
    for (int i = 0; i < 14; i++) hipEventCreate(&c->ev[i]);
    uint32_t n = c->n;
    size_t node_cap = 1 + 4ull * NLEVELS * (n / (c->cap + 1) + 1);
    c->dyn = cfg->dynamic_schema != 0;
    if (c->dyn) {
        /* schema_stack_capacity = 1.6 * grids_stack_capacity, grids_stack_capacity =
         * 4 * max(100000, N / grid_capacity) (workbench.cpp:36-41); grows on demand */
        const uint64_t grids = 4ull * std::max<uint64_t>(100000, n / c->cap);
        const uint32_t slots = (uint32_t)std::min<uint64_t>(grids * 8 / 5, 0x7FFFFFFFull);
        if (stc_schema_init(&c->sch, n, slots)) return bad("schema buffers");
        node_cap = std::max<size_t>(node_cap, c->sch.cap);
    }
    size_t front_cap = n / (c->cap + 1) + 2;
    if (c->xy_in.ensure(n) || c->codes.ensure(n) || c->idx.ensure(n) || c->codes_alt.ensure(n) ||
        c->idx_alt.ensure(n) || c->dlon.ensure(n) || c->posf.ensure(n) || c->lxy.ensure(n) || c->ld.ensure(n) ||
        c->nodes.ensure(node_cap) || c->frontier_a.ensure(front_cap) || c->frontier_b.ensure(front_cap) ||
        c->pt_leaf.ensure(n) || c->pt_lv.ensure(n) || c->pt_rank.ensure(n) || c->lp.ensure(n) ||
        c->lleaf.ensure(n) || c->lunits.ensure(n) || c->prange.ensure(n) ||
        c->lkeys.ensure(n) || c->ucount.ensure(n) || c->uoffset.ensure(n) || c->units.ensure(4ull * n) ||
        c->amb.ensure(1 << 16) || c->pcount.ensure(n) || c->moff[0].ensure(n + 1ull) || c->moff[1].ensure(n + 1ull) ||
        c->spill_big.ensure(n) || c->spill_huge.ensure(n) || c->spill_giant.ensure(n) ||
        c->rev_start.ensure(node_cap + 1) ||
        c->leaf_count.ensure(node_cap) || c->leaf_start.ensure(node_cap) || c->leaf_node.ensure(node_cap) ||
        c->temp.ensure(stc_radix_temp_words((uint32_t)std::max<size_t>(n, node_cap)) + 64))
        return bad("device allocation failed");
    if (hipMalloc((void **)&c->dcnt, sizeof(Counters)) != hipSuccess) return bad("counters");
    if (hipHostMalloc((void **)&c->hcnt, sizeof(Counters), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->hinit, sizeof(Counters), hipHostMallocDefault) != hipSuccess)
        return bad("pinned");
    /* no meetings yet: every segment of the current array is empty */
    uint64_t nb = cfg->num_meeting_buckets ? cfg->num_meeting_buckets : 2ull * n;
    if (c->pnew.ensure(nb) || c->meet[0].ensure(nb) || c->meet[1].ensure(nb))
        return bad("meeting arrays");
    if (hipMemsetAsync(c->moff[0].p, 0, 4ull * (n + 1), c->s) != hipSuccess) return bad("meeting offsets");
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
    (void)join_flush(c);
    wait_all_readers(c);
    if (c->s) hipStreamSynchronize(c->s);
    c->xy_in.release(); c->codes.release(); c->idx.release(); c->codes_alt.release(); c->idx_alt.release();
    c->temp.release(); c->dlon.release(); c->posf.release(); c->lxy.release(); c->ld.release(); c->nodes.release();
    c->frontier_a.release(); c->frontier_b.release(); c->leaf_count.release(); c->leaf_start.release();
    c->leaf_node.release(); c->pt_leaf.release(); c->pt_lv.release(); c->pt_rank.release(); c->lp.release();
    c->lleaf.release(); c->lunits.release(); c->urange.release(); c->prange.release();
    c->lkeys.release(); c->ucount.release(); c->uoffset.release(); c->units.release();
    c->amb.release(); c->amb_xy.release();
    c->rev_leaf.release(); c->rev_pid.release(); c->rev_leaf_alt.release(); c->rev_pid_alt.release();
    c->rev_start.release();
    c->rev.release(); c->rev_dl.release();
    c->pcount.release(); c->pnew.release(); c->acc.release();
    for (int i = 0; i < 2; i++) { c->meet[i].release(); c->moff[i].release(); }
    c->ptmp.release(); c->tbounds.release(); c->towner.release(); c->psrc.release(); c->pkrun.release();
    c->spill_big.release(); c->spill_huge.release(); c->spill_giant.release(); c->gsrt.release();
    c->units_in.release(); c->ukeys.release(); c->ukeys_alt.release();
    c->uidx.release(); c->uidx_alt.release(); c->utemp.release();
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
        release_block(c, r.blk);
        if (r.dev) (void)hipFree(r.dev);
    }
    c->runs.clear();
    for (auto &b : c->deferred) release_block(c, b);
    c->deferred.clear();
    c->pinned.clear();
    stc_query_bufs_free(c->qb);
    c->qb = nullptr;
    for (int i = 0; i < 14; i++) if (c->ev[i]) hipEventDestroy(c->ev[i]);
    stc_schema_free(&c->sch);
    if (c->s) hipStreamDestroy(c->s);
    if (c->qs) hipStreamDestroy(c->qs);
    if (c->fs) hipStreamDestroy(c->fs);
    for (int i = 0; i < 3; i++) if (c->fev[i]) hipEventDestroy(c->fev[i]);
    delete c;
    return STC_OK;
}

const char *stc_last_error(void *vc) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return g_last_fail.c_str();
    /* a copy per calling thread: the context's message can change under another thread's failure */
    static thread_local std::string last;
    std::lock_guard<std::mutex> lk(c->err_mu);
    last = c->err;
    return last.c_str();
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
    uint64_t candidates = 0;
};
static int step_tail(StcCtx *c, uint32_t t, StepGeom &g, stc_step_out *out);

/* the QTree rebuilt from the step's sorted cell codes (--disable_dynamic_schema,
 * geometry.cu:1629-1651) */
static int rebuild_tree(StcCtx *c, uint32_t *n_nodes, uint32_t *n_leaves) {
    const uint32_t n = c->n;
    hipStream_t s = c->s;
    Counters *cnt = c->dcnt;
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
    const Counters &h = *c->hcnt;
    if (h.bad_points) return fail(c, STC_E_INPUT, "point outside the map box");
    if (h.table_fail) return fail(c, STC_E_CAPACITY, "QTree node capacity exceeded");
    if (n <= c->cap) return fail(c, STC_E_ARG, "root does not split: too few objects for grid_capacity");
    *n_nodes = h.nodes;
    *n_leaves = h.leaves;
    return 0;
}

static int step_impl(StcCtx *c, const double2 *xy, uint32_t t, stc_step_out *out) {
    const uint32_t n = c->n;
    hipStream_t s = c->s;
    Counters *cnt = c->dcnt;
    memset(c->hinit, 0, sizeof(Counters));
    c->hinit->kv_count = c->kv_count;
    c->hinit->oversize_oid_count = c->oversize_oid_count;
    CK(hipMemcpyAsync(cnt, c->hinit, sizeof(Counters), hipMemcpyHostToDevice, s));
    CK(hipEventRecord(c->ev[0], s));

    /* 1-2: cells, Morton codes, dpm_lon; sort codes */
    hipLaunchKernelGGL(k_prep, dim3(blocks_for(n, 256)), dim3(256), 0, s, xy, n, c->map, c->codes.p, c->idx.p,
                       c->dlon.p, c->posf.p, cnt);
    CK(hipGetLastError());
    /* the order the partition walks the points in: Morton order, or with a kept schema the previous
     * step's leaf order (objects move a few metres a step, so it is as coherent, without the sort) */
    const bool reuse_order = c->dyn && c->sch.present && c->lp_valid;
    if (!reuse_order) CKRC(stc_radix_sort_u32(c->codes.p, c->idx.p, c->codes_alt.p, c->idx_alt.p, n, 28, c->temp.p, s));
    const uint32_t *walk = reuse_order ? (const uint32_t *)c->lp.p : (const uint32_t *)c->idx.p;

    /* 3: tree: rebuilt from the step's points, or the persistent dynamic schema */
    uint32_t n_nodes, n_leaves;
    if (c->dyn) {
        if (!c->sch.present) { /* the first step's points build it (process.cpp:401-402) */
            int rc = stc_schema_build(&c->sch, xy, c->idx.p, n, c->map, c->cap, c->cfg.reach_distance, c->temp.p,
                                      c->temp.cap, s);
            if (rc) return fail(c, rc, rc == STC_E_ARG ? "root does not split: too few objects for grid_capacity"
                                                       : std::string("schema build: ") + g_last_fail);
        }
        const uint32_t nl = c->sch.n_leaves;
        uint32_t *before = c->leaf_node.p;
        if (c->leaf_count.ensure(nl + 1) || c->leaf_start.ensure(nl + 1) || c->leaf_node.ensure(nl + 1) ||
            c->rev_start.ensure(nl + 1ull))
            return fail(c, STC_E_CAPACITY, "leaf buffers");
        if (c->leaf_node.p != before) c->sch.dirty = true;
        if (c->sch.dirty) CKRC(stc_schema_number_leaves(&c->sch, c->leaf_node.p, c->temp.p, c->temp.cap, s));
        n_nodes = c->sch.n_valid;
        n_leaves = nl;
    } else {
        CKRC(rebuild_tree(c, &n_nodes, &n_leaves));
    }
    Counters h;
    c->last_nodes = n_nodes;
    c->last_leaves = n_leaves;

    /* 4: partition */
    CK(hipMemsetAsync(c->leaf_count.p, 0, 4ull * n_leaves, s));
    if (!c->dyn)
        hipLaunchKernelGGL(k_leaf_nodes, dim3(blocks_for(n_nodes, 256)), dim3(256), 0, s, c->nodes.p, n_nodes,
                           c->leaf_node.p);
    hipLaunchKernelGGL(k_partition, dim3(blocks_for(n, 256)), dim3(256), 0, s, xy, walk, n,
                       (const Node *)tree_nodes(c), c->x_buffer, c->y_buffer, c->pt_leaf.p, c->pt_lv.p, c->leaf_count.p, cnt);
    CK(hipGetLastError());
    /* 5: leaves in id order */
    CKRC(stc_exclusive_scan_u32(c->leaf_count.p, c->leaf_start.p, n_leaves, nullptr, c->temp.p, s));
    CK(hipMemcpyAsync(c->lkeys.p, c->pt_leaf.p, 4ull * n, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_iota, dim3(blocks_for(n, 256)), dim3(256), 0, s, c->lp.p, n);
    CKRC(stc_radix_sort_u32(c->lkeys.p, c->lp.p, c->codes_alt.p, c->idx_alt.p, n,
                            std::max(1, ilog2_ceil(n_leaves)), c->temp.p, s));
    hipLaunchKernelGGL(k_gather, dim3(blocks_for(n, 256)), dim3(256), 0, s, xy, c->dlon.p, c->lp.p, c->pt_leaf.p,
                       c->leaf_start.p, n, c->lxy.p, c->ld.p, c->pt_rank.p, c->lleaf.p);
    c->lp_valid = true; /* lp: a permutation of the points, in this step's leaf order */
    hipLaunchKernelGGL(k_overflow, dim3(blocks_for(n_leaves, 256)), dim3(256), 0, s, c->leaf_count.p, n_leaves,
                       c->buf, cnt);
    CK(hipGetLastError());
    CK(hipEventRecord(c->ev[1], s));

    /* 6: border lookup (count, scan, emit) */
    double r = c->cfg.reach_distance;
    hipLaunchKernelGGL(k_lookup<false>, dim3(blocks_for(n, 128)), dim3(128), 0, s, (const uint32_t *)c->lp.p,
                       (const double2 *)c->lxy.p, (const double *)c->ld.p, n, (const Node *)tree_nodes(c),
                       c->pt_lv.p, r, c->band, c->ucount.p, (const uint32_t *)nullptr, (uint2 *)nullptr, cnt);
    CK(hipGetLastError());
    uint32_t *units_total = (uint32_t *)((char *)cnt + offsetof(Counters, units));
    CKRC(stc_exclusive_scan_u32(c->ucount.p, c->uoffset.p, n, units_total, c->temp.p, s));
    CKRC(sync_counters(c));
    h = *c->hcnt;
    if (h.bad_points) return fail(c, STC_E_INPUT, "point outside the map box");
    if (h.lookup_overflow) return fail(c, STC_E_CAPACITY, "border lookup: node stack depth exceeded");
    uint64_t n_units = (uint32_t)h.units;
    if (n_units >= (1u << 30)) return fail(c, STC_E_CAPACITY, "2^30 or more cross units in one step");
    if (c->units.ensure(n_units + 1)) return fail(c, STC_E_CAPACITY, "unit buffer allocation failed");
    hipLaunchKernelGGL(k_lookup<true>, dim3(blocks_for(n, 128)), dim3(128), 0, s, (const uint32_t *)c->lp.p,
                       (const double2 *)c->lxy.p, (const double *)c->ld.p, n, (const Node *)tree_nodes(c),
                       c->pt_lv.p, r, c->band, c->ucount.p, (const uint32_t *)c->uoffset.p, c->units.p, cnt);
    CK(hipGetLastError());
    CK(hipEventRecord(c->ev[2], s));

    /* 7: reverse cross units: per leaf, the border points with a cross unit to it, by pid */
    if (c->rev_leaf.ensure(n_units + 1) || c->rev_pid.ensure(n_units + 1) || c->rev_leaf_alt.ensure(n_units + 1) ||
        c->rev.ensure(n_units + 1) || c->rev_dl.ensure(n_units + 1) ||
        c->rev_pid_alt.ensure(n_units + 1) ||
        c->rev_start.ensure(n_leaves + 1ull) ||
        c->temp.ensure(stc_radix_temp_words((uint32_t)std::max<uint64_t>(std::max<uint64_t>(n_units, n),
                                                                          std::max<size_t>(c->nodes.cap, c->sch.cap))) + 64))
        return fail(c, STC_E_CAPACITY, "reverse unit buffers");
    if (n_units) {
        hipLaunchKernelGGL(k_units_split, dim3(blocks_for(n_units, 256)), dim3(256), 0, s, c->units.p,
                           (uint32_t)n_units, c->rev_leaf.p, c->rev_pid.p);
        CK(hipGetLastError());
        CKRC(stc_radix_sort_u32(c->rev_leaf.p, c->rev_pid.p, c->rev_leaf_alt.p, c->rev_pid_alt.p, (uint32_t)n_units,
                                std::max(1, ilog2_ceil(n_leaves)), c->temp.p, s));
        hipLaunchKernelGGL(k_rev_gather, dim3(blocks_for(n_units, 256)), dim3(256), 0, s, (const uint32_t *)c->rev_pid.p,
                           (uint32_t)n_units, xy, (const double *)c->dlon.p, c->rev.p,
                           c->rev_dl.p);
        CK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rev_starts, dim3(blocks_for(n_leaves + 1ull, 256)), dim3(256), 0, s,
                       (const uint32_t *)c->rev_leaf.p, (uint32_t)n_units, n_leaves, c->rev_start.p);
    CK(hipGetLastError());

    /* 8: the pairs owner by owner: count (a band combination reserves a slot),
     * the band decided on the host, segment offsets, write */
    hipLaunchKernelGGL(k_units_lorder, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const uint32_t *)c->lp.p,
                       (const uint32_t *)c->uoffset.p, (const uint32_t *)c->ucount.p, n, c->lunits.p);
    if (c->urange.ensure(n_units + 1)) return fail(c, STC_E_CAPACITY, "unit ranges");
    if (n_units)
        hipLaunchKernelGGL(k_unit_ranges, dim3(blocks_for(n_units, 256)), dim3(256), 0, s, (const uint2 *)c->units.p,
                           (uint32_t)n_units, (const uint32_t *)c->leaf_start.p, (const uint32_t *)c->leaf_count.p,
                           (const uint32_t *)c->lp.p, c->buf, c->urange.p);
    hipLaunchKernelGGL(k_point_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, (const uint32_t *)c->lp.p,
                       (const uint32_t *)c->lleaf.p, n, (const uint32_t *)c->leaf_start.p,
                       (const uint32_t *)c->leaf_count.p, (const uint32_t *)c->rev_start.p,
                       (const uint32_t *)c->rev_pid.p, c->buf, c->prange.p);
    CK(hipGetLastError());
    const PairGeo pg{xy, c->dlon.p, c->lxy.p, c->pt_leaf.p, c->pt_rank.p, c->leaf_start.p,
                     c->leaf_count.p, c->lp.p, c->ld.p, c->lleaf.p, c->lunits.p, c->urange.p, c->prange.p,
                     c->units.p, c->uoffset.p,
                     c->ucount.p, c->rev_start.p, c->rev_pid.p,
                     c->rev.p, c->rev_dl.p, n, c->buf, r, c->band, r * r * (1.0 - 1e-12),
                     r * r * (1.0 + 1e-12)};
    /* a multiple of 8 workgroups, for the XCD-aware owner order; at most 32k from 4M objects on (at 10M
     * objects a step is 0.5 ms shorter than with 16k, measured three times each on one box), else 16k
     * (with 32k, 64 batched streams of 48k-500k objects stepped 8 % slower) */
    const uint32_t pcap = n >= (1u << 22) ? 256u * 128u : 256u * 64u;
    const uint32_t pgrid = (uint32_t)std::min<uint64_t>((blocks_for((uint64_t)n * STC_WAVE, 256) + 7) & ~7ull, pcap);
    /* the last run's host copies overlap the collect pass, which computes more than it writes */
    if (queue_run_copies(c)) return fail(c, STC_E_HIP, std::string("run copies: ") + g_last_fail);
    if (c->pkrun.ensure(n)) return fail(c, STC_E_CAPACITY, "pair key runs");
    if (!c->psrc.p && c->psrc.ensure(std::max<size_t>((size_t)n * 16, 1u << 20)))
        return fail(c, STC_E_CAPACITY, "pair sources");
    for (int attempt = 0;; attempt++) {
        /* one clear for candidates .. pk_over (a fill launch can wait behind the run copies) */
        CK(hipMemsetAsync(&cnt->candidates, 0,
                          offsetof(Counters, pk_over) + sizeof(uint32_t) - offsetof(Counters, candidates), s));
        CK(hipEventRecord(c->ev[10], s));
        hipLaunchKernelGGL(k_pair_collect, dim3(pgrid), dim3(256), 0, s, pg, c->pcount.p, c->pkrun.p, c->psrc.p,
                           (uint64_t)std::min<size_t>(c->psrc.cap, 0xFFFFFFFFu), c->amb.p,
                           (uint32_t)std::min<size_t>(c->amb.cap, 0xFFFFFFFFu), c->spill_big.p, c->spill_huge.p, cnt);
        CK(hipGetLastError());
        CK(hipEventRecord(c->ev[11], s));
        CKRC(sync_counters(c));
        h = *c->hcnt;
        for (uint32_t k = 0; k < PK_SLICES; k++) {
            h.candidates += h.cand_part[k];
            h.pairs += h.pairs_part[k];
        }
        bool again = false;
        if (h.ambiguous > c->amb.cap) {
            if (c->amb.ensure(h.ambiguous + 1024)) return fail(c, STC_E_CAPACITY, "ambiguous list allocation failed");
            again = true;
        }
        if (h.pk_over) { /* a slice was short: grow every slice to the fullest one's reservations, with room */
            unsigned long long most = 0;
            for (uint32_t k = 0; k < PK_SLICES; k++) most = std::max(most, h.pk_used[k]);
            const unsigned long long want = most * PK_SLICES + most * PK_SLICES / 8;
            if (want >= 0xFFFFFFFFull) return fail(c, STC_E_CAPACITY, "more than 2^32 - 1 pair sources in one step");
            if (c->psrc.ensure(want)) return fail(c, STC_E_CAPACITY, "pair sources");
            again = true;
        }
        if (!again) break;
        if (attempt >= 2) return fail(c, STC_E_CAPACITY, "pair sources or ambiguous list still short after growing");
    }
    const uint64_t total = h.pairs, n_amb = h.ambiguous, n_cand = h.candidates;
    if (total >= 0xFFFFFFFFull) return fail(c, STC_E_CAPACITY, "more than 2^32 - 2 pairs in one step");
    std::vector<unsigned long long> accepted;
    if (n_amb) {
        /* decide the band with the C library's sin (geometry.h:40-48) */
        std::vector<uint2> ab(n_amb);
        std::vector<double4> pos(n_amb);
        if (c->amb_xy.ensure(n_amb)) return fail(c, STC_E_CAPACITY, "band position buffer");
        hipLaunchKernelGGL(k_amb_gather, dim3(blocks_for(n_amb, 256)), dim3(256), 0, s, xy,
                           (const uint2 *)c->amb.p, (uint32_t)n_amb, c->amb_xy.p);
        CK(hipGetLastError());
        CK(hipMemcpyAsync(ab.data(), c->amb.p, n_amb * sizeof(uint2), hipMemcpyDeviceToHost, s));
        CK(hipMemcpyAsync(pos.data(), c->amb_xy.p, n_amb * sizeof(double4), hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        for (uint64_t i = 0; i < n_amb; i++) {
            const uint2 e = ab[i];
            const double4 pq = pos[i];
            double dx = (pq.x - pq.z) / host_dpm_lon(pq.y);
            double dy = (pq.y - pq.w) / STC_DPM_LAT;
            if (std::sqrt(dx * dx + dy * dy) <= r)
                accepted.push_back(((unsigned long long)std::min(e.x, e.y) << 32) | std::max(e.x, e.y));
        }
        std::sort(accepted.begin(), accepted.end());
        accepted.erase(std::unique(accepted.begin(), accepted.end()), accepted.end());
    }
    if (c->acc.ensure(accepted.size() + 1)) return fail(c, STC_E_CAPACITY, "band pair list");
    if (!accepted.empty())
        CK(hipMemcpy(c->acc.p, accepted.data(), accepted.size() * sizeof(unsigned long long), hipMemcpyHostToDevice));
    const int nx = c->cur ^ 1;
    CKRC(stc_exclusive_scan_u32(c->pcount.p, c->moff[nx].p, n, c->moff[nx].p + n, c->temp.p, s));
    wait_readers(c, nx, false); /* queries of an earlier state may still read meet[nx] */
    if (c->pnew.ensure(total + 1) || c->meet[nx].ensure(total + 1))
        return fail(c, STC_E_CAPACITY, "pair and meeting arrays");
    CK(hipEventRecord(c->ev[6], s));
    /* the place pass runs faster with up to four times the collect pass's workgroups (measured at 10M
     * objects); both counts are multiples of 8 for the XCD order */
    const uint32_t wgrid = (uint32_t)std::min<uint64_t>((blocks_for((uint64_t)n * STC_WAVE, 256) + 7) & ~7ull, 65536u);
    hipLaunchKernelGGL(k_pair_place, dim3(wgrid), dim3(256), 0, s, pg, (const uint32_t *)c->moff[nx].p,
                       (const uint2 *)c->pkrun.p, (const uint32_t *)c->psrc.p, c->acc.p,
                       (uint32_t)accepted.size(), c->pnew.p, 32u - (uint32_t)__builtin_clz(c->n));
    CK(hipGetLastError());
    CK(hipEventRecord(c->ev[7], s));
    if (h.big) { /* segments longer than the place pass sorts: by waves, by blocks, in global scratch */
        hipLaunchKernelGGL(k_seg_sort_wave, dim3(blocks_for((uint64_t)h.big * STC_WAVE, 256)), dim3(256), 0, s, pg,
                           (const uint32_t *)c->spill_big.p, h.big, (const uint32_t *)c->moff[nx].p,
                           (const uint2 *)c->pkrun.p, (const uint32_t *)c->psrc.p, c->acc.p,
                           (uint32_t)accepted.size(), c->pnew.p);
        CK(hipGetLastError());
    }
    if (h.huge) { /* staged at their segments' places: room for all slots, allocated only when needed */
        if (c->ptmp.ensure(total + 1)) return fail(c, STC_E_CAPACITY, "long pair segments");
        hipLaunchKernelGGL(k_huge_expand, dim3(blocks_for((uint64_t)h.huge * STC_WAVE, 256)), dim3(256), 0, s, pg,
                           (const uint32_t *)c->spill_huge.p, h.huge, (const uint32_t *)c->moff[nx].p,
                           (const uint2 *)c->pkrun.p, (const uint32_t *)c->psrc.p, c->acc.p,
                           (uint32_t)accepted.size(), c->ptmp.p);
        CK(hipGetLastError());
        hipLaunchKernelGGL((k_seg_sort<256, SEG_CAP, false>), dim3(std::min<uint32_t>(h.huge, 4096)), dim3(256), 0, s,
                           (const uint32_t *)c->spill_huge.p, h.huge, (const uint32_t *)c->moff[nx].p,
                           (const PairSlot *)c->ptmp.p, c->pnew.p, c->spill_giant.p, &cnt->giant,
                           (unsigned long long *)nullptr, 0u, cnt);
        CK(hipGetLastError());
        CKRC(sync_counters(c));
        h = *c->hcnt;
    }
    if (h.giant) {
        std::vector<uint32_t> owners(h.giant), seg(2);
        CK(hipMemcpy(owners.data(), c->spill_giant.p, 4ull * h.giant, hipMemcpyDeviceToHost));
        uint32_t longest = 1;
        for (uint32_t o : owners) {
            CK(hipMemcpy(seg.data(), c->moff[nx].p + o, 8, hipMemcpyDeviceToHost));
            longest = std::max(longest, seg[1] - seg[0]);
        }
        uint32_t gcap = 1;
        while (gcap < longest) gcap <<= 1;
        const uint32_t hgrid = std::min<uint32_t>(h.giant, 64);
        if (c->gsrt.ensure((size_t)gcap * hgrid)) return fail(c, STC_E_CAPACITY, "sort scratch of a large segment");
        hipLaunchKernelGGL((k_seg_sort<256, 1, true>), dim3(hgrid), dim3(256), 0, s, (const uint32_t *)c->spill_giant.p,
                           h.giant, (const uint32_t *)c->moff[nx].p, (const PairSlot *)c->ptmp.p, c->pnew.p,
                           (uint32_t *)nullptr, (uint32_t *)nullptr, c->gsrt.p, gcap, cnt);
        CK(hipGetLastError());
    }
    c->plen = total;
    c->pairs_boxed = false;
    CK(hipEventRecord(c->ev[3], s));
    StepGeom g{n_nodes, n_leaves, n_units, 0, n_amb, 0, n_cand};
    return step_tail(c, t, g, out);
}

/* Steps 9-10, shared by the position step and the replay of persisted
 * meeting units: identify, rebuild when needed, flush, report. */
static int step_tail(StcCtx *c, uint32_t t, StepGeom &g, stc_step_out *out) {
    const uint32_t n = c->n;
    hipStream_t s = c->s;
    Counters *cnt = c->dcnt;
    /* 9: merge the step's pairs with the previous step's meetings; meetings
     * without a pair expire and the long ones become records. An asynchronous
     * flush of the previous step must be done first: the merge writes the
     * memtable and the per-object state that it reads and resets. */
    {
        const int rc = join_flush(c);
        if (rc) return rc;
    }
    if (queue_run_copies(c, false)) return fail(c, STC_E_HIP, std::string("run copies: ") + g_last_fail);
    const int cu = c->cur, nx = cu ^ 1;
    const uint64_t kv_before = c->kv_count;
    if (ensure_records(c, kv_before + 2 * std::max<uint64_t>(c->last_expired + c->last_expired / 4, 1 << 16) + 2))
        return fail(c, STC_E_CAPACITY, "record buffer allocation failed");
    const MergeArgs ma{t, c->cfg.min_meet_time, c->cfg.max_meet_time, c->end_time_min, c->map, c->last_t};
    CKRC(sync_counters(c));
    Counters h = *c->hcnt;
    if (h.table_fail) return fail(c, STC_E_HIP, "a pair segment was longer than its sort's scratch");
    const Counters before = h;
    /* tiles of the step's pairs and their ranges in the previous meetings */
    const uint32_t n_tiles = (uint32_t)std::max<uint64_t>((c->plen + MTILE - 1) / MTILE, 1);
    if (c->tbounds.ensure(n_tiles + 1ull) || c->towner.ensure(n_tiles)) return fail(c, STC_E_CAPACITY, "merge tile bounds");
    if (c->pairs_boxed) {
        hipLaunchKernelGGL(k_tile_bounds<PairHit>, dim3(blocks_for(n_tiles + 1ull, 256)), dim3(256), 0, s,
                           (const PairHit *)c->pbox.p, c->plen, (const Bucket *)c->meet[cu].p,
                           (const uint32_t *)c->moff[cu].p, (const uint32_t *)c->moff[nx].p, c->mlen, n_tiles,
                           (const uint32_t *)nullptr, c->tbounds.p);
    } else {
        hipLaunchKernelGGL(k_tile_owner, dim3(blocks_for(c->n, 256)), dim3(256), 0, s, (const uint32_t *)c->moff[nx].p,
                           c->n, c->towner.p);
        hipLaunchKernelGGL(k_tile_bounds<PairSlot>, dim3(blocks_for(n_tiles + 1ull, 256)), dim3(256), 0, s,
                           (const PairSlot *)c->pnew.p, c->plen, (const Bucket *)c->meet[cu].p,
                           (const uint32_t *)c->moff[cu].p, (const uint32_t *)c->moff[nx].p, c->mlen, n_tiles,
                           (const uint32_t *)c->towner.p, c->tbounds.p);
    }
    CK(hipGetLastError());
    const uint32_t mgrid = std::max<uint32_t>(std::min<uint32_t>((n_tiles + 3) / 4, 65536), 1);
    /* the merge writes meet[nx] and, after a flush emptied the memtable, records from index 0 */
    wait_readers(c, nx, kv_before == 0);
    for (int attempt = 0;; attempt++) {
        /* the merge rewrites the same buckets, so a re-run with more record room is exact */
        *c->hinit = before;
        c->hinit->new_keys = c->hinit->continuing = 0;
        c->hinit->live = c->hinit->expired = 0;
        c->hinit->kv_count = kv_before;
        CK(hipMemcpyAsync(cnt, c->hinit, sizeof(Counters), hipMemcpyHostToDevice, s));
        const MemtableDev mt{c->rec_keys, c->rec_boxes, c->rec_cap, c->same_pid_count, c->o_boxs, c->sids};
        CK(hipEventRecord(c->ev[8], s));
        if (c->pairs_boxed)
            hipLaunchKernelGGL(k_meet_merge<PairHit>, dim3(mgrid), dim3(256), 0, s, (const PairHit *)c->pbox.p,
                               c->plen, (const Bucket *)c->meet[cu].p, (const uint32_t *)c->tbounds.p, n_tiles,
                               c->meet[nx].p, ma, mt, (const float2 *)nullptr, (const uint32_t *)nullptr,
                               (const uint32_t *)nullptr, c->n, cnt);
        else
            hipLaunchKernelGGL(k_meet_merge<PairSlot>, dim3(mgrid), dim3(256), 0, s, (const PairSlot *)c->pnew.p,
                               c->plen, (const Bucket *)c->meet[cu].p, (const uint32_t *)c->tbounds.p, n_tiles,
                               c->meet[nx].p, ma, mt, (const float2 *)c->posf.p, (const uint32_t *)c->moff[nx].p,
                               (const uint32_t *)c->towner.p, c->n, cnt);
        CK(hipGetLastError());
        CK(hipEventRecord(c->ev[9], s));
        CKRC(sync_counters(c));
        h = *c->hcnt;
        if (h.table_fail) return fail(c, STC_E_HIP, "merge: the step's pairs are not sorted by key");
        if (h.kv_count <= c->rec_cap) break;
        if (attempt >= 2 || ensure_records(c, h.kv_count + h.kv_count / 4 + 2))
            return fail(c, STC_E_CAPACITY, "record buffer allocation failed");
    }
    /* per-object memtable state of the new records */
    if (h.kv_count > kv_before) {
        const MemtableDev mt{c->rec_keys, c->rec_boxes, c->rec_cap, c->same_pid_count, c->o_boxs, c->sids};
        hipLaunchKernelGGL(k_records_account, dim3(blocks_for(h.kv_count - kv_before, 256)), dim3(256), 0, s,
                           (const uint64_t *)c->rec_keys, (const float4 *)c->rec_boxes, kv_before, (uint64_t)h.kv_count,
                           c->cfg.max_meet_time, mt, cnt);
        CK(hipGetLastError());
        CKRC(sync_counters(c));
        h = *c->hcnt;
    }
    c->kv_count = h.kv_count;
    c->oversize_oid_count = h.oversize_oid_count;
    c->last_live = h.live;
    c->last_expired = h.expired;
    c->cur = nx;
    c->mlen = c->plen;
    g.pairs = h.live;
    CK(hipEventRecord(c->ev[4], s));

    /* 10: flush */
    uint32_t flushed = 0;
    if (c->kv_count > c->cfg.kv_restriction) {
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
        fa.oversize_oid_count = c->oversize_oid_count;
        fa.rec_keys = c->rec_keys;
        fa.rec_boxes = c->rec_boxes;
        fa.same_pid_count = c->same_pid_count;
        fa.o_boxs = c->o_boxs;
        fa.sids = c->sids;
        fa.reset_memtable = 1;
        if (c->cfg.async_flush) {
            /* the flush on its own stream and thread, after this step's records are complete */
            CK(hipEventRecord(c->fev[0], s));
            c->frc = 0;
            c->ferr.clear();
            c->fout = FlushOut();
            c->fout.acquire = pool_acquire;
            c->fout.acquire_ctx = &c->pinned;
            c->frun = true;
            c->fthr = std::thread([c, fa]() {
                if (hipSetDevice(c->device) != hipSuccess || hipStreamWaitEvent(c->fs, c->fev[0], 0) != hipSuccess ||
                    hipEventRecord(c->fev[1], c->fs) != hipSuccess) {
                    c->frc = STC_E_HIP;
                    c->ferr = "flush stream set-up";
                    return;
                }
                c->frc = stc_flush_run(&c->flushw, &fa, &c->fout, c->fs);
                if (c->frc) {
                    c->ferr = g_last_fail;
                } else if (hipEventRecord(c->fev[2], c->fs) != hipSuccess) {
                    c->frc = STC_E_HIP;
                    c->ferr = "flush end event";
                }
            });
        } else {
            RunHost run;
            FlushOut fo;
            fo.acquire = pool_acquire;
            fo.acquire_ctx = &c->pinned;
            int rc = stc_flush_run(&c->flushw, &fa, &fo, s);
            settle_deferred(c);
            if (rc) {
                release_block(c, fo.blk);
                return fail(c, rc, std::string("flush: ") + g_last_fail);
            }
            run.info = fo.info;
            run.ctfs.swap(fo.ctfs);
            run.lay = fo.lay;
            run.has_sorted = fo.has_sorted;
            run.blk = fo.blk;
            if (c->cfg.device_runs) CKRC(keep_run_on_device(c, run));
            c->runs.push_back(std::move(run));
        }
        /* the memtable's per-object state was reset inside the flush (reset_memtable) */
        c->kv_count = 0;
        c->oversize_oid_count = 0;
        c->end_time_min = t; /* process.cpp:271 with end_time_max = cur_time */
        c->epoch++;
        flushed = 1;
    }
    /* 11: the dynamic schema's update for the next step (geometry.cu:2159-2182) */
    const bool schema_step = c->dyn && g.nodes;
    if (schema_step) {
        CK(hipEventRecord(c->ev[12], s));
        int rc = stc_schema_update(&c->sch, c->leaf_count.p, c->cap, c->cfg.reach_distance,
                                   c->cfg.schema_update_delay, c->temp.p, c->temp.cap, s);
        if (rc) return fail(c, rc, std::string("schema update: ") + g_last_fail);
        CK(hipEventRecord(c->ev[13], s));
    }
    CK(hipEventRecord(c->ev[5], s));
    CK(hipEventSynchronize(c->ev[5]));

    c->stepped = true;
    c->last_t = t;
    if (!c->frun) publish(c); /* with a flush pending, the state is published when its run joins */
    if (out) {
        memset(out, 0, sizeof(*out));
        out->nodes = g.nodes;
        out->leaves = g.leaves;
        uint32_t border = 0;
        for (uint32_t k = 0; k < 64; k++) border += h.border_part[k];
        out->border_points = border;
        out->cross_units = g.units;
        out->candidates = g.candidates;
        out->pairs = g.pairs;
        out->active_meetings = h.live;
        out->expired = h.expired;
        out->kv_count = h.kv_count;
        out->overflow_points = h.overflow;
        out->ambiguous_pairs = g.ambiguous;
        out->ambiguous_nodes = h.ambiguous_nodes;
        out->table_buckets = c->mlen;
        out->table_occupied = h.live + h.expired;
        out->flushed = flushed;
        out->rebuilt = 0;
        hipEventElapsedTime(&out->ms_partition, c->ev[0], c->ev[1]);
        hipEventElapsedTime(&out->ms_filter, c->ev[1], c->ev[2]);
        hipEventElapsedTime(&out->ms_refine, c->ev[2], c->ev[3]);
        hipEventElapsedTime(&out->ms_meetings, c->ev[3], c->ev[4]);
        hipEventElapsedTime(&out->ms_flush, c->ev[4], c->ev[5]);
        hipEventElapsedTime(&out->ms_total, c->ev[0], c->ev[5]);
        hipEventElapsedTime(&out->ms_refine_kernel, c->ev[6], c->ev[7]);
        hipEventElapsedTime(&out->ms_identify_kernel, c->ev[8], c->ev[9]);
        out->continuing = h.continuing;
        out->new_keys = h.new_keys;
        if (g.units || g.nodes) hipEventElapsedTime(&out->ms_count_kernel, c->ev[10], c->ev[11]);
        if (schema_step) hipEventElapsedTime(&out->ms_schema, c->ev[12], c->ev[13]);
        if (flushed && !c->cfg.async_flush) {
            out->ms_flush_done = out->ms_flush;
        } else if (c->fms_pending) {
            out->ms_flush_done = c->flush_done_ms;
            c->fms_pending = false;
        }
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

/* Batched streams: one timestep of many independent contexts. Each context
 * steps on its own streams; up to `threads` host threads drive them, taking
 * the contexts largest first, so that the contexts' kernels and host round
 * trips overlap on the device instead of queueing behind one another. */
int stc_step_batch(void *const *ctxs, uint32_t n, const double *const *xy, const int32_t *on_device, uint32_t t,
                   uint32_t threads, stc_step_out *outs, int32_t *rcs) {
    if ((!ctxs || !xy) && n) return STC_E_ARG;
    for (uint32_t i = 0; i < n; i++) {
        if (!ctxs[i] || !xy[i]) return STC_E_ARG;
        for (uint32_t j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) {
                g_last_fail = "stc_step_batch: a context appears twice";
                return STC_E_ARG;
            }
    }
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return ((StcCtx *)ctxs[a])->n > ((StcCtx *)ctxs[b])->n; });
    std::vector<int> rc(n, STC_OK);
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t k; (k = next.fetch_add(1)) < n;) {
            const uint32_t i = order[k];
            StcCtx *c = (StcCtx *)ctxs[i];
            if (hipSetDevice(c->device) != hipSuccess) {
                rc[i] = fail(c, STC_E_HIP, "stc_step_batch: hipSetDevice failed");
                continue;
            }
            stc_step_out tmp;
            stc_step_out *o = outs ? outs + i : &tmp;
            rc[i] = on_device && on_device[i] ? stc_step_device(c, xy[i], t, o) : stc_step(c, xy[i], t, o);
        }
    };
    const uint32_t nt = std::max<uint32_t>(1, std::min<uint32_t>(threads, n));
    if (nt == 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        pool.reserve(nt);
        for (uint32_t k = 0; k < nt; k++) pool.emplace_back(work);
        for (auto &th : pool) th.join();
    }
    int first = STC_OK;
    for (uint32_t i = 0; i < n; i++) {
        if (rcs) rcs[i] = rc[i];
        if (rc[i] != STC_OK && first == STC_OK) first = rc[i];
    }
    return first;
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
    c->hinit->oversize_oid_count = c->oversize_oid_count;
    CK(hipMemcpyAsync(cnt, c->hinit, sizeof(Counters), hipMemcpyHostToDevice, s));
    CK(hipEventRecord(c->ev[0], s));
    CK(hipEventRecord(c->ev[1], s));
    CK(hipEventRecord(c->ev[2], s));
    if (c->units_in.ensure(nu + 1ull) || c->ukeys.ensure(nu + 1ull) || c->ukeys_alt.ensure(nu + 1ull) ||
        c->uidx.ensure(nu + 1ull) || c->uidx_alt.ensure(nu + 1ull) ||
        c->utemp.ensure(stc_radix_temp_words(std::max<uint32_t>(nu + 1, c->n)) + 64) || c->pbox.ensure(nu + 1ull) ||
        (wait_readers(c, c->cur ^ 1, false), c->meet[c->cur ^ 1].ensure(nu + 1ull)))
        return fail(c, STC_E_CAPACITY, "meeting unit buffers");
    if (nu) {
        CK(hipMemcpyAsync(c->units_in.p, units, (size_t)nu * sizeof(stc_meeting_unit), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_units_decode, dim3(blocks_for(nu, 256)), dim3(256), 0, s, c->units_in.p, nu, c->n,
                           c->ukeys.p, c->uidx.p, cnt);
        CK(hipGetLastError());
        CKRC(stc_radix_sort_u64(c->ukeys.p, c->uidx.p, c->ukeys_alt.p, c->uidx_alt.p, nu, 0,
                                32 + std::max(1, ilog2_ceil(c->n)), c->utemp.p, s));
        hipLaunchKernelGGL(k_units_gather, dim3(blocks_for(nu, 256)), dim3(256), 0, s, c->ukeys.p, c->uidx.p,
                           c->units_in.p, nu, c->pbox.p, cnt);
        CK(hipGetLastError());
    }
    CKRC(sync_counters(c));
    if (c->hcnt->bad_points)
        return fail(c, STC_E_INPUT, "meeting units must name distinct pairs of objects below num_objects "
                                    "by the Cantor key of (min, max)");
    /* the sorted pairs are already in owner segments: only the offsets are needed */
    const int nx = c->cur ^ 1;
    CK(hipEventRecord(c->ev[6], s));
    CK(hipMemsetAsync(c->pcount.p, 0, 4ull * c->n, s));
    if (nu) {
        hipLaunchKernelGGL(k_owner_hist, dim3(blocks_for(nu, 256)), dim3(256), 0, s, (const PairHit *)c->pbox.p, nu,
                           c->pcount.p);
        CK(hipGetLastError());
    }
    CKRC(stc_exclusive_scan_u32(c->pcount.p, c->moff[nx].p, c->n, c->moff[nx].p + c->n, c->utemp.p, s));
    CK(hipEventRecord(c->ev[7], s));
    c->plen = nu;
    c->pairs_boxed = true;
    CK(hipEventRecord(c->ev[3], s));
    StepGeom g{0, 0, 0, 0, 0, 0};
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
    std::vector<PairHit> ph;
    CKRC(step_pairs(c, ph));
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
    std::vector<PairHit> ph;
    CKRC(step_pairs(c, ph));
    std::vector<uint64_t> k(ph.size());
    for (size_t i = 0; i < ph.size(); i++) k[i] = ph[i].key;
    k.erase(std::unique(k.begin(), k.end()), k.end());
    *n = k.size();
    if (keys) memcpy(keys, k.data(), std::min<uint64_t>(cap, k.size()) * 8);
    return STC_OK;
}

int stc_get_active(void *vc, uint64_t *keys, uint32_t *start, uint32_t *end, float *mbr, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    std::vector<Bucket> tb(c->mlen);
    if (c->mlen) CK(hipMemcpy(tb.data(), c->meet[c->cur].p, c->mlen * sizeof(Bucket), hipMemcpyDeviceToHost));
    std::vector<Bucket> live; /* the array is sorted by key; tombstones trail each segment */
    for (const Bucket &b : tb)
        if (bucket_live(b)) live.push_back(b);
    *n = live.size();
    for (uint64_t i = 0; i < std::min<uint64_t>(cap, live.size()); i++) {
        if (keys) keys[i] = live[i].key;
        if (start) start[i] = live[i].start;
        if (end) end[i] = c->last_t; /* every live meeting was seen by the last step */
        if (mbr) for (int q = 0; q < 4; q++) mbr[4 * i + q] = o2f(live[i].mbr[q]);
    }
    return STC_OK;
}

int stc_get_records(void *vc, uint64_t *keys, float *boxes, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    JOIN_FLUSH(c);
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
    /* the leaves the last step used (a dynamic schema has been updated since,
     * but a leaf's slot keeps its box until a later step reuses it) */
    const uint32_t nl = c->last_leaves;
    std::vector<uint32_t> ln(nl), lc(nl);
    if (nl) {
        CK(hipMemcpy(ln.data(), c->leaf_node.p, nl * 4ull, hipMemcpyDeviceToHost));
        CK(hipMemcpy(lc.data(), c->leaf_count.p, nl * 4ull, hipMemcpyDeviceToHost));
    }
    struct L { double lo1, lo0, hi0, hi1; uint32_t cnt; };
    std::vector<L> ls(nl);
    const Node *tn = tree_nodes(c);
    for (uint32_t l = 0; l < nl; l++) {
        Node x;
        CK(hipMemcpy(&x, tn + ln[l], sizeof(Node), hipMemcpyDeviceToHost));
        ls[l] = L{x.lo1, x.lo0, x.hi0, x.hi1, lc[l]};
    }
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

int stc_build_schema(void *vc, const double *xy, int32_t xy_on_device) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !xy) return STC_E_ARG;
    if (!c->dyn) return fail(c, STC_E_STATE, "the context rebuilds its QTree every step (dynamic_schema = 0)");
    const double2 *src = (const double2 *)xy;
    if (!xy_on_device) {
        CK(hipMemcpyAsync(c->xy_in.p, xy, (size_t)c->n * 16, hipMemcpyHostToDevice, c->s));
        src = c->xy_in.p;
    }
    /* the points' Morton order keeps each wave of the build's counting on few nodes */
    hipLaunchKernelGGL(k_prep, dim3(blocks_for(c->n, 256)), dim3(256), 0, c->s, src, c->n, c->map, c->codes.p,
                       c->idx.p, c->dlon.p, (float2 *)nullptr, c->dcnt); /* posf keeps the last step's positions */
    CK(hipGetLastError());
    CKRC(stc_radix_sort_u32(c->codes.p, c->idx.p, c->codes_alt.p, c->idx_alt.p, c->n, 28, c->temp.p, c->s));
    int rc = stc_schema_build(&c->sch, src, c->idx.p, c->n, c->map, c->cap, c->cfg.reach_distance, c->temp.p,
                              c->temp.cap, c->s);
    if (rc) return fail(c, rc, rc == STC_E_ARG ? "root does not split: too few objects for grid_capacity"
                                               : std::string("schema build: ") + g_last_fail);
    CK(hipStreamSynchronize(c->s));
    return STC_OK;
}

int stc_set_schema(void *vc, const stc_qtnode *nodes, uint32_t n_nodes) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return STC_E_ARG;
    if (!c->dyn) return fail(c, STC_E_STATE, "the context rebuilds its QTree every step (dynamic_schema = 0)");
    int rc = stc_schema_set(&c->sch, nodes, n_nodes, c->s);
    if (rc == STC_E_ARG) return fail(c, rc, "not a QTree below node 0");
    if (rc) return fail(c, rc, std::string("schema upload: ") + g_last_fail);
    if (c->temp.ensure(stc_radix_temp_words(std::max<uint32_t>(c->n, c->sch.cap)) + 64))
        return fail(c, STC_E_CAPACITY, "scan scratch");
    return STC_OK;
}

int stc_get_schema(void *vc, stc_qtnode *nodes, uint64_t cap, uint64_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    if (!c->dyn) return fail(c, STC_E_STATE, "the context rebuilds its QTree every step (dynamic_schema = 0)");
    int rc = stc_schema_get(&c->sch, nodes, cap, n, c->s);
    if (rc == STC_E_STATE) return fail(c, rc, "no schema yet");
    if (rc) return fail(c, rc, std::string("schema download: ") + g_last_fail);
    return STC_OK;
}

int stc_num_runs(void *vc, uint32_t *n) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !n) return STC_E_ARG;
    JOIN_FLUSH(c);
    *n = (uint32_t)c->runs.size();
    return STC_OK;
}

int stc_get_run_info(void *vc, uint32_t run, stc_run_info *info) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || !info) return STC_E_ARG;
    JOIN_FLUSH(c);
    if (run >= c->runs.size()) return STC_E_ARG;
    *info = c->runs[run].info;
    return STC_OK;
}

int stc_get_run(void *vc, uint32_t run, stc_ctf_meta *ctfs, uint8_t *key_bytes, uint8_t *bitmaps, uint16_t *sids,
                uint64_t *o_keys, float *o_boxes, uint8_t *o_bitmap, uint64_t *sorted_keys, float *sorted_boxes) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return STC_E_ARG;
    JOIN_FLUSH(c);
    if (run >= c->runs.size()) return STC_E_ARG;
    RunHost &R = c->runs[run];
    if (queue_run_copies(c)) return fail(c, STC_E_HIP, std::string("run copies: ") + g_last_fail);
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
    JOIN_FLUSH(c);
    if (queue_run_copies(c)) return fail(c, STC_E_HIP, std::string("run copies: ") + g_last_fail);
    CK(hipStreamSynchronize(c->s));
    if (c->flushw.cs) CK(hipStreamSynchronize(c->flushw.cs));
    return STC_OK;
}

/* One query batch over the state the last completed step published: admission by free device
 * memory, a lease on that state's meeting array and memtable epoch for the batch's duration, the
 * batch on the query stream. `make` (when given) builds the batch from the published step instead
 * of `q`; returning false answers with no hits. */
extern "C++" {
template <class Make>
static int query_published(StcCtx *c, const stc_query *q, uint32_t nq, stc_hit *hits, uint64_t cap, uint64_t *n_hits,
                           uint64_t *counts, uint32_t *state_step, Make make) {
    std::lock_guard<std::mutex> ql(c->q_mu);
    /* admission: what the batch's query list, counters and hit buffer add to the query buffers */
    const uint64_t need = stc_query_bufs_growth(c->qb, nq, hits ? cap : 0);
    size_t free_b = 0, total_b = 0;
    if (need && hipMemGetInfo(&free_b, &total_b) == hipSuccess && need + STC_QUERY_RESERVE > free_b)
        return fail(c, STC_E_CAPACITY, "admission: the batch needs " + std::to_string(need) + " bytes of device memory, " +
                                           std::to_string(free_b) + " free");
    QueryState st;
    int mi;
    uint32_t ep;
    {
        std::lock_guard<std::mutex> lk(c->pub_mu);
        if (!c->pub_valid) return fail(c, STC_E_STATE, "no step processed");
        st = c->pub;
        mi = c->pub_meet;
        ep = c->pub_epoch;
        c->readers_meet[mi]++;
        c->readers_epoch[ep & 1]++;
        c->readers_total++;
    }
    std::vector<stc_query> built;
    bool run = true;
    if constexpr (!std::is_same<Make, std::nullptr_t>::value) {
        run = make(st.last_t, built);
        q = built.data();
        nq = (uint32_t)built.size();
    }
    int rc = 0;
    if (run) {
        if (!c->qb) c->qb = stc_query_bufs_new();
        rc = stc_query_run(c->qb, st, q, nq, hits, cap, n_hits, counts, c->qs);
    } else {
        *n_hits = 0;
    }
    {
        std::lock_guard<std::mutex> lk(c->pub_mu);
        c->readers_meet[mi]--;
        c->readers_epoch[ep & 1]--;
        c->readers_total--;
    }
    c->pub_cv.notify_all();
    if (rc) return fail(c, rc, std::string("query: ") + g_last_fail);
    if (state_step) *state_step = st.last_t;
    return STC_OK;
}
} /* extern "C++" */

int stc_live_search(void *vc, const uint32_t *pids, uint32_t n, stc_hit *hits, uint64_t cap, uint64_t *n_hits) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || (!pids && n) || !n_hits) return STC_E_ARG;
    /* live at t and established (t - start >= min_meet + 1): start <= t - min_meet - 1 and end >= t,
     * the overlap predicate of an object query with t0 = t, t1 = t - min_meet - 1, where t is the
     * step of the published state the batch answers */
    const uint32_t m = c->cfg.min_meet_time + 1;
    return query_published(c, nullptr, n, hits, cap, n_hits, nullptr, nullptr, [&](uint32_t t, std::vector<stc_query> &q) {
        if (t < m) return false;
        q.resize(n);
        for (uint32_t i = 0; i < n; i++) {
            memset(&q[i], 0, sizeof(stc_query));
            q[i].kind = STC_Q_OID;
            q[i].oid = pids[i];
            q[i].t0 = t;
            q[i].t1 = t - m;
        }
        return true;
    });
}

int stc_live_search_noalloc(void *vc, const uint32_t *pids, uint32_t n, stc_hit *hits, uint64_t cap,
                            uint64_t *n_hits) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || (!pids && n) || !n_hits || (!hits && cap)) return STC_E_ARG;
    if (n > NA_MAX_PIDS) return fail(c, STC_E_CAPACITY, "stc_live_search_noalloc: more than 16384 objects per call");
    if (!c->stepped) return fail(c, STC_E_STATE, "no step processed");
    uint8_t *sc = c->device >= 0 && c->device < 64 ? g_na_ptr[c->device] : nullptr;
    if (!sc) return fail(c, STC_E_STATE, "scratch not set up on this device");
    const uint32_t t = c->last_t, m = c->cfg.min_meet_time + 1;
    *n_hits = 0;
    if (t < m || n == 0) return STC_OK;
    hipStream_t s = c->s;
    CK(hipMemcpyAsync(sc + NA_HITS_AT, pids, (size_t)n * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_na_prep, dim3(1), dim3(1024), 0, s, sc, n);
    CK(hipGetLastError());
    if (c->mlen) {
        const uint32_t g = (uint32_t)std::min<uint64_t>(blocks_for(c->mlen, 256), 8192);
        hipLaunchKernelGGL(k_na_live, dim3(g), dim3(256), 0, s, (const Bucket *)c->meet[c->cur].p, c->mlen, t, t - m,
                           sc, n);
        CK(hipGetLastError());
    }
    unsigned long long total = 0;
    CK(hipMemcpyAsync(&total, sc, 8, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    if (total > NA_MAX_HITS) {
        *n_hits = total;
        return fail(c, STC_E_CAPACITY, "stc_live_search_noalloc: more than 262144 hits");
    }
    const uint64_t stored = std::min<uint64_t>(total, cap);
    if (stored) {
        CK(hipMemcpyAsync(hits, sc + NA_HITS_AT, stored * sizeof(stc_hit), hipMemcpyDeviceToHost, s));
        CK(hipStreamSynchronize(s));
        std::sort(hits, hits + stored, [](const stc_hit &a, const stc_hit &b) {
            if (a.query != b.query) return a.query < b.query;
            if (a.pid != b.pid) return a.pid < b.pid;
            if (a.target != b.target) return a.target < b.target;
            if (a.start != b.start) return a.start < b.start;
            return a.end < b.end;
        });
    }
    *n_hits = total;
    return STC_OK;
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
    if (!c || !dir || !prefix) return STC_E_ARG;
    JOIN_FLUSH(c);
    if (run >= c->runs.size()) return STC_E_ARG;
    RunHost &R = c->runs[run];
    if (queue_run_copies(c)) return fail(c, STC_E_HIP, std::string("run copies: ") + g_last_fail);
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

int stc_ctf_widths(stc_ctf_meta *ctf, const stc_box *map_mbr, uint32_t num_objects, uint32_t max_meet_time) {
    if (!ctf || !map_mbr || max_meet_time == 0) return STC_E_ARG;
    const double map[4] = {map_mbr->low[0], map_mbr->low[1], map_mbr->high[0], map_mbr->high[1]};
    stc_ctf_bits(ctf, map, num_objects, max_meet_time);
    return STC_OK;
}

/* Releases the host copies of all flushed runs. A run's copy to host memory
 * still completes (a copy not queued yet is queued with the next step, as it
 * would have been) before its pinned block is reused. */
int stc_drop_runs(void *vc) {
    StcCtx *c = (StcCtx *)vc;
    if (!c) return STC_E_ARG;
    if (c->frun) { /* the runs go when the pending flush joins, its own with them */
        c->drop_pending = true;
        return STC_OK;
    }
    drop_all_runs(c);
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
    JOIN_FLUSH(c);
    wait_all_readers(c); /* the merged runs' blocks are freed and the memtable buffers reused */
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
    settle_deferred(c);
    if (rc) {
        release_block(c, fo.blk);
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
        release_block(c, run.blk);
        if (run.dev) (void)hipFree(run.dev);
        return rc;
    }
    CK(hipStreamSynchronize(s));
    for (uint32_t r = first; r < first + count; r++) {
        release_block(c, c->runs[r].blk);
        (void)hipFree(c->runs[r].dev);
    }
    c->runs.erase(c->runs.begin() + first, c->runs.begin() + first + count);
    c->runs.insert(c->runs.begin() + first, std::move(run));
    publish(c);
    return STC_OK;
}

int stc_query_batch_at(void *vc, const stc_query *q, uint32_t nq, stc_hit *hits, uint64_t cap, uint64_t *n_hits,
                       uint64_t *counts, uint32_t *state_step) {
    StcCtx *c = (StcCtx *)vc;
    if (!c || (!q && nq) || !n_hits) return STC_E_ARG;
    for (uint32_t i = 0; i < nq; i++)
        if (q[i].kind < STC_Q_OID || q[i].kind > STC_Q_PAIR) return fail(c, STC_E_ARG, "unknown query kind");
    return query_published(c, q, nq, hits, cap, n_hits, counts, state_step, nullptr);
}

int stc_query_batch(void *vc, const stc_query *q, uint32_t nq, stc_hit *hits, uint64_t cap, uint64_t *n_hits,
                    uint64_t *counts) {
    return stc_query_batch_at(vc, q, nq, hits, cap, n_hits, counts, nullptr);
}

} /* extern "C" */
