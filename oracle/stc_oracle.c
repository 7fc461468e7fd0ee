/*
 * stc_oracle.c — CPU oracle (test infrastructure; see stc_oracle.h).
 *
 * Single-threaded, deterministic restatement of the reference GPU hot path.
 * Every stage cites the reference lines it follows. Data structures differ
 * from the reference (sorted arrays and merges instead of atomics and hash
 * probing) so that the oracle checks the product rather than mirroring it.
 *
 * Parity decisions taken here (each is also a test in tests/):
 *  P1 distance formula: dx scaled by 1/dpm_lon(lat of the probe), dy by
 *     1/dpm_lat, sqrt(dx*dx+dy*dy), no FMA contraction, glibc sin
 *     (src/cuda/cuda_util.cuh:41-53, src/geometry/geometry.h:40-48,
 *     src/util/util.h:59-66). Built with -ffp-contract=off.
 *  P2 QTree: the per-step rebuild of --disable_dynamic_schema
 *     (src/cuda/geometry.cu:1267-1393) restated over sparse, sorted cell
 *     counts instead of a dense 16384^2 histogram. Node boxes use the
 *     reference's exact expressions.
 *  P3 leaf overflow (src/cuda/geometry.cu:255): the reference keeps whichever
 *     points win an atomic race; the oracle keeps the lowest point ids and
 *     counts the overflow.
 *  P4 meetings: a meeting is a maximal run of steps in which the pair is
 *     found. The reference's 5-probe drop (geometry.cu:471) and its bucket
 *     reset on expiry (geometry.cu:736), which can break linear-probe chains
 *     and split a live meeting depending on thread order, are not reproduced.
 *  P5 MBR = float32 rounding of the min/max over both objects' positions over
 *     all steps of the meeting: the order-independent value of the racy
 *     updates at geometry.cu:474-490.
 *  P6 flush arithmetic follows geometry.cu:739-1107 and step_merge.cpp:132-186
 *     exactly, including the STR key that uses high[1] twice (geometry.cu:764)
 *     and the sid bits carried into packed keys (cuda_util.cuh:168).
 *     compat bit 0 reproduces the reductions that drop lanes 32-63 of each
 *     64-thread group (geometry.cu:865-888, 955-974); otherwise all records
 *     are reduced. Out-of-range float->int conversions saturate as on the
 *     reference GPU; host-side double->uint conversions in get_ctf_bits
 *     follow x86-64 cvttsd2si semantics. Bitmap writes that would leave the
 *     bitmap array are dropped (the reference writes out of bounds).
 */
#include "stc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <string.h>

typedef unsigned __int128 u128;

#define ONE_DIM 16384u
#define NLEVELS 15
#define OID_BIT 26
#define DURATION_BIT 32
#define END_BIT 32
#define SID_SHIFT (OID_BIT * 2 + DURATION_BIT + END_BIT) /* 116 */
#define BITMAP_EDGE 256u
#define BIT_COUNT (BITMAP_EDGE * BITMAP_EDGE)
#define BITMAP_BYTES (BIT_COUNT / 8)

static const double PI_REF = 3.14159265;                       /* util.h:38 */
static const double DPM_LAT = 360.0 / (40076.0 * 1000.0);      /* util.h:57 */

enum { T_INVALID = 0, T_BRANCH = 1, T_LEAF = 2 };

typedef struct node {
    double lo[2], hi[2];
    double mid_x, mid_y;
    uint32_t type;
    uint32_t child[4];
    uint32_t leaf;        /* leaf id (creation order) for leaves */
    uint32_t ovf, unf;    /* overflow_count, underflow_count of a dynamic schema */
} node;

typedef struct meeting {
    uint64_t key;         /* oid1 << 32 | oid2, oid1 < oid2 */
    uint32_t start, end;
    float mbr[4];
} meeting;

typedef struct pairhit {
    uint64_t key;
    float mbr[4];
} pairhit;

typedef struct record {
    u128 key;
    float box[4];
} record;

typedef struct run {
    orc_run_info info;
    orc_ctf *ctfs;
    uint8_t *key_bytes;
    uint8_t *bitmaps;
    uint16_t *sids;
    u128 *o_keys;
    float *o_boxes;
    uint8_t o_bitmap[BITMAP_BYTES];
    u128 *sorted_keys;
    float *sorted_boxes;
} run;

typedef struct orc {
    orc_config cfg;
    double x_buffer, y_buffer;
    char err[256];
    double last_flush_ms;           /* wall time of the last flush inside a step (a CPU flush baseline) */
    uint64_t last_flush_records;
    uint64_t counts[ORC_C_NUM];

    /* per step (with a dynamic schema the nodes persist; free slots are listed) */
    node *nodes; uint32_t n_nodes, cap_nodes;
    uint32_t *free_nodes; uint32_t n_free, cap_free;
    int have_schema;
    uint32_t n_leaves;
    uint32_t *leaf_node;        /* leaf id -> node index */
    uint32_t *pt_leaf;          /* point -> leaf id */
    uint8_t *pt_border;
    double *dlon;               /* per point dpm_lon(y) */
    uint32_t *leaf_start, *leaf_count, *leaf_members; /* CSR by leaf id, members sorted by pid */
    uint32_t *unit_pid, *unit_node; uint8_t *unit_inside; uint64_t n_units, cap_units;
    pairhit *pairs; uint64_t n_pairs, cap_pairs;

    /* meeting state */
    meeting *active; uint64_t n_active, cap_active;
    uint32_t end_time_min;

    /* memtable */
    record *recs; uint64_t n_recs, cap_recs;
    uint32_t *same_pid_count;
    uint32_t (*o_boxs)[4];      /* quantized per-object MBR: low0, low1, high0, high1 */
    uint16_t *sids;
    uint32_t sid_count, oversize_oid_count;
    int sticky;                 /* compaction flush: objects entering with sids 1 stay oversize */
    uint32_t last_t;            /* the last step */

    run *runs; uint32_t n_runs;
} orc;

/* ------------------------------------------------------------------ */
/* geometry (P1)                                                       */
/* ------------------------------------------------------------------ */

/* device form of degree_per_meter_longitude (cuda_util.cuh:41-44) */
double orc_dpm_lon(double lat) {
    return 360.0 / (sin((90 - fabs(lat)) * PI_REF / 180) * 40076.0 * 1000.0);
}

static inline double pt_dist(double x1, double y1, double x2, double y2, double dlon1) {
    double dx = x1 - x2;                 /* cuda_util.cuh:46-53 */
    double dy = y1 - y2;
    dx = dx / dlon1;
    dy = dy / DPM_LAT;
    return sqrt(dx * dx + dy * dy);
}

double orc_point_distance(double x1, double y1, double x2, double y2) {
    return pt_dist(x1, y1, x2, y2, orc_dpm_lon(y1));
}

static inline double box_dist(const double *lo, const double *hi, double px, double py, double dlon) {
    /* geometry.cu:33-42 */
    double dx = fmax(fabs(px - (lo[0] + hi[0]) / 2) - (hi[0] - lo[0]) / 2, 0.0);
    double dy = fmax(fabs(py - (lo[1] + hi[1]) / 2) - (hi[1] - lo[1]) / 2, 0.0);
    dy = dy / DPM_LAT;
    dx = dx / dlon;
    return sqrt(dx * dx + dy * dy);
}

double orc_box_distance(const double *b, double px, double py) {
    double lo[2] = {b[0], b[1]}, hi[2] = {b[2], b[3]};
    return box_dist(lo, hi, px, py, orc_dpm_lon(py));
}

/* ------------------------------------------------------------------ */
/* conversions                                                         */
/* ------------------------------------------------------------------ */

/* float/double -> unsigned as the reference GPU converts (saturating, NaN -> 0) */
static inline uint32_t sat_u32(double v) {
    if (!(v > 0)) return 0;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}
static inline uint64_t sat_u64(double v) {
    if (!(v > 0)) return 0;
    if (v >= 18446744073709551615.0) return ~0ull;
    return (uint64_t)v;
}
/* host double -> uint32 on x86-64 (cvttsd2si into a 64-bit register) */
static inline uint32_t x86_d2u32(double v) {
    if (!(v > -9223372036854775808.0 && v < 9223372036854775808.0)) return 0;
    return (uint32_t)(uint64_t)(int64_t)v;
}
/* host double -> int32 on x86-64 (cvttsd2si 32-bit; out of range -> INT_MIN) */
static inline int32_t x86_d2i32(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return (int32_t)0x80000000u;
    return (int32_t)v;
}

/* float_to_uint / uint_to_float (cuda_util.cuh:115-125) */
static inline uint32_t f2q(float xy) {
    xy += 180;
    return sat_u32((float)(xy * 100000));
}
static inline float q2f(uint32_t f) {
    return (float)f / 100000 - 180;
}

/* min_bits_to_store (util.h:527-533) */
uint32_t orc_min_bits(uint32_t v) {
    if (v == 0) return 1;
    return (uint32_t)log2((double)v) + 1;
}

/* ------------------------------------------------------------------ */
/* small utilities                                                     */
/* ------------------------------------------------------------------ */

static int grow(void **p, uint64_t *cap, uint64_t need, size_t elem) {
    if (need <= *cap) return 0;
    uint64_t nc = *cap ? *cap : 1024;
    while (nc < need) nc *= 2;
    void *q = realloc(*p, nc * elem);
    if (!q) return -1;
    *p = q;
    *cap = nc;
    return 0;
}

static void set_err(orc *o, const char *msg) {
    snprintf(o->err, sizeof(o->err), "%s", msg);
}

static uint32_t part1by1(uint32_t x) {
    x &= 0x0000ffff;
    x = (x | (x << 8)) & 0x00FF00FF;
    x = (x | (x << 4)) & 0x0F0F0F0F;
    x = (x | (x << 2)) & 0x33333333;
    x = (x | (x << 1)) & 0x55555555;
    return x;
}
static uint32_t morton(uint32_t x, uint32_t y) { return part1by1(x) | (part1by1(y) << 1); }
static uint32_t compact1by1(uint32_t x) {
    x &= 0x55555555;
    x = (x | (x >> 1)) & 0x33333333;
    x = (x | (x >> 2)) & 0x0F0F0F0F;
    x = (x | (x >> 4)) & 0x00FF00FF;
    x = (x | (x >> 8)) & 0x0000FFFF;
    return x;
}

static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}
static int cmp_pairhit(const void *a, const void *b) {
    uint64_t x = ((const pairhit *)a)->key, y = ((const pairhit *)b)->key;
    return x < y ? -1 : x > y;
}
static int cmp_record(const void *a, const void *b) {
    u128 x = ((const record *)a)->key, y = ((const record *)b)->key;
    return x < y ? -1 : x > y;
}

/* ------------------------------------------------------------------ */
/* lifecycle                                                           */
/* ------------------------------------------------------------------ */

void *orc_create(const orc_config *cfg) {
    orc *o = (orc *)calloc(1, sizeof(orc));
    if (!o) return NULL;
    o->cfg = *cfg;
    /* build_schema sets the border buffers from the map's lowest latitude
     * using the host formula (src/tracing/filter.cpp:30-31, util.h:59-66) */
    double absla = fabs(cfg->map_lo[1]);
    if (absla == 90) absla = 89;
    o->x_buffer = cfg->reach_distance * (360.0 / (sin((90 - absla) * PI_REF / 180) * 40076.0 * 1000.0));
    o->y_buffer = cfg->reach_distance * DPM_LAT;
    o->end_time_min = cfg->start_time + cfg->min_meet_time; /* process.cpp:404 */
    uint32_t n = cfg->num_objects;
    o->pt_leaf = (uint32_t *)calloc(n, sizeof(uint32_t));
    o->pt_border = (uint8_t *)calloc(n, 1);
    o->dlon = (double *)calloc(n, sizeof(double));
    o->leaf_members = (uint32_t *)calloc(n, sizeof(uint32_t));
    o->same_pid_count = (uint32_t *)calloc(n, sizeof(uint32_t));
    o->o_boxs = calloc(n, sizeof(*o->o_boxs));
    o->sids = (uint16_t *)calloc(n, sizeof(uint16_t));
    for (uint32_t i = 0; i < n; i++) { /* cuda_init_o_boxs (geometry.cu:1253-1263) */
        o->o_boxs[i][0] = o->o_boxs[i][1] = 100000000u;
        o->o_boxs[i][2] = o->o_boxs[i][3] = 0;
    }
    return o;
}

static void free_run(run *r) {
    free(r->ctfs); free(r->key_bytes); free(r->bitmaps); free(r->sids);
    free(r->o_keys); free(r->o_boxes); free(r->sorted_keys); free(r->sorted_boxes);
}

void orc_destroy(void *vo) {
    orc *o = (orc *)vo;
    if (!o) return;
    free(o->nodes); free(o->free_nodes); free(o->leaf_node); free(o->pt_leaf); free(o->pt_border); free(o->dlon);
    free(o->leaf_start); free(o->leaf_count); free(o->leaf_members);
    free(o->unit_pid); free(o->unit_node); free(o->unit_inside); free(o->pairs);
    free(o->active); free(o->recs); free(o->same_pid_count); free(o->o_boxs); free(o->sids);
    for (uint32_t i = 0; i < o->n_runs; i++) free_run(&o->runs[i]);
    free(o->runs);
    free(o);
}

const char *orc_last_error(void *o) { return ((orc *)o)->err; }
void orc_counts(void *o, uint64_t *out) { memcpy(out, ((orc *)o)->counts, sizeof(uint64_t) * ORC_C_NUM); }

/* ------------------------------------------------------------------ */
/* QTree rebuild (P2): geometry.cu:1267-1393                            */
/* ------------------------------------------------------------------ */

static uint32_t new_node(orc *o) {
    if (o->n_nodes == o->cap_nodes) {
        o->cap_nodes = o->cap_nodes ? o->cap_nodes * 2 : 4096;
        o->nodes = (node *)realloc(o->nodes, sizeof(node) * o->cap_nodes);
    }
    memset(&o->nodes[o->n_nodes], 0, sizeof(node));
    return o->n_nodes++;
}

/* box of block (x, y) at a level with xdim blocks per side (geometry.cu:1306-1311) */
static void block_box(const orc *o, node *nd, uint32_t x, uint32_t y, uint32_t xdim) {
    const double *lo = o->cfg.map_lo, *hi = o->cfg.map_hi;
    nd->lo[0] = lo[0] + (double)x * (hi[0] - lo[0]) / (double)xdim;
    nd->lo[1] = lo[1] + (double)y * (hi[1] - lo[1]) / (double)xdim;
    nd->hi[0] = lo[0] + (double)(x + 1) * (hi[0] - lo[0]) / (double)xdim;
    nd->hi[1] = lo[1] + (double)(y + 1) * (hi[1] - lo[1]) / (double)xdim;
    nd->mid_x = (nd->lo[0] + nd->hi[0]) / 2;
    nd->mid_y = (nd->lo[1] + nd->hi[1]) / 2;
}

/* leaf child i created by halving its parent (geometry.cu:1318-1334) */
static uint32_t halving_leaf(orc *o, uint32_t parent, int i) {
    uint32_t c = new_node(o);
    node *nd = &o->nodes[c];
    const node *p = &o->nodes[parent];
    double xhalf = p->mid_x - p->lo[0];
    double yhalf = p->mid_y - p->lo[1];
    nd->type = T_LEAF;
    nd->leaf = o->n_leaves++;
    nd->lo[0] = p->lo[0] + (double)(i % 2 == 1) * xhalf;
    nd->lo[1] = p->lo[1] + (double)(i / 2 == 1) * yhalf;
    nd->hi[0] = p->mid_x + (double)(i % 2 == 1) * xhalf;
    nd->hi[1] = p->mid_y + (double)(i / 2 == 1) * yhalf;
    nd->mid_x = (nd->lo[0] + nd->hi[0]) / 2;
    nd->mid_y = (nd->lo[1] + nd->hi[1]) / 2;
    return c;
}

typedef struct blk { uint32_t code, count, nodeid; } blk;

static int64_t find_blk(const blk *b, uint32_t n, uint32_t code) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if (b[mid].code < code) lo = mid + 1; else hi = mid;
    }
    return (lo < n && b[lo].code == code) ? (int64_t)lo : -1;
}

static int build_tree(orc *o, const double *xy) {
    uint32_t n = o->cfg.num_objects, cap = o->cfg.grid_capacity;
    const double *lo = o->cfg.map_lo, *hi = o->cfg.map_hi;
    o->n_nodes = 0;
    o->n_leaves = 0;
    new_node(o); /* node 0 is the root (cuda_clean_cells sets the stack index to 1) */

    /* cell histogram (cuda_build_qtree, geometry.cu:1270-1279), kept sparse */
    uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) {
        double fx = (xy[2 * i] - lo[0]) / (hi[0] - lo[0]) * ONE_DIM;
        double fy = (xy[2 * i + 1] - lo[1]) / (hi[1] - lo[1]) * ONE_DIM;
        if (!(fx > -1.0 && fx < (double)ONE_DIM && fy > -1.0 && fy < (double)ONE_DIM)) {
            free(codes);
            set_err(o, "point outside the map box");
            return -1;
        }
        codes[i] = morton((uint32_t)fx, (uint32_t)fy);
    }
    qsort(codes, n, sizeof(uint32_t), cmp_u32);

    blk *levels[NLEVELS];
    uint32_t nlev[NLEVELS];
    blk *b0 = (blk *)malloc(sizeof(blk) * (n ? n : 1));
    uint32_t nb = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (nb && b0[nb - 1].code == codes[i]) b0[nb - 1].count++;
        else { b0[nb].code = codes[i]; b0[nb].count = 1; b0[nb].nodeid = 0; nb++; }
    }
    free(codes);
    levels[0] = b0; nlev[0] = nb;
    for (int L = 1; L < NLEVELS; L++) {
        blk *prev = levels[L - 1];
        blk *cur = (blk *)malloc(sizeof(blk) * (nlev[L - 1] ? nlev[L - 1] : 1));
        uint32_t m = 0;
        for (uint32_t i = 0; i < nlev[L - 1]; i++) {
            uint32_t c = prev[i].code >> 2;
            if (m && cur[m - 1].code == c) cur[m - 1].count += prev[i].count;
            else { cur[m].code = c; cur[m].count = prev[i].count; cur[m].nodeid = 0; m++; }
        }
        levels[L] = cur; nlev[L] = m;
    }
    if (nlev[NLEVELS - 1] != 1 || levels[NLEVELS - 1][0].count <= cap) {
        for (int L = 0; L < NLEVELS; L++) free(levels[L]);
        set_err(o, "root does not split: too few objects for grid_capacity");
        return -1;
    }

    /* cuda_merge_qtree with gap = 1 << L (geometry.cu:1295-1393) */
    for (int L = 0; L < NLEVELS; L++) {
        uint32_t xdim = ONE_DIM >> L;
        for (uint32_t i = 0; i < nlev[L]; i++) {
            blk *bk = &levels[L][i];
            if (bk->count <= cap) continue;
            uint32_t x = compact1by1(bk->code), y = compact1by1(bk->code >> 1);
            uint32_t nd = (L == NLEVELS - 1) ? 0 : new_node(o);
            o->nodes[nd].type = T_BRANCH;
            block_box(o, &o->nodes[nd], x, y, xdim);
            for (int c = 0; c < 4; c++) {
                uint32_t child;
                if (L == 0) {
                    child = halving_leaf(o, nd, c);
                } else {
                    uint32_t sub = (bk->code << 2) | (uint32_t)((c & 1) | ((c >> 1) << 1));
                    int64_t k = find_blk(levels[L - 1], nlev[L - 1], sub);
                    if (k >= 0 && levels[L - 1][k].count > cap) child = levels[L - 1][k].nodeid;
                    else child = halving_leaf(o, nd, c);
                }
                o->nodes[nd].child[c] = child;
            }
            bk->nodeid = nd;
        }
    }
    for (int L = 0; L < NLEVELS; L++) free(levels[L]);

    free(o->leaf_node);
    o->leaf_node = (uint32_t *)malloc(sizeof(uint32_t) * (o->n_leaves ? o->n_leaves : 1));
    for (uint32_t i = 0; i < o->n_nodes; i++)
        if (o->nodes[i].type == T_LEAF) o->leaf_node[o->nodes[i].leaf] = i;
    return 0;
}

/* ------------------------------------------------------------------ */
/* dynamic schema (schema_mode 1, 2)                                    */
/* ------------------------------------------------------------------ */

/* host form of degree_per_meter_longitude (util.h:59-66) */
static double host_dpm_lon(double lat) {
    double absla = fabs(lat);
    if (absla == 90) absla = 89;
    return 360.0 / (sin((90 - absla) * PI_REF / 180) * 40076.0 * 1000.0);
}

static int take_node(orc *o, uint32_t *id) {
    if (o->n_free) { *id = o->free_nodes[--o->n_free]; return 0; }
    if (o->n_nodes == o->cap_nodes) {
        uint32_t nc = o->cap_nodes ? o->cap_nodes * 2 : 4096;
        node *nn = (node *)realloc(o->nodes, sizeof(node) * nc);
        if (!nn) return -1;
        o->nodes = nn;
        o->cap_nodes = nc;
    }
    *id = o->n_nodes++;
    return 0;
}

/* QTNode::split of the initial tree (QTree.h:157-185): a node holding at
 * least grid_capacity points whose half width in metres (box::width(true),
 * geometry.h:166-172, at its low latitude) exceeds reach_distance splits into
 * the exact halves of its box, and so on below. Points reach a node by the
 * descent rule (y > mid_y) * 2 + (x > mid_x), so the tree does not depend on
 * the insertion order of partitioner::build_schema's threads; it is built
 * here recursively over lists of point ids. */
static int schema_build_rec(orc *o, uint32_t nd, const double *xy, uint32_t *ids, uint32_t cnt, uint32_t *tmp,
                            int depth) {
    node b = o->nodes[nd];
    double width_m = (b.hi[0] - b.lo[0]) / host_dpm_lon(b.lo[1]);
    if (!(cnt >= o->cfg.grid_capacity && width_m / 2 > o->cfg.reach_distance)) return 0;
    if (depth > 60) { set_err(o, "initial schema deeper than 60 levels"); return -1; }
    const double lo0[4] = {b.lo[0], b.mid_x, b.lo[0], b.mid_x}, lo1[4] = {b.lo[1], b.lo[1], b.mid_y, b.mid_y};
    const double hi0[4] = {b.mid_x, b.hi[0], b.mid_x, b.hi[0]}, hi1[4] = {b.mid_y, b.mid_y, b.hi[1], b.hi[1]};
    uint32_t kids[4];
    for (int c = 0; c < 4; c++) {
        if (take_node(o, &kids[c])) { set_err(o, "out of memory"); return -1; }
        node *k = &o->nodes[kids[c]];
        memset(k, 0, sizeof(node));
        k->lo[0] = lo0[c]; k->lo[1] = lo1[c]; k->hi[0] = hi0[c]; k->hi[1] = hi1[c];
        k->mid_x = (k->lo[0] + k->hi[0]) / 2;
        k->mid_y = (k->lo[1] + k->hi[1]) / 2;
        k->type = T_LEAF;
    }
    o->nodes[nd].type = T_BRANCH;
    for (int c = 0; c < 4; c++) o->nodes[nd].child[c] = kids[c];
    /* distribute the ids by quadrant (stable) */
    uint32_t n4[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < cnt; i++) {
        uint32_t p = ids[i];
        n4[(xy[2 * p + 1] > b.mid_y) * 2 + (xy[2 * p] > b.mid_x)]++;
    }
    uint32_t at[4] = {0, n4[0], n4[0] + n4[1], n4[0] + n4[1] + n4[2]};
    uint32_t st[4] = {at[0], at[1], at[2], at[3]};
    for (uint32_t i = 0; i < cnt; i++) {
        uint32_t p = ids[i];
        tmp[at[(xy[2 * p + 1] > b.mid_y) * 2 + (xy[2 * p] > b.mid_x)]++] = p;
    }
    memcpy(ids, tmp, sizeof(uint32_t) * cnt);
    for (int c = 0; c < 4; c++)
        if (schema_build_rec(o, kids[c], xy, ids + st[c], n4[c], tmp + st[c], depth + 1)) return -1;
    return 0;
}

/* partitioner::build_schema (filter.cpp:27-93): the tree over the map box
 * from one timestep's points */
static int schema_build(orc *o, const double *xy) {
    uint32_t n = o->cfg.num_objects;
    o->n_nodes = 0;
    o->n_free = 0;
    new_node(o);
    node *r = &o->nodes[0];
    r->lo[0] = o->cfg.map_lo[0]; r->lo[1] = o->cfg.map_lo[1];
    r->hi[0] = o->cfg.map_hi[0]; r->hi[1] = o->cfg.map_hi[1];
    r->mid_x = (r->lo[0] + r->hi[0]) / 2;
    r->mid_y = (r->lo[1] + r->hi[1]) / 2;
    r->type = T_LEAF;
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!ids || !tmp) { free(ids); free(tmp); set_err(o, "out of memory"); return -1; }
    for (uint32_t i = 0; i < n; i++) ids[i] = i;
    int rc = schema_build_rec(o, 0, xy, ids, n, tmp, 0);
    free(ids);
    free(tmp);
    if (rc) return -1;
    if (o->nodes[0].type != T_BRANCH) { set_err(o, "root does not split: too few objects for grid_capacity"); return -1; }
    o->have_schema = 1;
    return 0;
}

int orc_build_schema(void *vo, const double *xy) {
    orc *o = (orc *)vo;
    if (o->cfg.schema_mode == 0) { set_err(o, "the schema is rebuilt every step (schema_mode 0)"); return -1; }
    return schema_build(o, xy);
}

/* leaf ids of the current schema, in node order */
static int schema_number_leaves(orc *o) {
    uint32_t nl = 0;
    for (uint32_t i = 0; i < o->n_nodes; i++)
        if (o->nodes[i].type == T_LEAF) nl++;
    free(o->leaf_node);
    o->leaf_node = (uint32_t *)malloc(sizeof(uint32_t) * (nl ? nl : 1));
    if (!o->leaf_node) { set_err(o, "out of memory"); return -1; }
    o->n_leaves = 0;
    for (uint32_t i = 0; i < o->n_nodes; i++)
        if (o->nodes[i].type == T_LEAF) {
            o->nodes[i].leaf = o->n_leaves;
            o->leaf_node[o->n_leaves++] = i;
        }
    return 0;
}

static uint32_t schema_valid_nodes(const orc *o) {
    uint32_t v = 0;
    for (uint32_t i = 0; i < o->n_nodes; i++) v += o->nodes[i].type != 0;
    return v;
}

/* The update after a step, from the step's leaf counts (all points of a
 * leaf, before the buffer cap). A leaf splits when it holds more than
 * grid_capacity points and is higher than 2 * reach_distance in metres
 * (GPU, cuda_update_schema_collect geometry.cu:1193-1212) or wider than that
 * at its low latitude (CPU, update_schema_unit filter.cpp:386-397), for
 * schema_update_delay consecutive steps; its children are the halving boxes
 * of cuda_update_schema_split (geometry.cu:1114-1158, filter.cpp:338-375).
 * A branch whose four children are leaves holding fewer than grid_capacity
 * points together merges back into a leaf (geometry.cu:1213-1232,
 * :1160-1191; filter.cpp:398-419). All decisions are taken on the state
 * before the update, then splits, then merges, as the three kernels do. */
static int schema_update(orc *o) {
    const uint32_t cap = o->cfg.grid_capacity, delay = o->cfg.schema_update_delay ? o->cfg.schema_update_delay : 1;
    const double r2 = o->cfg.reach_distance * 2;
    const uint32_t n = o->n_nodes;
    uint8_t *act = (uint8_t *)calloc(n ? n : 1, 1);
    if (!act) { set_err(o, "out of memory"); return -1; }
    for (uint32_t i = 0; i < n; i++) {
        node *nd = &o->nodes[i];
        if (nd->type == T_LEAF) {
            double extent = o->cfg.schema_mode == 1 ? (nd->hi[1] - nd->lo[1]) / DPM_LAT
                                                    : (nd->hi[0] - nd->lo[0]) / host_dpm_lon(nd->lo[1]);
            if (extent > r2 && o->leaf_count[nd->leaf] > cap) {
                if (++nd->ovf >= delay) { act[i] = 1; nd->ovf = 0; }
            } else {
                nd->ovf = 0;
            }
        } else if (nd->type == T_BRANCH) {
            uint32_t leafchild = 0, sum = 0;
            for (int c = 0; c < 4; c++) {
                const node *ch = &o->nodes[nd->child[c]];
                if (ch->type == T_LEAF) { leafchild++; sum += o->leaf_count[ch->leaf]; }
            }
            if (leafchild == 4 && sum < cap) {
                if (++nd->unf >= delay) { act[i] = 2; nd->unf = 0; }
            } else {
                nd->unf = 0;
            }
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        if (act[i] != 1) continue;
        uint32_t kids[4];
        for (int c = 0; c < 4; c++)
            if (take_node(o, &kids[c])) { free(act); set_err(o, "out of memory"); return -1; }
        node *p = &o->nodes[i];
        p->type = T_BRANCH;
        double xhalf = p->mid_x - p->lo[0], yhalf = p->mid_y - p->lo[1];
        for (int c = 0; c < 4; c++) {
            node *k = &o->nodes[kids[c]];
            memset(k, 0, sizeof(node));
            k->type = T_LEAF;
            k->lo[0] = p->lo[0] + (double)(c % 2 == 1) * xhalf;
            k->lo[1] = p->lo[1] + (double)(c / 2 == 1) * yhalf;
            k->hi[0] = p->mid_x + (double)(c % 2 == 1) * xhalf;
            k->hi[1] = p->mid_y + (double)(c / 2 == 1) * yhalf;
            k->mid_x = (k->lo[0] + k->hi[0]) / 2;
            k->mid_y = (k->lo[1] + k->hi[1]) / 2;
            p->child[c] = kids[c];
        }
    }
    for (uint32_t i = 0; i < n; i++) {
        if (act[i] != 2) continue;
        node *p = &o->nodes[i];
        for (int c = 0; c < 4; c++) {
            o->nodes[p->child[c]].type = 0;
            if (o->n_free == o->cap_free) {
                uint32_t nc = o->cap_free ? o->cap_free * 2 : 1024;
                uint32_t *nf = (uint32_t *)realloc(o->free_nodes, sizeof(uint32_t) * nc);
                if (!nf) { free(act); set_err(o, "out of memory"); return -1; }
                o->free_nodes = nf;
                o->cap_free = nc;
            }
            o->free_nodes[o->n_free++] = p->child[c];
        }
        p->type = T_LEAF;
    }
    free(act);
    return 0;
}

/* ------------------------------------------------------------------ */
/* partition, filtering, refinement                                    */
/* ------------------------------------------------------------------ */

static int push_unit(orc *o, uint32_t pid, uint32_t nd, uint8_t inside) {
    if (o->n_units == o->cap_units) {
        uint64_t c = o->cap_units;
        if (grow((void **)&o->unit_pid, &c, o->n_units + 1, 4)) return -1;
        c = o->cap_units;
        if (grow((void **)&o->unit_node, &c, o->n_units + 1, 4)) return -1;
        c = o->cap_units;
        if (grow((void **)&o->unit_inside, &o->cap_units, o->n_units + 1, 1)) return -1;
        (void)c;
    }
    o->unit_pid[o->n_units] = pid;
    o->unit_node[o->n_units] = nd;
    o->unit_inside[o->n_units] = inside;
    o->n_units++;
    return 0;
}

/* lookup from last_valid (cuda_filtering geometry.cu:344-386; lookup_rec filter.cpp:130-152) */
static int lookup(orc *o, uint32_t pid, double px, double py, uint32_t cur) {
    for (int i = 0; i < 4; i++) {
        uint32_t c = o->nodes[cur].child[i];
        const node *nd = &o->nodes[c];
        double dist = box_dist(nd->lo, nd->hi, px, py, o->dlon[pid]);
        if (dist <= o->cfg.reach_distance) {
            if (nd->type == T_LEAF) {
                if (py < nd->lo[1] || (py < nd->hi[1] && px < nd->lo[0])) {
                    if (push_unit(o, pid, c, 0)) return -1;
                }
            } else {
                if (lookup(o, pid, px, py, c)) return -1;
            }
        }
    }
    return 0;
}

static inline void round_box(float *m, double x, double y) {
    float fx = (float)x, fy = (float)y;
    if (fx < m[0]) m[0] = fx;
    if (fy < m[1]) m[1] = fy;
    if (fx > m[2]) m[2] = fx;
    if (fy > m[3]) m[3] = fy;
}

static int find_pairs(orc *o, const double *xy) {
    uint32_t n = o->cfg.num_objects;
    double r = o->cfg.reach_distance;
    uint32_t nl = o->n_leaves;
    o->n_units = 0;
    o->n_pairs = 0;
    uint64_t border = 0, overflow = 0, candidates = 0;

    for (uint32_t i = 0; i < n; i++) o->dlon[i] = orc_dpm_lon(xy[2 * i + 1]);

    /* partition (cuda_partition, geometry.cu:221-271) */
    uint32_t *last_valid = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    for (uint32_t p = 0; p < n; p++) {
        double px = xy[2 * p], py = xy[2 * p + 1];
        uint32_t cur = 0, lv = 0, depth = 0;
        for (;;) {
            const node *nd = &o->nodes[cur];
            int loc = (py > nd->mid_y) * 2 + (px > nd->mid_x);
            cur = nd->child[loc];
            const node *c = &o->nodes[cur];
            if (px + o->x_buffer < c->hi[0] && py + o->y_buffer < c->hi[1]) lv = cur;
            if (c->type == T_LEAF) break;
            if (++depth > 64) {
                free(last_valid);
                set_err(o, "QTree descent deeper than 64 levels (malformed tree)");
                return -1;
            }
        }
        o->pt_leaf[p] = o->nodes[cur].leaf;
        last_valid[p] = lv;
        o->pt_border[p] = (lv != cur);
    }
    /* leaf members in pid order (CSR); a leaf stores at most leaf_buffer of them */
    free(o->leaf_start); free(o->leaf_count);
    o->leaf_start = (uint32_t *)calloc(nl + 1, sizeof(uint32_t));
    o->leaf_count = (uint32_t *)calloc(nl + 1, sizeof(uint32_t));
    for (uint32_t p = 0; p < n; p++) o->leaf_count[o->pt_leaf[p]]++;
    for (uint32_t l = 0; l < nl; l++) o->leaf_start[l + 1] = o->leaf_start[l] + o->leaf_count[l];
    uint32_t *fill = (uint32_t *)calloc(nl + 1, sizeof(uint32_t));
    for (uint32_t p = 0; p < n; p++) {
        uint32_t l = o->pt_leaf[p];
        o->leaf_members[o->leaf_start[l] + fill[l]++] = p;
    }
    free(fill);
    for (uint32_t l = 0; l < nl; l++)
        if (o->leaf_count[l] > o->cfg.leaf_buffer) overflow += o->leaf_count[l] - o->cfg.leaf_buffer;

    /* check units: every point with its own leaf (inside), then border lookups */
    for (uint32_t p = 0; p < n; p++) {
        if (push_unit(o, p, o->leaf_node[o->pt_leaf[p]], 1)) goto oom;
    }
    uint64_t n_inside = o->n_units;
    for (uint32_t p = 0; p < n; p++) {
        if (!o->pt_border[p]) continue;
        border++;
        if (lookup(o, p, xy[2 * p], xy[2 * p + 1], last_valid[p])) goto oom;
    }
    o->counts[ORC_C_UNITS_INSIDE] = n_inside;
    o->counts[ORC_C_UNITS_CROSS] = o->n_units - n_inside;
    o->counts[ORC_C_BORDER] = border;
    free(last_valid);
    last_valid = NULL;

    /* refinement (cuda_refinement, geometry.cu:424-497) */
    for (uint64_t u = 0; u < o->n_units; u++) {
        uint32_t p = o->unit_pid[u];
        uint32_t l = o->nodes[o->unit_node[u]].leaf;
        uint32_t size = o->leaf_count[l] < o->cfg.leaf_buffer ? o->leaf_count[l] : o->cfg.leaf_buffer;
        const uint32_t *mem = o->leaf_members + o->leaf_start[l];
        double px = xy[2 * p], py = xy[2 * p + 1];
        for (uint32_t k = 0; k < size; k++) {
            uint32_t q = mem[k];
            if (!o->unit_inside[u] || p < q) {
                candidates++;
                double d = pt_dist(px, py, xy[2 * q], xy[2 * q + 1], o->dlon[p]);
                if (d <= r) {
                    uint32_t a = p < q ? p : q, b = p < q ? q : p;
                    if (grow((void **)&o->pairs, &o->cap_pairs, o->n_pairs + 1, sizeof(pairhit))) goto oom;
                    pairhit *h = &o->pairs[o->n_pairs++];
                    h->key = ((uint64_t)a << 32) | b;
                    h->mbr[0] = h->mbr[1] = 100000.0f;
                    h->mbr[2] = h->mbr[3] = -100000.0f;
                    round_box(h->mbr, px, py);
                    round_box(h->mbr, xy[2 * q], xy[2 * q + 1]);
                }
            }
        }
    }
    /* a pair can only be emitted twice through ulp-thin slivers between a
     * leaf box and the descent; keep one (the MBR is the same points) */
    qsort(o->pairs, o->n_pairs, sizeof(pairhit), cmp_pairhit);
    uint64_t m = 0;
    for (uint64_t i = 0; i < o->n_pairs; i++)
        if (m == 0 || o->pairs[m - 1].key != o->pairs[i].key) o->pairs[m++] = o->pairs[i];
    o->n_pairs = m;
    o->counts[ORC_C_CANDIDATES] = candidates;
    o->counts[ORC_C_OVERFLOW] = overflow;
    o->counts[ORC_C_PAIRS] = m;
    return 0;
oom:
    free(last_valid);
    set_err(o, "out of memory");
    return -1;
}

/* ------------------------------------------------------------------ */
/* meetings (P4, P5) and identification (geometry.cu:595-737)           */
/* ------------------------------------------------------------------ */

static int emit_meeting(orc *o, const meeting *mt, uint32_t t) {
    /* t - start >= min_meet_time + 1 (geometry.cu:665) */
    if (!(t - mt->start >= o->cfg.min_meet_time + 1)) return 0;
    /* MBR inside the map (box_contain, geometry.cu:667-670) */
    if (!((double)mt->mbr[0] >= o->cfg.map_lo[0] && (double)mt->mbr[2] <= o->cfg.map_hi[0] &&
          (double)mt->mbr[1] >= o->cfg.map_lo[1] && (double)mt->mbr[3] <= o->cfg.map_hi[1]))
        return 0;
    uint32_t duration = mt->end - mt->start;
    uint32_t pid = (uint32_t)(mt->key >> 32), target = (uint32_t)mt->key;
    if (duration > o->cfg.max_meet_time) { /* geometry.cu:708-712 */
        o->oversize_oid_count++;
        o->sids[pid] = 1;
        o->sids[target] = 1;
    }
    if (grow((void **)&o->recs, &o->cap_recs, o->n_recs + 2, sizeof(record))) return -1;
    for (int k = 0; k < 2; k++) {
        if (k == 1) { uint32_t s = pid; pid = target; target = s; }
        record *rc = &o->recs[o->n_recs++];
        rc->key = ((u128)pid << (OID_BIT + DURATION_BIT + END_BIT)) + ((u128)target << (DURATION_BIT + END_BIT)) +
                  ((u128)(uint32_t)(mt->end - mt->start) << END_BIT) + (u128)(uint32_t)(mt->end - o->end_time_min);
        memcpy(rc->box, mt->mbr, sizeof(rc->box));
        if (o->same_pid_count[pid]++ == 0) o->sid_count++;
        uint32_t *ob = o->o_boxs[pid]; /* atomic_mbr_update_by_mbr (geometry.cu:115-120) */
        uint32_t q;
        q = f2q(mt->mbr[0]); if (q < ob[0]) ob[0] = q;
        q = f2q(mt->mbr[2]); if (q > ob[2]) ob[2] = q;
        q = f2q(mt->mbr[1]); if (q < ob[1]) ob[1] = q;
        q = f2q(mt->mbr[3]); if (q > ob[3]) ob[3] = q;
    }
    return 0;
}

static int update_meetings(orc *o, uint32_t t) {
    meeting *na = NULL;
    uint64_t nna = 0, cap = 0;
    uint64_t i = 0, j = 0, expired = 0, rec0 = o->n_recs;
    if (grow((void **)&na, &cap, o->n_active + o->n_pairs + 1, sizeof(meeting))) goto oom;
    while (i < o->n_active || j < o->n_pairs) {
        if (j >= o->n_pairs || (i < o->n_active && o->active[i].key < o->pairs[j].key)) {
            expired++;
            if (emit_meeting(o, &o->active[i], t)) goto oom;
            i++;
        } else if (i >= o->n_active || o->pairs[j].key < o->active[i].key) {
            meeting *m = &na[nna++];
            m->key = o->pairs[j].key;
            m->start = m->end = t;
            memcpy(m->mbr, o->pairs[j].mbr, sizeof(m->mbr));
            j++;
        } else {
            meeting *m = &na[nna++];
            *m = o->active[i];
            m->end = t;
            const float *b = o->pairs[j].mbr;
            if (b[0] < m->mbr[0]) m->mbr[0] = b[0];
            if (b[1] < m->mbr[1]) m->mbr[1] = b[1];
            if (b[2] > m->mbr[2]) m->mbr[2] = b[2];
            if (b[3] > m->mbr[3]) m->mbr[3] = b[3];
            i++; j++;
        }
    }
    free(o->active);
    o->active = na;
    o->n_active = nna;
    o->cap_active = cap;
    o->counts[ORC_C_ACTIVE] = nna;
    o->counts[ORC_C_EXPIRED] = expired;
    o->counts[ORC_C_NEW_RECORDS] = o->n_recs - rec0;
    return 0;
oom:
    free(na);
    set_err(o, "out of memory");
    return -1;
}

/* ------------------------------------------------------------------ */
/* flush (P6): geometry.cu:1752-2137, step_merge.cpp:132-186             */
/* ------------------------------------------------------------------ */

/* CTF::get_ctf_bits (step_merge.cpp:132-186), host arithmetic */
void orc_ctf_bits(orc_ctf *c, const double *map, uint32_t num_objects, uint32_t max_meet_time) {
    c->flags = 0;
    c->id_bit = (uint16_t)orc_min_bits(num_objects);
    c->duration_bit = (uint16_t)orc_min_bits(max_meet_time - 1);
    c->end_bit = c->duration_bit;
    double mxl = (map[2] - map[0]) / 3000;
    double myl = (map[3] - map[1]) / 4200;
    float wx = c->mbr[2] - c->mbr[0];
    float wy = c->mbr[3] - c->mbr[1];
    uint32_t lxg = x86_d2u32((double)wx / mxl);
    uint32_t lyg = x86_d2u32((double)wy / myl);
    c->low_x_bit = (uint16_t)orc_min_bits((lxg > 1 ? lxg : 1) - 1);
    c->low_y_bit = (uint16_t)orc_min_bits((lyg > 1 ? lyg : 1) - 1);
    c->edge_bit = c->low_x_bit > c->low_y_bit ? c->low_x_bit : c->low_y_bit;
    c->mbr_bit = (uint16_t)(c->low_x_bit + c->low_y_bit + 2 * c->edge_bit);
    if (c->mbr_bit > 64) c->flags |= 1;
    uint32_t temp_bit = c->id_bit * 2u + c->duration_bit + c->end_bit + c->mbr_bit; /* eight_parallel */
    if (temp_bit > 128) c->flags |= 1;
    c->key_bit = (uint16_t)(temp_bit / 8 * 8);
    if (c->key_bit < temp_bit) c->key_bit += 8;
    int32_t w = x86_d2i32((double)wx / (map[2] - map[0]) * 300);
    int32_t h = x86_d2i32((double)wy / (map[3] - map[1]) * 420);
    int32_t xg = (w + 1) > 32 ? (w + 1) : 32;
    int32_t yg = (h + 1) > 32 ? (h + 1) : 32;
    c->x_grid = (uint16_t)xg;
    c->y_grid = (uint16_t)yg;
    if (c->x_grid >= 256 || c->y_grid >= 256) {
        c->x_grid = (uint16_t)((w + 1) < 255 ? (w + 1) : 255);
        c->y_grid = (uint16_t)((h + 1) < 255 ? (h + 1) : 255);
    }
    c->bitmap_size = ((uint32_t)c->x_grid * c->y_grid + 7) / 8;
    if (!(c->bitmap_size < BITMAP_BYTES)) c->flags |= 1;
}

/* a shift by 128 or more yields 0, as a clamped device shift would */
static inline u128 shl128(u128 v, uint32_t s) { return s < 128 ? (v << s) : 0; }

static inline uint32_t key_oid(u128 k) { return (uint32_t)((k >> (OID_BIT + DURATION_BIT + END_BIT)) & ((1ull << OID_BIT) - 1)); }
static inline uint32_t key_dur(u128 k) { return (uint32_t)((k >> END_BIT) & 0xFFFFFFFFull); }
static inline uint32_t key_end(u128 k) { return (uint32_t)(k & 0xFFFFFFFFull); }

typedef struct sortent { uint64_t key; uint32_t val; uint32_t pos; } sortent;
static int cmp_sortent(const void *a, const void *b) {
    const sortent *x = (const sortent *)a, *y = (const sortent *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos < y->pos ? -1 : x->pos > y->pos; /* stable, as the radix sort behind sort_by_key */
}
static void stable_sort_kv(uint64_t *keys, uint32_t *vals, uint64_t n) {
    if (n < 2) return;
    sortent *e = (sortent *)malloc(sizeof(sortent) * n);
    for (uint64_t i = 0; i < n; i++) { e[i].key = keys[i]; e[i].val = vals[i]; e[i].pos = (uint32_t)i; }
    qsort(e, n, sizeof(sortent), cmp_sortent);
    for (uint64_t i = 0; i < n; i++) { keys[i] = e[i].key; vals[i] = e[i].val; }
    free(e);
}

static void set_bit(uint8_t *bm, uint64_t total_bits, uint64_t bit) {
    if (bit < total_bits) bm[bit >> 3] |= (uint8_t)(1u << (bit & 7));
}

static int flush(orc *o, uint32_t t) {
    uint32_t N = o->cfg.num_objects, C = o->cfg.ctf_count;
    uint32_t split = (uint32_t)sqrt((double)C);
    const double *mlo = o->cfg.map_lo, *mhi = o->cfg.map_hi;
    double map4[4] = {mlo[0], mlo[1], mhi[0], mhi[1]};
    int compat = o->cfg.compat & 1;
    uint64_t total_bits = (uint64_t)C * BIT_COUNT;

    if (grow((void **)&o->runs, &(uint64_t){o->n_runs}, o->n_runs + 1, sizeof(run))) goto oom;
    run *R = &o->runs[o->n_runs];
    memset(R, 0, sizeof(run));
    R->ctfs = (orc_ctf *)calloc(C, sizeof(orc_ctf));
    R->bitmaps = (uint8_t *)calloc(total_bits / 8, 1);
    R->sids = (uint16_t *)calloc(N, sizeof(uint16_t));
    uint64_t *mid = (uint64_t *)calloc(N, sizeof(uint64_t));
    uint32_t *oids = (uint32_t *)calloc(N, sizeof(uint32_t));
    uint32_t *capacity = (uint32_t *)calloc(C, sizeof(uint32_t));
    if (!R->ctfs || !R->bitmaps || !R->sids || !mid || !oids || !capacity) goto oom;

    /* set_oid (geometry.cu:739-767) */
    for (uint32_t id = 0; id < N; id++) {
        oids[id] = id;
        if (o->same_pid_count[id] >= 1) {
            float low0 = q2f(o->o_boxs[id][0]), high0 = q2f(o->o_boxs[id][2]);
            float low1 = q2f(o->o_boxs[id][1]), high1 = q2f(o->o_boxs[id][3]);
            float area = (high1 - low1) * (high0 - low0);
            if ((o->sticky && o->sids[id] == 1) || (double)area > o->cfg.area_threshold) {
                o->oversize_oid_count++;
                o->sids[id] = 1;
                continue;
            }
            uint32_t am0 = o->o_boxs[id][0] / 2 + o->o_boxs[id][2] / 2;
            uint32_t am1 = o->o_boxs[id][3] / 2 + o->o_boxs[id][3] / 2; /* high[1] twice, as the reference */
            mid[id] = ((uint64_t)am0 << 32) + (uint64_t)am1;
        }
    }

    /* partition: records of oversize objects to the tail (geometry.cu:1773-1801) */
    uint64_t K = o->n_recs;
    record *front = (record *)malloc(sizeof(record) * (K ? K : 1));
    record *tail = (record *)malloc(sizeof(record) * (K ? K : 1));
    uint64_t nf = 0, nt = 0;
    for (uint64_t i = 0; i < K; i++) {
        if (o->sids[key_oid(o->recs[i].key)] != 1) front[nf++] = o->recs[i];
        else tail[nt++] = o->recs[i];
    }
    qsort(tail, nt, sizeof(record), cmp_record);
    R->info.oversize_count = (uint32_t)nt;

    /* get_o_limit (geometry.cu:926-984) */
    {
        uint32_t smin = 1u << 30, smax = 0;
        for (uint64_t i = 0; i < nt; i++) {
            if (compat && (i % 64) >= 32) continue;
            uint32_t end = key_end(tail[i].key) + o->end_time_min;
            uint32_t st = end - key_dur(tail[i].key);
            if (st < smin) smin = st;
            if (st > smax) smax = st;
        }
        R->info.o_start_min = smin;
        R->info.o_start_max = smax;
        R->info.o_end_min = o->end_time_min;
        R->info.o_end_max = t;
    }
    /* cuda_write_o_bitmap (geometry.cu:904-924): drawn into CTF 0's bitmap slot,
     * which is copied out and cleared; bits past the slot stay in CTF 1's slot */
    for (uint64_t k = 0; k < nt; k++) {
        const float *b = tail[k].box;
        uint32_t low0 = sat_u32(((double)b[0] - mlo[0]) / (mhi[0] - mlo[0]) * BITMAP_EDGE);
        uint32_t low1 = sat_u32(((double)b[1] - mlo[1]) / (mhi[1] - mlo[1]) * BITMAP_EDGE);
        uint32_t high0 = sat_u32(((double)b[2] - mlo[0]) / (mhi[0] - mlo[0]) * BITMAP_EDGE);
        uint32_t high1 = sat_u32(((double)b[3] - mlo[1]) / (mhi[1] - mlo[1]) * BITMAP_EDGE);
        for (uint64_t i = low0; i <= high0 && i < total_bits; i++)
            for (uint64_t j = low1; j <= high1 && i + j * BITMAP_EDGE < total_bits; j++)
                set_bit(R->bitmaps, total_bits, (uint32_t)(i + j * BITMAP_EDGE));
    }
    memcpy(R->o_bitmap, R->bitmaps, BITMAP_BYTES);
    memset(R->bitmaps, 0, BITMAP_BYTES);

    /* STR (geometry.cu:1858-1905) */
    stable_sort_kv(mid, oids, N);
    uint32_t z = (N - o->sid_count) + o->oversize_oid_count;
    for (uint32_t id = z; id < N; id++) mid[id] &= 0xFFFFFFFFull; /* narrow_xy_to_y */
    uint32_t xpc = (o->sid_count - o->oversize_oid_count) / split + 1;
    for (uint32_t i = 0; i < split; i++) {
        uint64_t a = (uint64_t)z + (uint64_t)xpc * i;
        uint64_t b = (i == split - 1) ? N : (uint64_t)z + (uint64_t)xpc * (i + 1);
        if (b > N) b = N;
        if (a < b) stable_sort_kv(mid + a, oids + a, b - a);
    }
    uint32_t pkc = xpc / split + 1;
    for (uint32_t id = z; id < N; id++) { /* make_sid (geometry.cu:779-795) */
        uint32_t xi = (id - z) / xpc;
        uint32_t yi = (id - z - xi * xpc) / pkc;
        o->sids[oids[id]] = (uint16_t)(xi * split + yi + 2);
    }
    /* write_sid_in_key and the main sort (geometry.cu:797-806, 1907) */
    for (uint64_t i = 0; i < nf; i++) {
        uint32_t oid = key_oid(front[i].key);
        front[i].key = (front[i].key & (((u128)1 << SID_SHIFT) - 1)) + ((u128)o->sids[oid] << SID_SHIFT);
    }
    qsort(front, nf, sizeof(record), cmp_record);
    R->info.kv_count = (uint32_t)nf;

    /* get_CTF_capacity (geometry.cu:808-818) */
    int bad_capacity = 0;
    for (uint32_t id = z; id < N; id++) {
        uint32_t oid = oids[id];
        uint32_t cid = (uint32_t)o->sids[oid] - 2;
        if (cid < C) capacity[cid] += o->same_pid_count[oid];
        else bad_capacity = 1;
    }

    /* cuda_get_limit (geometry.cu:820-902) + get_ctf_bits on the host */
    uint64_t key_index = 0;
    for (uint32_t c = 0; c < C; c++) {
        orc_ctf *ctf = &R->ctfs[c];
        float l0 = 100000.0f, l1 = 100000.0f, h0 = -100000.0f, h1 = -100000.0f;
        uint32_t smin = 1u << 30, smax = 0;
        for (uint64_t k = 0; k < capacity[c]; k++) {
            if (compat && (k % 64) >= 32) continue;
            uint64_t kid = key_index + k;
            if (kid >= nf) { bad_capacity = 1; break; }
            const float *b = front[kid].box;
            l0 = fminf(l0, b[0]); l1 = fminf(l1, b[1]);
            h0 = fmaxf(h0, b[2]); h1 = fmaxf(h1, b[3]);
            uint32_t end = key_end(front[kid].key) + o->end_time_min;
            uint32_t st = end - key_dur(front[kid].key);
            if (st < smin) smin = st;
            if (st > smax) smax = st;
        }
        ctf->kv_capacity = capacity[c];
        ctf->start_time_min = smin;
        ctf->start_time_max = smax;
        ctf->end_time_min = o->end_time_min;
        ctf->end_time_max = t;
        ctf->mbr[0] = l0; ctf->mbr[1] = l1; ctf->mbr[2] = h0; ctf->mbr[3] = h1;
        orc_ctf_bits(ctf, map4, N, o->cfg.max_meet_time);
        key_index += capacity[c];
    }

    /* cuda_write_bitmap (geometry.cu:986-1021) and write_key_and_mbr (1073-1107) */
    uint64_t total_bytes = 0, total_cap = 0;
    int raster_capped = 0;
    for (uint32_t c = 0; c < C; c++) total_bytes += (uint64_t)capacity[c] * R->ctfs[c].key_bit / 8;
    for (uint32_t c = 0; c < C; c++) total_cap += capacity[c];
    if (total_cap != nf) bad_capacity = 1;
    R->key_bytes = (uint8_t *)calloc(total_bytes ? total_bytes : 1, 1);
    if (!R->key_bytes) goto oom;
    key_index = 0;
    uint64_t bytes_index = 0;
    for (uint32_t c = 0; c < C; c++) {
        const orc_ctf *ctf = &R->ctfs[c];
        const float *m = ctf->mbr;
        uint64_t base_bit = (uint64_t)c * BIT_COUNT;
        uint32_t kb = ctf->key_bit / 8;
        float longer = fmaxf(m[2] - m[0], m[3] - m[1]);
        for (uint64_t k = 0; k < capacity[c]; k++) {
            uint64_t kid = key_index + k;
            if (kid >= nf) break;
            const float *b = front[kid].box;
            /* bitmap */
            uint32_t low0 = sat_u32((b[0] - m[0]) / (m[2] - m[0]) * (float)ctf->x_grid);
            uint32_t low1 = sat_u32((b[1] - m[1]) / (m[3] - m[1]) * (float)ctf->y_grid);
            uint32_t high0 = sat_u32((b[2] - m[0]) / (m[2] - m[0]) * (float)ctf->x_grid);
            uint32_t high1 = sat_u32((b[3] - m[1]) / (m[3] - m[1]) * (float)ctf->y_grid);
            if (ctf->x_grid) {
                uint64_t room = total_bits - base_bit; /* bits left in the array from this slot */
                uint32_t done = 0;   /* at most 2^20 cells per record (only reachable in compat mode) */
                for (uint64_t j = low1; j <= high1 && j * ctf->x_grid < room && done < (1u << 20); j++)
                    for (uint64_t i = low0; i <= high0 && i + j * ctf->x_grid < room && done < (1u << 20);
                         i++, done++)
                        set_bit(R->bitmaps, total_bits, base_bit + (uint32_t)(i + j * ctf->x_grid));
                if (done >= (1u << 20)) raster_capped = 1;
            }
            /* serialize_mbr, f_box form (geometry.cu:161-171) */
            uint64_t v0 = sat_u64((b[0] - m[0]) / (m[2] - m[0]) * (float)((1ull << ctf->low_x_bit) - 1));
            uint64_t v1 = sat_u64((b[1] - m[1]) / (m[3] - m[1]) * (float)((1ull << ctf->low_y_bit) - 1));
            uint64_t vx = sat_u64((b[2] - b[0]) / longer * (float)((1ull << ctf->edge_bit) - 1));
            uint64_t vy = sat_u64((b[3] - b[1]) / longer * (float)((1ull << ctf->edge_bit) - 1));
            uint32_t e = ctf->edge_bit, ly = ctf->low_y_bit;
            uint64_t vm = (ly + 2 * e < 64 ? (v0 << (ly + e + e)) : 0) + (2 * e < 64 ? (v1 << (e + e)) : 0) +
                          (e < 64 ? (vx << e) : 0) + vy;
            /* extract_fields keeps the sid bits in pid (cuda_util.cuh:163-169) */
            u128 key = front[kid].key;
            u128 end = key & 0xFFFFFFFFull;
            u128 dur = (key >> END_BIT) & 0xFFFFFFFFull;
            u128 target = (key >> (DURATION_BIT + END_BIT)) & ((1ull << OID_BIT) - 1);
            u128 pid = key >> (OID_BIT + DURATION_BIT + END_BIT);
            uint32_t s_mbr = ctf->mbr_bit, s_end = ctf->end_bit, s_dur = ctf->duration_bit, s_id = ctf->id_bit;
            u128 tk = shl128(pid, s_id + s_dur + s_end + s_mbr) + shl128(target, s_dur + s_end + s_mbr) +
                      shl128(dur, s_end + s_mbr) + shl128(end, s_mbr) + (u128)vm;
            uint8_t *dst = R->key_bytes + bytes_index + k * kb;
            for (uint32_t q = 0; q < kb && q < 16; q++) dst[q] = (uint8_t)(tk >> (8 * q));
        }
        key_index += capacity[c];
        bytes_index += (uint64_t)capacity[c] * ctf->key_bit / 8;
    }
    if (bad_capacity) for (uint32_t c = 0; c < C; c++) R->ctfs[c].flags |= 2;
    if (raster_capped) for (uint32_t c = 0; c < C; c++) R->ctfs[c].flags |= 8;
    if (!(nt < o->cfg.kv_restriction / 100)) for (uint32_t c = 0; c < C; c++) R->ctfs[c].flags |= 4;

    /* outputs */
    memcpy(R->sids, o->sids, sizeof(uint16_t) * N);
    R->o_keys = (u128 *)malloc(sizeof(u128) * (nt ? nt : 1));
    R->o_boxes = (float *)malloc(sizeof(float) * 4 * (nt ? nt : 1));
    R->sorted_keys = (u128 *)malloc(sizeof(u128) * (nf ? nf : 1));
    R->sorted_boxes = (float *)malloc(sizeof(float) * 4 * (nf ? nf : 1));
    for (uint64_t i = 0; i < nt; i++) { R->o_keys[i] = tail[i].key; memcpy(R->o_boxes + 4 * i, tail[i].box, 16); }
    for (uint64_t i = 0; i < nf; i++) { R->sorted_keys[i] = front[i].key; memcpy(R->sorted_boxes + 4 * i, front[i].box, 16); }
    R->info.ctf_count = C;
    R->info.num_objects = N;
    R->info.end_time_min = o->end_time_min;
    R->info.cur_time = t;
    R->info.sid_count = o->sid_count;
    R->info.oversize_oid_count = o->oversize_oid_count;
    R->info.key_bytes = total_bytes;
    o->n_runs++;

    /* reset (geometry.cu:2111-2131); the next run's end offsets are relative
     * to this flush's timestep (process.cpp:271 via end_time_max = cur_time) */
    o->n_recs = 0;
    o->sid_count = 0;
    o->oversize_oid_count = 0;
    memset(o->sids, 0, sizeof(uint16_t) * N);
    memset(o->same_pid_count, 0, sizeof(uint32_t) * N);
    for (uint32_t i = 0; i < N; i++) {
        o->o_boxs[i][0] = o->o_boxs[i][1] = 100000000u;
        o->o_boxs[i][2] = o->o_boxs[i][3] = 0;
    }
    o->end_time_min = t;
    free(front); free(tail); free(mid); free(oids); free(capacity);
    return 0;
oom:
    set_err(o, "out of memory in flush");
    return -1;
}

/* a step's flush, timed on the wall clock (the CPU flush baseline's unit) */
static int timed_flush(orc *o, uint32_t t) {
    struct timespec t0, t1;
    const uint64_t recs = o->n_recs;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int rc = flush(o, t);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    o->last_flush_ms = (double)(t1.tv_sec - t0.tv_sec) * 1e3 + (double)(t1.tv_nsec - t0.tv_nsec) / 1e6;
    o->last_flush_records = recs;
    return rc;
}

double orc_last_flush_ms(void *vo, uint64_t *records) {
    orc *o = (orc *)vo;
    if (records) *records = o->last_flush_records;
    return o->last_flush_ms;
}

/* ------------------------------------------------------------------ */
/* step                                                                */
/* ------------------------------------------------------------------ */

int orc_step(void *vo, const double *xy, uint32_t t) {
    orc *o = (orc *)vo;
    memset(o->counts, 0, sizeof(o->counts));
    if (o->cfg.schema_mode == 0) {
        if (build_tree(o, xy)) return -1;
        o->counts[ORC_C_NODES] = o->n_nodes;
    } else {
        /* the first step's points build the schema (process.cpp:401-402) */
        if (!o->have_schema && schema_build(o, xy)) return -1;
        if (schema_number_leaves(o)) return -1;
        o->counts[ORC_C_NODES] = schema_valid_nodes(o);
    }
    o->counts[ORC_C_LEAVES] = o->n_leaves;
    if (find_pairs(o, xy)) return -1;
    o->last_t = t;
    if (update_meetings(o, t)) return -1;
    o->counts[ORC_C_KV_COUNT] = o->n_recs;
    if (o->n_recs > o->cfg.kv_restriction) { /* geometry.cu:1752 */
        if (timed_flush(o, t)) return -1;
        o->counts[ORC_C_FLUSHED] = 1;
    }
    /* the schema update for the next step (geometry.cu:2159-2182; process.cpp:608-612) */
    if (o->cfg.schema_mode != 0 && schema_update(o)) return -1;
    return 0;
}

/* Persisted meeting units (workbench::dump_meetings / load_meetings,
 * workbench.cpp:446-470): one 32 B meeting_unit per pair found at a step,
 * key = Cantor(min pid, max pid) (geometry.cu:453-455), start = end = step,
 * mbr = float bbox of the two positions (geometry.cu:456-466). */
static uint64_t cantor(uint32_t a, uint32_t b) { return ((uint64_t)a + b) * ((uint64_t)a + b + 1) / 2 + b; }

/* inverse by exact integer search around the double estimate (cuda_util.cuh:55-70) */
static int uncantor(uint64_t z, uint32_t n, uint32_t *a, uint32_t *b) {
    uint64_t w = (uint64_t)floor((sqrt(8.0 * (double)z + 1) - 1) / 2);
    while (w > 0 && (u128)w * (w + 1) / 2 > z) w--;
    while ((u128)(w + 1) * (w + 2) / 2 <= z) w++;
    uint64_t y = z - (uint64_t)((u128)w * (w + 1) / 2), x = w - y;
    if (!(x < y && y < n)) return -1;
    *a = (uint32_t)x;
    *b = (uint32_t)y;
    return 0;
}

void orc_get_step_meetings(void *vo, orc_meeting_unit *u) {
    orc *o = (orc *)vo;
    for (uint64_t i = 0; i < o->n_pairs; i++) {
        uint32_t a = (uint32_t)(o->pairs[i].key >> 32), b = (uint32_t)o->pairs[i].key;
        u[i].key = cantor(a, b);
        u[i].start = u[i].end = o->last_t;
        memcpy(u[i].mbr, o->pairs[i].mbr, sizeof(u[i].mbr));
    }
}

/* a step from persisted units instead of positions (geometry.cu:1609-1622):
 * every unit's pair is inserted or extended with the unit's MBR */
int orc_step_meetings(void *vo, const orc_meeting_unit *u, uint64_t n, uint32_t t) {
    orc *o = (orc *)vo;
    memset(o->counts, 0, sizeof(o->counts));
    pairhit *p = (pairhit *)malloc(sizeof(pairhit) * (n ? n : 1));
    if (!p) { set_err(o, "out of memory"); return -1; }
    for (uint64_t i = 0; i < n; i++) {
        uint32_t a, b;
        if (uncantor(u[i].key, o->cfg.num_objects, &a, &b)) {
            free(p);
            set_err(o, "meeting unit key is not the Cantor key of two distinct objects below num_objects");
            return -1;
        }
        p[i].key = ((uint64_t)a << 32) | b;
        memcpy(p[i].mbr, u[i].mbr, sizeof(p[i].mbr));
    }
    qsort(p, n, sizeof(pairhit), cmp_pairhit);
    for (uint64_t i = 1; i < n; i++)
        if (p[i].key == p[i - 1].key) {
            free(p);
            set_err(o, "a pair appears twice among the step's meeting units");
            return -1;
        }
    free(o->pairs);
    o->pairs = p;
    o->n_pairs = n;
    o->cap_pairs = n ? n : 1;
    o->counts[ORC_C_PAIRS] = n;
    o->last_t = t;
    if (update_meetings(o, t)) return -1;
    o->counts[ORC_C_KV_COUNT] = o->n_recs;
    if (o->n_recs > o->cfg.kv_restriction) {
        if (timed_flush(o, t)) return -1;
        o->counts[ORC_C_FLUSHED] = 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* getters                                                             */
/* ------------------------------------------------------------------ */

uint64_t orc_num_pairs(void *o) { return ((orc *)o)->n_pairs; }
void orc_get_pairs(void *vo, uint64_t *keys) {
    orc *o = (orc *)vo;
    for (uint64_t i = 0; i < o->n_pairs; i++) keys[i] = o->pairs[i].key;
}

typedef struct leafsort { double lo1, lo0; uint32_t leaf; } leafsort;
static int cmp_leafsort(const void *a, const void *b) {
    const leafsort *x = (const leafsort *)a, *y = (const leafsort *)b;
    if (x->lo1 != y->lo1) return x->lo1 < y->lo1 ? -1 : 1;
    if (x->lo0 != y->lo0) return x->lo0 < y->lo0 ? -1 : 1;
    return 0;
}
static uint32_t *canonical_leaf_order(orc *o) {
    leafsort *s = (leafsort *)malloc(sizeof(leafsort) * (o->n_leaves ? o->n_leaves : 1));
    for (uint32_t l = 0; l < o->n_leaves; l++) {
        const node *nd = &o->nodes[o->leaf_node[l]];
        s[l].lo1 = nd->lo[1]; s[l].lo0 = nd->lo[0]; s[l].leaf = l;
    }
    qsort(s, o->n_leaves, sizeof(leafsort), cmp_leafsort);
    uint32_t *rank = (uint32_t *)malloc(sizeof(uint32_t) * (o->n_leaves ? o->n_leaves : 1));
    for (uint32_t i = 0; i < o->n_leaves; i++) rank[s[i].leaf] = i;
    free(s);
    return rank;
}

uint64_t orc_num_leaves(void *o) { return ((orc *)o)->n_leaves; }
void orc_get_leaves(void *vo, double *boxes, uint32_t *counts) {
    orc *o = (orc *)vo;
    uint32_t *rank = canonical_leaf_order(o);
    for (uint32_t l = 0; l < o->n_leaves; l++) {
        const node *nd = &o->nodes[o->leaf_node[l]];
        uint32_t r = rank[l];
        boxes[4 * r] = nd->lo[0]; boxes[4 * r + 1] = nd->lo[1];
        boxes[4 * r + 2] = nd->hi[0]; boxes[4 * r + 3] = nd->hi[1];
        counts[r] = o->leaf_count[l];
    }
    free(rank);
}
void orc_get_point_leaf(void *vo, uint32_t *leaf, uint8_t *border) {
    orc *o = (orc *)vo;
    uint32_t *rank = canonical_leaf_order(o);
    for (uint32_t p = 0; p < o->cfg.num_objects; p++) {
        leaf[p] = rank[o->pt_leaf[p]];
        border[p] = o->pt_border[p];
    }
    free(rank);
}
uint64_t orc_num_nodes(void *o) { return ((orc *)o)->n_nodes; }
void orc_get_nodes(void *vo, double *boxes, double *mids, uint32_t *type, uint32_t *children) {
    orc *o = (orc *)vo;
    for (uint32_t i = 0; i < o->n_nodes; i++) {
        const node *nd = &o->nodes[i];
        boxes[4 * i] = nd->lo[0]; boxes[4 * i + 1] = nd->lo[1];
        boxes[4 * i + 2] = nd->hi[0]; boxes[4 * i + 3] = nd->hi[1];
        mids[2 * i] = nd->mid_x; mids[2 * i + 1] = nd->mid_y;
        type[i] = nd->type;
        for (int c = 0; c < 4; c++) children[4 * i + c] = nd->child[c];
    }
}
uint64_t orc_num_units(void *o) { return ((orc *)o)->n_units; }
void orc_get_units(void *vo, uint32_t *pid, uint32_t *nodeidx, uint8_t *inside) {
    orc *o = (orc *)vo;
    memcpy(pid, o->unit_pid, 4 * o->n_units);
    memcpy(nodeidx, o->unit_node, 4 * o->n_units);
    memcpy(inside, o->unit_inside, o->n_units);
}

uint64_t orc_num_active(void *o) { return ((orc *)o)->n_active; }
void orc_get_active(void *vo, uint64_t *keys, uint32_t *start, uint32_t *end, float *mbr) {
    orc *o = (orc *)vo;
    for (uint64_t i = 0; i < o->n_active; i++) {
        keys[i] = o->active[i].key;
        start[i] = o->active[i].start;
        end[i] = o->active[i].end;
        memcpy(mbr + 4 * i, o->active[i].mbr, 16);
    }
}

uint64_t orc_num_records(void *o) { return ((orc *)o)->n_recs; }
void orc_get_records(void *vo, uint64_t *keys, float *boxes) {
    orc *o = (orc *)vo;
    for (uint64_t i = 0; i < o->n_recs; i++) {
        keys[2 * i] = (uint64_t)o->recs[i].key;
        keys[2 * i + 1] = (uint64_t)(o->recs[i].key >> 64);
        memcpy(boxes + 4 * i, o->recs[i].box, 16);
    }
}

uint32_t orc_num_runs(void *o) { return ((orc *)o)->n_runs; }
void orc_get_run_info(void *vo, uint32_t r, orc_run_info *info) {
    orc *o = (orc *)vo;
    if (r >= o->n_runs) { memset(info, 0, sizeof(*info)); return; }
    *info = o->runs[r].info;
}
void orc_get_run(void *vo, uint32_t r, orc_ctf *ctfs, uint8_t *key_bytes, uint8_t *bitmaps,
                 uint16_t *sids, uint64_t *o_keys, float *o_boxes, uint8_t *o_bitmap,
                 uint64_t *sorted_keys, float *sorted_boxes) {
    if (r >= ((orc *)vo)->n_runs) return;
    const run *R = &((orc *)vo)->runs[r];
    uint32_t C = R->info.ctf_count;
    if (ctfs) memcpy(ctfs, R->ctfs, sizeof(orc_ctf) * C);
    if (key_bytes) memcpy(key_bytes, R->key_bytes, R->info.key_bytes);
    if (bitmaps) memcpy(bitmaps, R->bitmaps, (uint64_t)C * BITMAP_BYTES);
    if (sids) memcpy(sids, R->sids, sizeof(uint16_t) * R->info.num_objects);
    if (o_keys) for (uint32_t i = 0; i < R->info.oversize_count; i++) {
        o_keys[2 * i] = (uint64_t)R->o_keys[i];
        o_keys[2 * i + 1] = (uint64_t)(R->o_keys[i] >> 64);
    }
    if (o_boxes) memcpy(o_boxes, R->o_boxes, 16ull * R->info.oversize_count);
    if (o_bitmap) memcpy(o_bitmap, R->o_bitmap, BITMAP_BYTES);
    if (sorted_keys) for (uint32_t i = 0; i < R->info.kv_count; i++) {
        sorted_keys[2 * i] = (uint64_t)R->sorted_keys[i];
        sorted_keys[2 * i + 1] = (uint64_t)(R->sorted_keys[i] >> 64);
    }
    if (sorted_boxes) memcpy(sorted_boxes, R->sorted_boxes, 16ull * R->info.kv_count);
}

/* ------------------------------------------------------------------ */
/* queries over the resident state (see stc_oracle.h)                  */
/* ------------------------------------------------------------------ */

typedef struct qhit_ctx {
    orc_hit *hits;
    uint64_t cap, n;
    uint64_t *counts;
    const orc_query *qs;
} qhit_ctx;

static void q_emit(qhit_ctx *h, uint32_t qi, uint32_t pid, uint32_t target, uint32_t start, uint32_t end,
                   uint32_t source, const double *mbr) {
    h->counts[qi]++;
    if (h->qs[qi].flags & 1u) return; /* counted only */
    if (h->n < h->cap) {
        orc_hit *x = &h->hits[h->n];
        x->query = qi; x->pid = pid; x->target = target; x->start = start; x->end = end; x->source = source;
        memcpy(x->mbr, mbr, sizeof(x->mbr));
    }
    h->n++;
}

static int q_time(const orc_query *q, uint32_t start, uint32_t end) { return start <= q->t1 && end >= q->t0; }
static int q_box(const orc_query *q, const double *m) {
    return m[0] <= q->box[2] && m[2] >= q->box[0] && m[1] <= q->box[3] && m[3] >= q->box[1];
}

/* one meeting as seen from its owner `pid` */
static void q_match(const orc_query *qs, uint32_t nq, qhit_ctx *h, uint32_t pid, uint32_t target, uint32_t start,
                    uint32_t end, uint32_t source, const double *mbr) {
    for (uint32_t qi = 0; qi < nq; qi++) {
        const orc_query *q = &qs[qi];
        if (!q_time(q, start, end)) continue;
        if ((q->kind == 1 && pid == q->oid) || (q->kind == 3 && pid == q->oid && target == q->oid2) ||
            (q->kind == 2 && pid < target && q_box(q, mbr)))
            q_emit(h, qi, pid, target, start, end, source, mbr);
    }
}

/* CTF::ctf_get_key_oid, parse_key and new_parse_mbr (step_merge.cpp:325-379)
 * on the little-endian packed key at p */
static void decode_packed(const orc_ctf *c, const uint8_t *p, uint32_t *pid, uint32_t *target, uint32_t *start,
                          uint32_t *end, double *box) {
    u128 k = 0;
    memcpy(&k, p, c->key_bit / 8);
    uint64_t vm = (uint64_t)(k & (((u128)1 << c->mbr_bit) - 1));
    u128 r = k >> c->mbr_bit;
    uint32_t e = (uint32_t)(r & ((1ull << c->end_bit) - 1));
    r >>= c->end_bit;
    uint32_t d = (uint32_t)(r & ((1ull << c->duration_bit) - 1));
    r >>= c->duration_bit;
    *target = (uint32_t)(r & ((1ull << c->id_bit) - 1));
    r >>= c->id_bit;
    *pid = (uint32_t)(r & ((1ull << c->id_bit) - 1));
    *end = e + c->end_time_min;
    *start = *end - d;
    uint64_t mask_edge = (1ull << c->edge_bit) - 1, mask_low_y = (1ull << c->low_y_bit) - 1,
             mask_low_x = (1ull << c->low_x_bit) - 1;
    uint64_t y = vm & mask_edge;
    vm >>= c->edge_bit;
    uint64_t x = vm & mask_edge;
    vm >>= c->edge_bit;
    uint64_t low1 = vm & mask_low_y;
    vm >>= c->low_y_bit;
    uint64_t low0 = vm & mask_low_x;
    float w0 = c->mbr[2] - c->mbr[0], w1 = c->mbr[3] - c->mbr[1];
    float longer = w0 > w1 ? w0 : w1;
    box[0] = (double)c->mbr[0] + (double)low0 / (double)mask_low_x * (double)w0;
    box[1] = (double)c->mbr[1] + (double)low1 / (double)mask_low_y * (double)w1;
    box[2] = box[0] + (double)x / (double)mask_edge * (double)longer;
    box[3] = box[1] + (double)y / (double)mask_edge * (double)longer;
}

uint32_t orc_end_time_min(void *o) { return ((orc *)o)->end_time_min; }

/* ------------------------------------------------------------------ */
/* compaction                                                          */
/* ------------------------------------------------------------------ */

/* CTF::new_parse_mbr_f_box (step_merge.cpp:381-401): each corner is the double
 * expression of new_parse_mbr stored to float, and the high corner adds to
 * the stored (float) low corner */
static void decode_fbox(const orc_ctf *c, const uint8_t *p, float *b) {
    u128 k = 0;
    memcpy(&k, p, c->key_bit / 8);
    uint64_t vm = (uint64_t)(k & (((u128)1 << c->mbr_bit) - 1));
    uint64_t mask_edge = (1ull << c->edge_bit) - 1, mask_low_y = (1ull << c->low_y_bit) - 1,
             mask_low_x = (1ull << c->low_x_bit) - 1;
    uint64_t y = vm & mask_edge;
    vm >>= c->edge_bit;
    uint64_t x = vm & mask_edge;
    vm >>= c->edge_bit;
    uint64_t low1 = vm & mask_low_y;
    vm >>= c->low_y_bit;
    uint64_t low0 = vm & mask_low_x;
    float w0 = c->mbr[2] - c->mbr[0], w1 = c->mbr[3] - c->mbr[1];
    float longer = w0 > w1 ? w0 : w1;
    b[0] = (float)((double)c->mbr[0] + (double)low0 / (double)mask_low_x * (double)w0);
    b[1] = (float)((double)c->mbr[1] + (double)low1 / (double)mask_low_y * (double)w1);
    b[2] = (float)((double)b[0] + (double)x / (double)mask_edge * (double)longer);
    b[3] = (float)((double)b[1] + (double)y / (double)mask_edge * (double)longer);
}

static int compact_add(orc *o, record *rc, uint32_t pid, uint32_t target, uint32_t dur, uint32_t end, uint32_t base,
                       const float *box) {
    if (pid >= o->cfg.num_objects || target >= o->cfg.num_objects) return -1;
    rc->key = ((u128)pid << (OID_BIT + DURATION_BIT + END_BIT)) + ((u128)target << (DURATION_BIT + END_BIT)) +
              ((u128)dur << END_BIT) + (u128)(uint32_t)(end - base);
    memcpy(rc->box, box, sizeof(rc->box));
    o->same_pid_count[pid]++;
    uint32_t *ob = o->o_boxs[pid], q;
    q = f2q(box[0]); if (q < ob[0]) ob[0] = q;
    q = f2q(box[2]); if (q > ob[2]) ob[2] = q;
    q = f2q(box[1]); if (q < ob[1]) ob[1] = q;
    q = f2q(box[3]); if (q > ob[3]) ob[3] = q;
    return 0;
}

/* Merge runs [first, first + count) into one run in their place: every record
 * of the inputs, decoded as tools/compaction.cpp:320-336 reads them, becomes a
 * memtable record whose end offset is relative to the earliest input base
 * (compaction.cpp:483-487); the memtable is flushed at the latest input's
 * flush time, with the objects oversize in any input kept oversize
 * (compaction.cpp:259-268). */
int orc_compact_runs(void *vo, uint32_t first, uint32_t count) {
    orc *o = (orc *)vo;
    if (count == 0 || (uint64_t)first + count > o->n_runs) { set_err(o, "run range out of bounds"); return -1; }
    if (count == 1) return 0;
    const uint32_t N = o->cfg.num_objects;
    uint64_t total = 0;
    uint32_t base = 0xFFFFFFFFu, tmax = 0;
    for (uint32_t r = first; r < first + count; r++) {
        const run *R = &o->runs[r];
        for (uint32_t c = 0; c < R->info.ctf_count; c++) total += R->ctfs[c].kv_capacity;
        total += R->info.oversize_count;
        if (R->info.end_time_min < base) base = R->info.end_time_min;
        if (R->info.cur_time > tmax) tmax = R->info.cur_time;
    }
    /* the memtable being filled by steps is set aside */
    record *s_recs = o->recs; uint64_t s_n = o->n_recs, s_cap = o->cap_recs;
    uint32_t *s_spc = o->same_pid_count; uint32_t (*s_obox)[4] = o->o_boxs; uint16_t *s_sids = o->sids;
    uint32_t s_sid = o->sid_count, s_over = o->oversize_oid_count, s_etm = o->end_time_min,
             s_kvr = o->cfg.kv_restriction;
    o->recs = (record *)malloc(sizeof(record) * (total ? total : 1));
    o->cap_recs = total ? total : 1;
    o->n_recs = 0;
    o->same_pid_count = (uint32_t *)calloc(N, sizeof(uint32_t));
    o->o_boxs = (uint32_t (*)[4])malloc(sizeof(uint32_t) * 4 * N);
    o->sids = (uint16_t *)calloc(N, sizeof(uint16_t));
    int rc = -1;
    if (!o->recs || !o->same_pid_count || !o->o_boxs || !o->sids) { set_err(o, "out of memory in compaction"); goto done; }
    for (uint32_t i = 0; i < N; i++) {
        o->o_boxs[i][0] = o->o_boxs[i][1] = 100000000u;
        o->o_boxs[i][2] = o->o_boxs[i][3] = 0;
    }
    for (uint32_t r = first; r < first + count; r++) {
        const run *R = &o->runs[r];
        uint64_t off = 0;
        for (uint32_t c = 0; c < R->info.ctf_count; c++) {
            const orc_ctf *ctf = &R->ctfs[c];
            uint32_t kb = ctf->key_bit / 8;
            for (uint32_t i = 0; i < ctf->kv_capacity; i++) {
                const uint8_t *p = R->key_bytes + off + (uint64_t)i * kb;
                uint32_t pid, target, start, end;
                double unused[4];
                float box[4];
                decode_packed(ctf, p, &pid, &target, &start, &end, unused);
                decode_fbox(ctf, p, box);
                if (compact_add(o, &o->recs[o->n_recs++], pid, target, end - start, end, base, box)) goto bad_id;
            }
            off += (uint64_t)ctf->kv_capacity * kb;
        }
        for (uint32_t i = 0; i < R->info.oversize_count; i++) {
            u128 k = R->o_keys[i];
            uint32_t pid = key_oid(k), target = (uint32_t)((k >> 64) & ((1ull << OID_BIT) - 1));
            if (compact_add(o, &o->recs[o->n_recs++], pid, target, key_dur(k), key_end(k) + R->info.end_time_min, base,
                            R->o_boxes + 4 * i))
                goto bad_id;
        }
        for (uint32_t id = 0; id < N; id++)
            if (R->sids[id] == 1) o->sids[id] = 1;
    }
    o->sid_count = 0;
    for (uint32_t id = 0; id < N; id++) o->sid_count += o->same_pid_count[id] ? 1 : 0;
    o->oversize_oid_count = 0; /* counted by the flush */
    o->end_time_min = base;
    if (count > 0xFFFFFFFFu / (s_kvr ? s_kvr : 1)) o->cfg.kv_restriction = 0xFFFFFFFFu;
    else o->cfg.kv_restriction = s_kvr * count;
    o->sticky = 1;
    rc = flush(o, tmax);
    o->sticky = 0;
    if (rc == 0) { /* the new run replaces its inputs */
        run merged = o->runs[o->n_runs - 1];
        for (uint32_t r = first; r < first + count; r++) free_run(&o->runs[r]);
        memmove(o->runs + first + 1, o->runs + first + count, sizeof(run) * (o->n_runs - 1 - (first + count)));
        o->runs[first] = merged;
        o->n_runs -= count;
    }
    goto done;
bad_id:
    set_err(o, "an input run holds a record whose object id is not below num_objects");
done:
    free(o->recs); free(o->same_pid_count); free(o->o_boxs); free(o->sids);
    o->recs = s_recs; o->n_recs = s_n; o->cap_recs = s_cap;
    o->same_pid_count = s_spc; o->o_boxs = s_obox; o->sids = s_sids;
    o->sid_count = s_sid; o->oversize_oid_count = s_over; o->end_time_min = s_etm;
    o->cfg.kv_restriction = s_kvr;
    return rc;
}

uint64_t orc_query_batch(void *vo, const orc_query *qs, uint32_t nq, orc_hit *hits, uint64_t cap, uint64_t *counts) {
    orc *o = (orc *)vo;
    qhit_ctx h = {hits, cap, 0, counts, qs};
    memset(counts, 0, sizeof(uint64_t) * nq);
    double m[4];
    /* live meetings: both directions */
    for (uint64_t i = 0; i < o->n_active; i++) {
        const meeting *mt = &o->active[i];
        for (int k = 0; k < 4; k++) m[k] = (double)mt->mbr[k];
        uint32_t a = (uint32_t)(mt->key >> 32), b = (uint32_t)mt->key;
        q_match(qs, nq, &h, a, b, mt->start, mt->end, 0, m);
        q_match(qs, nq, &h, b, a, mt->start, mt->end, 0, m);
    }
    /* memtable records */
    for (uint64_t i = 0; i < o->n_recs; i++) {
        const record *rc = &o->recs[i];
        uint32_t pid = (uint32_t)((rc->key >> 90) & ((1ull << OID_BIT) - 1));
        uint32_t target = (uint32_t)((rc->key >> 64) & ((1ull << OID_BIT) - 1));
        uint32_t end = key_end(rc->key) + o->end_time_min, start = end - key_dur(rc->key);
        for (int k = 0; k < 4; k++) m[k] = (double)rc->box[k];
        q_match(qs, nq, &h, pid, target, start, end, 1, m);
    }
    /* runs: every CTF record and every oversize record */
    for (uint32_t r = 0; r < o->n_runs; r++) {
        const run *R = &o->runs[r];
        uint64_t off = 0;
        for (uint32_t c = 0; c < R->info.ctf_count; c++) {
            const orc_ctf *ctf = &R->ctfs[c];
            uint32_t kb = ctf->key_bit / 8;
            for (uint32_t i = 0; i < ctf->kv_capacity; i++) {
                uint32_t pid, target, start, end;
                decode_packed(ctf, R->key_bytes + off + (uint64_t)i * kb, &pid, &target, &start, &end, m);
                q_match(qs, nq, &h, pid, target, start, end, 2 + r, m);
            }
            off += (uint64_t)ctf->kv_capacity * kb;
        }
        for (uint32_t i = 0; i < R->info.oversize_count; i++) {
            u128 k = R->o_keys[i];
            uint32_t pid = (uint32_t)((k >> 90) & ((1ull << OID_BIT) - 1));
            uint32_t target = (uint32_t)((k >> 64) & ((1ull << OID_BIT) - 1));
            uint32_t end = key_end(k) + R->info.end_time_min, start = end - key_dur(k);
            for (int j = 0; j < 4; j++) m[j] = (double)R->o_boxes[4 * i + j];
            q_match(qs, nq, &h, pid, target, start, end, 2 + r, m);
        }
    }
    return h.n;
}

/* ------------------------------------------------------------------ */
/* brute force pairs (independent of the tree)                         */
/* ------------------------------------------------------------------ */

typedef struct gcell { int64_t cell; uint32_t id; } gcell;
static int cmp_gcell(const void *a, const void *b) {
    const gcell *x = (const gcell *)a, *y = (const gcell *)b;
    if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
    return x->id < y->id ? -1 : x->id > y->id;
}

uint64_t orc_brute_pairs(const double *xy, uint32_t n, double r, int probe_low, uint64_t *keys, uint64_t cap) {
    if (n == 0) return 0;
    /* cells of 2r metres (with a margin) in degrees at the box's lowest latitude */
    double ymin = xy[1], ymax = xy[1], xmin = xy[0];
    for (uint32_t i = 0; i < n; i++) {
        if (xy[2 * i + 1] < ymin) ymin = xy[2 * i + 1];
        if (xy[2 * i + 1] > ymax) ymax = xy[2 * i + 1];
        if (xy[2 * i] < xmin) xmin = xy[2 * i];
    }
    double lat_hi = fabs(ymin) > fabs(ymax) ? fabs(ymin) : fabs(ymax);
    double cx = 2.5 * r * orc_dpm_lon(lat_hi), cy = 2.5 * r * DPM_LAT;
    gcell *g = (gcell *)malloc(sizeof(gcell) * n);
    for (uint32_t i = 0; i < n; i++) {
        int64_t gx = (int64_t)floor((xy[2 * i] - xmin) / cx);
        int64_t gy = (int64_t)floor((xy[2 * i + 1] - ymin) / cy);
        g[i].cell = gy * 4000003LL + gx;
        g[i].id = i;
    }
    qsort(g, n, sizeof(gcell), cmp_gcell);
    uint64_t cnt = 0;
    for (uint32_t a = 0; a < n; a++) {
        int64_t gx = (int64_t)floor((xy[2 * a] - xmin) / cx);
        int64_t gy = (int64_t)floor((xy[2 * a + 1] - ymin) / cy);
        for (int64_t dy = -1; dy <= 1; dy++)
            for (int64_t dx = -1; dx <= 1; dx++) {
                int64_t cell = (gy + dy) * 4000003LL + (gx + dx);
                uint32_t lo = 0, hi = n;
                while (lo < hi) {
                    uint32_t m = (lo + hi) / 2;
                    if (g[m].cell < cell) lo = m + 1; else hi = m;
                }
                for (uint32_t k = lo; k < n && g[k].cell == cell; k++) {
                    uint32_t b = g[k].id;
                    if (b <= a) continue;
                    uint32_t p = probe_low ? a : b, q = probe_low ? b : a;
                    double d = orc_point_distance(xy[2 * p], xy[2 * p + 1], xy[2 * q], xy[2 * q + 1]);
                    if (d <= r) {
                        if (cnt < cap) keys[cnt] = ((uint64_t)a << 32) | b;
                        cnt++;
                    }
                }
            }
    }
    free(g);
    return cnt;
}

/* the x86-64 conversions used by orc_ctf_bits, exposed for pinning */
uint32_t orc_conv_d2u32(double v) { return x86_d2u32(v); }
int32_t orc_conv_d2i32(double v) { return x86_d2i32(v); }
