/*
 * stc_oracle.h — CPU oracle for the per-timestep co-occurrence pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, single-threaded and
 * deterministically, what the reference's GPU path computes in
 * process_with_gpu (src/cuda/geometry.cu:1573-2218): the per-step QTree
 * rebuild, partition, border filtering, pair refinement, the meeting table,
 * meeting identification and the sort-based flush into a run of CTFs. It is
 * used as the checker by the tests, by __graft_entry__.smoke() and by the CPU
 * baseline leg of bench.py; the product never links or calls it.
 *
 * Parity decisions (see the file header of stc_oracle.c for the full list):
 *  - distances use the reference CPU formula order with no FMA contraction
 *    and glibc sin (src/geometry/geometry.h:40-48, src/util/util.h:59-66);
 *  - a meeting is a maximal run of consecutive timesteps in which the pair is
 *    found; the table never drops or splits a meeting;
 *  - a leaf stores its lowest `leaf_buffer` point ids when it overflows;
 *  - the flush follows the reference integer and float arithmetic exactly,
 *    with `compat` selecting the half-dropped CTF reduction of
 *    src/cuda/geometry.cu:865-888.
 */
#ifndef STC_ORACLE_H_
#define STC_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_config {
    uint32_t num_objects;
    uint32_t grid_capacity;   /* split threshold (config->grid_capacity) */
    uint32_t leaf_buffer;     /* stored points per leaf (grid_amplify * grid_capacity) */
    uint32_t min_meet_time;
    uint32_t max_meet_time;
    uint32_t kv_restriction;  /* flush when the memtable holds more records */
    uint32_t ctf_count;       /* must be a square */
    uint32_t start_time;
    uint32_t compat;          /* bit 0: reproduce the half-dropped CTF reductions */
    uint32_t schema_mode;     /* 0: QTree rebuilt every step (--disable_dynamic_schema,
                                    src/cuda/geometry.cu:1267-1393);
                                 1: dynamic schema: a tree built once from the first step's
                                    points (partitioner::build_schema, src/tracing/filter.cpp:27-93)
                                    and split/merged after every step by the GPU rules
                                    (cuda_update_schema_*, geometry.cu:1114-1234);
                                 2: as 1 with the CPU path's rules (update_schema_unit,
                                    filter.cpp:377-424) */
    double reach_distance;    /* metres */
    double area_threshold;    /* dProbFalse: object MBR area (deg^2) above which it is oversize */
    double map_lo[2];
    double map_hi[2];
    uint32_t schema_update_delay; /* config->schema_update_delay (config.h:38); 0 counts as 1 */
    uint32_t pad;
} orc_config;

/* CTF metadata, same fields and widths as the reference CTF minus its pointers
 * (src/tracing/step_merge.h:64-109). */
typedef struct orc_ctf {
    float mbr[4];             /* low0, low1, high0, high1 */
    uint32_t kv_capacity;
    uint32_t start_time_min, start_time_max;
    uint32_t end_time_min, end_time_max;
    uint32_t bitmap_size;
    uint16_t key_bit, id_bit, duration_bit, end_bit;
    uint16_t low_x_bit, low_y_bit, edge_bit, mbr_bit;
    uint16_t x_grid, y_grid;
    uint32_t flags;           /* bit 0: reference would fail an assert for this CTF */
} orc_ctf;

typedef struct orc_run_info {
    uint32_t kv_count;        /* records in the CTFs */
    uint32_t oversize_count;  /* records in the oversize buffer */
    uint32_t ctf_count;
    uint32_t num_objects;
    uint32_t o_start_min, o_start_max, o_end_min, o_end_max;
    uint32_t end_time_min;    /* time base of the end offsets in this run */
    uint32_t cur_time;        /* timestep at which the run was flushed */
    uint32_t sid_count, oversize_oid_count;
    uint64_t key_bytes;       /* packed key bytes over all CTFs */
} orc_run_info;

/* counters of the last processed step */
enum {
    ORC_C_NODES = 0, ORC_C_LEAVES, ORC_C_BORDER, ORC_C_UNITS_INSIDE, ORC_C_UNITS_CROSS,
    ORC_C_CANDIDATES, ORC_C_PAIRS, ORC_C_ACTIVE, ORC_C_EXPIRED, ORC_C_NEW_RECORDS,
    ORC_C_KV_COUNT, ORC_C_OVERFLOW, ORC_C_FLUSHED, ORC_C_NUM
};

void *orc_create(const orc_config *cfg);
/* dynamic schema: build the initial tree from xy now (otherwise the first
 * step builds it from its own points) */
int orc_build_schema(void *o, const double *xy);
void orc_destroy(void *o);
/* process one timestep; xy is N AoS (x, y) doubles in degrees. 0 ok, <0 error */
int orc_step(void *o, const double *xy, uint32_t t);
const char *orc_last_error(void *o);
void orc_counts(void *o, uint64_t *out /* ORC_C_NUM */);

/* pairs found in the last step, sorted, unique, key = oid1 << 32 | oid2 (oid1 < oid2) */
uint64_t orc_num_pairs(void *o);
void orc_get_pairs(void *o, uint64_t *keys);

/* leaves of the last step's QTree in canonical order (sorted by low y, low x);
 * boxes as low0, low1, high0, high1; member counts (before the buffer cap). */
uint64_t orc_num_leaves(void *o);
void orc_get_leaves(void *o, double *boxes, uint32_t *counts);
/* per point: canonical index of its leaf, and whether it is a border point */
void orc_get_point_leaf(void *o, uint32_t *leaf, uint8_t *border);
/* all node slots: boxes, type (0 free, 1 branch, 2 leaf), children as node
 * indices; node 0 is the root. With a dynamic schema these are the nodes
 * after the last step's update (the leaves above are the ones the step used). */
uint64_t orc_num_nodes(void *o);
void orc_get_nodes(void *o, double *boxes, double *mids, uint32_t *type, uint32_t *children);
/* check units of the last step: (pid, node index, inside) */
uint64_t orc_num_units(void *o);
void orc_get_units(void *o, uint32_t *pid, uint32_t *node, uint8_t *inside);

/* active meetings after the last step, sorted by key */
uint64_t orc_num_active(void *o);
void orc_get_active(void *o, uint64_t *keys, uint32_t *start, uint32_t *end, float *mbr);

/* memtable records (unsorted, emission order), keys as (lo64, hi64) pairs */
uint64_t orc_num_records(void *o);
void orc_get_records(void *o, uint64_t *keys, float *boxes);

/* flushed runs */
uint32_t orc_num_runs(void *o);
void orc_get_run_info(void *o, uint32_t run, orc_run_info *info);
void orc_get_run(void *o, uint32_t run, orc_ctf *ctfs, uint8_t *key_bytes, uint8_t *bitmaps,
                 uint16_t *sids, uint64_t *o_keys, float *o_boxes, uint8_t *o_bitmap,
                 uint64_t *sorted_keys, float *sorted_boxes);

/* helpers exposed for pinning tests */
double orc_dpm_lon(double lat);
double orc_point_distance(double x1, double y1, double x2, double y2);
double orc_box_distance(const double *box4, double px, double py);
uint32_t orc_min_bits(uint32_t v);
uint32_t orc_conv_d2u32(double v);
int32_t orc_conv_d2i32(double v);
void orc_ctf_bits(orc_ctf *ctf, const double *map4, uint32_t num_objects, uint32_t max_meet_time);
/* brute force: all unordered pairs (i < j) with distance measured from probe i
 * (probe_low = 1) or probe j (probe_low = 0) within r; returns count, writes
 * up to cap keys (unsorted). Uses a uniform grid to prune. */
/* Queries over the state after the last step (exact set semantics):
 *   kind 1, object history: meetings of `oid` (as record owner) whose
 *           [start, end] overlaps [t0, t1];
 *   kind 2, square range: meetings whose MBR intersects `box` (closed, in
 *           double) and whose time overlaps, each meeting once (pid < target);
 *   kind 3, pair lookup: meetings with pid == oid and target == oid2.
 * Sources: live meetings (start, end = last step, f32 MBR), memtable records
 * (end = offset + end_time_min, start = end - duration) and every run (keys
 * decoded as CTF::ctf_get_key_oid / parse_key / new_parse_mbr,
 * src/tracing/step_merge.cpp:325-379; oversize records unpacked). source: 0
 * live, 1 memtable, 2 + r run r. Hits beyond `cap` are counted, not stored. */
typedef struct orc_query {
    uint32_t kind, oid, oid2, t0, t1;
    uint32_t flags;   /* bit 0: count the query's hits, store none */
    double box[4];
} orc_query;
typedef struct orc_hit {
    uint32_t query, pid, target, start, end, source;
    double mbr[4];
} orc_hit;
uint64_t orc_query_batch(void *o, const orc_query *q, uint32_t nq, orc_hit *hits, uint64_t cap, uint64_t *counts);
uint32_t orc_end_time_min(void *o); /* time base of the memtable's end offsets */
/* meeting_unit (src/tracing/workbench.h:62-83) as persisted per second */
typedef struct orc_meeting_unit {
    uint64_t key;                /* Cantor(min pid, max pid) */
    uint32_t start, end;
    float mbr[4];
} orc_meeting_unit;
/* the last step's pairs as units, in pair order (orc_num_pairs of them) */
void orc_get_step_meetings(void *o, orc_meeting_unit *u);
/* a step from persisted units instead of positions */
int orc_step_meetings(void *o, const orc_meeting_unit *u, uint64_t n, uint32_t t);
/* merge runs [first, first + count) into one run in their place (tools/compaction.cpp) */
int orc_compact_runs(void *o, uint32_t first, uint32_t count);
/* wall-clock ms of the last flush a step ran, and its record count (the
 * single-thread CPU flush baseline: STR, record placement, widths, bitmaps,
 * key packing) */
double orc_last_flush_ms(void *o, uint64_t *records);

uint64_t orc_brute_pairs(const double *xy, uint32_t n, double r, int probe_low,
                         uint64_t *keys, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
