/*
 * stc_flush.h — sort-based flush of the memtable into a run of CTFs.
 */
#ifndef STC_FLUSH_H_
#define STC_FLUSH_H_

#include "stc_internal.h"
#include "../../include/stcdb.h"
#include <vector>

struct FlushWork {
    uint32_t n = 0, ctf_count = 0;
    bool str_full = false;                         /* STR sorts over the full 64-bit keys (a compact key overflowed) */
    /* device scratch, grown on demand */
    uint64_t *mid = nullptr, *mid_alt = nullptr;   /* per-object STR keys */
    uint32_t *oids = nullptr, *oids_alt = nullptr;
    uint32_t *capacity = nullptr;                  /* per-CTF record counts */
    uint32_t *ctf_base = nullptr;                  /* per-CTF first record */
    uint64_t *ctf_byte_base = nullptr;             /* per-CTF first packed byte */
    stc_ctf_meta *ctfs = nullptr;
    uint8_t *bitmaps = nullptr;                    /* ctf_count * 8192 */
    uint32_t *obj_base = nullptr, *obj_fill = nullptr; /* per-object record slots */
    uint8_t *part_tail = nullptr;                  /* per object: records go to the oversize buffer */
    uint32_t *obj_key = nullptr, *obj_id = nullptr, *obj_cnt = nullptr; /* object placement order */
    uint32_t *scan_temp = nullptr;
    size_t scan_temp_words = 0;
    uint32_t *flags = nullptr;                     /* device counters of the flush */
    /* record-sized scratch */
    uint64_t rec_cap = 0;
    uint64_t *keys_out = nullptr;                  /* sorted keys (lo, hi): CTF records, then the oversize tail */
    float4 *boxes_out = nullptr;
    uint64_t *tkeys = nullptr;                     /* records grouped by object, before ranking */
    float4 *tboxes = nullptr;
    uint8_t *key_bytes = nullptr;
    uint64_t key_bytes_cap = 0;
    /* run hand-off: the payload is copied to pinned host memory on `cs`,
     * overlapping the following steps; the next flush waits for `copied`
     * before it rewrites the buffers the copy reads. The copies are queued
     * only when stc_flush_release is called (the caller picks a point where
     * the device is busy computing rather than writing memory: copies to the
     * host slow every concurrent writer) or at the next flush. */
    uint16_t *sids_out = nullptr;                  /* sids as of the flush */
    uint4 *chunks = nullptr;                       /* (ctf, first, count) work items of the key/bitmap pass */
    size_t chunks_cap = 0;
    uint8_t *o_bitmap_dev = nullptr;               /* oversize bitmap as of the flush */
    hipStream_t cs = nullptr;
    hipEvent_t ready = nullptr, copied = nullptr;
    bool pending = false;                          /* copies queued on cs, maybe still running */
    struct Job {                                   /* copies of the last flush not yet queued */
        uint8_t *dst;
        size_t key_bytes_at, bitmaps_at, o_bitmap_at, sids_at, o_keys_at, o_boxes_at, sorted_keys_at,
            sorted_boxes_at, counters_at;
        uint64_t key_bytes, nf, n_tail;
        uint32_t ctf_count, n;
        bool has_sorted;
        hipEvent_t done;                           /* the pinned block's event */
    } job{};
    bool job_set = false;
};

struct FlushArgs {
    uint32_t n, ctf_count;
    uint64_t kv_count;
    uint32_t kv_restriction;
    uint32_t end_time_min, cur_time, max_meet_time;
    uint32_t oversize_oid_count;
    uint32_t compat;
    uint32_t keep_sorted;    /* copy the sorted CTF records back to the host */
    uint32_t reset_memtable;  /* clear the per-object memtable state once the run is built, before its copies start */
    uint32_t sticky_oversize; /* compaction: objects entering with sids 1 stay oversize (compaction.cpp:259-268) */
    double area_threshold;
    double map[4];
    uint64_t *rec_keys;      /* (lo, hi) per record */
    float4 *rec_boxes;
    uint32_t *same_pid_count;
    uint4 *o_boxs;
    uint16_t *sids;
};

/* Offsets of a run's payload inside one pinned host block. */
struct RunLayout {
    size_t key_bytes, bitmaps, o_bitmap, sids, o_keys, o_boxes, sorted_keys, sorted_boxes, counters, total;
};

/* A pinned host block and the event that completes the copies into it. */
struct PinnedBlock {
    uint8_t *p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
};

struct FlushOut {
    stc_run_info info;
    std::vector<stc_ctf_meta> ctfs;   /* final except flag 8, which is in the copied counters */
    RunLayout lay;
    bool has_sorted = false;
    /* called once the payload size is known; returns a block of at least
     * `bytes` whose event the flush records after its copies */
    PinnedBlock (*acquire)(void *ctx, size_t bytes) = nullptr;
    void *acquire_ctx = nullptr;
    PinnedBlock blk;
};

/* FlushCounters::raster_capped of the copied counters (flag 8 of every CTF) */
uint32_t stc_flush_raster_capped(const void *counters);

int stc_flush_init(FlushWork *w, uint32_t n, uint32_t ctf_count);
void stc_flush_free(FlushWork *w);
int stc_flush_run(FlushWork *w, const FlushArgs *a, FlushOut *out, hipStream_t s);
/* queue the last flush's copies to its pinned block behind the work queued on s so far (no-op when none wait) */
int stc_flush_release(FlushWork *w, hipStream_t s);
/* forget the copies into block `dst` (its run was dropped before they were queued) */
void stc_flush_cancel(FlushWork *w, const uint8_t *dst);

/* CTF::get_ctf_bits (src/tracing/step_merge.cpp:140-186), host arithmetic */
void stc_ctf_bits(stc_ctf_meta *c, const double *map, uint32_t num_objects, uint32_t max_meet_time);

#endif
