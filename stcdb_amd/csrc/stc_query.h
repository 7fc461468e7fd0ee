/*
 * stc_query.h — queries over the resident state (internal interface).
 *
 * The state after a step: live meetings in the table (end == last step),
 * memtable records, and the flushed runs kept in HBM (packed CTF keys, CTF
 * metadata, per-object sids, the oversize buffer). stc_query_run answers a
 * batch of object-history, square-range and pair queries with the set
 * semantics documented at stc_query_batch in include/stcdb.h.
 */
#ifndef STC_QUERY_H_
#define STC_QUERY_H_

#include "stc_internal.h"
#include "stc_table.h"
#include "../../include/stcdb.h"
#include <vector>

struct QueryRunDev {                 /* one run kept in HBM */
    const uint8_t *key_bytes;        /* packed CTF keys, CTF after CTF */
    const stc_ctf_meta *ctfs;        /* [ctf_count] */
    const uint64_t *byte_base;       /* [ctf_count + 1] first byte of each CTF */
    const uint32_t *rec_base;        /* [ctf_count + 1] first record of each CTF */
    const uint16_t *sids;            /* [num_objects] 0: no record, 1: oversize buffer, 2 + c: CTF c */
    const uint64_t *o_keys;          /* [2 * n_tail] (lo, hi) record keys of the oversize buffer */
    const float4 *o_boxes;           /* [n_tail] */
    uint32_t ctf_count, n_tail, end_time_min, source;
};

/* exact summary of a CTF's decoded records, for planning range queries */
struct CtfStats {
    double box[4];                   /* union of the decoded record boxes */
    uint32_t start_min, end_max;     /* over all records */
    uint32_t n_fwd;                  /* records with pid < target */
    uint32_t regular;                /* 1: every box has low <= high, every start <= end, no NaN */
};

struct QueryState {
    const Bucket *table;
    uint64_t nb;
    uint32_t last_t;
    uint32_t stepped;
    const uint64_t *rec_keys;        /* memtable records (lo, hi) */
    const float4 *rec_boxes;
    uint64_t n_rec;
    uint32_t end_time_min;           /* time base of the memtable's end offsets */
    uint32_t num_objects;
    double map[4];
    std::vector<QueryRunDev> runs;
    std::vector<const stc_ctf_meta *> run_ctfs;  /* host copies of each run's CTF metadata */
    std::vector<const CtfStats *> run_stats;     /* per run: [ctf_count] */
};

struct QueryBufs;                    /* device scratch of the query path, grown on demand */
QueryBufs *stc_query_bufs_new();
void stc_query_bufs_free(QueryBufs *b);
/* device bytes a batch of nq queries storing up to cap hits would add to b */
uint64_t stc_query_bufs_growth(const QueryBufs *b, uint32_t nq, uint64_t cap);
int stc_query_run(QueryBufs *b, const QueryState &st, const stc_query *q, uint32_t nq, stc_hit *hits, uint64_t cap,
                  uint64_t *n_hits, uint64_t *counts, hipStream_t s);
/* Compaction input: every record of run R (the kv_count records of its CTFs,
 * kv_count = the sum of their kv_capacity, then the oversize buffer) in
 * memtable form at keys/boxes[out0 ..], ends re-based on `base`, boxes decoded
 * as CTF::new_parse_mbr_f_box (step_merge.cpp:381-401); per object: record
 * counts, quantised MBRs, and sids[o] = 1 where R had o oversize. A record
 * whose ids are not below num_objects sets *bad and is not counted. */
int stc_compact_decode(const QueryRunDev &R, uint64_t kv_count, uint32_t num_objects, uint32_t base, uint64_t out0,
                       uint64_t *keys, float4 *boxes, uint32_t *same_pid_count, uint4 *o_boxs, uint16_t *sids,
                       uint32_t *bad, hipStream_t s);

/* decode every record of a run kept in HBM once and summarise each CTF */
int stc_query_ctf_stats(const QueryRunDev &run, const stc_ctf_meta *host_ctfs, std::vector<CtfStats> &out,
                        hipStream_t s);

#endif /* STC_QUERY_H_ */
