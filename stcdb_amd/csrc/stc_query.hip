/*
 * stc_query.hip — batched queries over the resident state.
 *
 * Object-history, square-range and pair queries (semantics at stc_query_batch in
 * include/stcdb.h) are answered from three sources in one pass each:
 *   - the meeting table: a grid-stride scan over the buckets, live buckets
 *     (end == last step) matched in both directions;
 *   - the memtable records: one thread per record;
 *   - the runs kept in HBM: object and pair queries look the object's CTF up
 *     through the run's sids and binary-search its packed keys, which a CTF
 *     stores sorted by pid (as CTF::search_SSTable does, step_merge.cpp:528);
 *     range queries scan the CTFs whose decoded boxes can reach the query box
 *     (one block per query and CTF) and the oversize buffer.
 * Object and pair queries are found per meeting through a small open-addressed
 * hash of the queried ids; range queries through a 64 x 64 grid of the map,
 * each (meeting, query) pair reported from the grid cell that holds the low
 * corner of their intersection, so exactly once.
 */
#include "stc_query.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#define QGRID 64
#define QLDS 4096 /* batches up to this many queries count in LDS, block by block */
#define MASK26 ((1u << 26) - 1u)

struct QDev {
    const stc_query *q;
    const uint32_t *oh_key;          /* queried ids, 0xFFFFFFFF = empty slot */
    const int32_t *oh_head;          /* first query of the id */
    const int32_t *q_next;           /* next query with the same id, -1 ends */
    uint32_t oh_bits;                /* 0: no object or pair queries */
    uint32_t n_range;
    const uint32_t *g_start;         /* [QGRID * QGRID + 1] range queries by grid cell */
    const uint32_t *g_list;
    double x0, y0, dx, dy;
    stc_hit *hits;
    uint64_t cap;                    /* 0: counts only */
    unsigned long long *n_hits;
    unsigned long long *counts;
    uint32_t nq;
    uint32_t lds;                    /* nq <= QLDS: per-block counters in LDS */
};

struct QueryBufs {
    stc_query *q = nullptr;
    uint32_t q_cap = 0;
    uint32_t *oh_key = nullptr;
    int32_t *oh_head = nullptr;
    uint32_t oh_cap = 0;
    int32_t *q_next = nullptr;
    uint32_t qn_cap = 0;
    uint32_t *oq = nullptr;          /* object and pair queries */
    uint32_t oq_cap = 0;
    uint32_t *g_start = nullptr;
    uint32_t *g_list = nullptr;
    uint32_t g_cap = 0;
    QueryRunDev *runs = nullptr;
    uint32_t runs_cap = 0;
    uint4 *cand = nullptr;
    uint64_t cand_cap = 0;
    stc_hit *hits = nullptr;
    uint64_t hits_cap = 0;
    unsigned long long *counters = nullptr; /* [0] total hits, [1 ..] per query */
    uint64_t counters_cap = 0;
};

QueryBufs *stc_query_bufs_new() { return new QueryBufs(); }

uint64_t stc_query_bufs_growth(const QueryBufs *b, uint32_t nq, uint64_t cap) {
    uint64_t need = 0;
    const uint64_t q_cap = b ? b->q_cap : 0, c_cap = b ? b->counters_cap : 0, h_cap = b ? b->hits_cap : 0;
    if (nq > q_cap) need += (uint64_t)nq * sizeof(stc_query) * 3 / 2;
    if (nq + 1ull > c_cap) need += (nq + 1ull) * 8 * 3 / 2;
    if (cap > h_cap) need += cap * sizeof(stc_hit) * 3 / 2;
    return need;
}

void stc_query_bufs_free(QueryBufs *b) {
    if (!b) return;
    void *ptrs[] = {b->q, b->oh_key, b->oh_head, b->q_next, b->oq, b->g_start, b->g_list, b->runs, b->cand,
                    b->hits, b->counters};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
}

template <typename T>
static int grow(T **p, uint64_t *cap, uint64_t need) {
    if (*p && *cap >= need) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    uint64_t nc = std::max<uint64_t>(need + need / 2, 64);
    if (hipMalloc((void **)p, nc * sizeof(T)) != hipSuccess) {
        *p = nullptr;
        return -1;
    }
    *cap = nc;
    return 0;
}
template <typename T>
static int grow32(T **p, uint32_t *cap, uint64_t need) {
    uint64_t c = *cap;
    int rc = grow(p, &c, need);
    *cap = (uint32_t)std::min<uint64_t>(c, 0xFFFFFFFFu);
    return rc;
}

/* the same arithmetic on the host (registration) and the device (lookup) */
__host__ __device__ inline int grid_cell(double v, double v0, double d) {
    double f = floor((v - v0) / d);
    if (!(f >= 0.0)) return 0;
    if (f > (double)(QGRID - 1)) return QGRID - 1;
    return (int)f;
}

__host__ __device__ inline uint32_t oid_slot(uint32_t oid, uint32_t bits) {
    return (uint32_t)(((uint64_t)oid * 0x9E3779B97F4A7C15ull) >> (64 - bits));
}

/* per-block query counters: cleared on entry, added to the global counts on
 * exit (one atomic per query and block instead of one per hit) */
__device__ inline void lc_clear(const QDev &Q, unsigned int *lc) {
    if (Q.lds)
        for (uint32_t i = threadIdx.x; i < Q.nq; i += blockDim.x) lc[i] = 0;
    __syncthreads();
}
__device__ inline void lc_flush(const QDev &Q, unsigned int *lc) {
    __syncthreads();
    if (Q.lds)
        for (uint32_t i = threadIdx.x; i < Q.nq; i += blockDim.x)
            if (lc[i]) atomicAdd(&Q.counts[i], (unsigned long long)lc[i]);
}

__device__ inline void q_emit(const QDev &Q, unsigned int *lc, uint32_t qi, uint32_t pid, uint32_t target,
                              uint32_t start, uint32_t end, uint32_t source, const double *m) {
    if (Q.lds) atomicAdd(&lc[qi], 1u);
    else atomicAdd(&Q.counts[qi], 1ull);
    if (!Q.cap || (Q.q[qi].flags & STC_QF_COUNT)) return;
    /* one reservation per call for all lanes that emit together: a counter
     * updated by every hit serialises on its address */
    const uint64_t act = __ballot(1);
    const int lane = threadIdx.x & (STC_WAVE - 1), leader = __ffsll((long long)act) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(Q.n_hits, (unsigned long long)__popcll(act));
    base = __shfl(base, leader, STC_WAVE);
    const unsigned long long k = base + (unsigned long long)__popcll(act & ((1ull << lane) - 1ull));
    if (k < Q.cap) {
        stc_hit h;
        h.query = qi;
        h.pid = pid;
        h.target = target;
        h.start = start;
        h.end = end;
        h.source = source;
        h.mbr[0] = m[0];
        h.mbr[1] = m[1];
        h.mbr[2] = m[2];
        h.mbr[3] = m[3];
        Q.hits[k] = h;
    }
}

__device__ inline bool q_time(const stc_query &q, uint32_t s, uint32_t e) { return s <= q.t1 && e >= q.t0; }
__device__ inline bool q_box(const stc_query &q, const double *m) {
    return m[0] <= q.box[2] && m[2] >= q.box[0] && m[1] <= q.box[3] && m[3] >= q.box[1];
}

/* object and pair queries of one meeting seen from pid */
__device__ inline void match_oid(const QDev &Q, unsigned int *lc, uint32_t pid, uint32_t target, uint32_t s,
                                 uint32_t e, uint32_t src, const double *m) {
    if (!Q.oh_bits) return;
    const uint32_t mask = (1u << Q.oh_bits) - 1u;
    uint32_t h = oid_slot(pid, Q.oh_bits);
    for (;;) {
        uint32_t k = Q.oh_key[h];
        if (k == pid) break;
        if (k == 0xFFFFFFFFu) return;
        h = (h + 1) & mask;
    }
    for (int32_t qi = Q.oh_head[h]; qi >= 0; qi = Q.q_next[qi]) {
        const stc_query &q = Q.q[qi];
        if (!q_time(q, s, e)) continue;
        if (q.kind == STC_Q_OID || (q.kind == STC_Q_PAIR && q.oid2 == target))
            q_emit(Q, lc, (uint32_t)qi, pid, target, s, e, src, m);
    }
}

/* range queries of one meeting (called once per meeting) */
__device__ inline void match_range(const QDev &Q, unsigned int *lc, uint32_t pid, uint32_t target, uint32_t s,
                                   uint32_t e, uint32_t src, const double *m) {
    if (!Q.n_range) return;
    const int cx0 = grid_cell(m[0], Q.x0, Q.dx), cx1 = grid_cell(m[2], Q.x0, Q.dx);
    const int cy0 = grid_cell(m[1], Q.y0, Q.dy), cy1 = grid_cell(m[3], Q.y0, Q.dy);
    for (int cy = cy0; cy <= cy1; cy++)
        for (int cx = cx0; cx <= cx1; cx++) {
            const uint32_t cell = (uint32_t)(cy * QGRID + cx);
            for (uint32_t j = Q.g_start[cell]; j < Q.g_start[cell + 1]; j++) {
                const uint32_t qi = Q.g_list[j];
                const stc_query &q = Q.q[qi];
                if (!q_time(q, s, e) || !q_box(q, m)) continue;
                if (grid_cell(fmax(m[0], q.box[0]), Q.x0, Q.dx) != cx || grid_cell(fmax(m[1], q.box[1]), Q.y0, Q.dy) != cy)
                    continue;
                q_emit(Q, lc, qi, pid, target, s, e, src, m);
            }
        }
}

/* CTF::ctf_get_key_oid, parse_key and new_parse_mbr (step_merge.cpp:325-379)
 * on the little-endian packed key at p */
__device__ inline unsigned __int128 load_packed(const uint8_t *p, uint32_t kb) {
    unsigned __int128 k = 0;
    for (uint32_t i = 0; i < kb && i < 16; i++) k |= (unsigned __int128)p[i] << (8 * i);
    return k;
}

__device__ inline uint32_t packed_pid(const stc_ctf_meta &c, unsigned __int128 k) {
    return (uint32_t)((k >> (c.id_bit + c.duration_bit + c.end_bit + c.mbr_bit)) & ((1ull << c.id_bit) - 1));
}

__device__ inline void decode_packed(const stc_ctf_meta &c, unsigned __int128 k, uint32_t *pid, uint32_t *target,
                                     uint32_t *start, uint32_t *end, double *box) {
    uint64_t vm = (uint64_t)(k & (((unsigned __int128)1 << c.mbr_bit) - 1));
    unsigned __int128 r = k >> c.mbr_bit;
    uint32_t e = (uint32_t)(r & ((1ull << c.end_bit) - 1));
    r >>= c.end_bit;
    uint32_t d = (uint32_t)(r & ((1ull << c.duration_bit) - 1));
    r >>= c.duration_bit;
    *target = (uint32_t)(r & ((1ull << c.id_bit) - 1));
    r >>= c.id_bit;
    *pid = (uint32_t)(r & ((1ull << c.id_bit) - 1));
    *end = e + c.end_time_min;
    *start = *end - d;
    const uint64_t mask_edge = (1ull << c.edge_bit) - 1, mask_low_y = (1ull << c.low_y_bit) - 1,
                   mask_low_x = (1ull << c.low_x_bit) - 1;
    uint64_t y = vm & mask_edge;
    vm >>= c.edge_bit;
    uint64_t x = vm & mask_edge;
    vm >>= c.edge_bit;
    uint64_t low1 = vm & mask_low_y;
    vm >>= c.low_y_bit;
    uint64_t low0 = vm & mask_low_x;
    const float w0 = c.mbr[2] - c.mbr[0], w1 = c.mbr[3] - c.mbr[1];
    const float longer = w0 > w1 ? w0 : w1;
    box[0] = (double)c.mbr[0] + (double)low0 / (double)mask_low_x * (double)w0;
    box[1] = (double)c.mbr[1] + (double)low1 / (double)mask_low_y * (double)w1;
    box[2] = box[0] + (double)x / (double)mask_edge * (double)longer;
    box[3] = box[1] + (double)y / (double)mask_edge * (double)longer;
}

__device__ inline void decode_tail(const uint64_t *o_keys, const float4 *o_boxes, uint32_t i, uint32_t base,
                                   uint32_t *pid, uint32_t *target, uint32_t *start, uint32_t *end, double *m) {
    uint64_t lo = o_keys[2 * i], hi = o_keys[2 * i + 1];
    *pid = (uint32_t)(hi >> 26) & MASK26;
    *target = (uint32_t)hi & MASK26;
    *end = (uint32_t)lo + base;
    *start = *end - (uint32_t)(lo >> 32);
    float4 b = o_boxes[i];
    m[0] = b.x;
    m[1] = b.y;
    m[2] = b.z;
    m[3] = b.w;
}

__global__ __launch_bounds__(256) void k_q_live(const Bucket *__restrict__ t, uint64_t nb, uint32_t last_t, QDev Q) {
    __shared__ unsigned int lc[QLDS];
    lc_clear(Q, lc);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
        Bucket b = t[i];
        if (!bucket_live(b)) continue;
        const uint32_t a = (uint32_t)(b.key >> 32), c = (uint32_t)b.key;
        double m[4] = {o2f(b.mbr[0]), o2f(b.mbr[1]), o2f(b.mbr[2]), o2f(b.mbr[3])};
        /* a live meeting ends at the step that built the array */
        match_oid(Q, lc, a, c, b.start, last_t, 0, m);
        match_oid(Q, lc, c, a, b.start, last_t, 0, m);
        match_range(Q, lc, a, c, b.start, last_t, 0, m);
    }
    lc_flush(Q, lc);
}

__global__ __launch_bounds__(256) void k_q_mem(const uint64_t *__restrict__ keys, const float4 *__restrict__ boxes,
                                               uint64_t n, uint32_t base, QDev Q) {
    __shared__ unsigned int lc[QLDS];
    lc_clear(Q, lc);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t pid, target, start, end;
        double m[4];
        decode_tail(keys, boxes, (uint32_t)i, base, &pid, &target, &start, &end, m);
        match_oid(Q, lc, pid, target, start, end, 1, m);
        if (pid < target) match_range(Q, lc, pid, target, start, end, 1, m);
    }
    lc_flush(Q, lc);
}

/* object and pair queries against one run: the object's CTF by its sid, then a
 * binary search of the CTF's pid-sorted packed keys; oversize objects (sid 1)
 * scan the oversize buffer */
__device__ inline void run_oid(const QDev &Q, unsigned int *lc, const QueryRunDev &R, uint32_t qi, uint32_t n_objects) {
    const stc_query q = Q.q[qi];
    if (q.oid >= n_objects) return;
    const uint16_t sid = R.sids[q.oid];
    uint32_t pid, target, start, end;
    double m[4];
    if (sid == 0) return;
    if (sid == 1) {
        for (uint32_t i = 0; i < R.n_tail; i++) {
            decode_tail(R.o_keys, R.o_boxes, i, R.end_time_min, &pid, &target, &start, &end, m);
            if (pid != q.oid || !q_time(q, start, end)) continue;
            if (q.kind == STC_Q_OID || target == q.oid2) q_emit(Q, lc, qi, pid, target, start, end, R.source, m);
        }
        return;
    }
    const uint32_t c = sid - 2u;
    if (c >= R.ctf_count) return;
    const stc_ctf_meta ctf = R.ctfs[c];
    const uint32_t kb = ctf.key_bit / 8;
    const uint8_t *base = R.key_bytes + R.byte_base[c];
    uint32_t lo = 0, hi = ctf.kv_capacity;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (packed_pid(ctf, load_packed(base + (uint64_t)mid * kb, kb)) < q.oid) lo = mid + 1;
        else hi = mid;
    }
    for (uint32_t i = lo; i < ctf.kv_capacity; i++) {
        unsigned __int128 k = load_packed(base + (uint64_t)i * kb, kb);
        if (packed_pid(ctf, k) != q.oid) break;
        decode_packed(ctf, k, &pid, &target, &start, &end, m);
        if (!q_time(q, start, end)) continue;
        if (q.kind == STC_Q_OID || target == q.oid2) q_emit(Q, lc, qi, pid, target, start, end, R.source, m);
    }
}

__global__ __launch_bounds__(256) void k_q_run_oid(const QueryRunDev *__restrict__ runs, uint32_t nruns,
                                                   const uint32_t *__restrict__ oq, uint32_t noq, uint32_t n_objects,
                                                   QDev Q) {
    __shared__ unsigned int lc[QLDS];
    lc_clear(Q, lc);
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < (uint64_t)noq * nruns) run_oid(Q, lc, runs[g / noq], oq[g % noq], n_objects);
    lc_flush(Q, lc);
}

/* range queries: one block per (query, run, CTF) candidate; CTF 0xFFFFFFFF is
 * the run's oversize buffer */
__global__ __launch_bounds__(256) void k_q_run_range(const QueryRunDev *__restrict__ runs,
                                                     const uint4 *__restrict__ cand, QDev Q) {
    __shared__ unsigned int lc[QLDS];
    lc_clear(Q, lc);
    const uint4 cd = cand[blockIdx.x];
    const uint32_t qi = cd.x;
    const stc_query q = Q.q[qi];
    const QueryRunDev R = runs[cd.y];
    uint32_t pid, target, start, end;
    double m[4];
    if (cd.z == 0xFFFFFFFFu) {
        for (uint32_t i = threadIdx.x; i < R.n_tail; i += blockDim.x) {
            decode_tail(R.o_keys, R.o_boxes, i, R.end_time_min, &pid, &target, &start, &end, m);
            if (pid < target && q_time(q, start, end) && q_box(q, m))
                q_emit(Q, lc, qi, pid, target, start, end, R.source, m);
        }
    } else {
        const stc_ctf_meta ctf = R.ctfs[cd.z];
        const uint32_t kb = ctf.key_bit / 8;
        const uint8_t *base = R.key_bytes + R.byte_base[cd.z];
        for (uint32_t i = threadIdx.x; i < ctf.kv_capacity; i += blockDim.x) {
            decode_packed(ctf, load_packed(base + (uint64_t)i * kb, kb), &pid, &target, &start, &end, m);
            if (pid < target && q_time(q, start, end) && q_box(q, m))
                q_emit(Q, lc, qi, pid, target, start, end, R.source, m);
        }
    }
    lc_flush(Q, lc);
}

/* one block per CTF: exact union box, time span and forward-record count */
#define STATS_CHUNK 4096u /* records of a CTF one block summarises */

/* one block per (CTF, chunk of STATS_CHUNK records): a partial summary,
 * combined per CTF on the host */
__global__ __launch_bounds__(256) void k_ctf_stats(QueryRunDev R, const uint2 *__restrict__ items,
                                                   CtfStats *out) {
    __shared__ double sb[4][256];
    __shared__ uint32_t su[4][256];
    const uint32_t c = items[blockIdx.x].x, first = items[blockIdx.x].y, tid = threadIdx.x;
    const stc_ctf_meta ctf = R.ctfs[c];
    const uint32_t kb = ctf.key_bit / 8;
    const uint8_t *base = R.key_bytes + R.byte_base[c];
    const uint32_t last = min(ctf.kv_capacity, first + STATS_CHUNK);
    double b0 = INFINITY, b1 = INFINITY, b2 = -INFINITY, b3 = -INFINITY;
    uint32_t smin = 0xFFFFFFFFu, emax = 0, fwd = 0, irregular = 0;
    for (uint32_t i = first + tid; i < last; i += blockDim.x) {
        uint32_t pid, target, start, end;
        double m[4];
        decode_packed(ctf, load_packed(base + (uint64_t)i * kb, kb), &pid, &target, &start, &end, m);
        if (!(m[0] <= m[2]) || !(m[1] <= m[3]) || start > end) irregular = 1; /* also catches NaN */
        b0 = fmin(b0, m[0]);
        b1 = fmin(b1, m[1]);
        b2 = fmax(b2, m[2]);
        b3 = fmax(b3, m[3]);
        smin = min(smin, start);
        emax = max(emax, end);
        fwd += pid < target;
    }
    sb[0][tid] = b0;
    sb[1][tid] = b1;
    sb[2][tid] = b2;
    sb[3][tid] = b3;
    su[0][tid] = smin;
    su[1][tid] = emax;
    su[2][tid] = fwd;
    su[3][tid] = irregular;
    __syncthreads();
    for (uint32_t o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) {
            sb[0][tid] = fmin(sb[0][tid], sb[0][tid + o]);
            sb[1][tid] = fmin(sb[1][tid], sb[1][tid + o]);
            sb[2][tid] = fmax(sb[2][tid], sb[2][tid + o]);
            sb[3][tid] = fmax(sb[3][tid], sb[3][tid + o]);
            su[0][tid] = min(su[0][tid], su[0][tid + o]);
            su[1][tid] = max(su[1][tid], su[1][tid + o]);
            su[2][tid] += su[2][tid + o];
            su[3][tid] |= su[3][tid + o];
        }
        __syncthreads();
    }
    if (tid == 0) {
        CtfStats st;
        st.box[0] = sb[0][0];
        st.box[1] = sb[1][0];
        st.box[2] = sb[2][0];
        st.box[3] = sb[3][0];
        st.start_min = su[0][0];
        st.end_max = su[1][0];
        st.n_fwd = su[2][0];
        st.regular = !su[3][0];
        out[blockIdx.x] = st;
    }
}

static uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)std::max<uint64_t>((n + b - 1) / b, 1); }

#define QCK(expr)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return stc_hip_fail(e_, #expr, __FILE__, __LINE__);                   \
    } while (0)

/* CTF::new_parse_mbr_f_box: the double expressions of new_parse_mbr stored to
 * float, the high corner computed from the stored (float) low corner */
__device__ inline float4 decode_fbox(const stc_ctf_meta &c, uint64_t vm) {
    const uint64_t mask_edge = (1ull << c.edge_bit) - 1, mask_low_y = (1ull << c.low_y_bit) - 1,
                   mask_low_x = (1ull << c.low_x_bit) - 1;
    uint64_t y = vm & mask_edge;
    vm >>= c.edge_bit;
    uint64_t x = vm & mask_edge;
    vm >>= c.edge_bit;
    uint64_t low1 = vm & mask_low_y;
    vm >>= c.low_y_bit;
    uint64_t low0 = vm & mask_low_x;
    const float w0 = c.mbr[2] - c.mbr[0], w1 = c.mbr[3] - c.mbr[1];
    const float longer = w0 > w1 ? w0 : w1;
    float4 b;
    b.x = (float)((double)c.mbr[0] + (double)low0 / (double)mask_low_x * (double)w0);
    b.y = (float)((double)c.mbr[1] + (double)low1 / (double)mask_low_y * (double)w1);
    b.z = (float)((double)b.x + (double)x / (double)mask_edge * (double)longer);
    b.w = (float)((double)b.y + (double)y / (double)mask_edge * (double)longer);
    return b;
}

/* A run's records are grouped by object, so the lanes of a wave mostly share
 * a few objects: their per-object counts and box bounds are combined across
 * each run of equal ids first, and the run's first lane makes the atomic
 * updates (64 updates of one address would be serialised). */
__global__ __launch_bounds__(256) void k_compact_decode(QueryRunDev R, uint64_t kv_count, uint32_t num_objects,
                                                        uint32_t base, uint64_t out0, uint64_t *__restrict__ keys,
                                                        float4 *__restrict__ boxes, uint32_t *__restrict__ same_pid_count,
                                                        uint4 *__restrict__ o_boxs, uint32_t *__restrict__ bad) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (STC_WAVE - 1);
    uint32_t pid = 0xFFFFFFFFu; /* no object: lanes past the end and bad records */
    uint32_t cnt = 0;
    uint4 q = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
    if (i < kv_count + R.n_tail) {
        uint32_t p, target, dur, end;
        float4 b;
        if (i < kv_count) {
            /* the record's CTF: the last c with rec_base[c] <= i */
            uint32_t lo = 0, hi = R.ctf_count;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) / 2;
                if (R.rec_base[mid] <= i) lo = mid;
                else hi = mid;
            }
            const stc_ctf_meta c = R.ctfs[lo];
            const uint32_t kb = c.key_bit / 8;
            unsigned __int128 k = load_packed(R.key_bytes + R.byte_base[lo] + (i - R.rec_base[lo]) * kb, kb);
            const uint64_t vm = (uint64_t)(k & (((unsigned __int128)1 << c.mbr_bit) - 1));
            unsigned __int128 r = k >> c.mbr_bit;
            end = (uint32_t)(r & ((1ull << c.end_bit) - 1)) + c.end_time_min;
            r >>= c.end_bit;
            dur = (uint32_t)(r & ((1ull << c.duration_bit) - 1));
            r >>= c.duration_bit;
            target = (uint32_t)(r & ((1ull << c.id_bit) - 1));
            r >>= c.id_bit;
            p = (uint32_t)(r & ((1ull << c.id_bit) - 1));
            b = decode_fbox(c, vm);
        } else {
            const uint64_t j = i - kv_count;
            const uint64_t lo = R.o_keys[2 * j], hi = R.o_keys[2 * j + 1];
            p = (uint32_t)(hi >> 26) & MASK26;
            target = (uint32_t)hi & MASK26;
            end = (uint32_t)lo + R.end_time_min;
            dur = (uint32_t)(lo >> 32);
            b = R.o_boxes[j];
        }
        keys[2 * (out0 + i)] = ((uint64_t)dur << 32) | (uint32_t)(end - base);
        keys[2 * (out0 + i) + 1] = ((uint64_t)p << 26) | target;
        boxes[out0 + i] = b;
        if (p >= num_objects || target >= num_objects) { /* not a key this context wrote */
            atomicOr(bad, 1u);
        } else {
            pid = p;
            cnt = 1;
            q = make_uint4(f2q(b.x), f2q(b.y), f2q(b.z), f2q(b.w));
        }
    }
    /* lane l ends up holding the combination over lanes [l, end of its run of equal ids) */
    for (int o = 1; o < STC_WAVE; o <<= 1) {
        const uint32_t op = __shfl_down(pid, o, STC_WAVE), oc = __shfl_down(cnt, o, STC_WAVE);
        const uint32_t o0 = __shfl_down(q.x, o, STC_WAVE), o1 = __shfl_down(q.y, o, STC_WAVE);
        const uint32_t o2 = __shfl_down(q.z, o, STC_WAVE), o3 = __shfl_down(q.w, o, STC_WAVE);
        if (lane + o < STC_WAVE && op == pid) {
            cnt += oc;
            q.x = min(q.x, o0);
            q.y = min(q.y, o1);
            q.z = max(q.z, o2);
            q.w = max(q.w, o3);
        }
    }
    const uint32_t prev = __shfl_up(pid, 1, STC_WAVE);
    if (pid != 0xFFFFFFFFu && (lane == 0 || prev != pid)) {
        atomicAdd(&same_pid_count[pid], cnt);
        atomicMin(&o_boxs[pid].x, q.x);
        atomicMax(&o_boxs[pid].z, q.z);
        atomicMin(&o_boxs[pid].y, q.y);
        atomicMax(&o_boxs[pid].w, q.w);
    }
}

__global__ void k_compact_sids(const uint16_t *__restrict__ run_sids, uint32_t n, uint16_t *__restrict__ sids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && run_sids[i] == 1) sids[i] = 1;
}

int stc_compact_decode(const QueryRunDev &R, uint64_t kv_count, uint32_t num_objects, uint32_t base, uint64_t out0,
                       uint64_t *keys, float4 *boxes, uint32_t *same_pid_count, uint4 *o_boxs, uint16_t *sids,
                       uint32_t *bad, hipStream_t s) {
    if (kv_count + R.n_tail)
        hipLaunchKernelGGL(k_compact_decode, dim3(blocks(kv_count + R.n_tail, 256)), dim3(256), 0, s, R, kv_count,
                           num_objects, base, out0, keys, boxes, same_pid_count, o_boxs, bad);
    QCK(hipGetLastError());
    hipLaunchKernelGGL(k_compact_sids, dim3(blocks(num_objects, 256)), dim3(256), 0, s, R.sids, num_objects, sids);
    QCK(hipGetLastError());
    return STC_OK;
}

int stc_query_ctf_stats(const QueryRunDev &run, const stc_ctf_meta *host_ctfs, std::vector<CtfStats> &out,
                        hipStream_t s) {
    out.assign(run.ctf_count, CtfStats());
    if (!run.ctf_count) return STC_OK;
    /* work items: every CTF in chunks of STATS_CHUNK records (an empty CTF still gets one) */
    std::vector<uint2> items;
    std::vector<uint32_t> first_item(run.ctf_count);
    for (uint32_t c = 0; c < run.ctf_count; c++) {
        first_item[c] = (uint32_t)items.size();
        const uint32_t cap = host_ctfs[c].kv_capacity;
        for (uint32_t k = 0; k == 0 || k < cap; k += STATS_CHUNK) items.push_back(make_uint2(c, k));
    }
    const uint32_t ni = (uint32_t)items.size();
    CtfStats *d = nullptr;
    uint2 *di = nullptr;
    QCK(hipMalloc((void **)&d, ni * sizeof(CtfStats)));
    hipError_t e = hipMalloc((void **)&di, ni * sizeof(uint2));
    std::vector<CtfStats> part(ni);
    if (e == hipSuccess) e = hipMemcpyAsync(di, items.data(), ni * sizeof(uint2), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_ctf_stats, dim3(ni), dim3(256), 0, s, run, (const uint2 *)di, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(part.data(), d, ni * sizeof(CtfStats), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (di) (void)hipFree(di);
    if (e != hipSuccess) return stc_hip_fail(e, "CTF statistics", __FILE__, __LINE__);
    for (uint32_t c = 0; c < run.ctf_count; c++) {
        const uint32_t k1 = c + 1 < run.ctf_count ? first_item[c + 1] : ni;
        CtfStats st = part[first_item[c]];
        for (uint32_t k = first_item[c] + 1; k < k1; k++) {
            const CtfStats &p = part[k];
            st.box[0] = std::fmin(st.box[0], p.box[0]);
            st.box[1] = std::fmin(st.box[1], p.box[1]);
            st.box[2] = std::fmax(st.box[2], p.box[2]);
            st.box[3] = std::fmax(st.box[3], p.box[3]);
            st.start_min = std::min(st.start_min, p.start_min);
            st.end_max = std::max(st.end_max, p.end_max);
            st.n_fwd += p.n_fwd;
            st.regular = st.regular && p.regular;
        }
        out[c] = st;
    }
    return STC_OK;
}

/* the decoded boxes of a CTF lie in [low, low + |extent| + longer edge] per
 * axis (new_parse_mbr): a conservative test, widened by 1e-6 degrees for the
 * float/double rounding of the bounds */
static bool ctf_may_intersect(const stc_ctf_meta &c, const double *box) {
    const double w0 = (double)(c.mbr[2] - c.mbr[0]), w1 = (double)(c.mbr[3] - c.mbr[1]);
    const double L = std::max(std::fabs(w0), std::fabs(w1)), eps = 1e-6;
    const double lo0 = (double)c.mbr[0] - std::fabs(w0) - eps, hi0 = (double)c.mbr[0] + std::fabs(w0) + L + eps;
    const double lo1 = (double)c.mbr[1] - std::fabs(w1) - eps, hi1 = (double)c.mbr[1] + std::fabs(w1) + L + eps;
    if (!(lo0 == lo0 && hi0 == hi0 && lo1 == lo1 && hi1 == hi1)) return true; /* NaN bounds: scan */
    return box[0] <= hi0 && box[2] >= lo0 && box[1] <= hi1 && box[3] >= lo1;
}

int stc_query_run(QueryBufs *B, const QueryState &st, const stc_query *q, uint32_t nq, stc_hit *hits, uint64_t cap,
                  uint64_t *n_hits, uint64_t *counts, hipStream_t s) {
    if (nq == 0) {
        *n_hits = 0;
        return STC_OK;
    }
    /* queries */
    if (grow32(&B->q, &B->q_cap, nq) || grow(&B->counters, &B->counters_cap, (uint64_t)nq + 1) ||
        grow(&B->hits, &B->hits_cap, std::max<uint64_t>(cap, 1)))
        return stc_hip_fail(hipErrorOutOfMemory, "query buffers", __FILE__, __LINE__);
    QCK(hipMemcpyAsync(B->q, q, nq * sizeof(stc_query), hipMemcpyHostToDevice, s));
    QCK(hipMemsetAsync(B->counters, 0, ((uint64_t)nq + 1) * sizeof(unsigned long long), s));

    /* hash of the queried ids (object and pair queries) */
    std::vector<uint32_t> oq;
    for (uint32_t i = 0; i < nq; i++)
        if (q[i].kind == STC_Q_OID || q[i].kind == STC_Q_PAIR) oq.push_back(i);
    uint32_t bits = 0;
    if (!oq.empty()) {
        bits = 6;
        while ((1ull << bits) < 2ull * oq.size()) bits++;
        const uint32_t size = 1u << bits, mask = size - 1;
        std::vector<uint32_t> key(size, 0xFFFFFFFFu);
        std::vector<int32_t> head(size, -1), next(nq, -1);
        for (size_t j = oq.size(); j-- > 0;) { /* insert in reverse: lists keep query order */
            uint32_t qi = oq[j], id = q[qi].oid, h = oid_slot(id, bits);
            while (key[h] != 0xFFFFFFFFu && key[h] != id) h = (h + 1) & mask;
            key[h] = id;
            next[qi] = head[h];
            head[h] = (int32_t)qi;
        }
        uint32_t hc = B->oh_cap, nc = B->oh_cap;
        if (grow32(&B->oh_key, &hc, size) || grow32(&B->oh_head, &nc, size) ||
            grow32(&B->q_next, &B->qn_cap, nq) || grow32(&B->oq, &B->oq_cap, oq.size()))
            return stc_hip_fail(hipErrorOutOfMemory, "query id hash", __FILE__, __LINE__);
        B->oh_cap = std::min(hc, nc);
        QCK(hipMemcpy(B->oh_key, key.data(), size * sizeof(uint32_t), hipMemcpyHostToDevice));
        QCK(hipMemcpy(B->oh_head, head.data(), size * sizeof(int32_t), hipMemcpyHostToDevice));
        QCK(hipMemcpy(B->q_next, next.data(), nq * sizeof(int32_t), hipMemcpyHostToDevice));
        QCK(hipMemcpy(B->oq, oq.data(), oq.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }

    /* grid of the range queries */
    const double x0 = st.map[0], y0 = st.map[1];
    const double dx = (st.map[2] - st.map[0]) / QGRID, dy = (st.map[3] - st.map[1]) / QGRID;
    std::vector<uint32_t> rq;
    for (uint32_t i = 0; i < nq; i++)
        if (q[i].kind == STC_Q_RANGE) rq.push_back(i);
    if (!rq.empty()) {
        std::vector<uint32_t> start(QGRID * QGRID + 1, 0), list;
        std::vector<std::vector<uint32_t>> cells(QGRID * QGRID);
        for (uint32_t qi : rq) {
            const double *b = q[qi].box;
            int cx0 = grid_cell(b[0], x0, dx), cx1 = grid_cell(b[2], x0, dx);
            int cy0 = grid_cell(b[1], y0, dy), cy1 = grid_cell(b[3], y0, dy);
            for (int cy = cy0; cy <= cy1; cy++)
                for (int cx = cx0; cx <= cx1; cx++) cells[cy * QGRID + cx].push_back(qi);
        }
        for (int c = 0; c < QGRID * QGRID; c++) {
            start[c + 1] = start[c] + (uint32_t)cells[c].size();
            list.insert(list.end(), cells[c].begin(), cells[c].end());
        }
        uint32_t gc = B->g_cap;
        if (grow32(&B->g_list, &gc, std::max<size_t>(list.size(), 1)))
            return stc_hip_fail(hipErrorOutOfMemory, "query grid", __FILE__, __LINE__);
        B->g_cap = gc;
        if (!B->g_start) QCK(hipMalloc((void **)&B->g_start, (QGRID * QGRID + 1) * sizeof(uint32_t)));
        QCK(hipMemcpy(B->g_start, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
        if (!list.empty()) QCK(hipMemcpy(B->g_list, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }

    QDev Q;
    Q.q = B->q;
    Q.oh_key = B->oh_key;
    Q.oh_head = B->oh_head;
    Q.q_next = B->q_next;
    Q.oh_bits = bits;
    Q.n_range = (uint32_t)rq.size();
    Q.g_start = B->g_start;
    Q.g_list = B->g_list;
    Q.x0 = x0;
    Q.y0 = y0;
    Q.dx = dx;
    Q.dy = dy;
    Q.hits = B->hits;
    Q.cap = hits ? cap : 0;
    Q.n_hits = B->counters;
    Q.counts = B->counters + 1;
    Q.nq = nq;
    Q.lds = nq <= QLDS;
    /* counts a planner adds without a kernel (CTFs wholly inside a range query) */
    std::vector<unsigned long long> planned(nq, 0);

    /* live meetings and memtable */
    if (st.stepped && st.table && st.nb) {
        uint32_t g = (uint32_t)std::min<uint64_t>(blocks(st.nb, 256), 8192);
        hipLaunchKernelGGL(k_q_live, dim3(g), dim3(256), 0, s, st.table, st.nb, st.last_t, Q);
        QCK(hipGetLastError());
    }
    if (st.n_rec) {
        hipLaunchKernelGGL(k_q_mem, dim3(std::min<uint32_t>(blocks(st.n_rec, 256), 8192)), dim3(256), 0, s,
                           st.rec_keys, st.rec_boxes, st.n_rec, st.end_time_min, Q);
        QCK(hipGetLastError());
    }
    /* runs */
    const uint32_t nruns = (uint32_t)st.runs.size();
    if (nruns) {
        if (grow32(&B->runs, &B->runs_cap, nruns))
            return stc_hip_fail(hipErrorOutOfMemory, "query runs", __FILE__, __LINE__);
        QCK(hipMemcpy(B->runs, st.runs.data(), nruns * sizeof(QueryRunDev), hipMemcpyHostToDevice));
        if (!oq.empty()) {
            hipLaunchKernelGGL(k_q_run_oid, dim3(blocks((uint64_t)oq.size() * nruns, 256)), dim3(256), 0, s, B->runs,
                               nruns, B->oq, (uint32_t)oq.size(), st.num_objects, Q);
            QCK(hipGetLastError());
        }
        if (!rq.empty()) {
            std::vector<uint4> cand;
            for (uint32_t qi : rq) {
                const stc_query &qq = q[qi];
                for (uint32_t r = 0; r < nruns; r++) {
                    const stc_ctf_meta *ctfs = st.run_ctfs[r];
                    const CtfStats *cs = st.run_stats[r];
                    for (uint32_t c = 0; c < st.runs[r].ctf_count; c++) {
                        if (!ctfs[c].kv_capacity) continue;
                        if (cs && cs[c].regular) {
                            const CtfStats &x = cs[c];
                            /* no record can match */
                            if (!(x.box[0] <= qq.box[2] && x.box[2] >= qq.box[0] && x.box[1] <= qq.box[3] &&
                                  x.box[3] >= qq.box[1]) || x.start_min > qq.t1 || x.end_max < qq.t0)
                                continue;
                            /* every forward record matches: count without decoding */
                            if ((!Q.cap || (qq.flags & STC_QF_COUNT)) && x.box[0] >= qq.box[0] &&
                                x.box[2] <= qq.box[2] && x.box[1] >= qq.box[1] &&
                                x.box[3] <= qq.box[3] && x.start_min >= qq.t0 && x.end_max <= qq.t1) {
                                planned[qi] += x.n_fwd;
                                continue;
                            }
                        } else if (!ctf_may_intersect(ctfs[c], qq.box)) {
                            continue;
                        }
                        cand.push_back(make_uint4(qi, r, c, 0));
                    }
                    if (st.runs[r].n_tail) cand.push_back(make_uint4(qi, r, 0xFFFFFFFFu, 0));
                }
            }
            if (!cand.empty()) {
                if (grow(&B->cand, &B->cand_cap, cand.size()))
                    return stc_hip_fail(hipErrorOutOfMemory, "query candidates", __FILE__, __LINE__);
                QCK(hipMemcpy(B->cand, cand.data(), cand.size() * sizeof(uint4), hipMemcpyHostToDevice));
                hipLaunchKernelGGL(k_q_run_range, dim3((uint32_t)cand.size()), dim3(256), 0, s, B->runs, B->cand, Q);
                QCK(hipGetLastError());
            }
        }
    }

    /* results */
    std::vector<unsigned long long> cnt((uint64_t)nq + 1);
    QCK(hipMemcpyAsync(cnt.data(), B->counters, cnt.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    QCK(hipStreamSynchronize(s));
    uint64_t total = 0; /* the hits of the queries that store theirs */
    for (uint32_t i = 0; i < nq; i++) {
        const uint64_t ci = cnt[i + 1] + planned[i];
        if (!(q[i].flags & STC_QF_COUNT)) total += ci;
        if (counts) counts[i] = ci;
    }
    *n_hits = total;
    uint64_t stored = Q.cap ? std::min<uint64_t>(cnt[0], cap) : 0;
    if (stored && hits) {
        QCK(hipMemcpy(hits, B->hits, stored * sizeof(stc_hit), hipMemcpyDeviceToHost));
        std::sort(hits, hits + stored, [](const stc_hit &a, const stc_hit &b) {
            if (a.query != b.query) return a.query < b.query;
            if (a.source != b.source) return a.source < b.source;
            if (a.pid != b.pid) return a.pid < b.pid;
            if (a.target != b.target) return a.target < b.target;
            if (a.start != b.start) return a.start < b.start;
            return a.end < b.end;
        });
    }
    return STC_OK;
}
