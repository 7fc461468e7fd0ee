/*
 * stc_flush.hip — sort-based flush of the memtable into one run (CTB) of
 * ctf_count CTFs (process_with_gpu's flush chain, src/cuda/geometry.cu:1752-2137).
 *
 * Order of the work, and what each step reproduces:
 *   F1 set_oid       per object: oversize test on the quantized object MBR
 *                    and the STR key (geometry.cu:739-767, including the
 *                    high[1]-twice y term of :764)
 *   F2 STR           stable radix sort of (STR key, oid) over all objects, then
 *                    one stable sort by (strip, y) that equals the per-strip
 *                    y sorts of geometry.cu:1858-1890; make_sid (:779-795)
 *   F3 CTF sizes     per-CTF record counts (get_CTF_capacity, :808-818) and
 *                    their prefix sums
 *   F4 placement     the reference sorts 128-bit keys (sid, pid, target,
 *                    duration, end) with thrust (:1907). Here objects are
 *                    ordered by (sid, oid) with a 12-bit stable radix sort,
 *                    each object gets the slot range of its records (prefix
 *                    sum of record counts), records are scattered to their
 *                    object's range, and each record moves to its rank inside
 *                    that range. Keys are unique, so the result is the sorted
 *                    order. Records of oversize objects form the sorted tail
 *                    (the oversize buffer, :1773-1813) the same way.
 *   F5 limits        per CTF: MBR and start-time bounds (cuda_get_limit,
 *                    :820-902); the oversize buffer's start bounds
 *                    (get_o_limit, :926-984). compat bit 0 reduces only the
 *                    records whose in-CTF index mod 64 is below 32, as the
 *                    reference's reduction does.
 *   F6 bits          CTF::get_ctf_bits on the host (step_merge.cpp:140-186)
 *   F7 bitmaps+keys  per CTF workgroup: rasterize record boxes into the CTF
 *                    bitmap staged in LDS (cuda_write_bitmap, :986-1021) and
 *                    pack each key to key_bit/8 little-endian bytes
 *                    (write_key_and_mbr, :1073-1107, serialize_mbr :161-171).
 *                    The oversize bitmap (:904-924) is drawn first into CTF 0's
 *                    slot, copied out and cleared, as the reference does.
 */
#include "stc_flush.h"
#include "stc_table.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#define SID_SHIFT_HI 52   /* sid << 116 is bit 52 of the high word */
#define OID_MASK ((1ull << 26) - 1)
#define BIT_COUNT 65536u
#define BITMAP_WORDS (BIT_COUNT / 32)
#define MAX_RASTER (1u << 20)

static inline uint32_t blocks_for(uint64_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

/* ------------------------------------------------------------------ */
/* host arithmetic of get_ctf_bits                                     */
/* ------------------------------------------------------------------ */

static inline uint32_t x86_d2u32(double v) {
    /* host double -> uint32 as x86-64 compilers emit it (cvttsd2si into a
     * 64-bit register, low 32 bits kept) */
    if (!(v > -9223372036854775808.0 && v < 9223372036854775808.0)) return 0;
    return (uint32_t)(uint64_t)(int64_t)v;
}
static inline int32_t x86_d2i32(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return (int32_t)0x80000000u;
    return (int32_t)v;
}
static inline uint32_t min_bits(uint32_t v) { /* min_bits_to_store (util.h:527-533) */
    if (v == 0) return 1;
    return (uint32_t)std::log2((double)v) + 1;
}

void stc_ctf_bits(stc_ctf_meta *c, const double *map, uint32_t num_objects, uint32_t max_meet_time) {
    c->flags &= ~1u;
    c->id_bit = (uint16_t)min_bits(num_objects);
    c->duration_bit = (uint16_t)min_bits(max_meet_time - 1);
    c->end_bit = c->duration_bit;
    double mxl = (map[2] - map[0]) / 3000;
    double myl = (map[3] - map[1]) / 4200;
    float wx = c->mbr[2] - c->mbr[0];
    float wy = c->mbr[3] - c->mbr[1];
    uint32_t lxg = x86_d2u32((double)wx / mxl);
    uint32_t lyg = x86_d2u32((double)wy / myl);
    c->low_x_bit = (uint16_t)min_bits((lxg > 1 ? lxg : 1) - 1);
    c->low_y_bit = (uint16_t)min_bits((lyg > 1 ? lyg : 1) - 1);
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
    if (!(c->bitmap_size < 8192)) c->flags |= 1;
}

/* ------------------------------------------------------------------ */
/* device helpers                                                      */
/* ------------------------------------------------------------------ */

struct FlushCounters {
    uint32_t oversize_oid_count; /* from identify, plus set_oid's area test */
    uint32_t sid_count;
    uint32_t z;                  /* objects without a CTF (no records or oversize) */
    uint32_t xpc, pkc, split;
    uint32_t n_tail;             /* records of oversize objects */
    uint32_t o_start_min, o_start_max;
    uint32_t bad;                /* inconsistent capacities (reference would misbehave) */
    uint32_t raster_capped;
    uint32_t str_wide;           /* an STR key fell outside the compact key's range (the flush is redone at full width) */
};

/* The STR keys in fewer bits: (mid x - base0 + 1) << by | (mid y - base1) for
 * an object with a CTF, 0 otherwise, where base0/base1 lie below every mid
 * coordinate a point inside the map can give and by covers the y range. The
 * map is monotone and one-to-one on that domain, so the stable sorts order the
 * objects exactly as over the full 64-bit keys, in fewer 8-bit passes. on = 0
 * keeps the full keys. */
struct StrCompact {
    uint32_t on, by;
    uint32_t base0, base1;
    uint64_t lim0, lim1;         /* x - base0 + 1 < lim0 and y - base1 < lim1 = 2^by */
    uint32_t str_bits, strip_bits; /* key bits of the two sorts (host side) */
};

/* float/double -> unsigned, saturating as the reference GPU's conversions do */
__device__ inline uint32_t sat_u32(double v) {
    if (!(v > 0)) return 0;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}
__device__ inline uint64_t sat_u64(double v) {
    if (!(v > 0)) return 0;
    if (v >= 18446744073709551615.0) return ~0ull;
    return (uint64_t)v;
}
__device__ inline float q2f(uint32_t f) { return (float)f / 100000 - 180; } /* uint_to_float */

/* ------------------------------------------------------------------ */
/* F1: set_oid                                                         */
/* ------------------------------------------------------------------ */

__global__ void k_set_oid(uint32_t n, const uint32_t *__restrict__ same_pid_count, const uint4 *__restrict__ o_boxs,
                          uint16_t *__restrict__ sids, uint64_t *__restrict__ mid, uint32_t *__restrict__ oids,
                          uint8_t *__restrict__ part_tail, double area_threshold, uint32_t sticky, StrCompact sc,
                          FlushCounters *fc) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    oids[id] = id;
    uint64_t m = 0;
    if (same_pid_count[id] >= 1) {
        uint4 ob = o_boxs[id];
        float low0 = q2f(ob.x), high0 = q2f(ob.z), low1 = q2f(ob.y), high1 = q2f(ob.w);
        float area = (high1 - low1) * (high0 - low0);
        if ((sticky && sids[id] == 1) || (double)area > area_threshold) {
            atomicAdd(&fc->oversize_oid_count, 1u);
            sids[id] = 1;
        } else {
            uint32_t am0 = ob.x / 2 + ob.z / 2;
            uint32_t am1 = ob.w / 2 + ob.w / 2; /* high[1] twice, as the reference */
            if (!sc.on) {
                m = ((uint64_t)am0 << 32) + (uint64_t)am1;
            } else {
                const uint64_t x = (uint64_t)am0 - sc.base0 + 1, y = (uint64_t)am1 - sc.base1;
                if (am0 < sc.base0 || am1 < sc.base1 || x >= sc.lim0 || y >= sc.lim1) atomicOr(&fc->str_wide, 1u);
                m = (x << sc.by) | (y & (sc.lim1 - 1));
            }
        }
    }
    mid[id] = m;
    part_tail[id] = (sids[id] == 1) ? 1 : 0; /* partition-time oversize flag (geometry.cu:1777-1780) */
}

/* sid_count: the objects with at least one record (geometry.cu:726-729 counts
 * them as their first record arrives) */
__global__ __launch_bounds__(256) void k_count_sids(const uint32_t *__restrict__ same_pid_count, uint32_t n,
                                                    FlushCounters *fc) {
    uint32_t mine = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        mine += same_pid_count[i] ? 1u : 0u;
    for (int o = STC_WAVE / 2; o > 0; o >>= 1) mine += __shfl_xor(mine, o, STC_WAVE);
    if ((threadIdx.x & (STC_WAVE - 1)) == 0 && mine) atomicAdd(&fc->sid_count, mine);
}

__global__ void k_str_prepare(FlushCounters *fc, uint32_t n, uint32_t split) {
    if (threadIdx.x || blockIdx.x) return;
    uint32_t z = (n - fc->sid_count) + fc->oversize_oid_count;
    uint32_t xpc = (fc->sid_count - fc->oversize_oid_count) / split + 1;
    fc->z = z;
    fc->xpc = xpc;
    fc->pkc = xpc / split + 1;
    fc->split = split;
}

/* narrow_xy_to_y + strip ranges: key = (strip + 1) << 32 | y for positions >= z
 * (with compact keys, (strip + 1) << by | y - base1) */
__global__ void k_strip_keys(uint64_t *mid, uint32_t n, StrCompact sc, const FlushCounters *fc) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    uint32_t z = fc->z;
    if (id < z) { mid[id] = 0; return; }
    uint32_t strip = (id - z) / fc->xpc;
    if (strip > fc->split - 1) strip = fc->split - 1;
    if (sc.on) mid[id] = ((uint64_t)(strip + 1) << sc.by) | (mid[id] & (sc.lim1 - 1));
    else mid[id] = ((uint64_t)(strip + 1) << 32) | (mid[id] & 0xFFFFFFFFull);
}

__global__ void k_make_sid(const uint32_t *oids, uint32_t n, uint16_t *sids, const FlushCounters *fc) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t z = fc->z;
    if (id >= n || id < z) return;
    uint32_t xi = (id - z) / fc->xpc;
    uint32_t yi = (id - z - xi * fc->xpc) / fc->pkc;
    sids[oids[id]] = (uint16_t)(xi * fc->split + yi + 2);
}

__global__ void k_ctf_capacity(const uint32_t *oids, uint32_t n, const uint16_t *sids,
                               const uint32_t *same_pid_count, uint32_t *capacity, uint32_t ctf_count,
                               FlushCounters *fc) {
    /* objects id >= z are in CTF order, so a wavefront spans one or two CTFs:
     * lanes of one CTF add their counts first and one lane updates it */
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (STC_WAVE - 1);
    bool on = id < n && id >= fc->z;
    uint32_t cid = 0, v = 0;
    if (on) {
        uint32_t oid = oids[id];
        cid = (uint32_t)sids[oid] - 2;
        v = same_pid_count[oid];
        if (cid >= ctf_count) {
            atomicAdd(&fc->bad, 1u);
            on = false;
        }
    }
    uint64_t act = __ballot(on);
    while (act) {
        int leader = __ffsll((unsigned long long)act) - 1;
        uint32_t lc = __shfl(cid, leader, STC_WAVE);
        bool mine = on && cid == lc;
        uint32_t sum = mine ? v : 0u;
        for (int o = STC_WAVE / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, STC_WAVE);
        if (lane == leader) atomicAdd(&capacity[lc], sum);
        act &= ~__ballot(mine);
        if (mine) on = false;
    }
}

/* object order for placement: objects by (sid, oid) — the CTF records, sid
 * being the key's top field — then the oversize objects by oid (key 4095, the
 * tail). Objects without records have no slots. */
__global__ void k_obj_keys(uint32_t n, const uint16_t *sids, const uint8_t *part_tail, const uint32_t *same_pid_count,
                           uint32_t *obj_key, uint32_t *obj_id, uint32_t *obj_cnt) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    uint32_t cnt = same_pid_count[id];
    uint32_t key = 0;
    if (cnt) key = part_tail[id] ? 4095u : (uint32_t)sids[id];
    obj_key[id] = key;
    obj_id[id] = id;
    obj_cnt[id] = cnt;
}

__global__ void k_gather_counts(const uint32_t *obj_id, const uint32_t *same_pid_count, uint32_t n, uint32_t *cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cnt[i] = same_pid_count[obj_id[i]];
}

__global__ void k_scatter_base(const uint32_t *obj_id, const uint32_t *scanned, uint32_t n, uint32_t *obj_base,
                               uint32_t *obj_fill) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t oid = obj_id[i];
    obj_base[oid] = scanned[i];
    obj_fill[oid] = 0;
}

__global__ void k_count_tail(const uint8_t *part_tail, const uint32_t *same_pid_count, uint32_t n, FlushCounters *fc) {
    uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t v = (id < n && part_tail[id]) ? same_pid_count[id] : 0;
    if (v) atomicAdd(&fc->n_tail, v);
}

/* records to their object's slot range; front keys get their CTF id (write_sid_in_key) */
__global__ void k_scatter_records(const uint64_t *__restrict__ keys, const float4 *__restrict__ boxes, uint64_t K,
                                  const uint32_t *__restrict__ obj_base, uint32_t *__restrict__ obj_fill,
                                  const uint16_t *__restrict__ sids, const uint8_t *__restrict__ part_tail,
                                  uint64_t *__restrict__ tkeys, float4 *__restrict__ tboxes) {
    uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    uint64_t lo = keys[2 * k], hi = keys[2 * k + 1];
    uint32_t oid = (uint32_t)((hi >> 26) & OID_MASK);
    uint32_t slot = obj_base[oid] + atomicAdd(&obj_fill[oid], 1u);
    if (!part_tail[oid]) hi = (hi & ((1ull << SID_SHIFT_HI) - 1)) + ((uint64_t)sids[oid] << SID_SHIFT_HI);
    tkeys[2 * (uint64_t)slot] = lo;
    tkeys[2 * (uint64_t)slot + 1] = hi;
    tboxes[slot] = boxes[k];
}

/* each record moves to its rank inside its object's range */
__global__ void k_rank_records(const uint64_t *__restrict__ tkeys, const float4 *__restrict__ tboxes, uint64_t K,
                               const uint32_t *__restrict__ obj_base, const uint32_t *__restrict__ same_pid_count,
                               uint64_t *__restrict__ okeys, float4 *__restrict__ oboxes) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= K) return;
    uint64_t lo = tkeys[2 * j], hi = tkeys[2 * j + 1];
    uint32_t oid = (uint32_t)((hi >> 26) & OID_MASK);
    uint32_t b = obj_base[oid], cnt = same_pid_count[oid];
    uint32_t rank = 0;
    for (uint32_t i = b; i < b + cnt; i++) {
        uint64_t l2 = tkeys[2 * (uint64_t)i], h2 = tkeys[2 * (uint64_t)i + 1];
        rank += (h2 < hi || (h2 == hi && l2 < lo)) ? 1u : 0u;
    }
    uint64_t dst = (uint64_t)b + rank;
    okeys[2 * dst] = lo;
    okeys[2 * dst + 1] = hi;
    oboxes[dst] = tboxes[j];
}

/* ------------------------------------------------------------------ */
/* F5: limits                                                          */
/* ------------------------------------------------------------------ */

__device__ inline float block_min_f(float v, float *sh) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_down(v, o, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) v = fminf(v, sh[i]);
        sh[0] = v;
    }
    __syncthreads();
    float r = sh[0];
    __syncthreads();
    return r;
}
__device__ inline float block_max_f(float v, float *sh) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_down(v, o, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) v = fmaxf(v, sh[i]);
        sh[0] = v;
    }
    __syncthreads();
    float r = sh[0];
    __syncthreads();
    return r;
}
__device__ inline uint32_t block_min_u(uint32_t v, uint32_t *sh) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_down((int)v, o, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) v = min(v, sh[i]);
        sh[0] = v;
    }
    __syncthreads();
    uint32_t r = sh[0];
    __syncthreads();
    return r;
}
__device__ inline uint32_t block_max_u(uint32_t v, uint32_t *sh) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_down((int)v, o, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); i++) v = max(v, sh[i]);
        sh[0] = v;
    }
    __syncthreads();
    uint32_t r = sh[0];
    __syncthreads();
    return r;
}

/* one workgroup per CTF (blockIdx.x < ctf_count) and one for the oversize
 * buffer (blockIdx.x == ctf_count) */
__global__ __launch_bounds__(256) void k_limits(const uint64_t *__restrict__ skeys, const float4 *__restrict__ sboxes,
                                                const uint32_t *__restrict__ capacity,
                                                const uint32_t *__restrict__ ctf_base, uint32_t ctf_count,
                                                stc_ctf_meta *ctfs, FlushCounters *fc, uint64_t K,
                                                uint32_t end_time_min, uint32_t cur_time, int compat) {
    __shared__ float shf[8];
    __shared__ uint32_t shu[8];
    uint32_t c = blockIdx.x;
    uint32_t n_tail = fc->n_tail;
    uint64_t first, count;
    if (c < ctf_count) {
        first = ctf_base[c];
        count = capacity[c];
    } else {
        first = K - n_tail;
        count = n_tail;
    }
    float l0 = 100000.0f, l1 = 100000.0f, h0 = -100000.0f, h1 = -100000.0f;
    uint32_t smin = 1u << 30, smax = 0;
    uint64_t total = 0;
    const uint64_t nf = K - n_tail;
    for (uint64_t k = threadIdx.x; k < count; k += blockDim.x) {
        if (c < ctf_count && (uint64_t)ctf_base[c] + k >= nf) break;
        if (compat && (k % 64) >= 32) continue;
        uint64_t kid = first + k;
        total = kid;
        float4 b = sboxes[kid];
        l0 = fminf(l0, b.x); l1 = fminf(l1, b.y);
        h0 = fmaxf(h0, b.z); h1 = fmaxf(h1, b.w);
        uint64_t lo = skeys[2 * kid];
        uint32_t end = (uint32_t)lo + end_time_min;
        uint32_t st = end - (uint32_t)(lo >> 32);
        smin = min(smin, st);
        smax = max(smax, st);
    }
    (void)total;
    l0 = block_min_f(l0, shf); l1 = block_min_f(l1, shf);
    h0 = block_max_f(h0, shf); h1 = block_max_f(h1, shf);
    smin = block_min_u(smin, shu); smax = block_max_u(smax, shu);
    if (threadIdx.x == 0) {
        if (c < ctf_count) {
            stc_ctf_meta m;
            memset(&m, 0, sizeof(m));
            m.mbr[0] = l0; m.mbr[1] = l1; m.mbr[2] = h0; m.mbr[3] = h1;
            m.kv_capacity = capacity[c];
            m.start_time_min = smin;
            m.start_time_max = smax;
            m.end_time_min = end_time_min;
            m.end_time_max = cur_time;
            ctfs[c] = m;
        } else {
            fc->o_start_min = smin;
            fc->o_start_max = smax;
        }
    }
}

/* oversize bitmap into CTF 0's slot (geometry.cu:904-924) */
__global__ void k_o_bitmap(const float4 *__restrict__ sboxes, uint64_t tail_first, const FlushCounters *fc,
                           uint32_t *__restrict__ words, uint64_t total_bits, double m0, double m1, double m2,
                           double m3) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= fc->n_tail) return;
    float4 b = sboxes[tail_first + k];
    uint32_t low0 = sat_u32(((double)b.x - m0) / (m2 - m0) * 256);
    uint32_t low1 = sat_u32(((double)b.y - m1) / (m3 - m1) * 256);
    uint32_t high0 = sat_u32(((double)b.z - m0) / (m2 - m0) * 256);
    uint32_t high1 = sat_u32(((double)b.w - m1) / (m3 - m1) * 256);
    for (uint64_t i = low0; i <= high0 && i < total_bits; i++)
        for (uint64_t j = low1; j <= high1 && i + j * 256 < total_bits; j++) {
            uint32_t bit = (uint32_t)(i + j * 256);
            atomicOr(&words[bit >> 5], 1u << (bit & 31));
        }
}

/* F7: bitmap and key packing, one workgroup per CTF */
#define BK_CHUNK 2048u /* records per block of k_bitmap_keys */

/* cuda_write_bitmap + write_key_and_mbr (geometry.cu:986-1021, 1073-1107) for
 * one chunk of one CTF's records: the chunk's bitmap bits are gathered in an
 * LDS copy of the CTF bitmap and OR-ed into it once, and its packed keys are
 * assembled in LDS and written as one contiguous byte range. Chunk = (ctf,
 * first record in the CTF, record count). */
__global__ __launch_bounds__(256) void k_bitmap_keys(const uint64_t *__restrict__ skeys,
                                                     const float4 *__restrict__ sboxes,
                                                     const uint4 *__restrict__ chunks,
                                                     const uint32_t *__restrict__ ctf_base,
                                                     const uint64_t *__restrict__ ctf_byte_base,
                                                     const stc_ctf_meta *__restrict__ ctfs, uint32_t ctf_count,
                                                     uint64_t nf, uint32_t *__restrict__ words,
                                                     uint8_t *__restrict__ key_bytes, FlushCounters *fcw) {
    __shared__ uint32_t bm[BITMAP_WORDS];
    __shared__ uint8_t kst[BK_CHUNK * 16];
    const uint4 ch = chunks[blockIdx.x];
    const uint32_t c = ch.x, k0 = ch.y, kn = ch.z;
    for (uint32_t i = threadIdx.x; i < BITMAP_WORDS; i += blockDim.x) bm[i] = 0;
    __syncthreads();
    stc_ctf_meta ctf = ctfs[c];
    const float m0 = ctf.mbr[0], m1 = ctf.mbr[1], m2 = ctf.mbr[2], m3 = ctf.mbr[3];
    const uint32_t xg = ctf.x_grid, yg = ctf.y_grid;
    const uint64_t total_bits = (uint64_t)ctf_count * BIT_COUNT;
    const uint64_t base_bit = (uint64_t)c * BIT_COUNT;
    const uint64_t room = total_bits - base_bit;
    const uint32_t kb = ctf.key_bit / 8;
    const uint32_t lx = ctf.low_x_bit, ly = ctf.low_y_bit, e = ctf.edge_bit;
    const uint32_t s_mbr = ctf.mbr_bit, s_end = ctf.end_bit, s_dur = ctf.duration_bit, s_id = ctf.id_bit;
    const float longer = fmaxf(m2 - m0, m3 - m1);
    const uint64_t first = ctf_base[c];
    /* records at or past nf are not in the CTF region (only when the capacities disagree with nf) */
    const uint32_t kend = (first + k0 >= nf) ? k0 : (uint32_t)min<uint64_t>(k0 + kn, nf - first);
    for (uint32_t k = k0 + threadIdx.x; k < kend; k += blockDim.x) {
        uint64_t kid = first + k;
        float4 b = sboxes[kid];
        /* cuda_write_bitmap */
        uint32_t low0 = sat_u32((b.x - m0) / (m2 - m0) * (float)xg);
        uint32_t low1 = sat_u32((b.y - m1) / (m3 - m1) * (float)yg);
        uint32_t high0 = sat_u32((b.z - m0) / (m2 - m0) * (float)xg);
        uint32_t high1 = sat_u32((b.w - m1) / (m3 - m1) * (float)yg);
        if (xg) {
            uint32_t done = 0;
            for (uint64_t j = low1; j <= high1 && j * xg < room && done < MAX_RASTER; j++)
                for (uint64_t i = low0; i <= high0 && i + j * xg < room && done < MAX_RASTER; i++, done++) {
                    uint32_t bit = (uint32_t)(i + j * xg);
                    if (bit < BIT_COUNT) atomicOr(&bm[bit >> 5], 1u << (bit & 31));
                    else {
                        uint64_t g = base_bit + bit;
                        atomicOr(&words[g >> 5], 1u << (g & 31));
                    }
                }
            if (done >= MAX_RASTER) atomicAdd(&fcw->raster_capped, 1u);
        }
        /* serialize_mbr (f_box form) */
        uint64_t v0 = sat_u64((b.x - m0) / (m2 - m0) * (float)((1ull << lx) - 1));
        uint64_t v1 = sat_u64((b.y - m1) / (m3 - m1) * (float)((1ull << ly) - 1));
        uint64_t vx = sat_u64((b.z - b.x) / longer * (float)((1ull << e) - 1));
        uint64_t vy = sat_u64((b.w - b.y) / longer * (float)((1ull << e) - 1));
        uint64_t vm = (ly + 2 * e < 64 ? (v0 << (ly + e + e)) : 0) + (2 * e < 64 ? (v1 << (e + e)) : 0) +
                      (e < 64 ? (vx << e) : 0) + vy;
        /* extract_fields keeps the sid bits in pid (cuda_util.cuh:163-169) */
        uint64_t lo = skeys[2 * kid], hi = skeys[2 * kid + 1];
        unsigned __int128 end = lo & 0xFFFFFFFFull;
        unsigned __int128 dur = lo >> 32;
        unsigned __int128 target = hi & OID_MASK;
        unsigned __int128 pid = hi >> 26;
        uint32_t sh1 = s_id + s_dur + s_end + s_mbr, sh2 = s_dur + s_end + s_mbr, sh3 = s_end + s_mbr;
        unsigned __int128 tk = (sh1 < 128 ? pid << sh1 : 0) + (sh2 < 128 ? target << sh2 : 0) +
                               (sh3 < 128 ? dur << sh3 : 0) + (s_mbr < 128 ? end << s_mbr : 0) +
                               (unsigned __int128)vm;
        uint8_t *dst = kst + (k - k0) * kb;
        for (uint32_t q = 0; q < kb && q < 16; q++) dst[q] = (uint8_t)(tk >> (8 * q));
    }
    __syncthreads();
    /* packed keys: one contiguous range, written byte-coalesced */
    const uint32_t nbytes = (kend - k0) * min(kb, 16u);
    uint8_t *gdst = key_bytes + ctf_byte_base[c] + (uint64_t)k0 * kb;
    if (kb <= 16)
        for (uint32_t i = threadIdx.x; i < nbytes; i += blockDim.x) gdst[i] = kst[i];
    uint32_t *gw = words + (uint64_t)c * BITMAP_WORDS;
    for (uint32_t i = threadIdx.x; i < BITMAP_WORDS; i += blockDim.x)
        if (bm[i]) atomicOr(&gw[i], bm[i]);
}

__global__ void k_fill_obox2(uint4 *a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = make_uint4(100000000u, 100000000u, 0u, 0u);
}

/* ------------------------------------------------------------------ */
/* host                                                                */
/* ------------------------------------------------------------------ */

#define FCK(expr)                                                                  \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) return stc_hip_fail(e_, #expr, __FILE__, __LINE__); \
    } while (0)
#define FRC(expr)                                                                  \
    do {                                                                           \
        int rc_ = (expr);                                                          \
        if (rc_) return rc_;                                                       \
    } while (0)

int stc_flush_init(FlushWork *w, uint32_t n, uint32_t ctf_count) {
    w->n = n;
    w->ctf_count = ctf_count;
    size_t N = n;
    if (hipMalloc((void **)&w->mid, N * 8) || hipMalloc((void **)&w->mid_alt, N * 8) ||
        hipMalloc((void **)&w->oids, N * 4) || hipMalloc((void **)&w->oids_alt, N * 4) ||
        hipMalloc((void **)&w->obj_base, N * 4) || hipMalloc((void **)&w->obj_fill, N * 4) ||
        hipMalloc((void **)&w->capacity, (ctf_count + 1) * 4) || hipMalloc((void **)&w->ctf_base, (ctf_count + 1) * 4) ||
        hipMalloc((void **)&w->ctf_byte_base, (ctf_count + 1) * 8) ||
        hipMalloc((void **)&w->ctfs, ctf_count * sizeof(stc_ctf_meta)) ||
        hipMalloc((void **)&w->bitmaps, (size_t)ctf_count * 8192) || hipMalloc((void **)&w->flags, 256) ||
        hipMalloc((void **)&w->part_tail, N) || hipMalloc((void **)&w->obj_key, N * 4) ||
        hipMalloc((void **)&w->obj_id, N * 4) || hipMalloc((void **)&w->obj_cnt, N * 4) ||
        hipMalloc((void **)&w->sids_out, N * 2) || hipMalloc((void **)&w->o_bitmap_dev, 8192))
        return STC_E_HIP;
    if (hipStreamCreateWithFlags(&w->cs, hipStreamNonBlocking) ||
        hipEventCreateWithFlags(&w->ready, hipEventDisableTiming) ||
        hipEventCreateWithFlags(&w->copied, hipEventDisableTiming))
        return STC_E_HIP;
    w->scan_temp_words = stc_radix_temp_words(n) + 64;
    if (hipMalloc((void **)&w->scan_temp, w->scan_temp_words * 4)) return STC_E_HIP;
    return 0;
}

void stc_flush_free(FlushWork *w) {
    w->job_set = false;
    if (w->cs) (void)hipStreamSynchronize(w->cs);
    if (w->cs) (void)hipStreamDestroy(w->cs);
    if (w->ready) (void)hipEventDestroy(w->ready);
    if (w->copied) (void)hipEventDestroy(w->copied);
    void *ptrs[] = {w->chunks, w->sids_out, w->o_bitmap_dev, w->mid, w->mid_alt, w->oids, w->oids_alt, w->obj_base, w->obj_fill, w->capacity, w->ctf_base,
                    w->ctf_byte_base, w->ctfs, w->bitmaps, w->flags, w->part_tail, w->obj_key, w->obj_id, w->obj_cnt,
                    w->scan_temp, w->keys_out, w->boxes_out, w->tkeys, w->tboxes, w->key_bytes};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    memset((void *)w, 0, sizeof(*w));
}

int stc_flush_release(FlushWork *w, hipStream_t s) {
    if (!w->job_set) return STC_OK;
    w->job_set = false;
    const FlushWork::Job &j = w->job;
    const FlushCounters *fc = (const FlushCounters *)w->flags;
    FCK(hipEventRecord(w->ready, s));
    FCK(hipStreamWaitEvent(w->cs, w->ready, 0));
    if (j.key_bytes) FCK(hipMemcpyAsync(j.dst + j.key_bytes_at, w->key_bytes, j.key_bytes, hipMemcpyDeviceToHost, w->cs));
    FCK(hipMemcpyAsync(j.dst + j.bitmaps_at, w->bitmaps, (size_t)j.ctf_count * 8192, hipMemcpyDeviceToHost, w->cs));
    FCK(hipMemcpyAsync(j.dst + j.o_bitmap_at, w->o_bitmap_dev, 8192, hipMemcpyDeviceToHost, w->cs));
    FCK(hipMemcpyAsync(j.dst + j.sids_at, w->sids_out, (size_t)j.n * 2, hipMemcpyDeviceToHost, w->cs));
    if (j.n_tail) {
        FCK(hipMemcpyAsync(j.dst + j.o_keys_at, w->keys_out + 2 * j.nf, j.n_tail * 16, hipMemcpyDeviceToHost, w->cs));
        FCK(hipMemcpyAsync(j.dst + j.o_boxes_at, w->boxes_out + j.nf, j.n_tail * 16, hipMemcpyDeviceToHost, w->cs));
    }
    if (j.has_sorted) {
        FCK(hipMemcpyAsync(j.dst + j.sorted_keys_at, w->keys_out, j.nf * 16, hipMemcpyDeviceToHost, w->cs));
        FCK(hipMemcpyAsync(j.dst + j.sorted_boxes_at, w->boxes_out, j.nf * 16, hipMemcpyDeviceToHost, w->cs));
    }
    FCK(hipMemcpyAsync(j.dst + j.counters_at, fc, sizeof(FlushCounters), hipMemcpyDeviceToHost, w->cs));
    FCK(hipEventRecord(w->copied, w->cs));
    FCK(hipEventRecord(j.done, w->cs));
    w->pending = true;
    return STC_OK;
}

void stc_flush_cancel(FlushWork *w, const uint8_t *dst) {
    if (w->job_set && w->job.dst == dst) w->job_set = false;
}

static int ensure_rec(FlushWork *w, uint64_t K) {
    if (K <= w->rec_cap) return 0;
    if (w->pending && hipEventSynchronize(w->copied)) return STC_E_HIP; /* a copy still reads them */
    uint64_t nc = std::max<uint64_t>(K + K / 4, 1 << 16);
    void *ptrs[] = {w->keys_out, w->boxes_out, w->tkeys, w->tboxes, w->key_bytes};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    w->keys_out = nullptr; w->boxes_out = nullptr; w->tkeys = nullptr; w->tboxes = nullptr; w->key_bytes = nullptr;
    if (hipMalloc((void **)&w->keys_out, nc * 16) || hipMalloc((void **)&w->boxes_out, nc * 16) ||
        hipMalloc((void **)&w->tkeys, nc * 16) || hipMalloc((void **)&w->tboxes, nc * 16) ||
        hipMalloc((void **)&w->key_bytes, nc * 16))
        return STC_E_HIP;
    w->rec_cap = nc;
    return 0;
}

int stc_flush_run(FlushWork *w, const FlushArgs *a, FlushOut *out, hipStream_t s) {
    const uint32_t n = a->n, C = a->ctf_count;
    const uint32_t split = (uint32_t)std::sqrt((double)C);
    const uint64_t K = a->kv_count;
    FRC(stc_flush_release(w, s));
    FRC(ensure_rec(w, K + 1));
    /* the previous run's copies read keys_out, key_bytes, bitmaps, sids_out and the counters */
    if (w->pending) FCK(hipStreamWaitEvent(s, w->copied, 0));
    FlushCounters *fc = (FlushCounters *)w->flags;
    FlushCounters init;
    memset(&init, 0, sizeof(init));
    init.oversize_oid_count = a->oversize_oid_count;
    init.sid_count = 0; /* counted below */
    FCK(hipMemcpyAsync(fc, &init, sizeof(init), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_count_sids, dim3(std::max<uint32_t>(std::min<uint32_t>(blocks_for(n, 256), 1024), 1)),
                       dim3(256), 0, s, a->same_pid_count, n, fc);
    FCK(hipMemsetAsync(w->capacity, 0, (C + 1) * 4, s));
    FCK(hipMemsetAsync(w->bitmaps, 0, (size_t)C * 8192, s));
    const uint32_t B = 256;

    /* compact STR keys from the map box: every record box lies inside it, so every object's quantised
     * mid point does (f2q is monotone), and x and y each need only the bits of the map's extent */
    StrCompact sc{};
    {
        const uint32_t qlo0 = f2q((float)a->map[0]), qlo1 = f2q((float)a->map[1]);
        const uint32_t qhi0 = f2q((float)a->map[2]), qhi1 = f2q((float)a->map[3]);
        sc.base0 = qlo0 ? qlo0 - 1 : 0; /* a mid point (x/2 + z/2) can fall one below a bound */
        sc.base1 = qlo1 ? qlo1 - 1 : 0;
        auto bits_for = [](uint64_t v) { uint32_t b = 0; while (b < 63 && (1ull << b) <= v) b++; return b; };
        const uint32_t bx = qhi0 >= sc.base0 ? bits_for((uint64_t)qhi0 - sc.base0 + 1) : 64;
        const uint32_t by = qhi1 >= sc.base1 ? bits_for((uint64_t)qhi1 - sc.base1) : 64;
        sc.on = !w->str_full && bx + by <= 56 && by + bits_for(split) <= 56;
        sc.by = by;
        sc.lim0 = sc.on ? 1ull << bx : 0;
        sc.lim1 = sc.on ? 1ull << by : 0;
        if (sc.on) {
            sc.str_bits = bx + by;
            sc.strip_bits = by + bits_for(split);
        }
    }
    /* F1 */
    hipLaunchKernelGGL(k_set_oid, dim3(blocks_for(n, B)), dim3(B), 0, s, n, a->same_pid_count, a->o_boxs, a->sids,
                       w->mid, w->oids, w->part_tail, a->area_threshold, a->sticky_oversize, sc, fc);
    hipLaunchKernelGGL(k_str_prepare, dim3(1), dim3(1), 0, s, fc, n, split);
    FCK(hipGetLastError());
    /* F2: STR over the full 64-bit key (x high, y low), stable from oid order */
    FRC(stc_radix_sort_u64(w->mid, w->oids, w->mid_alt, w->oids_alt, n, 0, sc.on ? (int)sc.str_bits : 64, w->scan_temp, s));
    hipLaunchKernelGGL(k_strip_keys, dim3(blocks_for(n, B)), dim3(B), 0, s, w->mid, n, sc, fc);
    FRC(stc_radix_sort_u64(w->mid, w->oids, w->mid_alt, w->oids_alt, n, 0, sc.on ? (int)sc.strip_bits : 40, w->scan_temp, s));
    hipLaunchKernelGGL(k_make_sid, dim3(blocks_for(n, B)), dim3(B), 0, s, w->oids, n, a->sids, fc);
    /* F3 */
    hipLaunchKernelGGL(k_ctf_capacity, dim3(blocks_for(n, B)), dim3(B), 0, s, w->oids, n, a->sids, a->same_pid_count,
                       w->capacity, C, fc);
    FCK(hipGetLastError());
    FRC(stc_exclusive_scan_u32(w->capacity, w->ctf_base, C, w->ctf_base + C, w->scan_temp, s));
    /* F4: object order (key, oid), slot ranges, scatter, rank */
    hipLaunchKernelGGL(k_obj_keys, dim3(blocks_for(n, B)), dim3(B), 0, s, n, a->sids, w->part_tail,
                       a->same_pid_count, w->obj_key, w->obj_id, w->obj_cnt);
    FRC(stc_radix_sort_u32(w->obj_key, w->obj_id, (uint32_t *)w->mid_alt, w->oids_alt, n, 12, w->scan_temp, s));
    hipLaunchKernelGGL(k_gather_counts, dim3(blocks_for(n, B)), dim3(B), 0, s, w->obj_id, a->same_pid_count, n,
                       w->obj_cnt);
    FRC(stc_exclusive_scan_u32(w->obj_cnt, w->obj_cnt, n, nullptr, w->scan_temp, s));
    hipLaunchKernelGGL(k_scatter_base, dim3(blocks_for(n, B)), dim3(B), 0, s, w->obj_id, w->obj_cnt, n, w->obj_base,
                       w->obj_fill);
    hipLaunchKernelGGL(k_count_tail, dim3(blocks_for(n, B)), dim3(B), 0, s, w->part_tail, a->same_pid_count, n, fc);
    if (K) {
        hipLaunchKernelGGL(k_scatter_records, dim3(blocks_for(K, B)), dim3(B), 0, s, a->rec_keys, a->rec_boxes, K,
                           w->obj_base, w->obj_fill, a->sids, w->part_tail, w->tkeys, w->tboxes);
        hipLaunchKernelGGL(k_rank_records, dim3(blocks_for(K, B)), dim3(B), 0, s, w->tkeys, w->tboxes, K,
                           w->obj_base, a->same_pid_count, w->keys_out, w->boxes_out);
    }
    FCK(hipGetLastError());
    /* F5 */
    hipLaunchKernelGGL(k_limits, dim3(C + 1), dim3(256), 0, s, w->keys_out, w->boxes_out, w->capacity, w->ctf_base,
                       C, w->ctfs, fc, K, a->end_time_min, a->cur_time, (int)(a->compat & 1));
    FCK(hipGetLastError());
    /* F6: bit widths on the host */
    FlushCounters hfc;
    std::vector<stc_ctf_meta> ctfs(C);
    std::vector<uint32_t> cap(C + 1), base(C + 1);
    FCK(hipMemcpyAsync(&hfc, fc, sizeof(hfc), hipMemcpyDeviceToHost, s));
    FCK(hipMemcpyAsync(ctfs.data(), w->ctfs, C * sizeof(stc_ctf_meta), hipMemcpyDeviceToHost, s));
    FCK(hipMemcpyAsync(cap.data(), w->capacity, C * 4, hipMemcpyDeviceToHost, s));
    FCK(hipMemcpyAsync(base.data(), w->ctf_base, (C + 1) * 4, hipMemcpyDeviceToHost, s));
    FCK(hipStreamSynchronize(s));
    if (hfc.str_wide && sc.on) { /* cannot happen for records inside the map; redone over the full keys */
        w->str_full = true;
        return stc_flush_run(w, a, out, s);
    }
    std::vector<uint64_t> byte_base(C + 1, 0);
    for (uint32_t c = 0; c < C; c++) {
        stc_ctf_bits(&ctfs[c], a->map, n, a->max_meet_time);
        byte_base[c + 1] = byte_base[c] + (uint64_t)cap[c] * ctfs[c].key_bit / 8;
    }
    const uint64_t n_tail = hfc.n_tail, nf = K - n_tail;
    uint32_t flags = 0;
    if (hfc.bad) flags |= 2;
    if (!(n_tail < a->kv_restriction / 100)) flags |= 4;
    if (base[C] != nf) flags |= 2;
    for (auto &m : ctfs) m.flags |= flags;
    if (byte_base[C] > w->rec_cap * 16) return stc_hip_fail(hipErrorOutOfMemory, "key bytes", __FILE__, __LINE__);
    FCK(hipMemcpyAsync(w->ctfs, ctfs.data(), C * sizeof(stc_ctf_meta), hipMemcpyHostToDevice, s));
    FCK(hipMemcpyAsync(w->ctf_byte_base, byte_base.data(), (C + 1) * 8, hipMemcpyHostToDevice, s));
    /* F7: oversize bitmap first (CTF 0's slot, then copied out and cleared) */
    uint32_t *words = (uint32_t *)w->bitmaps;
    uint64_t total_bits = (uint64_t)C * BIT_COUNT;
    if (n_tail)
        hipLaunchKernelGGL(k_o_bitmap, dim3(blocks_for(n_tail, B)), dim3(B), 0, s, w->boxes_out, nf, fc, words,
                           total_bits, a->map[0], a->map[1], a->map[2], a->map[3]);
    FCK(hipMemcpyAsync(w->o_bitmap_dev, w->bitmaps, 8192, hipMemcpyDeviceToDevice, s));
    FCK(hipMemsetAsync(w->bitmaps, 0, 8192, s));
    /* chunks of at most BK_CHUNK records, CTF by CTF */
    std::vector<uint4> chunks;
    for (uint32_t c = 0; c < C; c++)
        for (uint32_t k = 0; k < cap[c]; k += BK_CHUNK) chunks.push_back(make_uint4(c, k, std::min(BK_CHUNK, cap[c] - k), 0));
    if (!chunks.empty()) {
        if (chunks.size() > w->chunks_cap) {
            if (w->chunks) FCK(hipFree(w->chunks));
            w->chunks = nullptr;
            w->chunks_cap = 0;
            size_t nc = chunks.size() + chunks.size() / 2 + 64;
            FCK(hipMalloc((void **)&w->chunks, nc * sizeof(uint4)));
            w->chunks_cap = nc;
        }
        FCK(hipMemcpyAsync(w->chunks, chunks.data(), chunks.size() * sizeof(uint4), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_bitmap_keys, dim3((uint32_t)chunks.size()), dim3(256), 0, s, w->keys_out, w->boxes_out,
                           w->chunks, w->ctf_base, w->ctf_byte_base, w->ctfs, C, nf, words, w->key_bytes, fc);
        FCK(hipGetLastError());
    }
    FCK(hipMemcpyAsync(w->sids_out, a->sids, (size_t)n * 2, hipMemcpyDeviceToDevice, s));

    /* outputs: one pinned block, filled on the copy stream */
    out->ctfs = ctfs;
    out->has_sorted = a->keep_sorted && nf;
    RunLayout &L = out->lay;
    size_t off = 0;
    auto place = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    L.key_bytes = place(byte_base[C]);
    L.bitmaps = place((size_t)C * 8192);
    L.o_bitmap = place(8192);
    L.sids = place((size_t)n * 2);
    L.o_keys = place(n_tail * 16);
    L.o_boxes = place(n_tail * 16);
    L.sorted_keys = place(out->has_sorted ? nf * 16 : 0);
    L.sorted_boxes = place(out->has_sorted ? nf * 16 : 0);
    L.counters = place(sizeof(FlushCounters));
    L.total = off;
    out->blk = out->acquire(out->acquire_ctx, L.total);
    if (!out->blk.p) return stc_hip_fail(hipErrorOutOfMemory, "pinned run buffer", __FILE__, __LINE__);
    if (a->reset_memtable) { /* geometry.cu:2111-2131 */
        FCK(hipMemsetAsync(a->same_pid_count, 0, 4ull * n, s));
        FCK(hipMemsetAsync(a->sids, 0, 2ull * n, s));
        hipLaunchKernelGGL(k_fill_obox2, dim3(blocks_for(n, 256)), dim3(256), 0, s, a->o_boxs, n);
        FCK(hipGetLastError());
    }
    /* the copies to the pinned block wait for stc_flush_release */
    FlushWork::Job &j = w->job;
    j.dst = out->blk.p;
    j.key_bytes_at = L.key_bytes; j.bitmaps_at = L.bitmaps; j.o_bitmap_at = L.o_bitmap; j.sids_at = L.sids;
    j.o_keys_at = L.o_keys; j.o_boxes_at = L.o_boxes; j.sorted_keys_at = L.sorted_keys;
    j.sorted_boxes_at = L.sorted_boxes; j.counters_at = L.counters;
    j.key_bytes = byte_base[C];
    j.nf = nf;
    j.n_tail = n_tail;
    j.ctf_count = C;
    j.n = n;
    j.has_sorted = out->has_sorted;
    j.done = out->blk.ev;
    w->job_set = true;
    stc_run_info &info = out->info;
    memset(&info, 0, sizeof(info));
    info.kv_count = (uint32_t)nf;
    info.oversize_count = (uint32_t)n_tail;
    info.ctf_count = C;
    info.num_objects = n;
    info.o_start_min = hfc.o_start_min;
    info.o_start_max = hfc.o_start_max;
    info.o_end_min = a->end_time_min;
    info.o_end_max = a->cur_time;
    info.end_time_min = a->end_time_min;
    info.cur_time = a->cur_time;
    info.sid_count = hfc.sid_count;
    info.oversize_oid_count = hfc.oversize_oid_count;
    info.key_bytes = byte_base[C];
    return 0;
}

uint32_t stc_flush_raster_capped(const void *counters) {
    return ((const FlushCounters *)counters)->raster_capped;
}
