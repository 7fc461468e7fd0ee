/*
 * stc_radix.hip — stable LSD radix sort and prefix sums for gfx950.
 *
 * One pass per 8-bit digit, three launches per pass:
 *   histogram : each workgroup counts the digits of its tile of 4096 keys in LDS
 *               and writes them digit-major (digit * tiles + tile);
 *   scan      : exclusive prefix sum over that table (three-phase scan);
 *   scatter   : each workgroup ranks its tile stably, 256 keys per round, with
 *               per-wavefront peer masks built from 8 ballots (64-wide), then
 *               per-digit wave offsets in LDS; 4-byte keys are placed in LDS
 *               in digit order and written out in that order (coalesced),
 *               8-byte keys straight to their global positions.
 * Keys of equal digit keep their input order, so chaining passes from the
 * least significant digit gives a stable sort, as thrust::sort_by_key on
 * primitive keys does in the reference (src/cuda/geometry.cu:1860, 1878-1886).
 */
#include "stc_internal.h"

#define RS_THREADS 256
#define RS_ITEMS 16
#define RS_TILE (RS_THREADS * RS_ITEMS)
#define RS_BINS 256
#define SCAN_THREADS 1024
#define SCAN_ITEMS 4
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

size_t stc_radix_temp_words(uint32_t n) {
    uint32_t tiles = ceil_div(n ? n : 1, RS_TILE);
    size_t hist = (size_t)RS_BINS * tiles;
    size_t sums = ceil_div(hist, SCAN_TILE);
    size_t sums2 = ceil_div(sums, SCAN_TILE);
    /* histogram + scan partials (two levels) + total + slack */
    size_t scan_n = (n > hist ? n : hist);
    size_t s1 = ceil_div(scan_n, SCAN_TILE);
    size_t s2 = ceil_div(s1, SCAN_TILE);
    (void)sums; (void)sums2;
    return hist + s1 + s2 + 64;
}

/* ---------------- exclusive scan ---------------- */

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < STC_WAVE; o <<= 1) {
        uint32_t u = __shfl_up(v, o, STC_WAVE);
        if (lane >= o) v += u;
    }
    return v;
}

/* block-wide exclusive scan of one value per thread; returns the block total in *total */
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[SCAN_THREADS / STC_WAVE];
    int lane = threadIdx.x & (STC_WAVE - 1), w = threadIdx.x / STC_WAVE;
    uint32_t inc = wave_incl_scan(v, lane);
    if (lane == STC_WAVE - 1) wsum[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint32_t x = lane < (int)(blockDim.x / STC_WAVE) ? wsum[lane] : 0;
        uint32_t xi = wave_incl_scan(x, lane);
        if (lane < (int)(blockDim.x / STC_WAVE)) wsum[lane] = xi - x;
        if (lane == (int)(blockDim.x / STC_WAVE) - 1) *total = xi;
    }
    __syncthreads();
    uint32_t r = inc - v + wsum[w];
    __syncthreads();
    return r;
}

__global__ void k_scan_tiles(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *tile_sums) {
    __shared__ uint32_t tot;
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    uint32_t ex = block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0 && tile_sums) tile_sums[blockIdx.x] = tot;
}

__global__ void k_scan_add(uint32_t *out, uint32_t n, const uint32_t *tile_offsets) {
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t add = tile_offsets[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
        if (base + k < n) out[base + k] += add;
}

__global__ void k_scan_total(const uint32_t *in, const uint32_t *out, uint32_t n, uint32_t *total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = n ? out[n - 1] + in[n - 1] : 0;
}

/* in and out may alias for the first level; totals are recovered before
 * overwriting by scanning into `out` and keeping tile sums separately. */
static int scan_rec(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *temp, hipStream_t s) {
    uint32_t tiles = ceil_div(n, SCAN_TILE);
    if (tiles <= 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(SCAN_THREADS), 0, s, in, out, n, (uint32_t *)nullptr);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    uint32_t *sums = temp;
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(SCAN_THREADS), 0, s, in, out, n, sums);
    if (scan_rec(sums, sums, tiles, temp + tiles, s)) return -1;
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(SCAN_THREADS), 0, s, out, n, sums);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int stc_exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total_dev, uint32_t *temp,
                           hipStream_t s) {
    if (n == 0) {
        if (total_dev) hipMemsetAsync(total_dev, 0, 4, s);
        return 0;
    }
    /* keep the last input value so the total survives an in-place scan */
    uint32_t *last = temp;
    STC_CHECK(hipMemcpyAsync(last, in + n - 1, 4, hipMemcpyDeviceToDevice, s));
    if (scan_rec(in, out, n, temp + 16, s)) return stc_hip_fail(hipGetLastError(), "scan", __FILE__, __LINE__);
    if (total_dev) hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, s, last, out + n - 1, 1u, total_dev);
    STC_CHECK(hipGetLastError());
    return 0;
}

/* ---------------- radix passes ---------------- */

template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const K *keys, uint32_t n, int shift, uint32_t *hist,
                                                         uint32_t tiles) {
    __shared__ uint32_t h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; r++) {
        uint64_t i = base + (uint64_t)r * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(uint32_t)(keys[i] >> shift) & (RS_BINS - 1)], 1u);
    }
    __syncthreads();
    hist[(uint64_t)threadIdx.x * tiles + blockIdx.x] = h[threadIdx.x];
}

/* Each workgroup ranks its tile stably (256 keys per round, per-wavefront peer
 * masks from 8 ballots, per-digit wave offsets in LDS), places the keys in
 * LDS in digit order, and writes them out in that order: keys of one digit go
 * to consecutive global positions, so the writes are coalesced. */
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const K *kin, const uint32_t *vin, K *kout,
                                                            uint32_t *vout, uint32_t n, int shift,
                                                            const uint32_t *offs, uint32_t tiles) {
    __shared__ uint32_t run[RS_BINS];    /* keys of each digit ranked so far in this tile */
    __shared__ uint32_t base[RS_BINS];   /* global position of the tile's first key of each digit */
    __shared__ uint32_t lstart[RS_BINS]; /* position in the tile of its first key of each digit */
    __shared__ uint32_t wcnt[RS_THREADS / STC_WAVE][RS_BINS];
    __shared__ K sk[RS_TILE];
    __shared__ uint32_t sv[RS_TILE];
    __shared__ uint32_t tile_total;
    const int tid = threadIdx.x, lane = tid & (STC_WAVE - 1), w = tid / STC_WAVE;
    /* the tile's count of digit tid: the gap to the next entry of the scanned digit-major table */
    const uint64_t slot = (uint64_t)tid * tiles + blockIdx.x;
    const uint32_t b0 = offs[slot];
    const uint32_t b1 = slot + 1 < (uint64_t)RS_BINS * tiles ? offs[slot + 1] : n;
    base[tid] = b0;
    run[tid] = 0;
    lstart[tid] = block_excl_scan(b1 - b0, &tile_total);
#pragma unroll
    for (int q = 0; q < RS_THREADS / STC_WAVE; q++) wcnt[q][tid] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (STC_WAVE - lane));
    const uint64_t tile0 = (uint64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        uint64_t i = tile0 + (uint64_t)r * RS_THREADS + tid;
        bool valid = i < n;
        K k = 0;
        uint32_t v = 0, d = 0;
        if (valid) {
            k = kin[i];
            v = vin[i];
            d = (uint32_t)(k >> shift) & (RS_BINS - 1);
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[tid];
#pragma unroll
            for (int q = 0; q < RS_THREADS / STC_WAVE; q++) {
                uint32_t c = wcnt[q][tid];
                wcnt[q][tid] = acc;
                acc += c;
            }
            run[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            const uint32_t lp = lstart[d] + wcnt[w][d] + rank;
            sk[lp] = k;
            sv[lp] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RS_THREADS / STC_WAVE; q++) wcnt[q][tid] = 0;
        __syncthreads();
    }
    const uint32_t cnt = (uint32_t)min<uint64_t>(RS_TILE, n - tile0);
    for (uint32_t p = tid; p < cnt; p += RS_THREADS) {
        const K k = sk[p];
        const uint32_t d = (uint32_t)(k >> shift) & (RS_BINS - 1);
        const uint32_t pos = base[d] + (p - lstart[d]);
        kout[pos] = k;
        vout[pos] = sv[p];
    }
}

/* The same ranking, each key written straight to its global position: for
 * 8-byte keys, whose LDS tile would cost the kernel its occupancy. */
template <typename K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter_direct(const K *kin, const uint32_t *vin, K *kout,
                                                            uint32_t *vout, uint32_t n, int shift,
                                                            const uint32_t *offs, uint32_t tiles) {
    __shared__ uint32_t run[RS_BINS];
    __shared__ uint32_t base[RS_BINS];
    __shared__ uint32_t wcnt[RS_THREADS / STC_WAVE][RS_BINS];
    const int tid = threadIdx.x, lane = tid & (STC_WAVE - 1), w = tid / STC_WAVE;
    run[tid] = 0;
    base[tid] = offs[(uint64_t)tid * tiles + blockIdx.x];
#pragma unroll
    for (int q = 0; q < RS_THREADS / STC_WAVE; q++) wcnt[q][tid] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (STC_WAVE - lane));
    uint64_t tile0 = (uint64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        uint64_t i = tile0 + (uint64_t)r * RS_THREADS + tid;
        bool valid = i < n;
        K k = 0;
        uint32_t v = 0, d = 0;
        if (valid) {
            k = kin[i];
            v = vin[i];
            d = (uint32_t)(k >> shift) & (RS_BINS - 1);
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (d >> b) & 1u;
            uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            uint32_t acc = run[tid];
#pragma unroll
            for (int q = 0; q < RS_THREADS / STC_WAVE; q++) {
                uint32_t c = wcnt[q][tid];
                wcnt[q][tid] = acc;
                acc += c;
            }
            run[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            uint32_t pos = base[d] + wcnt[w][d] + rank;
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < RS_THREADS / STC_WAVE; q++) wcnt[q][tid] = 0;
        __syncthreads();
    }
}

template <typename K>
static int radix_sort(K *keys, uint32_t *vals, K *keys_alt, uint32_t *vals_alt, uint32_t n, int lo_bit,
                      int hi_bit, uint32_t *temp, hipStream_t s) {
    if (n <= 1 || hi_bit <= lo_bit) return 0;
    uint32_t tiles = ceil_div(n, RS_TILE);
    uint32_t *hist = temp;
    uint32_t *scan_temp = temp + (size_t)RS_BINS * tiles;
    K *src_k = keys, *dst_k = keys_alt;
    uint32_t *src_v = vals, *dst_v = vals_alt;
    int passes = 0;
    for (int shift = lo_bit; shift < hi_bit; shift += 8, passes++) {
        hipLaunchKernelGGL(k_rs_hist<K>, dim3(tiles), dim3(RS_THREADS), 0, s, (const K *)src_k, n, shift, hist,
                           tiles);
        STC_CHECK(hipGetLastError());
        int rc = stc_exclusive_scan_u32(hist, hist, RS_BINS * tiles, nullptr, scan_temp, s);
        if (rc) return rc;
        if constexpr (sizeof(K) == 4)
            hipLaunchKernelGGL(k_rs_scatter<K>, dim3(tiles), dim3(RS_THREADS), 0, s, (const K *)src_k,
                               (const uint32_t *)src_v, dst_k, dst_v, n, shift, (const uint32_t *)hist, tiles);
        else
            hipLaunchKernelGGL(k_rs_scatter_direct<K>, dim3(tiles), dim3(RS_THREADS), 0, s, (const K *)src_k,
                               (const uint32_t *)src_v, dst_k, dst_v, n, shift, (const uint32_t *)hist, tiles);
        STC_CHECK(hipGetLastError());
        K *tk = src_k; src_k = dst_k; dst_k = tk;
        uint32_t *tv = src_v; src_v = dst_v; dst_v = tv;
    }
    if (passes & 1) {
        STC_CHECK(hipMemcpyAsync(keys, keys_alt, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, s));
        STC_CHECK(hipMemcpyAsync(vals, vals_alt, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
    }
    return 0;
}

int stc_radix_sort_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt, uint32_t n,
                       int bits, uint32_t *temp, hipStream_t s) {
    return radix_sort<uint32_t>(keys, vals, keys_alt, vals_alt, n, 0, bits, temp, s);
}

int stc_radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt, uint32_t n,
                       int lo_bit, int hi_bit, uint32_t *temp, hipStream_t s) {
    return radix_sort<uint64_t>(keys, vals, keys_alt, vals_alt, n, lo_bit, hi_bit, temp, s);
}
