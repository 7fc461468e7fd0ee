/*
 * stc_flush.hip — sort-based flush of the memtable into one run (CTB) of
 * ctf_count CTFs (process_with_gpu's flush chain, src/cuda/geometry.cu:1752-2137).
 */
#include "stc_flush.h"

#include <cmath>
#include <cstring>

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

int stc_flush_init(FlushWork *w, uint32_t n, uint32_t ctf_count) {
    w->n = n;
    w->ctf_count = ctf_count;
    return 0;
}

void stc_flush_free(FlushWork *w) {
    (void)w;
}

int stc_flush_run(FlushWork *w, const FlushArgs *a, FlushOut *out, hipStream_t s) {
    (void)w; (void)a; (void)out; (void)s;
    return STC_E_STATE;
}
