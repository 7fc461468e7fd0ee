/*
 * stc_internal.h — shared declarations of the HIP implementation.
 */
#ifndef STC_INTERNAL_H_
#define STC_INTERNAL_H_

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define STC_WAVE 64

#define STC_CHECK(expr)                                                        \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) return stc_hip_fail(e_, #expr, __FILE__, __LINE__); \
    } while (0)

int stc_hip_fail(hipError_t e, const char *expr, const char *file, int line);

/* ---------------- radix sort (stc_radix.hip) ---------------- */
/* Stable LSD radix sort of (key, value) pairs by key bits [0, bits).
 * Sorts in place through the alternate buffers; after the call the sorted
 * data is in keys/vals (the alternates are scratch). temp must hold
 * stc_radix_temp_words(n) uint32 words. */
size_t stc_radix_temp_words(uint32_t n);
int stc_radix_sort_u32(uint32_t *keys, uint32_t *vals, uint32_t *keys_alt, uint32_t *vals_alt, uint32_t n,
                       int bits, uint32_t *temp, hipStream_t s);
int stc_radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt, uint32_t n,
                       int lo_bit, int hi_bit, uint32_t *temp, hipStream_t s);
/* exclusive prefix sum of n uint32 values (in may equal out); writes the total
 * to *total_dev when non-null. temp as for the radix sort. */
int stc_exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *total_dev, uint32_t *temp,
                           hipStream_t s);

#endif
