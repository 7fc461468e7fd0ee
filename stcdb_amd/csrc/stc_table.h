/*
 * stc_table.h — layout of the meeting table shared by the pipeline and the
 * query kernels.
 *
 * A bucket is the reference's meeting_unit (src/tracing/workbench.h:62-83):
 * 32 bytes of key, start, end and MBR. The MBR is stored as order-preserving
 * u32 images of the f32 bounds so that min/max updates are integer ops.
 * Keys are oid1 << 32 | oid2 with oid1 < oid2 < 2^26. Keys whose high word is
 * all ones mark free buckets (the low word tags the identify pass that freed
 * the bucket; a cleared table holds EMPTY_KEY); high word 0xFFFFFFFE marks a
 * tombstone, an occupied bucket that matches no key.
 */
#ifndef STC_TABLE_H_
#define STC_TABLE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#define EMPTY_KEY 0xFFFFFFFFFFFFFFFFull
#define FREE_KEY(tag) (0xFFFFFFFF00000000ull | (unsigned long long)(uint32_t)(tag))
#define TOMB_KEY(tag) (0xFFFFFFFE00000000ull | (unsigned long long)(uint32_t)(tag))
__host__ __device__ inline bool key_free(unsigned long long k) { return (k >> 32) == 0xFFFFFFFFull; }
__host__ __device__ inline bool key_tomb(unsigned long long k) { return (k >> 32) == 0xFFFFFFFEull; }

struct __align__(32) Bucket {
    unsigned long long key;
    uint32_t start, end;
    uint32_t mbr[4];          /* ordered ints of low0, low1, high0, high1 */
};

__host__ __device__ inline uint32_t f2o(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float o2f(uint32_t o) {
    uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    float f;
    memcpy(&f, &b, 4);
    return f;
}

/* float_to_uint (src/cuda/cuda_util.cuh:115-119): the quantised coordinate of
 * the per-object MBRs, saturating as the reference GPU converts */
__host__ __device__ inline uint32_t f2q(float xy) {
    xy += 180;
    float v = xy * 100000;
    if (!(v > 0.0f)) return 0;
    if (v >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

#endif /* STC_TABLE_H_ */
