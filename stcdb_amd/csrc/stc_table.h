/*
 * stc_table.h — layout of the live meetings, shared by the pipeline and the
 * query kernels.
 *
 * A bucket is the reference's meeting_unit (src/tracing/workbench.h:62-83)
 * without its end time: 28 bytes of key, start and MBR. The MBR is stored as
 * order-preserving u32 images of the f32 bounds so that min/max updates are
 * integer ops. Keys are oid1 << 32 | oid2 with oid1 < oid2 < 2^26.
 *
 * The live meetings after a step are one array of buckets sorted by key, one
 * segment per owner oid1 (a pair's smaller id). Bucket i of a step's array is
 * built from the step's pair slot i, so a slot that held a duplicate pair or
 * a rejected band pair leaves a dead bucket: key oid1 << 32 | DEAD_B (it sorts
 * after the owner's live pairs). The array has no free buckets: every bucket
 * is live or dead, and a live meeting was last seen at the step that built the
 * array, which is therefore its end time; the readers take it from the array's
 * step instead of from the bucket.
 */
#ifndef STC_TABLE_H_
#define STC_TABLE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#define DEAD_B 0xFFFFFFFFu /* partner word of a dead slot or bucket */

/* packed to 28 bytes (4-byte aligned): the merge streams two arrays of them per step */
struct __attribute__((packed, aligned(4))) Bucket {
    unsigned long long key;
    uint32_t start;
    uint32_t mbr[4];          /* ordered ints of low0, low1, high0, high1 */
};
static_assert(sizeof(Bucket) == 28, "a bucket is 28 bytes");

/* a bucket that holds a live meeting (not a dead slot) */
__host__ __device__ inline bool bucket_live(const Bucket &b) { return (uint32_t)b.key != DEAD_B; }

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
