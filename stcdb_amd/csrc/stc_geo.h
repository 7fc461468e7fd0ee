/*
 * stc_geo.h — distance arithmetic of the co-occurrence pipeline (host + device).
 *
 * The reference decides proximity with
 *     dx = (x1 - x2) / dpm_lon(y1);  dy = (y1 - y2) / dpm_lat;
 *     d  = sqrt(dx*dx + dy*dy)  <=  reach_distance
 * where dpm_lon(lat) = 360 / (sin((90 - |lat|) * PI / 180) * 40076 * 1000)
 * and PI = 3.14159265 (src/cuda/cuda_util.cuh:41-53, src/util/util.h:38, 59-66).
 * Every operation is a correctly rounded IEEE double operation once FMA
 * contraction is off, except sin. The device evaluates sin in double-double
 * arithmetic (about 104 significant bits) and rounds once (stc_sin_dd below),
 * which gives the correctly rounded sine. The C library's sin, which the CPU
 * path uses, differs from the correctly rounded value by one ulp for about
 * 0.1% of arguments; that moves a distance by at most a few ulp. The device
 * therefore decides every pair whose distance lies more than 16 ulp away from
 * r, and hands the pairs inside that band to the host, which re-evaluates them
 * with the C library's sin: pair decisions equal the CPU formula's. dpm_lon
 * depends only on the probe point, so it is computed once per point per step.
 *
 * Everything here must be compiled with -ffp-contract=off.
 */
#ifndef STC_GEO_H_
#define STC_GEO_H_

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define STC_GHD __host__ __device__ inline
#else
#define STC_GHD static inline
#endif

#define STC_PI_REF 3.14159265
#define STC_DPM_LAT (360.0 / (40076.0 * 1000.0))
#define STC_INV_DPM_LAT (40076.0 * 1000.0 / 360.0)

/* (-1)^k / (2k+1)! as unevaluated double-double sums, k = 0..17 */
#define STC_SIN_TERMS 18
#define STC_SIN_COEFFS                                                          \
    {0x1.0000000000000p+0, 0x0.0p+0},                                           \
    {-0x1.5555555555555p-3, -0x1.5555555555555p-57},                            \
    {0x1.1111111111111p-7, 0x1.1111111111111p-63},                              \
    {-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73},                           \
    {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},                            \
    {-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80},                            \
    {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87},                             \
    {-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97},                           \
    {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103},                            \
    {-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112},                          \
    {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120},                           \
    {-0x1.761b41316381ap-75, 0x1.3423c7d91404fp-130},                           \
    {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139},                           \
    {-0x1.d1ab1c2dccea3p-94, -0x1.054d0c78aea14p-149},                          \
    {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157},                           \
    {-0x1.434d2e783f5bcp-113, -0x1.0b87b91be9affp-167},                         \
    {0x1.3981254dd0d52p-123, -0x1.2b1f4c8015a2fp-177},                          \
    {-0x1.0dc59c716d91fp-133, -0x1.419e3fad3f031p-188}

typedef struct stc_dd { double hi, lo; } stc_dd;

STC_GHD stc_dd stc_two_sum(double a, double b) {
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    stc_dd r = {s, e};
    return r;
}
STC_GHD stc_dd stc_quick_two_sum(double a, double b) {
    double s = a + b;
    double e = b - (s - a);
    stc_dd r = {s, e};
    return r;
}
STC_GHD stc_dd stc_dd_add(stc_dd a, stc_dd b) {
    stc_dd s = stc_two_sum(a.hi, b.hi);
    stc_dd t = stc_two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = stc_quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return stc_quick_two_sum(s.hi, s.lo);
}
STC_GHD stc_dd stc_dd_mul(stc_dd a, stc_dd b) {
    double p = a.hi * b.hi;
    double e = fma(a.hi, b.hi, -p);
    e += a.hi * b.lo + a.lo * b.hi;
    return stc_quick_two_sum(p, e);
}

/* sin(x) for x in [0, pi/2], rounded once from a double-double Taylor sum */
STC_GHD double stc_sin_dd(double x) {
    const double c[STC_SIN_TERMS][2] = {STC_SIN_COEFFS};
    stc_dd xx = {x, 0.0};
    stc_dd x2 = stc_dd_mul(xx, xx);
    stc_dd s = {c[STC_SIN_TERMS - 1][0], c[STC_SIN_TERMS - 1][1]};
    for (int k = STC_SIN_TERMS - 2; k >= 0; k--) {
        stc_dd ck = {c[k][0], c[k][1]};
        s = stc_dd_add(stc_dd_mul(s, x2), ck);
    }
    s = stc_dd_mul(s, xx);
    return s.hi + s.lo;
}

/* degree_per_meter_longitude, device form (cuda_util.cuh:41-44) */
STC_GHD double stc_dpm_lon(double lat) {
    return 360.0 / (stc_sin_dd((90 - fabs(lat)) * STC_PI_REF / 180) * 40076.0 * 1000.0);
}

/* point distance with the probe's longitude scale (cuda_util.cuh:46-53) */
STC_GHD double stc_point_dist(double x1, double y1, double x2, double y2, double dlon1) {
    double dx = x1 - x2;
    double dy = y1 - y2;
    dx = dx / dlon1;
    dy = dy / STC_DPM_LAT;
    return sqrt(dx * dx + dy * dy);
}

/* distance from a point to a box (geometry.cu:33-42) */
STC_GHD double stc_box_dist(double lo0, double lo1, double hi0, double hi1, double px, double py, double dlon) {
    double dx = fmax(fabs(px - (lo0 + hi0) / 2) - (hi0 - lo0) / 2, 0.0);
    double dy = fmax(fabs(py - (lo1 + hi1) / 2) - (hi1 - lo1) / 2, 0.0);
    dy = dy / STC_DPM_LAT;
    dx = dx / dlon;
    return sqrt(dx * dx + dy * dy);
}

#endif /* STC_GEO_H_ */
