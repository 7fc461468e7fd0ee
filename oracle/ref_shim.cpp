// ref_shim.cpp — exports the reference's own geometry and key helpers with C
// linkage so tests can pin the oracle against them. TEST INFRASTRUCTURE ONLY.
// Wraps, without modification: Point::distance and box::distance
// (src/geometry/geometry.h:40-48, 174-185), degree_per_meter_longitude,
// cantorPairing, InverseCantorPairing1, min_bits_to_store and the get_key_*
// field extractors (src/util/util.h:59-78, 527-553).
#include "geometry/geometry.h"
#include <stdint.h>

extern "C" {
double ref_point_distance_geo(double x1, double y1, double x2, double y2) {
    Point a(x1, y1), b(x2, y2);
    return a.distance(b, true);
}
double ref_box_distance_geo(const double *b4, double px, double py) {
    box b(b4[0], b4[1], b4[2], b4[3]);
    Point p(px, py);
    return b.distance(p, true);
}
double ref_dpm_lon(double lat) { return degree_per_meter_longitude(lat); }
double ref_dpm_lat(void) { return degree_per_meter_latitude; }
uint64_t ref_cantor(uint32_t a, uint32_t b) { return cantorPairing(a, b); }
void ref_inverse_cantor(uint64_t z, uint32_t *a, uint32_t *b) {
    pair<uint, uint> p = InverseCantorPairing1(z);
    *a = p.first;
    *b = p.second;
}
uint32_t ref_min_bits(uint32_t v) { return min_bits_to_store(v); }
void ref_key_fields(uint64_t lo, uint64_t hi, uint32_t *out5) {
    __uint128_t k = ((__uint128_t)hi << 64) | lo;
    out5[0] = get_key_sid(k);
    out5[1] = get_key_oid(k);
    out5[2] = get_key_target(k);
    out5[3] = get_key_duration(k);
    out5[4] = get_key_end(k);
}
// how this compiler converts a double to the reference's `uint` and `int`
// (the conversions inside CTF::get_ctf_bits, step_merge.cpp:149-171)
uint32_t ref_conv_d2u(double v) { uint x = v; return x; }
int32_t ref_conv_d2i(double v) { int x = v; return x; }
}
