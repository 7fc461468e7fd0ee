/*
 * stc_gen.h — seeded synthetic trace generator (header-only, host + device).
 *
 * Objects walk a Manhattan street grid. Each object sits on a street segment
 * between two intersections: (ix, iy) is the last intersection passed, dir is
 * the heading (0:+x 1:+y 2:-x 3:-y) and off in [0, spacing) is the distance
 * travelled past (ix, iy). Every second an object advances by its constant
 * speed; when it reaches the next intersection it picks a new heading
 * uniformly among those that keep it on the grid (U-turns allowed) and spends
 * the remaining distance on the new segment. Speeds are below the spacing, so
 * at most one intersection is crossed per second.
 *
 * Determinism: all per-step decisions come from a counter-based hash of
 * (seed, object id, timestep), and the position arithmetic is plain IEEE
 * add/sub/mul on doubles. Compiled with floating-point contraction off, the
 * host and the GPU produce bit-identical coordinates. Initial placement
 * (which may use log/cos for hotspot starts) runs on the host only; the
 * device receives the initial state.
 *
 * The reference's own generator (src/tracing/generator.cpp:387-406) seeds
 * from time(NULL) and shares rand() across threads, so it cannot reproduce a
 * trace; this generator replaces it for all seeded workloads. Coordinates are
 * produced in degrees as the reference's Point (src/geometry/geometry.h:23-56)
 * expects: x = lon0 + X * dlon, y = lat0 + Y * dlat, with X, Y in metres.
 */
#ifndef STC_GEN_H_
#define STC_GEN_H_

#include <stdint.h>

#if defined(__HIPCC__)
#define STC_HD __host__ __device__ inline
#else
#define STC_HD static inline
#endif

typedef struct stc_walker {
    int32_t ix, iy;   /* last intersection passed (grid indices) */
    int32_t dir;      /* 0:+x 1:+y 2:-x 3:-y */
    int32_t pad;
    double off;       /* metres past (ix, iy) along dir, in [0, spacing) */
    double speed;     /* metres per second, constant per object */
} stc_walker;

typedef struct stc_gen_params {
    uint64_t seed;
    uint32_t num_objects;
    uint32_t grid_cells;     /* intersections run 0..grid_cells on each axis */
    double spacing;          /* metres between intersections */
    double vmin, vmax;       /* speed range, metres per second */
    uint32_t n_hotspots;     /* 0: all starts uniform */
    uint32_t pad;
    double hotspot_frac;     /* fraction of objects starting near a hotspot */
    double hotspot_sigma;    /* metres */
    double lon0, lat0;       /* degrees at street-grid origin (X = Y = 0) */
    double dlon, dlat;       /* degrees per metre along x and y */
} stc_gen_params;

STC_HD uint64_t stc_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* counter-based random 64-bit word for (seed, object, step, purpose) */
STC_HD uint64_t stc_rand(uint64_t seed, uint64_t oid, uint64_t step, uint64_t salt) {
    uint64_t h = stc_mix64(seed ^ (salt * 0xD1B54A32D192ED03ull));
    h = stc_mix64(h ^ oid);
    return stc_mix64(h ^ (step * 0x8CB92BA72F3D8DD7ull));
}

/* uniform double in [0, 1) from the top 53 bits (exact) */
STC_HD double stc_u01(uint64_t h) {
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

STC_HD int stc_dir_dx(int d) { return d == 0 ? 1 : (d == 2 ? -1 : 0); }
STC_HD int stc_dir_dy(int d) { return d == 1 ? 1 : (d == 3 ? -1 : 0); }

/* pick a heading uniformly among those leaving (ix, iy) onto the grid */
STC_HD int stc_pick_dir(int32_t ix, int32_t iy, int32_t gmax, uint64_t h) {
    int valid[4];
    int nv = 0;
    for (int d = 0; d < 4; d++) {
        int nx = ix + stc_dir_dx(d), ny = iy + stc_dir_dy(d);
        if (nx >= 0 && nx <= gmax && ny >= 0 && ny <= gmax) valid[nv++] = d;
    }
    return valid[(int)(h % (uint64_t)nv)];
}

/* advance one second (timestep t is the step being produced, t >= 1) */
STC_HD void stc_walk_step(stc_walker *w, const stc_gen_params *p, uint32_t oid, uint32_t t) {
    double o = w->off + w->speed;
    if (o >= p->spacing) {
        w->ix += stc_dir_dx(w->dir);
        w->iy += stc_dir_dy(w->dir);
        o = o - p->spacing;
        w->dir = stc_pick_dir(w->ix, w->iy, (int32_t)p->grid_cells, stc_rand(p->seed, oid, t, 2));
    }
    w->off = o;
}

/* position in degrees (AoS Point layout: x = longitude, y = latitude) */
STC_HD void stc_walk_pos(const stc_walker *w, const stc_gen_params *p, double *x, double *y) {
    double X = (double)w->ix * p->spacing + (double)stc_dir_dx(w->dir) * w->off;
    double Y = (double)w->iy * p->spacing + (double)stc_dir_dy(w->dir) * w->off;
    *x = p->lon0 + X * p->dlon;
    *y = p->lat0 + Y * p->dlat;
}

#include <math.h>

/* Initial state of objects [first, first + count). Uniform objects start on a
 * uniformly chosen intersection; hotspot objects start at the intersection
 * nearest to a Gaussian sample around one of n_hotspots uniform centres.
 * Heading, offset along the first segment and speed are uniform. Host only
 * (uses the C library's log/cos/sin). */
static inline void stc_gen_init(const stc_gen_params *p, stc_walker *w, uint32_t first, uint32_t count) {
    const uint64_t s = p->seed;
    const int32_t g = (int32_t)p->grid_cells;
    const double two_pi = 6.283185307179586;
    for (uint32_t k = 0; k < count; k++) {
        uint32_t oid = first + k;
        stc_walker *o = &w[k];
        int32_t ix, iy;
        if (p->n_hotspots > 0 && stc_u01(stc_rand(s, oid, 0, 3)) < p->hotspot_frac) {
            uint32_t h = (uint32_t)(stc_rand(s, oid, 0, 4) % p->n_hotspots);
            double cx = stc_u01(stc_rand(s, 0xFFFFFFFFull, h, 10)) * (double)g * p->spacing;
            double cy = stc_u01(stc_rand(s, 0xFFFFFFFFull, h, 11)) * (double)g * p->spacing;
            double u1 = stc_u01(stc_rand(s, oid, 0, 5));
            double u2 = stc_u01(stc_rand(s, oid, 0, 6));
            double rad = sqrt(-2.0 * log(1.0 - u1));
            double X = cx + p->hotspot_sigma * rad * cos(two_pi * u2);
            double Y = cy + p->hotspot_sigma * rad * sin(two_pi * u2);
            ix = (int32_t)floor(X / p->spacing + 0.5);
            iy = (int32_t)floor(Y / p->spacing + 0.5);
            ix = ix < 0 ? 0 : (ix > g ? g : ix);
            iy = iy < 0 ? 0 : (iy > g ? g : iy);
        } else {
            ix = (int32_t)(stc_rand(s, oid, 0, 12) % (uint64_t)(g + 1));
            iy = (int32_t)(stc_rand(s, oid, 0, 13) % (uint64_t)(g + 1));
        }
        o->ix = ix;
        o->iy = iy;
        o->dir = stc_pick_dir(ix, iy, g, stc_rand(s, oid, 0, 7));
        o->pad = 0;
        o->off = stc_u01(stc_rand(s, oid, 0, 8)) * p->spacing;
        o->speed = p->vmin + (p->vmax - p->vmin) * stc_u01(stc_rand(s, oid, 0, 9));
    }
}

#endif /* STC_GEN_H_ */
