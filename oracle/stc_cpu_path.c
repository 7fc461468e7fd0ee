/*
 * stc_cpu_path.c — the reference's CPU path, restated (test infrastructure:
 * the CPU baseline of the benchmark and a cross-check of the oracle).
 *
 * TEST INFRASTRUCTURE ONLY. When config->gpu is off, tracer::process runs per
 * timestep (src/tracing/process.cpp:452-459, 608-612):
 *   workbench::filter()         filter_unit, src/tracing/filter.cpp:210-313
 *   workbench::reachability()   reachability_unit, src/tracing/refinement.cpp:10-85
 *   workbench::update_meetings()  update_meetings_unit, src/tracing/meeting.cpp:32-101
 *   workbench::update_schema()  update_schema_unit, src/tracing/filter.cpp:377-449
 * each on config->num_threads pthreads that take batches of units from a
 * shared counter (query_context::next_batch, src/util/query_context.h:58-74:
 * batch = max(1, units / num_batchs), num_batchs 10000 for the filter and
 * 1000 otherwise). The tree is partitioner::build_schema's
 * (filter.cpp:27-93). This file restates those stages with the same data
 * layout (grids of grid_amplify * grid_capacity pids, 12-byte check units,
 * 32-byte meeting buckets in an open-addressed table of num_meeting_buckets
 * slots, key = Cantor(pid1, pid2), linear probing without a bound, expired
 * buckets reset) and the same thread structure. Differences, each the
 * race-free form of a reference race: a grid's slot and a bucket's key are
 * claimed with atomics where the reference checks and writes outside its
 * locks (filter.cpp:95-111, refinement.cpp:43-52); the schema update takes
 * all decisions first and then splits and merges (the reference's in-place
 * update can merge a parent on a child count it has just cleared).
 * The CPU path keeps no MBR (meeting buckets are host memory that
 * reachability_unit never fills) and has no flush.
 */
#include "stc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static const double PI_REF = 3.14159265;                  /* util.h:38 */
static const double DPM_LAT = 360.0 / (40076.0 * 1000.0); /* util.h:57 */
#define EMPTY_KEY (1ull << 62)                            /* ULL_MAX, util.h:36 */

enum { C_INVALID = 0, C_BRANCH = 1, C_LEAF = 2 };

typedef struct cnode {          /* QTSchema (QTree.h:49-62) */
    double lo[2], hi[2];
    double mid_x, mid_y;
    uint32_t type, child[4], grid;
    uint32_t ovf, unf;
} cnode;

typedef struct cunit {          /* checking_unit (workbench.h:55-60) */
    uint32_t pid, gid;
    uint8_t inside;
} cunit;

typedef struct cbucket {        /* meeting_unit (workbench.h:62-83) */
    _Atomic uint64_t key;
    uint32_t start, end;
    float mbr[4];               /* never written on the CPU path */
} cbucket;

typedef struct uvec { cunit *a; uint64_t n, cap; } uvec;
typedef struct kvec { uint64_t *a; uint64_t n, cap; } kvec;
typedef struct mvec { uint32_t (*a)[4]; uint64_t n, cap; } mvec;

typedef struct cpath {
    orc_config cfg;
    uint32_t nthreads;
    uint64_t B;
    int keep_pairs;             /* export each step's pair set (tests); the reference only counts them */
    double x_buffer, y_buffer;
    char err[256];
    /* schema */
    cnode *nodes;
    uint32_t n_nodes, cap_nodes;
    uint32_t *free_nodes, n_free, cap_free;
    int have_schema;
    /* grids (one per leaf, numbered each step) */
    uint32_t n_grids, buf;
    uint32_t *grid_node;
    _Atomic uint32_t *grid_counter;
    uint32_t *grids;
    /* check units: the N inside units, then the cross units */
    cunit *units;
    uint64_t n_units, cap_units;
    /* meetings */
    cbucket *buckets;
    const double *xy;
    uint32_t t;
    _Atomic int fail;
    /* per-thread outputs */
    uvec *tunits;
    kvec *tpairs;
    mvec *tmeet;
    uint64_t *tcand;
    uint64_t *ttaken, *tactive;
    /* last step */
    double ms[4];
    uint64_t counts[6];   /* units, candidates, pairs, taken, active, identified */
    uint64_t *pairs; uint64_t n_pairs;
    uint32_t (*meetings)[4]; uint64_t n_meetings;
} cpath;

static double now_ms(void) {
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1e3 + tv.tv_usec * 1e-3;
}

/* degree_per_meter_longitude (util.h:59-66) */
static double dpm_lon(double lat) {
    double a = fabs(lat);
    if (a == 90) a = 89;
    return 360.0 / (sin((90 - a) * PI_REF / 180) * 40076.0 * 1000.0);
}

static int push_u(uvec *v, cunit u) {
    if (v->n == v->cap) {
        uint64_t nc = v->cap ? v->cap * 2 : 4096;
        cunit *a = (cunit *)realloc(v->a, nc * sizeof(cunit));
        if (!a) return -1;
        v->a = a;
        v->cap = nc;
    }
    v->a[v->n++] = u;
    return 0;
}
static int push_k(kvec *v, uint64_t k) {
    if (v->n == v->cap) {
        uint64_t nc = v->cap ? v->cap * 2 : 4096;
        uint64_t *a = (uint64_t *)realloc(v->a, nc * 8);
        if (!a) return -1;
        v->a = a;
        v->cap = nc;
    }
    v->a[v->n++] = k;
    return 0;
}
static int push_m(mvec *v, uint32_t a0, uint32_t a1, uint32_t s, uint32_t e) {
    if (v->n == v->cap) {
        uint64_t nc = v->cap ? v->cap * 2 : 1024;
        uint32_t(*a)[4] = realloc(v->a, nc * sizeof(*a));
        if (!a) return -1;
        v->a = a;
        v->cap = nc;
    }
    v->a[v->n][0] = a0; v->a[v->n][1] = a1; v->a[v->n][2] = s; v->a[v->n][3] = e;
    v->n++;
    return 0;
}

/* ------------------------------------------------------------------ */
/* query_context: threads taking batches from one counter              */
/* ------------------------------------------------------------------ */

typedef struct qctx {
    cpath *c;
    uint64_t num_units, batch;
    _Atomic uint64_t counter;
    void (*fn)(cpath *, int tid, uint64_t start, uint64_t end);
} qctx;

typedef struct targ { qctx *q; int tid; } targ;

static void *worker(void *a) {
    targ *ta = (targ *)a;
    qctx *q = ta->q;
    for (;;) {
        uint64_t s = atomic_fetch_add(&q->counter, q->batch);
        if (s >= q->num_units) break;
        uint64_t e = s + q->batch < q->num_units ? s + q->batch : q->num_units;
        q->fn(q->c, ta->tid, s, e);
    }
    return NULL;
}

/* next_batch hands out max(1, num_units / num_batchs) units at a time */
static void run_threads(cpath *c, uint64_t num_units, uint64_t num_batchs,
                        void (*fn)(cpath *, int, uint64_t, uint64_t)) {
    qctx q;
    q.c = c;
    q.num_units = num_units;
    q.batch = num_units / num_batchs ? num_units / num_batchs : 1;
    atomic_init(&q.counter, 0);
    q.fn = fn;
    pthread_t th[256];
    targ ta[256];
    uint32_t nt = c->nthreads;
    for (uint32_t i = 0; i < nt; i++) {
        ta[i].q = &q;
        ta[i].tid = (int)i;
        pthread_create(&th[i], NULL, worker, &ta[i]);
    }
    for (uint32_t i = 0; i < nt; i++) pthread_join(th[i], NULL);
}

/* ------------------------------------------------------------------ */
/* schema                                                              */
/* ------------------------------------------------------------------ */

static int alloc_node(cpath *c, uint32_t *id) {
    if (c->n_free) { *id = c->free_nodes[--c->n_free]; return 0; }
    if (c->n_nodes == c->cap_nodes) {
        uint32_t nc = c->cap_nodes ? c->cap_nodes * 2 : 4096;
        cnode *nn = (cnode *)realloc(c->nodes, nc * sizeof(cnode));
        if (!nn) return -1;
        c->nodes = nn;
        c->cap_nodes = nc;
    }
    *id = c->n_nodes++;
    return 0;
}

static void set_box(cnode *k, double lo0, double lo1, double hi0, double hi1) {
    memset(k, 0, sizeof(*k));
    k->lo[0] = lo0; k->lo[1] = lo1; k->hi[0] = hi0; k->hi[1] = hi1;
    k->mid_x = (lo0 + hi0) / 2;
    k->mid_y = (lo1 + hi1) / 2;
    k->type = C_LEAF;
}

/* QTNode::insert/split (QTree.h:157-214): the tree does not depend on the
 * order of insertion, so one thread inserts the ids of a node at a time */
static int build_rec(cpath *c, uint32_t nd, const double *xy, uint32_t *ids, uint32_t n, uint32_t *tmp) {
    cnode b = c->nodes[nd];
    if (!(n >= c->cfg.grid_capacity && (b.hi[0] - b.lo[0]) / dpm_lon(b.lo[1]) / 2 > c->cfg.reach_distance)) return 0;
    uint32_t k[4];
    for (int i = 0; i < 4; i++)
        if (alloc_node(c, &k[i])) return -1;
    set_box(&c->nodes[k[0]], b.lo[0], b.lo[1], b.mid_x, b.mid_y);
    set_box(&c->nodes[k[1]], b.mid_x, b.lo[1], b.hi[0], b.mid_y);
    set_box(&c->nodes[k[2]], b.lo[0], b.mid_y, b.mid_x, b.hi[1]);
    set_box(&c->nodes[k[3]], b.mid_x, b.mid_y, b.hi[0], b.hi[1]);
    c->nodes[nd].type = C_BRANCH;
    memcpy(c->nodes[nd].child, k, sizeof(k));
    uint32_t cnt[4] = {0, 0, 0, 0}, at[4];
    for (uint32_t i = 0; i < n; i++) cnt[(xy[2 * ids[i] + 1] > b.mid_y) * 2 + (xy[2 * ids[i]] > b.mid_x)]++;
    at[0] = 0; at[1] = cnt[0]; at[2] = at[1] + cnt[1]; at[3] = at[2] + cnt[2];
    uint32_t st[4] = {at[0], at[1], at[2], at[3]};
    for (uint32_t i = 0; i < n; i++) {
        uint32_t p = ids[i];
        tmp[at[(xy[2 * p + 1] > b.mid_y) * 2 + (xy[2 * p] > b.mid_x)]++] = p;
    }
    memcpy(ids, tmp, n * sizeof(uint32_t));
    for (int i = 0; i < 4; i++)
        if (build_rec(c, k[i], xy, ids + st[i], cnt[i], tmp + st[i])) return -1;
    return 0;
}

static int build_schema(cpath *c, const double *xy) {
    uint32_t n = c->cfg.num_objects, root;
    c->n_nodes = 0;
    c->n_free = 0;
    if (alloc_node(c, &root)) return -1;
    set_box(&c->nodes[root], c->cfg.map_lo[0], c->cfg.map_lo[1], c->cfg.map_hi[0], c->cfg.map_hi[1]);
    uint32_t *ids = (uint32_t *)malloc(4ull * (n ? n : 1)), *tmp = (uint32_t *)malloc(4ull * (n ? n : 1));
    if (!ids || !tmp) { free(ids); free(tmp); return -1; }
    for (uint32_t i = 0; i < n; i++) ids[i] = i;
    int rc = build_rec(c, root, xy, ids, n, tmp);
    free(ids);
    free(tmp);
    c->have_schema = 1;
    return rc;
}

/* grids: one per leaf this step (the reference pops them from grids_stack) */
static int number_grids(cpath *c) {
    uint32_t ng = 0;
    for (uint32_t i = 0; i < c->n_nodes; i++)
        if (c->nodes[i].type == C_LEAF) ng++;
    if (ng > c->n_grids || !c->grid_node) {
        free(c->grid_node); free((void *)c->grid_counter); free(c->grids);
        c->grid_node = (uint32_t *)malloc(4ull * ng);
        c->grid_counter = (_Atomic uint32_t *)malloc(4ull * ng);
        c->grids = (uint32_t *)malloc(4ull * ng * c->buf);
        if (!c->grid_node || !c->grid_counter || !c->grids) return -1;
    }
    c->n_grids = 0;
    for (uint32_t i = 0; i < c->n_nodes; i++)
        if (c->nodes[i].type == C_LEAF) {
            c->nodes[i].grid = c->n_grids;
            c->grid_node[c->n_grids++] = i;
        }
    for (uint32_t g = 0; g < c->n_grids; g++) atomic_init(&c->grid_counter[g], 0); /* workbench::reset */
    return 0;
}

/* ------------------------------------------------------------------ */
/* filter (filter_unit, filter.cpp:210-288)                            */
/* ------------------------------------------------------------------ */

/* lookup_rec (filter.cpp:130-152) with box::distance (geometry.h:174-185) */
static void lookup_rec(cpath *c, int tid, uint32_t pid, double px, double py, uint32_t cur) {
    double r = c->cfg.reach_distance;
    for (int i = 0; i < 4; i++) {
        const cnode *ch = &c->nodes[c->nodes[cur].child[i]];
        double dx = fmax(fabs(px - (ch->lo[0] + ch->hi[0]) / 2) - (ch->hi[0] - ch->lo[0]) / 2, 0.0);
        double dy = fmax(fabs(py - (ch->lo[1] + ch->hi[1]) / 2) - (ch->hi[1] - ch->lo[1]) / 2, 0.0);
        dy = dy / DPM_LAT;
        dx = dx / dpm_lon(py);
        if (sqrt(dx * dx + dy * dy) <= r) {
            if (ch->type == C_LEAF) {
                if (py < ch->lo[1] || (py < ch->hi[1] && px < ch->lo[0])) {
                    cunit u = {pid, ch->grid, 0};
                    if (push_u(&c->tunits[tid], u)) atomic_store(&c->fail, 1);
                }
            } else {
                lookup_rec(c, tid, pid, px, py, c->nodes[cur].child[i]);
            }
        }
    }
}

static void filter_batch(cpath *c, int tid, uint64_t s, uint64_t e) {
    const double *xy = c->xy;
    for (uint64_t pid = s; pid < e; pid++) {
        double px = xy[2 * pid], py = xy[2 * pid + 1];
        uint32_t cur = 0, lv = 0;
        for (;;) {
            const cnode *nd = &c->nodes[cur];
            cur = nd->child[(py > nd->mid_y) * 2 + (px > nd->mid_x)];
            const cnode *ch = &c->nodes[cur];
            if (px + c->x_buffer < ch->hi[0] && py + c->y_buffer < ch->hi[1]) lv = cur;
            if (ch->type == C_LEAF) break;
        }
        /* workbench::insert (filter.cpp:95-111) */
        uint32_t g = c->nodes[cur].grid;
        uint32_t slot = atomic_fetch_add(&c->grid_counter[g], 1);
        if (slot < c->buf) c->grids[(uint64_t)g * c->buf + slot] = (uint32_t)pid;
        c->units[pid].pid = (uint32_t)pid;
        c->units[pid].gid = g;
        c->units[pid].inside = 1;
        if (lv != cur) lookup_rec(c, tid, (uint32_t)pid, px, py, lv);
    }
}

/* ------------------------------------------------------------------ */
/* reachability (reachability_unit, refinement.cpp:10-64)               */
/* ------------------------------------------------------------------ */

static void reach_batch(cpath *c, int tid, uint64_t s, uint64_t e) {
    const double *xy = c->xy;
    const double r = c->cfg.reach_distance;
    const uint64_t B = c->B;
    uint64_t cand = 0;
    for (uint64_t u = s; u < e; u++) {
        const cunit cu = c->units[u];
        uint32_t size = atomic_load_explicit(&c->grid_counter[cu.gid], memory_order_relaxed);
        if (size > c->buf) size = c->buf;
        const uint32_t *mem = c->grids + (uint64_t)cu.gid * c->buf;
        const double x1 = xy[2 * cu.pid], y1 = xy[2 * cu.pid + 1];
        const double dl = dpm_lon(y1);
        for (uint32_t i = 0; i < size; i++) {
            uint32_t q = mem[i];
            if (!(cu.pid < q || !cu.inside)) continue;
            cand++;
            /* Point::distance (geometry.h:40-48) */
            double dx = (x1 - xy[2 * q]) / dl, dy = (y1 - xy[2 * q + 1]) / DPM_LAT;
            if (!(sqrt(dx * dx + dy * dy) <= r)) continue;
            uint64_t a = cu.pid < q ? cu.pid : q, b = cu.pid < q ? q : cu.pid;
            uint64_t key = (a + b) * (a + b + 1) / 2 + b; /* cantorPairing (util.h:68-70) */
            uint64_t slot = key % B;
            for (uint64_t probe = 0;; probe++) {
                if (probe >= B) { atomic_store(&c->fail, 2); return; }
                cbucket *bk = &c->buckets[slot];
                uint64_t k = atomic_load_explicit(&bk->key, memory_order_acquire);
                if (k == key) { bk->end = c->t; break; }
                if (k == EMPTY_KEY) {
                    uint64_t exp = EMPTY_KEY;
                    if (atomic_compare_exchange_strong(&bk->key, &exp, key)) {
                        bk->start = c->t;
                        bk->end = c->t;
                        break;
                    }
                    if (exp == key) { bk->end = c->t; break; }
                }
                slot = (slot + 1) % B;
            }
        }
    }
    c->tcand[tid] += cand;
}

/* ------------------------------------------------------------------ */
/* update_meetings (update_meetings_unit, meeting.cpp:32-76)            */
/* ------------------------------------------------------------------ */

static uint32_t inv_cantor(uint64_t z, uint32_t *b) { /* InverseCantorPairing1 (util.h:72-78) */
    uint64_t w = (uint64_t)floor((sqrt(8.0 * (double)z + 1) - 1) / 2);
    while ((w + 1) * (w + 2) / 2 <= z) w++;
    while (w * (w + 1) / 2 > z) w--;
    uint64_t y = z - w * (w + 1) / 2;
    *b = (uint32_t)y;
    return (uint32_t)(w - y);
}

static void meet_batch(cpath *c, int tid, uint64_t s, uint64_t e) {
    const uint32_t t = c->t, m = c->cfg.min_meet_time;
    uint64_t taken = 0, active = 0;
    for (uint64_t bid = s; bid < e; bid++) {
        cbucket *bk = &c->buckets[bid];
        uint64_t key = atomic_load_explicit(&bk->key, memory_order_relaxed);
        if (key == EMPTY_KEY) continue;
        taken++;
        if (bk->end == t) { /* still active */
            active++;
            if (c->keep_pairs) {
                uint32_t b, a = inv_cantor(key, &b);
                if (push_k(&c->tpairs[tid], ((uint64_t)a << 32) | b)) atomic_store(&c->fail, 1);
            }
            continue;
        }
        if (t - bk->start >= m + 1) { /* batch_meet into h_meetings_ps (meeting.cpp:11-24) */
            uint32_t b, a = inv_cantor(key, &b);
            if (push_m(&c->tmeet[tid], a, b, bk->start, bk->end)) atomic_store(&c->fail, 1);
        }
        atomic_store_explicit(&bk->key, EMPTY_KEY, memory_order_relaxed); /* meeting_unit::reset */
    }
    c->ttaken[tid] += taken;
    c->tactive[tid] += active;
}

/* ------------------------------------------------------------------ */
/* update_schema (update_schema_unit, filter.cpp:377-424)               */
/* ------------------------------------------------------------------ */

static uint8_t *g_act; /* decisions of the running update (one update at a time per process) */

static void schema_batch(cpath *c, int tid, uint64_t s, uint64_t e) {
    (void)tid;
    const uint32_t cap = c->cfg.grid_capacity, delay = c->cfg.schema_update_delay ? c->cfg.schema_update_delay : 1;
    const double r2 = c->cfg.reach_distance * 2;
    for (uint64_t i = s; i < e; i++) {
        cnode *nd = &c->nodes[i];
        if (nd->type == C_LEAF) {
            if ((nd->hi[0] - nd->lo[0]) / dpm_lon(nd->lo[1]) > r2 &&
                atomic_load_explicit(&c->grid_counter[nd->grid], memory_order_relaxed) > cap) {
                if (++nd->ovf >= delay) { g_act[i] = 1; nd->ovf = 0; }
            } else {
                nd->ovf = 0;
            }
        } else if (nd->type == C_BRANCH) {
            uint32_t leafchild = 0, sum = 0;
            for (int k = 0; k < 4; k++) {
                const cnode *ch = &c->nodes[nd->child[k]];
                if (ch->type == C_LEAF) {
                    leafchild++;
                    sum += atomic_load_explicit(&c->grid_counter[ch->grid], memory_order_relaxed);
                }
            }
            if (leafchild == 4 && sum < cap) {
                if (++nd->unf >= delay) { g_act[i] = 2; nd->unf = 0; }
            } else {
                nd->unf = 0;
            }
        }
    }
}

static int update_schema(cpath *c) {
    uint32_t n = c->n_nodes;
    g_act = (uint8_t *)calloc(n ? n : 1, 1);
    if (!g_act) return -1;
    run_threads(c, n, 1000, schema_batch);
    for (uint32_t i = 0; i < n; i++) { /* split_node (filter.cpp:338-375) */
        if (g_act[i] != 1) continue;
        uint32_t k[4];
        for (int j = 0; j < 4; j++)
            if (alloc_node(c, &k[j])) { free(g_act); return -1; }
        cnode *p = &c->nodes[i];
        double xh = p->mid_x - p->lo[0], yh = p->mid_y - p->lo[1];
        for (int j = 0; j < 4; j++)
            set_box(&c->nodes[k[j]], p->lo[0] + (j % 2 == 1) * xh, p->lo[1] + (j / 2 == 1) * yh,
                    p->mid_x + (j % 2 == 1) * xh, p->mid_y + (j / 2 == 1) * yh);
        p->type = C_BRANCH;
        memcpy(p->child, k, sizeof(k));
    }
    for (uint32_t i = 0; i < n; i++) { /* merge_node (filter.cpp:316-336) */
        if (g_act[i] != 2) continue;
        cnode *p = &c->nodes[i];
        for (int j = 0; j < 4; j++) {
            c->nodes[p->child[j]].type = C_INVALID;
            if (c->n_free == c->cap_free) {
                uint32_t nc = c->cap_free ? c->cap_free * 2 : 1024;
                uint32_t *f = (uint32_t *)realloc(c->free_nodes, 4ull * nc);
                if (!f) { free(g_act); return -1; }
                c->free_nodes = f;
                c->cap_free = nc;
            }
            c->free_nodes[c->n_free++] = p->child[j];
        }
        p->type = C_LEAF;
    }
    free(g_act);
    g_act = NULL;
    return 0;
}

/* ------------------------------------------------------------------ */
/* API                                                                 */
/* ------------------------------------------------------------------ */

/* flags bit 0: export each step's pair set */
void *cpath_create(const orc_config *cfg, uint32_t num_threads, uint64_t num_buckets, uint32_t flags) {
    cpath *c = (cpath *)calloc(1, sizeof(cpath));
    if (!c) return NULL;
    c->cfg = *cfg;
    c->nthreads = num_threads < 1 ? 1 : (num_threads > 256 ? 256 : num_threads);
    c->B = num_buckets ? num_buckets : 2ull * cfg->num_objects; /* config.h:306-308 */
    c->buf = cfg->leaf_buffer;
    /* build_schema (filter.cpp:30-31): buffers from the map's lowest latitude */
    c->x_buffer = cfg->reach_distance * dpm_lon(cfg->map_lo[1]);
    c->y_buffer = cfg->reach_distance * DPM_LAT;
    c->buckets = (cbucket *)malloc(c->B * sizeof(cbucket));
    c->units = (cunit *)malloc(sizeof(cunit) * (cfg->num_objects ? cfg->num_objects : 1));
    c->tunits = (uvec *)calloc(c->nthreads, sizeof(uvec));
    c->tpairs = (kvec *)calloc(c->nthreads, sizeof(kvec));
    c->tmeet = (mvec *)calloc(c->nthreads, sizeof(mvec));
    c->tcand = (uint64_t *)calloc(c->nthreads, 8);
    c->ttaken = (uint64_t *)calloc(c->nthreads, 8);
    c->tactive = (uint64_t *)calloc(c->nthreads, 8);
    c->keep_pairs = flags & 1;
    if (!c->buckets || !c->units || !c->tunits || !c->tpairs || !c->tmeet || !c->tcand || !c->ttaken || !c->tactive) {
        free(c->buckets); free(c->units); free(c->tunits); free(c->tpairs); free(c->tmeet);
        free(c->tcand); free(c->ttaken); free(c->tactive); free(c);
        return NULL;
    }
    for (uint64_t i = 0; i < c->B; i++) { /* cuda_clean_buckets / meeting_unit::reset */
        atomic_init(&c->buckets[i].key, EMPTY_KEY);
        c->buckets[i].start = c->buckets[i].end = 0;
    }
    c->cap_units = cfg->num_objects;
    return c;
}

void cpath_destroy(void *h) {
    cpath *c = (cpath *)h;
    if (!c) return;
    for (uint32_t i = 0; i < c->nthreads; i++) { free(c->tunits[i].a); free(c->tpairs[i].a); free(c->tmeet[i].a); }
    free(c->tunits); free(c->tpairs); free(c->tmeet); free(c->tcand); free(c->ttaken); free(c->tactive);
    free(c->nodes); free(c->free_nodes); free(c->grid_node); free((void *)c->grid_counter); free(c->grids);
    free(c->units); free(c->buckets); free(c->pairs); free(c->meetings);
    free(c);
}

const char *cpath_last_error(void *h) { return ((cpath *)h)->err; }

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

int cpath_step(void *h, const double *xy, uint32_t t) {
    cpath *c = (cpath *)h;
    const uint32_t n = c->cfg.num_objects;
    c->xy = xy;
    c->t = t;
    atomic_store(&c->fail, 0);
    if (!c->have_schema && build_schema(c, xy)) { snprintf(c->err, sizeof c->err, "schema build failed"); return -1; }
    if (number_grids(c)) { snprintf(c->err, sizeof c->err, "out of memory"); return -1; }
    for (uint32_t i = 0; i < c->nthreads; i++) { c->tunits[i].n = 0; c->tpairs[i].n = 0; c->tmeet[i].n = 0; }
    memset(c->tcand, 0, 8ull * c->nthreads);
    memset(c->ttaken, 0, 8ull * c->nthreads);
    memset(c->tactive, 0, 8ull * c->nthreads);

    double t0 = now_ms();
    run_threads(c, n, 10000, filter_batch); /* workbench::filter, num_batchs 10000 (filter.cpp:300) */
    uint64_t cross = 0;                     /* batch_check: appended after the N inside units */
    for (uint32_t i = 0; i < c->nthreads; i++) cross += c->tunits[i].n;
    if (n + cross > c->cap_units) {
        cunit *u = (cunit *)realloc(c->units, sizeof(cunit) * (n + cross));
        if (!u) { snprintf(c->err, sizeof c->err, "out of memory"); return -1; }
        c->units = u;
        c->cap_units = n + cross;
    }
    uint64_t at = n;
    for (uint32_t i = 0; i < c->nthreads; i++) {
        memcpy(c->units + at, c->tunits[i].a, sizeof(cunit) * c->tunits[i].n);
        at += c->tunits[i].n;
    }
    c->n_units = at;
    double t1 = now_ms();
    run_threads(c, c->n_units, 1000, reach_batch); /* workbench::reachability */
    if (atomic_load(&c->fail) == 2) { snprintf(c->err, sizeof c->err, "meeting table full"); return -1; }
    double t2 = now_ms();
    run_threads(c, c->B, 1000, meet_batch); /* workbench::update_meetings */
    double t3 = now_ms();
    if (c->cfg.schema_mode != 0 && update_schema(c)) { snprintf(c->err, sizeof c->err, "schema update failed"); return -1; }
    double t4 = now_ms();
    if (atomic_load(&c->fail)) { snprintf(c->err, sizeof c->err, "out of memory"); return -1; }

    c->ms[0] = t1 - t0; c->ms[1] = t2 - t1; c->ms[2] = t3 - t2; c->ms[3] = t4 - t3;
    uint64_t np = 0, nm = 0, cand = 0, taken = 0, active = 0;
    for (uint32_t i = 0; i < c->nthreads; i++) {
        np += c->tpairs[i].n; nm += c->tmeet[i].n; cand += c->tcand[i]; taken += c->ttaken[i];
        active += c->tactive[i];
    }
    free(c->pairs);
    free(c->meetings);
    c->pairs = (uint64_t *)malloc(8 * (np ? np : 1));
    c->meetings = malloc(sizeof(*c->meetings) * (nm ? nm : 1));
    np = nm = 0;
    for (uint32_t i = 0; i < c->nthreads; i++) {
        memcpy(c->pairs + np, c->tpairs[i].a, 8 * c->tpairs[i].n);
        np += c->tpairs[i].n;
        memcpy(c->meetings + nm, c->tmeet[i].a, sizeof(*c->meetings) * c->tmeet[i].n);
        nm += c->tmeet[i].n;
    }
    qsort(c->pairs, np, 8, cmp_u64);
    c->n_pairs = np;
    c->n_meetings = nm;
    c->counts[0] = c->n_units; c->counts[1] = cand; c->counts[2] = active;
    c->counts[3] = taken; c->counts[4] = active; c->counts[5] = nm;
    return 0;
}

/* stage ms of the last step: filter, reachability, update_meetings, update_schema */
void cpath_stage_ms(void *h, double *ms) { memcpy(ms, ((cpath *)h)->ms, sizeof(double) * 4); }
/* units, candidates, pairs (active after the step), taken buckets, active, identified meetings */
void cpath_counts(void *h, uint64_t *out) { memcpy(out, ((cpath *)h)->counts, sizeof(uint64_t) * 6); }
uint64_t cpath_num_pairs(void *h) { return ((cpath *)h)->n_pairs; }
void cpath_get_pairs(void *h, uint64_t *keys) { cpath *c = (cpath *)h; memcpy(keys, c->pairs, 8 * c->n_pairs); }
uint64_t cpath_num_meetings(void *h) { return ((cpath *)h)->n_meetings; }
/* identified meetings of the last step: (pid1, pid2, start, end), in no order */
void cpath_get_meetings(void *h, uint32_t *m) {
    cpath *c = (cpath *)h;
    memcpy(m, c->meetings, sizeof(*c->meetings) * c->n_meetings);
}
uint64_t cpath_num_leaves(void *h) { return ((cpath *)h)->n_grids; }
/* leaves of the last step's tree (the tree that step used) and their counts */
void cpath_get_leaves(void *h, double *boxes, uint32_t *counts) {
    cpath *c = (cpath *)h;
    for (uint32_t g = 0; g < c->n_grids; g++) {
        const cnode *nd = &c->nodes[c->grid_node[g]];
        boxes[4 * g] = nd->lo[0]; boxes[4 * g + 1] = nd->lo[1];
        boxes[4 * g + 2] = nd->hi[0]; boxes[4 * g + 3] = nd->hi[1];
        counts[g] = atomic_load(&c->grid_counter[g]);
    }
}
