/*
 * stc_schema.h — the QTree nodes shared by the pipeline's kernels, and the
 * persistent QTree of the dynamic schema (stc_schema.hip).
 *
 * The reference keeps its QTree as an array of QTSchema nodes
 * (src/index/QTree.h:49-62). With the dynamic schema (its default,
 * src/util/config.h:35) the array lives across steps: it is built once on the
 * host from the first step's points (partitioner::build_schema,
 * src/tracing/filter.cpp:27-93) and after every step overfull leaves split and
 * underfull branches merge (cuda_update_schema_collect/split/merge,
 * src/cuda/geometry.cu:1114-1234, launched at :2159-2182); free node slots
 * return to a stack. Here the same tree is built and updated on the device.
 */
#ifndef STC_SCHEMA_H_
#define STC_SCHEMA_H_

#include "stc_internal.h"
#include "../../include/stcdb.h"

#define T_BRANCH 1u
#define T_LEAF 2u

struct Node {                 /* QTSchema (src/index/QTree.h:49-62) without its counters */
    double lo0, lo1, hi0, hi1;
    double mid_x, mid_y;
    uint32_t child[4];
    uint32_t type, leaf;      /* type 0 marks a free slot of a dynamic schema */
    uint32_t code, level;     /* Morton code of a rebuilt block; depth of a dynamic node */
};

struct MapBox { double lo0, lo1, hi0, hi1; };

/* The persistent tree: nodes in slots [0, hwm), free slots listed in fstack,
 * per slot the overflow and underflow counts of the update rule. */
struct SchemaState {
    Node *nodes = nullptr;
    uint32_t *nstate = nullptr;        /* overflow_count | underflow_count << 16 */
    uint32_t *fstack = nullptr;        /* free slots */
    uint32_t *fs = nullptr, *fm = nullptr, *rs = nullptr, *rm = nullptr; /* per-slot flags and ranks */
    uint32_t *pnode = nullptr;         /* initial build: each point's node (bit 31: settled) */
    uint32_t *cnt = nullptr;           /* initial build: points per node */
    uint32_t *dcount = nullptr;        /* device: splits, merges, ambiguous, leaves, ... */
    uint32_t *hcount = nullptr;        /* pinned copy */
    uint32_t *amb = nullptr;           /* initial build: nodes whose width test is within the guard band */
    uint32_t cap = 0;                  /* slots allocated */
    uint32_t n_points = 0;
    uint32_t hwm = 0, n_free = 0, n_valid = 0, n_leaves = 0;
    bool present = false;              /* a tree is installed */
    bool dirty = false;                /* leaf ids must be renumbered before the next step */
};

int stc_schema_init(SchemaState *S, uint32_t n_points, uint32_t slots);
void stc_schema_free(SchemaState *S);
/* partitioner::build_schema: the initial tree from one step's points (device);
 * order: the points in Morton order (a permutation of 0..n-1), or null */
int stc_schema_build(SchemaState *S, const double2 *xy, const uint32_t *order, uint32_t n, MapBox m, uint32_t cap,
                     double r, uint32_t *temp, size_t temp_words, hipStream_t s);
/* install a host-built tree (the schema cuda_create_device_bench copies in) */
int stc_schema_set(SchemaState *S, const stc_qtnode *nodes, uint32_t n, hipStream_t s);
int stc_schema_get(SchemaState *S, stc_qtnode *nodes, uint64_t cap, uint64_t *n, hipStream_t s);
/* leaf ids 0..n_leaves-1 in slot order; leaf_node[leaf] = slot */
int stc_schema_number_leaves(SchemaState *S, uint32_t *leaf_node, uint32_t *temp, size_t temp_words,
                             hipStream_t s);
/* the update after a step from its leaf counts (all points of each leaf) */
int stc_schema_update(SchemaState *S, const uint32_t *leaf_count, uint32_t cap, double r, uint32_t delay,
                      uint32_t *temp, size_t temp_words, hipStream_t s);

#endif
