/*
 * stc_schema.hip — the persistent QTree of the dynamic schema on gfx950.
 *
 * Initial build (partitioner::build_schema, src/tracing/filter.cpp:27-93,
 * with QTNode::insert/split, src/index/QTree.h:157-214). A node splits when
 * at least grid_capacity points reach it by the descent rule
 * (y > mid_y) * 2 + (x > mid_x) and half its width in metres at its low
 * latitude (box::width(true), src/geometry/geometry.h:166-172) exceeds
 * reach_distance; its children are the exact halves of its box. Which points
 * reach a node does not depend on the order the reference's threads insert
 * them, so the tree is built level by level: every unsettled point adds
 * itself to its node's count, the nodes of the level decide, the splitting
 * ones append four children, and their points descend one level. The width
 * test uses the host formula of dpm_lon (util.h:59-66); decisions within a
 * 16-ulp band of reach_distance are re-taken on the host with the C
 * library's sin.
 *
 * Update after each step (cuda_update_schema_collect, _split, _merge,
 * src/cuda/geometry.cu:1114-1234): a leaf holding more than grid_capacity
 * points and higher than 2 * reach_distance in metres splits into the
 * halving boxes lo + (i % 2) * (mid - lo), a branch whose four children are
 * leaves holding fewer than grid_capacity points together merges, each after
 * schema_update_delay consecutive steps. Flags and ranks come from prefix
 * sums, so the slots a split takes (from the free stack, then past the
 * high-water mark) and the order merged slots return to the stack are
 * deterministic; the reference takes them in atomic order, which changes slot
 * numbers but not the tree.
 */
#include "stc_schema.h"
#include "stc_geo.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#define SETTLED 0x80000000u
#define SB_MAX_DEPTH 60
#define SB_AMB_CAP 4096u

enum { DC_SPLITS = 0, DC_MERGES = 1, DC_AMB = 2, DC_LEAVES = 3, DC_WORDS = 8 };

/* degree_per_meter_longitude of the host (util.h:59-66), here with the
 * correctly rounded sine of stc_geo.h */
__device__ inline double dpm_lon_host_form(double lat) {
    double a = fabs(lat);
    if (a == 90) a = 89;
    return 360.0 / (stc_sin_dd((90 - a) * STC_PI_REF / 180) * 40076.0 * 1000.0);
}
static double host_dpm_lon_host_form(double lat) {
    double a = std::fabs(lat);
    if (a == 90) a = 89;
    return 360.0 / (std::sin((90 - a) * STC_PI_REF / 180) * 40076.0 * 1000.0);
}

/* ------------------------------------------------------------------ */
/* initial build                                                       */
/* ------------------------------------------------------------------ */

__global__ void k_sb_root(Node *nodes, uint32_t *nstate, MapBox m, uint32_t *cnt) {
    if (threadIdx.x || blockIdx.x) return;
    Node nd;
    nd.lo0 = m.lo0; nd.lo1 = m.lo1; nd.hi0 = m.hi0; nd.hi1 = m.hi1;
    nd.mid_x = (nd.lo0 + nd.hi0) / 2;
    nd.mid_y = (nd.lo1 + nd.hi1) / 2;
    for (int c = 0; c < 4; c++) nd.child[c] = 0;
    nd.type = T_LEAF;
    nd.leaf = 0;
    nd.code = 0;
    nd.level = 0;
    nodes[0] = nd;
    nstate[0] = 0;
    cnt[0] = 0;
}

/* every unsettled point adds itself to its node; points taken in Morton
 * order (`order`, or id order when null) share nodes along a wave, and each
 * run of lanes on one node adds its length once */
__global__ void k_sb_count(const uint32_t *__restrict__ pnode, const uint32_t *__restrict__ order, uint32_t n,
                           uint32_t *__restrict__ cnt) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & (STC_WAVE - 1);
    const uint32_t pn = j < n ? pnode[order ? order[j] : j] : SETTLED;
    const uint32_t prev = __shfl_up(pn, 1, STC_WAVE);
    const bool bnd = lane == 0 || prev != pn;
    const uint64_t bmask = __ballot(bnd);
    if (bnd && !(pn & SETTLED)) {
        const uint64_t later = lane == STC_WAVE - 1 ? 0ull : (bmask >> (lane + 1));
        const uint32_t end = later ? (uint32_t)(lane + __ffsll((long long)later)) : (uint32_t)STC_WAVE;
        atomicAdd(&cnt[pn], end - (uint32_t)lane);
    }
}

/* split decisions of the level's nodes [fb, fe) (QTNode::split, QTree.h:158-163) */
__global__ void k_sb_decide(const Node *__restrict__ nodes, const uint32_t *__restrict__ cnt, uint32_t fb,
                            uint32_t fe, uint32_t cap, double r, double band, uint32_t *__restrict__ flag,
                            uint32_t *__restrict__ amb, uint32_t *__restrict__ dcount) {
    const uint32_t i = fb + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fe) return;
    const Node nd = nodes[i];
    const double half = (nd.hi0 - nd.lo0) / dpm_lon_host_form(nd.lo1) / 2;
    const bool full = cnt[i] >= cap;
    flag[i - fb] = (full && half > r) ? 1u : 0u;
    if (full && fabs(half - r) <= band) {
        const uint32_t k = atomicAdd(&dcount[DC_AMB], 1u);
        if (k < SB_AMB_CAP) amb[k] = i;
    }
}

__global__ void k_sb_split(Node *__restrict__ nodes, uint32_t *__restrict__ nstate, uint32_t *__restrict__ cnt,
                           uint32_t fb, uint32_t fe, const uint32_t *__restrict__ flag,
                           const uint32_t *__restrict__ rank, uint32_t base) {
    const uint32_t i = fb + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fe || !flag[i - fb]) return;
    Node p = nodes[i];
    const uint32_t k0 = base + 4 * rank[i - fb];
    /* QTree.h:164-167: children bounded by the parent's mids */
    const double lo0[4] = {p.lo0, p.mid_x, p.lo0, p.mid_x}, lo1[4] = {p.lo1, p.lo1, p.mid_y, p.mid_y};
    const double hi0[4] = {p.mid_x, p.hi0, p.mid_x, p.hi0}, hi1[4] = {p.mid_y, p.mid_y, p.hi1, p.hi1};
    for (int c = 0; c < 4; c++) {
        Node k;
        k.lo0 = lo0[c]; k.lo1 = lo1[c]; k.hi0 = hi0[c]; k.hi1 = hi1[c];
        k.mid_x = (k.lo0 + k.hi0) / 2;
        k.mid_y = (k.lo1 + k.hi1) / 2;
        for (int j = 0; j < 4; j++) k.child[j] = 0;
        k.type = T_LEAF;
        k.leaf = 0;
        k.code = 0;
        k.level = p.level + 1;
        nodes[k0 + c] = k;
        nstate[k0 + c] = 0;
        cnt[k0 + c] = 0;
        p.child[c] = k0 + c;
    }
    p.type = T_BRANCH;
    nodes[i] = p;
}

__global__ void k_sb_descend(const double2 *__restrict__ xy, uint32_t *__restrict__ pnode, uint32_t n,
                             const Node *__restrict__ nodes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t pn = pnode[i];
    if (pn & SETTLED) return;
    const Node &nd = nodes[pn];
    if (nd.type == T_BRANCH) {
        const double2 q = xy[i];
        pnode[i] = nd.child[(q.y > nd.mid_y) * 2 + (q.x > nd.mid_x)];
    } else {
        pnode[i] = pn | SETTLED;
    }
}

/* ------------------------------------------------------------------ */
/* leaf numbering                                                      */
/* ------------------------------------------------------------------ */

__global__ void k_leaf_flag(const Node *__restrict__ nodes, uint32_t hwm, uint32_t *__restrict__ flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < hwm) flag[i] = nodes[i].type == T_LEAF;
}

__global__ void k_leaf_assign(Node *__restrict__ nodes, uint32_t hwm, const uint32_t *__restrict__ flag,
                              const uint32_t *__restrict__ rank, uint32_t *__restrict__ leaf_node) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hwm || !flag[i]) return;
    nodes[i].leaf = rank[i];
    leaf_node[rank[i]] = i;
}

/* ------------------------------------------------------------------ */
/* update after a step                                                 */
/* ------------------------------------------------------------------ */

/* cuda_update_schema_collect (geometry.cu:1193-1234) */
__global__ void k_schema_collect(const Node *__restrict__ nodes, uint32_t hwm, uint32_t *__restrict__ nstate,
                                 const uint32_t *__restrict__ leaf_count, uint32_t cap, double r2, uint32_t delay,
                                 uint32_t *__restrict__ fs, uint32_t *__restrict__ fm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hwm) return;
    const Node nd = nodes[i];
    const uint32_t st = nstate[i];
    uint32_t ovf = st & 0xFFFFu, unf = st >> 16, s = 0, m = 0;
    if (nd.type == T_LEAF) {
        if ((nd.hi1 - nd.lo1) / STC_DPM_LAT > r2 && leaf_count[nd.leaf] > cap) {
            ovf = ovf < 0xFFFFu ? ovf + 1 : ovf;
            if (ovf >= delay) { s = 1; ovf = 0; }
        } else {
            ovf = 0;
        }
    } else if (nd.type == T_BRANCH) {
        uint32_t leafchild = 0, sum = 0;
        for (int c = 0; c < 4; c++) {
            const Node &ch = nodes[nd.child[c]];
            if (ch.type == T_LEAF) {
                leafchild++;
                sum += leaf_count[ch.leaf];
            }
        }
        if (leafchild == 4 && sum < cap) {
            unf = unf < 0xFFFFu ? unf + 1 : unf;
            if (unf >= delay) { m = 1; unf = 0; }
        } else {
            unf = 0;
        }
    }
    nstate[i] = ovf | (unf << 16);
    fs[i] = s;
    fm[i] = m;
}

/* cuda_update_schema_split (geometry.cu:1114-1158): slot k of the split of
 * rank j (k = 4j + c) is the k-th from the top of the free stack, then past
 * the high-water mark */
__global__ void k_schema_split(Node *__restrict__ nodes, uint32_t *__restrict__ nstate, uint32_t hwm,
                               const uint32_t *__restrict__ fs, const uint32_t *__restrict__ rs,
                               const uint32_t *__restrict__ fstack, uint32_t n_free) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hwm || !fs[i]) return;
    Node p = nodes[i];
    const double xhalf = p.mid_x - p.lo0, yhalf = p.mid_y - p.lo1;
    for (int c = 0; c < 4; c++) {
        const uint32_t k = 4 * rs[i] + c;
        const uint32_t slot = k < n_free ? fstack[n_free - 1 - k] : hwm + (k - n_free);
        Node ch;
        ch.lo0 = p.lo0 + (double)(c % 2 == 1) * xhalf;
        ch.lo1 = p.lo1 + (double)(c / 2 == 1) * yhalf;
        ch.hi0 = p.mid_x + (double)(c % 2 == 1) * xhalf;
        ch.hi1 = p.mid_y + (double)(c / 2 == 1) * yhalf;
        ch.mid_x = (ch.lo0 + ch.hi0) / 2;
        ch.mid_y = (ch.lo1 + ch.hi1) / 2;
        for (int j = 0; j < 4; j++) ch.child[j] = 0;
        ch.type = T_LEAF;
        ch.leaf = 0;
        ch.code = 0;
        ch.level = p.level + 1;
        nodes[slot] = ch;
        nstate[slot] = 0;
        p.child[c] = slot;
    }
    p.type = T_BRANCH;
    nodes[i] = p;
}

/* cuda_update_schema_merge (geometry.cu:1160-1191) */
__global__ void k_schema_merge(Node *__restrict__ nodes, uint32_t hwm, const uint32_t *__restrict__ fm,
                               const uint32_t *__restrict__ rm, uint32_t *__restrict__ fstack, uint32_t n_free) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hwm || !fm[i]) return;
    for (int c = 0; c < 4; c++) {
        const uint32_t ch = nodes[i].child[c];
        nodes[ch].type = 0;
        fstack[n_free + 4 * rm[i] + c] = ch;
    }
    nodes[i].type = T_LEAF;
}

/* ------------------------------------------------------------------ */
/* host side                                                           */
/* ------------------------------------------------------------------ */

static inline uint32_t nblk(uint64_t n, uint32_t t) { return (uint32_t)std::max<uint64_t>((n + t - 1) / t, 1); }

template <typename T>
static int grow_copy(T *&p, uint32_t old_n, uint32_t new_n, bool copy, hipStream_t s) {
    T *q = nullptr;
    STC_CHECK(hipMalloc((void **)&q, (size_t)new_n * sizeof(T)));
    if (copy && p && old_n) STC_CHECK(hipMemcpyAsync(q, p, (size_t)old_n * sizeof(T), hipMemcpyDeviceToDevice, s));
    STC_CHECK(hipStreamSynchronize(s));
    if (p) STC_CHECK(hipFree(p));
    p = q;
    return 0;
}

static int reserve_slots(SchemaState *S, uint32_t need, hipStream_t s) {
    if (need <= S->cap) return 0;
    uint32_t nc = std::max<uint32_t>(need + need / 2, 1024);
    /* flags and ranks too: an update grows the slots between its scans and its splits */
    if (grow_copy(S->nodes, S->cap, nc, true, s) || grow_copy(S->nstate, S->cap, nc, true, s) ||
        grow_copy(S->fstack, S->cap, nc, true, s) || grow_copy(S->fs, S->cap, nc, true, s) ||
        grow_copy(S->fm, S->cap, nc, true, s) || grow_copy(S->rs, S->cap, nc, true, s) ||
        grow_copy(S->rm, S->cap, nc, true, s) || grow_copy(S->cnt, S->cap, nc, true, s))
        return STC_E_HIP;
    S->cap = nc;
    return 0;
}

int stc_schema_init(SchemaState *S, uint32_t n_points, uint32_t slots) {
    S->n_points = n_points;
    STC_CHECK(hipMalloc((void **)&S->pnode, (size_t)std::max<uint32_t>(n_points, 1) * 4));
    STC_CHECK(hipMalloc((void **)&S->dcount, DC_WORDS * 4));
    STC_CHECK(hipMalloc((void **)&S->amb, SB_AMB_CAP * 4));
    STC_CHECK(hipHostMalloc((void **)&S->hcount, DC_WORDS * 4, hipHostMallocDefault));
    S->cap = 0;
    return reserve_slots(S, slots, nullptr);
}

void stc_schema_free(SchemaState *S) {
    void *d[] = {S->nodes, S->nstate, S->fstack, S->fs, S->fm, S->rs, S->rm, S->pnode, S->cnt, S->dcount, S->amb};
    for (void *p : d)
        if (p) (void)hipFree(p);
    if (S->hcount) (void)hipHostFree(S->hcount);
    *S = SchemaState();
}

static int read_counts(SchemaState *S, hipStream_t s) {
    STC_CHECK(hipMemcpyAsync(S->hcount, S->dcount, DC_WORDS * 4, hipMemcpyDeviceToHost, s));
    STC_CHECK(hipStreamSynchronize(s));
    return 0;
}

int stc_schema_build(SchemaState *S, const double2 *xy, const uint32_t *order, uint32_t n, MapBox m, uint32_t cap,
                     double r, uint32_t *temp, size_t temp_words, hipStream_t s) {
    if (n != S->n_points) return STC_E_ARG;
    const double band = 16.0 * (std::nextafter(r, INFINITY) - r);
    hipLaunchKernelGGL(k_sb_root, dim3(1), dim3(64), 0, s, S->nodes, S->nstate, m, S->cnt);
    STC_CHECK(hipMemsetAsync(S->pnode, 0, (size_t)n * 4, s));
    STC_CHECK(hipGetLastError());
    uint32_t hwm = 1, fb = 0, fe = 1, splits = 0;
    for (int depth = 0; fb < fe; depth++) {
        if (depth > SB_MAX_DEPTH) return STC_E_CAPACITY;
        const uint32_t nf = fe - fb;
        STC_CHECK(hipMemsetAsync(S->dcount, 0, DC_WORDS * 4, s));
        hipLaunchKernelGGL(k_sb_count, dim3(nblk(n, 256)), dim3(256), 0, s, (const uint32_t *)S->pnode, order, n, S->cnt);
        hipLaunchKernelGGL(k_sb_decide, dim3(nblk(nf, 256)), dim3(256), 0, s, (const Node *)S->nodes,
                           (const uint32_t *)S->cnt, fb, fe, cap, r, band, S->fs, S->amb, S->dcount);
        STC_CHECK(hipGetLastError());
        { int rc_ = read_counts(S, s); if (rc_) return rc_; }
        const uint32_t n_amb = S->hcount[DC_AMB];
        if (n_amb) {
            /* the band: the width test with the C library's sin, as the host computes it */
            if (n_amb > SB_AMB_CAP) return STC_E_CAPACITY;
            std::vector<uint32_t> ids(n_amb);
            STC_CHECK(hipMemcpy(ids.data(), S->amb, n_amb * 4, hipMemcpyDeviceToHost));
            for (uint32_t id : ids) {
                Node nd;
                uint32_t f;
                STC_CHECK(hipMemcpy(&nd, S->nodes + id, sizeof(Node), hipMemcpyDeviceToHost));
                f = ((nd.hi0 - nd.lo0) / host_dpm_lon_host_form(nd.lo1) / 2 > r) ? 1u : 0u;
                STC_CHECK(hipMemcpy(S->fs + (id - fb), &f, 4, hipMemcpyHostToDevice));
            }
        }
        if (temp_words < stc_radix_temp_words(nf)) return STC_E_CAPACITY;
        int rc = stc_exclusive_scan_u32(S->fs, S->rs, nf, S->dcount + DC_SPLITS, temp, s);
        if (rc) return rc;
        { int rc_ = read_counts(S, s); if (rc_) return rc_; }
        const uint32_t ns = S->hcount[DC_SPLITS];
        if (!ns) break;
        if ((uint64_t)hwm + 4ull * ns > 0x7FFFFFFFull) return STC_E_CAPACITY;
        if ((rc = reserve_slots(S, hwm + 4 * ns, s))) return rc;
        hipLaunchKernelGGL(k_sb_split, dim3(nblk(nf, 256)), dim3(256), 0, s, S->nodes, S->nstate, S->cnt, fb, fe,
                           (const uint32_t *)S->fs, (const uint32_t *)S->rs, hwm);
        hipLaunchKernelGGL(k_sb_descend, dim3(nblk(n, 256)), dim3(256), 0, s, xy, S->pnode, n,
                           (const Node *)S->nodes);
        STC_CHECK(hipGetLastError());
        fb = hwm;
        fe = hwm + 4 * ns;
        hwm = fe;
        splits += ns;
    }
    if (!splits) return STC_E_ARG; /* the root does not split: too few objects for grid_capacity */
    S->hwm = hwm;
    S->n_free = 0;
    S->n_valid = hwm;
    S->n_leaves = 1 + 3 * splits;
    S->present = true;
    S->dirty = true;
    return 0;
}

int stc_schema_set(SchemaState *S, const stc_qtnode *q, uint32_t n, hipStream_t s) {
    if (!q || n == 0 || q[0].type != (int16_t)T_BRANCH) return STC_E_ARG;
    /* a tree below node 0: every valid node reached once, children of branches valid */
    std::vector<uint8_t> seen(n, 0);
    std::vector<uint32_t> stack{0};
    uint32_t reached = 0, leaves = 0;
    while (!stack.empty()) {
        const uint32_t v = stack.back();
        stack.pop_back();
        if (v >= n || seen[v] || (q[v].type != (int16_t)T_BRANCH && q[v].type != (int16_t)T_LEAF)) return STC_E_ARG;
        seen[v] = 1;
        reached++;
        if (q[v].type == (int16_t)T_LEAF) {
            leaves++;
            continue;
        }
        for (int c = 0; c < 4; c++) stack.push_back(q[v].children[c]);
    }
    uint32_t valid = 0;
    for (uint32_t i = 0; i < n; i++) valid += q[i].type == (int16_t)T_BRANCH || q[i].type == (int16_t)T_LEAF;
    if (valid != reached) return STC_E_ARG; /* a valid node outside the tree */
    int rc = reserve_slots(S, n, s);
    if (rc) return rc;
    std::vector<Node> nodes(n);
    std::vector<uint32_t> st(n), fstack;
    for (uint32_t i = 0; i < n; i++) {
        Node &d = nodes[i];
        d.lo0 = q[i].mbr[0]; d.lo1 = q[i].mbr[1]; d.hi0 = q[i].mbr[2]; d.hi1 = q[i].mbr[3];
        d.mid_x = q[i].mid_x;
        d.mid_y = q[i].mid_y;
        const bool valid_i = seen[i] != 0;
        for (int c = 0; c < 4; c++) d.child[c] = (valid_i && q[i].type == (int16_t)T_BRANCH) ? q[i].children[c] : 0;
        d.type = valid_i ? (uint32_t)q[i].type : 0u;
        d.leaf = 0;
        d.code = 0;
        d.level = (uint32_t)std::max<int16_t>(q[i].level, 0);
        st[i] = valid_i ? ((uint32_t)(uint16_t)q[i].overflow_count | ((uint32_t)(uint16_t)q[i].underflow_count << 16)) : 0;
        if (!valid_i) fstack.push_back(i);
    }
    STC_CHECK(hipMemcpyAsync(S->nodes, nodes.data(), (size_t)n * sizeof(Node), hipMemcpyHostToDevice, s));
    STC_CHECK(hipMemcpyAsync(S->nstate, st.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
    if (!fstack.empty())
        STC_CHECK(hipMemcpyAsync(S->fstack, fstack.data(), fstack.size() * 4, hipMemcpyHostToDevice, s));
    STC_CHECK(hipStreamSynchronize(s));
    S->hwm = n;
    S->n_free = (uint32_t)fstack.size();
    S->n_valid = reached;
    S->n_leaves = leaves;
    S->present = true;
    S->dirty = true;
    return 0;
}

int stc_schema_get(SchemaState *S, stc_qtnode *q, uint64_t cap, uint64_t *n, hipStream_t s) {
    if (!S->present) return STC_E_STATE;
    if (n) *n = S->hwm;
    if (!q || !cap) return 0;
    const uint32_t k = (uint32_t)std::min<uint64_t>(cap, S->hwm);
    std::vector<Node> nodes(k);
    std::vector<uint32_t> st(k);
    STC_CHECK(hipMemcpyAsync(nodes.data(), S->nodes, (size_t)k * sizeof(Node), hipMemcpyDeviceToHost, s));
    STC_CHECK(hipMemcpyAsync(st.data(), S->nstate, (size_t)k * 4, hipMemcpyDeviceToHost, s));
    STC_CHECK(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < k; i++) {
        stc_qtnode &o = q[i];
        const Node &d = nodes[i];
        memset(&o, 0, sizeof(o));
        o.type = (int16_t)d.type;
        o.level = (int16_t)d.level;
        o.grid_id = d.type == T_LEAF ? d.leaf : 0;
        o.overflow_count = (int16_t)(st[i] & 0xFFFFu);
        o.underflow_count = (int16_t)(st[i] >> 16);
        o.mid_x = d.mid_x;
        o.mid_y = d.mid_y;
        o.mbr[0] = d.lo0; o.mbr[1] = d.lo1; o.mbr[2] = d.hi0; o.mbr[3] = d.hi1;
        if (d.type == T_BRANCH)
            for (int c = 0; c < 4; c++) o.children[c] = d.child[c];
    }
    return 0;
}

int stc_schema_number_leaves(SchemaState *S, uint32_t *leaf_node, uint32_t *temp, size_t temp_words,
                             hipStream_t s) {
    if (temp_words < stc_radix_temp_words(S->hwm)) return STC_E_CAPACITY;
    hipLaunchKernelGGL(k_leaf_flag, dim3(nblk(S->hwm, 256)), dim3(256), 0, s, (const Node *)S->nodes, S->hwm, S->fs);
    STC_CHECK(hipGetLastError());
    int rc = stc_exclusive_scan_u32(S->fs, S->rs, S->hwm, nullptr, temp, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_leaf_assign, dim3(nblk(S->hwm, 256)), dim3(256), 0, s, S->nodes, S->hwm,
                       (const uint32_t *)S->fs, (const uint32_t *)S->rs, leaf_node);
    STC_CHECK(hipGetLastError());
    S->dirty = false;
    return 0;
}

int stc_schema_update(SchemaState *S, const uint32_t *leaf_count, uint32_t cap, double r, uint32_t delay,
                      uint32_t *temp, size_t temp_words, hipStream_t s) {
    const uint32_t hwm = S->hwm;
    if (temp_words < stc_radix_temp_words(hwm)) return STC_E_CAPACITY;
    STC_CHECK(hipMemsetAsync(S->dcount, 0, DC_WORDS * 4, s));
    hipLaunchKernelGGL(k_schema_collect, dim3(nblk(hwm, 256)), dim3(256), 0, s, (const Node *)S->nodes, hwm,
                       S->nstate, leaf_count, cap, 2 * r, delay ? delay : 1u, S->fs, S->fm);
    STC_CHECK(hipGetLastError());
    int rc = stc_exclusive_scan_u32(S->fs, S->rs, hwm, S->dcount + DC_SPLITS, temp, s);
    if (!rc) rc = stc_exclusive_scan_u32(S->fm, S->rm, hwm, S->dcount + DC_MERGES, temp, s);
    if (rc) return rc;
    { int rc_ = read_counts(S, s); if (rc_) return rc_; }
    const uint32_t ns = S->hcount[DC_SPLITS], nm = S->hcount[DC_MERGES];
    if (ns) {
        const uint32_t need = 4 * ns;
        const uint32_t new_hwm = hwm + (need > S->n_free ? need - S->n_free : 0);
        if ((rc = reserve_slots(S, new_hwm, s))) return rc;
        hipLaunchKernelGGL(k_schema_split, dim3(nblk(hwm, 256)), dim3(256), 0, s, S->nodes, S->nstate, hwm,
                           (const uint32_t *)S->fs, (const uint32_t *)S->rs, (const uint32_t *)S->fstack, S->n_free);
        STC_CHECK(hipGetLastError());
        S->n_free = need > S->n_free ? 0 : S->n_free - need;
        S->hwm = new_hwm;
    }
    if (nm) {
        hipLaunchKernelGGL(k_schema_merge, dim3(nblk(hwm, 256)), dim3(256), 0, s, S->nodes, hwm,
                           (const uint32_t *)S->fm, (const uint32_t *)S->rm, S->fstack, S->n_free);
        STC_CHECK(hipGetLastError());
        S->n_free += 4 * nm;
    }
    S->n_valid = S->n_valid + 4 * ns - 4 * nm;
    S->n_leaves = S->n_leaves + 3 * ns - 3 * nm;
    if (ns || nm) S->dirty = true;
    return 0;
}
