/*
 * stc_trace.hip — seeded synthetic traces on the host and in HBM.
 *
 * The walk itself is stc_gen.h; this file exposes it through the C ABI. A
 * device trace keeps one stc_walker per object in HBM and advances all of
 * them one second per call, writing AoS (x, y) points for the pipeline, so a
 * benchmark never moves points across PCIe.
 */
#include "stc_internal.h"
#include "stc_gen.h"
#include "../../include/stcdb.h"

#include <vector>

struct StcTrace {
    stc_gen_params p;
    stc_walker *w = nullptr;
    double2 *xy = nullptr;
    uint32_t t = 0;
    hipStream_t s = nullptr;
};

__global__ void k_walk(stc_walker *w, stc_gen_params p, uint32_t t, int advance, double2 *xy) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.num_objects) return;
    stc_walker o = w[i];
    if (advance) stc_walk_step(&o, &p, i, t);
    double x, y;
    stc_walk_pos(&o, &p, &x, &y);
    xy[i] = make_double2(x, y);
    if (advance) w[i] = o;
}

extern "C" {

int stc_gen_init_host(const stc_gen_params *p, stc_walker *w, uint32_t first, uint32_t count) {
    if (!p || !w) return STC_E_ARG;
    stc_gen_init(p, w, first, count);
    return STC_OK;
}

int stc_gen_step_host(const stc_gen_params *p, stc_walker *w, uint32_t count, uint32_t t) {
    if (!p || !w || t == 0) return STC_E_ARG;
    for (uint32_t i = 0; i < count; i++) stc_walk_step(&w[i], p, i, t);
    return STC_OK;
}

int stc_gen_pos_host(const stc_gen_params *p, const stc_walker *w, uint32_t count, double *xy) {
    if (!p || !w || !xy) return STC_E_ARG;
    for (uint32_t i = 0; i < count; i++) stc_walk_pos(&w[i], p, &xy[2 * i], &xy[2 * i + 1]);
    return STC_OK;
}

int stc_trace_create(const stc_gen_params *p, void **out) {
    if (!p || !out || p->num_objects == 0) return STC_E_ARG;
    StcTrace *tr = new StcTrace();
    tr->p = *p;
    size_t n = p->num_objects;
    std::vector<stc_walker> hw(n);
    stc_gen_init(p, hw.data(), 0, (uint32_t)n);
    if (hipStreamCreateWithFlags(&tr->s, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void **)&tr->w, n * sizeof(stc_walker)) != hipSuccess ||
        hipMalloc((void **)&tr->xy, n * sizeof(double2)) != hipSuccess ||
        hipMemcpy(tr->w, hw.data(), n * sizeof(stc_walker), hipMemcpyHostToDevice) != hipSuccess) {
        stc_trace_destroy(tr);
        return STC_E_HIP;
    }
    tr->t = 0;
    *out = tr;
    return STC_OK;
}

int stc_trace_next(void *vt, uint32_t t, const double **xy_dev) {
    StcTrace *tr = (StcTrace *)vt;
    if (!tr || !xy_dev) return STC_E_ARG;
    int advance;
    if (t == 0 && tr->t == 0) advance = 0;
    else if (t == tr->t + 1) advance = 1;
    else return STC_E_STATE; /* the walk only moves forward one second at a time */
    uint32_t n = tr->p.num_objects;
    hipLaunchKernelGGL(k_walk, dim3((n + 255) / 256), dim3(256), 0, tr->s, tr->w, tr->p, t, advance, tr->xy);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(tr->s) != hipSuccess) return STC_E_HIP;
    tr->t = t;
    *xy_dev = (const double *)tr->xy;
    return STC_OK;
}

int stc_trace_destroy(void *vt) {
    StcTrace *tr = (StcTrace *)vt;
    if (!tr) return STC_OK;
    if (tr->w) hipFree(tr->w);
    if (tr->xy) hipFree(tr->xy);
    if (tr->s) hipStreamDestroy(tr->s);
    delete tr;
    return STC_OK;
}

} /* extern "C" */
