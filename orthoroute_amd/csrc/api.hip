// C ABI entry points, error reporting and the handle (see include/orgpu.h).
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>

#include "common.h"
#include "engine.h"

static thread_local char g_err[1024] = "";

int orgpu_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int orgpu_fail_hip(hipError_t e, const char *what, const char *file, int line)
{
    snprintf(g_err, sizeof(g_err), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return e == hipErrorOutOfMemory ? ORGPU_ENOMEM : ORGPU_EHIP;
}

hipStream_t orgpu_stream(orgpu_t *h) { return h->stream; }

int orgpu_set_device(orgpu_t *h)
{
    HIP_CHECK_RET(hipSetDevice(h->device));
    return ORGPU_OK;
}

extern "C" const char *orgpu_last_error(void) { return g_err; }

extern "C" int orgpu_create(int device, void *hip_stream, orgpu_t **out)
{
    if (!out) return orgpu_fail(ORGPU_EINVAL, "create: null output pointer");
    *out = nullptr;
    int n = 0;
    HIP_CHECK_RET(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return orgpu_fail(ORGPU_EINVAL, "create: device %d of %d", device, n);
    HIP_CHECK_RET(hipSetDevice(device));
    orgpu_t *h = new (std::nothrow) orgpu_ctx();
    if (!h) return orgpu_fail(ORGPU_ENOMEM, "create: host allocation failed");
    h->device = device;
    h->stream = (hipStream_t)hip_stream;
    h->en.device = device;
    h->en.stream = h->stream;
    *out = h;
    return ORGPU_OK;
}

extern "C" int orgpu_destroy(orgpu_t *h)
{
    if (!h) return ORGPU_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    else hipDeviceSynchronize();
    engine_release(h->en);
    ref_release(h->en);
    solo_release(h->en);
    delete h;
    return ORGPU_OK;
}

extern "C" int orgpu_sync(orgpu_t *h)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "sync: null handle");
    HIP_CHECK_RET(hipSetDevice(h->device));
    HIP_CHECK_RET(hipStreamSynchronize(h->stream));
    return ORGPU_OK;
}

extern "C" int orgpu_reserve(orgpu_t *h, int64_t n_nodes, int32_t slots, int32_t n_buckets,
                             int64_t queue_capacity, int32_t max_path_len)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "reserve: null handle");
    HIP_CHECK_RET(hipSetDevice(h->device));
    HIP_CHECK_RET(hipStreamSynchronize(h->stream));
    return engine_reserve(h->en, n_nodes, slots, n_buckets, queue_capacity, max_path_len);
}

extern "C" int orgpu_route(orgpu_t *h, const orgpu_graph *g, const orgpu_net_req *nets, int32_t n_nets,
                           const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                           float keepout_cost,
                           const orgpu_sssp_params *params, int32_t *path_nodes, int32_t *path_edges,
                           int64_t path_capacity, int64_t *path_offset, float *path_dist,
                           orgpu_sssp_stats *stats)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "route: null handle");
    if (n_nets < 0) return orgpu_fail(ORGPU_EINVAL, "route: negative request count");
    HIP_CHECK_RET(hipSetDevice(h->device));
    if (h->en.cap_slots == 0) return orgpu_fail(ORGPU_EINVAL, "route: call orgpu_reserve first");
    if (n_nets == 0) {
        if (path_offset) path_offset[0] = 0;
        return ORGPU_OK;
    }
    if (!path_nodes) return orgpu_fail(ORGPU_EINVAL, "route: null path output");
    return engine_route(h->en, g, nets, n_nets, roi_bitmaps, n_roi, keepout_owner, keepout_cost, params,
                        path_nodes, path_edges, path_capacity, path_offset, path_dist, stats);
}

extern "C" int orgpu_reserve_solo(orgpu_t *h, int64_t n_nodes, int32_t X, int32_t Y, int32_t L, int32_t slots,
                                  int32_t list_cap)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "reserve_solo: null handle");
    HIP_CHECK_RET(hipSetDevice(h->device));
    HIP_CHECK_RET(hipStreamSynchronize(h->stream));
    return engine_reserve_solo(h->en, n_nodes, X, Y, L, slots, list_cap);
}

extern "C" int orgpu_route_async(orgpu_t *h, const orgpu_graph *g, const orgpu_net_req *reqs, int32_t n,
                                 const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                                 float keepout_cost, const orgpu_sssp_params *params, int32_t *path_nodes,
                                 int32_t *path_edges, int64_t capacity, uint32_t *path_off, int32_t *path_len,
                                 float *path_dist, uint32_t *used)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "route_async: null handle");
    if (!g || !params || n < 0 || (n > 0 && (!reqs || !path_nodes || !path_off || !path_len || !path_dist)))
        return orgpu_fail(ORGPU_EINVAL, "route_async: null argument or negative count");
    if (!g->indptr || !g->indices || !g->cost || !g->edge_kind || g->n_nodes < 1)
        return orgpu_fail(ORGPU_EINVAL, "route_async: bad graph");
    if (!(params->delta > 0.0f)) return orgpu_fail(ORGPU_EINVAL, "route_async: delta must be > 0");
    if (!(params->h_xy >= 0.0f && params->h_z >= 0.0f && params->h_xy < 3.0e38f && params->h_z < 3.0e38f))
        return orgpu_fail(ORGPU_EINVAL, "route_async: h_xy / h_z must be finite and >= 0");
    if ((params->h_xy > 0.0f || params->h_z > 0.0f) &&
        (g->X <= 0 || g->Y <= 0 || g->L <= 0 || (int64_t)g->X * g->Y * g->L != g->n_nodes))
        return orgpu_fail(ORGPU_EINVAL, "route_async: goal-directed search needs the lattice shape");
    if (keepout_owner && !(keepout_cost >= 1.0e6f))
        return orgpu_fail(ORGPU_EINVAL, "route_async: keepout_cost must dominate path costs (>= 1e6)");
    if (params->slots < 1 || params->list_cap < 0)
        return orgpu_fail(ORGPU_EINVAL, "route_async: slots %d / list_cap %d", params->slots, params->list_cap);
    if (params->waves != 0 && params->waves != 1 && params->waves != 2 && params->waves != 4)
        return orgpu_fail(ORGPU_EINVAL, "route_async: waves %d is not 0, 1, 2 or 4", params->waves);
    HIP_CHECK_RET(hipSetDevice(h->device));
    return engine_route_async(h->en, g, reqs, n, roi_bitmaps, n_roi, keepout_owner, params, path_nodes, path_edges,
                              capacity, path_off, path_len, path_dist, used);
}

extern "C" int orgpu_route_ms(orgpu_t *h, const orgpu_graph *g, const orgpu_ms_req *reqs, int32_t n_reqs,
                              const orgpu_ms_point *seeds, int64_t n_seeds, const orgpu_ms_point *targets,
                              int64_t n_targets, const uint32_t *roi_bitmaps, int32_t n_roi,
                              const int32_t *keepout_owner, float keepout_cost, const orgpu_sssp_params *params,
                              int32_t *path_nodes, int32_t *path_edges, int64_t path_capacity, int64_t *path_offset,
                              double *path_total, int32_t *path_target, orgpu_sssp_stats *stats)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "route_ms: null handle");
    if (n_reqs < 0) return orgpu_fail(ORGPU_EINVAL, "route_ms: negative request count");
    HIP_CHECK_RET(hipSetDevice(h->device));
    if (h->en.cap_slots == 0) return orgpu_fail(ORGPU_EINVAL, "route_ms: call orgpu_reserve first");
    if (n_reqs == 0) {
        if (path_offset) path_offset[0] = 0;
        return ORGPU_OK;
    }
    if (!path_nodes) return orgpu_fail(ORGPU_EINVAL, "route_ms: null path output");
    MsCall ms{reqs, seeds, targets, n_seeds, n_targets, path_total, path_target};
    return engine_route(h->en, g, nullptr, n_reqs, roi_bitmaps, n_roi, keepout_owner, keepout_cost, params,
                        path_nodes, path_edges, path_capacity, path_offset, nullptr, stats, &ms);
}