// A5/A6: per-edge congestion state in HBM -- usage counts, present, present
// EMA, the fused per-iteration cost update (base/present/history/jitter +
// via pooling + via hard block), history, and nonzero extraction for the
// host's exact reductions.
//
// Every kernel is a streaming pass (HBM bound).  Float arithmetic follows the
// reference's numpy float32 expressions operation by operation (the library is
// built with FP contraction off), so results are bit-identical to
// EdgeAccountant (UPF:799-1043) and the CPU via code (UPF:3291-3387,
// 3483-3546).
#include <string.h>

#include <algorithm>

#include "common.h"
#include "engine.h"

static int ensure_sparse(Engine &en, int64_t cap)
{
    if (en.sp_cap >= cap) return ORGPU_OK;
    if (en.sp_idx) { (void)hipFree(en.sp_idx); (void)hipFree(en.sp_val); }
    en.sp_idx = nullptr; en.sp_val = nullptr; en.sp_cap = 0;
    HIP_CHECK_RET(hipMalloc((void **)&en.sp_idx, cap * sizeof(int32_t)));
    HIP_CHECK_RET(hipMalloc((void **)&en.sp_val, cap * sizeof(float)));
    if (!en.sp_count) HIP_CHECK_RET(hipMalloc((void **)&en.sp_count, sizeof(unsigned long long)));
    if (!en.h_sp_count) HIP_CHECK_RET(hipHostMalloc((void **)&en.h_sp_count, sizeof(unsigned long long), hipHostMallocDefault));
    en.sp_cap = cap;
    return ORGPU_OK;
}

static int ensure_upload(Engine &en, int64_t n)
{
    if (en.up_cap >= n) return ORGPU_OK;
    if (en.up_buf) (void)hipFree(en.up_buf);
    en.up_buf = nullptr; en.up_cap = 0;
    int64_t cap = std::max<int64_t>(n, 1 << 16);
    HIP_CHECK_RET(hipMalloc((void **)&en.up_buf, cap * sizeof(int32_t)));
    en.up_cap = cap;
    return ORGPU_OK;
}

// wave-aggregated append of (idx, val) into the sparse staging buffers
__device__ __forceinline__ void sparse_push(bool want, int32_t idx, float val, int32_t *oidx, float *oval,
                                            unsigned long long *count, int64_t cap)
{
    const uint64_t m = __ballot(want);
    if (!m) return;
    const uint32_t lane = __lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    unsigned long long base = 0;
    if ((int)lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(m));
    base = __shfl(base, leader);
    if (want) {
        const unsigned long long pos = base + (unsigned long long)__popcll(m & ((1ull << lane) - 1ull));
        if ((int64_t)pos < cap) { oidx[pos] = idx; oval[pos] = val; }
    }
}

__global__ void k_scatter(int32_t *count, const int32_t *ids, int64_t n, int32_t delta, int64_t E)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = ids[i];
        if (e >= 0 && e < E) atomicAdd(&count[e], delta);
    }
}

__global__ void k_present(int64_t E, const int32_t *count, float *present)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x)
        present[e] = (float)count[e];
}

__global__ void k_penalty(int64_t E, float *present, const int32_t *ids, int64_t n, float pen)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = ids[i];
        if (e >= 0 && e < E) present[e] = present[e] + pen;
    }
}

// present_ema update; nonzeros of max(0, ema - 1)
__global__ void __launch_bounds__(256)
k_ema(int64_t E, const int32_t *count, float *ema, int first, float beta, float omb, int32_t *oidx, float *oval,
      unsigned long long *ocount, int64_t cap)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e0 = (int64_t)blockIdx.x * blockDim.x; e0 < E; e0 += stride) {
        const int64_t e = e0 + threadIdx.x;
        bool want = false;
        float over = 0.0f;
        if (e < E) {
            const float p = (float)count[e];
            const float old = first ? p : ema[e];
            const float a = beta * p;
            const float b = omb * old;
            const float v = a + b;
            ema[e] = v;
            over = fmaxf(0.0f, v - 1.0f);
            want = over > 0.0f;
        }
        sparse_push(want, (int32_t)e, over, oidx, oval, ocount, cap);
    }
}

struct CostArgs {
    orgpu_cost_params p;
    int64_t N;
    const int32_t *indptr, *indices;
    const uint8_t *kind;
    const float *base, *ema, *hist;
    float *total;
};

// One thread per node: the row's edges share the source (x, y, z), which
// gives edge_layer (= z) and the via column without extra loads.
__global__ void __launch_bounds__(256)
k_cost(CostArgs a)
{
    const orgpu_cost_params &p = a.p;
    const int64_t plane = (int64_t)p.X * p.Y;
    const int Lm2 = p.L - 2;
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < a.N; u += (int64_t)gridDim.x * blockDim.x) {
        const int z = (int)(u / plane);
        const int64_t r = u - (int64_t)z * plane;
        const int y = (int)(r / p.X);
        const int x = (int)(r - (int64_t)y * p.X);
        const float lb = p.layer_bias[z];
        const int64_t col = (int64_t)x * p.Y + y;   // numpy [x][y] layout
        const int r0 = a.indptr[u], r1 = a.indptr[u + 1];
        for (int e = r0; e < r1; ++e) {
            const uint8_t k = a.kind[e];
            const float over = fmaxf(0.0f, a.ema[e] - 1.0f);
            const float peb = k == 0 ? lb : 1.0f;
            const float adj = ((a.base[e] * p.via_cost_mult) * p.base_cost_weight) * peb;
            float inv = 1.0f;
            if (k == 0) inv = peb != 0.0f ? 1.0f / fmaxf(peb, 0.5f) : 1.0f;
            const float pt = (p.pres_fac * inv) * over;
            const float hw = p.hist_weight * a.hist[e];
            float tot = (adj + pt) + hw;
            if (p.add_jitter) {
                const float je = fmodf((float)e, 9973.0f) * p.jitter_scale;
                tot = tot + je;
            }
            if (k != 0 && (p.via_pooling || p.hard_block)) {
                const int zv = (int)((int64_t)a.indices[e] / plane);
                int zlo = min(z, zv), zhi = max(z, zv);
                zlo = min(max(zlo, 1), Lm2);
                zhi = min(max(zhi, 1), Lm2);
                if (p.via_pooling) {
                    float pen = 0.0f;
                    if (p.col_pres) pen = pen + p.col_weight * p.col_pres[col];
                    if (p.seg_prefix && p.seg_z > 0) {
                        const int hi = zhi - 2, lo = zlo - 2;
                        const float ph = (hi >= 0 && hi < p.seg_z) ? p.seg_prefix[col * p.seg_z + hi] : 0.0f;
                        const float pl = (lo >= 0 && lo < p.seg_z) ? p.seg_prefix[col * p.seg_z + lo] : 0.0f;
                        const float sp = (ph - pl) * (zhi > zlo ? 1.0f : 0.0f);
                        pen = pen + p.seg_weight * sp;
                    }
                    if (p.via_discount) {
                        const double d = p.via_discount[zlo * p.L + zhi];
                        if (p.discount_f64) pen = (float)((double)pen * d);
                        else pen = pen * (float)d;
                    }
                    if (pen > 0.0f) tot = tot + p.pres_fac * pen;
                }
                if (p.hard_block) {
                    bool blk = p.col_use && p.col_use[col] >= p.col_cap;
                    if (!blk && p.seg_use && p.seg_z > 0) {
                        for (int zz = zlo; zz < zhi; ++zz) {
                            const int s = zz - 1;
                            if (s >= 0 && s < p.seg_z && p.seg_use[col * p.seg_z + s] >= p.seg_cap) { blk = true; break; }
                        }
                    }
                    if (blk) tot = __builtin_inff();
                }
            }
            a.total[e] = tot;
        }
    }
}

__global__ void k_history(int64_t E, const float *present, const float *base, float *hist, float decay, float gain,
                          float cap_mult, int use_cap)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += (int64_t)gridDim.x * blockDim.x) {
        float h = hist[e] * decay;
        const float over = fmaxf(0.0f, present[e] - 1.0f);
        float inc = gain * over;
        if (use_cap) inc = fminf(inc, cap_mult * base[e]);
        hist[e] = h + inc;
    }
}

__global__ void __launch_bounds__(256)
k_over_sparse(int64_t E, const float *present, const float *hist, int mode, float hs, float thr, int32_t *oidx,
              float *oval, unsigned long long *ocount, int64_t cap)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e0 = (int64_t)blockIdx.x * blockDim.x; e0 < E; e0 += stride) {
        const int64_t e = e0 + threadIdx.x;
        bool want = false;
        float over = 0.0f;
        if (e < E) {
            over = fmaxf(0.0f, present[e] - 1.0f);
            if (mode == 0) want = over > 0.0f;
            else want = (over + hs * hist[e]) > thr;
        }
        sparse_push(want, (int32_t)e, over, oidx, oval, ocount, cap);
    }
}

static int grid_for(int64_t n, int64_t per = 256)
{
    int64_t g = (n + per - 1) / per;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

static int read_sparse(Engine &en, orgpu_sparse_out *out)
{
    hipStream_t s = en.stream;
    HIP_CHECK_RET(hipMemcpyAsync(en.h_sp_count, en.sp_count, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK_RET(hipStreamSynchronize(s));
    const int64_t n = (int64_t)*en.h_sp_count;
    out->count = n;
    const int64_t m = std::min<int64_t>(std::min<int64_t>(n, out->capacity), en.sp_cap);
    if (m > 0) {
        HIP_CHECK_RET(hipMemcpyAsync(out->idx, en.sp_idx, m * sizeof(int32_t), hipMemcpyDeviceToHost, s));
        HIP_CHECK_RET(hipMemcpyAsync(out->val, en.sp_val, m * sizeof(float), hipMemcpyDeviceToHost, s));
        HIP_CHECK_RET(hipStreamSynchronize(s));
    }
    if (n > out->capacity || n > en.sp_cap)
        return orgpu_fail(ORGPU_ECAPACITY, "sparse output: %lld nonzeros exceed capacity %lld", (long long)n,
                          (long long)std::min<int64_t>(out->capacity, en.sp_cap));
    return ORGPU_OK;
}

extern "C" int orgpu_usage_scatter(orgpu_t *h, int32_t *count, int64_t n_edges, const int32_t *edge_ids, int64_t n,
                                   int32_t delta)
{
    if (!h || !count || (n > 0 && !edge_ids) || n < 0) return orgpu_fail(ORGPU_EINVAL, "usage_scatter: bad argument");
    if (n == 0) return ORGPU_OK;
    HIP_CHECK_RET(hipSetDevice(h->device));
    int rc = ensure_upload(h->en, n);
    if (rc) return rc;
    hipStream_t s = h->stream;
    HIP_CHECK_RET(hipMemcpyAsync(h->en.up_buf, edge_ids, n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scatter, dim3(grid_for(n)), dim3(256), 0, s, count, h->en.up_buf, n, delta, n_edges);
    HIP_CHECK_RET(hipGetLastError());
    // the pageable host array may be reused as soon as we return
    HIP_CHECK_RET(hipStreamSynchronize(s));
    return ORGPU_OK;
}

extern "C" int orgpu_present_build(orgpu_t *h, int64_t n_edges, const int32_t *count, float *present,
                                   const int32_t *penalty_edges, int64_t n_penalty, float penalty)
{
    if (!h || !count || !present || n_edges < 0 || n_penalty < 0 || (n_penalty > 0 && !penalty_edges))
        return orgpu_fail(ORGPU_EINVAL, "present_build: bad argument");
    HIP_CHECK_RET(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    hipLaunchKernelGGL(k_present, dim3(grid_for(n_edges)), dim3(256), 0, s, n_edges, count, present);
    HIP_CHECK_RET(hipGetLastError());
    if (n_penalty > 0) {
        int rc = ensure_upload(h->en, n_penalty);
        if (rc) return rc;
        HIP_CHECK_RET(hipMemcpyAsync(h->en.up_buf, penalty_edges, n_penalty * sizeof(int32_t), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_penalty, dim3(grid_for(n_penalty)), dim3(256), 0, s, n_edges, present, h->en.up_buf,
                           n_penalty, penalty);
        HIP_CHECK_RET(hipGetLastError());
        HIP_CHECK_RET(hipStreamSynchronize(s));
    }
    return ORGPU_OK;
}

extern "C" int orgpu_present_ema(orgpu_t *h, int64_t n_edges, const int32_t *count, float *present_ema, int first,
                                 float beta, float one_minus_beta, orgpu_sparse_out *over_out)
{
    if (!h || !count || !present_ema || !over_out || n_edges < 0) return orgpu_fail(ORGPU_EINVAL, "present_ema: bad argument");
    HIP_CHECK_RET(hipSetDevice(h->device));
    Engine &en = h->en;
    int rc = ensure_sparse(en, std::max<int64_t>(n_edges, 1));
    if (rc) return rc;
    hipStream_t s = h->stream;
    HIP_CHECK_RET(hipMemsetAsync(en.sp_count, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_ema, dim3(grid_for(n_edges)), dim3(256), 0, s, n_edges, count, present_ema, first, beta,
                       one_minus_beta, en.sp_idx, en.sp_val, en.sp_count, en.sp_cap);
    HIP_CHECK_RET(hipGetLastError());
    return read_sparse(en, over_out);
}

extern "C" int orgpu_cost_update(orgpu_t *h, const orgpu_cost_params *p, int64_t n_nodes, int64_t n_edges,
                                 const int32_t *indptr, const int32_t *indices, const uint8_t *edge_kind,
                                 const float *base_cost, const float *present_ema, const float *history,
                                 float *total_cost)
{
    if (!h || !p || !indptr || !indices || !edge_kind || !base_cost || !present_ema || !history || !total_cost)
        return orgpu_fail(ORGPU_EINVAL, "cost_update: null argument");
    if (!p->layer_bias) return orgpu_fail(ORGPU_EINVAL, "cost_update: layer_bias is required");
    if ((int64_t)p->X * p->Y * p->L != n_nodes || p->L < 1 || p->L > ORGPU_MAX_LAYERS)
        return orgpu_fail(ORGPU_EINVAL, "cost_update: shape %d x %d x %d does not match %lld nodes", p->X, p->Y, p->L,
                          (long long)n_nodes);
    if (p->seg_z < 0 || p->seg_z > p->L) return orgpu_fail(ORGPU_EINVAL, "cost_update: bad seg_z %d", p->seg_z);
    (void)n_edges;
    HIP_CHECK_RET(hipSetDevice(h->device));
    CostArgs a;
    a.p = *p; a.N = n_nodes; a.indptr = indptr; a.indices = indices; a.kind = edge_kind;
    a.base = base_cost; a.ema = present_ema; a.hist = history; a.total = total_cost;
    hipLaunchKernelGGL(k_cost, dim3(grid_for(n_nodes)), dim3(256), 0, h->stream, a);
    HIP_CHECK_RET(hipGetLastError());
    return ORGPU_OK;
}

extern "C" int orgpu_history_update(orgpu_t *h, int64_t n_edges, const float *present, const float *base_cost,
                                    float *history, float decay, float gain, float cap_mult, int use_cap)
{
    if (!h || !present || !base_cost || !history || n_edges < 0) return orgpu_fail(ORGPU_EINVAL, "history_update: bad argument");
    HIP_CHECK_RET(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_history, dim3(grid_for(n_edges)), dim3(256), 0, h->stream, n_edges, present, base_cost, history,
                       decay, gain, cap_mult, use_cap);
    HIP_CHECK_RET(hipGetLastError());
    return ORGPU_OK;
}

extern "C" int orgpu_overuse_sparse(orgpu_t *h, int64_t n_edges, const float *present, const float *history, int mode,
                                    float hist_scale, float threshold, orgpu_sparse_out *out)
{
    if (!h || !present || !out || n_edges < 0 || (mode == 1 && !history) || mode < 0 || mode > 1)
        return orgpu_fail(ORGPU_EINVAL, "overuse_sparse: bad argument");
    HIP_CHECK_RET(hipSetDevice(h->device));
    Engine &en = h->en;
    int rc = ensure_sparse(en, std::max<int64_t>(n_edges, 1));
    if (rc) return rc;
    hipStream_t s = h->stream;
    HIP_CHECK_RET(hipMemsetAsync(en.sp_count, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_over_sparse, dim3(grid_for(n_edges)), dim3(256), 0, s, n_edges, present, history, mode,
                       hist_scale, threshold, en.sp_idx, en.sp_val, en.sp_count, en.sp_cap);
    HIP_CHECK_RET(hipGetLastError());
    return read_sparse(en, out);
}

// ---------------------------------------------------------------------------
// Via pooling pressure (UPF:3626-3637): per via column, and per segment of
// the column, pres = 0.6 * max(0, use - cap) + 0.4 * pres in float32 (each
// product rounded, then the sum: numpy's float32 expression with the Python
// constants taken as float32), then the running sum of the column's segment
// pressures along z in float32 (numpy's cumsum along the last axis).  One
// thread per column.  Also adds sum max(0, use - cap) over columns and
// segments to *over (the via part of the iteration's overuse sum).
// Columns are taken grid-stride, so any column count fits the grid.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_via_pressure(int64_t n_cols, int32_t seg_z, const int16_t *col_use, int32_t col_cap, const int16_t *seg_use,
               int32_t seg_cap, float *col_pres, float *seg_pres, float *seg_prefix, unsigned long long *over)
{
    long long tot = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_cols; c += (int64_t)gridDim.x * blockDim.x) {
        const float k6 = 0.6f, k4 = 0.4f;
        const int co = max(0, (int)col_use[c] - col_cap);
        tot += co;
        const float a = k6 * (float)co;
        const float b = k4 * col_pres[c];
        col_pres[c] = a + b;
        float acc = 0.0f;
        for (int32_t z = 0; z < seg_z; ++z) {
            const int64_t i = c * seg_z + z;
            const int so = max(0, (int)seg_use[i] - seg_cap);
            tot += so;
            const float pa = k6 * (float)so;
            const float pb = k4 * seg_pres[i];
            const float pv = pa + pb;
            seg_pres[i] = pv;
            acc = acc + pv;
            seg_prefix[i] = acc;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)tot, o, 64), hi = __shfl_xor((uint32_t)((unsigned long long)tot >> 32), o, 64);
        tot += (long long)(((unsigned long long)hi << 32) | lo);
    }
    if ((threadIdx.x & 63) == 0 && tot) atomicAdd(over, (unsigned long long)tot);
}

extern "C" int orgpu_via_pressure(orgpu_t *h, int64_t n_cols, int32_t seg_z, const int16_t *col_use, int32_t col_cap,
                                  const int16_t *seg_use, int32_t seg_cap, float *col_pres, float *seg_pres,
                                  float *seg_prefix, unsigned long long *over)
{
    if (!h || n_cols < 0 || seg_z < 0 || seg_z > ORGPU_MAX_LAYERS || !over ||
        (n_cols > 0 && (!col_use || !col_pres)) || (n_cols > 0 && seg_z > 0 && (!seg_use || !seg_pres || !seg_prefix)))
        return orgpu_fail(ORGPU_EINVAL, "via_pressure: bad argument");
    HIP_CHECK_RET(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    HIP_CHECK_RET(hipMemsetAsync(over, 0, sizeof(unsigned long long), s));
    if (n_cols == 0) return ORGPU_OK;
    hipLaunchKernelGGL(k_via_pressure, dim3(grid_for(n_cols)), dim3(256), 0, s, n_cols, seg_z, col_use, col_cap,
                       seg_use, seg_cap, col_pres, seg_pres, seg_prefix, over);
    HIP_CHECK_RET(hipGetLastError());
    return ORGPU_OK;
}

// ---------------------------------------------------------------------------
// Per committed path, the sum of a per-edge quantity over its edge list in
// path order: float64, one addition after another from 0.0, each term the
// float32 value widened exactly -- the value of Python's
// sum(float(v[e]) for e in edges).  One wave per path: its lanes load 64
// edge ids and values at a time (coalesced ids, gathered values), then every
// lane runs the same serial chain of additions over the 64 values taken
// lane by lane, so the order of the additions is the path order.  An edge id
// outside [0, n_edges) is not read; its path's sum becomes NaN.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_segment_sums(int64_t n_seg, const int64_t *offs, const int32_t *edges, int64_t n_edges, const float *dense,
               double *out)
{
    const int64_t seg = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;   // uniform per wave
    const int lane = threadIdx.x & 63;
    if (seg >= n_seg) return;
    const int64_t a = offs[seg], b = offs[seg + 1];
    double acc = 0.0;
    int bad = 0;
    for (int64_t base = a; base < b; base += 64) {
        const int64_t k = base + lane;
        float v = 0.0f;
        if (k < b) {
            const int32_t e = edges[k];
            if (e >= 0 && (int64_t)e < n_edges) v = dense[e];
            else bad = 1;
        }
        const int m = (int)((b - base) < 64 ? (b - base) : 64);
        for (int j = 0; j < m; ++j) acc += (double)__shfl(v, j, 64);
    }
    if (__ballot(bad)) acc = __builtin_nan("");
    if (lane == 0) out[seg] = acc;
}

extern "C" int orgpu_segment_sums(orgpu_t *h, int64_t n_edges, const float *dense, const int32_t *edges,
                                  const int64_t *offs, int64_t n_seg, double *out)
{
    if (!h || n_seg < 0 || n_edges < 0 || (n_seg > 0 && (!dense || !offs || !out)))
        return orgpu_fail(ORGPU_EINVAL, "segment_sums: bad argument");
    if (n_seg == 0) return ORGPU_OK;
    if (n_seg > ((int64_t)1 << 40)) return orgpu_fail(ORGPU_EINVAL, "segment_sums: %lld segments", (long long)n_seg);
    HIP_CHECK_RET(hipSetDevice(h->device));
    const int64_t blocks = (n_seg + 3) / 4;   // four waves (paths) per 256-thread block
    hipLaunchKernelGGL(k_segment_sums, dim3((unsigned)blocks), dim3(256), 0, h->stream, n_seg, offs, edges, n_edges,
                       dense, out);
    HIP_CHECK_RET(hipGetLastError());
    return ORGPU_OK;
}

// ---------------------------------------------------------------------------
// Lower bounds of the iteration's costs for the goal-directed search: the
// cheapest planar edge, and the cheapest via edge per layer of span
// (cost / |z_dst - z_src|).  Non-negative floats order like their bit
// patterns, so the minima are taken with integer atomicMin on the bits.
// Infinite (blocked) and NaN costs are skipped.
// ---------------------------------------------------------------------------
__global__ void k_cost_bounds(int64_t n_edges, const int32_t *indices, const float *cost, const uint8_t *edge_kind,
                              const uint8_t *edge_layer, int32_t plane, uint32_t *out_bits, uint32_t *hist)
{
    __shared__ uint32_t s_hist[64];
    if (threadIdx.x < 64) s_hist[threadIdx.x] = 0;
    __syncthreads();
    float m_xy = __builtin_inff(), m_z = __builtin_inff();
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_edges; e += (int64_t)gridDim.x * blockDim.x) {
        const float c = cost[e];
        if (!(c >= 0.0f) || !(c < __builtin_inff())) continue;
        if (edge_kind[e] == 0) {
            m_xy = fminf(m_xy, c);
            if (hist) {
                // bin = binary exponent + 32, clamped to [0, 63] (0 also takes zeros)
                const int ex = c > 0.0f ? (int)((__float_as_uint(c) >> 23) & 0xFF) - 127 + 32 : 0;
                atomicAdd(&s_hist[ex < 0 ? 0 : (ex > 63 ? 63 : ex)], 1u);
            }
        } else {
            const int span = abs(indices[e] / plane - (int)edge_layer[e]);
            if (span > 0) m_z = fminf(m_z, c / (float)span);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        m_xy = fminf(m_xy, __shfl_xor(m_xy, o, 64));
        m_z = fminf(m_z, __shfl_xor(m_z, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMin(&out_bits[0], __float_as_uint(m_xy));
        atomicMin(&out_bits[1], __float_as_uint(m_z));
    }
    __syncthreads();
    if (hist && threadIdx.x < 64 && s_hist[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_hist[threadIdx.x]);
}

extern "C" int orgpu_cost_bounds(orgpu_t *h, const orgpu_graph *g, const uint8_t *edge_layer, float *out,
                                 uint32_t *hist)
{
    if (!h || !g || !g->indices || !g->cost || !g->edge_kind || !edge_layer || !out || g->n_edges < 0)
        return orgpu_fail(ORGPU_EINVAL, "cost_bounds: bad argument");
    if (g->X <= 0 || g->Y <= 0) return orgpu_fail(ORGPU_EINVAL, "cost_bounds: lattice shape missing");
    HIP_CHECK_RET(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    HIP_CHECK_RET(hipMemsetAsync(out, 0x7F, 2 * sizeof(float), s));   // 0x7F7F7F7F: a large finite float
    if (hist) HIP_CHECK_RET(hipMemsetAsync(hist, 0, 64 * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_cost_bounds, dim3(grid_for(std::max<int64_t>(g->n_edges, 1))), dim3(256), 0, s, g->n_edges,
                       g->indices, g->cost, g->edge_kind, edge_layer, g->X * g->Y, (uint32_t *)out, hist);
    HIP_CHECK_RET(hipGetLastError());
    return ORGPU_OK;
}
