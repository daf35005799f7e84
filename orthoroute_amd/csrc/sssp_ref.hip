// A3 reference-order search: SimpleDijkstra.find_path_roi (UPF:1801-1887)
// with the reference's own arithmetic and pop order, one wavefront per
// search, many searches per launch.
//
// The reference pops a binary heap of (float64 alt, roi index) tuples, skips
// visited nodes, relaxes the popped node's CSR row in order and accepts an
// edge when float32(alt) < dist[v] (dist is float32; NumPy 2 compares a
// Python float with a float32 scalar in float32), storing dist[v] =
// float32(alt) and parent[v] = u; it stops when dst pops.  The pop order is
// that of an exact priority queue on (alt, roi index), so any exact priority
// queue reproduces it.  Here:
//  * a node's roi index is represented by a key that orders like it: 2v + 2
//    for the ascending part of roi_nodes, the request's exceptions for the
//    nodes the reference's list holds out of order;
//  * the queue is a 64-ary heap in global memory: a pop's sift-down loads
//    the 64 children of a level with one coalesced wave load and takes the
//    minimum with wave shuffles; a push loads its whole ancestor chain at
//    once (one lane per level) and shifts the larger ancestors down in one
//    store;
//  * an accepted node is pushed once per popped row: a later parallel edge
//    to the same node is accepted only with a strictly smaller float32
//    value, and the smaller entry pops first, so the reference's extra
//    entries only ever pop as stale ones;
//  * per search, node state {dist f32, parent i32, stamp, visited stamp}
//    (16 B per node) is valid when its stamp equals the search's id, so
//    nothing is cleared between searches.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "common.h"
#include "engine.h"

namespace {

struct RefNode {
    float dist;
    int32_t parent;
    uint32_t stamp;    // search id that set dist / parent
    uint32_t vstamp;   // search id that visited the node
};

struct HeapEnt {
    double alt;
    uint32_t key;      // roi-position key of the node
    int32_t node;
};

struct RefCtl {
    uint32_t next;     // next request to take
    uint32_t pad;
};

struct RefArgs {
    int64_t N;
    const int32_t *indptr, *indices;
    const float *cost;
    const uint8_t *kind;
    const uint32_t *roi;
    int64_t roi_words;
    int32_t n_roi;
    const int32_t *ko_owner;
    float ko_cost;
    const orgpu_ref_req *reqs;
    int32_t n;
    RefNode *state;    // [slots][N]
    HeapEnt *heap;     // [slots][hcap]
    int64_t hcap;
    uint32_t sid_base;
    RefCtl *ctl;
    int32_t *out_nodes, *out_edges;
    int64_t out_cap;
    int64_t *out_off;
    int32_t *out_len;
    float *out_dist;
    unsigned long long *counters;
    unsigned long long *cursor;   // output entries handed out
};

__device__ __forceinline__ bool hlt(double a, uint32_t ka, double b, uint32_t kb)
{
    return a < b || (a == b && ka < kb);
}

// lane holding the smallest (alt, key), lowest lane on equal pairs; the
// same value in every lane
__device__ __forceinline__ int wave_argmin(double a, uint32_t k)
{
    int l = (int)__lane_id();
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double a2 = __shfl_xor(a, o, 64);
        const uint32_t k2 = __shfl_xor(k, o, 64);
        const int l2 = __shfl_xor(l, o, 64);
        if (hlt(a2, k2, a, k) || (a2 == a && k2 == k && l2 < l)) { a = a2; k = k2; l = l2; }
    }
    return l;
}

__device__ __forceinline__ uint32_t node_key(const orgpu_ref_req &q, int32_t v)
{
    uint32_t k = 2u * (uint32_t)v + 2u;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if (i < q.n_rank && q.rank_node[i] == v) k = q.rank_key[i];
    return k;
}

// Stores by one lane become visible to the wave's later loads of other lanes
// once they have completed (the workgroup is the wave, on one CU's L1).
__device__ __forceinline__ void wave_sync_mem() { __threadfence_block(); }

// Removes and returns the heap's root (hn > 0).  All lanes take part.
__device__ HeapEnt heap_pop(HeapEnt *heap, int64_t &hn)
{
    const int lane = (int)__lane_id();
    const HeapEnt top = heap[0];
    const int64_t n = hn - 1;
    if (n > 0) {
        const HeapEnt x = heap[n];
        int64_t pos = 0;
        for (;;) {
            const int64_t c0 = pos * 64 + 1;
            if (c0 >= n) break;
            const int64_t ci = c0 + lane;
            double ca = __builtin_inf();
            uint32_t ck = 0xFFFFFFFFu;
            int32_t cn = -1;
            if (ci < n) {
                const HeapEnt c = heap[ci];
                ca = c.alt; ck = c.key; cn = c.node;
            }
            const int w = wave_argmin(ca, ck);
            const double ba = __shfl(ca, w, 64);
            const uint32_t bk = __shfl(ck, w, 64);
            const int32_t bn = __shfl(cn, w, 64);
            if (!hlt(ba, bk, x.alt, x.key)) break;
            if (lane == 0) heap[pos] = HeapEnt{ba, bk, bn};
            pos = c0 + w;
        }
        if (lane == 0) heap[pos] = x;
    }
    hn = n;
    wave_sync_mem();
    return top;
}

// Inserts (a, k, node); the caller has checked the capacity.  The ancestors
// of the new position lie on a chain whose keys do not increase going up;
// lane l loads ancestor l, the ancestors larger than the new entry (a prefix
// of the lanes) move one level down and the entry takes the highest freed
// place.
__device__ void heap_push(HeapEnt *heap, int64_t &hn, double a, uint32_t k, int32_t node)
{
    const int lane = (int)__lane_id();
    const int64_t i = hn;
    hn = i + 1;
    int64_t my_pos = -1;
    int depth = 0;
    for (int64_t q = i; q > 0; ++depth) {
        q = (q - 1) / 64;
        if (depth == lane) my_pos = q;
    }
    bool larger = false;
    HeapEnt e{0.0, 0u, 0};
    if (lane < depth) {
        e = heap[my_pos];
        larger = hlt(a, k, e.alt, e.key);
    }
    const uint64_t m = __ballot(larger);
    const int j = (int)__builtin_ctzll(~m);          // ancestors 0..j-1 move down
    const int64_t below = __shfl(my_pos, lane > 0 ? lane - 1 : 0, 64);
    if (larger) heap[lane == 0 ? i : below] = e;
    const int64_t dest = j == 0 ? i : __shfl(my_pos, j - 1, 64);
    if (lane == 0) heap[dest] = HeapEnt{a, k, node};
    wave_sync_mem();
}

__device__ __forceinline__ bool roi_has(const RefArgs &a, const uint32_t *roi, int32_t v)
{
    return !roi || ((roi[v >> 5] >> (v & 31)) & 1u);
}

// One wavefront per workgroup; each takes requests from a shared counter.
__global__ void __launch_bounds__(64)
k_sssp_ref(RefArgs a)
{
    const int lane = (int)__lane_id();
    const int slot = blockIdx.x;
    RefNode *st = a.state + (int64_t)slot * a.N;
    HeapEnt *heap = a.heap + (int64_t)slot * a.hcap;
    unsigned long long pops = 0, scanned = 0, written = 0;
    uint32_t errs = 0;
    for (;;) {
        uint32_t r = 0;
        if (lane == 0) r = atomicAdd(&a.ctl->next, 1u);
        r = __shfl(r, 0, 64);
        if (r >= (uint32_t)a.n) break;
        const orgpu_ref_req q = a.reqs[r];
        const uint32_t sid = a.sid_base + r + 1u;
        int32_t len = 0;
        float dd = __builtin_inff();
        int64_t off = 0;
        const bool valid = q.src >= 0 && q.src < a.N && q.dst >= 0 && q.dst < a.N && q.n_rank >= 0 &&
                           q.n_rank <= 8 && q.roi_index < a.n_roi && (q.roi_index < 0 || a.roi) &&
                           (!q.use_keepout || a.ko_owner);
        const uint32_t *roi = valid && q.roi_index >= 0 ? a.roi + (int64_t)q.roi_index * a.roi_words : nullptr;
        if (!valid) {
            len = -4;
        } else if (roi_has(a, roi, q.src) && roi_has(a, roi, q.dst)) {
            int64_t hn = 0;
            if (lane == 0) st[q.src] = RefNode{0.0f, -1, sid, st[q.src].vstamp};
            wave_sync_mem();
            heap_push(heap, hn, 0.0, node_key(q, q.src), q.src);
            bool found = false;
            while (hn > 0) {
                const HeapEnt t = heap_pop(heap, hn);
                const int32_t u = t.node;
                const uint32_t uv = st[u].vstamp;
                if (uv == sid) continue;              // stale entry
                if (lane == 0) st[u].vstamp = sid;
                ++pops;
                if (u == q.dst) { found = true; break; }
                const double du = t.alt;
                const int row = a.indptr[u], row_end = a.indptr[u + 1];
                const bool ko_node = q.use_keepout && a.ko_owner[u] >= 0 && a.ko_owner[u] != q.net_id;
                scanned += (unsigned long long)(row_end - row);
                bool overflow = false;
                for (int k0 = row; k0 < row_end; k0 += 64) {
                    const int e = k0 + lane;
                    const int cnt = min(64, row_end - k0);
                    bool ok = false;
                    int32_t v = -1;
                    double alt = 0.0;
                    float dv = __builtin_inff();
                    uint32_t vvs = 0;
                    if (e < row_end) {
                        v = a.indices[e];
                        const float w = (ko_node && a.kind[e] == 0) ? a.ko_cost : a.cost[e];
                        if (v != u && roi_has(a, roi, v)) {
                            const RefNode s = st[v];
                            if (s.vstamp != sid) {
                                ok = true;
                                vvs = s.vstamp;
                                dv = s.stamp == sid ? s.dist : __builtin_inff();
                                alt = du + (double)w;
                            }
                        }
                    }
                    // the row's edges into the same node, in row order: the
                    // first lane of a node applies them one after another
                    bool first = ok, acc = false;
                    float run = dv;
                    double best = 0.0;
                    for (int j = 0; j < cnt; ++j) {
                        const int32_t vj = __shfl(v, j, 64);
                        const int okj = __shfl((int)ok, j, 64);
                        const double aj = __shfl(alt, j, 64);
                        if (!ok || !okj || vj != v) continue;
                        if (j < lane) {
                            first = false;
                        } else {
                            const float fa = (float)aj;
                            if (fa < run) { run = fa; best = aj; acc = true; }
                        }
                    }
                    acc = acc && first;
                    if (acc) st[v] = RefNode{run, u, sid, vvs};
                    uint64_t m = __ballot(acc);
                    if ((int64_t)__popcll(m) + hn > a.hcap) { overflow = true; break; }
                    while (m) {
                        const int l = (int)__builtin_ctzll(m);
                        m &= m - 1;
                        const double pa = __shfl(best, l, 64);
                        const int32_t pv = __shfl(v, l, 64);
                        heap_push(heap, hn, pa, node_key(q, pv), pv);
                    }
                }
                if (overflow) { len = -2; break; }
            }
            if (found) {
                dd = st[q.dst].dist;
                // parent chain dst -> src into the heap's storage (the heap
                // is no longer needed): one lane, one dependent load per hop
                int32_t *stage = reinterpret_cast<int32_t *>(heap);
                const int64_t stage_cap = a.hcap * 4;
                int32_t n = 0;
                if (lane == 0) {
                    int32_t c = q.dst;
                    while (c != -1) {
                        if (n >= stage_cap) { n = -3; break; }
                        stage[n++] = c;
                        c = st[c].parent;
                    }
                }
                n = __shfl(n, 0, 64);
                if (n < 0) {
                    len = -3;
                } else if (n > 1) {
                    unsigned long long o = 0;
                    if (lane == 0) o = atomicAdd(a.cursor, (unsigned long long)n);
                    o = __shfl(o, 0, 64);
                    if ((int64_t)o + n > a.out_cap) {
                        len = -1;
                    } else {
                        wave_sync_mem();
                        off = (int64_t)o;
                        len = n;
                        written += (unsigned long long)n;
                        // node t of the path is stage[n-1-t]; the edge leaving
                        // it is the first CSR match of the next node
                        for (int t = lane; t < n; t += 64) {
                            const int32_t x = stage[n - 1 - t];
                            a.out_nodes[off + t] = x;
                            int32_t ee = -1;
                            if (t + 1 < n) {
                                const int32_t y = stage[n - 2 - t];
                                for (int e = a.indptr[x]; e < a.indptr[x + 1]; ++e)
                                    if (a.indices[e] == y) { ee = e; break; }
                            }
                            if (a.out_edges) a.out_edges[off + t] = ee;
                        }
                    }
                }
            }
        }
        if (len < 0) errs |= 1u << (-len);
        if (lane == 0) {
            a.out_off[r] = off;
            a.out_len[r] = len;
            if (a.out_dist) a.out_dist[r] = dd;
        }
    }
    if (lane == 0 && a.counters) {
        atomicAdd(&a.counters[0], pops);
        atomicAdd(&a.counters[1], scanned);
        atomicAdd(&a.counters[2], written);
        if (errs) atomicOr(&a.counters[3], (unsigned long long)errs);
    }
}

__global__ void k_ref_begin(RefCtl *ctl, unsigned long long *cursor)
{
    ctl->next = 0u;
    *cursor = 0ull;
}

__global__ void k_ref_clear(uint4 *p, int64_t n16)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

}  // namespace

// Scratch of one reference-order call: the caller's reservation (handle) or
// another fixed region, with the search ids already used in it.
struct RefScratch {
    RefNode *state;
    HeapEnt *heap;
    RefCtl *ctl;
    unsigned long long *cursor;
    int64_t nodes, hcap;
    int32_t slots;
    uint32_t *sid;
};

static int ref_launch(hipStream_t s, RefScratch &sc, const orgpu_graph *g, const orgpu_ref_req *reqs, int32_t n,
                      const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner, float keepout_cost,
                      int32_t slots, int32_t *path_nodes, int32_t *path_edges, int64_t path_capacity,
                      int64_t *path_offset, int32_t *path_len, float *path_dist, uint64_t *counters)
{
    if (!g || !g->indptr || !g->indices || !g->cost || !g->edge_kind)
        return orgpu_fail(ORGPU_EINVAL, "route_ref: null graph array");
    if (n < 0) return orgpu_fail(ORGPU_EINVAL, "route_ref: negative request count");
    if (n == 0) return ORGPU_OK;
    if (!reqs || !path_nodes || !path_offset || !path_len)
        return orgpu_fail(ORGPU_EINVAL, "route_ref: null request or output array");
    if (g->n_nodes < 1 || g->n_nodes > sc.nodes)
        return orgpu_fail(ORGPU_EINVAL, "route_ref: %lld nodes outside the reserved %lld", (long long)g->n_nodes,
                          (long long)sc.nodes);
    if (n_roi > 0 && !roi_bitmaps) return orgpu_fail(ORGPU_EINVAL, "route_ref: null ROI rows");
    if (keepout_owner && !(keepout_cost > 0.0f)) return orgpu_fail(ORGPU_EINVAL, "route_ref: keepout_cost must be > 0");
    if (path_capacity < 0) return orgpu_fail(ORGPU_EINVAL, "route_ref: negative path capacity");
    int32_t k = slots > 0 ? slots : sc.slots;
    if (k > sc.slots) return orgpu_fail(ORGPU_EINVAL, "route_ref: %d slots outside the reserved %d", k, sc.slots);
    k = std::min(k, n);
    // search ids sid+1 .. sid+n; when they would wrap, every stamp is cleared
    if ((uint64_t)*sc.sid + (uint64_t)n + 1ull >= 0xFFFFFFFFull) {
        const int64_t n16 = (int64_t)sc.slots * sc.nodes;   // one uint4 per RefNode
        hipLaunchKernelGGL(k_ref_clear, dim3(2048), dim3(256), 0, s, reinterpret_cast<uint4 *>(sc.state), n16);
        HIP_CHECK_RET(hipGetLastError());
        *sc.sid = 0;
    }
    RefArgs a;
    memset(&a, 0, sizeof(a));
    a.N = g->n_nodes;
    a.indptr = g->indptr; a.indices = g->indices; a.cost = g->cost; a.kind = g->edge_kind;
    a.roi = roi_bitmaps; a.roi_words = (g->n_nodes + 31) / 32; a.n_roi = roi_bitmaps ? n_roi : 0;
    a.ko_owner = keepout_owner; a.ko_cost = keepout_cost;
    a.reqs = reqs; a.n = n;
    a.state = sc.state; a.heap = sc.heap; a.hcap = sc.hcap;
    a.sid_base = *sc.sid;
    a.ctl = sc.ctl; a.cursor = sc.cursor;
    a.out_nodes = path_nodes; a.out_edges = path_edges; a.out_cap = path_capacity;
    a.out_off = path_offset; a.out_len = path_len; a.out_dist = path_dist;
    a.counters = reinterpret_cast<unsigned long long *>(counters);
    hipLaunchKernelGGL(k_ref_begin, dim3(1), dim3(1), 0, s, sc.ctl, sc.cursor);
    hipLaunchKernelGGL(k_sssp_ref, dim3(k), dim3(64), 0, s, a);
    HIP_CHECK_RET(hipGetLastError());
    *sc.sid += (uint32_t)n;
    return ORGPU_OK;
}

void ref_release(Engine &en)
{
    if (en.ref_state) (void)hipFree(en.ref_state);
    if (en.ref_heap) (void)hipFree(en.ref_heap);
    if (en.ref_ctl) (void)hipFree(en.ref_ctl);
    en.ref_state = en.ref_heap = en.ref_ctl = nullptr;
    en.ref_nodes = en.ref_hcap = 0;
    en.ref_slots = 0;
    en.ref_sid = 0;
}

#define REF_MAX_SLOTS 4096

extern "C" int orgpu_reserve_ref(orgpu_t *h, int64_t n_nodes, int32_t slots, int64_t heap_cap)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "reserve_ref: null handle");
    if (n_nodes < 1 || n_nodes >= (1ll << 30)) return orgpu_fail(ORGPU_EINVAL, "reserve_ref: bad node count");
    if (slots < 1 || slots > REF_MAX_SLOTS)
        return orgpu_fail(ORGPU_EINVAL, "reserve_ref: slots %d outside [1, %d]", slots, REF_MAX_SLOTS);
    if (heap_cap < 64 || heap_cap >= (1ll << 31))
        return orgpu_fail(ORGPU_EINVAL, "reserve_ref: heap capacity %lld outside [64, 2^31)", (long long)heap_cap);
    Engine &en = h->en;
    HIP_CHECK_RET(hipSetDevice(h->device));
    if (en.ref_state && en.ref_nodes >= n_nodes && en.ref_slots >= slots && en.ref_hcap >= heap_cap) return ORGPU_OK;
    HIP_CHECK_RET(hipStreamSynchronize(h->stream));
    ref_release(en);
    const size_t sb = (size_t)slots * (size_t)n_nodes * sizeof(RefNode);
    const size_t hb = (size_t)slots * (size_t)heap_cap * sizeof(HeapEnt);
    HIP_CHECK_RET(hipMalloc(&en.ref_state, sb));
    HIP_CHECK_RET(hipMalloc(&en.ref_heap, hb));
    HIP_CHECK_RET(hipMalloc(&en.ref_ctl, 256));
    // stamps start at 0: search ids are >= 1
    HIP_CHECK_RET(hipMemsetAsync(en.ref_state, 0, sb, h->stream));
    HIP_CHECK_RET(hipStreamSynchronize(h->stream));
    en.ref_nodes = n_nodes; en.ref_slots = slots; en.ref_hcap = heap_cap; en.ref_sid = 0;
    return ORGPU_OK;
}

extern "C" int orgpu_route_ref(orgpu_t *h, const orgpu_graph *g, const orgpu_ref_req *reqs, int32_t n,
                               const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                               float keepout_cost, int32_t slots, int32_t *path_nodes, int32_t *path_edges,
                               int64_t path_capacity, int64_t *path_offset, int32_t *path_len, float *path_dist,
                               uint64_t *counters)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "route_ref: null handle");
    HIP_CHECK_RET(hipSetDevice(h->device));
    Engine &en = h->en;
    if (!en.ref_state) return orgpu_fail(ORGPU_EINVAL, "route_ref: call orgpu_reserve_ref first");
    RefScratch sc;
    sc.state = static_cast<RefNode *>(en.ref_state);
    sc.heap = static_cast<HeapEnt *>(en.ref_heap);
    sc.ctl = static_cast<RefCtl *>(en.ref_ctl);
    sc.cursor = reinterpret_cast<unsigned long long *>(static_cast<char *>(en.ref_ctl) + 64);
    sc.nodes = en.ref_nodes; sc.hcap = en.ref_hcap; sc.slots = en.ref_slots;
    sc.sid = &en.ref_sid;
    return ref_launch(h->stream, sc, g, reqs, n, roi_bitmaps, n_roi, keepout_owner, keepout_cost, slots, path_nodes,
                      path_edges, path_capacity, path_offset, path_len, path_dist, counters);
}

// ---------------------------------------------------------------------------
// The same search for callers that cannot allocate: its scratch is the
// file-scope array below (one copy per GPU, held until the process ends).
// Layout: [control 256 B][node state: slots * N * 16 B][heaps: slots *
// heap_cap * 16 B].  In host code `scratch` names a host-side stand-in, so
// its device address is looked up once per device (hipGetSymbolAddress) and
// only that address is used.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[ORGPU_REF_NOALLOC_BYTES];

#define NOALLOC_MAX_DEVICES 64
static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[NOALLOC_MAX_DEVICES];
static uint32_t g_noalloc_sid[NOALLOC_MAX_DEVICES];
static int64_t g_noalloc_state_bytes[NOALLOC_MAX_DEVICES];   // layout the stamps were last cleared for

extern "C" int orgpu_route_ref_noalloc(int device, void *hip_stream, const orgpu_graph *g, const orgpu_ref_req *reqs,
                                       int32_t n, const uint32_t *roi_bitmaps, int32_t n_roi,
                                       const int32_t *keepout_owner, float keepout_cost, int32_t slots,
                                       int64_t heap_cap, int32_t *path_nodes, int32_t *path_edges,
                                       int64_t path_capacity, int64_t *path_offset, int32_t *path_len,
                                       float *path_dist, uint64_t *counters)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return orgpu_fail(ORGPU_EINVAL, "route_ref_noalloc: bad device %d", device);
    if (!g || g->n_nodes < 1) return orgpu_fail(ORGPU_EINVAL, "route_ref_noalloc: null or empty graph");
    if (slots < 1 || heap_cap < 64) return orgpu_fail(ORGPU_EINVAL, "route_ref_noalloc: slots >= 1 and heap_cap >= 64 required");
    const uint64_t state_bytes = (uint64_t)slots * (uint64_t)g->n_nodes * sizeof(RefNode);
    const uint64_t heap_bytes = (uint64_t)slots * (uint64_t)heap_cap * sizeof(HeapEnt);
    if (256ull + state_bytes + heap_bytes > (uint64_t)ORGPU_REF_NOALLOC_BYTES)
        return orgpu_fail(ORGPU_EINVAL, "route_ref_noalloc: %d slots x (%lld nodes + %lld heap entries) x 16 B "
                          "exceed the static scratch of %llu bytes", slots, (long long)g->n_nodes,
                          (long long)heap_cap, (unsigned long long)ORGPU_REF_NOALLOC_BYTES);
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    int prev = 0;
    HIP_CHECK_RET(hipGetDevice(&prev));
    if (prev != device) HIP_CHECK_RET(hipSetDevice(device));
    int rc = ORGPU_OK;
    hipStream_t s = (hipStream_t)hip_stream;
    if (!g_noalloc_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) rc = orgpu_fail_hip(e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                       "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        else { g_noalloc_ptr[device] = p; g_noalloc_state_bytes[device] = -1; }
    }
    if (rc == ORGPU_OK) {
        unsigned char *base = static_cast<unsigned char *>(g_noalloc_ptr[device]);
        // node records must hold stamps of earlier searches only: when the
        // layout changes, bytes that held heap entries may become records,
        // so the records are cleared (a fill kernel) and the ids restart
        if (g_noalloc_state_bytes[device] != (int64_t)state_bytes) {
            hipLaunchKernelGGL(k_ref_clear, dim3(1024), dim3(256), 0, s, reinterpret_cast<uint4 *>(base + 256),
                               (int64_t)(state_bytes / 16));
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) rc = orgpu_fail_hip(e, "k_ref_clear", __FILE__, __LINE__);
            g_noalloc_state_bytes[device] = (int64_t)state_bytes;
            g_noalloc_sid[device] = 0;
        }
        if (rc == ORGPU_OK) {
            RefScratch sc;
            sc.ctl = reinterpret_cast<RefCtl *>(base);
            sc.cursor = reinterpret_cast<unsigned long long *>(base + 64);
            sc.state = reinterpret_cast<RefNode *>(base + 256);
            sc.heap = reinterpret_cast<HeapEnt *>(base + 256 + state_bytes);
            sc.nodes = g->n_nodes; sc.hcap = heap_cap; sc.slots = slots;
            sc.sid = &g_noalloc_sid[device];
            rc = ref_launch(s, sc, g, reqs, n, roi_bitmaps, n_roi, keepout_owner, keepout_cost, slots, path_nodes,
                            path_edges, path_capacity, path_offset, path_len, path_dist, counters);
        }
    }
    if (prev != device) (void)hipSetDevice(prev);
    return rc;
}
