// A2: per-net ROI extraction (bidirectional BFS + truncation), one
// 1024-thread workgroup per net, many nets per launch.
//
// Reference: ROIExtractor.extract_roi_bfs (UPF:1636-1734) and the must-keep
// re-add in _route_all (UPF:4683-4700).  The BFS is level-synchronous and the
// set it marks does not depend on the order nodes are visited within a level,
// so the workgroup expands each frontier in parallel.  Marks: the src wave
// sets bit A (the output row), the dst wave sets bit B (scratch).  A node the
// other wave already holds is not re-marked -- touching it only flags `met`,
// which ends the search after the current step, so the reference's mark 3 is
// never observed.
//
// Truncation keeps the max_nodes lowest node ids (the reference sorts the ROI
// ascending), minus the nodes its must-keep swap pushes out (positions 0/1 when
// src/dst fall beyond the cut), then src and dst are re-added.
#include <algorithm>
#include <vector>

#include "common.h"
#include "engine.h"

#define RWG 1024

struct RoiArgs {
    int64_t N;
    int64_t words;
    const int32_t *indptr, *indices;
    const orgpu_roi_req *reqs;
    int n;
    uint32_t *rows;       // [n_roi][words]
    int32_t n_roi;
    uint32_t *scratch_b;  // [slots][words]
    int32_t *list_a, *list_b;  // [slots][N]
    orgpu_roi_info *info;  // device [n]
    uint32_t *next;        // next request to take (zero at launch)
};

__device__ __forceinline__ bool bit_get(const uint32_t *bm, int32_t v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

__device__ __forceinline__ int wave_scan_incl(int x)
{
    const int lane = (int)__lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    return x;
}

// position of the r-th (0-based) set bit of w; w has more than r bits set
__device__ __forceinline__ int nth_set_bit(uint32_t w, int r)
{
    int pos = 0, c;
    c = __popc(w & 0xFFFFu); if (r >= c) { r -= c; w >>= 16; pos += 16; }
    c = __popc(w & 0xFFu);   if (r >= c) { r -= c; w >>= 8;  pos += 8; }
    c = __popc(w & 0xFu);    if (r >= c) { r -= c; w >>= 4;  pos += 4; }
    c = __popc(w & 0x3u);    if (r >= c) { r -= c; w >>= 2;  pos += 2; }
    if (r >= (int)(w & 1u)) pos += 1;
    return pos;
}

__global__ void __launch_bounds__(RWG)
k_roi(RoiArgs a)
{
    __shared__ int s_na, s_nb, s_met;   // s_na + s_nb = nodes seen
    __shared__ int s_fa0, s_fa1, s_fb0, s_fb1, s_depth;
    __shared__ int s_wsum[RWG / 64];
    __shared__ int s_cut, s_min0, s_min1, s_kept_delta;
    const int tid = threadIdx.x;
    uint32_t *B = a.scratch_b + (int64_t)blockIdx.x * a.words;
    int32_t *LA = a.list_a + (int64_t)blockIdx.x * a.N;
    int32_t *LB = a.list_b + (int64_t)blockIdx.x * a.N;
    // nets are taken one at a time from a shared counter, so a workgroup
    // that drew small ROIs takes more of them
    __shared__ int s_next;
    for (;;) {
        if (tid == 0) s_next = (int)atomicAdd(a.next, 1u);
        __syncthreads();
        const int r = s_next;
        __syncthreads();
        if (r >= a.n) break;
        const orgpu_roi_req q = a.reqs[r];
        uint32_t *A = a.rows + (int64_t)q.roi_index * a.words;
        for (int64_t w = tid; w < a.words; w += RWG) A[w] = 0u;
        __syncthreads();
        if (tid == 0) {
            LA[0] = q.src; LB[0] = q.dst;
            A[q.src >> 5] |= 1u << (q.src & 31);
            if (q.dst != q.src) B[q.dst >> 5] |= 1u << (q.dst & 31);
            s_na = 1; s_nb = q.dst != q.src ? 1 : 0;
            s_fa0 = 0; s_fa1 = 1; s_fb0 = 0; s_fb1 = s_nb;
            s_met = 0; s_depth = 0;
        }
        __syncthreads();
        const double stop = (double)q.max_nodes * 1.5;
        while (true) {
            const bool go = s_depth < q.max_depth && (s_fa1 > s_fa0 || s_fb1 > s_fb0) && !s_met;
            __syncthreads();
            if (!go) break;
            for (int wave = 0; wave < 2; ++wave) {
                if (wave == 1 && s_met) break;
                const int f0 = wave == 0 ? s_fa0 : s_fb0;
                const int f1 = wave == 0 ? s_fa1 : s_fb1;
                uint32_t *mine = wave == 0 ? A : B;
                const uint32_t *other = wave == 0 ? B : A;
                int32_t *L = wave == 0 ? LA : LB;
                int *cnt = wave == 0 ? &s_na : &s_nb;
                // a thread's edges go four at a time with their loads and
                // bitmap atomics issued together (the marks a level leaves do
                // not depend on the order its edges are visited)
                for (int i = f0 + tid; i < f1; i += RWG) {
                    const int32_t u = L[i];
                    const int e0 = a.indptr[u], e1 = a.indptr[u + 1];
                    for (int e = e0; e < e1; e += 4) {
                        int32_t v[4];
                        bool ok[4];
                        uint32_t old[4], m[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            ok[q] = e + q < e1;
                            v[q] = ok[q] ? a.indices[e + q] : 0;
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (ok[q] && bit_get(other, v[q])) { s_met = 1; ok[q] = false; }
                        // a node this wave already holds needs no atomic: bits
                        // of `mine` are only set during a net's search, and
                        // its clears (at the start of the net for A, at the
                        // end of the previous net for B) are this workgroup's
                        // own stores
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            m[q] = 1u << (v[q] & 31);
                            if (ok[q] && (mine[v[q] >> 5] & m[q])) ok[q] = false;
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            old[q] = ok[q] ? atomicOr(&mine[v[q] >> 5], m[q]) : m[q];
                        // newly marked nodes join the list: one LDS atomic
                        // per wave and edge position (the active lanes that
                        // marked a node take consecutive entries)
                        const uint64_t below = (1ull << __lane_id()) - 1ull;
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const bool fresh = !(old[q] & m[q]);
                            const uint64_t bal = __ballot(fresh);
                            if (bal == 0ull) continue;
                            const int leader = __ffsll((unsigned long long)bal) - 1;
                            int base = 0;
                            if ((int)__lane_id() == leader) base = atomicAdd(cnt, __popcll(bal));
                            base = __shfl(base, leader);
                            if (fresh) L[base + __popcll(bal & below)] = v[q];
                        }
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    if (wave == 0) { s_fa0 = f1; s_fa1 = s_na; }
                    else { s_fb0 = f1; s_fb1 = s_nb; }
                }
                __syncthreads();
            }
            if (tid == 0) s_depth = s_depth + 1;
            __syncthreads();
            if ((double)(s_na + s_nb) > stop) break;
        }
        // seen = A | B: move the dst wave into the output row, clear scratch.
        // Word by word with plain stores: neither row changes during this
        // loop except through these stores, and every thread that writes a
        // word writes the same value A | B, so the stores need no atomics
        // (the workgroup's own stores and loads share one L1 and L2).
        const int nb_total = s_nb;
        for (int i = tid; i < nb_total; i += RWG) {
            const int32_t w = LB[i] >> 5;
            A[w] = A[w] | B[w];
        }
        __syncthreads();
        // clear the scratch words the dst wave used (whole words: every bit
        // of B belongs to this net)
        for (int i = tid; i < nb_total; i += RWG) {
            const int32_t v = LB[i];
            B[v >> 5] = 0u;
        }
        __syncthreads();
        const long long seen = (long long)s_na + s_nb;
        long long kept = seen;
        if (seen > q.max_nodes) {
            // Rank the seen nodes by id in one pass over the row: thread t
            // counts the set bits of its contiguous word range, a block scan
            // gives each range's first rank, and the ranges that hold ranks
            // 1 and 2 (the two lowest ids: positions 0 and 1 of the sorted
            // ROI) and max_nodes (the cut) locate those bits.
            const int64_t per = (a.words + RWG - 1) / RWG;
            const int64_t w0 = std::min<int64_t>((int64_t)tid * per, a.words);
            const int64_t w1 = std::min<int64_t>(w0 + per, a.words);
            int c = 0;
            for (int64_t w = w0; w < w1; ++w) c += __popc(A[w]);
            const int incl = wave_scan_incl(c);
            if ((tid & 63) == 63) s_wsum[tid >> 6] = incl;
            if (tid == 0) { s_cut = -1; s_min0 = 0x7FFFFFFF; s_min1 = 0x7FFFFFFF; s_kept_delta = 0; }
            __syncthreads();
            long long excl = incl - c;
            for (int j = 0; j < (tid >> 6); ++j) excl += s_wsum[j];
            // id of the r-th (1-based) seen node when this range holds it
            auto find = [&](long long r) -> int32_t {
                if (!(excl < r && r <= excl + c)) return -1;
                int need = (int)(r - excl);
                for (int64_t w = w0; w < w1; ++w) {
                    const uint32_t x = A[w];
                    const int pc = __popc(x);
                    if (need <= pc) return (int32_t)(w * 32 + nth_set_bit(x, need - 1));
                    need -= pc;
                }
                return -1;
            };
            const int32_t f_cut = find(q.max_nodes), f0 = find(1), f1 = find(2);
            if (f_cut >= 0) s_cut = f_cut;
            if (f0 >= 0) s_min0 = f0;
            if (f1 >= 0) s_min1 = f1;
            __syncthreads();
            const int32_t cut = s_cut;
            // the bits above the cut leave the row (this thread's words)
            for (int64_t w = w0; w < w1; ++w) {
                if (w * 32 + 31 <= cut) continue;
                const int64_t keep_bits = (int64_t)cut - w * 32 + 1;   // ids w*32 .. cut stay
                const uint32_t keep = keep_bits <= 0 ? 0u : (1u << keep_bits) - 1u;
                const uint32_t x = A[w];
                if (x & ~keep) A[w] = x & keep;
            }
            __syncthreads();
            if (tid == 0) {
                kept = q.max_nodes;
                const int32_t k0 = min(q.src, q.dst), k1 = max(q.src, q.dst);
                // must-keep swaps push positions 0 / 1 out of the kept range
                if (k0 > cut && (A[s_min0 >> 5] >> (s_min0 & 31) & 1u)) { A[s_min0 >> 5] &= ~(1u << (s_min0 & 31)); kept--; }
                if (k1 > cut && (A[s_min1 >> 5] >> (s_min1 & 31) & 1u)) { A[s_min1 >> 5] &= ~(1u << (s_min1 & 31)); kept--; }
                s_kept_delta = (int)(kept - q.max_nodes);
            }
            __syncthreads();
            kept = q.max_nodes + s_kept_delta;
        }
        if (tid == 0) {
            // _route_all re-adds src and dst when missing
            const int32_t ks[2] = {q.src, q.dst};
            if (q.readd) {
                for (int j = 0; j < 2; ++j) {
                    const int32_t v = ks[j];
                    if (!((A[v >> 5] >> (v & 31)) & 1u)) { A[v >> 5] |= 1u << (v & 31); kept++; }
                }
            }
            if (a.info) {
                a.info[r].seen = seen;
                a.info[r].kept = kept;
                a.info[r].depth = s_depth;
                a.info[r].met = s_met;
                a.info[r].src_in = (A[q.src >> 5] >> (q.src & 31)) & 1u;
                a.info[r].dst_in = (A[q.dst >> 5] >> (q.dst & 31)) & 1u;
            }
        }
        __syncthreads();
    }
}

static int ensure_roi_scratch(Engine &en, int64_t N, int slots)
{
    const int64_t words = (N + 31) / 32;
    if (en.roi_n >= N && en.roi_slots >= slots) return ORGPU_OK;
    if (en.roi_b) { (void)hipFree(en.roi_b); (void)hipFree(en.roi_la); (void)hipFree(en.roi_lb); }
    en.roi_b = nullptr; en.roi_la = nullptr; en.roi_lb = nullptr; en.roi_n = 0; en.roi_slots = 0;
    HIP_CHECK_RET(hipMalloc((void **)&en.roi_b, (size_t)slots * words * sizeof(uint32_t)));
    HIP_CHECK_RET(hipMemsetAsync(en.roi_b, 0, (size_t)slots * words * sizeof(uint32_t), en.stream));
    HIP_CHECK_RET(hipMalloc((void **)&en.roi_la, (size_t)slots * N * sizeof(int32_t)));
    HIP_CHECK_RET(hipMalloc((void **)&en.roi_lb, (size_t)slots * N * sizeof(int32_t)));
    en.roi_n = N; en.roi_slots = slots;
    return ORGPU_OK;
}

static int ensure_roi_staging(Engine &en, int64_t n)
{
    if (en.roi_req_cap >= n) return ORGPU_OK;
    if (en.roi_req) { (void)hipFree(en.roi_req); (void)hipFree(en.roi_info); }
    en.roi_req = nullptr; en.roi_info = nullptr; en.roi_req_cap = 0;
    const int64_t cap = std::max<int64_t>(n, 1024);
    HIP_CHECK_RET(hipMalloc((void **)&en.roi_req, sizeof(orgpu_roi_req) * cap));
    HIP_CHECK_RET(hipMalloc((void **)&en.roi_info, sizeof(orgpu_roi_info) * cap));
    if (!en.roi_next) HIP_CHECK_RET(hipMalloc((void **)&en.roi_next, sizeof(uint32_t)));
    en.roi_req_cap = cap;
    return ORGPU_OK;
}

extern "C" int orgpu_roi_bfs(orgpu_t *h, const orgpu_graph *g, const orgpu_roi_req *reqs, int32_t n,
                             uint32_t *roi_bitmaps, int32_t n_roi, orgpu_roi_info *info)
{
    if (!h || !g || n < 0 || (n > 0 && (!reqs || !roi_bitmaps))) return orgpu_fail(ORGPU_EINVAL, "roi_bfs: bad argument");
    if (n == 0) return ORGPU_OK;
    const int64_t N = g->n_nodes;
    if (N <= 0 || N >= KEY_MAX_NODES || !g->indptr || !g->indices) return orgpu_fail(ORGPU_EINVAL, "roi_bfs: bad graph");
    std::vector<int> row_user(n_roi > 0 ? n_roi : 0, -1);
    for (int i = 0; i < n; ++i) {
        const orgpu_roi_req &q = reqs[i];
        if (q.src < 0 || q.src >= N || q.dst < 0 || q.dst >= N || q.roi_index < 0 || q.roi_index >= n_roi ||
            q.max_nodes < 1 || q.max_depth < 0)
            return orgpu_fail(ORGPU_EINVAL, "roi_bfs: request %d invalid (src %d dst %d row %d of %d)", i, q.src, q.dst,
                              q.roi_index, n_roi);
        if (row_user[q.roi_index] >= 0)
            return orgpu_fail(ORGPU_EINVAL, "roi_bfs: requests %d and %d share bitmap row %d", row_user[q.roi_index], i,
                              q.roi_index);
        row_user[q.roi_index] = i;
    }
    HIP_CHECK_RET(hipSetDevice(h->device));
    Engine &en = h->en;
    // concurrent workgroups, each with ~9 bytes of scratch per node (up to
    // the handle's ROI scratch limit, 32 GB by default): a net's BFS is a
    // chain of short levels, so throughput comes from many nets in flight
    // (four workgroups per CU at 1024)
    const int slots = std::min<int>(n, (int)std::max<int64_t>(1, std::min<int64_t>(1024, en.roi_limit / (N * 9 + 1))));
    int rc = ensure_roi_scratch(en, N, slots);
    if (!rc) rc = ensure_roi_staging(en, n);
    if (rc) return rc;
    hipStream_t s = h->stream;
    orgpu_roi_req *dreq = en.roi_req;
    orgpu_roi_info *dinfo = en.roi_info;
    uint32_t *dnext = en.roi_next;
    HIP_CHECK_RET(hipMemcpyAsync(dreq, reqs, sizeof(orgpu_roi_req) * n, hipMemcpyHostToDevice, s));
    HIP_CHECK_RET(hipMemsetAsync(dnext, 0, sizeof(uint32_t), s));
    RoiArgs a;
    a.N = N; a.words = (N + 31) / 32; a.indptr = g->indptr; a.indices = g->indices;
    a.reqs = dreq; a.n = n; a.rows = roi_bitmaps; a.n_roi = n_roi;
    a.scratch_b = en.roi_b; a.list_a = en.roi_la; a.list_b = en.roi_lb; a.info = dinfo; a.next = dnext;
    hipLaunchKernelGGL(k_roi, dim3(slots), dim3(RWG), 0, s, a);
    HIP_CHECK_RET(hipGetLastError());
    if (info) HIP_CHECK_RET(hipMemcpyAsync(info, dinfo, sizeof(orgpu_roi_info) * n, hipMemcpyDeviceToHost, s));
    HIP_CHECK_RET(hipStreamSynchronize(s));
    return ORGPU_OK;
}

extern "C" int orgpu_set_roi_scratch(orgpu_t *h, int64_t max_bytes)
{
    if (!h) return orgpu_fail(ORGPU_EINVAL, "set_roi_scratch: null handle");
    if (max_bytes < 1) return orgpu_fail(ORGPU_EINVAL, "set_roi_scratch: limit %lld must be positive", (long long)max_bytes);
    Engine &en = h->en;
    en.roi_limit = max_bytes;
    if (en.roi_b && (int64_t)en.roi_slots * (en.roi_n * 9 + 1) > max_bytes) {
        // drop scratch above the new limit; the next call sizes it again
        HIP_CHECK_RET(hipSetDevice(h->device));
        HIP_CHECK_RET(hipStreamSynchronize(h->stream));
        (void)hipFree(en.roi_b); (void)hipFree(en.roi_la); (void)hipFree(en.roi_lb);
        en.roi_b = nullptr; en.roi_la = nullptr; en.roi_lb = nullptr; en.roi_n = 0; en.roi_slots = 0;
    }
    return ORGPU_OK;
}
