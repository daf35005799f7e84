// A3/A4: batched exact SSSP (bucketed delta-stepping on packed 64-bit keys)
// and the backtrace, for K concurrent searches ("slots") on one CSR graph.
//
// Reference behaviour replaced: SimpleDijkstra.find_path_roi (UPF:1801-1887),
// the owner-aware keepout (UPF:3150-3228), the parent walk (UPF:1880-1887)
// and _path_to_edges (UPF:5463-5476).  The result is the FP32 shortest path
// with the lowest-predecessor-index tie-break (see common.h for the key).
//
// Design (one HIP stream, no host sync inside a round):
//  * state: keys u64[K][NP]; a work item is the u64 slot*NP + node, where the
//    per-slot stride NP is N rounded up to a whole chunk (2048 items), so a
//    chunk belongs to one slot and a batch gets its slot by one division per
//    chunk.
//  * buckets are bitmaps over all items (one bit per (slot, node)): a ring of
//    nb bitmaps for buckets cur .. cur+nb-1, an OUT bitmap for re-relaxations
//    inside the current bucket and an overflow bitmap for buckets >= cur+nb.
//    Setting a bit is the push (and its own de-duplication); a round scans
//    its bucket's bitmap in item order, so a wavefront's items are neighbours
//    on the lattice and their CSR rows and neighbour keys share cache lines.
//  * one round = a k_round launch (every workgroup expands a share of the
//    bucket's listed chunks and leaves its counts in its own partials entry)
//    followed by a one-workgroup k_control launch that reduces the partials,
//    swaps OUT in, advances the bucket, detects finished slots, and schedules
//    an overflow redistribution or goes idle.  The host queues rounds blindly
//    and polls a status block every few dozen rounds.
//  * every slot shares the global bucket index; a slot seeded at bucket c
//    measures its distances from c (per-slot base), so finished slots are
//    refilled with the next request while the others keep running.
//  * a wavefront takes the set bits of 64 bitmap words (2048 items) at a
//    time (ballot-free prefix scan + bit select), gathers them into 64-item
//    batches, and each lane relaxes its item's CSR row with 64-bit
//    atomicMins; improvements inside the current bucket go to the wave's
//    LDS queue and are expanded in the same round.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"

#define WG 256
#define WAVES_PER_WG (WG / 64)
#define MAX_SLOTS 2048          // slots of the bucket-round engine
#define RESERVE_MAX_SLOTS 4096  // slots a handle may reserve (one-wave engine)
#define CHUNK_ITEMS 2048
#define MAX_NB MAX_RING

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ bool test_and_set_bit(uint32_t *bm, uint32_t item)
{
    const uint32_t m = 1u << (item & 31);
    return (atomicOr(&bm[item >> 5], m) & m) != 0;
}

// inclusive wave scan
__device__ __forceinline__ int wave_scan(int x)
{
    const int lane = (int)lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    return x;
}

__device__ __forceinline__ int32_t wave_sum_i32(int32_t x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)x, o, 64), hi = __shfl_xor((uint32_t)(x >> 32), o, 64);
        x += ((uint64_t)hi << 32) | lo;
    }
    return x;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)x, o, 64), hi = __shfl_xor((uint32_t)(x >> 32), o, 64);
        const uint64_t y = ((uint64_t)hi << 32) | lo;
        x = y < x ? y : x;
    }
    return x;
}

// Lane of the batch element t: the highest lane j whose exclusive prefix
// excl_j is <= t (lanes with a zero count share the next lane's prefix, so
// the highest such lane is the one that owns element t).
__device__ __forceinline__ int owner_lane(int excl, int t)
{
    int j = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        const int cand = j + step;
        const int ec = __shfl(excl, cand);
        if (ec <= t) j = cand;
    }
    return j;
}

// position of the r-th (0-based) set bit of w; w has more than r bits set
__device__ __forceinline__ uint32_t select_bit(uint32_t w, uint32_t r)
{
    uint32_t pos = 0, c;
    c = __popc(w & 0xFFFFu); if (r >= c) { r -= c; w >>= 16; pos += 16; }
    c = __popc(w & 0xFFu);   if (r >= c) { r -= c; w >>= 8;  pos += 8; }
    c = __popc(w & 0xFu);    if (r >= c) { r -= c; w >>= 4;  pos += 4; }
    c = __popc(w & 0x3u);    if (r >= c) { r -= c; w >>= 2;  pos += 2; }
    if (r >= (w & 1u)) pos += 1;
    return pos;
}

// Appends chunk ch to chunk-list storage sid (its region's bin).
__device__ __forceinline__ void list_append(const RoundArgs &a, uint32_t sid, uint32_t ch)
{
    const uint32_t bin = (uint32_t)((uint64_t)ch * LIST_BINS / a.chunks);
    const uint64_t cell = (uint64_t)sid * LIST_BINS + bin;
    const uint32_t pos = atomicAdd(&a.list_cnt[cell], 1u);
    a.lists[cell * a.bin_cap + pos] = ch;
}

// Marks bit `item` of bucket bitmap `buf`; returns true when it was clear.
// The first bit of a word also marks the word's chunk (64 words) in the
// bucket's summary bitmap, and the chunk whose summary bit this turns on is
// appended to the bucket's chunk list: the chunks a round must visit, each
// listed once.  list_sid: the bucket's chunk-list storage.
__device__ __forceinline__ bool push_bit(const RoundArgs &a, uint32_t buf, uint32_t list_sid, uint64_t item)
{
    const uint32_t m = 1u << (item & 31);
    const uint32_t old = atomicOr(&a.bits[(uint64_t)buf * a.bitmap_words + (item >> 5)], m);
    if (old & m) return false;
    if (old == 0u) {
        const uint32_t ch = (uint32_t)(item >> 11);
        const uint32_t cm = 1u << (ch & 31);
        if (!(atomicOr(&a.summ[(uint64_t)buf * a.summ_words + (ch >> 5)], cm) & cm)) list_append(a, list_sid, ch);
    }
    return true;
}

// push_bit for a relaxation push: during a relaxation round the target
// bitmaps (OUT, later ring buckets, overflow) only gain bits, so a set bit
// seen by a plain load is set and the atomic can be skipped (a stale copy can
// only miss bits, and then the atomic runs).
__device__ __forceinline__ bool push_bit_relax(const RoundArgs &a, uint32_t buf, uint32_t list_sid, uint64_t item)
{
    const uint32_t m = 1u << (item & 31);
    if (a.bits[(uint64_t)buf * a.bitmap_words + (item >> 5)] & m) return false;
    return push_bit(a, buf, list_sid, item);
}

// A key of (slot, node) left KEY_INF: records the node's chunk of
// 2^TOUCH_SHIFT keys so that reusing the slot resets only the chunks its
// search touched.  1024-key chunks (8 KB) keep the reset close to the keys
// a search set (it covers a compact region on each of many layers, a few
// hundred nodes of every lattice row it crosses) while first touches, one
// atomic each, stay rare.
#define TOUCH_SHIFT 10
#define TOUCH_KEYS (1u << TOUCH_SHIFT)
__device__ __forceinline__ void touch_key(const RoundArgs &a, uint32_t slot, uint32_t node)
{
    const uint32_t ch = node >> TOUCH_SHIFT;
    const uint32_t m = 1u << (ch & 31);
    uint32_t *w = a.touch_bits + (uint64_t)slot * a.touch_words + (ch >> 5);
    if (*w & m) return;
    if (atomicOr(w, m) & m) return;
    const uint32_t pos = atomicAdd(&a.touch_n[slot], 1u);
    a.touch_list[(uint64_t)slot * a.touch_cap + pos] = ch;
}

// Workgroups are dispatched round-robin over the 8 XCDs.  Each XCD takes one
// contiguous eighth of the round's chunk list (chunks are in item order, so
// a slot's keys and the lattice rows its frontier touches stay in that XCD's
// L2), and inside it the waves stride over the chunks.
struct ChunkWalk {
    uint32_t first, end, stride;
};

__device__ __forceinline__ ChunkWalk chunk_walk(uint32_t n, uint32_t wave)
{
    const uint32_t g = gridDim.x, b = blockIdx.x;
    ChunkWalk w;
    if (g % 8 == 0) {
        const uint32_t xcd = b % 8, per = g / 8, local = b / 8;
        w.first = (uint32_t)((uint64_t)n * xcd / 8) + local * WAVES_PER_WG + wave;
        w.end = (uint32_t)((uint64_t)n * (xcd + 1) / 8);
        w.stride = per * WAVES_PER_WG;
    } else {
        w.first = b * WAVES_PER_WG + wave;
        w.end = n;
        w.stride = g * WAVES_PER_WG;
    }
    return w;
}

// A wave's queue of items it pushed into the current bucket: expanded later
// in the same round (a round then covers several hops of a bucket instead of
// one).  Ring of LQ_CAP entries in LDS; tail is advanced by the lanes with an
// LDS atomic, head only at wave-uniform points.
#define LQ_CAP 1024
struct LocalQ {
    uint32_t *buf;      // node
    uint16_t *sbuf;     // slot
    uint32_t *tail;
    uint32_t head;
};

// After a batch (wave-uniform): lanes that found the queue full pushed to OUT
// instead, so the tail is pulled back to the last stored entry.
__device__ __forceinline__ void lq_clamp(const LocalQ &lq)
{
    const uint32_t t = __hip_atomic_load(lq.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (t - lq.head > LQ_CAP && __lane_id() == 0)
        __hip_atomic_store(lq.tail, lq.head + LQ_CAP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// Lower bound of the remaining cost from node v to the slot's destination
// (goal-directed bucketing; 0 when the call runs in plain Dijkstra order).
__device__ __forceinline__ float heur(const RoundArgs &a, const SlotInfo &si, uint32_t v)
{
    if ((a.h_xy == 0.0f && a.h_z == 0.0f) || si.ms) return 0.0f;
    const int z = (int)(v / (uint32_t)a.plane);
    const int r = (int)(v - (uint32_t)z * (uint32_t)a.plane);
    const int y = r / a.X;
    const int x = r - y * a.X;
    return a.h_xy * (float)(abs(x - si.dx) + abs(y - si.dy)) + a.h_z * (float)abs(z - si.dz);
}

// Expands one item per lane (lanes with has=false only take part in the
// wave's collective steps): consumes it, and if its key still lies in the
// current bucket relaxes its out-edges.  A lane walks its own CSR row four
// edges at a time.  The loads are arranged in as few dependent levels as the
// data allows, because a sparse bucket is expanded hop by hop through the
// wave's local queue and each hop costs the latency of its dependent chain:
// (1) slot, own key and CSR row bounds; (2) the four edges' targets and
// costs; (3) their ROI bits and target keys; (4) the four atomicMins; then
// the pushes, and the touched-chunk records last.
__device__ __forceinline__ void relax_item(const RoundArgs &a, uint32_t slot, uint32_t u, bool has, bool from_bitmap,
                                           const LocalQ &lq, uint64_t cur, uint32_t nb,
                                           const uint32_t *s_ring, const uint32_t *s_lid, uint32_t s_out, uint32_t s_ovf,
                                           int32_t *s_slot_delta, int32_t *s_buf_delta,
                                           unsigned long long &my_items, unsigned long long &my_edges,
                                           uint64_t &my_ovf_min)
{
    if (!has) return;
    const uint64_t NP = (uint64_t)a.NP;
    if (from_bitmap) atomicAdd(&s_slot_delta[slot], -1);
    const SlotInfo si = a.slots[slot];
    uint64_t *keys = a.keys + (uint64_t)slot * NP;
    const uint64_t ku = __hip_atomic_load(&keys[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int row = a.indptr[u];
    const int row_end = a.indptr[u + 1];
    if (si.state != 1) return;
    if (ku == KEY_INF) return;
    {
        // Plain order: an item whose key has dropped below the current bucket
        // since its push was already expanded there (stale copy).  Goal-
        // directed order: FP32 rounding of g + h can put a neighbour one
        // bucket below the bucket that pushed it (into OUT), so everything at
        // or below the current bucket is expanded.
        const uint64_t b = si.base + key_bucket_rel(ku, a.inv_delta, heur(a, si, u));
        const bool goal = a.h_xy != 0.0f || a.h_z != 0.0f;
        if (b > cur || (b < cur && !goal)) return;
    }
    // a target of a multi-source search is settled but not expanded
    if (si.ms && ((a.tbits[(uint64_t)slot * a.tbits_words + (u >> 5)] >> (u & 31)) & 1u)) return;
    const int deg = row_end - row;
    my_items++;
    my_edges += (unsigned long long)deg;
    const uint32_t ko0 = key_ko(ku);
    const float d0 = key_dist(ku);
    // keepout: edges of other nets' via barrels (planar edges leaving an
    // owned node) count as keepout edges instead of adding their cost
    const bool ko_node = si.use_keepout && a.ko_owner && a.ko_owner[u] >= 0 && a.ko_owner[u] != si.net_id;
    const uint32_t *roi = si.roi_index >= 0 ? a.roi + (uint64_t)si.roi_index * a.roi_words : nullptr;
    for (int k0 = 0; k0 < deg; k0 += 4) {
        uint32_t v[4];
        float c[4];
        bool ok[4], planar_ko[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            ok[q] = k0 + q < deg;
            planar_ko[q] = false;
            if (ok[q]) {
                v[q] = (uint32_t)a.indices[row + k0 + q];
                c[q] = a.cost[row + k0 + q];
                if (ko_node) planar_ko[q] = a.kind[row + k0 + q] == 0;
            }
        }
        // target keys are read with the ROI bits (every target id is a
        // valid node); an edge leaving the ROI drops its key unused
        uint32_t rw[4];
        uint64_t kv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!ok[q]) continue;
            if (roi) rw[q] = roi[v[q] >> 5];
            kv[q] = keys[v[q]];
        }
        uint64_t nk[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!ok[q]) continue;
            if (roi && !((rw[q] >> (v[q] & 31)) & 1u)) { ok[q] = false; continue; }
            uint32_t ko = ko0;
            float nd = d0;
            if (planar_ko[q]) ko = ko_add(ko);
            else nd = nd + c[q];
            nk[q] = key_pack(ko, nd, u + 1u);
            ok[q] = nd < __builtin_inff() && nk[q] < kv[q];
        }
        uint64_t old[4];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ok[q]) old[q] = atomicMin((unsigned long long *)&keys[v[q]], (unsigned long long)nk[q]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (!ok[q] || !(nk[q] < old[q] && key_cost(nk[q]) < key_cost(old[q]))) continue;
            const uint64_t b = si.base + key_bucket_rel(nk[q], a.inv_delta, heur(a, si, v[q]));
            if (b <= cur) {
                // same bucket: this wave expands it later in the round (its
                // local queue), unless the queue is full
                const uint32_t pos = atomicAdd(lq.tail, 1u);
                if (pos - lq.head < LQ_CAP) {
                    __hip_atomic_store(&lq.buf[pos & (LQ_CAP - 1)], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    __hip_atomic_store(&lq.sbuf[pos & (LQ_CAP - 1)], (uint16_t)slot, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
                    continue;
                }
            }
            uint32_t buf;
            if (b <= cur) buf = s_out;
            else if (b < cur + nb) buf = s_ring[b % nb];
            else buf = s_ovf;
            if (buf == s_ovf && b < my_ovf_min) my_ovf_min = b;
            if (push_bit_relax(a, buf, s_lid[buf], (uint64_t)slot * NP + v[q])) {
                atomicAdd(&s_slot_delta[slot], 1);
                atomicAdd(&s_buf_delta[buf], 1);
            }
        }
        // a key that left KEY_INF records its chunk for the slot's reset
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (ok[q] && old[q] == KEY_INF) touch_key(a, slot, v[q]);
    }
}

// Entry li of a binned chunk list (binoff: exclusive prefix of the bins).
__device__ __forceinline__ uint32_t list_at(const uint32_t *lst, const uint32_t *binoff, uint64_t bin_cap, uint32_t li)
{
    uint32_t b = 0;
#pragma unroll
    for (uint32_t step = LIST_BINS / 2; step > 0; step >>= 1)
        if (binoff[b + step] <= li) b += step;
    return lst[(uint64_t)b * bin_cap + (li - binoff[b])];
}

// One round: expand (RELAX) or redistribute (REDIST) every item of the
// round's input bucket, visiting the chunks the control kernel listed.  Each
// workgroup leaves its bucket-count changes and counters in its own partials
// entry (plain stores); the control kernel that follows reduces them.
__global__ void __launch_bounds__(WG)
k_round(RoundArgs a)
{
    Ctrl *c = a.ctrl;
    const uint32_t mode = c->mode;
    if (mode == MODE_IDLE) return;
    const uint64_t cur = c->cur;
    const uint32_t nb = a.nb;

    __shared__ int32_t s_slot_delta[MAX_SLOTS];   // pending-item change per slot
    __shared__ int32_t s_buf_delta[MAX_BUFS];     // set-bit change per bitmap
    __shared__ uint32_t s_ring[MAX_RING];
    __shared__ uint32_t s_lid[MAX_BUFS];          // bitmap -> chunk-list storage
    __shared__ uint32_t s_out, s_ovf, s_in, s_nlist, s_spare;
    __shared__ uint32_t s_binoff[LIST_BINS + 1];  // exclusive prefix of the input list's bins
    __shared__ unsigned long long s_items, s_edges, s_ovf_min;
    __shared__ uint32_t s_clear[WAVES_PER_WG][64]; // redistribution: bits leaving each word
    __shared__ uint32_t s_ibuf[WAVES_PER_WG][128]; // relaxation: items gathered for the next batch (node,
    __shared__ uint16_t s_ibuf_s[WAVES_PER_WG][128]; //  slot)
    __shared__ uint32_t s_lq[WAVES_PER_WG][LQ_CAP]; // relaxation: this wave's same-bucket items (node,
    __shared__ uint16_t s_lq_s[WAVES_PER_WG][LQ_CAP]; //  slot)
    __shared__ uint32_t s_lq_tail[WAVES_PER_WG];
    for (int i = threadIdx.x; i < a.K; i += WG) s_slot_delta[i] = 0;
    if (threadIdx.x < MAX_BUFS) s_buf_delta[threadIdx.x] = 0;
    if (threadIdx.x < WAVES_PER_WG) s_lq_tail[threadIdx.x] = 0;
    if (threadIdx.x < nb) s_ring[threadIdx.x] = c->ring[threadIdx.x];
    if (threadIdx.x < nb + 3) s_lid[threadIdx.x] = c->list_id[threadIdx.x];
    if (threadIdx.x == 0) {
        s_out = c->out_buf; s_ovf = c->ovf_buf;
        s_in = mode == MODE_RELAX ? c->ring[cur % nb] : c->ovf_buf;
        s_spare = c->spare_list;
        s_items = 0; s_edges = 0; s_ovf_min = ~0ull;
    }
    __syncthreads();
    // chunks of the input bucket (no chunk joins this list during the round):
    // bin sizes -> exclusive prefix in LDS
    const uint32_t in_sid = s_lid[s_in];
    if (threadIdx.x < 64) {
        const int cnt = threadIdx.x < LIST_BINS ? (int)a.list_cnt[(uint64_t)in_sid * LIST_BINS + threadIdx.x] : 0;
        const int inc = wave_scan(cnt);
        s_binoff[threadIdx.x] = (uint32_t)(inc - cnt);
        if (threadIdx.x == 63) { s_binoff[LIST_BINS] = (uint32_t)inc; s_nlist = (uint32_t)inc; }
    }
    __syncthreads();
    const uint32_t *clist = a.lists + (uint64_t)in_sid * LIST_BINS * a.bin_cap;

    const uint32_t in_buf = s_in;
    uint32_t *bin = a.bits + (uint64_t)in_buf * a.bitmap_words;
    const uint32_t lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const uint64_t W = a.bitmap_words;
    const uint64_t NP = (uint64_t)a.NP;
    const uint32_t cps = a.cps;
    unsigned long long my_items = 0, my_edges = 0;
    // lowest overflow bucket this lane saw a key move into.  It counts every
    // such key, also one whose item already sits in the overflow bitmap (its
    // bit was set): that item's bucket just dropped, and the bucket jump out
    // of an empty ring must not pass it.
    uint64_t my_ovf_min = ~0ull;
    uint32_t *clr = s_clear[wave];

    const ChunkWalk cw = chunk_walk(s_nlist, wave);
    if (mode == MODE_RELAX) {
        // Items of consecutive chunks are gathered into full 64-item batches
        // (a chunk of a sparse bucket holds a few dozen), so each batch's
        // dependent loads and atomics serve 64 items at once.
        uint32_t *ibuf = s_ibuf[wave];
        uint16_t *ibuf_s = s_ibuf_s[wave];
        int nbuf = 0;
        LocalQ lq{s_lq[wave], s_lq_s[wave], &s_lq_tail[wave], 0u};
        for (uint32_t li = cw.first;; li += cw.stride) {
            const bool more = li < cw.end;
            if (more) {
                const uint64_t ch = list_at(clist, s_binoff, a.bin_cap, li);
                const uint64_t wi = ch * 64 + lane;
                const uint32_t word = wi < W ? bin[wi] : 0u;
                // a relaxation round consumes its whole bucket; nothing else
                // sets bits in it during the round (pushes go to OUT, later
                // ring buckets or overflow), and every chunk of a summary
                // word that has a bit set is listed and consumed in this
                // round, so the word is cleared whole (by each of them)
                if (lane == 0) a.summ[(uint64_t)in_buf * a.summ_words + (ch >> 5)] = 0u;
                if (__ballot(word != 0u) == 0ull) continue;
                if (word) bin[wi] = 0u;
                const int cnt = __popc(word);
                const int incl = wave_scan(cnt);
                const int total = __shfl(incl, 63);
                const int excl = incl - cnt;
                // the chunk's slot and its first node (chunks never straddle slots)
                const uint32_t cslot = (uint32_t)(ch / cps);
                const uint32_t cnode = (uint32_t)(ch - (uint64_t)cslot * cps) * CHUNK_ITEMS;
                for (int b0 = 0; b0 < total; b0 += 64) {
                    const int t = b0 + (int)lane;
                    const int j = owner_lane(excl, t);
                    const uint32_t wj = __shfl(word, j);
                    const int ej = __shfl(excl, j);
                    const int k = min(64, total - b0);
                    if (t < total) {
                        const uint32_t bit = select_bit(wj, (uint32_t)(t - ej));
                        __hip_atomic_store(&ibuf[nbuf + (int)lane], cnode + (uint32_t)j * 32u + bit, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WAVEFRONT);
                        __hip_atomic_store(&ibuf_s[nbuf + (int)lane], (uint16_t)cslot, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WAVEFRONT);
                    }
                    nbuf += k;
                    if (nbuf >= 64) {
                        const uint32_t nd = __hip_atomic_load(&ibuf[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        const uint32_t sl = __hip_atomic_load(&ibuf_s[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        const uint32_t rest = __hip_atomic_load(&ibuf[64 + lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        const uint16_t rest_s = __hip_atomic_load(&ibuf_s[64 + lane], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_WAVEFRONT);
                        __hip_atomic_store(&ibuf[lane], rest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        __hip_atomic_store(&ibuf_s[lane], rest_s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                        nbuf -= 64;
                        relax_item(a, sl, nd, true, true, lq, cur, nb, s_ring, s_lid, s_out, s_ovf, s_slot_delta,
                                   s_buf_delta, my_items, my_edges, my_ovf_min);
                        lq_clamp(lq);
                    }
                }
            } else {
                if (nbuf > 0) {
                    const uint32_t nd = __hip_atomic_load(&ibuf[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    const uint32_t sl = __hip_atomic_load(&ibuf_s[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
                    relax_item(a, sl, nd, (int)lane < nbuf, true, lq, cur, nb, s_ring, s_lid, s_out, s_ovf, s_slot_delta,
                               s_buf_delta, my_items, my_edges, my_ovf_min);
                    lq_clamp(lq);
                }
                break;
            }
        }
        // expand the wave's own same-bucket items: several hops of the bucket
        // in one round; what is left after local_batches batches goes to OUT
        for (int nbat = 0;; ++nbat) {
            const uint32_t tail = __hip_atomic_load(lq.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            if (tail == lq.head) break;
            if (nbat == a.local_batches) {
                for (uint32_t i = lq.head + lane; i < tail; i += 64) {
                    const uint32_t nd = __hip_atomic_load(&lq.buf[i & (LQ_CAP - 1)], __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_WAVEFRONT);
                    const uint32_t sl = __hip_atomic_load(&lq.sbuf[i & (LQ_CAP - 1)], __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_WAVEFRONT);
                    if (push_bit(a, s_out, s_lid[s_out], (uint64_t)sl * NP + nd)) {
                        atomicAdd(&s_slot_delta[sl], 1);
                        atomicAdd(&s_buf_delta[s_out], 1);
                    }
                }
                break;
            }
            const uint32_t n = min(64u, tail - lq.head);
            uint32_t nd = 0, sl = 0;
            if (lane < n) {
                nd = __hip_atomic_load(&lq.buf[(lq.head + lane) & (LQ_CAP - 1)], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
                sl = __hip_atomic_load(&lq.sbuf[(lq.head + lane) & (LQ_CAP - 1)], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WAVEFRONT);
            }
            lq.head += n;
            relax_item(a, sl, nd, lane < n, false, lq, cur, nb, s_ring, s_lid, s_out, s_ovf, s_slot_delta, s_buf_delta,
                       my_items, my_edges, my_ovf_min);
            lq_clamp(lq);
        }
    } else {
        for (uint32_t li = cw.first; li < cw.end; li += cw.stride) {
            const uint64_t ch = list_at(clist, s_binoff, a.bin_cap, li);
            const uint64_t wi = ch * 64 + lane;
            const uint32_t word = wi < W ? bin[wi] : 0u;
            if (__ballot(word != 0u) == 0ull) continue;
            __hip_atomic_store(&clr[lane], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            const int cnt = __popc(word);
            const int incl = wave_scan(cnt);
            const int total = __shfl(incl, 63);
            const int excl = incl - cnt;
            const uint32_t slot = (uint32_t)(ch / cps);
            const uint32_t cnode = (uint32_t)(ch - (uint64_t)slot * cps) * CHUNK_ITEMS;
            for (int b0 = 0; b0 < total; b0 += 64) {
                const int t = b0 + (int)lane;
                const int j = owner_lane(excl, t);
                const uint32_t wj = __shfl(word, j);
                const int ej = __shfl(excl, j);
                if (t < total) {
                    const uint32_t bit = select_bit(wj, (uint32_t)(t - ej));
                    const uint32_t node = cnode + (uint32_t)j * 32u + bit;
                    const uint64_t item = (uint64_t)slot * NP + node;
                    const SlotInfo si = a.slots[slot];
                    bool leaves = true;   // the item leaves the overflow bitmap
                    bool live = false;
                    uint32_t dst_buf = 0;
                    if (si.state == 1) {
                        const uint64_t ku = __hip_atomic_load(&a.keys[item], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (ku != KEY_INF) {
                            const uint64_t b = si.base + key_bucket_rel(ku, a.inv_delta, heur(a, si, node));
                            if (b >= cur) {
                                if (b < cur + nb) {
                                    dst_buf = s_ring[b % nb];
                                    live = push_bit(a, dst_buf, s_lid[dst_buf], item);
                                } else {
                                    leaves = false;
                                    if (b < my_ovf_min) my_ovf_min = b;
                                }
                            }
                        }
                    }
                    if (leaves) {
                        atomicOr(&clr[j], 1u << bit);
                        if (!live) atomicAdd(&s_slot_delta[slot], -1);
                        atomicAdd(&s_buf_delta[in_buf], -1);
                        if (live) atomicAdd(&s_buf_delta[dst_buf], 1);
                    }
                }
            }
            // no other writer touches the overflow bitmap during the round
            const uint32_t m = __hip_atomic_load(&clr[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
            const uint32_t left = word & ~m;
            if (m) bin[wi] = left;
            // a chunk that keeps items moves to the spare list (the overflow
            // bucket's list after the round); an emptied one leaves the summary
            const bool any_left = __ballot(left != 0u) != 0ull;
            if (lane == 0) {
                if (!any_left) {
                    atomicAnd(&a.summ[(uint64_t)in_buf * a.summ_words + (ch >> 5)], ~(1u << (ch & 31)));
                } else {
                    list_append(a, s_spare, (uint32_t)ch);
                }
            }
        }
    }
    {
        const uint64_t m = wave_min_u64(my_ovf_min);
        if (lane == 0 && m != ~0ull) atomicMin(&s_ovf_min, (unsigned long long)m);
    }
    atomicAdd(&s_items, my_items);
    atomicAdd(&s_edges, my_edges);
    __syncthreads();
    // pending items per slot: a workgroup's chunks belong to few slots
    for (int s = threadIdx.x; s < a.K; s += WG) {
        const int dlt = s_slot_delta[s];
        if (dlt != 0) atomicAdd(&a.slots[s].pending, dlt);
    }
    RoundPart *pp = &a.parts[blockIdx.x];
    if (threadIdx.x < nb + 3) pp->buf_delta[threadIdx.x] = s_buf_delta[threadIdx.x];
    if (threadIdx.x == 0) {
        pp->items = s_items;
        pp->edges = s_edges;
        pp->ovf_min = s_ovf_min;
    }
}

#define CTRL_WG 1024
// targets a multi-source slot may have for the early completion test
#define MS_EARLY_TARGETS 256

// Runs after each round (one workgroup): reduces the round's partials,
// advances the bucket window, detects finished slots, schedules an overflow
// redistribution or goes idle, and retires the consumed bucket's chunk list.
// The control block's counts and ring are staged through LDS for the
// serial decision.
__global__ void __launch_bounds__(CTRL_WG)
k_control(RoundArgs a, int round_grid)
{
    Ctrl *c = a.ctrl;
    const uint32_t nb = a.nb;
    const uint32_t nbuf = nb + 3;
    __shared__ int32_t s_cnt[MAX_BUFS];
    __shared__ uint32_t s_ring[MAX_RING];
    __shared__ uint32_t s_mode0, s_out, s_ovf;
    __shared__ uint64_t s_cur, s_ovf_min0;
    __shared__ unsigned long long s_items, s_edges, s_ovf_min;
    __shared__ uint64_t s_newcur;
    __shared__ uint32_t s_mode, s_in_cleared, s_in_list, s_zero_list;
    if (threadIdx.x < nbuf) s_cnt[threadIdx.x] = (int32_t)c->count[threadIdx.x];
    if (threadIdx.x < nb) s_ring[threadIdx.x] = c->ring[threadIdx.x];
    if (threadIdx.x == 0) {
        s_mode0 = c->mode; s_out = c->out_buf; s_ovf = c->ovf_buf;
        s_cur = c->cur; s_ovf_min0 = c->ovf_min;
        s_items = 0; s_edges = 0; s_ovf_min = ~0ull;
    }
    __syncthreads();
    const uint32_t mode = s_mode0;
    if (mode == MODE_IDLE) return;
    const uint64_t cur = s_cur;
    {
        // each wave sums its threads' partials with shuffles; one LDS atomic
        // per wave and counter
        const bool lead = (threadIdx.x & 63) == 0;
        for (uint32_t b = 0; b < nbuf; ++b) {
            int32_t d = 0;
            for (int g = threadIdx.x; g < round_grid; g += CTRL_WG) d += a.parts[g].buf_delta[b];
            d = wave_sum_i32(d);
            if (lead && d) atomicAdd(&s_cnt[b], d);
        }
        unsigned long long it = 0, ed = 0, om = ~0ull;
        for (int g = threadIdx.x; g < round_grid; g += CTRL_WG) {
            const RoundPart *pp = &a.parts[g];
            it += pp->items; ed += pp->edges;
            om = pp->ovf_min < om ? pp->ovf_min : om;
        }
        it = wave_sum_u64(it);
        ed = wave_sum_u64(ed);
        om = wave_min_u64(om);
        if (lead) {
            if (it) atomicAdd(&s_items, it);
            if (ed) atomicAdd(&s_edges, ed);
            if (om != ~0ull) atomicMin(&s_ovf_min, om);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t ovf_min = s_ovf_min < s_ovf_min0 ? s_ovf_min : s_ovf_min0;
        uint64_t newcur = cur;
        uint32_t m = 0xFFFFFFFFu;   // decided below unless set here
        s_in_cleared = 0xFFFFFFFFu;
        s_zero_list = 0xFFFFFFFFu;
        if (mode == MODE_RELAX) {
            // the round consumed (and cleared) its bucket
            const uint32_t pos = (uint32_t)(cur % nb);
            const uint32_t in = s_ring[pos];
            s_cnt[in] = 0;
            s_in_cleared = in;
            s_in_list = c->list_id[in];
            if (s_cnt[s_out] > 0) {
                s_ring[pos] = s_out;
                s_out = in;
                m = MODE_RELAX;
            } else {
                newcur = cur + 1;
            }
        } else {
            // the overflow bucket's surviving chunks were listed in the spare
            // storage: it becomes the overflow list, the old one the spare
            const uint32_t old = c->list_id[s_ovf];
            c->list_id[s_ovf] = c->spare_list;
            c->spare_list = old;
            s_zero_list = old;
        }
        // (a redistribution moved what fits the window out of the overflow
        // bitmap in place; its count is already adjusted)
        int32_t tot = s_cnt[s_out];
        for (uint32_t p = 0; p < nb; ++p) tot += s_cnt[s_ring[p]];
        if (m == 0xFFFFFFFFu) {
            const int32_t ovf = s_cnt[s_ovf];
            if (tot == 0) {
                if (ovf == 0) {
                    m = MODE_IDLE;
                } else {
                    // nothing inside the window: jump to the lowest overflow
                    // bucket and redistribute
                    if (ovf_min > newcur) newcur = ovf_min;
                    m = MODE_REDIST;
                }
            } else if (mode == MODE_RELAX && ovf > 0 && newcur % nb == 0) {
                // the window slid by a full ring: pull overflow items into it
                m = MODE_REDIST;
            } else {
                m = MODE_RELAX;
            }
        }
        if (m == MODE_REDIST) ovf_min = ~0ull;
        s_mode = m;
        s_newcur = newcur;
        c->items += s_items;
        c->edges += s_edges;
        c->ovf_min = ovf_min;
        c->out_buf = s_out;
        c->cur = newcur;
        c->rounds++;
    }
    __syncthreads();
    const uint64_t newcur = s_newcur;
    const uint32_t m = s_mode;
    if (threadIdx.x < nbuf) c->count[threadIdx.x] = (uint32_t)s_cnt[threadIdx.x];
    if (threadIdx.x < nb) c->ring[threadIdx.x] = s_ring[threadIdx.x];
    // the consumed bucket: its round cleared the summary words of its listed
    // chunks; empty its list
    if (s_in_cleared != 0xFFFFFFFFu && threadIdx.x < LIST_BINS)
        a.list_cnt[(uint64_t)s_in_list * LIST_BINS + threadIdx.x] = 0u;
    if (s_zero_list != 0xFFFFFFFFu && threadIdx.x < LIST_BINS)
        a.list_cnt[(uint64_t)s_zero_list * LIST_BINS + threadIdx.x] = 0u;
    // slot completion: dst settled (its bucket is below the new current
    // bucket), or no pending items left (dst unreachable)
    for (int s = threadIdx.x; s < a.K; s += CTRL_WG) {
        SlotInfo *si = &a.slots[s];
        if (si->state != 1) continue;
        // no set bits left: the search is exhausted and every key final
        const bool exhausted = si->pending <= 0;
        // goal-directed order settles the destination one bucket later: the
        // margin absorbs FP32 rounding of g + h on its equal-cost predecessors
        const uint64_t margin = (a.h_xy != 0.0f || a.h_z != 0.0f) ? 1u : 0u;
        if (si->ms) {
            // A multi-source search is done when its frontier is empty, or
            // once the current bucket has passed the bucket of the best
            // (keepout edges, dist + target cost) among its reached targets
            // (at most MS_EARLY_TARGETS are scanned): every node with a dist
            // up to that total is then settled, so no cheaper or equal target
            // is left to find and the keys on the paths are final.  The
            // backtrace picks the target.
            bool done = exhausted;
            if (!done && si->t_n <= MS_EARLY_TARGETS) {
                uint64_t best = ~0ull;
                const uint64_t *ks = a.keys + (uint64_t)s * (uint64_t)a.NP;
                for (int t = 0; t < si->t_n; ++t) {
                    const orgpu_ms_point tp = a.ms_tgt[si->t_off + t];
                    const uint64_t k = ks[tp.node];
                    if (k == KEY_INF) continue;
                    // float(dist + cost) >= every float dist <= that total
                    const float tot = (float)((double)key_dist(k) + (double)tp.cost);
                    const uint64_t b = key_bucket_rel(key_pack(key_ko(k), tot, 0u), a.inv_delta);
                    best = b < best ? b : best;
                }
                done = best != ~0ull && si->base + best + margin < newcur;
            }
            if (done) { si->state = 2; atomicSub(&c->active, 1u); }
            continue;
        }
        const uint64_t kd = a.keys[(uint64_t)s * (uint64_t)a.NP + (uint64_t)si->dst];
        const bool reached = kd != KEY_INF;
        if (reached && (exhausted || si->base + key_bucket_rel(kd, a.inv_delta) + margin < newcur)) {
            si->state = 2; atomicSub(&c->active, 1u);
        } else if (exhausted) {
            si->state = 3; atomicSub(&c->active, 1u);
        }
    }
    if (threadIdx.x == 0) c->mode = m;
}

// ---------------------------------------------------------------------------
// slot management
// ---------------------------------------------------------------------------

__global__ void k_fill_keys(uint64_t *keys, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        keys[i] = KEY_INF;
}

struct SeedReq {
    int32_t slot, src, dst, net_id, roi_index, use_keepout, req, ms;
    int32_t s_off, s_n, t_off, t_n;   // multi-source: seeds / targets in the call's arrays
};

// Resets the keys of the slots about to be seeded: only the chunks their
// previous searches touched (each workgroup fills whole chunks), and clears
// those chunks' touch bits (whole words: every set bit of the slot's record
// is a listed chunk).  k_seed, which runs after this, restarts the lists.
// Grid (G, n): request blockIdx.y gets G workgroups that stride over the
// chunks its slot's previous search touched (a few slots with many chunks
// and many slots with few both spread over the chip).
__global__ void k_reset_keys(RoundArgs a, const SeedReq *reqs, int n)
{
    const uint64_t N = (uint64_t)a.N;
    const uint64_t NP = (uint64_t)a.NP;
    const int r = (int)blockIdx.y;
    if (r >= n) return;
    const uint32_t s = (uint32_t)reqs[r].slot;
    const uint32_t cnt = a.touch_n[s];
    const uint32_t *lst = a.touch_list + (uint64_t)s * a.touch_cap;
    uint64_t *ks = a.keys + (uint64_t)s * NP;
    uint32_t *tb = a.touch_bits + (uint64_t)s * a.touch_words;
    for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {
        const uint32_t ch = lst[i];
        const uint64_t lo = (uint64_t)ch * TOUCH_KEYS;
        const uint64_t hi = lo + TOUCH_KEYS < N ? lo + TOUCH_KEYS : N;
        for (uint64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) ks[v] = KEY_INF;
        if (threadIdx.x == 0) tb[ch >> 5] = 0u;
    }
}

// Marks finished slots free (their queued items, if any, keep draining).
__global__ void k_release(RoundArgs a, const int32_t *list, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SlotInfo *si = &a.slots[list[i]];
    si->state = 0;
    if (si->ms) {
        uint32_t *row = a.tbits + (uint64_t)list[i] * a.tbits_words;
        for (int t = 0; t < si->t_n; ++t) {
            const uint32_t v = (uint32_t)a.ms_tgt[si->t_off + t].node;
            atomicAnd(&row[v >> 5], ~(1u << (v & 31)));
        }
        si->ms = 0;
    }
}

__global__ void k_seed(RoundArgs a, const SeedReq *reqs, int n, const orgpu_ms_point *ms_seed)
{
    Ctrl *c = a.ctrl;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const SeedReq r = reqs[i];
        SlotInfo *si = &a.slots[r.slot];
        si->src = r.src; si->dst = r.dst; si->net_id = r.net_id;
        si->roi_index = r.roi_index; si->use_keepout = r.use_keepout;
        si->req = r.req;
        si->ms = r.ms; si->t_off = r.t_off; si->t_n = r.t_n;
        if (a.plane > 0 && a.X > 0 && !r.ms) {
            const int z = r.dst / a.plane, rr = r.dst - z * a.plane;
            si->dz = z; si->dy = rr / a.X; si->dx = rr - (rr / a.X) * a.X;
        } else {
            si->dx = si->dy = si->dz = 0;
        }
        // the slot's keys and touch bits were reset (k_reset_keys): start a
        // new touch list
        a.touch_n[r.slot] = 0u;
        // a slot is reseeded only once its previous search has no set bits
        // left, so its items' bits are clear here
        if (!r.ms) {
            // the source (g = 0) lands in the current bucket: its bucket under
            // the goal-directed order, floor(h(src) / delta), is taken off the
            // slot's base (the bucket index starts at 2^32, so this never wraps)
            si->base = c->cur - key_bucket_rel(key_pack(0, 0.0f, KEY_NOPRED), a.inv_delta, heur(a, *si, (uint32_t)r.src));
            si->state = 1;
            const uint64_t item = (uint64_t)r.slot * (uint64_t)a.NP + (uint64_t)r.src;
            a.keys[item] = key_pack(0, 0.0f, KEY_NOPRED);
            const uint32_t buf = c->ring[c->cur % a.nb];
            si->pending = 1;
            touch_key(a, (uint32_t)r.slot, (uint32_t)r.src);
            if (push_bit(a, buf, c->list_id[buf], item)) atomicAdd(&c->count[buf], 1u);
        } else {
            si->base = c->cur;
            si->state = 1;
            uint32_t *trow = a.tbits + (uint64_t)r.slot * a.tbits_words;
            for (int t = 0; t < r.t_n; ++t) {
                const uint32_t v = (uint32_t)a.ms_tgt[r.t_off + t].node;
                atomicOr(&trow[v >> 5], 1u << (v & 31));
            }
            const uint32_t *roi = r.roi_index >= 0 ? a.roi + (uint64_t)r.roi_index * a.roi_words : nullptr;
            int pend = 0;
            for (int t = 0; t < r.s_n; ++t) {
                const orgpu_ms_point sp = ms_seed[r.s_off + t];
                const uint32_t v = (uint32_t)sp.node;
                if (roi && !((roi[v >> 5] >> (v & 31)) & 1u)) continue;
                const uint64_t item = (uint64_t)r.slot * (uint64_t)a.NP + v;
                const uint64_t k = key_pack(0, sp.cost, KEY_NOPRED);
                const uint64_t old = atomicMin((unsigned long long *)&a.keys[item], (unsigned long long)k);
                if (old == KEY_INF) touch_key(a, (uint32_t)r.slot, v);
                if (k < old && key_cost(k) < key_cost(old)) {
                    const uint64_t b = si->base + key_bucket_rel(k, a.inv_delta);
                    uint32_t buf;
                    if (b < c->cur + a.nb) {
                        buf = c->ring[b % a.nb];
                    } else {
                        buf = c->ovf_buf;
                        atomicMin((unsigned long long *)&c->ovf_min, (unsigned long long)b);
                    }
                    if (push_bit(a, buf, c->list_id[buf], item)) { atomicAdd(&c->count[buf], 1u); ++pend; }
                }
            }
            si->pending = pend;
        }
        atomicAdd(&c->active, 1u);
    }
    __syncthreads();
    if (i == 0) {
        __threadfence();
        c->mode = c->mode == MODE_IDLE ? MODE_RELAX : c->mode;
    }
}

// One wave per finished slot.  Lane 0 picks the end node (the destination,
// or the chosen target of a multi-source search) and walks the predecessor
// fields back to the source into the staging row (one dependent key load per
// hop); then the lanes find each hop's CSR edge id (first match in the sorted
// row, UPF:5463-5476) and pack the path src -> dst into the shared output.
// len = 0: unreachable; -1: broken chain or longer than the staging row;
// -2: the path crosses KEY_KO_MAX or more keepout edges (the key's count stops there).
__global__ void __launch_bounds__(64)
k_backtrace(RoundArgs a, const int32_t *slot_list, int n, int32_t *stage_nodes, int32_t *stage_edges,
            int32_t *lens, float *dists, int32_t max_len, double *ms_total, int32_t *ms_choice,
            int32_t *pack, uint32_t *pack_n, uint32_t *pack_off, int64_t pack_half)
{
    __shared__ int s_len;
    __shared__ uint32_t s_off;
    const int i = blockIdx.x;
    if (i >= n) return;
    const int lane = (int)threadIdx.x;
    const int s = slot_list[i];
    const SlotInfo si = a.slots[s];
    int32_t *pn = stage_nodes + (int64_t)s * max_len;
    int32_t *pe = stage_edges + (int64_t)s * max_len;
    const uint64_t *ks = a.keys + (uint64_t)s * (uint64_t)a.NP;
    if (lane == 0) {
        int len = 0;
        int32_t end = si.dst;
        bool found = si.state == 2;
        if (found && si.ms) {
            // the cheapest target by (keepout edges, dist + target cost), then
            // dist, node index, entry order
            int best = -1;
            uint32_t bko = 0;
            double bt = 0.0;
            float bd = 0.0f;
            int32_t bn = 0;
            for (int t = 0; t < si.t_n; ++t) {
                const orgpu_ms_point tp = a.ms_tgt[si.t_off + t];
                const uint64_t k = ks[tp.node];
                if (k == KEY_INF) continue;
                const uint32_t ko = key_ko(k);
                const float d = key_dist(k);
                const double tot = (double)d + (double)tp.cost;
                const bool better = best < 0 || ko < bko ||
                                    (ko == bko && (tot < bt || (tot == bt && (d < bd || (d == bd && tp.node < bn)))));
                if (better) { best = t; bko = ko; bt = tot; bd = d; bn = tp.node; }
            }
            if (ms_choice) ms_choice[s] = best;
            if (ms_total) ms_total[s] = best >= 0 ? bt : (double)__builtin_inff();
            found = best >= 0;
            end = bn;
        }
        if (!found) {
            dists[s] = __builtin_inff();
        } else if (key_ko(ks[end]) >= KEY_KO_MAX) {
            // the key's keepout count stops at KEY_KO_MAX: past it, paths would
            // no longer be ordered by their number of keepout edges
            dists[s] = key_dist(ks[end]);
            len = -2;
        } else {
            dists[s] = key_dist(ks[end]);
            int32_t cur = end;
            while (true) {
                if (len >= max_len) { len = -1; break; }
                pn[len++] = cur;
                if (!si.ms && cur == si.src) break;
                const uint32_t pf = key_pred(ks[cur]);
                if (pf == KEY_NOPRED) {
                    if (!si.ms) len = -1;   // a two-pin chain must end at its source
                    break;                  // a multi-source chain ends at a seed
                }
                cur = (int32_t)(pf - 1u);
            }
        }
        lens[s] = len;
        s_len = len;
        if (len > 0) {
            const uint32_t off = atomicAdd(pack_n, (uint32_t)len);
            pack_off[s] = off;
            s_off = off;
        }
    }
    __syncthreads();
    const int len = s_len;
    if (len <= 0) return;
    // edge entering pn[t] (from pn[t + 1]): first CSR edge pn[t+1] -> pn[t]
    for (int t = lane; t + 1 < len; t += 64) {
        const int32_t v = pn[t], u = pn[t + 1];
        const int r0 = a.indptr[u], r1 = a.indptr[u + 1];
        int ee = -1;
        for (int e = r0; e < r1; ++e)
            if (a.indices[e] == v) { ee = e; break; }
        pe[t] = ee;
    }
    __syncthreads();
    // pack src -> dst: node t is pn[len-1-t], the edge leaving it pe[len-2-t]
    int32_t *on = pack + s_off;
    int32_t *oe = pack + pack_half + s_off;
    for (int t = lane; t < len; t += 64) {
        on[t] = pn[len - 1 - t];
        oe[t] = t + 1 < len ? pe[len - 2 - t] : -1;
    }
}

// Optional Dijkstra-minimal work count: nodes whose (keepout, distance) is at
// most the destination's, and the sum of their out-degrees.
__global__ void k_count_minimal(RoundArgs a, int slot, unsigned long long *out)
{
    const SlotInfo si = a.slots[slot];
    const uint64_t *ks = a.keys + (uint64_t)slot * (uint64_t)a.NP;
    const uint64_t lim = key_cost(ks[si.dst]);
    unsigned long long nn = 0, ne = 0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < a.N; v += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = ks[v];
        if (k != KEY_INF && key_cost(k) <= lim) { nn++; ne += (unsigned long long)(a.indptr[v + 1] - a.indptr[v]); }
    }
    for (int o = 32; o > 0; o >>= 1) { nn += __shfl_down(nn, o, 64); ne += __shfl_down(ne, o, 64); }
    if (lane_id() == 0) { atomicAdd(&out[0], nn); atomicAdd(&out[1], ne); }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
#include "engine.h"

#define MAX_GRID 8192

static int launch_round(hipStream_t s, const RoundArgs &a, int grid)
{
    hipLaunchKernelGGL(k_round, dim3(grid), dim3(WG), 0, s, a);
    hipLaunchKernelGGL(k_control, dim3(1), dim3(CTRL_WG), 0, s, a, grid);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// rpp rounds on the engine stream: launch by launch (mode 0), or as one
// HIP graph of the rpp round pairs (mode 1), captured on a private stream
// the first time and again whenever the round arguments change.  The graph
// holds kernel nodes only.
static int launch_rounds(Engine &en, const RoundArgs &a, int grid, int rpp, int mode)
{
    if (mode != 1) {
        for (int r = 0; r < rpp; ++r) launch_round(en.stream, a, grid);
        HIP_CHECK_RET(hipGetLastError());
        return ORGPU_OK;
    }
    if (!en.round_exec || en.round_exec_rpp != rpp || en.round_exec_grid != grid ||
        memcmp(en.round_exec_args, &a, sizeof(a)) != 0) {
        if (en.round_exec) { (void)hipGraphExecDestroy(en.round_exec); en.round_exec = nullptr; }
        if (!en.capture_stream) HIP_CHECK_RET(hipStreamCreateWithFlags(&en.capture_stream, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        HIP_CHECK_RET(hipStreamBeginCapture(en.capture_stream, hipStreamCaptureModeThreadLocal));
        for (int r = 0; r < rpp; ++r) launch_round(en.capture_stream, a, grid);
        const hipError_t le = hipGetLastError();
        HIP_CHECK_RET(hipStreamEndCapture(en.capture_stream, &g));
        if (le != hipSuccess) { (void)hipGraphDestroy(g); return orgpu_fail_hip(le, "round capture", __FILE__, __LINE__); }
        const hipError_t ie = hipGraphInstantiate(&en.round_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) { en.round_exec = nullptr; return orgpu_fail_hip(ie, "hipGraphInstantiate", __FILE__, __LINE__); }
        memcpy(en.round_exec_args, &a, sizeof(a));
        en.round_exec_rpp = rpp;
        en.round_exec_grid = grid;
    }
    HIP_CHECK_RET(hipGraphLaunch(en.round_exec, en.stream));
    return ORGPU_OK;
}

int engine_reserve(Engine &en, int64_t n_nodes, int32_t slots, int32_t n_buckets, int64_t qcap, int32_t max_path)
{
    if (slots < 1 || slots > RESERVE_MAX_SLOTS)
        return orgpu_fail(ORGPU_EINVAL, "reserve: slots %d outside [1, %d]", slots, RESERVE_MAX_SLOTS);
    if (n_buckets < 1 || n_buckets > MAX_NB) return orgpu_fail(ORGPU_EINVAL, "reserve: n_buckets %d outside [1, %d]", n_buckets, MAX_NB);
    if (n_nodes < 1 || n_nodes >= KEY_MAX_NODES) return orgpu_fail(ORGPU_EINVAL, "reserve: bad node count");
    if (max_path < 2) max_path = 1 << 16;
    // work items slot * NP + node: NP is n_nodes rounded up to whole chunks.
    // Keys, touch records and path staging are kept for every slot; the
    // bucket-round engine's bitmaps and slot records for at most MAX_SLOTS
    // of them (the one-wave engine needs none of those).
    const uint64_t np = ((uint64_t)n_nodes + CHUNK_ITEMS - 1) / CHUNK_ITEMS * CHUNK_ITEMS;
    const int32_t bslots = std::min(slots, MAX_SLOTS);
    const uint64_t keys_n = np * (uint64_t)slots;
    const uint64_t items = np * (uint64_t)bslots;
    if (keys_n / CHUNK_ITEMS >= (1ull << 32))
        return orgpu_fail(ORGPU_EINVAL, "reserve: slots*nodes = %llu exceeds 2^43 work items", (unsigned long long)keys_n);
    (void)qcap;   // frontiers are bitmaps: no item queues to size
    if (en.cap_nodes >= n_nodes && en.cap_slots >= slots && en.cap_nb >= n_buckets &&
        en.cap_path >= max_path)
        return ORGPU_OK;
    engine_release(en);
    const uint64_t words = (items + 31) / 32;
    const uint64_t chunks = (words + 63) / 64;
    const uint64_t summ_words = (chunks + 31) / 32;
    const int nbufs = n_buckets + 3;
    // keys are allocated by the first search call, at the stride its engine
    // lays them out with (engine_keys_alloc): a handle whose first call runs
    // the one-wave engine allocates them once, at that engine's tiled stride
    en.keys = nullptr;
    en.keys_cap = 0;
    en.keys_n = -1;
    HIP_CHECK_RET(hipMalloc((void **)&en.bits, (uint64_t)nbufs * words * sizeof(uint32_t)));
    HIP_CHECK_RET(hipMemsetAsync(en.bits, 0, (uint64_t)nbufs * words * sizeof(uint32_t), en.stream));
    HIP_CHECK_RET(hipMalloc((void **)&en.summ, (uint64_t)nbufs * summ_words * sizeof(uint32_t)));
    const uint64_t bin_cap = (chunks + LIST_BINS - 1) / LIST_BINS + 1;
    HIP_CHECK_RET(hipMalloc((void **)&en.lists, (uint64_t)(nbufs + 1) * LIST_BINS * bin_cap * sizeof(uint32_t)));
    HIP_CHECK_RET(hipMalloc((void **)&en.list_cnt, (uint64_t)(nbufs + 1) * LIST_BINS * sizeof(uint32_t)));
    {
        const uint64_t tcap = ((uint64_t)n_nodes + TOUCH_KEYS - 1) / TOUCH_KEYS, twords = (tcap + 31) / 32;
        HIP_CHECK_RET(hipMalloc((void **)&en.touch_bits, (uint64_t)slots * twords * sizeof(uint32_t)));
        HIP_CHECK_RET(hipMalloc((void **)&en.touch_list, (uint64_t)slots * tcap * sizeof(uint32_t)));
        HIP_CHECK_RET(hipMalloc((void **)&en.touch_n, (uint64_t)slots * sizeof(uint32_t)));
    }
    HIP_CHECK_RET(hipMalloc((void **)&en.parts, (size_t)MAX_GRID * sizeof(RoundPart)));
    HIP_CHECK_RET(hipMalloc((void **)&en.ctrl, sizeof(Ctrl)));
    HIP_CHECK_RET(hipMalloc((void **)&en.slots, sizeof(SlotInfo) * slots));
    HIP_CHECK_RET(hipMalloc((void **)&en.seed, sizeof(SeedReq) * slots));
    HIP_CHECK_RET(hipMalloc((void **)&en.slot_list, sizeof(int32_t) * slots));
    HIP_CHECK_RET(hipMalloc((void **)&en.stage_nodes, (size_t)slots * max_path * sizeof(int32_t)));
    HIP_CHECK_RET(hipMalloc((void **)&en.stage_edges, (size_t)slots * max_path * sizeof(int32_t)));
    HIP_CHECK_RET(hipMalloc((void **)&en.pack, (size_t)2 * slots * max_path * sizeof(int32_t)));
    HIP_CHECK_RET(hipMalloc((void **)&en.pack_n, sizeof(uint32_t)));
    HIP_CHECK_RET(hipMalloc((void **)&en.pack_off, sizeof(uint32_t) * slots));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_pack_off, sizeof(uint32_t) * slots, hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_pack_n, sizeof(uint32_t), hipHostMallocDefault));
    HIP_CHECK_RET(hipMalloc((void **)&en.lens, sizeof(int32_t) * slots));
    HIP_CHECK_RET(hipMalloc((void **)&en.dists, sizeof(float) * slots));
    HIP_CHECK_RET(hipMalloc((void **)&en.counters, sizeof(unsigned long long) * 2));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_slots, sizeof(SlotInfo) * slots, hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_seed, sizeof(SeedReq) * slots, hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_slot_list, sizeof(int32_t) * slots, hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_lens, sizeof(int32_t) * slots, hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_dists, sizeof(float) * slots, hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_stage, (size_t)2 * slots * max_path * sizeof(int32_t), hipHostMallocDefault));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_counters, sizeof(unsigned long long) * 2, hipHostMallocDefault));
    HIP_CHECK_RET(hipEventCreate(&en.ev_poll));
    HIP_CHECK_RET(hipEventCreate(&en.ev_t0));
    HIP_CHECK_RET(hipEventCreate(&en.ev_t1));
    HIP_CHECK_RET(hipEventCreate(&en.ev_r0));
    HIP_CHECK_RET(hipEventCreate(&en.ev_r1));
    en.cap_nodes = n_nodes; en.cap_slots = slots; en.cap_bslots = bslots; en.cap_nb = n_buckets; en.cap_path = max_path;
    en.words = words;
    en.summ_words = summ_words;
    en.chunks = chunks;
    return ORGPU_OK;
}

void engine_release(Engine &en)
{
    solo_release(en);
    if (en.keys) hipFree(en.keys);
    if (en.bits) hipFree(en.bits);
    if (en.summ) hipFree(en.summ);
    if (en.lists) hipFree(en.lists);
    if (en.list_cnt) hipFree(en.list_cnt);
    if (en.touch_bits) hipFree(en.touch_bits);
    if (en.touch_list) hipFree(en.touch_list);
    if (en.touch_n) hipFree(en.touch_n);
    if (en.ms_seed) hipFree(en.ms_seed);
    if (en.ms_tgt) hipFree(en.ms_tgt);
    if (en.tbits) hipFree(en.tbits);
    if (en.ms_total) hipFree(en.ms_total);
    if (en.ms_choice) hipFree(en.ms_choice);
    if (en.h_ms_total) hipHostFree(en.h_ms_total);
    if (en.h_ms_choice) hipHostFree(en.h_ms_choice);
    if (en.parts) hipFree(en.parts);
    if (en.ctrl) hipFree(en.ctrl);
    if (en.slots) hipFree(en.slots);
    if (en.seed) hipFree(en.seed);
    if (en.slot_list) hipFree(en.slot_list);
    if (en.stage_nodes) hipFree(en.stage_nodes);
    if (en.pack) hipFree(en.pack);
    if (en.pack_n) hipFree(en.pack_n);
    if (en.pack_off) hipFree(en.pack_off);
    if (en.h_pack_off) hipHostFree(en.h_pack_off);
    if (en.h_pack_n) hipHostFree(en.h_pack_n);
    if (en.stage_edges) hipFree(en.stage_edges);
    if (en.lens) hipFree(en.lens);
    if (en.dists) hipFree(en.dists);
    if (en.counters) hipFree(en.counters);
    if (en.h_ctrl) hipHostFree(en.h_ctrl);
    if (en.h_slots) hipHostFree(en.h_slots);
    if (en.h_seed) hipHostFree(en.h_seed);
    if (en.h_slot_list) hipHostFree(en.h_slot_list);
    if (en.h_lens) hipHostFree(en.h_lens);
    if (en.h_dists) hipHostFree(en.h_dists);
    if (en.h_stage) hipHostFree(en.h_stage);
    if (en.h_counters) hipHostFree(en.h_counters);
    if (en.ev_poll) hipEventDestroy(en.ev_poll);
    if (en.ev_t0) hipEventDestroy(en.ev_t0);
    if (en.ev_t1) hipEventDestroy(en.ev_t1);
    if (en.ev_r0) hipEventDestroy(en.ev_r0);
    if (en.ev_r1) hipEventDestroy(en.ev_r1);
    if (en.round_exec) hipGraphExecDestroy(en.round_exec);
    if (en.capture_stream) hipStreamDestroy(en.capture_stream);
    if (en.sp_idx) hipFree(en.sp_idx);
    if (en.sp_val) hipFree(en.sp_val);
    if (en.sp_count) hipFree(en.sp_count);
    if (en.h_sp_count) hipHostFree(en.h_sp_count);
    if (en.up_buf) hipFree(en.up_buf);
    if (en.roi_b) hipFree(en.roi_b);
    if (en.roi_la) hipFree(en.roi_la);
    if (en.roi_lb) hipFree(en.roi_lb);
    if (en.roi_req) hipFree(en.roi_req);
    if (en.roi_info) hipFree(en.roi_info);
    if (en.roi_next) hipFree(en.roi_next);
    if (en.roi_T) hipFree(en.roi_T);
    Engine fresh;
    fresh.stream = en.stream;
    fresh.device = en.device;
    fresh.roi_limit = en.roi_limit;
    // the reference-order search scratch has its own reserve / release
    fresh.ref_state = en.ref_state; fresh.ref_heap = en.ref_heap; fresh.ref_ctl = en.ref_ctl;
    fresh.ref_nodes = en.ref_nodes; fresh.ref_hcap = en.ref_hcap; fresh.ref_slots = en.ref_slots;
    fresh.ref_sid = en.ref_sid;
    en = fresh;
}

// Resets the chunks listed in the touch records of every slot (grid (G,
// slots): slot blockIdx.y) and clears their touch bits.
__global__ void k_flush_touched(RoundArgs a)
{
    const uint64_t N = (uint64_t)a.N;
    const uint32_t s = blockIdx.y;
    const uint32_t cnt = a.touch_n[s];
    const uint32_t *lst = a.touch_list + (uint64_t)s * a.touch_cap;
    uint64_t *ks = a.keys + (uint64_t)s * (uint64_t)a.NP;
    uint32_t *tb = a.touch_bits + (uint64_t)s * a.touch_words;
    for (uint32_t i = blockIdx.x; i < cnt; i += gridDim.x) {
        const uint32_t ch = lst[i];
        const uint64_t lo = (uint64_t)ch * TOUCH_KEYS;
        const uint64_t hi = lo + TOUCH_KEYS < N ? lo + TOUCH_KEYS : N;
        for (uint64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) ks[v] = KEY_INF;
        if (threadIdx.x == 0) tb[ch >> 5] = 0u;
    }
}

// The key buffer holds at least `need` keys (grown when smaller; every key
// is refilled with KEY_INF by the next call that uses them)
int engine_keys_alloc(Engine &en, uint64_t need)
{
    if (need <= (uint64_t)en.keys_cap && en.keys) return ORGPU_OK;
    if (en.keys) hipFree(en.keys);
    en.keys = nullptr;
    en.keys_cap = 0;
    HIP_CHECK_RET(hipMalloc((void **)&en.keys, (size_t)need * sizeof(uint64_t)));
    en.keys_cap = (int64_t)need;
    en.keys_n = -1;
    return ORGPU_OK;
}

int engine_prepare_keys(Engine &en, int64_t N)
{
    hipStream_t s = en.stream;
    if (N <= 0 || N > en.cap_nodes) return orgpu_fail(ORGPU_EINVAL, "prepare_keys: node count outside the reserve");
    RoundArgs a;
    memset(&a, 0, sizeof(a));
    a.N = N;
    a.NP = (N + CHUNK_ITEMS - 1) / CHUNK_ITEMS * CHUNK_ITEMS;
    a.keys = en.keys;
    a.touch_cap = (N + TOUCH_KEYS - 1) / TOUCH_KEYS;
    a.touch_words = (a.touch_cap + 31) / 32;
    a.touch_bits = en.touch_bits; a.touch_list = en.touch_list; a.touch_n = en.touch_n;
    if (en.keys_n != N) {
        hipLaunchKernelGGL(k_fill_keys, dim3(4096), dim3(256), 0, s, en.keys, en.keys_cap);
        HIP_CHECK_RET(hipGetLastError());
        en.keys_n = N;
    } else if (en.touch_dirty) {
        hipLaunchKernelGGL(k_flush_touched, dim3(64, en.cap_slots), dim3(256), 0, s, a);
        HIP_CHECK_RET(hipGetLastError());
    } else {
        return ORGPU_OK;
    }
    HIP_CHECK_RET(hipMemsetAsync(en.touch_n, 0, sizeof(uint32_t) * en.cap_slots, s));
    HIP_CHECK_RET(hipMemsetAsync(en.touch_bits, 0, sizeof(uint32_t) * (size_t)en.cap_slots * a.touch_words, s));
    en.touch_dirty = false;
    return ORGPU_OK;
}

int engine_route(Engine &en, const orgpu_graph *g, const orgpu_net_req *nets, int32_t n_nets,
                 const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner, float keepout_cost,
                 const orgpu_sssp_params *p,
                 int32_t *path_nodes, int32_t *path_edges, int64_t path_capacity, int64_t *path_offset,
                 float *path_dist, orgpu_sssp_stats *stats, const MsCall *ms)
{
    if (!g || !p || (n_nets > 0 && !nets && !ms) || !path_offset) return orgpu_fail(ORGPU_EINVAL, "route: null argument");
    if (!g->indptr || !g->indices || !g->cost || !g->edge_kind) return orgpu_fail(ORGPU_EINVAL, "route: null graph array");
    const int64_t N = g->n_nodes;
    const int K = p->slots;
    if (N <= 0 || N > en.cap_nodes || K < 1 || K > en.cap_slots || p->n_buckets < 1 || p->n_buckets > en.cap_nb)
        return orgpu_fail(ORGPU_EINVAL, "route: request (N=%lld, K=%d, nb=%d) outside the reserved capacity "
                          "(N=%lld, K=%d, nb=%d)", (long long)N, K, p->n_buckets, (long long)en.cap_nodes,
                          en.cap_slots, en.cap_nb);
    if (!(p->delta > 0.0f)) return orgpu_fail(ORGPU_EINVAL, "route: delta must be > 0");
    if (p->local_batches < 0 || p->local_batches > 4096)
        return orgpu_fail(ORGPU_EINVAL, "route: local_batches %d outside [0, 4096]", p->local_batches);
    if (!(p->h_xy >= 0.0f && p->h_z >= 0.0f && p->h_xy < 3.0e38f && p->h_z < 3.0e38f))
        return orgpu_fail(ORGPU_EINVAL, "route: h_xy %g / h_z %g must be finite and >= 0", (double)p->h_xy, (double)p->h_z);
    if ((p->h_xy > 0.0f || p->h_z > 0.0f) && (g->X <= 0 || g->Y <= 0 || g->L <= 0 ||
                                              (int64_t)g->X * g->Y * g->L != g->n_nodes))
        return orgpu_fail(ORGPU_EINVAL, "route: goal-directed search needs the lattice shape (X*Y*L == n_nodes)");
    // keepout edges are ranked by count ahead of the remaining cost, which is
    // what a penalty far above every path cost (the reference's 1e9) does
    if (keepout_owner && !(keepout_cost >= 1.0e6f))
        return orgpu_fail(ORGPU_EINVAL, "route: keepout_cost %g must dominate path costs (>= 1e6)", (double)keepout_cost);
    for (int i = 0; i < n_nets && !ms; ++i) {
        const orgpu_net_req &r = nets[i];
        if (r.src < 0 || r.src >= N || r.dst < 0 || r.dst >= N)
            return orgpu_fail(ORGPU_EINVAL, "route: request %d endpoint outside [0, %lld)", i, (long long)N);
        if (r.roi_index >= 0 && (!roi_bitmaps || r.roi_index >= n_roi))
            return orgpu_fail(ORGPU_EINVAL, "route: request %d ROI row %d outside [0, %d)", i, r.roi_index, n_roi);
        if (r.use_keepout && !keepout_owner) return orgpu_fail(ORGPU_EINVAL, "route: request %d needs keepout_owner", i);
    }
    if (p->engine != 0 && p->engine != 1) return orgpu_fail(ORGPU_EINVAL, "route: engine %d is not 0 or 1", p->engine);
    if (p->list_cap < 0) return orgpu_fail(ORGPU_EINVAL, "route: list_cap %d < 0", p->list_cap);
    if (p->waves != 0 && p->waves != 1 && p->waves != 2 && p->waves != 4)
        return orgpu_fail(ORGPU_EINVAL, "route: waves %d is not 0, 1, 2 or 4", p->waves);
    if (p->engine == 1 && !ms && !p->count_minimal)
        return engine_route_solo(en, g, nets, n_nets, roi_bitmaps, n_roi, keepout_owner, keepout_cost, p, path_nodes,
                                 path_edges, path_capacity, path_offset, path_dist, stats);
    if (K > en.cap_bslots)
        return orgpu_fail(ORGPU_EINVAL, "route: %d slots on the bucket-round engine (at most %d)", K, en.cap_bslots);
    hipStream_t s = en.stream;
    if (ms) {
        if ((ms->n_seeds > 0 && !ms->seeds) || (ms->n_targets > 0 && !ms->targets) || (n_nets > 0 && !ms->reqs) ||
            ms->n_seeds < 0 || ms->n_targets < 0)
            return orgpu_fail(ORGPU_EINVAL, "route_ms: null argument");
        for (int64_t i = 0; i < ms->n_seeds; ++i) {
            const orgpu_ms_point &q = ms->seeds[i];
            if (q.node < 0 || q.node >= N || !(q.cost >= 0.0f && q.cost < 3.0e38f))
                return orgpu_fail(ORGPU_EINVAL, "route_ms: seed %lld (node %d, cost %g) invalid", (long long)i, q.node,
                                  (double)q.cost);
        }
        for (int64_t i = 0; i < ms->n_targets; ++i) {
            const orgpu_ms_point &q = ms->targets[i];
            if (q.node < 0 || q.node >= N || !(q.cost >= 0.0f && q.cost < 3.0e38f))
                return orgpu_fail(ORGPU_EINVAL, "route_ms: target %lld (node %d, cost %g) invalid", (long long)i, q.node,
                                  (double)q.cost);
        }
        for (int i = 0; i < n_nets; ++i) {
            const orgpu_ms_req &r = ms->reqs[i];
            if (r.seed_off < 0 || r.seed_n < 0 || (int64_t)r.seed_off + r.seed_n > ms->n_seeds ||
                r.target_off < 0 || r.target_n < 0 || (int64_t)r.target_off + r.target_n > ms->n_targets)
                return orgpu_fail(ORGPU_EINVAL, "route_ms: request %d seed/target range outside the arrays", i);
            if (r.roi_index >= 0 && (!roi_bitmaps || r.roi_index >= n_roi))
                return orgpu_fail(ORGPU_EINVAL, "route_ms: request %d ROI row %d outside [0, %d)", i, r.roi_index, n_roi);
            if (r.use_keepout && !keepout_owner) return orgpu_fail(ORGPU_EINVAL, "route_ms: request %d needs keepout_owner", i);
        }
        // device copies of the seeds / targets, per-slot target bitmaps and outputs
        const int64_t tw = (N + 31) / 32;
        if (en.ms_seed_cap < std::max<int64_t>(ms->n_seeds, 1)) {
            if (en.ms_seed) hipFree(en.ms_seed);
            en.ms_seed_cap = std::max<int64_t>(ms->n_seeds, 1024);
            HIP_CHECK_RET(hipMalloc((void **)&en.ms_seed, sizeof(orgpu_ms_point) * en.ms_seed_cap));
        }
        if (en.ms_tgt_cap < std::max<int64_t>(ms->n_targets, 1)) {
            if (en.ms_tgt) hipFree(en.ms_tgt);
            en.ms_tgt_cap = std::max<int64_t>(ms->n_targets, 1024);
            HIP_CHECK_RET(hipMalloc((void **)&en.ms_tgt, sizeof(orgpu_ms_point) * en.ms_tgt_cap));
        }
        if (en.tbits_cap < (int64_t)K * tw) {
            if (en.tbits) hipFree(en.tbits);
            en.tbits_cap = (int64_t)en.cap_slots * ((en.cap_nodes + 31) / 32);
            HIP_CHECK_RET(hipMalloc((void **)&en.tbits, sizeof(uint32_t) * en.tbits_cap));
        }
        if (!en.ms_total) {
            HIP_CHECK_RET(hipMalloc((void **)&en.ms_total, sizeof(double) * MAX_SLOTS));
            HIP_CHECK_RET(hipMalloc((void **)&en.ms_choice, sizeof(int32_t) * MAX_SLOTS));
            HIP_CHECK_RET(hipHostMalloc((void **)&en.h_ms_total, sizeof(double) * MAX_SLOTS, hipHostMallocDefault));
            HIP_CHECK_RET(hipHostMalloc((void **)&en.h_ms_choice, sizeof(int32_t) * MAX_SLOTS, hipHostMallocDefault));
        }
        if (ms->n_seeds > 0)
            HIP_CHECK_RET(hipMemcpyAsync(en.ms_seed, ms->seeds, sizeof(orgpu_ms_point) * ms->n_seeds,
                                         hipMemcpyHostToDevice, s));
        if (ms->n_targets > 0)
            HIP_CHECK_RET(hipMemcpyAsync(en.ms_tgt, ms->targets, sizeof(orgpu_ms_point) * ms->n_targets,
                                         hipMemcpyHostToDevice, s));
        HIP_CHECK_RET(hipMemsetAsync(en.tbits, 0, sizeof(uint32_t) * (size_t)K * tw, s));
    }
    const int nb = p->n_buckets;
    RoundArgs a;
    memset(&a, 0, sizeof(a));   // padding included: a captured round graph is reused when the bytes match
    a.ctrl = en.ctrl; a.slots = en.slots; a.K = K; a.N = N; a.nb = (uint32_t)nb;
    a.NP = (N + CHUNK_ITEMS - 1) / CHUNK_ITEMS * CHUNK_ITEMS;
    a.cps = (uint32_t)(a.NP / CHUNK_ITEMS);
    a.inv_delta = 1.0f / p->delta;
    a.h_xy = p->h_xy; a.h_z = p->h_z;
    a.local_batches = p->local_batches > 0 ? p->local_batches : 64;
    a.X = g->X; a.plane = g->X * g->Y;
    a.indptr = g->indptr; a.indices = g->indices; a.cost = g->cost; a.kind = g->edge_kind;
    a.roi = roi_bitmaps; a.roi_words = (N + 31) / 32;
    a.ko_owner = keepout_owner;
    a.keys = en.keys;
    a.bits = en.bits; a.bitmap_words = (uint64_t)a.NP * (uint64_t)K / 32;
    a.summ = en.summ; a.summ_words = ((a.bitmap_words + 63) / 64 + 31) / 32;
    a.ms_tgt = ms ? en.ms_tgt : nullptr;
    a.tbits = ms ? en.tbits : nullptr;
    a.tbits_words = (N + 31) / 32;
    a.touch_cap = (N + TOUCH_KEYS - 1) / TOUCH_KEYS;
    a.touch_words = (a.touch_cap + 31) / 32;
    a.touch_bits = en.touch_bits; a.touch_list = en.touch_list; a.touch_n = en.touch_n;
    a.lists = en.lists;
    a.list_cnt = en.list_cnt;
    a.chunks = (a.bitmap_words + 63) / 64;
    a.bin_cap = (a.chunks + LIST_BINS - 1) / LIST_BINS + 1;
    a.parts = en.parts;

    // control block init
    Ctrl hc;
    memset(&hc, 0, sizeof(hc));
    hc.mode = MODE_IDLE;
    hc.cur = 1ull << 32;   // room below for goal-directed slot bases
    hc.ovf_min = ~0ull;
    for (int i = 0; i < nb; ++i) hc.ring[i] = (uint32_t)i;
    hc.out_buf = (uint32_t)nb; hc.ovf_buf = (uint32_t)nb + 1; hc.ovf_next_buf = (uint32_t)nb + 2;
    for (int i = 0; i < nb + 3; ++i) hc.list_id[i] = (uint32_t)i;
    hc.spare_list = (uint32_t)nb + 3;
    *en.h_ctrl = hc;
    HIP_CHECK_RET(hipMemcpyAsync(en.ctrl, en.h_ctrl, sizeof(Ctrl), hipMemcpyHostToDevice, s));
    HIP_CHECK_RET(hipMemsetAsync(en.slots, 0, sizeof(SlotInfo) * K, s));
    // queue bitmaps must be clear: they are cleared as items are consumed, but
    // a previous call may have stopped with items left (finished slots)
    HIP_CHECK_RET(hipMemsetAsync(en.bits, 0, (size_t)(nb + 3) * a.bitmap_words * sizeof(uint32_t), s));
    HIP_CHECK_RET(hipMemsetAsync(en.summ, 0, (size_t)(nb + 3) * a.summ_words * sizeof(uint32_t), s));
    HIP_CHECK_RET(hipMemsetAsync(en.list_cnt, 0, (size_t)(nb + 4) * LIST_BINS * sizeof(uint32_t), s));
    // Keys are reset per slot from the touch records of its last search; those
    // records describe the layout slot*NP of the node count they were made for.
    {
        const int rk = engine_keys_alloc(en, (uint64_t)en.cap_slots * (uint64_t)a.NP);
        if (rk) return rk;
        a.keys = en.keys;
    }
    if (en.keys_n != N) {
        hipLaunchKernelGGL(k_fill_keys, dim3(4096), dim3(256), 0, s, en.keys, en.keys_cap);
        HIP_CHECK_RET(hipGetLastError());
        HIP_CHECK_RET(hipMemsetAsync(en.touch_n, 0, sizeof(uint32_t) * en.cap_slots, s));
        HIP_CHECK_RET(hipMemsetAsync(en.touch_bits, 0, sizeof(uint32_t) * (size_t)en.cap_slots * a.touch_words, s));
        en.keys_n = N;
    }

    int grid = p->grid > 0 ? p->grid : 1024;
    grid = std::min((grid + 7) / 8 * 8, MAX_GRID);
    const int rpp = p->rounds_per_poll > 0 ? p->rounds_per_poll : 32;
    std::vector<int> slot_req(K, -1);
    int next = 0, done_nets = 0;
    int64_t out_pos = 0;
    path_offset[0] = 0;
    std::vector<int64_t> lens(n_nets, 0), offs(n_nets, 0);
    unsigned long long min_nodes = 0, min_edges = 0;
    HIP_CHECK_RET(hipEventRecord(en.ev_t0, s));

    // slot_drained[k]: the slot's previous search has no queued items left
    std::vector<char> slot_drained(K, 1);
    int last_seeded = 0;
    auto seed_free_slots = [&]() -> int {
        int ns = 0;
        last_seeded = 0;
        for (int k = 0; k < K && next < n_nets; ++k) {
            if (slot_req[k] >= 0 || !slot_drained[k]) continue;
            if (ms) {
                const orgpu_ms_req &r = ms->reqs[next];
                en.h_seed[ns] = SeedReq{k, -1, -1, r.net_id, r.roi_index, r.use_keepout, next, 1,
                                        r.seed_off, r.seed_n, r.target_off, r.target_n};
            } else {
                const orgpu_net_req &r = nets[next];
                en.h_seed[ns] = SeedReq{k, r.src, r.dst, r.net_id, r.roi_index, r.use_keepout, next, 0, 0, 0, 0, 0};
            }
            slot_req[k] = next;
            ++next; ++ns;
        }
        if (ns == 0) return 0;
        en.touch_dirty = true;
        HIP_CHECK_RET(hipMemcpyAsync(en.seed, en.h_seed, sizeof(SeedReq) * ns, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_reset_keys, dim3(std::max(1, std::min(512, 2048 / ns)), ns), dim3(256), 0, s, a, en.seed, ns);
        hipLaunchKernelGGL(k_seed, dim3((ns + 255) / 256), dim3(256), 0, s, a, en.seed, ns, ms ? en.ms_seed : nullptr);
        HIP_CHECK_RET(hipGetLastError());
        last_seeded = ns;
        return 0;
    };
    int rc = seed_free_slots();
    if (rc) return rc;

    int64_t launches = 0;
    double round_ms = 0.0;
    while (done_nets < n_nets) {
        HIP_CHECK_RET(hipEventRecord(en.ev_r0, s));
        rc = launch_rounds(en, a, grid, rpp, p->launch_mode);
        if (rc) return rc;
        HIP_CHECK_RET(hipEventRecord(en.ev_r1, s));
        launches += rpp;
        HIP_CHECK_RET(hipMemcpyAsync(en.h_ctrl, en.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
        HIP_CHECK_RET(hipMemcpyAsync(en.h_slots, en.slots, sizeof(SlotInfo) * K, hipMemcpyDeviceToHost, s));
        HIP_CHECK_RET(hipEventRecord(en.ev_poll, s));
        HIP_CHECK_RET(hipEventSynchronize(en.ev_poll));
        {
            float bms = 0.0f;
            HIP_CHECK_RET(hipEventElapsedTime(&bms, en.ev_r0, en.ev_r1));
            round_ms += bms;
        }
        int nf = 0, searching = 0;
        for (int k = 0; k < K; ++k) {
            if (slot_req[k] < 0) {
                // free slot: reusable once the device reports no queued items
                slot_drained[k] = en.h_slots[k].pending <= 0;
                continue;
            }
            const uint32_t st = en.h_slots[k].state;
            if (st == 2 || st == 3) en.h_slot_list[nf++] = k;
            else ++searching;
        }
        if (en.h_ctrl->rounds > (1ull << 32))
            return orgpu_fail(ORGPU_EHIP, "route: no completion after %llu rounds",
                              (unsigned long long)en.h_ctrl->rounds);
        if (nf == 0 && searching > 0 && en.h_ctrl->mode == MODE_IDLE)
            return orgpu_fail(ORGPU_EHIP, "route: engine idle while %d searches are unfinished", searching);
        if (nf == 0) {
            if (searching == 0) {
                rc = seed_free_slots();
                if (rc) return rc;
                // the status snapshot predates this seeding; with nothing
                // seeded, nothing searching and the engine idle, the
                // remaining requests can never finish
                if (last_seeded == 0 && en.h_ctrl->mode == MODE_IDLE && next >= n_nets)
                    return orgpu_fail(ORGPU_EHIP, "route: engine idle with %d unfinished requests",
                                      n_nets - done_nets);
            }
            continue;
        }
        HIP_CHECK_RET(hipMemcpyAsync(en.slot_list, en.h_slot_list, sizeof(int32_t) * nf, hipMemcpyHostToDevice, s));
        HIP_CHECK_RET(hipMemsetAsync(en.pack_n, 0, sizeof(uint32_t), s));
        const int64_t half = (int64_t)en.cap_slots * en.cap_path;
        hipLaunchKernelGGL(k_backtrace, dim3(nf), dim3(64), 0, s, a, en.slot_list, nf,
                           en.stage_nodes, en.stage_edges, en.lens, en.dists, en.cap_path,
                           ms ? en.ms_total : nullptr, ms ? en.ms_choice : nullptr,
                           en.pack, en.pack_n, en.pack_off, half);
        HIP_CHECK_RET(hipGetLastError());
        HIP_CHECK_RET(hipMemcpyAsync(en.h_lens, en.lens, sizeof(int32_t) * K, hipMemcpyDeviceToHost, s));
        HIP_CHECK_RET(hipMemcpyAsync(en.h_dists, en.dists, sizeof(float) * K, hipMemcpyDeviceToHost, s));
        HIP_CHECK_RET(hipMemcpyAsync(en.h_pack_off, en.pack_off, sizeof(uint32_t) * K, hipMemcpyDeviceToHost, s));
        HIP_CHECK_RET(hipMemcpyAsync(en.h_pack_n, en.pack_n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        if (ms) {
            HIP_CHECK_RET(hipMemcpyAsync(en.h_ms_total, en.ms_total, sizeof(double) * K, hipMemcpyDeviceToHost, s));
            HIP_CHECK_RET(hipMemcpyAsync(en.h_ms_choice, en.ms_choice, sizeof(int32_t) * K, hipMemcpyDeviceToHost, s));
        }
        HIP_CHECK_RET(hipStreamSynchronize(s));
        for (int i = 0; i < nf; ++i) {
            const int k = en.h_slot_list[i];
            if (en.h_lens[k] == -2)
                return orgpu_fail(ORGPU_ECAPACITY, "route: request %d: its path crosses 255 or more keepout edges, "
                                  "which the search key cannot count", slot_req[k]);
            if (en.h_lens[k] < 0)
                return orgpu_fail(ORGPU_ECAPACITY, "route: request %d: predecessor chain broken or longer than %d",
                                  slot_req[k], en.cap_path);
        }
        // every finished path, packed src -> dst, comes back in one copy
        // (nodes at h_stage, edges at h_stage + half)
        const uint32_t packed = *en.h_pack_n;
        if (packed > 0) {
            HIP_CHECK_RET(hipMemcpyAsync(en.h_stage, en.pack, sizeof(int32_t) * packed, hipMemcpyDeviceToHost, s));
            if (path_edges)
                HIP_CHECK_RET(hipMemcpyAsync(en.h_stage + half, en.pack + half, sizeof(int32_t) * packed,
                                             hipMemcpyDeviceToHost, s));
            HIP_CHECK_RET(hipStreamSynchronize(s));
        }
        for (int i = 0; i < nf; ++i) {
            const int k = en.h_slot_list[i];
            const int req = slot_req[k];
            const int len = en.h_lens[k];
            const int32_t *h = en.h_stage + (len > 0 ? en.h_pack_off[k] : 0u);
            if (p->count_minimal && en.h_slots[k].state == 2 && !ms) {
                HIP_CHECK_RET(hipMemsetAsync(en.counters, 0, sizeof(unsigned long long) * 2, s));
                hipLaunchKernelGGL(k_count_minimal, dim3(512), dim3(256), 0, s, a, k, en.counters);
                HIP_CHECK_RET(hipMemcpyAsync(en.h_counters, en.counters, sizeof(unsigned long long) * 2,
                                             hipMemcpyDeviceToHost, s));
                HIP_CHECK_RET(hipStreamSynchronize(s));
                min_nodes += en.h_counters[0];
                min_edges += en.h_counters[1];
            }
            if (len > 0) {
                if (out_pos + len > path_capacity)
                    return orgpu_fail(ORGPU_ECAPACITY, "route: path output capacity %lld exceeded", (long long)path_capacity);
                memcpy(path_nodes + out_pos, h, sizeof(int32_t) * len);
                if (path_edges) memcpy(path_edges + out_pos, h + half, sizeof(int32_t) * len);
            }
            lens[req] = len;
            offs[req] = out_pos;
            out_pos += len;
            if (path_dist) path_dist[req] = en.h_dists[k];
            if (ms && ms->path_total) ms->path_total[req] = en.h_ms_total[k];
            if (ms && ms->path_target) ms->path_target[req] = en.h_ms_choice[k];
            slot_req[k] = -1;
            en.h_slots[k].state = 0;
            ++done_nets;
        }
        // mark the finished slots free on the device before reseeding
        hipLaunchKernelGGL(k_release, dim3((nf + 255) / 256), dim3(256), 0, s, a, en.slot_list, nf);
        HIP_CHECK_RET(hipGetLastError());
        for (int i = 0; i < nf; ++i) {
            const int k = en.h_slot_list[i];
            slot_drained[k] = en.h_slots[k].pending <= 0;
        }
        rc = seed_free_slots();
        if (rc) return rc;
    }
    HIP_CHECK_RET(hipEventRecord(en.ev_t1, s));
    HIP_CHECK_RET(hipMemcpyAsync(en.h_ctrl, en.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
    HIP_CHECK_RET(hipStreamSynchronize(s));
    // outputs in request order
    {
        std::vector<int32_t> tmp_n(out_pos), tmp_e(out_pos);
        std::copy(path_nodes, path_nodes + out_pos, tmp_n.begin());
        if (path_edges) std::copy(path_edges, path_edges + out_pos, tmp_e.begin());
        int64_t pos = 0;
        for (int i = 0; i < n_nets; ++i) {
            path_offset[i] = pos;
            std::copy(tmp_n.begin() + offs[i], tmp_n.begin() + offs[i] + lens[i], path_nodes + pos);
            if (path_edges) std::copy(tmp_e.begin() + offs[i], tmp_e.begin() + offs[i] + lens[i], path_edges + pos);
            pos += lens[i];
        }
        path_offset[n_nets] = pos;
    }
    if (stats) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, en.ev_t0, en.ev_t1);
        stats->rounds = (int64_t)en.h_ctrl->rounds;
        stats->items = (int64_t)en.h_ctrl->items;
        stats->edges = (int64_t)en.h_ctrl->edges;
        stats->minimal_nodes = (int64_t)min_nodes;
        stats->minimal_edges = (int64_t)min_edges;
        stats->kernel_ms = ms;
        stats->launches = launches;
        stats->round_ms = round_ms;
        stats->retried = 0;
    }
    return ORGPU_OK;
}
