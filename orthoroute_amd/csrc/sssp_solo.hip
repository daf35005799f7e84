// A3/A4, second engine: one wavefront per search ("solo" searches).
//
// Same contract as the bucket-round engine of sssp.hip (the FP32 shortest
// path with the lowest-predecessor tie-break, keys packed as in common.h,
// goal-directed buckets with the one-bucket margin at the destination), for
// two-pin requests.  Reference behaviour replaced: SimpleDijkstra.
// find_path_roi (UPF:1801-1887), the owner-aware keepout (UPF:3150-3228),
// the parent walk (UPF:1880-1887) and _path_to_edges (UPF:5463-5476).
//
// Design:
//  * a persistent grid of one-wave workgroups; workgroup w owns key slot w
//    (keys u64[NP] of its own) and takes requests from a global counter
//    until none are left.  No workgroup ever waits for another one.
//  * the wave alone reads and writes its slot's keys: a candidate that
//    passes a plain-load filter (and then its target's ROI bit) meets the
//    wave's other candidates for the same node in an LDS table (one slot per
//    node, claimed by linear probing, LDS 64-bit min), and the winner writes
//    the key with a plain store.  No global atomics: every global atomic,
//    workgroup-scope ones included, is performed at the memory side here.
//  * the frontier is a radix heap of per-wave lists in device memory
//    (entries (key cost, node) in the key's own bit layout; chains of chunks
//    of the slot's arena): list 0 holds the current bucket, list (i, d) the
//    entries whose bucket agrees with the current one above hex digit i and
//    has digit d there.  Expanding the next bucket empties the lowest
//    non-empty list into lower ones, so an entry moves at most once per
//    digit, however far ahead its bucket lies (keepout edges put whole
//    buckets 2^32 apart).  An entry whose cost no longer equals its node's key is
//    stale and skipped (every strict improvement pushes a new entry).
//    Appends are wave-aggregated (one ballot per target list), so list
//    traffic is coalesced and costs no atomics.
//  * keys a search set are recorded per chunk of 2^touch_shift keys in an
//    LDS bitmap and reset to KEY_INF by the wave when its search ends, so
//    every search starts on a clean slot.
//  * when the destination is settled (its bucket plus the margin lies below
//    the next bucket to expand) or the frontier is empty, lane 0 walks the
//    predecessor fields back to the source, the lanes find each hop's first
//    CSR edge and the path is packed src -> dst into the shared output.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "engine.h"

#define SOLO_LEVELS 10              // hex digits of a bucket index (8-bit keepout count and 32 bits)
#define SOLO_NL (1 + SOLO_LEVELS * 16)   // current bucket + one list per (digit, value)
#define SOLO_TOUCH_WORDS 1024       // LDS touch bitmap: 32768 key chunks
#define SOLO_DD_BITS 9
#define SOLO_DD (1u << SOLO_DD_BITS)  // LDS relaxation table entries (>= 4 x 64 candidates)
#define SOLO_MAX_STEPS (1u << 22)  // expansion steps of 64 entries per search

// request outcomes (out_len) besides a path length >= 0
#define SOLO_BROKEN (-1)            // predecessor chain broken or longer than the staging row
#define SOLO_KO_SATURATED (-2)      // path crosses KEY_KO_MAX or more keepout edges
#define SOLO_RETRY (-3)             // list or output capacity exceeded: route it on the other engine
#define SOLO_INVALID (-4)           // endpoint outside the graph, ROI row outside the rows given, or a
                                    // keepout request without the owner map

struct SoloArgs {
    const int32_t *indptr;
    const int32_t *indices;
    const float *cost;
    const uint8_t *kind;
    const uint32_t *roi;           // [n_roi][roi_words]
    int64_t roi_words;
    const int32_t *ko_owner;       // [N] or null
    int64_t N, NP;                 // NP: a slot's key stride (the tiled index space, rounded up)
    int32_t X, Y, plane;           // node = z*plane + y*X + x
    int32_t tiled;                 // keys in 2x2x4 (x, y, z) tiles of 16 (one 128-byte line)
    int32_t TX, TY;                // tiles along x and y
    uint64_t m_plane, m_x;         // division by plane and by X (magic_div)
    uint32_t s_plane, s_x;
    int32_t lookahead;             // buckets past the current one an under-full step takes entries from
    float inv_delta, h_xy, h_z;
    const orgpu_net_req *reqs;
    int32_t n_reqs;
    uint32_t *ctl;                 // [0] next request, [1] packed path entries used
    uint64_t *keys;                // [slots][NP]
    uint64_t *arena;               // [slots][nch][SOLO_CH] bucket-list entries
    uint32_t *links;               // [slots][2][nch]: chunk links, free stack
    uint32_t nch;                  // chunks per slot
    uint32_t touch_shift;          // keys per touch chunk = 2^touch_shift
    int32_t *stage;                // [slots][max_path] predecessor walk
    int32_t max_path;
    int32_t *pack_nodes;           // packed paths src -> dst, pack_cap entries
    int32_t *pack_edges;           // first CSR edge of each hop (-1 after the last node), or null
    uint64_t pack_cap;
    int32_t n_roi;                 // ROI rows (requests are checked against them and N)
    int32_t *out_len;
    uint32_t *out_off;
    float *out_dist;
    unsigned long long *counters;  // [0] items expanded, [1] edges scanned
    const uint4 *recs;             // [NP][2] adjacency records in key order (null: CSR rows only)
};

__device__ __forceinline__ uint64_t kload(const uint64_t *p)
{
    return __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void kstore(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// orders this wave's earlier key stores before its later key loads
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t wave_min64(uint64_t x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)x, o, 64), hi = __shfl_xor((uint32_t)(x >> 32), o, 64);
        const uint64_t y = ((uint64_t)hi << 32) | lo;
        x = y < x ? y : x;
    }
    return x;
}

__device__ __forceinline__ int wave_max_i32(int x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
    return x;
}

// frontier entry: the key's (keepout, cost) fields with the node in the
// predecessor field
__device__ __forceinline__ uint64_t entry_pack(uint64_t key, uint32_t node)
{
    return (key & ~(uint64_t)KEY_PRED_MASK) | (uint64_t)node;
}

// Bucket lists of one search: chains of SOLO_CH-entry chunks in the slot's
// arena.  Chunks come from a free stack (chunks a list has consumed) or
// from a bump pointer, so any list may hold up to the whole arena.  The
// chain links and the free stack live in device memory; the per-list state
// (head / tail chunk and position, entry count) and the allocator counters
// in LDS.  Every operation is wave-uniform.
#define SOLO_CH 256
#define SOLO_NONE 0xFFFFFFFFu
struct SoloLists {
    uint64_t *arena;       // [nch][SOLO_CH]
    uint32_t *link;        // [nch] next chunk of the chain
    uint32_t *freestk;     // [nch]
    uint32_t nch;
    uint32_t *hc;                        // LDS [SOLO_NL]: head chunk
    uint16_t *hp;                        //   position in it (< SOLO_CH)
    uint32_t *tc;                        //   tail chunk
    uint16_t *tp;                        //   entries in it (<= SOLO_CH)
    uint32_t *cnt;                       //   entries of the list
    uint32_t *al;                        // LDS [2]: free-stack top, bump pointer
    uint32_t *mask;                      // LDS [SOLO_LEVELS]: non-empty lists of each digit

    __device__ __forceinline__ bool empty(uint32_t t) const { return cnt[t] == 0u; }
    __device__ __forceinline__ uint32_t alloc() const
    {
        const uint32_t top = al[0];
        if (top > 0u) { al[0] = top - 1u; return freestk[top - 1u]; }
        const uint32_t b = al[1];
        if (b < nch) { al[1] = b + 1u; return b; }
        return SOLO_NONE;
    }
    __device__ __forceinline__ void release(uint32_t c) const
    {
        const uint32_t top = al[0];
        freestk[top] = c;
        al[0] = top + 1u;
    }
    __device__ __forceinline__ void reset(uint32_t lane) const
    {
        for (uint32_t t = lane; t < SOLO_NL; t += 64) {
            hc[t] = SOLO_NONE; tc[t] = SOLO_NONE; hp[t] = 0u; tp[t] = 0u; cnt[t] = 0u;
        }
        if (lane < 2) al[lane] = 0u;
        if (lane < SOLO_LEVELS) mask[lane] = 0u;
    }
    // Takes up to maxn (<= 64) entries from the head of list t: lane
    // base + k gets entry k, for k < n.  The chunks it empties go back to the
    // free stack; the entries are read before any later append can reuse
    // them.
    __device__ __forceinline__ uint64_t pop(uint32_t t, uint32_t lane, uint32_t base, uint32_t maxn,
                                            uint32_t &n_out) const
    {
        const uint32_t c = cnt[t];
        const uint32_t n = c < maxn ? c : maxn;
        n_out = n;
        const uint32_t h = hc[t], p = hp[t];
        const bool cross = p + n > SOLO_CH;
        const uint32_t nx = cross ? link[h] : SOLO_NONE;
        uint64_t e = 0ull;
        if (lane >= base && lane < base + n) {
            const uint32_t pos = p + (lane - base);
            e = pos < SOLO_CH ? arena[(uint64_t)h * SOLO_CH + pos] : arena[(uint64_t)nx * SOLO_CH + (pos - SOLO_CH)];
        }
        consume(t, c, n, h, p, cross, nx);
        return e;
    }
    // pop() for up to SOLO_CH entries, for the whole wave: entry k goes to
    // out[base + k], for k < the count returned
    __device__ __forceinline__ uint32_t pop_to(uint32_t t, uint32_t lane, uint32_t base, uint32_t maxn,
                                               uint64_t *out) const
    {
        const uint32_t c = cnt[t];
        const uint32_t n = c < maxn ? c : maxn;
        const uint32_t h = hc[t], p = hp[t];
        const bool cross = p + n > SOLO_CH;
        const uint32_t nx = cross ? link[h] : SOLO_NONE;
        for (uint32_t k = lane; k < n; k += 64) {
            const uint32_t pos = p + k;
            out[base + k] = pos < SOLO_CH ? arena[(uint64_t)h * SOLO_CH + pos]
                                          : arena[(uint64_t)nx * SOLO_CH + (pos - SOLO_CH)];
        }
        consume(t, c, n, h, p, cross, nx);
        return n;
    }
    // list state after n of its c entries were taken from position p of
    // head chunk h (nx: the next chunk when they cross into it)
    __device__ __forceinline__ void consume(uint32_t t, uint32_t c, uint32_t n, uint32_t h, uint32_t p, bool cross,
                                            uint32_t nx) const
    {
        const uint32_t left = c - n;
        cnt[t] = left;
        if (left == 0u) {
            if (t > 0u) mask[(t - 1u) >> 4] &= ~(1u << ((t - 1u) & 15u));
            release(h);
            if (cross) release(nx);
            hc[t] = SOLO_NONE; tc[t] = SOLO_NONE; hp[t] = 0u; tp[t] = 0u;
        } else if (p + n >= SOLO_CH) {
            // (a consumed head chunk always has a successor here: entries remain)
            const uint32_t nh = cross ? nx : link[h];
            release(h);
            hc[t] = nh;
            hp[t] = p + n - SOLO_CH;
        } else {
            hp[t] = p + n;
        }
    }
    // Moves every entry of list t to the empty list 0 (their buckets all
    // equal the new current bucket).
    __device__ __forceinline__ void take(uint32_t t, uint32_t lane) const
    {
        if (lane == 0) {
            hc[0] = hc[t]; hp[0] = hp[t]; tc[0] = tc[t]; tp[0] = tp[t]; cnt[0] = cnt[t];
            hc[t] = SOLO_NONE; tc[t] = SOLO_NONE; hp[t] = 0u; tp[t] = 0u; cnt[t] = 0u;
            mask[(t - 1u) >> 4] &= ~(1u << ((t - 1u) & 15u));
        }
    }
};

// Appends each active (lane, j) entry e[j] to its list t[j], one pass per
// distinct target list over all four candidates of every lane (a step's
// pushes mostly go to one to three lists).  Returns false (wave-uniform)
// when the arena is exhausted; the entries that did not fit are dropped and
// the caller abandons the search.
__device__ __forceinline__ bool lists_append4(const SoloLists &L, bool act[4], const uint32_t t[4], const uint64_t e[4])
{
    bool ok = true;
    for (;;) {
        const bool anyl = act[0] || act[1] || act[2] || act[3];
        const uint64_t m = __ballot(anyl);
        if (m == 0ull) break;
        const uint32_t tf = act[0] ? t[0] : act[1] ? t[1] : act[2] ? t[2] : t[3];
        const uint32_t t0 = (uint32_t)__shfl((int)tf, __builtin_ctzll(m));
        uint64_t mm[4];
        uint32_t off[4];
        uint32_t cnt = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            mm[j] = __ballot(act[j] && t[j] == t0);
            off[j] = cnt;
            cnt += (uint32_t)__popcll(mm[j]);
        }
        uint32_t tc = L.tc[t0], tp = L.tp[t0];
        bool fits = true;
        if (tc == SOLO_NONE) {
            tc = L.alloc();
            if (tc == SOLO_NONE) fits = false;
            else { L.hc[t0] = tc; L.hp[t0] = 0u; tp = 0u; }
        }
        uint32_t c2 = SOLO_NONE;
        if (fits && tp + cnt > SOLO_CH) {
            c2 = L.alloc();
            if (c2 == SOLO_NONE) { fits = false; L.tc[t0] = tc; L.tp[t0] = tp; }
            else L.link[tc] = c2;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const bool mine = act[j] && t[j] == t0;
            if (mine && fits) {
                const uint32_t pos = tp + off[j] + lanes_below(mm[j]);
                const uint64_t at = pos < SOLO_CH ? (uint64_t)tc * SOLO_CH + pos : (uint64_t)c2 * SOLO_CH + (pos - SOLO_CH);
                L.arena[at] = e[j];
            }
            if (mine) act[j] = false;
        }
        if (!fits) { ok = false; continue; }
        if (c2 != SOLO_NONE) { L.tc[t0] = c2; L.tp[t0] = tp + cnt - SOLO_CH; }
        else { L.tc[t0] = tc; L.tp[t0] = tp + cnt; }
        L.cnt[t0] += cnt;
        if (t0 > 0u) L.mask[(t0 - 1u) >> 4] |= 1u << ((t0 - 1u) & 15u);
    }
    return ok;
}

__device__ __forceinline__ bool lists_append(const SoloLists &L, bool act, uint32_t t, uint64_t e)
{
    bool a4[4] = {act, false, false, false};
    const uint32_t t4[4] = {t, 0u, 0u, 0u};
    const uint64_t e4[4] = {e, 0ull, 0ull, 0ull};
    return lists_append4(L, a4, t4, e4);
}

// List of bucket b relative to the current bucket cur: 0 for b <= cur,
// else (i, d) = the highest hex digit where b differs from cur and b's value
// there.
__device__ __forceinline__ uint32_t radix_list(uint64_t b, uint64_t cur)
{
    if (b <= cur) return 0u;
    const int i = (63 - __builtin_clzll(b ^ cur)) >> 2;
    return 1u + (uint32_t)i * 16u + ((uint32_t)(b >> (4 * i)) & 15u);
}

struct Cell {
    int x, y, z;
};

// floor(v / d) for v < 2^28 as (v * m) >> sh with m = ceil(2^sh / d),
// sh = 28 + ceil(log2 d) (solo_magic): exact, since m * d - 2^sh < d and
// v < 2^28 keep the error term below 1 / d
__device__ __forceinline__ uint32_t magic_div(uint32_t v, uint64_t m, uint32_t sh)
{
    return (uint32_t)(((uint64_t)v * m) >> sh);
}

__device__ __forceinline__ Cell node_cell(const SoloArgs &a, uint32_t v)
{
    Cell c;
    c.z = (int)magic_div(v, a.m_plane, a.s_plane);
    const uint32_t r = v - (uint32_t)c.z * (uint32_t)a.plane;
    c.y = (int)magic_div(r, a.m_x, a.s_x);
    c.x = (int)(r - (uint32_t)c.y * (uint32_t)a.X);
    return c;
}

// index of a node's key inside its slot: a tile of 2 x 2 cells on 4 layers
// shares one 128-byte line, so most of a node's lattice neighbours (the
// next cell along its layer, the layers above and below) share its line
__device__ __forceinline__ uint32_t tile_index(const SoloArgs &a, const Cell &c)
{
    const uint32_t t = ((uint32_t)(c.z >> 2) * (uint32_t)a.TY + (uint32_t)(c.y >> 1)) * (uint32_t)a.TX + (uint32_t)(c.x >> 1);
    return t * 16u + ((uint32_t)(c.z & 3) << 2) + ((uint32_t)(c.y & 1) << 1) + (uint32_t)(c.x & 1);
}

__device__ __forceinline__ uint32_t key_at(const SoloArgs &a, uint32_t v)
{
    return a.tiled ? tile_index(a, node_cell(a, v)) : v;
}

// lower bound of the remaining cost from cell c to the destination
__device__ __forceinline__ float solo_heur(const SoloArgs &a, int dx, int dy, int dz, const Cell &c)
{
    if (a.h_xy == 0.0f && a.h_z == 0.0f) return 0.0f;
    return a.h_xy * (float)(abs(c.x - dx) + abs(c.y - dy)) + a.h_z * (float)abs(c.z - dz);
}

// Adjacency records.  On a lattice most nodes have at most six edges whose
// targets are u -+ 1, u -+ X or u + dz * plane; such a node's row is kept,
// in the keys' tile order, as one 32-byte record: six FP32 costs, six
// one-byte target codes, the degree and a bit per planar edge (edge kind 0,
// for keepouts), so an expanded node costs one cache line instead of its row
// bounds, target indices and costs.  A node with another degree or target
// keeps its CSR row (degree byte SOLO_REC_CSR).  Records are written from
// the call's CSR and costs before each search launch.
#define SOLO_REC_MAXDEG 6
#define SOLO_REC_CSR 0xFFu
// target codes: 1..4 = u - X, u - 1, u + 1, u + X; 0x80 | dz = u + dz * plane
// (dz a signed 7-bit layer step)
__device__ __forceinline__ uint32_t rec_code(int64_t dv, int64_t X, int64_t plane)
{
    if (dv == -X) return 1u;
    if (dv == -1) return 2u;
    if (dv == 1) return 3u;
    if (dv == X) return 4u;
    if (dv != 0 && dv % plane == 0 && dv / plane >= -63 && dv / plane <= 63)
        return 0x80u | ((uint32_t)(dv / plane) & 0x7Fu);
    return 0u;
}

__device__ __forceinline__ uint32_t rec_target(uint32_t u, uint32_t code, uint32_t X, uint32_t plane)
{
    if (code & 0x80u) {
        const int dz = (int)(code << 25) >> 25;
        return (uint32_t)((int)u + dz * (int)plane);
    }
    return code == 1u ? u - X : code == 2u ? u - 1u : code == 3u ? u + 1u : u + X;
}

// one thread per node; ctl[2] counts the nodes left on their CSR rows
__global__ void __launch_bounds__(256)
k_solo_adjacency(SoloArgs a, uint4 *recs)
{
    uint32_t n_csr = 0;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < a.N; n += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t u = (uint32_t)n;
        const uint32_t t = tile_index(a, node_cell(a, u));
        const int row = a.indptr[u];
        const int deg = a.indptr[u + 1] - row;
        uint32_t cw[SOLO_REC_MAXDEG] = {0u, 0u, 0u, 0u, 0u, 0u};
        uint32_t code = 0u, kmask = 0u;
        bool fits = deg <= SOLO_REC_MAXDEG;
        for (int k = 0; fits && k < deg; ++k) {
            const uint32_t ck = rec_code((int64_t)a.indices[row + k] - (int64_t)u, a.X, a.plane);
            if (ck == 0u) { fits = false; break; }
            cw[k] = __float_as_uint(a.cost[row + k]);
            if (k < 4) code |= ck << (8 * k);
            else kmask |= ck << (8 * (k - 4));
            if (a.kind[row + k] == 0) kmask |= 1u << (24 + k);
        }
        uint4 r0, r1;
        if (fits) {
            r0 = make_uint4(cw[0], cw[1], cw[2], cw[3]);
            r1 = make_uint4(cw[4], cw[5], code, kmask | ((uint32_t)deg << 16));
        } else {
            r0 = make_uint4(0u, 0u, 0u, 0u);
            r1 = make_uint4(0u, 0u, 0u, SOLO_REC_CSR << 16);
            ++n_csr;
        }
        recs[2ull * t] = r0;
        recs[2ull * t + 1] = r1;
    }
    if (n_csr) atomicAdd(&a.ctl[2], n_csr);
}

// The adjacency of expanded node u (key index kidx): its record, or its CSR
// row (row, deg) when it has none or records are off for the launch
__device__ __forceinline__ void adj_load(const SoloArgs &a, bool use_rec, uint32_t u, uint32_t kidx, uint4 &r0,
                                         uint4 &r1, int &row, int &deg, bool &csr)
{
    csr = true;
    if (use_rec) {
        const uint4 *p = a.recs + 2ull * kidx;
        r0 = p[0];
        r1 = p[1];
        const uint32_t d = (r1.w >> 16) & 0xFFu;
        csr = d == SOLO_REC_CSR;
        deg = (int)d;
    }
    if (csr) {
        row = a.indptr[u];
        deg = a.indptr[u + 1] - row;
    }
}

__device__ __forceinline__ uint32_t u4_at(const uint4 &r, int j)
{
    return j == 0 ? r.x : j == 1 ? r.y : j == 2 ? r.z : r.w;
}

// edge k0e + j of the node (j a constant of the unrolled caller; a record
// has at most six edges, so k0e is 0 or 4 there): target, cost, and
// whether it is a planar edge (kind 0) when keepouts are counted
__device__ __forceinline__ void adj_edge(const SoloArgs &a, bool csr, const uint4 &r0, const uint4 &r1, int row,
                                         uint32_t u, int k0e, int j, bool ko, uint32_t &v, float &c, bool &planar)
{
    if (csr) {
        v = (uint32_t)a.indices[row + k0e + j];
        c = a.cost[row + k0e + j];
        planar = ko && a.kind[row + k0e + j] == 0;
        return;
    }
    const bool lo = k0e == 0;
    c = __uint_as_float(lo ? u4_at(r0, j) : u4_at(r1, j & 1));
    const uint32_t code = ((lo ? r1.z : r1.w) >> (8 * (lo ? j : (j & 1)))) & 0xFFu;
    v = rec_target(u, code, (uint32_t)a.X, (uint32_t)a.plane);
    planar = ko && ((r1.w >> (24 + k0e + j)) & 1u);
}

__global__ void __launch_bounds__(64)
k_solo(SoloArgs a)
{
    __shared__ uint32_t s_touch[SOLO_TOUCH_WORDS];
    __shared__ uint32_t s_hc[SOLO_NL], s_tc[SOLO_NL], s_cnt[SOLO_NL], s_al[2];
    __shared__ uint16_t s_hp[SOLO_NL], s_tp[SOLO_NL];      // (two-byte positions: 8 two-wave workgroups per CU)
    __shared__ uint32_t s_mask[SOLO_LEVELS];
    __shared__ uint32_t s_ddtag[SOLO_DD];          // relaxation table: node + 1 (0 = free)
    __shared__ uint64_t s_ddkey[SOLO_DD];          // its best candidate key
    const uint32_t lane = threadIdx.x;
    const uint32_t slot = blockIdx.x;
    uint64_t *keys = a.keys + (uint64_t)slot * (uint64_t)a.NP;
    SoloLists L{a.arena + (uint64_t)slot * (uint64_t)a.nch * SOLO_CH, a.links + (uint64_t)slot * 2u * a.nch,
                a.links + (uint64_t)slot * 2u * a.nch + a.nch, a.nch, s_hc, s_hp, s_tc, s_tp, s_cnt, s_al, s_mask};
    int32_t *stage = a.stage + (uint64_t)slot * (uint64_t)a.max_path;
    const uint32_t tshift = a.touch_shift;
    for (int i = lane; i < SOLO_TOUCH_WORDS; i += 64) s_touch[i] = 0u;
    for (int i = lane; i < SOLO_DD; i += 64) { s_ddtag[i] = 0u; s_ddkey[i] = KEY_INF; }
    const bool goal = a.h_xy != 0.0f || a.h_z != 0.0f;
    const uint64_t margin = goal ? 1u : 0u;
    // records, unless more than one node in eight keeps its CSR row
    const bool use_rec = a.recs && a.tiled && (uint64_t)a.ctl[2] * 8u <= (uint64_t)a.N;
    unsigned long long my_items = 0, my_edges = 0;

    for (;;) {
        uint32_t r = 0;
        if (lane == 0) r = atomicAdd(&a.ctl[0], 1u);
        r = (uint32_t)__shfl((int)r, 0);
        if (r >= (uint32_t)a.n_reqs) break;
        const orgpu_net_req q = a.reqs[r];
        if (q.src < 0 || (int64_t)q.src >= a.N || q.dst < 0 || (int64_t)q.dst >= a.N || q.roi_index >= a.n_roi ||
            (q.use_keepout && !a.ko_owner)) {
            if (lane == 0) {
                a.out_len[r] = SOLO_INVALID;
                a.out_off[r] = 0u;
                a.out_dist[r] = __builtin_inff();
            }
            continue;
        }
        const uint32_t src = (uint32_t)q.src, dst = (uint32_t)q.dst;
        int dx = 0, dy = 0, dz = 0;
        if (goal) {
            const Cell cd = node_cell(a, dst);
            dx = cd.x; dy = cd.y; dz = cd.z;
        }
        const uint32_t *roi = q.roi_index >= 0 ? a.roi + (uint64_t)q.roi_index * (uint64_t)a.roi_words : nullptr;
        const bool use_ko = q.use_keepout && a.ko_owner;
        L.reset(lane);

        // seed the source in its own bucket
        const uint64_t k0 = key_pack(0, 0.0f, KEY_NOPRED);
        const bool lat = a.tiled || goal;
        const uint32_t kdst = key_at(a, dst);
        uint64_t cur = key_bucket_rel(k0, a.inv_delta, lat ? solo_heur(a, dx, dy, dz, node_cell(a, src)) : 0.0f);
        if (lane == 0) {
            const uint32_t ks = key_at(a, src);
            kstore(&keys[ks], k0);
            atomicOr(&s_touch[(ks >> tshift) >> 5], 1u << ((ks >> tshift) & 31));
        }
        bool ok = lists_append(L, lane == 0, 0u, entry_pack(k0, src));
        wave_fence();
        bool reached = false;
        // a bound on the expansion steps of one search, so that every wave
        // ends whatever the input (a search that reaches it is handed back)
        uint32_t steps = 0;

        while (ok) {
            // expand the current bucket to its fixed point (entries pushed
            // into it meanwhile are appended to the same list)
            const uint32_t lc = 0u;
            while (ok && !L.empty(lc)) {
                if (++steps > SOLO_MAX_STEPS) { ok = false; break; }
                uint32_t n = 0;
                uint64_t e = L.pop(lc, lane, 0u, 64u, n);
                // an under-full step also takes entries of the next buckets
                // of this 16-bucket block (up to a.lookahead ahead): expanding
                // a node early is exact (it is expanded again if its key
                // drops later) and it fills lanes that would idle
                if (n < 64u && a.lookahead > 0) {
                    const uint32_t d0 = (uint32_t)(cur & 15u);
                    const uint32_t dl = min(15u, d0 + (uint32_t)a.lookahead);
                    // the non-empty lists of digits d0+1 .. dl
                    uint32_t mk = L.mask[0] & (((2u << dl) - 1u) & ~((2u << d0) - 1u));
                    while (mk && n < 64u) {
                        const uint32_t d = (uint32_t)__builtin_ctz(mk);
                        mk &= mk - 1u;
                        uint32_t n2 = 0;
                        const uint64_t e2 = L.pop(1u + d, lane, n, 64u - n, n2);
                        if (lane >= n && lane < n + n2) e = e2;
                        n += n2;
                    }
                }
                bool has = lane < n;
                const uint32_t u = (uint32_t)(e & KEY_PRED_MASK);
                uint64_t ku = 0;
                int row = 0, deg = 0;
                bool csr = true;
                uint4 r0 = make_uint4(0u, 0u, 0u, 0u), r1 = r0;
                int32_t own = -1;
                Cell cu = Cell{0, 0, 0};
                if (has) {
                    if (lat) cu = node_cell(a, u);
                    const uint32_t ki = a.tiled ? tile_index(a, cu) : u;
                    ku = kload(&keys[ki]);
                    adj_load(a, use_rec, u, ki, r0, r1, row, deg, csr);
                    if (use_ko) own = a.ko_owner[u];
                }
                // stale: a cheaper entry of the node was pushed since
                has = has && key_cost(ku) == key_cost(e);
                if (!has) deg = 0;
                if (has) { my_items++; my_edges += (unsigned long long)deg; }
                const uint32_t ko0 = key_ko(e);
                const float d0 = key_dist(e);
                const bool ko_node = has && use_ko && own >= 0 && own != q.net_id;
                const int maxdeg = wave_max_i32(deg);
                for (int k0e = 0; k0e < maxdeg; k0e += 4) {
                    uint32_t v[4];
                    float c[4];
                    bool go[4], pko[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        go[j] = k0e + j < deg;
                        pko[j] = false;
                        v[j] = 0u;
                        c[j] = 0.0f;
                        if (go[j]) adj_edge(a, csr, r0, r1, row, u, k0e, j, ko_node, v[j], c[j], pko[j]);
                    }
                    // target keys first; the ROI bit is read only for a
                    // candidate that would improve its target (an edge to a
                    // settled neighbour, the common case, costs one load)
                    uint64_t kv[4], nk[4];
                    uint32_t kix[4];
                    Cell cv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        kv[j] = KEY_INF;
                        nk[j] = KEY_INF;
                        kix[j] = v[j];
                        cv[j] = Cell{0, 0, 0};
                        if (!go[j]) continue;
                        if (lat) {
                            cv[j] = node_cell(a, v[j]);
                            if (a.tiled) kix[j] = tile_index(a, cv[j]);
                        }
                        kv[j] = kload(&keys[kix[j]]);
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (!go[j]) continue;
                        uint32_t ko = ko0;
                        float nd = d0;
                        if (pko[j]) ko = ko_add(ko);
                        else nd = nd + c[j];
                        nk[j] = key_pack(ko, nd, u + 1u);
                        go[j] = nd < __builtin_inff() && nk[j] < kv[j];
                    }
                    if (roi) {
                        uint32_t rw[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) rw[j] = go[j] ? roi[v[j] >> 5] : 0u;
#pragma unroll
                        for (int j = 0; j < 4; ++j) go[j] = go[j] && ((rw[j] >> (v[j] & 31)) & 1u);
                    }
                    // relax: only this wave writes the slot's keys, so the
                    // lanes' candidates for one node meet in the wave's LDS
                    // table (claim a slot per node by linear probing, LDS
                    // 64-bit min), and the winner writes the key with a plain
                    // store: no global atomics
                    uint32_t pos[4];
                    bool won[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        pos[j] = 0u;
                        if (!go[j]) continue;
                        uint32_t p = (v[j] * 2654435761u) >> (32 - SOLO_DD_BITS);
                        for (;;) {
                            const uint32_t o = atomicCAS(&s_ddtag[p], 0u, v[j] + 1u);
                            if (o == 0u || o == v[j] + 1u) break;
                            p = (p + 1u) & (SOLO_DD - 1u);
                        }
                        pos[j] = p;
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (go[j]) atomicMin((unsigned long long *)&s_ddkey[pos[j]], (unsigned long long)nk[j]);
#pragma unroll
                    for (int j = 0; j < 4; ++j) won[j] = go[j] && s_ddkey[pos[j]] == nk[j];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (!won[j]) continue;
                        s_ddtag[pos[j]] = 0u;
                        s_ddkey[pos[j]] = KEY_INF;
                        kstore(&keys[kix[j]], nk[j]);
                    }
                    // touched chunks and pushes of strict improvements
                    bool push[4];
                    uint32_t tl4[4];
                    uint64_t e4[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (won[j] && kv[j] == KEY_INF) {
                            const uint32_t ch = kix[j] >> tshift;
                            atomicOr(&s_touch[ch >> 5], 1u << (ch & 31));
                        }
                        push[j] = won[j] && key_cost(nk[j]) < key_cost(kv[j]);
                        tl4[j] = push[j] ? radix_list(key_bucket_rel(nk[j], a.inv_delta,
                                                                      solo_heur(a, dx, dy, dz, cv[j])), cur) : 0u;
                        e4[j] = entry_pack(nk[j], v[j]);
                    }
                    ok = lists_append4(L, push, tl4, e4) && ok;
                }
                wave_fence();
            }
            if (!ok) break;
            // next bucket to expand: the lowest non-empty list's lower bound
            // (the current bucket's digits above i, digit d, zeros below)
            uint64_t next = ~0ull;
            uint32_t tl = 0u;
            for (uint32_t i = 0; i < SOLO_LEVELS; ++i) {
                const uint32_t m = L.mask[i];
                if (m) {
                    const uint32_t d = (uint32_t)__builtin_ctz(m);
                    const uint64_t low = (16ull << (4 * i)) - 1ull;
                    next = (cur & ~low) | ((uint64_t)d << (4 * i));
                    tl = 1u + i * 16u + d;
                    break;
                }
            }
            const uint64_t kd = kload(&keys[kdst]);
            if (kd != KEY_INF && (next == ~0ull || key_bucket_rel(kd, a.inv_delta) + margin < next)) {
                reached = true;
                break;
            }
            if (next == ~0ull) break;   // frontier empty: dst unreachable
            // the list's live entries move to lower lists relative to the new
            // current bucket (they agree with it above digit i); a list of
            // the lowest digit holds that bucket alone and becomes list 0
            cur = next;
            if (tl <= 16u) {
                L.take(tl, lane);
                continue;
            }
            while (ok && !L.empty(tl)) {
                uint32_t n = 0;
                const uint64_t e = L.pop(tl, lane, 0u, 64u, n);
                bool has = lane < n;
                const uint32_t v = (uint32_t)(e & KEY_PRED_MASK);
                const Cell cc = has && lat ? node_cell(a, v) : Cell{0, 0, 0};
                const uint64_t kv = has ? kload(&keys[a.tiled ? tile_index(a, cc) : v]) : 0ull;
                has = has && key_cost(kv) == key_cost(e);
                const uint32_t t = has ? radix_list(key_bucket_rel(kv, a.inv_delta, solo_heur(a, dx, dy, dz, cc)), cur) : 0u;
                ok = lists_append(L, has, t, e) && ok;
            }
        }

        // result: walk the predecessors (lane 0), then pack src -> dst
        int32_t len = 0;
        float dist = __builtin_inff();
        if (!ok) {
            len = SOLO_RETRY;
        } else if (reached) {
            const uint64_t kd = kload(&keys[kdst]);
            dist = key_dist(kd);
            if (key_ko(kd) >= KEY_KO_MAX) {
                len = SOLO_KO_SATURATED;
            } else if (lane == 0) {
                uint32_t v = dst;
                for (;;) {
                    if (len >= a.max_path) { len = SOLO_BROKEN; break; }
                    stage[len++] = (int32_t)v;
                    if (v == src) break;
                    const uint32_t pf = key_pred(kload(&keys[key_at(a, v)]));
                    if (pf == KEY_NOPRED) { len = SOLO_BROKEN; break; }
                    v = pf - 1u;
                }
            }
            len = __shfl(len, 0);
        }
        uint32_t off = 0;
        if (len > 0) {
            if (lane == 0) off = atomicAdd(&a.ctl[1], (uint32_t)len);
            off = (uint32_t)__shfl((int)off, 0);
            if ((uint64_t)off + (uint64_t)len > a.pack_cap) len = SOLO_RETRY;
        }
        if (len > 0) {
            wave_fence();
            int32_t *on = a.pack_nodes + off;
            int32_t *oe = a.pack_edges ? a.pack_edges + off : nullptr;
            // node t of the packed path is stage[len-1-t]; the edge leaving
            // it is the first CSR edge stage[len-1-t] -> stage[len-2-t]
            for (int t = lane; t < len; t += 64) {
                const int32_t un = stage[len - 1 - t];
                on[t] = un;
                if (!oe) continue;
                int ee = -1;
                if (t + 1 < len) {
                    const int32_t vn = stage[len - 2 - t];
                    const int r0 = a.indptr[un], r1 = a.indptr[un + 1];
                    for (int e2 = r0; e2 < r1; ++e2)
                        if (a.indices[e2] == vn) { ee = e2; break; }
                }
                oe[t] = ee;
            }
        }
        if (lane == 0) {
            a.out_len[r] = len;
            a.out_off[r] = off;
            a.out_dist[r] = dist;
        }
        // reset every key chunk the search touched, and the lists
        for (int w0 = 0; w0 < SOLO_TOUCH_WORDS; w0 += 64) {
            uint32_t word = s_touch[w0 + lane];
            uint64_t any = __ballot(word != 0u);
            while (any) {
                const int l = __builtin_ctzll(any);
                uint32_t wl = (uint32_t)__shfl((int)word, l);
                while (wl) {
                    const int bit = __builtin_ctz(wl);
                    wl &= wl - 1u;
                    const uint64_t lo = ((uint64_t)(w0 + l) * 32u + (uint64_t)bit) << tshift;
                    const uint64_t hi = min(lo + (1ull << tshift), (uint64_t)a.NP);
                    for (uint64_t k = lo + lane; k < hi; k += 64) keys[k] = KEY_INF;
                }
                any &= any - 1ull;
            }
            s_touch[w0 + lane] = 0u;
        }
        wave_fence();
    }
    // counters
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        my_items += __shfl_xor(my_items, o, 64);
        my_edges += __shfl_xor(my_edges, o, 64);
    }
    if (lane == 0) {
        atomicAdd(&a.counters[0], my_items);
        atomicAdd(&a.counters[1], my_edges);
    }
}

// ---------------------------------------------------------------------------
// W wavefronts per search (W = 2 or 4): the same search as k_solo, for calls
// with fewer searches in flight than the GPU holds waves (a large board
// whose keys fit fewer slots, a batch shorter than the slots).  Wave 0 alone
// owns the bucket lists: it takes up to 64*W entries per step into LDS, and
// the waves append their pushes one wave at a time.  Every wave expands
// 64 entries; candidates for one node meet in the workgroup's LDS table,
// and block barriers separate the table's claim / minimum, its read and its
// clear.  Every barrier is reached by all W waves: each loop condition is
// read from LDS by all of them between two barriers.
// ---------------------------------------------------------------------------
template <int W>
__global__ void __launch_bounds__(64 * W)
k_solo_mw(SoloArgs a)
{
    static_assert(W == 2 || W == 4, "2 or 4 waves per search");
    constexpr uint32_t NT = 64u * W;                 // entries per step (<= SOLO_CH)
    constexpr uint32_t DD = SOLO_DD * (uint32_t)W;   // >= 2 x the 4*NT candidates of a pass
    constexpr uint32_t DD_BITS = SOLO_DD_BITS + (W == 2 ? 1 : 2);
    __shared__ uint32_t s_touch[SOLO_TOUCH_WORDS];
    __shared__ uint32_t s_hc[SOLO_NL], s_tc[SOLO_NL], s_cnt[SOLO_NL], s_al[2];
    __shared__ uint16_t s_hp[SOLO_NL], s_tp[SOLO_NL];      // (two-byte positions: 8 two-wave workgroups per CU)
    __shared__ uint32_t s_mask[SOLO_LEVELS];
    __shared__ uint32_t s_ddtag[DD];
    __shared__ uint64_t s_ddkey[DD];
    __shared__ uint64_t s_ent[NT];                   // a step's entries
    __shared__ uint32_t s_ctl[4];                    // [0] request, [1] entries this step, [2] lists ok
    __shared__ int32_t s_deg[W];                     // per wave: largest degree of a live entry
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t slot = blockIdx.x;
    uint64_t *keys = a.keys + (uint64_t)slot * (uint64_t)a.NP;
    SoloLists L{a.arena + (uint64_t)slot * (uint64_t)a.nch * SOLO_CH, a.links + (uint64_t)slot * 2u * a.nch,
                a.links + (uint64_t)slot * 2u * a.nch + a.nch, a.nch, s_hc, s_hp, s_tc, s_tp, s_cnt, s_al, s_mask};
    int32_t *stage = a.stage + (uint64_t)slot * (uint64_t)a.max_path;
    const uint32_t tshift = a.touch_shift;
    for (uint32_t i = tid; i < SOLO_TOUCH_WORDS; i += NT) s_touch[i] = 0u;
    for (uint32_t i = tid; i < DD; i += NT) { s_ddtag[i] = 0u; s_ddkey[i] = KEY_INF; }
    const bool goal = a.h_xy != 0.0f || a.h_z != 0.0f;
    const uint64_t margin = goal ? 1u : 0u;
    const bool lat = a.tiled || goal;
    const bool use_rec = a.recs && a.tiled && (uint64_t)a.ctl[2] * 8u <= (uint64_t)a.N;
    unsigned long long my_items = 0, my_edges = 0;
    __syncthreads();

    for (;;) {
        if (tid == 0) s_ctl[0] = atomicAdd(&a.ctl[0], 1u);
        __syncthreads();
        const uint32_t r = s_ctl[0];
        __syncthreads();                             // read by all before the next request overwrites it
        if (r >= (uint32_t)a.n_reqs) break;
        const orgpu_net_req q = a.reqs[r];
        if (q.src < 0 || (int64_t)q.src >= a.N || q.dst < 0 || (int64_t)q.dst >= a.N || q.roi_index >= a.n_roi ||
            (q.use_keepout && !a.ko_owner)) {
            if (tid == 0) {
                a.out_len[r] = SOLO_INVALID;
                a.out_off[r] = 0u;
                a.out_dist[r] = __builtin_inff();
            }
            continue;
        }
        const uint32_t src = (uint32_t)q.src, dst = (uint32_t)q.dst;
        int dx = 0, dy = 0, dz = 0;
        if (goal) {
            const Cell cd = node_cell(a, dst);
            dx = cd.x; dy = cd.y; dz = cd.z;
        }
        const uint32_t *roi = q.roi_index >= 0 ? a.roi + (uint64_t)q.roi_index * (uint64_t)a.roi_words : nullptr;
        const bool use_ko = q.use_keepout && a.ko_owner;
        const uint64_t k0 = key_pack(0, 0.0f, KEY_NOPRED);
        const uint32_t kdst = key_at(a, dst);
        uint64_t cur = key_bucket_rel(k0, a.inv_delta, lat ? solo_heur(a, dx, dy, dz, node_cell(a, src)) : 0.0f);
        if (wave == 0) {
            L.reset(lane);
            if (lane == 0) {
                const uint32_t ks = key_at(a, src);
                kstore(&keys[ks], k0);
                atomicOr(&s_touch[(ks >> tshift) >> 5], 1u << ((ks >> tshift) & 31));
            }
            const bool okw = lists_append(L, lane == 0, 0u, entry_pack(k0, src));
            if (lane == 0) s_ctl[2] = okw ? 1u : 0u;
        }
        __syncthreads();
        bool ok = s_ctl[2] != 0u;
        bool reached = false;
        uint32_t steps = 0;

        while (ok) {
            // expand the current bucket to its fixed point
            for (;;) {
                const bool more = ok && !L.empty(0u);
                __syncthreads();                     // all have read the list state
                if (!more) break;
                if (++steps > SOLO_MAX_STEPS) { ok = false; break; }
                if (wave == 0) {
                    uint32_t n = L.pop_to(0u, lane, 0u, NT, s_ent);
                    if (n < NT && a.lookahead > 0) {
                        const uint32_t d0 = (uint32_t)(cur & 15u);
                        const uint32_t dl = min(15u, d0 + (uint32_t)a.lookahead);
                        uint32_t mk = L.mask[0] & (((2u << dl) - 1u) & ~((2u << d0) - 1u));
                        while (mk && n < NT) {
                            const uint32_t d = (uint32_t)__builtin_ctz(mk);
                            mk &= mk - 1u;
                            n += L.pop_to(1u + d, lane, n, NT - n, s_ent);
                        }
                    }
                    if (lane == 0) s_ctl[1] = n;
                }
                __syncthreads();                     // the step's entries
                const uint32_t n = s_ctl[1];
                bool has = tid < n;
                const uint64_t e = has ? s_ent[tid] : 0ull;
                const uint32_t u = (uint32_t)(e & KEY_PRED_MASK);
                uint64_t ku = 0;
                int row = 0, deg = 0;
                bool csr = true;
                uint4 r0 = make_uint4(0u, 0u, 0u, 0u), r1 = r0;
                int32_t own = -1;
                Cell cu = Cell{0, 0, 0};
                if (has) {
                    if (lat) cu = node_cell(a, u);
                    const uint32_t ki = a.tiled ? tile_index(a, cu) : u;
                    ku = kload(&keys[ki]);
                    adj_load(a, use_rec, u, ki, r0, r1, row, deg, csr);
                    if (use_ko) own = a.ko_owner[u];
                }
                has = has && key_cost(ku) == key_cost(e);
                if (!has) deg = 0;
                if (has) { my_items++; my_edges += (unsigned long long)deg; }
                const uint32_t ko0 = key_ko(e);
                const float d0 = key_dist(e);
                const bool ko_node = has && use_ko && own >= 0 && own != q.net_id;
                const int wdeg = wave_max_i32(deg);
                if (lane == 0) s_deg[wave] = wdeg;
                __syncthreads();
                int maxdeg = 0;
#pragma unroll
                for (int w = 0; w < W; ++w) maxdeg = max(maxdeg, s_deg[w]);
                for (int k0e = 0; k0e < maxdeg; k0e += 4) {
                    uint32_t v[4];
                    float c[4];
                    bool go[4], pko[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        go[j] = k0e + j < deg;
                        pko[j] = false;
                        v[j] = 0u;
                        c[j] = 0.0f;
                        if (go[j]) adj_edge(a, csr, r0, r1, row, u, k0e, j, ko_node, v[j], c[j], pko[j]);
                    }
                    uint64_t kv[4], nk[4];
                    uint32_t kix[4];
                    Cell cv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        kv[j] = KEY_INF;
                        nk[j] = KEY_INF;
                        kix[j] = v[j];
                        cv[j] = Cell{0, 0, 0};
                        if (!go[j]) continue;
                        if (lat) {
                            cv[j] = node_cell(a, v[j]);
                            if (a.tiled) kix[j] = tile_index(a, cv[j]);
                        }
                        kv[j] = kload(&keys[kix[j]]);
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (!go[j]) continue;
                        uint32_t ko = ko0;
                        float nd = d0;
                        if (pko[j]) ko = ko_add(ko);
                        else nd = nd + c[j];
                        nk[j] = key_pack(ko, nd, u + 1u);
                        go[j] = nd < __builtin_inff() && nk[j] < kv[j];
                    }
                    if (roi) {
                        uint32_t rw[4];
#pragma unroll
                        for (int j = 0; j < 4; ++j) rw[j] = go[j] ? roi[v[j] >> 5] : 0u;
#pragma unroll
                        for (int j = 0; j < 4; ++j) go[j] = go[j] && ((rw[j] >> (v[j] & 31)) & 1u);
                    }
                    // the workgroup's candidates for one node meet in the
                    // LDS table: claim and minimum, barrier, read, barrier,
                    // clear and store by the winner
                    uint32_t pos[4];
                    bool won[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        pos[j] = 0u;
                        if (!go[j]) continue;
                        uint32_t p = (v[j] * 2654435761u) >> (32 - DD_BITS);
                        for (;;) {
                            const uint32_t o = atomicCAS(&s_ddtag[p], 0u, v[j] + 1u);
                            if (o == 0u || o == v[j] + 1u) break;
                            p = (p + 1u) & (DD - 1u);
                        }
                        pos[j] = p;
                    }
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (go[j]) atomicMin((unsigned long long *)&s_ddkey[pos[j]], (unsigned long long)nk[j]);
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < 4; ++j) won[j] = go[j] && s_ddkey[pos[j]] == nk[j];
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (!won[j]) continue;
                        s_ddtag[pos[j]] = 0u;
                        s_ddkey[pos[j]] = KEY_INF;
                        kstore(&keys[kix[j]], nk[j]);
                    }
                    bool push[4];
                    uint32_t tl4[4];
                    uint64_t e4[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        if (won[j] && kv[j] == KEY_INF) {
                            const uint32_t ch = kix[j] >> tshift;
                            atomicOr(&s_touch[ch >> 5], 1u << (ch & 31));
                        }
                        push[j] = won[j] && key_cost(nk[j]) < key_cost(kv[j]);
                        tl4[j] = push[j] ? radix_list(key_bucket_rel(nk[j], a.inv_delta,
                                                                      solo_heur(a, dx, dy, dz, cv[j])), cur) : 0u;
                        e4[j] = entry_pack(nk[j], v[j]);
                    }
                    // the waves append one at a time (the lists' state is
                    // the workgroup's); the last barrier also orders this
                    // pass's table clears and key stores before the next
                    for (int w = 0; w < W; ++w) {
                        if ((int)wave == w) {
                            const bool okw = lists_append4(L, push, tl4, e4);
                            if (!okw && lane == 0) s_ctl[2] = 0u;
                        }
                        __syncthreads();
                    }
                }
                ok = s_ctl[2] != 0u;
            }
            if (!ok) break;
            uint64_t next = ~0ull;
            uint32_t tl = 0u;
            for (uint32_t i = 0; i < SOLO_LEVELS; ++i) {
                const uint32_t m = L.mask[i];
                if (m) {
                    const uint32_t d = (uint32_t)__builtin_ctz(m);
                    const uint64_t low = (16ull << (4 * i)) - 1ull;
                    next = (cur & ~low) | ((uint64_t)d << (4 * i));
                    tl = 1u + i * 16u + d;
                    break;
                }
            }
            const uint64_t kd = kload(&keys[kdst]);
            if (kd != KEY_INF && (next == ~0ull || key_bucket_rel(kd, a.inv_delta) + margin < next)) {
                reached = true;
                break;
            }
            if (next == ~0ull) break;
            cur = next;
            __syncthreads();                         // all have read the list state
            if (wave == 0) {
                if (tl <= 16u) {
                    L.take(tl, lane);
                } else {
                    bool okw = true;
                    while (okw && !L.empty(tl)) {
                        uint32_t n = 0;
                        const uint64_t e = L.pop(tl, lane, 0u, 64u, n);
                        bool has = lane < n;
                        const uint32_t v = (uint32_t)(e & KEY_PRED_MASK);
                        const Cell cc = has && lat ? node_cell(a, v) : Cell{0, 0, 0};
                        const uint64_t kv = has ? kload(&keys[a.tiled ? tile_index(a, cc) : v]) : 0ull;
                        has = has && key_cost(kv) == key_cost(e);
                        const uint32_t t = has ? radix_list(key_bucket_rel(kv, a.inv_delta,
                                                                           solo_heur(a, dx, dy, dz, cc)), cur) : 0u;
                        okw = lists_append(L, has, t, e) && okw;
                    }
                    if (!okw && lane == 0) s_ctl[2] = 0u;
                }
            }
            __syncthreads();
            ok = s_ctl[2] != 0u;
        }

        // result (wave 0): walk the predecessors, pack src -> dst
        if (wave == 0) {
            int32_t len = 0;
            float dist = __builtin_inff();
            if (!ok) {
                len = SOLO_RETRY;
            } else if (reached) {
                const uint64_t kd = kload(&keys[kdst]);
                dist = key_dist(kd);
                if (key_ko(kd) >= KEY_KO_MAX) {
                    len = SOLO_KO_SATURATED;
                } else if (lane == 0) {
                    uint32_t v = dst;
                    for (;;) {
                        if (len >= a.max_path) { len = SOLO_BROKEN; break; }
                        stage[len++] = (int32_t)v;
                        if (v == src) break;
                        const uint32_t pf = key_pred(kload(&keys[key_at(a, v)]));
                        if (pf == KEY_NOPRED) { len = SOLO_BROKEN; break; }
                        v = pf - 1u;
                    }
                }
                len = __shfl(len, 0);
            }
            uint32_t off = 0;
            if (len > 0) {
                if (lane == 0) off = atomicAdd(&a.ctl[1], (uint32_t)len);
                off = (uint32_t)__shfl((int)off, 0);
                if ((uint64_t)off + (uint64_t)len > a.pack_cap) len = SOLO_RETRY;
            }
            if (len > 0) {
                wave_fence();
                int32_t *on = a.pack_nodes + off;
                int32_t *oe = a.pack_edges ? a.pack_edges + off : nullptr;
                for (int t = lane; t < len; t += 64) {
                    const int32_t un = stage[len - 1 - t];
                    on[t] = un;
                    if (!oe) continue;
                    int ee = -1;
                    if (t + 1 < len) {
                        const int32_t vn = stage[len - 2 - t];
                        const int r0 = a.indptr[un], r1 = a.indptr[un + 1];
                        for (int e2 = r0; e2 < r1; ++e2)
                            if (a.indices[e2] == vn) { ee = e2; break; }
                    }
                    oe[t] = ee;
                }
            }
            if (lane == 0) {
                a.out_len[r] = len;
                a.out_off[r] = off;
                a.out_dist[r] = dist;
            }
        }
        __syncthreads();                             // the walk has read the keys
        // reset every key chunk the search touched (the waves split the words)
        for (uint32_t w0 = wave * 64u; w0 < SOLO_TOUCH_WORDS; w0 += NT) {
            uint32_t word = s_touch[w0 + lane];
            uint64_t any = __ballot(word != 0u);
            while (any) {
                const int l = __builtin_ctzll(any);
                uint32_t wl = (uint32_t)__shfl((int)word, l);
                while (wl) {
                    const int bit = __builtin_ctz(wl);
                    wl &= wl - 1u;
                    const uint64_t lo = ((uint64_t)(w0 + l) * 32u + (uint64_t)bit) << tshift;
                    const uint64_t hi = min(lo + (1ull << tshift), (uint64_t)a.NP);
                    for (uint64_t k = lo + lane; k < hi; k += 64) keys[k] = KEY_INF;
                }
                any &= any - 1ull;
            }
            s_touch[w0 + lane] = 0u;
        }
        __syncthreads();
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        my_items += __shfl_xor(my_items, o, 64);
        my_edges += __shfl_xor(my_edges, o, 64);
    }
    if (lane == 0) {
        atomicAdd(&a.counters[0], my_items);
        atomicAdd(&a.counters[1], my_edges);
    }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

// m = ceil(2^sh / d), sh = 28 + ceil(log2 d): magic_div(v, m, sh) = v / d
// for every v < 2^28 (the node-index range)
static void solo_magic(uint32_t d, uint64_t *m, uint32_t *sh)
{
    uint32_t lg = 0;
    while ((1ull << lg) < (uint64_t)d) ++lg;
    *sh = 28u + lg;
    *m = ((1ull << *sh) + (uint64_t)d - 1ull) / (uint64_t)d;
}

// the call's control words: zeroed by a kernel (a stream a caller captures
// records no memset node)
__global__ void k_solo_begin(uint32_t *ctl)
{
    if (threadIdx.x < 16) ctl[threadIdx.x] = 0u;
}

__global__ void k_solo_used(const uint32_t *ctl, uint32_t *used)
{
    if (threadIdx.x == 0) *used = ctl[1];
}

// Key layout of a graph: 2x2x4 tiles on a lattice (X*Y*L nodes in node
// order z*X*Y + y*X + x), node order otherwise; a slot's stride NP covers
// the index space in whole 2048-key chunks, touched in chunks of 2^tshift.
struct SoloLayout {
    bool tiled;
    int32_t TX, TY;
    int64_t NP;
    uint32_t tshift;
};

static SoloLayout solo_layout(int64_t N, int32_t X, int32_t Y, int32_t L)
{
    SoloLayout y;
    y.tiled = X > 0 && Y > 0 && L > 0 && (int64_t)X * Y * L == N;
    y.TX = (X + 1) / 2;
    y.TY = (Y + 1) / 2;
    const int64_t span = y.tiled ? (int64_t)y.TX * y.TY * ((L + 3) / 4) * 16 : N;
    y.NP = (span + 2047) / 2048 * 2048;
    y.tshift = 6;
    while ((((uint64_t)y.NP + (1ull << y.tshift) - 1) >> y.tshift) > (uint64_t)SOLO_TOUCH_WORDS * 32u) ++y.tshift;
    return y;
}

// bucket-list storage per search: list_cap entries (0: 2^20), in chunks
static uint32_t solo_chunks(int32_t list_cap)
{
    const uint64_t want = list_cap > 0 ? (uint64_t)list_cap : (1ull << 20);
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(2, (want + SOLO_CH - 1) / SOLO_CH), 1u << 24);
}

static int solo_reserve_arena(Engine &en, int32_t slots, uint32_t nch)
{
    if (en.solo_arena == nullptr || en.solo_slots < slots || en.solo_nch != nch) {
        if (en.solo_arena) hipFree(en.solo_arena);
        if (en.solo_links) hipFree(en.solo_links);
        en.solo_arena = nullptr; en.solo_links = nullptr;
        HIP_CHECK_RET(hipMalloc((void **)&en.solo_arena, (size_t)slots * nch * SOLO_CH * sizeof(uint64_t)));
        HIP_CHECK_RET(hipMalloc((void **)&en.solo_links, (size_t)slots * 2 * nch * sizeof(uint32_t)));
        en.solo_slots = slots;
        en.solo_nch = nch;
    }
    if (!en.solo_ctl) {
        HIP_CHECK_RET(hipMalloc((void **)&en.solo_ctl, 64));
        HIP_CHECK_RET(hipHostMalloc((void **)&en.h_solo_ctl, 64, hipHostMallocDefault));
    }
    return ORGPU_OK;
}

// keys for `slots` searches at stride NP (the tiled index space can exceed
// the node count the handle reserved for); every key is refilled with
// KEY_INF by the next engine_prepare_keys
static int solo_reserve_keys(Engine &en, int32_t slots, int64_t NP)
{
    return engine_keys_alloc(en, std::max<uint64_t>((uint64_t)slots * (uint64_t)NP, (uint64_t)en.keys_cap));
}

// adjacency records for a key stride of NP
static int solo_reserve_recs(Engine &en, int64_t NP)
{
    if (en.solo_recs_cap >= NP) return ORGPU_OK;
    if (en.solo_recs) hipFree(en.solo_recs);
    en.solo_recs = nullptr;
    en.solo_recs_cap = 0;
    HIP_CHECK_RET(hipMalloc((void **)&en.solo_recs, (size_t)NP * 2 * sizeof(uint4)));
    en.solo_recs_cap = NP;
    return ORGPU_OK;
}

static int solo_reserve_requests(Engine &en, int32_t n_reqs)
{
    if (en.solo_req_cap >= n_reqs) return ORGPU_OK;
    if (en.solo_req) hipFree(en.solo_req);
    if (en.solo_out) hipFree(en.solo_out);
    if (en.h_solo_out) hipHostFree(en.h_solo_out);
    en.solo_req = nullptr; en.solo_out = nullptr; en.h_solo_out = nullptr;
    const int32_t cap = std::max(n_reqs, 4096);
    HIP_CHECK_RET(hipMalloc((void **)&en.solo_req, sizeof(orgpu_net_req) * cap));
    HIP_CHECK_RET(hipMalloc((void **)&en.solo_out, (size_t)3 * cap * sizeof(int32_t)));
    HIP_CHECK_RET(hipHostMalloc((void **)&en.h_solo_out, (size_t)3 * cap * sizeof(int32_t), hipHostMallocDefault));
    en.solo_req_cap = cap;
    return ORGPU_OK;
}

void solo_release(Engine &en)
{
    if (en.solo_arena) hipFree(en.solo_arena);
    if (en.solo_links) hipFree(en.solo_links);
    if (en.solo_req) hipFree(en.solo_req);
    if (en.solo_out) hipFree(en.solo_out);
    if (en.h_solo_out) hipHostFree(en.h_solo_out);
    if (en.solo_ctl) hipFree(en.solo_ctl);
    if (en.h_solo_ctl) hipHostFree(en.h_solo_ctl);
    if (en.solo_recs) hipFree(en.solo_recs);
    en.solo_recs = nullptr; en.solo_recs_cap = 0;
    en.solo_arena = nullptr; en.solo_links = nullptr; en.solo_req = nullptr; en.solo_out = nullptr;
    en.h_solo_out = nullptr;
    en.solo_ctl = nullptr; en.h_solo_ctl = nullptr;
    en.solo_slots = 0; en.solo_nch = 0; en.solo_req_cap = 0;
}

// The kernel arguments of a call (everything but requests and outputs).
static void solo_args(SoloArgs &a, const Engine &en, const orgpu_graph *g, const SoloLayout &lay,
                      const orgpu_sssp_params *p, const uint32_t *roi_bitmaps, int32_t n_roi,
                      const int32_t *keepout_owner)
{
    const int64_t N = g->n_nodes;
    memset(&a, 0, sizeof(a));
    a.indptr = g->indptr; a.indices = g->indices; a.cost = g->cost; a.kind = g->edge_kind;
    a.roi = roi_bitmaps; a.roi_words = (N + 31) / 32;
    a.n_roi = roi_bitmaps ? n_roi : 0;
    a.ko_owner = keepout_owner;
    a.N = N; a.NP = lay.NP;
    a.X = g->X; a.Y = g->Y; a.plane = g->X * g->Y;
    a.tiled = lay.tiled ? 1 : 0; a.TX = lay.TX; a.TY = lay.TY;
    if (lay.tiled || p->h_xy > 0.0f || p->h_z > 0.0f) {
        solo_magic((uint32_t)a.plane, &a.m_plane, &a.s_plane);
        solo_magic((uint32_t)a.X, &a.m_x, &a.s_x);
    }
    a.lookahead = p->lookahead == 0 ? 15 : (p->lookahead < 0 ? 0 : std::min(p->lookahead, 15));
    a.inv_delta = 1.0f / p->delta;
    a.h_xy = p->h_xy; a.h_z = p->h_z;
    a.ctl = en.solo_ctl;
    a.keys = en.keys;
    a.arena = en.solo_arena; a.links = en.solo_links; a.nch = en.solo_nch;
    a.touch_shift = lay.tshift;
    a.stage = en.stage_nodes; a.max_path = en.cap_path;
    a.counters = (unsigned long long *)(en.solo_ctl + 4);
    a.recs = lay.tiled && g->edge_kind && en.solo_recs && en.solo_recs_cap >= lay.NP ? en.solo_recs : nullptr;
}

// writes the call's adjacency records (after k_solo_begin, before the search)
static void solo_build_recs(const SoloArgs &a, hipStream_t s)
{
    if (!a.recs) return;
    int64_t blocks = (a.N + 255) / 256;
    if (blocks > 65535 * 8) blocks = 65535 * 8;
    hipLaunchKernelGGL(k_solo_adjacency, dim3((unsigned)blocks), dim3(256), 0, s, a, const_cast<uint4 *>(a.recs));
}

// wavefronts per search: params->waves, or automatic.  A call whose
// searches all start at once (n <= slots) ends with its longest search, so
// up to 1024 of them (as many four-wave workgroups as the GPU holds at once)
// take four waves each; a call whose slots refill is bound by the memory
// system's request rate, where two waves per search measured best.
static int solo_waves(const orgpu_sssp_params *p, int slots, int n)
{
    if (p->waves == 1 || p->waves == 2 || p->waves == 4) return p->waves;
    return n <= slots && n <= 1024 ? 4 : 2;
}

static void solo_launch(const SoloArgs &a, int grid, int waves, hipStream_t s)
{
    if (waves == 4) hipLaunchKernelGGL(k_solo_mw<4>, dim3(grid), dim3(256), 0, s, a);
    else if (waves == 2) hipLaunchKernelGGL(k_solo_mw<2>, dim3(grid), dim3(128), 0, s, a);
    else hipLaunchKernelGGL(k_solo, dim3(grid), dim3(64), 0, s, a);
}

int engine_reserve_solo(Engine &en, int64_t n_nodes, int32_t X, int32_t Y, int32_t L, int32_t slots, int32_t list_cap)
{
    if (en.cap_slots == 0) return orgpu_fail(ORGPU_EINVAL, "reserve_solo: call orgpu_reserve first");
    if (slots < 1 || slots > en.cap_slots || n_nodes < 1 || n_nodes > en.cap_nodes || list_cap < 0)
        return orgpu_fail(ORGPU_EINVAL, "reserve_solo: (nodes %lld, slots %d, list_cap %d) outside the handle's "
                          "reserve (nodes %lld, slots %d)", (long long)n_nodes, slots, list_cap,
                          (long long)en.cap_nodes, en.cap_slots);
    const SoloLayout lay = solo_layout(n_nodes, X, Y, L);
    int rc = solo_reserve_keys(en, slots, lay.NP);
    if (rc) return rc;
    rc = solo_reserve_arena(en, slots, solo_chunks(list_cap));
    if (rc) return rc;
    rc = solo_reserve_recs(en, lay.NP);
    if (rc) return rc;
    return engine_prepare_keys(en, n_nodes);
}

int engine_route_async(Engine &en, const orgpu_graph *g, const orgpu_net_req *d_reqs, int32_t n,
                       const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                       const orgpu_sssp_params *p, int32_t *path_nodes, int32_t *path_edges, int64_t capacity,
                       uint32_t *path_off, int32_t *path_len, float *path_dist, uint32_t *used)
{
    const int64_t N = g->n_nodes;
    const int K = std::min(p->slots, en.cap_slots);
    const SoloLayout lay = solo_layout(N, g->X, g->Y, g->L);
    if (!en.solo_arena || en.solo_slots < K || en.solo_nch != solo_chunks(p->list_cap) ||
        (uint64_t)K * (uint64_t)lay.NP > (uint64_t)en.keys_cap)
        return orgpu_fail(ORGPU_EINVAL, "route_async: call orgpu_reserve_solo for this graph, %d slots and list_cap %d "
                          "first", K, p->list_cap);
    if (capacity < 0 || capacity > (int64_t)UINT32_MAX)
        return orgpu_fail(ORGPU_EINVAL, "route_async: capacity %lld outside [0, 2^32)", (long long)capacity);
    hipStream_t s = en.stream;
    int rc = engine_prepare_keys(en, N);
    if (rc) return rc;
    SoloArgs a;
    solo_args(a, en, g, lay, p, roi_bitmaps, n_roi, keepout_owner);
    a.reqs = d_reqs; a.n_reqs = n;
    a.pack_nodes = path_nodes; a.pack_edges = path_edges; a.pack_cap = (uint64_t)capacity;
    a.out_len = path_len; a.out_off = path_off; a.out_dist = path_dist;
    hipLaunchKernelGGL(k_solo_begin, dim3(1), dim3(64), 0, s, en.solo_ctl);
    if (n > 0) {
        solo_build_recs(a, s);
        solo_launch(a, std::min(K, n), solo_waves(p, K, n), s);
    }
    if (used) hipLaunchKernelGGL(k_solo_used, dim3(1), dim3(64), 0, s, en.solo_ctl, used);
    HIP_CHECK_RET(hipGetLastError());
    return ORGPU_OK;
}

// Two-pin requests on the solo engine, blocking (orgpu_route with engine 1).
// Arguments were checked by engine_route; outputs as engine_route writes
// them.  Requests the engine gives back for capacity are routed on the
// bucket-round engine.
int engine_route_solo(Engine &en, const orgpu_graph *g, const orgpu_net_req *nets, int32_t n_nets,
                      const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner, float keepout_cost,
                      const orgpu_sssp_params *p, int32_t *path_nodes, int32_t *path_edges, int64_t path_capacity,
                      int64_t *path_offset, float *path_dist, orgpu_sssp_stats *stats)
{
    const int64_t N = g->n_nodes;
    const int K = std::min(p->slots, en.cap_slots);
    hipStream_t s = en.stream;
    path_offset[0] = 0;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (n_nets == 0) return ORGPU_OK;
    const SoloLayout lay = solo_layout(N, g->X, g->Y, g->L);
    int rc = solo_reserve_keys(en, K, lay.NP);
    if (rc) return rc;
    rc = solo_reserve_arena(en, K, solo_chunks(p->list_cap));
    if (rc) return rc;
    rc = solo_reserve_requests(en, n_nets);
    if (rc) return rc;
    if (lay.tiled) {
        rc = solo_reserve_recs(en, lay.NP);
        if (rc) return rc;
    }
    rc = engine_prepare_keys(en, N);
    if (rc) return rc;
    const int64_t cap = en.solo_req_cap;
    const uint64_t half = (uint64_t)en.cap_slots * (uint64_t)en.cap_path;

    SoloArgs a;
    solo_args(a, en, g, lay, p, roi_bitmaps, n_roi, keepout_owner);
    a.reqs = en.solo_req; a.n_reqs = n_nets;
    a.pack_nodes = en.pack; a.pack_edges = en.pack + half; a.pack_cap = half;
    a.out_len = en.solo_out;
    a.out_off = (uint32_t *)(en.solo_out + cap);
    a.out_dist = (float *)(en.solo_out + 2 * cap);

    HIP_CHECK_RET(hipMemcpyAsync(en.solo_req, nets, sizeof(orgpu_net_req) * n_nets, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_solo_begin, dim3(1), dim3(64), 0, s, en.solo_ctl);
    solo_build_recs(a, s);
    HIP_CHECK_RET(hipEventRecord(en.ev_t0, s));
    solo_launch(a, std::min(K, n_nets), solo_waves(p, K, n_nets), s);
    HIP_CHECK_RET(hipGetLastError());
    HIP_CHECK_RET(hipEventRecord(en.ev_r1, s));
    HIP_CHECK_RET(hipMemcpyAsync(en.h_solo_ctl, en.solo_ctl, 64, hipMemcpyDeviceToHost, s));
    for (int part = 0; part < 3; ++part)
        HIP_CHECK_RET(hipMemcpyAsync(en.h_solo_out + part * cap, en.solo_out + part * cap,
                                     sizeof(int32_t) * (size_t)n_nets, hipMemcpyDeviceToHost, s));
    HIP_CHECK_RET(hipStreamSynchronize(s));
    const uint32_t used = (uint32_t)std::min<uint64_t>(en.h_solo_ctl[1], half);
    if (used > 0) {
        HIP_CHECK_RET(hipMemcpyAsync(en.h_stage, en.pack, sizeof(int32_t) * used, hipMemcpyDeviceToHost, s));
        if (path_edges)
            HIP_CHECK_RET(hipMemcpyAsync(en.h_stage + half, en.pack + half, sizeof(int32_t) * used,
                                         hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK_RET(hipEventRecord(en.ev_t1, s));
    HIP_CHECK_RET(hipStreamSynchronize(s));
    const int32_t *h_len = en.h_solo_out;
    const uint32_t *h_off = (const uint32_t *)(en.h_solo_out + cap);
    const float *h_dist = (const float *)(en.h_solo_out + 2 * cap);
    const unsigned long long *h_cnt = (const unsigned long long *)(en.h_solo_ctl + 4);
    float kms = 0.0f, rms = 0.0f;
    hipEventElapsedTime(&kms, en.ev_t0, en.ev_t1);
    hipEventElapsedTime(&rms, en.ev_t0, en.ev_r1);

    std::vector<int32_t> retry;
    for (int i = 0; i < n_nets; ++i) {
        if (h_len[i] == SOLO_KO_SATURATED)
            return orgpu_fail(ORGPU_ECAPACITY, "route: request %d: its path crosses 255 or more keepout edges, "
                              "which the search key cannot count", i);
        if (h_len[i] == SOLO_BROKEN)
            return orgpu_fail(ORGPU_ECAPACITY, "route: request %d: predecessor chain broken or longer than %d", i,
                              en.cap_path);
        if (h_len[i] == SOLO_INVALID)
            return orgpu_fail(ORGPU_EINVAL, "route: request %d outside the graph or the ROI rows", i);
        if (h_len[i] == SOLO_RETRY) retry.push_back(i);
    }
    // requests given back for capacity: routed on the bucket-round engine
    // (which reuses the pinned staging, so the packed paths are kept first)
    orgpu_sssp_stats st2;
    memset(&st2, 0, sizeof(st2));
    std::vector<int64_t> off2(retry.size() + 1, 0);
    std::vector<float> dist2(retry.size());
    std::vector<int32_t> n2, e2, keep_n, keep_e;
    const int32_t *stage_n = en.h_stage, *stage_e = en.h_stage + half;
    if (!retry.empty()) {
        keep_n.assign(en.h_stage, en.h_stage + used);
        if (path_edges) keep_e.assign(en.h_stage + half, en.h_stage + half + used);
        stage_n = keep_n.data();
        stage_e = keep_e.data();
        std::vector<orgpu_net_req> sub(retry.size());
        for (size_t j = 0; j < retry.size(); ++j) sub[j] = nets[retry[j]];
        orgpu_sssp_params p2 = *p;
        p2.engine = 0;
        p2.slots = std::min(p->slots, en.cap_bslots);
        n2.resize((size_t)std::max<int64_t>(path_capacity, 1));
        if (path_edges) e2.resize(n2.size());
        rc = engine_route(en, g, sub.data(), (int32_t)sub.size(), roi_bitmaps, n_roi, keepout_owner, keepout_cost,
                          &p2, n2.data(), path_edges ? e2.data() : nullptr, path_capacity, off2.data(), dist2.data(),
                          &st2);
        if (rc) return rc;
    }
    int64_t pos = 0;
    size_t j = 0;
    for (int i = 0; i < n_nets; ++i) {
        path_offset[i] = pos;
        const int32_t *src_n = nullptr, *src_e = nullptr;
        int64_t len = 0;
        float d = 0.0f;
        if (j < retry.size() && retry[j] == i) {
            len = off2[j + 1] - off2[j];
            src_n = n2.data() + off2[j];
            if (path_edges) src_e = e2.data() + off2[j];
            d = dist2[j];
            ++j;
        } else {
            len = h_len[i];
            src_n = stage_n + h_off[i];
            src_e = stage_e + h_off[i];
            d = h_dist[i];
        }
        if (len > 0) {
            if (pos + len > path_capacity)
                return orgpu_fail(ORGPU_ECAPACITY, "route: path output capacity %lld exceeded", (long long)path_capacity);
            memcpy(path_nodes + pos, src_n, sizeof(int32_t) * len);
            if (path_edges) memcpy(path_edges + pos, src_e, sizeof(int32_t) * len);
            pos += len;
        }
        if (path_dist) path_dist[i] = d;
    }
    path_offset[n_nets] = pos;
    if (stats) {
        stats->rounds = st2.rounds;
        stats->items = (int64_t)h_cnt[0] + st2.items;
        stats->edges = (int64_t)h_cnt[1] + st2.edges;
        stats->kernel_ms = kms + st2.kernel_ms;
        stats->launches = 1 + st2.launches;
        stats->round_ms = rms + st2.round_ms;
        stats->retried = (int64_t)retry.size();
    }
    return ORGPU_OK;
}
