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
// ascending), with the must-keeps (src, dst and up to two portal seeds) beyond
// the cut swapped into the list's first positions (the ids there leave); the
// caller's re-add then puts back the seeds, src and dst when asked.
#include <algorithm>
#include <vector>

#include "common.h"
#include "engine.h"

#ifndef ORGPU_ROI_WG
#define ORGPU_ROI_WG 1024
#endif
#define RWG ORGPU_ROI_WG

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

// The reference's list rules on a finished ROI row A holding `seen` marked
// nodes (UPF:1700-1734, 4659-4700), then the request's info record.  The
// list is the seen nodes ascending, the must-keeps it lacks appended; when it
// is longer than max_nodes, each must-keep beyond the cut trades places with
// the list's position p (p = its rank among the must-keeps by position), so
// it stays and the node there leaves, and the list is cut.  With readd the
// keeps, then src and dst, are appended when missing.  Called by every
// thread of the workgroup.
struct RoiShared {
    int wsum[RWG / 64];
    int cut;        // node at sorted position max_nodes - 1 (the last one kept)
    int mins[4];    // nodes at sorted positions 0..3
};

__device__ __forceinline__ bool row_has(const uint32_t *A, int32_t v) { return (A[v >> 5] >> (v & 31)) & 1u; }
__device__ __forceinline__ void row_set(uint32_t *A, int32_t v) { A[v >> 5] |= 1u << (v & 31); }
__device__ __forceinline__ void row_clear(uint32_t *A, int32_t v) { A[v >> 5] &= ~(1u << (v & 31)); }

__device__ void roi_finish(uint32_t *A, int64_t words, const orgpu_roi_req &q, long long seen, int depth, int met,
                           orgpu_roi_info *info, RoiShared &sh)
{
    const int tid = threadIdx.x;
    // must-keeps in the reference's order: src, dst, then the keeps not yet listed
    int32_t mk[4];
    int nmk = 0;
    {
        const int32_t cand[4] = {q.src, q.dst, q.keep[0], q.keep[1]};
        for (int i = 0; i < 4; ++i) {
            const int32_t v = cand[i];
            bool dup = v < 0;
            for (int j = 0; j < nmk; ++j) dup |= mk[j] == v;
            if (!dup) mk[nmk++] = v;
        }
    }
    bool mseen[4] = {false, false, false, false};
    int nmiss = 0;
    for (int j = 0; j < nmk; ++j) {
        mseen[j] = row_has(A, mk[j]);
        nmiss += !mseen[j];
    }
    __syncthreads();   // every thread has read the must-keep bits before any change
    const long long R = seen, total = seen + nmiss;
    const bool cut_seen = R > q.max_nodes;
    if (total > q.max_nodes) {
        // Rank the seen nodes by id.  Each wave owns a contiguous chunk of
        // the row and its lanes read consecutive words.  Pass 1 counts the
        // chunk's bits; the prefix of the chunk totals gives each chunk's
        // first rank; pass 2, in the chunks that hold a wanted rank (1-4:
        // the list's first positions, and max_nodes: the cut), scans again
        // with a running prefix to locate those bits.
        const int wid = tid >> 6, lane = tid & 63;
        const int64_t per = (words + RWG / 64 - 1) / (RWG / 64);
        const int64_t c0 = std::min<int64_t>((int64_t)wid * per, words);
        const int64_t c1 = std::min<int64_t>(c0 + per, words);
        int c = 0;
        for (int64_t w = c0 + lane; w < c1; w += 64) c += __popc(A[w]);
        const int tot = __shfl(wave_scan_incl(c), 63, 64);
        if (lane == 0) sh.wsum[wid] = tot;
        if (tid == 0) { sh.cut = -1; for (int k = 0; k < 4; ++k) sh.mins[k] = -1; }
        __syncthreads();
        long long base = 0;
        for (int j = 0; j < wid; ++j) base += sh.wsum[j];
        const long long targets[5] = {cut_seen ? (long long)q.max_nodes : -1, 1, 2, 3, 4};
        bool mine = false;
        for (int k = 0; k < 5; ++k) mine |= base < targets[k] && targets[k] <= base + tot;
        if (mine) {
            long long run = base;
            for (int64_t s0 = c0; s0 < c1 && run < base + tot; s0 += 64) {
                const int64_t w = s0 + lane;
                const uint32_t x = w < c1 ? A[w] : 0u;
                const int pc = __popc(x);
                const int incl = wave_scan_incl(pc);
                const long long lo = run + incl - pc;   // seen nodes before this lane's word
                for (int k = 0; k < 5; ++k) {
                    const long long r = targets[k];
                    if (lo < r && r <= lo + pc) {
                        const int32_t id = (int32_t)(w * 32 + nth_set_bit(x, (int)(r - lo) - 1));
                        if (k == 0) sh.cut = id;
                        else sh.mins[k - 1] = id;
                    }
                }
                run += __shfl(incl, 63, 64);
            }
        }
        __syncthreads();
        if (cut_seen) {
            // the bits above the cut leave the row (the words from the cut's on)
            const int32_t cut = sh.cut;
            for (int64_t w = std::max<int64_t>(0, (int64_t)cut >> 5) + tid; w < words; w += RWG) {
                const int64_t keep_bits = (int64_t)cut - w * 32 + 1;   // ids w*32 .. cut stay
                const uint32_t keep = keep_bits <= 0 ? 0u : keep_bits >= 32 ? ~0u : (1u << keep_bits) - 1u;
                const uint32_t x = A[w];
                if (x & ~keep) A[w] = x & keep;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        int32_t front[4] = {-1, -1, -1, -1}, left[4] = {-1, -1, -1, -1}, app[4] = {-1, -1, -1, -1};
        int na = 0;
        // the must-keeps in list order: the seen ones by id, then the missing
        // ones at positions R, R+1, ...; `beyond` = past the cut
        int32_t kn[4], miss[4];
        bool beyond[4];
        int nk = 0, nm = 0;
        for (int j = 0; j < nmk; ++j) {
            if (!mseen[j]) continue;
            int at = nk++;
            while (at > 0 && kn[at - 1] > mk[j]) { kn[at] = kn[at - 1]; beyond[at] = beyond[at - 1]; --at; }
            kn[at] = mk[j];
            beyond[at] = cut_seen && mk[j] > sh.cut;
        }
        long long pos = R;
        for (int j = 0; j < nmk; ++j) {
            if (mseen[j]) continue;
            kn[nk] = mk[j];
            miss[nm++] = mk[j];
            beyond[nk] = pos >= q.max_nodes;
            if (!beyond[nk]) { row_set(A, mk[j]); if (na < 4) app[na++] = mk[j]; }
            ++nk; ++pos;
        }
        long long kept = total;
        if (total > q.max_nodes) {
            for (int s2 = 0; s2 < nk; ++s2) {
                // a must-keep past the cut trades places with position s2; when
                // s2 itself is past the cut both stay out
                if (!beyond[s2] || s2 >= q.max_nodes) continue;
                // position s2 holds the s2-th lowest seen id, or an appended
                // must-keep when the seen part is shorter
                const int32_t out = s2 < R ? sh.mins[s2] : miss[s2 - R];
                if (out >= 0) row_clear(A, out);
                row_set(A, kn[s2]);
                front[s2] = kn[s2];
                left[s2] = out;
            }
            kept = q.max_nodes;
        }
        if (q.readd) {
            // _route_all appends the portal seeds, then src and dst, when missing
            const int32_t ra[4] = {q.keep[0], q.keep[1], q.src, q.dst};
            for (int j = 0; j < 4; ++j) {
                const int32_t v = ra[j];
                if (v < 0 || row_has(A, v)) continue;
                row_set(A, v);
                kept++;
                if (na < 4) app[na++] = v;
            }
        }
        if (info) {
            info->seen = seen;
            info->kept = kept;
            info->depth = depth;
            info->met = met;
            info->src_in = row_has(A, q.src);
            info->dst_in = row_has(A, q.dst);
            for (int j = 0; j < 4; ++j) {
                info->front[j] = front[j];
                info->left[j] = left[j];
                info->appended[j] = app[j];
            }
        }
    }
    __syncthreads();
}

__global__ void __launch_bounds__(RWG)
k_roi(RoiArgs a)
{
    __shared__ int s_na, s_nb, s_met;   // s_na + s_nb = nodes seen
    __shared__ int s_fa0, s_fa1, s_fb0, s_fb1, s_depth;
    __shared__ RoiShared sh;
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
        roi_finish(A, a.words, q, (long long)s_na + s_nb, s_depth, s_met, a.info ? a.info + r : nullptr, sh);
    }
}


// ---------------------------------------------------------------------------
// ROI on the lattice itself, in closed form.
//
// Every layer of the routing lattice is a full X x Y grid (planar edges along
// its direction, or none) and vias join layer pairs at every cell, so the hop
// distance from (x0, y0, z0) to (x, y, z) is
//     dx + dy + T[z0][need][z],   need = (dx > 0) | (dy > 0) << 1,
// with T[z0][need][z] the fewest via hops from layer z0 to layer z through
// layers that include a planar layer allowing x moves if dx > 0 and one
// allowing y moves if dy > 0 (a breadth-first search over (layer, covered
// directions) on the host).  The bidirectional BFS of the reference then
// marks ball_src(ks) | ball_dst(kd): before the waves meet each wave's marks
// are a ball; the src step j meets iff the src-dst distance D = 2j - 1 and
// the dst step j iff D = 2j; the early stop needs only the ball sizes, which
// follow from counting grid columns by planar distance (clipped to the
// board) times layers by via hops.  So a net costs a few thousand integer
// operations plus one pass over the words of its ROI's bounding box,
// instead of a traversal of ~75,000 nodes' edges.
// ---------------------------------------------------------------------------
#define ROI_MAXD 1024      // deepest level the closed form evaluates
#define ROI_TK 256         // via-hop values; 255 = unreachable

struct LatRoiArgs {
    int32_t X, Y, L;
    int64_t N, words;
    const uint8_t *T;      // [L][4][L]
    const orgpu_roi_req *reqs;
    int n;
    uint32_t *rows;
    int32_t n_roi;
    orgpu_roi_info *info;
    uint32_t *next;
};

__device__ __forceinline__ int lat_hops(int x0, int y0, const uint8_t *Tz, int L, int x, int y, int z)
{
    const int dx = abs(x - x0), dy = abs(y - y0);
    const int t = Tz[((dx > 0) | ((dy > 0) << 1)) * L + z];
    return t == 255 ? 0x3FFFFFFF : dx + dy + t;
}

__global__ void __launch_bounds__(RWG)
k_roi_lattice(LatRoiArgs a)
{
    __shared__ int s_next;
    __shared__ uint8_t s_T[2][4 * ORGPU_MAX_LAYERS];
    __shared__ int s_zc[2][4][ROI_TK];
    __shared__ int s_h[2][ROI_MAXD + 1];
    __shared__ int s_kmax, s_ks, s_kd, s_depth, s_met;
    __shared__ unsigned long long s_seen;
    __shared__ RoiShared sh;
    const int tid = threadIdx.x;
    const int lane = (int)__lane_id();
    const int L = a.L, X = a.X, Y = a.Y;
    const int64_t plane = (int64_t)X * Y;
    for (;;) {
        if (tid == 0) s_next = (int)atomicAdd(a.next, 1u);
        __syncthreads();
        const int r = s_next;
        __syncthreads();
        if (r >= a.n) break;
        const orgpu_roi_req q = a.reqs[r];
        uint32_t *A = a.rows + (int64_t)q.roi_index * a.words;
        const int z0 = (int)(q.src / plane), z1 = (int)(q.dst / plane);
        const int x0 = (int)((q.src % plane) % X), y0 = (int)((q.src % plane) / X);
        const int x1 = (int)((q.dst % plane) % X), y1 = (int)((q.dst % plane) / X);
        const int md = min(q.max_depth, ROI_MAXD);
        for (int i = tid; i < 4 * L; i += RWG) {
            s_T[0][i] = a.T[(int64_t)z0 * 4 * L + i];
            s_T[1][i] = a.T[(int64_t)z1 * 4 * L + i];
        }
        for (int i = tid; i < 2 * 4 * ROI_TK; i += RWG) (&s_zc[0][0][0])[i] = 0;
        for (int i = tid; i < 2 * (ROI_MAXD + 1); i += RWG) (&s_h[0][0])[i] = 0;
        if (tid == 0) { s_seen = 0ull; s_kmax = 0; }
        __syncthreads();
        // layers by via hops: zc[w][need][t] = #{z : T[zw][need][z] == t}
        for (int i = tid; i < 2 * 4 * L; i += RWG) {
            const int w = i / (4 * L), rem = i - w * 4 * L;
            const int t = s_T[w][rem];
            if (t != 255) {
                atomicAdd(&s_zc[w][rem / L][t], 1);
                atomicMax(&s_kmax, t);
            }
        }
        __syncthreads();
        // ball sizes: h[w][j] = nodes at distance exactly j (j <= md) from
        // src (w = 0) or dst (w = 1): columns at planar distance m, by
        // whether dx and dy are nonzero, times layers at via distance j - m
        const int kmax = s_kmax;
        for (int i = tid; i < 2 * (md + 1); i += RWG) {
            const int w = i / (md + 1), m = i - w * (md + 1);
            const int cx = w ? x1 : x0, cy = w ? y1 : y0;
            auto nx = [&](int d) { return (cx - d >= 0 ? 1 : 0) + (cx + d <= X - 1 ? 1 : 0); };
            auto ny = [&](int d) { return (cy - d >= 0 ? 1 : 0) + (cy + d <= Y - 1 ? 1 : 0); };
            int cnt[4];
            cnt[0] = m == 0 ? 1 : 0;
            cnt[1] = m > 0 ? nx(m) : 0;
            cnt[2] = m > 0 ? ny(m) : 0;
            int c3 = 0;
            for (int d = 1; d < m; ++d) c3 += nx(d) * ny(m - d);
            cnt[3] = c3;
            for (int c = 0; c < 4; ++c) {
                if (!cnt[c]) continue;
                for (int t = 0; t <= kmax && m + t <= md; ++t) {
                    const int z = s_zc[w][c][t];
                    if (z) atomicAdd(&s_h[w][m + t], cnt[c] * z);
                }
            }
        }
        __syncthreads();
        if (tid == 0) {
            // the reference's level loop (UPF:1653-1700) on the ball sizes
            const int D = lat_hops(x0, y0, s_T[0], L, x1, y1, z1);
            const bool same = q.src == q.dst;
            const double stop = (double)q.max_nodes * 1.5;
            long long seen = same ? 1 : 2;
            int fa = 1, fb = 1, depth = 0, met = 0, ks = 0, kd = 0;
            while (depth < md && (fa > 0 || fb > 0) && !met) {
                const int j = depth + 1;
                if (fa > 0) {
                    if (!same && D == 2 * j - 1) met = 1;
                    fa = s_h[0][j];
                    seen += fa;
                }
                ks = j;
                if (!met) {
                    if (same) {
                        // both waves start at the node: the dst wave's first
                        // step touches the src wave's marks
                        if (s_h[0][1] > 0) met = 1;
                        fb = 0;
                    } else if (fb > 0) {
                        if (D == 2 * j) met = 1;
                        fb = s_h[1][j];
                        seen += fb;
                    }
                    kd = j;
                }
                depth = j;
                if ((double)seen > stop) break;
            }
            s_ks = ks; s_kd = kd; s_depth = depth; s_met = met;
        }
        for (int64_t w = tid; w < a.words; w += RWG) A[w] = 0u;
        __syncthreads();
        // the row: words over the bounding box of both balls, all layers;
        // a wave takes one (layer, y) row segment, each half-wave one word,
        // one lane per node.  A word shared with the previous segment (rows
        // shorter than a word apart) belongs to that segment only.
        const int ks = s_ks, kd = s_kd;
        const int xlo = max(0, min(x0 - ks, x1 - kd)), xhi = min(X - 1, max(x0 + ks, x1 + kd));
        const int ylo = max(0, min(y0 - ks, y1 - kd)), yhi = min(Y - 1, max(y0 + ks, y1 + kd));
        const int nyb = yhi - ylo + 1;
        const int nrows = L * nyb;
        unsigned long long my_seen = 0;
        for (int row = tid >> 6; row < nrows; row += RWG / 64) {
            const int z = row / nyb, y = ylo + (row - z * nyb);
            const int64_t base = (int64_t)z * plane + (int64_t)y * X;
            int64_t wa = (base + xlo) >> 5;
            const int64_t wb = (base + xhi) >> 5;
            int64_t pbase = -1;
            if (y > ylo) pbase = base - X;
            else if (z > 0) pbase = (int64_t)(z - 1) * plane + (int64_t)yhi * X;
            if (pbase >= 0) wa = max(wa, ((pbase + xhi) >> 5) + 1);
            for (int64_t wp = wa; wp <= wb; wp += 2) {
                const int64_t w = wp + (lane >> 5);
                const int64_t node = w * 32 + (lane & 31);
                bool in = false;
                if (w <= wb && node < a.N) {
                    const int nz = (int)(node / plane);
                    const int64_t rem = node - (int64_t)nz * plane;
                    const int ny_ = (int)(rem / X), nx_ = (int)(rem - (int64_t)ny_ * X);
                    in = lat_hops(x0, y0, s_T[0], L, nx_, ny_, nz) <= ks ||
                         lat_hops(x1, y1, s_T[1], L, nx_, ny_, nz) <= kd;
                }
                const uint64_t bal = __ballot(in);
                if ((lane & 31) == 0 && w <= wb) {
                    const uint32_t word = lane < 32 ? (uint32_t)bal : (uint32_t)(bal >> 32);
                    A[w] = word;
                    my_seen += (unsigned long long)__popc(word);
                }
            }
        }
        if (my_seen) atomicAdd(&s_seen, my_seen);
        __syncthreads();
        roi_finish(A, a.words, q, (long long)s_seen, s_depth, s_met, a.info ? a.info + r : nullptr, sh);
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
            q.max_nodes < 1 || q.max_depth < 0 || q.keep[0] < -1 || q.keep[0] >= N || q.keep[1] < -1 || q.keep[1] >= N)
            return orgpu_fail(ORGPU_EINVAL, "roi_bfs: request %d invalid (src %d dst %d keeps %d %d row %d of %d, "
                              "max_nodes %lld)", i, q.src, q.dst, q.keep[0], q.keep[1], q.roi_index, n_roi,
                              (long long)q.max_nodes);
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

// Fewest via hops between layers covering the needed planar directions
// (bit 0: x moves on an 'h' or 'hv' layer, bit 1: y moves on a 'v' or 'hv'
// layer): T[z0][need][z], 255 where no layer sequence exists.  Each via
// pair joins its two layers in both directions.
static int lattice_hop_table(const orgpu_lattice_desc *d, std::vector<uint8_t> &T)
{
    const int L = d->L;
    std::vector<int> cov(L, 0);
    for (int z = 0; z < L; ++z)
        if (d->planar[z]) cov[z] = d->dir[z] == 1 ? 1 : d->dir[z] == 0 ? 2 : 3;
    std::vector<std::vector<int>> adj(L);
    for (int i = 0; i < d->npairs; ++i) {
        const int f = d->pair_from[i], t = d->pair_to[i];
        adj[f].push_back(t);
        adj[t].push_back(f);
    }
    T.assign((size_t)L * 4 * L, 255);
    std::vector<int> dist((size_t)L * 4);
    std::vector<int> cur, nxt;
    for (int z0 = 0; z0 < L; ++z0) {
        std::fill(dist.begin(), dist.end(), -1);
        cur.assign(1, z0 * 4 + cov[z0]);
        dist[z0 * 4 + cov[z0]] = 0;
        while (!cur.empty()) {
            nxt.clear();
            for (int st : cur) {
                const int z = st >> 2, m = st & 3;
                for (int b : adj[z]) {
                    const int ns = b * 4 + (m | cov[b]);
                    if (dist[ns] < 0) { dist[ns] = dist[st] + 1; nxt.push_back(ns); }
                }
            }
            cur.swap(nxt);
        }
        for (int z = 0; z < L; ++z)
            for (int m = 0; m < 4; ++m) {
                const int dd = dist[z * 4 + m];
                if (dd < 0) continue;
                if (dd > 254) return orgpu_fail(ORGPU_EINVAL, "roi_lattice: via distance %d too large", dd);
                for (int need = 0; need < 4; ++need)
                    if ((need & ~m) == 0) {
                        uint8_t &e = T[((size_t)z0 * 4 + need) * L + z];
                        if (dd < e) e = (uint8_t)dd;
                    }
            }
    }
    return ORGPU_OK;
}

extern "C" int orgpu_roi_lattice(orgpu_t *h, const orgpu_lattice_desc *d, const orgpu_roi_req *reqs, int32_t n,
                                 uint32_t *roi_bitmaps, int32_t n_roi, orgpu_roi_info *info)
{
    if (!h || !d || n < 0 || (n > 0 && (!reqs || !roi_bitmaps))) return orgpu_fail(ORGPU_EINVAL, "roi_lattice: bad argument");
    if (n == 0) return ORGPU_OK;
    int64_t N = 0, E = 0;
    int rc = orgpu_lattice_count(d, &N, &E);
    if (rc) return rc;
    if (N >= KEY_MAX_NODES) return orgpu_fail(ORGPU_EINVAL, "roi_lattice: lattice too large");
    for (int i = 0; i < d->npairs; ++i)
        if (!d->pair_from || !d->pair_to || d->pair_from[i] < 0 || d->pair_from[i] >= d->L || d->pair_to[i] < 0 ||
            d->pair_to[i] >= d->L)
            return orgpu_fail(ORGPU_EINVAL, "roi_lattice: via pair %d outside the layers", i);
    std::vector<int> row_user(n_roi > 0 ? n_roi : 0, -1);
    for (int i = 0; i < n; ++i) {
        const orgpu_roi_req &q = reqs[i];
        if (q.src < 0 || q.src >= N || q.dst < 0 || q.dst >= N || q.roi_index < 0 || q.roi_index >= n_roi ||
            q.max_nodes < 1 || q.max_depth < 0 || q.keep[0] < -1 || q.keep[0] >= N || q.keep[1] < -1 || q.keep[1] >= N)
            return orgpu_fail(ORGPU_EINVAL, "roi_lattice: request %d invalid (src %d dst %d keeps %d %d row %d of %d, "
                              "max_nodes %lld)", i, q.src, q.dst, q.keep[0], q.keep[1], q.roi_index, n_roi,
                              (long long)q.max_nodes);
        if (q.max_depth > ROI_MAXD)
            return orgpu_fail(ORGPU_EINVAL, "roi_lattice: request %d depth %d above %d (use orgpu_roi_bfs)", i,
                              q.max_depth, ROI_MAXD);
        if (row_user[q.roi_index] >= 0)
            return orgpu_fail(ORGPU_EINVAL, "roi_lattice: requests %d and %d share bitmap row %d", row_user[q.roi_index],
                              i, q.roi_index);
        row_user[q.roi_index] = i;
    }
    std::vector<uint8_t> T;
    rc = lattice_hop_table(d, T);
    if (rc) return rc;
    HIP_CHECK_RET(hipSetDevice(h->device));
    Engine &en = h->en;
    rc = ensure_roi_staging(en, n);
    if (rc) return rc;
    if (!en.roi_T) HIP_CHECK_RET(hipMalloc((void **)&en.roi_T, (size_t)ORGPU_MAX_LAYERS * 4 * ORGPU_MAX_LAYERS));
    hipStream_t s = h->stream;
    HIP_CHECK_RET(hipMemcpyAsync(en.roi_T, T.data(), T.size(), hipMemcpyHostToDevice, s));
    HIP_CHECK_RET(hipMemcpyAsync(en.roi_req, reqs, sizeof(orgpu_roi_req) * n, hipMemcpyHostToDevice, s));
    HIP_CHECK_RET(hipMemsetAsync(en.roi_next, 0, sizeof(uint32_t), s));
    LatRoiArgs a;
    a.X = d->X; a.Y = d->Y; a.L = d->L; a.N = N; a.words = (N + 31) / 32;
    a.T = en.roi_T; a.reqs = en.roi_req; a.n = n; a.rows = roi_bitmaps; a.n_roi = n_roi;
    a.info = en.roi_info; a.next = en.roi_next;
    hipLaunchKernelGGL(k_roi_lattice, dim3(std::min(n, 1024)), dim3(RWG), 0, s, a);
    HIP_CHECK_RET(hipGetLastError());
    if (info) HIP_CHECK_RET(hipMemcpyAsync(info, en.roi_info, sizeof(orgpu_roi_info) * n, hipMemcpyDeviceToHost, s));
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
