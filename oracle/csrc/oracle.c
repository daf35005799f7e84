/*
 * CPU oracle for the PathFinder inner loop.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the reference algorithms of
 * orthoroute/algorithms/manhattan/unified_pathfinder.py (cited as UPF:line).
 * It is the checker the GPU path is compared against; nothing in the product
 * package links it.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg load it.
 *
 *   orc_lattice_*   Lattice3D.build_graph + CSRGraph.finalize   UPF:1195-1338, 666-792
 *   orc_sssp_canon  FP32 shortest path, packed (dist, lowest pred) semantics
 *                   (the GPU contract; same search space as SimpleDijkstra)
 *   orc_sssp_ref    SimpleDijkstra.find_path_roi, emulated exactly       UPF:1801-1887
 *   orc_roi_bfs     ROIExtractor.extract_roi_bfs                         UPF:1636-1734
 *   orc_ms_ref      SimpleDijkstra.find_path_multisource_multisink, exact UPF:1889-1989
 *   orc_ms_canon    the same search in the canonical FP32 semantics (GPU contract)
 *
 * Built with gcc -O2 (SSE float arithmetic, no excess precision, no FMA
 * contraction: -ffp-contract=off).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define DIR_V 0
#define DIR_H 1
#define DIR_HV 2

/* ------------------------------------------------------------------------ */
/* Lattice                                                                   */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t src, dst;
    float cost;
} edge_t;

/* Number of directed edges the reference inserts (UPF:1206-1222). */
int64_t orc_lattice_edge_count(int X, int Y, int L, const uint8_t *planar,
                               const uint8_t *dir, int npairs)
{
    int64_t e = 0;
    for (int z = 0; z < L; ++z) {
        if (!planar[z]) continue;
        if (dir[z] == DIR_H || dir[z] == DIR_HV) e += 2LL * Y * (X - 1);
        if (dir[z] == DIR_V || dir[z] == DIR_HV) e += 2LL * X * (Y - 1);
    }
    e += 2LL * X * Y * npairs;
    return e;
}

/*
 * Insertion order follows UPF:1228-1276 exactly: planar layers ascending,
 * H edges y-major then x (u->v then v->u), V edges x-major then y; then for
 * every (x, y) in x-major order and every via pair in the given order,
 * u->v then v->u with cost via_cost * (1 + 0.15 * (span - 1)) computed in
 * double and stored as float32 (the reference's structured 'f4' field).
 * CSRGraph.finalize then sorts the structured records with
 * sort(order='src', kind='mergesort') (UPF:723).
 */
int orc_lattice_build(int X, int Y, int L, const uint8_t *planar,
                      const uint8_t *dir, float pitch, double via_cost,
                      const int32_t *pair_from, const int32_t *pair_to,
                      int npairs, int32_t *indptr, int32_t *indices,
                      float *costs, uint8_t *edge_layer, uint8_t *edge_kind)
{
    const int64_t N = (int64_t)X * Y * L;
    const int64_t plane = (int64_t)X * Y;
    const int64_t E = orc_lattice_edge_count(X, Y, L, planar, dir, npairs);
    edge_t *ed = (edge_t *)malloc(sizeof(edge_t) * (size_t)(E > 0 ? E : 1));
    if (!ed) return -1;
    int64_t k = 0;
#define NODE(x, y, z) ((int32_t)((int64_t)(z) * plane + (int64_t)(y) * X + (x)))
    for (int z = 0; z < L; ++z) {
        if (!planar[z]) continue;
        if (dir[z] == DIR_H || dir[z] == DIR_HV) {
            for (int y = 0; y < Y; ++y)
                for (int x = 0; x < X - 1; ++x) {
                    int32_t u = NODE(x, y, z), v = NODE(x + 1, y, z);
                    ed[k++] = (edge_t){u, v, pitch};
                    ed[k++] = (edge_t){v, u, pitch};
                }
        }
        if (dir[z] == DIR_V || dir[z] == DIR_HV) {
            for (int x = 0; x < X; ++x)
                for (int y = 0; y < Y - 1; ++y) {
                    int32_t u = NODE(x, y, z), v = NODE(x, y + 1, z);
                    ed[k++] = (edge_t){u, v, pitch};
                    ed[k++] = (edge_t){v, u, pitch};
                }
        }
    }
    for (int x = 0; x < X; ++x)
        for (int y = 0; y < Y; ++y)
            for (int p = 0; p < npairs; ++p) {
                int zf = pair_from[p], zt = pair_to[p];
                int span = abs(zt - zf);
                float c = (float)(via_cost * (1.0 + 0.15 * (double)(span - 1)));
                int32_t u = NODE(x, y, zf), v = NODE(x, y, zt);
                ed[k++] = (edge_t){u, v, c};
                ed[k++] = (edge_t){v, u, c};
            }
#undef NODE
    /* stable counting sort by source */
    int64_t *cnt = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    if (!cnt) { free(ed); return -1; }
    for (int64_t i = 0; i < E; ++i) cnt[ed[i].src + 1]++;
    for (int64_t n = 0; n < N; ++n) cnt[n + 1] += cnt[n];
    for (int64_t n = 0; n <= N; ++n) indptr[n] = (int32_t)cnt[n];
    for (int64_t i = 0; i < E; ++i) {
        int64_t pos = cnt[ed[i].src]++;
        indices[pos] = ed[i].dst;
        costs[pos] = ed[i].cost;
        int32_t sl = (int32_t)(ed[i].src / plane), dl = (int32_t)(ed[i].dst / plane);
        edge_layer[pos] = (uint8_t)sl;
        edge_kind[pos] = (uint8_t)(sl != dl);
    }
    free(cnt);
    free(ed);
    /* numpy's sort(order='src') on the structured (src, dst, cost) records
     * breaks ties with the remaining fields in dtype order (dst, then cost):
     * each row ends up sorted by (dst, cost).  Equal records are identical,
     * so stability among them is unobservable. */
    for (int64_t n = 0; n < N; ++n) {
        for (int64_t i = indptr[n] + 1; i < indptr[n + 1]; ++i) {
            int32_t d = indices[i];
            float c = costs[i];
            uint8_t el = edge_layer[i], ek = edge_kind[i];
            int64_t j = i - 1;
            while (j >= indptr[n] && (indices[j] > d || (indices[j] == d && costs[j] > c))) {
                indices[j + 1] = indices[j]; costs[j + 1] = costs[j];
                edge_layer[j + 1] = edge_layer[j]; edge_kind[j + 1] = edge_kind[j];
                --j;
            }
            indices[j + 1] = d; costs[j + 1] = c; edge_layer[j + 1] = el; edge_kind[j + 1] = ek;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Binary heaps                                                              */
/* ------------------------------------------------------------------------ */

typedef struct { float d; int32_t n; } hf_t;       /* canon: (fp32 dist, node) */
typedef struct { double d; int32_t n; } hd_t;      /* ref: (float64 alt, roi idx) */

typedef struct { hf_t *a; int64_t n, cap; } heapf_t;
typedef struct { hd_t *a; int64_t n, cap; } heapd_t;

static int hf_less(hf_t x, hf_t y) { return x.d < y.d || (x.d == y.d && x.n < y.n); }
static int hd_less(hd_t x, hd_t y) { return x.d < y.d || (x.d == y.d && x.n < y.n); }

#define HEAP_IMPL(NAME, H, T, LESS)                                          \
    static int NAME##_push(H *h, T v) {                                      \
        if (h->n == h->cap) {                                                \
            int64_t nc = h->cap ? h->cap * 2 : 1024;                         \
            T *na = (T *)realloc(h->a, sizeof(T) * (size_t)nc);              \
            if (!na) return -1;                                              \
            h->a = na; h->cap = nc;                                          \
        }                                                                    \
        int64_t i = h->n++;                                                  \
        while (i > 0) {                                                      \
            int64_t p = (i - 1) / 2;                                         \
            if (!LESS(v, h->a[p])) break;                                    \
            h->a[i] = h->a[p]; i = p;                                        \
        }                                                                    \
        h->a[i] = v;                                                         \
        return 0;                                                            \
    }                                                                        \
    static T NAME##_pop(H *h) {                                              \
        T top = h->a[0];                                                     \
        T last = h->a[--h->n];                                               \
        int64_t i = 0;                                                       \
        for (;;) {                                                           \
            int64_t c = 2 * i + 1;                                           \
            if (c >= h->n) break;                                            \
            if (c + 1 < h->n && LESS(h->a[c + 1], h->a[c])) c++;             \
            if (!LESS(h->a[c], last)) break;                                 \
            h->a[i] = h->a[c]; i = c;                                        \
        }                                                                    \
        if (h->n > 0) h->a[i] = last;                                        \
        return top;                                                          \
    }

HEAP_IMPL(hf, heapf_t, hf_t, hf_less)
HEAP_IMPL(hd, heapd_t, hd_t, hd_less)

/* ------------------------------------------------------------------------ */
/* Canonical FP32 SSSP (the contract the GPU kernels implement)              */
/* ------------------------------------------------------------------------ */

/*
 * Search space: nodes with roi[n] != 0 (roi == NULL: the whole graph), the
 * same restriction SimpleDijkstra applies through global_to_roi (UPF:1858-
 * 1862).  Edge weight: costs[e], replaced by keepout_cost when the source
 * node carries a keepout owner other than net_id and the edge is planar
 * (UPF:3195-3221: every edge except a via in the same column; via edges of
 * this lattice never change column, so "planar" == edge_kind 0).
 *
 * Semantics: dist[v] = least fixpoint of min_u fl32(dist[u] + w(u, v));
 * pred[v] = lowest node index u attaining it.  Equivalent to keeping the
 * 64-bit key (float bits(dist) << 32 | pred) and relaxing with min.
 * The search stops once every node with dist <= dist[dst] has been expanded.
 *
 * Returns 1 if dst was reached, 0 otherwise.  *expanded / *traversed count
 * settled nodes and the out-edges examined from them (for GTEPS).
 */
int orc_sssp_canon(int64_t N, const int32_t *indptr, const int32_t *indices,
                   const float *costs, const uint8_t *edge_kind,
                   const uint8_t *roi, const int32_t *keepout_owner,
                   int32_t net_id, float keepout_cost, int32_t src,
                   int32_t dst, float *dist, int32_t *pred, uint8_t *nko,
                   int64_t *expanded, int64_t *traversed)
{
    (void)keepout_cost;
    /* Path order: (number of keepout edges, fp32 cost of the other edges),
     * then the lowest predecessor.  A keepout edge costs the reference 1e9
     * (UPF:3182), far above any path cost, so its float64 sums order paths
     * the same way; counting keepout edges apart keeps the fp32 cost exact. */
    uint8_t *ko = nko;
    for (int64_t i = 0; i < N; ++i) { dist[i] = INFINITY; pred[i] = -1; ko[i] = 255; }
    uint8_t *done = (uint8_t *)calloc((size_t)N, 1);
    if (!done) return -1;
    heapf_t h = {0};
    int64_t nexp = 0, ntrav = 0;
    dist[src] = 0.0f;
    ko[src] = 0;
    /* heap key: keepout count folded into the float's order by popping
     * (ko, d) lexicographically -- one heap per keepout level would do too;
     * here the heap compares (ko, d, node) through a combined float key is
     * not possible, so levels are processed in increasing ko order */
    int reached = 0;
    for (int level = 0; level < 255 && !(reached); ++level) {
        /* seed this level with every node that sits at it */
        h.n = 0;
        for (int64_t i = 0; i < N; ++i)
            if (ko[i] == level && !done[i]) hf_push(&h, (hf_t){dist[i], (int32_t)i});
        if (h.n == 0) continue;
        while (h.n > 0) {
            hf_t t = hf_pop(&h);
            int32_t u = t.n;
            if (done[u] || ko[u] != level || t.d != dist[u]) continue;
            if (reached && dist[u] > dist[dst]) break;
            done[u] = 1;
            nexp++;
            if (u == dst) reached = 1;
            float du = dist[u];
            int kou = keepout_owner && keepout_owner[u] >= 0 && keepout_owner[u] != net_id;
            for (int32_t e = indptr[u]; e < indptr[u + 1]; ++e) {
                int32_t v = indices[e];
                ntrav++;
                if (roi && !roi[v]) continue;
                int nk = level;
                float nd;
                if (kou && edge_kind[e] == 0) { nk = level < 254 ? level + 1 : 254; nd = du; }
                else { nd = du + costs[e]; if (!(nd < INFINITY)) continue; }
                int better = nk < ko[v] || (nk == ko[v] && nd < dist[v]);
                if (better) {
                    ko[v] = (uint8_t)nk;
                    dist[v] = nd;
                    pred[v] = u;
                    if (nk == level) hf_push(&h, (hf_t){nd, v});
                } else if (nk == ko[v] && nd == dist[v] && u < pred[v]) {
                    pred[v] = u;
                }
            }
        }
    }
    free(h.a);
    free(done);
    if (expanded) *expanded = nexp;
    if (traversed) *traversed = ntrav;
    return reached;
}

/* Walk pred from dst back to src; returns path length (0 on failure/cycle).
 * out must hold max_len entries; the path is written src .. dst. */
int64_t orc_backtrace(const int32_t *pred, int32_t src, int32_t dst,
                      int32_t *out, int64_t max_len)
{
    int64_t n = 0;
    int32_t cur = dst;
    while (cur != -1) {
        if (n >= max_len) return 0;
        out[n++] = cur;
        if (cur == src) break;
        cur = pred[cur];
    }
    if (n == 0 || out[n - 1] != src) return 0;
    for (int64_t i = 0, j = n - 1; i < j; ++i, --j) {
        int32_t t = out[i]; out[i] = out[j]; out[j] = t;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Reference Dijkstra, emulated exactly (UPF:1801-1887)                      */
/* ------------------------------------------------------------------------ */

/*
 * roi_nodes / global_to_roi as the reference builds them.  Heap entries are
 * (float64 alt, roi index), popped in tuple order like heapq.  dist is float32;
 * `alt < dist[v]` compares float32(alt) against it (NumPy 2 weak-scalar rule),
 * and dist[v] stores float32(alt).  Exit when dst pops.  Writes the path (global
 * ids, src..dst) into out and returns its length, or 0 for "None" (unreached
 * or length <= 1).  keepout as in orc_sssp_canon (the reference writes the
 * keepout cost into the cost array for the duration of the search).
 */
int64_t orc_sssp_ref(int64_t N, const int32_t *indptr, const int32_t *indices,
                     const float *costs, const uint8_t *edge_kind,
                     int64_t R, const int32_t *roi_nodes,
                     const int32_t *global_to_roi, const int32_t *keepout_owner,
                     int32_t net_id, float keepout_cost, int32_t src,
                     int32_t dst, int32_t *out, int64_t max_len,
                     double *path_alt)
{
    (void)N;
    int32_t rs = global_to_roi[src], rd = global_to_roi[dst];
    if (rs < 0 || rd < 0) return 0;
    float *dist = (float *)malloc(sizeof(float) * (size_t)R);
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)R);
    uint8_t *visited = (uint8_t *)calloc((size_t)R, 1);
    double *alts = (double *)malloc(sizeof(double) * (size_t)R);
    if (!dist || !parent || !visited || !alts) return -1;
    for (int64_t i = 0; i < R; ++i) { dist[i] = INFINITY; parent[i] = -1; alts[i] = INFINITY; }
    dist[rs] = 0.0f;
    alts[rs] = 0.0;
    heapd_t h = {0};
    hd_push(&h, (hd_t){0.0, rs});
    while (h.n > 0) {
        hd_t t = hd_pop(&h);
        int32_t ur = t.n;
        if (visited[ur]) continue;
        visited[ur] = 1;
        alts[ur] = t.d;
        if (ur == rd) break;
        int32_t ug = roi_nodes[ur];
        int ko = keepout_owner && keepout_owner[ug] >= 0 && keepout_owner[ug] != net_id;
        for (int32_t e = indptr[ug]; e < indptr[ug + 1]; ++e) {
            int32_t vg = indices[e];
            int32_t vr = global_to_roi[vg];
            if (vr < 0 || visited[vr]) continue;
            float w = (ko && edge_kind[e] == 0) ? keepout_cost : costs[e];
            double alt = t.d + (double)w;
            if ((float)alt < dist[vr]) {
                dist[vr] = (float)alt;
                parent[vr] = ur;
                hd_push(&h, (hd_t){alt, vr});
            }
        }
    }
    int64_t n = 0;
    if (isfinite(dist[rd])) {
        int32_t cur = rd;
        while (cur != -1 && n < max_len) {
            out[n++] = roi_nodes[cur];
            cur = parent[cur];
        }
        if (cur != -1) n = 0;
        for (int64_t i = 0, j = n - 1; i < j; ++i, --j) {
            int32_t tt = out[i]; out[i] = out[j]; out[j] = tt;
        }
        if (path_alt) *path_alt = alts[rd];
    }
    free(h.a); free(dist); free(parent); free(visited); free(alts);
    return n > 1 ? n : 0;
}

/* ------------------------------------------------------------------------ */
/* Multi-source / multi-sink search (UPF:1889-1989)                          */
/* ------------------------------------------------------------------------ */

/*
 * Exact emulation of find_path_multisource_multisink.  Seeds (node, cost)
 * inside the ROI set dist = float32(cost) and push (cost, roi index) -- a
 * later seed of the same node overwrites an earlier one, as the reference's
 * loop does.  Targets inside the ROI form a dict: a later entry for the same
 * node replaces the earlier one.  The heap pops (float64 key, roi index);
 * a popped target is recorded when du + target_cost < best (strict, so the
 * first found wins ties) and is not expanded; the search runs until the heap
 * is empty.  Relaxation compares float32(alt) < dist like SimpleDijkstra.
 * Writes the path (seed .. target, global ids) and returns its length, or 0
 * for "None" (no target reached, or a path of one node).
 */
int64_t orc_ms_ref(int64_t N, const int32_t *indptr, const int32_t *indices, const float *costs,
                   int64_t R, const int32_t *roi_nodes, const int32_t *global_to_roi,
                   int64_t n_seed, const int32_t *seed_node, const double *seed_cost,
                   int64_t n_tgt, const int32_t *tgt_node, const double *tgt_cost,
                   int32_t *out, int64_t max_len, double *best_total)
{
    (void)N;
    float *dist = (float *)malloc(sizeof(float) * (size_t)R);
    int32_t *parent = (int32_t *)malloc(sizeof(int32_t) * (size_t)R);
    uint8_t *visited = (uint8_t *)calloc((size_t)R, 1);
    uint8_t *is_t = (uint8_t *)calloc((size_t)R, 1);
    double *tcost = (double *)malloc(sizeof(double) * (size_t)R);
    if (!dist || !parent || !visited || !is_t || !tcost) return -1;
    for (int64_t i = 0; i < R; ++i) { dist[i] = INFINITY; parent[i] = -1; }
    heapd_t h = {0};
    for (int64_t i = 0; i < n_seed; ++i) {
        int32_t r = global_to_roi[seed_node[i]];
        if (r < 0) continue;
        dist[r] = (float)seed_cost[i];
        hd_push(&h, (hd_t){seed_cost[i], r});
    }
    int64_t nt = 0;
    for (int64_t i = 0; i < n_tgt; ++i) {
        int32_t r = global_to_roi[tgt_node[i]];
        if (r < 0) continue;
        is_t[r] = 1;
        tcost[r] = tgt_cost[i];
        nt++;
    }
    int32_t reached = -1;
    double best = INFINITY;
    if (h.n > 0 && nt > 0) {
        while (h.n > 0) {
            hd_t t = hd_pop(&h);
            int32_t ur = t.n;
            if (visited[ur]) continue;
            visited[ur] = 1;
            if (is_t[ur]) {
                double total = t.d + tcost[ur];
                if (total < best) { best = total; reached = ur; }
                continue;
            }
            int32_t ug = roi_nodes[ur];
            for (int32_t e = indptr[ug]; e < indptr[ug + 1]; ++e) {
                int32_t vr = global_to_roi[indices[e]];
                if (vr < 0 || visited[vr]) continue;
                double alt = t.d + (double)costs[e];
                if ((float)alt < dist[vr]) {
                    dist[vr] = (float)alt;
                    parent[vr] = ur;
                    hd_push(&h, (hd_t){alt, vr});
                }
            }
        }
    }
    int64_t n = 0;
    if (reached >= 0) {
        int32_t cur = reached;
        while (cur != -1 && n < max_len) { out[n++] = roi_nodes[cur]; cur = parent[cur]; }
        if (cur != -1) n = 0;
        for (int64_t i = 0, j = n - 1; i < j; ++i, --j) { int32_t tt = out[i]; out[i] = out[j]; out[j] = tt; }
    }
    if (best_total) *best_total = best;
    free(h.a); free(dist); free(parent); free(visited); free(is_t); free(tcost);
    return n > 1 ? n : 0;
}

/*
 * Canonical multi-source / multi-sink search (the GPU contract).  Search
 * space: nodes with roi[n] != 0 (NULL: whole graph).  Seeds inside it start
 * at key (cost, no predecessor), the lowest cost winning for a repeated node.
 * Keys follow orc_sssp_canon (fp32 distance, lowest predecessor on ties);
 * target nodes are settled but never expanded, and the search runs to
 * exhaustion.  The chosen target entry minimises (double(dist) +
 * double(target cost), dist, node, entry index).  Returns the entry index or
 * -1; dist/pred describe the whole search.
 */
int64_t orc_ms_canon(int64_t N, const int32_t *indptr, const int32_t *indices, const float *costs,
                     const uint8_t *roi, int64_t n_seed, const int32_t *seed_node, const float *seed_cost,
                     int64_t n_tgt, const int32_t *tgt_node, const float *tgt_cost,
                     float *dist, int32_t *pred, double *best_total, int64_t *expanded, int64_t *traversed)
{
    uint8_t *done = (uint8_t *)calloc((size_t)N, 1);
    uint8_t *is_t = (uint8_t *)calloc((size_t)N, 1);
    if (!done || !is_t) return -2;
    for (int64_t i = 0; i < N; ++i) { dist[i] = INFINITY; pred[i] = -1; }
    heapf_t h = {0};
    for (int64_t i = 0; i < n_seed; ++i) {
        int32_t v = seed_node[i];
        if (roi && !roi[v]) continue;
        if (seed_cost[i] < dist[v]) { dist[v] = seed_cost[i]; hf_push(&h, (hf_t){seed_cost[i], v}); }
    }
    for (int64_t i = 0; i < n_tgt; ++i)
        if (!roi || roi[tgt_node[i]]) is_t[tgt_node[i]] = 1;
    int64_t nexp = 0, ntrav = 0;
    while (h.n > 0) {
        hf_t t = hf_pop(&h);
        int32_t u = t.n;
        if (done[u] || t.d != dist[u]) continue;
        done[u] = 1;
        nexp++;
        if (is_t[u]) continue;
        float du = dist[u];
        for (int32_t e = indptr[u]; e < indptr[u + 1]; ++e) {
            int32_t v = indices[e];
            ntrav++;
            if (roi && !roi[v]) continue;
            float nd = du + costs[e];
            if (!(nd < INFINITY)) continue;
            if (nd < dist[v]) {
                dist[v] = nd; pred[v] = u;
                hf_push(&h, (hf_t){nd, v});
            } else if (nd == dist[v] && pred[v] != -1 && u < pred[v]) {
                pred[v] = u;
            }
        }
    }
    int64_t best = -1;
    double bt = INFINITY;
    for (int64_t i = 0; i < n_tgt; ++i) {
        int32_t v = tgt_node[i];
        if ((roi && !roi[v]) || !(dist[v] < INFINITY)) continue;
        double tot = (double)dist[v] + (double)tgt_cost[i];
        int better = best < 0 || tot < bt ||
                     (tot == bt && (dist[v] < dist[tgt_node[best]] ||
                                    (dist[v] == dist[tgt_node[best]] && v < tgt_node[best])));
        if (better) { best = i; bt = tot; }
    }
    if (best_total) *best_total = bt;
    if (expanded) *expanded = nexp;
    if (traversed) *traversed = ntrav;
    free(h.a); free(done); free(is_t);
    return best;
}

/* ------------------------------------------------------------------------ */
/* ROI BFS (UPF:1636-1734)                                                   */
/* ------------------------------------------------------------------------ */

/*
 * Bidirectional BFS: per depth, the src wave (mark 1) expands, then - unless
 * the waves met - the dst wave (mark 2).  A wave touching a node seen with a
 * different mark sets it to 3 and flags `met`.  Stops at max_depth, when both
 * queues are empty, on `met`, or when more than 1.5 * max_nodes nodes are
 * seen.  roi = ascending seen nodes; must-keeps (src, dst) appended if absent;
 * if larger than max_nodes, must-keeps at positions >= max_nodes are swapped
 * to the front (position = their rank among must-keeps found) and the list is
 * truncated.  Writes roi_out (capacity N) and returns its length.
 * seen_out (optional, N bytes) receives the final marks.
 */
int64_t orc_roi_bfs(int64_t N, const int32_t *indptr, const int32_t *indices,
                    int32_t src, int32_t dst, int max_depth, int64_t max_nodes,
                    const int32_t *keep, int nkeep, int readd,
                    int32_t *roi_out, uint8_t *seen_out, int *depth_out,
                    int *met_out)
{
    uint8_t *seen = (uint8_t *)calloc((size_t)N, 1);
    int32_t *qs = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int32_t *qd = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    int32_t *nq = (int32_t *)malloc(sizeof(int32_t) * (size_t)N);
    if (!seen || !qs || !qd || !nq) return -1;
    int64_t ns = 1, nd = 1, nseen = 0;
    qs[0] = src; qd[0] = dst;
    seen[src] = 1;
    seen[dst] = 2;
    nseen = (src == dst) ? 1 : 2;
    int depth = 0, met = 0;
    while (depth < max_depth && (ns > 0 || nd > 0) && !met) {
        for (int wave = 0; wave < 2; ++wave) {
            if (wave == 1 && met) break;
            int32_t *q = wave == 0 ? qs : qd;
            int64_t qn = wave == 0 ? ns : nd;
            uint8_t mark = (uint8_t)(wave == 0 ? 1 : 2);
            int64_t nn = 0;
            int mflag = 0;
            for (int64_t i = 0; i < qn; ++i) {
                int32_t u = q[i];
                for (int32_t e = indptr[u]; e < indptr[u + 1]; ++e) {
                    int32_t v = indices[e];
                    if (seen[v] == 0) {
                        seen[v] = mark;
                        nq[nn++] = v;
                        nseen++;
                    } else if (seen[v] != mark) {
                        seen[v] = 3;
                        mflag = 1;
                    }
                }
            }
            memcpy(q, nq, sizeof(int32_t) * (size_t)nn);
            if (wave == 0) ns = nn; else nd = nn;
            if (mflag) met = 1;
        }
        depth++;
        if ((double)nseen > (double)max_nodes * 1.5) break;
    }
    /* roi_nodes = the seen nodes ascending; must-keeps it lacks appended
     * (src, dst, then the portal seeds not yet listed, UPF:1700-1708) */
    int64_t R = 0;
    for (int64_t n = 0; n < N; ++n)
        if (seen[n]) roi_out[R++] = (int32_t)n;
    int32_t mk[8];
    int nmk = 0;
    {
        int32_t cand[8];
        int nc = 0;
        cand[nc++] = src; cand[nc++] = dst;
        for (int i = 0; i < nkeep && nc < 8; ++i) cand[nc++] = keep[i];
        for (int i = 0; i < nc; ++i) {
            int dup = cand[i] < 0;
            for (int j = 0; j < nmk; ++j) dup |= mk[j] == cand[i];
            if (!dup) mk[nmk++] = cand[i];
        }
    }
    for (int j = 0; j < nmk; ++j)
        if (!seen[mk[j]]) roi_out[R++] = mk[j];
    /* truncation: the must-keeps, in list order, each beyond the cut trades
     * places with position (its rank among them) (UPF:1711-1727) */
    if (R > max_nodes) {
        int64_t kept_idx[8];
        int nk = 0;
        for (int64_t i = 0; i < R && nk < 8; ++i)
            for (int j = 0; j < nmk; ++j)
                if (roi_out[i] == mk[j]) { kept_idx[nk++] = i; break; }
        for (int t = 0; t < nk; ++t) {
            if (kept_idx[t] >= max_nodes) {
                int32_t x = roi_out[kept_idx[t]];
                roi_out[kept_idx[t]] = roi_out[t];
                roi_out[t] = x;
            }
        }
        R = max_nodes;
    }
    /* _route_all: portal seeds, then src and dst, appended when missing
     * (UPF:4659-4696) */
    if (readd) {
        int32_t ra[8];
        int nr = 0;
        for (int i = 0; i < nkeep && nr < 6; ++i) ra[nr++] = keep[i];
        ra[nr++] = src; ra[nr++] = dst;
        for (int i = 0; i < nr; ++i) {
            if (ra[i] < 0) continue;
            int have = 0;
            for (int64_t k = 0; k < R && !have; ++k) have = roi_out[k] == ra[i];
            if (!have) roi_out[R++] = ra[i];
        }
    }
    if (seen_out) memcpy(seen_out, seen, (size_t)N);
    if (depth_out) *depth_out = depth;
    if (met_out) *met_out = met;
    free(seen); free(qs); free(qd); free(nq);
    return R;
}
