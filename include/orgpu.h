/*
 * orgpu: C ABI of the MI355X (gfx950) PathFinder inner loop.
 *
 * Plain pointers and sizes only.  Every large array is caller-owned device
 * memory (allocated by the caller, e.g. a torch tensor on cuda:0); it must
 * stay alive until the handle's stream has been synchronised.  The handle owns
 * only its scratch (search keys, bucket bitmaps, path staging), sized by
 * orgpu_reserve.  Calls are asynchronous on the handle's stream unless the
 * comment says otherwise.  Return value: 0 on success, a negative ORGPU_E*
 * code on failure (detail in orgpu_last_error(), thread-local).  "No path" is
 * not an error: the path length is 0 and the Python layer maps it to None.
 *
 * One handle = one HIP stream = one host thread at a time.  Several handles
 * may coexist (serving slots).  No C++ exception crosses this boundary.
 *
 * Reference interfaces each entry point replaces are cited as
 * UPF = orthoroute/algorithms/manhattan/unified_pathfinder.py.
 */
#ifndef ORGPU_H
#define ORGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORGPU_OK 0
#define ORGPU_EINVAL -1    /* invalid argument */
#define ORGPU_ENOMEM -2    /* device allocation failed */
#define ORGPU_EHIP -3      /* HIP runtime error */
#define ORGPU_ECAPACITY -4 /* reserved capacity exceeded (queue, path length, slots) */

#define ORGPU_MAX_LAYERS 64

typedef struct orgpu_ctx orgpu_t;

/* ---- handle -------------------------------------------------------------- */

/* Creates a handle on `device` bound to `hip_stream` (a hipStream_t; NULL =
 * the null stream).  Replaces nothing in the reference: there the CuPy
 * module-level state of CUDADijkstra (UPF:2205-2218) plays this role. */
int orgpu_create(int device, void *hip_stream, orgpu_t **out);
int orgpu_destroy(orgpu_t *h);
const char *orgpu_last_error(void);
/* Blocks until all work queued on the handle's stream has finished. */
int orgpu_sync(orgpu_t *h);

/* ---- A1 lattice ------------------------------------------------------------ */

/* Lattice3D(bounds, pitch, layers) + build_graph(via_cost) (UPF:1075-1338,
 * 647-792).  node = z*X*Y + y*X + x.  dir[z]: 0 = 'v', 1 = 'h', 2 = 'hv'.
 * planar[z] != 0: layer z carries planar edges.  Via pairs (from, to) are
 * given in the order the reference iterates them (UPF:1266-1276); each pair
 * yields u->v and v->u with cost via_cost*(1+0.15*(|to-from|-1)). */
typedef struct orgpu_lattice_desc {
    int32_t X, Y, L;
    float pitch;
    double via_cost;
    uint8_t planar[ORGPU_MAX_LAYERS];
    uint8_t dir[ORGPU_MAX_LAYERS];
    int32_t npairs;
    const int32_t *pair_from; /* host, npairs */
    const int32_t *pair_to;   /* host, npairs */
} orgpu_lattice_desc;

/* Host-only: node and edge counts (UPF:1206-1222). */
int orgpu_lattice_count(const orgpu_lattice_desc *d, int64_t *n_nodes, int64_t *n_edges);

/* CSR arrays exactly as CSRGraph.finalize leaves them (rows sorted by
 * (dst, cost), UPF:723): indptr i32[N+1], indices i32[E], base_cost f32[E],
 * edge_layer u8[E] (source layer), edge_kind u8[E] (1 = via). */
int orgpu_lattice_build(orgpu_t *h, const orgpu_lattice_desc *d, int32_t *indptr,
                        int32_t *indices, float *base_cost, uint8_t *edge_layer,
                        uint8_t *edge_kind);

/* ---- A2/A3/A4 search -------------------------------------------------------- */

typedef struct orgpu_graph {
    int64_t n_nodes, n_edges;
    const int32_t *indptr;    /* device i32[N+1] */
    const int32_t *indices;   /* device i32[E]  */
    const float *cost;        /* device f32[E]: this iteration's total_cost */
    const uint8_t *edge_kind; /* device u8[E]  */
    int32_t X, Y, L;          /* lattice shape (node = z*X*Y + y*X + x) */
} orgpu_graph;

/* One search request: SimpleDijkstra.find_path_roi(src, dst, costs, roi,
 * global_to_roi) (UPF:1801-1887) with the per-net owner-aware keepout of
 * _apply_owner_aware_via_keepouts (UPF:3150-3228) folded into the edge
 * weight.  roi_index < 0: whole graph; otherwise row roi_index of the ROI
 * bitmap array passed to orgpu_route.  use_keepout != 0: out-edges of a node
 * whose keepout owner is neither -1 nor net_id, other than via edges, cost
 * the keepout penalty (1e9 in the reference) instead of their cost. */
typedef struct orgpu_net_req {
    int32_t src, dst;
    int32_t net_id;
    int32_t roi_index;
    int32_t use_keepout;
    int32_t reserved;
} orgpu_net_req;

typedef struct orgpu_sssp_params {
    float delta;            /* bucket width (cost units) */
    int32_t n_buckets;      /* bucket ring size (>= 1) */
    int32_t slots;          /* concurrent searches K (1 = plain sequential) */
    int32_t rounds_per_poll;/* kernel rounds queued between host status polls */
    int32_t grid;           /* workgroups per round launch (0 = auto) */
    int32_t count_minimal;  /* 1: count Dijkstra-minimal settled nodes/edges per net */
    /* Goal-directed search (0, 0 = plain Dijkstra order).  Lower bounds of the
     * cost of one planar step (h_xy) and of one layer of via span (h_z) under
     * the call's costs: buckets are taken on g + h with
     * h(v) = h_xy*(|dx|+|dy|) + h_z*|dz| to the destination.  With h a lower
     * bound of the remaining cost, every node on any equal-cost shortest path
     * is still expanded before the destination is declared settled, so paths
     * and tie-breaks are those of the plain search. */
    float h_xy, h_z;
    /* Same-bucket items a wave pushes are expanded by that wave within the
     * round, up to local_batches batches of 64 (0 = 64); the rest wait for
     * the next round. */
    int32_t local_batches;
    /* 0: the rounds between two polls are launched one by one; 1: they are
     * replayed as one HIP graph (captured on the first call and again when
     * the round arguments change).  Same kernels and results either way. */
    int32_t launch_mode;
    /* Search engine for two-pin requests: 0 = bucket rounds over all slots
     * (every request kind); 1 = one wavefront per search, a persistent grid
     * of min(slots, n_nets) one-wave workgroups that take requests in order
     * from a device counter, each owning one key slot (same paths, dists and
     * edge ids).  Multi-source requests and count_minimal use engine 0. */
    int32_t engine;
    /* engine 1: bucket-list entries one search may hold at once (8 bytes
     * each, per slot; 0 = 2^20).  A search that would exceed it is handed to
     * engine 0 (counted in orgpu_sssp_stats.retried). */
    int32_t list_cap;
    /* engine 1: a step with fewer than 64 entries of the current bucket also
     * expands entries of up to this many following buckets (same paths: a
     * node expanded early is expanded again if its key drops); 0 = 15, < 0 =
     * none, at most 15. */
    int32_t lookahead;
    /* engine 1: wavefronts that share one search, 1, 2 or 4 (0 = automatic:
     * 4 when all n_nets searches start at once, n_nets <= slots and
     * n_nets <= 1024, else 2).  Same paths; more waves per search shorten the
     * longest searches, which end a call whose slots do not refill. */
    int32_t waves;
} orgpu_sssp_params;

typedef struct orgpu_sssp_stats {
    int64_t rounds;          /* round launches executed */
    int64_t items;           /* frontier items expanded (incl. re-expansions) */
    int64_t edges;           /* edges relaxed (incl. re-relaxations) */
    int64_t minimal_nodes;   /* Dijkstra-minimal settled nodes (count_minimal) */
    int64_t minimal_edges;   /* their out-edges (count_minimal) */
    double kernel_ms;        /* device time from first seed to last path (HIP events) */
    int64_t launches;        /* round kernels launched (idle ones return at once) */
    double round_ms;         /* device time of the round kernels alone (HIP events
                                around each back-to-back batch of launches; engine 1:
                                the search kernel) */
    int64_t retried;         /* engine 1: requests handed to engine 0 for capacity */
} orgpu_sssp_stats;

/* Reserves scratch for graphs of up to n_nodes nodes, `slots` concurrent
 * searches (at most 4096; the bucket-round engine runs at most 2048 of them)
 * and `n_buckets` ring buckets (at most 32): with
 * NP = n_nodes rounded up to a multiple of 2048, (n_buckets+3) bucket
 * bitmaps of slots*NP bits and path staging 2*slots*max_path_len ints; the
 * keys (8 bytes per slot and node: slots*NP for the bucket-round engine, the
 * tiled index space for the one-wave engine) are allocated by the first
 * search call, at the stride its engine needs, and grown once if a later
 * call's engine needs more.  queue_capacity is ignored (the frontiers are
 * bitmaps; the argument stays for callers of the earlier queue form).
 * Blocks (allocation). */
int orgpu_reserve(orgpu_t *h, int64_t n_nodes, int32_t slots, int32_t n_buckets,
                  int64_t queue_capacity, int32_t max_path_len);

/* Routes n_nets requests (host array) to completion: K = params->slots
 * searches run concurrently and a finished slot is refilled with the next
 * request.  Writes each path (src..dst, global node ids) and the CSR edge id
 * of each hop (first match in the row, UPF:5463-5476) into the host arrays
 * path_nodes / path_edges at path_offset[i] (host i64[n_nets+1], filled by
 * this call); capacity = path_capacity entries each.  path_len 0 = no path.
 * roi_bitmaps: device u32[n_roi][ceil(N/32)], bit set = node in ROI.
 * keepout_owner: device i32[N] or NULL.  Blocks until done. */
int orgpu_route(orgpu_t *h, const orgpu_graph *g, const orgpu_net_req *nets, int32_t n_nets,
                const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                float keepout_cost,
                const orgpu_sssp_params *params, int32_t *path_nodes, int32_t *path_edges,
                int64_t path_capacity, int64_t *path_offset, float *path_dist,
                orgpu_sssp_stats *stats);

/* One-wave engine, asynchronous form.  Replaces the same reference code as
 * orgpu_route (SimpleDijkstra.find_path_roi per net with the owner-aware
 * keepout, UPF:1801-1887, 3150-3241; the parent walk and _path_to_edges,
 * UPF:1880-1887, 5463-5476), for callers that keep requests and results on
 * the device.
 *
 * orgpu_reserve_solo: after orgpu_reserve (nodes, slots, path length), the
 * one-wave engine's own scratch for a graph of n_nodes nodes (a lattice
 * when X*Y*L == n_nodes: its keys are then laid out in tiles) and `slots`
 * concurrent searches with bucket lists of list_cap entries each (0 = 2^20).
 * Blocks (allocation).
 *
 * orgpu_route_async: enqueues the n requests of the device array `reqs` on
 * the handle's stream and returns at once; allocates nothing (it needs
 * orgpu_reserve_solo for this graph, params->slots and params->list_cap).
 * After orgpu_sync(h), for request i (device arrays): path_len[i] nodes of
 * its path src..dst at path_nodes[path_off[i]], the first CSR edge of each
 * hop at path_edges[path_off[i]] (-1 after the last node; path_edges may be
 * NULL), path_dist[i] its cost; *used (device u32, may be NULL) the entries
 * written.  path_len[i]: 0 = no path, -1 = broken chain or longer than the
 * reserved path length, -2 = 31 or more keepout edges, -3 = given back (its
 * bucket lists or the output capacity ran out: route it with orgpu_route),
 * -4 = invalid request (endpoint outside the graph, ROI row >= n_roi, or
 * use_keepout without keepout_owner).  Paths are those of orgpu_route. */
int orgpu_reserve_solo(orgpu_t *h, int64_t n_nodes, int32_t X, int32_t Y, int32_t L, int32_t slots,
                       int32_t list_cap);
int orgpu_route_async(orgpu_t *h, const orgpu_graph *g, const orgpu_net_req *reqs, int32_t n,
                      const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                      float keepout_cost, const orgpu_sssp_params *params, int32_t *path_nodes,
                      int32_t *path_edges, int64_t capacity, uint32_t *path_off, int32_t *path_len,
                      float *path_dist, uint32_t *used);

/* Multi-source / multi-sink search: SimpleDijkstra.find_path_multisource_
 * multisink (UPF:1889-1989) and the GPU seam find_path_fullgraph_gpu_seeds
 * (UPF:4558-4565).  Each request seeds seeds[seed_off .. +seed_n) at their
 * entry costs and ends on the cheapest of targets[target_off .. +target_n)
 * by (distance + target cost); targets are settled but never expanded, and
 * the search runs until its frontier is empty (the reference does not stop
 * at the first target).  Seeds and targets outside the request's ROI row
 * are ignored.  Among equal totals the lower distance, then the lower node
 * index, then the earlier entry wins. */
typedef struct orgpu_ms_point {
    int32_t node;
    float cost;             /* finite, >= 0 */
} orgpu_ms_point;

typedef struct orgpu_ms_req {
    int32_t seed_off, seed_n;
    int32_t target_off, target_n;
    int32_t net_id;
    int32_t roi_index;      /* < 0: whole graph */
    int32_t use_keepout;
    int32_t reserved;
} orgpu_ms_req;

/* As orgpu_route, for multi-source / multi-sink requests.  Paths run from a
 * seed to the chosen target; path_total[i] (host f64[n], optional) receives
 * dist + target cost, path_target[i] (host i32[n], optional) the chosen
 * target's index within the request's target list (-1: none reached).  A
 * chosen target that is itself a seed gives a path of one node. */
int orgpu_route_ms(orgpu_t *h, const orgpu_graph *g, const orgpu_ms_req *reqs, int32_t n_reqs,
                   const orgpu_ms_point *seeds, int64_t n_seeds, const orgpu_ms_point *targets,
                   int64_t n_targets, const uint32_t *roi_bitmaps, int32_t n_roi,
                   const int32_t *keepout_owner, float keepout_cost, const orgpu_sssp_params *params,
                   int32_t *path_nodes, int32_t *path_edges, int64_t path_capacity, int64_t *path_offset,
                   double *path_total, int32_t *path_target, orgpu_sssp_stats *stats);

/* ---- A3 reference-order search ------------------------------------------------ */

/* One request of the reference-order search: SimpleDijkstra.find_path_roi
 * (UPF:1801-1887) with the reference's own arithmetic and pop order, for
 * callers that need its exact node sequences (and so its negotiation
 * trajectory) rather than the canonical tie-break of orgpu_route.  The heap
 * holds (float64 alt, position of the node in the reference's roi_nodes
 * list); the position order is given as a key: 2*v + 2 for a node v of an
 * ascending list, with up to eight exceptions (rank_node[i] has key
 * rank_key[i]) for nodes the list holds out of order (orgpu_roi_info.front /
 * appended: keys 2*left[p] + 2 and 2*N + 2 + j; a portal net's re-added
 * src/dst follow its appended nodes).  roi_index < 0: whole graph
 * (roi_nodes = arange(N)).  use_keepout as in orgpu_net_req (a keepout edge
 * costs keepout_cost, summed in float64 like the reference's 1e9). */
typedef struct orgpu_ref_req {
    int32_t src, dst;
    int32_t net_id;
    int32_t roi_index;
    int32_t use_keepout;
    int32_t n_rank;           /* 0..8 */
    int32_t rank_node[8];
    uint32_t rank_key[8];
} orgpu_ref_req;

/* Reserves the reference-order search scratch: per concurrent search
 * (`slots`, at most 4096) 16 bytes per node of a graph of up to n_nodes
 * nodes and a heap of heap_cap 16-byte entries (the heap also stages the
 * path, so paths up to 4*heap_cap nodes).  Blocks (allocation). */
int orgpu_reserve_ref(orgpu_t *h, int64_t n_nodes, int32_t slots, int64_t heap_cap);

/* Runs n reference-order searches (reqs: DEVICE array), `slots` at a time
 * (0 = the reserved count), one wavefront per search.  Asynchronous on the
 * handle's stream; every output is a DEVICE array:
 *   path_offset i64[n], path_len i32[n]: request i's path (src..dst, global
 *   ids) is path_nodes[path_offset[i] .. +path_len[i]) and the CSR edge id of
 *   each hop (first match in the row, UPF:5463-5476) path_edges[same];
 *   path_len 0 = None (no path, or a path of one node as the reference
 *   returns None for), < 0 = error for that request: -1 path output
 *   capacity, -2 heap capacity, -3 path longer than the staging row,
 *   -4 invalid request (endpoint, ROI row or keepout array);
 *   path_dist f32[n]: dist[dst] (float32, as the reference stores it);
 *   counters u64[4] (may be NULL): += pops, += edges scanned, += path
 *   entries written, and word 3 = OR of the error bits (1 << -len).
 * The caller zeroes counters[2] (the output cursor) before the call. */
int orgpu_route_ref(orgpu_t *h, const orgpu_graph *g, const orgpu_ref_req *reqs, int32_t n,
                    const uint32_t *roi_bitmaps, int32_t n_roi, const int32_t *keepout_owner,
                    float keepout_cost, int32_t slots, int32_t *path_nodes, int32_t *path_edges,
                    int64_t path_capacity, int64_t *path_offset, int32_t *path_len, float *path_dist,
                    uint64_t *counters);

/* orgpu_route_ref for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  No handle: the search scratch is a
 * fixed static device array of ORGPU_REF_NOALLOC_BYTES per GPU (256 bytes of
 * control, then slots * n_nodes * 16 bytes of node state, then slots *
 * heap_cap * 16 bytes of heaps), held by each GPU the library runs on until
 * the process ends; a request that does not fit returns ORGPU_EINVAL.
 * Calls nothing that allocates.  Runs on `hip_stream` of `device`; inputs,
 * outputs and their meaning as orgpu_route_ref.  Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU (other streams or other
 * host threads before the first has finished on the device) race. */
#define ORGPU_REF_NOALLOC_BYTES (64ull << 20)
int orgpu_route_ref_noalloc(int device, void *hip_stream, const orgpu_graph *g, const orgpu_ref_req *reqs,
                            int32_t n, const uint32_t *roi_bitmaps, int32_t n_roi,
                            const int32_t *keepout_owner, float keepout_cost, int32_t slots,
                            int64_t heap_cap, int32_t *path_nodes, int32_t *path_edges,
                            int64_t path_capacity, int64_t *path_offset, int32_t *path_len,
                            float *path_dist, uint64_t *counters);

/* ---- A2 ROI extraction -------------------------------------------------------- */

/* ROIExtractor.extract_roi_bfs (UPF:1636-1734) plus the must-keep re-add of
 * _route_all (UPF:4659-4700): bidirectional level-synchronous BFS from src
 * (wave 1) and dst (wave 2) until max_depth, both frontiers empty, the waves
 * meeting, or more than 1.5*max_nodes nodes seen; then the reference's list
 * rules: roi_nodes = the seen nodes ascending, must-keeps it lacks appended
 * (src, dst, then keep[0], keep[1] -- the portal seeds of a portal net,
 * UPF:1700-1708), truncation to max_nodes with the must-keeps beyond the cut
 * swapped to the front (UPF:1711-1727); with readd, the keeps and then src
 * and dst appended when missing (UPF:4659-4696).  The ROI is written as
 * bitmap row roi_index (bit set = in ROI). */
typedef struct orgpu_roi_req {
    int32_t src, dst;
    int32_t max_depth;
    int32_t roi_index;
    int64_t max_nodes;  /* >= 1 */
    int32_t readd;      /* 1: re-add keeps, src, dst (main search); 0: as extract_roi returns it (fallback) */
    int32_t reserved;
    int32_t keep[2];    /* further must-keep nodes (-1: none) */
} orgpu_roi_req;

typedef struct orgpu_roi_info {
    int64_t seen;     /* nodes the BFS marked */
    int64_t kept;     /* nodes in the final ROI */
    int32_t depth;    /* BFS depth reached */
    int32_t met;      /* waves met */
    int32_t src_in;   /* src in the final ROI */
    int32_t dst_in;   /* dst in the final ROI */
    /* Where the reference's roi_nodes list holds nodes out of ascending order
     * (its position order breaks heap ties, UPF:1841-1875): front[p] is the
     * must-keep that the truncation swapped into position p (p = 0..3),
     * which node left[p] held before (-1, -1: no swap at p, UPF:1716-1727);
     * appended[j] is the j-th node kept after the sorted part (must-keeps
     * appended before the truncation, then the re-adds; -1: none). */
    int32_t front[4];
    int32_t left[4];
    int32_t appended[4];
} orgpu_roi_info;

/* reqs: host array; roi_bitmaps: device u32[n_roi][ceil(N/32)], rows written
 * completely; info: host array (may be NULL).  Blocks. */
int orgpu_roi_bfs(orgpu_t *h, const orgpu_graph *g, const orgpu_roi_req *reqs, int32_t n,
                  uint32_t *roi_bitmaps, int32_t n_roi, orgpu_roi_info *info);

/* The same ROI as orgpu_roi_bfs (rows, info and the reference's rules,
 * UPF:1636-1734 and 4683-4700) for requests on the lattice that `d` describes
 * (the graph orgpu_lattice_build makes from it), computed in closed form: the
 * hop distance between two nodes is dx + dy plus the fewest via hops between
 * their layers through planar layers of the needed directions, so each
 * wave's marks are a ball whose size per level, meeting step and early stop
 * follow without traversing edges.  max_depth at most 1024.  Needs no ROI
 * scratch.  Blocks. */
int orgpu_roi_lattice(orgpu_t *h, const orgpu_lattice_desc *d, const orgpu_roi_req *reqs, int32_t n,
                      uint32_t *roi_bitmaps, int32_t n_roi, orgpu_roi_info *info);

/* Upper bound in bytes on the handle's ROI BFS scratch (about 9 bytes per
 * node per workgroup slot; default 32 GB).  No counterpart in the reference,
 * whose GPU pools have fixed sizes (cuda_dijkstra.py:2322-2352); serving
 * callers that keep several boards resident on one GPU lower it so each job's
 * footprint stays bounded.  Scratch above a lowered limit is released. */
int orgpu_set_roi_scratch(orgpu_t *h, int64_t max_bytes);

/* ---- A5/A6 congestion state (EdgeAccountant in HBM) ------------------------- */

/* Nonzero entries of a per-edge quantity, copied to host arrays (idx, val)
 * of `capacity` entries, in no particular order; *count receives the number of
 * nonzeros (may exceed capacity: then only `capacity` were written and the
 * call returns ORGPU_ECAPACITY). */
typedef struct orgpu_sparse_out {
    int32_t *idx;      /* host */
    float *val;        /* host */
    int64_t capacity;
    int64_t count;     /* out */
} orgpu_sparse_out;

/* canonical usage count[e] += delta for every edge id in the host array
 * (commit_path / clear_path / _rebuild_usage_from_committed_nets,
 * UPF:821-836, 4971-4987).  count: device i32[E]. */
int orgpu_usage_scatter(orgpu_t *h, int32_t *count, int64_t n_edges, const int32_t *edge_ids,
                        int64_t n, int32_t delta);

/* present[e] = float(count[e]) (refresh_from_canonical, UPF:814-819), then
 * present[e] += penalty once per edge of the host list (the barrel-conflict
 * penalty, UPF:3863-3871; numpy fancy += applies once per unique index). */
int orgpu_present_build(orgpu_t *h, int64_t n_edges, const int32_t *count, float *present,
                        const int32_t *penalty_edges, int64_t n_penalty, float penalty);

/* present_ema = beta*present + (1-beta)*present_ema with present = count
 * (update_present_ema, UPF:963-971; first != 0 starts from
 * present_ema = present, UPF:3650-3652).  beta and one_minus_beta are the
 * float32 constants numpy would use.  Also reports the nonzeros of
 * max(0, present_ema - 1) (input of _compute_layer_bias, UPF:4825-4878). */
int orgpu_present_ema(orgpu_t *h, int64_t n_edges, const int32_t *count, float *present_ema,
                      int first, float beta, float one_minus_beta, orgpu_sparse_out *over_out);

/* Per-iteration cost (EdgeAccountant.update_costs UPF:973-1043, then the
 * via pooling penalties UPF:3291-3387 and, when hard_block != 0, the via
 * capacity hard block UPF:3483-3546), one fused pass.  All float32 constants
 * are the values numpy would use; via_discount[zlo*L+zhi] is the leave-hot-
 * layer factor (double; applied in double when discount_f64, else rounded to
 * float first). */
typedef struct orgpu_cost_params {
    int32_t X, Y, L;
    float via_cost_mult, base_cost_weight, pres_fac, hist_weight, jitter_scale;
    int32_t add_jitter;
    const float *layer_bias;      /* device f32[L] */
    int32_t via_pooling;          /* apply pooling penalties */
    int32_t seg_z;                /* number of via segments per column (L-2, 0 = none) */
    const float *col_pres;        /* device f32[X*Y]   ([x][y]) */
    const float *seg_prefix;      /* device f32[X*Y*seg_z] ([x][y][s]) */
    float col_weight, seg_weight;
    const double *via_discount;   /* device f64[L*L] */
    int32_t discount_f64;
    int32_t hard_block;
    const int16_t *col_use;       /* device i16[X*Y] */
    const int16_t *seg_use;       /* device i16[X*Y*seg_z] */
    int32_t col_cap, seg_cap;
} orgpu_cost_params;

int orgpu_cost_update(orgpu_t *h, const orgpu_cost_params *p, int64_t n_nodes, int64_t n_edges,
                      const int32_t *indptr, const int32_t *indices, const uint8_t *edge_kind,
                      const float *base_cost, const float *present_ema, const float *history,
                      float *total_cost);

/* history = history*decay; history += min(gain*max(0, present-1), cap_mult*base)
 * (EdgeAccountant.update_history with use_raw_present, UPF:898-961; the cap
 * is skipped when use_cap == 0). */
int orgpu_history_update(orgpu_t *h, int64_t n_edges, const float *present, const float *base_cost,
                         float *history, float decay, float gain, float cap_mult, int use_cap);

/* Nonzeros of max(0, present-1) (mode 0), or the hot-edge set of
 * _build_hotset, over + 0.1*history > 0.5, with val = over (mode 1,
 * UPF:5021-5024). */
int orgpu_overuse_sparse(orgpu_t *h, int64_t n_edges, const float *present, const float *history,
                         int mode, float hist_scale, float threshold, orgpu_sparse_out *out);

/* Via pooling pressure (UPF:3626-3637), per via column c of n_cols and per
 * segment z of its seg_z: col_pres[c] = 0.6*max(0, col_use[c] - col_cap) +
 * 0.4*col_pres[c], the same for seg_pres[c*seg_z + z] with seg_cap, and
 * seg_prefix = the running sum of seg_pres along z (float32, in the order
 * numpy's float32 expressions and cumsum take).  *over (device u64) is set
 * to the sum of max(0, use - cap) over columns and segments (the via part of
 * the overuse sum, UPF:838-876).  All arrays are device memory in
 * [column][segment] order.  Asynchronous on the handle's stream. */
int orgpu_via_pressure(orgpu_t *h, int64_t n_cols, int32_t seg_z, const int16_t *col_use, int32_t col_cap,
                       const int16_t *seg_use, int32_t seg_cap, float *col_pres, float *seg_pres,
                       float *seg_prefix, unsigned long long *over);

/* Per committed path, the sum of a per-edge float32 quantity over the path's
 * edge list in path order, in float64, one addition after another from 0.0:
 * the value of Python's sum(float(v[e]) for e in edges) in the impact scores
 * of _build_hotset (UPF:5060-5072: the values outside the hot set are zero
 * here, and adding +0.0 leaves such a sum unchanged) and
 * _rip_top_k_offenders (UPF:5300-5335).  Path s covers
 * edges[offs[s] .. offs[s+1]); an edge id outside [0, n_edges) makes its
 * path's sum NaN (it is not read).  dense: device f32[n_edges]; edges:
 * device i32; offs: device i64[n_seg+1], non-decreasing; out: device
 * f64[n_seg].  Asynchronous on the handle's stream. */
int orgpu_segment_sums(orgpu_t *h, int64_t n_edges, const float *dense, const int32_t *edges,
                       const int64_t *offs, int64_t n_seg, double *out);

/* Lower bounds for the goal-directed search (orgpu_sssp_params.h_xy/h_z):
 * out[0] = min cost over planar edges, out[1] = min over via edges of
 * cost / layer span, both over finite non-negative costs (device f32[2],
 * caller-owned; 0x7F7F7F7F when a kind has no finite edge).  hist (device
 * u32[64] or null): planar finite costs counted by binary exponent, bin =
 * exponent + 32 clamped to [0, 63] -- the scale from which the caller picks
 * the bucket width.  No reference counterpart: the reference's Dijkstra is
 * neither goal-directed nor bucketed. */
int orgpu_cost_bounds(orgpu_t *h, const orgpu_graph *g, const uint8_t *edge_layer, float *out, uint32_t *hist);

#ifdef __cplusplus
}
#endif

#endif /* ORGPU_H */
