"""Device side of the router: HBM-resident lattice and search calls.

PyTorch is used only to own device memory (tensors on cuda:0) and to provide
the HIP stream; every computation runs in liborgpu.so's hand-written gfx950
kernels through the C ABI in include/orgpu.h.  There is no CPU fallback: on a
machine without the library or a GPU these calls raise.
"""
import ctypes

import numpy as np

from . import _abi
from .lattice import lattice_spec

KEEPOUT_COST = 1e9  # UPF:3182


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("orthoroute_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


MAX_SLOTS = 2048       # concurrent searches of the bucket-round engine (engine 0)
SOLO_MAX_SLOTS = 4096  # concurrent searches of the one-wave engine (engine 1)
SOLO_ARENA = 1 << 20   # engine 1: bucket-list entries per search (the library default)
CHUNK_ITEMS = 2048     # a search's keys and bitmap bits are laid out per whole chunk
_DEVICE_BYTES = None


def device_memory():
    """Total memory of the current device in bytes (0 without a GPU)."""
    global _DEVICE_BYTES
    if _DEVICE_BYTES is None:
        import torch
        _DEVICE_BYTES = int(torch.cuda.mem_get_info()[1]) if torch.cuda.is_available() else 0
    return _DEVICE_BYTES


def slot_bytes(n_nodes, n_buckets=8, engine=0):
    """Engine memory one concurrent search takes on a graph of n_nodes: an
    8-byte key per node (the per-slot stride rounds n_nodes up to a whole
    chunk), its touched-chunk records, and then n_buckets + 3 bucket-bitmap
    bits per node (engine 0) or its bucket-list arena (engine 1)."""
    np_ = -(-int(n_nodes) // CHUNK_ITEMS) * CHUNK_ITEMS
    common = 8 * np_ + 8 * (np_ // 1024 + 1)
    if engine == 1:
        return common + 8 * SOLO_ARENA + 8 * (SOLO_ARENA // 256)
    return common + (int(n_buckets) + 3) * np_ // 8


def max_slots(n_nodes, budget_bytes=None, n_buckets=8, engine=0):
    """Concurrent searches a graph of n_nodes allows: as many as the
    engine's slot memory (slot_bytes each) fits in budget_bytes (default 70 %
    of the device's memory), at most MAX_SLOTS (engine 0) or SOLO_MAX_SLOTS
    (engine 1, whose handle still holds the bucket engine's bitmaps for up to
    MAX_SLOTS of its slots, for searches it hands back)."""
    if budget_bytes is None:
        budget_bytes = 0.7 * device_memory()
    if engine != 1:
        return max(1, min(MAX_SLOTS, int(budget_bytes // slot_bytes(n_nodes, n_buckets))))
    s1 = slot_bytes(n_nodes, n_buckets, 1)
    bm = slot_bytes(n_nodes, n_buckets, 0) - slot_bytes(n_nodes, n_buckets, 1) + 8 * SOLO_ARENA + 8 * (SOLO_ARENA // 256)
    k = int(budget_bytes // (s1 + bm))
    if k > MAX_SLOTS:
        k = int((budget_bytes - MAX_SLOTS * bm) // s1)
    return max(1, min(SOLO_MAX_SLOTS, k))


class Context:
    """One orgpu handle bound to a HIP stream on a device: `stream` (a
    torch.cuda.Stream) or else torch's current stream there.  Torch work on
    the handle's buffers must run on the same stream (the serving loop gives
    each job its own stream and makes it current in the job's thread)."""

    def __init__(self, device=0, stream=None):
        torch = _torch()
        self.torch = torch
        self.device = torch.device("cuda", device)
        self.lib = _abi.load()
        if stream is None:
            with torch.cuda.device(self.device):
                stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        h = ctypes.c_void_p()
        _abi.check(self.lib.orgpu_create(device, ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(h)))
        self.handle = h
        self._reserved = None

    def close(self):
        if self.handle:
            self.lib.orgpu_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _abi.check(self.lib.orgpu_sync(self.handle))

    def set_roi_scratch(self, max_bytes):
        """Bound the handle's ROI BFS scratch (bytes)."""
        _abi.check(self.lib.orgpu_set_roi_scratch(self.handle, int(max_bytes)))

    # ---- A1 ---------------------------------------------------------------
    def build_lattice(self, X, Y, L, policy="reference", pitch=1.0, via_cost=3.0):
        return DeviceLattice(self, X, Y, L, policy, pitch, via_cost)

    # ---- A3 ---------------------------------------------------------------
    def reserve(self, n_nodes, slots, n_buckets, queue_capacity=0, max_path_len=1 << 16):
        key = (int(n_nodes), int(slots), int(n_buckets), int(queue_capacity), int(max_path_len))
        if self._reserved is not None and all(a >= b for a, b in zip(self._reserved, key)):
            return
        _abi.check(self.lib.orgpu_reserve(self.handle, key[0], key[1], key[2], key[3], key[4]))
        self._reserved = key


def _lattice_desc(X, Y, L, policy, pitch, via_cost):
    planar, dirs, pairs = lattice_spec(L, policy)
    d = _abi.LatticeDesc()
    d.X, d.Y, d.L = X, Y, L
    d.pitch = pitch
    d.via_cost = via_cost
    for z in range(L):
        d.planar[z] = planar[z]
        d.dir[z] = dirs[z]
    pf = np.ascontiguousarray([p[0] for p in pairs], dtype=np.int32)
    pt = np.ascontiguousarray([p[1] for p in pairs], dtype=np.int32)
    d.npairs = len(pairs)
    d.pair_from = pf.ctypes.data_as(ctypes.c_void_p)
    d.pair_to = pt.ctypes.data_as(ctypes.c_void_p)
    return d, (pf, pt)


def lattice_counts(X, Y, L, policy="reference"):
    d, keep = _lattice_desc(X, Y, L, policy, 1.0, 3.0)
    n, e = ctypes.c_int64(), ctypes.c_int64()
    _abi.check(_abi.load().orgpu_lattice_count(ctypes.byref(d), ctypes.byref(n), ctypes.byref(e)))
    return n.value, e.value


class DeviceLattice:
    """CSR of the lattice in HBM (indptr, indices, base_cost, edge_layer, edge_kind)."""

    def __init__(self, ctx, X, Y, L, policy="reference", pitch=1.0, via_cost=3.0):
        torch = ctx.torch
        self.ctx = ctx
        self.X, self.Y, self.L, self.policy = int(X), int(Y), int(L), policy
        self.pitch, self.via_cost = float(pitch), float(via_cost)
        d, keep = _lattice_desc(self.X, self.Y, self.L, policy, self.pitch, self.via_cost)
        self.desc, self._desc_keep = d, keep      # for the closed-form ROI
        n, e = ctypes.c_int64(), ctypes.c_int64()
        _abi.check(ctx.lib.orgpu_lattice_count(ctypes.byref(d), ctypes.byref(n), ctypes.byref(e)))
        self.N, self.E = n.value, e.value
        dev = ctx.device
        self.indptr = torch.empty(self.N + 1, dtype=torch.int32, device=dev)
        self.indices = torch.empty(max(self.E, 1), dtype=torch.int32, device=dev)
        self.base_cost = torch.empty(max(self.E, 1), dtype=torch.float32, device=dev)
        self.edge_layer = torch.empty(max(self.E, 1), dtype=torch.uint8, device=dev)
        self.edge_kind = torch.empty(max(self.E, 1), dtype=torch.uint8, device=dev)
        _abi.check(ctx.lib.orgpu_lattice_build(ctx.handle, ctypes.byref(d), _ptr(self.indptr),
                                               _ptr(self.indices), _ptr(self.base_cost),
                                               _ptr(self.edge_layer), _ptr(self.edge_kind)))

    def graph(self, cost):
        g = _abi.Graph()
        g.n_nodes, g.n_edges = self.N, self.E
        g.indptr, g.indices = self.indptr.data_ptr(), self.indices.data_ptr()
        g.cost, g.edge_kind = cost.data_ptr(), self.edge_kind.data_ptr()
        g.X, g.Y, g.L = self.X, self.Y, self.L
        return g


class DeviceGraph:
    """A caller's CSR (indptr i32[N+1], indices i32[E], numpy or device
    tensors) in HBM with the per-edge layer and kind the search reads.  With
    the lattice shape (X, Y, L), node = z*X*Y + y*X + x gives each edge's
    layers; without it the graph is one layer of N nodes (every edge planar,
    no goal-directed bounds), as the reference's full-graph GPU search lays
    it out (cuda_dijkstra.py:5552-5556)."""

    def __init__(self, ctx, indptr, indices, shape=None):
        torch = ctx.torch
        self.ctx = ctx
        ip = indptr.detach().cpu().numpy() if hasattr(indptr, "detach") else np.asarray(indptr)
        ix = indices.detach().cpu().numpy() if hasattr(indices, "detach") else np.asarray(indices)
        ip = np.ascontiguousarray(ip, dtype=np.int32)
        ix = np.ascontiguousarray(ix, dtype=np.int32)
        self.N, self.E = len(ip) - 1, int(ip[-1])
        pad = int(self.E == 0)
        if self.N < 1 or len(ix) < self.E or ip[0] != 0 or np.any(np.diff(ip) < 0):
            raise ValueError("indptr/indices do not form a CSR graph")
        if self.E and (ix[:self.E].min() < 0 or ix[:self.E].max() >= self.N):
            raise ValueError("CSR column index outside [0, N)")
        if shape is None:
            self.X, self.Y, self.L = self.N, 1, 1
        else:
            self.X, self.Y, self.L = (int(v) for v in shape)
            if self.X * self.Y * self.L != self.N:
                raise ValueError(f"lattice shape {shape} does not hold {self.N} nodes")
        plane = self.X * self.Y
        src_z = np.repeat((np.arange(self.N, dtype=np.int64) // plane).astype(np.int32), np.diff(ip))
        dst_z = ix[:self.E] // plane
        dev = ctx.device
        self.indptr = torch.from_numpy(ip).to(dev)
        self.indices = torch.from_numpy(np.concatenate([ix[:self.E], np.zeros(pad, np.int32)])).to(dev)
        self.edge_layer = torch.from_numpy(np.concatenate([src_z.astype(np.uint8),
                                                           np.zeros(pad, np.uint8)])).to(dev)
        self.edge_kind = torch.from_numpy(np.concatenate([(src_z != dst_z).astype(np.uint8),
                                                          np.zeros(pad, np.uint8)])).to(dev)

    def graph(self, cost):
        g = _abi.Graph()
        g.n_nodes, g.n_edges = self.N, self.E
        g.indptr, g.indices = self.indptr.data_ptr(), self.indices.data_ptr()
        g.cost, g.edge_kind = cost.data_ptr(), self.edge_kind.data_ptr()
        g.X, g.Y, g.L = self.X, self.Y, self.L
        return g


class SearchParams:
    """Search launch parameters.  h_xy / h_z: lower bounds of one planar step
    and of one layer of via span under the call's costs (cost_bounds); both 0
    runs the plain Dijkstra bucket order.  Paths are the same either way."""

    def __init__(self, slots=1, delta=1.0, n_buckets=8, rounds_per_poll=32, grid=0, count_minimal=False,
                 h_xy=0.0, h_z=0.0, local_batches=0, graph_rounds=False, engine=0, list_cap=0, lookahead=0,
                 waves=0):
        self.slots, self.delta, self.n_buckets = int(slots), float(delta), int(n_buckets)
        self.rounds_per_poll, self.grid, self.count_minimal = int(rounds_per_poll), int(grid), bool(count_minimal)
        self.h_xy, self.h_z = float(h_xy), float(h_z)
        self.local_batches = int(local_batches)
        self.graph_rounds = bool(graph_rounds)   # replay a poll's rounds as one captured HIP graph
        # two-pin searches: 0 = bucket rounds over all slots, 1 = one wavefront
        # per search (list_cap: entries per bucket list, 0 = default)
        self.engine, self.list_cap = int(engine), int(list_cap)
        self.lookahead = int(lookahead)   # engine 1: buckets an under-full step looks ahead (0 = 15, < 0 = none)
        self.waves = int(waves)           # engine 1: wavefronts per search, 1, 2 or 4 (0 = by searches in flight)

    def c(self):
        p = _abi.SsspParams()
        p.delta, p.n_buckets, p.slots = self.delta, self.n_buckets, self.slots
        p.rounds_per_poll, p.grid, p.count_minimal = self.rounds_per_poll, self.grid, int(self.count_minimal)
        p.h_xy, p.h_z = self.h_xy, self.h_z
        p.local_batches = self.local_batches
        p.launch_mode = int(self.graph_rounds)
        p.engine, p.list_cap, p.lookahead, p.waves = self.engine, self.list_cap, self.lookahead, self.waves
        return p


def cost_bounds(ctx, lat, cost, with_scale=False):
    """Lower bounds (h_xy, h_z) of the costs for the goal-directed search:
    the cheapest planar edge and the cheapest via edge per layer of span,
    shrunk by 2^-10 so that FP32 rounding of g + h can never exceed a true
    remaining cost; 0 where no finite edge of the kind exists.  with_scale:
    also return the median planar cost to within a factor of 2 (the upper
    edge of the binary-exponent bin that holds the median)."""
    torch = ctx.torch
    out = torch.empty(2, dtype=torch.float32, device=ctx.device)
    hist = torch.empty(64, dtype=torch.int32, device=ctx.device) if with_scale else None
    g = lat.graph(cost)
    _abi.check(ctx.lib.orgpu_cost_bounds(ctx.handle, ctypes.byref(g), _ptr(lat.edge_layer), _ptr(out), _ptr(hist)))
    b = out.cpu().numpy().astype(np.float64)
    res = []
    for v in b:
        res.append(0.0 if not np.isfinite(v) or v > 1e30 or v <= 0.0 else float(np.float32(v * (1.0 - 2.0 ** -10))))
    if not with_scale:
        return res[0], res[1]
    h = hist.cpu().numpy().astype(np.int64)
    med = 0.0
    if h.sum() > 0:
        k = int(np.searchsorted(np.cumsum(h), (h.sum() + 1) // 2))
        med = float(2.0 ** (k - 32 + 1))
    return res[0], res[1], med


def edge_list_sums(ctx, n_edges, lists, idx, val):
    """Per edge list (host integer arrays), the sum over its edges, in list
    order, of the per-edge quantity that is val at the edge ids idx and zero
    elsewhere: float64, one addition after another from 0.0, each term the
    float32 value widened -- the value of Python's sum(float(v[e]) for e in
    edges) -- computed on the device (orgpu_segment_sums, one wave per list).
    Returns float64[len(lists)]; raises on an edge id outside [0, n_edges)."""
    torch = ctx.torch
    n = len(lists)
    out = np.zeros(n, dtype=np.float64)
    if n == 0:
        return out
    lens = np.fromiter((len(e) for e in lists), dtype=np.int64, count=n)
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    if offs[-1] == 0 or len(idx) == 0:
        return out
    flat = np.concatenate(lists)
    lo, hi = int(flat.min()), int(flat.max())
    if lo < 0 or hi >= n_edges:
        raise ValueError(f"edge_list_sums: edge id {lo if lo < 0 else hi} outside [0, {n_edges})")
    idx = np.asarray(idx, dtype=np.int64)
    if len(idx) and (int(idx.min()) < 0 or int(idx.max()) >= n_edges):
        raise ValueError("edge_list_sums: value index outside the edge range")
    dev = ctx.device
    with torch.cuda.stream(ctx.stream):
        dense = torch.zeros(max(int(n_edges), 1), dtype=torch.float32, device=dev)
        dense[torch.from_numpy(idx).to(dev)] = torch.from_numpy(np.asarray(val, dtype=np.float32)).to(dev)
        e_d = torch.from_numpy(flat.astype(np.int32, copy=False)).to(dev)
        o_d = torch.from_numpy(offs).to(dev)
        res = torch.empty(n, dtype=torch.float64, device=dev)
        _abi.check(ctx.lib.orgpu_segment_sums(ctx.handle, int(n_edges), _ptr(dense), _ptr(e_d), _ptr(o_d), n,
                                              _ptr(res)))
        return res.cpu().numpy()


def route(ctx, lat, cost, requests, params, roi_bitmaps=None, keepout_owner=None, path_capacity=None):
    """Route requests = list of (src, dst, net_id, roi_index, use_keepout) on
    the lattice with edge costs `cost` (device f32[E]).  Returns
    (list of (nodes, edges, dist) or None per request, stats dict)."""
    nodes, edges, offs, dist, stats = route_flat(ctx, lat, cost, requests, params, roi_bitmaps, keepout_owner,
                                                 path_capacity)
    out = []
    ol = offs.tolist()
    dl = dist.tolist()
    for i in range(len(requests)):
        a, b = ol[i], ol[i + 1]
        out.append((nodes[a:b], edges[a:b - 1], dl[i]) if b - a > 1 else None)
    return out, stats


def route_flat(ctx, lat, cost, requests, params, roi_bitmaps=None, keepout_owner=None, path_capacity=None):
    """route() without the per-request list: (nodes, edges, offsets,
    dist, stats), request i's path is nodes[offsets[i]:offsets[i+1]] (a
    path of one node or none means no path) and its edges
    edges[offsets[i]:offsets[i+1]-1]."""
    n = len(requests)
    ctx.reserve(lat.N, params.slots, params.n_buckets)
    # orgpu_net_req rows (src, dst, net_id, roi_index, use_keepout, reserved)
    reqs = np.zeros((max(n, 1), 6), dtype=np.int32)
    if n:
        reqs[:n, :5] = np.asarray(requests, dtype=np.int64)
    if path_capacity is None:
        path_capacity = max(1024, n * 4 * (lat.X + lat.Y + 2 * lat.L))
    nodes = np.empty(path_capacity, dtype=np.int32)
    edges = np.empty(path_capacity, dtype=np.int32)
    offs = np.zeros(n + 1, dtype=np.int64)
    dist = np.zeros(max(n, 1), dtype=np.float32)
    st = _abi.SsspStats()
    g = lat.graph(cost)
    n_roi = 0 if roi_bitmaps is None else int(roi_bitmaps.shape[0])
    if roi_bitmaps is not None and roi_bitmaps.shape[1] * 32 < lat.N:
        raise ValueError("roi bitmap rows are shorter than the node count")
    rc = ctx.lib.orgpu_route(ctx.handle, ctypes.byref(g), reqs.ctypes.data_as(ctypes.c_void_p), n,
                             _ptr(roi_bitmaps), n_roi, _ptr(keepout_owner), ctypes.c_float(KEEPOUT_COST),
                             ctypes.byref(params.c()), nodes.ctypes.data_as(ctypes.c_void_p),
                             edges.ctypes.data_as(ctypes.c_void_p), path_capacity,
                             offs.ctypes.data_as(ctypes.c_void_p), dist.ctypes.data_as(ctypes.c_void_p),
                             ctypes.byref(st))
    _abi.check(rc)
    # views into compact copies of the used output (the capacity-sized
    # buffers are released; a kept path holds only its call's paths)
    used = int(offs[n])
    nodes, edges = nodes[:used].copy(), edges[:used].copy()
    stats = {"rounds": st.rounds, "items": st.items, "edges": st.edges,
             "minimal_nodes": st.minimal_nodes, "minimal_edges": st.minimal_edges,
             "kernel_ms": st.kernel_ms, "launches": st.launches, "round_ms": st.round_ms,
             "retried": st.retried}
    return nodes, edges, offs, dist[:n], stats


def solo_waves(waves, slots, n):
    """Wavefronts per search the one-wave engine runs n searches with on
    `slots` slots, as orgpu_route picks them: params.waves when it is 1, 2
    or 4; else 4 when all n start at once and n <= 1024, otherwise 2."""
    if int(waves) in (1, 2, 4):
        return int(waves)
    return 4 if n <= slots and n <= 1024 else 2


def reserve_solo(ctx, lat, slots, list_cap=0):
    """The one-wave engine's scratch for `slots` concurrent searches on
    lattice `lat` (orgpu_reserve_solo, after the handle's orgpu_reserve)."""
    ctx.reserve(lat.N, slots, 8)
    _abi.check(ctx.lib.orgpu_reserve_solo(ctx.handle, int(lat.N), int(lat.X), int(lat.Y), int(lat.L), int(slots),
                                          int(list_cap)))


def route_async(ctx, lat, cost, reqs_d, params, out, roi_bitmaps=None, keepout_owner=None):
    """Enqueues the searches of the device request rows reqs_d (int32
    [n, 6]: src, dst, net_id, roi_index, use_keepout, 0) on the handle's
    stream (orgpu_route_async) and returns at once.  out: a dict of device
    tensors from route_async_outputs.  After ctx.sync(): out["len"][i] nodes
    of request i's path at out["nodes"][out["off"][i]:] (edges likewise),
    out["dist"][i] its cost, out["used"] the entries written; a negative
    length is the code of include/orgpu.h."""
    n = int(reqs_d.shape[0])
    g = lat.graph(cost)
    n_roi = 0 if roi_bitmaps is None else int(roi_bitmaps.shape[0])
    _abi.check(ctx.lib.orgpu_route_async(ctx.handle, ctypes.byref(g), _ptr(reqs_d), n, _ptr(roi_bitmaps), n_roi,
                                         _ptr(keepout_owner), ctypes.c_float(KEEPOUT_COST), ctypes.byref(params.c()),
                                         _ptr(out["nodes"]), _ptr(out.get("edges")), int(out["nodes"].numel()),
                                         _ptr(out["off"]), _ptr(out["len"]), _ptr(out["dist"]), _ptr(out["used"])))


def route_async_outputs(ctx, n, capacity, edges=True):
    """Device result buffers for route_async: n requests, `capacity` packed
    path entries."""
    torch = ctx.torch
    dev = ctx.device
    out = {"nodes": torch.empty(max(int(capacity), 1), dtype=torch.int32, device=dev),
           "off": torch.zeros(max(n, 1), dtype=torch.int32, device=dev),
           "len": torch.zeros(max(n, 1), dtype=torch.int32, device=dev),
           "dist": torch.zeros(max(n, 1), dtype=torch.float32, device=dev),
           "used": torch.zeros(1, dtype=torch.int32, device=dev)}
    if edges:
        out["edges"] = torch.empty(max(int(capacity), 1), dtype=torch.int32, device=dev)
    return out


MS_POINT = np.dtype([("node", "<i4"), ("cost", "<f4")])
# orgpu_roi_req / orgpu_roi_info rows (the ctypes layouts of _abi.RoiReq / RoiInfo)
ROI_REQ = np.dtype({"names": ["src", "dst", "max_depth", "roi_index", "max_nodes", "readd", "reserved", "keep"],
                    "formats": ["<i4", "<i4", "<i4", "<i4", "<i8", "<i4", "<i4", ("<i4", 2)],
                    "offsets": [0, 4, 8, 12, 16, 24, 28, 32], "itemsize": 40})
ROI_INFO = np.dtype({"names": ["seen", "kept", "depth", "met", "src_in", "dst_in", "front", "left", "appended"],
                     "formats": ["<i8", "<i8", "<i4", "<i4", "<i4", "<i4", ("<i4", 4), ("<i4", 4), ("<i4", 4)],
                     "offsets": [0, 8, 16, 20, 24, 28, 32, 48, 64], "itemsize": 80})
# orgpu_ref_req rows
REF_REQ = np.dtype({"names": ["src", "dst", "net_id", "roi_index", "use_keepout", "n_rank", "rank_node", "rank_key"],
                    "formats": ["<i4", "<i4", "<i4", "<i4", "<i4", "<i4", ("<i4", 8), ("<u4", 8)],
                    "offsets": [0, 4, 8, 12, 16, 20, 24, 56], "itemsize": 88})


def _ms_points(pts):
    """(node, cost) pairs -> orgpu_ms_point rows; pts is a sequence of pairs
    or a (nodes, costs) pair of arrays."""
    if isinstance(pts, tuple) and len(pts) == 2 and isinstance(pts[0], np.ndarray):
        out = np.empty(len(pts[0]), dtype=MS_POINT)
        out["node"], out["cost"] = pts[0], pts[1]
        return out
    out = np.empty(len(pts), dtype=MS_POINT)
    if len(pts):
        a = np.asarray(pts, dtype=np.float64).reshape(-1, 2)
        out["node"], out["cost"] = a[:, 0], a[:, 1]
    return out


def route_ms(ctx, lat, cost, requests, params, roi_bitmaps=None, keepout_owner=None, path_capacity=None):
    """Multi-source / multi-sink searches.  requests = list of (seeds,
    targets, net_id, roi_index, use_keepout) with seeds / targets lists of
    (node, entry cost) pairs, or (nodes, costs) array pairs.  Returns (list
    of (nodes, edges, total, target_entry) or None per request, stats
    dict); a chosen target that is itself a seed gives a one-node path."""
    n = len(requests)
    ctx.reserve(lat.N, params.slots, params.n_buckets)
    reqs = np.zeros((max(n, 1), 8), dtype=np.int32)
    sl, tl = [], []
    so = to = 0
    for i, (sd, tg, nid, roi, ko) in enumerate(requests):
        sp, tp = _ms_points(sd), _ms_points(tg)
        reqs[i] = (so, len(sp), to, len(tp), int(nid), int(roi), int(ko), 0)
        so += len(sp)
        to += len(tp)
        sl.append(sp)
        tl.append(tp)
    sarr = np.concatenate(sl) if sl else np.empty(0, MS_POINT)
    tarr = np.concatenate(tl) if tl else np.empty(0, MS_POINT)
    if len(sarr) == 0:
        sarr = np.zeros(1, MS_POINT)
    if len(tarr) == 0:
        tarr = np.zeros(1, MS_POINT)
    if path_capacity is None:
        path_capacity = max(1024, n * 4 * (lat.X + lat.Y + 2 * lat.L))
    nodes = np.empty(path_capacity, dtype=np.int32)
    edges = np.empty(path_capacity, dtype=np.int32)
    offs = np.zeros(n + 1, dtype=np.int64)
    total = np.zeros(max(n, 1), dtype=np.float64)
    choice = np.zeros(max(n, 1), dtype=np.int32)
    st = _abi.SsspStats()
    g = lat.graph(cost)
    n_roi = 0 if roi_bitmaps is None else int(roi_bitmaps.shape[0])
    if roi_bitmaps is not None and roi_bitmaps.shape[1] * 32 < lat.N:
        raise ValueError("roi bitmap rows are shorter than the node count")
    vp = ctypes.c_void_p
    rc = ctx.lib.orgpu_route_ms(ctx.handle, ctypes.byref(g), reqs.ctypes.data_as(vp), n,
                                sarr.ctypes.data_as(vp), so, tarr.ctypes.data_as(vp), to,
                                _ptr(roi_bitmaps), n_roi, _ptr(keepout_owner), ctypes.c_float(KEEPOUT_COST),
                                ctypes.byref(params.c()), nodes.ctypes.data_as(vp), edges.ctypes.data_as(vp),
                                path_capacity, offs.ctypes.data_as(vp), total.ctypes.data_as(vp),
                                choice.ctypes.data_as(vp), ctypes.byref(st))
    _abi.check(rc)
    used = int(offs[n])
    nodes, edges = nodes[:used].copy(), edges[:used].copy()
    out = []
    ol = offs.tolist()
    for i in range(n):
        a, b = ol[i], ol[i + 1]
        out.append((nodes[a:b], edges[a:b - 1], float(total[i]), int(choice[i])) if b > a else None)
    stats = {"rounds": st.rounds, "items": st.items, "edges": st.edges,
             "kernel_ms": st.kernel_ms, "launches": st.launches, "round_ms": st.round_ms}
    return out, stats


ROI_LATTICE_MAX_DEPTH = 1024


def roi_bfs(ctx, lat, requests, rows, readd=True, method="auto", raw=False, keeps=None):
    """Batched ROI extraction.  requests: list of (src, dst, max_depth,
    max_nodes, row); rows: device u32[n_roi][ceil(N/32)] written in place.
    keeps: per request up to two further must-keep nodes (a portal net's
    seeds), an (n, 2) array with -1 for none.
    readd: re-add src/dst after truncation (the main search of _route_all) or
    not (its wider-ROI fallback).  method: "bfs" traverses the CSR;
    "lattice" computes the same ROI in closed form from the lattice shape
    (DeviceLattice only); "auto" takes "lattice" where it applies.
    Returns a list of (seen, kept, depth, met, src_in, dst_in), or with
    raw=True the ROI_INFO records themselves (which also say where the
    reference's roi_nodes list holds nodes out of ascending order)."""
    n = len(requests)
    if n == 0:
        return []
    q = np.asarray(requests, dtype=np.int64).reshape(n, 5)
    reqs = np.zeros(n, dtype=ROI_REQ)
    reqs["src"], reqs["dst"], reqs["max_depth"] = q[:, 0], q[:, 1], q[:, 2]
    reqs["max_nodes"], reqs["roi_index"] = q[:, 3], q[:, 4]
    reqs["readd"] = int(bool(readd))
    reqs["keep"] = -1 if keeps is None else np.asarray(keeps, dtype=np.int64).reshape(n, 2)
    info = np.zeros(n, dtype=ROI_INFO)
    vp = ctypes.c_void_p
    closed = getattr(lat, "desc", None) is not None and int(q[:, 2].max()) <= ROI_LATTICE_MAX_DEPTH
    if method == "lattice" and not closed:
        raise ValueError("closed-form ROI needs a DeviceLattice and depths up to 1024")
    if method == "lattice" or (method == "auto" and closed):
        _abi.check(ctx.lib.orgpu_roi_lattice(ctx.handle, ctypes.byref(lat.desc), reqs.ctypes.data_as(vp), n,
                                             _ptr(rows), int(rows.shape[0]), info.ctypes.data_as(vp)))
    else:
        g = lat.graph(lat.base_cost)
        _abi.check(ctx.lib.orgpu_roi_bfs(ctx.handle, ctypes.byref(g), reqs.ctypes.data_as(vp), n,
                                         _ptr(rows), int(rows.shape[0]), info.ctypes.data_as(vp)))
    if raw:
        return info
    return list(zip(info["seen"].tolist(), info["kept"].tolist(), info["depth"].tolist(), info["met"].tolist(),
                    (info["src_in"] != 0).tolist(), (info["dst_in"] != 0).tolist()))


def rank_exceptions(info, n_nodes):
    """(node, key) pairs of the reference-order search for an ROI whose
    ROI_INFO record is `info`: the keys order nodes like their positions in
    the reference's roi_nodes list (UPF:1716-1727, 4689-4696), 2*v + 2 for
    the ascending part; a must-keep swapped into position p takes the key of
    the node that left it, and the j-th appended node comes after every
    other (2*N + 2 + j)."""
    out = []
    for p in range(4):
        if info["front"][p] >= 0:
            out.append((int(info["front"][p]), 2 * int(info["left"][p]) + 2))
    for j in range(4):
        if info["appended"][j] >= 0:
            out.append((int(info["appended"][j]), 2 * int(n_nodes) + 2 + j))
    return out


def ref_requests(requests, ranks=None):
    """orgpu_ref_req rows from (src, dst, net_id, roi_index, use_keepout)
    rows and, per request, a list of (node, key) rank exceptions."""
    q = np.asarray(requests, dtype=np.int64).reshape(-1, 5)
    rows = np.zeros(len(q), dtype=REF_REQ)
    rows["src"], rows["dst"], rows["net_id"] = q[:, 0], q[:, 1], q[:, 2]
    rows["roi_index"], rows["use_keepout"] = q[:, 3], q[:, 4]
    if ranks is not None:
        for i, ex in enumerate(ranks):
            if not ex:
                continue
            if len(ex) > 8:
                raise ValueError(f"request {i}: {len(ex)} rank exceptions, at most eight are supported (an ROI "
                                 "holds at most four swapped-in and six appended must-keeps, and only a "
                                 "truncated ROI swaps)")
            rows["n_rank"][i] = len(ex)
            for k, (v, key) in enumerate(ex):
                rows["rank_node"][i, k] = v
                rows["rank_key"][i, k] = key
    return rows


REF_ERRORS = {-1: "path output capacity", -2: "heap capacity", -3: "path longer than the staging row",
              -4: "invalid request"}


def ref_slots_for(n_nodes, heap_cap, budget=32 << 30):
    """Concurrent reference-order searches whose scratch (16 B per node and
    16 B per heap entry each) fits `budget` bytes, at most 1024."""
    per = 16 * int(n_nodes) + 16 * int(heap_cap)
    return max(1, min(1024, int(budget) // per))


def route_ref(ctx, lat, cost, requests, roi_bitmaps=None, keepout_owner=None, ranks=None, slots=None,
              heap_cap=None, path_capacity=None):
    """Reference-order searches (SimpleDijkstra.find_path_roi's own pop
    order and float64 / float32 arithmetic).  requests: (src, dst, net_id,
    roi_index, use_keepout) rows; ranks: per request the (node, key) rank
    exceptions of its ROI (rank_exceptions), or None.  Returns (nodes,
    edges, offsets, dist, stats) like route_flat: request i's path is
    nodes[offsets[i]:offsets[i+1]] (fewer than two nodes = None)."""
    torch = ctx.torch
    rows = ref_requests(requests, ranks)
    n = len(rows)
    dev = ctx.device
    if heap_cap is None:
        heap_cap = int(min(max(4096, lat.N), 1 << 20))
    cap = ref_slots_for(lat.N, heap_cap)
    slots = cap if slots is None else min(int(slots), cap)
    slots = max(1, min(slots, max(n, 1)))
    key = (int(lat.N), slots, int(heap_cap))
    have = getattr(ctx, "_ref_reserved", None)
    if have is None or not all(a >= b for a, b in zip(have, key)):
        _abi.check(ctx.lib.orgpu_reserve_ref(ctx.handle, key[0], key[1], key[2]))
        ctx._ref_reserved = key
    if n == 0:
        return (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(1, np.int64), np.zeros(0, np.float32),
                {"pops": 0, "edges": 0, "kernel_ms": 0.0})
    if path_capacity is None:
        path_capacity = max(1024, n * 4 * (lat.X + lat.Y + 2 * lat.L))
    n_roi = 0 if roi_bitmaps is None else int(roi_bitmaps.shape[0])
    if roi_bitmaps is not None and roi_bitmaps.shape[1] * 32 < lat.N:
        raise ValueError("roi bitmap rows are shorter than the node count")
    g = lat.graph(cost)
    d_req = torch.from_numpy(rows.view(np.uint8)).to(dev)
    for _ in range(4):
        nodes = torch.empty(max(path_capacity, 1), dtype=torch.int32, device=dev)
        edges = torch.empty(max(path_capacity, 1), dtype=torch.int32, device=dev)
        off = torch.zeros(n, dtype=torch.int64, device=dev)
        ln = torch.zeros(n, dtype=torch.int32, device=dev)
        dist = torch.zeros(n, dtype=torch.float32, device=dev)
        cnt = torch.zeros(4, dtype=torch.int64, device=dev)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(ctx.stream)
        _abi.check(ctx.lib.orgpu_route_ref(ctx.handle, ctypes.byref(g), _ptr(d_req), n, _ptr(roi_bitmaps), n_roi,
                                           _ptr(keepout_owner), ctypes.c_float(KEEPOUT_COST), slots, _ptr(nodes),
                                           _ptr(edges), int(path_capacity), _ptr(off), _ptr(ln), _ptr(dist),
                                           _ptr(cnt)))
        e1.record(ctx.stream)
        ctx.sync()
        lens = ln.cpu().numpy()
        if (lens == -1).any():
            path_capacity *= 2        # the output ran out: again with room for every path
            continue
        break
    bad = lens[lens < 0]
    if len(bad):
        raise _abi.OrgpuError(_abi.ORGPU_ECAPACITY if bad[0] != -4 else _abi.ORGPU_EINVAL,
                              f"route_ref: {REF_ERRORS.get(int(bad[0]), int(bad[0]))} "
                              f"({len(bad)} of {n} requests)")
    offs_dev = off.cpu().numpy()
    c = cnt.cpu().numpy()
    used = int(c[2])
    allnodes = nodes[:used].cpu().numpy()
    alledges = edges[:used].cpu().numpy()
    # compact, in request order
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    idx = np.repeat(offs_dev - offs[:-1], lens) + np.arange(int(offs[-1]), dtype=np.int64)
    stats = {"pops": int(c[0]), "edges": int(c[1]), "kernel_ms": float(e0.elapsed_time(e1)), "slots": slots}
    return allnodes[idx], alledges[idx], offs, dist.cpu().numpy(), stats


class Sparse:
    """Host buffers receiving the nonzeros of a per-edge quantity."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.idx = np.empty(max(self.capacity, 1), dtype=np.int32)
        self.val = np.empty(max(self.capacity, 1), dtype=np.float32)

    def c(self):
        s = _abi.SparseOut()
        s.idx = self.idx.ctypes.data_as(ctypes.c_void_p)
        s.val = self.val.ctypes.data_as(ctypes.c_void_p)
        s.capacity = self.capacity
        s.count = 0
        return s

    def fetch(self, fn, *args):
        """Call fn(*args, &out); returns (idx, val) sorted by edge index."""
        s = self.c()
        rc = fn(*args, ctypes.byref(s))
        if rc == -4 and s.count > self.capacity:
            self.__init__(int(s.count * 1.25) + 1024)
            s = self.c()
            rc = fn(*args, ctypes.byref(s))
        _abi.check(rc)
        n = int(s.count)
        idx = self.idx[:n].astype(np.int64)
        val = self.val[:n].copy()
        order = np.argsort(idx, kind="stable")
        return idx[order], val[order]


class EdgeState:
    """EdgeAccountant state in HBM (UPF:799-1043): canonical usage counts,
    present, present_ema, history and the iteration's total cost."""

    def __init__(self, ctx, lat):
        torch = ctx.torch
        self.ctx, self.lat = ctx, lat
        E, dev = max(lat.E, 1), ctx.device
        self.count = torch.zeros(E, dtype=torch.int32, device=dev)
        self.present = torch.zeros(E, dtype=torch.float32, device=dev)
        self.present_ema = torch.zeros(E, dtype=torch.float32, device=dev)
        self.history = torch.zeros(E, dtype=torch.float32, device=dev)
        self.total_cost = torch.zeros(E, dtype=torch.float32, device=dev)
        self.sparse = Sparse(1 << 16)

    def scatter(self, edge_ids, delta):
        ids = np.ascontiguousarray(edge_ids, dtype=np.int32)
        if len(ids) == 0:
            return
        _abi.check(self.ctx.lib.orgpu_usage_scatter(self.ctx.handle, _ptr(self.count), self.lat.E,
                                                    ids.ctypes.data_as(ctypes.c_void_p), len(ids), int(delta)))

    def build_present(self, penalty_edges=None, penalty=0.0):
        pe = np.ascontiguousarray(penalty_edges if penalty_edges is not None else [], dtype=np.int32)
        _abi.check(self.ctx.lib.orgpu_present_build(self.ctx.handle, self.lat.E, _ptr(self.count), _ptr(self.present),
                                                    pe.ctypes.data_as(ctypes.c_void_p), len(pe),
                                                    ctypes.c_float(np.float32(penalty))))

    def update_ema(self, first, beta, one_minus_beta):
        L = self.ctx.lib
        return self.sparse.fetch(L.orgpu_present_ema, self.ctx.handle, self.lat.E, _ptr(self.count),
                                 _ptr(self.present_ema), int(first), ctypes.c_float(beta),
                                 ctypes.c_float(one_minus_beta))

    def overuse_sparse(self, mode=0, hist_scale=0.1, threshold=0.5):
        L = self.ctx.lib
        return self.sparse.fetch(L.orgpu_overuse_sparse, self.ctx.handle, self.lat.E, _ptr(self.present),
                                 _ptr(self.history), int(mode), ctypes.c_float(hist_scale),
                                 ctypes.c_float(threshold))

    def update_costs(self, params):
        lat = self.lat
        _abi.check(self.ctx.lib.orgpu_cost_update(self.ctx.handle, ctypes.byref(params), lat.N, lat.E,
                                                  _ptr(lat.indptr), _ptr(lat.indices), _ptr(lat.edge_kind),
                                                  _ptr(lat.base_cost), _ptr(self.present_ema), _ptr(self.history),
                                                  _ptr(self.total_cost)))

    def update_history(self, decay, gain, cap_mult, use_cap=True):
        _abi.check(self.ctx.lib.orgpu_history_update(self.ctx.handle, self.lat.E, _ptr(self.present),
                                                     _ptr(self.lat.base_cost), _ptr(self.history),
                                                     ctypes.c_float(decay), ctypes.c_float(gain),
                                                     ctypes.c_float(cap_mult), int(bool(use_cap))))
