"""Solver seam: the object the reference keeps in ``SimpleDijkstra.gpu_solver``
(UPF:2205-2218) and calls for ROI searches (``find_path_roi_gpu``, UPF:1814)
and for portal nets on the full graph (``find_path_fullgraph_gpu_seeds``,
UPF:4558-4565), plus a batched entry (``find_paths_batch``) that routes many
nets in one engine call.

Every search runs in liborgpu.so's exact delta-stepping engine on the GPU
(canonical FP32 order: lowest predecessor index among equal-cost paths).
There is no CPU fallback: without the library or a GPU the constructor
raises, and the reference's caller then takes its own CPU path
(UPF:1819-1821).
"""
import numpy as np

from .device import (Context, DeviceGraph, SearchParams, cost_bounds, max_slots, route as device_route,
                     route_ms as device_route_ms)
from .steiner import grow_trees


def _bits_from_nodes(nodes, N):
    mask = np.zeros(N, dtype=np.uint8)
    mask[np.asarray(nodes, dtype=np.int64)] = 1
    return _bits_from_mask(mask)


def _bits_from_mask(mask):
    b = np.packbits(np.asarray(mask, dtype=np.uint8), bitorder="little")
    return np.concatenate([b, np.zeros((-len(b)) % 4, np.uint8)]).view(np.uint32)


class GpuSolver:
    """GPU shortest-path solver over one CSR graph.

    indptr / indices: the router's CSR (``CSRGraph.indptr/indices``, numpy or
    device tensors); x_steps, y_steps, layers: the lattice shape
    (node = z*X*Y + y*X + x), which gives each edge's layers for the keepout
    rule and the goal-directed bounds; without it the graph is searched in
    plain Dijkstra bucket order.  base_costs is accepted for the reference's
    constructor shape; every call passes its own costs."""

    def __init__(self, indptr, indices, base_costs=None, x_steps=None, y_steps=None, layers=None, *,
                 ctx=None, device=0, slots=64):
        self.ctx = ctx or Context(device)
        shape = None if x_steps is None else (int(x_steps), int(y_steps), int(layers))
        self.graph = DeviceGraph(self.ctx, indptr, indices, shape)
        self.N = self.graph.N
        self.slots = int(slots)
        self.goal_directed = shape is not None
        self.last_stats = None

    # ---- helpers -------------------------------------------------------------
    def _costs(self, costs):
        torch = self.ctx.torch
        if isinstance(costs, torch.Tensor):
            c = costs.to(self.ctx.device, torch.float32).contiguous()
        else:
            if hasattr(costs, "get"):
                costs = costs.get()
            c = torch.from_numpy(np.ascontiguousarray(costs, dtype=np.float32)).to(self.ctx.device)
        if c.numel() < self.graph.E:
            raise ValueError(f"costs hold {c.numel()} entries, the graph has {self.graph.E} edges")
        return c

    def _params(self, c, slots):
        """Bucket width from the median edge cost; goal-directed bounds when
        the lattice shape is known."""
        hx, hz, med = cost_bounds(self.ctx, self.graph, c, with_scale=True)
        if not self.goal_directed:
            hx = hz = 0.0
        delta = max(2.4 * med, 1e-3) if med > 0 else 1.0
        return SearchParams(slots=max(1, min(slots, self.slots, max_slots(self.N))), delta=delta, n_buckets=8,
                            h_xy=hx, h_z=hz)

    def _rows(self, rows):
        torch = self.ctx.torch
        return torch.from_numpy(np.ascontiguousarray(np.stack(rows))).to(self.ctx.device)

    # ---- reference entry points ---------------------------------------------
    def find_path_roi_gpu(self, src, dst, costs, roi_nodes, global_to_roi):
        """Shortest src -> dst path inside the ROI (CDJ:4909-4996 interface):
        a list of global node ids, or None when dst is unreachable or equal to
        src."""
        g2r = global_to_roi.get() if hasattr(global_to_roi, "get") else global_to_roi
        g2r = np.asarray(g2r)
        if g2r[src] < 0 or g2r[dst] < 0:
            return None
        res = self.find_paths_batch(costs, [(int(src), int(dst))], roi_spec=[_bits_from_mask(g2r >= 0)])
        return res[0]

    def find_path_fullgraph_gpu_seeds(self, costs, src_seeds, dst_targets, ub_hint=None, *,
                                      allowed_bitmap=None, use_bitmap=False):
        """Multi-source / multi-sink search on the full graph (CDJ:5467-5751
        interface): seeds and targets at cost 0, nodes restricted to the set
        bits of allowed_bitmap when use_bitmap.  Returns the path from a seed
        to the cheapest target (lowest node id among equal costs), or None
        when none is reachable or the best exceeds ub_hint."""
        src = np.asarray(src_seeds, dtype=np.int64).ravel()
        dst = np.asarray(dst_targets, dtype=np.int64).ravel()
        if len(src) == 0 or len(dst) == 0:
            return None
        rows = None
        if use_bitmap and allowed_bitmap is not None:
            bm = np.ascontiguousarray(allowed_bitmap, dtype=np.uint32).ravel()
            if len(bm) * 32 < self.N:
                raise ValueError("allowed_bitmap is shorter than the node count")
            rows = [bm[:(self.N + 31) // 32]]
        out = self.find_paths_multisource(costs, [([(int(v), 0.0) for v in src], [(int(v), 0.0) for v in dst])],
                                          roi_spec=rows)
        r = out[0]
        if r is None:
            return None
        path, total = r[0], r[1]
        if ub_hint is not None and total > ub_hint:
            return None
        return path

    # ---- batched entries ---------------------------------------------------------
    def find_paths_batch(self, costs, nets, roi_spec=None, keepout_spec=None):
        """Route many two-pin nets in one engine call on one cost snapshot.

        nets: (src, dst) or (src, dst, net_id) per net.  roi_spec: None (whole
        graph) or one entry per net: None, a node list, or a packed u32
        bitmap row (bit set = in the ROI).  keepout_spec: None or the node
        owner array i32[N] (-1 free): planar edges leaving a node owned by
        another net count as keepout edges, which a path avoids before any
        cost (UPF:3150-3228).  Returns one path (list of node ids) or None
        per net."""
        torch = self.ctx.torch
        c = self._costs(costs)
        rows, reqs = [], []
        for i, n in enumerate(nets):
            s, d = int(n[0]), int(n[1])
            nid = int(n[2]) if len(n) > 2 else i
            ri = -1
            if roi_spec is not None and roi_spec[i] is not None:
                r = np.asarray(roi_spec[i])
                ri = len(rows)
                rows.append(r.astype(np.uint32) if r.dtype == np.uint32 else _bits_from_nodes(r, self.N))
            reqs.append((s, d, nid, ri, int(keepout_spec is not None)))
        ko = None
        if keepout_spec is not None:
            ko = keepout_spec if isinstance(keepout_spec, torch.Tensor) else torch.from_numpy(
                np.ascontiguousarray(keepout_spec, dtype=np.int32))
            ko = ko.to(self.ctx.device, torch.int32).contiguous()
        res, self.last_stats = device_route(self.ctx, self.graph, c, reqs, self._params(c, len(reqs)),
                                            roi_bitmaps=self._rows(rows) if rows else None, keepout_owner=ko)
        return [None if r is None else [int(v) for v in r[0]] for r in res]

    def find_paths_multisource(self, costs, requests, roi_spec=None):
        """Many multi-source / multi-sink searches in one engine call.
        requests: (seeds, targets) per search, each a list of (node, entry
        cost); roi_spec as in find_paths_batch.  Returns (path, total,
        target entry) or None per search; the path may be a single node when
        the cheapest target is itself a seed."""
        c = self._costs(costs)
        rows, reqs = [], []
        for i, (sd, tg) in enumerate(requests):
            ri = -1
            if roi_spec is not None and roi_spec[i] is not None:
                r = np.asarray(roi_spec[i])
                ri = len(rows)
                rows.append(r.astype(np.uint32) if r.dtype == np.uint32 else _bits_from_nodes(r, self.N))
            reqs.append((list(sd), list(tg), i, ri, 0))
        p = self._params(c, len(reqs))
        p.h_xy = p.h_z = 0.0
        res, self.last_stats = device_route_ms(self.ctx, self.graph, c, reqs, p,
                                               roi_bitmaps=self._rows(rows) if rows else None)
        return [None if r is None else ([int(v) for v in r[0]], r[2], r[3]) for r in res]

    def route_trees(self, costs, nets, roi_spec=None):
        """Multi-terminal nets as trees (orthoroute_amd.steiner.grow_trees):
        nets is a list of pin lists; roi_spec as in find_paths_batch, one
        entry per net.  Returns grow_trees' per-net dicts."""
        c = self._costs(costs)
        rows, row_of = [], []
        for i in range(len(nets)):
            r = None if roi_spec is None else roi_spec[i]
            if r is None:
                row_of.append(-1)
                continue
            r = np.asarray(r)
            row_of.append(len(rows))
            rows.append(r.astype(np.uint32) if r.dtype == np.uint32 else _bits_from_nodes(r, self.N))
        p = self._params(c, len(nets))
        trees, self.last_stats = grow_trees(self.ctx, self.graph, c, nets, p,
                                            roi_rows=self._rows(rows) if rows else None, row_of=row_of)
        return trees
