"""PathFinder negotiated-congestion router on MI355X.

Drop-in for orthoroute/algorithms/manhattan/unified_pathfinder.py (UPF): the
same classes and entry points (PathFinderRouter / UnifiedPathFinder,
PathFinderConfig, GPUConfig, Lattice3D, CSRGraph, EdgeAccountant,
ROIExtractor, SimpleDijkstra) and the same negotiation schedule, with the
per-edge state (usage counts, present, present EMA, history, total cost) and
every per-edge / per-node pass running in hand-written gfx950 kernels
(liborgpu.so, include/orgpu.h).  The host keeps the per-net bookkeeping
(paths, edge lists, via columns, ownership) and the small decision
reductions, which it evaluates exactly as numpy would on the dense arrays
(orthoroute_amd.reductions).

Shortest paths are the FP32 packed-key semantics of the kernels: path cost
within 1e-5 of the reference's float64 Dijkstra, ties broken by the lowest
predecessor index.

Batching: within one negotiation iteration every net sees the same cost
snapshot (UPF:3722-3741 recomputes costs once per iteration; keepouts are
fixed per iteration), so all nets of the iteration are searched together,
K at a time, and committed afterwards in the reference's order.  The only
order-dependent step, the wider-ROI fallback for the first five failures
(UPF:4751-4781), is replayed in that order after the batch.  `batch_slots=1`
gives the plain one-net-at-a-time form.
"""
import ctypes
import gc
import logging
import math
import os
import random
import threading
import time
from collections import defaultdict
from typing import Dict, List, Optional, Set, Tuple

import numpy as np

from ... import _abi
from ...device import (Context, DeviceLattice, EdgeState, SearchParams, cost_bounds, edge_list_sums, max_slots,
                       rank_exceptions, roi_bfs, route as device_route, route_flat as device_route_flat,
                       route_ms as device_route_ms, route_ref as device_route_ref)
from ...autotune import BucketTuner, base_width, default_target
from ...lattice import DIR_H, DIR_HV, DIR_V, layer_directions, reference_via_pairs
from ...reductions import sum_f32
from .board_analyzer import analyze_board_characteristics
from .pad_escape_planner import PadEscapePlanner, Portal  # noqa: F401  (Portal: the reference's export)
from ...steiner import grow_trees
from .parameter_derivation import apply_derived_parameters, derive_routing_parameters

logger = logging.getLogger(__name__)
f32 = np.float32

ROI_THRESHOLD_STEPS = 125      # UPF:4617
MAX_ROI_NODES = 50_000         # UPF:1659
KEEPOUT_COST = 1e9             # UPF:3182


class NodePath(np.ndarray):
    """A routed net's node ids (int32) as the negotiation keeps them between
    iterations: an array view of the search output, true when non-empty like
    the reference's list.  route_multiple_nets hands back lists."""

    def __bool__(self):
        return self.size > 0


class GeometryPayload:
    def __init__(self, tracks, vias):
        self.tracks = tracks
        self.vias = vias


def _dedupe(items, key):
    """items in order, later ones with an earlier one's key dropped."""
    seen, out = set(), []
    for it in items:
        k = key(it)
        if k not in seen:
            seen.add(k)
            out.append(it)
    return out


class GPUConfig:
    """UPF:549-561 (the flags of the reference's CuPy path; kept for callers)."""
    GPU_MODE = True
    DEBUG_INVARIANTS = True
    USE_PERSISTENT_KERNEL = False
    USE_GPU_COMPACTION = True
    USE_DELTA_STEPPING = True
    DELTA_VALUE = 0.5


class PathFinderConfig:
    """UPF:564-634 defaults, plus this build's execution knobs:
    lattice_policy ('reference' | 'adjacent'), batch_slots (concurrent
    searches per launch), sssp_delta / sssp_buckets (bucket width and ring
    size; None = autotuned), max_iterations_cap, count_minimal,
    goal_directed."""
    max_iterations: int = 40
    pres_fac_init: float = 1.0
    pres_fac_mult: float = 1.10
    pres_fac_max: float = 64.0
    hist_gain: float = 0.2
    base_cost_weight: float = 0.3
    grid_pitch: float = 0.4
    via_cost: float = 0.7
    portal_discount: float = 0.4
    span_alpha: float = 0.15
    iter1_always_connect: bool = True
    portal_enabled: bool = True
    portal_delta_min: int = 3
    portal_delta_max: int = 12
    portal_delta_pref: int = 6
    portal_x_snap_max: float = 0.5
    portal_via_discount: float = 0.15
    portal_retarget_patience: int = 3
    stagnation_patience: int = 5
    use_gpu: bool = True
    batch_size: int = 32
    layer_count: int = 6
    strict_drc: bool = False
    mode: str = "near_far"
    roi_parallel: bool = False
    per_net_budget_s: float = 5.0
    max_roi_nodes: int = 750000
    delta_multiplier: float = 4.0
    adaptive_delta: bool = True
    strict_capacity: bool = True
    reroute_only_offenders: bool = True
    layer_shortfall_percentile: float = 95.0
    layer_shortfall_cap: int = 16
    enable_profiling: bool = False
    enable_instrumentation: bool = False
    strict_overuse_block: bool = False
    hist_cost_weight: float = 10.0
    log_iteration_details: bool = False
    acc_fac: float = 0.0
    phase_block_after: int = 2
    congestion_multiplier: float = 1.0
    max_search_nodes: int = 2000000
    layer_names = ("F.Cu", "In1.Cu", "In2.Cu", "In3.Cu", "In4.Cu", "B.Cu")   # UPF:632
    hotset_cap: int = 512
    allowed_via_spans = None
    # execution knobs of this build
    lattice_policy: str = "reference"
    batch_slots: int = 1024
    sssp_delta = None
    sssp_buckets = None
    sssp_local_batches = 0         # same-bucket batches per wave per round (0 = library default)
    sssp_grid = 0                  # workgroups per round launch (0 = library default)
    sssp_rounds_per_poll = 32      # rounds queued between status polls / slot refills
    sssp_graph_rounds = False      # replay the rounds between polls as one captured HIP graph
    sssp_engine = "auto"           # two-pin searches: 0 = bucket rounds over all slots, 1 = one
                                   # wavefront per search, "auto" = 1 when at least sssp_solo_min
                                   # searches run at once, else 0 (same paths either way)
    sssp_solo_min = 1024
    sssp_solo_nodes = 2_000_000    # "auto": small batches of a board up to this many nodes, in a
    sssp_solo_density = 0.001      # negotiation of at least this many nets per node, may move to
                                   # engine 1 (_small_engine_predict, _small_engine_record)
    sssp_list_cap = 0              # engine 1: bucket-list entries per search (0 = library default)
    sssp_lookahead = 0             # engine 1: buckets an under-full step looks ahead (0 = library
                                   # default, < 0 = none)
    sssp_solo_waves = 0            # engine 1: wavefronts per search, 1, 2 or 4 (0 = library default:
                                   # 4 when a batch's searches all start at once, else 2)
    sssp_target_buckets = None     # widen the bucket so a batch's deep searches span about this many
                                   # (None: 128 from 1 M nodes up, 32 below)
    max_iterations_cap = None
    iteration_geometry = True      # build the provisional tracks and vias that iteration_cb receives
                                   # each iteration (False: it receives empty lists)
    count_minimal = False          # also count Dijkstra-minimal nodes/edges per search
    goal_directed = True           # bucket on g + lower bound to the destination (same paths)
    sssp_order = "canonical"       # "canonical": FP32 costs, lowest-predecessor tie-break (batched
                                   # delta-stepping); "reference": the reference heap's float64 sums,
                                   # float32 acceptance and (alt, roi position) pop order, one
                                   # wavefront per search -- the reference's own trajectories


# ----------------------------------------------------------------------------
# geometry / lattice
# ----------------------------------------------------------------------------

class KiCadGeometry:
    """pathfinder/kicad_geometry.py:12-72: grid snapped to pitch multiples."""

    def __init__(self, bounds, pitch=0.4, layer_count=2):
        self.min_x, self.min_y, self.max_x, self.max_y = bounds
        self.pitch = pitch
        self.grid_min_x = round(self.min_x / pitch) * pitch
        self.grid_min_y = round(self.min_y / pitch) * pitch
        self.grid_max_x = round(self.max_x / pitch) * pitch
        self.grid_max_y = round(self.max_y / pitch) * pitch
        self.x_steps = int((self.grid_max_x - self.grid_min_x) / pitch) + 1
        self.y_steps = int((self.grid_max_y - self.grid_min_y) / pitch) + 1
        self.layer_count = layer_count

    def lattice_to_world(self, x, y):
        return (self.grid_min_x + x * self.pitch, self.grid_min_y + y * self.pitch)

    def world_to_lattice(self, wx, wy):
        return (round((wx - self.grid_min_x) / self.pitch), round((wy - self.grid_min_y) / self.pitch))

    def node_index(self, x, y, z):
        return z * (self.x_steps * self.y_steps) + y * self.x_steps + x

    def index_to_coords(self, n):
        plane = self.x_steps * self.y_steps
        z = n // plane
        r = n % plane
        return (r % self.x_steps, r // self.x_steps, z)


class CSRGraph:
    """The lattice CSR in HBM (UPF:647-792): indptr, indices, base_costs,
    edge_layer, edge_kind are torch tensors on the GPU."""

    def __init__(self, dlat: DeviceLattice):
        self.dlat = dlat
        self.indptr = dlat.indptr
        self.indices = dlat.indices
        self.base_costs = dlat.base_cost
        self.edge_layer = dlat.edge_layer
        self.edge_kind = dlat.edge_kind
        self.num_nodes, self.num_edges = dlat.N, dlat.E


class Lattice3D:
    """UPF:1075-1338: H/V discipline, via pairs, and build_graph on the GPU."""

    def __init__(self, bounds, pitch, layers, policy="reference"):
        self.bounds = bounds
        self.pitch = pitch
        self.layers = layers
        self.policy = policy
        self.geom = KiCadGeometry(bounds, pitch, layer_count=layers)
        self.x_steps = self.geom.x_steps
        self.y_steps = self.geom.y_steps
        self.num_nodes = self.x_steps * self.y_steps * layers
        self.layer_dir = ["hv", "hv"] if layers == 2 else ["v" if z == 0 else ("h" if z % 2 == 1 else "v")
                                                           for z in range(layers)]

    def get_legal_axis(self, layer):
        if layer >= len(self.layer_dir):
            return "h" if layer % 2 == 1 else "v"
        return self.layer_dir[layer]

    def is_legal_planar_edge(self, fx, fy, fl, tx, ty, tl):
        if fl != tl:
            return True
        dx, dy = abs(tx - fx), abs(ty - fy)
        if dx + dy != 1:
            return False
        d = self.get_legal_axis(fl)
        if d == "hv":
            return True
        if d == "h":
            return dy == 0 and dx == 1
        return dx == 0 and dy == 1

    def get_legal_via_pairs(self, layer_count):
        return set(reference_via_pairs(layer_count))

    def node_idx(self, x, y, z):
        return self.geom.node_index(x, y, z)

    def idx_to_coord(self, idx):
        return self.geom.index_to_coords(int(idx))

    def world_to_lattice(self, x_mm, y_mm):
        return self.geom.world_to_lattice(x_mm, y_mm)

    def build_graph(self, via_cost, allowed_via_spans=None, use_gpu=True, ctx=None) -> CSRGraph:
        ctx = ctx or Context(0)
        dlat = ctx.build_lattice(self.x_steps, self.y_steps, self.layers, self.policy,
                                 pitch=float(self.pitch), via_cost=float(via_cost))
        return CSRGraph(dlat)


# ----------------------------------------------------------------------------
# accountant
# ----------------------------------------------------------------------------

class EdgeAccountant:
    """UPF:799-1043 with the arrays in HBM.  `count` is the canonical usage
    store (UPF:807); present/present_ema/history/total_cost are float32 as in
    the reference; capacity is 1 for every edge (UPF:811) and is not stored."""

    def __init__(self, ctx, dlat):
        self.E = dlat.E
        self.state = EdgeState(ctx, dlat)
        self.present = self.state.present
        self.history = self.state.history

    @property
    def present_ema(self):
        return self.state.present_ema

    @property
    def total_cost(self):
        return self.state.total_cost

    def commit_path(self, edge_indices):
        self.state.scatter(edge_indices, +1)

    def clear_path(self, edge_indices):
        self.state.scatter(edge_indices, -1)

    def refresh_from_canonical(self):
        self.state.build_present()

    def compute_overuse(self, router_instance=None):
        self.state.build_present()
        idx, val = self.state.overuse_sparse(0)
        s = int(sum_f32(self.E, idx, val))
        col = seg = 0
        if router_instance is not None:
            col = int(np.maximum(0, router_instance.via_col_use - router_instance.via_col_cap).sum())
            if getattr(router_instance, "_segZ", 0) > 0:
                seg = int(np.maximum(0, router_instance.via_seg_use - router_instance.via_seg_cap).sum())
        return s + col + seg, int(len(idx))


# ----------------------------------------------------------------------------
# ROI / search front ends (API compatibility: single-net calls)
# ----------------------------------------------------------------------------

class ROIExtractor:
    """UPF:1345-1734.  extract_roi runs the GPU BFS for one net and returns
    (roi_nodes ascending, global_to_roi) like the reference."""

    def __init__(self, graph: CSRGraph, use_gpu=True, lattice=None, ctx=None):
        self.graph = graph
        self.lattice = lattice
        self.ctx = ctx or graph.dlat.ctx
        self.N = graph.num_nodes
        self.max_roi_nodes = MAX_ROI_NODES

    def roi_depth(self, src, dst, stagnation_bonus=0.0):
        sx, sy, _ = self.lattice.idx_to_coord(src)
        dx, dy, _ = self.lattice.idx_to_coord(dst)
        md = abs(dx - sx) + abs(dy - sy)
        radius = max(60, int(md * 0.75 + stagnation_bonus * 2.0))
        return min(int(radius + stagnation_bonus * 2.0), 150)

    def roi_depth_many(self, src, dst, bonus):
        """roi_depth for arrays of requests (same float64 arithmetic)."""
        X, plane = self.lattice.x_steps, self.lattice.x_steps * self.lattice.y_steps
        rs, rd = np.asarray(src, np.int64) % plane, np.asarray(dst, np.int64) % plane
        md = np.abs(rd % X - rs % X) + np.abs(rd // X - rs // X)
        b2 = np.asarray(bonus, np.float64) * 2.0
        radius = np.maximum(60, np.floor(md * 0.75 + b2).astype(np.int64))
        return np.minimum(np.floor(radius + b2).astype(np.int64), 150)

    def extract_roi(self, src, dst, initial_radius=40, stagnation_bonus=0.0, portal_seeds=None):
        torch = self.ctx.torch
        words = (self.N + 31) // 32
        rows = torch.zeros((1, words), dtype=torch.int32, device=self.ctx.device)
        depth = self.roi_depth(src, dst, stagnation_bonus)
        roi_bfs(self.ctx, self.graph.dlat, [(src, dst, depth, self.max_roi_nodes, 0)], rows, readd=False)
        bits = np.unpackbits(rows.cpu().numpy().view(np.uint8), bitorder="little")[:self.N]
        roi = np.nonzero(bits)[0].astype(np.int32)
        g2r = np.full(self.N, -1, dtype=np.int32)
        g2r[roi] = np.arange(len(roi), dtype=np.int32)
        return roi, g2r


class SimpleDijkstra:
    """UPF:1787-1989 surface: find_path_roi(src, dst, costs, roi_nodes,
    global_to_roi) runs one GPU search restricted to the ROI."""

    def __init__(self, graph: CSRGraph, lattice=None, ctx=None):
        self.graph = graph
        self.ctx = ctx or graph.dlat.ctx
        self.N = graph.num_nodes
        self.plane_size = lattice.x_steps * lattice.y_steps if lattice else None
        self.params = SearchParams(slots=1, delta=1.0, n_buckets=8)

    def find_path_roi(self, src, dst, costs, roi_nodes, global_to_roi):
        g2r = np.asarray(global_to_roi)
        if g2r[src] < 0 or g2r[dst] < 0:
            return None
        torch = self.ctx.torch
        mask = (g2r >= 0).astype(np.uint8)
        bits = np.packbits(mask, bitorder="little")
        bits = np.concatenate([bits, np.zeros((-len(bits)) % 4, np.uint8)]).view(np.int32)
        rows = torch.from_numpy(bits.reshape(1, -1)).to(self.ctx.device)
        if not isinstance(costs, torch.Tensor):
            costs = torch.from_numpy(np.ascontiguousarray(costs, dtype=np.float32)).to(self.ctx.device)
        res, _ = device_route(self.ctx, self.graph.dlat, costs, [(src, dst, -1, 0, 0)], self.params, roi_bitmaps=rows)
        r = res[0]
        return None if r is None else [int(v) for v in r[0]]

    def find_path_multisource_multisink(self, src_seeds, dst_targets, costs, roi_nodes, global_to_roi):
        """UPF:1889-1989: cheapest path from any seed (node, entry cost) to
        any target (node, exit cost) inside the ROI; targets are settled but
        not expanded and the search runs until its frontier is empty.
        Returns (path, entry_layer, exit_layer), or None when no target is
        reached or the best path is a single node.  Ties between targets go
        to the lower dist, then the lower node id."""
        g2r = np.asarray(global_to_roi.get() if hasattr(global_to_roi, "get") else global_to_roi)
        seeds = [(int(v), float(c)) for v, c in src_seeds if g2r[int(v)] >= 0]
        targets = [(int(v), float(c)) for v, c in dst_targets if g2r[int(v)] >= 0]
        if not seeds or not targets:
            return None
        torch = self.ctx.torch
        bits = np.packbits((g2r >= 0).astype(np.uint8), bitorder="little")
        bits = np.concatenate([bits, np.zeros((-len(bits)) % 4, np.uint8)]).view(np.int32)
        rows = torch.from_numpy(bits.reshape(1, -1)).to(self.ctx.device)
        if not isinstance(costs, torch.Tensor):
            costs = torch.from_numpy(np.ascontiguousarray(costs, dtype=np.float32)).to(self.ctx.device)
        res, _ = device_route_ms(self.ctx, self.graph.dlat, costs, [(seeds, targets, -1, 0, 0)], self.params,
                                 roi_bitmaps=rows)
        r = res[0]
        if r is None or len(r[0]) <= 1:
            return None
        path = [int(v) for v in r[0]]
        if self.plane_size:
            return path, path[0] // self.plane_size, path[-1] // self.plane_size
        return path, 0, 0


# ----------------------------------------------------------------------------
# router
# ----------------------------------------------------------------------------

class PathFinderRouter:
    """PathFinder negotiated congestion routing (UPF:1996-5914 surface)."""

    def __init__(self, config: PathFinderConfig = None, use_gpu: bool = None):
        self.config = config or PathFinderConfig()
        if use_gpu is not None:
            self.config.use_gpu = use_gpu
        env = os.getenv("USE_GPU")
        if env is not None:
            self.config.use_gpu = env == "1"
        self.lattice: Optional[Lattice3D] = None
        self.graph: Optional[CSRGraph] = None
        self.accounting: Optional[EdgeAccountant] = None
        self.solver: Optional[SimpleDijkstra] = None
        self.roi_extractor: Optional[ROIExtractor] = None
        self.ctx: Optional[Context] = None
        self.pad_to_node: Dict[str, int] = {}
        self.net_paths: Dict[str, List[int]] = {}
        self.iteration = 0
        self._negotiation_ran = False
        self._geometry_payload = GeometryPayload([], [])
        self.locked_nets: Set[str] = set()
        self.net_clean_streak: Dict[str, int] = defaultdict(int)
        self._net_to_edges: Dict[str, np.ndarray] = {}
        self._net_nodes: Dict[str, np.ndarray] = {}
        # multi-terminal nets routed as trees inside the negotiation
        # (multi_terminal_enabled): pins per net, and per routed tree its
        # branches (node arrays) and their edge arrays
        self._tree_pins: Dict[str, List[int]] = {}
        self._tree_branches: Dict[str, Tuple[list, list]] = {}
        self._last_depth_hint = None               # bucket sizing (speed only)
        self._ko_nodes = np.zeros(0, np.int64)      # via keepouts: node, first owner
        self._ko_owner = np.zeros(0, np.int64)
        self.portals = {}
        self.net_pad_ids: Dict[str, Tuple[str, str]] = {}
        self.net_portal_failures: Dict[str, int] = defaultdict(int)
        self.net_portal_layers: Dict[str, Tuple[int, int]] = {}
        self.escape_planner: Optional[PadEscapePlanner] = None
        self._escape_tracks, self._escape_vias = [], []
        self.stagnation_counter = 0
        self.full_graph_fallback_count = 0
        self._last_ripped: Set[str] = set()
        self._freeze_pres_fac_until = 0
        self.stats = defaultdict(float)
        self.iteration_log = []

    # ---------------- initialisation
    def initialize_graph(self, board) -> bool:
        """UPF:2088-2252 for a Board (pads on nets give the bounds)."""
        bounds = self._calc_bounds(board)
        layers = getattr(board, "layer_count", None) or len(getattr(board, "layers", []) or []) \
            or self.config.layer_count
        self.config.layer_count = int(layers)
        names = self.config.layer_names
        if not names or len(names) < self.config.layer_count:
            # layer names from the board (a list of names, or Layer objects)
            board_names = getattr(board, "layer_names", None) or getattr(board, "layers", None) or []
            board_names = [getattr(n, "name", n) for n in board_names]
            self.config.layer_names = (board_names or
                                       (["F.Cu"] + [f"In{i}.Cu" for i in range(1, self.config.layer_count - 1)]
                                        + ["B.Cu"]))
        self._init_lattice(bounds, self.config.grid_pitch, self.config.layer_count)
        self._map_pads(board)
        # the escape planner is built once the pads are mapped; it seeds the
        # router's generator (UPF:2237-2240)
        self.escape_planner = PadEscapePlanner(self.lattice, self.config, self.pad_to_node,
                                               random_seed=getattr(self.config, "escape_random_seed", 42),
                                               rng=self._rng)
        return True

    def initialize_lattice(self, x_steps, y_steps, layers, pitch=1.0, via_cost=None):
        """Synthetic boards: the lattice a board whose pads span
        (3*pitch, 3*pitch)..((x_steps-4)*pitch, (y_steps-4)*pitch) produces."""
        if via_cost is not None:
            self.config.via_cost = via_cost
        self.config.grid_pitch = pitch
        self.config.layer_count = layers
        bounds = (0.0, 0.0, (x_steps - 1) * pitch, (y_steps - 1) * pitch)
        self._init_lattice(bounds, pitch, layers)
        assert (self.lattice.x_steps, self.lattice.y_steps) == (x_steps, y_steps)
        return True

    def _init_lattice(self, bounds, pitch, layers):
        t0 = time.perf_counter()
        self.ctx = self.ctx or Context(0)
        self.lattice = Lattice3D(bounds, pitch, layers, getattr(self.config, "lattice_policy", "reference"))
        self.graph = self.lattice.build_graph(self.config.via_cost, ctx=self.ctx)
        # per-lattice engine choices start over
        for k in ("_rows_fit", "_small_engine"):
            self.__dict__.pop(k, None)
        self.N = self.lattice.num_nodes
        self.accounting = EdgeAccountant(self.ctx, self.graph.dlat)
        X, Y, L = self.lattice.x_steps, self.lattice.y_steps, layers
        self._Nx, self._Ny, self._Nz = X, Y, L
        self.layer_bias = np.ones(L, dtype=np.float32)
        self.via_col_cap = np.full((X, Y), int(getattr(self.config, "via_column_capacity", 4)), dtype=np.int16)
        self.via_col_use = np.zeros((X, Y), dtype=np.int16)
        self._segZ = max(L - 2, 0)
        if self._segZ > 0:
            cap = int(getattr(self.config, "via_segment_capacity", 2))
            self.via_seg_cap = np.full((X, Y, self._segZ), cap, dtype=np.int8)
            self.via_seg_use = np.zeros((X, Y, self._segZ), dtype=np.int16)
        # via pooling state in HBM ([x][y] and [x][y][segment] like the host
        # arrays): usage copied from the host's recount, pressure EMAs and
        # the segment prefix updated by orgpu_via_pressure, read by the cost
        # update
        torch, dev = self.ctx.torch, self.ctx.device
        self._via_dev = {
            "col_use": torch.zeros(X * Y, dtype=torch.int16, device=dev),
            "col_pres": torch.zeros(X * Y, dtype=torch.float32, device=dev),
            "seg_use": torch.zeros(max(X * Y * self._segZ, 1), dtype=torch.int16, device=dev),
            "seg_pres": torch.zeros(max(X * Y * self._segZ, 1), dtype=torch.float32, device=dev),
            "seg_prefix": torch.zeros(max(X * Y * self._segZ, 1), dtype=torch.float32, device=dev),
            "over": torch.zeros(1, dtype=torch.int64, device=dev),
        }
        self.node_owner = np.full(self.N, -1, dtype=np.int32)
        self.net_id_map = {}
        self.next_net_id = 0
        self.solver = SimpleDijkstra(self.graph, self.lattice, self.ctx)
        self.roi_extractor = ROIExtractor(self.graph, lattice=self.lattice, ctx=self.ctx)
        # PadEscapePlanner(random_seed=42) seeds Python's random at the end of
        # initialize_graph (UPF:2239-2240, pad_escape_planner.py:124); the
        # negotiation's draws continue from that state
        self._rng = random.Random(getattr(self.config, "escape_random_seed", 42))
        self._via_rank = None
        self.ctx.sync()
        self.stats["lattice_build_s"] = time.perf_counter() - t0

    def _calc_bounds(self, board):
        margin = 3.0
        pads = []
        for net in getattr(board, "nets", []) or []:
            if len(getattr(net, "pads", [])) >= 2:
                pads.extend(p for p in net.pads if getattr(p, "position", None) is not None)
        if pads:
            xs = [p.position.x for p in pads]
            ys = [p.position.y for p in pads]
            return (min(xs) - margin, min(ys) - margin, max(xs) + margin, max(ys) + margin)
        if hasattr(board, "_kicad_bounds"):
            b = board._kicad_bounds
            return (b[0] - margin, b[1] - margin, b[2] + margin, b[3] + margin)
        return (0, 0, 100, 100)

    def _pad_key(self, pad, comp=None):
        cid = getattr(pad, "component_id", None) or (getattr(comp, "id", None) if comp else None) \
            or "GENERIC_COMPONENT"
        if cid == "GENERIC_COMPONENT" and hasattr(pad, "position"):
            return f"{cid}_{pad.id}@{int(round(pad.position.x * 1000))},{int(round(pad.position.y * 1000))}"
        return f"{cid}_{pad.id}"

    def _get_pad_layer(self, pad):
        names = self.config.layer_names or []
        if getattr(pad, "layer", None):
            if str(pad.layer) in names:
                return names.index(str(pad.layer))
        if getattr(pad, "layers", None):
            if str(pad.layers[0]) in names:
                return names.index(str(pad.layers[0]))
        return 0

    def _map_pads(self, board):
        def snap(x_mm, y_mm, layer):
            x, y = self.lattice.world_to_lattice(x_mm, y_mm)
            x = max(0, min(x, self.lattice.x_steps - 1))
            y = max(0, min(y, self.lattice.y_steps - 1))
            return self.lattice.node_idx(x, y, layer)
        for comp in getattr(board, "components", []) or []:
            for pad in getattr(comp, "pads", []) or []:
                self.pad_to_node[self._pad_key(pad, comp)] = snap(pad.position.x, pad.position.y,
                                                                   self._get_pad_layer(pad))
        for pad in getattr(board, "pads", []) or []:
            k = self._pad_key(pad)
            if k not in self.pad_to_node:
                self.pad_to_node[k] = snap(pad.position.x, pad.position.y, self._get_pad_layer(pad))
        for net in getattr(board, "nets", []) or []:
            for pad in getattr(net, "pads", []) or []:
                k = self._pad_key(pad)
                if k not in self.pad_to_node and getattr(pad, "position", None) is not None:
                    self.pad_to_node[k] = snap(pad.position.x, pad.position.y, self._get_pad_layer(pad))

    # ---------------- legacy no-op API (UPF:5588-5594, 5619)
    def map_all_pads(self, board):
        return None

    def prepare_routing_runtime(self):
        return None

    def precompute_all_pad_escapes(self, board, nets_to_route=None):
        """UPF:5619-5650: plans the pads' portals (PadEscapePlanner), routes
        nets between them from now on and keeps the escape stubs for
        emit_geometry.  Returns (escape tracks, escape vias); nothing before
        initialize_graph(board) has built the planner."""
        if self.escape_planner is None:
            logger.error("precompute_all_pad_escapes: call initialize_graph(board) first")
            return [], []
        result = self.escape_planner.precompute_all_pad_escapes(board, nets_to_route)
        self.portals = dict(self.escape_planner.portals)
        self._escape_tracks, self._escape_vias = result
        return result

    # ---------------- small helpers
    def _get_net_id(self, name):
        if name not in self.net_id_map:
            self.net_id_map[name] = self.next_net_id
            self.next_net_id += 1
        return self.net_id_map[name]

    def _coords(self, n):
        return self.lattice.idx_to_coord(n)

    def _via_hops(self, path):
        """(x, y, z_lo, z_hi) of every via hop of a node path."""
        p = np.asarray(path, dtype=np.int64)
        plane = self._Nx * self._Ny
        z = p // plane
        r = p % plane
        hop = (r[1:] == r[:-1]) & (z[1:] != z[:-1])
        if not hop.any():
            return np.zeros((0, 4), np.int64)
        rr = r[:-1][hop]
        za, zb = z[:-1][hop], z[1:][hop]
        return np.stack([rr % self._Nx, rr // self._Nx, np.minimum(za, zb), np.maximum(za, zb)], axis=1)

    # ---------------- requests
    def route_multiple_nets(self, requests: List, progress_cb=None, iteration_cb=None) -> Dict:
        """UPF:2530-2598.  Nets with more than two pads are set aside as
        multi-terminal nets and not routed, as in the reference, whose
        PathFinderConfig has no multi_terminal_enabled attribute.  With
        config.multi_terminal_enabled set they are routed inside the
        negotiation as trees (as are requests (name, [three or more pin
        nodes])): each iteration a scheduled tree net is ripped up and regrown
        on the iteration's costs by multi-source searches from the tree to
        its unconnected pins (orthoroute_amd.steiner), its edges are committed
        like a path's, each branch marks its via barrels, and net_paths holds
        the tree's nodes in first-seen order.  A tree that cannot reach every
        pin counts as a failed net.  Trees search the whole graph without
        keepouts and take no wider-ROI fallback (the reference defines no
        multi-terminal search that runs; this is this build's definition)."""
        multi = [q for q in requests if hasattr(q, "pads") and len(q.pads) > 2]
        if multi and not getattr(self.config, "multi_terminal_enabled", False):
            requests = [q for q in requests if not (hasattr(q, "pads") and len(q.pads) > 2)]
        tasks = self._parse_requests(requests)
        if not tasks:
            self._negotiation_ran = True
            return {}
        # Python's cyclic collector stays off for the negotiation: the loop
        # makes no reference cycles, and a full collection walks every
        # object the process holds (on the 8,000-net backplane two or three
        # of them cost 50-100 ms each per negotiation).  It is switched back
        # on at the end if it was on.
        gc_on = gc.isenabled()
        gc.disable()
        try:
            return self._pathfinder_negotiation(tasks, progress_cb, iteration_cb)
        finally:
            if gc_on:
                gc.enable()
            self._paths_as_lists()

    def _paths_as_lists(self):
        """net_paths values as the reference's lists of node ids (the
        negotiation keeps them as NodePath arrays between iterations); in
        place, so a result dict that holds net_paths sees the lists."""
        for n, p in self.net_paths.items():
            if isinstance(p, np.ndarray):
                self.net_paths[n] = p.tolist()

    def _parse_requests(self, requests):
        """UPF:2600-2676: a Board net's task is its first two pads' nodes, or
        with portals on both pads the portals' entry-layer nodes.  A tree net
        (multi_terminal_enabled: more than two distinct pad nodes, or a
        request (name, [pins])) is a task of its first two pins, with its
        pins kept for the tree search."""
        tasks = {}
        trees_on = bool(getattr(self.config, "multi_terminal_enabled", False))
        for req in requests:
            if trees_on and hasattr(req, "name") and hasattr(req, "pads") and len(req.pads) > 2:
                pins = []
                for pad in req.pads:
                    k = self._pad_key(pad)
                    if k in self.pad_to_node and self.pad_to_node[k] not in pins:
                        pins.append(int(self.pad_to_node[k]))
                if len(pins) > 2:
                    self._tree_pins[req.name] = pins
                    tasks[req.name] = (pins[0], pins[1])
                    continue
            if trees_on and isinstance(req, tuple) and len(req) == 2 and not isinstance(req[1], (int, np.integer)):
                name, nodes = req
                pins = []
                for v in nodes:
                    if int(v) not in pins:
                        pins.append(int(v))
                if len(pins) > 2:
                    self._tree_pins[name] = pins
                    tasks[name] = (pins[0], pins[1])
                elif len(pins) == 2:
                    tasks[name] = (pins[0], pins[1])
                elif pins:
                    self.net_paths[name] = [pins[0]]
                continue
            if hasattr(req, "name") and hasattr(req, "pads"):
                if len(req.pads) < 2:
                    continue
                k1, k2 = self._pad_key(req.pads[0]), self._pad_key(req.pads[1])
                if k1 in self.pad_to_node and k2 in self.pad_to_node:
                    s, d = self.pad_to_node[k1], self.pad_to_node[k2]
                    self.net_pad_ids[req.name] = (k1, k2)
                    if self.config.portal_enabled and k1 in self.portals and k2 in self.portals:
                        p1, p2 = self.portals[k1], self.portals[k2]
                        s = self.lattice.node_idx(p1.x_idx, p1.y_idx, p1.entry_layer)
                        d = self.lattice.node_idx(p2.x_idx, p2.y_idx, p2.entry_layer)
                        self.net_portal_layers[req.name] = (p1.entry_layer, p2.entry_layer)
                    if s != d:
                        tasks[req.name] = (s, d)
                    else:
                        self.net_paths[req.name] = [s]
            elif isinstance(req, tuple) and len(req) == 3:
                name, s, d = req
                if isinstance(s, (int, np.integer)) and isinstance(d, (int, np.integer)):
                    s, d = int(s), int(d)
                    if s != d:
                        tasks[name] = (s, d)
                    else:
                        self.net_paths[name] = [s]
        return tasks

    # ---------------- per-iteration device steps
    def _update_present_ema(self, first):
        idx, val = self.accounting.state.update_ema(first, f32(0.40), f32(1.0 - 0.40))
        return idx, val

    def _edge_layers(self, idx):
        if len(idx) == 0:
            return np.zeros(0, np.int64)
        torch = self.ctx.torch
        t = torch.from_numpy(np.asarray(idx, dtype=np.int64)).to(self.ctx.device)
        return self.graph.edge_layer[t].cpu().numpy().astype(np.int64)

    def _compute_layer_bias(self, ema_idx, ema_val, alpha=0.88, max_boost=1.80):
        """UPF:4825-4878 (bincount in float64 over the nonzeros, same order)."""
        L = self._Nz
        layers = self._edge_layers(ema_idx)
        per = np.bincount(layers, weights=ema_val.astype(np.float64), minlength=L)
        maxv = float(per.max()) if len(per) else 0.0
        if maxv <= 1e-9:
            raw = np.ones(L, dtype=np.float32)
        else:
            raw = 1.0 + (1.5 if L <= 2 else 0.75) * (per / maxv)
        if not hasattr(self, "_layer_bias_ema"):
            self._layer_bias_ema = raw.astype(np.float32)
        else:
            self._layer_bias_ema = (alpha * self._layer_bias_ema + (1.0 - alpha) * raw).astype(np.float32)
        self._layer_bias_ema = np.clip(self._layer_bias_ema, 1.0, max_boost)
        return self._layer_bias_ema

    def _via_discount_table(self):
        """Leave-hot-layer factor of the via pooling penalty (UPF:3356-3367)
        per (z_lo, z_hi), in the dtype self.layer_bias has at this point."""
        k = float(getattr(self.config, "via_hot_layer_discount", 0.20))
        b = self.layer_bias
        L = self._Nz
        zl = np.repeat(np.arange(L), L)
        zh = np.tile(np.arange(L), L)
        d = (1.0 - k * np.maximum(b[zl], 0.0)) * (1.0 + 0.5 * k * np.maximum(b[zh], 0.0))
        return d.astype(np.float64), b.dtype == np.float64

    def _update_costs(self, pres_fac, hist_weight, vcm, bcw, layer_bias, it, with_vias=True):
        torch = self.ctx.torch
        dev = self.ctx.device
        keep = []

        def dptr(a):
            t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
            keep.append(t)
            return ctypes.c_void_p(t.data_ptr())
        p = _abi.CostParams()
        p.X, p.Y, p.L = self._Nx, self._Ny, self._Nz
        p.via_cost_mult, p.base_cost_weight = f32(vcm), f32(bcw)
        p.pres_fac, p.hist_weight = f32(pres_fac), f32(hist_weight)
        p.jitter_scale, p.add_jitter = f32(1e-6), 1
        p.layer_bias = dptr(np.asarray(layer_bias, dtype=np.float32))
        if with_vias:
            p.via_pooling = 1
            p.seg_z = self._segZ
            vd = self._via_dev
            p.col_pres = ctypes.c_void_p(vd["col_pres"].data_ptr())
            if self._segZ > 0:
                p.seg_prefix = ctypes.c_void_p(vd["seg_prefix"].data_ptr())
            p.col_weight = f32(float(getattr(self.config, "via_column_weight", 1.0)))
            p.seg_weight = f32(float(getattr(self.config, "via_segment_weight", 1.0)))
            disc, is64 = self._via_discount_table()
            p.via_discount = dptr(disc)
            p.discount_f64 = int(is64)
            if it > 1:
                p.hard_block = 1
                p.col_use = ctypes.c_void_p(vd["col_use"].data_ptr())
                if self._segZ > 0:
                    p.seg_use = ctypes.c_void_p(vd["seg_use"].data_ptr())
                p.col_cap, p.seg_cap = 4, 2
        self.accounting.state.update_costs(p)
        self.ctx.sync()
        # the iteration's cost lower bounds for the goal-directed search, and
        # its typical planar cost (bucket width scale)
        hx, hz, med = cost_bounds(self.ctx, self.graph.dlat, self.accounting.total_cost, with_scale=True)
        self._hbounds = (hx, hz)
        self._cost_scale = med

    # ---------------- committed-net arrays
    def _flat(self):
        """Flat arrays over the nets that hold a path, in net_paths order:
        per edge (edge id, src node, dst node, net id) and per via hop
        (x, y, z_lo, z_hi, net id).  The result is kept until a committed
        path changes: the cache holds the (name, path, edges) objects it was
        built from and is reused only when every one is the same object."""
        items = []
        trees = getattr(self, "_tree_branches", {})
        n2n, n2e = self._net_nodes, self._net_to_edges
        for name, path in self.net_paths.items():
            if len(path) < 2:
                continue
            tb = trees.get(name)
            if tb is not None:
                # a tree: each branch is a path of the net
                items.extend((name, b, b, e) for b, e in zip(*tb) if len(b) > 1)
                continue
            items.append((name, path, n2n.get(name), n2e.get(name)))
        cache = getattr(self, "_flat_cache", None)
        if cache is not None and len(cache[0]) == len(items) and all(
                a[1] is b[1] and a[2] is b[2] and a[3] is b[3] and a[0] == b[0] for a, b in zip(cache[0], items)):
            return cache[1]
        nodes_l, edges_l, ids = [], [], []
        for name, path, nodes, e in items:
            if nodes is None:
                nodes = np.asarray(path, dtype=np.int64)
            if e is None or len(e) != len(nodes) - 1:
                e = np.full(len(nodes) - 1, -1, np.int64)
            nodes_l.append(nodes); edges_l.append(e); ids.append(self._get_net_id(name))
        z0 = np.zeros(0, np.int64)
        if not nodes_l:
            out = (z0, z0, z0, z0, np.zeros((0, 4), np.int64), z0)
        else:
            # node and edge ids keep the engine's int32 (int64 when a path
            # held int64); every value is below N or E
            lens = np.fromiter((len(n) for n in nodes_l), dtype=np.int64, count=len(nodes_l))
            alln = np.concatenate(nodes_l)
            inner = np.ones(len(alln) - 1, dtype=bool)
            inner[np.cumsum(lens)[:-1] - 1] = False      # pairs that straddle two paths
            S = alln[:-1][inner]
            D = alln[1:][inner]
            E = np.concatenate(edges_l)
            NS = np.repeat(np.asarray(ids, np.int64), lens - 1)
            H, NH = self._hops_of(S, D, NS)
            if len(E) == 0 or int(E.min()) >= 0:
                out = (E, S, D, NS, H, NH)
            else:
                ok = E >= 0
                out = (E[ok], S[ok], D[ok], NS[ok], H, NH)
        self._flat_cache = (items, out)
        return out

    def _hops_of(self, S, D, NS):
        """Via hops (x, y, z_lo, z_hi) among node pairs S->D, with their nets.
        S->D are lattice edges: a planar step moves the index by 1 or X, a
        via by a nonzero multiple of X*Y, so |D - S| >= X*Y marks the vias."""
        plane = self._Nx * self._Ny
        hop = np.abs(D - S) >= plane
        if not hop.any():
            return np.zeros((0, 4), np.int64), NS[hop]
        s, d = S[hop].astype(np.int64), D[hop].astype(np.int64)
        zs, r = np.divmod(s, plane)
        zd = d // plane
        H = np.stack([r % self._Nx, r // self._Nx, np.minimum(zs, zd), np.maximum(zs, zd)], axis=1)
        return H, NS[hop]

    # ---------------- via usage / ownership (UPF:2678-2904)
    def _rebuild_via_usage_from_committed(self, flat):
        """Column/segment via usage, keepout owners (first owner wins) and the
        node ownership map from every committed path."""
        _, _, _, _, H, NH = flat
        X, Y, L = self._Nx, self._Ny, self._Nz
        plane = X * Y
        self.via_col_use.fill(0)
        if self._segZ > 0:
            self.via_seg_use.fill(0)
        if len(H):
            x, y, zlo, zhi = H[:, 0], H[:, 1], H[:, 2], H[:, 3]
            np.add.at(self.via_col_use, (x, y), 1)
            if self._segZ > 0:
                a = np.clip(zlo, 1, L - 2)
                b = np.clip(zhi, 1, L - 2)
                n = np.maximum(b - a, 0)
                if n.sum():
                    rep = np.repeat(np.arange(len(n)), n)
                    off = np.arange(int(n.sum())) - np.repeat(np.cumsum(n) - n, n)
                    np.add.at(self.via_seg_use, (x[rep], y[rep], (a[rep] - 1) + off), 1)
            # barrel nodes (z_lo..z_hi inclusive), in committed order
            span = zhi - zlo + 1
            rep = np.repeat(np.arange(len(span)), span)
            zz = np.repeat(zlo, span) + (np.arange(int(span.sum())) - np.repeat(np.cumsum(span) - span, span))
            nodes = zz * plane + y[rep] * X + x[rep]
            owner = NH[rep]
            # keepouts: first committed owner of each (z, x, y)
            u, first = np.unique(nodes, return_index=True)
            self._ko_nodes, self._ko_owner = u, owner[first]
            # node ownership: last committed owner of each barrel node wins;
            # every keepout cell is a barrel node, so the reference's
            # keepout fill of unowned cells (UPF:2890-2904) changes nothing
            ur, last_rev = np.unique(nodes[::-1], return_index=True)
            self.node_owner.fill(-1)
            self.node_owner[ur] = owner[::-1][last_rev]
        else:
            self.node_owner.fill(-1)
            self._ko_nodes = np.zeros(0, np.int64)
            self._ko_owner = np.zeros(0, np.int64)

    def _via_pressure(self):
        """Via pooling pressure of the iteration (UPF:3626-3637) on the
        device: the usage just recounted on the host goes to HBM, where
        orgpu_via_pressure updates the column and segment pressure EMAs and
        the segment prefix the cost update reads, and sums the via overuse."""
        torch = self.ctx.torch
        vd = self._via_dev
        with torch.cuda.stream(self.ctx.stream):
            vd["col_use"].copy_(torch.from_numpy(np.ascontiguousarray(self.via_col_use).reshape(-1)))
            if self._segZ > 0:
                vd["seg_use"].copy_(torch.from_numpy(np.ascontiguousarray(self.via_seg_use).reshape(-1)))
        seg_cap = int(self.via_seg_cap.flat[0]) if self._segZ > 0 else 0
        _abi.check(self.ctx.lib.orgpu_via_pressure(
            self.ctx.handle, self._Nx * self._Ny, self._segZ, vd["col_use"].data_ptr(), int(self.via_col_cap.flat[0]),
            vd["seg_use"].data_ptr(), seg_cap, vd["col_pres"].data_ptr(), vd["seg_pres"].data_ptr(),
            vd["seg_prefix"].data_ptr(), vd["over"].data_ptr()))

    def _mark_barrels_many(self, paths, nids):
        """_mark_barrels for several paths in order (a later path's barrel
        overwrites an earlier one's owner, as the per-path calls would)."""
        if not paths:
            return
        lens = np.fromiter((len(p) for p in paths), dtype=np.int64, count=len(paths))
        # node ids stay below 2^25, so their differences fit the paths' int32
        allp = np.concatenate(paths)
        ends = np.cumsum(lens)
        # via hops among the consecutive pairs inside each path (a planar
        # step moves the index by 1 or X, a via by a multiple of X*Y; the
        # pairs across two paths are dropped), with the net of each
        hop = np.abs(allp[1:] - allp[:-1]) >= self._Nx * self._Ny
        hop[ends[:-1] - 1] = False
        i = np.nonzero(hop)[0]
        NS = np.asarray(nids, np.int64)[np.searchsorted(ends, i, side="right")]
        H, NH = self._hops_of(allp[i], allp[i + 1], NS)
        if len(H) == 0:
            return
        X, plane = self._Nx, self._Nx * self._Ny
        span = H[:, 3] - H[:, 2] + 1
        rep = np.repeat(np.arange(len(span)), span)
        zz = np.repeat(H[:, 2], span) + (np.arange(int(span.sum())) - np.repeat(np.cumsum(span) - span, span))
        nodes = zz * plane + H[rep, 1] * X + H[rep, 0]
        owner = NH[rep]
        ur, last_rev = np.unique(nodes[::-1], return_index=True)
        self.node_owner[ur] = owner[::-1][last_rev]

    def _mark_barrels(self, path, nid):
        X, plane = self._Nx, self._Nx * self._Ny
        for x, y, zlo, zhi in self._via_hops(path):
            zs = np.arange(int(zlo), int(zhi) + 1)
            self.node_owner[zs * plane + int(y) * X + int(x)] = nid

    def _keepout_owner_array(self):
        own = np.full(self.N, -1, dtype=np.int32)
        own[self._ko_nodes] = self._ko_owner
        return own

    def _detect_barrel_conflicts(self, flat):
        """UPF:5478-5569: committed edges touching a node owned by another net."""
        E, S, D, NE, _, _ = flat
        if len(E) == 0:
            return np.zeros(0, np.int64), 0
        so = self.node_owner[S]
        do = self.node_owner[D]
        m = ((so != -1) & (so != NE)) | ((do != -1) & (do != NE))
        ce = E[m]
        return ce, int(len(ce))

    # ---------------- bookkeeping (UPF:4989-5007)
    def _update_net_edge_tracking(self, nid, edges, path=None):
        self._net_to_edges.pop(nid, None)
        # int32 arrays from the engine are kept as they are (every consumer
        # indexes or concatenates them); other inputs become int64
        self._net_to_edges[nid] = edges if isinstance(edges, np.ndarray) and edges.dtype == np.int32 \
            else np.asarray(edges, dtype=np.int64)
        if path is not None:
            self._net_nodes[nid] = path if isinstance(path, np.ndarray) and path.dtype == np.int32 \
                else np.asarray(path, dtype=np.int64)

    def _clear_net_edge_tracking(self, nid):
        self._net_to_edges.pop(nid, None)
        self._net_nodes.pop(nid, None)
        getattr(self, "_tree_branches", {}).pop(nid, None)

    # ---------------- ordering / hotset / rip (UPF:4401-4439, 5009-5140, 5300-5335)
    def _dense(self, idx, val, slot=None):
        """val scattered into a dense float32[E] (zero elsewhere).  With a
        slot name the array is that slot's persistent buffer: only the
        entries its previous contents set are cleared, instead of allocating
        and zero-filling E floats on every call; the previous contents of the
        slot are gone."""
        if slot is None:
            d = np.zeros(self.accounting.E, dtype=np.float32)
            d[idx] = val
            return d
        bufs = self.__dict__.setdefault("_dense_bufs", {})
        ent = bufs.get(slot)
        if ent is None or len(ent[0]) != self.accounting.E:
            ent = [np.zeros(self.accounting.E, dtype=np.float32), np.zeros(0, np.int64)]
            bufs[slot] = ent
        d, prev = ent
        d[prev] = 0.0
        idx = np.asarray(idx, dtype=np.int64)
        d[idx] = val
        ent[1] = idx
        return d

    @staticmethod
    def _path_sums(segs, dense):
        """Per edge list, Python's sum() of dense[e] over its edges in path
        order (sequential float64), for non-negative values: adding +0.0
        leaves such a sum unchanged, so only the nonzero values are added,
        every list's j-th one at step j."""
        out = np.zeros(len(segs), dtype=np.float64)
        if not segs:
            return out
        lens = np.fromiter((len(e) for e in segs), dtype=np.int64, count=len(segs))
        flat = np.concatenate(segs).astype(np.int64) if lens.sum() else np.zeros(0, np.int64)
        vals = dense[flat]
        nzm = vals != 0
        if not nzm.any():
            return out
        v = vals[nzm].astype(np.float64)
        sg = np.repeat(np.arange(len(segs)), lens)[nzm]
        cnt = np.bincount(sg, minlength=len(segs))
        pos = np.arange(len(v)) - (np.cumsum(cnt) - cnt)[sg]
        # lists with about as many nonzeros share a zero-padded matrix whose
        # rows are summed left to right by cumsum (an accumulate is
        # sequential; the padding adds +0.0)
        cls = np.zeros(len(segs), dtype=np.int64)
        has = cnt > 0
        cls[has] = np.ceil(np.log2(cnt[has])).astype(np.int64)
        by = np.argsort(cls[sg], kind="stable")          # nonzeros grouped by class, list order kept
        kk = cls[sg][by]
        ks = np.unique(kk)
        bounds = np.searchsorted(kk, np.append(ks, ks[-1] + 1))
        for j, k in enumerate(ks.tolist()):
            idx = by[bounds[j]:bounds[j + 1]]
            members, row = np.unique(sg[idx], return_inverse=True)
            m = np.zeros((len(members), 1 << k), dtype=np.float64)
            m[row, pos[idx]] = v[idx]
            out[members] = np.cumsum(m, axis=1)[:, -1]
        return out

    def _order_nets_by_difficulty(self, tasks, over_dense):
        """UPF:4401-4439: score = manhattan*3*(1 + mean overuse on the net's
        previous path), sorted descending with the net name as tie-break.
        Vectorised; the congestion term is the sequential float64 sum the
        reference's sum() takes, and it is only evaluated for nets whose
        previous path crosses an overused edge (it is 0.0 otherwise)."""
        names = list(tasks)
        if not names:
            return []
        sd = np.asarray([tasks[n] for n in names], dtype=np.int64)
        X, plane = self._Nx, self._Nx * self._Ny
        zs, rs = np.divmod(sd[:, 0], plane)
        zd, rd = np.divmod(sd[:, 1], plane)
        dist = np.abs(rd % X - rs % X) + np.abs(rd // X - rs // X) + np.abs(zd - zs)
        cong = np.zeros(len(names), dtype=np.float64)
        with_edges = [i for i, n in enumerate(names) if n in self._net_to_edges]
        if with_edges and over_dense is not None and len(over_dense):
            # the reference's sum() over the net's previous path, in path order
            segs = [self._net_to_edges[names[i]] for i in with_edges]
            lens = np.fromiter((len(e) for e in segs), dtype=np.int64, count=len(segs))
            acc = self._path_sums(segs, over_dense)
            hitn = acc > 0
            cong[np.asarray(with_edges)[hitn]] = acc[hitn] / np.maximum(1, lens[hitn])
        score = (dist * 3).astype(np.float64) * (cong + 1.0)
        # descending (score, name): names are unique, so this is the order
        # of the reference's reverse tuple sort
        by_name = sorted(range(len(names)), key=names.__getitem__)
        name_rank = np.empty(len(names), dtype=np.int64)
        name_rank[by_name] = np.arange(len(names))
        order = np.lexsort((name_rank, score))[::-1]
        rot = self._rng.randint(0, min(5, len(names) // 10))
        ordered = [names[i] for i in order.tolist()]
        if rot > 0:
            ordered = ordered[rot:] + ordered[:rot]
        return ordered

    def _build_hotset(self, tasks, hot_idx, hot_over, flat, ripped=None):
        ripped = ripped or set()
        if not hasattr(self, "_net_clean_iters"):
            self._net_clean_iters = {}
        if len(hot_idx) == 0:
            unrouted = {n for n in tasks if not len(self.net_paths.get(n, ()))}
            return unrouted | ripped
        E, _, _, NE, _, _ = flat
        # hot edges: a 1 marker in a slot buffer (an edge can be hot with
        # zero overuse)
        hot_mask = self._dense(hot_idx, 1.0, slot="hot_mask")
        id_to_name = {v: k for k, v in self.net_id_map.items()}
        offenders = {id_to_name[int(i)] for i in np.unique(NE[hot_mask[E] != 0])} if len(E) else set()
        for n in tasks:
            if n in offenders:
                self._net_clean_iters[n] = 0
            else:
                self._net_clean_iters[n] = self._net_clean_iters.get(n, 0) + 1
        thr = int(getattr(self.config, "freeze_after_clean", 3))
        offenders -= {n for n in offenders if self._net_clean_iters.get(n, 0) >= thr}
        offenders |= ripped
        unrouted = {n for n in tasks if not len(self.net_paths.get(n, ()))}
        scored = [n for n in offenders if n in self._net_to_edges]
        # impact: the sum of the hot edges' overuse along the net's path
        # (UPF:5071), on the device for every offender at once
        sums = edge_list_sums(self.ctx, self.accounting.E, [self._net_to_edges[n] for n in scored],
                              hot_idx, hot_over)
        scores = list(zip(sums.tolist(), scored))
        for n in unrouted:
            if n not in offenders:
                scores.append((0.0, n))
        scores.sort(reverse=True)
        cap = min(self.config.hotset_cap, 100)
        rng = random.Random(42 + self.iteration)
        k_top = int(0.6 * cap)
        top = [n for _, n in scores[:k_top]]
        rest = [n for _, n in scores[k_top:]]
        k_rand = min(cap - k_top, len(rest))
        sample = rng.sample(rest, k_rand) if rest else []
        if not hasattr(self, "_last_reroute_iter"):
            self._last_reroute_iter = {}
        cool = [n for n in top + sample if self.iteration - self._last_reroute_iter.get(n, -999) > 1]
        for n in cool:
            self._last_reroute_iter[n] = self.iteration
        return set(cool)

    def _rip_top_k_offenders(self, o_idx, o_val, k=20):
        cand = [n for n, path in self.net_paths.items()
                if len(path) and n not in self.locked_nets and n in self._net_to_edges]
        imps = edge_list_sums(self.ctx, self.accounting.E, [self._net_to_edges[n] for n in cand], o_idx, o_val)
        scores = [(v, n) for v, n in zip(imps.tolist(), cand) if v > 0]
        scores.sort(reverse=True)
        victims = {n for _, n in scores[:k]}
        for n in victims:
            if self.net_paths.get(n) and n in self._net_to_edges:
                self.accounting.clear_path(self._net_to_edges[n])
                self.net_paths[n] = []
                self._clear_net_edge_tracking(n)
                self.net_clean_streak[n] = 0
        return victims

    # ---------------- batched search of one iteration (UPF:4441-4823)
    def _engine_for(self, n_searches):
        """Search engine of a batch of n two-pin searches: the one-wave engine
        fills the GPU with many searches (a workgroup each); a few searches
        run on the bucket-round engine, which spreads each over every CU,
        unless the board has at most sssp_solo_nodes nodes, the negotiation
        at least sssp_solo_density nets per node (a crowded board, whose late
        searches are deep; on lightly loaded boards the bucket engine's few
        rounds are the faster way), it has no multi-terminal trees, the
        board's rows fit the one-wave engine's adjacency records, and a
        small batch has moved them to the one-wave engine
        (_small_engine_record; _search_batch also predicts the rounds of each
        such batch, _small_engine_predict).  The plain form (one slot) stays
        on the bucket-round engine."""
        e = self.config.sssp_engine
        if e == "auto":
            slots = int(self.config.batch_slots)
            n = min(int(n_searches), slots)
            if n >= int(self.config.sssp_solo_min):
                return 1
            if slots > 1 and self.N <= int(getattr(self.config, "sssp_solo_nodes", 0) or 0):
                return int(self.__dict__.get("_small_engine", 0))
            return 0
        return int(e or 0)

    def _tuner_for(self, engine):
        """The bucket-target tuner of this negotiation's batches on `engine`
        (each engine has its own: their best widths differ), or None outside
        a negotiation."""
        key = self.__dict__.get("_tuner_key")
        if key is None:
            return None
        t = self._tuners.get(int(engine))
        if t is None:
            t = self._tuners[int(engine)] = BucketTuner(f"{key}:{int(engine)}", self.N)
        return t

    def _rows_fit_records(self):
        """True when at most one node in eight has more than six edges, so
        that the one-wave engine reads adjacency records rather than CSR rows
        (the reference via rules give most nodes of a 6-layer board 7-10
        edges: cfg1 in canonical order took 1.37 s on the one-wave engine's
        CSR path against 0.95 s on the bucket engine)."""
        v = self.__dict__.get("_rows_fit")
        if v is None:
            ip = self.graph.dlat.indptr
            v = bool(int(((ip[1:] - ip[:-1]) > 6).sum().item()) * 8 <= self.N)
            self._rows_fit = v
        return v

    def _small_solo_time(self):
        """A conservative one-wave engine time for a batch of searches that
        each cover the board (about N/256 steps of a four-wave workgroup at
        about 20 us a step), and at least 1/4 s: below that the bucket
        engine's batches are cheap whatever the board (lightly loaded boards,
        short searches, where it is the faster engine)."""
        return max((self.N / 256.0) * 20e-6, 0.25)

    def _small_engine_record(self, engine, seconds):
        """Moves a negotiation's small batches (fewer than sssp_solo_min
        searches, several slots, a board of at most sssp_solo_nodes nodes) to
        the one-wave engine for good when a bucket-engine batch took more than
        twice _small_solo_time: its rounds were the deep, launch-bound ones
        of a congested board (a 100,820-node board with 2,520 nets: 40
        iterations in 5.8 s on the one-wave engine against 46.6 s; single
        bucket-engine batches there took up to 30 s).  Paths are the same on
        either engine."""
        if engine == 0 and seconds > 2.0 * self._small_solo_time() and not self.__dict__.get("_small_engine"):
            self._small_engine = 1
            self.stats["small_engine_switch_measured"] += 1

    def _small_engine_predict(self, depth_hint, slots):
        """The engine of a small batch before it runs: the one-wave engine when
        the bucket engine's rounds for the batch's depth estimate (the deepest
        search spans depth / bucket width buckets, each a round of about
        0.1 ms) would take more than twice _small_solo_time."""
        if depth_hint is None:
            return 0
        width = self._search_params(slots, depth_hint, engine=0).delta
        return 1 if (float(depth_hint) / max(width, 1e-9)) * 1e-4 > 2.0 * self._small_solo_time() else 0

    def _search_params(self, slots, depth=None, engine=0):
        """depth: the batch's estimated search depth in cost units beyond
        each source's lower bound (None = unknown); with no fixed delta
        configured, the bucket is widened so that this depth spans about
        config.sssp_target_buckets buckets.  engine: the search engine the
        batch runs on (its base bucket width differs)."""
        d = self.config.sssp_delta
        nb = self.config.sssp_buckets
        if d is None or nb is None:
            d2, nb2 = base_width(self, getattr(self, "_cost_scale", 0.0), engine), 8
            target = self.config.sssp_target_buckets
            if target is None:
                # a negotiation's tuner (one per engine) picks the target of
                # batches with a depth estimate; otherwise the starting default
                tuner = self._tuner_for(engine)
                if tuner is not None and depth is not None:
                    target = tuner.target(ratio=float(depth) / d2)
                else:
                    target = default_target(self.N)
            if depth is not None and target:
                d2 = max(d2, float(depth) / float(target))
            d = d if d is not None else d2
            nb = nb if nb is not None else nb2
        # goal-directed order (same paths) unless the Dijkstra-minimal work is
        # being counted, which needs the plain order's settled set
        hx, hz = (0.0, 0.0)
        if self.config.goal_directed and not self.config.count_minimal:
            hx, hz = getattr(self, "_hbounds", (0.0, 0.0))
        return SearchParams(slots=slots, delta=float(d), n_buckets=int(nb),
                            rounds_per_poll=int(self.config.sssp_rounds_per_poll or 32),
                            grid=int(self.config.sssp_grid or 0), count_minimal=bool(self.config.count_minimal),
                            h_xy=hx, h_z=hz, local_batches=int(self.config.sssp_local_batches or 0),
                            graph_rounds=bool(getattr(self.config, "sssp_graph_rounds", False)),
                            engine=int(engine), list_cap=int(getattr(self.config, "sssp_list_cap", 0) or 0),
                            lookahead=int(getattr(self.config, "sssp_lookahead", 0) or 0),
                            waves=int(getattr(self.config, "sssp_solo_waves", 0) or 0))

    def _depth_hint(self, ordered, sd):
        """How far past their first bucket the searches of a batch of nets
        reach: the 90th percentile, over the nets that hold a path, of that
        path's cost on this iteration's costs minus the goal-directed lower
        bound at the net's source.  The current path is a feasible path, so
        its cost bounds the search's final distance from above.  None when no
        net holds a finite-cost path (e.g. iteration 1).  Paths are
        identical for any bucket width; this only sets the width."""
        n2e = self._net_to_edges
        have = [i for i, nid in enumerate(ordered) if nid in n2e]
        if not have:
            return None
        edges = [n2e[ordered[i]] for i in have]
        lens = np.fromiter((len(e) for e in edges), dtype=np.int64, count=len(edges))
        if not lens.sum():
            return None
        torch = self.ctx.torch
        flat = torch.from_numpy(np.concatenate(edges).astype(np.int64)).to(self.ctx.device)
        c = self.accounting.total_cost[flat].double().cpu().numpy()
        starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
        ok = lens > 0
        cost = np.add.reduceat(c, starts[ok]) if ok.any() else np.zeros(0)
        hx, hz = getattr(self, "_hbounds", (0.0, 0.0))
        X, plane = self._Nx, self._Nx * self._Ny
        s, t = sd[np.asarray(have)[ok], 0], sd[np.asarray(have)[ok], 1]
        rs, rt = s % plane, t % plane
        h = hx * (np.abs(rt % X - rs % X) + np.abs(rt // X - rs // X)) + hz * np.abs(t // plane - s // plane)
        gap = cost - h
        gap = gap[np.isfinite(gap)]
        return float(np.quantile(gap, 0.9)) if len(gap) else None

    def _search_batch(self, reqs, keepout_owner=None, readd=True, depth_hint=None):
        """reqs: per request (src, dst, net_int, roi, stagnation_bonus), as a
        list of tuples or an (n, 5) array, optionally followed by (roi_src,
        roi_dst, keep0, keep1, add0, add1) for a portal net (n, 11): its ROI
        grows from roi_src/roi_dst (the portals' entry-layer nodes) and
        keeps keep0/keep1 (its seeds, -1 none), and with readd add0/add1 (the
        task's nodes; -1 none) are appended after the other re-added nodes
        when missing (UPF:4619-4700).  roi False = whole graph.  A request
        whose src or dst is not in its ROI finds no path.
        depth_hint: the batch's search depth estimate (see _depth_hint).  Returns per
        request (nodes, edges) as int32 arrays, or None."""
        torch = self.ctx.torch
        dlat = self.graph.dlat
        R = np.asarray(reqs, dtype=np.float64)
        R = R.reshape(len(R), -1) if R.size else np.zeros((0, 5))
        n = len(R)
        out = [None] * n
        src, dst = R[:, 0].astype(np.int64), R[:, 1].astype(np.int64)
        nid, use_roi, bonus = R[:, 2].astype(np.int64), R[:, 3] != 0, R[:, 4]
        if R.shape[1] >= 11:
            rsrc, rdst = R[:, 5].astype(np.int64), R[:, 6].astype(np.int64)
            keeps, adds = R[:, 7:9].astype(np.int64), R[:, 9:11].astype(np.int64)
        else:
            rsrc, rdst, keeps, adds = src, dst, None, None
        roi_at = np.nonzero(use_roi)[0]
        row_of = np.full(n, -1, dtype=np.int64)
        row_of[roi_at] = np.arange(len(roi_at))
        rows = None
        keep = np.ones(n, dtype=bool)
        ref_order = self.config.sssp_order == "reference"
        info = None
        added = {}
        t0 = time.perf_counter()
        if len(roi_at):
            words = (self.N + 31) // 32
            rows = torch.empty((len(roi_at), words), dtype=torch.int32, device=self.ctx.device)
            depth = self.roi_extractor.roi_depth_many(rsrc[roi_at], rdst[roi_at], bonus[roi_at])
            rq = np.stack([rsrc[roi_at], rdst[roi_at], depth, np.full(len(roi_at), MAX_ROI_NODES, np.int64),
                           np.arange(len(roi_at))], axis=1)
            kp = keeps[roi_at] if keeps is not None and (keeps[roi_at] >= 0).any() else None
            info = roi_bfs(self.ctx, dlat, rq, rows, readd=readd, raw=True, keeps=kp)
            if readd and adds is not None:
                added = self._append_roi_nodes(rows, roi_at, adds[roi_at], rsrc[roi_at], rdst[roi_at])
            same = (src[roi_at] == rsrc[roi_at]) & (dst[roi_at] == rdst[roi_at])
            ok = (info["src_in"] != 0) & (info["dst_in"] != 0)
            if not same.all():
                at = np.nonzero(~same)[0]
                ok[at] = self._roi_has(rows, at, src[roi_at[at]]) & self._roi_has(rows, at, dst[roi_at[at]])
            keep[roi_at[~ok]] = False
        self.stats["roi_s"] += time.perf_counter() - t0
        todo = np.nonzero(keep)[0]
        if not len(todo):
            return out
        eng = 0 if ref_order else self._engine_for(len(todo))
        # (not with multi-terminal trees in the negotiation: there the
        # two-pin keepout searches around the trees measured slower on the
        # one-wave engine, cfg3 with its 200 trees 1.42 against 1.15 s)
        small = (not ref_order and not getattr(self, "_tree_pins", None) and self._rows_fit_records()
                 and self.config.sssp_engine == "auto"
                 and int(self.config.batch_slots) > 1
                 and min(len(todo), int(self.config.batch_slots)) < int(self.config.sssp_solo_min)
                 and self.N <= int(getattr(self.config, "sssp_solo_nodes", 0) or 0)
                 and self.__dict__.get("_negotiation_nets", 0)
                 >= float(getattr(self.config, "sssp_solo_density", 0.0)) * self.N)
        if small and eng == 0:
            eng = self._small_engine_predict(depth_hint, max(1, min(int(self.config.batch_slots), len(todo))))
            if eng == 1:
                # the negotiation's later small batches (the wider-ROI
                # fallback's included, which have no depth estimate) follow
                self._small_engine = 1
                self.stats["small_engine_switch_predicted"] += 1
        slots = max(1, min(int(self.config.batch_slots), len(todo), max_slots(self.N, engine=eng)))
        params = self._search_params(slots, depth_hint, engine=eng)
        sreqs = np.stack([src[todo], dst[todo], nid[todo], row_of[todo],
                          np.full(len(todo), int(keepout_owner is not None), np.int64)], axis=1)
        ko = None
        if keepout_owner is not None:
            ko = torch.from_numpy(keepout_owner).to(self.ctx.device)
        t0 = time.perf_counter()
        if ref_order:
            # the reference's pop order needs each ROI's list order: the
            # ascending node order, with the exceptions its ROI record names
            ranks = []
            for i in todo.tolist():
                if row_of[i] < 0:
                    ranks.append(None)
                    continue
                ex = rank_exceptions(info[row_of[i]], self.N)
                na = int((info["appended"][row_of[i]] >= 0).sum())
                ex += [(v, 2 * self.N + 2 + na + k) for k, v in enumerate(added.get(int(row_of[i]), ()))]
                ranks.append(ex)
            nodes, edges, offs, _, rst = device_route_ref(self.ctx, dlat, self.accounting.total_cost, sreqs[:, :5],
                                                          roi_bitmaps=rows, keepout_owner=ko, ranks=ranks,
                                                          slots=int(self.config.batch_slots))
            self.stats["sssp_s"] += time.perf_counter() - t0
            self.stats["sssp_kernel_ms"] += rst["kernel_ms"]
            self.stats["sssp_edges"] += rst["edges"]
            self.stats["sssp_items"] += rst["pops"]
            self.stats["sssp_searches"] += len(sreqs)
            ol = offs.tolist()
            for j, i in enumerate(todo.tolist()):
                a, b = ol[j], ol[j + 1]
                if b - a > 1:
                    out[i] = (nodes[a:b], edges[a:b - 1])
            return out
        nodes, edges, offs, _, st = device_route_flat(self.ctx, dlat, self.accounting.total_cost, sreqs, params,
                                                      roi_bitmaps=rows, keepout_owner=ko)
        dt = time.perf_counter() - t0
        self.stats["sssp_s"] += dt
        tuner = self._tuner_for(params.engine)
        if (tuner is not None and depth_hint is not None and self.config.sssp_target_buckets is None
                and self.config.sssp_delta is None):
            tuner.record(dt)
        if small:
            self._small_engine_record(params.engine, dt)
        self.stats["sssp_kernel_ms"] += st["kernel_ms"]
        self.stats["sssp_rounds"] += st["rounds"]
        self.stats["sssp_edges"] += st["edges"]
        self.stats["sssp_items"] += st["items"]
        self.stats["sssp_launches"] += st["launches"]
        self.stats["sssp_round_ms"] += st["round_ms"]
        self.stats["sssp_minimal_nodes"] += st["minimal_nodes"]
        self.stats["sssp_minimal_edges"] += st["minimal_edges"]
        self.stats["sssp_retried"] += st["retried"]
        self.stats["sssp_searches"] += len(sreqs)
        if params.engine == 1:
            # the one-wave engine's own share (searches it handed back to the
            # bucket engine included: st["retried"])
            self.stats["solo_kernel_ms"] += st["round_ms"]
            self.stats["solo_items"] += st["items"]
            self.stats["solo_edges"] += st["edges"]
            self.stats["solo_launches"] += st["launches"]
        t0 = time.perf_counter()
        ol = offs.tolist()
        for j, i in enumerate(todo.tolist()):
            a, b = ol[j], ol[j + 1]
            if b - a > 1:
                out[i] = (nodes[a:b], edges[a:b - 1])
        self.stats["search_out_s"] += time.perf_counter() - t0
        return out

    def _route_trees(self, names):
        """Grows the trees of the tree nets `names` on this iteration's costs
        (orthoroute_amd.steiner.grow_trees: every net still growing takes one
        multi-source step per engine call, whole graph, no keepouts).
        Returns {name: (branches, branch edge arrays, tree nodes) or None when
        a pin stays unreached}."""
        if not names:
            return {}
        slots = max(1, min(int(self.config.batch_slots), len(names), max_slots(self.N)))
        t0 = time.perf_counter()
        grown, st = grow_trees(self.ctx, self.graph.dlat, self.accounting.total_cost,
                               [self._tree_pins[n] for n in names], self._search_params(slots))
        self.stats["tree_s"] += time.perf_counter() - t0
        self.stats["sssp_searches"] += st["searches"]
        self.stats["sssp_edges"] += st["edges"]
        self.stats["sssp_items"] += st["items"]
        out = {}
        for n, t in zip(names, grown):
            if t["unreached"]:
                out[n] = None
                continue
            branches = [np.asarray(b, dtype=np.int32) for b in t["branches"]]
            bedges, off = [], 0
            for b in branches:
                k = len(b) - 1
                bedges.append(np.asarray(t["edges"][off:off + k], dtype=np.int64))
                off += k
            out[n] = (branches, bedges, np.asarray(t["nodes"], dtype=np.int32))
        return out

    def _roi_has(self, rows, row_idx, nodes):
        """Membership bits of nodes[k] in ROI row row_idx[k] (bool array)."""
        torch = self.ctx.torch
        nodes = np.asarray(nodes, dtype=np.int64)
        r = torch.from_numpy(np.asarray(row_idx, dtype=np.int64)).to(self.ctx.device)
        w = torch.from_numpy(nodes >> 5).to(self.ctx.device)
        words = rows[r, w].cpu().numpy().view(np.uint32).astype(np.int64)
        return ((words >> (nodes & 31)) & 1).astype(bool)

    def _append_roi_nodes(self, rows, roi_at, adds, rsrc, rdst):
        """A portal net's task nodes that its ROI lacks, appended after the
        ROI's own re-added nodes (UPF:4683-4696: the reference appends src,
        then dst, when missing).  They differ from the ROI's endpoints only
        after a portal retarget, so this runs for few requests.  Returns
        {row: [appended nodes in order]}."""
        out = {}
        cand = np.nonzero(((adds[:, 0] >= 0) & (adds[:, 0] != rsrc)) | ((adds[:, 1] >= 0) & (adds[:, 1] != rdst)))[0]
        torch = self.ctx.torch
        for j in cand.tolist():
            for v in adds[j].tolist():
                if v < 0 or self._roi_has(rows, [j], [v])[0]:
                    continue
                w = int(rows[j, v >> 5].item()) & 0xFFFFFFFF
                w |= 1 << (v & 31)
                rows[j, v >> 5] = torch.tensor(np.array([w], np.uint32).view(np.int32)[0], dtype=torch.int32)
                out.setdefault(j, []).append(v)
        return out

    def _portal_nodes(self, nid):
        """(seed at the source portal, seed at the target portal, source
        entry node, target entry node) of a net routed between its pads'
        portals (UPF:4509-4524, 4624-4639; a portal's seed is its F.Cu node,
        UPF:2495-2513), or None.  A portal whose row a retarget has moved
        off the lattice is not used."""
        if not self.config.portal_enabled or not self.portals:
            return None
        ids = self.net_pad_ids.get(nid)
        if ids is None:
            return None
        sp, dp = self.portals.get(ids[0]), self.portals.get(ids[1])
        if sp is None or dp is None:
            return None
        Y = self.lattice.y_steps
        if not (0 <= sp.y_idx < Y and 0 <= dp.y_idx < Y):
            return None
        ni = self.lattice.node_idx
        return (ni(sp.x_idx, sp.y_idx, 0), ni(dp.x_idx, dp.y_idx, 0),
                ni(sp.x_idx, sp.y_idx, sp.entry_layer), ni(dp.x_idx, dp.y_idx, dp.entry_layer))

    def _retarget_portals_for_net(self, nid):
        """UPF:4929-4969: after repeated failures a net's portals first flip
        to the pad's other side, then move to the preferred escape length."""
        ids = self.net_pad_ids.get(nid)
        if ids is None:
            return
        for pid in ids:
            p = self.portals.get(pid)
            if p is None:
                continue
            if p.retarget_count == 0:
                p.direction = -p.direction
                p.y_idx = p.y_idx - 2 * p.direction * p.delta_steps
                p.retarget_count += 1
            elif p.retarget_count == 1:
                new = self.config.portal_delta_pref
                if new != p.delta_steps:
                    p.y_idx += p.direction * (new - p.delta_steps)
                    p.delta_steps = new
                    p.retarget_count += 1

    def _route_all(self, tasks, all_tasks=None, pres_fac=1.0, iteration=0, over_map=None):
        # host phases of the pass (seconds, accumulated in self.stats):
        # route_req_s until the searches start, route_post_s from their end
        # to the commit (reordering, trees, fallbacks), route_commit_s the
        # commit and barrel marks
        t_start = time.perf_counter()
        if all_tasks is None:
            all_tasks = tasks
        self.full_graph_fallback_count = 0
        bonus = self.stagnation_counter * 0.6
        if over_map is None:
            over_map = np.zeros(self.accounting.E, dtype=np.float32)
        total = len(tasks)
        # tree nets of this call (multi_terminal_enabled): grown below, not
        # searched as two-pin nets
        tree_pins = getattr(self, "_tree_pins", {})
        trees = {nm for nm in tasks if nm in tree_pins}
        use_ko = total < 100 and len(self._ko_nodes) > 0
        ko_owner = self._keepout_owner_array() if use_ko else None
        # The searches of an iteration read one cost snapshot, so their
        # results do not depend on the routing order.  For a large batch they
        # run while a helper thread computes that order (started once the
        # requests are built, so that it takes the GIL only while the library
        # calls have released it), longest first by Manhattan distance, the
        # order's leading term: slots refill in request order, and long
        # searches started last would hold a refill batch open.  Results and
        # requests are then put in routing order, which the fallback and the
        # commit follow.
        helper = None
        if total > 1024 and not use_ko:
            box = {}

            def order_job():
                try:
                    box["ordered"] = self._order_nets_by_difficulty(tasks, over_map)
                except BaseException as exc:     # re-raised below
                    box["error"] = exc
            helper = threading.Thread(target=order_job, name="net-order")
            names = [nm for nm in tasks if nm not in trees]
            st = np.fromiter((v for nid in names for v in tasks[nid]), dtype=np.int64,
                             count=2 * len(names)).reshape(-1, 2)
            X_, plane_ = self._Nx, self._Nx * self._Ny
            zs_, rs_ = np.divmod(st[:, 0], plane_)
            zd_, rd_ = np.divmod(st[:, 1], plane_)
            dist_ = np.abs(rd_ % X_ - rs_ % X_) + np.abs(rd_ // X_ - rs_ // X_) + np.abs(zd_ - zs_)
            by_len = np.argsort(-dist_, kind="stable")
            search_order = [names[i] for i in by_len.tolist()]
            sd = st[by_len]
        else:
            ordered = self._order_nets_by_difficulty(tasks, over_map)
            search_order = [nm for nm in ordered if nm not in trees] if trees else ordered
            sd = np.fromiter((v for nid in search_order for v in tasks[nid]), dtype=np.int64,
                             count=2 * len(search_order)).reshape(-1, 2)
        n = len(search_order)
        X, plane = self._Nx, self._Nx * self._Ny
        rs, rd = sd[:, 0] % plane, sd[:, 1] % plane
        md = np.abs(rd % X - rs % X) + np.abs(rd // X - rs // X)
        reqs = np.empty((n, 11), dtype=np.float64)
        reqs[:, 0], reqs[:, 1] = sd[:, 0], sd[:, 1]
        reqs[:, 2] = [self._get_net_id(nid) for nid in search_order] if use_ko else -1
        reqs[:, 3] = md < ROI_THRESHOLD_STEPS
        reqs[:, 4] = bonus
        reqs[:, 5], reqs[:, 6] = sd[:, 0], sd[:, 1]
        reqs[:, 7:11] = -1
        # a net with portals on both pads searches from its source portal's
        # seed to its target portal's (a multi-source search with one seed
        # each, UPF:4732-4738, is this two-pin search) in an ROI grown from
        # the portals' entry nodes that keeps both seeds
        portal = {}
        if self.config.portal_enabled and self.portals:
            for i, nid in enumerate(search_order):
                pn = self._portal_nodes(nid)
                if pn is not None:
                    portal[nid] = pn
                    reqs[i, 0], reqs[i, 1], reqs[i, 5], reqs[i, 6] = pn
                    reqs[i, 7], reqs[i, 8] = pn[0], pn[1]
                    reqs[i, 9], reqs[i, 10] = sd[i, 0], sd[i, 1]
        # a batch that the engine runs at once (no slot refills) is only as
        # fast as its deepest searches: size its buckets by their depth (the
        # last estimate when none of the batch's nets holds a path)
        hint = None
        if n <= 1024:
            hint = self._depth_hint(search_order, sd)
            if hint is None:
                hint = self._last_depth_hint
            self._last_depth_hint = hint
        self.stats["route_req_s"] += time.perf_counter() - t_start
        if helper is not None:
            helper.start()
        try:
            results = self._search_batch(reqs, ko_owner, readd=True, depth_hint=hint)
        finally:
            if helper is not None:
                helper.join()
        t_post = time.perf_counter()
        if helper is not None:
            if "error" in box:
                raise box["error"]
            ordered = box["ordered"]
        if helper is not None or trees:
            # results and requests in routing order (a tree net's entry is
            # filled by the tree search)
            pos = {nm: i for i, nm in enumerate(search_order)}
            perm = np.fromiter((pos.get(nm, -1) for nm in ordered), dtype=np.int64, count=len(ordered))
            rows = np.full((len(ordered), reqs.shape[1]), -1.0)
            have = perm >= 0
            rows[have] = reqs[perm[have]]
            results = [results[i] if i >= 0 else None for i in perm.tolist()]
            reqs = rows
        if trees:
            grown = self._route_trees([nm for nm in ordered if nm in trees])
            for j, nm in enumerate(ordered):
                if nm in trees:
                    results[j] = grown[nm]
        # a portal net whose portal search failed searches its task's nodes
        # in the same ROI (UPF:4740-4748)
        if portal:
            pf = [i for i, (nid, r) in enumerate(zip(ordered, results))
                  if nid in portal and (r is None or len(r[0]) <= 1)]
            if pf:
                preqs = reqs[pf].copy()
                preqs[:, 0], preqs[:, 1] = preqs[:, 9], preqs[:, 10]
                for i, r in zip(pf, self._search_batch(preqs, ko_owner, readd=True)):
                    results[i] = r
        # wider-ROI fallback for the first five failures, in routing order;
        # a portal net's grows from its entry nodes without its seeds and
        # searches seed to seed (UPF:4753-4781)
        fb = [i for i, r in enumerate(results) if (r is None or len(r[0]) <= 1) and ordered[i] not in trees][:5]
        self.full_graph_fallback_count = len(fb)
        if fb:
            freqs = np.array([(reqs[i, 0], reqs[i, 1], -1, 1, bonus * 2, reqs[i, 5], reqs[i, 6], -1, -1, -1, -1)
                              for i in fb], dtype=np.float64)
            fres = self._search_batch(freqs, None, readd=False)
            for i, r in zip(fb, fres):
                results[i] = r
        t_commit = time.perf_counter()
        self.stats["route_post_s"] += t_commit - t_post
        # commit in routing order (UPF:4783-4812); usage counts are integer
        # adds, so the per-net +1/-1 scatters run as two batched launches
        # (the per-net bookkeeping of _clear_net_edge_tracking /
        # _update_net_edge_tracking, inlined: this loop runs over every net)
        dec, inc, marked, mark_ids = [], [], [], []
        paths, n2e, n2n, get_id = self.net_paths, self._net_to_edges, self._net_nodes, self._get_net_id
        id_map = self.net_id_map
        tb = self.__dict__.setdefault("_tree_branches", {})
        for nid, r in zip(ordered, results):
            held = paths.get(nid)
            if held is not None and len(held):
                old = n2e.pop(nid, None)
                if old is not None:
                    dec.append(old)
                n2n.pop(nid, None)
                tb.pop(nid, None)
            if nid in trees:
                if r is None:
                    paths[nid] = []
                    continue
                branches, bedges, nodes = r
                edges = np.concatenate(bedges) if bedges else np.zeros(0, np.int64)
                inc.append(edges)
                paths[nid] = nodes.view(NodePath)
                nid_int = get_id(nid)
                for b in branches:
                    marked.append(b)
                    mark_ids.append(nid_int)
                n2e[nid] = edges
                tb[nid] = (branches, bedges)
                continue
            if r is not None and len(r[0]) > 1:
                path, edges = r
                inc.append(edges)
                paths[nid] = path.view(NodePath)
                marked.append(path)
                k = id_map.get(nid)
                mark_ids.append(k if k is not None else get_id(nid))
                n2e.pop(nid, None)
                n2e[nid] = edges
                n2n[nid] = path
            else:
                paths[nid] = []
                n2e.pop(nid, None)
                n2n.pop(nid, None)
        if self.net_pad_ids and self.config.portal_enabled:
            # portal failures and retargeting (UPF:4800-4812); a net's portals
            # are its own pads', so retargeting after the batch is the same as
            # between nets
            plane = self._Nx * self._Ny
            patience = int(self.config.portal_retarget_patience)
            for nid, r in zip(ordered, results):
                if nid in trees:
                    continue
                if r is not None and len(r[0]) > 1:
                    if nid in portal:
                        self.net_portal_layers[nid] = (int(r[0][0]) // plane, int(r[0][-1]) // plane)
                elif nid in self.net_pad_ids:
                    self.net_portal_failures[nid] += 1
                    if self.net_portal_failures[nid] >= patience:
                        self._retarget_portals_for_net(nid)
                        self.net_portal_failures[nid] = 0
        if dec:
            self.accounting.clear_path(np.concatenate(dec))
        if inc:
            self.accounting.commit_path(np.concatenate(inc))
        self._mark_barrels_many(marked, mark_ids)
        routed = sum(1 for p in self.net_paths.values() if len(p))
        self.stats["route_commit_s"] += time.perf_counter() - t_commit
        return routed, len(all_tasks) - routed

    # ---------------- negotiation (UPF:3548-4259)
    def _negotiation_setup(self, tasks):
        """Board analysis, derived parameters and the layer-count schedule
        (UPF:3548-3600): returns (pres_fac, pres_fac_mult, pres_fac_max,
        hist_gain, hmult)."""
        cfg = self.config
        ch = analyze_board_characteristics(self.lattice, tasks)
        apply_derived_parameters(cfg, derive_routing_parameters(ch))
        self.h_layers, self.v_layers = ch.h_layers, ch.v_layers
        n_sig = self._Nz - 2
        pres_fac = float(cfg.pres_fac_init)
        pres_fac_mult = float(cfg.pres_fac_mult)
        hist_gain = float(cfg.hist_gain)
        if n_sig <= 12:
            pres_fac_max, hmult = 32.0, 1.2
        elif n_sig <= 20:
            pres_fac_max, hmult = 64.0, 1.0
        else:
            pres_fac_max, hmult = 128.0, 0.8
        pres_fac_mult = float(os.getenv("ORTHO_PRES_FAC_MULT", pres_fac_mult))
        pres_fac_max = float(os.getenv("ORTHO_PRES_FAC_MAX", pres_fac_max))
        hist_gain = float(os.getenv("ORTHO_HIST_GAIN", hist_gain))
        return pres_fac, pres_fac_mult, pres_fac_max, hist_gain, hmult

    def _begin_iteration(self, it, pres_fac, hmult):
        """STEPS 0-2 of an iteration (UPF:3640-3745): usage refresh, via
        usage and pooling pressure, overuse snapshot, present EMA, layer bias
        and the cost update.  Returns (flat committed arrays, dense
        iteration-start overuse)."""
        cfg = self.config
        E = self.accounting.E
        self.accounting.refresh_from_canonical()
        flat = self._flat()
        self._rebuild_via_usage_from_committed(flat)
        self._via_pressure()
        # over = max(0, present - 1) at iteration start (counts)
        p_idx, p_val = self.accounting.state.overuse_sparse(0)
        over_map = self._dense(p_idx, p_val, slot="over_map")
        ema_idx, ema_val = self._update_present_ema(first=(it == 1))
        lb = self._compute_layer_bias(ema_idx, ema_val)
        vcm = 1.0
        if pres_fac >= 64:
            if self._via_rank is None:
                k = self.graph.edge_kind[:E].to(dtype=self.ctx.torch.int64)
                self._via_rank = (self.ctx.torch.cumsum(k, 0) - k).cpu().numpy()
                self._n_via = int(k.sum().item())
            kind = self.graph.edge_kind[:E]
            kk = kind[self.ctx.torch.from_numpy(p_idx).to(self.ctx.device)].cpu().numpy() if len(p_idx) else np.zeros(0, np.uint8)
            vsel = kk == 1
            via_over = float(sum_f32(self._n_via, self._via_rank[p_idx[vsel]], p_val[vsel]))
            tot_over = float(sum_f32(E, p_idx, p_val))
            if tot_over > 0:
                vcm = 1.5 if via_over / tot_over > 0.7 else 0.5
        self._update_costs(pres_fac, cfg.hist_cost_weight * hmult, vcm, cfg.base_cost_weight, lb, it)
        return flat, over_map

    def _pathfinder_negotiation(self, tasks, progress_cb=None, iteration_cb=None):
        cfg = self.config
        t_start = time.perf_counter()
        pres_fac, pres_fac_mult, pres_fac_max, hist_gain, hmult = self._negotiation_setup(tasks)
        n_sig = self._Nz - 2
        best = float("inf")
        stagnant = 0
        self._negotiation_ran = True
        self._last_depth_hint = None
        # the bucket target of the batches with a depth estimate, tuned on
        # this negotiation's own search times (speed only: paths are the same)
        slots_key = int(min(int(cfg.batch_slots), max_slots(self.N)) > 1)
        self._tuner_key = f"{self._Nx}x{self._Ny}x{self._Nz}:{self.lattice.policy}:{slots_key}"
        self._tuners = {}
        self._negotiation_nets = len(tasks)
        self._net_failure_count = {}
        self._excluded_nets = set()
        routed = failed = 0
        over_sum = over_cnt = 0
        jam = None
        E = self.accounting.E
        for it in range(1, cfg.max_iterations + 1):
            t_it = time.perf_counter()
            self.iteration = it
            # STEP 0-2: usage counts are maintained incrementally in HBM
            flat, over_map = self._begin_iteration(it, pres_fac, hmult)
            if it > 10 and it % 10 == 0 and self._excluded_nets:
                self._excluded_nets.clear()
                self._net_failure_count.clear()
            if it > 4:
                for n in tasks:
                    if n in self._last_ripped:
                        continue
                    if not len(self.net_paths.get(n, ())):
                        self._net_failure_count[n] = self._net_failure_count.get(n, 0) + 1
                    else:
                        self._net_failure_count[n] = 0
                for n, f in self._net_failure_count.items():
                    if f >= 5 and n not in self._excluded_nets:
                        self._excluded_nets.add(n)
            if cfg.reroute_only_offenders and it > 1:
                h_idx, h_over = self.accounting.state.overuse_sparse(1, 0.1, 0.5)
                offenders = self._build_hotset(tasks, h_idx, h_over, flat, ripped=self._last_ripped)
                offenders -= self._excluded_nets
                sub = {k: v for k, v in tasks.items() if k in offenders and k not in self._excluded_nets}
                self._last_ripped = set()
            else:
                sub = {k: v for k, v in tasks.items() if k not in self._excluded_nets}
            if it > 1:
                ids = list(sub.keys())
                self._rng.seed(42 + it)
                self._rng.shuffle(ids)
                sub = {k: sub[k] for k in ids}
            t_r = time.perf_counter()
            routed, failed = self._route_all(sub, all_tasks=tasks, pres_fac=pres_fac, iteration=it, over_map=over_map)
            self.stats["route_s"] += time.perf_counter() - t_r
            # present = counts, then the barrel-conflict penalty (UPF:3833-3877)
            if it <= 50 or it % 10 == 0:
                ce, cc = self._detect_barrel_conflicts(self._flat())
            else:
                ce, cc = np.zeros(0, np.int32), getattr(self, "_last_barrel_conflict_count", 0)
            self._last_barrel_conflict_count = cc
            if cc > 0 and it <= 50:
                self.accounting.state.build_present(np.unique(ce), min(10.0 * pres_fac, 100.0))
            else:
                self.accounting.state.build_present()
            o_idx, o_val = self.accounting.state.overuse_sparse(0)
            over_sum = int(sum_f32(E, o_idx, o_val))
            over_cnt = int(len(o_idx))
            # the via part: column and segment overuse of the usage recounted
            # at the iteration's start (orgpu_via_pressure's sum)
            over_sum += int(self._via_dev["over"].item())
            by_layer = None
            if over_sum > 0 and len(o_idx):
                lay = self._edge_layers(o_idx)
                by_layer = {}
                for z in range(1, self._Nz + 1):
                    m = lay == z
                    if np.any(m):
                        by_layer[z] = float(np.sum(o_val[m]))
            if by_layer and sum(by_layer.values()) > 0:
                lo = np.array([by_layer.get(z, 0.0) for z in range(self._Nz)])
                mean = np.mean(lo[lo > 0]) + 1e-9
                pressure = lo / mean
                if n_sig <= 12:
                    a, bmin, bmax = 0.20, 0.60, 1.80
                elif n_sig <= 20:
                    a, bmin, bmax = 0.12, 0.75, 1.50
                else:
                    a, bmin, bmax = 0.08, 0.85, 1.20
                target = 1.0 + a * (pressure - 1.0)
                self.layer_bias = 0.85 * self.layer_bias + 0.15 * target
                self.layer_bias = np.clip(self.layer_bias, bmin, bmax)
                tot_l = sum(by_layer.values())
                if tot_l > 0:
                    mi, mo = max(by_layer.items(), key=lambda x: x[1])
                    pct = (mo / tot_l) * 100
                    if jam is None:
                        jam = {"layer": None, "count": 0, "boost_until": 0}
                    if pct >= 60.0:
                        if jam["layer"] == mi:
                            jam["count"] += 1
                        else:
                            jam = {"layer": mi, "count": 1, "boost_until": 0}
                        if jam["count"] >= 3 and it > jam["boost_until"]:
                            old = self.layer_bias[mi]
                            self.layer_bias[mi] = min(old * 1.8, bmax)
                            jam["boost_until"] = it + 3
                    else:
                        jam = {"layer": None, "count": 0, "boost_until": 0}
            self.iteration_log.append({"it": it, "over_sum": over_sum, "over_cnt": over_cnt, "pres_fac": pres_fac,
                                       "routed": routed, "failed": failed, "rerouted": len(sub),
                                       "t": time.perf_counter() - t_it})
            if over_sum == 0:
                unrouted = {n for n in tasks if not len(self.net_paths.get(n, ()))}
                if not unrouted:
                    break
                pres_fac = min(pres_fac, 10.0)
            hg = hist_gain
            if it >= 8 and over_sum > 0:
                hg = min(1.2, hist_gain * 1.25)
            use_cap = os.getenv("ORTHO_NO_HISTORY_CAP", "0") == "0"
            decay = 0.98 if it >= 10 else 1.0
            self.accounting.state.update_history(f32(decay), f32(hg), f32(15.0 if use_cap else 1.0), use_cap)
            if progress_cb:
                try:
                    progress_cb(it, cfg.max_iterations, f"Iteration {it}")
                except Exception:
                    pass
            if iteration_cb:
                try:
                    # the current routing state as geometry (UPF:4066-4117;
                    # no escape geometry to merge in this build)
                    prov_tracks, prov_vias = (self._generate_geometry_from_paths()
                                              if cfg.iteration_geometry else ([], []))
                    iteration_cb(it, prov_tracks, prov_vias, over_sum, over_cnt)
                except Exception as e:
                    logger.warning(f"[ITER {it}] iteration callback failed: {e}")
            if failed == 0 and over_sum == 0 and self._last_barrel_conflict_count == 0:
                self.stats["negotiation_s"] = time.perf_counter() - t_start
                return {"success": True, "paths": self.net_paths, "converged": True}
            if over_sum < best:
                best = over_sum
                stagnant = 0
            else:
                stagnant += 1
            if stagnant == 2:
                pres_fac = min(pres_fac * 1.5, pres_fac_max)
            if stagnant >= cfg.stagnation_patience:
                self.stagnation_counter += 1
                self._last_ripped = self._rip_top_k_offenders(o_idx, o_val, k=20)
                self._freeze_pres_fac_until = it + 2
                stagnant = 0
                continue
            if it > self._freeze_pres_fac_until:
                pres_fac = min(pres_fac * pres_fac_mult, pres_fac_max)
        if 0 < over_sum <= 100:
            r = self._detail_pass(tasks, over_sum, over_cnt, o_idx, o_val)
            if r["success"]:
                self.stats["negotiation_s"] = time.perf_counter() - t_start
                return r
        self.stats["negotiation_s"] = time.perf_counter() - t_start
        success = failed == 0 and over_sum == 0
        excluded = getattr(self, "_excluded_nets", set())
        return {"success": success, "converged": success, "paths": self.net_paths,
                "barrel_conflicts": getattr(self, "_last_barrel_conflict_count", 0),
                "excluded_nets": len(excluded), "excluded_net_ids": list(excluded),
                "error_code": None if success else "ROUTING-FAILED",
                "message": "Complete" if success else f"{failed} unrouted, {over_cnt} overused",
                "overuse_sum": over_sum, "overuse_edges": over_cnt, "failed_nets": failed,
                "layer_recommendation": self._analyze_layer_requirements(failed, over_cnt, over_sum)}

    def _analyze_layer_requirements(self, failed, over_edges, over_sum):
        """UPF:4261-4304."""
        L = self.lattice.layers
        fail_rate = failed / max(1, len(self.net_pad_ids))
        if fail_rate > 0.4 and over_edges > 200:
            add = max(4, int(failed / 50))
            return {"needs_more": True, "additional": add, "recommended_total": L + add,
                    "reason": f"High failure rate ({fail_rate * 100:.1f}%) with {over_edges} congested edges"}
        if over_sum > 800 and over_edges > 400:
            return {"needs_more": True, "additional": 6, "recommended_total": L + 6,
                    "reason": f"Severe congestion ({over_sum} conflicts across {over_edges} edges)"}
        if fail_rate > 0.3:
            add = 4 if fail_rate > 0.35 else 2
            return {"needs_more": True, "additional": add, "recommended_total": L + add,
                    "reason": f"Moderate routing failure ({fail_rate * 100:.1f}%)"}
        return {"needs_more": False, "additional": 0, "recommended_total": L,
                "reason": f"Failure rate ({fail_rate * 100:.1f}%) and overuse ({over_edges} edges) acceptable"}

    def _detail_pass(self, tasks, initial_overuse, initial_edges, o_idx, o_val):
        """UPF:4306-4399.  The reference picks the first 60 of a Python set of
        net names (order depends on the string hash seed); sorted names are used
        here."""
        conflict = np.zeros(self.accounting.E, dtype=bool)
        conflict[o_idx] = True
        nets = set()
        for n, path in self.net_paths.items():
            e = self._net_to_edges.get(n)
            if path and e is not None and conflict[e].any():
                nets.add(n)
        if not nets:
            return {"success": False, "error_code": "NO-CONFLICT-NETS"}
        ctasks = {n: tasks[n] for n in sorted(nets) if n in tasks}
        cfg = self.config
        pres_fac = cfg.pres_fac_max * 0.5
        best = initial_overuse
        E = self.accounting.E
        for _ in range(1, 11):
            self.accounting.state.build_present()
            ones = np.ones(self._Nz, dtype=np.float32)
            self._update_costs(pres_fac, cfg.hist_cost_weight * 1.5, 0.3, cfg.base_cost_weight, ones, 0,
                               with_vias=False)
            h_idx, h_over = self.accounting.state.overuse_sparse(1, 0.1, 0.5)
            hs = self._build_hotset(ctasks, h_idx, h_over, self._flat())
            hs = set(sorted(hs)[:60]) or nets
            sub = {k: v for k, v in ctasks.items() if k in hs}
            p_idx, p_val = self.accounting.state.overuse_sparse(0)
            old = self.stagnation_counter
            self.stagnation_counter += 3
            self._route_all(sub, all_tasks=tasks, pres_fac=pres_fac, over_map=self._dense(p_idx, p_val))
            self.stagnation_counter = old
            self.accounting.state.build_present()
            s, _ = self.accounting.compute_overuse(self)
            if s == 0:
                return {"success": True, "paths": self.net_paths}
            best = min(best, s)
            # update_history(gain*2, decay 0.98) on present_ema (UPF:4385-4390)
            self._history_from_ema(f32(0.98), f32(cfg.hist_gain * 2.0))
            pres_fac = min(pres_fac * 1.5, cfg.pres_fac_max)
        return {"success": False, "error_code": "DETAIL-INCOMPLETE", "overuse_sum": best}

    def _history_from_ema(self, decay, gain):
        st = self.accounting.state
        torch = self.ctx.torch
        saved = st.present.clone()
        st.present.copy_(st.present_ema)
        st.update_history(decay, gain, f32(15.0), True)
        st.present.copy_(saved)

    # ---------------- geometry (UPF:5677-5914 surface)
    def emit_geometry(self, board):
        """UPF:5677-5747: tracks and vias of the committed paths, after the
        pad escape stubs when there are any (escapes first, then routes,
        duplicates by rounded geometry dropped); the export payload only when
        the routing has no overuse (edge or via), the counts either way."""
        tracks, vias = self._generate_geometry_from_paths()
        if self._escape_tracks:
            tracks = _dedupe(list(self._escape_tracks) + tracks,
                             lambda t: (t["net"], t["layer"], round(t["x1"], 3), round(t["y1"], 3),
                                        round(t["x2"], 3), round(t["y2"], 3), round(t["width"], 3)))
            vias = _dedupe(list(self._escape_vias) + vias,
                           lambda v: (v["net"], round(v["x"], 3), round(v["y"], 3), v.get("from_layer"),
                                      v.get("to_layer"), round(v.get("drill", 0), 3), round(v.get("diameter", 0), 3)))
        self._provisional_geometry = GeometryPayload(tracks, vias)
        over_sum, _ = self.accounting.compute_overuse(router_instance=self)
        self._geometry_payload = GeometryPayload(tracks, vias) if over_sum == 0 else GeometryPayload([], [])
        return len(tracks), len(vias)

    def get_geometry_payload(self):
        """UPF:5898-5910: the export payload, or the provisional geometry
        when there is no export payload (overuse remains)."""
        prov = getattr(self, "_provisional_geometry", None)
        if not self._geometry_payload.tracks and not self._geometry_payload.vias and prov is not None \
                and (prov.tracks or prov.vias):
            return prov
        return self._geometry_payload

    def get_provisional_geometry(self):
        return getattr(self, "_provisional_geometry", GeometryPayload([], []))

    def _generate_geometry_from_paths(self):
        """UPF:5749-5830 with _segment_world / _via_world (UPF:5596-5617,
        5652-5675): every path becomes its maximal straight track runs (a run
        ends at a direction change, before a via and at the path's end) and
        one via record per layer change, in path order, nets in net_paths
        order.  A net's records depend on its path alone, so they are kept
        per net and recomputed only for nets whose path object changed since
        the last call (a late negotiation iteration reroutes a hundred of
        thousands of nets); the records of an unchanged net are the same
        objects as before.  A tree net's records are its branches' records,
        branch after branch."""
        items = []      # (cache key, net, path)
        trees = getattr(self, "_tree_branches", {})
        for net, p in self.net_paths.items():
            if not p:
                continue
            tb = trees.get(net)
            if tb is None:
                items.append((net, net, p))
            else:
                items.extend((("branch", net, i), net, b) for i, b in enumerate(tb[0]))
        if not items:
            return [], []
        g = self.lattice.geom
        names = list(self.config.layer_names or [])
        sig = (float(g.pitch), float(g.grid_min_x), float(g.grid_min_y), self.config.grid_pitch * 0.6,
               tuple(names), self.lattice.policy, self._Nx, self._Ny, self._Nz)
        cache = getattr(self, "_geo_cache", None)
        if cache is None or cache[0] != sig:
            cache = (sig, {})
        per = cache[1]
        todo = [(key, net, p) for key, net, p in items if key not in per or per[key][0] is not p]
        if todo:
            recs = self._geometry_records([(net, p) for _, net, p in todo])
            for (key, _, p), rec in zip(todo, recs):
                per[key] = (p, rec[0], rec[1])
        keep = {key: per[key] for key, _, _ in items}
        self._geo_cache = (sig, keep)
        tracks = [t for key, _, _ in items for t in keep[key][1]]
        vias = [v for key, _, _ in items for v in keep[key][2]]
        return tracks, vias

    def _geometry_records(self, items):
        """(tracks, vias) per (net, path) item, vectorised over all items.
        Planar hops on the outer layers are refused on the reference lattice,
        whose outer layers carry no planar edges.  A path with a hop that is
        not a lattice move (the reference skips such hops without advancing
        its walk) goes through the per-hop walk instead."""
        X, L = self._Nx, self._Nz
        plane = X * self._Ny
        g = self.lattice.geom
        pitch, ox, oy = float(g.pitch), float(g.grid_min_x), float(g.grid_min_y)
        width = self.config.grid_pitch * 0.6
        names = list(self.config.layer_names or [])
        lname = [names[z] if z < len(names) else f"L{z}" for z in range(L)]
        guard = self.lattice.policy == "reference"
        haxis = np.array([self.lattice.get_legal_axis(z) == "h" for z in range(L)])

        lens = np.fromiter((len(p) for _, p in items), dtype=np.int64, count=len(items))
        allp = np.concatenate([np.asarray(p, dtype=np.int64) for _, p in items])
        inner = np.ones(max(len(allp) - 1, 0), dtype=bool)
        inner[np.cumsum(lens)[:-1] - 1] = False
        a, b = allp[:-1][inner], allp[1:][inner]
        hp = np.repeat(np.arange(len(items)), lens - 1)
        za, ra = np.divmod(a, plane)
        zb, rb = np.divmod(b, plane)
        xa, ya = ra % X, ra // X
        xb, yb = rb % X, rb // X
        planar = za == zb
        outer = planar & ((za == 0) | (za == L - 1)) if guard else np.zeros(len(a), dtype=bool)
        run = planar & ~outer
        dx, dy = xb - xa, yb - ya
        legal = (np.abs(dx) + np.abs(dy) == 1) & np.where(haxis[za], dy == 0, dx == 0)
        bad_paths = np.unique(hp[run & ~legal])
        ok = ~np.isin(hp, bad_paths) if len(bad_paths) else np.ones(len(a), dtype=bool)

        def world(x, y):
            wx = ox + np.asarray(x, dtype=np.float64) * pitch
            wy = oy + np.asarray(y, dtype=np.float64) * pitch
            return ox + np.round((wx - ox) / pitch) * pitch, oy + np.round((wy - oy) / pitch) * pitch

        # runs: consecutive planar hops of one path in one direction
        r_idx = np.nonzero(run & ok)[0]
        dcode = (np.sign(dx[r_idx]) + 1) * 3 + (np.sign(dy[r_idx]) + 1)
        start = np.ones(len(r_idx), dtype=bool)
        if len(r_idx) > 1:
            start[1:] = (r_idx[1:] != r_idx[:-1] + 1) | (dcode[1:] != dcode[:-1]) | (hp[r_idx[1:]] != hp[r_idx[:-1]])
        first = r_idx[start]
        last = r_idx[np.append(np.nonzero(start)[0][1:] - 1, len(r_idx) - 1)] if len(r_idx) else r_idx
        x1, y1 = world(xa[first], ya[first])
        x2, y2 = world(xb[last], yb[last])
        v_idx = np.nonzero(~planar & ok)[0]
        vx, vy = world(xa[v_idx], ya[v_idx])
        vlo, vhi = np.minimum(za[v_idx], zb[v_idx]), np.maximum(za[v_idx], zb[v_idx])

        net_names = [net for net, _ in items]
        t_path, v_path = hp[first], hp[v_idx]
        t_rec = [{"net": net_names[k], "layer": lname[z], "x1": p1, "y1": q1, "x2": p2, "y2": q2, "width": width}
                 for k, z, p1, q1, p2, q2 in zip(t_path.tolist(), za[first].tolist(), x1.tolist(), y1.tolist(),
                                                 x2.tolist(), y2.tolist())]
        v_rec = [{"net": net_names[k], "x": p1, "y": q1, "from_layer": lname[lo], "to_layer": lname[hi],
                  "diameter": 0.25, "drill": 0.15}
                 for k, p1, q1, lo, hi in zip(v_path.tolist(), vx.tolist(), vy.tolist(), vlo.tolist(), vhi.tolist())]
        # records are in item order: split them per item
        n = len(items)
        tb = np.searchsorted(t_path, np.arange(n + 1)).tolist()
        vb = np.searchsorted(v_path, np.arange(n + 1)).tolist()
        out = [(t_rec[tb[k]:tb[k + 1]], v_rec[vb[k]:vb[k + 1]]) for k in range(n)]
        for k in bad_paths.tolist():
            net, p = items[k]
            out[k] = self._walk_geometry(net, p, lname, guard, width, world)
        return out

    def _walk_geometry(self, net, path, lname, guard, width, world):
        """The per-hop form of UPF:5756-5830 for one path."""
        tracks, vias = [], []
        L = self._Nz

        def seg(u, v, z):
            (ux, uy, _), (vx_, vy_, _) = self._coords(u), self._coords(v)
            p1, q1 = world(ux, uy)
            p2, q2 = world(vx_, vy_)
            return {"net": net, "layer": lname[z], "x1": float(p1), "y1": float(q1), "x2": float(p2),
                    "y2": float(q2), "width": width}

        run_start = prev = path[0]
        prev_dir = None
        prev_layer = self._coords(prev)[2]
        for node in path[1:]:
            x0, y0, z0 = self._coords(prev)
            x1, y1, z1 = self._coords(node)
            if guard and z0 == z1 and (z0 == 0 or z0 == L - 1):
                prev = run_start = node
                prev_layer = z1
                continue
            if z0 == z1:
                if abs(x1 - x0) + abs(y1 - y0) != 1:
                    continue
                if self.lattice.get_legal_axis(z0) == "h":
                    if y1 != y0:
                        continue
                elif x1 != x0:
                    continue
            if z1 != z0:
                if prev != run_start:
                    tracks.append(seg(run_start, prev, prev_layer))
                wx, wy = world(x0, y0)
                vias.append({"net": net, "x": float(wx), "y": float(wy), "from_layer": lname[min(z0, z1)],
                             "to_layer": lname[max(z0, z1)], "diameter": 0.25, "drill": 0.15})
                run_start = node
                prev_dir = None
            else:
                d = ((x1 > x0) - (x1 < x0), (y1 > y0) - (y1 < y0))
                if prev_dir is not None and d != prev_dir:
                    tracks.append(seg(run_start, prev, prev_layer))
                    run_start = prev
                prev_dir = d
            prev = node
            prev_layer = z1
        if prev != run_start:
            tracks.append(seg(run_start, prev, prev_layer))
        return tracks, vias


UnifiedPathFinder = PathFinderRouter
