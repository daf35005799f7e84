"""CPU oracle for the PathFinder inner loop -- TEST INFRASTRUCTURE ONLY.

This package restates the reference algorithms of
orthoroute/algorithms/manhattan/unified_pathfinder.py (UPF) on the CPU so the
GPU path can be checked against them.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline; the product
package ``orthoroute_amd`` never imports it.

Parts and what pins them (each checked in tests/test_oracle_*.py):

  lattice_spec / build_lattice   UPF:1091-1106, 1136-1181, 1195-1338, 666-792
        pinned by the reference's own CSR arrays (tests/golden/lattice_*.npz,
        via_pairs.json) and by closed forms (edge counts, degree, symmetry).
  sssp_canon                     the GPU contract (FP32 packed-key semantics)
        pinned by scipy.sparse.csgraph.dijkstra distances, brute force on tiny
        graphs, the tight-edge / lowest-pred invariants, and path cost within
        1e-5 of the reference Dijkstra.
  sssp_ref                       SimpleDijkstra.find_path_roi UPF:1801-1887
        pinned by the reference's own paths (tests/golden/dijkstra.npz).
  roi_bfs                        ROIExtractor.extract_roi(_bfs) UPF:1619-1734
        pinned by the reference's own ROIs (tests/golden/roi_bfs.npz).
  ms_ref                         find_path_multisource_multisink UPF:1889-1989
        pinned by the reference's own paths (tests/golden/multisource.npz).
  ms_canon                       multi-source search in the GPU contract's
        semantics: pinned by brute force over all simple paths on tiny graphs
        and by the total cost of ms_ref (within 1e-5 relative).
  steiner_canon                  multi-terminal trees by repeated ms_canon
        attachments (the operation the build defines for the reference's
        unreachable multi-terminal router, MT:557-820): pinned by scipy
        multi-source distances (each branch costs the tree-to-pin distance
        of the pin it attaches, and that pin is the nearest one left) and by
        the two-pin case reducing to sssp_canon.
  oracle.costs, oracle.negotiation
        pinned by tests/golden/costs.npz and the reference's full negotiation
        trajectories (tests/golden/negotiation*.npz).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

DIR_V, DIR_H, DIR_HV = 0, 1, 2


def build(force=False):
    """Compile oracle/csrc/oracle.c with gcc (the oracle Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i32, i64, f32, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_float, ctypes.c_double
        _lib.orc_lattice_edge_count.restype = i64
        _lib.orc_lattice_edge_count.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int]
        _lib.orc_lattice_build.restype = ctypes.c_int
        _lib.orc_lattice_build.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, f32, f64,
                                           P, P, ctypes.c_int, P, P, P, P, P]
        _lib.orc_sssp_canon.restype = ctypes.c_int
        _lib.orc_sssp_canon.argtypes = [i64, P, P, P, P, P, P, i32, f32, i32, i32, P, P, P, P, P]
        _lib.orc_backtrace.restype = i64
        _lib.orc_backtrace.argtypes = [P, i32, i32, P, i64]
        _lib.orc_sssp_ref.restype = i64
        _lib.orc_sssp_ref.argtypes = [i64, P, P, P, P, i64, P, P, P, i32, f32, i32, i32, P, i64, P]
        _lib.orc_ms_ref.restype = i64
        _lib.orc_ms_ref.argtypes = [i64, P, P, P, i64, P, P, i64, P, P, i64, P, P, P, i64, P]
        _lib.orc_ms_canon.restype = i64
        _lib.orc_ms_canon.argtypes = [i64, P, P, P, P, i64, P, P, i64, P, P, P, P, P, P, P]
        _lib.orc_roi_bfs.restype = i64
        _lib.orc_roi_bfs.argtypes = [i64, P, P, i32, i32, ctypes.c_int, i64, P, ctypes.c_int, ctypes.c_int,
                                     P, P, P, P]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# Lattice
# --------------------------------------------------------------------------

def reference_via_pairs(L):
    """Via layer pairs in the iteration order of the reference's Python set
    (Lattice3D.get_legal_via_pairs, UPF:1136-1181): ordered pairs among the
    routing layers 1..L-2, then (0, z) and (z, 0); {(0,1),(1,0)} for L == 2."""
    if L == 2:
        pairs = {(0, 1), (1, 0)}
        return [tuple(p) for p in pairs]
    routing = list(range(1, L - 1))
    s = set()
    for z1 in routing:
        for z2 in routing:
            if z1 != z2:
                s.add((z1, z2))
    for z in routing:
        s.add((0, z))
        s.add((z, 0))
    return [tuple(p) for p in s]


def lattice_spec(L, policy="reference"):
    """(planar[L] u8, dir[L] u8, pairs) for a lattice policy.

    reference: UPF:1091-1106 directions (L==2: 'hv','hv'; else z=0 'v',
      odd 'h', even 'v'); planar edges on z in [1, L-2] (all layers if L<=2)
      (UPF:1210); via pairs from reference_via_pairs.
    adjacent: same directions, planar edges on every layer, vias only between
      adjacent layers (z, z+1), one pair per adjacency.
    """
    if L == 2:
        dirs = [DIR_HV, DIR_HV]
    else:
        dirs = [DIR_V if z == 0 else (DIR_H if z % 2 == 1 else DIR_V) for z in range(L)]
    if policy == "reference":
        planar = [1] * L if L <= 2 else [0] + [1] * (L - 2) + [0]
        pairs = reference_via_pairs(L)
    elif policy == "adjacent":
        planar = [1] * L
        pairs = [(z, z + 1) for z in range(L - 1)]
    else:
        raise ValueError(f"unknown lattice policy {policy!r}")
    return (np.array(planar, dtype=np.uint8), np.array(dirs, dtype=np.uint8), pairs)


class Lattice:
    """CSR arrays of the X*Y*L Manhattan lattice, node = z*X*Y + y*X + x."""

    def __init__(self, X, Y, L, policy="reference", pitch=1.0, via_cost=3.0):
        self.X, self.Y, self.L = int(X), int(Y), int(L)
        self.policy = policy
        self.N = self.X * self.Y * self.L
        planar, dirs, pairs = lattice_spec(self.L, policy)
        self.planar, self.dirs, self.pairs = planar, dirs, pairs
        pf = np.array([p[0] for p in pairs], dtype=np.int32)
        pt = np.array([p[1] for p in pairs], dtype=np.int32)
        lb = lib()
        E = lb.orc_lattice_edge_count(self.X, self.Y, self.L, _p(planar), _p(dirs), len(pairs))
        self.E = int(E)
        self.indptr = np.zeros(self.N + 1, dtype=np.int32)
        self.indices = np.zeros(self.E, dtype=np.int32)
        self.costs = np.zeros(self.E, dtype=np.float32)
        self.edge_layer = np.zeros(self.E, dtype=np.uint8)
        self.edge_kind = np.zeros(self.E, dtype=np.uint8)
        rc = lb.orc_lattice_build(self.X, self.Y, self.L, _p(planar), _p(dirs), float(pitch),
                                  float(via_cost), _p(pf), _p(pt), len(pairs), _p(self.indptr),
                                  _p(self.indices), _p(self.costs), _p(self.edge_layer),
                                  _p(self.edge_kind))
        if rc != 0:
            raise MemoryError("oracle lattice build failed")

    def coords(self, n):
        plane = self.X * self.Y
        n = np.asarray(n)
        return n % plane % self.X, n % plane // self.X, n // plane

    def node(self, x, y, z):
        return (z * self.Y + y) * self.X + x


# --------------------------------------------------------------------------
# SSSP
# --------------------------------------------------------------------------

KEEPOUT_COST = np.float32(1e9)   # UPF:3182


def sssp_canon(lat, costs, src, dst, roi_mask=None, keepout_owner=None, net_id=-1):
    """FP32 SSSP with lowest-predecessor tie-break; returns (path list or None,
    dist f32[N], pred i32[N], expanded, traversed).  Paths are ordered by
    (number of keepout edges, fp32 cost of the other edges)."""
    N = lat.N
    dist = np.empty(N, dtype=np.float32)
    pred = np.empty(N, dtype=np.int32)
    costs = np.ascontiguousarray(costs, dtype=np.float32)
    rm = None if roi_mask is None else np.ascontiguousarray(roi_mask, dtype=np.uint8)
    ko = None if keepout_owner is None else np.ascontiguousarray(keepout_owner, dtype=np.int32)
    nko = np.empty(N, dtype=np.uint8)
    ex, tr = ctypes.c_int64(0), ctypes.c_int64(0)
    r = lib().orc_sssp_canon(N, _p(lat.indptr), _p(lat.indices), _p(costs), _p(lat.edge_kind),
                             _p(rm), _p(ko), int(net_id), float(KEEPOUT_COST), int(src), int(dst),
                             _p(dist), _p(pred), _p(nko), ctypes.byref(ex), ctypes.byref(tr))
    if r < 0:
        raise MemoryError("oracle sssp failed")
    path = None
    if r == 1:
        out = np.empty(N, dtype=np.int32)
        n = lib().orc_backtrace(_p(pred), int(src), int(dst), _p(out), N)
        if n > 1:
            path = out[:n].tolist()
    return path, dist, pred, ex.value, tr.value


def sssp_ref(lat, costs, src, dst, roi_nodes, global_to_roi=None, keepout_owner=None, net_id=-1):
    """Exact emulation of SimpleDijkstra.find_path_roi; returns (path or None, alt)."""
    roi_nodes = np.ascontiguousarray(roi_nodes, dtype=np.int32)
    if global_to_roi is None:
        global_to_roi = np.full(lat.N, -1, dtype=np.int32)
        global_to_roi[roi_nodes] = np.arange(len(roi_nodes), dtype=np.int32)
    g2r = np.ascontiguousarray(global_to_roi, dtype=np.int32)
    costs = np.ascontiguousarray(costs, dtype=np.float32)
    ko = None if keepout_owner is None else np.ascontiguousarray(keepout_owner, dtype=np.int32)
    out = np.empty(max(len(roi_nodes), 1), dtype=np.int32)
    alt = ctypes.c_double(np.inf)
    n = lib().orc_sssp_ref(lat.N, _p(lat.indptr), _p(lat.indices), _p(costs), _p(lat.edge_kind),
                           len(roi_nodes), _p(roi_nodes), _p(g2r), _p(ko), int(net_id),
                           float(KEEPOUT_COST), int(src), int(dst), _p(out), len(out),
                           ctypes.byref(alt))
    if n < 0:
        raise MemoryError("oracle ref sssp failed")
    return (out[:n].tolist() if n > 1 else None), alt.value


# --------------------------------------------------------------------------
# ROI
# --------------------------------------------------------------------------

def roi_radius(lat, src, dst, stagnation_bonus=0.0):
    """ROIExtractor.extract_roi radius rule (UPF:1619-1634) and the BFS depth
    cap (UPF:1653): radius = max(60, int(dist*0.75 + 2*bonus)),
    depth = min(int(radius + 2*bonus), 150)."""
    sx, sy, _ = lat.coords(src)
    dx, dy, _ = lat.coords(dst)
    md = abs(int(dx) - int(sx)) + abs(int(dy) - int(sy))
    radius = max(60, int(md * 0.75 + stagnation_bonus * 2.0))
    return radius, min(int(radius + stagnation_bonus * 2.0), 150)


def roi_bfs(lat, src, dst, stagnation_bonus=0.0, max_nodes=50_000, keep=(), readd=False):
    """ROI node list in the reference's order (extract_roi -> extract_roi_bfs,
    UPF:1619-1734): keep = further must-keep nodes (a portal net's seeds);
    readd: then the keeps, src and dst appended when missing, as _route_all
    does for the main search (UPF:4659-4696)."""
    _, depth = roi_radius(lat, src, dst, stagnation_bonus)
    out = np.empty(lat.N, dtype=np.int32)
    seen = np.empty(lat.N, dtype=np.uint8)
    d, m = ctypes.c_int(0), ctypes.c_int(0)
    kp = np.ascontiguousarray([int(k) for k in keep], dtype=np.int32) if len(keep) else np.zeros(1, np.int32)
    R = lib().orc_roi_bfs(lat.N, _p(lat.indptr), _p(lat.indices), int(src), int(dst), int(depth),
                          int(max_nodes), _p(kp), len(keep), int(bool(readd)), _p(out), _p(seen), ctypes.byref(d),
                          ctypes.byref(m))
    if R < 0:
        raise MemoryError("oracle roi failed")
    return out[:R].copy()


def ms_ref(lat, costs, seeds, targets, roi_nodes, global_to_roi=None):
    """Exact emulation of find_path_multisource_multisink (UPF:1889-1989):
    seeds / targets are lists of (node, cost).  Returns (path or None, best
    total)."""
    roi_nodes = np.ascontiguousarray(roi_nodes, dtype=np.int32)
    if global_to_roi is None:
        global_to_roi = np.full(lat.N, -1, dtype=np.int32)
        global_to_roi[roi_nodes] = np.arange(len(roi_nodes), dtype=np.int32)
    g2r = np.ascontiguousarray(global_to_roi, dtype=np.int32)
    sn = np.ascontiguousarray([a for a, _ in seeds], dtype=np.int32)
    sc = np.ascontiguousarray([b for _, b in seeds], dtype=np.float64)
    tn = np.ascontiguousarray([a for a, _ in targets], dtype=np.int32)
    tc = np.ascontiguousarray([b for _, b in targets], dtype=np.float64)
    costs = np.ascontiguousarray(costs, dtype=np.float32)
    out = np.empty(max(len(roi_nodes), 1), dtype=np.int32)
    best = ctypes.c_double(np.inf)
    n = lib().orc_ms_ref(lat.N, _p(lat.indptr), _p(lat.indices), _p(costs), len(roi_nodes), _p(roi_nodes), _p(g2r),
                         len(sn), _p(sn), _p(sc), len(tn), _p(tn), _p(tc), _p(out), len(out), ctypes.byref(best))
    if n < 0:
        raise MemoryError("oracle ms_ref failed")
    return (out[:n].tolist() if n > 1 else None), best.value


def ms_canon(lat, costs, seeds, targets, roi_mask=None):
    """Canonical multi-source / multi-sink search (the GPU contract): returns
    (path or None, chosen target entry index or -1, best total (float64),
    dist f32[N], pred i32[N], expanded, traversed)."""
    sn = np.ascontiguousarray([a for a, _ in seeds], dtype=np.int32)
    sc = np.ascontiguousarray([b for _, b in seeds], dtype=np.float32)
    tn = np.ascontiguousarray([a for a, _ in targets], dtype=np.int32)
    tc = np.ascontiguousarray([b for _, b in targets], dtype=np.float32)
    costs = np.ascontiguousarray(costs, dtype=np.float32)
    rm = None if roi_mask is None else np.ascontiguousarray(roi_mask, dtype=np.uint8)
    dist = np.empty(lat.N, dtype=np.float32)
    pred = np.empty(lat.N, dtype=np.int32)
    best = ctypes.c_double(np.inf)
    ex, tr = ctypes.c_int64(0), ctypes.c_int64(0)
    k = lib().orc_ms_canon(lat.N, _p(lat.indptr), _p(lat.indices), _p(costs), _p(rm), len(sn), _p(sn), _p(sc),
                           len(tn), _p(tn), _p(tc), _p(dist), _p(pred), ctypes.byref(best), ctypes.byref(ex),
                           ctypes.byref(tr))
    if k < -1:
        raise MemoryError("oracle ms_canon failed")
    path = None
    if k >= 0:
        cur, p = int(tn[k]), []
        while cur != -1:
            p.append(cur)
            cur = int(pred[cur])
        p.reverse()
        path = p if len(p) > 1 else None
    return path, int(k), best.value, dist, pred, ex.value, tr.value


def steiner_canon(lat, costs, pins, roi_mask=None):
    """Grow a tree over `pins` (node ids): start at pins[0]; each step runs
    ms_canon with every tree node seeded at cost 0 and every unconnected pin
    as a target (cost 0), attaches the chosen pin by its path and adds the
    path's nodes to the tree.  Returns (branches, order, unreached, nodes) as
    orthoroute_amd.steiner.grow_trees does."""
    pins = [int(v) for v in pins]
    tree, in_tree = pins[:1], set(pins[:1])
    left = list(range(1, len(pins)))
    branches, order, unreached = [], [0] if pins else [], []
    while left:
        path, k, _, *_ = ms_canon(lat, costs, [(v, 0.0) for v in tree], [(pins[j], 0.0) for j in left], roi_mask)
        if k < 0:
            unreached = sorted(left)
            break
        j = left.pop(k)
        order.append(j)
        br = path if path is not None else [pins[j]]
        branches.append(br)
        for v in br:
            if v not in in_tree:
                in_tree.add(v)
                tree.append(v)
    return branches, order, unreached, tree


def path_cost(lat, costs, path):
    """float64 sum over the first-match CSR edge of each hop (UPF:5463-5476)."""
    if not path or len(path) < 2:
        return None
    tot = 0.0
    for u, v in zip(path, path[1:]):
        s, e = lat.indptr[u], lat.indptr[u + 1]
        hit = np.nonzero(lat.indices[s:e] == v)[0]
        if len(hit) == 0:
            return None
        tot += float(costs[s + hit[0]])
    return tot


def path_min_cost(lat, costs, path):
    """float64 sum over the cheapest parallel edge of each hop (what a
    shortest-path search minimises)."""
    if not path or len(path) < 2:
        return None
    tot = 0.0
    for u, v in zip(path, path[1:]):
        s, e = lat.indptr[u], lat.indptr[u + 1]
        m = lat.indices[s:e] == v
        if not m.any():
            return None
        tot += float(np.min(costs[s:e][m]))
    return tot
