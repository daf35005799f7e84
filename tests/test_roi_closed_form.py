"""Closed-form ROI on the routing lattice, checked on the CPU.

The ROI kernel for the router's own lattice does not traverse edges: the hop
distance from a node (x0, y0, z0) to (x, y, z) on a lattice whose layers are
full grids with vias at every cell is

    dx + dy + T[z0][need][z],   need = (dx > 0) | (dy > 0) << 1,

with T the fewest via hops from layer z0 to layer z that visit a planar
layer allowing x moves when dx > 0 and one allowing y moves when dy > 0
(a breadth-first search over (layer, covered directions) states).  The
bidirectional BFS of the reference (UPF:1636-1734) then marks
ball_src(ks) | ball_dst(kd), where the step at which the waves meet follows
from the src-dst distance D (src step j meets iff D = 2j - 1, dst step j iff
D = 2j) and the early stop from the ball sizes.

This file checks both facts against plain BFS: the distance formula against
scipy's unweighted shortest paths from every node of small lattices (both
via policies, 1 to 8 layers), and the level rules against the oracle's
line-by-line BFS (its marks, depth and meet flag) on random nets."""
import ctypes

import numpy as np
import pytest

import oracle
from orthoroute_amd.lattice import DIR_H, DIR_HV, DIR_V, lattice_spec

scipy_sparse = pytest.importorskip("scipy.sparse")
csgraph = pytest.importorskip("scipy.sparse.csgraph")

UNREACH = 255


def layer_table(L, policy):
    """T[z0, need, z]: fewest via hops from layer z0 to layer z covering the
    directions in `need` (bit 0: x moves, bit 1: y moves); 255 = none."""
    planar, dirs, pairs = lattice_spec(L, policy)
    cov = [0] * L
    for z in range(L):
        if planar[z]:
            cov[z] = {DIR_H: 1, DIR_V: 2, DIR_HV: 3}[dirs[z]]
    adj = [[] for _ in range(L)]
    for a, b in pairs:          # each pair adds a via in both directions
        adj[a].append(b)
        adj[b].append(a)
    T = np.full((L, 4, L), UNREACH, dtype=np.int64)
    for z0 in range(L):
        dist = {(z0, cov[z0]): 0}
        frontier = [(z0, cov[z0])]
        while frontier:
            nxt = []
            for z, m in frontier:
                for b in adj[z]:
                    s = (b, m | cov[b])
                    if s not in dist:
                        dist[s] = dist[(z, m)] + 1
                        nxt.append(s)
            frontier = nxt
        for (z, m), d in dist.items():
            for need in range(4):
                if need & ~m == 0:
                    T[z0, need, z] = min(T[z0, need, z], d)
    return T


def lattice_dist(X, Y, L, T, s):
    """Hop distance from node s to every node (int64, UNREACH+ = none)."""
    plane = X * Y
    z0, r = divmod(int(s), plane)
    y0, x0 = divmod(r, X)
    n = np.arange(X * Y * L)
    z, r = np.divmod(n, plane)
    y, x = np.divmod(r, X)
    dx, dy = np.abs(x - x0), np.abs(y - y0)
    need = (dx > 0).astype(np.int64) | ((dy > 0).astype(np.int64) << 1)
    t = T[z0, need, z]
    return np.where(t >= UNREACH, np.iinfo(np.int64).max // 4, dx + dy + t)


def bfs_dist(lat, s):
    g = scipy_sparse.csr_matrix((np.ones(lat.E), lat.indices, lat.indptr), shape=(lat.N, lat.N))
    d = csgraph.shortest_path(g, unweighted=True, indices=[int(s)])[0]
    return d


@pytest.mark.parametrize("policy", ["reference", "adjacent"])
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 6, 8])
def test_distance_formula_equals_bfs(policy, L):
    for X, Y in [(5, 4), (7, 6)]:
        lat = oracle.Lattice(X, Y, L, policy)
        T = layer_table(L, policy)
        for s in range(lat.N):
            ref = bfs_dist(lat, s)
            got = lattice_dist(X, Y, L, T, s).astype(np.float64)
            got[got > 1e15] = np.inf
            np.testing.assert_array_equal(got, ref, err_msg=f"{policy} {X}x{Y}x{L} from {s}")


def roi_levels(hs, ht, D, same, max_depth, max_nodes):
    """The reference's level loop (UPF:1653-1700) on ball sizes: returns
    (ks, kd, depth, met).  hs[j] / ht[j]: nodes at distance exactly j from
    src / dst; D: src-dst distance (None if unreachable); same: src == dst."""
    def h(a, j):
        return int(a[j]) if j < len(a) else 0
    stop = max_nodes * 1.5
    na, nb = 1, (0 if same else 1)
    fa, fb = 1, 1
    depth, met, seen = 0, 0, (1 if same else 2)
    ks = kd = 0
    while depth < max_depth and (fa > 0 or fb > 0) and not met:
        j = depth + 1
        if fa > 0:
            if not same and D is not None and D == 2 * j - 1:
                met = 1
            fa = h(hs, j)
            seen += fa
        ks = j
        if not met:
            if same:
                # both waves start at the node: the dst wave's first step
                # touches the src wave's marks
                if h(hs, 1) > 0:
                    met = 1
                fb = 0
            elif fb > 0:
                if D is not None and D == 2 * j:
                    met = 1
                fb = h(ht, j)
                seen += fb
            kd = j
        depth = j
        if seen > stop:
            break
    return ks, kd, depth, met


def _oracle_bfs(lat, s, d, depth, max_nodes):
    out = np.empty(lat.N, dtype=np.int32)
    seen = np.empty(lat.N, dtype=np.uint8)
    dd, mm = ctypes.c_int(0), ctypes.c_int(0)
    R = oracle.lib().orc_roi_bfs(lat.N, oracle._p(lat.indptr), oracle._p(lat.indices), int(s), int(d), int(depth),
                                 int(max_nodes), None, 0, 0, oracle._p(out), oracle._p(seen), ctypes.byref(dd),
                                 ctypes.byref(mm))
    assert R >= 0
    return seen != 0, dd.value, mm.value


@pytest.mark.parametrize("policy", ["reference", "adjacent"])
@pytest.mark.parametrize("shape", [(20, 16, 4), (24, 20, 6), (30, 9, 3), (12, 12, 2)])
def test_level_rules_equal_oracle_bfs(policy, shape):
    X, Y, L = shape
    lat = oracle.Lattice(X, Y, L, policy)
    T = layer_table(L, policy)
    rng = np.random.default_rng(X * 100 + L)
    cases = []
    for _ in range(60):
        s, d = (int(v) for v in rng.integers(0, lat.N, 2))
        depth = int(rng.choice([0, 1, 2, 3, 5, 8, 20, 150]))
        mx = int(rng.choice([5, 40, 300, 50_000]))
        cases.append((s, d, depth, mx))
    cases += [(7, 7, 150, 50_000), (7, 7, 0, 50_000), (3, lat.N - 1, 150, 50_000)]
    for s, d, depth, mx in cases:
        ds, dd = lattice_dist(X, Y, L, T, s), lattice_dist(X, Y, L, T, d)
        big = np.iinfo(np.int64).max // 4
        hs = np.bincount(ds[ds < big])
        ht = np.bincount(dd[dd < big])
        D = int(ds[d]) if ds[d] < big else None
        ks, kd, dep, met = roi_levels(hs, ht, D, s == d, depth, mx)
        mask = (ds <= ks) | (dd <= kd)
        o_mask, o_depth, o_met = _oracle_bfs(lat, s, d, depth, mx)
        assert (dep, met) == (o_depth, o_met), f"{policy} {shape} {s}->{d} depth {depth} max {mx}"
        np.testing.assert_array_equal(mask, o_mask, err_msg=f"{policy} {shape} {s}->{d} depth {depth} max {mx}")
