"""The multi-terminal tree oracle (oracle.steiner_canon) against what the
mathematics fixes: every branch is a shortest path from the tree built so
far to the pin it attaches (scipy multi-source Dijkstra, float64), that pin is
the nearest unconnected one, branches start on the tree and only add new
nodes, and a two-pin net reduces to the canonical two-pin search."""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra as sp_dijkstra

import oracle


def costs_for(o, rng):
    c = (o.costs * (0.3 + 2 * rng.random(o.E))).astype(np.float32)
    c[(o.edge_kind == 1) & (rng.random(o.E) < 0.1)] = np.inf
    return c


def csr_matrix(o, c, mask=None):
    src = np.repeat(np.arange(o.N), np.diff(o.indptr))
    keep = np.isfinite(c)
    if mask is not None:
        keep &= mask[src].astype(bool) & mask[o.indices].astype(bool)
    # duplicate (u, v) entries: keep the cheapest (csr_matrix would sum them)
    key = src[keep].astype(np.int64) * o.N + o.indices[keep]
    w = c[keep].astype(np.float64)
    order = np.lexsort((w, key))
    key, w = key[order], w[order]
    first = np.ones(len(key), bool)
    first[1:] = key[1:] != key[:-1]
    key, w = key[first], w[first]
    return sp.csr_matrix((w, (key // o.N, key % o.N)), shape=(o.N, o.N))


@pytest.mark.parametrize("seed,roi", [(0, False), (1, False), (2, True), (3, True)])
def test_branches_are_shortest_tree_to_pin_paths(seed, roi):
    rng = np.random.default_rng(seed)
    o = oracle.Lattice(14, 12, 5, "reference")
    c = costs_for(o, rng)
    mask = None
    if roi:
        xs, ys = np.arange(o.N) % 14, (np.arange(o.N) // 14) % 12
        mask = ((xs > 1) & (ys < 10)).astype(np.uint8)
    A = csr_matrix(o, c, mask)
    for net in range(6):
        npins = int(rng.integers(3, 9))
        pins = [int(v) for v in rng.choice(o.N, npins, replace=False)]
        if mask is not None:
            pins = [v for v in pins if mask[v]] or pins[:1]
        branches, order, unreached, nodes = oracle.steiner_canon(o, c, pins, mask)
        tree = [pins[0]]
        left = list(range(1, len(pins)))
        for br, j in zip(branches, order[1:]):
            d = sp_dijkstra(A, indices=np.asarray(tree), min_only=True)
            want = min(d[pins[k]] for k in left)
            assert br[0] in tree and br[-1] == pins[j]
            assert not any(v in tree for v in br[1:])
            assert np.isclose(d[pins[j]], want, rtol=1e-5, atol=1e-6)
            cost = oracle.path_min_cost(o, c, br) if len(br) > 1 else 0.0
            assert np.isclose(cost, want, rtol=1e-5, atol=1e-6), (net, j, cost, want)
            left.remove(j)
            tree += [v for v in br if v not in tree]
        d = sp_dijkstra(A, indices=np.asarray(tree), min_only=True)
        assert sorted(left) == unreached
        assert all(not np.isfinite(d[pins[k]]) for k in unreached)
        assert nodes == tree


def test_two_pins_reduce_to_the_two_pin_search():
    rng = np.random.default_rng(9)
    o = oracle.Lattice(16, 12, 6, "reference")
    c = costs_for(o, rng)
    for _ in range(20):
        s, d = (int(v) for v in rng.choice(o.N, 2, replace=False))
        branches, order, unreached, _ = oracle.steiner_canon(o, c, [s, d])
        path = oracle.sssp_canon(o, c, s, d)[0]
        if path is None:
            assert unreached == [1] and branches == []
        else:
            assert branches == [path] and order == [0, 1]
