"""GPU multi-terminal trees (orthoroute_amd.steiner.grow_trees through
GpuSolver.route_trees and MultiTerminalNetRouter, MT:1005-1160 surface)
against the oracle's tree (oracle.steiner_canon): same branches, attachment
order and unreached pins, bit for bit."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def costs_for(o, rng):
    c = (o.costs * (0.3 + 2 * rng.random(o.E))).astype(np.float32)
    c[(o.edge_kind == 1) & (rng.random(o.E) < 0.1)] = np.inf
    return c


def random_nets(o, rng, n):
    nets = []
    for i in range(n):
        k = int(rng.integers(2, 9))
        pins = [int(v) for v in rng.choice(o.N, k, replace=False)]
        if rng.random() < 0.15:
            pins.append(pins[1])          # a pin listed twice
        nets.append(pins)
    return nets


@pytest.mark.parametrize("shape,policy,roi", [((20, 16, 6), "reference", False), ((26, 22, 4), "adjacent", True)])
def test_route_trees_matches_oracle(shape, policy, roi):
    from orthoroute_amd.device import Context
    from orthoroute_amd.solver import GpuSolver
    X, Y, L = shape
    say(f"[steiner] {shape} {policy} roi={roi}")
    rng = np.random.default_rng(41)
    o = oracle.Lattice(X, Y, L, policy)
    c = costs_for(o, rng)
    s = GpuSolver(o.indptr, o.indices, o.costs, X, Y, L, ctx=Context(0), slots=16)
    nets = random_nets(o, rng, 30)
    masks = None
    if roi:
        xs, ys = np.arange(o.N) % X, (np.arange(o.N) // X) % Y
        masks = [((xs >= (i % 3)) & (ys < Y - (i % 4))).astype(np.uint8) for i in range(len(nets))]
    trees = s.route_trees(c, nets, roi_spec=None if masks is None else [np.nonzero(m)[0] for m in masks])
    for i, (pins, t) in enumerate(zip(nets, trees)):
        branches, order, unreached, nodes = oracle.steiner_canon(o, c, pins, None if masks is None else masks[i])
        assert t["order"] == order, f"net {i}"
        assert t["branches"] == branches, f"net {i}"
        assert t["unreached"] == unreached, f"net {i}"
        assert t["nodes"] == nodes, f"net {i}"
        exp_edges = []
        for br in branches:
            for u, v in zip(br, br[1:]):
                a, b = o.indptr[u], o.indptr[u + 1]
                exp_edges.append(int(a + np.nonzero(o.indices[a:b] == v)[0][0]))
        assert t["edges"].tolist() == exp_edges, f"net {i}"


def test_multi_terminal_router(monkeypatch):
    from orthoroute_amd.algorithms.manhattan.multi_terminal import MultiTerminalNetRouter
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    say("[steiner] MultiTerminalNetRouter")
    X, Y, L = 24, 20, 6
    cfg = PathFinderConfig()
    cfg.batch_slots = 8
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    o = oracle.Lattice(X, Y, L, "reference")
    rng = np.random.default_rng(43)
    nets = [(f"N{i}", pins) for i, pins in enumerate(random_nets(o, rng, 12))]
    nets.append(("GND", [int(v) for v in rng.choice(o.N, 3, replace=False)]))
    res = MultiTerminalNetRouter(r).route_all_nets(nets)
    base = r.graph.dlat.base_cost.cpu().numpy()[:o.E]
    full = 0
    for name, pins in nets:
        seen, uniq = set(), []
        for v in pins:
            if v not in seen:
                seen.add(v)
                uniq.append(v)
        _, _, unreached, nodes = oracle.steiner_canon(o, base, uniq)
        assert res[name] == nodes, name
        full += not unreached
    v = res["_validation"]
    assert v["total_nets"] == len(nets) and v["fully_connected"] == full
