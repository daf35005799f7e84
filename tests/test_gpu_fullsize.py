"""Parity at the benchmark's full sizes: the 600x400x18 backplane (4.32 M
nodes, 16.8 M edges) and the 1000x800x24 stress board (19.2 M nodes, 75.2 M
edges) routed by the batched GPU engine on their iteration-1 cost snapshots
against the oracle's ROI + canonical search for a sample of the same nets,
and cfg3's multi-terminal trees (300x300x12) against the oracle's trees."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("which", ["cfg2", "cfg4"])
def test_fullsize_sample_matches_oracle(which):
    from orthoroute_amd import workloads
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import (PathFinderConfig, PathFinderRouter,
                                                                        ROI_THRESHOLD_STEPS)
    X, Y, L, reqs = workloads.CONFIGS[which]()
    say(f"[fullsize] {which} {X}x{Y}x{L}, {len(reqs)} nets")
    cfg = PathFinderConfig()
    cfg.lattice_policy = "adjacent"
    cfg.batch_slots = 512
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    tasks = r._parse_requests(reqs)
    pres_fac, _, _, _, hmult = r._negotiation_setup(tasks)
    r.iteration = 1
    r._begin_iteration(1, pres_fac, hmult)
    # the first 16 nets of each kind: ROI-limited and whole-graph searches
    short, long_ = [], []
    for name, (s, d) in tasks.items():
        sx, sy, _ = r._coords(s)
        dx, dy, _ = r._coords(d)
        near = abs(dx - sx) + abs(dy - sy) < ROI_THRESHOLD_STEPS
        (short if near else long_).append((s, d, -1, near, 0.0))
    sample = short[:16] + long_[:6] if which == "cfg2" else short[:24] + long_[:2]
    say(f"[fullsize] {len(sample)} sampled searches ({sum(1 for q in sample if q[3])} ROI-limited)")
    got = r._search_batch(sample)
    lat = oracle.Lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    costs = r.accounting.total_cost[:r.accounting.E].cpu().numpy()
    for i, ((s, d, _, near, bonus), g) in enumerate(zip(sample, got)):
        mask = None
        if near:
            mask = np.zeros(lat.N, np.uint8)
            mask[oracle.roi_bfs(lat, s, d, stagnation_bonus=bonus)] = 1
            mask[s] = mask[d] = 1
        path = oracle.sssp_canon(lat, costs, s, d, roi_mask=mask)[0]
        assert (None if g is None else g[0].tolist()) == path, f"net {i} ({s}->{d}, roi={near})"


def test_cfg4_1024_concurrent_searches_match_plain_and_oracle():
    """The stress board with 1,024 searches in flight at once (64-bit work
    items; 19.2 M nodes x 1,024 slots = 19.7 G items): 1,100 of its nets
    batched equal the same nets searched one at a time, and a sample equals
    the oracle's ROI + canonical search."""
    from orthoroute_amd import workloads
    from orthoroute_amd.device import max_slots
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import (PathFinderConfig, PathFinderRouter,
                                                                        ROI_THRESHOLD_STEPS)
    X, Y, L, reqs = workloads.CONFIGS["cfg4"]()
    cfg = PathFinderConfig()
    cfg.lattice_policy = "adjacent"
    cfg.batch_slots = 1024
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    assert max_slots(r.N) >= 1024, "the device does not hold 1,024 searches of this board"
    tasks = r._parse_requests(reqs)
    pres_fac, _, _, _, hmult = r._negotiation_setup(tasks)
    r.iteration = 1
    r._begin_iteration(1, pres_fac, hmult)
    sample = []
    for name, (s, d) in list(tasks.items())[:1100]:
        sx, sy, _ = r._coords(s)
        dx, dy, _ = r._coords(d)
        sample.append((s, d, -1, abs(dx - sx) + abs(dy - sy) < ROI_THRESHOLD_STEPS, 0.0))
    say(f"[fullsize] cfg4 {len(sample)} searches, 1024 at once ({sum(1 for q in sample if q[3])} ROI-limited)")
    batched = r._search_batch(sample)
    assert r.ctx._reserved[1] == 1024
    cfg.batch_slots = 1
    say("[fullsize] cfg4 the same searches one at a time")
    plain = r._search_batch(sample)
    for i, (b, p) in enumerate(zip(batched, plain)):
        assert (None if b is None else b[0].tolist()) == (None if p is None else p[0].tolist()), f"net {i}"
    # about a quarter of the stress board's nets find no path inside their ROI
    assert sum(1 for b in batched if b is not None) > 700
    lat = oracle.Lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    costs = r.accounting.total_cost[:r.accounting.E].cpu().numpy()
    picks = [i for i, q in enumerate(sample) if q[3]][:10] + [i for i, q in enumerate(sample) if not q[3]][:2]
    for i in picks:
        s, d, _, near, bonus = sample[i]
        mask = None
        if near:
            mask = np.zeros(lat.N, np.uint8)
            mask[oracle.roi_bfs(lat, s, d, stagnation_bonus=bonus)] = 1
            mask[s] = mask[d] = 1
        path = oracle.sssp_canon(lat, costs, s, d, roi_mask=mask)[0]
        g = batched[i]
        assert (None if g is None else g[0].tolist()) == path, f"net {i} ({s}->{d}, roi={near})"


def test_cfg3_multi_terminal_sample_matches_oracle():
    from orthoroute_amd import workloads
    from orthoroute_amd.algorithms.manhattan.multi_terminal import MultiTerminalNetRouter
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    X, Y, L, nets = workloads.cfg3_multi()
    say(f"[fullsize] cfg3 multi-terminal {X}x{Y}x{L}")
    cfg = PathFinderConfig()
    cfg.lattice_policy = "adjacent"
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    mtr = MultiTerminalNetRouter(r)
    res = mtr.route_all_nets(nets)
    assert res["_validation"]["total_nets"] == len(nets)
    lat = oracle.Lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    base = r.graph.dlat.base_cost[:lat.E].cpu().numpy()
    for name, pins in nets[:3]:
        branches, order, unreached, nodes = oracle.steiner_canon(lat, base, pins)
        t = mtr.trees[name]
        assert (t["branches"], t["order"], t["unreached"]) == (branches, order, unreached), name
        assert res[name] == nodes
