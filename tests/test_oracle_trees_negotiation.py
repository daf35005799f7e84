"""Multi-terminal nets as trees inside the negotiation (this build's
definition: the reference's multi-terminal router never runs) on the oracle:
every committed tree spans all its pins through its branches, the usage the
negotiation keeps equals the committed trees' and paths' edges, a tree net
of two pins is the two-pin net, and without multi_terminal_enabled the
tree nets are not routed (as in the reference, UPF:2548-2585)."""
import numpy as np
import pytest

import oracle
from oracle.negotiation import Config, OracleRouter


def board(seed, X=18, Y=16, n_two=14, n_tree=6):
    rng = np.random.default_rng(seed)
    cells = rng.permutation(X * Y)
    reqs, k = [], 0
    for i in range(n_two):
        reqs.append((f"N{i:02d}", int(cells[k]), int(cells[k + 1])))
        k += 2
    for i in range(n_tree):
        m = int(rng.integers(3, 6))
        reqs.append((f"T{i:02d}", [int(c) for c in cells[k:k + m]]))
        k += m
    return reqs


def run(reqs, X=18, Y=16, L=4, trees=True, cap=12):
    cfg = Config(grid_pitch=1.0, via_cost=3.0, multi_terminal_enabled=trees)
    r = OracleRouter(X, Y, L, "reference", pitch=1.0, via_cost=3.0, sssp="canon", config=cfg)
    rec = []
    res = r.route_multiple_nets(reqs, max_iterations=cap,
                                iteration_cb=lambda it, t, v, s, c: rec.append((it, s, c, r.history.copy())))
    return r, res, rec


def spans(lat, branches, pins):
    """The branches' edges connect every pin (union-find over the hops)."""
    parent = {}

    def find(a):
        while parent.setdefault(a, a) != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    ip, ix = lat.indptr, lat.indices
    for b in branches:
        for u, v in zip(b, b[1:]):
            assert v in ix[ip[u]:ip[u + 1]].tolist(), "a branch hop is not an edge"
            parent[find(u)] = find(v)
    roots = {find(p) for p in pins}
    return len(roots) == 1


@pytest.mark.parametrize("seed", [1, 2])
def test_trees_in_negotiation_are_spanning_and_accounted(seed):
    reqs = board(seed)
    r, res, rec = run(reqs)
    assert len(rec) >= 1
    trees = [q for q in reqs if len(q) == 2]
    routed = 0
    for name, pins in trees:
        p = res["paths"].get(name)
        if not p:
            continue
        routed += 1
        br = r.tree_branches[name]
        assert spans(r.lat, br, pins)
        assert set(pins) <= set(p)
        # first-seen order: the first pin, then each branch's new nodes
        assert p[0] == pins[0] and len(set(p)) == len(p)
        assert sorted(r.net_to_edges[name]) == sorted(e for b in br for e in r.path_to_edges(b))
    # (rip-up can leave a net without a path at the cap)
    assert routed >= len(trees) // 2
    # the usage the negotiation keeps is the committed nets' edges
    use = {}
    for name, e in r.net_to_edges.items():
        if res["paths"].get(name):
            for x in e:
                use[x] = use.get(x, 0) + 1
    assert use == {k: v for k, v in r.canonical.items() if v}


def test_tree_net_of_two_pins_is_the_two_pin_net():
    reqs = board(3, n_tree=0)
    as_lists = [(n, [s, d]) for n, s, d in reqs]
    r1, res1, rec1 = run(reqs)
    r2, res2, rec2 = run(as_lists)
    assert [x[:3] for x in rec1] == [x[:3] for x in rec2]
    for a, b in zip(rec1, rec2):
        assert np.array_equal(a[3].view(np.uint32), b[3].view(np.uint32))
    assert {k: list(v) for k, v in res1["paths"].items()} == {k: list(v) for k, v in res2["paths"].items()}


def test_tree_nets_not_routed_without_multi_terminal():
    reqs = board(4)
    r, res, rec = run(reqs, trees=False)
    assert not any(n.startswith("T") for n in res["paths"])
    r2, res2, rec2 = run([q for q in reqs if len(q) == 3], trees=False)
    assert [x[:3] for x in rec] == [x[:3] for x in rec2]


def test_tree_growth_matches_steiner_oracle_on_first_iteration():
    """Iteration 1 grows every tree on the first cost snapshot exactly as
    oracle.steiner_canon does on those costs."""
    reqs = board(5)
    cfg = Config(grid_pitch=1.0, via_cost=3.0, multi_terminal_enabled=True)
    r = OracleRouter(18, 16, 4, "reference", pitch=1.0, via_cost=3.0, sssp="canon", config=cfg)
    snap = {}
    orig = r.route_all

    def wrap(tasks, all_tasks, pres_fac, iteration=0):
        if not snap:
            snap["c"] = r.total_cost.copy()
        return orig(tasks, all_tasks, pres_fac, iteration)
    r.route_all = wrap
    res = r.route_multiple_nets(reqs, max_iterations=1)
    for name, pins in [q for q in reqs if len(q) == 2]:
        _, _, unreached, nodes = oracle.steiner_canon(r.lat, snap["c"], pins)
        assert (list(res["paths"].get(name) or []) == list(nodes)) or unreached
