"""Edge cases of the search entry points on the GPU, each against the oracle:
an empty request list; a request whose source is its destination (no path:
the reference returns a path only when it has more than one node,
UPF:1880-1887); and a multi-source search with more targets than early
completion scans (such a search runs until its frontier is empty, and must
still pick the reference's target, UPF:1934-1966)."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from orthoroute_amd.device import Context
    return Context(0)


def test_empty_and_self_requests(ctx):
    import torch
    from orthoroute_amd.device import SearchParams, route
    X, Y, L = 12, 10, 4
    say(f"[edge] empty and src==dst requests on {X}x{Y}x{L}")
    lat = ctx.build_lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    o = oracle.Lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    cost = (o.costs * np.float32(0.3)).astype(np.float32)
    cost_d = torch.from_numpy(cost).to(ctx.device)
    params = SearchParams(slots=2, delta=1.0, n_buckets=8)
    res, _ = route(ctx, lat, cost_d, [], params)
    assert res == []
    far = X * Y - 1
    reqs = [(5, 5, 0, -1, 0), (5, far, 1, -1, 0), (far, far, 2, -1, 0), (far, 5, 3, -1, 0)]
    res, _ = route(ctx, lat, cost_d, reqs, params)
    assert res[0] is None and res[2] is None
    for i in (1, 3):
        s, d = reqs[i][:2]
        path, *_ = oracle.sssp_canon(o, cost, s, d)
        assert res[i] is not None and res[i][0].tolist() == path, f"request {i}"


def test_multisource_with_more_targets_than_early_completion_scans(ctx):
    import torch
    from orthoroute_amd.device import SearchParams, route_ms
    X, Y, L = 20, 16, 4
    say(f"[edge] multi-source with up to 300 targets on {X}x{Y}x{L}")
    lat = ctx.build_lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    o = oracle.Lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    rng = np.random.default_rng(17)
    c = (o.costs * (0.2 + 3 * rng.random(o.E))).astype(np.float32)
    cost_d = torch.from_numpy(c).to(ctx.device)
    reqs = []
    for i, nt in enumerate((300, 257, 256, 9)):
        nodes = rng.choice(o.N, 3 + nt, replace=False)
        seeds = [(int(v), float(np.float32(rng.random() * 2))) for v in nodes[:3]]
        targets = [(int(v), float(np.float32(rng.random() * 3))) for v in nodes[3:]]
        reqs.append((seeds, targets, i, -1, 0))
    params = SearchParams(slots=3, delta=0.5, n_buckets=8)
    got, _ = route_ms(ctx, lat, cost_d, reqs, params)
    for i, (seeds, targets, *_rest) in enumerate(reqs):
        path, k, best, *_ = oracle.ms_canon(o, c, seeds, targets)
        assert k >= 0 and got[i] is not None, f"request {i}"
        nodes, edges, total, choice = got[i]
        assert (choice, total) == (k, best), f"request {i}: target {choice} total {total!r} vs {k} {best!r}"
        exp = path if path is not None else [targets[k][0]]
        assert nodes.tolist() == exp, f"request {i}: path differs"
