"""GPU parity of the multi-source / multi-sink search (orgpu_route_ms) with
the oracle's canonical search (oracle.ms_canon, pinned against
SimpleDijkstra.find_path_multisource_multisink, UPF:1889-1989, by the
multisource golden vectors): same path, same chosen target, same total."""
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


def pack_rows(masks, N):
    words = (N + 31) // 32
    out = np.zeros((len(masks), words), np.uint32)
    for i, m in enumerate(masks):
        b = np.packbits(m.astype(np.uint8), bitorder="little")
        b = np.concatenate([b, np.zeros(words * 4 - len(b), np.uint8)])
        out[i] = b.view(np.uint32)
    return out


def random_costs(o, rng):
    c = (o.costs * (0.2 + 3 * rng.random(o.E))).astype(np.float32)
    c[(o.edge_kind == 1) & (rng.random(o.E) < 0.1)] = np.inf
    return c


def random_requests(o, rng, n, n_roi):
    reqs = []
    for i in range(n):
        ns = int(rng.integers(0, 7))
        nt = int(rng.integers(0, 9))
        seeds = [(int(rng.integers(0, o.N)), float(np.float32(rng.random() * 4 * (rng.random() < 0.6))))
                 for _ in range(ns)]
        targets = [(int(rng.integers(0, o.N)), float(np.float32(rng.random() * 3 * (rng.random() < 0.5))))
                   for _ in range(nt)]
        if ns and nt and rng.random() < 0.2:
            targets.append((seeds[0][0], 0.5))            # a target that is also a seed
        if ns and rng.random() < 0.2:
            seeds.append((seeds[-1][0], seeds[-1][1] * 0.5))   # the same node seeded twice
        roi = int(rng.integers(-1, n_roi)) if n_roi else -1
        reqs.append((seeds, targets, i, roi, 0))
    return reqs


def check_against_oracle(o, c, reqs, masks, got):
    for i, (seeds, targets, _, roi, _) in enumerate(reqs):
        m = masks[roi] if roi >= 0 else None
        path, k, best, *_ = oracle.ms_canon(o, c, seeds, targets, m)
        if k < 0:
            assert got[i] is None, f"request {i}: GPU found a path, oracle none"
            continue
        assert got[i] is not None, f"request {i}: oracle reached target entry {k}, GPU none"
        nodes, edges, total, choice = got[i]
        assert choice == k, f"request {i}: target entry {choice} vs oracle {k}"
        assert total == best, f"request {i}: total {total!r} vs oracle {best!r}"
        exp = path if path is not None else [targets[k][0]]
        assert nodes.tolist() == exp, f"request {i}: path differs"
        for t, (u, v) in enumerate(zip(exp, exp[1:])):
            s, e = o.indptr[u], o.indptr[u + 1]
            assert edges[t] == s + int(np.nonzero(o.indices[s:e] == v)[0][0])


@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
@pytest.mark.parametrize("slots", [1, 5])
@pytest.mark.parametrize("case", [("reference", 18, 14, 6, 0), ("adjacent", 31, 23, 4, 1), ("reference", 40, 36, 8, 2)])
def test_route_ms_matches_oracle(ctx, case, slots, goal):
    import torch
    from orthoroute_amd.device import SearchParams, route_ms
    policy, X, Y, L, seed = case
    say(f"[route_ms] {policy} {X}x{Y}x{L} slots={slots} goal={goal} seed={seed}")
    rng = np.random.default_rng(100 + seed)
    o = oracle.Lattice(X, Y, L, policy)
    g = ctx.build_lattice(X, Y, L, policy)
    c = random_costs(o, rng)
    cost_d = torch.from_numpy(c).to(ctx.device)
    # ROI rows: two boxes over all layers, one random half of the nodes
    xs = np.arange(o.N) % X
    ys = (np.arange(o.N) // X) % Y
    masks = [(xs < X // 2) & (ys < Y // 2 + 2), (xs >= X // 3) & (ys >= Y // 4), rng.random(o.N) < 0.5]
    rows = torch.from_numpy(pack_rows(masks, o.N)).to(ctx.device)
    reqs = random_requests(o, rng, 14, len(masks))
    params = SearchParams(slots=slots, delta=1.5, n_buckets=8, h_xy=0.05 if goal else 0.0, h_z=0.1 if goal else 0.0)
    got, st = route_ms(ctx, g, cost_d, reqs, params, roi_bitmaps=rows)
    say(f"  rounds={st['rounds']} items={st['items']}")
    check_against_oracle(o, c, reqs, masks, got)


def test_route_ms_calls_in_a_row_and_after_two_pin_searches(ctx):
    """Multi-source calls interleaved with two-pin calls on the same handle:
    target marks and keys of one call never leak into the next."""
    import torch
    from orthoroute_amd.device import SearchParams, route, route_ms
    say("[route_ms] interleaved with route")
    rng = np.random.default_rng(7)
    X, Y, L = 24, 20, 6
    o = oracle.Lattice(X, Y, L, "reference")
    g = ctx.build_lattice(X, Y, L, "reference")
    c = random_costs(o, rng)
    cost_d = torch.from_numpy(c).to(ctx.device)
    params = SearchParams(slots=3, delta=1.0, n_buckets=4)
    for rep in range(3):
        reqs = random_requests(o, rng, 9, 0)
        got, _ = route_ms(ctx, g, cost_d, reqs, params)
        check_against_oracle(o, c, reqs, [], got)
        pairs = [(int(a), int(b), i, -1, 0) for i, (a, b) in enumerate(rng.integers(0, o.N, (6, 2)))]
        res, _ = route(ctx, g, cost_d, pairs, params)
        for (s, d, *_), r in zip(pairs, res):
            path = oracle.sssp_canon(o, c, s, d)[0]
            assert (None if r is None else r[0].tolist()) == (path if path and len(path) > 1 else None)


def test_route_ms_rejects_bad_input(ctx):
    import torch
    from orthoroute_amd._abi import OrgpuError
    from orthoroute_amd.device import SearchParams, route_ms
    X, Y, L = 8, 6, 4
    g = ctx.build_lattice(X, Y, L, "reference")
    o = oracle.Lattice(X, Y, L, "reference")
    cost_d = torch.from_numpy(o.costs.copy()).to(ctx.device)
    params = SearchParams(slots=2)
    with pytest.raises(OrgpuError):
        route_ms(ctx, g, cost_d, [([(o.N, 0.0)], [(0, 0.0)], 0, -1, 0)], params)
    with pytest.raises(OrgpuError):
        route_ms(ctx, g, cost_d, [([(0, -1.0)], [(1, 0.0)], 0, -1, 0)], params)
    with pytest.raises(OrgpuError):
        route_ms(ctx, g, cost_d, [([(0, 0.0)], [(1, 0.0)], 0, 3, 0)], params)
    got, _ = route_ms(ctx, g, cost_d, [], params)
    assert got == []
