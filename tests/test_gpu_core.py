"""GPU parity: lattice builder and batched SSSP vs the CPU oracle (bit-exact)."""
import os
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


def to_np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("policy", ["reference", "adjacent"])
@pytest.mark.parametrize("shape", [(5, 4, 2), (6, 5, 3), (8, 6, 4), (12, 10, 6), (7, 6, 8), (96, 96, 6), (33, 17, 12)])
def test_lattice_bit_exact(ctx, policy, shape):
    X, Y, L = shape
    say(f"[lattice] {policy} {shape}")
    g = ctx.build_lattice(X, Y, L, policy, pitch=1.0, via_cost=3.0)
    o = oracle.Lattice(X, Y, L, policy, pitch=1.0, via_cost=3.0)
    np.testing.assert_array_equal(to_np(g.indptr), o.indptr)
    np.testing.assert_array_equal(to_np(g.indices)[:o.E], o.indices)
    np.testing.assert_array_equal(to_np(g.base_cost)[:o.E].view(np.uint32), o.costs.view(np.uint32))
    np.testing.assert_array_equal(to_np(g.edge_layer)[:o.E], o.edge_layer)
    np.testing.assert_array_equal(to_np(g.edge_kind)[:o.E], o.edge_kind)


def pack_bits(mask):
    b = np.packbits(mask.astype(np.uint8), bitorder="little")
    pad = (-len(b)) % 4
    return np.concatenate([b, np.zeros(pad, np.uint8)]).view(np.uint32)


def make_costs(o, rng, kind="jitter"):
    if kind == "congested":   # late-iteration-like: heavy-tailed present/history terms
        c = (o.costs * np.float32(0.3)).astype(np.float32)
        c += (np.arange(o.E, dtype=np.float32) % 9973) * np.float32(1e-6)
        hot = rng.random(o.E) < 0.2
        c[hot] += (rng.pareto(1.2, hot.sum()) * 20.0).astype(np.float32)
        blk = (o.edge_kind == 1) & (rng.random(o.E) < 0.05)
        c[blk] = np.inf
        return c
    if kind == "jitter":   # reference-like: 0.3*base + (e % 9973)*1e-6 + congestion
        c = (o.costs * np.float32(0.3)).astype(np.float32)
        c += (np.arange(o.E, dtype=np.float32) % 9973) * np.float32(1e-6)
        hot = rng.random(o.E) < 0.05
        c[hot] += rng.integers(1, 20, hot.sum()).astype(np.float32) * np.float32(0.7)
    elif kind == "ties":
        c = (o.costs * np.float32(0.5) * rng.integers(1, 3, o.E).astype(np.float32)).astype(np.float32)
    else:
        c = (o.costs * (0.2 + 3 * rng.random(o.E))).astype(np.float32)
    blk = (o.edge_kind == 1) & (rng.random(o.E) < 0.15)
    c[blk] = np.inf
    return c


def run_case(ctx, X, Y, L, policy, slots, n_nets, seed, cost_kind, roi_frac=0.5, keepout=False, dn=None,
             goal=False, graph=False, engine=0, list_cap=0, lookahead=0, waves=0):
    import torch
    from orthoroute_amd.device import SearchParams, cost_bounds, route
    say(f"[route] {policy} {X}x{Y}x{L} slots={slots} nets={n_nets} seed={seed} costs={cost_kind} keepout={keepout}")
    rng = np.random.default_rng(seed)
    o = oracle.Lattice(X, Y, L, policy)
    g = ctx.build_lattice(X, Y, L, policy)
    c = make_costs(o, rng, cost_kind)
    cost_d = torch.from_numpy(c).to(ctx.device)
    owner = None
    if keepout:
        owner = np.full(o.N, -1, np.int32)
        sel = rng.random(o.N) < 0.08
        owner[sel] = rng.integers(0, 6, sel.sum())
    reqs, rois, masks = [], [], []
    for i in range(n_nets):
        s, d = rng.integers(0, o.N, 2)
        while s == d:
            d = rng.integers(0, o.N)
        roi_idx = -1
        mask = None
        if rng.random() < roi_frac:
            roi = oracle.roi_bfs(o, int(s), int(d), max_nodes=int(rng.choice([300, 2000, 50_000])))
            mask = np.zeros(o.N, np.uint8); mask[roi] = 1; mask[s] = 1; mask[d] = 1
            roi_idx = len(rois)
            rois.append(pack_bits(mask))
        masks.append(mask)
        reqs.append((int(s), int(d), int(i % 6), roi_idx, int(keepout)))
    roi_d = torch.from_numpy(np.stack(rois)).to(ctx.device) if rois else None
    own_d = torch.from_numpy(owner).to(ctx.device) if owner is not None else None
    if dn is None:
        dn = (float(rng.choice([0.3, 1.0, 4.0])), int(rng.choice([1, 4, 8])))
    hx, hz = cost_bounds(ctx, g, cost_d) if goal else (0.0, 0.0)
    say(f"[route]   delta={dn[0]} n_buckets={dn[1]} h=({hx}, {hz}) graph={graph} engine={engine} waves={waves}")
    p = SearchParams(slots=slots, delta=dn[0], n_buckets=dn[1], rounds_per_poll=16, h_xy=hx, h_z=hz,
                     graph_rounds=graph, engine=engine, list_cap=list_cap, lookahead=lookahead, waves=waves)
    res, stats = route(ctx, g, cost_d, reqs, p, roi_bitmaps=roi_d, keepout_owner=own_d)
    for i, (s, d, nid, roi_idx, ko) in enumerate(reqs):
        path, dist, pred, _, _ = oracle.sssp_canon(o, c, s, d, roi_mask=masks[i],
                                                   keepout_owner=owner if ko else None, net_id=nid)
        got = None if res[i] is None else res[i][0].tolist()
        assert got == path, f"net {i} ({s}->{d}) roi={roi_idx}: gpu {got} oracle {path}"
        if path:
            # first-match CSR edge per hop
            exp_e = []
            for u, v in zip(path, path[1:]):
                row = o.indices[o.indptr[u]:o.indptr[u + 1]]
                exp_e.append(int(o.indptr[u] + np.nonzero(row == v)[0][0]))
            assert res[i][1].tolist() == exp_e
            assert np.float32(res[i][2]) == dist[d]
    return stats


@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
@pytest.mark.parametrize("policy", ["reference", "adjacent"])
@pytest.mark.parametrize("slots", [1, 7])
def test_route_matches_oracle(ctx, policy, slots, goal):
    for seed, kind in [(1, "jitter"), (2, "ties"), (3, "random")]:
        run_case(ctx, 24, 20, 6, policy, slots, 24, seed + 10 * slots, kind, goal=goal)


@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
@pytest.mark.parametrize("slots", [1, 7])
def test_route_graph_rounds_match_oracle(ctx, slots, goal):
    """Rounds replayed as a captured HIP graph (re-captured when the bucket
    width or ring size changes between calls on the same handle): the same
    paths as the oracle, call after call."""
    for seed, kind, dn in [(4, "jitter", (1.0, 8)), (5, "ties", (0.3, 4)), (6, "random", (1.0, 8))]:
        run_case(ctx, 24, 20, 6, "reference", slots, 24, seed + 10 * slots, kind, dn=dn, goal=goal, graph=True)


@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
def test_route_keepout_matches_oracle(ctx, goal):
    run_case(ctx, 20, 18, 6, "reference", 5, 20, 77, "jitter", keepout=True, goal=goal)
    run_case(ctx, 20, 18, 6, "adjacent", 1, 10, 78, "random", keepout=True, goal=goal)


@pytest.mark.parametrize("engine", [0, 1])
def test_route_keepout_saturation_fails_loudly(ctx, engine):
    """The key counts keepout edges up to 255: a search whose best path
    crosses that many (every node of the adjacent lattice owned by another
    net, a destination 260 planar steps away) raises instead of returning a
    path ordered by cost alone; the same request without keepouts routes, and
    200 keepout edges route and equal the oracle.  Both search engines."""
    import torch
    from orthoroute_amd._abi import OrgpuError
    from orthoroute_amd.device import SearchParams, route
    X, Y, L = 300, 3, 2
    g = ctx.build_lattice(X, Y, L, "adjacent")
    o = oracle.Lattice(X, Y, L, "adjacent")
    cost_d = torch.from_numpy(o.costs.astype(np.float32)).to(ctx.device)
    own_d = torch.full((o.N,), 5, dtype=torch.int32, device=ctx.device)
    p = SearchParams(slots=1, delta=1.0, n_buckets=8, engine=engine)
    src, dst = X + 2, X + 262         # row 1 of layer 0, 260 steps apart
    res, _ = route(ctx, g, cost_d, [(src, dst, 1, -1, 0)], p, keepout_owner=own_d)
    assert res[0] is not None and len(res[0][0]) >= 261
    with pytest.raises(OrgpuError, match="255 or more keepout edges"):
        route(ctx, g, cost_d, [(src, dst, 1, -1, 1)], p, keepout_owner=own_d)
    # 200 keepout edges stay below the limit and equal the oracle
    res, _ = route(ctx, g, cost_d, [(src, src + 200, 1, -1, 1)], p, keepout_owner=own_d)
    owner = np.full(o.N, 5, np.int32)
    path = oracle.sssp_canon(o, o.costs.astype(np.float32), src, src + 200, keepout_owner=owner, net_id=1)[0]
    assert res[0][0].tolist() == path


def test_cost_bounds(ctx):
    # the cheapest planar edge and the cheapest via edge per layer of span,
    # recomputed by brute force over the oracle lattice
    import torch
    from orthoroute_amd.device import cost_bounds
    for policy, seed in [("reference", 31), ("adjacent", 32)]:
        X, Y, L = 18, 15, 7
        rng = np.random.default_rng(seed)
        o = oracle.Lattice(X, Y, L, policy)
        g = ctx.build_lattice(X, Y, L, policy)
        c = make_costs(o, rng, "random")
        hx, hz = cost_bounds(ctx, g, torch.from_numpy(c).to(ctx.device))
        plane = X * Y
        src = np.repeat(np.arange(o.N), np.diff(o.indptr))
        span = np.abs(o.indices // plane - src // plane)
        fin = np.isfinite(c)
        exp_xy = c[(o.edge_kind == 0) & fin].min()
        vs = (o.edge_kind == 1) & fin
        exp_z = (c[vs] / span[vs].astype(np.float32)).min()
        shrink = lambda v: float(np.float32(np.float64(v) * (1.0 - 2.0 ** -10)))
        assert hx == shrink(exp_xy) and hz == shrink(exp_z), (policy, hx, hz, exp_xy, exp_z)
        # planar cost scale: the binary-exponent bin holding the median
        _, _, med = cost_bounds(ctx, g, torch.from_numpy(c).to(ctx.device), with_scale=True)
        pl = c[(o.edge_kind == 0) & fin].astype(np.float64)
        ex = np.where(pl > 0, np.frexp(pl)[1] - 1 + 32, 0).clip(0, 63)
        hist = np.bincount(ex, minlength=64)
        k = int(np.searchsorted(np.cumsum(hist), (hist.sum() + 1) // 2))
        assert med == 2.0 ** (k - 31), (policy, med, k)


def test_route_unreachable_and_refill(ctx):
    # B.Cu (top layer) of the reference lattice has no edges: unreachable
    # destinations must finish as "no path" without blocking refills
    stats = run_case(ctx, 16, 14, 6, "reference", 3, 30, 5, "jitter", roi_frac=0.0)
    assert stats["rounds"] > 0


@pytest.mark.parametrize("nb", [1, 2])
def test_route_overflow_heavy(ctx, nb):
    # narrow buckets and a short ring: most keys land in the overflow queue and
    # many improve again while queued there (lower overflow bucket, item
    # already present); the bucket jump out of an empty ring must not pass them
    for seed, kind in [(21, "random"), (22, "jitter")]:
        run_case(ctx, 24, 20, 6, "adjacent", 4, 16, seed, kind, roi_frac=0.0, dn=(0.05, nb))
        run_case(ctx, 24, 20, 6, "adjacent", 4, 16, seed, kind, roi_frac=0.0, dn=(0.05, nb), goal=True)


@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
@pytest.mark.parametrize("delta", [0.3, 6.0, 60.0, 600.0])
def test_route_congested_costs_any_delta(ctx, delta, goal):
    # paths must not depend on the bucket width, also when costs span four
    # orders of magnitude (late negotiation iterations)
    for policy, seed in [("adjacent", 41), ("reference", 42)]:
        run_case(ctx, 30, 26, 6, policy, 16, 40, seed, "congested", roi_frac=0.3, dn=(delta, 8), goal=goal)
