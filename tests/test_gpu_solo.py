"""GPU parity of the one-wave-per-search engine (orgpu_sssp_params.engine
= 1), with one, two and four wavefronts per search: paths, first-match edge
ids and FP32 distances bit-equal to the canonical oracle (sssp_canon) and to
the bucket-round engine, including ROIs, keepouts, congested costs, capacity
hand-back to the bucket engine and the edge cases of the two-pin search."""
import sys

import numpy as np
import pytest

import oracle
from test_gpu_core import make_costs, pack_bits, run_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from orthoroute_amd.device import Context
    return Context(0)


WAVES = [1, 2, 4]


@pytest.mark.parametrize("waves", WAVES)
@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
@pytest.mark.parametrize("policy", ["reference", "adjacent"])
@pytest.mark.parametrize("slots", [1, 7, 64])
def test_solo_matches_oracle(ctx, policy, slots, goal, waves):
    for seed, kind in [(1, "jitter"), (2, "ties"), (3, "random"), (4, "congested")]:
        st = run_case(ctx, 24, 20, 6, policy, slots, 24, seed + 10 * slots, kind, goal=goal, engine=1,
                      waves=waves)
        assert st["retried"] == 0


@pytest.mark.parametrize("waves", WAVES)
@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
@pytest.mark.parametrize("policy", ["reference", "adjacent"])
def test_solo_keepout_matches_oracle(ctx, goal, waves, policy):
    """Owner-aware keepouts: on the reference via rules most nodes have more
    than six edges (the launch reads CSR rows), on adjacent-layer vias every
    node has an adjacency record (planar edges marked in its record)."""
    for seed, kind in [(21, "jitter"), (22, "ties"), (23, "congested")]:
        run_case(ctx, 20, 16, 6, policy, 16, 30, seed, kind, keepout=True, goal=goal, engine=1, waves=waves)


@pytest.mark.parametrize("waves", [1, 4])
@pytest.mark.parametrize("lookahead", [-1, 0, 15], ids=["no-lookahead", "default", "lookahead15"])
@pytest.mark.parametrize("dn", [(0.05, 1), (0.3, 8), (4.0, 32), (100.0, 8)], ids=["narrow", "mid", "wide", "one"])
def test_solo_bucket_widths(ctx, dn, lookahead, waves):
    """Bucket widths from far below one edge cost (many radix-list moves) to
    wider than every path (one bucket), with and without early expansion of
    the next buckets' entries: the same paths."""
    run_case(ctx, 30, 24, 6, "adjacent", 16, 40, 31, "congested", dn=dn, goal=True, engine=1, lookahead=lookahead,
             waves=waves)
    run_case(ctx, 30, 24, 6, "adjacent", 16, 40, 32, "jitter", dn=dn, goal=False, engine=1, lookahead=lookahead,
             waves=waves)


@pytest.mark.parametrize("waves", WAVES)
def test_solo_capacity_hands_back(ctx, waves):
    """Bucket-list storage of two 256-entry chunks per search runs out on most
    searches of this board: those searches are routed by the bucket engine and
    every path still equals the oracle's."""
    st = run_case(ctx, 40, 32, 6, "adjacent", 16, 32, 41, "jitter", roi_frac=0.0, dn=(0.3, 8), goal=False,
                  engine=1, list_cap=1, waves=waves)
    say(f"[solo] waves {waves}: retried {st['retried']} of 32")
    assert st["retried"] > 0


def test_solo_equals_bucket_engine(ctx):
    """A larger board, 400 nets with ROIs: both engines return the same
    nodes, edge ids and distances."""
    import torch
    from orthoroute_amd.device import SearchParams, cost_bounds, route_flat
    X, Y, L = 120, 96, 8
    rng = np.random.default_rng(51)
    o = oracle.Lattice(X, Y, L, "adjacent")
    g = ctx.build_lattice(X, Y, L, "adjacent")
    c = make_costs(o, rng, "jitter")
    cost_d = torch.from_numpy(c).to(ctx.device)
    reqs, rois = [], []
    plane = X * Y
    for i in range(400):
        s, d = (int(v) for v in rng.integers(0, plane, 2))
        roi_idx = -1
        if i % 2 == 0:
            roi = oracle.roi_bfs(o, s, d, max_nodes=20_000)
            mask = np.zeros(o.N, np.uint8); mask[roi] = 1; mask[s] = 1; mask[d] = 1
            roi_idx = len(rois)
            rois.append(pack_bits(mask))
        reqs.append((s, d, i % 5, roi_idx, 0))
    roi_d = torch.from_numpy(np.stack(rois)).to(ctx.device)
    hx, hz = cost_bounds(ctx, g, cost_d)
    out = []
    for eng, waves in ((0, 0), (1, 1), (1, 2), (1, 4)):
        p = SearchParams(slots=128, delta=1.0, n_buckets=8, h_xy=hx, h_z=hz, engine=eng, waves=waves)
        nodes, edges, offs, dist, st = route_flat(ctx, g, cost_d, reqs, p, roi_bitmaps=roi_d)
        say(f"[solo] engine {eng} waves {waves}: {st['kernel_ms']:.2f} ms, {st['items']} items, "
            f"{st['edges']} edges, retried {st['retried']}")
        out.append((nodes, edges, offs, dist))
    for o2 in out[1:]:
        for a, b in zip(out[0], o2):
            np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("waves", [1, 4])
def test_solo_edge_cases(ctx, waves):
    """Source equal to destination (a one-node path), a destination outside
    the ROI (no path), an empty request list, and a call after a bucket-engine
    call on the same handle (pending touch records are flushed first)."""
    import torch
    from orthoroute_amd.device import SearchParams, route, route_flat
    X, Y, L = 16, 12, 4
    o = oracle.Lattice(X, Y, L, "reference")
    g = ctx.build_lattice(X, Y, L, "reference")
    rng = np.random.default_rng(61)
    c = make_costs(o, rng, "jitter")
    cost_d = torch.from_numpy(c).to(ctx.device)
    p1 = SearchParams(slots=4, delta=1.0, n_buckets=8, engine=1, waves=waves)
    nodes, edges, offs, dist, st = route_flat(ctx, g, cost_d, [], p1)
    assert len(offs) == 1 and offs[0] == 0
    mask = np.zeros(o.N, np.uint8)
    mask[:X] = 1                      # first row of layer 0 only
    roi_d = torch.from_numpy(pack_bits(mask)[None, :]).to(ctx.device)
    reqs = [(5, 5, 0, -1, 0), (0, X * Y + 3, 1, 0, 0), (0, X - 1, 2, 0, 0), (3, 77, 3, -1, 0)]
    # a bucket-engine call first leaves touch records on its slots
    p0 = SearchParams(slots=4, delta=1.0, n_buckets=8, engine=0)
    res0, _ = route(ctx, g, cost_d, reqs, p0, roi_bitmaps=roi_d)
    res1, _ = route(ctx, g, cost_d, reqs, p1, roi_bitmaps=roi_d)
    masks = [None, mask, mask, None]
    for i, (s, d, nid, ri, ko) in enumerate(reqs):
        path, dd, *_ = oracle.sssp_canon(o, c, s, d, roi_mask=masks[i])
        for res in (res0, res1):
            got = None if res[i] is None else res[i][0].tolist()
            exp = path if path is not None and len(path) > 1 else None
            assert got == exp, f"request {i}: {got} vs {exp}"
    # and the bucket engine again after the solo engine
    res2, _ = route(ctx, g, cost_d, reqs, p0, roi_bitmaps=roi_d)
    assert [None if r is None else r[0].tolist() for r in res2] == [None if r is None else r[0].tolist() for r in res0]


def test_route_async_matches_route():
    """orgpu_route_async (device requests and results, nothing allocated,
    nothing waited for; four waves per search) returns the paths, edge ids
    and distances of orgpu_route on the one-wave engine with one wave each; invalid requests come back as -4, an
    output too small as -3 (given back), and a call before
    orgpu_reserve_solo fails."""
    import torch
    from orthoroute_amd._abi import OrgpuError
    from orthoroute_amd.device import (Context, SearchParams, cost_bounds, reserve_solo, route_async,
                                       route_async_outputs, route_flat)
    ctx2 = Context(0)
    X, Y, L = 60, 48, 6
    rng = np.random.default_rng(71)
    o = oracle.Lattice(X, Y, L, "adjacent")
    g = ctx2.build_lattice(X, Y, L, "adjacent")
    c = make_costs(o, rng, "jitter")
    cost_d = torch.from_numpy(c).to(ctx2.device)
    owner = np.full(o.N, -1, np.int32)
    sel = rng.random(o.N) < 0.05
    owner[sel] = rng.integers(0, 6, sel.sum())
    own_d = torch.from_numpy(owner).to(ctx2.device)
    reqs, rois = [], []
    for i in range(300):
        s, d = (int(v) for v in rng.integers(0, o.N, 2))
        roi_idx = -1
        if i % 3 == 0:
            roi = oracle.roi_bfs(o, s, d, max_nodes=4000)
            mask = np.zeros(o.N, np.uint8); mask[roi] = 1; mask[s] = 1; mask[d] = 1
            roi_idx = len(rois)
            rois.append(pack_bits(mask))
        reqs.append((s, d, i % 6, roi_idx, int(i % 7 == 0)))
    roi_d = torch.from_numpy(np.stack(rois)).to(ctx2.device)
    hx, hz = cost_bounds(ctx2, g, cost_d)
    p = SearchParams(slots=64, delta=0.5, n_buckets=8, h_xy=hx, h_z=hz, engine=1, waves=4)
    p1 = SearchParams(slots=64, delta=0.5, n_buckets=8, h_xy=hx, h_z=hz, engine=1, waves=1)
    out = route_async_outputs(ctx2, len(reqs), 300 * 400)
    req_d = torch.from_numpy(np.concatenate([np.asarray(reqs, np.int32), np.zeros((len(reqs), 1), np.int32)], 1)).to(
        ctx2.device)
    with pytest.raises(OrgpuError):
        route_async(ctx2, g, cost_d, req_d, p, out, roi_bitmaps=roi_d, keepout_owner=own_d)
    nodes, edges, offs, dist, st = route_flat(ctx2, g, cost_d, reqs, p1, roi_bitmaps=roi_d, keepout_owner=own_d)
    reserve_solo(ctx2, g, 64)
    for call in range(2):
        n = len(reqs) if call == 0 else 120     # a second, smaller call on the same reserve
        route_async(ctx2, g, cost_d, req_d[:n], p, out, roi_bitmaps=roi_d, keepout_owner=own_d)
        ctx2.sync()
        ln = out["len"].cpu().numpy()[:n]
        off = out["off"].cpu().numpy()[:n]
        dd = out["dist"].cpu().numpy()[:n]
        an, ae = out["nodes"].cpu().numpy(), out["edges"].cpu().numpy()
        assert int(out["used"].cpu().numpy()[0]) == int(ln[ln > 0].sum())
        for i in range(n):
            exp = nodes[offs[i]:offs[i + 1]]
            assert ln[i] == len(exp), (call, i)
            np.testing.assert_array_equal(an[off[i]:off[i] + ln[i]], exp)
            np.testing.assert_array_equal(ae[off[i]:off[i] + ln[i]], edges[offs[i]:offs[i + 1]])
            if len(exp):
                assert dd[i] == dist[i]
    # invalid requests: endpoint outside the graph, ROI row past the rows,
    # keepout without the owner map
    bad = torch.tensor([[o.N, 0, 0, -1, 0, 0], [0, 5, 0, len(rois), 0, 0], [3, 9, 0, -1, 1, 0], [3, 9, 0, -1, 0, 0]],
                       dtype=torch.int32, device=ctx2.device)
    route_async(ctx2, g, cost_d, bad, p, out, roi_bitmaps=roi_d)
    ctx2.sync()
    assert out["len"].cpu().numpy()[:4].tolist()[:3] == [-4, -4, -4]
    assert out["len"].cpu().numpy()[3] > 0
    # an output too small for the paths: given back (-3), nothing written past it
    small = route_async_outputs(ctx2, len(reqs), 40)
    route_async(ctx2, g, cost_d, req_d, p, small, roi_bitmaps=roi_d, keepout_owner=own_d)
    ctx2.sync()
    sl = small["len"].cpu().numpy()[:len(reqs)]
    assert (sl == -3).sum() > 0 and sl[sl > 0].sum() <= 40
    ctx2.close()


def test_solo_waves_rejected(ctx):
    """A wavefront count other than 0, 1, 2 or 4 is an argument error."""
    import torch
    from orthoroute_amd._abi import OrgpuError
    from orthoroute_amd.device import SearchParams, route_flat
    o = oracle.Lattice(8, 8, 2, "reference")
    g = ctx.build_lattice(8, 8, 2, "reference")
    cost_d = torch.from_numpy(make_costs(o, np.random.default_rng(3), "jitter")).to(ctx.device)
    with pytest.raises(OrgpuError):
        route_flat(ctx, g, cost_d, [(0, 9, 0, -1, 0)], SearchParams(slots=2, engine=1, waves=3))


@pytest.mark.parametrize("frac", [0.05, 0.3], ids=["mixed", "csr-launch"])
@pytest.mark.parametrize("waves", WAVES)
def test_solo_adjacency_records_on_irregular_graph(ctx, waves, frac):
    """A lattice CSR with extra edges on a fraction of its nodes (a target
    off the lattice pattern, or more than six edges): those nodes keep their
    CSR rows while the others read adjacency records (5 %), or the launch
    reads every row from the CSR (30 %, above one node in eight).  Paths,
    first-match edge ids and distances equal the oracle's on that graph."""
    import types
    import torch
    from orthoroute_amd.device import DeviceGraph, SearchParams, route
    X, Y, L = 30, 24, 6
    o = oracle.Lattice(X, Y, L, "adjacent")
    rng = np.random.default_rng(81 + int(100 * frac))
    plane, N = X * Y, o.N
    rows = [list(o.indices[o.indptr[u]:o.indptr[u + 1]]) for u in range(N)]
    pick = rng.random(N) < frac
    for u in np.nonzero(pick)[0].tolist():
        if u % 2:
            rows[u].append(int(rng.integers(0, N)))                     # off the lattice pattern
        else:
            z = u // plane
            rows[u] += [((z + k) % L) * plane + u % plane for k in (1, 2, 3, 4, 5) if (z + k) % L != z]
    indptr = np.zeros(N + 1, np.int32)
    indptr[1:] = np.cumsum([len(r) for r in rows])
    indices = np.concatenate([np.asarray(r, np.int32) for r in rows])
    src_z = np.repeat(np.arange(N) // plane, np.diff(indptr))
    kind = (src_z != indices // plane).astype(np.uint8)
    og = types.SimpleNamespace(N=N, indptr=indptr, indices=indices, edge_kind=kind)
    c = (1.0 + rng.random(len(indices)) * 3.0).astype(np.float32)
    g = DeviceGraph(ctx, indptr, indices, shape=(X, Y, L))
    cost_d = torch.from_numpy(c).to(ctx.device)
    reqs = []
    for i in range(40):
        s, d = (int(v) for v in rng.integers(0, N, 2))
        reqs.append((s, d, i % 4, -1, 0))
    # plain bucket order: the off-lattice edges void the lattice lower bounds
    p = SearchParams(slots=16, delta=0.5, n_buckets=8, engine=1, waves=waves)
    res, st = route(ctx, g, cost_d, reqs, p)
    say(f"[solo] irregular graph, {pick.sum()} of {N} nodes changed, waves {waves}: retried {st['retried']}")
    for i, (s, d, nid, ri, ko) in enumerate(reqs):
        path, dist, *_ = oracle.sssp_canon(og, c, s, d)
        got = None if res[i] is None else res[i][0].tolist()
        assert got == (path if path is not None and len(path) > 1 else None), f"request {i}"
        if got:
            exp_e = [int(indptr[u] + np.nonzero(indices[indptr[u]:indptr[u + 1]] == v)[0][0])
                     for u, v in zip(path, path[1:])]
            assert res[i][1].tolist() == exp_e
            assert np.float32(res[i][2]) == dist[d]
