"""GPU parity of the one-wavefront-per-search engine (orgpu_sssp_params.engine
= 1): paths, first-match edge ids and FP32 distances bit-equal to the
canonical oracle (sssp_canon) and to the bucket-round engine, including ROIs,
keepouts, congested costs, capacity hand-back to the bucket engine and the
edge cases of the two-pin search."""
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


@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
@pytest.mark.parametrize("policy", ["reference", "adjacent"])
@pytest.mark.parametrize("slots", [1, 7, 64])
def test_solo_matches_oracle(ctx, policy, slots, goal):
    for seed, kind in [(1, "jitter"), (2, "ties"), (3, "random"), (4, "congested")]:
        st = run_case(ctx, 24, 20, 6, policy, slots, 24, seed + 10 * slots, kind, goal=goal, engine=1)
        assert st["retried"] == 0


@pytest.mark.parametrize("goal", [False, True], ids=["dijkstra", "goal"])
def test_solo_keepout_matches_oracle(ctx, goal):
    for seed, kind in [(21, "jitter"), (22, "ties"), (23, "congested")]:
        run_case(ctx, 20, 16, 6, "reference", 16, 30, seed, kind, keepout=True, goal=goal, engine=1)


@pytest.mark.parametrize("lookahead", [-1, 0, 15], ids=["no-lookahead", "default", "lookahead15"])
@pytest.mark.parametrize("dn", [(0.05, 1), (0.3, 8), (4.0, 32), (100.0, 8)], ids=["narrow", "mid", "wide", "one"])
def test_solo_bucket_widths(ctx, dn, lookahead):
    """Bucket widths from far below one edge cost (many radix-list moves) to
    wider than every path (one bucket), with and without early expansion of
    the next buckets' entries: the same paths."""
    run_case(ctx, 30, 24, 6, "adjacent", 16, 40, 31, "congested", dn=dn, goal=True, engine=1, lookahead=lookahead)
    run_case(ctx, 30, 24, 6, "adjacent", 16, 40, 32, "jitter", dn=dn, goal=False, engine=1, lookahead=lookahead)


def test_solo_capacity_hands_back(ctx):
    """Bucket-list storage of two 256-entry chunks per search runs out on most
    searches of this board: those searches are routed by the bucket engine and
    every path still equals the oracle's."""
    st = run_case(ctx, 40, 32, 6, "adjacent", 16, 32, 41, "jitter", roi_frac=0.0, dn=(0.3, 8), goal=False,
                  engine=1, list_cap=1)
    say(f"[solo] retried {st['retried']} of 32")
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
    for eng in (0, 1):
        p = SearchParams(slots=128, delta=1.0, n_buckets=8, h_xy=hx, h_z=hz, engine=eng)
        nodes, edges, offs, dist, st = route_flat(ctx, g, cost_d, reqs, p, roi_bitmaps=roi_d)
        say(f"[solo] engine {eng}: {st['kernel_ms']:.2f} ms, {st['items']} items, {st['edges']} edges, "
            f"retried {st['retried']}")
        out.append((nodes, edges, offs, dist))
    for a, b in zip(out[0], out[1]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


def test_solo_edge_cases(ctx):
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
    p1 = SearchParams(slots=4, delta=1.0, n_buckets=8, engine=1)
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
