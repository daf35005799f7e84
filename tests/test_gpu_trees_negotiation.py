"""Multi-terminal nets as trees inside the GPU negotiation
(multi_terminal_enabled) against the oracle's negotiation with the same
definition: per-iteration overuse and history bits, every net's final path
or tree, each tree's branches, and the emitted geometry (a tree's records
are its branches' records)."""
import numpy as np
import pytest

from oracle import geometry as G
from oracle.negotiation import Config, OracleRouter

pytestmark = pytest.mark.gpu

CASES = [
    # name, X, Y, L, policy, two-pin nets, tree nets, seed, cap
    ("small", 18, 16, 4, "reference", 14, 6, 1, 12),
    ("mid", 40, 36, 6, "reference", 40, 15, 2, 20),
    ("adjacent", 36, 30, 6, "adjacent", 36, 12, 3, 15),
]


def board(X, Y, n_two, n_tree, seed):
    rng = np.random.default_rng(seed)
    cells = rng.permutation(X * Y)
    reqs, k = [], 0
    for i in range(n_two):
        reqs.append((f"N{i:02d}", int(cells[k]), int(cells[k + 1])))
        k += 2
    for i in range(n_tree):
        m = int(rng.integers(3, 7))
        reqs.append((f"T{i:02d}", [int(c) for c in cells[k:k + m]]))
        k += m
    return reqs


def run_gpu(X, Y, L, policy, reqs, cap, slots):
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    cfg = PathFinderConfig()
    cfg.lattice_policy = policy
    cfg.batch_slots = slots
    cfg.max_iterations_cap = cap
    cfg.multi_terminal_enabled = True
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    rec = []

    def cb(it, t, v, s, c):
        rec.append((it, s, c, r.accounting.history.cpu().numpy()[:r.accounting.E].copy()))
    res = r.route_multiple_nets(reqs, iteration_cb=cb)
    return r, res, rec


def run_oracle(X, Y, L, policy, reqs, cap):
    cfg = Config(grid_pitch=1.0, via_cost=3.0, multi_terminal_enabled=True)
    r = OracleRouter(X, Y, L, policy, pitch=1.0, via_cost=3.0, sssp="canon", config=cfg)
    rec = []
    res = r.route_multiple_nets(reqs, max_iterations=cap,
                                iteration_cb=lambda it, t, v, s, c: rec.append((it, s, c, r.history.copy())))
    return r, res, rec


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("slots", [1, 64])
def test_tree_negotiation_matches_oracle(case, slots):
    name, X, Y, L, policy, n_two, n_tree, seed, cap = case
    print(f"[trees] {name} {X}x{Y}x{L} {policy} two-pin={n_two} trees={n_tree} slots={slots}", flush=True)
    reqs = board(X, Y, n_two, n_tree, seed)
    g, g_res, g_rec = run_gpu(X, Y, L, policy, reqs, cap, slots)
    o, o_res, o_rec = run_oracle(X, Y, L, policy, reqs, cap)
    assert [x[:3] for x in g_rec] == [x[:3] for x in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for q in reqs:
        assert list(g_res["paths"].get(q[0], [])) == list(o_res["paths"].get(q[0], [])), q[0]
    trees = [q[0] for q in reqs if len(q) == 2]
    routed = [n for n in trees if g_res["paths"].get(n)]
    assert routed, "no tree routed"
    for n in routed:
        assert [b.tolist() for b in g._tree_branches[n][0]] == o.tree_branches[n], n
    # geometry: a tree's records are its branches' records
    tracks, vias = g._generate_geometry_from_paths()
    geo = g.lattice.geom
    paths = {}
    order = []
    for n, p in o_res["paths"].items():
        if not p:
            continue
        if n in o.tree_branches:
            for i, b in enumerate(o.tree_branches[n]):
                paths[(n, i)] = b
                order.append((n, i))
        else:
            paths[n] = p
            order.append(n)
    et, ev = [], []
    for key in order:
        net = key[0] if isinstance(key, tuple) else key
        t, v = G.generate({net: paths[key]}, X, Y, L, list(g.config.layer_names)[:max(L, 6)], float(geo.pitch),
                          float(geo.grid_min_x), float(geo.grid_min_y), guard_outer=(policy == "reference"))
        et.extend(t)
        ev.extend(v)
    for e in et:
        e["width"] = g.config.grid_pitch * 0.6
    assert tracks == et
    assert vias == ev
