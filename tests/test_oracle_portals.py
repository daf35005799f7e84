"""Pad escapes and portal routing against the reference's own run
(tests/golden/portals.json, written by make_golden.py from
PadEscapePlanner.precompute_all_pad_escapes and route_multiple_nets on a
board with two connectors, a dense row, off-column pads and a through-hole
pad):

  * the escape planner (host code) gives the reference's portals, escape
    tracks and leaves Python's generator in the reference's state;
  * the oracle's negotiation with those portals reproduces the reference's
    trajectory (overuse, history / present_ema bits, routing order and
    every net's path, iteration for iteration);
  * a multi-source search with one seed and one target settles like the
    two-pin search between them (the oracle routes portal nets that way);
  * emit_geometry's merge (escapes first, duplicates dropped) over the
    reference's final paths gives the reference's payload.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

import oracle
from oracle import geometry as G
from oracle.negotiation import OraclePortal, OracleRouter

LAYER_NAMES = ["F.Cu", "In1.Cu", "In2.Cu", "In3.Cu", "In4.Cu", "B.Cu"]


@pytest.fixture(scope="module")
def fx(golden_dir):
    with open(os.path.join(golden_dir, "portals.json")) as f:
        return json.load(f)


def board_from(fx):
    from orthoroute_amd.board import Board, Component, Coordinate, Net, Pad
    pads, comps = {}, []
    for c in fx["comps"]:
        ps = [Pad(id=p["id"], component_id=c["id"], net_id=None, position=Coordinate(p["x"], p["y"]),
                  size=(p["w"], p["h"]), drill_size=p["drill"]) for p in c["pads"]]
        for q in ps:
            pads[q.id] = q
        comps.append(Component(id=c["id"], reference=c["id"], value="", footprint="", position=ps[0].position,
                               pads=ps))
    nets = [Net(id=n, name=n, pads=[pads[p] for p in pids]) for n, pids in fx["nets"]]
    return Board(id="P", name="portals", components=comps, nets=nets, layer_count=6)


def calc_bounds(fx):
    """The router's lattice bounds: the net pads' extent plus 3 mm."""
    xy = {p["id"]: (p["x"], p["y"]) for c in fx["comps"] for p in c["pads"]}
    pts = [xy[p] for _, pids in fx["nets"] if len(pids) >= 2 for p in pids]
    xs, ys = [p[0] for p in pts], [p[1] for p in pts]
    return (min(xs) - 3.0, min(ys) - 3.0, max(xs) + 3.0, max(ys) + 3.0)


def pad_id(comp, pad):
    return f"{comp}_{pad}"


def test_planner_matches_reference(fx):
    from orthoroute_amd.algorithms.manhattan.pad_escape_planner import PadEscapePlanner
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import Lattice3D, PathFinderConfig
    board = board_from(fx)
    lat = Lattice3D(calc_bounds(fx), 1.0, 6)
    assert (lat.x_steps, lat.y_steps) == (fx["X"], fx["Y"])
    cfg = PathFinderConfig()
    cfg.grid_pitch = 1.0
    rng = random.Random()
    pl = PadEscapePlanner(lat, cfg, dict(fx["pad_to_node"]), random_seed=42, rng=rng)
    tracks, vias = pl.precompute_all_pad_escapes(board)
    got = {k: [p.x_idx, p.y_idx, p.pad_layer, p.delta_steps, p.direction, p.pad_x, p.pad_y, p.entry_layer]
           for k, p in pl.portals.items()}
    assert list(got) == list(fx["portals"])          # same pads, same order
    assert got == fx["portals"]
    assert tracks == fx["escape_tracks"]
    assert vias == fx["escape_vias"] == []
    st = rng.getstate()
    assert [st[0], list(st[1]), st[2]] == fx["rng_state"]
    # the board exercises: both escape directions, collision shortening
    # (lengths below the first draw), 45-degree stubs, the through-hole pad
    # (the reference reads pad.drill, which domain pads do not have)
    dirs = {v[4] for v in got.values()}
    assert dirs == {-1, 1}
    assert any(abs(v[5] - (fx["grid_min_x"] + v[0])) > 0.01 for v in got.values())
    assert "U1_U1_3" in got
    assert any(t["x1"] != t["x2"] for t in tracks)


def test_planner_is_deterministic_and_seeded(fx):
    """The generator is reseeded at construction: a planner built on a used
    generator plans the same portals."""
    from orthoroute_amd.algorithms.manhattan.pad_escape_planner import PadEscapePlanner
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import Lattice3D, PathFinderConfig
    board = board_from(fx)
    lat = Lattice3D(calc_bounds(fx), 1.0, 6)
    cfg = PathFinderConfig()
    cfg.grid_pitch = 1.0
    rng = random.Random(7)
    rng.random()
    pl = PadEscapePlanner(lat, cfg, dict(fx["pad_to_node"]), rng=rng)
    pl.precompute_all_pad_escapes(board)
    assert {k: [p.x_idx, p.y_idx, p.delta_steps, p.entry_layer] for k, p in pl.portals.items()} == \
        {k: [v[0], v[1], v[3], v[7]] for k, v in fx["portals"].items()}


def run_oracle_portals(fx, sssp="ref"):
    X, Y, L = fx["X"], fx["Y"], fx["L"]
    r = OracleRouter(X, Y, L, "reference", pitch=1.0, via_cost=3.0, sssp=sssp, bounds=calc_bounds(fx))
    for k, v in fx["portals"].items():
        r.portals[k] = OraclePortal(x_idx=v[0], y_idx=v[1], entry_layer=v[7], direction=v[4], delta_steps=v[3])
    st = fx["rng_state"]
    r.rng.setstate((st[0], tuple(st[1]), st[2]))
    nets = [(name, [pad_id(p.split("_")[0], p) for p in pids]) for name, pids in fx["nets"]]
    rec, orders, paths = [], [], []
    orig = r.order_nets

    def order_wrap(tasks):
        o = orig(tasks)
        orders.append(list(o))
        return o
    r.order_nets = order_wrap

    def cb(it, tracks, vias, s, c):
        rec.append((it, s, c, digest(r.history), digest(r.present_ema)))
        paths.append({k: [int(v) for v in p] for k, p in r.net_paths.items()})
    res = r.route_board_nets(nets, fx["pad_to_node"], max_iterations=int(fx["tasks_cap"]), iteration_cb=cb)
    return r, res, rec, orders, paths


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def test_oracle_portal_negotiation_matches_reference(fx):
    r, res, rec, orders, paths = run_oracle_portals(fx)
    assert {k: list(v) for k, v in r.net_pad_ids.items()} == fx["net_pad_ids"]
    assert [list(x[:3]) for x in rec] == fx["over"]
    assert orders == fx["order"]
    for i, (p, q) in enumerate(zip(paths, fx["paths_iter"])):
        assert p == q, f"paths after iteration {i + 1}"
    assert [x[3] for x in rec] == fx["hist_digest"]
    assert [x[4] for x in rec] == fx["ema_digest"]
    assert {k: [int(v) for v in p] for k, p in res["paths"].items()} == fx["final_paths"]
    assert bool(res.get("success", False)) == fx["success"]
    got = {k: [p.x_idx, p.y_idx, p.delta_steps, p.direction, p.retarget_count] for k, p in r.portals.items()}
    assert got == fx["retargeted"]
    # portal nets really searched seed to seed: every routed portal net's
    # path runs between its portals' F.Cu nodes
    plane = fx["X"] * fx["Y"]
    n_portal = 0
    for name, (a, b) in r.net_pad_ids.items():
        p = res["paths"].get(name)
        if a in r.portals and b in r.portals and p:
            n_portal += 1
            assert p[0] // plane == 0 and p[-1] // plane == 0
    assert n_portal >= 10


def test_oracle_portal_retarget_rule():
    """_retarget_portals_for_net (UPF:4929-4969) as written: the first
    retarget negates the direction and moves the portal by twice its length
    against the NEW direction (so a portal above its pad moves further up,
    not to the pad's other side), the second moves it to the preferred
    length, later ones change nothing."""
    r = OracleRouter(12, 30, 4, "reference", pitch=1.0, via_cost=3.0)
    r.portals = {"a": OraclePortal(3, 14, 1, 1, 4), "b": OraclePortal(8, 10, 2, -1, 3)}
    r.net_pad_ids = {"n": ("a", "b")}
    r.retarget_portals("n")
    a, b = r.portals["a"], r.portals["b"]
    assert (a.y_idx, a.direction, a.delta_steps, a.retarget_count) == (22, -1, 4, 1)
    assert (b.y_idx, b.direction, b.delta_steps, b.retarget_count) == (4, 1, 3, 1)
    r.retarget_portals("n")
    assert (a.y_idx, a.direction, a.delta_steps, a.retarget_count) == (20, -1, 6, 2)
    assert (b.y_idx, b.direction, b.delta_steps, b.retarget_count) == (7, 1, 6, 2)
    r.retarget_portals("n")
    assert (a.y_idx, a.retarget_count, b.y_idx, b.retarget_count) == (20, 2, 7, 2)
    # a portal already at the preferred length is not counted again
    r.portals = {"c": OraclePortal(3, 14, 1, 1, 6, retarget_count=1)}
    r.net_pad_ids = {"m": ("c", "c")}
    r.retarget_portals("m")
    assert (r.portals["c"].y_idx, r.portals["c"].retarget_count) == (14, 1)


@pytest.mark.parametrize("seed", range(6))
def test_single_seed_multisource_is_two_pin(seed):
    """find_path_multisource_multisink with one seed and one target (both
    cost 0) returns find_path_roi's path between them on the same ROI list,
    including tie-heavy costs and appended (out-of-order) ROI nodes."""
    rng = np.random.default_rng(seed)
    X, Y, L = 14, 12, 6
    lat = oracle.Lattice(X, Y, L, "reference", pitch=1.0, via_cost=3.0)
    base = lat.costs.astype(np.float32)
    if seed % 2:
        costs = (base * np.float32(0.5) * rng.integers(1, 3, lat.E).astype(np.float32)).astype(np.float32)
    else:
        costs = (base * (1 + 3 * rng.random(lat.E))).astype(np.float32)
    plane = X * Y
    for _ in range(12):
        s = int(rng.integers(0, plane))               # F.Cu seeds, like portal seeds
        d = int(rng.integers(0, plane))
        if s == d:
            continue
        es, ed = s + plane * int(rng.integers(1, L - 1)), d + plane * int(rng.integers(1, L - 1))
        roi = oracle.roi_bfs(lat, es, ed, stagnation_bonus=float(rng.integers(0, 3)), max_nodes=600, keep=(s, d),
                             readd=True)
        ms, _ = oracle.ms_ref(lat, costs, [(s, 0.0)], [(d, 0.0)], roi)
        tp, _ = oracle.sssp_ref(lat, costs, s, d, roi)
        assert (ms is None) == (tp is None)
        if ms is not None:
            assert list(ms) == list(tp)


def test_escape_merge_matches_reference_payload(fx):
    """emit_geometry (UPF:5677-5747) over the reference's final paths: the
    escape stubs, then the routed tracks and vias, duplicates by rounded
    geometry dropped, is the reference's payload."""
    paths = {k: v for k, v in fx["final_paths"].items() if v}
    t, v = G.generate(paths, fx["X"], fx["Y"], fx["L"], LAYER_NAMES, 1.0, fx["grid_min_x"], fx["grid_min_y"])

    def dedupe(items, key):
        seen, out = set(), []
        for it in items:
            k = key(it)
            if k not in seen:
                seen.add(k)
                out.append(it)
        return out
    tracks = dedupe(fx["escape_tracks"] + t, lambda x: (x["net"], x["layer"], round(x["x1"], 3), round(x["y1"], 3),
                                                          round(x["x2"], 3), round(x["y2"], 3), round(x["width"], 3)))
    vias = dedupe(fx["escape_vias"] + v, lambda x: (x["net"], round(x["x"], 3), round(x["y"], 3), x.get("from_layer"),
                                                      x.get("to_layer"), round(x.get("drill", 0), 3),
                                                      round(x.get("diameter", 0), 3)))
    assert tracks == fx["payload_tracks"]
    assert vias == fx["payload_vias"]
    assert fx["geometry_counts"] == [len(tracks), len(vias)]
