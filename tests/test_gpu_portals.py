"""Portal routing on the GPU (pad escapes planned on the host, nets searched
between their portals' F.Cu seeds in ROIs grown from the portals' entry
nodes that keep both seeds):

  * in the reference's search order, the router reproduces the reference's
    own run on the portal board (tests/golden/portals.json): pad mapping,
    portals, escape stubs, per-iteration overuse, history and present_ema
    bits, routing order and every net's path, and the merged geometry
    payload;
  * in the canonical order (FP32, lowest predecessor), batched, it equals
    the oracle's negotiation with the same portals and search semantics.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from oracle.negotiation import OraclePortal, OracleRouter

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


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


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def run_router(fx, order, slots=1024):
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    cfg = PathFinderConfig()
    cfg.grid_pitch = 1.0
    cfg.via_cost = 3.0
    cfg.max_iterations_cap = int(fx["tasks_cap"])
    cfg.sssp_order = order
    cfg.batch_slots = slots
    r = PathFinderRouter(cfg)
    board = board_from(fx)
    r.initialize_graph(board)
    tracks, vias = r.precompute_all_pad_escapes(board)
    rec, orders, paths = [], [], []
    orig = r._order_nets_by_difficulty

    def order_wrap(tasks, over):
        o = orig(tasks, over)
        orders.append(list(o))
        return o
    r._order_nets_by_difficulty = order_wrap

    def cb(it, t, v, s, c):
        E = r.accounting.E
        rec.append((it, s, c, r.accounting.history.cpu().numpy()[:E].copy(),
                    r.accounting.present_ema.cpu().numpy()[:E].copy()))
        paths.append({k: [int(x) for x in p] for k, p in r.net_paths.items()})
    res = r.route_multiple_nets(board.nets, iteration_cb=cb)
    return r, board, (tracks, vias), res, rec, orders, paths


def test_portal_negotiation_reference_order_reproduces_reference(fx):
    say(f"[portals] reference order {fx['X']}x{fx['Y']}x{fx['L']} nets={len(fx['nets'])} cap={fx['tasks_cap']}")
    r, board, (tracks, vias), res, rec, orders, paths = run_router(fx, "reference")
    assert (r.lattice.x_steps, r.lattice.y_steps, r._Nz) == (fx["X"], fx["Y"], fx["L"])
    assert {k: int(v) for k, v in r.pad_to_node.items()} == fx["pad_to_node"]
    assert tracks == fx["escape_tracks"] and vias == fx["escape_vias"]
    assert {k: list(v) for k, v in r.net_pad_ids.items()} == fx["net_pad_ids"]
    assert [[x[0], x[1], x[2]] for x in rec] == fx["over"]
    assert orders == fx["order"]
    for i, (p, q) in enumerate(zip(paths, fx["paths_iter"])):
        assert p == q, f"paths after iteration {i + 1}"
    assert [digest(x[3]) for x in rec] == fx["hist_digest"]
    assert [digest(x[4]) for x in rec] == fx["ema_digest"]
    assert {k: [int(v) for v in p] for k, p in res["paths"].items()} == fx["final_paths"]
    assert bool(res.get("success", False)) == fx["success"]
    assert {k: [p.x_idx, p.y_idx, p.delta_steps, p.direction, p.retarget_count]
            for k, p in r.portals.items()} == fx["retargeted"]
    n_tracks, n_vias = r.emit_geometry(board)
    payload = r.get_geometry_payload()
    assert [n_tracks, n_vias] == fx["geometry_counts"]
    assert payload.tracks == fx["payload_tracks"]
    assert payload.vias == fx["payload_vias"]


@pytest.mark.parametrize("slots", [1, 1024])
def test_portal_negotiation_canonical_matches_oracle(fx, slots):
    say(f"[portals] canonical order slots={slots}")
    r, board, _, res, rec, orders, paths = run_router(fx, "canonical", slots)
    bounds = r.lattice.bounds
    o = OracleRouter(fx["X"], fx["Y"], fx["L"], "reference", pitch=1.0, via_cost=3.0, sssp="canon", bounds=bounds)
    for k, v in fx["portals"].items():
        o.portals[k] = OraclePortal(x_idx=v[0], y_idx=v[1], entry_layer=v[7], direction=v[4], delta_steps=v[3])
    st = fx["rng_state"]
    o.rng.setstate((st[0], tuple(st[1]), st[2]))
    o_rec = []
    nets = [(name, [f"{p.split('_')[0]}_{p}" for p in pids]) for name, pids in fx["nets"]]
    o_res = o.route_board_nets(nets, fx["pad_to_node"], max_iterations=int(fx["tasks_cap"]),
                               iteration_cb=lambda it, t, v, s, c: o_rec.append((it, s, c, o.history.copy())))
    assert [x[:3] for x in rec] == [x[:3] for x in o_rec]
    for a, b in zip(rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    assert {k: list(v) for k, v in res["paths"].items()} == {k: list(v) for k, v in o_res["paths"].items()}
    assert {k: [p.y_idx, p.retarget_count] for k, p in r.portals.items()} == \
        {k: [p.y_idx, p.retarget_count] for k, p in o.portals.items()}
