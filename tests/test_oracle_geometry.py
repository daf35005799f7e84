"""Geometry emission oracle (oracle.geometry, UPF:5596-5675, 5749-5830)
against hand-worked answers and invariants of the track/via records."""
import os

import numpy as np

from oracle import geometry as G

X, Y, L = 10, 8, 6
P = X * Y
NAMES = ["F.Cu", "In1.Cu", "In2.Cu", "In3.Cu", "In4.Cu", "B.Cu"]


def n(x, y, z):
    return z * P + y * X + x


def run(paths, guard=True, pitch=0.5, ox=1.0, oy=2.0):
    return G.generate(paths, X, Y, L, NAMES, pitch, ox, oy, guard_outer=guard)


def test_known_answer_runs_corners_and_vias():
    # layer 1 is 'h', layer 2 is 'v' (UPF:1091-1106)
    path = [n(0, 0, 0), n(0, 0, 1), n(1, 0, 1), n(2, 0, 1), n(3, 0, 1),   # via 0->1, run east
            n(3, 0, 2), n(3, 1, 2), n(3, 2, 2),                           # via 1->2, run north
            n(3, 2, 4), n(2, 2, 4)]                                       # via 2->4, one step west... on 'v'? no
    tracks, vias = run({"A": path})
    # layer 4 is 'v': the west step (dx=1) is skipped, so the walk ends at (3,2,4)
    assert [(t["layer"], t["x1"], t["y1"], t["x2"], t["y2"]) for t in tracks] == [
        ("In1.Cu", 1.0, 2.0, 2.5, 2.0), ("In2.Cu", 2.5, 2.0, 2.5, 3.0)]
    assert [(v["x"], v["y"], v["from_layer"], v["to_layer"]) for v in vias] == [
        (1.0, 2.0, "F.Cu", "In1.Cu"), (2.5, 2.0, "In1.Cu", "In2.Cu"), (2.5, 3.0, "In2.Cu", "In4.Cu")]
    assert all(t["width"] == 0.5 * 0.6 for t in tracks)
    assert all(v["diameter"] == 0.25 and v["drill"] == 0.15 for v in vias)


def test_known_answer_direction_change_on_hv_free_layer_and_guard():
    # a U-turn on an 'h' layer splits the run at the turning node; planar
    # hops on layer 0 are refused and restart the walk at their far node
    path = [n(2, 3, 0), n(3, 3, 0), n(3, 3, 1), n(4, 3, 1), n(5, 3, 1), n(4, 3, 1), n(3, 3, 1)]
    tracks, vias = run({"B": path}, pitch=1.0, ox=0.0, oy=0.0)
    assert [(t["x1"], t["x2"]) for t in tracks] == [(3.0, 5.0), (5.0, 3.0)]
    assert [(v["x"], v["from_layer"]) for v in vias] == [(3.0, "F.Cu")]
    # without the guard the first planar hop on layer 0 ('v' axis) is an
    # off-axis move and is skipped without advancing: the via then starts at
    # (2,3,0)->(3,3,1), which is not a lattice move but is not checked
    tracks2, vias2 = run({"B": path}, guard=False, pitch=1.0, ox=0.0, oy=0.0)
    assert [(v["x"], v["y"]) for v in vias2] == [(2.0, 3.0)]


def random_walk(rng, steps):
    x, y, z = int(rng.integers(0, X)), int(rng.integers(0, Y)), int(rng.integers(0, L))
    p = [n(x, y, z)]
    for _ in range(steps):
        r = rng.random()
        if r < 0.2:
            z = int(np.clip(z + rng.choice([-2, -1, 1, 2]), 0, L - 1))
        elif G.axis(L, z) == "h":
            x = int(np.clip(x + rng.choice([-1, 1]), 0, X - 1))
        else:
            y = int(np.clip(y + rng.choice([-1, 1]), 0, Y - 1))
        if n(x, y, z) != p[-1]:
            p.append(n(x, y, z))
    return p


def test_invariants_on_lattice_walks():
    rng = np.random.default_rng(3)
    paths = {f"N{i}": random_walk(rng, 40) for i in range(60)}
    tracks, vias = run(paths, guard=False, pitch=1.0, ox=0.0, oy=0.0)
    # every track is axis-aligned and non-empty; along one net, planar unit
    # steps covered by tracks equal the walk's planar hops
    for t in tracks:
        assert (t["x1"] == t["x2"]) != (t["y1"] == t["y2"])
    for net, p in paths.items():
        hops = sum(1 for a, b in zip(p, p[1:]) if a // P == b // P)
        cover = sum(abs(t["x2"] - t["x1"]) + abs(t["y2"] - t["y1"]) for t in tracks if t["net"] == net)
        assert cover == hops, net
        nvia = sum(1 for a, b in zip(p, p[1:]) if a // P != b // P)
        assert nvia == sum(1 for v in vias if v["net"] == net)
    # consecutive tracks of one leg (no via between them) never continue in
    # the same direction: runs are maximal
    by_net = {}
    for t in tracks:
        by_net.setdefault(t["net"], []).append(t)
    for net, ts in by_net.items():
        via_at = {(v["x"], v["y"]) for v in vias if v["net"] == net}
        for a, b in zip(ts, ts[1:]):
            if a["layer"] == b["layer"] and (a["x2"], a["y2"]) == (b["x1"], b["y1"]) and \
                    (a["x2"], a["y2"]) not in via_at:
                da = (np.sign(a["x2"] - a["x1"]), np.sign(a["y2"] - a["y1"]))
                db = (np.sign(b["x2"] - b["x1"]), np.sign(b["y2"] - b["y1"]))
                assert da != db


def test_geometry_matches_reference_records(golden_dir):
    """The reference's own _generate_geometry_from_paths output
    (tests/golden/geometry.json): the reference-routed final paths of a
    24x22x6 board and crafted paths through every special case of the walk."""
    import json
    with open(os.path.join(golden_dir, "geometry.json")) as f:
        d = json.load(f)
    t, v = G.generate(d["paths"], d["X"], d["Y"], d["L"], d["layer_names"], d["pitch"], d["grid_min_x"],
                      d["grid_min_y"], guard_outer=True)
    for e in t:
        e["width"] = d["grid_pitch"] * 0.6
    assert t == d["tracks"]
    assert v == d["vias"]
