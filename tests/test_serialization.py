""".ORP / .ORS files (orthoroute_amd.serialization) against the reference's own
writers and reader on the same synthetic board and geometry
(tests/golden/serialization.json, written by tests/golden/make_golden.py)."""
import gzip
import json
import os

import pytest

from orthoroute_amd import serialization as S
from orthoroute_amd.board import Board, Component, Coordinate, Layer, Net, Pad


@pytest.fixture(scope="module")
def gold(golden_dir):
    with open(os.path.join(golden_dir, "serialization.json")) as f:
        return json.load(f)


def jsonish(x):
    return json.loads(json.dumps(x))


def build_board(spec):
    pad_obj, comps = {}, []
    for c in spec["components"]:
        pads = [Pad(id=p["id"], component_id=c["id"], net_id=None, position=Coordinate(p["x"], p["y"]),
                    size=(p["w"], p["h"]), drill_size=p["drill"], layer=p["layer"]) for p in c["pads"]]
        pad_obj.update({p.id: p for p in pads})
        comps.append(Component(id=c["id"], reference=c["ref"], value="", footprint="",
                               position=Coordinate(c["x"], c["y"]), pads=pads))
    nets = [Net(id=nid, name=name, pads=[pad_obj[p] for p in pids]) for nid, name, pids in spec["nets"]]
    layers = [Layer(name=n, type="signal", stackup_position=i) for i, n in enumerate(["F.Cu", "In1.Cu", "In2.Cu", "B.Cu"])]
    return Board(id="B1", name="synthetic.kicad_pcb", components=comps, nets=nets, layers=layers, layer_count=4)


class Geo:
    def __init__(self, tracks, vias):
        self.tracks, self.vias = tracks, vias


def test_orp_writer_matches_reference(gold, tmp_path):
    board = build_board(gold["board_spec"])
    S.export_board_to_orp(board, tmp_path / "b.orp")
    with gzip.open(tmp_path / "b.orp", "rt") as f:
        got = json.load(f)
    got["metadata"].pop("export_timestamp")
    assert got == gold["orp"]


def test_orp_reader_builds_the_reference_board(gold, tmp_path):
    p = tmp_path / "plain.orp"
    p.write_text(json.dumps(gold["orp"]))            # uncompressed files are read too
    b = S.convert_orp_to_board(S.import_board_from_orp(p))
    view = {
        "id": b.id, "name": b.name, "layer_count": b.layer_count, "thickness": b.thickness,
        "layer_names": b.layer_names, "kicad_bounds": list(b._kicad_bounds),
        "components": [{"id": c.id, "reference": c.reference, "x": c.position.x, "y": c.position.y,
                        "layer": c.layer, "pads": [q.id for q in c.pads]} for c in b.components],
        "nets": [{"id": n.id, "name": n.name, "netclass": n.netclass, "pads": [q.id for q in n.pads]}
                 for n in b.nets],
        "pads": {q.id: {"component_id": q.component_id, "net_id": q.net_id, "x": q.position.x, "y": q.position.y,
                        "size": list(q.size), "drill": q.drill_size, "layer": q.layer}
                 for c in b.components for q in c.pads},
    }
    assert jsonish(view) == gold["orp_board"]
    assert b.validate_integrity() == []


def test_ors_writer_reader_match_reference(gold, tmp_path):
    g = gold["geometry_in"]
    S.export_solution_to_ors(Geo(g["tracks"], g["vias"]), g["metrics"], g["metadata"], tmp_path / "s.ors")
    with gzip.open(tmp_path / "s.ors", "rt") as f:
        got = json.load(f)
    got["metadata"].pop("export_timestamp")
    assert got == gold["ors"]
    geo, meta = S.import_solution_from_ors(tmp_path / "s.ors")
    pl = S.convert_ors_to_geometry_payload(geo)
    assert jsonish({"tracks": pl.tracks, "vias": pl.vias}) == gold["payload"]
    assert meta["statistics"] == gold["ors"]["statistics"]


def test_format_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        S.import_board_from_orp(tmp_path / "missing.orp")
    bad = tmp_path / "bad.orp"
    for doc, msg in (({"board": {}}, "Missing format_version"), ({"format_version": "2.0"}, "Unsupported"),
                     ({"format_version": "1.0", "metadata": {}}, "Missing required sections")):
        bad.write_text(json.dumps(doc))
        with pytest.raises(ValueError, match=msg):
            S.import_board_from_orp(bad)
    with pytest.raises(ValueError):
        S.export_solution_to_ors(None, [], {}, tmp_path / "x.ors")


def test_router_geometry_round_trip(tmp_path):
    """Tracks and vias the router emits survive .ORS export and import."""
    import numpy as np
    from test_host_logic import geometry_router, lattice_walk
    rng = np.random.default_rng(5)
    r = geometry_router(12, 10, 6, "reference", 0.5, (1.0, 1.0))
    r.net_paths = {f"N{i}": lattice_walk(rng, 12, 10, 6, 30, False) for i in range(20)}
    tracks, vias = r._generate_geometry_from_paths()
    S.export_solution_to_ors(Geo(tracks, vias), [{"iteration": 1, "overuse": 0}], {"converged": True},
                             tmp_path / "r.ors", compress=False)
    geo, meta = S.import_solution_from_ors(tmp_path / "r.ors")
    assert meta["statistics"]["total_tracks"] == len(tracks) and meta["statistics"]["total_vias"] == len(vias)
    pl = S.convert_ors_to_geometry_payload(geo)
    got = sorted((t["net_id"], t["layer"], t["start"], t["end"]) for t in pl.tracks)
    exp = sorted((t["net"], t["layer"], (t["x1"], t["y1"]), (t["x2"], t["y2"])) for t in tracks)
    assert got == exp
    assert sorted((v["net_id"], v["position"]) for v in pl.vias) == sorted((v["net"], (v["x"], v["y"])) for v in vias)
