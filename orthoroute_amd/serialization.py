""".ORP (board) and .ORS (solution) files: gzip-compressed JSON, format 1.0.

The headless flow reads a board exported from KiCad as .ORP, routes it and
writes the tracks and vias as .ORS (main.py:287-622).  Section names, keys,
defaults and conversions follow the reference's readers and writers
(orthoroute/infrastructure/serialization/orp_exporter.py: 39-92, 94-540,
543-590, 769-987; ors_exporter.py: 18-97, 99-340, 342-453), including their
fallbacks: a track needs 'x1'/'y1' or 'start'/'end' keys, a via's diameter is
read from 'size' (0.4 when absent), and the wirelength statistic sums only
tracks that carry 'start'/'end' or 'start_x'... keys.
"""
import gzip
import json
from datetime import datetime, timezone
from pathlib import Path
from typing import Any, Dict, List, Tuple
from uuid import uuid4

from .board import Board, Component, Coordinate, Layer, Net, Pad

FORMAT_VERSION = "1.0"
ORP_SECTIONS = ["metadata", "board", "pads", "nets", "layers", "drc_rules", "grid_parameters"]
ORS_SECTIONS = ["metadata", "geometry", "iteration_metrics", "statistics"]


def _timestamp():
    return datetime.now(timezone.utc).replace(tzinfo=None).isoformat() + "Z"


def _write(data, filepath, compress):
    out = Path(filepath)
    out.parent.mkdir(parents=True, exist_ok=True)
    opener = (lambda: gzip.open(out, "wt", encoding="utf-8")) if compress else \
        (lambda: open(out, "w", encoding="utf-8"))
    with opener() as f:
        json.dump(data, f, indent=2, ensure_ascii=False)


def _read(filepath, kind, sections):
    path = Path(filepath)
    if not path.exists():
        raise FileNotFoundError(f"{kind} file not found: {path}")
    try:
        with gzip.open(path, "rt", encoding="utf-8") as f:
            data = json.load(f)
    except (gzip.BadGzipFile, OSError):
        with open(path, "r", encoding="utf-8") as f:
            data = json.load(f)
    version = data.get("format_version")
    if not version:
        raise ValueError(f"Missing format_version in {kind} file")
    if version != FORMAT_VERSION:
        raise ValueError(f"Unsupported {kind} format version: {version}")
    missing = [s for s in sections if s not in data]
    if missing:
        raise ValueError(f"Missing required sections: {', '.join(missing)}")
    return data


# ---------------------------------------------------------------------------
# .ORP
# ---------------------------------------------------------------------------

def export_board_to_orp(board: Board, filepath: str, compress: bool = True) -> None:
    """Write a Board as .ORP (the reference's Board-object branch)."""
    issues = board.validate_integrity()
    if issues:
        raise ValueError(f"Board validation failed: {', '.join(issues)}")
    b = board.get_bounds()
    pads = []
    for comp in board.components:
        for pad in comp.pads:
            net = board.get_net(pad.net_id) if pad.net_id else None
            pads.append({"id": pad.id, "component_id": pad.component_id, "component_ref": comp.reference,
                         "net_id": pad.net_id, "net_name": net.name if net else None,
                         "position": {"x": pad.position.x, "y": pad.position.y},
                         "size": {"width": pad.size[0], "height": pad.size[1]},
                         "drill_size": pad.drill_size, "layer": pad.layer, "shape": pad.shape, "angle": pad.angle})
    nets = []
    for net in board.nets:
        nb = net.get_bounds()
        nets.append({"id": net.id, "name": net.name, "netclass": net.netclass, "pad_count": len(net.pads),
                     "is_routable": net.is_routable,
                     "terminals": [{"pad_id": p.id, "position": {"x": p.position.x, "y": p.position.y},
                                    "layer": p.layer} for p in net.pads],
                     "bounds": {"x_min": nb.min_x, "y_min": nb.min_y, "x_max": nb.max_x, "y_max": nb.max_y}})
    data = {
        "format_version": FORMAT_VERSION,
        "metadata": {"filename": Path(filepath).name, "board_name": board.name, "board_id": board.id,
                     "export_timestamp": _timestamp(), "orthoroute_version": "0.1.0"},
        "board": {"bounds": {"x_min": b.min_x, "y_min": b.min_y, "x_max": b.max_x, "y_max": b.max_y,
                             "width": b.width, "height": b.height},
                  "layer_count": board.layer_count, "thickness": board.thickness},
        "pads": pads,
        "nets": nets,
        "layers": [{"name": l.name, "type": l.type, "stackup_position": l.stackup_position,
                    "thickness": l.thickness, "material": l.material, "is_routing_layer": l.is_routing_layer}
                   for l in board.layers],
        "drc_rules": {"default": {"track_width": 0.25, "clearance": 0.2, "via_diameter": 0.6, "via_drill": 0.3,
                                  "netclass": "Default"},
                      "netclasses": {},
                      "min_values": {"track_width": 0.1, "clearance": 0.1, "via_diameter": 0.2, "via_drill": 0.1}},
        "grid_parameters": {"grid_pitch": 0.4, "expansion_margin": 3.0,
                            "coordinate_system": "PCB coordinates in millimeters, origin at board bounds minimum",
                            "board_bounds_used": {"x_min": b.min_x, "y_min": b.min_y, "x_max": b.max_x,
                                                  "y_max": b.max_y}},
    }
    _write(data, filepath, compress)


def import_board_from_orp(filepath: str) -> Dict[str, Any]:
    """Read a .ORP file (gzip or plain JSON) and check its version and
    sections."""
    return _read(filepath, "ORP", ORP_SECTIONS)


def convert_orp_to_board(orp: Dict[str, Any]) -> Board:
    """ORP dictionary -> Board: pads grouped into components by
    component_id (pads without one go to 'ORPHANED_PADS'), a component's
    position is its pads' centroid and its layer its first pad's, nets take
    their pads from the terminals' pad ids, default layer stack when the file
    lists none; _kicad_bounds and layer_names are set for the router."""
    meta = orp.get("metadata", {})
    bdata = orp.get("board", {})
    layer_count = bdata.get("layer_count", 2)
    layers = [Layer(name=l.get("name", "F.Cu"), type=l.get("type", "signal"),
                    stackup_position=l.get("stackup_position", 0), thickness=l.get("thickness", 0.035),
                    material=l.get("material", "copper")) for l in orp.get("layers", [])]
    if not layers:
        names = ["F.Cu", "B.Cu"] if layer_count == 2 else \
            ["F.Cu"] + [f"In{i}.Cu" for i in range(1, layer_count - 1)] + ["B.Cu"]
        layers = [Layer(name=n, type="signal", stackup_position=i) for i, n in enumerate(names)]
    pads_data = orp.get("pads", [])
    by_id, by_comp = {}, {}
    for pd in pads_data:
        pos, size = pd.get("position", {}), pd.get("size", {})
        pad = Pad(id=pd.get("id", str(uuid4())), component_id=pd.get("component_id", ""), net_id=pd.get("net_id"),
                  position=Coordinate(pos.get("x", 0.0), pos.get("y", 0.0)),
                  size=(size.get("width", 0.0), size.get("height", 0.0)), drill_size=pd.get("drill_size"),
                  layer=pd.get("layer", "F.Cu"), shape=pd.get("shape", "circle"), angle=pd.get("angle", 0.0))
        by_id[pad.id] = pad
        by_comp.setdefault(pd.get("component_id", ""), []).append(pad)
    ref_of = {}
    for pd in pads_data:
        ref_of.setdefault(pd.get("component_id"), pd.get("component_ref", pd.get("component_id")))
    components = []
    for cid, cpads in by_comp.items():
        if not cid:
            cid, ref = "ORPHANED_PADS", "ORPHANED"
            for p in cpads:
                p.component_id = cid
        else:
            ref = ref_of.get(cid, "U?")
        n = len(cpads)
        centroid = Coordinate(sum(p.position.x for p in cpads) / n, sum(p.position.y for p in cpads) / n)
        components.append(Component(id=cid, reference=ref, value="", footprint="", position=centroid, angle=0.0,
                                    layer=cpads[0].layer, pads=cpads))
    nets = []
    for nd in orp.get("nets", []):
        nid = nd.get("id", str(uuid4()))
        npads = []
        for t in nd.get("terminals", []):
            pid = t.get("pad_id")
            if pid and pid in by_id:
                by_id[pid].net_id = nid
                npads.append(by_id[pid])
        nets.append(Net(id=nid, name=nd.get("name", ""), netclass=nd.get("netclass", "Default"), pads=npads))
    board = Board(id=meta.get("board_id", str(uuid4())), name=meta.get("board_name", "Unknown Board"),
                  components=components, nets=nets, layers=layers, thickness=bdata.get("thickness", 1.6),
                  layer_count=layer_count)
    bd = bdata.get("bounds", {})
    if bd:
        board._kicad_bounds = (bd.get("x_min", 0.0), bd.get("y_min", 0.0), bd.get("x_max", 0.0), bd.get("y_max", 0.0))
    board.layer_names = [l.name for l in layers]
    return board


# ---------------------------------------------------------------------------
# .ORS
# ---------------------------------------------------------------------------

def _geometry_section(geometry) -> Dict[str, Any]:
    by_net, usage = {}, {}

    def entry(net_id):
        if net_id not in by_net:
            by_net[net_id] = {"net_id": net_id, "tracks": [], "vias": []}
        return by_net[net_id]

    for t in geometry.tracks:
        net_id = str(t.get("net_id") or t.get("net", "unknown"))
        if "start" in t and "end" in t:
            (sx, sy), (ex, ey) = t["start"], t["end"]
        elif "x1" in t and "y1" in t:
            sx, sy, ex, ey = t.get("x1", 0.0), t.get("y1", 0.0), t.get("x2", 0.0), t.get("y2", 0.0)
        else:
            sx, sy = t.get("start_x", 0.0), t.get("start_y", 0.0)
            ex, ey = t.get("end_x", 0.0), t.get("end_y", 0.0)
        layer = t.get("layer", 0)
        entry(net_id)["tracks"].append({"layer": layer, "start": {"x": float(sx), "y": float(sy)},
                                        "end": {"x": float(ex), "y": float(ey)},
                                        "width": float(t.get("width", 0.15))})
        usage[layer] = usage.get(layer, 0) + 1
    for v in geometry.vias:
        net_id = str(v.get("net_id") or v.get("net", "unknown"))
        if "position" in v:
            p = v["position"]
            px, py = p if isinstance(p, (list, tuple)) else (p.get("x", 0.0), p.get("y", 0.0))
        elif "x" in v and "y" in v:
            px, py = v.get("x", 0.0), v.get("y", 0.0)
        else:
            px, py = 0.0, 0.0
        entry(net_id)["vias"].append({"position": {"x": float(px), "y": float(py)},
                                      "from_layer": v.get("from_layer", 0), "to_layer": v.get("to_layer", 1),
                                      "diameter": float(v.get("size", 0.4)), "drill": float(v.get("drill", 0.2))})
    all_tracks = [dict(t, net_id=n["net_id"]) for n in by_net.values() for t in n["tracks"]]
    all_vias = [dict(v, net_id=n["net_id"]) for n in by_net.values() for v in n["vias"]]
    return {"by_net": by_net, "all_tracks": all_tracks, "all_vias": all_vias, "layer_usage": usage}


def _metrics_section(metrics: List[Dict[str, Any]]) -> List[Dict[str, Any]]:
    return [{"iteration": m.get("iteration", 0),
             "overuse_count": m.get("overuse_count", m.get("overuse", 0)),
             "nets_routed": m.get("nets_routed", m.get("routed", 0)),
             "overflow_cost": m.get("overflow_cost", m.get("overflow", 0.0)),
             "wirelength": m.get("wirelength", 0.0),
             "via_count": m.get("via_count", m.get("vias", 0)),
             "iteration_time_seconds": m.get("iteration_time", m.get("time", 0.0))} for m in metrics]


def _statistics_section(geometry, metrics, metadata) -> Dict[str, Any]:
    wl = 0.0
    for t in geometry.tracks:
        if "start" in t and "end" in t:
            (sx, sy), (ex, ey) = t["start"], t["end"]
        else:
            sx, sy = t.get("start_x", 0.0), t.get("start_y", 0.0)
            ex, ey = t.get("end_x", 0.0), t.get("end_y", 0.0)
        wl += ((ex - sx) ** 2 + (ey - sy) ** 2) ** 0.5
    routed = metadata.get("nets_routed", 0)
    if routed == 0:
        routed = len({str(t.get("net_id")) for t in geometry.tracks if t.get("net_id")})
    over, overflow = 0, 0.0
    if metrics:
        last = metrics[-1]
        over = last.get("overuse_count", last.get("overuse", 0))
        overflow = last.get("overflow_cost", last.get("overflow", 0.0))
    return {"total_tracks": len(geometry.tracks), "total_vias": len(geometry.vias), "total_wirelength_mm": wl,
            "nets_routed": routed, "final_overuse_count": over, "final_overflow_cost": overflow,
            "converged": metadata.get("converged", over == 0), "iterations_completed": len(metrics)}


def export_solution_to_ors(geometry, iteration_metrics: List[Dict[str, Any]], metadata: Dict[str, Any],
                           filepath: str, compress: bool = True) -> None:
    """Write a routing solution (a geometry payload with .tracks / .vias,
    per-iteration metrics, run metadata) as .ORS."""
    if geometry is None:
        raise ValueError("Geometry cannot be None")
    if not hasattr(geometry, "tracks") or not hasattr(geometry, "vias"):
        raise ValueError("Geometry must have 'tracks' and 'vias' attributes")
    data = {
        "format_version": FORMAT_VERSION,
        "metadata": {"export_timestamp": _timestamp(),
                     "orthoroute_version": metadata.get("orthoroute_version", "0.1.0"),
                     "board_name": metadata.get("board_name", "unknown"),
                     "total_iterations": metadata.get("total_iterations", 0),
                     "converged": metadata.get("converged", False),
                     "total_time_seconds": metadata.get("total_time", 0.0),
                     "notes": metadata.get("notes", "")},
        "geometry": _geometry_section(geometry),
        "iteration_metrics": _metrics_section(iteration_metrics),
        "statistics": _statistics_section(geometry, iteration_metrics, metadata),
    }
    _write(data, filepath, compress)


def import_solution_from_ors(filepath: str) -> Tuple[Dict[str, Any], Dict[str, Any]]:
    """Read a .ORS file: (geometry section, {metadata, statistics,
    iteration_metrics})."""
    data = _read(filepath, "ORS", ORS_SECTIONS)
    return data["geometry"], {"metadata": data["metadata"], "statistics": data["statistics"],
                              "iteration_metrics": data["iteration_metrics"]}


class SolutionGeometry:
    def __init__(self, tracks, vias):
        self.tracks = tracks
        self.vias = vias


def convert_ors_to_geometry_payload(geometry_data: Dict[str, Any]) -> SolutionGeometry:
    """Flat track / via lists of an .ORS geometry section back in the
    payload form (positions as tuples, the via diameter under 'size')."""
    tracks = [{"net_id": t.get("net_id"), "layer": t.get("layer"), "start": (t["start"]["x"], t["start"]["y"]),
               "end": (t["end"]["x"], t["end"]["y"]), "width": t.get("width")}
              for t in geometry_data.get("all_tracks", [])]
    vias = [{"net_id": v.get("net_id"), "position": (v["position"]["x"], v["position"]["y"]),
             "from_layer": v.get("from_layer"), "to_layer": v.get("to_layer"), "size": v.get("diameter"),
             "drill": v.get("drill")} for v in geometry_data.get("all_vias", [])]
    return SolutionGeometry(tracks, vias)
