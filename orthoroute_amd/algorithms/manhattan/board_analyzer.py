"""Board characteristics that drive the negotiation schedule.

Reference: orthoroute/algorithms/manhattan/board_analyzer.py:58-302
(analyze_board_characteristics, _assign_hv_layers_by_demand).  Same inputs,
same arithmetic (Python floats), same result fields.
"""
from dataclasses import dataclass, field
from typing import Dict, List, Set, Tuple


@dataclass
class BoardCharacteristics:
    signal_layers: List[int]
    h_layers: Set[int]
    v_layers: Set[int]
    layer_count: int
    board_width_mm: float
    board_height_mm: float
    usable_area_mm2: float
    grid_pitch_mm: float
    total_horizontal_channels: int
    total_vertical_channels: int
    net_count: int
    total_hpwl_mm: float
    avg_net_hpwl_mm: float
    demand_horizontal_mm: float
    demand_vertical_mm: float
    demand_h_percentage: float
    congestion_ratio: float
    density_nets_per_mm2: float
    routing_complexity: str
    allowed_via_spans: Set[Tuple[int, int]] = field(default_factory=set)
    via_flexibility: float = 1.0


def _assign_hv_layers_by_demand(signal_layers, tasks, lattice, anchor_top_is_h=True):
    if len(signal_layers) <= 2:
        return set(signal_layers), set(signal_layers), 0.5
    if not tasks:
        h = set(l for l in signal_layers if (l % 2 == 1) == anchor_top_is_h)
        return h, set(signal_layers) - h, 0.5
    ddx = ddy = 0.0
    for _, (s, d) in tasks.items():
        sx, sy, _ = lattice.idx_to_coord(s)
        dx, dy, _ = lattice.idx_to_coord(d)
        ddx += abs(dx - sx)
        ddy += abs(dy - sy)
    tot = ddx + ddy
    h_pct = 0.5 if tot < 1e-6 else ddx / tot
    n = len(signal_layers)
    h_goal = max(1, min(n - 1, int(round(n * h_pct))))
    h_layers, v_layers = set(), set()
    want_h = anchor_top_is_h
    for layer in signal_layers:
        if want_h and len(h_layers) < h_goal:
            h_layers.add(layer)
        else:
            v_layers.add(layer)
        want_h = not want_h
    while len(h_layers) < h_goal and v_layers:
        mid = sorted(v_layers)[len(v_layers) // 2]
        v_layers.remove(mid)
        h_layers.add(mid)
    while len(h_layers) > h_goal and h_layers:
        mid = sorted(h_layers)[len(h_layers) // 2]
        h_layers.remove(mid)
        v_layers.add(mid)
    return h_layers, v_layers, h_pct


def analyze_board_characteristics(lattice, tasks: Dict, board_data=None) -> BoardCharacteristics:
    L = lattice.layers
    pitch = lattice.pitch
    if hasattr(lattice, "bounds"):
        b = lattice.bounds
        w, hgt = b[2] - b[0], b[3] - b[1]
    else:
        w, hgt = lattice.x_steps * pitch, lattice.y_steps * pitch
    signal = list(range(L)) if L <= 2 else list(range(1, L - 1))
    h_layers, v_layers, _ = _assign_hv_layers_by_demand(signal, tasks, lattice)
    cx, cy = int(w / pitch), int(hgt / pitch)
    tot_h = len(h_layers) * cx * cy
    tot_v = len(v_layers) * cy * cx
    hpwl = dem_h = dem_v = 0.0
    for _, (s, d) in tasks.items():
        sx, sy, _ = lattice.idx_to_coord(s)
        dx, dy, _ = lattice.idx_to_coord(d)
        ddx = abs(dx - sx) * pitch
        ddy = abs(dy - sy) * pitch
        hpwl += ddx + ddy
        dem_h += ddx
        dem_v += ddy
    n = len(tasks)
    rho = (hpwl * 1.3) / max(1.0, (tot_h + tot_v) * pitch * 0.75)
    if rho < 0.6:
        cx_s = "SPARSE"
    elif rho < 0.9:
        cx_s = "NORMAL"
    elif rho < 1.2:
        cx_s = "TIGHT"
    else:
        cx_s = "DENSE"
    spans = set((i, j) for i in range(L) for j in range(i + 1, L))
    flex = len(spans) / max(1, L * (L - 1) // 2)
    return BoardCharacteristics(
        signal_layers=signal, h_layers=h_layers, v_layers=v_layers, layer_count=len(signal),
        board_width_mm=w, board_height_mm=hgt, usable_area_mm2=w * hgt, grid_pitch_mm=pitch,
        total_horizontal_channels=tot_h, total_vertical_channels=tot_v, net_count=n,
        total_hpwl_mm=hpwl, avg_net_hpwl_mm=hpwl / max(1, n), demand_horizontal_mm=dem_h,
        demand_vertical_mm=dem_v, demand_h_percentage=dem_h / max(1.0, dem_h + dem_v),
        congestion_ratio=rho, density_nets_per_mm2=n / max(1.0, w * hgt), routing_complexity=cx_s,
        allowed_via_spans=spans, via_flexibility=flex)
