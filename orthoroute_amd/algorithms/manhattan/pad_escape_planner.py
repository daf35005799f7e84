"""Pad escape planning: a portal per routable SMD pad.

Reference: PadEscapePlanner.precompute_all_pad_escapes
(orthoroute/algorithms/manhattan/pad_escape_planner.py:127-486, cited PEP)
with _plan_column_escapes (PEP:487-687), _try_create_portal (PEP:689-752),
the local DRC (PEP:817-958) and _emit_portal_escape_geometry (PEP:960-1035).

A portal is the F.Cu lattice point at the end of a short vertical escape
stub from the pad; a net whose two pads both have portals is routed between
its portals (the router seeds its search there).  Per column of pads (pads
snapped to a lattice column, sorted by row):
  1. each pad escapes toward the larger gap to its column neighbours (an
     isolated pad by the parity of x + y), with a random length in
     [3, min(12, half that gap, the board edge)];
  2. overlapping escapes in the column (1-step buffer) are shortened by 2
     steps, alternating which of a pair, until none overlap;
  3. each escape gets a random entry layer and is accepted at the first
     length, shortest last, then in the opposite direction longest last, whose
     portal point and three stub samples keep the pad clearance to every
     other pad within 3 mm.
The random draws come from one generator seeded once (the reference seeds
Python's global generator with 42 when the planner is built, PEP:122-124),
in the reference's order: a column's lengths, then its entry layers.
The planner is host code: it runs once per board before routing.
"""
import random
from dataclasses import dataclass

PAD_CLEARANCE_MM = 0.15       # pathfinder/config.py:57
MIN_STEPS, MAX_STEPS = 3, 12  # PEP:533-534
DRC_RADIUS_MM = 3.0           # PEP:714, 721
CELL_MM = 5.0                 # spatial index cell (PEP:830)


@dataclass
class Portal:
    """PEP:55-84: where a pad's escape stub meets the routing grid."""
    x_idx: int
    y_idx: int
    pad_layer: int
    delta_steps: int
    direction: int
    pad_x: float
    pad_y: float
    entry_layer: int = 1
    score: float = 0.0
    retarget_count: int = 0


def pad_key(pad, comp=None):
    """The reference's pad id (PEP:754-764, UPF:2308-2319)."""
    cid = getattr(pad, "component_id", None) or (getattr(comp, "id", None) if comp else None) \
        or "GENERIC_COMPONENT"
    if cid == "GENERIC_COMPONENT" and hasattr(pad, "position"):
        return f"{cid}_{pad.id}@{int(round(pad.position.x * 1000))},{int(round(pad.position.y * 1000))}"
    return f"{cid}_{pad.id}"


class _PadIndex:
    """Pads hashed into 5 mm cells for the local clearance test.  A point is
    clear when no other pad whose centre lies within 3 mm on both axes has
    it inside its clearance-expanded box (PEP:846-958)."""

    def __init__(self, geoms):
        self.geoms = geoms
        self.cells = {}
        for pid, g in geoms.items():
            self.cells.setdefault((int(g["x"] / CELL_MM), int(g["y"] / CELL_MM)), []).append(pid)
        self.reach = int(DRC_RADIUS_MM / CELL_MM) + 1

    def clear(self, x, y, own):
        cx, cy = int(x / CELL_MM), int(y / CELL_MM)
        for i in range(cx - self.reach, cx + self.reach + 1):
            for j in range(cy - self.reach, cy + self.reach + 1):
                for pid in self.cells.get((i, j), ()):
                    if pid == own:
                        continue
                    g = self.geoms[pid]
                    px, py = g["x"], g["y"]
                    if abs(px - x) > DRC_RADIUS_MM or abs(py - y) > DRC_RADIUS_MM:
                        continue
                    hw, hh = g["width"] / 2.0, g["height"] / 2.0
                    if px - hw - PAD_CLEARANCE_MM <= x <= px + hw + PAD_CLEARANCE_MM and \
                            py - hh - PAD_CLEARANCE_MM <= y <= py + hh + PAD_CLEARANCE_MM:
                        return False
        return True


class PadEscapePlanner:
    """PEP:87-1035.  rng: the generator the planner draws from (the router
    passes its own, so that the draws the negotiation makes later follow the
    planner's, as with the reference's shared global generator); it is
    seeded with random_seed here."""

    def __init__(self, lattice, config, pad_to_node, random_seed=42, rng=None):
        self.lattice = lattice
        self.config = config
        self.pad_to_node = pad_to_node
        self.portals = {}
        self.random_seed = random_seed
        self.rng = rng if rng is not None else random.Random()
        self.rng.seed(random_seed)

    # ---- pad geometry (PEP:160-200, 771-815)
    def _pad_geometries(self, board):
        gui = getattr(board, "_gui_pads", None) or []
        geoms = {}
        if gui:
            # GUI pad records matched to pad ids carrying coordinates (x, y in um)
            keyed = []
            for pid in self.pad_to_node:
                if "@" not in pid:
                    continue
                try:
                    xs, ys = pid.split("@")[1].split(",")
                    keyed.append((pid, int(xs) / 1000.0, int(ys) / 1000.0))
                except ValueError:
                    continue
            for rec in gui:
                x, y = rec["x"], rec["y"]
                for pid, px, py in keyed:
                    if abs(px - x) < 0.01 and abs(py - y) < 0.01:
                        geoms[pid] = {"x": x, "y": y, "width": rec["width"], "height": rec["height"]}
                        break
            return geoms

        def size_mm(pad):
            # the reference reads pad.size as internal units (nm) here
            if hasattr(pad, "size"):
                sz = pad.size
                sx = sz.x if hasattr(sz, "x") else sz[0]
                sy = sz.y if hasattr(sz, "y") else sz[1]
                return sx / 1_000_000.0, sy / 1_000_000.0
            return 0.5, 0.5
        for comp in getattr(board, "components", []) or []:
            for pad in getattr(comp, "pads", []) or []:
                w, h = size_mm(pad)
                geoms[pad_key(pad, comp)] = {"x": pad.position.x, "y": pad.position.y, "width": w, "height": h}
        for pad in getattr(board, "pads", []) or []:
            pid = pad_key(pad)
            if pid not in geoms:
                w, h = size_mm(pad)
                geoms[pid] = {"x": pad.position.x, "y": pad.position.y, "width": w, "height": h}
        return geoms

    def _snap(self, pad, comp):
        """(pad id, x_idx, y_idx, x, y, layer) or None when the pad is a
        through-hole pad or lies off its column by more than half a step."""
        if getattr(pad, "drill", 0.0) > 0:
            return None
        px, py = pad.position.x, pad.position.y
        x, y = self.lattice.world_to_lattice(px, py)
        x = max(0, min(x, self.lattice.x_steps - 1))
        y = max(0, min(y, self.lattice.y_steps - 1))
        sx, _ = self.lattice.geom.lattice_to_world(x, 0)
        if abs(px - sx) / self.config.grid_pitch > self.config.portal_x_snap_max:
            return None
        return pad_key(pad, comp), x, y, px, py, 0

    def precompute_all_pad_escapes(self, board, nets_to_route=None):
        """Portals for the first two pads of every net (PEP:127-486).
        Returns (escape tracks, escape vias); the vias are empty (the router
        places the via where its path leaves F.Cu)."""
        geoms = self._pad_geometries(board)
        if nets_to_route is None:
            nets_to_route = list(getattr(board, "nets", []) or [])
        routable, net_pads = set(), {}
        for net in nets_to_route:
            if not hasattr(net, "name") or not hasattr(net, "pads") or len(net.pads) < 2:
                continue
            a, b = pad_key(net.pads[0]), pad_key(net.pads[1])
            if a in self.pad_to_node and b in self.pad_to_node:
                routable.update((a, b))
                net_pads[net.name] = (a, b)
        self.portals.clear()
        pads = []
        for comp in getattr(board, "components", []) or []:
            for pad in getattr(comp, "pads", []) or []:
                if pad_key(pad, comp) in routable:
                    s = self._snap(pad, comp)
                    if s is not None:
                        pads.append(s)
        for pad in getattr(board, "pads", []) or []:
            if pad_key(pad) in routable:
                s = self._snap(pad, None)
                if s is not None:
                    pads.append(s)
        pads.sort(key=lambda p: (p[1], p[2], p[0]))
        index = _PadIndex(geoms)
        columns = {}
        for pid, x, y, px, py, layer in pads:
            columns.setdefault(x, []).append((pid, y, px, py, layer))
        for x in sorted(columns):
            self._plan_column(x, columns[x], index)
        owner = {}
        for name, (a, b) in net_pads.items():
            owner[a] = name
            owner[b] = name
        tracks = []
        for pid, portal in self.portals.items():
            tracks.extend(self._escape_geometry(owner.get(pid, f"PAD_{pid}"), portal))
        return tracks, []

    # ---- one column (PEP:487-687)
    def _plan_column(self, x, col, index):
        ny = self.lattice.y_steps
        plans = []
        for i, (pid, y, px, py, layer) in enumerate(col):
            below = y - col[i - 1][1] if i > 0 else None
            above = col[i + 1][1] - y if i + 1 < len(col) else None
            if above is not None and below is not None:
                d = 1 if above > below else -1
            elif above is not None:
                d = 1
            elif below is not None:
                d = -1
            else:
                d = -1 if (x + y) % 2 == 0 else 1
            edge_room = ny - 1 - y if d > 0 else y
            gap = above if d > 0 else below
            room = gap // 2 if gap is not None else MAX_STEPS
            hi = max(min(MAX_STEPS, room, edge_room), MIN_STEPS)
            plans.append([pid, y, d, self.rng.randint(MIN_STEPS, hi), px, py, layer])
        # overlapping escapes (with a one-step buffer) are shortened by 2
        # steps, the first of a pair on even passes, the second on odd ones
        for it in range(100):
            hit = False
            for i in range(len(plans)):
                for j in range(i + 1, len(plans)):
                    a, b = plans[i], plans[j]
                    ea, eb = a[1] + a[2] * a[3], b[1] + b[2] * b[3]
                    lo_a, hi_a = min(a[1], ea), max(a[1], ea)
                    lo_b, hi_b = min(b[1], eb), max(b[1], eb)
                    if hi_a + 1 < lo_b or hi_b + 1 < lo_a:
                        continue
                    hit = True
                    victim = a if it % 2 == 0 else b
                    if victim[3] > MIN_STEPS:
                        victim[3] = max(MIN_STEPS, victim[3] - 2)
            if not hit:
                break
        layers = list(range(1, self.lattice.layers)) or [1, 2]
        for pid, y, d, delta, px, py, layer in plans:
            entry = self.rng.choice(layers)
            portal = None
            for t in range(delta, MIN_STEPS - 1, -1):
                portal = self._try(x, y, d, t, pid, px, py, layer, entry, index)
                if portal:
                    break
            if portal is None:
                room = ny - 1 - y if -d > 0 else y
                for t in range(MIN_STEPS, min(MAX_STEPS, room) + 1):
                    portal = self._try(x, y, -d, t, pid, px, py, layer, entry, index)
                    if portal:
                        break
            if portal is not None:
                self.portals[pid] = portal

    def _try(self, x, y, d, delta, pid, px, py, layer, entry, index):
        """PEP:689-752: the portal at delta steps in direction d, when the
        portal point and the stub at 1/4, 1/2, 3/4 keep clear of other pads."""
        yp = y + d * delta
        wx, wy = self.lattice.geom.lattice_to_world(x, yp)
        if not index.clear(wx, wy, pid):
            return None
        for t in (0.25, 0.5, 0.75):
            if not index.clear(px + t * (wx - px), py + t * (wy - py), pid):
                return None
        return Portal(x_idx=x, y_idx=yp, pad_layer=layer, delta_steps=delta, direction=d, pad_x=px, pad_y=py,
                      entry_layer=entry, score=0.0, retarget_count=0)

    def _escape_geometry(self, net, portal):
        """PEP:960-1035: the F.Cu stub from the pad to the portal point,
        vertical then at 45 degrees when the pad is off its column."""
        names = list(getattr(self.config, "layer_names", None) or [])
        lname = names[portal.pad_layer] if portal.pad_layer < len(names) else f"L{portal.pad_layer}"
        wx, wy = self.lattice.geom.lattice_to_world(portal.x_idx, portal.y_idx)
        dx, dy = wx - portal.pad_x, wy - portal.pad_y
        width = self.config.grid_pitch * 0.6

        def seg(x1, y1, x2, y2):
            return {"net": net, "layer": lname, "x1": x1, "y1": y1, "x2": x2, "y2": y2, "width": width,
                    "escape": True}
        if abs(dx) > 0.01:
            dy45 = (1 if dy > 0 else -1) * abs(dx)
            mid_y = portal.pad_y + (dy - dy45)
            out = []
            if abs(dy - dy45) > 0.01:
                out.append(seg(portal.pad_x, portal.pad_y, portal.pad_x, mid_y))
            out.append(seg(portal.pad_x, mid_y, wx, wy))
            return out
        return [seg(portal.pad_x, portal.pad_y, wx, wy)]
