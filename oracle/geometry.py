"""Geometry emission oracle -- TEST INFRASTRUCTURE ONLY.

A per-hop restatement of _generate_geometry_from_paths, _segment_world and
_via_world (UPF:5596-5617, 5652-5675, 5749-5830): each path becomes straight
track runs (flushed at a direction change, before a via and at the end) and
one via record per layer change.  Hops that are not lattice moves (planar
Manhattan step != 1, or against the layer's axis) are skipped without
advancing the walk, and planar hops on the outer layers are refused (the
walk restarts at their far node), as in the reference.  Pinned by
hand-worked known answers and invariants in tests/test_oracle_geometry.py.
"""


def axis(L, z):
    """Lattice3D layer direction (UPF:1091-1106): both 'hv' at L=2; layer 0
    'v', odd 'h', even 'v'."""
    if L == 2:
        return "hv"
    return "v" if z == 0 else ("h" if z % 2 == 1 else "v")


def generate(paths, X, Y, L, layer_names, pitch, origin_x, origin_y, guard_outer=True):
    """paths: {net: [node ids]}.  Returns (tracks, vias) as lists of dicts
    with the reference's keys."""
    plane = X * Y

    def coord(n):
        return n % X, (n % plane) // X, n // plane

    def world(x, y):
        wx, wy = origin_x + x * pitch, origin_y + y * pitch
        return (origin_x + round((wx - origin_x) / pitch) * pitch,
                origin_y + round((wy - origin_y) / pitch) * pitch)

    def lname(z):
        return layer_names[z] if z < len(layer_names) else f"L{z}"

    def seg(a, b, layer, net):
        ax, ay, _ = coord(a)
        bx, by, _ = coord(b)
        (x1, y1), (x2, y2) = world(ax, ay), world(bx, by)
        return {"net": net, "layer": lname(layer), "x1": x1, "y1": y1, "x2": x2, "y2": y2, "width": pitch * 0.6}

    def via(at, net, z0, z1):
        x, y, _ = coord(at)
        wx, wy = world(x, y)
        lo, hi = min(z0, z1), max(z0, z1)
        return {"net": net, "x": wx, "y": wy, "from_layer": lname(lo), "to_layer": lname(hi),
                "diameter": 0.25, "drill": 0.15}

    tracks, vias = [], []
    for net, path in paths.items():
        if not path:
            continue
        run_start = prev = path[0]
        prev_dir = None
        prev_layer = coord(prev)[2]
        for node in path[1:]:
            x0, y0, z0 = coord(prev)
            x1, y1, z1 = coord(node)
            if guard_outer and z0 == z1 and (z0 == 0 or z0 == L - 1):
                prev = node
                prev_layer = z1
                run_start = node
                continue
            dx, dy = abs(x1 - x0), abs(y1 - y0)
            if z0 == z1:
                if dx + dy != 1:
                    continue
                if axis(L, z0) == "h":
                    if dy != 0:
                        continue
                elif dx != 0:
                    continue
            if z1 != z0:
                if prev != run_start:
                    tracks.append(seg(run_start, prev, prev_layer, net))
                vias.append(via(prev, net, z0, z1))
                run_start = node
                prev_dir = None
            else:
                d = ((x1 > x0) - (x1 < x0), (y1 > y0) - (y1 < y0))
                if prev_dir is not None and d != prev_dir:
                    tracks.append(seg(run_start, prev, prev_layer, net))
                    run_start = prev
                prev_dir = d
            prev = node
            prev_layer = z1
        if prev != run_start:
            tracks.append(seg(run_start, prev, prev_layer, net))
    return tracks, vias
