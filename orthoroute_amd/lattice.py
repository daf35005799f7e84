"""Lattice policies of the 3D Manhattan routing grid.

Reference: Lattice3D (UPF = orthoroute/algorithms/manhattan/unified_pathfinder.py,
UPF:1075-1338) and KiCadGeometry (pathfinder/kicad_geometry.py:15-72).
Node index n = z*X*Y + y*X + x.  Layer directions (UPF:1091-1106): a 2-layer
board allows both axes on both layers; otherwise layer 0 is vertical, odd
layers horizontal, even layers vertical.

Two edge policies:
  reference  what the reference builds: planar edges on layers 1..L-2 only
             (all layers when L <= 2, UPF:1210), and a via between every
             ordered pair of routing layers plus layer 0 <-> every routing
             layer (UPF:1136-1181).  Each ordered pair is visited once per
             direction, so every via appears twice per direction (UPF:1266-1276).
  adjacent   planar edges on every layer, vias only between adjacent layers,
             one edge per direction.
"""
DIR_V, DIR_H, DIR_HV = 0, 1, 2
POLICIES = ("reference", "adjacent")


def reference_via_pairs(L):
    """Ordered via layer pairs in the order the reference iterates its Python
    set of pairs (UPF:1136-1181); CPython's set of small-int tuples iterates in
    a deterministic order that depends only on the insertion sequence, which is
    reproduced here."""
    if L == 2:
        return list({(0, 1), (1, 0)})
    routing = list(range(1, L - 1))
    pairs = set()
    for a in routing:
        for b in routing:
            if a != b:
                pairs.add((a, b))
    for z in routing:
        pairs.add((0, z))
        pairs.add((z, 0))
    return list(pairs)


def layer_directions(L):
    if L == 2:
        return [DIR_HV, DIR_HV]
    return [DIR_V if z == 0 else (DIR_H if z % 2 == 1 else DIR_V) for z in range(L)]


def lattice_spec(L, policy="reference"):
    """(planar flags, directions, ordered via pairs) for an L-layer lattice."""
    dirs = layer_directions(L)
    if policy == "reference":
        planar = [1] * L if L <= 2 else [0] + [1] * (L - 2) + [0]
        pairs = reference_via_pairs(L)
    elif policy == "adjacent":
        planar = [1] * L
        pairs = [(z, z + 1) for z in range(L - 1)]
    else:
        raise ValueError(f"unknown lattice policy {policy!r} (expected one of {POLICIES})")
    return planar, dirs, pairs


def edge_count(X, Y, L, policy="reference"):
    planar, dirs, pairs = lattice_spec(L, policy)
    e = 0
    for z in range(L):
        if not planar[z]:
            continue
        if dirs[z] in (DIR_H, DIR_HV):
            e += 2 * Y * (X - 1)
        if dirs[z] in (DIR_V, DIR_HV):
            e += 2 * X * (Y - 1)
    return e + 2 * X * Y * len(pairs)


def node_index(X, Y, x, y, z):
    return (z * Y + y) * X + x


def node_coords(X, Y, n):
    plane = X * Y
    z, r = divmod(int(n), plane)
    y, x = divmod(r, X)
    return x, y, z
