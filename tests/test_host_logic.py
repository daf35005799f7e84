"""Host-side bookkeeping of the router (no GPU): the vectorised forms against
straightforward per-net loops that restate the reference's semantics.

  ordering   UPF:4401-4439 (score, sort, rotation)
  via usage  UPF:2678-2762 (column / segment counts, first keepout owner)
  ownership  UPF:2764-2837, 2880-2904 (last barrel writer wins)
  conflicts  UPF:5478-5569 (edges touching a node owned by another net)
"""
import random

import numpy as np
import pytest

from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderRouter


def bare_router(X, Y, L):
    r = PathFinderRouter.__new__(PathFinderRouter)
    r._Nx, r._Ny, r._Nz = X, Y, L
    r.N = X * Y * L
    r._segZ = max(L - 2, 0)
    r.via_col_use = np.zeros((X, Y), np.int16)
    r.via_seg_use = np.zeros((X, Y, r._segZ), np.int16)
    r.node_owner = np.full(r.N, -1, np.int32)
    r.net_paths, r._net_nodes, r._net_to_edges = {}, {}, {}
    r.net_id_map, r.next_net_id = {}, 0
    r._ko_nodes = np.zeros(0, np.int64)
    r._ko_owner = np.zeros(0, np.int64)
    return r


def random_paths(r, n_nets, rng):
    """Random lattice walks (planar steps and via hops of span 1-3)."""
    X, Y, L = r._Nx, r._Ny, r._Nz
    plane = X * Y
    for i in range(n_nets):
        x, y, z = rng.integers(0, X), rng.integers(0, Y), rng.integers(0, L)
        path = [int(z * plane + y * X + x)]
        for _ in range(rng.integers(1, 25)):
            if rng.random() < 0.3:
                z2 = int(np.clip(z + rng.choice([-3, -2, -1, 1, 2, 3]), 0, L - 1))
                if z2 == z:
                    continue
                z = z2
            elif rng.random() < 0.5:
                x = int(np.clip(x + rng.choice([-1, 1]), 0, X - 1))
            else:
                y = int(np.clip(y + rng.choice([-1, 1]), 0, Y - 1))
            n = int(z * plane + y * X + x)
            if n != path[-1]:
                path.append(n)
        name = f"N{i:03d}"
        if len(path) < 2:
            r.net_paths[name] = []
            continue
        r.net_paths[name] = path
        r._net_nodes[name] = np.asarray(path, np.int64)
        r._net_to_edges[name] = rng.integers(0, 10_000, len(path) - 1).astype(np.int64)
        r._get_net_id(name)


def loop_via_usage(r):
    """The reference's per-path loops: column/segment use, keepouts
    (first owner wins), then node ownership (last barrel wins)."""
    X, Y, L = r._Nx, r._Ny, r._Nz
    plane = X * Y
    col = np.zeros((X, Y), np.int64)
    seg = np.zeros((X, Y, max(L - 2, 0)), np.int64)
    keep = {}
    owner = np.full(r.N, -1, np.int64)
    for name, path in r.net_paths.items():
        if not path or len(path) < 2:
            continue
        nid = r.net_id_map[name]
        for a, b in zip(path, path[1:]):
            za, ra = divmod(a, plane)
            zb, rb = divmod(b, plane)
            if ra != rb or za == zb:
                continue
            x, y = ra % X, ra // X
            lo, hi = min(za, zb), max(za, zb)
            col[x, y] += 1
            s0, s1 = max(1, min(lo, L - 2)), max(1, min(hi, L - 2))
            for zz in range(s0, s1):
                seg[x, y, zz - 1] += 1
            for zz in range(lo, hi + 1):
                keep.setdefault(zz * plane + y * X + x, nid)
                owner[zz * plane + y * X + x] = nid
    return col, seg, keep, owner


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_via_usage_keepouts_and_ownership(seed):
    rng = np.random.default_rng(seed)
    r = bare_router(13, 9, 6)
    random_paths(r, 40, rng)
    col, seg, keep, owner = loop_via_usage(r)
    r._rebuild_via_usage_from_committed(r._flat())
    np.testing.assert_array_equal(r.via_col_use, col)
    np.testing.assert_array_equal(r.via_seg_use, seg)
    assert dict(zip(r._ko_nodes.tolist(), r._ko_owner.tolist())) == keep
    np.testing.assert_array_equal(r.node_owner, owner)


@pytest.mark.parametrize("seed", [3, 4])
def test_mark_barrels_many_equals_per_path_marks(seed):
    rng = np.random.default_rng(seed)
    r = bare_router(11, 8, 5)
    random_paths(r, 30, rng)
    paths = [np.asarray(p, np.int64) for p in r.net_paths.values() if p]
    ids = [r.net_id_map[n] for n, p in r.net_paths.items() if p]
    r.node_owner[:] = -1
    for p, i in zip(paths, ids):
        r._mark_barrels(p.tolist(), i)
    expect = r.node_owner.copy()
    r.node_owner[:] = -1
    r._mark_barrels_many(paths, ids)
    np.testing.assert_array_equal(r.node_owner, expect)


def test_barrel_conflicts():
    rng = np.random.default_rng(5)
    r = bare_router(10, 10, 4)
    random_paths(r, 30, rng)
    r._rebuild_via_usage_from_committed(r._flat())
    got, n = r._detect_barrel_conflicts(r._flat())
    exp = []
    for name, path in r.net_paths.items():
        if not path or len(path) < 2:
            continue
        nid = r.net_id_map[name]
        for (a, b), e in zip(zip(path, path[1:]), r._net_to_edges[name]):
            oa, ob = r.node_owner[a], r.node_owner[b]
            if (oa != -1 and oa != nid) or (ob != -1 and ob != nid):
                exp.append(int(e))
    assert n == len(exp) and got.tolist() == exp


@pytest.mark.parametrize("seed,n_over", [(6, 200), (7, 200), (8, 4000)])
def test_order_nets_matches_loop_form(seed, n_over):
    rng = np.random.default_rng(seed)
    r = bare_router(37, 23, 5)
    plane = 37 * 23
    tasks = {}
    for i in range(300):
        s, d = rng.integers(0, r.N, 2)
        tasks[f"N{i:03d}"] = (int(s), int(d))
        if rng.random() < 0.7:
            r._net_to_edges[f"N{i:03d}"] = rng.integers(0, 5000, rng.integers(0, 30)).astype(np.int64)
    over = np.zeros(5000, np.float32)
    over[rng.integers(0, 5000, n_over)] = (rng.random(n_over) * 3).astype(np.float32)
    r._rng = random.Random(seed)
    got = r._order_nets_by_difficulty(tasks, over)

    def coords(n):
        z, rr = divmod(n, plane)
        return rr % 37, rr // 37, z
    scores = []
    for nid, (s, d) in tasks.items():
        sx, sy, sz = coords(s)
        dx, dy, dz = coords(d)
        dist = abs(dx - sx) + abs(dy - sy) + abs(dz - sz)
        cong = 0.0
        if nid in r._net_to_edges:
            e = r._net_to_edges[nid]
            cong = sum(float(over[int(x)]) for x in e) / max(1, len(e))
        scores.append((dist * 3 * (cong + 1), nid))
    scores.sort(reverse=True)
    rot = random.Random(seed).randint(0, min(5, len(scores) // 10))
    exp = [n for _, n in scores]
    if rot > 0:
        exp = exp[rot:] + exp[:rot]
    assert got == exp


def geometry_router(X, Y, L, policy, pitch, origin):
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import Lattice3D, PathFinderConfig
    r = bare_router(X, Y, L)
    r.config = PathFinderConfig()
    r.config.grid_pitch = pitch
    r.config.layer_names = ["F.Cu"] + [f"In{i}.Cu" for i in range(1, L - 1)] + ["B.Cu"]
    ox, oy = origin
    r.lattice = Lattice3D((ox, oy, ox + (X - 1) * pitch, oy + (Y - 1) * pitch), pitch, L, policy)
    assert (r.lattice.x_steps, r.lattice.y_steps) == (X, Y)
    return r


def lattice_walk(rng, X, Y, L, steps, irregular):
    from oracle.geometry import axis
    x, y, z = int(rng.integers(0, X)), int(rng.integers(0, Y)), int(rng.integers(0, L))
    p = [z * X * Y + y * X + x]
    for _ in range(steps):
        u = rng.random()
        if irregular and u < 0.04:
            x = int(np.clip(x + rng.integers(-2, 3), 0, X - 1))      # not a lattice move
            y = int(np.clip(y + rng.integers(-2, 3), 0, Y - 1))
        elif u < 0.2:
            z = int(np.clip(z + rng.choice([-2, -1, 1, 2]), 0, L - 1))
        elif axis(L, z) == "h" or (axis(L, z) == "hv" and rng.random() < 0.5):
            x = int(np.clip(x + rng.choice([-1, 1]), 0, X - 1))
        else:
            y = int(np.clip(y + rng.choice([-1, 1]), 0, Y - 1))
        n = z * X * Y + y * X + x
        if n != p[-1]:
            p.append(n)
    return p


@pytest.mark.parametrize("policy,L,irregular", [("reference", 6, False), ("reference", 6, True),
                                                ("adjacent", 5, True), ("reference", 2, False)])
def test_geometry_matches_per_hop_oracle(policy, L, irregular):
    """The vectorised emission equals the per-hop restatement of
    UPF:5749-5830 (oracle.geometry) record for record."""
    from oracle import geometry as G
    rng = np.random.default_rng(L * 10 + irregular)
    X, Y, pitch, origin = 17, 13, 0.5, (2.0, 3.0)
    r = geometry_router(X, Y, L, policy, pitch, origin)
    r.net_paths = {f"N{i:03d}": lattice_walk(rng, X, Y, L, int(rng.integers(0, 60)), irregular) for i in range(80)}
    r.net_paths["EMPTY"] = []
    got_t, got_v = r._generate_geometry_from_paths()
    g = r.lattice.geom
    exp_t, exp_v = G.generate(r.net_paths, X, Y, L, r.config.layer_names, g.pitch, g.grid_min_x, g.grid_min_y,
                              guard_outer=(policy == "reference"))
    for t in exp_t:
        t["width"] = r.config.grid_pitch * 0.6
    assert got_t == exp_t
    assert got_v == exp_v
