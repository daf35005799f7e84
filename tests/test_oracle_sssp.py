"""Oracle SSSP and ROI BFS vs the reference's own outputs, scipy and brute force."""
import itertools
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra as sp_dijkstra

import oracle


def _golden(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def test_roi_bfs_matches_reference(golden_dir):
    g = _golden(golden_dir, "roi_bfs.npz")
    lat = oracle.Lattice(int(g["X"]), int(g["Y"]), int(g["L"]))
    off = 0
    for i in range(len(g["src"])):
        n = int(g["roi_len"][i])
        ref = g["roi"][off:off + n]
        off += n
        got = oracle.roi_bfs(lat, int(g["src"][i]), int(g["dst"][i]),
                             stagnation_bonus=float(g["bonus"][i]), max_nodes=int(g["max_nodes"][i]))
        np.testing.assert_array_equal(got, ref, err_msg=f"case {i}")


def _roi_for(lat, src, dst, mode):
    if mode == 0:
        return oracle.roi_bfs(lat, src, dst)
    return np.arange(lat.N, dtype=np.int32)


def test_ref_dijkstra_matches_reference_paths(golden_dir):
    g = _golden(golden_dir, "dijkstra.npz")
    lat = oracle.Lattice(int(g["X"]), int(g["Y"]), int(g["L"]))
    off = 0
    for i in range(len(g["src"])):
        s, src = divmod(int(g["src"][i]), 1_000_000)
        dst = int(g["dst"][i])
        n = int(g["path_len"][i])
        ref = g["paths"][off:off + n].tolist()
        off += n
        roi = _roi_for(lat, src, dst, int(g["mode"][i]))
        path, _ = oracle.sssp_ref(lat, g["snaps"][s], src, dst, roi)
        assert (path or []) == ref, f"case {i}"


def test_canon_cost_within_1e5_of_reference(golden_dir):
    g = _golden(golden_dir, "dijkstra.npz")
    lat = oracle.Lattice(int(g["X"]), int(g["Y"]), int(g["L"]))
    off = 0
    for i in range(len(g["src"])):
        s, src = divmod(int(g["src"][i]), 1_000_000)
        dst = int(g["dst"][i])
        n = int(g["path_len"][i])
        ref = g["paths"][off:off + n].tolist()
        off += n
        costs = g["snaps"][s]
        roi = _roi_for(lat, src, dst, int(g["mode"][i]))
        mask = np.zeros(lat.N, np.uint8); mask[roi] = 1
        path, dist, pred, _, _ = oracle.sssp_canon(lat, costs, src, dst, roi_mask=mask)
        if not ref:
            assert path is None
            continue
        c_ref = oracle.path_min_cost(lat, costs, ref)
        c_can = oracle.path_min_cost(lat, costs, path)
        assert abs(c_can - c_ref) <= 1e-5 * c_ref, (i, c_can, c_ref)


def _scipy_dist(lat, costs, src, mask=None):
    rows = np.repeat(np.arange(lat.N), np.diff(lat.indptr))
    keep = np.isfinite(costs)
    if mask is not None:
        keep &= (mask[rows] > 0) & (mask[lat.indices] > 0)
    # parallel edges: keep the minimum (csr_matrix would sum duplicates)
    r, c, w = rows[keep], lat.indices[keep], costs[keep].astype(np.float64)
    order = np.lexsort((w, c, r))
    r, c, w = r[order], c[order], w[order]
    first = np.ones(len(r), bool)
    first[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
    m = sp.csr_matrix((w[first], (r[first], c[first])), shape=(lat.N, lat.N))
    return sp_dijkstra(m, directed=True, indices=src)


@pytest.mark.parametrize("policy", ["reference", "adjacent"])
def test_canon_distances_match_scipy(policy):
    lat = oracle.Lattice(14, 11, 6, policy)
    rng = np.random.default_rng(5)
    costs = (lat.costs * (0.3 + 2 * rng.random(lat.E))).astype(np.float32)
    costs[(lat.edge_kind == 1) & (rng.random(lat.E) < 0.1)] = np.inf
    for trial in range(6):
        src = int(rng.integers(0, lat.N))
        d64 = _scipy_dist(lat, costs, src)
        far = np.where(np.isfinite(d64))[0]
        dst = int(far[np.argmax(d64[far])])
        path, dist, pred, _, _ = oracle.sssp_canon(lat, costs, src, dst)
        assert path is not None and path[0] == src and path[-1] == dst
        # every node at least as close as dst is settled: fp32 distance within
        # accumulated rounding of the float64 distance
        settled = d64 <= d64[dst] * (1 - 1e-5)
        np.testing.assert_allclose(dist[settled], d64[settled], rtol=2e-6, atol=1e-6)
        # invariants: every hop is tight in fp32, pred is the lowest tight pred
        for v in path[1:]:
            u = pred[v]
            s, e = lat.indptr[u], lat.indptr[u + 1]
            ws = costs[s:e][lat.indices[s:e] == v]
            assert np.any(dist[u] + ws == dist[v])
        for v in path[1:]:
            tight = []
            rows = np.repeat(np.arange(lat.N), np.diff(lat.indptr))
            for e_ in np.nonzero(lat.indices == v)[0]:
                u = rows[e_]
                if np.float32(dist[u] + costs[e_]) == dist[v]:
                    tight.append(u)
            assert pred[v] == min(tight)


def test_canon_brute_force_tiny():
    """Enumerate all simple paths on tiny random digraphs."""
    rng = np.random.default_rng(3)

    class G:
        pass
    for trial in range(40):
        n = 7
        adj = []
        for u in range(n):
            nb = sorted(set(rng.integers(0, n, 3).tolist()) - {u})
            adj.append(nb)
        g = G()
        g.N = n
        g.indptr = np.zeros(n + 1, np.int32)
        g.indptr[1:] = np.cumsum([len(a) for a in adj])
        g.indices = np.array([v for a in adj for v in a], dtype=np.int32)
        g.edge_kind = np.zeros(len(g.indices), np.uint8)
        w = rng.integers(1, 4, len(g.indices)).astype(np.float32) * np.float32(0.5)
        best = {}
        def dfs(u, cost, seen, path):
            if u == n - 1:
                key = (cost, )
                if (n - 1) not in best or cost < best[n - 1][0]:
                    best[n - 1] = (cost, list(path))
                return
            for e in range(g.indptr[u], g.indptr[u + 1]):
                v = int(g.indices[e])
                if v not in seen:
                    seen.add(v); path.append(v)
                    dfs(v, cost + float(w[e]), seen, path)
                    path.pop(); seen.discard(v)
        dfs(0, 0.0, {0}, [0])
        path, dist, pred, _, _ = oracle.sssp_canon(g, w, 0, n - 1)
        if n - 1 not in best:
            assert path is None
        else:
            assert dist[n - 1] == np.float32(best[n - 1][0])


# ---------------------------------------------------------------------------
# multi-source / multi-sink (UPF:1889-1989)
# ---------------------------------------------------------------------------

def _ms_cases(g):
    rows = g["rows"]
    off = 0
    for row in rows:
        s, mode, so, sn, to, tn, plen = (int(v) for v in row)
        seeds = list(zip(g["seed_node"][so:so + sn].tolist(), g["seed_cost"][so:so + sn].tolist()))
        targets = list(zip(g["target_node"][to:to + tn].tolist(), g["target_cost"][to:to + tn].tolist()))
        ref = g["paths"][off:off + plen].tolist()
        off += plen
        yield s, mode, seeds, targets, ref


def test_ms_ref_matches_reference_paths(golden_dir):
    g = _golden(golden_dir, "multisource.npz")
    lat = oracle.Lattice(int(g["X"]), int(g["Y"]), int(g["L"]))
    n = 0
    for i, (s, mode, seeds, targets, ref) in enumerate(_ms_cases(g)):
        roi = _roi_for(lat, seeds[0][0], targets[0][0], mode) if mode == 1 else \
            oracle.roi_bfs(lat, seeds[0][0], targets[0][0], max_nodes=50_000)
        path, _ = oracle.ms_ref(lat, g["snaps"][s], seeds, targets, roi)
        assert (path or []) == ref, f"case {i}"
        n += 1
    assert n == len(g["rows"])


def test_ms_canon_total_within_1e5_of_reference(golden_dir):
    g = _golden(golden_dir, "multisource.npz")
    lat = oracle.Lattice(int(g["X"]), int(g["Y"]), int(g["L"]))
    for i, (s, mode, seeds, targets, ref) in enumerate(_ms_cases(g)):
        roi = np.arange(lat.N, dtype=np.int32) if mode == 1 else \
            oracle.roi_bfs(lat, seeds[0][0], targets[0][0], max_nodes=50_000)
        _, best_ref = oracle.ms_ref(lat, g["snaps"][s], seeds, targets, roi)
        mask = np.zeros(lat.N, np.uint8)
        mask[roi] = 1
        path, k, best, dist, pred, _, _ = oracle.ms_canon(lat, g["snaps"][s], seeds, targets, roi_mask=mask)
        assert (k >= 0) == np.isfinite(best_ref), f"case {i}"
        if k >= 0:
            assert abs(best - best_ref) <= 1e-5 * max(1.0, abs(best_ref)), (i, best, best_ref)
            if ref and path is not None:
                # the canonical path is a valid seed -> target walk of that total
                assert path[0] in [a for a, _ in seeds] and path[-1] == targets[k][0]


def test_ms_canon_brute_force_tiny():
    """All simple paths from any seed to any target (targets are not passed
    through: they are never expanded) on tiny random digraphs."""
    rng = np.random.default_rng(17)

    class G:
        pass
    for trial in range(60):
        n = 8
        adj = [sorted(set(rng.integers(0, n, 3).tolist()) - {u}) for u in range(n)]
        g = G()
        g.N = n
        g.indptr = np.zeros(n + 1, np.int32)
        g.indptr[1:] = np.cumsum([len(a) for a in adj])
        g.indices = np.array([v for a in adj for v in a], dtype=np.int32)
        w = rng.integers(1, 5, len(g.indices)).astype(np.float32) * np.float32(0.5)
        perm = rng.permutation(n)
        seeds = [(int(perm[0]), float(rng.integers(0, 3) * 0.5)), (int(perm[1]), float(rng.integers(0, 3) * 0.5))]
        targets = [(int(perm[2]), float(rng.integers(0, 3) * 0.5)), (int(perm[3]), float(rng.integers(0, 3) * 0.5))]
        tset = {t for t, _ in targets}
        best = np.inf

        def dfs(u, cost, seen):
            nonlocal best
            for t, tc in targets:
                if u == t:
                    best = min(best, cost + tc)
            if u in tset:
                return
            for e in range(g.indptr[u], g.indptr[u + 1]):
                v = int(g.indices[e])
                if v not in seen:
                    seen.add(v)
                    dfs(v, cost + float(w[e]), seen)
                    seen.discard(v)
        for s, sc in seeds:
            dfs(s, sc, {s})
        path, k, tot, dist, pred, _, _ = oracle.ms_canon(g, w, seeds, targets)
        if not np.isfinite(best):
            assert k < 0
        else:
            assert k >= 0 and tot == best, (trial, tot, best)
            # the path walks CSR edges from a seed, every hop tight
            assert path[0] in [s for s, _ in seeds]
            for a, b in zip(path, path[1:]):
                row = range(g.indptr[a], g.indptr[a + 1])
                assert any(int(g.indices[e]) == b and np.float32(dist[a] + w[e]) == dist[b] for e in row)


def test_roi_with_portal_seeds_matches_reference(golden_dir):
    """extract_roi with portal seeds as must-keeps: the reference's lists
    (roi_keep.npz), order included, with truncations that swap must-keeps to
    the front and seeds the BFS never reaches."""
    g = _golden(golden_dir, "roi_keep.npz")
    lat = oracle.Lattice(int(g["X"]), int(g["Y"]), int(g["L"]))
    so = ro = 0
    for i in range(len(g["src"])):
        k = int(g["seed_len"][i])
        seeds = g["seeds"][so:so + k].tolist()
        so += k
        n = int(g["roi_len"][i])
        ref = g["roi"][ro:ro + n]
        ro += n
        got = oracle.roi_bfs(lat, int(g["src"][i]), int(g["dst"][i]), stagnation_bonus=float(g["bonus"][i]),
                             max_nodes=int(g["max_nodes"][i]), keep=seeds)
        np.testing.assert_array_equal(got, ref, err_msg=f"case {i}")


def test_canon_counts_many_keepout_edges_like_the_reference():
    """Paths that cross far more than 31 keepout edges: the canonical search
    counts every one (up to 254) and its path has the reference's keepout
    count and cost (the reference sums 1e9 per keepout edge in float64), on a
    lattice where every node belongs to another net and on one where a
    keepout-free detour exists but costs more."""
    X, Y, L = 300, 3, 2
    o = oracle.Lattice(X, Y, L, "adjacent")
    costs = o.costs.astype(np.float32)
    for owner_kind in ("all", "rows"):
        owner = np.full(o.N, 5, np.int32)
        if owner_kind == "rows":
            owner[2 * X:3 * X] = -1          # row 2 of layer 0 is free of keepouts
        src = X + 2
        for steps in (40, 120, 200):
            dst = src + steps
            path, dist, pred, *_ = oracle.sssp_canon(o, costs, src, dst, keepout_owner=owner, net_id=1)
            rpath, ralt = oracle.sssp_ref(o, costs, src, dst, np.arange(o.N), keepout_owner=owner, net_id=1)
            assert path is not None and rpath is not None

            def ko_cost(p):
                k, c = 0, 0.0
                for u, v in zip(p, p[1:]):
                    e = o.indptr[u] + int(np.nonzero(o.indices[o.indptr[u]:o.indptr[u + 1]] == v)[0][0])
                    if owner[u] >= 0 and owner[u] != 1 and o.edge_kind[e] == 0:
                        k += 1
                    else:
                        c += float(costs[e])
                return k, c
            kc, cc = ko_cost(path)
            kr, cr = ko_cost(rpath)
            assert kc == kr, (owner_kind, steps, kc, kr)
            assert abs(cc - cr) <= 1e-5 * max(cr, 1.0)
            if owner_kind == "all":
                assert kc == steps                  # every planar step leaves an owned node
