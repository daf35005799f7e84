"""GPU parity of the solver seam (orthoroute_amd.solver.GpuSolver, the
reference's SimpleDijkstra.gpu_solver, UPF:2205-2218, 1814, 4558-4565) and of
SimpleDijkstra.find_path_multisource_multisink (UPF:1889-1989) with the
oracle's canonical searches."""
import sys
import types

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lattice_costs(o, rng):
    c = (o.costs * (0.2 + 3 * rng.random(o.E))).astype(np.float32)
    c[(o.edge_kind == 1) & (rng.random(o.E) < 0.1)] = np.inf
    return c


@pytest.fixture(scope="module")
def setup():
    from orthoroute_amd.device import Context
    from orthoroute_amd.solver import GpuSolver
    ctx = Context(0)
    X, Y, L = 30, 24, 6
    o = oracle.Lattice(X, Y, L, "reference")
    s = GpuSolver(o.indptr, o.indices, o.costs, X, Y, L, ctx=ctx, slots=16)
    return ctx, o, s


def test_find_path_roi_gpu(setup):
    ctx, o, s = setup
    say("[solver] find_path_roi_gpu")
    rng = np.random.default_rng(31)
    c = lattice_costs(o, rng)
    for t in range(12):
        src, dst = (int(v) for v in rng.integers(0, o.N, 2))
        roi = oracle.roi_bfs(o, src, dst) if t % 3 else np.arange(o.N, dtype=np.int32)
        g2r = np.full(o.N, -1, np.int32)
        g2r[roi] = np.arange(len(roi), dtype=np.int32)
        path = oracle.sssp_canon(o, c, src, dst, roi_mask=g2r >= 0)[0]
        costs = ctx.torch.from_numpy(c).to(ctx.device) if t % 2 else c
        assert s.find_path_roi_gpu(src, dst, costs, roi, g2r) == path, f"case {t}"


def test_find_paths_batch_with_rois_and_keepouts(setup):
    ctx, o, s = setup
    say("[solver] find_paths_batch")
    rng = np.random.default_rng(32)
    c = lattice_costs(o, rng)
    owner = np.full(o.N, -1, np.int32)
    owner[rng.integers(0, o.N, o.N // 8)] = rng.integers(0, 5, o.N // 8)
    nets, rois = [], []
    for i in range(40):
        src, dst = (int(v) for v in rng.integers(0, o.N, 2))
        nets.append((src, dst, i % 5))
        rois.append(None if i % 4 == 0 else oracle.roi_bfs(o, src, dst))
    got = s.find_paths_batch(c, nets, roi_spec=rois, keepout_spec=owner)
    for i, ((src, dst, nid), roi) in enumerate(zip(nets, rois)):
        mask = None
        if roi is not None:
            mask = np.zeros(o.N, np.uint8)
            mask[roi] = 1
        path = oracle.sssp_canon(o, c, src, dst, roi_mask=mask, keepout_owner=owner, net_id=nid)[0]
        assert got[i] == path, f"net {i}"


def test_find_path_fullgraph_gpu_seeds(setup):
    ctx, o, s = setup
    say("[solver] find_path_fullgraph_gpu_seeds")
    rng = np.random.default_rng(33)
    c = lattice_costs(o, rng)
    for t in range(10):
        src = rng.integers(0, o.N, int(rng.integers(1, 6))).astype(np.int32)
        dst = rng.integers(0, o.N, int(rng.integers(1, 6))).astype(np.int32)
        allowed = rng.random(o.N) < 0.85
        allowed[src] = True
        allowed[dst] = True
        bm = np.packbits(allowed.astype(np.uint8), bitorder="little")
        bm = np.concatenate([bm, np.zeros((-len(bm)) % 4, np.uint8)]).view(np.uint32)
        use = bool(t % 2)
        path, k, best, *_ = oracle.ms_canon(o, c, [(int(v), 0.0) for v in src], [(int(v), 0.0) for v in dst],
                                            allowed if use else None)
        got = s.find_path_fullgraph_gpu_seeds(c, src, dst, allowed_bitmap=bm, use_bitmap=use)
        if k < 0:
            assert got is None, f"case {t}"
        else:
            assert got == (path if path is not None else [int(dst[k])]), f"case {t}"
            if best > 0:
                assert s.find_path_fullgraph_gpu_seeds(c, src, dst, ub_hint=best * 0.5, allowed_bitmap=bm,
                                                       use_bitmap=use) is None


def test_generic_csr_without_lattice_shape(setup):
    """A CSR that is not a lattice: one layer of N nodes, plain bucket order."""
    from orthoroute_amd.solver import GpuSolver
    ctx, _, _ = setup
    say("[solver] generic CSR")
    rng = np.random.default_rng(34)
    N = 3000
    src = rng.integers(0, N, 5 * N)
    dst = (src + rng.integers(1, 40, 5 * N)) % N
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    indptr = np.zeros(N + 1, np.int32)
    np.add.at(indptr, src + 1, 1)
    indptr = np.cumsum(indptr).astype(np.int32)
    indices = dst.astype(np.int32)
    c = (0.5 + rng.random(len(indices)) * 4).astype(np.float32)
    g = types.SimpleNamespace(N=N, indptr=indptr, indices=indices, edge_kind=np.zeros(len(indices), np.uint8))
    s = GpuSolver(indptr, indices, c, ctx=ctx, slots=8)
    pairs = [tuple(int(v) for v in rng.integers(0, N, 2)) for _ in range(20)]
    got = s.find_paths_batch(c, pairs)
    for (a, b), p in zip(pairs, got):
        assert p == oracle.sssp_canon(g, c, a, b)[0]


def test_simple_dijkstra_multisource(setup):
    """The router's own SimpleDijkstra.find_path_multisource_multisink."""
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    say("[solver] SimpleDijkstra multisource")
    X, Y, L = 20, 16, 6
    r = PathFinderRouter(PathFinderConfig())
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    o = oracle.Lattice(X, Y, L, "reference")
    rng = np.random.default_rng(35)
    c = lattice_costs(o, rng)
    for t in range(10):
        seeds = [(int(rng.integers(0, o.N)), float(np.float32(rng.random() * 2))) for _ in range(3)]
        targets = [(int(rng.integers(0, o.N)), float(np.float32(rng.random() * 2))) for _ in range(4)]
        roi = oracle.roi_bfs(o, seeds[0][0], targets[0][0]) if t % 2 else np.arange(o.N, dtype=np.int32)
        g2r = np.full(o.N, -1, np.int32)
        g2r[roi] = np.arange(len(roi), dtype=np.int32)
        path, k, *_ = oracle.ms_canon(o, c, seeds, targets, g2r >= 0)
        got = r.solver.find_path_multisource_multisink(seeds, targets, c, roi, g2r)
        if path is None:
            assert got is None, f"case {t}"
        else:
            assert got == (path, path[0] // (X * Y), path[-1] // (X * Y)), f"case {t}"
