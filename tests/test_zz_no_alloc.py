"""orgpu_route_ref_noalloc (static device scratch, no allocation) against
orgpu_route_ref (handle scratch) through the C ABI: identical paths, edge ids
and float32 distances on two calls in a row, the second with a smaller input
(a smaller lattice, fewer requests, fewer slots), and a clear error for a
request larger than the static scratch."""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def noalloc_call(ctx, lat, cost_d, rows, roi_d, own_d, slots, heap_cap, cap):
    """One call of the static-scratch entry point; outputs compacted to
    request order like device.route_ref."""
    import torch
    from orthoroute_amd import _abi
    n = len(rows)
    dev = ctx.device
    d_req = torch.from_numpy(rows.view(np.uint8)).to(dev)
    nodes = torch.empty(cap, dtype=torch.int32, device=dev)
    edges = torch.empty(cap, dtype=torch.int32, device=dev)
    off = torch.zeros(n, dtype=torch.int64, device=dev)
    ln = torch.zeros(n, dtype=torch.int32, device=dev)
    dist = torch.zeros(n, dtype=torch.float32, device=dev)
    cnt = torch.zeros(4, dtype=torch.int64, device=dev)
    g = lat.graph(cost_d)
    n_roi = 0 if roi_d is None else int(roi_d.shape[0])
    rc = ctx.lib.orgpu_route_ref_noalloc(0, ctypes.c_void_p(ctx.stream.cuda_stream), ctypes.byref(g), _p(d_req), n,
                                         _p(roi_d), n_roi, _p(own_d), ctypes.c_float(1e9), slots, heap_cap,
                                         _p(nodes), _p(edges), cap, _p(off), _p(ln), _p(dist), _p(cnt))
    _abi.check(rc)
    ctx.sync()
    lens = ln.cpu().numpy()
    assert (lens >= 0).all(), f"request errors {lens[lens < 0]}"
    o = off.cpu().numpy()
    allnodes, alledges = nodes.cpu().numpy(), edges.cpu().numpy()
    offs = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    idx = np.repeat(o - offs[:-1], lens) + np.arange(int(offs[-1]), dtype=np.int64)
    return allnodes[idx], alledges[idx], offs, dist.cpu().numpy()


def case(ctx, X, Y, L, policy, n, seed):
    import torch
    from orthoroute_amd.device import rank_exceptions, ref_requests, roi_bfs
    rng = np.random.default_rng(seed)
    o = oracle.Lattice(X, Y, L, policy)
    lat = ctx.build_lattice(X, Y, L, policy)
    c = (o.costs * np.float32(0.3)).astype(np.float32)
    c += (np.arange(o.E, dtype=np.float32) % 9973) * np.float32(1e-6)
    hot = rng.random(o.E) < 0.1
    c[hot] += rng.integers(1, 9, hot.sum()).astype(np.float32)
    cost_d = torch.from_numpy(c).to(ctx.device)
    owner = np.full(o.N, -1, np.int32)
    sel = rng.random(o.N) < 0.05
    owner[sel] = rng.integers(0, 4, sel.sum())
    own_d = torch.from_numpy(owner).to(ctx.device)
    spec = [(int(rng.integers(0, o.N)), int(rng.integers(0, o.N))) for _ in range(n)]
    roi_d = torch.zeros((n, (o.N + 31) // 32), dtype=torch.int32, device=ctx.device)
    rq = [(s, d, oracle.roi_radius(o, s, d)[1], int(rng.choice([200, 50_000])), j) for j, (s, d) in enumerate(spec)]
    info = roi_bfs(ctx, lat, rq, roi_d, readd=True, raw=True)
    reqs = [(s, d, j % 4, j if j % 3 else -1, int(j % 5 == 0)) for j, (s, d) in enumerate(spec)]
    ranks = [rank_exceptions(info[j], o.N) if j % 3 else None for j in range(n)]
    return o, lat, c, cost_d, owner, own_d, roi_d, reqs, ranks, ref_requests(reqs, ranks)


@pytest.fixture(scope="module")
def ctx():
    from orthoroute_amd.device import Context
    return Context(0)


def test_noalloc_equals_route_ref_two_calls(ctx):
    from orthoroute_amd.device import route_ref
    for step, (X, Y, L, policy, n, seed, slots, heap_cap) in enumerate([
            (28, 24, 6, "reference", 120, 1, 16, 8192),
            (14, 12, 4, "adjacent", 40, 2, 3, 1024)]):
        o, lat, c, cost_d, owner, own_d, roi_d, reqs, ranks, rows = case(ctx, X, Y, L, policy, n, seed)
        need = 256 + slots * (o.N + heap_cap) * 16
        say(f"[noalloc] call {step + 1}: {X}x{Y}x{L} {policy}, {n} requests, {slots} slots, heap {heap_cap}, "
            f"{need} of {64 << 20} scratch bytes")
        a_nodes, a_edges, a_offs, a_dist = noalloc_call(ctx, lat, cost_d, rows, roi_d, own_d, slots, heap_cap,
                                                        cap=n * 4 * (X + Y + 2 * L))
        b_nodes, b_edges, b_offs, b_dist, _ = route_ref(ctx, lat, cost_d, reqs, roi_bitmaps=roi_d,
                                                        keepout_owner=own_d, ranks=ranks, slots=slots,
                                                        heap_cap=heap_cap)
        np.testing.assert_array_equal(a_offs, b_offs)
        np.testing.assert_array_equal(a_nodes, b_nodes)
        np.testing.assert_array_equal(a_edges, b_edges)
        np.testing.assert_array_equal(a_dist.view(np.uint32), b_dist.view(np.uint32))
        # and the oracle's emulation of the reference heap on a sample
        for j in range(0, n, 7):
            s, d, nid, r, ko = reqs[j]
            lst = np.arange(o.N, dtype=np.int32)
            if r >= 0:
                row = roi_d[r].cpu().numpy().view(np.uint32)
                mask = ((row[np.arange(o.N) >> 5] >> (np.arange(o.N) & 31)) & 1).astype(bool)
                lst = np.nonzero(mask)[0].astype(np.int32)
                order = dict(ranks[j] or [])
                keys = np.array([order.get(int(v), 2 * int(v) + 2) for v in lst], dtype=np.int64)
                lst = lst[np.argsort(keys, kind="stable")]
            path, _ = oracle.sssp_ref(o, c, s, d, lst, keepout_owner=owner if ko else None, net_id=nid)
            assert a_nodes[a_offs[j]:a_offs[j + 1]].tolist() == (path or []), f"call {step + 1} request {j}"


def test_noalloc_rejects_oversize(ctx):
    import torch
    from orthoroute_amd import _abi
    o, lat, c, cost_d, owner, own_d, roi_d, reqs, ranks, rows = case(ctx, 14, 12, 4, "adjacent", 4, 3)
    d_req = torch.from_numpy(rows.view(np.uint8)).to(ctx.device)
    out = torch.empty(1024, dtype=torch.int32, device=ctx.device)
    off = torch.zeros(4, dtype=torch.int64, device=ctx.device)
    ln = torch.zeros(4, dtype=torch.int32, device=ctx.device)
    g = lat.graph(cost_d)
    slots = (64 << 20) // ((o.N + (1 << 20)) * 16) + 1
    rc = ctx.lib.orgpu_route_ref_noalloc(0, ctypes.c_void_p(ctx.stream.cuda_stream), ctypes.byref(g), _p(d_req), 4,
                                         _p(roi_d), int(roi_d.shape[0]), None, ctypes.c_float(1e9), slots, 1 << 20,
                                         _p(out), _p(out), 1024, _p(off), _p(ln), None, None)
    assert rc == _abi.ORGPU_EINVAL
    assert b"static scratch" in ctx.lib.orgpu_last_error()
