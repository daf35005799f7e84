"""GPU parity of the ROI BFS and the congestion-cost kernels against the oracle
and the reference's own recorded outputs (bit-exact)."""
import os
import sys

import numpy as np
import pytest

import oracle
from oracle.negotiation import OracleRouter

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    from orthoroute_amd.device import Context
    return Context(0)


def unpack(row, N):
    return np.unpackbits(row.view(np.uint8), bitorder="little")[:N].astype(bool)


@pytest.mark.parametrize("method", ["bfs", "lattice"])
def test_roi_matches_reference_golden(ctx, golden_dir, method):
    import torch
    from orthoroute_amd.device import roi_bfs
    g = np.load(os.path.join(golden_dir, "roi_bfs.npz"))
    X, Y, L = int(g["X"]), int(g["Y"]), int(g["L"])
    lat = ctx.build_lattice(X, Y, L, "reference")
    o = oracle.Lattice(X, Y, L, "reference")
    n = len(g["src"])
    words = (lat.N + 31) // 32
    rows = torch.zeros((n, words), dtype=torch.int32, device=ctx.device)
    reqs = []
    for i in range(n):
        s, d, b = int(g["src"][i]), int(g["dst"][i]), float(g["bonus"][i])
        _, depth = oracle.roi_radius(o, s, d, b)
        reqs.append((s, d, depth, int(g["max_nodes"][i]), i))
    say(f"[roi] {n} golden cases ({method})")
    info = roi_bfs(ctx, lat, reqs, rows, method=method)
    rows_h = rows.cpu().numpy().view(np.uint32)
    off = 0
    for i in range(n):
        k = int(g["roi_len"][i])
        ref = set(g["roi"][off:off + k].tolist()) | {reqs[i][0], reqs[i][1]}
        off += k
        got = set(np.nonzero(unpack(rows_h[i], lat.N))[0].tolist())
        assert got == ref, f"case {i}: {len(got)} vs {len(ref)}"
        assert info[i][1] == len(ref)


@pytest.mark.parametrize("method", ["bfs", "lattice"])
def test_roi_random_truncation(ctx, method):
    import torch
    from orthoroute_amd.device import roi_bfs
    rng = np.random.default_rng(4)
    # the (180, 150, 8) shape spans several 4096-word chunks of the ROI
    # bitmap; 1,100 nets on the last shape exceed the 1,024 concurrent
    # workgroups, so workgroups run several nets in turn on their scratch
    for policy, (X, Y, L), n in [("reference", (40, 30, 8), 40), ("adjacent", (50, 44, 6), 40),
                                 ("adjacent", (180, 150, 8), 40), ("reference", (30, 24, 6), 1100)]:
        lat = ctx.build_lattice(X, Y, L, policy)
        o = oracle.Lattice(X, Y, L, policy)
        words = (lat.N + 31) // 32
        rows = torch.zeros((n, words), dtype=torch.int32, device=ctx.device)
        reqs = []
        for i in range(n):
            s, d = (int(v) for v in rng.choice(lat.N, 2, replace=False))
            mx = int(rng.choice([50, 300, 2000, 50_000]))
            depth = int(rng.integers(0, 80 if X < 100 else 160))
            reqs.append((s, d, depth, mx, i))
        say(f"[roi] random {policy} {X}x{Y}x{L} ({method})")
        roi_bfs(ctx, lat, reqs, rows, method=method)
        rows_h = rows.cpu().numpy().view(np.uint32)
        for i, (s, d, depth, mx, _) in enumerate(reqs):
            out = np.empty(o.N, np.int32)
            R = oracle.lib().orc_roi_bfs(o.N, oracle._p(o.indptr), oracle._p(o.indices), s, d, depth, mx, None, 0, 0,
                                         oracle._p(out), None, None, None)
            ref = set(out[:R].tolist()) | {s, d}
            got = set(np.nonzero(unpack(rows_h[i], lat.N))[0].tolist())
            assert got == ref, f"{policy} case {i} (s={s} d={d} depth={depth} max={mx})"


def test_roi_lattice_equals_bfs_and_oracle(ctx):
    """The closed-form ROI against the BFS kernel (rows and every info field)
    and the oracle's BFS (marks, depth, meet flag) on every layer count from
    1 to 8, both via policies, with depth caps from 0 to 300, node caps from
    3 to 50,000, with and without the re-add, and src == dst (oracle only:
    both of its waves start at the node)."""
    import ctypes
    import torch
    from orthoroute_amd.device import roi_bfs
    rng = np.random.default_rng(11)
    for policy in ("reference", "adjacent"):
        for L in range(1, 9):
            X, Y = int(rng.integers(20, 70)), int(rng.integers(15, 60))
            lat = ctx.build_lattice(X, Y, L, policy)
            o = oracle.Lattice(X, Y, L, policy)
            n = 48
            words = (lat.N + 31) // 32
            reqs = []
            for i in range(n):
                s, d = (int(v) for v in rng.choice(lat.N, 2, replace=False))
                if i % 12 == 0:
                    d = s
                depth = int(rng.choice([0, 1, 2, 7, 25, 60, 150, 300]))
                mx = int(rng.choice([3, 50, 600, 5000, 50_000]))
                reqs.append((s, d, depth, mx, i))
            for readd in (True, False):
                say(f"[roi] lattice vs bfs {policy} {X}x{Y}x{L} readd={readd}")
                r1 = torch.zeros((n, words), dtype=torch.int32, device=ctx.device)
                r2 = torch.full((n, words), -1, dtype=torch.int32, device=ctx.device)
                i1 = roi_bfs(ctx, lat, reqs, r1, readd=readd, method="bfs")
                i2 = roi_bfs(ctx, lat, reqs, r2, readd=readd, method="lattice")
                h1, h2 = r1.cpu().numpy(), r2.cpu().numpy()
                for i, (s, d, depth, mx, _) in enumerate(reqs):
                    out = np.empty(o.N, np.int32)
                    seen = np.empty(o.N, np.uint8)
                    dd, mm = ctypes.c_int(0), ctypes.c_int(0)
                    R = oracle.lib().orc_roi_bfs(o.N, oracle._p(o.indptr), oracle._p(o.indices), s, d, depth, mx, None, 0, 0,
                                                 oracle._p(out), oracle._p(seen), ctypes.byref(dd), ctypes.byref(mm))
                    ref = set(out[:R].tolist()) | ({s, d} if readd else set())
                    got = set(np.nonzero(unpack(h2[i], lat.N))[0].tolist())
                    tag = f"{policy} {X}x{Y}x{L} case {i} (s={s} d={d} depth={depth} max={mx} readd={readd})"
                    assert got == ref, tag
                    assert (i2[i][2], i2[i][3]) == (dd.value, mm.value), tag
                    assert i2[i][0] == int((seen != 0).sum()), tag
                    if s != d:
                        np.testing.assert_array_equal(h1[i], h2[i], err_msg=tag)
                        assert i1[i] == i2[i], tag


def _dev(ctx, a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)


def test_cost_kernels_match_reference_golden(ctx, golden_dir):
    """EdgeAccountant.update_costs / update_present_ema / update_history as the
    reference computed them (tests/golden/costs.npz)."""
    import ctypes
    import torch
    from orthoroute_amd import _abi
    from orthoroute_amd.device import EdgeState
    g = np.load(os.path.join(golden_dir, "costs.npz"))
    X, Y, L = int(g["X"]), int(g["Y"]), int(g["L"])
    lat = ctx.build_lattice(X, Y, L, "reference")
    st = EdgeState(ctx, lat)
    st.present_ema.copy_(_dev(ctx, g["ema0"]))
    st.history.copy_(_dev(ctx, g["hist0"]))
    lb = _dev(ctx, g["layer_bias"].astype(np.float32))

    def params(pf, hw, vcm, bcw, bias):
        p = _abi.CostParams()
        p.X, p.Y, p.L = X, Y, L
        p.via_cost_mult, p.base_cost_weight = np.float32(vcm), np.float32(bcw)
        p.pres_fac, p.hist_weight = np.float32(pf), np.float32(hw)
        p.jitter_scale, p.add_jitter = np.float32(1e-6), 1
        p.layer_bias = ctypes.c_void_p(bias.data_ptr())
        return p
    say("[costs] update_costs with layer bias")
    st.update_costs(params(3.7, 12.0, 0.5, 0.3, lb))
    np.testing.assert_array_equal(st.total_cost.cpu().numpy()[:lat.E].view(np.uint32), g["total_a"].view(np.uint32))
    ones = torch.ones(L, dtype=torch.float32, device=ctx.device)
    st.update_costs(params(1.0, 10.0, 1.0, 0.01, ones))
    np.testing.assert_array_equal(st.total_cost.cpu().numpy()[:lat.E].view(np.uint32), g["total_b"].view(np.uint32))
    # present_ema from present (= integer counts in the golden)
    st.count.copy_(_dev(ctx, g["present0"].astype(np.int32)))
    idx, val = st.update_ema(False, np.float32(0.4), np.float32(1.0 - 0.4))
    np.testing.assert_array_equal(st.present_ema.cpu().numpy()[:lat.E].view(np.uint32), g["ema1"].view(np.uint32))
    over = np.maximum(0, g["ema1"] - 1).astype(np.float32)
    np.testing.assert_array_equal(idx, np.nonzero(over > 0)[0])
    np.testing.assert_array_equal(val, over[over > 0])
    # history with raw present, cap 15*base, decay 0.98
    st.build_present()
    st.update_history(np.float32(0.98), np.float32(0.25), np.float32(15.0), True)
    np.testing.assert_array_equal(st.history.cpu().numpy()[:lat.E].view(np.uint32), g["hist1"].view(np.uint32))


def test_via_pooling_and_hard_block_match_oracle(ctx):
    import ctypes
    import torch
    from orthoroute_amd import _abi
    from orthoroute_amd.device import EdgeState
    rng = np.random.default_rng(21)
    for policy, (X, Y, L) in [("reference", (14, 11, 8)), ("adjacent", (16, 9, 6))]:
        for f64 in (False, True):
            say(f"[via] {policy} f64={f64}")
            o = OracleRouter(X, Y, L, policy, pitch=1.0, via_cost=3.0)
            E = o.E
            o.present_ema = (rng.random(E) * 3).astype(np.float32)
            o.history = (rng.random(E) * 2).astype(np.float32)
            o.via_col_pres = (rng.random((X, Y)) * (rng.random((X, Y)) < 0.3) * 2).astype(np.float32)
            o.via_seg_pres = (rng.random((X, Y, L - 2)) * (rng.random((X, Y, L - 2)) < 0.3)).astype(np.float32)
            np.cumsum(o.via_seg_pres, axis=2, out=o.via_seg_prefix)
            o.via_col_use = rng.integers(0, 6, (X, Y)).astype(np.int16)
            o.via_seg_use = rng.integers(0, 3, (X, Y, L - 2)).astype(np.int16)
            lb = (1 + 0.8 * rng.random(L)).astype(np.float32)
            o.layer_bias = (0.75 + rng.random(L)).astype(np.float64 if f64 else np.float32)
            pf = 5.3
            o.update_costs(o.lat.costs, pf, 9.6, vcm=0.5, bcw=0.3, edge_layer=o.lat.edge_layer,
                           layer_bias=lb, edge_kind=o.lat.edge_kind)
            o.apply_via_pooling(pf)
            o.block_via_edges()
            exp = o.total_cost
            lat = ctx.build_lattice(X, Y, L, policy)
            st = EdgeState(ctx, lat)
            st.present_ema.copy_(_dev(ctx, o.present_ema))
            st.history.copy_(_dev(ctx, o.history))
            # leave-hot-layer discount table in the layer_bias dtype
            k = 0.20
            b = o.layer_bias
            disc = np.ones((L, L), np.float64)
            for zl in range(L):
                for zh in range(L):
                    d = (1.0 - k * np.maximum(b[zl:zl + 1], 0.0)) * (1.0 + 0.5 * k * np.maximum(b[zh:zh + 1], 0.0))
                    disc[zl, zh] = float(d[0])
            keep = [_dev(ctx, lb), _dev(ctx, o.via_col_pres), _dev(ctx, o.via_seg_prefix), _dev(ctx, disc),
                    _dev(ctx, o.via_col_use), _dev(ctx, o.via_seg_use)]
            p = _abi.CostParams()
            p.X, p.Y, p.L = X, Y, L
            p.via_cost_mult, p.base_cost_weight = np.float32(0.5), np.float32(0.3)
            p.pres_fac, p.hist_weight = np.float32(pf), np.float32(9.6)
            p.jitter_scale, p.add_jitter = np.float32(1e-6), 1
            p.layer_bias = ctypes.c_void_p(keep[0].data_ptr())
            p.via_pooling, p.seg_z = 1, L - 2
            p.col_pres, p.seg_prefix = ctypes.c_void_p(keep[1].data_ptr()), ctypes.c_void_p(keep[2].data_ptr())
            p.col_weight, p.seg_weight = np.float32(1.0), np.float32(1.0)
            p.via_discount, p.discount_f64 = ctypes.c_void_p(keep[3].data_ptr()), int(f64)
            p.hard_block = 1
            p.col_use, p.seg_use = ctypes.c_void_p(keep[4].data_ptr()), ctypes.c_void_p(keep[5].data_ptr())
            p.col_cap, p.seg_cap = 4, 2
            st.update_costs(p)
            got = st.total_cost.cpu().numpy()[:E]
            np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


@pytest.mark.parametrize("method", ["lattice", "bfs"])
@pytest.mark.parametrize("readd", [False, True])
def test_roi_must_keeps_match_oracle(method, readd):
    """Portal seeds as must-keeps on both ROI kernels: the row equals the
    oracle list's set, and the front swaps and appended nodes it reports are
    the oracle list's own (truncations at 150-1,000 nodes, seeds outside the
    ball, seeds equal to an endpoint)."""
    import torch
    from orthoroute_amd.device import Context, roi_bfs
    ctx = Context(0)
    X, Y, L = 30, 24, 6
    lat = ctx.build_lattice(X, Y, L, "reference")
    o = oracle.Lattice(X, Y, L, "reference")
    rng = np.random.default_rng(17)
    plane = X * Y
    reqs, keeps, lists = [], [], []
    for j in range(60):
        s, d = (int(v) for v in rng.integers(plane, o.N, 2))
        kind = j % 3
        if kind == 0:
            kp = [s % plane, d % plane]
        elif kind == 1:
            kp = [int(rng.integers(0, o.N)), -1]
        else:
            kp = [s, int(rng.integers(0, o.N))]
        mx = int(rng.choice([3, 5, 150, 300, 1000, 50_000]))
        b = float(rng.choice([0.0, 0.6]))
        _, depth = oracle.roi_radius(o, s, d, b)
        reqs.append((s, d, depth, mx, j))
        keeps.append(kp)
        lists.append(oracle.roi_bfs(o, s, d, stagnation_bonus=b, max_nodes=mx, keep=[k for k in kp if k >= 0],
                                    readd=readd))
    rows = torch.zeros((len(reqs), (o.N + 31) // 32), dtype=torch.int32, device=ctx.device)
    info = roi_bfs(ctx, lat, reqs, rows, readd=readd, method=method, raw=True, keeps=np.asarray(keeps))
    h = rows.cpu().numpy()
    for j, lst in enumerate(lists):
        got = set(np.nonzero(unpack(h[j], o.N))[0].tolist())
        assert got == set(lst.tolist()), f"request {j}: keeps {keeps[j]}"
        assert int(info["kept"][j]) == len(lst)
        for p in range(4):
            if info["front"][j][p] >= 0:
                assert lst[p] == info["front"][j][p], f"request {j}: front {p}"
        app = [int(v) for v in info["appended"][j] if v >= 0]
        # appended nodes sit after every node that stayed in sorted order
        assert lst[len(lst) - len(app):].tolist() == app, f"request {j}: appended {app}"
