"""Oracle lattice vs the reference's own CSR arrays and closed forms."""
import json
import os

import numpy as np
import pytest

import oracle


@pytest.mark.parametrize("shape", [(5, 4, 2), (6, 5, 3), (8, 6, 4), (12, 10, 6), (7, 6, 8)])
def test_lattice_matches_reference_golden(golden_dir, shape):
    X, Y, L = shape
    g = np.load(os.path.join(golden_dir, f"lattice_{X}x{Y}x{L}.npz"))
    lat = oracle.Lattice(X, Y, L, "reference", pitch=1.0, via_cost=3.0)
    np.testing.assert_array_equal(lat.indptr, g["indptr"])
    np.testing.assert_array_equal(lat.indices, g["indices"])
    np.testing.assert_array_equal(lat.costs.view(np.uint32), g["costs"].astype(np.float32).view(np.uint32))
    np.testing.assert_array_equal(lat.edge_layer, g["edge_layer"])
    np.testing.assert_array_equal(lat.edge_kind, g["edge_kind"])


def test_via_pair_order_matches_reference(golden_dir):
    with open(os.path.join(golden_dir, "via_pairs.json")) as f:
        ref = json.load(f)
    for L, pairs in ref.items():
        assert [list(p) for p in oracle.reference_via_pairs(int(L))] == pairs


def closed_form_edges(X, Y, L, policy):
    n = L - 2
    if policy == "reference":
        if L == 2:
            planar = 2 * Y * (X - 1) + 2 * X * (Y - 1)
            return 2 * planar + 2 * X * Y * 2
        h = sum(1 for z in range(1, L - 1) if z % 2 == 1)
        v = (L - 2) - h
        return h * 2 * Y * (X - 1) + v * 2 * X * (Y - 1) + X * Y * 2 * (n * (n - 1) + 2 * n)
    h = sum(1 for z in range(L) if z % 2 == 1)
    v = L - h
    return h * 2 * Y * (X - 1) + v * 2 * X * (Y - 1) + 2 * (L - 1) * X * Y


@pytest.mark.parametrize("policy", ["reference", "adjacent"])
@pytest.mark.parametrize("shape", [(8, 6, 4), (12, 10, 6), (9, 7, 5), (96, 96, 6)])
def test_lattice_closed_forms(policy, shape):
    X, Y, L = shape
    lat = oracle.Lattice(X, Y, L, policy)
    assert lat.E == closed_form_edges(X, Y, L, policy)
    # survey table: cfg1 96x96x6 has 441,600 (reference) / 201,600 (adjacent) edges
    if shape == (96, 96, 6):
        assert lat.E == {"reference": 441_600, "adjacent": 201_600}[policy]
    src = np.repeat(np.arange(lat.N), np.diff(lat.indptr))
    dst = lat.indices
    # symmetry: the multiset of (u, v) equals the multiset of (v, u)
    a = np.sort(src.astype(np.int64) * lat.N + dst)
    b = np.sort(dst.astype(np.int64) * lat.N + src)
    np.testing.assert_array_equal(a, b)
    xs, ys, zs = lat.coords(src)
    xd, yd, zd = lat.coords(dst)
    via = zs != zd
    np.testing.assert_array_equal(via.astype(np.uint8), lat.edge_kind)
    np.testing.assert_array_equal(zs.astype(np.uint8), lat.edge_layer)
    assert np.all((xs[via] == xd[via]) & (ys[via] == yd[via]))
    planar = ~via
    step = np.abs(xs - xd) + np.abs(ys - yd)
    assert np.all(step[planar] == 1)
    # H/V discipline: odd layers move in x only, even layers in y only (L > 2)
    if L > 2:
        odd = (zs % 2 == 1) & planar
        assert np.all(ys[odd] == yd[odd])
        even = (zs % 2 == 0) & planar
        assert np.all(xs[even] == xd[even])
    # via cost 3*(1+0.15*(span-1)), planar cost = pitch
    span = np.abs(zs - zd)
    exp = np.where(via, (3.0 * (1.0 + 0.15 * (span - 1))).astype(np.float32), np.float32(1.0))
    np.testing.assert_array_equal(lat.costs, exp.astype(np.float32))
