"""The device impact sums of the hotset and the rip-up (orgpu_segment_sums
through device.edge_list_sums): per edge list, Python's
sum(float(v[e]) for e in edges if e in hot) -- float64, one addition after
another in list order -- against that loop as the oracle's hotset and rip-up
write it, the router's vectorised host form, and sums whose value the order
of the additions fixes."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    from orthoroute_amd.device import Context
    return Context(0)


def python_sums(lists, idx, val):
    """The oracle's impact loop (oracle/negotiation.py build_hotset)."""
    v = dict(zip(np.asarray(idx).tolist(), np.asarray(val, np.float32).tolist()))
    return [sum(float(np.float32(v[e])) for e in lst.tolist() if e in v) for lst in lists]


@pytest.mark.parametrize("seed,n_lists,n_edges,n_hot", [(1, 300, 5000, 400), (2, 2000, 200000, 20000),
                                                        (3, 50, 1000, 1000)])
def test_edge_list_sums_match_python_sum(ctx, seed, n_lists, n_edges, n_hot):
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderRouter
    from orthoroute_amd.device import edge_list_sums
    say(f"[segment sums] seed={seed} lists={n_lists} edges={n_edges} hot={n_hot}")
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 700, n_lists)
    lens[::17] = 0                                     # empty lists
    lens[1::29] = 64                                   # exactly one wave chunk
    lists = [rng.integers(0, n_edges, k).astype(np.int32) for k in lens]
    idx = rng.choice(n_edges, n_hot, replace=False)
    val = (rng.random(n_hot) * rng.choice([1e-3, 1.0, 37.0, 1e4], n_hot)).astype(np.float32)
    val[::11] = 0.0                                    # hot with zero overuse
    got = edge_list_sums(ctx, n_edges, lists, idx, val)
    exp = python_sums(lists, idx, val)
    assert got.tolist() == exp
    dense = np.zeros(n_edges, np.float32)
    dense[idx] = val
    host = PathFinderRouter._path_sums(lists, dense)
    assert got.tolist() == host.tolist()


def test_edge_list_sums_add_in_list_order(ctx):
    """2^53 + 1 rounds back to 2^53 (ties to even), so only the
    left-to-right order gives 2^53 for [2^53, 1, 1]; a reversed or pairwise
    order gives 2^53 + 2."""
    from orthoroute_amd.device import edge_list_sums
    big = np.float32(2.0 ** 53)
    idx = np.array([0, 1, 2], np.int64)
    val = np.array([big, 1.0, 1.0], np.float32)
    lists = [np.array([0, 1, 2], np.int32), np.array([1, 2, 0], np.int32),
             np.array([0] + [1] * 100, np.int32), np.array([1] * 100 + [0], np.int32)]
    got = edge_list_sums(ctx, 8, lists, idx, val)
    assert got.tolist() == [2.0 ** 53, 2.0 ** 53 + 2.0, 2.0 ** 53, 2.0 ** 53 + 100.0]


def test_edge_list_sums_reject_bad_ids(ctx):
    import torch
    from orthoroute_amd import _abi
    from orthoroute_amd.device import _ptr, edge_list_sums
    with pytest.raises(ValueError):
        edge_list_sums(ctx, 10, [np.array([1, 10], np.int32)], [1], [1.0])
    with pytest.raises(ValueError):
        edge_list_sums(ctx, 10, [np.array([1, -1], np.int32)], [1], [1.0])
    # the kernel itself reads no id outside the range and returns NaN for it
    dev = ctx.device
    dense = torch.ones(10, dtype=torch.float32, device=dev)
    e = torch.tensor([1, 2, 12, 3, 4], dtype=torch.int32, device=dev)
    offs = torch.tensor([0, 2, 5], dtype=torch.int64, device=dev)
    out = torch.empty(2, dtype=torch.float64, device=dev)
    _abi.check(ctx.lib.orgpu_segment_sums(ctx.handle, 10, _ptr(dense), _ptr(e), _ptr(offs), 2, _ptr(out)))
    r = out.cpu().numpy()
    assert r[0] == 2.0 and np.isnan(r[1])
