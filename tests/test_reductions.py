"""Host reductions reproduce numpy's float32 sum / float64 bincount exactly."""
import numpy as np
import pytest

from orthoroute_amd.reductions import bincount_f64, sum_f32


@pytest.mark.parametrize("seed", range(6))
def test_sparse_sum_matches_numpy(seed):
    rng = np.random.default_rng(seed)
    for n in [1, 7, 8, 127, 128, 129, 1000, 8191, 8192, 8193, 16384, 16385, 24577, 50_000, 300_000, 441_600]:
        a = np.zeros(n, np.float32)
        k = max(1, int(n * rng.choice([0.00001, 0.001, 0.05, 0.5])))
        idx = rng.choice(n, min(k, n), replace=False)
        a[idx] = (rng.random(len(idx)) * rng.choice([1, 37.5, 1e4], len(idx))).astype(np.float32)
        ref = np.sum(a)
        got = sum_f32(n, idx, a[idx])
        assert got.view(np.uint32) == ref.view(np.uint32), (n, got, ref)


def test_sparse_sum_distinguishes_orders():
    # a case where sequential and pairwise float32 sums differ: the helper must
    # follow numpy, not a plain running sum
    rng = np.random.default_rng(9)
    n = 40_000
    a = (rng.random(n) * 1e4).astype(np.float32)
    seq = np.float32(0)
    for x in a:
        seq = np.float32(seq + x)
    assert seq != np.sum(a)
    assert sum_f32(n, np.arange(n), a) == np.sum(a)


def test_bincount():
    rng = np.random.default_rng(3)
    k = rng.integers(0, 7, 1000)
    w = rng.random(1000).astype(np.float32)
    np.testing.assert_array_equal(bincount_f64(k, w, 9), np.bincount(k, weights=w, minlength=9))
