"""Host reductions that reproduce the reference's numpy arithmetic exactly.

The negotiation loop makes decisions on float sums of per-edge arrays
(overuse totals, per-layer loads).  Those arrays live in HBM here; the GPU
returns only their nonzero entries, and these helpers rebuild the exact value
numpy would have produced on the dense array:

* ``sum_f32(n, idx, val)`` == ``np.sum(a)`` for the float32 array ``a`` of
  length n that is zero except ``a[idx] = val``.  NumPy reduces a contiguous
  float32 array in buffers of 8192 elements, adding each buffer's pairwise sum
  (8 accumulators below 129 elements, halving above, split points multiples
  of 8) to a float32 running total that starts at 0.  Zeros contribute exactly
  nothing, so only blocks holding nonzeros need evaluating.
* ``bincount_f64(keys, val, n)`` == ``np.bincount(keys, weights=val, minlength=n)``
  (float64, sequential in index order).
"""
import numpy as np

_BUF = 8192
_PW = 128
_f32 = np.float32


def _pw_dense(a):
    """numpy's float32 pairwise sum of a dense block (len <= 128)."""
    n = len(a)
    if n < 8:
        r = _f32(0.0)
        for x in a:
            r = _f32(r + x)
        return r
    r = a[:8].astype(np.float32).copy()
    i = 8
    stop = n - (n % 8)
    while i < stop:
        r = (r + a[i:i + 8]).astype(np.float32)
        i += 8
    res = _f32(_f32(_f32(r[0] + r[1]) + _f32(r[2] + r[3])) + _f32(_f32(r[4] + r[5]) + _f32(r[6] + r[7])))
    while i < n:
        res = _f32(res + a[i])
        i += 1
    return res


def sum_f32(n, idx, val):
    """np.sum of a length-n float32 array that is zero except a[idx] = val
    (idx without repeats)."""
    idx = np.asarray(idx, dtype=np.int64)
    val = np.asarray(val, dtype=np.float32)
    if len(idx) == 0:
        return _f32(0.0)
    if len(idx) > 2048:
        # many nonzeros: materialising the dense array and letting numpy sum
        # it is the reference's own operation and cheaper than the block walk
        a = np.zeros(n, dtype=np.float32)
        a[idx] = val
        return np.sum(a)
    order = np.argsort(idx, kind="stable")
    idx, val = idx[order], val[order]
    total = _f32(0.0)
    # buffer boundaries that hold nonzeros
    bufs = np.unique(idx // _BUF)
    for b in bufs:
        lo = int(b) * _BUF
        m = min(_BUF, n - lo)
        p0 = int(np.searchsorted(idx, lo, side="left"))
        p1 = int(np.searchsorted(idx, lo + m, side="left"))
        total = _f32(total + _pw_sparse(lo, m, idx, val, p0, p1))
    return total


def _pw_sparse(lo, n, idx, val, p0, p1):
    """Pairwise sum of positions [lo, lo+n) whose nonzeros are idx[p0:p1]."""
    if p0 == p1:
        return _f32(0.0)
    if n <= _PW:
        block = np.zeros(n, dtype=np.float32)
        block[idx[p0:p1] - lo] = val[p0:p1]
        return _pw_dense(block)
    n2 = n // 2
    n2 -= n2 % 8
    mid = p0 + int(np.searchsorted(idx[p0:p1], lo + n2, side="left"))
    return _f32(_pw_sparse(lo, n2, idx, val, p0, mid) +
                _pw_sparse(lo + n2, n - n2, idx, val, mid, p1))


def bincount_f64(keys, val, n):
    """np.bincount(keys, weights=val, minlength=n): float64, index order."""
    return np.bincount(np.asarray(keys, dtype=np.int64), weights=np.asarray(val, dtype=np.float64),
                       minlength=n)
