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
  nothing, so only buffers holding nonzeros need evaluating, and numpy itself
  evaluates them.
* ``bincount_f64(keys, val, n)`` == ``np.bincount(keys, weights=val, minlength=n)``
  (float64, sequential in index order).
"""
import numpy as np

_BUF = 8192
_f32 = np.float32


def sum_f32(n, idx, val):
    """np.sum of a length-n float32 array that is zero except a[idx] = val
    (idx without repeats).  Only the 8192-element buffers that hold nonzeros
    are materialised: full buffers as rows of one 2-D array, whose row sums
    (numpy's pairwise sum of each contiguous row) are the buffer sums, and a
    final partial buffer at its own length (its pairwise split points depend
    on its length, so it is not zero-padded).  The buffer sums are then added
    in order to a float32 total that starts at 0, as numpy does; empty
    buffers add exactly 0."""
    idx = np.asarray(idx, dtype=np.int64)
    val = np.asarray(val, dtype=np.float32)
    if len(idx) == 0:
        return _f32(0.0)
    if n <= _BUF:
        a = np.zeros(n, dtype=np.float32)
        a[idx] = val
        return np.sum(a)
    buf = idx // _BUF
    ub, row = np.unique(buf, return_inverse=True)
    last = (n - 1) // _BUF
    tail = n - last * _BUF                      # length of the last buffer
    sums = np.zeros(len(ub), dtype=np.float32)
    full = ub != last if tail < _BUF else np.ones(len(ub), dtype=bool)
    if full.any():
        rows_of_full = np.cumsum(full) - 1      # row of each full buffer in the block matrix
        sel = full[row]
        M = np.zeros((int(full.sum()), _BUF), dtype=np.float32)
        M[rows_of_full[row[sel]], idx[sel] - buf[sel] * _BUF] = val[sel]
        sums[full] = M.sum(axis=1)
    if not full.all():
        t = np.zeros(tail, dtype=np.float32)
        sel = buf == last
        t[idx[sel] - last * _BUF] = val[sel]
        sums[~full] = np.sum(t)
    # sequential float32 running total over the buffers, in buffer order
    return np.add.accumulate(sums, dtype=np.float32)[-1]


def bincount_f64(keys, val, n):
    """np.bincount(keys, weights=val, minlength=n): float64, index order."""
    return np.bincount(np.asarray(keys, dtype=np.int64), weights=np.asarray(val, dtype=np.float64),
                       minlength=n)
