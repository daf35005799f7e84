"""orgpu_via_pressure (the via pooling pressure EMAs, the segment prefix and
the via overuse sum) bit for bit against the numpy float32 expressions of
the negotiation's iteration start (oracle/negotiation.py route_multiple_nets,
UPF:3626-3637), over several iterations of changing usage, with pressures
large and small enough that a fused multiply-add or a reordered sum would
change bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("X,Y,segz,seed", [(37, 23, 4, 1), (120, 90, 16, 2), (64, 64, 0, 3), (5, 3, 22, 4)])
def test_via_pressure_matches_numpy(X, Y, segz, seed):
    import torch
    from orthoroute_amd import _abi
    from orthoroute_amd.device import Context
    print(f"[via pressure] {X}x{Y} segments={segz} seed={seed}", flush=True)
    ctx = Context(0)
    dev = ctx.device
    rng = np.random.default_rng(seed)
    col_cap, seg_cap = 4, 2
    col_pres = (rng.random((X, Y)) * 1e3).astype(np.float32)
    seg_pres = (rng.random((X, Y, segz)) * rng.choice([1e-3, 1.0, 7e4], (X, Y, segz))).astype(np.float32)
    seg_prefix = np.zeros((X, Y, segz), np.float32)
    d = {k: torch.from_numpy(v.reshape(-1).copy()).to(dev) for k, v in
         (("col_pres", col_pres), ("seg_pres", seg_pres), ("seg_prefix", seg_prefix))}
    if segz == 0:
        for k in ("seg_pres", "seg_prefix"):
            d[k] = torch.zeros(1, dtype=torch.float32, device=dev)
    over_d = torch.zeros(1, dtype=torch.int64, device=dev)
    for it in range(4):
        col_use = rng.integers(0, 9, (X, Y)).astype(np.int16)
        seg_use = rng.integers(0, 5, (X, Y, segz)).astype(np.int16)
        cu = torch.from_numpy(col_use.reshape(-1).copy()).to(dev)
        su = torch.from_numpy(seg_use.reshape(-1).copy()).to(dev) if segz else torch.zeros(1, dtype=torch.int16,
                                                                                             device=dev)
        _abi.check(ctx.lib.orgpu_via_pressure(ctx.handle, X * Y, segz, cu.data_ptr(), col_cap, su.data_ptr(), seg_cap,
                                              d["col_pres"].data_ptr(), d["seg_pres"].data_ptr(),
                                              d["seg_prefix"].data_ptr(), over_d.data_ptr()))
        # the iteration start's numpy expressions
        over = np.maximum(0, col_use - np.int16(col_cap)).astype(np.float32)
        col_pres = 0.6 * over + 0.4 * col_pres
        exp_over = int(np.maximum(0, col_use - np.int16(col_cap)).sum())
        if segz:
            over = np.maximum(0, seg_use - np.int8(seg_cap)).astype(np.float32)
            seg_pres = 0.6 * over + 0.4 * seg_pres
            np.cumsum(seg_pres, axis=2, out=seg_prefix)
            exp_over += int(np.maximum(0, seg_use - np.int8(seg_cap)).sum())
        assert col_pres.dtype == np.float32 and seg_pres.dtype == np.float32
        np.testing.assert_array_equal(d["col_pres"].cpu().numpy().view(np.uint32),
                                      col_pres.reshape(-1).view(np.uint32))
        if segz:
            np.testing.assert_array_equal(d["seg_pres"].cpu().numpy().view(np.uint32),
                                          seg_pres.reshape(-1).view(np.uint32))
            np.testing.assert_array_equal(d["seg_prefix"].cpu().numpy().view(np.uint32),
                                          seg_prefix.reshape(-1).view(np.uint32))
        assert int(over_d.item()) == exp_over
    ctx.close()
