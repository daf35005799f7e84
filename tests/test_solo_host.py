"""Host-side rules of the one-wave search engine, on the CPU.

* Node coordinates: the kernel divides a node index (below 2^28) by the
  layer plane X*Y and by X as (v * m) >> sh with m = ceil(2^sh / d),
  sh = 28 + ceil(log2 d) (sssp_solo.hip, solo_magic / magic_div).  Checked
  against true integer division for every divisor up to 3000, the benchmark
  boards' planes and widths, random divisors up to 2^28, and dividends at
  the multiples of d and its neighbours, 0 and 2^28 - 1.
* Key tiles: the 2x2x4 tile index is a bijection from a lattice's cells
  into [0, tiles * 16) (no two nodes share a key), and the tile of 16 keys
  is one 128-byte line.
* Bucket lists: the radix-list rule puts a bucket above the current one in
  the list of the highest hex digit where they differ, and every entry of
  that list agrees with the new current bucket above that digit once the
  list is chosen (so moving it only ever goes to lower lists).
"""
import random

import numpy as np


def magic(d):
    lg = 0
    while (1 << lg) < d:
        lg += 1
    sh = 28 + lg
    return ((1 << sh) + d - 1) // d, sh


def test_magic_division_is_exact():
    rng = random.Random(7)
    divisors = list(range(1, 3001)) + [600 * 400, 600, 1000 * 800, 1000, 300 * 300, 96 * 96, (1 << 28) - 1]
    divisors += [rng.randrange(1, 1 << 28) for _ in range(500)]
    for d in divisors:
        m, sh = magic(d)
        assert (1 << 28) * m < (1 << 64)            # the 64-bit product never overflows
        vs = {0, 1, d - 1, d, d + 1, (1 << 28) - 1}
        vs.update(rng.randrange(0, 1 << 28) for _ in range(50))
        vs.update(k * d + o for k in range(1, 40) for o in (-1, 0, 1))
        for v in vs:
            if 0 <= v < (1 << 28):
                assert (v * m) >> sh == v // d, (d, v)


def tile_index(x, y, z, X, Y):
    TX, TY = (X + 1) // 2, (Y + 1) // 2
    t = ((z >> 2) * TY + (y >> 1)) * TX + (x >> 1)
    return t * 16 + ((z & 3) << 2) + ((y & 1) << 1) + (x & 1)


def test_tile_index_is_a_bijection_into_the_padded_space():
    for X, Y, L in [(1, 1, 1), (5, 4, 2), (6, 5, 3), (7, 6, 8), (24, 20, 6), (33, 17, 12), (16, 14, 18)]:
        z, y, x = np.meshgrid(np.arange(L), np.arange(Y), np.arange(X), indexing="ij")
        idx = tile_index(x.ravel(), y.ravel(), z.ravel(), X, Y)
        span = ((X + 1) // 2) * ((Y + 1) // 2) * ((L + 3) // 4) * 16
        assert len(np.unique(idx)) == X * Y * L
        assert idx.min() >= 0 and idx.max() < span
        # a tile (16 keys: one aligned 128-byte line) holds cells that differ
        # by one step along x, y or up to three layers
        tiles = idx // 16
        for t in np.unique(tiles)[:50]:
            sel = tiles == t
            assert np.ptp(x.ravel()[sel]) <= 1 and np.ptp(y.ravel()[sel]) <= 1 and np.ptp(z.ravel()[sel]) <= 3


def radix_list(b, cur):
    if b <= cur:
        return 0
    i = ((b ^ cur).bit_length() - 1) >> 2
    return 1 + i * 16 + ((b >> (4 * i)) & 15)


def test_radix_lists_only_move_down():
    rng = random.Random(3)
    for _ in range(2000):
        cur = rng.randrange(0, 1 << 37)
        b = cur + rng.choice([0, 1, 2, 15, 16, 17, 255, 4096, 1 << 20, 1 << 32, 5 << 32]) + rng.randrange(0, 3)
        t = radix_list(b, cur)
        if t == 0:
            assert b <= cur
            continue
        i, d = (t - 1) // 16, (t - 1) % 16
        assert (b >> (4 * i)) & 15 == d and d > (cur >> (4 * i)) & 15
        assert b >> (4 * (i + 1)) == cur >> (4 * (i + 1))
        # choosing that list sets the current bucket to its lower bound
        low = (1 << (4 * (i + 1))) - 1
        new = (cur & ~low) | (d << (4 * i))
        assert new <= b
        assert radix_list(b, new) < t


def test_wave_count_rule():
    """Wavefronts per search: an explicit 1, 2 or 4 is kept; otherwise four
    when every search of the call starts at once (n <= slots, n <= 1024) and
    two when slots refill or more than 1024 run at once."""
    from orthoroute_amd.device import solo_waves
    assert [solo_waves(w, 64, 300) for w in (1, 2, 4)] == [1, 2, 4]
    assert solo_waves(0, 1024, 1024) == 4
    assert solo_waves(0, 2048, 1024) == 4
    assert solo_waves(0, 2048, 1025) == 2
    assert solo_waves(0, 1024, 40000) == 2
    assert solo_waves(0, 64, 65) == 2
    assert solo_waves(0, 64, 1) == 4


def test_small_batch_engine_switch_thresholds():
    """A negotiation's small batches move to the one-wave engine for good once
    a bucket-engine batch took more than twice a board-covering one-wave
    batch (N/256 steps of 20 us, at least 1/4 s); one-wave batches and quick
    bucket-engine batches leave the choice alone."""
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    r = PathFinderRouter(PathFinderConfig())
    r.N = 100_000                        # 7.8 ms of one-wave steps: the 1/4 s floor holds
    assert r._small_solo_time() == 0.25
    r._small_engine_record(0, 0.45)
    r._small_engine_record(1, 5.0)
    assert not r.__dict__.get("_small_engine")
    r._small_engine_record(0, 0.55)
    assert r._small_engine == 1 and r.stats["small_engine_switch_measured"] == 1
    r._small_engine_record(1, 9.0)      # no way back
    assert r._small_engine == 1
    big = PathFinderRouter(PathFinderConfig())
    big.N = 1_920_000                    # 150 ms of one-wave steps
    assert abs(big._small_solo_time() - 0.25) < 1e-12
    big.N = 4_000_000
    assert abs(big._small_solo_time() - 4_000_000 / 256 * 20e-6) < 1e-12
