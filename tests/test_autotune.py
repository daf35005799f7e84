"""The online bucket-target tuner on synthetic cost curves: it moves to the
cheapest target whether that lies above or below its start, holds it, and
re-probes the neighbours now and then."""
import math

import numpy as np
import pytest

from orthoroute_amd import autotune
from orthoroute_amd.autotune import TARGETS, BucketTuner


def run(cost_of, n=30, start_nodes=2_000_000, key=None, noise=0.0, seed=0):
    autotune._SETTLED.clear()
    rng = np.random.default_rng(seed)
    t = BucketTuner(key or "k", start_nodes)
    seen = []
    for _ in range(n):
        tg = t.target()
        seen.append(tg)
        t.record(cost_of(tg) * (1.0 + noise * rng.standard_normal()))
    return t, seen


@pytest.mark.parametrize("best", TARGETS)
def test_tuner_finds_the_cheapest_target(best):
    def cost(tg):   # each x4 step away from `best` costs 1.7x (measured: 1.75x / 2.3x per step on the backplane)
        return 1.7 ** abs(math.log(tg / best, 4))
    t, seen = run(cost)
    assert TARGETS[t.best] == best
    # after the search the target in use is mostly the best one
    assert seen[-12:].count(best) >= 9


def test_tuner_climbs_with_drifting_costs():
    """Costs that grow 3x over the negotiation (congestion building up) do
    not stop the climb to a target four steps away."""
    t, seen = run(lambda tg: 1.8 ** abs(math.log(tg / 8192, 4)), n=30)
    assert TARGETS[t.best] == 8192


def test_tuner_holds_under_noise_and_flat_costs():
    t, seen = run(lambda tg: 1.0, noise=0.05)
    assert TARGETS[t.best] == autotune.default_target(2_000_000)
    assert seen[-12:].count(128) >= 9


def test_tuner_remembers_the_settled_target_per_key():
    t, _ = run(lambda tg: 1.7 ** abs(math.log(tg / 8192, 4)), key="board")
    assert autotune._SETTLED["board"] == 8192
    assert BucketTuner("board", 2_000_000).target() == 8192
    assert BucketTuner("other", 2_000_000).target() == 128
    assert BucketTuner("small", 50_000).target() == 32


def test_tuner_starts_at_the_base_width_target_unless_settled():
    """With nothing settled for its key, the first target is the one nearest
    (log scale) to the depth / base-width ratio it is given; a settled
    target, or no ratio, keeps the start."""
    autotune._SETTLED.clear()
    assert BucketTuner("a", 2_000_000).target(ratio=31_627.0) == 32768
    assert BucketTuner("b", 2_000_000).target(ratio=2_300.0) == 2048
    assert BucketTuner("c", 2_000_000).target(ratio=3.0) == 32
    assert BucketTuner("d", 2_000_000).target() == 128
    assert BucketTuner("e", 2_000_000).target(ratio=float("inf")) == 128
    t = BucketTuner("f", 2_000_000)
    assert t.target(ratio=9_000.0) == 8192
    assert t.target(ratio=40.0) == 8192          # only the first call moves the start
    autotune._SETTLED["g"] = 512
    assert BucketTuner("g", 2_000_000).target(ratio=31_627.0) == 512
