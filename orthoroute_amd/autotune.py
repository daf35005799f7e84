"""Search-parameter selection (bucket width delta, ring size) for the SSSP.

The bucket width trades re-relaxation (wide buckets: more Bellman-Ford rounds
inside a bucket) against round count (narrow buckets: more kernel launches).
`choose_params` returns a cached choice per lattice shape; `tune` measures
candidate (delta, n_buckets) pairs on a sample of the caller's searches and
records the fastest.  Results are identical for every choice: only speed
changes.
"""
import json
import os
import time

_CACHE = {}
_CACHE_FILE = os.environ.get("ORGPU_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "orgpu_tune.json"))


def _key(router, slots):
    lat = router.lattice
    return f"{lat.x_steps}x{lat.y_steps}x{lat.layers}:{lat.policy}:{int(slots > 1)}"


def _load():
    if _CACHE or not os.path.exists(_CACHE_FILE):
        return
    try:
        with open(_CACHE_FILE) as f:
            _CACHE.update(json.load(f))
    except (OSError, ValueError):
        pass


def choose_params(router, slots, cost_scale=0.0):
    """(delta, n_buckets) for this lattice; a measured entry when one exists,
    otherwise a width of a few typical planar steps.  cost_scale: the
    iteration's median planar cost rounded up to a power of two (0 = unknown,
    then the base cost of a step); congestion raises it in later iterations,
    and the width follows so that a search keeps a similar number of
    buckets."""
    _load()
    k = _key(router, slots)
    if k in _CACHE:
        d, nb = _CACHE[k]
        return float(d), int(nb)
    if cost_scale > 0.0:
        return max(2.4 * float(cost_scale), 1e-3), 8
    step = float(router.config.base_cost_weight) * float(router.lattice.pitch)
    return max(step * 4.0, 1e-3), 8


def tune(router, sample_reqs, candidates=None, slots=None):
    """Time candidate (delta, n_buckets) on sample_reqs (list of
    (src, dst, net_id, use_roi, bonus)) with the router's current costs and
    keep the fastest for this lattice shape."""
    slots = slots or router.config.batch_slots
    step = float(router.config.base_cost_weight) * float(router.lattice.pitch)
    candidates = candidates or [(step * m, nb) for m in (1, 2, 4, 8, 16) for nb in (4, 8, 16)]
    best, best_t = None, float("inf")
    saved = (router.config.sssp_delta, router.config.sssp_buckets)
    try:
        for d, nb in candidates:
            router.config.sssp_delta, router.config.sssp_buckets = d, nb
            t0 = time.perf_counter()
            router._search_batch(sample_reqs)
            t = time.perf_counter() - t0
            if t < best_t:
                best, best_t = (d, nb), t
    finally:
        router.config.sssp_delta, router.config.sssp_buckets = saved
    _CACHE[_key(router, slots)] = list(best)
    try:
        os.makedirs(os.path.dirname(_CACHE_FILE), exist_ok=True)
        with open(_CACHE_FILE, "w") as f:
            json.dump(_CACHE, f)
    except OSError:
        pass
    return best, best_t
