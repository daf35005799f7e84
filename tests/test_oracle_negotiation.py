"""Oracle negotiation (reference Dijkstra semantics) reproduces the reference's
own routing trajectories recorded in tests/golden/negotiation*.npz."""
import os

import numpy as np
import pytest

from oracle.negotiation import OracleRouter


def load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


@pytest.mark.parametrize("name", ["negotiation_easy.npz", "negotiation_l4.npz", "negotiation.npz"])
def test_trajectory_matches_reference(golden_dir, name):
    g = load(golden_dir, name)
    X, Y, L = int(g["X"]), int(g["Y"]), int(g["L"])
    r = OracleRouter(X, Y, L, "reference", pitch=1.0, via_cost=3.0, sssp="ref")
    reqs = [(f"N{i:03d}", int(s), int(d)) for i, (s, d) in enumerate(zip(g["src"], g["dst"]))]
    rec = {"hist": [], "ema": [], "over": [], "lb": []}
    orders, costs = [], []
    orig = r.order_nets

    def order_wrap(tasks):
        o = orig(tasks)
        orders.append(list(o))
        costs.append(r.total_cost.copy())
        return o
    r.order_nets = order_wrap

    def cb(it, tracks, vias, s, c):
        rec["over"].append((it, s, c))
        rec["hist"].append(r.history.copy())
        rec["ema"].append(r.present_ema.copy())
        rec["lb"].append(np.asarray(r.layer_bias, dtype=np.float64).copy())
    res = r.route_multiple_nets(reqs, max_iterations=int(g["cap"]), iteration_cb=cb)
    names = [q[0] for q in reqs]
    # per-iteration overuse sums/counts
    assert [tuple(x) for x in g["over"].tolist()] == rec["over"]
    # net routing order and the cost snapshot every _route_all saw
    off = 0
    assert len(orders) == len(g["order_len"])
    for i, n in enumerate(g["order_len"]):
        exp = [names[j] for j in g["order"][off:off + n]]
        off += n
        assert orders[i] == exp, f"order differs at call {i}"
        np.testing.assert_array_equal(costs[i].view(np.uint32), g["total_cost"][i].view(np.uint32),
                                      err_msg=f"total_cost differs at call {i}")
    for i in range(len(rec["hist"])):
        np.testing.assert_array_equal(rec["hist"][i].view(np.uint32), g["history"][i].view(np.uint32))
        np.testing.assert_array_equal(rec["ema"][i].view(np.uint32), g["ema"][i].view(np.uint32))
        np.testing.assert_array_equal(rec["lb"][i], g["layer_bias"][i])
    off = 0
    for i, n in enumerate(g["final_len"]):
        exp = g["final_nodes"][off:off + n].tolist()
        off += n
        assert list(res["paths"].get(names[i], [])) == exp, f"final path of {names[i]}"
    assert bool(res.get("success", False)) == bool(g["success"])


def _digest(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def run_oracle_compact(g, sssp="ref"):
    import json
    X, Y, L = int(g["X"]), int(g["Y"]), int(g["L"])
    width = 4 if len(g["src"]) >= 300 else 3
    reqs = [(f"N{i:0{width}d}", int(s), int(d)) for i, (s, d) in enumerate(zip(g["src"], g["dst"]))]
    idx = {q[0]: i for i, q in enumerate(reqs)}
    env = json.loads(str(g["env"]))
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        r = OracleRouter(X, Y, L, "reference", pitch=1.0, via_cost=3.0, sssp=sssp)
        orders, cost_digests, rec = [], [], []
        orig = r.order_nets

        def order_wrap(tasks):
            o = orig(tasks)
            orders.append([idx[nm] for nm in o])
            cost_digests.append(_digest(r.total_cost))
            return o
        r.order_nets = order_wrap

        def cb(it, tracks, vias, s, c):
            rec.append(((it, s, c), _digest(r.history), _digest(r.present_ema)))
        res = r.route_multiple_nets(reqs, max_iterations=int(g["cap"]), iteration_cb=cb)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return reqs, r, res, rec, orders, cost_digests


@pytest.mark.parametrize("name", ["negotiation_l16.npz", "negotiation_l24.npz"])
def test_many_layer_trajectory_matches_reference(golden_dir, name):
    """Boards with 14 and 22 signal layers take the schedule branches a
    board of 12 or fewer does not (pres_fac_max 64 / 128, history weight
    x1.0 / x0.8, UPF:3577-3585), and with the reference's pres_fac ramp
    override (ORTHO_PRES_FAC_MULT, UPF:3588) pres_fac passes 64 within the
    cap, where via annealing runs (UPF:3665-3687)."""
    g = load(golden_dir, name)
    reqs, r, res, rec, orders, cost_digests = run_oracle_compact(g)
    assert [tuple(x) for x in g["over"].tolist()] == [x[0] for x in rec]
    assert [x[1] for x in rec] == g["hist_digest"].tolist()
    assert [x[2] for x in rec] == g["ema_digest"].tolist()
    assert cost_digests == g["cost_digest"].tolist()
    off = 0
    for i, n in enumerate(g["order_len"]):
        assert orders[i] == g["order"][off:off + n].tolist(), f"order differs at call {i}"
        off += n
    np.testing.assert_array_equal(r.history.view(np.uint32), g["final_history"].view(np.uint32))
    off = 0
    for i, n in enumerate(g["final_len"]):
        assert list(res["paths"].get(reqs[i][0], [])) == g["final_nodes"][off:off + n].tolist()
        off += n


# Where the canonical search order (FP32 costs, lowest-predecessor ties: the
# GPU's batched engine) departs from the reference's own runs.  Path costs
# agree within 1e-5 relative, but on congested boards an equal-cost
# different path changes the usage, so the trajectories part.
CANON_DIVERGENCE = {
    # name: (first iteration whose overuse or history differs, final paths equal, nets)
    "negotiation_easy.npz": (None, 14, 14),
    "negotiation_l4.npz": (2, 5, 45),
    "negotiation.npz": (4, 31, 60),
}


@pytest.mark.parametrize("name", sorted(CANON_DIVERGENCE))
def test_canonical_order_divergence_is_recorded(golden_dir, name):
    g = load(golden_dir, name)
    X, Y, L = int(g["X"]), int(g["Y"]), int(g["L"])
    r = OracleRouter(X, Y, L, "reference", pitch=1.0, via_cost=3.0, sssp="canon")
    reqs = [(f"N{i:03d}", int(s), int(d)) for i, (s, d) in enumerate(zip(g["src"], g["dst"]))]
    rec = []
    res = r.route_multiple_nets(reqs, max_iterations=int(g["cap"]),
                                iteration_cb=lambda it, t, v, s, c: rec.append(((it, s, c), r.history.copy())))
    first = None
    for i, (over, hist) in enumerate(rec):
        if i >= len(g["over"]) or tuple(g["over"][i]) != over or \
                not np.array_equal(hist.view(np.uint32), g["history"][i].view(np.uint32)):
            first = over[0]
            break
    off, equal = 0, 0
    for i, n in enumerate(g["final_len"]):
        equal += list(res["paths"].get(reqs[i][0], [])) == g["final_nodes"][off:off + n].tolist()
        off += n
    assert (first, equal, len(reqs)) == CANON_DIVERGENCE[name]


def test_cfg1_trajectory_matches_reference(golden_dir):
    """The full 96x96x6 board with its 300 nets, 30 iterations: the
    reference's recorded run (overuse, routing orders, total_cost / history /
    present_ema digests, final arrays and paths)."""
    g = load(golden_dir, "trajectory_cfg1.npz")
    reqs, r, res, rec, orders, cost_digests = run_oracle_compact(g)
    assert [tuple(x) for x in g["over"].tolist()] == [x[0] for x in rec]
    assert [x[1] for x in rec] == g["hist_digest"].tolist()
    assert [x[2] for x in rec] == g["ema_digest"].tolist()
    assert cost_digests == g["cost_digest"].tolist()
    off = 0
    for i, n in enumerate(g["order_len"]):
        assert orders[i] == g["order"][off:off + n].tolist(), f"order differs at call {i}"
        off += n
    np.testing.assert_array_equal(r.history.view(np.uint32), g["final_history"].view(np.uint32))
    off = 0
    for i, n in enumerate(g["final_len"]):
        assert list(res["paths"].get(reqs[i][0], [])) == g["final_nodes"][off:off + n].tolist()
        off += n
