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
