"""Multi-terminal nets as trees grown by batched multi-source searches.

The reference's multi-terminal router (orthoroute/algorithms/manhattan/
multi_terminal.py, MT:557-820, MT:1005-1160) attaches pads one at a time to
the routed tree by two-pin searches from chosen tree nodes; its pair search
calls a method the router does not define, so that path never runs.  This
module defines the operation as the multi-source search it describes: per
net the tree starts at its first pin, and each step seeds every tree node at
cost 0 and targets every pin not yet connected
(``SimpleDijkstra.find_path_multisource_multisink`` semantics, UPF:1889-1989,
in the canonical FP32 order of orgpu_route_ms), attaches the cheapest pin by
its path and adds the path's nodes to the tree.  All nets still growing take
one step together in one engine call, on one cost snapshot.
"""
import numpy as np

from .device import route_ms


def grow_trees(ctx, graph, cost, nets, params, roi_rows=None, row_of=None):
    """nets: list of pin lists (node ids).  cost: device f32[E].  params: a
    SearchParams (its goal-directed bounds are not used: the searches have
    several targets).  roi_rows / row_of: optional u32 bitmap rows and the
    row of each net (-1 = whole graph).

    Returns (trees, stats).  Per net: "branches" (node lists from a tree node
    to the attached pin, in attachment order), "edges" (CSR edge ids of the
    branches), "order" (pin indices in attachment order, first pin first),
    "unreached" (pin indices never connected), "nodes" (tree nodes in
    first-seen order).  A pin already in the tree attaches by a one-node
    branch; targets are never expanded, so no branch passes through a pin
    that is still unconnected.  Ties between pins go to the lower distance,
    then the lower node id, then the earlier pin."""
    p = type(params)(slots=params.slots, delta=params.delta, n_buckets=params.n_buckets,
                     rounds_per_poll=params.rounds_per_poll, grid=params.grid,
                     local_batches=params.local_batches)
    state = []
    for pins in nets:
        pins = [int(v) for v in pins]
        state.append({"pins": pins, "tree": pins[:1], "in_tree": set(pins[:1]),
                      "left": list(range(1, len(pins))), "branches": [], "edges": [],
                      "order": [0] if pins else [], "unreached": []})
    stats = {"rounds": 0, "items": 0, "edges": 0, "kernel_ms": 0.0, "steps": 0, "searches": 0}
    while True:
        active = [i for i, st in enumerate(state) if st["left"]]
        if not active:
            break
        reqs = []
        for i in active:
            st = state[i]
            tree = np.asarray(st["tree"], dtype=np.int32)
            tgt = np.asarray([st["pins"][j] for j in st["left"]], dtype=np.int32)
            reqs.append(((tree, np.zeros(len(tree), np.float32)), (tgt, np.zeros(len(tgt), np.float32)), i,
                         -1 if row_of is None else int(row_of[i]), 0))
        res, rs = route_ms(ctx, graph, cost, reqs, p, roi_bitmaps=roi_rows)
        stats["steps"] += 1
        stats["searches"] += len(reqs)
        for k in ("rounds", "items", "edges", "kernel_ms"):
            stats[k] += rs[k]
        for i, r in zip(active, res):
            st = state[i]
            if r is None:
                st["unreached"].extend(st["left"])
                st["left"] = []
                continue
            nodes, edges, _, t = r
            j = st["left"].pop(t)
            st["order"].append(j)
            st["branches"].append(nodes.tolist())
            st["edges"].extend(edges.tolist())
            for v in nodes.tolist():
                if v not in st["in_tree"]:
                    st["in_tree"].add(v)
                    st["tree"].append(v)
    trees = [{"branches": st["branches"], "edges": np.asarray(st["edges"], dtype=np.int64),
              "order": st["order"], "unreached": sorted(st["unreached"]), "nodes": st["tree"]} for st in state]
    return trees, stats
