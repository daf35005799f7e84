"""Geometry emission on the GPU router: the router's vectorised emission of
the reference's recorded input paths equals the reference's records
(tests/golden/geometry.json); a negotiation routed on the GPU in reference
order emits the reference's geometry of its own routing; a canonical-order
GPU negotiation emits the oracle's geometry of the oracle's paths."""
import json
import os

import numpy as np
import pytest

from oracle import geometry as G

pytestmark = pytest.mark.gpu


def _router(X, Y, L, order="canonical", policy="reference", cap=30):
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    cfg = PathFinderConfig()
    cfg.lattice_policy = policy
    cfg.sssp_order = order
    cfg.max_iterations_cap = cap
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    return r


@pytest.fixture(scope="module")
def fixture(golden_dir):
    with open(os.path.join(golden_dir, "geometry.json")) as f:
        return json.load(f)


def _names(r, d):
    r.config.layer_names = list(d["layer_names"])
    r.config.grid_pitch = d["grid_pitch"]


def test_router_emission_equals_reference_records(fixture):
    d = fixture
    r = _router(d["X"], d["Y"], d["L"])
    _names(r, d)
    g = r.lattice.geom
    assert (float(g.pitch), float(g.grid_min_x), float(g.grid_min_y)) == (d["pitch"], d["grid_min_x"], d["grid_min_y"])
    r.net_paths = {k: list(v) for k, v in d["paths"].items()}
    tracks, vias = r._generate_geometry_from_paths()
    assert tracks == d["tracks"]
    assert vias == d["vias"]
    # again with some paths changed: per-net records are recomputed only for those
    r.net_paths["C_run"] = list(d["paths"]["C_corner"])
    t2, v2 = r._generate_geometry_from_paths()
    exp_t, exp_v = G.generate(r.net_paths, d["X"], d["Y"], d["L"], d["layer_names"], d["pitch"], d["grid_min_x"],
                              d["grid_min_y"], guard_outer=True)
    for e in exp_t:
        e["width"] = d["grid_pitch"] * 0.6
    assert (t2, v2) == (exp_t, exp_v)


def test_reference_order_routing_emits_reference_geometry(fixture, golden_dir):
    """negotiation_easy routed on the GPU in reference order: its final
    paths are the reference's, and so is their geometry."""
    d = fixture
    g = np.load(os.path.join(golden_dir, "negotiation_easy.npz"))
    reqs = [(f"N{i:03d}", int(s), int(t)) for i, (s, t) in enumerate(zip(g["src"], g["dst"]))]
    r = _router(d["X"], d["Y"], d["L"], order="reference", cap=int(g["cap"]))
    _names(r, d)
    res = r.route_multiple_nets(reqs)
    routed = {f"R{i:03d}": list(map(int, res["paths"].get(q[0], []))) for i, q in enumerate(reqs)}
    assert routed == {k: v for k, v in d["paths"].items() if k.startswith("R")}
    r.net_paths = routed
    tracks, vias = r._generate_geometry_from_paths()
    assert tracks == [t for t in d["tracks"] if t["net"].startswith("R")]
    assert vias == [v for v in d["vias"] if v["net"].startswith("R")]


@pytest.mark.parametrize("policy", ["reference", "adjacent"])
def test_canonical_routing_geometry_equals_oracle(policy):
    """A canonical-order GPU negotiation (28x24x6, 40 nets) against the
    oracle's negotiation with the same search semantics: the same final
    paths, and the router's emission equals the oracle's per-hop walk."""
    from oracle.negotiation import OracleRouter
    X, Y, L = 28, 24, 6
    rng = np.random.default_rng(8)
    cells = rng.permutation(X * Y)[:80]
    reqs = [(f"N{i:03d}", int(cells[2 * i]), int(cells[2 * i + 1])) for i in range(40)]
    r = _router(X, Y, L, policy=policy, cap=15)
    res = r.route_multiple_nets(reqs)
    o = OracleRouter(X, Y, L, policy, pitch=1.0, via_cost=3.0, sssp="canon")
    ores = o.route_multiple_nets(reqs, max_iterations=15)
    gp = {nm: list(map(int, res["paths"].get(nm, []))) for nm, *_ in reqs}
    op = {nm: list(map(int, ores["paths"].get(nm, []))) for nm, *_ in reqs}
    assert gp == op
    r.net_paths = gp
    tracks, vias = r._generate_geometry_from_paths()
    geo = r.lattice.geom
    et, ev = G.generate(op, X, Y, L, r.config.layer_names or [], float(geo.pitch), float(geo.grid_min_x),
                        float(geo.grid_min_y), guard_outer=policy == "reference")
    for e in et:
        e["width"] = r.config.grid_pitch * 0.6
    assert tracks == et
    assert vias == ev
