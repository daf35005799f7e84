"""End-to-end GPU negotiation vs the CPU oracle with the same (FP32,
lowest-predecessor) search semantics: identical per-iteration overuse, history
and final paths, for the plain (one net per launch) and batched forms."""
import os

import numpy as np
import pytest

from oracle.negotiation import OracleRouter

pytestmark = pytest.mark.gpu

CASES = [
    # name, X, Y, L, policy, nets, seed, cap
    ("easy", 24, 22, 6, "reference", 14, 5, 30),
    ("l4", 16, 14, 4, "reference", 45, 9, 25),
    ("dense", 20, 20, 6, "reference", 60, 7, 30),
    ("adjacent", 24, 20, 6, "adjacent", 50, 3, 25),
]


def requests(X, Y, n, seed):
    rng = np.random.default_rng(seed)
    cells = rng.permutation(X * Y)[: 2 * n]
    return [(f"N{i:03d}", int(cells[2 * i]), int(cells[2 * i + 1])) for i in range(n)]


def run_oracle(X, Y, L, policy, reqs, cap):
    r = OracleRouter(X, Y, L, policy, pitch=1.0, via_cost=3.0, sssp="canon")
    rec = []
    res = r.route_multiple_nets(reqs, max_iterations=cap,
                                iteration_cb=lambda it, t, v, s, c: rec.append((it, s, c, r.history.copy())))
    return res, rec


def run_gpu(X, Y, L, policy, reqs, cap, slots, delta=1.0, target_buckets=None, graph=False, engine="auto"):
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    cfg = PathFinderConfig()
    cfg.lattice_policy = policy
    cfg.batch_slots = slots
    cfg.max_iterations_cap = cap
    if delta is not None:
        cfg.sssp_delta, cfg.sssp_buckets = delta, 8
    if target_buckets is not None:
        cfg.sssp_target_buckets = target_buckets
    cfg.sssp_graph_rounds = graph
    cfg.sssp_engine = engine
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    rec = []

    def cb(it, t, v, s, c):
        rec.append((it, s, c, r.accounting.history.cpu().numpy()[:r.accounting.E].copy()))
    res = r.route_multiple_nets(reqs, iteration_cb=cb)
    return res, rec


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("slots", [1, 64])
def test_negotiation_matches_oracle(case, slots):
    name, X, Y, L, policy, n, seed, cap = case
    print(f"[router] {name} {X}x{Y}x{L} {policy} nets={n} slots={slots}", flush=True)
    reqs = requests(X, Y, n, seed)
    o_res, o_rec = run_oracle(X, Y, L, policy, reqs, cap)
    g_res, g_rec = run_gpu(X, Y, L, policy, reqs, cap, slots)
    assert [r[:3] for r in g_rec] == [r[:3] for r in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for nm, *_ in reqs:
        assert list(g_res["paths"].get(nm, [])) == list(o_res["paths"].get(nm, [])), nm
    assert bool(g_res.get("success")) == bool(o_res.get("success"))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("delta", [1.0, None], ids=["fixed", "tuned"])
def test_negotiation_solo_engine_matches_oracle(case, delta):
    """Every search of the negotiation on the one-wave engine (keepout
    iterations included): the oracle's trajectory, history bits and paths."""
    name, X, Y, L, policy, n, seed, cap = case
    print(f"[router] {name} engine 1 delta={delta}", flush=True)
    reqs = requests(X, Y, n, seed)
    o_res, o_rec = run_oracle(X, Y, L, policy, reqs, cap)
    g_res, g_rec = run_gpu(X, Y, L, policy, reqs, cap, 64, delta=delta, engine=1)
    assert [r[:3] for r in g_rec] == [r[:3] for r in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for nm, *_ in reqs:
        assert list(g_res["paths"].get(nm, [])) == list(o_res["paths"].get(nm, [])), nm


@pytest.mark.parametrize("shape", [(12, 10, 16, 30, 16), (10, 8, 24, 24, 24)], ids=["l16", "l24"])
def test_negotiation_many_layers_matches_oracle(shape, monkeypatch):
    """14 and 22 signal layers with a fast pres_fac ramp (the reference's
    ORTHO_PRES_FAC_MULT override): the pres_fac_max 64 / 128 and history
    weight branches and via annealing; batched GPU run vs the oracle with
    the same search semantics."""
    X, Y, L, n, seed = shape
    monkeypatch.setenv("ORTHO_PRES_FAC_MULT", "1.6")
    print(f"[router] {X}x{Y}x{L} nets={n}", flush=True)
    reqs = requests(X, Y, n, seed)
    o_res, o_rec = run_oracle(X, Y, L, "reference", reqs, 30)
    g_res, g_rec = run_gpu(X, Y, L, "reference", reqs, 30, 1024, delta=None)
    assert [r[:3] for r in g_rec] == [r[:3] for r in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for nm, *_ in reqs:
        assert list(g_res["paths"].get(nm, [])) == list(o_res["paths"].get(nm, [])), nm


@pytest.mark.parametrize("policy,engine", [("reference", "auto"), ("adjacent", "auto"), ("reference", 1)])
def test_cfg1_full_negotiation_matches_oracle(policy, engine):
    """The whole cfg1 board: 96x96x6, its 300 seeded top-layer nets, 30
    iterations, batched on the GPU vs the oracle with the same search
    semantics (per-iteration overuse, history bits, final paths); also
    with every search on the one-wave engine."""
    from orthoroute_amd import workloads
    X, Y, L, reqs = workloads.cfg1()
    print(f"[router] cfg1 {X}x{Y}x{L} {policy} nets={len(reqs)} cap=30 engine={engine}", flush=True)
    g_res, g_rec = run_gpu(X, Y, L, policy, reqs, 30, 1024, delta=None, engine=engine)
    print(f"[router] cfg1 gpu done: final overuse {g_rec[-1][1]}", flush=True)
    o_res, o_rec = run_oracle(X, Y, L, policy, reqs, 30)
    assert [r[:3] for r in g_rec] == [r[:3] for r in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for nm, *_ in reqs:
        assert list(g_res["paths"].get(nm, [])) == list(o_res["paths"].get(nm, [])), nm


def test_negotiation_graph_rounds_matches_oracle():
    """The whole negotiation with every poll's rounds replayed as a captured
    HIP graph: the oracle's trajectory, history bits and paths."""
    name, X, Y, L, policy, n, seed, cap = CASES[2]
    print(f"[router] {name} graph rounds", flush=True)
    reqs = requests(X, Y, n, seed)
    o_res, o_rec = run_oracle(X, Y, L, policy, reqs, cap)
    g_res, g_rec = run_gpu(X, Y, L, policy, reqs, cap, 64, delta=None, graph=True)
    assert [r[:3] for r in g_rec] == [r[:3] for r in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for nm, *_ in reqs:
        assert list(g_res["paths"].get(nm, [])) == list(o_res["paths"].get(nm, [])), nm


def test_negotiation_large_batch_matches_oracle():
    """An iteration of more than 1,024 nets takes the path where the
    searches start in task order while a helper thread computes the routing
    order: the oracle's trajectory, history bits and final paths."""
    X, Y, L, n, cap = 80, 72, 4, 1100, 3
    print(f"[router] large batch {X}x{Y}x{L} nets={n}", flush=True)
    reqs = requests(X, Y, n, 31)
    o_res, o_rec = run_oracle(X, Y, L, "adjacent", reqs, cap)
    g_res, g_rec = run_gpu(X, Y, L, "adjacent", reqs, cap, 1024, delta=None)
    assert [r[:3] for r in g_rec] == [r[:3] for r in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for nm, *_ in reqs:
        assert list(g_res["paths"].get(nm, [])) == list(o_res["paths"].get(nm, [])), nm


@pytest.mark.parametrize("target", [0, 2, 128])
def test_depth_sized_buckets_keep_results(target):
    """The router's bucket width (the lattice's default, widened by the
    batch's depth estimate towards `target` buckets; 0 = not widened) sets
    only the speed: the per-iteration overuse, history and final paths equal
    the oracle's for every choice."""
    name, X, Y, L, policy, n, seed, cap = CASES[2]
    print(f"[router] depth-sized buckets target={target} on {name}", flush=True)
    reqs = requests(X, Y, n, seed)
    o_res, o_rec = run_oracle(X, Y, L, policy, reqs, cap)
    g_res, g_rec = run_gpu(X, Y, L, policy, reqs, cap, 64, delta=None, target_buckets=target)
    assert [r[:3] for r in g_rec] == [r[:3] for r in o_rec]
    for a, b in zip(g_rec, o_rec):
        np.testing.assert_array_equal(a[3].view(np.uint32), b[3].view(np.uint32), err_msg=f"history it {a[0]}")
    for nm, *_ in reqs:
        assert list(g_res["paths"].get(nm, [])) == list(o_res["paths"].get(nm, [])), nm


def test_headless_orp_to_ors(tmp_path):
    """A board written as .ORP, routed headless, the solution written as .ORS:
    the file holds the router's own geometry and one metrics row per
    iteration."""
    from orthoroute_amd import serialization as S
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig
    from orthoroute_amd.board import Board, Component, Coordinate, Layer, Net, Pad
    from orthoroute_amd.headless import route_orp_file
    print("[router] headless ORP -> ORS", flush=True)
    rng = np.random.default_rng(8)
    pads, comps = [], []
    for c in range(6):
        cp = [Pad(id=f"P{c}_{k}", component_id=f"C{c}", net_id=None,
                  position=Coordinate(float(2 + 3 * c), float(2 + 2 * k)), size=(0.5, 0.5)) for k in range(4)]
        pads += cp
        comps.append(Component(id=f"C{c}", reference=f"U{c}", value="", footprint="",
                               position=Coordinate(2.0 + 3 * c, 5.0), pads=cp))
    order = rng.permutation(len(pads))
    nets = [Net(id=f"N{i}", name=f"SIG{i}", pads=[pads[order[2 * i]], pads[order[2 * i + 1]]]) for i in range(8)]
    layers = [Layer(name=n, type="signal", stackup_position=i) for i, n in enumerate(["F.Cu", "In1.Cu", "In2.Cu", "B.Cu"])]
    board = Board(id="B", name="t.kicad_pcb", components=comps, nets=nets, layers=layers, layer_count=4)
    S.export_board_to_orp(board, tmp_path / "t.orp")
    cfg = PathFinderConfig()
    cfg.grid_pitch = 0.5
    cfg.max_iterations_cap = 20
    router, result = route_orp_file(tmp_path / "t.orp", tmp_path / "t.ors", config=cfg)
    geo, meta = S.import_solution_from_ors(tmp_path / "t.ors")
    payload = router.get_geometry_payload()
    assert meta["statistics"]["total_tracks"] == len(payload.tracks)
    assert meta["statistics"]["total_vias"] == len(payload.vias)
    assert len(meta["iteration_metrics"]) >= 1
    assert meta["metadata"]["converged"] == bool(result["converged"])
    if result["success"]:
        # a converged routing exports its geometry; one with overuse exports none
        assert len(payload.tracks) + len(payload.vias) > 0
    routed = [n for n, p in result["paths"].items() if p]
    assert len(routed) >= 1
