"""Headless routing of one board file: .ORP in, .ORS out (the steps of the
reference's headless mode, main.py:287-622): read and convert the board,
build the lattice, route every net with the negotiation loop, emit
geometry and write the solution with per-iteration metrics."""
import time

from .algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
from .serialization import convert_orp_to_board, export_solution_to_ors, import_board_from_orp


def route_orp_file(orp_path, ors_path, config: PathFinderConfig = None, compress: bool = True):
    """Route the board in `orp_path` and write the solution to `ors_path`.
    Returns (router, result dict of route_multiple_nets)."""
    t0 = time.time()
    orp = import_board_from_orp(orp_path)
    board = convert_orp_to_board(orp)
    cfg = config or PathFinderConfig()
    pitch = (orp.get("grid_parameters") or {}).get("grid_pitch")
    if config is None and pitch:
        cfg.grid_pitch = float(pitch)
    router = PathFinderRouter(cfg)
    router.initialize_graph(board)
    # the escape planner reads pad outlines from GUI-style pad records
    # (main.py:494-500: x, y, width, height in mm)
    board._gui_pads = [{"x": p["position"]["x"], "y": p["position"]["y"], "width": p["size"]["width"],
                        "height": p["size"]["height"]} for p in orp.get("pads", [])
                       if p.get("position") and p.get("size")]
    router.precompute_all_pad_escapes(board)
    router.prepare_routing_runtime()
    metrics = []

    def on_iteration(it, tracks, vias, over_sum, over_cnt):
        # the reference's headless record (main.py:513-537): fields the
        # callback cannot see keep their neutral values
        metrics.append({"iteration": it, "timestamp": time.strftime("%Y-%m-%d %H:%M:%S"), "duration_s": 0.0,
                        "overuse": over_cnt, "barrel_conflicts": 0, "routed_nets": len(tracks), "failed_nets": 0,
                        "total_edges": len(tracks) + len(vias), "pres_fac": 1.0, "pres_fac_mult": 1.0,
                        "hist_gain": 0.0, "hist_cost_weight": 0.0, "via_penalty": 3.0, "hotset_size": 0,
                        "stagnant_iters": 0, "stagnation_events": 0, "plateau_kick_applied": False})

    result = router.route_multiple_nets(board.nets, iteration_cb=on_iteration)
    n_tracks, n_vias = router.emit_geometry(board)
    meta = {"orthoroute_version": "0.1.0", "total_time": time.time() - t0, "iterations": len(metrics),
            "converged": bool(result.get("converged", False)), "barrel_conflicts": result.get("barrel_conflicts", 0),
            "nets_routed": result.get("nets_routed", 0), "wirelength": result.get("wirelength", 0.0),
            "via_count": n_vias, "track_count": n_tracks, "overflow": result.get("overflow", 0)}
    export_solution_to_ors(router.get_geometry_payload(), metrics, meta, ors_path, compress=compress)
    return router, result
