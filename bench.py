#!/usr/bin/env python3
"""PathFinder routing benchmark on one MI355X.

Workload: the 600x400x18 backplane (4,320,000 lattice nodes, alternating H/V
layers with vias between adjacent layers, 16,782,000 directed edges, FP32
costs, via cost 3.0); 20 connectors of 2 rows x 200 pins at 2-cell pitch at
seeded positions; 8,000 two-pin nets joining pin k of connector i to pin k of
connector j (orthoroute_amd.workloads.cfg2).

One step = one full PathFinder routing pass over the board: every net is
(re)routed on the iteration-1 cost snapshot -- ROI extraction for the short
nets, the batched SSSP, backtrace, commit of the usage counts -- followed by
the present / overuse accounting.  Inputs (lattice, costs, requests) are
resident in HBM before the timed region.

value = SSSP GTEPS of the whole step: Dijkstra-minimal edges (out-edges of the
nodes a label-setting search must settle, i.e. every node whose final cost is
at most the destination's) per second of step wall time.  The searches run in
goal-directed bucket order and scan far fewer edges than that (reported as
edges_scanned_per_step); the minimal count is the fixed measure of the work a
Dijkstra search of the same requests must do.  It comes from an untimed
counting pass in plain Dijkstra order (or from --minimal-counts).

Also reported: plain (one search at a time) vs batched on a sample of the same
searches; the k_round roofline (bytes of the expanded items and scanned edges
per launch over the HIP-event time of the round launches); the CPU oracle's
GTEPS on a bounded sample of the same searches (single core), with its paths
compared against the GPU's; geometry emission; cfg3's multi-terminal trees;
the whole negotiation to zero overuse on a board the schedule solves, batched
and plain, against the oracle; and one routing pass of the 1000x800x24
stress board (cfg4), with its own plain vs batched sample.
"""
import argparse
import json
import os
import sys
import time

METRIC = "PathFinder time-to-zero-overuse (s) + SSSP GTEPS, batched vs plain, 1 GPU"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of a label-setting search (no re-expansions, no frontier
# or de-duplication traffic): per settled node its CSR row bounds (8) and own
# key (8); per scanned edge its target id (4), cost (4) and target key (8)
BYTES_PER_NODE = 16
BYTES_PER_EDGE = 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--slots", type=int, default=2048,
                    help="concurrent searches of the batched form (clamped to what 70 %% of the device memory holds)")
    ap.add_argument("--plain-sample", type=int, default=1024,
                    help="searches timed in the plain vs batched comparison (the batched form holds up to "
                         "--slots of them at once)")
    ap.add_argument("--stress-nets", type=int, default=40000,
                    help="nets of the 1000x800x24 stress board routed in one timed pass (0: skip)")
    ap.add_argument("--stress-plain", type=int, default=1024,
                    help="searches of the stress board timed plain vs batched")
    ap.add_argument("--rounds-per-poll", type=int, default=32, help="SSSP rounds queued between slot refills")
    ap.add_argument("--ttz-nets", type=int, default=120,
                    help="nets of the 96x96x6 board whose full negotiation is timed to zero overuse (0: skip)")
    ap.add_argument("--mt-nets", type=int, default=200,
                    help="cfg3 multi-terminal nets routed as trees after the timed steps (0: skip)")
    ap.add_argument("--cfg2-iterations", type=int, default=8,
                    help="iteration cap of the full backplane negotiation timed batched and plain (0: skip)")
    ap.add_argument("--cfg2-plain", type=int, default=1, help="also run the backplane negotiation plain (0: no)")
    ap.add_argument("--cfg3-tree-iterations", type=int, default=5,
                    help="iteration cap of the cfg3 negotiation with its multi-terminal nets as trees (0: skip)")
    ap.add_argument("--cfg3-tree-plain", type=int, default=1, help="also run that negotiation plain (0: no)")
    ap.add_argument("--cfg3-iterations", type=int, default=5,
                    help="iteration cap of cfg3's 1,400 two-pin nets negotiated batched and plain (0: skip)")
    ap.add_argument("--ref-iterations", type=int, default=30,
                    help="iteration cap of the reference-order negotiation of the full 96x96x6 board, checked "
                         "against the reference's recorded trajectory (0: skip)")
    ap.add_argument("--cost-reps", type=int, default=5, help="timed repetitions of each cost-update kernel (0: skip)")
    ap.add_argument("--graph-rounds", action="store_true", help="replay the rounds between polls as HIP graphs")
    ap.add_argument("--engine", default="auto", choices=["0", "1", "auto"],
                    help="two-pin search engine: 0 = bucket rounds over all slots, 1 = one wavefront per "
                         "search, auto = 1 when at least 1024 searches run at once")
    ap.add_argument("--list-cap", type=int, default=0, help="engine 1: entries per bucket list (0 = default)")
    ap.add_argument("--waves", type=int, default=0, choices=[0, 1, 2, 4],
                    help="engine 1: wavefronts per search (0 = by the searches in flight)")
    ap.add_argument("--serve-jobs", type=int, default=16, help="serving stream: jobs (0 = skip)")
    ap.add_argument("--serve-max-nodes", type=float, default=4e6, help="serving stream: largest board (nodes)")
    ap.add_argument("--serve-active", type=int, default=4, help="serving slots (jobs on the GPU at once)")
    ap.add_argument("--serve-nets", default="density", choices=["density", "cfg5"],
                    help="nets per serving job: density = 0.75 X Y / (X + Y) (boards the schedule mostly "
                         "solves), cfg5 = log-uniform 500-30,000")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0, help="serving: HBM the running jobs may hold")
    ap.add_argument("--serve-extras", type=int, default=1,
                    help="serving: also the burst capacity runs and the alone-vs-mix re-runs (0: no)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU oracle sample budget (0 = skip)")
    ap.add_argument("--workload", default="cfg2")
    ap.add_argument("--delta", type=float, default=0.0, help="bucket width (0 = the router's choice)")
    ap.add_argument("--buckets", type=int, default=0, help="ring buckets (0 = the router's choice)")
    ap.add_argument("--local-batches", type=int, default=0, help="same-bucket batches per wave per round")
    ap.add_argument("--grid", type=int, default=0, help="workgroups per round launch")
    ap.add_argument("--minimal-counts", default="", metavar="JSON",
                    help="take the Dijkstra-minimal node/edge counts of the step from this file (written by an "
                         "earlier run with --save-minimal-counts) instead of an untimed counting pass; with "
                         "--warmup 0 --plain-sample 0 the timed step's launches are then the only SSSP launches")
    ap.add_argument("--save-minimal-counts", default="", metavar="JSON")
    ap.add_argument("--access-probe", type=int, default=0, metavar="N",
                    help="print the device's random 8-byte load and 64-bit atomicMin rates over N accesses and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run 5 plain device copies of BYTES bytes (HBM counter calibration) and exit")
    return ap.parse_args()


def copy_probe(nbytes):
    """Five plain device-to-device copies of a known size: the reference point
    that scales the profiler's FETCH_SIZE / WRITE_SIZE readings to bytes."""
    import torch
    src = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "launches": 5}), flush=True)


def access_probe(n):
    """Random-access rates of the device: n 8-byte loads (gather) and n
    64-bit atomic minimums (scatter_reduce amin) at uniformly random indices
    of arrays from 1 MiB to 1 GiB; prints one JSON line per array size."""
    import torch
    dev = "cuda"
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    for nbytes in (1 << 20, 8 << 20, 64 << 20, 1 << 30):
        m = nbytes // 8
        keys = torch.randint(0, 1 << 62, (m,), device=dev, dtype=torch.int64, generator=g)
        idx = torch.randint(0, m, (n,), device=dev, dtype=torch.int64, generator=g)
        vals = torch.randint(0, 1 << 62, (n,), device=dev, dtype=torch.int64, generator=g)
        out = {"array_bytes": nbytes, "accesses": n}
        for label, fn in (("gather", lambda: keys[idx]),
                          ("atomic_min", lambda: keys.scatter_reduce_(0, idx, vals, "amin"))):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                fn()
            torch.cuda.synchronize()
            out[label + "_G_per_s"] = round(3 * n / (time.perf_counter() - t0) / 1e9, 2)
        print(json.dumps(out), flush=True)


def engine_arg(v):
    return v if v == "auto" else int(v)


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be >= 1")
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.access_probe > 0:
        access_probe(args.access_probe)
        return

    import numpy as np
    import torch
    from orthoroute_amd import workloads
    from orthoroute_amd.device import max_slots, solo_waves
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter

    X, Y, L, reqs = workloads.CONFIGS[args.workload]()
    cfg = PathFinderConfig()
    cfg.lattice_policy = "adjacent"
    cfg.batch_slots = args.slots
    if args.delta > 0:
        cfg.sssp_delta = args.delta
    if args.buckets > 0:
        cfg.sssp_buckets = args.buckets
    cfg.sssp_local_batches = args.local_batches
    cfg.sssp_grid = args.grid
    cfg.sssp_rounds_per_poll = args.rounds_per_poll
    cfg.sssp_graph_rounds = args.graph_rounds
    cfg.sssp_engine, cfg.sssp_list_cap = engine_arg(args.engine), args.list_cap
    cfg.sssp_solo_waves = args.waves
    r = PathFinderRouter(cfg)
    t0 = time.perf_counter()
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    torch.cuda.synchronize()
    log(f"[bench] lattice {X}x{Y}x{L} N={r.N} E={r.accounting.E} built in {time.perf_counter() - t0:.2f}s")
    tasks = r._parse_requests(reqs)
    pres_fac, _, _, _, hmult = r._negotiation_setup(tasks)
    r.iteration = 1
    _, over_map = r._begin_iteration(1, pres_fac, hmult)
    torch.cuda.synchronize()
    st = r.accounting.state

    def step():
        routed, failed = r._route_all(tasks, over_map=over_map)
        st.build_present()
        o_idx, _ = st.overuse_sparse(0)
        return routed, failed, len(o_idx)

    def snap():
        return {k: float(v) for k, v in r.stats.items()}

    n_warm = args.warmup
    if args.minimal_counts:
        with open(args.minimal_counts) as f:
            mc = json.load(f)[args.workload]
        min_nodes, min_edges = float(mc["minimal_nodes"]), float(mc["minimal_edges"])
        log(f"[bench] minimal counts from {args.minimal_counts}: nodes={min_nodes:.4g} edges={min_edges:.4g}")
    else:
        # untimed counting pass (plain Dijkstra order, on the bucket engine):
        # Dijkstra-minimal nodes/edges of this step's searches.  It warms
        # nothing the timed steps use (the one-wave engine's scratch is
        # allocated and every slot's keys filled on that engine's first
        # call), so the warmup steps follow it
        cfg.count_minimal = True
        s0 = snap()
        t0 = time.perf_counter()
        routed, failed, over_cnt = step()
        s1 = snap()
        cfg.count_minimal = False
        min_nodes = s1.get("sssp_minimal_nodes", 0) - s0.get("sssp_minimal_nodes", 0)
        min_edges = s1.get("sssp_minimal_edges", 0) - s0.get("sssp_minimal_edges", 0)
        log(f"[bench] counting pass {time.perf_counter() - t0:.2f}s routed={routed} failed={failed} "
            f"overused_edges={over_cnt} minimal nodes={min_nodes:.4g} edges={min_edges:.4g}")
        if args.save_minimal_counts:
            with open(args.save_minimal_counts, "w") as f:
                json.dump({args.workload: {"minimal_nodes": int(min_nodes), "minimal_edges": int(min_edges)}}, f)
    for w in range(max(0, n_warm)):
        t0 = time.perf_counter()
        step()
        log(f"[bench] warmup step {time.perf_counter() - t0:.2f}s")

    torch.cuda.synchronize()
    s0 = snap()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        routed, failed, over_cnt = step()
        log(f"[bench] step {k + 1}/{args.steps} {time.perf_counter() - ts:.2f}s routed={routed} failed={failed}")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s1 = snap()
    d = {k: s1.get(k, 0.0) - s0.get(k, 0.0) for k in s1}
    K = args.steps
    gteps = min_edges * K / wall / 1e9
    # the step's dominant kernel: the one-wave search kernel when the pass
    # runs on engine 1 (the wider-ROI fallback of the first five failures,
    # a handful of searches, runs on the bucket engine beside it), else the
    # bucket round kernel
    step_engine = r._engine_for(len(tasks))
    pre = "solo_" if step_engine == 1 else "sssp_"
    launches = d.get(pre + "launches", 0.0)
    round_ms = d.get("solo_kernel_ms" if step_engine == 1 else "sssp_round_ms", 0.0)
    avg_launch_us = 1e3 * round_ms / max(launches, 1)
    # algorithmic bytes of the work that kernel's launches performed
    # (expanded items and scanned edges, re-expansions included); the
    # Dijkstra-minimal equivalent is reported beside it
    alg_bytes = BYTES_PER_NODE * d.get(pre + "items", 0) + BYTES_PER_EDGE * d.get(pre + "edges", 0)
    achieved = alg_bytes / max(round_ms * 1e-3, 1e-12) / 1e9
    dij_bytes = (BYTES_PER_NODE * min_nodes + BYTES_PER_EDGE * min_edges) * K
    dij_equiv = dij_bytes / max(round_ms * 1e-3, 1e-12) / 1e9
    log(f"[bench] {K} steps in {wall:.2f}s; rounds kernel {round_ms / 1e3:.2f}s over {launches:.0f} launches "
        f"({avg_launch_us:.1f} us avg); edges scanned {d.get('sssp_edges', 0):.4g}")

    # plain (one search at a time) vs batched on the same sample of searches
    sreqs = sample_reqs(r, tasks, args.plain_sample)
    timing, paths, same = plain_vs_batched(r, sreqs, args.slots)

    cpu = None
    if args.cpu_seconds > 0:
        creqs = sreqs[:48] or sample_reqs(r, tasks, 48)
        cpu = cpu_baseline(r, creqs, paths["batched"][:len(creqs)] if sreqs else [None] * len(creqs),
                           args.cpu_seconds, X, Y, L, compare=bool(sreqs))

    geo = geometry(r)
    costk = cost_kernels(r, pres_fac, hmult, args.cost_reps) if args.cost_reps > 0 else None
    # the sections below build boards of their own: the backplane's search
    # scratch (each slot's keys and bucket bitmaps) is released first
    eng_line = r._engine_for(len(tasks))
    sp_line = r._search_params(args.slots, engine=eng_line)
    slots_line = min(args.slots, max_slots(r.N, engine=eng_line))
    r.ctx.close()
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    neg2 = negotiation(args, "cfg2", args.cfg2_iterations, bool(args.cfg2_plain)) if args.cfg2_iterations > 0 else None
    neg3 = negotiation(args, "cfg3", args.cfg3_iterations, True) if args.cfg3_iterations > 0 else None
    neg3t = negotiation(args, "cfg3", args.cfg3_tree_iterations, bool(args.cfg3_tree_plain), trees=True) \
        if args.cfg3_tree_iterations > 0 else None
    refo = reference_order(args) if args.ref_iterations > 0 else None
    mt = multi_terminal(args) if args.mt_nets > 0 else None
    ttz = time_to_zero_overuse(args) if args.ttz_nets > 0 else None
    stress_line = stress(args) if args.stress_nets > 0 else None
    serve = serving(args) if args.serve_jobs > 0 else None

    # HBM traffic per launch of the dominant kernel comes from a committed
    # rocprofv3 --pmc profile and is reported only when that profile
    # measured this build's kernel source (the digest it records equals the
    # digest of the source file now)
    # (engine 1 runs k_solo with one wave per search, k_solo_mw<W> with W)
    kname, kfile, ksrc = "k_round", "k_round", "sssp.hip"
    if step_engine == 1:
        w = solo_waves(args.waves, slots_line, len(tasks))
        kname, kfile, ksrc = ("k_solo", "k_solo", "sssp_solo.hip") if w == 1 else \
            (f"k_solo_mw<{w}>", f"k_solo_mw{w}", "sssp_solo.hip")
    traffic, traffic_note = None, "no counter profile"
    here = os.path.dirname(os.path.abspath(__file__))
    prof = os.path.join(here, "profiles", f"{kfile}_cfg2_traffic.json")
    if os.path.exists(prof):
        import hashlib
        with open(prof) as f:
            tj = json.load(f)
        with open(os.path.join(here, "orthoroute_amd", "csrc", ksrc), "rb") as f:
            src_digest = hashlib.sha256(f.read()).hexdigest()
        if tj.get("kernel_source_sha256") == src_digest:
            traffic = tj.get("hbm_bytes_per_launch")
            traffic_note = f"rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel source ({src_digest[:12]}), " \
                           f"copy-calibrated: profiles/{kfile}_cfg2_traffic.json"
        else:
            traffic_note = "the counter profile measured another version of the kernel source: not reported"

    line = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall / K, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload} backplane {X}x{Y}x{L}, adjacent-layer vias, {r.N} nodes, "
                        f"{r.accounting.E} edges, {len(tasks)} two-pin nets; one step = one full routing "
                        f"pass on the iteration-1 cost snapshot",
            "nodes": r.N, "edges": r.accounting.E, "nets": len(tasks),
            "engine": step_engine, "waves": solo_waves(args.waves, slots_line, len(tasks)) if step_engine == 1
            else None, "slots": slots_line, "delta": sp_line.delta, "n_buckets": sp_line.n_buckets,
            "local_batches": args.local_batches or 64, "grid": args.grid or 1024,
            "rounds_per_poll": args.rounds_per_poll,
        },
        "nets_per_s": round(len(tasks) * K / wall, 1),
        "routed": routed, "failed": failed,
        "minimal_edges_per_step": int(min_edges), "minimal_nodes_per_step": int(min_nodes),
        "value_kind": "Dijkstra-minimal equivalent: out-edges of the nodes a label-setting search of the same "
                      "requests settles, per second of step wall time (the searches scan fewer, see "
                      "edges_scanned_per_step)",
        "edges_scanned_per_step": int(d.get("sssp_edges", 0) / K),
        "scanned_edges_per_s_wall": round(d.get("sssp_edges", 0) / wall / 1e9, 3),
        "scanned_edges_per_s_rounds": round(d.get("sssp_edges", 0) / max(round_ms * 1e-3, 1e-12) / 1e9, 3),
        "scanned_unit": "G edges/s",
        "sssp_s_per_step": round(d.get("sssp_s", 0) / K, 3),
        "sssp_kernel_gteps": round(min_edges * K / max(d.get("sssp_kernel_ms", 0) * 1e-3, 1e-9) / 1e9, 4),
        "roi_s_per_step": round(d.get("roi_s", 0) / K, 3),
        "plain_vs_batched": {"searches": len(sreqs), "plain_s": round(timing["plain"], 3),
                             "batched_s": round(timing["batched"], 3),
                             "speedup": round(timing["plain"] / max(timing["batched"], 1e-9), 2),
                             "identical_paths": same},
        "roofline": {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_note": traffic_note,
                     "avg_launch_us": round(avg_launch_us, 2), "launches": int(launches),
                     "alg_bytes_per_launch": round(alg_bytes / max(launches, 1), 1),
                     "dijkstra_minimal_equiv_GBs": round(dij_equiv, 1)},
        "cpu_baseline": cpu,
        "geometry": geo,
        "cost_kernels": costk,
        "negotiation_cfg2": neg2,
        "negotiation_cfg3": neg3,
        "negotiation_cfg3_trees": neg3t,
        "reference_order": refo,
        "multi_terminal": mt,
        "time_to_zero_overuse": ttz,
        "stress": stress_line,
        "serving": serve,
    }
    print(json.dumps(line), flush=True)


def cost_kernels(r, pres_fac, hmult, reps):
    """HIP-event times of the per-iteration congestion kernels on the
    backplane (16.8 M edges): the cost update (k_cost, 22 algorithmic bytes
    per edge: base, EMA, history, total f32 and layer / kind u8), the
    history update (k_history, 20 B: present, base, history in and out plus
    the decay term) and the present EMA (k_ema, 12 B: count, EMA in and out;
    its call also compacts the EMA's nonzeros and copies them to the host,
    which the time includes).  Run after the timed steps; the cost update
    rewrites the same total_cost from unchanged inputs."""
    import numpy as np
    import torch
    st = r.accounting.state
    E = r.accounting.E
    out = {}

    def timed(label, fn, nbytes):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = 1e3 * e0.elapsed_time(e1) / reps
        gbs = nbytes * E / (us * 1e-6) / 1e9
        out[label] = {"us": round(us, 1), "bytes_per_edge": nbytes, "GBs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4)}
        if gbs > HBM_PEAK_GBS:
            out[label]["note"] = (f"above the HBM peak: the {reps} back-to-back launches stream the same "
                                  f"{nbytes * E / 2**20:.0f} MiB, part of which the 256 MB MALL holds between them")

    orig = st.update_costs

    def update_timed(p):
        # the cost parameters point at buffers that live only during this call
        timed("k_cost", lambda: orig(p), 22)
    st.update_costs = update_timed
    try:
        r._update_costs(pres_fac, r.config.hist_cost_weight * hmult, 1.0, r.config.base_cost_weight,
                        r._layer_bias_ema, 1)
    finally:
        del st.update_costs
    timed("k_history", lambda: st.update_history(1.0, 0.2, 15.0), 20)
    timed("k_ema_with_compaction", lambda: st.update_ema(False, np.float32(0.4), np.float32(0.6)), 12)
    # this call's time is the kernel plus the wait for it and the copy of the
    # EMA's nonzeros to the host: not a kernel roofline
    out["k_ema_with_compaction"]["frac"] = None
    out["k_ema_with_compaction"]["note"] = "includes the stream sync and the device-to-host copy of the nonzeros"
    log(f"[bench] cost kernels: " + ", ".join(f"{k} {v['us']} us ({v['GBs']} GB/s)" for k, v in out.items()))
    out["edges"] = E
    return out


def negotiation(args, which, cap, with_plain, trees=False):
    """route_multiple_nets on a whole board from entry to exit (lattice and
    cost state resident), batched and, if asked, plain (one slot: searches one
    after another), with the iteration cap `cap`: time to zero overuse when
    the schedule gets there, else the time to the cap; both forms must give
    the same per-iteration overuse and final paths.  cfg2: the backplane's
    8,000 nets; cfg3: the BGA board's 1,400 two-pin ball-to-perimeter nets,
    with trees=True also its 200 multi-terminal nets as trees inside the
    negotiation (one negotiated board)."""
    import torch
    from orthoroute_amd import workloads
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    X, Y, L, reqs = workloads.CONFIGS[which]()
    n_two, n_tree = len(reqs), 0
    if trees:
        mt = workloads.cfg3_multi()[3]
        reqs = list(reqs) + list(mt)
        n_tree = len(mt)

    def run(slots):
        cfg = PathFinderConfig()
        cfg.iteration_geometry = False     # the iteration callbacks below read only the overuse
        cfg.lattice_policy = "adjacent"
        cfg.batch_slots = slots
        cfg.max_iterations_cap = cap
        cfg.sssp_graph_rounds = args.graph_rounds
        cfg.sssp_engine, cfg.sssp_list_cap = engine_arg(args.engine), args.list_cap
        cfg.sssp_solo_waves = args.waves
        cfg.multi_terminal_enabled = trees
        r = PathFinderRouter(cfg)
        r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
        torch.cuda.synchronize()
        rec, t_it = [], []
        t0 = time.perf_counter()

        def cb(it, t, v, s, c):
            rec.append((it, s, c))
            t_it.append(time.perf_counter() - t0)
        res = r.route_multiple_nets(reqs, iteration_cb=cb)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        paths = {nm: list(map(int, res["paths"].get(nm, []))) for nm, *_ in reqs}
        stats = dict(r.stats)
        r.ctx.close()
        del r
        torch.cuda.empty_cache()
        return dt, res, rec, t_it, paths, stats
    dt, res, rec, t_it, paths, stats = run(args.slots)
    zero = next((t for (it, s, c), t in zip(rec, t_it) if s == 0), None)
    log(f"[bench] {which} negotiation batched: {len(rec)} iterations in {dt:.2f}s, overuse at exit {rec[-1][1]}, "
        f"zero overuse at {zero}")
    kind = f"{n_two} two-pin nets" + (f" and {n_tree} multi-terminal nets as trees" if n_tree else "")
    out = {"workload": f"{which} {X}x{Y}x{L}, adjacent-layer vias, {kind}, iteration cap {cap}",
           "iterations": len(rec), "time_to_zero_overuse_s": None if zero is None else round(zero, 3),
           "reached_zero_overuse": zero is not None, "converged": bool(res.get("converged")),
           "time_to_exit_s": round(dt, 3), "overuse_at_exit": int(rec[-1][1]),
           "overuse_per_iteration": [int(x[1]) for x in rec],
           "iteration_s": [round(b - a, 3) for a, b in zip([0.0] + t_it[:-1], t_it)],
           "sssp_s": round(stats.get("sssp_s", 0.0), 3), "roi_s": round(stats.get("roi_s", 0.0), 3),
           "tree_s": round(stats.get("tree_s", 0.0), 3),
           "routed_trees": sum(1 for q in reqs[n_two:] if res["paths"].get(q[0])),
           "searches": int(stats.get("sssp_searches", 0)),
           "small_batches_to_one_wave_engine": ("predicted" if stats.get("small_engine_switch_predicted") else
                                                "measured" if stats.get("small_engine_switch_measured") else None)}
    if with_plain:
        pdt, pres, prec, pt_it, ppaths, _ = run(1)
        same = prec == rec and ppaths == paths
        log(f"[bench] {which} negotiation plain: {len(prec)} iterations in {pdt:.2f}s; same trajectory={same}")
        out["plain"] = {"slots": 1, "time_to_exit_s": round(pdt, 3), "same_trajectory_and_paths": bool(same),
                        "speedup_batched": round(pdt / max(dt, 1e-9), 2)}
    return out


def reference_order(args):
    """The reference-order search (float64 sums, float32 acceptance,
    (alt, roi position) pops: SimpleDijkstra's own choices) driving the whole
    negotiation of the 96x96x6 board with its 300 nets, against the
    reference's recorded run of the same board (tests/golden/
    trajectory_cfg1.npz: per-iteration overuse and sha256 digests of the
    history / present_ema arrays, final paths).  The reference's own time for
    that run (CPython, one core of the build container) is in the fixture and
    reported beside it, labelled as such; the reference does not run here."""
    import hashlib
    import numpy as np
    import torch
    from orthoroute_amd import workloads
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    fx = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "trajectory_cfg1.npz")
    if not os.path.exists(fx):
        return None
    g = np.load(fx)
    X, Y, L, reqs = workloads.cfg1()
    cap = min(args.ref_iterations, int(g["cap"]))

    def dig(t):
        return hashlib.sha256(np.ascontiguousarray(t.cpu().numpy()[:E], np.float32).tobytes()).hexdigest()

    def run(order):
        nonlocal E
        cfg = PathFinderConfig()
        cfg.iteration_geometry = False     # the iteration callbacks below read only the overuse
        cfg.lattice_policy = "reference"
        cfg.max_iterations_cap = cap
        cfg.sssp_order = order
        r = PathFinderRouter(cfg)
        r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
        E = r.accounting.E
        torch.cuda.synchronize()
        rec = []
        t0 = time.perf_counter()
        cbt = [0.0]

        def cb(it, t, v, s, c):
            t1 = time.perf_counter()
            rec.append(((it, s, c), dig(r.accounting.history), dig(r.accounting.present_ema)))
            cbt[0] += time.perf_counter() - t1
        res = r.route_multiple_nets(reqs, iteration_cb=cb)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0 - cbt[0]
        return dt, res, rec, dict(r.stats)
    E = 0
    dt, res, rec, st = run("reference")
    n = min(len(rec), len(g["over"]))
    same_over = [x[0] for x in rec[:n]] == [tuple(v) for v in g["over"].tolist()[:n]]
    same_hist = [x[1] for x in rec[:n]] == g["hist_digest"].tolist()[:n]
    same_ema = [x[2] for x in rec[:n]] == g["ema_digest"].tolist()[:n]
    same_paths = None
    if cap == int(g["cap"]):
        off, ok = 0, True
        for i, k in enumerate(g["final_len"]):
            ok &= list(map(int, res["paths"].get(reqs[i][0], []))) == g["final_nodes"][off:off + k].tolist()
            off += k
        same_paths = bool(ok)
    cdt, cres, crec, _ = run("canonical")
    log(f"[bench] reference order cfg1: {len(rec)} iterations in {dt:.2f}s (canonical {cdt:.2f}s); equal to the "
        f"reference: overuse {same_over}, history {same_hist}, ema {same_ema}, final paths {same_paths}")
    first_div = next((a[0][0] for a, b in zip(crec, rec) if a[0] != b[0] or a[1] != b[1]), None)
    return {"workload": f"cfg1 96x96x6 (reference via rules), {len(reqs)} nets, iteration cap {cap}",
            "iterations": len(rec), "s": round(dt, 3), "searches": int(st.get("sssp_searches", 0)),
            "equal_to_reference": {"overuse_per_iteration": bool(same_over), "history_digests": bool(same_hist),
                                   "present_ema_digests": bool(same_ema), "final_paths": same_paths},
            "reference_cpu": {"s": round(float(g["t_route_s"]), 1), "cores": 1, "kind": "reference",
                              "where": "the reference's route_multiple_nets (CPython) recorded by "
                                       "tests/golden/make_golden.py under taskset -c 0 in the build container",
                              "cpu": str(g["cpu"]), "speedup_gpu": round(float(g["t_route_s"]) / max(dt, 1e-9), 1)},
            "canonical_order": {"s": round(cdt, 3), "first_iteration_differing_from_reference_order": first_div}}


def sample_reqs(r, tasks, n):
    """The first n two-pin nets of a router's task list as search requests
    (src, dst, -1, use ROI, 0), with the router's ROI rule."""
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import ROI_THRESHOLD_STEPS
    out = []
    for name, (s, dd) in list(tasks.items())[:n]:
        sx, sy, _ = r._coords(s)
        dx, dy, _ = r._coords(dd)
        out.append((s, dd, -1, abs(dx - sx) + abs(dy - sy) < ROI_THRESHOLD_STEPS, 0.0))
    return out


def plain_vs_batched(r, sreqs, slots):
    """The same searches on the router's current costs in the plain form
    (one slot: one search after another) and the batched form; returns the
    times, the paths and whether the paths are identical."""
    import torch
    timing = {"plain": 0.0, "batched": 0.0}
    paths = {"plain": [], "batched": []}
    for label, k in ((("plain", 1), ("batched", slots)) if sreqs else ()):
        r.config.batch_slots = k
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = r._search_batch(sreqs)
        torch.cuda.synchronize()
        timing[label] = time.perf_counter() - t0
        paths[label] = [None if o is None else o[0].tolist() for o in out]
        log(f"[bench] {label} sample of {len(sreqs)} searches: {timing[label]:.2f}s")
    r.config.batch_slots = slots
    return timing, paths, paths["plain"] == paths["batched"]


def stress(args):
    """The 1000x800x24 stress board (adjacent-layer vias, 19.2 M nodes,
    75.2 M edges) with 40,000 two-pin nets whose boxes have lognormal
    half-widths: one full routing pass on the iteration-1 costs (ROI
    extraction, batched SSSP, backtrace, commit, overuse accounting), timed
    after an untimed warm-up pass, and the plain form against the batched
    form on the first --stress-plain nets."""
    import torch
    from orthoroute_amd import workloads
    from orthoroute_amd.device import max_slots
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    X, Y, L, reqs = workloads.cfg4(n_nets=args.stress_nets)
    import gc
    gc.collect()
    torch.cuda.empty_cache()
    slots = min(args.slots, 1024)      # 1,024 searches of 19.2 M nodes: 184 GB of keys and bitmaps
    cfg = PathFinderConfig()
    cfg.lattice_policy = "adjacent"
    cfg.batch_slots = slots
    cfg.sssp_graph_rounds = args.graph_rounds
    cfg.sssp_engine, cfg.sssp_list_cap = engine_arg(args.engine), args.list_cap
    cfg.sssp_solo_waves = args.waves
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    tasks = r._parse_requests(reqs)
    pres_fac, _, _, _, hmult = r._negotiation_setup(tasks)
    r.iteration = 1
    _, over_map = r._begin_iteration(1, pres_fac, hmult)
    st = r.accounting.state

    def step():
        routed, failed = r._route_all(tasks, over_map=over_map)
        st.build_present()
        return routed, failed

    step()
    torch.cuda.synchronize()
    s0 = {k: float(v) for k, v in r.stats.items()}
    t0 = time.perf_counter()
    routed, failed = step()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    d = {k: float(v) - s0.get(k, 0.0) for k, v in r.stats.items()}
    log(f"[bench] stress {X}x{Y}x{L}: {len(tasks)} nets in {t:.2f}s, routed={routed} failed={failed}")
    sreqs = sample_reqs(r, tasks, args.stress_plain)
    timing, _, same = plain_vs_batched(r, sreqs, slots)
    r.ctx.close()
    return {"workload": f"cfg4 stress board {X}x{Y}x{L}, adjacent-layer vias, {r.N} nodes, {r.accounting.E} edges, "
                        f"{len(tasks)} two-pin nets; one routing pass on the iteration-1 cost snapshot",
            "pass_s": round(t, 3), "nets_per_s": round(len(tasks) / max(t, 1e-9), 1), "routed": routed,
            "failed": failed, "slots": min(slots, max_slots(r.N)),
            "sssp_s": round(d.get("sssp_s", 0.0), 3), "roi_s": round(d.get("roi_s", 0.0), 3),
            "edges_scanned": int(d.get("sssp_edges", 0.0)),
            "plain_vs_batched": {"searches": len(sreqs), "plain_s": round(timing["plain"], 3),
                                 "batched_s": round(timing["batched"], 3),
                                 "speedup": round(timing["plain"] / max(timing["batched"], 1e-9), 2),
                                 "identical_paths": same}}


def time_to_zero_overuse(args):
    """The whole negotiation (route_multiple_nets from entry to its exit) on
    a board it solves: the 96x96x6 lattice with the reference's via rules
    and the first --ttz-nets nets of the 300-net generator (the full 300 do
    not reach zero overuse under the reference's schedule), in the batched
    form (all of an iteration's nets in one engine call) and the plain form
    (one slot: the searches run one after another).  The oracle runs the
    same schedule with the same search semantics on the CPU; its
    per-iteration overuse and its final paths are compared, and so are the
    plain form's."""
    import torch
    from orthoroute_amd import workloads
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle.negotiation import OracleRouter
    X, Y, L, reqs = workloads.cfg1(n_nets=args.ttz_nets)

    def gpu_run(slots=1024):
        cfg = PathFinderConfig()
        cfg.iteration_geometry = False     # the iteration callbacks below read only the overuse
        cfg.lattice_policy = "reference"
        cfg.batch_slots = slots
        cfg.sssp_graph_rounds = args.graph_rounds
        cfg.sssp_engine, cfg.sssp_list_cap = engine_arg(args.engine), args.list_cap
        cfg.sssp_solo_waves = args.waves
        r = PathFinderRouter(cfg)
        r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
        torch.cuda.synchronize()
        rec = []
        t0 = time.perf_counter()
        res = r.route_multiple_nets(reqs, iteration_cb=lambda it, t, v, s, c: rec.append((it, s, c)))
        torch.cuda.synchronize()
        return time.perf_counter() - t0, res, rec
    gpu_run()                                     # warm-up (kernels, allocations)
    t_gpu, res, rec = gpu_run()
    log(f"[bench] time to zero overuse: {len(rec)} iterations, success={res.get('success')} in {t_gpu:.3f}s")
    # the plain form: one search per engine launch sequence (one slot)
    gpu_run(1)
    t_plain, pres, prec = gpu_run(1)
    plain_same = prec == rec and all(list(pres["paths"].get(nm, [])) == list(res["paths"].get(nm, []))
                                     for nm, *_ in reqs)
    log(f"[bench] plain form: {len(prec)} iterations in {t_plain:.3f}s; same trajectory and paths={plain_same}")
    o = OracleRouter(X, Y, L, "reference", pitch=1.0, via_cost=3.0, sssp="canon")
    orec = []
    t0 = time.perf_counter()
    ores = o.route_multiple_nets(reqs, max_iterations=len(rec) + 5,
                                 iteration_cb=lambda it, t, v, s, c: orec.append((it, s, c)))
    t_cpu = time.perf_counter() - t0
    same_traj = orec == rec
    same_paths = all(list(res["paths"].get(nm, [])) == list(ores["paths"].get(nm, [])) for nm, *_ in reqs)
    log(f"[bench] oracle negotiation {t_cpu:.2f}s; trajectory equal={same_traj}, paths equal={same_paths}")
    return {"workload": f"cfg1 lattice 96x96x6 (reference via rules), first {len(reqs)} of the 300 nets",
            "converged": bool(res.get("success")), "iterations": len(rec), "s": round(t_gpu, 4),
            "plain": {"s": round(t_plain, 4), "slots": 1, "same_trajectory_and_paths": bool(plain_same),
                      "speedup_batched": round(t_plain / max(t_gpu, 1e-9), 2)},
            "cpu_oracle": {"s": round(t_cpu, 3), "cores": 1, "kind": "port",
                           "trajectory_equal": bool(same_traj), "paths_equal": bool(same_paths)}}


def serving(args):
    """The serving loop on a Poisson stream of routing jobs (cfg5's arrival
    process, mean gap 0.5 s; boards log-uniform from 1e5 nodes to
    --serve-max-nodes, 4-24 layers, adjacent-layer vias; each job's nets
    drawn like cfg1 at a density the reference's schedule mostly solves,
    0.75 * X * Y / (X + Y) nets; iteration cap 40).  Reports p50 / p99
    time to zero overuse and latency, jobs per hour, how far the jobs'
    device work overlapped (sum of per-job routing time / makespan), the
    capacity when all jobs arrive at once (with the serving slots and with
    one), and re-runs the three quickest jobs alone to compare their
    trajectories and paths with the mix."""
    import torch
    from orthoroute_amd import workloads
    from orthoroute_amd.serving import RoutingServer, replay, run_job, summarize
    jobs = workloads.cfg5(n_jobs=args.serve_jobs, n_min=1e5, n_max=args.serve_max_nodes,
                          net_density=0.75 if args.serve_nets == "density" else None, max_iterations=40)
    torch.cuda.synchronize()
    budget = int(args.serve_budget_gib * 2 ** 30)

    def done(job, r):
        if isinstance(r, BaseException):
            log(f"[bench] serving: job {job.job_id} failed: {r}")
        else:
            log(f"[bench] serving: job {job.job_id} ({r.nodes} nodes, {r.nets} nets) done after {r.iterations} "
                f"iterations in {r.route_s:.2f}s, zero overuse at {r.zero_overuse_s}")
    srv = RoutingServer(hbm_budget=budget, max_active=args.serve_active, roi_limit=1 << 30, on_done=done)
    t0 = time.perf_counter()
    res = replay(srv, jobs)
    srv.close()
    makespan = time.perf_counter() - t0
    summ = summarize(res, makespan)
    busy = sum(r.route_s for r in res if r.error is None)
    log(f"[bench] serving: {summ['completed']}/{len(jobs)} jobs in {makespan:.2f}s, "
        f"{summ['reached_zero_overuse']} reached zero overuse, {summ['converged']} converged, "
        f"p50/p99 time to zero overuse {summ['p50_time_to_zero_overuse_s']}/{summ['p99_time_to_zero_overuse_s']}s, "
        f"peak {srv.peak_active} jobs on the GPU")
    # capacity: the same jobs all at once, with the serving slots and with one
    # (jobs strictly one after another)
    import dataclasses
    burst = [dataclasses.replace(j, arrival_s=0.0) for j in jobs]
    cap = {}
    for active in ((args.serve_active, 1) if args.serve_extras else ()):
        bsrv = RoutingServer(hbm_budget=budget, max_active=active, roi_limit=1 << 30)
        tb = time.perf_counter()
        bres = replay(bsrv, burst)
        bsrv.close()
        mk = time.perf_counter() - tb
        same = all(b.trajectory == r.trajectory for b, r in zip(bres, res))
        cap[active] = {"slots": active, "makespan_s": round(mk, 3),
                       "jobs_per_hour": round(3600.0 * len(jobs) / max(mk, 1e-9), 1),
                       "same_trajectories_as_stream": bool(same)}
    if cap:
        log(f"[bench] serving burst: {len(jobs)} jobs in {cap[args.serve_active]['makespan_s']}s with "
            f"{args.serve_active} slots, {cap[1]['makespan_s']}s with one")
    quick = sorted((r for r in res if r.error is None), key=lambda r: r.route_s)[:3] if args.serve_extras else []
    same = []
    for r in quick:
        job = jobs[r.job_id]
        ares, atraj, _, _ = run_job(job, roi_limit=1 << 30)
        same.append(atraj == r.trajectory and all(
            [int(v) for v in ares["paths"].get(nm, [])] == [int(v) for v in r.paths.get(nm, [])]
            for nm, _, _ in job.nets))
    summ.update({
        "workload": f"cfg5-shaped Poisson stream: {len(jobs)} jobs, mean gap 0.5 s, boards "
                    f"{min(j.nodes for j in jobs)}-{max(j.nodes for j in jobs)} nodes, 4-24 layers, "
                    f"{min(len(j.nets) for j in jobs)}-{max(len(j.nets) for j in jobs)} nets per job, cap 40 iterations",
        "serving_slots": args.serve_active, "peak_jobs_on_gpu": srv.peak_active,
        "peak_admitted_gib": round(srv.peak_used / 2 ** 30, 2), "hbm_budget_gib": args.serve_budget_gib,
        "sum_route_s": round(busy, 3), "overlap": round(busy / max(makespan, 1e-9), 2),
        "alone_equals_mix": f"{sum(same)}/{len(same)}" if same else None,
        "burst": cap.get(args.serve_active), "burst_one_slot": cap.get(1),
        "burst_speedup": round(cap[1]["makespan_s"] / max(cap[args.serve_active]["makespan_s"], 1e-9), 2)
        if cap else None,
        "jobs": [{"id": r.job_id, "nodes": r.nodes, "layers": jobs[r.job_id].L, "nets": r.nets,
                  "iterations": r.iterations, "converged": r.converged,
                  "zero_overuse_s": None if r.zero_overuse_s is None else round(r.zero_overuse_s, 3),
                  "route_s": round(r.route_s, 3), "latency_s": round(r.latency_s, 3)}
                 for r in res if r.error is None],
    })
    return summ


def geometry(r):
    """Track / via emission (UPF:5749-5830) of the step's committed paths,
    and the oracle's per-hop walk over the first 500 of them (compared)."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import geometry as G
    t0 = time.perf_counter()
    tracks, vias = r._generate_geometry_from_paths()
    t = time.perf_counter() - t0
    hops = sum(max(len(p) - 1, 0) for p in r.net_paths.values())
    sample = dict(list((n, list(map(int, p))) for n, p in r.net_paths.items() if p)[:500])
    g = r.lattice.geom
    t1 = time.perf_counter()
    et, ev = G.generate(sample, r._Nx, r._Ny, r._Nz, r.config.layer_names or [], g.pitch, g.grid_min_x,
                        g.grid_min_y, guard_outer=r.lattice.policy == "reference")
    tc = time.perf_counter() - t1
    width = r.config.grid_pitch * 0.6
    for e in et:
        e["width"] = width
    names = set(sample)
    same = [x for x in tracks if x["net"] in names] == et and [x for x in vias if x["net"] in names] == ev
    shops = sum(len(p) - 1 for p in sample.values())
    log(f"[bench] geometry: {len(tracks)} tracks, {len(vias)} vias from {hops} hops in {t:.3f}s; "
        f"oracle sample {shops} hops in {tc:.2f}s, equal={same}")
    return {"paths": sum(1 for p in r.net_paths.values() if p), "hops": hops, "tracks": len(tracks),
            "vias": len(vias), "s": round(t, 4), "hops_per_s": round(hops / max(t, 1e-9), 1),
            "cpu_oracle": {"paths": len(sample), "hops_per_s": round(shops / max(tc, 1e-9), 1), "cores": 1,
                           "records_equal": bool(same)}}


def multi_terminal(args):
    """cfg3's multi-terminal nets (300x300x12, adjacent-layer vias, 3-8 pins
    each) routed as trees by batched multi-source searches on the base
    costs; the oracle's tree for the first nets, about 5 s of CPU, compared
    with the GPU's."""
    import numpy as np
    import torch
    from orthoroute_amd import workloads
    from orthoroute_amd.algorithms.manhattan.multi_terminal import MultiTerminalNetRouter
    from orthoroute_amd.algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    X, Y, L, nets = workloads.cfg3_multi()
    nets = nets[:args.mt_nets]
    cfg = PathFinderConfig()
    cfg.lattice_policy = "adjacent"
    cfg.batch_slots = args.slots
    r = PathFinderRouter(cfg)
    r.initialize_lattice(X, Y, L, pitch=1.0, via_cost=3.0)
    MultiTerminalNetRouter(r).route_all_nets(nets[:8])      # warm-up
    torch.cuda.synchronize()
    mtr = MultiTerminalNetRouter(r)
    t0 = time.perf_counter()
    res = mtr.route_all_nets(nets)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    pins = sum(len(p) for _, p in nets)
    attach = sum(len(mtr.trees[n]["order"]) - 1 for n, _ in nets)
    log(f"[bench] multi-terminal: {len(nets)} nets, {pins} pins, {attach} attachments in {t:.3f}s")
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    lat = oracle.Lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    base = r.graph.dlat.base_cost[:lat.E].cpu().numpy()
    t1 = time.perf_counter()
    checked = equal = 0
    for name, p in nets:
        br, order, unreached, nodes = oracle.steiner_canon(lat, base, p)
        tr = mtr.trees[name]
        equal += int(tr["branches"] == br and tr["order"] == order and tr["unreached"] == unreached)
        checked += 1
        if time.perf_counter() - t1 > 5.0:
            break
    tc = time.perf_counter() - t1
    log(f"[bench] multi-terminal oracle: {checked} nets in {tc:.2f}s, {equal}/{checked} trees equal")
    return {"workload": f"cfg3 multi-terminal nets {X}x{Y}x{L}, adjacent-layer vias, {len(nets)} nets, {pins} pins, "
                        f"base costs",
            "trees_s": round(t, 4), "attachments": attach, "steps": mtr.stats.get("steps"),
            "attachments_per_s": round(attach / max(t, 1e-9), 1),
            "fully_connected": res["_validation"]["fully_connected"],
            "cpu_oracle": {"nets": checked, "s": round(tc, 3), "cores": 1,
                           "attachments_per_s": round(sum(len(mtr.trees[n]["order"]) - 1 for n, _ in nets[:checked])
                                                      / max(tc, 1e-9), 2),
                           "trees_equal_to_gpu": f"{equal}/{checked}"}}


def cpu_baseline(r, sreqs, gpu_paths, budget_s, X, Y, L, compare=True):
    """The oracle's label-setting search (C, one thread) on the same searches
    and the same cost snapshot, for about budget_s seconds."""
    import numpy as np
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle
    t0 = time.perf_counter()
    lat = oracle.Lattice(X, Y, L, "adjacent", pitch=1.0, via_cost=3.0)
    costs = r.accounting.total_cost[:r.accounting.E].cpu().numpy()
    log(f"[bench] cpu oracle lattice + costs in {time.perf_counter() - t0:.2f}s")
    done = traversed = 0
    equal = 0
    t_search = 0.0
    for (s, d, _, use_roi, bonus), gp in zip(sreqs, gpu_paths):
        t1 = time.perf_counter()
        mask = None
        if use_roi:
            roi = oracle.roi_bfs(lat, s, d, stagnation_bonus=bonus)
            mask = np.zeros(lat.N, np.uint8)
            mask[roi] = 1
            mask[s] = mask[d] = 1          # the router re-adds both endpoints
        path, _, _, _, tr = oracle.sssp_canon(lat, costs, s, d, roi_mask=mask)
        t_search += time.perf_counter() - t1
        done += 1
        traversed += tr
        equal += int((path is None and gp is None) or (path is not None and gp is not None and
                                                         list(path) == list(gp)))
        if t_search > budget_s:
            break
    log(f"[bench] cpu oracle {done} searches in {t_search:.2f}s, {equal}/{done} paths equal to the GPU's")
    # the same searches over a pool of host threads (the oracle's C calls
    # release the GIL): nets are independent within an iteration
    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(16, os.cpu_count() or 1))

    def one(req):
        s, d, _, use_roi, bonus = req
        mask = None
        if use_roi:
            mask = np.zeros(lat.N, np.uint8)
            mask[oracle.roi_bfs(lat, s, d, stagnation_bonus=bonus)] = 1
            mask[s] = mask[d] = 1
        return oracle.sssp_canon(lat, costs, s, d, roi_mask=mask)[4]
    t1 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        tr_par = sum(ex.map(one, list(sreqs)[:done]))
    t_par = time.perf_counter() - t1
    log(f"[bench] cpu oracle, {workers} threads: {done} searches in {t_par:.2f}s")
    return {"value": round(traversed / max(t_search, 1e-9) / 1e9, 5), "unit": "GTEPS", "cores": 1,
            "kind": "port",
            "sample": f"the first {done} nets of the request list on the step's cost snapshot: oracle ROI BFS "
                      f"(nets under 125 steps) + label-setting search in C on one thread, {t_search:.1f}s",
            "searches_per_s": round(done / max(t_search, 1e-9), 3),
            "paths_equal_to_gpu": f"{equal}/{done}" if compare else "not compared",
            "threads": {"cores": workers, "value": round(tr_par / max(t_par, 1e-9) / 1e9, 5),
                        "searches_per_s": round(done / max(t_par, 1e-9), 3), "s": round(t_par, 2)}}


if __name__ == "__main__":
    main()
