"""Serving loop: routing jobs for boards of uneven size, admitted while
others are on the GPU.

A job is one board (an X x Y x L lattice) and its two-pin nets, routed by
`PathFinderRouter.route_multiple_nets` to zero overuse (or its iteration
cap).  The reference routes one board per process run (main.py:287-622,
headless ORP -> ORS); this loop keeps several boards resident in HBM at once
and overlaps their negotiations on one GPU:

* Admission is first come, first served by HBM footprint: the head of the
  queue starts once a serving slot is free and its footprint (lattice CSR,
  edge state, search and ROI scratch at the job's search-slot count) fits
  beside the jobs already running.  A job larger than the whole budget is
  rejected.  Jobs that arrive while others run are admitted as slots and
  memory free up (continuous batching across jobs), not after the queue
  drains.
* Each admitted job runs in its own host thread with its own HIP stream and
  orgpu handle, so its kernels interleave with the other jobs' kernels on
  the device.  Every library call releases the GIL while it waits on its
  stream.
* Results do not depend on the mix: a job's negotiation reads only its own
  lattice, cost state and nets (the router's random draws use per-router
  generators), and the searches are exact, so a job returns the same paths
  and per-iteration overuse alone and beside others.

Times per job, from entering route_multiple_nets with the lattice built:
`zero_overuse_s` to the end of the first iteration with zero overuse,
`route_s` to the negotiation's exit (zero overuse with every net routed, or
the iteration cap: a net the reference's ROI rules cannot route keeps the
loop going after the overuse is gone); `latency_s` runs from arrival to the
result.
"""
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .device import SOLO_ARENA, Context, lattice_counts, max_slots

MAX_PATH = 1 << 16          # path staging length the router reserves per search slot
RING_BUCKETS = 8            # bucket ring the router's search uses by default
TOUCH_KEYS = 1024           # keys per touched-chunk record of a search slot
SOLO_MIN = 1024             # searches at once from which the router's batches run on the one-wave engine
SOLO_NODES = 2_000_000      # boards up to this many nodes run every batch on it (several slots)


@dataclass
class Job:
    """One routing job: a board lattice and its two-pin nets in the
    reference's request tuple form (name, src, dst) (UPF:2660-2669)."""
    job_id: int
    X: int
    Y: int
    L: int
    nets: List[Tuple[str, int, int]]
    arrival_s: float = 0.0
    policy: str = "adjacent"
    via_cost: float = 3.0
    max_iterations: int = 40

    @property
    def nodes(self):
        return self.X * self.Y * self.L


@dataclass
class JobResult:
    job_id: int
    nodes: int
    edges: int
    nets: int
    slots: int
    footprint: int
    converged: bool
    iterations: int
    overuse_sum: int
    paths: dict
    trajectory: list
    arrival_s: float
    admit_s: float
    done_s: float
    route_s: float
    zero_overuse_s: Optional[float] = None
    error: Optional[str] = None

    @property
    def latency_s(self):
        return self.done_s - self.arrival_s

    @property
    def wait_s(self):
        return self.admit_s - self.arrival_s


def footprint(X, Y, L, policy, n_nets, slots, roi_limit):
    """Device bytes a job holds while it runs: the CSR (indptr, indices,
    base cost, layer, kind), the edge state (count, present, EMA, history,
    total cost), the keepout owner map, the search scratch of `slots`
    concurrent searches (keys, bucket bitmaps, touch records, path staging;
    with the one-wave engine, which the router uses when at least SOLO_MIN
    searches run at once or the board has at most SOLO_NODES nodes, keys
    over the tiled index space and each search's
    bucket-list arena, and its adjacency records) and the ROI BFS scratch and
    bitmap rows.  An upper estimate of what the engine and the router
    allocate."""
    N, E = lattice_counts(X, Y, L, policy)
    solo = min(int(n_nets), int(slots)) >= SOLO_MIN or (int(slots) > 1 and N <= SOLO_NODES)
    csr = 4 * (N + 1) + 10 * E
    state = 20 * E + 4 * N
    per_slot = 8 * N + (RING_BUCKETS + 3) * N // 8 + 4 * (N // TOUCH_KEYS + 1) + 4 * MAX_PATH * 4 + 64
    if solo:
        tiled = ((X + 1) // 2) * ((Y + 1) // 2) * ((L + 3) // 4) * 16
        per_slot += 8 * (max(tiled, N) - N + 2048) + 8 * SOLO_ARENA + 8 * (SOLO_ARENA // 256)
    search = slots * per_slot + (RING_BUCKETS + 4) * (N * slots // 2048 + 64) * 4 // 32 + (1 << 20)
    if solo:
        # the one-wave engine's adjacency records: 32 B per key index
        tiled = ((X + 1) // 2) * ((Y + 1) // 2) * ((L + 3) // 4) * 16
        search += 32 * (max(tiled, N) + 2048)
    roi_slots = max(1, min(1024, int(n_nets), roi_limit // (9 * N + 1)))
    roi = roi_slots * 9 * N + int(n_nets) * ((N + 31) // 32) * 4
    return int(1.05 * (csr + state + search + roi)) + (64 << 20)


def fit_slots(job, slots, budget, roi_limit):
    """The most concurrent searches, at most `slots`, whose job footprint
    fits `budget` bytes (1 when none does: the job is then rejected)."""
    fp = lambda k: footprint(job.X, job.Y, job.L, job.policy, len(job.nets), k, roi_limit)
    if fp(slots) <= budget:
        return slots
    lo, hi = 1, slots
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if fp(mid) <= budget:
            lo = mid
        else:
            hi = mid - 1
    return lo


def choose_slots(job, cap=1024, budget_bytes=None):
    """Concurrent searches for a job: all its nets at once when the slot
    memory (device.slot_bytes per search) fits budget_bytes (default 70 % of
    the device's memory), at most `cap`."""
    return max(1, min(int(cap), len(job.nets), max_slots(job.nodes, budget_bytes)))


def run_job(job, device=0, slots=None, roi_limit=1 << 30, stream=None):
    """Route one job to completion on its own stream; returns
    (result dict of route_multiple_nets, per-iteration (it, over_sum,
    over_cnt), route seconds, (N, E)).  The result dict also carries
    "_iteration_s": seconds from the start to the end of each iteration."""
    from .algorithms.manhattan.unified_pathfinder import PathFinderConfig, PathFinderRouter
    import torch
    torch.cuda.set_device(device)
    stream = stream or torch.cuda.Stream(device=device)
    traj = []
    with torch.cuda.stream(stream):
        ctx = Context(device, stream=stream)
        try:
            ctx.set_roi_scratch(roi_limit)
            cfg = PathFinderConfig()
            cfg.lattice_policy = job.policy
            cfg.batch_slots = int(slots or choose_slots(job))
            cfg.max_iterations_cap = int(job.max_iterations)
            cfg.iteration_geometry = False     # the progress callback reads only the overuse
            r = PathFinderRouter(cfg)
            r.ctx = ctx
            r.initialize_lattice(job.X, job.Y, job.L, pitch=1.0, via_cost=job.via_cost)
            ne = (r.N, r.graph.dlat.E)
            stream.synchronize()
            t0 = time.perf_counter()
            t_it = []

            def on_iteration(it, tracks, vias, over_sum, over_cnt):
                traj.append((it, over_sum, over_cnt))
                t_it.append(time.perf_counter() - t0)
            res = r.route_multiple_nets(job.nets, iteration_cb=on_iteration)
            stream.synchronize()
            dt = time.perf_counter() - t0
            res["_iteration_s"] = t_it
            del r
        finally:
            stream.synchronize()
            ctx.close()
    return res, traj, dt, ne


class RoutingServer:
    """First-come-first-served admission by HBM footprint over `max_active`
    serving slots (host thread + HIP stream + orgpu handle each).

    hbm_budget: device bytes the running jobs may hold together (default:
    80 % of the device's free memory at start).  roi_limit: ROI BFS scratch
    per job (bytes).  slot_cap: most concurrent searches per job.  on_done:
    called with (job, JobResult or the exception) as each job ends."""

    def __init__(self, device=0, hbm_budget=None, max_active=4, roi_limit=1 << 30, slot_cap=1024, runner=None,
                 on_done=None):
        self.device = int(device)
        self._on_done = on_done          # on_done(job, JobResult or exception), from the job's thread
        if runner is None or hbm_budget is None:
            import torch
            if not torch.cuda.is_available():
                raise RuntimeError("RoutingServer needs a ROCm GPU")
            if hbm_budget is None:
                free, _ = torch.cuda.mem_get_info(self.device)
                hbm_budget = int(0.8 * free)
        # runner(job, device, slots, roi_limit) -> (result dict, trajectory,
        # route seconds, (N, E)); run_job routes on the GPU
        self._runner = runner or run_job
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.roi_limit = int(roi_limit)
        self.slot_cap = int(slot_cap)
        self._cv = threading.Condition()
        self._queue = []                 # (job, future, footprint, slots, arrival perf time)
        self._active = 0
        self._used = 0
        self._closed = False
        self.t0 = time.perf_counter()
        self.peak_active = 0
        self.peak_used = 0
        self.admit_log = []              # (t, job_id, active after admission, bytes in use)
        self._admitter = threading.Thread(target=self._admit_loop, name="orgpu-admit", daemon=True)
        self._admitter.start()

    def now(self):
        return time.perf_counter() - self.t0

    def submit(self, job):
        """Queue a job; returns a Future of its JobResult."""
        fut = Future()
        # as many searches at once as the board allows and the budget holds
        slots = fit_slots(job, choose_slots(job, self.slot_cap), self.budget, self.roi_limit)
        fp = footprint(job.X, job.Y, job.L, job.policy, len(job.nets), slots, self.roi_limit)
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            self._queue.append((job, fut, fp, slots, self.now()))
            self._cv.notify_all()
        return fut

    def close(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
            if wait:
                while self._active or self._queue:
                    self._cv.wait()
        self._admitter.join(timeout=5)

    def _admit_loop(self):
        while True:
            with self._cv:
                while True:
                    if self._queue:
                        job, fut, fp, slots, t_arr = self._queue[0]
                        if fp > self.budget:
                            self._queue.pop(0)
                            fut.set_exception(MemoryError(
                                f"job {job.job_id}: footprint {fp / 2**30:.2f} GiB exceeds the "
                                f"server budget {self.budget / 2**30:.2f} GiB"))
                            self._cv.notify_all()
                            continue
                        if self._active < self.max_active and self._used + fp <= self.budget:
                            break
                    elif self._closed:
                        return
                    self._cv.wait()
                self._queue.pop(0)
                self._active += 1
                self._used += fp
                self.peak_active = max(self.peak_active, self._active)
                self.peak_used = max(self.peak_used, self._used)
                t_admit = self.now()
                self.admit_log.append((t_admit, job.job_id, self._active, self._used))
            th = threading.Thread(target=self._run, args=(job, fut, fp, slots, t_arr, t_admit),
                                  name=f"orgpu-job-{job.job_id}", daemon=True)
            th.start()

    def _run(self, job, fut, fp, slots, t_arr, t_admit):
        try:
            res, traj, dt, (N, E) = self._runner(job, self.device, slots, self.roi_limit)
            t_it = res.get("_iteration_s") or []
            zero = next((t for (it, s, c), t in zip(traj, t_it) if s == 0 and c == 0), None)
            if zero is None and res.get("converged"):
                # the converging iteration exits before its callback (UPF:4023-4027)
                zero = dt
            out = JobResult(job.job_id, N, E, len(job.nets), slots, fp, bool(res.get("converged")),
                            len(traj), int(res.get("overuse_sum") or 0), res.get("paths", {}), traj,
                            t_arr, t_admit, self.now(), dt, zero)
        except BaseException as exc:     # reported through the future
            out = None
            err = exc
        with self._cv:
            self._active -= 1
            self._used -= fp
            self._cv.notify_all()
        if self._on_done is not None:
            try:
                self._on_done(job, out if out is not None else err)
            except Exception:
                pass
        if out is not None:
            fut.set_result(out)
        else:
            fut.set_exception(err)


def replay(server, jobs, time_scale=1.0):
    """Submit each job at its arrival time (seconds from now, times
    time_scale) and wait for all; returns the JobResults in job order
    (a rejected or failed job gives a JobResult with `error`)."""
    futs = []
    start = time.perf_counter()
    for job in sorted(jobs, key=lambda j: j.arrival_s):
        delay = start + job.arrival_s * time_scale - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        futs.append((job, server.submit(job)))
    out = {}
    for job, f in futs:
        try:
            out[job.job_id] = f.result()
        except Exception as exc:
            out[job.job_id] = JobResult(job.job_id, job.nodes, 0, len(job.nets), 0, 0, False, 0, 0, {}, [],
                                        job.arrival_s, float("nan"), float("nan"), float("nan"), error=str(exc))
    return [out[j.job_id] for j in jobs]


def summarize(results, makespan_s):
    """p50 / p99 of time to zero overuse (jobs that reached it), of the
    negotiation's run time and of latency; throughput in jobs per hour."""
    ok = [r for r in results if r.error is None]
    conv = [r for r in ok if r.converged]
    zero = [r.zero_overuse_s for r in ok if r.zero_overuse_s is not None]

    def pct(v, q):
        return round(float(np.percentile(np.asarray(v, dtype=np.float64), q)), 4) if len(v) else None
    route = [r.route_s for r in ok]
    lat = [r.latency_s for r in ok]
    return {"jobs": len(results), "completed": len(ok), "reached_zero_overuse": len(zero),
            "converged": len(conv), "failed": len(results) - len(ok),
            "p50_time_to_zero_overuse_s": pct(zero, 50), "p99_time_to_zero_overuse_s": pct(zero, 99),
            "p50_route_s": pct(route, 50), "p99_route_s": pct(route, 99),
            "p50_latency_s": pct(lat, 50), "p99_latency_s": pct(lat, 99),
            "p50_wait_s": pct([r.wait_s for r in ok], 50),
            "makespan_s": round(makespan_s, 3),
            "jobs_per_hour": round(3600.0 * len(ok) / max(makespan_s, 1e-9), 1)}
