"""Serving loop host logic on the CPU: admission by HBM footprint, serving
slots, first-come-first-served order, continuous admission while other jobs
run, rejection of jobs larger than the budget, and the serving workload
generator.  The GPU runner is replaced by a timed stand-in; the GPU test of
the same loop is tests/test_gpu_serving.py."""
import threading
import time

import numpy as np
import pytest

from orthoroute_amd import workloads
from orthoroute_amd.serving import Job, RoutingServer, choose_slots, footprint, replay, summarize


def _job(jid, X=40, Y=40, L=4, n=10, arrival=0.0):
    return Job(jid, X, Y, L, [(f"n{i}", i, i + 1) for i in range(n)], arrival)


class Recorder:
    """Stand-in runner: sleeps per job and records what was running."""

    def __init__(self, server_ref, durations):
        self.server_ref = server_ref
        self.durations = durations
        self.lock = threading.Lock()
        self.running = {}
        self.max_bytes = 0
        self.max_running = 0
        self.order = []

    def __call__(self, job, device, slots, roi_limit):
        srv = self.server_ref[0]
        fp = footprint(job.X, job.Y, job.L, job.policy, len(job.nets), slots, roi_limit)
        with self.lock:
            self.running[job.job_id] = fp
            self.order.append(job.job_id)
            self.max_bytes = max(self.max_bytes, sum(self.running.values()))
            self.max_running = max(self.max_running, len(self.running))
            assert sum(self.running.values()) <= srv.budget
        time.sleep(self.durations.get(job.job_id, 0.02))
        with self.lock:
            del self.running[job.job_id]
        dur = self.durations.get(job.job_id, 0.02)
        traj = [(1, 5, 2), (2, 0, 0)]
        return {"converged": True, "overuse_sum": 0, "paths": {nm: [s, d] for nm, s, d in job.nets},
                "_iteration_s": [dur / 2, dur]}, traj, dur, (job.nodes, 4 * job.nodes)


def test_footprint_grows_with_board_nets_and_slots():
    base = footprint(64, 64, 6, "adjacent", 100, 16, 1 << 30)
    assert footprint(128, 128, 6, "adjacent", 100, 16, 1 << 30) > base
    assert footprint(64, 64, 6, "adjacent", 400, 16, 1 << 30) > base
    assert footprint(64, 64, 6, "adjacent", 100, 64, 1 << 30) > base
    # the search keys alone: 8 bytes per node per slot
    assert footprint(64, 64, 6, "adjacent", 100, 64, 1 << 30) - base >= 48 * 8 * 64 * 64 * 6


def test_choose_slots_respects_memory_budget():
    from orthoroute_amd.device import max_slots, slot_bytes
    small = _job(0, n=30)
    assert choose_slots(small, budget_bytes=16 << 30) == 30
    big = Job(1, 2000, 2000, 24, [("a", 0, 1)] * 2000)   # 96 M nodes
    k = choose_slots(big, budget_bytes=64 << 30)
    assert k == max_slots(big.nodes, 64 << 30) and 1 < k < 2000
    # keys (8 B) and 11 bucket-bitmap bits per node and search, per whole chunk of 2048 nodes
    assert k * slot_bytes(big.nodes) <= 64 << 30 < (k + 1) * slot_bytes(big.nodes)
    assert slot_bytes(1) == slot_bytes(2048) and slot_bytes(2049) > slot_bytes(2048)
    assert max_slots(10, budget_bytes=1 << 50) == 2048


def test_admission_budget_slots_and_fifo():
    # the footprint at the slot count the server picks for these jobs
    fps = [footprint(40, 40, 4, "adjacent", 10, choose_slots(_job(0)), 1 << 20)] * 6
    budget = int(2.5 * fps[0])          # two jobs fit at once
    ref = [None]
    rec = Recorder(ref, {i: 0.05 for i in range(6)})
    srv = RoutingServer(hbm_budget=budget, max_active=4, roi_limit=1 << 20, runner=rec)
    ref[0] = srv
    jobs = [_job(i) for i in range(6)]
    res = replay(srv, jobs)
    srv.close()
    assert all(r.error is None for r in res)
    assert rec.max_running == 2 and srv.peak_active == 2
    assert srv.peak_used <= budget
    assert rec.order == list(range(6))                       # first come, first served
    assert [a[1] for a in srv.admit_log] == list(range(6))


def test_serving_slots_cap_concurrency():
    ref = [None]
    rec = Recorder(ref, {i: 0.05 for i in range(5)})
    srv = RoutingServer(hbm_budget=1 << 40, max_active=3, roi_limit=1 << 20, runner=rec)
    ref[0] = srv
    replay(srv, [_job(i) for i in range(5)])
    srv.close()
    assert rec.max_running == 3


def test_late_arrival_is_admitted_while_others_run():
    """A job that arrives while a long one runs starts at once (a free slot
    and memory) and finishes first: admission does not wait for the queue
    to drain."""
    ref = [None]
    rec = Recorder(ref, {0: 0.6, 1: 0.05})
    srv = RoutingServer(hbm_budget=1 << 40, max_active=2, roi_limit=1 << 20, runner=rec)
    ref[0] = srv
    res = replay(srv, [_job(0, arrival=0.0), _job(1, arrival=0.1)])
    srv.close()
    a, b = res
    assert b.admit_s < a.done_s and b.done_s < a.done_s
    assert b.wait_s < 0.1


def test_oversize_job_is_rejected_and_the_rest_run():
    ref = [None]
    rec = Recorder(ref, {})
    small = footprint(40, 40, 4, "adjacent", 10, choose_slots(_job(0)), 1 << 20)
    srv = RoutingServer(hbm_budget=2 * small, max_active=2, roi_limit=1 << 20, runner=rec)
    ref[0] = srv
    res = replay(srv, [_job(0), _job(1, X=400, Y=400, L=8, n=200), _job(2)])
    srv.close()
    assert res[0].error is None and res[2].error is None
    assert res[1].error is not None and "exceeds the server budget" in res[1].error
    s = summarize(res, 1.0)
    assert s["completed"] == 2 and s["failed"] == 1 and s["converged"] == 2
    # time to zero overuse: the end of the first iteration with none
    assert s["reached_zero_overuse"] == 2 and res[0].zero_overuse_s == 0.02
    assert s["p50_time_to_zero_overuse_s"] == 0.02


def test_runner_error_frees_the_slot():
    ref = [None]

    def boom(job, device, slots, roi_limit):
        if job.job_id == 0:
            raise RuntimeError("device lost")
        return Recorder(ref, {})(job, device, slots, roi_limit)
    srv = RoutingServer(hbm_budget=1 << 40, max_active=1, roi_limit=1 << 20, runner=boom)
    ref[0] = srv
    res = replay(srv, [_job(0), _job(1)])
    srv.close()
    assert "device lost" in res[0].error and res[1].error is None


def test_cfg5_generator_is_seeded_and_in_range():
    a = workloads.cfg5(n_jobs=12, n_min=1e5, n_max=2e6)
    b = workloads.cfg5(n_jobs=12, n_min=1e5, n_max=2e6)
    assert [(j.X, j.L, j.nets[:3], j.arrival_s) for j in a] == [(j.X, j.L, j.nets[:3], j.arrival_s) for j in b]
    arr = np.array([j.arrival_s for j in a])
    assert arr[0] == 0.0 and np.all(np.diff(arr) > 0)
    for j in a:
        assert 4 <= j.L <= 24 and j.X == j.Y
        assert 0.5e5 <= j.nodes <= 2.5e6
        assert 1 <= len(j.nets) <= 30000
        plane = j.X * j.Y
        pins = [p for _, s, d in j.nets for p in (s, d)]
        assert all(0 <= p < plane for p in pins) and len(set(pins)) == len(pins)
    d = workloads.cfg5(n_jobs=4, net_density=0.75)
    for j in d:
        assert len(j.nets) == max(1, min(int(round(0.75 * j.X * j.Y / (2 * j.X))), j.X * j.Y // 2))


def test_large_job_gets_fewer_searches_instead_of_rejection():
    """A job whose footprint at the slot cap exceeds the budget is admitted
    with as many concurrent searches as the budget holds (the job then runs
    alone), and is rejected only when one search does not fit."""
    from orthoroute_amd.serving import fit_slots
    big = _job(0, X=600, Y=600, L=8, n=2000)
    full = footprint(big.X, big.Y, big.L, big.policy, len(big.nets), 1024, 1 << 20)
    budget = full // 3
    k = fit_slots(big, 1024, budget, 1 << 20)
    assert 1 <= k < 1024
    assert footprint(big.X, big.Y, big.L, big.policy, len(big.nets), k, 1 << 20) <= budget
    assert footprint(big.X, big.Y, big.L, big.policy, len(big.nets), k + 1, 1 << 20) > budget
    ref = [None]
    rec = Recorder(ref, {0: 0.02})
    srv = RoutingServer(hbm_budget=budget, max_active=2, roi_limit=1 << 20, runner=rec)
    ref[0] = srv
    res = replay(srv, [big])
    srv.close()
    # (choose_slots caps the searches by the device's memory first: one
    # search without a GPU)
    assert res[0].error is None and res[0].slots == fit_slots(big, choose_slots(big, 1024), budget, 1 << 20)
    tiny = RoutingServer(hbm_budget=1 << 20, max_active=1, roi_limit=1 << 20, runner=rec)
    ref[0] = tiny
    res = replay(tiny, [big])
    tiny.close()
    assert res[0].error is not None and "exceeds" in res[0].error


def test_footprint_counts_the_one_wave_engine_on_small_boards():
    """A board of at most SOLO_NODES nodes runs every batch of several slots
    on the one-wave engine: its footprint holds each slot's bucket-list
    arena; one slot (the plain form) does not."""
    from orthoroute_amd.device import SOLO_ARENA
    from orthoroute_amd.serving import SOLO_NODES
    assert 100 * 100 * 4 <= SOLO_NODES
    one = footprint(100, 100, 4, "adjacent", 10, 1, 1 << 20)
    ten = footprint(100, 100, 4, "adjacent", 10, 10, 1 << 20)
    assert ten - one >= 10 * 8 * SOLO_ARENA
