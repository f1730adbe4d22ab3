"""CPU tests of the serving loop's host logic (orthoroute_amd.serve) with a
stand-in executor: FIFO admission by footprint and slots, continuous
batching at iteration boundaries, stop rules, rejection of jobs larger than
the budget, latency bookkeeping.  The device executor is covered by the
GPU tests."""
from types import SimpleNamespace

import numpy as np
import pytest

from orthoroute_amd.netgen import plain_64, uneven
from orthoroute_amd.router import RouterConfig, footprint_estimate, split_waves
from orthoroute_amd.serve import RouteJob, RoutingServer, poisson_jobs


class FakeRunner:
    """Overuse falls by one per iteration from a per-board start value; an
    iteration is 'in flight' for two polls."""
    log = []

    def __init__(self, netlist, cfg, stream):
        self.netlist, self.cfg, self.stream = netlist, cfg, stream
        self.left = int(netlist.meta.get("start_overuse", 3))
        self.inflight = 0
        self.it = 0
        FakeRunner.log.append(("admit", netlist.name, stream))

    def launch_iteration(self):
        assert self.inflight == 0
        self.inflight = 2

    def ready(self):
        self.inflight -= 1
        return self.inflight <= 0

    def finish_iteration(self):
        self.it += 1
        self.inflight = 0
        self.left = max(0, self.left - 1)
        return SimpleNamespace(overuse=self.left, failed=0, routed=self.netlist.n, gpu_ms=1.0)

    def digest(self):
        return f"{self.netlist.name}:{self.it}"


def _job(i, arrival, start_overuse, side=64, n=100, max_it=50):
    nl = uneven(side, side, 4, n, seed=i)
    nl.meta["start_overuse"] = start_overuse
    nl.name = f"job{i}"
    return RouteJob(i, nl, arrival, max_it)


def test_footprint_estimate_counts_lattice_arena_and_workspace():
    nl = plain_64()
    fp = footprint_estimate(nl)
    N = nl.spec.N
    assert fp >= N * 72 + int(nl.roi_nodes().sum()) * 4
    assert fp < N * 72 + int(nl.roi_nodes().sum()) * 40 + (4 << 20)


def test_split_waves_respects_cap_and_order():
    R = np.array([100, 2000, 50, 700, 64], np.int64)
    waves = [np.array([1, 3, 0, 2, 4])]
    parts = split_waves(waves, R, 2200)
    assert np.concatenate(parts).tolist() == [1, 3, 0, 2, 4]
    for p in parts:
        assert int(((R[p] + 63) // 64 * 64).sum()) <= 2200
    with pytest.raises(ValueError):
        split_waves(waves, R, 1000)


def test_server_runs_every_job_to_its_stop_rule():
    FakeRunner.log = []
    jobs = [_job(0, 0.0, 3), _job(1, 0.0, 0), _job(2, 0.001, 9, max_it=4)]
    srv = RoutingServer(RouterConfig(), hbm_budget=10 ** 12, runner=FakeRunner, streams=["s0", "s1"],
                        idle_sleep=0.0)
    summ = srv.run(jobs, time_limit=30)
    assert summ["completed"] == 3 and summ["rejected"] == 0
    by_id = {j.job_id: j for j in srv.finished}
    assert by_id[0].iterations == 3 and by_id[0].converged
    assert by_id[1].iterations == 1 and by_id[1].converged      # zero overuse after the first pass
    assert by_id[2].iterations == 4 and not by_id[2].converged  # iteration limit
    assert by_id[2].overuse_history == [8, 7, 6, 5]
    # only two slots: the third job waits for one and reuses its stream
    assert srv.peak_active == 2
    assert {s for _, _, s in FakeRunner.log} <= {"s0", "s1"}
    for j in srv.finished:
        assert j.finished_at >= j.admitted_at >= j.arrival - 1e-9
        assert j.latency >= 0
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0


def test_admission_by_footprint_is_fifo_and_rejects_oversize():
    FakeRunner.log = []
    small = [_job(i, 0.0, 2) for i in range(4)]
    fp = footprint_estimate(small[0].netlist)
    huge = _job(9, 0.0, 1, side=512, n=50)
    budget = int(fp * 2.5)          # room for two small jobs at once
    assert footprint_estimate(huge.netlist) > budget
    srv = RoutingServer(RouterConfig(), hbm_budget=budget, runner=FakeRunner, streams=list("abcdefgh"),
                        idle_sleep=0.0)
    summ = srv.run(small + [huge], time_limit=30)
    assert summ["rejected"] == 1 and summ["completed"] == 4
    assert srv.peak_active == 2 and srv.peak_in_use <= budget
    admitted = [name for ev, name, _ in FakeRunner.log if ev == "admit"]
    assert admitted == ["job0", "job1", "job2", "job3"]
    assert [j for j in srv.finished if j.rejected][0].job_id == 9


def test_poisson_arrivals_are_seeded_and_increasing():
    a = poisson_jobs(6, 4.0, seed=3, make=lambda rng: plain_64())
    b = poisson_jobs(6, 4.0, seed=3, make=lambda rng: plain_64())
    ta = [j.arrival for j in a]
    assert ta == [j.arrival for j in b]
    assert all(x < y for x, y in zip(ta, ta[1:]))
