"""Serving stream: continuous batching of routing jobs on one GPU.

Routing jobs (a board + its nets) arrive over time.  The server keeps each
admitted job's lattice, cost, history and occupancy state resident in HBM
(one DevicePathFinder per job) and runs PathFinder on it until zero overuse
or its iteration limit.

* Admission: FIFO.  The job at the head enters as soon as a slot (a HIP
  stream of default priority) is free and its HBM footprint
  (router.footprint_estimate) fits in the budget next to the jobs already
  resident.  A job larger than the whole budget is rejected.
* Continuous batching: every resident job always has one iteration in
  flight on its own stream.  The host collects whichever finishes first,
  records it and immediately enqueues that job's next iteration, so jobs
  that arrive while others run join at the next iteration boundary and
  their kernels overlap on the device.  A finished job frees its slot and
  memory at once, which admits the next waiting job.
* Each job's result does not depend on the mix: a job's kernels read and
  write only its own buffers on its own stream, and the job's iteration
  sequence is the same as alone (checked by digest in the tests).

Reported per job: arrival, admission and finish times, iterations, final
overuse, result digest; over the stream: throughput and p50/p99 latency.

Start the process with at least as many HIP hardware queues as slots
(GPU_MAX_HW_QUEUES, read when the HIP runtime starts; HIP's default is 4):
streams that share a hardware queue execute in submission order, which ties
a short job's iterations to a long job's.  Each job's launch path copies
through pinned buffers on its own stream only, so no job waits on another
job's work from the host side.
This replaces nothing in the reference (its engine routes one board per
process, UP:6163-7534); the PathFinder inside each job is the device path.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from .netgen import Netlist, serving_job
from .router import RouterConfig, footprint_estimate


@dataclass
class RouteJob:
    job_id: int
    netlist: Netlist
    arrival: float                    # seconds from the start of the stream
    max_iterations: int = 50
    admitted_at: float | None = None
    finished_at: float | None = None
    iterations: int = 0
    overuse_history: list = field(default_factory=list)
    failed: int = 0
    converged: bool = False
    rejected: str = ""
    digest: str = ""
    footprint: int = 0
    nets_routed: int = 0
    gpu_ms: float = 0.0
    # called as on_finish(job, runner) when the job retires, before its
    # router is dropped (for example boardio.ors_writer: paths -> .ORS)
    on_finish: Callable | None = None

    @property
    def latency(self) -> float:
        return float("nan") if self.finished_at is None else self.finished_at - self.arrival

    @property
    def queued(self) -> float:
        return float("nan") if self.admitted_at is None else self.admitted_at - self.arrival


def poisson_jobs(n_jobs: int, rate: float, seed: int = 0, max_iterations: int = 50,
                 make: Callable = serving_job) -> list[RouteJob]:
    """n_jobs boards with exponential inter-arrival times (rate jobs/s)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        t += float(rng.exponential(1.0 / rate))
        jobs.append(RouteJob(i, make(rng), t, max_iterations))
    return jobs


def _device_runner(netlist: Netlist, cfg: RouterConfig, stream):
    from .router import DevicePathFinder
    return DevicePathFinder(netlist, cfg, stream=stream)


class RoutingServer:
    """Continuous-batching server over one GPU (see the module docstring).

    ``runner(netlist, cfg, stream)`` builds a job's executor; the default is
    a DevicePathFinder.  It must offer launch_iteration(), ready(),
    finish_iteration() -> stats with .overuse/.failed/.routed/.gpu_ms,
    converged and digest()."""

    def __init__(self, cfg: RouterConfig | None = None, hbm_budget: int | None = None, max_slots: int = 8,
                 device="cuda", runner: Callable | None = None, streams: list | None = None,
                 clock: Callable[[], float] = time.perf_counter, idle_sleep: float = 2e-4,
                 progress: Callable[[str], None] | None = None):
        self.cfg = cfg or RouterConfig(iterations=50, stop_at_zero_overuse=True)
        self.runner = runner or _device_runner
        if streams is None:
            import torch
            streams = [torch.cuda.Stream(device) for _ in range(max_slots)]
        if hbm_budget is None:
            import torch
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget = int(free * 0.85)
        self.budget = int(hbm_budget)
        self.streams = list(streams)
        self.free_slots = list(range(len(self.streams)))[::-1]
        self.pending: deque[RouteJob] = deque()
        self.active: dict[int, tuple] = {}
        self.finished: list[RouteJob] = []
        self.in_use = 0
        self.peak_in_use = 0
        self.peak_active = 0
        self.clock = clock
        self.idle_sleep = idle_sleep
        self.progress = progress      # called with a line per retired job
        self._t0 = None

    # -- bookkeeping ---------------------------------------------------------
    def now(self) -> float:
        return self.clock() - self._t0

    def submit(self, job: RouteJob) -> None:
        job.footprint = footprint_estimate(job.netlist, self.cfg)
        if job.footprint > self.budget:
            job.rejected = f"footprint {job.footprint} B exceeds the {self.budget} B budget"
            job.finished_at = None
            self.finished.append(job)
            return
        self.pending.append(job)

    def _job_cfg(self, job: RouteJob) -> RouterConfig:
        import dataclasses
        return dataclasses.replace(self.cfg, iterations=job.max_iterations, stop_at_zero_overuse=True,
                                   collect_stats=False)

    def _admit(self) -> bool:
        admitted = False
        while self.pending and self.free_slots:
            job = self.pending[0]
            if self.in_use + job.footprint > self.budget:
                break  # FIFO: the head waits for memory
            self.pending.popleft()
            slot = self.free_slots.pop()
            runner = self.runner(job.netlist, self._job_cfg(job), self.streams[slot])
            job.admitted_at = self.now()
            self.in_use += job.footprint
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            self.active[job.job_id] = (job, runner, slot)
            self.peak_active = max(self.peak_active, len(self.active))
            self._launch(runner)
            admitted = True
        return admitted

    def _retire(self, job: RouteJob, runner, slot: int) -> None:
        job.digest = runner.digest()
        job.finished_at = self.now()
        if job.on_finish is not None:
            job.on_finish(job, runner)
        del self.active[job.job_id]
        self.in_use -= job.footprint
        self.free_slots.append(slot)
        self.finished.append(job)
        if self.progress is not None:
            self.progress(f"job {job.job_id} {job.netlist.name} done at {job.finished_at:.1f} s: "
                          f"latency {job.latency:.1f} s, {job.iterations} iterations, "
                          f"{len(self.finished)} finished, {len(self.active)} resident, {len(self.pending)} waiting")

    def _poll(self) -> bool:
        progressed = False
        for jid in list(self.active):
            job, runner, slot = self.active[jid]
            if not runner.ready():
                continue
            st = runner.finish_iteration()
            progressed = True
            job.iterations += 1
            job.overuse_history.append(int(st.overuse))
            job.failed = int(st.failed)
            job.nets_routed += int(st.routed)
            job.gpu_ms += float(st.gpu_ms)
            if (st.overuse == 0 and st.failed == 0) or job.iterations >= job.max_iterations:
                job.converged = st.overuse == 0 and st.failed == 0
                self._retire(job, runner, slot)
                del runner
            else:
                self._launch(runner)
        return progressed

    def _launch(self, runner) -> None:
        """Enqueue the runner's next iteration, telling it how many jobs
        share the device now (a runner with ``concurrent_jobs`` picks launch
        shapes that leave room for the others; results do not depend on
        it)."""
        if hasattr(runner, "concurrent_jobs"):
            runner.concurrent_jobs = max(1, len(self.active))
        runner.launch_iteration()

    # -- the loop ------------------------------------------------------------
    def run(self, jobs: list[RouteJob], time_limit: float | None = None) -> dict:
        """Feed ``jobs`` at their arrival times and serve until all finish."""
        jobs = sorted(jobs, key=lambda j: j.arrival)
        self._t0 = self.clock()
        i = 0
        while i < len(jobs) or self.pending or self.active:
            now = self.now()
            if time_limit is not None and now > time_limit:
                raise TimeoutError(f"serving stream exceeded {time_limit} s")
            while i < len(jobs) and jobs[i].arrival <= now:
                self.submit(jobs[i])
                i += 1
            busy = self._admit()
            busy = self._poll() or busy
            if not busy:
                if not self.active and not self.pending and i < len(jobs):
                    time.sleep(max(0.0, min(jobs[i].arrival - self.now(), 0.05)))
                else:
                    time.sleep(self.idle_sleep)
        return self.summary()

    def summary(self) -> dict:
        done = [j for j in self.finished if not j.rejected]
        lat = np.array([j.latency for j in done], np.float64)
        makespan = max((j.finished_at for j in done), default=0.0)
        return {
            "jobs": len(self.finished), "completed": len(done),
            "rejected": sum(1 for j in self.finished if j.rejected),
            "converged": sum(1 for j in done if j.converged),
            "makespan_s": makespan,
            "jobs_per_s": len(done) / makespan if makespan > 0 else 0.0,
            "nets_routed_per_s": sum(j.nets_routed for j in done) / makespan if makespan > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "mean_queue_s": float(np.mean([j.queued for j in done])) if done else None,
            "peak_resident_jobs": self.peak_active,
            "peak_footprint_gb": self.peak_in_use / 1e9,
            "budget_gb": self.budget / 1e9,
        }
