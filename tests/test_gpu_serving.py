"""GPU tests of the serving stream and the captured launch sequence.

* Every job served in a mix (continuous batching: several jobs resident,
  their iterations in flight together on separate streams, a slot limit
  that makes later arrivals wait) ends with the same result as the same job
  routed alone: identical digest (paths, fp32 path costs, occupancy, node
  use) and identical per-iteration overuse.  Tolerance: none (bit-exact).
* An iteration whose wave launches are captured in one hipGraph equals the
  directly launched iteration, over several iterations, with the program's
  own eager work and synchronisations between them.
"""
import dataclasses

import numpy as np
import pytest
import torch

from orthoroute_amd.netgen import plain_64, uneven
from orthoroute_amd.router import DevicePathFinder, RouterConfig
from orthoroute_amd.serve import RouteJob, RoutingServer

pytestmark = pytest.mark.gpu


def _boards():
    return [uneven(64, 64, 6, 150, seed=21), uneven(96, 80, 8, 300, seed=22), plain_64(),
            uneven(48, 48, 4, 80, seed=23), uneven(128, 96, 6, 400, seed=24)]


def _alone(nl, cfg, max_it):
    pf = DevicePathFinder(nl, dataclasses.replace(cfg, iterations=max_it), stream=torch.cuda.Stream())
    hist = []
    for _ in range(max_it):
        st = pf.run_iteration()
        hist.append(st.overuse)
        if st.overuse == 0 and st.failed == 0:
            break
    return pf.digest(), hist


def test_served_jobs_equal_alone(device):
    cfg = RouterConfig(iterations=10, stop_at_zero_overuse=True, delta=4.0, block=256)
    boards = _boards()
    max_it = 10
    print("alone runs", flush=True)
    want = [_alone(nl, cfg, max_it) for nl in boards]
    jobs = [RouteJob(i, nl, 0.02 * i, max_it) for i, nl in enumerate(boards)]
    print("served mix", flush=True)
    srv = RoutingServer(cfg, max_slots=3)
    summ = srv.run(jobs, time_limit=300)
    assert summ["completed"] == len(boards) and summ["rejected"] == 0
    assert srv.peak_active >= 2          # jobs really were resident together
    for j in srv.finished:
        digest, hist = want[j.job_id]
        assert j.overuse_history == hist, j.netlist.name
        assert j.digest == digest, j.netlist.name
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0


def test_graph_captured_iteration_equals_direct(device):
    nl = uneven(96, 80, 8, 300, seed=22)
    base = RouterConfig(iterations=4, delta=4.0, block=256)
    a = DevicePathFinder(nl, dataclasses.replace(base, graph=False))
    b = DevicePathFinder(nl, dataclasses.replace(base, graph=True))
    for it in range(4):
        print(f"iteration {it + 1}", flush=True)
        sa = a.run_iteration()
        sb = b.run_iteration()
        torch.cuda.synchronize()
        assert sb.extra["graph"] and not sa.extra["graph"]
        for f in ("routed", "failed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed",
                  "path_hops"):
            assert getattr(sa, f) == getattr(sb, f), (it, f)
        assert a.digest() == b.digest()
    pa, pb = a.paths(), b.paths()
    assert all(np.array_equal(x, y) for x, y in zip(pa, pb))
