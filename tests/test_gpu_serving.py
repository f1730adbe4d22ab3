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


@pytest.mark.parametrize("portals", [0, 1])
def test_board_file_routes_like_the_oracle(device, tmp_path, portals):
    """A board read from an ORP file (two-pin pads, or portal terminals on
    every inner layer) routed on the device: its ORS geometry equals the
    oracle's for the same board, four iterations."""
    from oracle.oracle import OraclePathFinder
    from orthoroute_amd.boardio import load_orp, netlist_to_orp, orp_to_board, save_orp, solution_to_ors
    save_orp(netlist_to_orp(uneven(96, 80, 8, 300, seed=31), name="file_board"), tmp_path / "b.orp")
    board = orp_to_board(load_orp(tmp_path / "b.orp"), portals=portals)
    cfg = RouterConfig(iterations=4, delta=16.0, block=1024, lds_nodes=16384)
    dev = DevicePathFinder(board.netlist, cfg, device=device)
    orc = OraclePathFinder(board.netlist, cfg)
    for _ in range(4):
        print(f"[board] portals={portals} iteration {dev.iteration + 1}", flush=True)
        sd = dev.run_iteration("batched")
        so = orc.run_iteration()
        assert sd.overuse == so["edge_over"] + so["node_over"] and sd.routed == so["routed"]
    got = solution_to_ors(board, dev.paths(), dev.history, 0.0)
    want = solution_to_ors(board, orc.paths(), [], 0.0)
    assert got["geometry"] == want["geometry"]
    assert got["statistics"]["total_wirelength_mm"] == want["statistics"]["total_wirelength_mm"]
