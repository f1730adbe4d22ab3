"""The full-size workloads against the CPU oracle on the GPU, bit-exact.

* 1000x1000x12 SSSP query (the microbenchmark's first query: costs U[1,4)
  per physical edge, a seeded 5% x10, present factor 0.5; whole box, no
  early exit): every node's fp32 distance bits and parent equal the
  oracle's heap Dijkstra, on the tile engine and on the grid-wide state
  machine.
* 800x800x20 board, 8000 lognormal nets (seed 3), iterations 1 to 5:
  every path, fp32 path cost, occupancy, node use, via column, history and
  the settled / relaxed counts equal the oracle's sequential run.
* Serving stream: boards drawn from the stream's own generator
  (serving_job: 256x256x16, 512x512x8, 128x128x32, 1024x1024x24), each
  capped at 3 iterations, served together (resident at once, a slot limit
  that makes the last wait) end bit-identical to the same job routed alone.
"""
import dataclasses

import numpy as np
import pytest
import torch

from oracle.oracle import OraclePathFinder, sssp_box as oracle_sssp_box
from orthoroute_amd.lattice import LatticeSpec
from orthoroute_amd.netgen import serving_job, uneven_800
from orthoroute_amd.router import DevicePathFinder, LatticeState, RouterConfig, Workspace
from orthoroute_amd.serve import RouteJob, RoutingServer
from orthoroute_amd.sssp import sssp_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_query():
    """The microbenchmark lattice and its first query, plus the oracle's
    arrays for it (one full heap Dijkstra on 12 M nodes, ~8 s on one core)."""
    spec = LatticeSpec(1000, 1000, 12)
    base = spec.random_edge_costs(1)
    rng = np.random.default_rng(101)
    src, dst = (int(v) for v in rng.integers(0, spec.XY, size=2))
    from oracle.oracle import lattice_weights, params_from
    N = spec.N
    w = lattice_weights(spec, base, np.zeros(4 * N, np.float32), np.zeros(4 * N, np.int32), np.zeros(N, np.float32),
                        np.zeros(N, np.int32), params_from(RouterConfig(), 0.5))
    print("[c1] oracle heap Dijkstra on 12 M nodes", flush=True)
    od, op, _, _, _ = oracle_sssp_box(spec, w, (0, 0, spec.X, spec.Y), src, -1, True)
    return spec, base, w, src, dst, od, op


@pytest.mark.parametrize("tiles,queue", [(True, True), (True, False), (False, False)])
def test_sssp_1000x1000x12_full_query_equals_oracle(device, c1_query, tiles, queue):
    spec, base, w, src, dst, od, op = c1_query
    lat = LatticeState(spec, base=base, device=device)
    lat.reprice(RouterConfig().params(0.5))
    assert np.array_equal(lat.wcost.cpu().numpy().view(np.uint32), w.view(np.uint32))
    ws = Workspace(device)
    print(f"[c1] device query, tiles={tiles} queue={queue}", flush=True)
    dist, parent, rc = sssp_box(lat, (0, 0, spec.X, spec.Y), src, dst, full=True, delta=16.0, block=256, ws=ws,
                                grid=not tiles, tiles=tiles, queue=queue)
    assert rc == 0
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), od.view(np.uint32))
    assert np.array_equal(parent.cpu().numpy(), op)


@pytest.mark.parametrize("queue", [False, True])
def test_uneven_800_iterations_1_to_5_equal_oracle(device, queue):
    nl = uneven_800()
    cfg = RouterConfig(iterations=5, delta=16.0, block=1024, lds_nodes=16384, tile_queue=queue)
    dev = DevicePathFinder(nl, cfg, device=device)
    orc = OraclePathFinder(nl, cfg)
    for it in range(5):
        print(f"[c3] iteration {it + 1} queue={queue}", flush=True)
        sd = dev.run_iteration("batched")
        so = orc.run_iteration()
        for k in ("routed", "failed", "edge_over", "edge_over_count", "node_over", "node_over_count", "via_hops",
                  "barrels", "settled", "relaxed", "path_hops"):
            assert getattr(sd, k) == so[k], (it + 1, k)
        st = dev.lat.to_numpy()
        for k in ("occ", "node_use", "via_col"):
            assert np.array_equal(st[k], getattr(orc, k)), (it + 1, k)
        for k in ("hist", "node_hist"):
            assert np.array_equal(st[k].view(np.uint32), getattr(orc, k).view(np.uint32)), (it + 1, k)
        for i, (a, b) in enumerate(zip(dev.paths(), orc.paths())):
            assert np.array_equal(a, b), (it + 1, i)
        assert np.array_equal(dev.path_costs().view(np.uint32), orc.path_cost.view(np.uint32))


SERVING_SEEDS = (168, 75, 117, 39)   # serving_job boards 256x256x16, 512x512x8, 128x128x32, 1024x1024x24


@pytest.mark.parametrize("queue", [False, True])
def test_serving_stream_boards_served_equal_alone(device, queue):
    boards = [serving_job(np.random.default_rng(s)) for s in SERVING_SEEDS]
    assert [(b.spec.X, b.spec.L) for b in boards] == [(256, 16), (512, 8), (128, 32), (1024, 24)]
    cfg = RouterConfig(iterations=3, stop_at_zero_overuse=True, delta=16.0, block=1024, lds_nodes=16384,
                       tile_queue=queue)
    max_it = 3
    want = []
    for nl in boards:
        print(f"[c4] alone {nl.name}", flush=True)
        pf = DevicePathFinder(nl, dataclasses.replace(cfg, iterations=max_it), stream=torch.cuda.Stream())
        hist = []
        for _ in range(max_it):
            st = pf.run_iteration()
            hist.append(st.overuse)
            if st.overuse == 0 and st.failed == 0:
                break
        want.append((pf.digest(), hist))
        del pf
    print("[c4] served mix", flush=True)
    jobs = [RouteJob(i, nl, 0.01 * i, max_it) for i, nl in enumerate(boards)]
    srv = RoutingServer(cfg, max_slots=3)
    summ = srv.run(jobs, time_limit=300)
    assert summ["completed"] == len(boards) and summ["rejected"] == 0
    assert srv.peak_active >= 2
    for j in srv.finished:
        digest, hist = want[j.job_id]
        assert j.overuse_history == hist, j.netlist.name
        assert j.digest == digest, j.netlist.name
