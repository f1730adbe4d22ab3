"""Host-thread build of the kernel sources (tests/cpu_build) under
AddressSanitizer: the route kernels -- one launch per net, and the
dependency-driven persistent launch with sequential and concurrent
workgroups -- run on the CPU and must match the oracle bit for bit without a
single out-of-bounds access."""
import shutil

import numpy as np
import pytest

from oracle.oracle import OraclePathFinder
from orthoroute_amd.netgen import uneven
from orthoroute_amd.router import RouterConfig

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def driver():
    from tests.cpu_build.harness import build_driver
    return build_driver()


@pytest.mark.parametrize("mode,groups,conc,lds_nodes", [(1, 0, 1, 12288), (0, 3, 3, 12288), (1, 0, 1, 0),
                                                        (2, 0, 3, 0), (2, 0, 2, 12288)])
def test_route_kernels_on_host_threads(driver, tmp_path, monkeypatch, mode, groups, conc, lds_nodes):
    """mode 1: one launch per net; 0: the dependency-driven launch; 2: the
    wave sequence of orx_route_waves, whose waves get their global workspace
    range cleared chip-wide before they run (lds_nodes=0 keeps every net's
    keys there)."""
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", str(conc))
    nl = uneven(24, 20, 3, 10, seed=4)
    cfg = RouterConfig(iterations=2, lds_nodes=lds_nodes)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=mode, max_groups=groups)
    o = OraclePathFinder(nl, cfg)
    for _ in range(2):
        d = c.run_iteration(timeout=300)
        od = o.run_iteration()
        for k in ("routed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        assert np.array_equal(c.occ, o.occ)
        assert np.array_equal(c.hist.view(np.uint32), o.hist.view(np.uint32))


def test_congested_board_many_threshold_phases(driver, tmp_path, monkeypatch):
    """A small board packed with nets, three iterations: history and present
    costs spread the distances far beyond delta, so the searches run many
    threshold phases with the adaptive next threshold -- still bit-exact."""
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", "2")
    nl = uneven(16, 12, 3, 24, seed=8)
    cfg = RouterConfig(iterations=3, delta=1.0, lds_nodes=0)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=2)
    o = OraclePathFinder(nl, cfg)
    phases = 0
    for _ in range(3):
        d = c.run_iteration(timeout=600)
        od = o.run_iteration()
        for k in ("routed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        phases += d["phases"]
    assert phases > 100


def _bfs_layer_max(spec, box, src):
    """Largest hop layer of a breadth-first search from src inside the box."""
    x0, y0, w, h = box
    nb = spec.neighbours()
    x, y, z = spec.coords(np.arange(spec.N))
    inside = (x >= x0) & (x < x0 + w) & (y >= y0) & (y < y0 + h)
    seen = np.zeros(spec.N, bool)
    seen[src] = True
    layer = np.array([src])
    best = 1
    while len(layer):
        cand = nb[layer].reshape(-1)
        cand = np.unique(cand[cand >= 0])
        cand = cand[inside[cand] & ~seen[cand]]
        seen[cand] = True
        layer = cand
        best = max(best, len(layer))
    return best


def test_box_search_near_queue_spills_past_the_lds_ring(driver, tmp_path):
    """A pure frontier search (delta = inf) whose rounds hold more nodes than
    the 2048-entry LDS ring of the near queue: dist and parent arrays equal
    the oracle's, with every access in bounds."""
    from oracle.oracle import sssp_box as oracle_sssp_box
    from orthoroute_amd.lattice import LatticeSpec
    from tests.cpu_build.harness import sssp_box_host_threads
    spec = LatticeSpec(184, 182, 8)
    w = spec.base_costs()
    box = (2, 1, 180, 180)
    src = int(spec.node(92, 91, 0))
    assert _bfs_layer_max(spec, box, src) > 2048
    dist, parent, stats = sssp_box_host_threads(spec, w, box, src, -1, tmp_path, grid=False, delta=float("inf"),
                                                block=256, timeout=900)
    od, op, _, _, _ = oracle_sssp_box(spec, w, box, src)
    assert np.array_equal(dist.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(parent, op)


@pytest.mark.parametrize("delta", [1.0, 4.0, float("inf")])
def test_grid_search_on_host_threads(driver, tmp_path, delta):
    """The grid-wide state machine (captured step graph replayed until done;
    each launch takes its step from the launch before): bit-exact with the
    oracle over many threshold phases (delta 1), a few (4) and one (inf)."""
    from oracle.oracle import sssp_box as oracle_sssp_box
    from orthoroute_amd.lattice import LatticeSpec
    from tests.cpu_build.harness import sssp_box_host_threads
    spec = LatticeSpec(26, 20, 3)
    w = spec.random_edge_costs(9)
    box = (3, 2, 20, 16)
    src = int(spec.node(10, 10, 0))
    dist, parent, stats = sssp_box_host_threads(spec, w, box, src, -1, tmp_path, grid=True, delta=delta, block=256)
    od, op, _, _, _ = oracle_sssp_box(spec, w, box, src)
    assert np.array_equal(dist.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(parent, op)
    if delta == 1.0:
        assert stats[4] > 10          # threshold phases (splits) taken


def test_grid_search_stops_at_the_sink_on_host_threads(driver, tmp_path):
    """The grid search with a sink stops once the sink is final: its distance
    and its parent chain equal the oracle's full search."""
    from oracle.oracle import sssp_box as oracle_sssp_box
    from orthoroute_amd.lattice import LatticeSpec
    from tests.cpu_build.harness import sssp_box_host_threads
    spec = LatticeSpec(26, 20, 3)
    w = spec.random_edge_costs(5)
    box = (1, 1, 24, 18)
    src, dst = int(spec.node(3, 4, 0)), int(spec.node(20, 15, 1))
    dist, parent, _ = sssp_box_host_threads(spec, w, box, src, dst, tmp_path, grid=True, delta=2.0, block=256)
    od, op, _, _, _ = oracle_sssp_box(spec, w, box, src)
    x0, y0, bw, bh = box

    def local(n):
        z, rem = divmod(int(n), spec.X * spec.Y)
        y, x = divmod(rem, spec.X)
        return (z * bh + (y - y0)) * bw + (x - x0)

    rd = local(dst)
    assert dist[rd].view(np.uint32) == od[rd].view(np.uint32)
    chain, v = [], rd
    while v >= 0 and len(chain) <= len(dist):
        chain.append(v)
        p = parent[v]
        v = local(p) if p >= 0 else -1
    ochain, v = [], rd
    while v >= 0 and len(ochain) <= len(od):
        ochain.append(v)
        p = op[v]
        v = local(p) if p >= 0 else -1
    assert chain == ochain and chain[-1] == local(src)


@pytest.mark.parametrize("mode,lds_nodes", [(2, 0), (1, 12288)])
def test_portal_nets_on_host_threads(driver, tmp_path, monkeypatch, mode, lds_nodes):
    """Multi-terminal nets (seeds and targets on every inner layer of two
    escape cells per pad) mixed with two-pin nets: paths, fp32 totals,
    occupancy, history and search counts equal the oracle's."""
    from orthoroute_amd.netgen import with_portals
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", "2")
    nl = with_portals(uneven(24, 20, 4, 10, seed=4), candidates=2, fraction=0.7, seed=1)
    assert 0 < int((nl.terms.counts()[0] > 0).sum()) < nl.n
    cfg = RouterConfig(iterations=2, lds_nodes=lds_nodes)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=mode)
    o = OraclePathFinder(nl, cfg)
    for _ in range(2):
        d = c.run_iteration(timeout=300)
        od = o.run_iteration()
        for k in ("routed", "failed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed",
                  "path_hops"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        assert np.array_equal(c.path_cost.view(np.uint32), o.path_cost.view(np.uint32))
        assert np.array_equal(c.occ, o.occ)
        assert np.array_equal(c.node_use, o.node_use)
        assert np.array_equal(c.hist.view(np.uint32), o.hist.view(np.uint32))


def test_nets_without_a_path_commit_nothing(driver, tmp_path):
    """A two-pin net whose sink is walled off, and a multi-terminal net whose
    best target is one of its seeds: both end with ORX_NO_PATH, an empty
    path and nothing committed, as in the oracle; the other nets route."""
    from orthoroute_amd.netgen import Netlist, Terminals
    from tests.cpu_build.harness import CpuBuildPathFinder
    nl2 = uneven(24, 20, 3, 6, seed=9)
    spec = nl2.spec
    base = spec.base_costs()
    nb = spec.neighbours()
    sink = int(nl2.dst[0])
    for s in range(4):                      # every arc into net 0's sink
        u = int(nb[sink, s])
        if u >= 0:
            base[4 * u + (3 - s)] = np.inf
    n = nl2.n
    a = int(spec.node(*[int(v[0]) for v in spec.coords(np.asarray([nl2.src[1]]))][:2], 1))
    c2 = int(spec.node(*[int(v[0]) for v in spec.coords(np.asarray([nl2.dst[1]]))][:2], 1))
    seed_ptr = np.zeros(n + 1, np.int32)
    seed_ptr[2:] = 2                        # net 1: seeds {src pad, a}, targets {a, dst cell on layer 1}
    T = Terminals(seed_ptr, np.asarray([int(nl2.src[1]), a], np.int32), np.zeros(2, np.float32),
                  seed_ptr.copy(), np.asarray([a, c2], np.int32), np.zeros(2, np.float32))
    nl = Netlist(spec, nl2.src, nl2.dst, nl2.box, "nopath", dict(nl2.meta), T)
    cfg = RouterConfig(iterations=1)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=2, base=base)
    o = OraclePathFinder(nl, cfg, base=base)
    d = c.run_iteration(timeout=300)
    od = o.run_iteration()
    assert d["failed"] == od["failed"] == 2
    assert d["status"][0] == 1 and d["status"][1] == 1      # ORX_NO_PATH
    assert c.path_len[0] == 0 and c.path_len[1] == 0
    for a_, b_ in zip(c.paths(), o.paths()):
        assert np.array_equal(a_, b_)
    assert np.array_equal(c.occ, o.occ)
    assert np.array_equal(c.node_use, o.node_use)
    assert np.array_equal(c.via_col, o.via_col)


def test_roi_bfs_on_host_threads(driver, tmp_path):
    """The device BFS region of interest (count / scan / scatter / finalize
    kernels included) on host threads: ROI list, global_to_roi, depth and
    meeting level equal the oracle's, marks and map left clean by the
    release, every access in bounds.  Balls on a lattice, blocked arcs,
    portal seeds and the truncation rule."""
    from oracle.oracle import roi_bfs
    from orthoroute_amd.lattice import LatticeSpec
    from tests.cpu_build.harness import roi_bfs_host_threads
    spec = LatticeSpec(40, 34, 4)
    ip, ix, _, _ = spec.to_csr()
    N = spec.N
    rng = np.random.default_rng(9)
    blocked = rng.random(len(ix)) < 0.15
    cases = [dict(initial_radius=6), dict(initial_radius=40, stagnation_bonus=1.3),
             dict(initial_radius=30, portal_seeds=[(5, 0.1), (N - 3, 0.2)], max_nodes=300),
             dict(initial_radius=12, blocked=blocked), dict(initial_radius=80, max_nodes=100000)]
    cases.append(dict(initial_radius=40, same=True))      # src == dst
    for i, kw in enumerate(cases):
        src, dst = (int(v) for v in rng.choice(N, 2, replace=False))
        if kw.pop("same", False):
            dst = src
        roi, g2r, out, clean = roi_bfs_host_threads(ip, ix, N, src, dst, tmp_path / f"c{i}", **kw)
        oroi, og2r, info = roi_bfs(ip, ix, N, src, dst, **kw)
        assert np.array_equal(roi, oroi), i
        assert np.array_equal(g2r, og2r), i
        assert out[1] == info["depth"] and out[3] == info["seen"] == out[5], i
        assert (out[2] if out[2] >= 0 else None) == info["meeting_depth"], i
        assert out[4] == 0 and clean.tolist() == [0, 0], i


@pytest.mark.parametrize("X,Y,L,n,seed,conc,cap", [
    (40, 34, 6, 16, 4, 1, -1),     # class 6, every step launch
    (40, 34, 6, 16, 4, 3, 1),      # concurrent workgroups; one step, the rest in the drain
    (30, 26, 5, 12, 7, 2, 0),      # generic class (5 layers): everything in the drain
    (34, 30, 20, 6, 2, 2, -1),     # class 20 (the 800 x 800 x 20 tiles)
])
def test_tile_engine_on_host_threads(driver, tmp_path, monkeypatch, X, Y, L, n, seed, conc, cap):
    """The tile engine (orx_route_tiles): boxes of several tiles, step
    budgets from ample to none (the single-workgroup drain finishes the
    search), workgroups run one at a time or concurrently -- paths, fp32
    costs, occupancy, history and settled / relaxed counts equal the
    oracle's, with no out-of-bounds access."""
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", str(conc))
    nl = uneven(X, Y, L, n, seed=seed)
    cfg = RouterConfig(iterations=2)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=3, extra_meta={"tile_steps_cap": cap})
    o = OraclePathFinder(nl, cfg)
    for _ in range(2):
        d = c.run_iteration(timeout=600)
        od = o.run_iteration()
        for k in ("routed", "failed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed",
                  "path_hops"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        assert np.array_equal(c.path_cost.view(np.uint32), o.path_cost.view(np.uint32))
        assert np.array_equal(c.occ, o.occ)
        assert np.array_equal(c.hist.view(np.uint32), o.hist.view(np.uint32))


def test_tile_engine_phase_cap_on_host_threads(driver, tmp_path, monkeypatch):
    """Routing-wave visits capped at three phases (a visit that still changes
    keys stops and its tile is visited again, rescanned from scratch, in the
    next step) equal the oracle's.  The default cap (four) runs in every
    other routing test here, the uncapped visit in the whole-box queries."""
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", "2")
    monkeypatch.setenv("ORX_TILE_PHASE_CAP", "3")
    nl = uneven(30, 26, 6, 12, seed=7)
    cfg = RouterConfig(iterations=2)
    # one step launch, then the single-workgroup drain (capped visits go back on its list)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=3, extra_meta={"tile_steps_cap": 1})
    o = OraclePathFinder(nl, cfg)
    for _ in range(2):
        d = c.run_iteration(timeout=600)
        od = o.run_iteration()
        for k in ("routed", "failed", "edge_over", "node_over", "settled", "relaxed", "path_hops"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        assert np.array_equal(c.path_cost.view(np.uint32), o.path_cost.view(np.uint32))
        assert np.array_equal(c.occ, o.occ)
        assert np.array_equal(c.hist.view(np.uint32), o.hist.view(np.uint32))


@pytest.mark.parametrize("X,Y,L,n,seed,conc", [
    (40, 34, 6, 16, 4, 1),     # class 6, workgroups one at a time (the first one does every visit)
    (40, 34, 6, 16, 4, 4),     # concurrent workgroups taking tiles from the ring
    (30, 26, 5, 12, 7, 3),     # generic class (5 layers)
    (34, 30, 20, 6, 2, 3),     # class 20 (the 800 x 800 x 20 tiles)
])
def test_tile_engine_queue_form_on_host_threads(driver, tmp_path, monkeypatch, X, Y, L, n, seed, conc):
    """The queue form of the tile engine (ORX_F_ASYNC: one persistent launch
    per wave, tiles visited from a ring as they are queued, a state word per
    tile): paths, fp32 costs, occupancy, history and counts equal the
    oracle's, with workgroups one at a time or concurrent, and no
    out-of-bounds access."""
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", str(conc))
    nl = uneven(X, Y, L, n, seed=seed)
    cfg = RouterConfig(iterations=2)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=3, extra_meta={"flags": 1 | 32})
    o = OraclePathFinder(nl, cfg)
    for _ in range(2):
        d = c.run_iteration(timeout=600)
        od = o.run_iteration()
        for k in ("routed", "failed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed",
                  "path_hops"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        assert np.array_equal(c.path_cost.view(np.uint32), o.path_cost.view(np.uint32))
        assert np.array_equal(c.occ, o.occ)
        assert np.array_equal(c.hist.view(np.uint32), o.hist.view(np.uint32))


def test_tile_queue_deadline_commits_nothing(driver, tmp_path, monkeypatch):
    """The queue form's bounded wait: with a deadline far below a visit's
    time (ORX_TILE_RUN_SECONDS), waiting workgroups give up and flag the
    launch; every net of such a wave then fails with ORX_E_OVERFLOW and
    commits nothing, the launch still ends, and the nets of the waves before
    the first failed one equal the oracle's (later waves see the costs
    without the failed nets' paths, so they may route differently; which
    waves miss the deadline depends on the host's load)."""
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", "4")
    monkeypatch.setenv("ORX_TILE_RUN_SECONDS", "0.000001")
    nl = uneven(40, 34, 6, 16, seed=4)
    cfg = RouterConfig(iterations=1)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=3, extra_meta={"flags": 1 | 32})
    o = OraclePathFinder(nl, cfg)
    d = c.run_iteration(timeout=600)
    o.run_iteration()
    st = d["status"]
    assert set(np.unique(st).tolist()) <= {0, -3}
    assert (st == -3).sum() >= 1
    failed = np.flatnonzero(st == -3)
    assert (c.path_len[failed] == 0).all()
    first_failed_wave = int(c.last_wave[failed].min())
    for k, (a, b) in enumerate(zip(c.paths(), o.paths())):
        if st[k] == 0 and c.last_wave[k] < first_failed_wave:
            assert np.array_equal(a, b), k
    # occupancy is exactly what the committed paths hold
    occ = np.zeros_like(c.occ)
    nb = nl.spec.neighbours()
    for pth in c.paths():
        for u, v in zip(pth[:-1], pth[1:]):
            s = int(np.flatnonzero(nb[u] == v)[0])
            occ[4 * u + s] += 1
            occ[4 * v + (3 - s)] += 1
    assert np.array_equal(occ, c.occ)


@pytest.mark.parametrize("conc", [1, 4])
def test_tile_queue_box_query_on_host_threads(driver, tmp_path, monkeypatch, conc):
    """A whole-box query on the queue form (orx_sssp_box with ORX_F_TILES |
    ORX_F_ASYNC) over a box of many tiles: dist bits and parents equal the
    oracle's heap Dijkstra."""
    from oracle.oracle import sssp_box as oracle_sssp_box
    from orthoroute_amd.lattice import LatticeSpec
    from tests.cpu_build.harness import sssp_box_host_threads
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", str(conc))
    spec = LatticeSpec(70, 52, 6)
    w = spec.random_edge_costs(5)
    box = (2, 3, 66, 47)
    src = int(spec.node(40, 20, 0))
    dist, parent, stats = sssp_box_host_threads(spec, w, box, src, -1, tmp_path, grid=False, flags=16 | 32)
    od, op, _, _, _ = oracle_sssp_box(spec, w, box, src)
    assert np.array_equal(dist.view(np.uint32), od.view(np.uint32))
    np.testing.assert_array_equal(parent, op)


def test_tile_engine_portal_nets_on_host_threads(driver, tmp_path, monkeypatch):
    """Multi-terminal nets on the tile engine: targets are sinks (never
    expanded, also across a tile boundary), the best target by total."""
    from orthoroute_amd.netgen import with_portals
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", "2")
    nl = with_portals(uneven(36, 30, 6, 10, seed=4), candidates=2, fraction=0.7, seed=1)
    cfg = RouterConfig(iterations=2)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=3)
    o = OraclePathFinder(nl, cfg)
    for _ in range(2):
        d = c.run_iteration(timeout=600)
        od = o.run_iteration()
        for k in ("routed", "failed", "edge_over", "node_over", "settled", "relaxed", "path_hops"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        assert np.array_equal(c.path_cost.view(np.uint32), o.path_cost.view(np.uint32))


@pytest.mark.parametrize("mode", [2, 3])
def test_broken_plan_is_refused(driver, tmp_path, mode):
    """A plan whose workspace offset of one slot is off its 64-node
    alignment: that net is refused (ORX_E_ARG, no path, nothing committed)
    on both engines, and nothing reads or writes outside the buffers."""
    from tests.cpu_build.harness import CpuBuildPathFinder
    nl = uneven(24, 20, 3, 6, seed=9)
    cfg = RouterConfig(iterations=1)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=mode, extra_meta={"bad_slot": 1})
    d = c.run_iteration(timeout=300)
    st = d["status"]
    assert (st == -1).sum() == 1
    k = int(np.flatnonzero(st == -1)[0])
    assert c.path_len[k] == 0
    assert d["failed"] >= 1


def _lattice_csr(X, Y, L, seed):
    from orthoroute_amd.lattice import LatticeSpec, in_rank_and_transpose
    spec = LatticeSpec(X, Y, L)
    ip, ix = spec.to_csr()[:2]
    rng = np.random.default_rng(seed)
    costs = rng.uniform(1.0, 4.0, len(ix)).astype(np.float32)
    costs[rng.random(len(ix)) < 0.05] *= 10
    rank, in_ip, in_src = in_rank_and_transpose(spec.N, ip, ix)
    return spec, np.asarray(ip), np.asarray(ix), costs, rank, in_ip, in_src


@pytest.mark.parametrize("conc,roi", [(1, False), (4, True)])
def test_csr_whole_chip_search_on_host_threads(driver, tmp_path, monkeypatch, conc, roi):
    """The whole-chip form of the CSR seam search (ORX_F_GRID: grid-wide
    near-far steps as a device state machine) and the parallel parent walk,
    on a lattice CSR with random costs: multi-seed / multi-target with seed
    and target costs, node penalties, whole graph and an ROI map.  Keys'
    distance bits, the chosen target and the path equal the oracle's."""
    from oracle.oracle import csr_path, sssp_csr
    from tests.cpu_build.harness import csr_search_host_threads
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", str(conc))
    spec, ip, ix, costs, rank, in_ip, in_src = _lattice_csr(16, 14, 3, 3)
    rng = np.random.default_rng(7)
    seeds = rng.choice(spec.N, 5, replace=False).astype(np.int32)
    targets = rng.choice(spec.N, 7, replace=False).astype(np.int32)
    sc = rng.uniform(0, 3, len(seeds)).astype(np.float32)
    tc = rng.uniform(0, 3, len(targets)).astype(np.float32)
    roi_nodes = g2r = None
    R = spec.N
    if roi:
        keep = np.sort(np.unique(np.concatenate([rng.choice(spec.N, spec.N * 3 // 4, replace=False), seeds,
                                                 targets]))).astype(np.int32)
        roi_nodes = keep
        g2r = np.full(spec.N, -1, np.int32)
        g2r[keep] = np.arange(len(keep), dtype=np.int32)
        R = len(keep)
    pen = rng.uniform(0, 0.5, R).astype(np.float32)
    keys, best, n, path = csr_search_host_threads(spec.N, ip, ix, rank, in_ip, in_src, costs, seeds, targets,
                                                  tmp_path, seed_cost=sc, tg_cost=tc, roi_nodes=roi_nodes, g2r=g2r,
                                                  pen=pen, grid=True, walk=True)
    rc, obest, od, op = sssp_csr(spec.N, ip, ix, costs, rank, in_ip, in_src, roi_nodes=roi_nodes, g2r=g2r, pen=pen,
                                 seeds=seeds, seed_cost=sc, targets=targets, tg_cost=tc)
    assert best == obest
    opath = csr_path(obest, op, roi_nodes, g2r)
    assert path == opath
    # every key the oracle settled below the chosen target's total is final
    dist = (keys >> np.uint64(32)).astype(np.uint32).view(np.float32)
    fin = np.isfinite(od) & (od < od[(g2r[obest] if roi else obest)])
    assert np.array_equal(dist[fin].view(np.uint32), od[fin].view(np.uint32))


def test_csr_whole_chip_full_search_and_one_lane_walk_agree(driver, tmp_path, monkeypatch):
    """A single-source whole-graph search to completion (no targets) on the
    whole-chip form: every node's distance bits equal the oracle's; the
    parallel walk and the one-lane walk give the same path to a far node."""
    from oracle.oracle import csr_path, sssp_csr
    from tests.cpu_build.harness import csr_search_host_threads
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", "3")
    spec, ip, ix, costs, rank, in_ip, in_src = _lattice_csr(24, 18, 3, 5)
    src = np.array([spec.node(3, 4, 0)], np.int32)
    far = np.array([spec.node(22, 16, 2)], np.int32)
    k_grid, b1, n1, p1 = csr_search_host_threads(spec.N, ip, ix, rank, in_ip, in_src, costs, src, far,
                                                 tmp_path / "a", grid=True, walk=True, seed_pen=False)
    k_one, b2, n2, p2 = csr_search_host_threads(spec.N, ip, ix, rank, in_ip, in_src, costs, src, far,
                                                tmp_path / "b", grid=False, walk=False, seed_pen=False)
    assert b1 == b2 == int(far[0]) and p1 == p2 and n1 == n2 == len(p1) > 33
    rc, ob, od, op = sssp_csr(spec.N, ip, ix, costs, rank, in_ip, in_src, seeds=src, targets=far, seed_pen=False)
    assert p1 == csr_path(ob, op)
    k_all, _, _, _ = csr_search_host_threads(spec.N, ip, ix, rank, in_ip, in_src, costs, src, np.zeros(0, np.int32),
                                             tmp_path / "c", grid=True, walk=False, seed_pen=False)
    rc, _, od_all, _ = sssp_csr(spec.N, ip, ix, costs, rank, in_ip, in_src, seeds=src, targets=(), seed_pen=False)
    dist = (k_all >> np.uint64(32)).astype(np.uint32).view(np.float32)
    assert np.array_equal(dist.view(np.uint32), od_all.view(np.uint32))
