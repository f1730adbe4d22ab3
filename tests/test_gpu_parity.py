"""GPU parity: the HIP path against the CPU oracle, bit-exact.

Integer and index results (paths, occupancy, node use, via columns, overuse,
settled / relaxed counts) must be identical; fp32 results (path costs,
history, distances) must be bitwise identical too, because both sides apply
the same fp32 operations in the same order.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OraclePathFinder, sssp_box as oracle_sssp_box
from orthoroute_amd.netgen import plain_64, uneven, backplane_400
from orthoroute_amd.router import DevicePathFinder, LatticeState, RouterConfig
from orthoroute_amd.lattice import LatticeSpec
from orthoroute_amd.sssp import sssp_box

pytestmark = pytest.mark.gpu

# the search engines: the tile engine on every wave, in step form and in
# queue form (one persistent launch per wave); the default hybrid (waves of
# small ROIs one workgroup per net, the rest on the tile engine); and one
# workgroup per net throughout
ENGINES = {"tiles": dict(engine="tiles", tile_queue=False, small_wave_nodes=0),
           "tiles16": dict(engine="tiles", tile_queue=False, small_wave_nodes=0, tile_groups=16),
           "queue": dict(engine="tiles", tile_queue=True, small_wave_nodes=0),
           "hybrid": dict(engine="tiles", small_wave_nodes=4096),
           "wave": dict(engine="wave")}

STAT_KEYS = ["routed", "failed", "edge_over", "edge_over_count", "node_over", "node_over_count", "via_hops",
             "barrels", "settled", "relaxed", "path_hops"]


def _assert_same_iteration(sd, so):
    got = {k: getattr(sd, k) for k in STAT_KEYS}
    want = {k: so[k] for k in STAT_KEYS}
    assert got == want, f"iteration {so['iteration']}: {got} != {want}"


def _assert_same_state(dev: DevicePathFinder, orc: OraclePathFinder):
    st = dev.lat.to_numpy()
    for k in ("occ", "node_use", "via_col"):
        np.testing.assert_array_equal(st[k], getattr(orc, k), err_msg=k)
    for k in ("hist", "node_hist"):
        assert np.array_equal(st[k].view(np.uint32), getattr(orc, k).view(np.uint32)), k
    dp, op = dev.paths(), orc.paths()
    assert len(dp) == len(op)
    for i, (a, b) in enumerate(zip(dp, op)):
        assert np.array_equal(a, b), f"net {i}: path differs"
    dc = dev.path_costs()
    assert np.array_equal(dc.view(np.uint32), orc.path_cost.view(np.uint32))


def _run_both(nl, cfg, iters, mode, device):
    dev = DevicePathFinder(nl, cfg, device=device)
    orc = OraclePathFinder(nl, cfg)
    for _ in range(iters):
        print(f"[parity] {nl.name} mode={mode} iteration {dev.iteration + 1}", flush=True)
        sd = dev.run_iteration(mode)
        so = orc.run_iteration()
        _assert_same_iteration(sd, so)
        _assert_same_state(dev, orc)
    return dev, orc


@pytest.mark.parametrize("engine", ["tiles", "queue", "hybrid", "wave"])
@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_plain_64_ten_iterations_bit_exact(device, mode, engine):
    nl = plain_64()
    cfg = RouterConfig(iterations=10, **ENGINES[engine])
    dev, orc = _run_both(nl, cfg, 10, mode, device)
    assert dev.history[-1].overuse == orc.history[-1]["edge_over"] + orc.history[-1]["node_over"]


@pytest.mark.parametrize("astar", [True, False])
@pytest.mark.parametrize("lds_nodes", [0, 12288])
@pytest.mark.parametrize("delta", [1.0, 3.0, float("inf")])
def test_uneven_small_bit_exact(device, lds_nodes, delta, astar):
    nl = uneven(96, 80, 8, 300, seed=11)
    cfg = RouterConfig(iterations=4, delta=delta, lds_nodes=lds_nodes, astar=astar, engine="wave")
    _run_both(nl, cfg, 4, "batched", device)


@pytest.mark.parametrize("head_nets", [0, 1, 37, 128])
def test_split_planning_bit_exact(device, head_nets):
    """The head of the order planned and launched before the rest (its waves
    after the head's) routes like the sequential order, for any head size."""
    nl = uneven(96, 80, 8, 300, seed=13)
    cfg = RouterConfig(iterations=3, head_nets=head_nets)
    _run_both(nl, cfg, 3, "batched", device)


@pytest.mark.parametrize("lds_nodes", [0, 16384])
def test_wide_bucket_large_block_bit_exact(device, lds_nodes):
    """The tuned launch shape (wide bucket, 1024 lanes) on a congested board."""
    nl = uneven(96, 80, 8, 300, seed=12)
    cfg = RouterConfig(iterations=3, delta=16.0, block=1024, lds_nodes=lds_nodes, engine="wave")
    _run_both(nl, cfg, 3, "batched", device)


@pytest.mark.parametrize("engine", ["tiles", "tiles16", "queue", "hybrid", "wave"])
@pytest.mark.parametrize("block", [256, 512, 1024])
def test_block_sizes_bit_exact(device, block, engine):
    nl = uneven(64, 64, 6, 150, seed=5)
    cfg = RouterConfig(iterations=3, block=block, **ENGINES[engine])
    _run_both(nl, cfg, 3, "batched", device)


@pytest.mark.parametrize("engine", ["tiles", "queue", "hybrid", "wave"])
def test_largest_rois_bit_exact(device, engine):
    """The 16 largest nets of the 800x800x20 board (ROIs of 0.5-2.5 M nodes)
    against the oracle: on the tile engine, hundreds of tiles per net over
    tens of steps; on the per-net engine, keys and bitmaps in the global
    workspace, the near queue spilling past its LDS ring."""
    from orthoroute_amd.netgen import uneven_800
    nl = uneven_800()
    big = np.argsort(-nl.roi_nodes(), kind="stable")[:16]
    sub = nl.select(np.sort(big), "largest16")
    assert sub.roi_nodes().min() > 400_000
    cfg = RouterConfig(iterations=2, delta=16.0, block=1024, lds_nodes=16384, **ENGINES[engine])
    _run_both(sub, cfg, 2, "batched", device)


def test_backplane_subset_bit_exact(device):
    nl = backplane_400()
    sub = nl.subset(200)
    cfg = RouterConfig(iterations=2)
    _run_both(sub, cfg, 2, "batched", device)


@pytest.mark.parametrize("lds_nodes", [0, 12288])
def test_sssp_box_random_weights(device, lds_nodes):
    spec = LatticeSpec(20, 17, 5)
    rng = np.random.default_rng(7)
    w = rng.uniform(1.0, 4.0, 4 * spec.N).astype(np.float32)
    w[spec.neighbours().reshape(-1) < 0] = np.inf
    lat = LatticeState(spec, device=device)
    lat.wcost.copy_(torch.from_numpy(w))
    box = (2, 1, 15, 14)
    for src in (spec.node(5, 5, 0), spec.node(10, 9, 3)):
        dist, parent, rc = sssp_box(lat, box, int(src), full=True, lds_nodes=lds_nodes)
        od, op, _, _, _ = oracle_sssp_box(spec, w, box, int(src))
        assert np.array_equal(dist.cpu().numpy().view(np.uint32), od.view(np.uint32))
        np.testing.assert_array_equal(parent.cpu().numpy(), op)


@pytest.mark.skip(reason="orx_route_dag is switched off on the device: its first gfx950 run faulted "
                         "(illegal address) and the cause is not found; covered by tests/test_cpu_build.py")
@pytest.mark.parametrize("max_groups", [1, 3, 0])
def test_dependency_launch_any_group_count(device, max_groups):
    """The persistent launch gives the sequential result for any number of
    resident workgroups (1 = fully serial claim order)."""
    nl = uneven(80, 70, 6, 200, seed=13)
    cfg = RouterConfig(iterations=3)
    dev = DevicePathFinder(nl, cfg, device=device)
    dev.max_groups = max_groups
    orc = OraclePathFinder(nl, cfg)
    for _ in range(3):
        sd = dev.run_iteration("dag")
        so = orc.run_iteration()
        _assert_same_iteration(sd, so)
        _assert_same_state(dev, orc)


@pytest.mark.parametrize("delta", [1.0, 4.0, float("inf")])
def test_sssp_grid_random_weights(device, delta):
    """Grid-wide search (ORX_F_GRID): dist and parent arrays bit-exact."""
    from orthoroute_amd import _lib
    spec = LatticeSpec(60, 50, 6)
    w = spec.random_edge_costs(3)
    lat = LatticeState(spec, device=device)
    lat.wcost.copy_(torch.from_numpy(w))
    box = (3, 2, 50, 45)
    stats = torch.zeros(8, dtype=torch.int64, device=device)
    for src in (spec.node(10, 10, 0), spec.node(40, 30, 5)):
        dist, parent, rc = sssp_box(lat, box, int(src), full=True, delta=delta, grid=True, stats=stats)
        od, op, _, ns, nr = oracle_sssp_box(spec, w, box, int(src))
        assert np.array_equal(dist.cpu().numpy().view(np.uint32), od.view(np.uint32))
        np.testing.assert_array_equal(parent.cpu().numpy(), op)


@pytest.mark.parametrize("delta", [2.0, float("inf")])
def test_sssp_grid_stops_at_the_sink(device, delta):
    """Grid-wide search with a sink: it stops once the sink is final; the
    sink's distance and parent chain equal the oracle's full search."""
    spec = LatticeSpec(60, 50, 6)
    w = spec.random_edge_costs(8)
    lat = LatticeState(spec, device=device)
    lat.wcost.copy_(torch.from_numpy(w))
    box = (2, 3, 55, 44)
    x0, y0, bw, bh = box

    def local(n):
        z, rem = divmod(int(n), spec.X * spec.Y)
        y, x = divmod(rem, spec.X)
        return (z * bh + (y - y0)) * bw + (x - x0)

    for src, dst in ((spec.node(5, 6, 0), spec.node(50, 40, 3)), (spec.node(30, 20, 5), spec.node(31, 20, 5))):
        print(f"[grid-sink] delta {delta} src {int(src)} dst {int(dst)}", flush=True)
        dist, parent, rc = sssp_box(lat, box, int(src), int(dst), full=False, delta=delta, grid=True)
        od, op, _, _, _ = oracle_sssp_box(spec, w, box, int(src))
        dist, parent = dist.cpu().numpy(), parent.cpu().numpy()
        rd = local(dst)
        assert rc == 0 and dist[rd].view(np.uint32) == od[rd].view(np.uint32)
        chain, ochain = [rd], [rd]
        while parent[chain[-1]] >= 0 and len(chain) <= len(dist):
            chain.append(local(parent[chain[-1]]))
        while op[ochain[-1]] >= 0 and len(ochain) <= len(od):
            ochain.append(local(op[ochain[-1]]))
        assert chain == ochain and chain[-1] == local(src)


@pytest.mark.parametrize("lds_nodes,block", [(0, 1024), (16384, 1024), (12288, 256)])
def test_portal_nets_bit_exact(device, lds_nodes, block):
    """Multi-terminal nets (find_path_multisource_multisink semantics): every
    pad a seed or target on each inner layer of two escape cells, mixed with
    two-pin nets, three iterations."""
    from orthoroute_amd.netgen import with_portals
    nl = with_portals(uneven(96, 80, 8, 300, seed=13), candidates=2, fraction=0.8, seed=3)
    cfg = RouterConfig(iterations=3, delta=16.0, block=block, lds_nodes=lds_nodes)
    _run_both(nl, cfg, 3, "batched", device)


def test_portal_nets_plain_mode_bit_exact(device):
    from orthoroute_amd.netgen import with_portals
    nl = with_portals(uneven(64, 64, 6, 150, seed=5), candidates=3, fraction=1.0, seed=4)
    cfg = RouterConfig(iterations=2)
    _run_both(nl, cfg, 2, "plain", device)


def test_portal_largest_rois_bit_exact(device):
    """Portal form of the 16 largest nets of the 800x800x20 board (54
    targets per net, search state in the global workspace)."""
    from orthoroute_amd.netgen import uneven_800, with_portals
    nl = uneven_800()
    big = np.argsort(-nl.roi_nodes(), kind="stable")[:16]
    sub = with_portals(nl.select(np.sort(big), "largest16"), candidates=3, seed=5)
    cfg = RouterConfig(iterations=1, delta=16.0, block=1024, lds_nodes=16384)
    _run_both(sub, cfg, 1, "batched", device)


def test_nets_without_a_path_commit_nothing(device):
    """A two-pin net whose sink is walled off and a multi-terminal net whose
    best target is one of its seeds end with ORX_NO_PATH and commit nothing,
    as in the oracle."""
    from orthoroute_amd.netgen import Netlist, Terminals
    nl2 = uneven(24, 20, 3, 6, seed=9)
    spec = nl2.spec
    base = spec.base_costs()
    nb = spec.neighbours()
    sink = int(nl2.dst[0])
    for s in range(4):
        u = int(nb[sink, s])
        if u >= 0:
            base[4 * u + (3 - s)] = np.inf
    n = nl2.n
    sx, sy, _ = spec.coords(np.asarray([nl2.src[1]]))
    tx, ty, _ = spec.coords(np.asarray([nl2.dst[1]]))
    a, c2 = int(spec.node(int(sx[0]), int(sy[0]), 1)), int(spec.node(int(tx[0]), int(ty[0]), 1))
    seed_ptr = np.zeros(n + 1, np.int32)
    seed_ptr[2:] = 2
    T = Terminals(seed_ptr, np.asarray([int(nl2.src[1]), a], np.int32), np.zeros(2, np.float32),
                  seed_ptr.copy(), np.asarray([a, c2], np.int32), np.zeros(2, np.float32))
    nl = Netlist(spec, nl2.src, nl2.dst, nl2.box, "nopath", dict(nl2.meta), T)
    cfg = RouterConfig(iterations=1)
    dev = DevicePathFinder(nl, cfg, device=device, base=base)
    orc = OraclePathFinder(nl, cfg, base=base)
    sd = dev.run_iteration("batched")
    so = orc.run_iteration()
    _assert_same_iteration(sd, so)
    _assert_same_state(dev, orc)
    st = dev.status()
    assert st[0] == 1 and st[1] == 1 and sd.failed == 2


def test_hotset_mode_bit_exact(device):
    """Hotset mode: offenders and unrouted nets rerouted (from the device's
    per-net edge and node overuse), every net after a stalled iteration;
    twelve iterations against the oracle's identical policy."""
    nl = uneven(96, 80, 8, 300, seed=11)
    cfg = RouterConfig(iterations=12, hotset=True, delta=16.0, block=1024, lds_nodes=16384)
    dev, orc = _run_both(nl, cfg, 12, "batched", device)
    assert min(s.routed for s in dev.history[1:]) < nl.n


def test_long_nets_on_the_full_board_bit_exact(device):
    """The reference's long-net rule (nets of at least 60 steps here route on
    the whole board): 26 of 120 nets search the full 160x120x6 lattice on
    the tile engine, beside boxed short nets; 3 iterations equal the
    oracle's sequential run."""
    from orthoroute_amd.netgen import long_nets_full_board
    nl = long_nets_full_board(uneven(160, 120, 6, 120, seed=21), 60)
    assert nl.meta["full_board_nets"] == 26
    cfg = RouterConfig(iterations=3, delta=16.0, block=1024, lds_nodes=16384)
    _run_both(nl, cfg, 3, "batched", device)
