"""CPU tests of the host side: wave planning, launch-plan flattening, the
present-factor schedule, net order, and the C ABI's exported symbols (the
library loads without a GPU; no compute call is made)."""
import re
from pathlib import Path

import numpy as np
import pytest

from orthoroute_amd import _lib
from orthoroute_amd.netgen import backplane_400, plain_64, uneven
from orthoroute_amd.router import _flatten_waves
from orthoroute_amd.waves import (check_disjoint, color_waves, group_waves, net_distance, net_order, plan_waves,
                                  pres_schedule, resolve_pres_fac_max)

REPO = Path(__file__).resolve().parent.parent


def _boards():
    yield plain_64()
    yield uneven(96, 80, 6, 300, seed=5)
    bp = backplane_400()
    yield bp.select(np.arange(0, bp.n, 7))


@pytest.mark.parametrize("k", [0, 1, 2])
def test_native_waves_equal_python_colouring(k):
    nl = list(_boards())[k]
    rng = np.random.default_rng(k)
    order = rng.permutation(nl.n).astype(np.int32)
    R = nl.roi_nodes()
    want = group_waves(color_waves(nl.box, order, nl.spec.X, nl.spec.Y), order, R)
    got = plan_waves(nl.box, order, nl.spec.X, nl.spec.Y, R)
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_tiled_planner_equals_cell_planner(seed):
    """orx_plan_waves (tiled) gives exactly the waves of orx_plan_iteration
    (per cell), on boards whose sides are not multiples of the tile."""
    from orthoroute_amd.waves import plan_iteration
    nl = uneven(203, 157, 4, 900, seed=seed)
    order = np.random.default_rng(seed).permutation(nl.n).astype(np.int32)
    R = nl.roi_nodes()
    wave, sched, _, _ = plan_iteration(nl.box, order, nl.spec.X, nl.spec.Y, R)
    got = plan_waves(nl.box, order, nl.spec.X, nl.spec.Y, R)
    assert np.array_equal(np.concatenate(got), sched)
    assert sum(len(w) for w in got) == nl.n and len(got) == int(wave.max()) + 1


@pytest.mark.parametrize("seed", [4, 9, 12])
def test_tiled_planner_on_dense_serving_boards_and_subsets(seed):
    """Boards whose small boxes cover each tile partly many times (the tiles'
    stamp lists fill and spill into cells), full and hotset-like orders."""
    from orthoroute_amd.netgen import serving_job
    nl = serving_job(np.random.default_rng(seed), 0)
    R = nl.roi_nodes()
    rng = np.random.default_rng(seed)
    for order in (rng.permutation(nl.n), rng.permutation(nl.n)[: nl.n // 3]):
        order = order.astype(np.int32)
        want = group_waves(color_waves(nl.box, order, nl.spec.X, nl.spec.Y), order, R)
        got = plan_waves(nl.box, order, nl.spec.X, nl.spec.Y, R)
        assert len(got) == len(want)
        assert all(np.array_equal(a, b) for a, b in zip(got, want))


def test_waves_are_disjoint_and_keep_sequential_order():
    nl = uneven(120, 120, 4, 400, seed=11)
    order = np.random.default_rng(3).permutation(nl.n).astype(np.int32)
    waves = plan_waves(nl.box, order, nl.spec.X, nl.spec.Y, nl.roi_nodes())
    assert sorted(np.concatenate(waves).tolist()) == list(range(nl.n))
    wave_of = np.empty(nl.n, np.int64)
    for i, w in enumerate(waves):
        assert check_disjoint(nl.box, w)
        wave_of[w] = i
    pos = np.empty(nl.n, np.int64)
    pos[order] = np.arange(nl.n)
    b = nl.box.astype(np.int64)
    x0, y0, x1, y1 = b[:, 0], b[:, 1], b[:, 0] + b[:, 2], b[:, 1] + b[:, 3]
    ov = (x0[:, None] < x1[None]) & (x0[None] < x1[:, None]) & (y0[:, None] < y1[None]) & (y0[None] < y1[:, None])
    np.fill_diagonal(ov, False)
    i, j = np.nonzero(ov & (pos[:, None] < pos[None, :]))
    # an overlapping pair routes in routing order: the earlier net in an earlier wave
    assert (wave_of[i] < wave_of[j]).all()


def test_flatten_waves_offsets_and_sizes():
    R = np.array([10, 64, 65, 1, 200, 7], np.int64)
    waves = [np.array([2, 0], np.int32), np.array([4], np.int32), np.array([1, 5, 3], np.int32)]
    flat, off, ptr, wmax, wws, need = _flatten_waves(waves, R)
    assert flat.tolist() == [2, 0, 4, 1, 5, 3]
    assert ptr.tolist() == [0, 2, 3, 6]
    assert off.tolist() == [0, 128, 0, 0, 64, 128]
    assert wmax.tolist() == [65, 200, 64]
    assert wws.tolist() == [128 + 64, 256, 64 + 64 + 64]
    assert need == max(128 + 64, 256, 64 + 64 + 64)
    assert (off % 64 == 0).all()
    e = _flatten_waves([], R)
    assert e[2].tolist() == [0] and e[5] == 0


def test_pres_schedule_is_exact_fp32_and_capped():
    s = pres_schedule(12, 0.5, 1.5, 32.0)
    assert s.dtype == np.float32
    want = [0.5 * 1.5 ** k for k in range(12)]
    for got, w in zip(s, want):
        assert float(got) == min(w, 32.0)  # dyadic rationals: exact while 3^k < 2^24
    assert resolve_pres_fac_max(6) == 64.0          # configured 64 beats the 4-signal-layer floor
    assert resolve_pres_fac_max(24, configured=1.0) == 128.0


def test_net_order_hardest_first_with_rotation():
    nl = plain_64()
    d = net_distance(nl.spec, nl.src, nl.dst)
    o1 = net_order(1, d, np.zeros(nl.n, np.int64), np.zeros(nl.n, np.int32))
    assert sorted(o1.tolist()) == list(range(nl.n))
    assert np.array_equal(o1, net_order(1, d, np.zeros(nl.n, np.int64), np.zeros(nl.n, np.int32)))
    # congestion raises difficulty: a net with overuse moves towards the front
    over = np.zeros(nl.n, np.int64)
    easiest = int(np.argmin(np.where(d > 0, d, d.max() + 1)))
    over[easiest] = 10 ** 6
    hops = np.ones(nl.n, np.int32)
    o2 = net_order(2, d, over, hops)
    # hardest first, then the whole order rotated left by random.Random(42 + it)
    import random
    rot = random.Random(44).randint(0, min(5, nl.n // 10))
    assert list(o2).index(easiest) == (nl.n - rot) % nl.n


def _declared_functions():
    text = (REPO / "include" / "orthoroute_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(orx_\w+)\s*\(", text, flags=re.M)))


def test_capi_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert "orx_route_wave" in names and "orx_sssp_csr" in names and "orx_route_waves" in names
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound
    lib = _lib.load()
    for n in names:
        assert getattr(lib, n) is not None
    assert lib.orx_version() >= 10000


@pytest.mark.parametrize("dims", [(7, 5, 4), (16, 9, 1), (3, 3, 6)])
def test_device_built_base_costs_equal_host_array(dims):
    """LatticeState builds the default base costs with torch ops on the
    device; the array must equal the host construction (run here on CPU)."""
    from orthoroute_amd.lattice import LatticeSpec
    spec = LatticeSpec(*dims)
    got = spec.base_costs_device("cpu").numpy()
    want = spec.base_costs()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    m = spec.base_cost_minima().reshape(-1, 4)
    b = want.reshape(-1, 4)
    assert m[:, 1:3][np.isfinite(m[:, 1:3])].min() == b[:, 1:3][np.isfinite(b[:, 1:3])].min()
    via = b[:, [0, 3]][np.isfinite(b[:, [0, 3]])]
    mv = m[:, [0, 3]][np.isfinite(m[:, [0, 3]])]
    assert (via.size == 0 and mv.size == 0) or mv.min() == via.min()


def test_hotset_filter_keeps_offenders_in_order():
    from orthoroute_amd.router import hotset_filter
    order = np.array([4, 2, 0, 3, 1, 5], np.int32)
    edge_over = np.array([0, 2, 0, 0, 0, 1])
    node_over = np.array([0, 0, 0, 1, 0, 0])
    hops = np.array([5, 7, 0, 9, 4, 3])          # net 2 has no path
    assert hotset_filter(order, edge_over, node_over, hops).tolist() == [2, 3, 1, 5]
    assert hotset_filter(order, np.zeros(6), np.zeros(6), np.ones(6)).tolist() == []


def test_native_planner_on_a_subset_order():
    """A hotset routes a subset of the nets: the planner's waves over that
    subset equal the reference colouring of the same subset, and a net id
    outside the netlist is rejected."""
    from orthoroute_amd.waves import color_waves, group_waves, plan_waves
    nl = uneven(96, 80, 6, 300, seed=2)
    rng = np.random.default_rng(5)
    order = rng.permutation(nl.n)[:170].astype(np.int32)
    R = nl.roi_nodes()
    got = plan_waves(nl.box, order, nl.spec.X, nl.spec.Y, R)
    want = group_waves(color_waves(nl.box, order, nl.spec.X, nl.spec.Y), order, R)
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert a.tolist() == b.tolist()
    with pytest.raises(ValueError):
        plan_waves(nl.box, np.array([0, nl.n], np.int32), nl.spec.X, nl.spec.Y, R)


def test_adaptive_delta_widens_only_after_many_thin_phases():
    """Per-net engine: delta grows x4 after an iteration with more than 8
    threshold phases per net and rounds under a quarter of a workgroup (the
    congested case); an iteration whose searches end in their first bucket
    leaves it, and it stops at delta_max.  The tile engine has no buckets:
    it never widens."""
    from types import SimpleNamespace

    from orthoroute_amd.router import DevicePathFinder, IterationStats, RouterConfig

    def stats(phases, items, rounds, routed=100):
        return IterationStats(iteration=1, pres_fac=0.5, routed=routed, failed=0, edge_over=0, edge_over_count=0,
                              node_over=0, node_over_count=0, via_hops=0, barrels=0, items=items, rounds=rounds,
                              phases=phases)

    tiles = SimpleNamespace(cfg=RouterConfig(delta=16.0, block=1024, engine="tiles"), delta_mult=1.0)
    DevicePathFinder._adapt_delta(tiles, stats(phases=2700, items=48_000, rounds=1000))
    assert tiles.delta_mult == 1.0
    job = SimpleNamespace(cfg=RouterConfig(delta=16.0, block=1024, engine="wave"), delta_mult=1.0)
    DevicePathFinder._adapt_delta(job, stats(phases=5, items=60_000, rounds=1000))      # A*: first bucket
    assert job.delta_mult == 1.0
    DevicePathFinder._adapt_delta(job, stats(phases=2700, items=48_000, rounds=1000))   # thin phases
    assert job.delta_mult == 4.0
    DevicePathFinder._adapt_delta(job, stats(phases=2700, items=900_000, rounds=1000))  # full rounds
    assert job.delta_mult == 4.0
    for _ in range(20):
        DevicePathFinder._adapt_delta(job, stats(phases=2700, items=48_000, rounds=1000))
    assert 16.0 * job.delta_mult == 1048576.0
    off = SimpleNamespace(cfg=RouterConfig(delta=16.0, block=1024, adapt_delta=False), delta_mult=1.0)
    DevicePathFinder._adapt_delta(off, stats(phases=2700, items=48_000, rounds=1000))
    assert off.delta_mult == 1.0


@pytest.mark.parametrize("X,Y,L", [(7, 5, 2), (9, 6, 4), (64, 64, 6)])
def test_lattice_arc_counts(X, Y, L):
    """Directed arc counts of the lattice CSR.  The two-layer case is the
    reference's own known answer (tests/test_lattice.py:131-146: lateral
    2*xs*(ys-1) + 2*ys*(xs-1), vias 2*xs*ys); with planar arcs on every layer
    (the ORTHO_LATERAL_ALL form, unified_pathfinder.py:1691-1701) each V
    layer adds 2*X*(Y-1), each H layer 2*Y*(X-1), each via level 2*X*Y.
    64x64x6 gives 89,344 arcs."""
    from orthoroute_amd.lattice import LatticeSpec
    spec = LatticeSpec(X, Y, L)
    ip, ix, _, _ = spec.to_csr()
    z = np.repeat(np.arange(L), X * Y)
    src = np.repeat(np.arange(spec.N), np.diff(ip))
    via = (ix // (X * Y)) != z[src]
    n_v = (L + 1) // 2                       # layers 0, 2, ... are vertical
    n_h = L // 2
    assert int((~via).sum()) == n_v * 2 * X * (Y - 1) + n_h * 2 * Y * (X - 1)
    assert int(via.sum()) == 2 * X * Y * (L - 1)
    if (X, Y, L) == (64, 64, 6):
        assert len(ix) == 89_344


def test_long_nets_get_the_whole_board():
    """long_nets_full_board: |dx| + |dy| >= steps -> box (0, 0, X, Y); shorter
    nets keep their bounding box + margin; the threshold is inclusive."""
    import numpy as np
    from orthoroute_amd.lattice import LatticeSpec
    from orthoroute_amd.netgen import Netlist, long_nets_full_board, roi_boxes
    spec = LatticeSpec(50, 40, 4)
    src = np.array([spec.node(1, 1, 0), spec.node(5, 5, 0), spec.node(10, 3, 0)], np.int32)
    dst = np.array([spec.node(21, 1, 0), spec.node(5, 24, 0), spec.node(12, 4, 0)], np.int32)
    nl = Netlist(spec, src, dst, roi_boxes(spec, src, dst, 4), "t")
    out = long_nets_full_board(nl, 20)          # lengths 20, 19, 3
    assert out.box[0].tolist() == [0, 0, 50, 40]
    assert out.box[1].tolist() == nl.box[1].tolist() and out.box[2].tolist() == nl.box[2].tolist()
    assert out.meta["full_board_nets"] == 1
