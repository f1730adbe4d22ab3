"""CPU checks of the oracle PathFinder driver: determinism, bookkeeping
invariants, A*/Dijkstra equivalence and convergence on small boards."""
import numpy as np
import pytest

from oracle.oracle import OraclePathFinder
from orthoroute_amd.netgen import plain_64, uneven
from orthoroute_amd.router import RouterConfig


def _recount(o):
    """Occupancy, node use and via columns recomputed from the paths."""
    spec = o.spec
    occ = np.zeros_like(o.occ)
    use = np.zeros_like(o.node_use)
    col = np.zeros_like(o.via_col)
    nb = spec.neighbours()
    for p in o.paths():
        use[p] += 1
        prev_vert = False
        for a, b in zip(p[:-1], p[1:]):
            s = int(np.nonzero(nb[a] == b)[0][0])
            occ[4 * a + s] += 1
            occ[4 * b + 3 - s] += 1
            vert = s in (0, 3)
            if vert and not prev_vert:
                col[a % spec.XY] += 1
            prev_vert = vert
    return occ, use, col


def test_bookkeeping_matches_paths_and_paths_are_valid():
    nl = uneven(60, 50, 6, 120, seed=8)
    o = OraclePathFinder(nl, RouterConfig(iterations=3))
    for _ in range(3):
        o.run_iteration()
    occ, use, col = _recount(o)
    np.testing.assert_array_equal(occ, o.occ)
    np.testing.assert_array_equal(use, o.node_use)
    np.testing.assert_array_equal(col, o.via_col)
    spec = nl.spec
    for k, p in enumerate(o.paths()):
        assert p[0] == nl.src[k] and p[-1] == nl.dst[k]
        assert len(set(p.tolist())) == len(p)  # simple path
        x, y, _ = spec.coords(p)
        x0, y0, w, h = nl.box[k]
        assert (x >= x0).all() and (x < x0 + w).all() and (y >= y0).all() and (y < y0 + h).all()


def test_deterministic_across_runs():
    nl = plain_64()
    runs = []
    for _ in range(2):
        o = OraclePathFinder(nl, RouterConfig(iterations=3))
        for _ in range(3):
            o.run_iteration()
        runs.append(o)
    for a, b in zip(runs[0].paths(), runs[1].paths()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("make", [plain_64, lambda: uneven(120, 100, 8, 250, seed=2)])
def test_astar_and_dijkstra_give_identical_results(make):
    nl = make()
    res = []
    for astar in (False, True):
        o = OraclePathFinder(nl, RouterConfig(iterations=4, astar=astar))
        hist = [o.run_iteration() for _ in range(4)]
        res.append((o, hist))
    (a, ha), (b, hb) = res
    for x, y in zip(a.paths(), b.paths()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.occ, b.occ)
    assert np.array_equal(a.hist.view(np.uint32), b.hist.view(np.uint32))
    assert np.array_equal(a.path_cost.view(np.uint32), b.path_cost.view(np.uint32))
    assert [h["edge_over"] for h in ha] == [h["edge_over"] for h in hb]
    assert sum(h["settled"] for h in hb) < sum(h["settled"] for h in ha)


def test_overuse_falls_on_plain_64():
    o = OraclePathFinder(plain_64(), RouterConfig(iterations=10))
    over = [(d := o.run_iteration())["edge_over"] + d["node_over"] for _ in range(10)]
    assert over[-1] < over[0] / 50


@pytest.mark.parametrize("threads", [1, 4])
def test_wave_parallel_oracle_equals_sequential(threads):
    """The all-core CPU baseline (nets of a disjoint-ROI wave on host
    threads) gives the sequential order's result, bit for bit."""
    from orthoroute_amd.netgen import uneven
    nl = uneven(90, 70, 6, 260, seed=17)
    cfg = RouterConfig(iterations=3)
    a, b = OraclePathFinder(nl, cfg), OraclePathFinder(nl, cfg)
    for _ in range(3):
        da = a.run_iteration()
        db = b.run_iteration_waves(threads)
        for k in ("routed", "failed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed",
                  "path_hops"):
            assert da[k] == db[k], k
        for x, y in zip(a.paths(), b.paths()):
            assert np.array_equal(x, y)
        assert np.array_equal(a.occ, b.occ)
        assert np.array_equal(a.hist.view(np.uint32), b.hist.view(np.uint32))


def test_hotset_reroutes_offenders_and_converges():
    """Hotset mode (offenders and unrouted nets; every net after an
    iteration that did not lower the overuse) on 64x64x6: zero overuse
    within 30 iterations, fewer reroutes than rerouting every net, and a
    net outside the hotset keeps its path untouched."""
    nl = plain_64()
    o = OraclePathFinder(nl, RouterConfig(iterations=30, hotset=True))
    routed, paths_before = [], None
    for it in range(30):
        order = o.order_for(it + 1)
        if it > 0 and 0 < len(order) < nl.n:
            paths_before = (set(range(nl.n)) - set(order.tolist()), [p.copy() for p in o.paths()])
        d = o.run_iteration(order)
        if paths_before is not None and it > 0 and 0 < len(order) < nl.n:
            keep, before = paths_before
            for k in keep:
                assert np.array_equal(o.paths()[k], before[k])
        routed.append(d["routed"])
        if d["edge_over"] + d["node_over"] == 0:
            break
    assert d["edge_over"] + d["node_over"] == 0
    assert sum(routed) < nl.n * len(routed)


def test_hotset_is_exactly_the_offender_set():
    """The oracle's hotset order, iteration by iteration, against the
    offender rule of _build_hotset (unified_pathfinder.py:9143-9171) worked
    out here from the committed paths alone: a net is rerouted when a hop
    of its path sits on an arc over capacity, a node of its path is over
    capacity, or it has no path; the order is the iteration's base order
    (net_order over per-net overuse recounted from the paths) restricted to
    those nets; after an iteration that did not lower the overuse every net
    is rerouted.  The same rule without its node term differs from the
    oracle on this board, so a dropped term fails here."""
    from orthoroute_amd.waves import net_order
    nl = plain_64()
    cfg = RouterConfig(iterations=14, hotset=True)
    o = OraclePathFinder(nl, cfg)
    nb = o.spec.neighbours()
    totals, node_term_mattered, checked = [], False, 0
    for it in range(1, 15):
        order = o.order_for(it)
        if it > 1:
            occ, use = o.occ, o.node_use
            over = np.zeros(nl.n, np.int64)
            node_over = np.zeros(nl.n, np.int64)
            hops = np.zeros(nl.n, np.int32)
            for k, p in enumerate(o.paths()):
                hops[k] = max(len(p) - 1, 0)
                node_over[k] = int(np.maximum(use[p] - cfg.node_cap, 0).sum()) if len(p) else 0
                for a, b in zip(p[:-1], p[1:]):
                    s = int(np.nonzero(nb[a] == b)[0][0])
                    over[k] += max(int(occ[4 * a + s]) - cfg.edge_cap, 0)
            base = net_order(it, o.distance, over, hops)
            stalled = len(totals) >= 2 and totals[-1] >= totals[-2]
            if stalled:
                want = base
            else:
                want = np.asarray([k for k in base if over[k] > 0 or node_over[k] > 0 or hops[k] == 0], np.int32)
                edge_only = [k for k in base if over[k] > 0 or hops[k] == 0]
                node_term_mattered |= edge_only != want.tolist()
            assert np.array_equal(order, want), it
            checked += 1
        d = o.run_iteration(order)
        totals.append(d["edge_over"] + d["node_over"])
        if totals[-1] == 0:
            break
    assert checked >= 5
    assert node_term_mattered
