"""Pins the oracle's multi-terminal route: find_path_multisource_multisink
(unified_pathfinder.py:2550-2659) run on a net's box inside a PathFinder
iteration.  Seeds start at fl(seed_cost + q_seed), targets are sinks, and
the best target minimises (fl(dist + target_cost), node id).

Each check is chosen so that a plausible mistake fails it:
  * scipy.sparse.csgraph.dijkstra in float64 over a super-source, on dyadic
    costs (exact): a dropped seed cost or seed node term, a target cost
    added twice or not at all, targets that expand;
  * a corridor whose only way to the cheap target crosses a dear one
    (targets must not expand);
  * the reduction to the two-pin route when q = 0 and one seed and one
    target cost 0 (equal paths, costs, occupancy and search counts);
  * a pure-Python canonical search (independent code for keys, parents
    and the target choice) on tie-heavy tiny boxes;
  * a best target that is itself a seed gives no path (UP:2647-2648).
"""
import heapq

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

from oracle.oracle import OraclePathFinder
from orthoroute_amd.lattice import LatticeSpec
from orthoroute_amd.netgen import Netlist, Terminals, uneven
from orthoroute_amd.router import RouterConfig


def _net(spec, box, seeds, scost, targets, tcost):
    T = Terminals(np.asarray([0, len(seeds)], np.int32), np.asarray(seeds, np.int32),
                  np.asarray(scost, np.float32), np.asarray([0, len(targets)], np.int32),
                  np.asarray(targets, np.int32), np.asarray(tcost, np.float32))
    return Netlist(spec, np.asarray([seeds[0]], np.int32), np.asarray([targets[0]], np.int32),
                   np.asarray([box], np.int32), "one", {}, T)


def _box_nodes(spec, box):
    x0, y0, w, h = box
    return [int(spec.node(x0 + lx, y0 + ly, z)) for z in range(spec.L) for ly in range(h) for lx in range(w)]


def _check_path(spec, path, seeds, targets, box):
    nb = spec.neighbours()
    inside = set(_box_nodes(spec, box))
    assert int(path[0]) in set(seeds)
    for u, v in zip(path[:-1], path[1:]):
        assert int(v) in nb[int(u)], "not a lattice move"
        assert int(u) in inside and int(v) in inside
    assert not (set(map(int, path[:-1])) & set(targets)), "a path ran through a target"


@pytest.mark.parametrize("trial", range(12))
def test_total_matches_float64_dijkstra_over_a_super_source(trial):
    rng = np.random.default_rng(1000 + trial)
    spec = LatticeSpec(9, 7, 4)
    base = spec.base_costs() * rng.choice(np.float32([1.0, 1.25, 1.5, 2.0]), 4 * spec.N)
    base = base.astype(np.float32)
    box = (1, 0, 7, 7)
    nodes = _box_nodes(spec, box)
    picks = rng.choice(nodes, 7, replace=False)
    seeds, targets = [int(v) for v in picks[:3]], [int(v) for v in picks[3:]]
    scost = rng.choice([0.0, 0.5, 1.25, 2.0], 3)
    tcost = rng.choice([0.0, 0.25, 0.75, 2.0], 4)
    nl = _net(spec, box, seeds, scost, targets, tcost)
    cfg = RouterConfig(iterations=1, astar=bool(trial % 2))
    o = OraclePathFinder(nl, cfg, base=base)
    o.node_hist[:] = rng.integers(0, 9, spec.N).astype(np.float32) / np.float32(4.0)
    st = o.run_iteration()

    # float64 Dijkstra: super-source -> seed at cost + q(seed); targets have no out-arcs
    q = o.node_hist.astype(np.float64)   # use = 0 on a fresh board: q = node_hist_w * node_hist
    idx = {v: i for i, v in enumerate(nodes)}
    nb = spec.neighbours()
    rows, cols, vals = [], [], []
    tset = set(targets)
    for u in nodes:
        if u in tset:
            continue
        for s in range(4):
            v = int(nb[u, s])
            if v < 0 or v not in idx or not np.isfinite(base[4 * u + s]):
                continue
            rows.append(idx[u]); cols.append(idx[v]); vals.append(float(base[4 * u + s]) + q[v])
    S = len(nodes)
    for sd, c in zip(seeds, scost):
        rows.append(S); cols.append(idx[sd]); vals.append(float(c) + q[sd])
    g = sp.csr_matrix((vals, (rows, cols)), shape=(S + 1, S + 1))
    d = dijkstra(g, indices=S)
    tot = [(d[idx[t]] + float(c), t) for t, c in zip(targets, tcost) if np.isfinite(d[idx[t]])]
    if not tot:
        assert st["failed"] == 1
        return
    best_total, best = min(tot)
    assert st["routed"] == 1
    path = o.paths()[0]
    assert int(path[-1]) == best
    assert float(o.path_cost[0]) == best_total
    _check_path(spec, path, seeds, targets, box)


def test_targets_are_sinks():
    """One corridor (5x1 cells, layer 1 horizontal; layer 0 runs along y and
    has no moves): the 0-cost target at x=4 lies behind the 100-cost target
    at x=2, so the route must end at x=2 with total 2 + 100."""
    spec = LatticeSpec(5, 1, 2)
    s, a, b = spec.node(0, 0, 1), spec.node(2, 0, 1), spec.node(4, 0, 1)
    nl = _net(spec, (0, 0, 5, 1), [s], [0.0], [a, b], [100.0, 0.0])
    o = OraclePathFinder(nl, RouterConfig(iterations=1))
    st = o.run_iteration()
    assert st["routed"] == 1
    assert [int(v) for v in o.paths()[0]] == [s, spec.node(1, 0, 1), a]
    assert float(o.path_cost[0]) == 102.0


def test_best_target_that_is_a_seed_gives_no_path():
    spec = LatticeSpec(6, 6, 3)
    a, b, c = spec.node(1, 1, 1), spec.node(4, 4, 1), spec.node(4, 1, 1)
    nl = _net(spec, (0, 0, 6, 6), [a, b], [0.0, 0.0], [a, c], [0.0, 0.0])
    o = OraclePathFinder(nl, RouterConfig(iterations=1))
    st = o.run_iteration()
    assert st["failed"] == 1 and st["routed"] == 0
    assert o.path_len[0] == 0
    assert int(o.occ.sum()) == 0 and int(o.node_use.sum()) == 0


def test_single_seed_and_target_at_cost_zero_equal_the_two_pin_route():
    nl2 = uneven(24, 20, 3, 12, seed=7)
    n = nl2.n
    T = Terminals(np.arange(n + 1, dtype=np.int32), nl2.src.astype(np.int32), np.zeros(n, np.float32),
                  np.arange(n + 1, dtype=np.int32), nl2.dst.astype(np.int32), np.zeros(n, np.float32))
    nlm = Netlist(nl2.spec, nl2.src, nl2.dst, nl2.box, "multi", dict(nl2.meta), T)
    cfg = RouterConfig(iterations=3, node_pres_w=0.0, node_hist_w=0.0)   # q = 0 everywhere
    a, b = OraclePathFinder(nl2, cfg), OraclePathFinder(nlm, cfg)
    for _ in range(3):
        da, db = a.run_iteration(), b.run_iteration()
        assert {k: v for k, v in da.items()} == {k: v for k, v in db.items()}
        for p, r in zip(a.paths(), b.paths()):
            assert np.array_equal(p, r)
        assert np.array_equal(a.path_cost.view(np.uint32), b.path_cost.view(np.uint32))
        assert np.array_equal(a.occ, b.occ)


def _py_multi(spec, wfun, q, box, seeds, scost, targets, tcost):
    """Canonical multi-seed search in pure Python: key (dist, hops, in-slot),
    ties of equal keys irrelevant (a key fixes the parent), targets never
    expanded; returns (best target, total, path) or None."""
    x0, y0, w, h = box
    nb = spec.neighbours()

    def inside(v):
        x, y, _ = spec.coords(np.asarray([v]))
        return x0 <= int(x[0]) < x0 + w and y0 <= int(y[0]) < y0 + h

    key, heap, done = {}, [], set()
    for sd, c in zip(seeds, scost):
        k = (np.float32(np.float32(c) + np.float32(q[sd])), 0, 255)
        if sd not in key or k < key[sd]:
            key[sd] = k
            heapq.heappush(heap, (k, sd))
    tset = set(targets)
    while heap:
        k, u = heapq.heappop(heap)
        if u in done or key[u] != k:
            continue
        done.add(u)
        if u in tset:
            continue
        for s in range(4):
            v = int(nb[u, s])
            if v < 0 or not inside(v) or v in done:
                continue
            wt = wfun(u, s, v)
            if not np.isfinite(wt):
                continue
            nk = (np.float32(k[0] + wt), k[1] + 1, 3 - s)
            if v not in key or nk < key[v]:
                key[v] = nk
                heapq.heappush(heap, (nk, v))
    best = None
    for t, c in zip(targets, tcost):
        if t not in key:
            continue
        tot = np.float32(key[t][0] + np.float32(c))
        if best is None or (tot, t) < best[:2]:
            best = (tot, t)
    if best is None:
        return None
    path, v = [best[1]], best[1]
    while key[v][2] != 255:
        v = int(nb[v, key[v][2]])
        path.append(v)
    return best[1], best[0], path[::-1]


@pytest.mark.parametrize("trial", range(8))
def test_tie_heavy_boxes_match_a_pure_python_search(trial):
    rng = np.random.default_rng(77 + trial)
    spec = LatticeSpec(8, 8, 3)
    base = spec.base_costs()
    box = (1, 1, 6, 6)
    nodes = _box_nodes(spec, box)
    picks = rng.choice(nodes, 6, replace=False)
    seeds, targets = [int(v) for v in picks[:3]], [int(v) for v in picks[3:]]
    scost = rng.choice([0.0, 1.0], 3)
    tcost = rng.choice([0.0, 1.0], 3)
    nl = _net(spec, box, seeds, scost, targets, tcost)
    o = OraclePathFinder(nl, RouterConfig(iterations=1), base=base)
    st = o.run_iteration()
    q = np.zeros(spec.N, np.float32)
    ref = _py_multi(spec, lambda u, s, v: np.float32(base[4 * u + s]), q, box, seeds, scost, targets, tcost)
    if ref is None or len(ref[2]) < 2:
        assert st["failed"] == 1
        return
    assert st["routed"] == 1
    assert [int(v) for v in o.paths()[0]] == ref[2]
    assert float(o.path_cost[0]) == float(ref[1])
