"""Pins the CPU oracle's search against what the mathematics fixes.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong slot / reverse-slot mapping, a transposed axis, a wrong
tie-break) fails it:
  * brute-force fp32 Bellman-Ford fixpoint over the box (order-free);
  * scipy.sparse.csgraph.dijkstra in float64 on dyadic costs (exact);
  * closed-form distances on an uncongested lattice;
  * the pure-Python heapq twin (independent code for keys and parents);
  * fixpoint invariants: d(v) <= fl(d(u)+w), equality + hops-1 on the parent
    arc, the parent is the lowest-id tight predecessor among those with
    minimal hops, parents acyclic, path cost equals the fp32 fold.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import dijkstra

from oracle.oracle import py_box_dijkstra, sssp_box
from orthoroute_amd.lattice import LatticeSpec


def _box_nodes(spec, box):
    x0, y0, w, h = box
    out = []
    for z in range(spec.L):
        for ly in range(h):
            for lx in range(w):
                out.append(int(spec.node(x0 + lx, y0 + ly, z)))
    return np.asarray(out, np.int64)


def _bellman_ford_fp32(spec, wcost, box, src):
    """Order-free fp32 fixpoint of (dist, hops), then the canonical parent."""
    nodes = _box_nodes(spec, box)
    inbox = {int(g): i for i, g in enumerate(nodes)}
    nb = spec.neighbours()
    R = len(nodes)
    d = np.full(R, np.inf, np.float32)
    hp = np.full(R, np.iinfo(np.int64).max, np.int64)
    d[inbox[src]] = 0.0
    hp[inbox[src]] = 0
    changed = True
    while changed:
        changed = False
        for i, u in enumerate(nodes):
            if not np.isfinite(d[i]):
                continue
            for s in range(4):
                v = int(nb[u, s])
                if v < 0 or v not in inbox:
                    continue
                w = np.float32(wcost[4 * u + s])
                if np.isinf(w):
                    continue
                nd = np.float32(d[i] + w)
                j = inbox[v]
                if (nd, hp[i] + 1) < (d[j], hp[j]):
                    d[j], hp[j] = nd, hp[i] + 1
                    changed = True
    par = np.full(R, -1, np.int64)
    for j, v in enumerate(nodes):
        if not np.isfinite(d[j]) or hp[j] == 0:
            continue
        cands = []
        for t in range(4):
            u = int(nb[v, t])
            if u < 0 or u not in inbox:
                continue
            i = inbox[u]
            w = np.float32(wcost[4 * u + (3 - t)])
            if np.isfinite(d[i]) and np.float32(d[i] + w) == d[j] and hp[i] + 1 == hp[j]:
                cands.append(u)
        par[j] = min(cands)
    return d, hp, par


def _random_weights(spec, rng, lo=1.0, hi=4.0, dyadic=False):
    if dyadic:
        w = rng.integers(4, 16, 4 * spec.N).astype(np.float32) / np.float32(4.0)
    else:
        w = rng.uniform(lo, hi, 4 * spec.N).astype(np.float32)
    w[spec.neighbours().reshape(-1) < 0] = np.inf
    return w


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_bellman_ford_fixpoint(seed):
    spec = LatticeSpec(7, 6, 3)
    rng = np.random.default_rng(seed)
    w = _random_weights(spec, rng)
    box = (1, 0, 5, 6)
    src = int(spec.node(2, 3, 0))
    d, par, hops, _, _ = sssp_box(spec, w, box, src)
    bd, bh, bp = _bellman_ford_fp32(spec, w, box, src)
    assert np.array_equal(d.view(np.uint32), bd.view(np.uint32))
    reach = np.isfinite(bd)
    np.testing.assert_array_equal(hops[reach], bh[reach])
    np.testing.assert_array_equal(par, bp)


def test_oracle_ties_pick_min_hops_then_lowest_id():
    # all weights equal: many equal-distance paths, parent must be canonical
    spec = LatticeSpec(6, 6, 4)
    w = np.ones(4 * spec.N, np.float32)
    w[spec.neighbours().reshape(-1) < 0] = np.inf
    box = (0, 0, 6, 6)
    src = int(spec.node(1, 1, 0))
    d, par, hops, _, _ = sssp_box(spec, w, box, src)
    bd, bh, bp = _bellman_ford_fp32(spec, w, box, src)
    np.testing.assert_array_equal(par, bp)


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_matches_python_twin(seed):
    spec = LatticeSpec(9, 8, 4)
    rng = np.random.default_rng(seed)
    w = _random_weights(spec, rng)
    box = (1, 2, 7, 5)
    src = int(spec.node(3, 4, 1))
    d, par, hops, _, _ = sssp_box(spec, w, box, src)
    twin = py_box_dijkstra(spec, w, box, src)
    nodes = _box_nodes(spec, box)
    for i, g in enumerate(nodes):
        if int(g) in twin:
            td, th, tp = twin[int(g)]
            assert np.float32(td) == d[i] and th == hops[i] and tp == par[i]
        else:
            assert np.isinf(d[i]) and par[i] == -1


def test_oracle_matches_scipy_on_dyadic_costs():
    spec = LatticeSpec(12, 10, 4)
    rng = np.random.default_rng(9)
    w = _random_weights(spec, rng, dyadic=True)
    box = (0, 0, 12, 10)
    src = int(spec.node(4, 7, 0))
    d, _, _, _, _ = sssp_box(spec, w, box, src)
    nb = spec.neighbours()
    rows, cols, vals = [], [], []
    for u in range(spec.N):
        for s in range(4):
            if nb[u, s] >= 0:
                rows.append(u), cols.append(nb[u, s]), vals.append(float(w[4 * u + s]))
    G = sp.csr_matrix((vals, (rows, cols)), shape=(spec.N, spec.N))
    ref = dijkstra(G, directed=True, indices=src)
    np.testing.assert_array_equal(d.astype(np.float64), ref)


def test_oracle_closed_form_uncongested():
    # planar 1, via 3: layer 0 is V (y), layer 1 H (x)
    spec = LatticeSpec(10, 10, 4)
    w = spec.base_costs()
    src = int(spec.node(2, 3, 0))
    d, _, _, _, _ = sssp_box(spec, w, (0, 0, 10, 10), src)

    def at(x, y, z):
        return d[(z * 10 + y) * 10 + x]

    assert at(2, 8, 0) == 5.0            # same column on a V layer: |dy|
    assert at(7, 3, 1) == 3.0 + 5.0      # up one via, along x on H layer
    assert at(7, 8, 0) == 5.0 + 3.0 + 5.0 + 3.0  # y on layer 0, via, x on layer 1, via down
    assert at(2, 3, 3) == 9.0            # three vias straight up


def test_oracle_fp32_fold_and_parent_invariants():
    spec = LatticeSpec(8, 8, 4)
    rng = np.random.default_rng(21)
    w = _random_weights(spec, rng, 1.0, 1000.0)
    box = (0, 0, 8, 8)
    src = int(spec.node(0, 0, 0))
    d, par, hops, _, _ = sssp_box(spec, w, box, src)
    nb = spec.neighbours()
    for v in range(spec.N):
        for t in range(4):
            u = int(nb[v, t])
            if u < 0:
                continue
            assert d[v] <= np.float32(d[u] + w[4 * u + (3 - t)])
        if par[v] >= 0:
            u = int(par[v])
            s = [s for s in range(4) if nb[u, s] == v][0]
            assert np.float32(d[u] + w[4 * u + s]) == d[v]
            assert hops[u] + 1 == hops[v]
    # walk a path from the far corner and re-fold its cost
    v = int(spec.node(7, 7, 3))
    path = [v]
    while par[v] >= 0:
        v = int(par[v])
        path.append(v)
        assert len(path) <= spec.N
    path = path[::-1]
    assert path[0] == src
    acc = np.float32(0.0)
    for a, b in zip(path, path[1:]):
        s = [s for s in range(4) if nb[a, s] == b][0]
        acc = np.float32(acc + w[4 * a + s])
    assert acc == d[int(spec.node(7, 7, 3))]


def test_oracle_early_exit_matches_full_on_sink():
    spec = LatticeSpec(10, 9, 5)
    rng = np.random.default_rng(5)
    w = _random_weights(spec, rng)
    box = (0, 0, 10, 9)
    src, dst = int(spec.node(1, 1, 0)), int(spec.node(8, 7, 0))
    dfull, pfull, _, nset_full, _ = sssp_box(spec, w, box, src, dst, full=True)
    dex, pex, _, nset, nrel = sssp_box(spec, w, box, src, dst, full=False)
    r = (0 * 9 + 7) * 10 + 8
    assert dex[r] == dfull[r]
    # settled count = nodes with key <= key(dst) (label-setting pops until the sink)
    assert nset <= nset_full and nrel > 0
