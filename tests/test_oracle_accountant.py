"""Pins the oracle's accountant (commit / rip-up, overuse, history, cost)
with the reference's own known answers, restated for the lattice form:

  * tests/test_lattice.py:255-276 -- a hop walked forward and backward
    commits the same physical edge; raw overuse (2, 2), canonical (1, 1);
  * tests/test_via_accounting.py:18-70 -- commit sets every arc of the path,
    commits stack, a rip-up reverses a commit, overuse counts the excess
    over capacity (three uses of one edge: total 2, count 1), no overuse at
    capacity;
  * tests/unit/test_edge_accountant.py -- history with decay (10 * 0.5 + 2 * 1
    = 7.0, 205-221), the history cap (min(990, 5 * 2) = 10, 225-247), raw
    present (10 - 1 = 9.0, 268-280), the cost formula base * 0.01 +
    1.5 * overuse + 1.0 * history (3.51, 311-330), overuse over three
    overused edges (3, 3, 447-466).

The reference's EMA, jitter, layer bias and canonical-dict refresh are not
part of this path's cost form and have no counterpart here.
"""
import numpy as np
import pytest

from oracle.oracle import OracleParams, apply_path, lattice_weights, overuse, update_history
from orthoroute_amd.lattice import LatticeSpec


def _state(spec):
    return (np.zeros(4 * spec.N, np.int32), np.zeros(spec.N, np.int32), np.zeros(spec.XY, np.int32))


def _params(pres=0.0, base_w=1.0, hist_w=1.0, gain=1.0, decay=1.0, cap_mult=0.0):
    return OracleParams(pres, base_w, hist_w, 0.0, 1.0, 1, 1, gain, decay, cap_mult, 1.0, 0.0, 0.0, 0.0)


def _raw_overuse(occ, cap=1):
    over = np.maximum(occ - cap, 0)
    return int(over.sum()), int((over > 0).sum())


@pytest.fixture
def spec():
    return LatticeSpec(5, 4, 3)


def test_opposite_traversals_share_one_physical_edge(spec):
    """T/test_lattice.py:255-276: commit a->b and b->a."""
    occ, use, col = _state(spec)
    a, b = int(spec.node(2, 3, 1)), int(spec.node(3, 3, 1))   # layer 1 is horizontal
    apply_path(spec, [a, b], +1, occ, use, col)
    fwd = set(np.flatnonzero(occ).tolist())
    occ2, use2, col2 = _state(spec)
    apply_path(spec, [b, a], +1, occ2, use2, col2)
    assert set(np.flatnonzero(occ2).tolist()) == fwd and len(fwd) == 2   # both arcs, same edge
    apply_path(spec, [b, a], +1, occ, use, col)
    assert _raw_overuse(occ) == (2, 2)
    o = overuse(spec, occ, use, col)
    assert (o["edge_over"], o["edge_over_count"]) == (1, 1)


def test_commit_stack_and_rip_up(spec):
    """T/test_via_accounting.py:18-41: commit sets the path's arcs (both
    directions) and nodes, commits stack, a rip-up reverses a commit."""
    occ, use, col = _state(spec)
    p1 = [int(spec.node(1, 0, 0)), int(spec.node(1, 1, 0)), int(spec.node(1, 2, 0))]   # layer 0: vertical
    p2 = [int(spec.node(1, 1, 0)), int(spec.node(1, 2, 0)), int(spec.node(1, 3, 0))]
    apply_path(spec, p1, +1, occ, use, col)
    assert int(occ.sum()) == 4 and occ.max() == 1          # 2 hops x 2 arcs
    assert use[p1].tolist() == [1, 1, 1] and int(use.sum()) == 3
    apply_path(spec, p2, +1, occ, use, col)
    assert occ.max() == 2 and int((occ == 2).sum()) == 2    # the shared hop, both arcs
    apply_path(spec, p2, -1, occ, use, col)
    apply_path(spec, p1, -1, occ, use, col)
    assert not occ.any() and not use.any() and not col.any()


def test_overuse_counts_the_excess(spec):
    """T/test_via_accounting.py:52-65: at capacity nothing; one edge used
    three times: total 2, count 1.  Node overuse likewise."""
    occ, use, col = _state(spec)
    a, b, c = (int(spec.node(x, 2, 1)) for x in (0, 1, 2))
    apply_path(spec, [a, b, c], +1, occ, use, col)
    o = overuse(spec, occ, use, col)
    assert (o["edge_over"], o["edge_over_count"], o["node_over"], o["node_over_count"]) == (0, 0, 0, 0)
    apply_path(spec, [b, c], +1, occ, use, col)
    apply_path(spec, [c, b], +1, occ, use, col)
    o = overuse(spec, occ, use, col)
    assert (o["edge_over"], o["edge_over_count"]) == (2, 1)
    assert (o["node_over"], o["node_over_count"]) == (4, 2)      # b and c used three times


def test_overuse_over_three_overused_edges(spec):
    """T/unit/test_edge_accountant.py:447-466: three edges at use 2,
    capacity 1: (3, 3)."""
    occ, use, col = _state(spec)
    hops = [(int(spec.node(0, 0, 1)), int(spec.node(1, 0, 1))),
            (int(spec.node(2, 1, 1)), int(spec.node(3, 1, 1))),
            (int(spec.node(4, 3, 0)), int(spec.node(4, 2, 0)))]
    for _ in range(2):
        for h in hops:
            apply_path(spec, list(h), +1, occ, use, col)
    o = overuse(spec, occ, use, col)
    assert (o["edge_over"], o["edge_over_count"]) == (3, 3)


def test_via_runs_count_one_barrel_each(spec):
    """A vertical run of two vias is one via-column use (UP:4500-4512) and
    two via hops."""
    occ, use, col = _state(spec)
    p = [int(spec.node(2, 1, 0)), int(spec.node(2, 1, 1)), int(spec.node(2, 1, 2)), int(spec.node(2, 2, 2))]
    apply_path(spec, p, +1, occ, use, col)
    o = overuse(spec, occ, use, col)
    assert o["via_hops"] == 2 and o["barrels"] == 1 and col[spec.node(2, 1, 0) % spec.XY] == 1


def test_history_decay_cap_and_raw_present():
    """T/unit/test_edge_accountant.py: 205-221 (10 * 0.5 + 2 * 1 = 7; edges at
    or under capacity only decay), 225-247 (min(10 * 99, 5 * 2) = 10),
    268-280 (raw present 10, capacity 1: 9)."""
    base = np.full(3, 2.0, np.float32)
    hist = np.array([10.0, 10.0, 10.0], np.float32)
    occ = np.array([2, 1, 0], np.int32)
    update_history(base, hist, occ, np.zeros(1, np.float32), np.zeros(1, np.int32), _params(gain=2.0, decay=0.5))
    assert hist.tolist() == [7.0, 5.0, 5.0]
    hist = np.zeros(1, np.float32)
    update_history(base[:1], hist, np.array([100], np.int32), np.zeros(1, np.float32), np.zeros(1, np.int32),
                   _params(gain=10.0, cap_mult=5.0))
    assert hist.tolist() == [10.0]
    hist = np.zeros(1, np.float32)
    update_history(base[:1], hist, np.array([10], np.int32), np.zeros(1, np.float32), np.zeros(1, np.int32),
                   _params())
    assert hist.tolist() == [9.0]


def test_cost_formula(spec):
    """T/unit/test_edge_accountant.py:311-330: base 1.0 * base_w 0.01 +
    pres 1.5 * overuse 1 + hist_w 1.0 * history 2.0 = 3.51; with no use and
    no history the weight is the base term alone.  Here the overuse is the
    prospective one (use + 1 - capacity, UP:1386-1422): one net already on
    the edge."""
    N = spec.N
    base = spec.base_costs()
    hist = np.zeros(4 * N, np.float32)
    occ = np.zeros(4 * N, np.int32)
    u = int(spec.node(1, 1, 1))
    s = 2                                                     # planar +x arc of a horizontal layer
    hist[4 * u + s] = 2.0
    occ[4 * u + s] = 1
    w = lattice_weights(spec, base, hist, occ, np.zeros(N, np.float32), np.zeros(N, np.int32),
                        _params(pres=1.5, base_w=0.01))
    assert abs(float(w[4 * u + s]) - 3.51) < 1e-6
    v = int(spec.node(3, 2, 1))
    assert w[4 * v + s] == np.float32(np.float32(1.0) * np.float32(0.01))
    assert np.isinf(w[4 * int(spec.node(0, 0, 0)) + 0])       # no via below layer 0
