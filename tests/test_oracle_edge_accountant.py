"""Pins oracle/accountant.py (the CPU restatement of EdgeAccountant and of the
via capacity kernels' CPU forms) with the reference's known answers and with
closed forms, on the CPU.

Known answers, tests/unit/test_edge_accountant.py: construction (67-85),
commit 1/1/1 and repeated arcs 2/1/1 (86-101), EMA 6.0 / between 5 and 10 /
smoothing / beta effect / steady 5.0 (110-171), history +1.0 on overused arcs
only (181-197), decay 10 * 0.5 + 2 * 1 = 7.0 (199-217), cap min(990, 10) = 10
(219-244), EMA-based 4.0 (246-262) and raw 9.0 (264-276), decay 98.0 and
100 * 0.98^10 (278-298), cost 0.01 + 1.5 + 2.0 = 3.51 (305-330), base only
0.01 (332-343), jitter range (345-360), via multiplier 0.02 (362-376), layer
bias 0.005 / 0.01 / 0.015 (378-412) and uniform vias (414-438), overuse (3, 3)
/ (0, 0) / count 3 (447-475).
"""
import numpy as np
import pytest

from oracle.accountant import (OracleEdgeAccountant, oracle_hard_block, oracle_via_metadata,
                               oracle_via_penalties)
from orthoroute_amd.lattice import LatticeSpec


@pytest.fixture
def acc():
    return OracleEdgeAccountant(10)


@pytest.fixture
def used():
    a = OracleEdgeAccountant(10)
    a.present[:3] = 2.0
    a.present[3:6] = 1.0
    a.present[6:] = 0.5
    a.present_ema = a.present.copy()
    return a


def test_construction(acc):
    assert acc.E == 10 and len(acc.present) == 10 and len(acc.capacity) == 10
    assert not acc.present.any() and not acc.present_ema.any() and not acc.history.any()
    assert (acc.capacity == 1.0).all()


def test_commit(acc):
    acc.commit_path([0, 1, 2])
    assert list(acc.present[:4]) == [1.0, 1.0, 1.0, 0.0]
    b = OracleEdgeAccountant(10)
    b.commit_path([0, 1, 0, 2])
    assert list(b.present[:3]) == [2.0, 1.0, 1.0]
    assert b.canonical == {0: 2, 1: 1, 2: 1}


def test_clear_floors_and_drops_dict_entries(acc):
    acc.commit_path([4, 4, 5])
    acc.clear_path([4, 5, 5, 7])
    assert acc.canonical == {4: 1}
    assert list(acc.present[4:8]) == [1.0, 0.0, 0.0, 0.0]
    acc.present[9] = -3.0           # the floor covers the whole array
    acc.clear_path([4])
    assert acc.present[9] == 0.0 and acc.canonical == {}
    assert acc.verify_present_matches_canonical()


def test_batched_clears_equal_sequential_clears():
    rng = np.random.default_rng(5)
    paths = [rng.integers(0, 40, size=rng.integers(1, 12)) for _ in range(30)]
    a, b = OracleEdgeAccountant(40), OracleEdgeAccountant(40)
    for p in paths:
        a.commit_path(p)
        b.commit_path(p)
    drop = paths[::2] + [rng.integers(0, 40, size=9)]
    for p in drop:
        a.clear_path(p)
    b.clear_path(np.concatenate(drop))
    assert np.array_equal(a.present, b.present) and a.canonical == b.canonical


def test_present_ema(acc):
    acc.present[0] = 10.0
    acc.update_present_ema(beta=0.6)
    assert np.isclose(acc.present_ema[0], 6.0)
    e1 = acc.present_ema[0]
    acc.present[0] = 0.0
    acc.update_present_ema(beta=0.6)
    assert 0.0 < acc.present_ema[0] < e1
    s = OracleEdgeAccountant(1)
    s.present[0] = s.present_ema[0] = 5.0
    s.update_present_ema(beta=0.6)
    assert np.isclose(s.present_ema[0], 5.0)
    hi, lo = OracleEdgeAccountant(1), OracleEdgeAccountant(1)
    hi.present[0] = lo.present[0] = 10.0
    hi.update_present_ema(beta=0.9)
    lo.update_present_ema(beta=0.1)
    assert hi.present_ema[0] > lo.present_ema[0]


def test_history_known_answers(used):
    used.update_history(gain=1.0, decay_factor=1.0)
    assert used.history[0] == 1.0 and used.history[3] == 0.0 and used.history[6] == 0.0
    a = OracleEdgeAccountant(10)
    a.present[:3] = 2.0
    a.present_ema = a.present.copy()
    a.history[0] = 10.0
    a.update_history(gain=2.0, decay_factor=0.5)
    assert np.isclose(a.history[0], 7.0, atol=0.01)


def test_history_cap(used):
    used.present_ema[0] = 100.0
    used.update_history(gain=10.0, base_costs=np.full(10, 2.0, np.float32), history_cap_multiplier=5.0,
                        decay_factor=1.0)
    assert used.history[0] == 10.0


def test_history_ema_and_raw(used):
    used.present[0], used.present_ema[0] = 10.0, 5.0
    used.update_history(gain=1.0, use_raw_present=False, decay_factor=1.0)
    assert np.isclose(used.history[0], 4.0)
    b = OracleEdgeAccountant(10)
    b.present[0], b.present_ema[0] = 10.0, 5.0
    b.update_history(gain=1.0, use_raw_present=True, decay_factor=1.0)
    assert np.isclose(b.history[0], 9.0)


def test_history_decay(acc):
    acc.history[0] = 100.0
    acc.update_history(gain=1.0, decay_factor=0.98)
    assert np.isclose(acc.history[0], 98.0)
    for _ in range(9):
        acc.update_history(gain=1.0, decay_factor=0.98)
    assert np.isclose(acc.history[0], 100.0 * 0.98 ** 10, atol=0.5)


def test_cost_known_answers(used):
    base = np.ones(10, np.float32)
    used.history[0] = 2.0
    used.update_costs(base, pres_fac=1.5, hist_weight=1.0, add_jitter=False)
    assert np.isclose(used.total_cost[0], 0.01 + 1.5 + 2.0, atol=1e-6)
    a = OracleEdgeAccountant(10)
    a.update_costs(base, pres_fac=1.0, hist_weight=1.0, add_jitter=False)
    assert np.allclose(a.total_cost, 0.01, atol=1e-7)
    a.update_costs(base, pres_fac=0.0, hist_weight=0.0, via_cost_multiplier=2.0, add_jitter=False)
    assert np.allclose(a.total_cost, 0.02, atol=1e-7)


def test_cost_jitter_closed_form():
    n = 20000
    a = OracleEdgeAccountant(n)
    a.update_costs(np.zeros(n, np.float32), pres_fac=1.0, hist_weight=0.0)
    want = (np.float32(np.arange(n) % 9973) * np.float32(1e-6)).astype(np.float32)
    assert np.array_equal(a.total_cost, want)
    assert 0 < a.total_cost[:10].max() - a.total_cost[:10].min() < 1e-3
    assert a.total_cost[9973] == 0.0 and a.total_cost[9974] == np.float32(1e-6)


def test_cost_layer_bias():
    a = OracleEdgeAccountant(10)
    base = np.ones(10, np.float32)
    layer = np.array([0, 0, 1, 1, 2, 2, 3, 3, 4, 4], np.int32)
    bias = np.array([0.5, 1.0, 1.5, 1.0, 1.0], np.float32)
    a.update_costs(base, pres_fac=0.0, hist_weight=0.0, add_jitter=False, edge_layer=layer,
                   layer_bias_per_layer=bias, edge_kind=np.zeros(10, np.int32))
    assert np.isclose(a.total_cost[0], 0.005) and np.isclose(a.total_cost[2], 0.01)
    assert np.isclose(a.total_cost[4], 0.015)
    a.update_costs(base, pres_fac=0.0, hist_weight=0.0, add_jitter=False, edge_layer=layer,
                   layer_bias_per_layer=bias, edge_kind=np.array([0, 1] * 5, np.int32))
    assert np.allclose(a.total_cost[1::2], a.total_cost[1])


def test_layer_bias_scales_present_term_and_live_costs():
    """pres_fac * bias scales the present term and is the live scale."""
    a = OracleEdgeAccountant(4)
    a.present_ema[:] = [3.0, 3.0, 0.0, 0.0]
    a.present[:] = [1.0, 0.0, 2.0, 0.0]
    a.update_costs(np.zeros(4, np.float32), pres_fac=2.0, hist_weight=0.0, add_jitter=False,
                   edge_layer=np.array([0, 1, 0, 1]), layer_bias_per_layer=np.array([0.5, 2.0], np.float32),
                   edge_kind=np.array([0, 0, 1, 0]))
    assert list(a.total_cost) == [2.0, 8.0, 0.0, 0.0]        # (2 * b) * (3 - 1)
    a.begin_live_present_costs()
    # static part 0; live = s * max(0, present + 1 - 1) = s * present
    assert list(a.total_cost) == [1.0, 0.0, 4.0, 0.0]
    a.present[1] = 3.0
    a.refresh_live_present_costs([1])
    assert a.total_cost[1] == 12.0


def test_overuse_known_answers(used, acc):
    assert used.compute_overuse() == (3, 3)
    assert acc.compute_overuse() == (0, 0)
    mask = np.zeros(10, bool)
    mask[1:] = True
    assert used.compute_overuse(mask) == (2, 2)


def test_cost_balance_ratio(used):
    used.history[:] = 0.5
    assert np.isclose(used.cost_balance_ratio(2.0, 1.0), (2.0 * 5.0) / 3.0)


def test_refresh_from_canonical(acc):
    acc.commit_path([1, 1, 3])
    acc.present[:] = 7.0
    assert not acc.verify_present_matches_canonical()
    acc.refresh_from_canonical()
    assert list(acc.present[:4]) == [0.0, 2.0, 0.0, 1.0] and acc.verify_present_matches_canonical()


def _via_case():
    spec = LatticeSpec(4, 3, 5)
    indptr, indices, arc_ids, _ = spec.to_csr()
    kind = (arc_ids % 4 == 0) | (arc_ids % 4 == 3)      # slots 0 and 3 are vias
    return spec, indptr, indices, kind


def test_via_metadata_matches_lattice():
    spec, indptr, indices, kind = _via_case()
    m = oracle_via_metadata(indptr, indices, kind, spec.X, spec.Y, spec.L)
    src = np.repeat(np.arange(spec.N), np.diff(indptr))
    for i, e in enumerate(m["indices"]):
        u, v = int(src[e]), int(indices[e])
        (xu, yu, zu), (xv, yv, zv) = ([int(c[0]) for c in spec.coords(np.array([w]))] for w in (u, v))
        assert (xu, yu) == (xv, yv) == tuple(m["xy_coords"][i])
        assert m["z_lo"][i] == min(max(min(zu, zv), 1), spec.L - 2)
        assert m["z_hi"][i] == min(max(max(zu, zv), 1), spec.L - 2)
    assert len(m["indices"]) == 2 * spec.XY * (spec.L - 1)


def test_via_kernels_hand_computed():
    meta = {"indices": np.array([5, 2, 7]), "xy_coords": np.array([[0, 1], [1, 0], [1, 1]]),
            "z_lo": np.array([1, 1, 2]), "z_hi": np.array([3, 2, 2])}
    col_use = np.array([[0, 2], [0, 0]], np.int16)
    col_cap = np.array([[1, 2], [1, 1]], np.int16)
    seg_use = np.zeros((2, 2, 3), np.int8)
    seg_cap = np.ones((2, 2, 3), np.int8)
    seg_use[1, 0, 0] = 1                                   # column (1, 0), segment 0
    cost = np.zeros(8, np.float32)
    # arc 5: column (0, 1) full; arc 2: segment z=1 -> 0 full; arc 7: span empty
    assert oracle_hard_block(meta, col_use, col_cap, seg_use, seg_cap, cost, 3) == 2
    assert np.isinf(cost[5]) and np.isinf(cost[2]) and cost[7] == 0.0
    cost = np.zeros(8, np.float32)
    assert oracle_hard_block(meta, col_use, col_cap, None, None, cost, 3) == 1
    col_pres = np.array([[0.0, 1.0], [0.5, 0.0]], np.float32)
    seg_pres = np.zeros((2, 2, 3), np.float32)
    seg_pres[0, 1, :] = [2.0, 4.0, 8.0]
    cost = np.ones(8, np.float32)
    n = oracle_via_penalties(meta, col_pres, seg_pres, 3.0, 0.25, cost, 3)
    # arc 5: 1*3 + (2 + 4) * 0.25 = 4.5; arc 2: 0.5*3 = 1.5; arc 7: 0 (skipped)
    assert n == 2 and cost[5] == 5.5 and cost[2] == 2.5 and cost[7] == 1.0
