"""The reference's own known-answer vectors for this path, restated against
the CPU oracle (and, in test_gpu_seam.py, against the HIP path).

Sources (bbenchoff/OrthoRoute):
  tests/test_gpu_fallback.py:104-117  ROI-CSR folds the entered node's penalty: [7.0, 3.0]
  tests/test_gpu_fallback.py:120-145  seed pays its node penalty: path [1, 3]
  tests/test_gpu_fallback.py:148-173  fp32 absorption chain, seed cost 16e6: path [3, 2, 1, 0]
  tests/test_gpu_fallback.py:176-268  random 6x6 grid, seed 20260727, 8 trials,
                                       path cost == float64 heapq cost (rel 1e-5)
"""
import heapq

import numpy as np
import pytest

from oracle.oracle import csr_path, sssp_csr
from orthoroute_amd.lattice import in_rank_and_transpose


def _graph(indptr, indices):
    N = len(indptr) - 1
    rank, in_indptr, in_src = in_rank_and_transpose(N, indptr, indices)
    return N, np.asarray(indptr, np.int32), np.asarray(indices, np.int32), rank, in_indptr, in_src


def test_roi_csr_prices_destination_node():
    # 2 nodes, arcs 0->1 (cost 2) and 1->0 (cost 3); penalty of node 1 = 5.
    # The folded ROI weights are [2+5, 3+0] = [7, 3]; the path 0 -> 1 costs 7.
    N, ip, ix, rk, iip, isrc = _graph([0, 1, 2], [1, 0])
    costs = np.array([2.0, 3.0], np.float32)
    pen = np.array([0.0, 5.0], np.float32)
    rc, best, dist, parent = sssp_csr(N, ip, ix, costs, rk, iip, isrc, roi_nodes=np.array([0, 1]),
                                      g2r=np.array([0, 1]), pen=pen, seeds=[0], targets=[1], seed_pen=False)
    assert rc == 0 and best == 1
    assert dist[1] == np.float32(7.0)
    rc, best, dist, parent = sssp_csr(N, ip, ix, costs, rk, iip, isrc, roi_nodes=np.array([0, 1]),
                                      g2r=np.array([0, 1]), pen=pen, seeds=[1], targets=[0], seed_pen=False)
    assert dist[0] == np.float32(3.0)


def test_fullgraph_prices_source_seed_node():
    N, ip, ix, rk, iip, isrc = _graph([0, 1, 2, 2, 2], [3, 3])
    rc, best, dist, parent = sssp_csr(
        N, ip, ix, np.array([1.0, 2.0], np.float32), rk, iip, isrc,
        pen=np.array([100.0, 0.0, 0.0, 0.0], np.float32), seeds=[0, 1], seed_cost=np.zeros(2, np.float32),
        targets=[3], tg_cost=np.zeros(1, np.float32))
    assert rc == 0
    assert csr_path(best, parent) == [1, 3]


def test_fullgraph_high_cost_rounding_does_not_cycle_parents():
    N, ip, ix, rk, iip, isrc = _graph([0, 1, 3, 5, 6], [1, 0, 2, 1, 3, 2])
    rc, best, dist, parent = sssp_csr(
        N, ip, ix, np.full(6, 0.4, np.float32), rk, iip, isrc, seeds=[3],
        seed_cost=np.array([16_000_000.0], np.float32), targets=[0], tg_cost=np.zeros(1, np.float32))
    assert rc == 0
    assert csr_path(best, parent) == [3, 2, 1, 0]


def _grid6():
    width = height = 6
    rows = [[] for _ in range(36)]
    for y in range(height):
        for x in range(width):
            node = y * width + x
            for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                nx, ny = x + dx, y + dy
                if 0 <= nx < width and 0 <= ny < height:
                    rows[node].append(ny * width + nx)
    indptr = np.zeros(37, np.int32)
    for node, nbrs in enumerate(rows):
        indptr[node + 1] = indptr[node] + len(nbrs)
    indices = np.asarray([v for r in rows for v in r], np.int32)
    return indptr, indices


def reference_cost(indptr, indices, costs, penalty, source, target):
    """float64 heapq Dijkstra priced like the reference test (entered-node penalty)."""
    n = len(indptr) - 1
    distance = np.full(n, np.inf)
    distance[source] = float(penalty[source])
    queue = [(distance[source], source)]
    while queue:
        cur, node = heapq.heappop(queue)
        if cur != distance[node]:
            continue
        if node == target:
            return cur
        for e in range(indptr[node], indptr[node + 1]):
            nb = int(indices[e])
            cand = cur + float(costs[e]) + float(penalty[nb])
            if cand < distance[nb]:
                distance[nb] = cand
                heapq.heappush(queue, (cand, nb))
    return np.inf


def returned_cost(indptr, indices, path, costs, penalty):
    total = float(penalty[path[0]])
    for s, t in zip(path, path[1:]):
        off = np.flatnonzero(indices[indptr[s]:indptr[s + 1]] == t)
        assert off.size == 1
        total += float(costs[indptr[s] + int(off[0])]) + float(penalty[t])
    return total


def random_grid_cases():
    """The 8 trials of test_gpu_fullgraph_matches_reference_dijkstra_on_random_costs."""
    indptr, indices = _grid6()
    rng = np.random.default_rng(20260727)
    cases = []
    for _ in range(8):
        costs = rng.uniform(0.05, 8.0, len(indices)).astype(np.float32)
        penalty = rng.uniform(0.0, 3.0, 36).astype(np.float32)
        penalty[rng.random(36) < 0.7] = 0.0
        source, target = rng.choice(36, size=2, replace=False).astype(np.int32)
        cases.append((costs, penalty, int(source), int(target)))
    return indptr, indices, cases


def test_random_6x6_grid_matches_float64_dijkstra():
    indptr, indices, cases = random_grid_cases()
    N, ip, ix, rk, iip, isrc = _graph(indptr, indices)
    for costs, penalty, source, target in cases:
        rc, best, dist, parent = sssp_csr(N, ip, ix, costs, rk, iip, isrc, pen=penalty, seeds=[source],
                                          seed_cost=np.zeros(1, np.float32), targets=[target],
                                          tg_cost=np.zeros(1, np.float32))
        path = csr_path(best, parent)
        assert path and path[0] == source and path[-1] == target
        assert returned_cost(indptr, indices, path, costs, penalty) == pytest.approx(
            reference_cost(indptr, indices, costs, penalty, source, target), rel=1e-5, abs=1e-5)
