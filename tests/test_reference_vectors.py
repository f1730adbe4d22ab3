"""The reference's own known-answer vectors for this path, restated against
the CPU oracle (and, in test_gpu_seam.py, against the HIP path).  The
vectors are data in tests/golden/gpu_fallback_vectors.json.

Sources (bbenchoff/OrthoRoute):
  tests/test_gpu_fallback.py:104-117  ROI-CSR folds the entered node's penalty: [7.0, 3.0]
  tests/test_gpu_fallback.py:120-145  seed pays its node penalty: path [1, 3]
  tests/test_gpu_fallback.py:148-173  fp32 absorption chain, seed cost 16e6: path [3, 2, 1, 0]
  tests/test_gpu_fallback.py:176-268  random 6x6 grid, seed 20260727, 8 trials,
                                       path cost == float64 heapq cost (rel 1e-5)
"""
import heapq
import json
from pathlib import Path

import numpy as np
import pytest

from oracle.oracle import csr_path, sssp_csr
from orthoroute_amd.lattice import in_rank_and_transpose

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "gpu_fallback_vectors.json").read_text())


def f32(v):
    return np.asarray(v, np.float32)


def _graph(indptr, indices):
    N = len(indptr) - 1
    rank, in_indptr, in_src = in_rank_and_transpose(N, indptr, indices)
    return N, np.asarray(indptr, np.int32), np.asarray(indices, np.int32), rank, in_indptr, in_src


def test_roi_csr_prices_destination_node():
    # 2 nodes, arcs 0->1 (cost 2) and 1->0 (cost 3); penalty of node 1 = 5.
    # The folded ROI weights are [2+5, 3+0] = [7, 3]; the path 0 -> 1 costs 7.
    v = GOLDEN["roi_csr_prices_destination_node"]
    N, ip, ix, rk, iip, isrc = _graph(v["indptr"], v["indices"])
    roi = np.asarray(v["roi_nodes"])
    for q in v["queries"]:
        rc, best, dist, parent = sssp_csr(N, ip, ix, f32(v["costs"]), rk, iip, isrc, roi_nodes=roi, g2r=roi,
                                          pen=f32(v["node_penalty"]), seeds=[q["src"]], targets=[q["dst"]],
                                          seed_pen=False)
        assert rc == 0 and best == q["dst"]
        assert dist[q["dst"]] == np.float32(q["dist"])
        assert csr_path(best, parent) == q["path"]


def test_fullgraph_prices_source_seed_node():
    v = GOLDEN["fullgraph_prices_source_seed_node"]
    N, ip, ix, rk, iip, isrc = _graph(v["indptr"], v["indices"])
    rc, best, dist, parent = sssp_csr(
        N, ip, ix, f32(v["costs"]), rk, iip, isrc, pen=f32(v["node_penalty"]), seeds=v["seeds"],
        seed_cost=f32(v["seed_costs"]), targets=v["targets"], tg_cost=f32(v["target_costs"]))
    assert rc == 0
    assert csr_path(best, parent) == v["path"]


def test_fullgraph_high_cost_rounding_does_not_cycle_parents():
    v = GOLDEN["fullgraph_high_cost_rounding"]
    N, ip, ix, rk, iip, isrc = _graph(v["indptr"], v["indices"])
    rc, best, dist, parent = sssp_csr(
        N, ip, ix, f32(v["costs"]), rk, iip, isrc, seeds=v["seeds"], seed_cost=f32(v["seed_costs"]),
        targets=v["targets"], tg_cost=f32(v["target_costs"]))
    assert rc == 0
    assert csr_path(best, parent) == v["path"]


def _grid6():
    g = GOLDEN["random_6x6_grid"]
    width, height = g["width"], g["height"]
    rows = [[] for _ in range(width * height)]
    for y in range(height):
        for x in range(width):
            node = y * width + x
            for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                nx, ny = x + dx, y + dy
                if 0 <= nx < width and 0 <= ny < height:
                    rows[node].append(ny * width + nx)
    indptr = np.zeros(width * height + 1, np.int32)
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
    g = GOLDEN["random_6x6_grid"]
    indptr, indices = _grid6()
    n = g["width"] * g["height"]
    rng = np.random.default_rng(g["seed"])
    cases = []
    for _ in range(g["trials"]):
        costs = rng.uniform(*g["cost_range"], len(indices)).astype(np.float32)
        penalty = rng.uniform(*g["penalty_range"], n).astype(np.float32)
        penalty[rng.random(n) < g["penalty_zero_fraction"]] = 0.0
        source, target = rng.choice(n, size=2, replace=False).astype(np.int32)
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
        rel = GOLDEN["random_6x6_grid"]["rel"]
        assert returned_cost(indptr, indices, path, costs, penalty) == pytest.approx(
            reference_cost(indptr, indices, costs, penalty, source, target), rel=rel, abs=rel)
