"""GPU tests of the drop-in solver seam (orthoroute_amd.solver.HipDijkstra,
generic CSR kernel) against the reference's own vectors and the oracle.

Vectors restated from bbenchoff/OrthoRoute tests/test_gpu_fallback.py:
104-117 (ROI-CSR prices the entered node: [7.0, 3.0]), 120-145 (seed pays
its penalty: [1, 3]), 148-173 (fp32 absorption chain: [3, 2, 1, 0]),
176-268 (random 6x6 grid, seed 20260727, cost equals float64 heapq,
rel 1e-5); the data is in tests/golden/gpu_fallback_vectors.json.  Each
runs on the one-workgroup search and on the whole-chip form (grid_min 0).
"""
import numpy as np
import pytest
import torch

from oracle.oracle import csr_path, sssp_csr
from orthoroute_amd.lattice import LatticeSpec, in_rank_and_transpose
from orthoroute_amd.solver import HipDijkstra, try_attach_hip
from tests.test_reference_vectors import GOLDEN, f32, random_grid_cases, reference_cost, returned_cost

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid_min", [65536, 0])
def test_roi_prices_destination_node(device, grid_min):
    v = GOLDEN["roi_csr_prices_destination_node"]
    hip = HipDijkstra(indptr=np.array(v["indptr"]), indices=np.array(v["indices"]), device=device,
                      grid_min=grid_min)
    costs, pen = f32(v["costs"]), f32(v["node_penalty"])
    roi = np.array(v["roi_nodes"], np.int32)
    for q in v["queries"]:
        assert hip.find_path_roi_gpu(q["src"], q["dst"], costs, roi, roi, node_penalty=pen) == q["path"]
        # the ROI-local search pays the entered node's penalty (2 + 5 into node 1)
        res = hip.find_path_multisource_multisink([(q["src"], 0.0)], [(q["dst"], 0.0)], costs, roi, roi,
                                                  node_penalty=pen)
        assert res[0] == q["path"]


@pytest.mark.parametrize("grid_min", [65536, 0])
def test_fullgraph_prices_source_seed_node(device, grid_min):
    v = GOLDEN["fullgraph_prices_source_seed_node"]
    hip = HipDijkstra(indptr=np.array(v["indptr"]), indices=np.array(v["indices"]), device=device,
                      grid_min=grid_min)
    path = hip.find_path_fullgraph_gpu_seeds(
        costs=torch.from_numpy(f32(v["costs"])).to(device), src_seeds=np.array(v["seeds"], np.int32),
        dst_targets=np.array(v["targets"], np.int32), src_seed_costs=f32(v["seed_costs"]),
        dst_target_costs=f32(v["target_costs"]), node_penalty=torch.from_numpy(f32(v["node_penalty"])).to(device))
    assert path == v["path"]


@pytest.mark.parametrize("grid_min", [65536, 0])
def test_fullgraph_high_cost_rounding_does_not_cycle_parents(device, grid_min):
    v = GOLDEN["fullgraph_high_cost_rounding"]
    hip = HipDijkstra(indptr=np.array(v["indptr"]), indices=np.array(v["indices"]), device=device,
                      grid_min=grid_min)
    path = hip.find_path_fullgraph_gpu_seeds(
        costs=torch.from_numpy(f32(v["costs"])).to(device), src_seeds=np.array(v["seeds"], np.int32),
        dst_targets=np.array(v["targets"], np.int32), src_seed_costs=f32(v["seed_costs"]),
        dst_target_costs=f32(v["target_costs"]))
    assert path == v["path"]


@pytest.mark.parametrize("grid_min", [65536, 0])
def test_fullgraph_matches_reference_dijkstra_on_random_costs(device, grid_min):
    indptr, indices, cases = random_grid_cases()
    hip = HipDijkstra(indptr=indptr, indices=indices, device=device, grid_min=grid_min)
    for costs, penalty, source, target in cases:
        path = hip.find_path_fullgraph_gpu_seeds(
            costs=torch.from_numpy(costs).to(device), src_seeds=np.array([source], np.int32),
            dst_targets=np.array([target], np.int32), src_seed_costs=np.zeros(1, np.float32),
            dst_target_costs=np.zeros(1, np.float32), node_penalty=torch.from_numpy(penalty).to(device))
        assert path and path[0] == source and path[-1] == target
        rel = GOLDEN["random_6x6_grid"]["rel"]
        assert returned_cost(indptr, indices, path, costs, penalty) == pytest.approx(
            reference_cost(indptr, indices, costs, penalty, source, target), rel=rel, abs=rel)


@pytest.mark.parametrize("grid_min", [65536, 0])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_roi_search_matches_oracle_on_lattice_csr(device, seed, grid_min):
    """Random costs and penalties on a lattice CSR with a random ROI: the
    path equals the oracle's (canonical tie-break), bit for bit."""
    spec = LatticeSpec(14, 12, 4)
    rng = np.random.default_rng(seed)
    indptr, indices, _, _ = spec.to_csr()
    E = len(indices)
    costs = rng.integers(2, 12, E).astype(np.float32) / np.float32(2.0)   # many equal-cost ties
    roi_nodes = np.sort(rng.choice(spec.N, size=spec.N * 3 // 4, replace=False)).astype(np.int32)
    g2r = np.full(spec.N, -1, np.int32)
    g2r[roi_nodes] = np.arange(len(roi_nodes), dtype=np.int32)
    pen = rng.integers(0, 4, len(roi_nodes)).astype(np.float32)
    rank, in_indptr, in_src = in_rank_and_transpose(spec.N, indptr, indices)
    hip = HipDijkstra(indptr=indptr, indices=indices, device=device, grid_min=grid_min)
    for _ in range(5):
        s, t = rng.choice(roi_nodes, 2, replace=False)
        got = hip.find_path_roi_gpu(int(s), int(t), costs, roi_nodes, g2r, node_penalty=pen)
        rc, best, dist, parent = sssp_csr(spec.N, indptr, indices, costs, rank, in_indptr, in_src,
                                          roi_nodes=roi_nodes, g2r=g2r, pen=pen, seeds=[int(s)],
                                          targets=[int(t)], seed_pen=False)
        want = csr_path(best, parent, roi_nodes, g2r) if rc == 0 else None
        assert got == (want if want and len(want) > 1 else None)


def test_try_attach_hip_grafts_solver(device):
    class Solver:  # a SimpleDijkstra stand-in with its two methods
        def __init__(self):
            self.cpu_calls = 0

        def find_path_roi(self, *a, **k):
            self.cpu_calls += 1
            return None

        def find_path_multisource_multisink(self, *a, **k):
            self.cpu_calls += 1
            return None

    class Graph:
        indptr = np.array([0, 1, 2, 2, 2])
        indices = np.array([3, 3])

    s = Solver()
    assert try_attach_hip(s, Graph(), device=device)
    roi = np.arange(4, dtype=np.int32)
    path = s.find_path_roi(1, 3, np.array([1.0, 2.0], np.float32), roi, roi)
    assert path == [1, 3] and s.cpu_calls == 0
    assert s.gpu_solver is s.hip_solver


def test_fullgraph_seeds_on_a_400x400x16_board_equals_oracle(device):
    """find_path_fullgraph_gpu_seeds on the whole-chip form over the CSR of
    a 400x400x16 lattice (2.56 M nodes, 9.9 M arcs; U[1,4) costs, 5% x10,
    node penalties), six priced seeds on one side of the board, six priced
    targets on the other: the path equals the oracle's, and with no targets
    left the parallel walk agrees with the one-lane walk on the same keys."""
    spec = LatticeSpec(400, 400, 16)
    indptr, indices, _, _ = spec.to_csr()
    rng = np.random.default_rng(11)
    E = len(indices)
    costs = rng.uniform(1.0, 4.0, E).astype(np.float32)
    costs[rng.random(E) < 0.05] *= np.float32(10.0)
    pen = rng.uniform(0.0, 0.5, spec.N).astype(np.float32)
    seeds = np.array([spec.node(int(x), int(y), int(z)) for x, y, z in
                      zip(rng.integers(5, 40, 6), rng.integers(0, 400, 6), rng.integers(0, 16, 6))], np.int32)
    targets = np.array([spec.node(int(x), int(y), int(z)) for x, y, z in
                        zip(rng.integers(360, 395, 6), rng.integers(0, 400, 6), rng.integers(0, 16, 6))], np.int32)
    sc = rng.uniform(0, 2, 6).astype(np.float32)
    tc = rng.uniform(0, 2, 6).astype(np.float32)
    print("[seam] 400x400x16 whole-chip search", flush=True)
    hip = HipDijkstra(indptr=indptr, indices=indices, device=device)
    got = hip.find_path_fullgraph_gpu_seeds(costs=torch.from_numpy(costs).to(device), src_seeds=seeds,
                                           dst_targets=targets, src_seed_costs=sc, dst_target_costs=tc,
                                           node_penalty=torch.from_numpy(pen).to(device))
    print("[seam] oracle", flush=True)
    rank, in_indptr, in_src = in_rank_and_transpose(spec.N, indptr, indices)
    rc, best, dist, parent = sssp_csr(spec.N, indptr, indices, costs, rank, in_indptr, in_src, pen=pen,
                                      seeds=seeds, seed_cost=sc, targets=targets, tg_cost=tc)
    assert rc == 0
    want = csr_path(best, parent)
    assert got == want and len(got) > 350
