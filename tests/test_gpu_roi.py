"""GPU: the device BFS region of interest (extract_roi_bfs,
unified_pathfinder.py:2257-2384) against the oracle, exact: ROI node list
(order included), global_to_roi, depth, meeting level -- over balls,
blocked arcs, portal seeds, truncation and src == dst, with buffers reused
across calls; and the device ROI handed straight to find_path_roi_gpu."""
import numpy as np
import pytest
import torch

from oracle.oracle import roi_bfs
from orthoroute_amd.lattice import LatticeSpec
from orthoroute_amd.roi import DeviceRoiBfs
from orthoroute_amd.solver import HipDijkstra

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lattice_solver(device):
    spec = LatticeSpec(120, 100, 8)
    ip, ix, _, _ = spec.to_csr()
    return spec, ip, ix, HipDijkstra(indptr=ip, indices=ix, device=device)


def test_device_roi_equals_the_oracle(lattice_solver):
    spec, ip, ix, solver = lattice_solver
    N = spec.N
    rng = np.random.default_rng(17)
    blocked = rng.random(len(ix)) < 0.1
    plain, with_blocked = DeviceRoiBfs(solver), DeviceRoiBfs(solver, blocked=blocked)
    cases = []
    for i in range(24):
        kw = {"initial_radius": int(rng.choice([4, 10, 40, 80])), "stagnation_bonus": float(rng.choice([0.0, 0.7, 3.1]))}
        if i % 3 == 0:
            kw["portal_seeds"] = [(int(v), 0.5) for v in rng.choice(N, 3, replace=False)]
        if i % 4 == 1:
            kw["max_nodes"] = int(rng.choice([50, 400, 3000]))
        cases.append((i % 5 == 2, kw))
    cases += [(False, {"initial_radius": 40}), (True, {"initial_radius": 3, "stagnation_bonus": 3.1})]
    for i, (use_blocked, kw) in enumerate(cases):
        src, dst = (int(v) for v in rng.choice(N, 2, replace=False))
        if i >= 24:
            dst = src                  # the last two: src == dst
        print(f"[roi] case {i} src {src} dst {dst} {kw} blocked={use_blocked}", flush=True)
        r = with_blocked if use_blocked else plain
        max_nodes = kw.pop("max_nodes", 50_000)
        r.max_roi_nodes = max_nodes
        roi, g2r = r.extract_roi_bfs(src, dst, **kw)
        oroi, og2r, info = roi_bfs(ip, ix, N, src, dst, max_nodes=max_nodes,
                                   blocked=blocked if use_blocked else None, **kw)
        assert np.array_equal(roi.cpu().numpy(), oroi), i
        assert np.array_equal(g2r.cpu().numpy(), og2r), i
        assert r.last["depth"] == info["depth"] and r.last["meeting_depth"] == info["meeting_depth"], i
        assert r.last["seen"] == info["seen"], i


def test_device_roi_feeds_find_path_roi_gpu(lattice_solver):
    spec, ip, ix, solver = lattice_solver
    N = spec.N
    costs = spec.random_edge_costs(3)[spec.to_csr()[2]] if hasattr(spec, "random_edge_costs") else np.ones(len(ix))
    costs = np.asarray(costs, np.float32)
    r = DeviceRoiBfs(solver)
    rng = np.random.default_rng(4)
    for _ in range(6):
        src, dst = (int(v) for v in rng.choice(N, 2, replace=False))
        roi, g2r = r.extract_roi_bfs(src, dst, initial_radius=60)
        a = solver.find_path_roi_gpu(src, dst, costs, roi, g2r)
        b = solver.find_path_roi_gpu(src, dst, costs, roi.cpu().numpy(), g2r.cpu().numpy())
        assert a == b
        assert a is None or (a[0] == src and a[-1] == dst)
