"""Host side of the solver seam (no GPU): the graft's fallback rules and the
argument checks that run before any device work."""
import numpy as np
import pytest

from orthoroute_amd.solver import check_search_args, graft


class _CpuSolver:
    """Stands in for SimpleDijkstra: its find_path_roi consults gpu_solver
    first, as unified_pathfinder.py:2465-2480 does for large ROIs."""

    def __init__(self):
        self.gpu_solver = None
        self.gpu_calls_seen = 0
        self.cpu_calls = 0

    def find_path_roi(self, src, dst, costs, roi_nodes, global_to_roi, node_penalty=None):
        if self.gpu_solver is not None:
            self.gpu_calls_seen += 1
            self.gpu_solver.find_path_roi_gpu(src, dst, costs, roi_nodes, global_to_roi, node_penalty)
        self.cpu_calls += 1
        return [src, dst]

    def find_path_multisource_multisink(self, src_seeds, dst_targets, costs, roi_nodes, global_to_roi,
                                        node_penalty=None):
        self.cpu_calls += 1
        return [src_seeds[0][0], dst_targets[0][0]], 0, 0


class _Accel:
    def __init__(self, mode):
        self.mode = mode
        self.calls = 0

    def find_path_roi_gpu(self, *a, **k):
        self.calls += 1
        if self.mode == "raise":
            raise RuntimeError("device fault")
        return None if self.mode == "none" else [7, 8, 9]

    def find_path_multisource_multisink(self, *a, **k):
        self.calls += 1
        if self.mode == "raise":
            raise ValueError("bad shape")
        return None if self.mode == "none" else ([7, 9], 0, 0)


@pytest.mark.parametrize("mode", ["raise", "none"])
def test_graft_falls_back_to_the_cpu_method_once(mode):
    """An accelerator that raises or finds nothing hands the net to the CPU
    method; the CPU method then runs with gpu_solver unset (no second GPU
    search), and gpu_solver is back in place afterwards."""
    solver, acc = _CpuSolver(), _Accel(mode)
    graft(solver, acc)
    assert solver.gpu_solver is acc
    assert solver.find_path_roi(1, 2, None, None, None) == [1, 2]
    assert acc.calls == 1 and solver.gpu_calls_seen == 0 and solver.cpu_calls == 1
    assert solver.gpu_solver is acc
    res = solver.find_path_multisource_multisink([(3, 0.0)], [(4, 0.0)], None, None, None)
    assert res == ([3, 4], 0, 0) and acc.calls == 2 and solver.cpu_calls == 2


def test_graft_uses_the_accelerator_result():
    solver, acc = _CpuSolver(), _Accel("ok")
    graft(solver, acc)
    assert solver.find_path_roi(1, 2, None, None, None) == [7, 8, 9]
    assert solver.find_path_multisource_multisink([(3, 0.0)], [(4, 0.0)], None, None, None) == ([7, 9], 0, 0)
    assert solver.cpu_calls == 0


def _args(**over):
    N, E, R = 10, 24, 4
    a = dict(N=N, E=E, R=R, costs=np.ones(E, np.float32), roi_nodes=np.array([1, 2, 3, 4]),
             g2r=np.full(N, -1, np.int32), pen=np.zeros(R, np.float32), allowed=None, seeds=[1], seed_cost=None,
             targets=[4], tg_cost=None)
    a.update(over)
    return a


def test_search_args_accepts_consistent_shapes():
    check_search_args(**_args())
    check_search_args(**_args(R=10, roi_nodes=None, g2r=None, pen=np.zeros(10, np.float32),
                              allowed=np.zeros(1, np.uint32), seed_cost=[0.5], tg_cost=np.zeros(1)))


@pytest.mark.parametrize("over", [
    dict(costs=np.ones(23, np.float32)),                 # arc count
    dict(pen=np.zeros(3, np.float32)),                   # ROI-local penalty length
    dict(g2r=np.full(9, -1, np.int32)),                  # N-sized map
    dict(roi_nodes=np.array([1, 2, 3, 10])),             # ROI node outside the graph
    dict(roi_nodes=np.array([1, 2, 3])),                 # R mismatch
    dict(seeds=[10]),                                    # seed id out of range
    dict(targets=[-1]),                                  # target id out of range
    dict(seed_cost=[0.0, 1.0]),                          # seed costs vs seeds
    dict(tg_cost=[]),                                    # target costs vs targets
    dict(R=10, roi_nodes=None, g2r=None, pen=np.zeros(10), allowed=np.zeros(0, np.uint32)),  # bitmap words
    dict(roi_nodes=None),                                # ROI without its map
])
def test_search_args_rejects_mismatches(over):
    with pytest.raises(ValueError):
        check_search_args(**_args(**over))
