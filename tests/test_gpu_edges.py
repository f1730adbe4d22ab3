"""GPU parity on the edge cases of the route path, against the oracle,
bit-exact, on every engine: a net whose source is its sink, boxes of one or
two cells, boxes clipped by every board edge and corner, two nets sharing
both pins (a node held twice), and an empty netlist."""
import numpy as np
import pytest

from oracle.oracle import OraclePathFinder
from orthoroute_amd.lattice import LatticeSpec
from orthoroute_amd.netgen import Netlist, roi_boxes
from orthoroute_amd.router import DevicePathFinder, RouterConfig

pytestmark = pytest.mark.gpu

ENGINES = {"tiles": dict(engine="tiles", small_wave_nodes=0),
           "hybrid": dict(engine="tiles"),
           "wave": dict(engine="wave")}


def _edge_board() -> Netlist:
    spec = LatticeSpec(40, 30, 6)
    n = spec.node
    pins = [((5, 5), (5, 5)),        # source is the sink
            ((0, 0), (3, 2)),        # lower-left corner
            ((39, 29), (36, 29)),    # upper-right corner, along the top edge
            ((10, 10), (10, 11)),    # one cell apart
            ((0, 15), (39, 15)),     # the full width, both side edges
            ((20, 0), (20, 29)),     # the full height, top and bottom edges
            ((7, 20), (8, 20)),      # a 2 x 1 box (set below)
            ((30, 12), (30, 12)),    # source is the sink in a 1 x 1 box (set below)
            ((12, 3), (25, 8)),      # these two share both pins
            ((12, 3), (25, 8))]
    src = np.array([n(a, b, 0) for (a, b), _ in pins], np.int32)
    dst = np.array([n(a, b, 0) for _, (a, b) in pins], np.int32)
    box = roi_boxes(spec, src, dst, 4)
    box[6] = [7, 20, 2, 1]
    box[7] = [30, 12, 1, 1]
    return Netlist(spec, src, dst, box, "edge_cases_40x30x6")


@pytest.mark.parametrize("engine", list(ENGINES))
def test_edge_case_nets_equal_the_oracle(device, engine):
    nl = _edge_board()
    cfg = RouterConfig(iterations=4, **ENGINES[engine])
    dev = DevicePathFinder(nl, cfg, device=device)
    orc = OraclePathFinder(nl, cfg)
    for it in range(4):
        print(f"[edges] engine={engine} iteration {it + 1}", flush=True)
        sd = dev.run_iteration("batched")
        so = orc.run_iteration()
        for k in ("routed", "failed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed",
                  "path_hops"):
            assert getattr(sd, k) == so[k], (it, k)
        for i, (a, b) in enumerate(zip(dev.paths(), orc.paths())):
            assert np.array_equal(a, b), (it, i)
        assert np.array_equal(dev.path_costs().view(np.uint32), orc.path_cost.view(np.uint32))
        st = dev.lat.to_numpy()
        for k in ("occ", "node_use", "via_col"):
            np.testing.assert_array_equal(st[k], getattr(orc, k), err_msg=k)
    lens = [len(p) for p in dev.paths()]
    assert lens[0] == 1 and lens[7] == 1          # a single node, no hop
    assert dev.history[-1].node_over > 0 or dev.history[0].node_over > 0   # the shared pins


def test_empty_netlist_runs(device):
    spec = LatticeSpec(16, 12, 4)
    empty = np.zeros(0, np.int32)
    nl = Netlist(spec, empty, empty.copy(), np.zeros((0, 4), np.int32), "empty")
    dev = DevicePathFinder(nl, RouterConfig(iterations=2), device=device)
    for _ in range(2):
        st = dev.run_iteration("batched")
        assert st.routed == 0 and st.failed == 0 and st.overuse == 0 and st.waves == 0
    assert dev.paths() == []
