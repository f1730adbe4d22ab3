"""Full-size lockstep checks on the GPU: batched disjoint-ROI waves against
the sequential per-net order (the plain one-net-per-launch form), iteration
by iteration, bit-exact (paths, fp32 path costs, occupancy, node use, via
columns, overuse).

* 400x400x16 backplane (two 40x40 pin arrays joined by a seeded
  permutation, 1600 long nets + 400 short ones): 2 iterations.  The board
  is congested beyond repair (1600 nets through a 40-row band), its waves
  are nearly one net each, and an iteration takes 6-20 s in either form,
  so the lockstep is kept to the first two iterations.
* 800x800x20 board, 8000 lognormal-length nets (seed 3): 10 iterations.

The kernels themselves are checked against the CPU oracle in
tests/test_gpu_parity.py (including the largest ROIs of the 800x800x20
board and a backplane subset) and at full size in tests/test_gpu_configs.py
(800x800x20 iterations 1-5, the 1000x1000x12 query); these tests pin
batching against sequential order on the device over more iterations.
"""
import numpy as np
import pytest

from orthoroute_amd.netgen import backplane_400, uneven_800
from orthoroute_amd.router import DevicePathFinder, RouterConfig

pytestmark = pytest.mark.gpu

FIELDS = ("routed", "failed", "edge_over", "edge_over_count", "node_over", "node_over_count", "via_hops",
          "barrels", "path_hops")


def _lockstep(nl, iterations, device):
    cfg = RouterConfig(iterations=iterations, delta=16.0, block=1024, lds_nodes=16384, collect_stats=False)
    seq = DevicePathFinder(nl, cfg, device=device)
    bat = DevicePathFinder(nl, cfg, device=device)
    for it in range(iterations):
        ss = seq.run_iteration("plain")
        sb = bat.run_iteration("batched")
        print(f"[full-size] {nl.name} iteration {it + 1}: overuse {sb.overuse}, plain {ss.seconds * 1e3:.0f} ms, "
              f"batched {sb.seconds * 1e3:.0f} ms ({sb.waves} waves)", flush=True)
        for f in FIELDS:
            assert getattr(ss, f) == getattr(sb, f), (it + 1, f)
        assert seq.digest() == bat.digest(), it + 1
        vs, vb = seq.lat.via_col.cpu().numpy(), bat.lat.via_col.cpu().numpy()
        assert np.array_equal(vs, vb)
    return seq, bat


def test_backplane_400_batched_equals_sequential(device):
    _lockstep(backplane_400(), 2, device)


def test_uneven_800_ten_iterations_batched_equals_sequential(device):
    _lockstep(uneven_800(), 10, device)
