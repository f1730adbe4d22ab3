"""Host-thread build of the kernel sources (tests/cpu_build) under
AddressSanitizer: the route kernels -- one launch per net, and the
dependency-driven persistent launch with sequential and concurrent
workgroups -- run on the CPU and must match the oracle bit for bit without a
single out-of-bounds access."""
import shutil

import numpy as np
import pytest

from oracle.oracle import OraclePathFinder
from orthoroute_amd.netgen import uneven
from orthoroute_amd.router import RouterConfig

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def driver():
    from tests.cpu_build.harness import build_driver
    return build_driver()


@pytest.mark.parametrize("mode,groups,conc", [(1, 0, 1), (0, 3, 3)])
def test_route_kernels_on_host_threads(driver, tmp_path, monkeypatch, mode, groups, conc):
    from tests.cpu_build.harness import CpuBuildPathFinder
    monkeypatch.setenv("HIPCPU_CONCURRENT_WG", str(conc))
    nl = uneven(24, 20, 3, 10, seed=4)
    cfg = RouterConfig(iterations=2)
    c = CpuBuildPathFinder(nl, cfg, tmp_path, mode=mode, max_groups=groups)
    o = OraclePathFinder(nl, cfg)
    for _ in range(2):
        d = c.run_iteration(timeout=300)
        od = o.run_iteration()
        for k in ("routed", "edge_over", "node_over", "via_hops", "barrels", "settled", "relaxed"):
            assert d[k] == od[k], k
        for a, b in zip(c.paths(), o.paths()):
            assert np.array_equal(a, b)
        assert np.array_equal(c.occ, o.occ)
        assert np.array_equal(c.hist.view(np.uint32), o.hist.view(np.uint32))
