"""The allocation-free entry point (orx_sssp_box_noalloc) through the C ABI:
two calls in a row on one GPU, the second on a smaller box, each equal to
the allocating orx_sssp_box on the same query -- fp32 distance bits and
parents identical (integer / index results, tolerance zero) -- and a box
above the scratch limit refused with ORX_E_ARG."""
import ctypes as C

import numpy as np
import pytest
import torch

from orthoroute_amd import _lib
from orthoroute_amd.lattice import LatticeSpec
from orthoroute_amd.router import LatticeState, Workspace

pytestmark = pytest.mark.gpu


def _query(lat, box, src, dst, noalloc, ws=None):
    x0, y0, w, h = box
    R = w * h * lat.spec.L
    dist = torch.full((R,), -1.0, dtype=torch.float32, device=lat.device)
    parent = torch.full((R,), -7, dtype=torch.int32, device=lat.device)
    cfg = _lib.OrxRouteCfg(4.0, 256, 12288, _lib.ORX_F_FULL, None)
    stream = torch.cuda.current_stream().cuda_stream
    if noalloc:
        rc = _lib.call("orx_sssp_box_noalloc", C.byref(lat.struct), x0, y0, w, h, src, dst, dist.data_ptr(),
                       parent.data_ptr(), C.byref(cfg), stream, allow=(_lib.ORX_OK, _lib.ORX_NO_PATH))
    else:
        ws.ensure(R)
        rc = _lib.call("orx_sssp_box", C.byref(lat.struct), x0, y0, w, h, src, dst, dist.data_ptr(),
                       parent.data_ptr(), C.byref(ws.struct), C.byref(cfg), stream,
                       allow=(_lib.ORX_OK, _lib.ORX_NO_PATH))
    return rc, dist.cpu().numpy(), parent.cpu().numpy()


def test_noalloc_equals_allocating_entry_twice(device):
    spec = LatticeSpec(120, 100, 6)
    rng = np.random.default_rng(3)
    w4 = rng.uniform(1.0, 4.0, 4 * spec.N).astype(np.float32)
    w4[spec.neighbours().reshape(-1) < 0] = np.inf
    lat = LatticeState(spec, device=device)
    lat.wcost.copy_(torch.from_numpy(w4))
    ws = Workspace(device)
    # first call: a 100 x 100 box (60 000 nodes, near the 65 536 limit); then
    # a smaller one, whose ROI reuses the front of the same scratch
    for box, src, dst in (((10, 0, 100, 100), spec.node(40, 50, 0), spec.node(100, 90, 3)),
                          ((5, 5, 30, 20), spec.node(10, 10, 2), spec.node(30, 20, 0))):
        assert box[2] * box[3] * spec.L <= _lib.ORX_NOALLOC_MAX_NODES
        rc_a, d_a, p_a = _query(lat, box, int(src), int(dst), noalloc=False, ws=ws)
        rc_b, d_b, p_b = _query(lat, box, int(src), int(dst), noalloc=True)
        assert rc_a == rc_b == _lib.ORX_OK
        assert np.array_equal(d_a.view(np.uint32), d_b.view(np.uint32))
        assert np.array_equal(p_a, p_b)
        assert np.isfinite(d_b).sum() > 0


def test_noalloc_refuses_beyond_its_scratch(device):
    spec = LatticeSpec(120, 100, 6)
    lat = LatticeState(spec, device=device)
    lat.reprice(__import__("orthoroute_amd.router", fromlist=["RouterConfig"]).RouterConfig().params(0.5))
    with pytest.raises(_lib.HipError) as e:
        _query(lat, (0, 0, 120, 100), int(spec.node(1, 1, 0)), -1, noalloc=True)
    assert e.value.code == _lib.ORX_E_ARG
