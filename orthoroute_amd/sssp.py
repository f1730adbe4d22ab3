"""Single-query SSSP entry points (array-reporting form).

``sssp_box`` runs one query over an ROI box of the lattice on the current
entering weights (``LatticeState.wcost``) and returns ROI-local distance and
parent arrays -- the microbenchmark form of find_path_roi_gpu
(cuda_dijkstra.py:4951-5040) with the whole ROI settled.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .router import TILE_RESERVE, LatticeState, Workspace, _stream_handle


def sssp_box(lat: LatticeState, box, src: int, dst: int = -1, *, full: bool = True, delta: float = 1.0,
             block: int = 256, lds_nodes: int = 12288, ws: Workspace | None = None,
             stats: torch.Tensor | None = None, stream=None, grid: bool = False, tiles: bool = False,
             queue: bool = False):
    """Returns (dist[R] fp32, parent[R] int32 global ids, status) as device tensors.
    tiles=True runs the tile-step engine (the whole box to its fixpoint),
    queue=True its queue form (one persistent launch); otherwise ROIs above
    2^18 nodes (or grid=True) run as the grid-wide state machine and smaller
    ones in one workgroup."""
    _lib.init_device()
    x0, y0, w, h = (int(v) for v in box)
    R = w * h * lat.spec.L
    if ws is None:
        ws = Workspace(lat.device)
    ws.ensure(R + (TILE_RESERVE if tiles else 0))
    dist = torch.empty(R, dtype=torch.float32, device=lat.device)
    parent = torch.empty(R, dtype=torch.int32, device=lat.device)
    flags = (_lib.ORX_F_FULL if full else 0) | (_lib.ORX_F_STATS if stats is not None else 0) | \
        (_lib.ORX_F_GRID if grid else 0) | (_lib.ORX_F_TILES if tiles else 0) | \
        (_lib.ORX_F_ASYNC if tiles and queue else 0)
    cfg = _lib.OrxRouteCfg(float(delta), int(block), int(lds_nodes), flags,
                           stats.data_ptr() if stats is not None else None)
    rc = _lib.call("orx_sssp_box", C.byref(lat.struct), x0, y0, w, h, int(src), int(dst), dist.data_ptr(),
                   parent.data_ptr(), C.byref(ws.struct), C.byref(cfg), _stream_handle(stream),
                   allow=(_lib.ORX_OK, _lib.ORX_NO_PATH))
    return dist, parent, rc
