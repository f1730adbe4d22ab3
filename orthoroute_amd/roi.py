"""Bidirectional BFS region of interest on the GPU.

``DeviceRoiBfs.extract_roi_bfs`` is ROIExtractor.extract_roi_bfs
(unified_pathfinder.py:2257-2384) for the CSR graph of a ``HipDijkstra``:
same arguments, same derived limits (max_depth, post-meeting halo,
max_roi_nodes and its 1.5x stop), same node order and must-keep truncation.
The ROI node list and the global_to_roi map stay in device memory, where
``HipDijkstra.find_path_roi_gpu`` reads them; the N-sized map is allocated
once and only the previous ROI's entries are reset between calls (the
reference allocates one per net, UP:2381-2382).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


class DeviceRoiBfs:
    def __init__(self, solver, blocked=None, max_roi_nodes: int = 50_000):
        """``solver``: a HipDijkstra (its device CSR is used); ``blocked``:
        optional bool per CSR arc (the graph's blocked_edges)."""
        self.solver = solver
        self.N = N = int(solver.N)
        self.max_roi_nodes = int(max_roi_nodes)
        dev = solver.device
        self.device = dev
        self.marks = torch.zeros((N + 15) // 16, dtype=torch.int32, device=dev)
        self.qcap = N
        self.queue = torch.empty(4 * N, dtype=torch.int32, device=dev)
        self.blocks = torch.empty(max(1, (N + 4095) // 4096), dtype=torch.int32, device=dev)
        self.cap = N + _lib.MAX_TARGETS + 2
        self.roi = torch.empty(self.cap, dtype=torch.int32, device=dev)
        self.g2r = torch.full((N,), -1, dtype=torch.int32, device=dev)
        self.out = torch.zeros(8, dtype=torch.int64, device=dev)
        self.keep = torch.empty(_lib.MAX_TARGETS + 2, dtype=torch.int32, device=dev)
        self._out_h = torch.zeros(8, dtype=torch.int64, pin_memory=True)
        self._keep_h = torch.zeros(_lib.MAX_TARGETS + 2, dtype=torch.int32, pin_memory=True)
        self.blocked = None
        if blocked is not None:
            b = np.asarray(blocked, bool)
            if b.shape != (solver.E,):
                raise ValueError("blocked must hold one flag per CSR arc")
            self.blocked = torch.from_numpy(b.astype(np.uint8)).to(dev)
        self._live = False
        self.last = {}

    def extract_roi_bfs(self, src: int, dst: int, initial_radius: int = 40, stagnation_bonus: float = 0.0,
                        portal_seeds: list | None = None):
        """(roi_nodes, global_to_roi) as device int32 tensors; roi_nodes is a
        view valid until the next call."""
        src, dst = int(src), int(dst)
        if not (0 <= src < self.N and 0 <= dst < self.N):
            raise ValueError("src / dst outside the graph")
        keep = [src, dst]
        for node, _ in (portal_seeds or []):
            node = int(node)
            if not 0 <= node < self.N:
                raise ValueError("portal seed outside the graph")
            if node not in keep:
                keep.append(node)
        if len(keep) > len(self._keep_h):
            raise ValueError(f"at most {_lib.MAX_TARGETS} portal seeds")
        max_depth = min(int(initial_radius + stagnation_bonus * 2.0), 80)
        halo = min(16, 2 + int(max(0.0, stagnation_bonus) / 0.6))
        stream = torch.cuda.current_stream(self.device)
        s = stream.cuda_stream
        with torch.cuda.stream(stream):
            if self._live:
                _lib.call("orx_roi_release", self.roi.data_ptr(), self.out.data_ptr(), self.cap,
                          self.g2r.data_ptr(), s)
                self._live = False
            self._keep_h[:len(keep)] = torch.tensor(keep, dtype=torch.int32)
            self.keep[:len(keep)].copy_(self._keep_h[:len(keep)], non_blocking=True)
            _lib.call("orx_roi_bfs", C.byref(self.solver.graph),
                      None if self.blocked is None else self.blocked.data_ptr(), src, dst, max_depth, halo,
                      self.max_roi_nodes, self.keep.data_ptr(), len(keep), self.marks.data_ptr(),
                      self.queue.data_ptr(), self.qcap, self.blocks.data_ptr(), self.roi.data_ptr(), self.cap,
                      self.g2r.data_ptr(), self.out.data_ptr(), s)
            self._live = True
            self._out_h.copy_(self.out, non_blocking=True)
        stream.synchronize()
        o = self._out_h.tolist()
        if o[4]:
            raise _lib.HipError("orx_roi_bfs", _lib.ORX_E_OVERFLOW, f"ROI buffers overflowed ({o[4]})")
        if o[3] != o[5]:
            raise _lib.HipError("orx_roi_bfs", _lib.ORX_E_ARG, f"seen {o[3]} != compacted {o[5]}")
        self.last = {"depth": o[1], "meeting_depth": None if o[2] < 0 else o[2], "seen": o[3]}
        return self.roi[:o[0]], self.g2r
