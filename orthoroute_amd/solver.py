"""Drop-in accelerator for the engine's solver seam.

``HipDijkstra`` exposes the methods the reference engine calls on its
``gpu_solver`` slot and on ``SimpleDijkstra``
(unified_pathfinder.py:2451-2659, 3099-3120, 8083-8093):

* ``find_path_roi_gpu`` / ``find_path_roi`` (cuda_dijkstra.py:4951-5040):
  source at distance 0, ROI-local node penalty paid when a node is entered;
* ``find_path_multisource_multisink`` (unified_pathfinder.py:2550-2659):
  seeds pay their own penalty, targets are sinks, returns
  (path, entry_layer, exit_layer);
* ``find_path_fullgraph_gpu_seeds`` (cuda_dijkstra.py:5516-5537): whole graph,
  seed / target costs, global node penalty, optional allowed bitmap.

All return global node ids (source first) or ``None`` when no path exists,
like the reference.  ``try_attach_hip`` grafts the accelerator onto a
SimpleDijkstra the way try_attach_metal does (metal_dijkstra.py:44-89): the
wrapped methods call the HIP solver first and hand a ``None`` result, or
any exception it raises, to the solver's own original method, so a net the
caller could route is never dropped.  Every query's shapes and node ids
are checked on the host first (``check_search_args``).  Inside this package there is no CPU search: when liborx_hip.so or
the GPU is missing the constructor raises.
"""
from __future__ import annotations

import ctypes as C
import logging
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .lattice import in_rank_and_transpose

log = logging.getLogger(__name__)


def _host(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        return a.detach().cpu().numpy()
    if hasattr(a, "get"):
        return a.get()
    return np.asarray(a)


def _in_roi(global_to_roi, nodes) -> np.ndarray:
    """Membership of ``nodes`` in an ROI map; a device map is gathered at
    those nodes only (no copy of the N-sized map)."""
    idx = np.asarray(nodes, np.int64)
    if isinstance(global_to_roi, torch.Tensor) and global_to_roi.is_cuda:
        t = torch.from_numpy(idx).to(global_to_roi.device)
        return (global_to_roi[t] >= 0).cpu().numpy()
    return _host(global_to_roi)[idx] >= 0


def _n(a) -> int:
    return int(a.numel()) if isinstance(a, torch.Tensor) else len(np.asarray(_host(a)).reshape(-1))


def check_search_args(N: int, E: int, R: int, costs, roi_nodes, g2r, pen, allowed, seeds, seed_cost, targets,
                      tg_cost) -> None:
    """Shapes and ranges of one seam query, checked on the host before any
    device work (the kernel would index seed_cost[i], tg_cost[i], pen[r],
    g2r[gid] and allowed[gid >> 5] with them).  Raises ValueError like the
    reference does for mismatched terminal costs and node penalties
    (cuda_dijkstra.py:5566-5570, 5594-5598)."""
    if costs is None or _n(costs) != E:
        raise ValueError(f"costs has {0 if costs is None else _n(costs)} entries, graph has {E} arcs")
    if (roi_nodes is None) != (g2r is None):
        raise ValueError("roi_nodes and global_to_roi go together")
    if roi_nodes is not None:
        if _n(roi_nodes) != R:
            raise ValueError(f"roi_nodes has {_n(roi_nodes)} entries, expected {R}")
        if _n(g2r) != N:
            raise ValueError(f"global_to_roi has {_n(g2r)} entries, graph has {N} nodes")
        rn = np.asarray(_host(roi_nodes)).reshape(-1)
        if rn.size and (rn.min() < 0 or rn.max() >= N):
            raise ValueError("roi_nodes holds a node id outside the graph")
    elif R != N:
        raise ValueError("a whole-graph query has R = N")
    if pen is not None and _n(pen) != R:
        raise ValueError(f"node_penalty has {_n(pen)} entries, expected {R} "
                         f"({'ROI-local' if roi_nodes is not None else 'global'})")
    if allowed is not None and _n(allowed) < (N + 31) // 32:
        raise ValueError(f"allowed_bitmap has {_n(allowed)} words, needs {(N + 31) // 32}")
    for name, ids, cost in (("seed", seeds, seed_cost), ("target", targets, tg_cost)):
        ids = np.asarray(ids, np.int64).reshape(-1)
        if ids.size and (ids.min() < 0 or ids.max() >= N):
            raise ValueError(f"a {name} node id is outside the graph")
        if cost is not None and _n(cost) != ids.size:
            raise ValueError(f"{name} costs have {_n(cost)} entries for {ids.size} {name}s")


# ROIs of at least this many nodes run the whole-chip form of the search
# (ORX_F_GRID) and the parallel parent walk; smaller ones one workgroup.
GRID_MIN_NODES = 65536


class HipDijkstra:
    """SSSP over a CSR graph on the GPU (generic form of the lattice kernels).
    Small ROIs run in one workgroup; ROIs of at least ``grid_min`` nodes (the
    full-graph seeds search on a board) run on the whole chip.  ``delta`` is
    the near-far bucket width: results do not depend on it; on the
    400x400x16 lattice CSR a full-graph query takes 29.5 / 19.0 / 13.5 /
    12.0 ms at 1 / 4 / 16 / 64."""

    def __init__(self, graph=None, lattice=None, *, indptr=None, indices=None, device="cuda",
                 delta: float = 64.0, block: int = 256, grid_min: int = GRID_MIN_NODES):
        _lib.init_device()
        if graph is not None:
            indptr = graph.indptr
            indices = graph.indices
        ip = _host(indptr).astype(np.int64)
        ix = _host(indices).astype(np.int64)
        self.N = len(ip) - 1
        self.E = len(ix)
        if self.N >= 2**24:
            raise ValueError("graphs above 2^24 nodes are not supported by the packed key")
        rank, in_indptr, in_src = in_rank_and_transpose(self.N, ip, ix)
        self.device = torch.device(device)
        dev = self.device
        self.indptr = torch.from_numpy(ip.astype(np.int32)).to(dev)
        self.indices = torch.from_numpy(ix.astype(np.int32)).to(dev)
        self.in_rank = torch.from_numpy(rank).to(dev)
        self.in_indptr = torch.from_numpy(in_indptr).to(dev)
        self.in_src = torch.from_numpy(in_src).to(dev)
        self.graph = _lib.OrxCsr(self.N, self.E, self.indptr.data_ptr(), self.indices.data_ptr(),
                                 self.in_rank.data_ptr(), self.in_indptr.data_ptr(), self.in_src.data_ptr())
        self.plane_size = (lattice.x_steps * lattice.y_steps) if lattice is not None and hasattr(lattice, "x_steps") \
            else None
        self.delta = float(delta)
        self.block = int(block)
        self.grid_min = int(grid_min)
        self._cap = 0
        self._key = self._scratch = self._path = None
        self._small = torch.zeros(4, dtype=torch.int32, device=dev)  # best, len

    # -- buffers -------------------------------------------------------------
    def _ensure(self, R: int):
        if R <= self._cap:
            return
        R = int(R)
        self._key = torch.empty(R, dtype=torch.int64, device=self.device)
        self._scratch = torch.empty(self._scratch_words(R), dtype=torch.int32, device=self.device)
        self._path = torch.empty(R, dtype=torch.int32, device=self.device)
        self._cap = R

    @staticmethod
    def _scratch_words(R: int) -> int:
        """Search scratch: 4R queue words + 4 bitmaps, then the whole-chip
        form's 64 control words at an 8-byte boundary."""
        base = 4 * R + 4 * ((R + 31) // 32)
        return (base + 1) // 2 * 2 + 64

    def _dev(self, a, dtype) -> Optional[torch.Tensor]:
        if a is None:
            return None
        if isinstance(a, torch.Tensor):
            return a.to(self.device, dtype).contiguous()
        return torch.from_numpy(np.ascontiguousarray(_host(a)).astype(
            {torch.float32: np.float32, torch.int32: np.int32}[dtype])).to(self.device)

    # -- core ----------------------------------------------------------------
    def _search(self, costs, R, roi_nodes, g2r, pen, allowed, seeds, seed_cost, targets, tg_cost,
                seed_pen: bool) -> Optional[List[int]]:
        check_search_args(self.N, self.E, R, costs, roi_nodes, g2r, pen, allowed, seeds, seed_cost, targets,
                          tg_cost)
        self._ensure(R)
        d_costs = self._dev(costs, torch.float32)
        d_seeds = self._dev(np.asarray(seeds, np.int32), torch.int32)
        d_sc = self._dev(seed_cost, torch.float32) if seed_cost is not None else None
        d_tg = self._dev(np.asarray(targets, np.int32), torch.int32)
        d_tc = self._dev(tg_cost, torch.float32) if tg_cost is not None else None
        d_roi = self._dev(roi_nodes, torch.int32) if roi_nodes is not None else None
        d_g2r = self._dev(g2r, torch.int32) if g2r is not None else None
        d_pen = self._dev(pen, torch.float32) if pen is not None else None
        d_allowed = None
        if allowed is not None:
            d_allowed = torch.from_numpy(np.ascontiguousarray(_host(allowed)).astype(np.uint32).view(np.int32)).to(
                self.device)
        grid = R >= self.grid_min
        flags = (_lib.ORX_F_SEED_PEN if seed_pen else 0) | (_lib.ORX_F_GRID if grid else 0)
        cfg = _lib.OrxRouteCfg(self.delta, self.block, 0, flags, None)
        stream = torch.cuda.current_stream().cuda_stream
        best = self._small[0:1]
        plen = self._small[1:2]
        p = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
        sw = int(self._scratch.numel())
        _lib.call("orx_sssp_csr", C.byref(self.graph), d_costs.data_ptr(), int(R), p(d_roi), p(d_g2r), p(d_pen),
                  p(d_allowed), len(seeds), d_seeds.data_ptr(), p(d_sc), len(targets), d_tg.data_ptr(), p(d_tc),
                  self._key.data_ptr(), self._scratch.data_ptr(), sw, best.data_ptr(),
                  C.byref(cfg), stream)
        if grid:
            _lib.call("orx_backtrace_csr_roi", C.byref(self.graph), self._key.data_ptr(), int(R), p(d_roi),
                      p(d_g2r), best.data_ptr(), self._path.data_ptr(), int(self._cap), plen.data_ptr(),
                      self._scratch.data_ptr(), sw, stream)
        else:
            _lib.call("orx_backtrace_csr", C.byref(self.graph), self._key.data_ptr(), p(d_g2r), best.data_ptr(),
                      self._path.data_ptr(), int(self._cap), plen.data_ptr(), stream)
        n = int(plen.item())
        if n == _lib.ORX_NO_PATH:
            return None
        if n < 0:
            raise _lib.HipError("orx_backtrace_csr", n, "parent walk failed")
        path = self._path[:n].cpu().numpy().astype(np.int64).tolist()
        return path if len(path) > 1 else None

    # -- seam methods --------------------------------------------------------
    def find_path_roi_gpu(self, src: int, dst: int, costs, roi_nodes, global_to_roi,
                          node_penalty=None) -> Optional[List[int]]:
        if not _in_roi(global_to_roi, [int(src), int(dst)]).all():
            return None
        R = len(roi_nodes)
        return self._search(costs, R, roi_nodes, global_to_roi, node_penalty, None, [int(src)], None, [int(dst)],
                            None, seed_pen=False)

    find_path_roi = find_path_roi_gpu

    def find_path_multisource_multisink(self, src_seeds: Sequence[Tuple[int, float]],
                                        dst_targets: Sequence[Tuple[int, float]], costs, roi_nodes,
                                        global_to_roi, node_penalty=None):
        src_seeds, dst_targets = list(src_seeds), list(dst_targets)
        s_in = _in_roi(global_to_roi, [int(n) for n, _ in src_seeds]) if src_seeds else np.zeros(0, bool)
        t_in = _in_roi(global_to_roi, [int(n) for n, _ in dst_targets]) if dst_targets else np.zeros(0, bool)
        seeds = [(int(n), float(c)) for (n, c), ok in zip(src_seeds, s_in) if ok]
        tgs = [(int(n), float(c)) for (n, c), ok in zip(dst_targets, t_in) if ok]
        if not seeds or not tgs:
            return None
        path = self._search(costs, len(roi_nodes), roi_nodes, global_to_roi, node_penalty, None,
                            [n for n, _ in seeds], np.asarray([c for _, c in seeds], np.float32),
                            [n for n, _ in tgs], np.asarray([c for _, c in tgs], np.float32), seed_pen=True)
        if path is None:
            return None
        if self.plane_size:
            return path, path[0] // self.plane_size, path[-1] // self.plane_size
        return path, 0, 0

    def find_path_fullgraph_gpu_seeds(self, costs, src_seeds, dst_targets, ub_hint=None, *,
                                      src_seed_costs=None, dst_target_costs=None, node_penalty=None,
                                      allowed_bitmap=None, use_bitmap=False) -> Optional[List[int]]:
        seeds = np.asarray(_host(src_seeds), np.int32).reshape(-1)
        tgs = np.asarray(_host(dst_targets), np.int32).reshape(-1)
        if seeds.size == 0 or tgs.size == 0:
            return None
        sc = np.zeros(seeds.size, np.float32) if src_seed_costs is None else \
            np.asarray(_host(src_seed_costs), np.float32).reshape(-1)
        tc = np.zeros(tgs.size, np.float32) if dst_target_costs is None else \
            np.asarray(_host(dst_target_costs), np.float32).reshape(-1)
        allowed = allowed_bitmap if use_bitmap else None
        return self._search(costs, self.N, None, None, node_penalty, allowed, seeds.tolist(), sc, tgs.tolist(), tc,
                            seed_pen=True)


def graft(solver, hip) -> None:
    """Wrap ``solver.find_path_roi`` / ``find_path_multisource_multisink`` so
    the accelerator ``hip`` answers first, and set ``solver.gpu_solver``
    (unified_pathfinder.py:3099-3120).  A ``None`` from the accelerator, or
    any exception it raises (logged), hands the query to the solver's own
    original method, so a net the CPU could route is never dropped
    (metal_dijkstra.py:374-387 catches the same way).  The original method
    runs with ``gpu_solver`` unset, so it does not search on the GPU a
    second time (SimpleDijkstra.find_path_roi tries gpu_solver first for
    large ROIs, UP:2465-2480)."""
    orig_roi = solver.find_path_roi
    orig_multi = getattr(solver, "find_path_multisource_multisink", None)

    def on_cpu(fn, *args, **kw):
        saved = getattr(solver, "gpu_solver", None)
        solver.gpu_solver = None
        try:
            return fn(*args, **kw)
        finally:
            solver.gpu_solver = saved

    def roi(src, dst, costs, roi_nodes, global_to_roi, node_penalty=None):
        try:
            path = hip.find_path_roi_gpu(src, dst, costs, roi_nodes, global_to_roi, node_penalty=node_penalty)
        except Exception as e:   # noqa: BLE001 -- any accelerator failure goes to the CPU path
            log.warning("[HIP] find_path_roi failed (%s); CPU fallback", e)
            path = None
        if path is not None:
            return path
        return on_cpu(orig_roi, src, dst, costs, roi_nodes, global_to_roi, node_penalty=node_penalty)

    def multi(src_seeds, dst_targets, costs, roi_nodes, global_to_roi, node_penalty=None):
        try:
            res = hip.find_path_multisource_multisink(src_seeds, dst_targets, costs, roi_nodes, global_to_roi,
                                                      node_penalty=node_penalty)
        except Exception as e:   # noqa: BLE001
            log.warning("[HIP] find_path_multisource_multisink failed (%s); CPU fallback", e)
            res = None
        if res is not None:
            return res
        return on_cpu(orig_multi, src_seeds, dst_targets, costs, roi_nodes, global_to_roi,
                      node_penalty=node_penalty)

    solver.find_path_roi = roi
    if orig_multi is not None:
        solver.find_path_multisource_multisink = multi
    solver.gpu_solver = hip
    solver.hip_solver = hip


def try_attach_hip(solver, graph, lattice=None, *, device="cuda") -> bool:
    """Graft the HIP solver onto a SimpleDijkstra-like object (the
    ORTHO_BACKEND=hip counterpart of try_attach_metal, metal_dijkstra.py:44-89).
    Returns False (solver untouched) when the library or GPU is unavailable;
    see ``graft`` for the wrapping."""
    try:
        hip = HipDijkstra(graph, lattice, device=device)
    except (_lib.HipUnavailable, RuntimeError, OSError) as e:
        log.info("[HIP] accelerator not attached: %s", e)
        return False
    graft(solver, hip)
    log.info("[HIP] gfx950 accelerator attached")
    return True
