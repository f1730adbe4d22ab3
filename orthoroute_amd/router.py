"""Device-resident PathFinder driver (the host side of the hot path).

Restates the negotiation loop of UnifiedPathFinder._pathfinder_negotiation
(unified_pathfinder.py:6163-7391) in its plain form, with every array
resident in HBM and every per-net step on the GPU:

  per iteration k (1-based):
    1. costs for pres_fac p_k from base, history and occupancy
       (update_costs + begin_live_present_costs, UP:6411-6418, 7940-7941)
    2. net order = _order_nets_by_difficulty (UP:7840-7901)
    3. route: batched -> disjoint-ROI waves, one launch per wave;
              plain   -> one launch per net, in order (the baseline form)
       each net: rip-up, SSSP in its ROI box, backtrace, commit, live
       re-price (UP:7949-8430)
    4. overuse + history (compute_overuse UP:6575, update_history
       UP:7161-7167), then p_{k+1} = min(p_k * mult, pres_fac_max) (UP:7379-7383)

Buffers are PyTorch-ROCm tensors; kernels are called through the C ABI
(liborx_hip.so).  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .lattice import LatticeSpec
from .netgen import Netlist
from .waves import (check_schedule, color_waves, group_waves, net_distance, net_order, plan_iteration,
                    pres_schedule, resolve_pres_fac_max)


@dataclass
class RouterConfig:
    """Cost schedule and launch parameters (defaults: the plain form)."""
    iterations: int = 10
    pres_init: float = 0.5
    pres_mult: float = 1.5
    pres_max: float | None = None     # None -> layer-dependent ceiling (UP:801-816)
    base_w: float = 1.0
    hist_w: float = 1.0
    node_pres_w: float = 1.0
    node_hist_w: float = 1.0
    edge_cap: int = 1
    node_cap: int = 1
    hist_gain: float = 1.0
    hist_decay: float = 1.0
    hist_cap_mult: float = 0.0
    node_hist_gain: float = 1.0
    delta: float = 1.0
    block: int = 256
    lds_nodes: int = 12288
    astar: bool = True
    stop_at_zero_overuse: bool = False
    collect_stats: bool = True

    def schedule(self, L: int) -> np.ndarray:
        pmax = self.pres_max if self.pres_max is not None else resolve_pres_fac_max(L)
        return pres_schedule(self.iterations, self.pres_init, self.pres_mult, pmax)

    def params(self, pres: float, astar=(0.0, 0.0, 0.0)) -> _lib.OrxParams:
        p = np.float32(pres)
        return _lib.OrxParams(
            pres_fac=float(p), base_w=self.base_w, hist_w=self.hist_w,
            node_pres=float(np.float32(self.node_pres_w) * p), node_hist_w=self.node_hist_w,
            edge_cap=self.edge_cap, node_cap=self.node_cap, hist_gain=self.hist_gain,
            hist_decay=self.hist_decay, hist_cap_mult=self.hist_cap_mult,
            node_hist_gain=self.node_hist_gain,
            astar_cp=astar[0], astar_cv=astar[1], astar_gmax=astar[2])

    def astar_bounds(self, base: np.ndarray) -> tuple[float, float, float]:
        """(cp, cv, gmax) of the sink-directed heuristic, (0, 0, 0) when off."""
        weights = (self.base_w, self.hist_w, self.node_pres_w, self.node_hist_w, self.hist_gain,
                   self.node_hist_gain, self.hist_decay)
        if not self.astar or min(weights) < 0 or self.pres_init < 0 or self.pres_mult < 0:
            return (0.0, 0.0, 0.0)
        return astar_bounds(base, self.base_w)


@dataclass
class IterationStats:
    iteration: int
    pres_fac: float
    routed: int
    failed: int
    edge_over: int
    edge_over_count: int
    node_over: int
    node_over_count: int
    via_hops: int
    barrels: int
    settled: int = 0
    relaxed: int = 0
    items: int = 0
    rounds: int = 0
    phases: int = 0
    path_hops: int = 0
    waves: int = 0
    seconds: float = 0.0
    gpu_ms: float = 0.0
    route_ms: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def overuse(self) -> int:
        return self.edge_over + self.node_over


ASTAR_MARGIN_BITS = 8


def _round_down_f32(x: float) -> np.float32:
    f = np.float32(x)
    if float(f) > x:
        f = np.nextafter(f, np.float32(0))
    return f


def astar_bounds(base: np.ndarray, base_w: float, margin_bits: int = ASTAR_MARGIN_BITS):
    """Per-move lower bounds of the entering weight for the A* heuristic.

    Every entering weight is fl(fl(fl(b*base_w) + ...) + ...) + q with
    non-negative terms, so it is at least fl(b*base_w).  The bounds used are
    c' <= c_min * (1 - 2^-margin_bits); the margin m = c_min - c' covers half
    an fp32 ulp of every distance below gmax = 2^(24 + floor(log2 m)), which
    makes f = g + h monotone along tight arcs there."""
    b = np.asarray(base, np.float32).reshape(-1, 4)
    planar = b[:, 1:3][np.isfinite(b[:, 1:3])]
    via = b[:, [0, 3]][np.isfinite(b[:, [0, 3]])]
    if planar.size == 0:
        return (0.0, 0.0, 0.0)
    bw = np.float32(base_w)
    cmin_p = float(np.float32(planar.min()) * bw)
    cmin_v = float(np.float32(via.min()) * bw) if via.size else cmin_p
    if cmin_p <= 0 or cmin_v <= 0:
        return (0.0, 0.0, 0.0)
    k = 1.0 - 2.0 ** -margin_bits
    cp = _round_down_f32(cmin_p * k)
    cv = _round_down_f32(cmin_v * k)
    m = min(cmin_p - float(cp), cmin_v - float(cv))
    gmax = 2.0 ** min(24 + int(np.floor(np.log2(m))), 120)
    return (float(cp), float(cv), float(gmax))


def _stream_handle(stream: torch.cuda.Stream | None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class LatticeState:
    """Lattice arrays resident in HBM (see orx_lattice in include/orthoroute_hip.h)."""

    def __init__(self, spec: LatticeSpec, base: np.ndarray | None = None, device="cuda"):
        self.spec = spec
        self.device = torch.device(device)
        N = spec.N
        b = spec.base_costs() if base is None else np.asarray(base, np.float32)
        if b.shape != (4 * N,):
            raise ValueError("base must have 4*N entries")
        dev = self.device
        self.base = torch.from_numpy(b).to(dev)
        self.hist = torch.zeros(4 * N, dtype=torch.float32, device=dev)
        self.occ = torch.zeros(4 * N, dtype=torch.int32, device=dev)
        self.wcost = torch.empty(4 * N, dtype=torch.float32, device=dev)
        self.node_hist = torch.zeros(N, dtype=torch.float32, device=dev)
        self.node_use = torch.zeros(N, dtype=torch.int32, device=dev)
        self.via_col = torch.zeros(spec.XY, dtype=torch.int32, device=dev)
        self.struct = _lib.OrxLattice(
            spec.X, spec.Y, spec.L, spec.hmask,
            self.base.data_ptr(), self.hist.data_ptr(), self.occ.data_ptr(), self.wcost.data_ptr(),
            self.node_hist.data_ptr(), self.node_use.data_ptr(), self.via_col.data_ptr())

    def nbytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in
                   (self.base, self.hist, self.occ, self.wcost, self.node_hist, self.node_use, self.via_col))

    def reprice(self, params: _lib.OrxParams, stream=None):
        _lib.call("orx_lattice_reprice", C.byref(self.struct), C.byref(params), _stream_handle(stream))

    def to_numpy(self) -> dict:
        return {k: getattr(self, k).cpu().numpy() for k in
                ("hist", "occ", "wcost", "node_hist", "node_use", "via_col")}


class NetTables:
    """Per-net device tables of one job (orx_nets)."""

    def __init__(self, netlist: Netlist, device="cuda"):
        dev = torch.device(device)
        n = netlist.n
        self.n = n
        R = netlist.roi_nodes()
        if R.max(initial=0) >= 2**24:
            raise ValueError("an ROI box exceeds 2^24 nodes")
        cap = R.astype(np.int64)
        off = np.zeros(n, np.int64)
        if n:
            off[1:] = np.cumsum(cap)[:-1]
        self.arena_size = int(cap.sum())
        self.roi_nodes = R
        self.box = torch.from_numpy(np.ascontiguousarray(netlist.box.reshape(-1), np.int32)).to(dev)
        self.src = torch.from_numpy(netlist.src.astype(np.int32)).to(dev)
        self.dst = torch.from_numpy(netlist.dst.astype(np.int32)).to(dev)
        self.path_arena = torch.zeros(max(1, self.arena_size), dtype=torch.int32, device=dev)
        self.path_off_np = off
        self.path_off = torch.from_numpy(off).to(dev)
        self.path_cap = torch.from_numpy(cap.astype(np.int32)).to(dev)
        self.path_len = torch.zeros(n, dtype=torch.int32, device=dev)
        self.path_cost = torch.full((n,), float("inf"), dtype=torch.float32, device=dev)
        self.status = torch.zeros(n, dtype=torch.int32, device=dev)
        # per-net search workspace offsets (dependency-driven launch) and
        # completion flags stamped with the iteration epoch
        r64 = (R.astype(np.int64) + 63) // 64 * 64
        ws_off = np.zeros(n, np.int64)
        if n:
            ws_off[1:] = np.cumsum(r64)[:-1]
        self.ws_nodes = int(r64.sum())
        self.ws_off = torch.from_numpy(ws_off).to(dev)
        self.done = torch.full((max(1, n),), -1, dtype=torch.int32, device=dev)
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        self.struct = _lib.OrxNets(
            n, self.box.data_ptr(), self.src.data_ptr(), self.dst.data_ptr(), self.path_arena.data_ptr(),
            self.path_off.data_ptr(), self.path_cap.data_ptr(), self.path_len.data_ptr(),
            self.path_cost.data_ptr(), self.status.data_ptr())

    def nbytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in
                   (self.box, self.src, self.dst, self.path_arena, self.path_off, self.path_cap,
                    self.path_len, self.path_cost, self.status, self.ws_off, self.done))

    def paths(self) -> list[np.ndarray]:
        arena = self.path_arena.cpu().numpy()
        lens = self.path_len.cpu().numpy()
        return [arena[o:o + l].copy() for o, l in zip(self.path_off_np, lens)]


class Workspace:
    """Search scratch of a wave (orx_workspace), grown on demand."""

    def __init__(self, device="cuda"):
        self.device = torch.device(device)
        self.nodes = 0
        self.keys = self.queues = self.bitmaps = None
        self.struct = None

    def ensure(self, nodes: int):
        nodes = int((nodes + 63) // 64 * 64)
        if nodes <= self.nodes:
            return
        nodes = max(nodes, int(self.nodes * 1.25))
        nodes = (nodes + 63) // 64 * 64
        dev = self.device
        self.keys = torch.empty(nodes, dtype=torch.int64, device=dev)
        self.queues = torch.empty(4 * nodes, dtype=torch.int32, device=dev)
        self.bitmaps = torch.empty(3 * nodes // 32, dtype=torch.int32, device=dev)
        self.nodes = nodes
        self.struct = _lib.OrxWorkspace(self.keys.data_ptr(), self.queues.data_ptr(),
                                        self.bitmaps.data_ptr(), nodes)

    def nbytes(self) -> int:
        return self.nodes * (8 + 16) + 3 * self.nodes // 8


def _wave_offsets(roi_nodes: np.ndarray, nets: np.ndarray) -> tuple[np.ndarray, int]:
    r = (roi_nodes[nets] + 63) // 64 * 64
    off = np.zeros(len(nets), np.int64)
    if len(nets):
        off[1:] = np.cumsum(r)[:-1]
    return off, int(r.sum())


class DevicePathFinder:
    """PathFinder on one routing job, state resident on the GPU."""

    def __init__(self, netlist: Netlist, cfg: RouterConfig | None = None, device="cuda",
                 stream: torch.cuda.Stream | None = None, base: np.ndarray | None = None):
        _lib.init_device()
        self.netlist = netlist
        self.cfg = cfg or RouterConfig()
        self.spec = netlist.spec
        self.device = torch.device(device)
        self.stream = stream
        self.lat = LatticeState(self.spec, base, device)
        self.nets = NetTables(netlist, device)
        self.ws = Workspace(device)
        self.distance = net_distance(self.spec, netlist.src, netlist.dst)
        self.schedule = self.cfg.schedule(self.spec.L)
        self.astar = self.cfg.astar_bounds(self.spec.base_costs() if base is None else base)
        self.iteration = 0
        self.max_groups = 0
        self.history: list[IterationStats] = []
        self._route_stats = torch.zeros(8, dtype=torch.int64, device=self.device)
        self._iter_out = torch.zeros(8, dtype=torch.int64, device=self.device)
        self._over = torch.zeros(max(1, netlist.n), dtype=torch.int64, device=self.device)
        self._hops = torch.zeros(max(1, netlist.n), dtype=torch.int32, device=self.device)

    # -- helpers -----------------------------------------------------------
    def _s(self) -> int:
        return _stream_handle(self.stream)

    def footprint_bytes(self) -> int:
        return self.lat.nbytes() + self.nets.nbytes() + self.ws.nbytes()

    def route_cfg(self, flags: int = 0) -> _lib.OrxRouteCfg:
        c = self.cfg
        fl = flags | (_lib.ORX_F_STATS if c.collect_stats else 0)
        return _lib.OrxRouteCfg(float(c.delta), int(c.block), int(c.lds_nodes), fl, self._route_stats.data_ptr())

    def order_for(self, iteration: int) -> np.ndarray:
        """_order_nets_by_difficulty from the device occupancy (one readback)."""
        n = self.netlist.n
        if iteration > 1 and n:
            _lib.call("orx_net_overuse", C.byref(self.lat.struct), C.byref(self.nets.struct),
                      self.cfg.edge_cap, self._over.data_ptr(), self._hops.data_ptr(), self._s())
            over = self._over[:n].cpu().numpy()
            hops = self._hops[:n].cpu().numpy()
        else:
            over = np.zeros(n, np.int64)
            hops = np.zeros(n, np.int32)
        return net_order(iteration, self.distance, over, hops)

    def plan(self, order: np.ndarray, mode: str) -> list[np.ndarray]:
        """Launch groups: one per net (plain) or one per disjoint-ROI wave."""
        if mode == "plain":
            return [np.asarray([k], np.int32) for k in order]
        if mode not in ("batched", "waves"):
            raise ValueError("mode must be 'batched' ('waves') or 'plain'")
        wave = color_waves(self.netlist.box, order, self.spec.X, self.spec.Y)
        return group_waves(wave, order, self.nets.roi_nodes)

    # -- one iteration -----------------------------------------------------
    def run_iteration(self, mode: str = "batched", order: np.ndarray | None = None,
                      waves: list[np.ndarray] | None = None) -> IterationStats:
        if self.iteration >= len(self.schedule):
            raise RuntimeError("schedule exhausted; raise cfg.iterations")
        it = self.iteration + 1
        pres = float(self.schedule[it - 1])
        params = self.cfg.params(pres, self.astar)
        stream = self.stream if self.stream is not None else torch.cuda.current_stream()
        t0 = time.perf_counter()
        if order is None:
            order = self.order_for(it)
        if mode == "dag":
            raise RuntimeError("the dependency-driven launch (orx_route_dag) is switched off on the device")
        if waves is None:
            waves = self.plan(order, mode)
        # upload the wave plan once
        offs = []
        need = 0
        for w in waves:
            off, total = _wave_offsets(self.nets.roi_nodes, w)
            offs.append(off)
            need = max(need, total)
        self.ws.ensure(need)
        flat_nets = np.concatenate(waves) if waves else np.zeros(0, np.int32)
        flat_off = np.concatenate(offs) if offs else np.zeros(0, np.int64)
        d_nets = torch.from_numpy(flat_nets.astype(np.int32)).to(self.device, non_blocking=False)
        d_off = torch.from_numpy(flat_off.astype(np.int64)).to(self.device, non_blocking=False)
        self._route_stats.zero_()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        evr0 = torch.cuda.Event(enable_timing=True)
        evr1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        self.lat.reprice(params, stream)
        rcfg = self.route_cfg()
        evr0.record(stream)
        s = self._s()
        pos = 0
        for w in waves:
            K = len(w)
            maxR = int(self.nets.roi_nodes[w].max())
            _lib.call("orx_route_wave", C.byref(self.lat.struct), C.byref(params), C.byref(self.nets.struct),
                      d_nets.data_ptr() + 4 * pos, d_off.data_ptr() + 8 * pos, K, maxR,
                      C.byref(self.ws.struct), C.byref(rcfg), s)
            pos += K
        evr1.record(stream)
        _lib.call("orx_lattice_end_iteration", C.byref(self.lat.struct), C.byref(params),
                  self._iter_out.data_ptr(), s)
        ev1.record(stream)
        out = self._iter_out.cpu().numpy()
        rs = self._route_stats.cpu().numpy()
        stream.synchronize()
        dt = time.perf_counter() - t0
        st = IterationStats(
            iteration=it, pres_fac=pres, routed=int(rs[5]), failed=int(rs[6]),
            edge_over=int(out[0]), edge_over_count=int(out[1]), node_over=int(out[2]),
            node_over_count=int(out[3]), via_hops=int(out[4]), barrels=int(out[5]),
            settled=int(rs[0]), relaxed=int(rs[1]), items=int(rs[2]), rounds=int(rs[3]), phases=int(rs[4]),
            path_hops=int(rs[7]), waves=len(waves), seconds=dt, gpu_ms=ev0.elapsed_time(ev1),
            route_ms=evr0.elapsed_time(evr1))
        self.iteration = it
        self.history.append(st)
        return st

    def _run_dag(self, it, pres, params, order, stream, t0) -> IterationStats:
        """One persistent dependency-driven launch for the whole iteration."""
        nl = self.netlist
        wave, sched, dep_ptr, dep_ids = plan_iteration(nl.box, order, self.spec.X, self.spec.Y, self.nets.roi_nodes)
        check_schedule(sched, dep_ptr, dep_ids)
        self.ws.ensure(self.nets.ws_nodes)
        d_sched = torch.from_numpy(sched).to(self.device)
        d_ptr = torch.from_numpy(dep_ptr).to(self.device)
        d_ids = torch.from_numpy(dep_ids if len(dep_ids) else np.zeros(1, np.int32)).to(self.device)
        self._route_stats.zero_()
        ev0, ev1, evr0, evr1 = (torch.cuda.Event(enable_timing=True) for _ in range(4))
        ev0.record(stream)
        self.lat.reprice(params, stream)
        rcfg = self.route_cfg()
        s = self._s()
        evr0.record(stream)
        _lib.call("orx_route_dag", C.byref(self.lat.struct), C.byref(params), C.byref(self.nets.struct),
                  d_sched.data_ptr(), nl.n, d_ptr.data_ptr(), d_ids.data_ptr(), self.nets.done.data_ptr(), it,
                  self.nets.counter.data_ptr(), self.nets.ws_off.data_ptr(), int(self.nets.roi_nodes.max()),
                  C.byref(self.ws.struct), C.byref(rcfg), int(self.max_groups), s)
        evr1.record(stream)
        _lib.call("orx_lattice_end_iteration", C.byref(self.lat.struct), C.byref(params),
                  self._iter_out.data_ptr(), s)
        ev1.record(stream)
        out = self._iter_out.cpu().numpy()
        rs = self._route_stats.cpu().numpy()
        stream.synchronize()
        dt = time.perf_counter() - t0
        st = IterationStats(
            iteration=it, pres_fac=pres, routed=int(rs[5]), failed=int(rs[6]),
            edge_over=int(out[0]), edge_over_count=int(out[1]), node_over=int(out[2]),
            node_over_count=int(out[3]), via_hops=int(out[4]), barrels=int(out[5]),
            settled=int(rs[0]), relaxed=int(rs[1]), items=int(rs[2]), rounds=int(rs[3]), phases=int(rs[4]),
            path_hops=int(rs[7]), waves=int(wave.max()) + 1 if len(wave) else 0, seconds=dt,
            gpu_ms=ev0.elapsed_time(ev1), route_ms=evr0.elapsed_time(evr1),
            extra={"launches": 1, "deps": int(len(dep_ids))})
        self.iteration = it
        self.history.append(st)
        return st

    def run(self, iterations: int | None = None, mode: str = "batched") -> list[IterationStats]:
        n_it = iterations or self.cfg.iterations
        out = []
        for _ in range(n_it):
            st = self.run_iteration(mode)
            out.append(st)
            if self.cfg.stop_at_zero_overuse and st.overuse == 0 and st.failed == 0:
                break
        return out

    # -- results -------------------------------------------------------------
    def paths(self) -> list[np.ndarray]:
        return self.nets.paths()

    def path_costs(self) -> np.ndarray:
        return self.nets.path_cost.cpu().numpy()

    def status(self) -> np.ndarray:
        return self.nets.status.cpu().numpy()
