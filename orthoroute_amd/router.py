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
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .lattice import LatticeSpec
from .netgen import Netlist
from .waves import (check_schedule, net_distance, net_order, plan_iteration, plan_waves, pres_schedule,
                    resolve_pres_fac_max)


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
    graph: bool = False               # capture each iteration's wave launches in one hipGraph
    ws_cap_nodes: int = 1 << 24       # search workspace bound: larger waves launch in parts
    hotset: bool = False              # reroute only offenders (hotset_filter) unless the overuse stalled
    adapt_delta: bool = True          # widen delta x4 after an iteration of many thin threshold phases
    delta_max: float = 1048576.0
    engine: str = "tiles"             # "tiles": tile-step search (orx_route_tiles); "wave": one workgroup per net
    tile_queue: bool = False          # tile engine, queue form: one persistent launch per wave (ORX_F_ASYNC)
    small_wave_nodes: int = 16384     # tile engine: waves whose largest ROI is at most this run one workgroup per net
    tile_groups: int = 0              # tile engine: cap on the workgroups of a launch (0: the whole chip)
    head_nets: int = 128              # batched: plan and launch the first this many nets of the order before the rest

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
        dev = self.device
        if base is None:
            self.base = spec.base_costs_device(dev)
        else:
            b = np.asarray(base, np.float32)
            if b.shape != (4 * N,):
                raise ValueError("base must have 4*N entries")
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
        # multi-terminal nets: seed / target tables (orx_nets.seed_ptr ...)
        self.terms = []
        tp = [0] * 6
        if netlist.terms is not None:
            T = netlist.terms
            T.validate(n, netlist.spec.N)
            zero = np.float32(0.0)   # + 0 turns -0.0 into +0.0 (keys order by fp32 bits)
            arrs = [(T.seed_ptr, np.int32), (T.seed_node, np.int32), (T.seed_cost + zero, np.float32),
                    (T.tgt_ptr, np.int32), (T.tgt_node, np.int32), (T.tgt_cost + zero, np.float32)]
            self.terms = [torch.from_numpy(np.ascontiguousarray(a, t)).to(dev) if len(a) else
                          torch.zeros(1, dtype=torch.int32 if t == np.int32 else torch.float32, device=dev)
                          for a, t in arrs]
            tp = [t.data_ptr() for t in self.terms]
        self.struct = _lib.OrxNets(
            n, self.box.data_ptr(), self.src.data_ptr(), self.dst.data_ptr(), self.path_arena.data_ptr(),
            self.path_off.data_ptr(), self.path_cap.data_ptr(), self.path_len.data_ptr(),
            self.path_cost.data_ptr(), self.status.data_ptr(), *tp)

    def nbytes(self) -> int:
        return sum(t.numel() * t.element_size() for t in
                   (self.box, self.src, self.dst, self.path_arena, self.path_off, self.path_cap,
                    self.path_len, self.path_cost, self.status, self.ws_off, self.done, *self.terms))

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
        nodes = int((max(int(nodes), TILE_MIN_WS) + 63) // 64 * 64)
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


TILE_X = 16           # tile width of the tile engine (cells along x)
TILE_RESERVE = 512    # workspace nodes the tile engine keeps for its step counters
TILE_MIN_WS = 1024    # smallest workspace the tile engine accepts


def tile_height(L: int) -> int:
    """Tile height (cells along y) of the tile engine for L layers: 16 for
    the layer counts 6, 8, 12, 16 and 20, else 10 (at most 5120 nodes)."""
    return 16 if L in (6, 8, 12, 16, 20) else 10


def tile_steps(box: np.ndarray, L: int) -> np.ndarray:
    """Step budget of each net on the tile engine: its tile count along x
    plus along y (about the tile distance from a corner to the opposite
    one; a search still active after its wave's budget finishes in the
    single-workgroup drain)."""
    b = np.asarray(box).reshape(-1, 4).astype(np.int64)
    ty = tile_height(L)
    return ((b[:, 2] + TILE_X - 1) // TILE_X + (b[:, 3] + ty - 1) // ty).astype(np.int32)


def _wave_offsets(roi_nodes: np.ndarray, nets: np.ndarray) -> tuple[np.ndarray, int]:
    r = (roi_nodes[nets] + 63) // 64 * 64
    off = np.zeros(len(nets), np.int64)
    if len(nets):
        off[1:] = np.cumsum(r)[:-1]
    return off, int(r.sum())


def hotset_filter(order: np.ndarray, edge_over: np.ndarray, node_over: np.ndarray,
                  hops: np.ndarray) -> np.ndarray:
    """The nets of ``order`` (kept in that order) that the hotset reroutes:
    those whose committed path holds an over-capacity edge or node, and
    those without a path (_build_hotset, unified_pathfinder.py:9143-9171:
    offenders on present overuse plus path-node conflicts, and unrouted
    nets (0 hops); its caps, freezing and random exploration are not part
    of it).  The router applies it from iteration 2 on, except after an
    iteration that did not lower the overuse, when every net is rerouted."""
    hot = (np.asarray(edge_over) > 0) | (np.asarray(node_over) > 0) | (np.asarray(hops) == 0)
    return order[hot[order]]


def split_waves(waves: list[np.ndarray], roi_nodes: np.ndarray, cap_nodes: int) -> list[np.ndarray]:
    """Split every wave whose 64-aligned ROI total exceeds ``cap_nodes`` into
    consecutive parts that fit (nets of one wave are independent, so any
    split gives the same result).  Every ROI must fit on its own."""
    if not waves:
        return []
    sizes = np.fromiter((len(w) for w in waves), np.int64, len(waves))
    if sizes.min() > 0:
        # common case, one pass: no wave exceeds the cap
        r64_all = (roi_nodes[np.concatenate(waves)].astype(np.int64) + 63) // 64 * 64
        starts = np.zeros(len(waves), np.int64)
        np.cumsum(sizes[:-1], out=starts[1:])
        if int(np.add.reduceat(r64_all, starts).max()) <= cap_nodes:
            return list(waves)
    out = []
    for w in waves:
        r64 = (roi_nodes[w].astype(np.int64) + 63) // 64 * 64
        if len(w) == 0 or int(r64.sum()) <= cap_nodes:
            out.append(w)
            continue
        if int(r64.max()) > cap_nodes:
            raise ValueError("an ROI exceeds the workspace cap")
        start, acc = 0, 0
        for i, r in enumerate(r64.tolist()):
            if acc + r > cap_nodes:
                out.append(w[start:i])
                start, acc = i, 0
            acc += r
        out.append(w[start:])
    return out


def footprint_estimate(netlist: Netlist, cfg: RouterConfig | None = None) -> int:
    """HBM bytes a DevicePathFinder on ``netlist`` holds at most: lattice
    arrays (72 B/node + 4 B/column), path arena (4 B per ROI node of every
    net), per-net tables, and the search workspace of the largest launch
    (24.375 B per node, bounded by the workspace cap)."""
    cfg = cfg or RouterConfig()
    spec = netlist.spec
    R = netlist.roi_nodes().astype(np.int64)
    r64 = (R + 63) // 64 * 64
    ws_nodes = min(int(r64.sum()), max(int(cfg.ws_cap_nodes), int(r64.max(initial=64))))
    ws_nodes = int(ws_nodes * 1.25) // 64 * 64 + 64  # Workspace.ensure grows by 1.25x
    lattice = spec.N * 72 + spec.XY * 4
    nets = int(R.sum()) * 4 + netlist.n * 64
    return lattice + nets + ws_nodes * 24 + 3 * ws_nodes // 8 + (1 << 20)


def _flatten_waves(waves: list[np.ndarray], roi_nodes: np.ndarray):
    """Launch plan of orx_route_waves: net ids and 64-aligned workspace
    offsets (restarting at 0 in every wave), wave_ptr[n_waves+1],
    wave_max[n_waves] (largest ROI of each wave), wave_ws[n_waves] (the
    workspace nodes each wave occupies) and the workspace size the largest
    wave needs."""
    nw = len(waves)
    sizes = np.fromiter((len(w) for w in waves), np.int64, nw)
    wave_ptr = np.zeros(nw + 1, np.int32)
    np.cumsum(sizes, out=wave_ptr[1:])
    if nw == 0:
        return (np.zeros(0, np.int32), np.zeros(0, np.int64), wave_ptr, np.zeros(0, np.int64),
                np.zeros(0, np.int64), 0)
    flat = np.concatenate(waves).astype(np.int32)
    R = roi_nodes[flat].astype(np.int64)
    r64 = (R + 63) // 64 * 64
    incl = np.cumsum(r64)
    start = np.repeat(incl[wave_ptr[:-1]] - r64[wave_ptr[:-1]], sizes)
    off = (incl - r64 - start).astype(np.int64)
    wave_max = np.maximum.reduceat(R, wave_ptr[:-1]) if len(flat) else np.zeros(nw, np.int64)
    totals = np.add.reduceat(r64, wave_ptr[:-1]).astype(np.int64)
    return flat, off, wave_ptr, wave_max.astype(np.int64), totals, int(totals.max())


@dataclass
class _Pending:
    """An iteration enqueued by launch_iteration and not yet collected."""
    it: int
    pres: float
    waves: int
    t0: float
    events: tuple
    done: torch.cuda.Event
    graph: object
    keep: tuple        # device buffers the launches read
    nets: int


class DevicePathFinder:
    """PathFinder on one routing job, state resident on the GPU."""

    def __init__(self, netlist: Netlist, cfg: RouterConfig | None = None, device="cuda",
                 stream: torch.cuda.Stream | None = None, base: np.ndarray | None = None):
        _lib.init_device()
        self.netlist = netlist
        self.cfg = cfg or RouterConfig()
        self.spec = netlist.spec
        self.device = torch.device(device)
        # every torch op and kernel of this job runs on one stream, so jobs on
        # different streams overlap on the device and never race
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        n = netlist.n
        with torch.cuda.stream(self.stream):
            self.lat = LatticeState(self.spec, base, device)
            self.nets = NetTables(netlist, device)
            self.ws = Workspace(device)
            self._route_stats = torch.zeros(8, dtype=torch.int64, device=self.device)
            self._iter_out = torch.zeros(8, dtype=torch.int64, device=self.device)
            self._over = torch.zeros(max(1, n), dtype=torch.int64, device=self.device)
            self._node_over = torch.zeros(max(1, n), dtype=torch.int64, device=self.device)
            self._hops = torch.zeros(max(1, n), dtype=torch.int32, device=self.device)
            self._plan_nets = torch.empty(max(1, n), dtype=torch.int32, device=self.device)
            self._plan_off = torch.empty(max(1, n), dtype=torch.int64, device=self.device)
        # pinned host staging: the launch path copies asynchronously on this
        # job's stream and never waits for other streams' work
        self._iter_out_h = torch.zeros(8, dtype=torch.int64, pin_memory=True)
        self._route_stats_h = torch.zeros(8, dtype=torch.int64, pin_memory=True)
        self._over_h = torch.zeros(max(1, n), dtype=torch.int64, pin_memory=True)
        self._node_over_h = torch.zeros(max(1, n), dtype=torch.int64, pin_memory=True)
        self._hops_h = torch.zeros(max(1, n), dtype=torch.int32, pin_memory=True)
        self._plan_nets_h = torch.empty(max(1, n), dtype=torch.int32, pin_memory=True)
        self._plan_off_h = torch.empty(max(1, n), dtype=torch.int64, pin_memory=True)
        self.stream.synchronize()
        self.distance = net_distance(self.spec, netlist.src, netlist.dst)
        self._tile_steps = tile_steps(netlist.box, self.spec.L)
        self.schedule = self.cfg.schedule(self.spec.L)
        self.astar = self.cfg.astar_bounds(self.spec.base_cost_minima() if base is None else base)
        self.iteration = 0
        self.max_groups = 0
        # jobs sharing the GPU with this one (the serving loop sets it).  A
        # tile-step workgroup holds a whole CU's LDS, and such launches from
        # one job stall the dispatch of the others' kernels; so while jobs
        # share the device, every wave runs one workgroup per net (the same
        # fixpoint, so the same result)
        self.concurrent_jobs = 1
        self.delta_mult = 1.0     # adaptive widening of cfg.delta (adapt_delta)
        self.history: list[IterationStats] = []
        self._pending: _Pending | None = None
        self._staged_for = -1     # iteration whose order input is staged in _over_h / _hops_h

    # -- helpers -----------------------------------------------------------
    def _s(self) -> int:
        return _stream_handle(self.stream)

    def footprint_bytes(self) -> int:
        return self.lat.nbytes() + self.nets.nbytes() + self.ws.nbytes()

    _STATE = (("lat", ("hist", "occ", "wcost", "node_hist", "node_use", "via_col")),
              ("nets", ("path_arena", "path_len", "path_cost", "status")))

    def snapshot(self) -> dict:
        """Device copy of every array an iteration mutates, plus the host
        iteration counter (for re-running an iteration, e.g. to time it
        under other launch parameters)."""
        if self._pending is not None:
            raise RuntimeError("an iteration is in flight")
        snap = {"iteration": self.iteration, "history": len(self.history), "delta_mult": self.delta_mult}
        with torch.cuda.stream(self.stream):
            for owner, names in self._STATE:
                obj = getattr(self, owner)
                for n in names:
                    snap[(owner, n)] = getattr(obj, n).clone()
        return snap

    def restore(self, snap: dict) -> None:
        if self._pending is not None:
            raise RuntimeError("an iteration is in flight")
        with torch.cuda.stream(self.stream):
            for owner, names in self._STATE:
                obj = getattr(self, owner)
                for n in names:
                    getattr(obj, n).copy_(snap[(owner, n)])
        self.iteration = snap["iteration"]
        self.delta_mult = snap["delta_mult"]
        self._staged_for = -1
        del self.history[snap["history"]:]

    @property
    def delta(self) -> float:
        """Bucket width of the next launch: cfg.delta, widened by adapt_delta."""
        return float(self.cfg.delta) * self.delta_mult

    def route_cfg(self, flags: int = 0) -> _lib.OrxRouteCfg:
        c = self.cfg
        fl = flags | (_lib.ORX_F_STATS if c.collect_stats else 0) | \
            (_lib.ORX_F_ASYNC if c.engine == "tiles" and c.tile_queue else 0)
        return _lib.OrxRouteCfg(self.delta, int(c.block), int(c.lds_nodes), fl, self._route_stats.data_ptr(),
                                int(c.tile_groups))

    def _adapt_delta(self, st: "IterationStats") -> None:
        """Widen delta x4 (up to delta_max) after an iteration whose searches
        ran more than 8 threshold phases per net with rounds under a quarter
        of a workgroup: the weights have outgrown the bucket width (a
        congested board late in the negotiation), so a narrow bucket only
        yields thin bands of nodes per round.  Boards whose searches end in
        their first bucket (A* toward the sink) never trigger it.  Results do
        not depend on delta; only time does."""
        c = self.cfg
        nets = st.routed + st.failed
        if c.engine != "wave" or not c.adapt_delta or nets == 0 or st.rounds == 0 or math.isinf(c.delta):
            return
        if st.phases > 8 * nets and 4 * st.items < st.rounds * c.block and \
                float(c.delta) * self.delta_mult * 4.0 <= c.delta_max:
            self.delta_mult *= 4.0

    def _stage_net_overuse(self) -> None:
        """Enqueue the per-net overuse / hop counts of the committed paths and
        their copy into pinned host memory on this job's stream."""
        n = self.netlist.n
        with torch.cuda.stream(self.stream):
            _lib.call("orx_net_overuse", C.byref(self.lat.struct), C.byref(self.nets.struct),
                      self.cfg.edge_cap, self.cfg.node_cap, self._over.data_ptr(), self._node_over.data_ptr(),
                      self._hops.data_ptr(), self._s())
            self._over_h[:n].copy_(self._over[:n], non_blocking=True)
            self._node_over_h[:n].copy_(self._node_over[:n], non_blocking=True)
            self._hops_h[:n].copy_(self._hops[:n], non_blocking=True)

    def order_for(self, iteration: int) -> np.ndarray:
        """_order_nets_by_difficulty from the device occupancy.  The per-net
        overuse of the previous iteration is normally staged by that
        iteration's launch already (no extra round trip here)."""
        n = self.netlist.n
        if iteration > 1 and n:
            if self._staged_for != iteration:
                self._stage_net_overuse()
                self.stream.synchronize()
            over = self._over_h[:n].numpy().copy()
            hops = self._hops_h[:n].numpy().copy()
        else:
            over = np.zeros(n, np.int64)
            hops = np.zeros(n, np.int32)
        order = net_order(iteration, self.distance, over, hops)
        h = self.history
        stalled = len(h) >= 2 and h[-1].overuse >= h[-2].overuse
        if self.cfg.hotset and iteration > 1 and not stalled:
            order = hotset_filter(order, over, self._node_over_h[:n].numpy(), hops)
        return order

    def plan(self, order: np.ndarray, mode: str) -> list[np.ndarray]:
        """Launch groups: one per net (plain) or one per disjoint-ROI wave
        (native planner; inside a wave the largest ROI comes first)."""
        if mode == "plain":
            return [np.asarray([k], np.int32) for k in order]
        if mode not in ("batched", "waves"):
            raise ValueError("mode must be 'batched' ('waves') or 'plain'")
        waves = plan_waves(self.netlist.box, order, self.spec.X, self.spec.Y, self.nets.roi_nodes)
        return split_waves(waves, self.nets.roi_nodes, self.ws_cap())

    def _wave_steps(self, flat_nets: np.ndarray, wave_ptr: np.ndarray) -> np.ndarray:
        """Tile-engine step budget of each wave: its largest net's."""
        if len(flat_nets) == 0:
            return np.zeros(0, np.int32)
        per = self._tile_steps[flat_nets]
        return np.maximum.reduceat(per, wave_ptr[:-1]).astype(np.int32)

    def _route_calls(self, params, rcfg, flat_nets, wave_ptr, wave_max, wave_ws, d_nets, d_off) -> list:
        """The entry-point calls that route an iteration's waves, in order:
        orx_route_tiles for the tile engine, orx_route_waves for one
        workgroup per net.  On the tile engine, each run of consecutive waves
        whose largest ROI is at most cfg.small_wave_nodes goes to
        orx_route_waves (a launch or two per wave instead of a tile step
        sequence; every engine reaches the same fixpoint, so the result is
        the same)."""
        c = self.cfg
        if c.engine not in ("tiles", "wave"):
            raise ValueError("engine must be 'tiles' or 'wave'")
        nw = len(wave_ptr) - 1
        if nw <= 0:
            return []
        if c.engine == "wave" or self.concurrent_jobs > 1:
            small = np.ones(nw, bool)
        else:
            small = wave_max <= int(c.small_wave_nodes) if c.small_wave_nodes > 0 else np.zeros(nw, bool)
        wave_steps = self._wave_steps(flat_nets, wave_ptr) if not small.all() else None
        lat, pr, nets, ws, rc = (C.byref(self.lat.struct), C.byref(params), C.byref(self.nets.struct),
                                 C.byref(self.ws.struct), C.byref(rcfg))
        calls = []
        i = 0
        while i < nw:
            j = i + 1
            while j < nw and small[j] == small[i]:
                j += 1
            a = int(wave_ptr[i])
            ptr = np.ascontiguousarray(wave_ptr[i:j + 1] - a, np.int32)
            wws = np.ascontiguousarray(wave_ws[i:j], np.int64)
            dn, do = d_nets.data_ptr() + 4 * a, d_off.data_ptr() + 8 * a
            if small[i]:
                wmax = np.ascontiguousarray(wave_max[i:j], np.int64)
                calls.append(("orx_route_waves", (lat, pr, nets, dn, do, j - i, ptr.ctypes.data, wmax.ctypes.data,
                                                  wws.ctypes.data, ws, rc), (ptr, wmax, wws)))
            else:
                wst = np.ascontiguousarray(wave_steps[i:j], np.int32)
                calls.append(("orx_route_tiles", (lat, pr, nets, dn, do, j - i, ptr.ctypes.data, wws.ctypes.data,
                                                  wst.ctypes.data, ws, rc), (ptr, wws, wst)))
            i = j
        self._call_keep = calls      # host arrays stay alive while the calls run
        return [(e, a) for e, a, _ in calls]

    def ws_cap(self) -> int:
        """Workspace nodes one launch may use (at least the largest ROI)."""
        r64 = (int(self.nets.roi_nodes.max(initial=1)) + 63) // 64 * 64
        return max(int(self.cfg.ws_cap_nodes), r64)

    # -- one iteration -----------------------------------------------------
    def run_iteration(self, mode: str = "batched", order: np.ndarray | None = None,
                      waves: list[np.ndarray] | None = None, graph: bool | None = None) -> IterationStats:
        """One PathFinder iteration, synchronous."""
        return self.finish_iteration(self.launch_iteration(mode, order, waves, graph))

    def launch_iteration(self, mode: str = "batched", order: np.ndarray | None = None,
                         waves: list[np.ndarray] | None = None, graph: bool | None = None) -> "_Pending":
        """Enqueue one iteration on this job's stream and return without
        waiting for it: re-price, the waves (one launch each, or one captured
        hipGraph of them when ``graph``), then overuse + history.  The host
        work (order readback, planning, upload) happens here; the device work
        overlaps whatever other streams run.  finish_iteration collects it."""
        if self._pending is not None:
            raise RuntimeError("an iteration of this job is still in flight; call finish_iteration first")
        if self.iteration >= len(self.schedule):
            raise RuntimeError("schedule exhausted; raise cfg.iterations")
        if mode == "dag":
            raise RuntimeError("the dependency-driven launch (orx_route_dag) is switched off on the device")
        graph = self.cfg.graph if graph is None else graph
        it = self.iteration + 1
        pres = float(self.schedule[it - 1])
        params = self.cfg.params(pres, self.astar)
        stream = self.stream
        t0 = time.perf_counter()
        if order is None:
            order = self.order_for(it)
        # Split planning: the head of the order (the hardest nets, whose
        # waves take the most device time) is planned and launched first, and
        # the rest is planned while the device routes the head.  The rest's
        # waves start after the head's, so every net still comes after each
        # earlier net whose box it overlaps: the same results.
        rest = None
        if waves is None:
            h = int(self.cfg.head_nets)
            if mode in ("batched", "waves") and not graph and h > 0 and len(order) > 2 * h:
                waves, rest = self.plan(order[:h], mode), order[h:]
            else:
                waves = self.plan(order, mode)
        s = self._s()
        ev = tuple(torch.cuda.Event(enable_timing=True) for _ in range(4))
        d_nets, d_off = self._plan_nets, self._plan_off
        reserve = TILE_RESERVE if self.cfg.engine == "tiles" else 0

        def stage(part: list[np.ndarray], base: int):
            """Upload a part of the plan at offset ``base`` of the plan
            buffers; its launch calls."""
            flat_nets, flat_off, wave_ptr, wave_max, wave_ws, need = _flatten_waves(part, self.nets.roi_nodes)
            m = len(flat_nets)
            if base + m > len(self._plan_nets_h):
                raise ValueError("a launch plan lists more nets than the job has")
            self._plan_nets_h.numpy()[base:base + m] = flat_nets
            self._plan_off_h.numpy()[base:base + m] = flat_off
            self.ws.ensure(need + reserve)
            if m:
                d_nets[base:base + m].copy_(self._plan_nets_h[base:base + m], non_blocking=True)
                d_off[base:base + m].copy_(self._plan_off_h[base:base + m], non_blocking=True)
            return self._route_calls(params, rcfg, flat_nets, wave_ptr, wave_max, wave_ws, d_nets[base:],
                                     d_off[base:]), m

        n_waves = len(waves)
        with torch.cuda.stream(stream):
            self._route_stats.zero_()
            ev[0].record(stream)
            self.lat.reprice(params, stream)
            rcfg = self.route_cfg()
            ev[1].record(stream)
            exec_graph = None
            calls, m = stage(waves, 0)
            if graph and len(waves):
                cs = C.c_void_p()
                _lib.call("orx_graph_begin", C.byref(cs))
                g = C.c_void_p()
                try:
                    for entry, route_args in calls:
                        _lib.call(entry, *route_args, cs)
                except Exception:
                    # end the capture (which frees the stream) and drop the graph
                    if _lib.load().orx_graph_end(cs, C.byref(g)) == _lib.ORX_OK:
                        _lib.load().orx_graph_destroy(g)
                    raise
                _lib.call("orx_graph_end", cs, C.byref(g))
                exec_graph = g
                try:
                    _lib.call("orx_graph_launch", exec_graph, s)
                except Exception:
                    _lib.load().orx_graph_destroy(exec_graph)
                    raise
            else:
                for entry, route_args in calls:
                    _lib.call(entry, *route_args, s)
                if rest is not None:
                    tail = self.plan(rest, mode)
                    calls, m2 = stage(tail, m)
                    for entry, route_args in calls:
                        _lib.call(entry, *route_args, s)
                    n_waves += len(tail)
                    m += m2
            ev[2].record(stream)
            _lib.call("orx_lattice_end_iteration", C.byref(self.lat.struct), C.byref(params),
                      self._iter_out.data_ptr(), s)
            ev[3].record(stream)
            self._iter_out_h.copy_(self._iter_out, non_blocking=True)
            self._route_stats_h.copy_(self._route_stats, non_blocking=True)
            if self.netlist.n and it < len(self.schedule):
                self._stage_net_overuse()     # input of the next iteration's order
            done = torch.cuda.Event()
            done.record(stream)
        self._pending = _Pending(it, pres, n_waves, t0, ev, done, exec_graph, (d_nets, d_off), m)
        return self._pending

    def ready(self) -> bool:
        """True when the iteration in flight (if any) has finished on the device."""
        return self._pending is None or self._pending.done.query()

    def finish_iteration(self, pending: "_Pending | None" = None) -> IterationStats:
        p = pending or self._pending
        if p is None or p is not self._pending:
            raise RuntimeError("no iteration of this job in flight")
        p.done.synchronize()
        if p.graph is not None:
            _lib.call("orx_graph_destroy", p.graph)
        dt = time.perf_counter() - p.t0
        out = self._iter_out_h.numpy().copy()
        rs = self._route_stats_h.numpy().copy()
        ev0, evr0, evr1, ev1 = p.events
        st = IterationStats(
            iteration=p.it, pres_fac=p.pres, routed=int(rs[5]), failed=int(rs[6]),
            edge_over=int(out[0]), edge_over_count=int(out[1]), node_over=int(out[2]),
            node_over_count=int(out[3]), via_hops=int(out[4]), barrels=int(out[5]),
            settled=int(rs[0]), relaxed=int(rs[1]), items=int(rs[2]), rounds=int(rs[3]), phases=int(rs[4]),
            path_hops=int(rs[7]), waves=p.waves, seconds=dt, gpu_ms=ev0.elapsed_time(ev1),
            route_ms=evr0.elapsed_time(evr1), extra={"graph": p.graph is not None, "launched_nets": p.nets})
        self._pending = None
        self.iteration = p.it
        self._staged_for = p.it + 1 if self.netlist.n and p.it < len(self.schedule) else -1
        self.history.append(st)
        self._adapt_delta(st)
        return st

    @property
    def converged(self) -> bool:
        """Zero overuse and no failed net after the last iteration."""
        return bool(self.history) and self.history[-1].overuse == 0 and self.history[-1].failed == 0

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
        self._adapt_delta(st)
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
        with torch.cuda.stream(self.stream):
            return self.nets.paths()

    def digest(self) -> str:
        """SHA-256 over every net's path (node ids), its fp32 cost and the
        final occupancy: equal digests mean bit-identical results."""
        import hashlib
        h = hashlib.sha256()
        with torch.cuda.stream(self.stream):
            lens = self.nets.path_len.cpu().numpy()
            arena = self.nets.path_arena.cpu().numpy()
            cost = self.nets.path_cost.cpu().numpy()
            occ = self.lat.occ.cpu().numpy()
            use = self.lat.node_use.cpu().numpy()
        for o, l in zip(self.nets.path_off_np, lens):
            h.update(arena[o:o + l].tobytes())
            h.update(b"|")
        h.update(cost.tobytes())
        h.update(occ.tobytes())
        h.update(use.tobytes())
        return h.hexdigest()

    def path_costs(self) -> np.ndarray:
        with torch.cuda.stream(self.stream):
            return self.nets.path_cost.cpu().numpy()

    def status(self) -> np.ndarray:
        with torch.cuda.stream(self.stream):
            return self.nets.status.cpu().numpy()
