"""Net ordering and disjoint-ROI wave colouring.

Net order per iteration restates _order_nets_by_difficulty
(unified_pathfinder.py:7840-7901): difficulty = Manhattan distance (x, y, z)
* (pin_degree + 1) * (congestion + 1), congestion = mean overuse over the
net's previously committed arcs, hardest first with ties by net id
(descending), then a rotation by random.Random(42 + iteration).randint(0,
min(5, n // 10)).  Every net is rerouted every iteration.

Wave colouring: a net's wave is 1 + the largest wave of any EARLIER net (in
this iteration's order) whose ROI box overlaps its own.  Nets of one wave have
pairwise disjoint boxes and every overlapping pair keeps its sequential
order, so routing wave by wave (nets of a wave in parallel) gives exactly the
sequential result: a net only reads and writes arcs and nodes inside its own
box.
"""
from __future__ import annotations

import random

import numpy as np


def resolve_pres_fac_max(L: int, configured: float = 64.0) -> float:
    """Layer-dependent ceiling (unified_pathfinder.py:801-816) over L-2 signal layers."""
    signal = L - 2
    floor = 32.0 if signal <= 12 else (64.0 if signal <= 20 else 128.0)
    return max(float(configured), floor)


def pres_schedule(iters: int, init: float = 0.5, mult: float = 1.5, pmax: float = 64.0) -> np.ndarray:
    """fp32 present factors per iteration: p1 = init, p(k+1) = min(fl(p(k)*mult), pmax)."""
    out = np.empty(iters, np.float32)
    p = np.float32(init)
    for i in range(iters):
        out[i] = p
        p = np.minimum(np.float32(p * np.float32(mult)), np.float32(pmax))
    return out


def net_distance(spec, src: np.ndarray, dst: np.ndarray) -> np.ndarray:
    sx, sy, sz = spec.coords(src)
    tx, ty, tz = spec.coords(dst)
    return (np.abs(tx - sx) + np.abs(ty - sy) + np.abs(tz - sz)).astype(np.int64)


def net_order(iteration: int, distance: np.ndarray, over_sum: np.ndarray, hops: np.ndarray) -> np.ndarray:
    """Routing order of one iteration (1-based), hardest first."""
    n = len(distance)
    congestion = over_sum.astype(np.float64) / np.maximum(1, hops).astype(np.float64)
    difficulty = (distance * 3).astype(np.float64) * (congestion + 1.0)
    ids = np.arange(n)
    order = np.lexsort((ids, difficulty))[::-1]  # descending by (difficulty, id)
    rot = random.Random(42 + int(iteration)).randint(0, min(5, n // 10))
    if rot > 0:
        order = np.concatenate([order[rot:], order[:rot]])
    return order.astype(np.int32)


def color_waves(box: np.ndarray, order: np.ndarray, X: int, Y: int) -> np.ndarray:
    """Wave index (0-based) of every net in ``order`` (-1 for the others)."""
    stamp = np.full((Y, X), -1, dtype=np.int32)
    wave = np.full(len(box), -1, np.int32)
    for k in order:
        x0, y0, w, h = (int(v) for v in box[k])
        win = stamp[y0:y0 + h, x0:x0 + w]
        wk = int(win.max()) + 1
        wave[k] = wk
        win[...] = wk
    return wave


def group_waves(wave: np.ndarray, order: np.ndarray, roi_nodes: np.ndarray):
    """Lists of net ids per wave; inside a wave the largest ROI comes first
    (long jobs start first on the device)."""
    n_waves = int(wave[order].max()) + 1 if len(order) else 0
    buckets = [[] for _ in range(n_waves)]
    for k in order:
        buckets[wave[k]].append(int(k))
    out = []
    for b in buckets:
        b = np.asarray(b, np.int32)
        b = b[np.argsort(-roi_nodes[b], kind="stable")]
        out.append(b)
    return out


def check_disjoint(box: np.ndarray, nets: np.ndarray) -> bool:
    """True when the ROI boxes of ``nets`` are pairwise disjoint."""
    b = box[nets].astype(np.int64)
    x0, y0 = b[:, 0], b[:, 1]
    x1, y1 = x0 + b[:, 2], y0 + b[:, 3]
    ox = (x0[:, None] < x1[None, :]) & (x0[None, :] < x1[:, None])
    oy = (y0[:, None] < y1[None, :]) & (y0[None, :] < y1[:, None])
    o = ox & oy
    np.fill_diagonal(o, False)
    return not bool(o.any())


def plan_iteration(box: np.ndarray, order: np.ndarray, X: int, Y: int, roi_nodes: np.ndarray):
    """Native planner (orx_plan_iteration in liborx_hip.so): returns
    (wave[n], sched[n], dep_ptr[n+1], dep_ids[]) for the dependency-driven
    launch.  Runs on the host only."""
    import ctypes as C

    from . import _lib
    lib = _lib.load()
    n = len(order)
    box = np.ascontiguousarray(box.reshape(-1), np.int32)
    if len(box) != 4 * n:
        raise ValueError("plan_iteration needs an order of every net (a permutation)")
    order = np.ascontiguousarray(order, np.int32)
    rn = np.ascontiguousarray(roi_nodes, np.int64)
    wave = np.empty(n, np.int32)
    sched = np.empty(n, np.int32)
    dep_ptr = np.empty(n + 1, np.int32)
    cap = max(16, 8 * n)
    while True:
        dep_ids = np.empty(cap, np.int32)
        total = lib.orx_plan_iteration(box.ctypes.data, order.ctypes.data, n, X, Y, rn.ctypes.data,
                                       wave.ctypes.data, sched.ctypes.data, dep_ptr.ctypes.data,
                                       dep_ids.ctypes.data, cap)
        if total < 0:
            raise ValueError(f"orx_plan_iteration rejected its input ({total})")
        if total <= cap:
            return wave, sched, dep_ptr, dep_ids[:total]
        cap = int(total)


def plan_waves(box: np.ndarray, order: np.ndarray, X: int, Y: int, roi_nodes: np.ndarray) -> list[np.ndarray]:
    """The disjoint-ROI waves of one iteration from the native planner: the
    same waves and in-wave order as group_waves(color_waves(...)), computed
    in C++ on a tiled grid (orx_plan_waves)."""
    from . import _lib
    lib = _lib.load()
    n = len(order)
    if n == 0:
        return []
    box = np.ascontiguousarray(box.reshape(-1), np.int32)
    n_nets = len(box) // 4
    order = np.ascontiguousarray(order, np.int32)
    rn = np.ascontiguousarray(roi_nodes, np.int64)
    if len(rn) != n_nets:
        raise ValueError("roi_nodes and box describe different net counts")
    wave = np.full(n_nets, -1, np.int32)
    sched = np.empty(n, np.int32)
    nw = lib.orx_plan_waves(box.ctypes.data, n_nets, order.ctypes.data, n, X, Y, rn.ctypes.data,
                            wave.ctypes.data, sched.ctypes.data)
    if nw < 0:
        raise ValueError(f"orx_plan_waves rejected its input ({nw})")
    ws = wave[sched]
    cuts = np.flatnonzero(np.diff(ws)) + 1
    return np.split(sched, cuts)


def check_schedule(sched: np.ndarray, dep_ptr: np.ndarray, dep_ids: np.ndarray) -> None:
    """Every dependency must come earlier in the schedule (the launch's
    progress guarantee rests on it)."""
    n = len(sched)
    pos = np.full(n, -1, np.int64)
    pos[sched] = np.arange(n)
    if (pos < 0).any():
        raise ValueError("schedule is not a permutation of the nets")
    owner = np.repeat(np.arange(n), np.diff(dep_ptr))
    if len(dep_ids) and not (pos[dep_ids] < pos[owner]).all():
        raise ValueError("a dependency is scheduled after its dependant")
