"""Python side of the CPU parity oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker; the product path (orthoroute_amd) never
imports it.  It wraps liborx_oracle.so (oracle/orx_oracle.c, a C restatement
of the reference CPU path; see that file's header for the reference lines it
follows) and adds pure-Python twins for tiny cases.

Parity status of each oracle function (how it is pinned, see tests/):
  sssp_box / pathfinder SSSP  -- pinned: brute-force Bellman-Ford fixpoint on
      tiny lattices, scipy.sparse.csgraph.dijkstra (float64, exact for dyadic
      costs), closed-form uncongested distances, fixpoint invariants, and the
      pure-Python heapq twin below;
  sssp_csr -- pinned: the reference's own vectors (tests/test_gpu_fallback.py:
      seed pricing [1, 3], fp32 absorption chain [3, 2, 1, 0], random 6x6 grid
      seed 20260727 vs float64 heapq, ROI-CSR folded weights [7.0, 3.0]);
  accountant (commit / clear / overuse / history / cost) -- pinned:
      tests/unit/test_edge_accountant.py, tests/test_via_accounting.py,
      tests/test_lattice.py:255-276 vectors, restated;
  iteration host logic (difficulty order, present-factor schedule, A* bounds,
      wave colouring) -- restated below, not imported from the product;
      pinned by hand-worked orders (the congestion term included), exact
      fp32 schedules, A*-off runs giving the same state, and the product's
      own functions checked against these (tests/test_oracle_order.py).
"""
from __future__ import annotations

import ctypes as C
import heapq
import random
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liborx_oracle.so"

KEY_INF = 0xFFFFFFFFFFFFFFFF
ROOT_SLOT = 0xFF


class OracleParams(C.Structure):
    _fields_ = [
        ("pres_fac", C.c_float), ("base_w", C.c_float), ("hist_w", C.c_float),
        ("node_pres", C.c_float), ("node_hist_w", C.c_float),
        ("edge_cap", C.c_int32), ("node_cap", C.c_int32),
        ("hist_gain", C.c_float), ("hist_decay", C.c_float),
        ("hist_cap_mult", C.c_float), ("node_hist_gain", C.c_float),
        ("astar_cp", C.c_float), ("astar_cv", C.c_float), ("astar_gmax", C.c_float),
    ]


class OracleTerms(C.Structure):
    """Seeds and targets of multi-terminal nets (orx_terms in orx_oracle.c)."""
    _fields_ = [("seed_ptr", C.c_void_p), ("seed_node", C.c_void_p), ("seed_cost", C.c_void_p),
                ("tgt_ptr", C.c_void_p), ("tgt_node", C.c_void_p), ("tgt_cost", C.c_void_p)]


def terms_struct(terms):
    """(ctypes struct or None, arrays kept alive) for a netlist's Terminals."""
    if terms is None:
        return None, ()
    keep = (np.ascontiguousarray(terms.seed_ptr, np.int32), np.ascontiguousarray(terms.seed_node, np.int32),
            np.ascontiguousarray(terms.seed_cost, np.float32), np.ascontiguousarray(terms.tgt_ptr, np.int32),
            np.ascontiguousarray(terms.tgt_node, np.int32), np.ascontiguousarray(terms.tgt_cost, np.float32))
    return OracleTerms(*(a.ctypes.data for a in keep)), keep


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    P = C.c_void_p
    lib.orx_oracle_pathfinder.restype = C.c_int
    lib.orx_oracle_pathfinder.argtypes = [
        C.c_int32, C.c_int32, C.c_int32, C.c_uint32, P, C.c_int32, P, P, P, C.c_int32, P, P,
        C.POINTER(OracleParams), P, P, P, P, P, P, P, P, P, P, P, C.c_int64, C.POINTER(OracleTerms)]
    lib.orx_oracle_pathfinder_waves.restype = C.c_int
    lib.orx_oracle_pathfinder_waves.argtypes = [
        C.c_int32, C.c_int32, C.c_int32, C.c_uint32, P, C.c_int32, P, P, P, P, P, C.c_int32, C.c_float,
        C.POINTER(OracleParams), P, P, P, P, P, P, P, P, P, P, P, C.c_int32, C.POINTER(OracleTerms)]
    lib.orx_oracle_sssp_box.restype = C.c_int
    lib.orx_oracle_sssp_box.argtypes = [
        C.c_int32, C.c_int32, C.c_int32, C.c_uint32, P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
        C.c_int64, C.c_int64, C.c_int32, P, P, P, P]
    lib.orx_oracle_sssp_csr.restype = C.c_int
    lib.orx_oracle_sssp_csr.argtypes = [
        C.c_int64, P, P, P, P, P, P, C.c_int64, P, P, P, P, C.c_int32, C.c_int32, P, P, C.c_int32, P, P, P, P, P,
        P]
    lib.orx_oracle_overuse.restype = None
    lib.orx_oracle_overuse.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, C.c_int32, C.c_int32, P]
    lib.orx_oracle_update_history.restype = None
    lib.orx_oracle_update_history.argtypes = [C.c_int64, P, P, P, C.c_int64, P, P, C.POINTER(OracleParams)]
    lib.orx_oracle_lattice_weights.restype = None
    lib.orx_oracle_lattice_weights.argtypes = [
        C.c_int32, C.c_int32, C.c_int32, C.c_uint32, P, P, P, P, P, C.POINTER(OracleParams), P]
    lib.orx_oracle_apply_path.restype = C.c_int
    lib.orx_oracle_apply_path.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32, P, C.c_int64, C.c_int32, P,
                                          P, P]
    lib.orx_oracle_net_overuse.restype = None
    lib.orx_oracle_net_overuse.argtypes = [
        C.c_int32, C.c_int32, C.c_int32, C.c_uint32, P, C.c_int32, C.c_int32, P, P, P, P, P]
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


def params_from(cfg, pres: float) -> OracleParams:
    p = np.float32(pres)
    return OracleParams(float(p), cfg.base_w, cfg.hist_w, float(np.float32(cfg.node_pres_w) * p),
                        cfg.node_hist_w, cfg.edge_cap, cfg.node_cap, cfg.hist_gain, cfg.hist_decay,
                        cfg.hist_cap_mult, cfg.node_hist_gain)


# -- host logic of one iteration, restated from the reference ----------------

def lattice_distance(X: int, Y: int, src, dst) -> np.ndarray:
    """Manhattan distance in lattice coordinates between each net's source and
    sink (unified_pathfinder.py:7851-7854; node = (z*Y + y)*X + x)."""
    out = np.zeros(len(src), np.int64)
    for k, (s, t) in enumerate(zip(src, dst)):
        sz, r = divmod(int(s), X * Y)
        sy, sx = divmod(r, X)
        tz, r = divmod(int(t), X * Y)
        ty, tx = divmod(r, X)
        out[k] = abs(tx - sx) + abs(ty - sy) + abs(tz - sz)
    return out


def difficulty_order(iteration: int, distance, over_sum, hops) -> np.ndarray:
    """Hardest-first routing order of iteration ``iteration`` (1-based),
    _order_nets_by_difficulty (unified_pathfinder.py:7840-7900) without barrel
    roles: difficulty = distance x (pin degree 2 + 1) x (congestion + 1), where
    congestion is the mean overuse max(0, present - cap) over the arcs of the
    net's previous path (``over_sum`` / ``hops``; 0 with no path); sorted on
    (difficulty, net id) descending, then rotated left by
    random.Random(42 + iteration).randint(0, min(5, n // 10))."""
    scores = []
    for k in range(len(distance)):
        congestion = float(over_sum[k]) / max(1, int(hops[k]))
        scores.append((int(distance[k]) * (2 + 1) * (congestion + 1), k))
    scores.sort(reverse=True)
    ordered = [k for _, k in scores]
    rotation = random.Random(42 + int(iteration)).randint(0, min(5, len(scores) // 10))
    if rotation > 0:
        ordered = ordered[rotation:] + ordered[:rotation]
    return np.asarray(ordered, np.int32)


def present_factor_max(L: int, configured: float = 64.0) -> float:
    """resolve_pres_fac_max (unified_pathfinder.py:801-814) over the L - 2
    signal layers: the configured ceiling, at least 32 / 64 / 128 for up to
    12 / 20 / more signal layers."""
    signal = L - 2
    floor = 32.0 if signal <= 12 else 64.0 if signal <= 20 else 128.0
    return max(float(configured), floor)


def present_factors(iterations: int, init: float, mult: float, pmax: float) -> np.ndarray:
    """fp32 present factor of each iteration: init, then min(p x mult, pmax)
    (unified_pathfinder.py:7379-7383 at work scale 1)."""
    out = []
    p = np.float32(init)
    for _ in range(iterations):
        out.append(p)
        p = min(np.float32(p * np.float32(mult)), np.float32(pmax))
    return np.asarray(out, np.float32)


def astar_lower_bounds(base: np.ndarray, base_w: float, margin_bits: int = 8) -> tuple[float, float, float]:
    """(cp, cv, gmax) of the sink-directed search: the cheapest finite planar
    (slots 1, 2) and via (slots 0, 3) entering weight fl(b x base_w), shrunk
    by 2^-margin_bits and rounded down in fp32, and the distance below which
    that margin covers half an fp32 ulp, 2^(24 + floor(log2 margin)).  Paths
    do not depend on them (an A*-off run gives the same state); the search
    counts do."""
    b = np.asarray(base, np.float32).reshape(-1, 4)
    bw = np.float32(base_w)
    mins = []
    for cols in ((1, 2), (0, 3)):
        v = b[:, cols].ravel()
        v = v[np.isfinite(v)]
        mins.append(float(np.float32(v.min()) * bw) if v.size else None)
    if mins[0] is None:
        return (0.0, 0.0, 0.0)
    if mins[1] is None:
        mins[1] = mins[0]
    if min(mins) <= 0:
        return (0.0, 0.0, 0.0)
    shrunk = []
    for m in mins:
        f = np.float32(m * (1.0 - 2.0 ** -margin_bits))
        while float(f) > m * (1.0 - 2.0 ** -margin_bits):
            f = np.nextafter(f, np.float32(0))
        shrunk.append(f)
    margin = min(m - float(f) for m, f in zip(mins, shrunk))
    gmax = 2.0 ** min(24 + int(np.floor(np.log2(margin))), 120)
    return (float(shrunk[0]), float(shrunk[1]), float(gmax))


def disjoint_waves(box: np.ndarray, order: np.ndarray) -> list[np.ndarray]:
    """Nets of ``order`` in waves of pairwise disjoint boxes: each net goes one
    wave after the latest earlier net whose box overlaps its own, so nets
    that share cells keep their order and a wave's nets can route in any
    order.  Inside a wave nets keep the routing order."""
    b = np.asarray(box, np.int64).reshape(-1, 4)
    order = np.asarray(order, np.int64)
    wave = np.full(len(order), -1, np.int64)
    x0, y0 = b[order, 0], b[order, 1]
    x1, y1 = x0 + b[order, 2], y0 + b[order, 3]
    for i in range(len(order)):
        hit = (x0[:i] < x1[i]) & (x0[i] < x1[:i]) & (y0[:i] < y1[i]) & (y0[i] < y1[:i])
        wave[i] = int(wave[:i][hit].max()) + 1 if hit.any() else 0
    n_waves = int(wave.max()) + 1 if len(order) else 0
    return [order[wave == w].astype(np.int32) for w in range(n_waves)]


class OraclePathFinder:
    """Sequential CPU PathFinder over box ROIs.  Its host logic (order,
    present factors, A* bounds) is the restatement above, not the product's."""

    def __init__(self, netlist, cfg, base: np.ndarray | None = None):
        self.lib = load()
        self.netlist = netlist
        self.cfg = cfg
        spec = netlist.spec
        self.spec = spec
        N = spec.N
        self.base = (spec.base_costs() if base is None else np.asarray(base, np.float32)).copy()
        self.hist = np.zeros(4 * N, np.float32)
        self.occ = np.zeros(4 * N, np.int32)
        self.node_hist = np.zeros(N, np.float32)
        self.node_use = np.zeros(N, np.int32)
        self.via_col = np.zeros(spec.XY, np.int32)
        R = netlist.roi_nodes().astype(np.int64)
        self.path_cap = R
        self.path_off = np.zeros(netlist.n, np.int64)
        if netlist.n:
            self.path_off[1:] = np.cumsum(R)[:-1]
        self.arena = np.zeros(max(1, int(R.sum())), np.int32)
        self.path_len = np.zeros(netlist.n, np.int32)
        self.path_cost = np.full(netlist.n, np.inf, np.float32)
        self.box = np.ascontiguousarray(netlist.box.reshape(-1), np.int32)
        self.src = netlist.src.astype(np.int32)
        self.dst = netlist.dst.astype(np.int32)
        self.distance = lattice_distance(spec.X, spec.Y, netlist.src, netlist.dst)
        pmax = cfg.pres_max if cfg.pres_max is not None else present_factor_max(spec.L)
        self.schedule = present_factors(cfg.iterations, cfg.pres_init, cfg.pres_mult, pmax)
        weights = (cfg.base_w, cfg.hist_w, cfg.node_pres_w, cfg.node_hist_w, cfg.hist_gain, cfg.node_hist_gain,
                   cfg.hist_decay)
        astar_ok = cfg.astar and min(weights) >= 0 and cfg.pres_init >= 0 and cfg.pres_mult >= 0
        self.astar = astar_lower_bounds(self.base, cfg.base_w) if astar_ok else (0.0, 0.0, 0.0)
        self.iteration = 0
        self.history = []
        self.terms, self._terms_keep = terms_struct(getattr(netlist, "terms", None))

    def net_overuse(self):
        n = self.netlist.n
        over = np.zeros(n, np.int64)
        hops = np.zeros(n, np.int32)
        s = self.spec
        self.lib.orx_oracle_net_overuse(s.X, s.Y, s.L, s.hmask, _p(self.occ), self.cfg.edge_cap, n,
                                        _p(self.arena), _p(self.path_off), _p(self.path_len), _p(over), _p(hops))
        return over, hops

    def order_for(self, iteration: int) -> np.ndarray:
        if iteration > 1:
            over, hops = self.net_overuse()
        else:
            over = np.zeros(self.netlist.n, np.int64)
            hops = np.zeros(self.netlist.n, np.int32)
        order = difficulty_order(iteration, self.distance, over, hops)
        h = self.history
        stalled = len(h) >= 2 and (h[-1]["edge_over"] + h[-1]["node_over"] >= h[-2]["edge_over"] + h[-2]["node_over"])
        if getattr(self.cfg, "hotset", False) and iteration > 1 and not stalled:
            # offenders only (_build_hotset, unified_pathfinder.py:9143-9171):
            # a committed path on an over-capacity edge or node, or no path;
            # every net after an iteration that did not lower the overuse
            node_over = np.array([int(np.maximum(self.node_use[p] - self.cfg.node_cap, 0).sum())
                                  for p in self.paths()], np.int64)
            keep = [k for k in order if over[k] > 0 or node_over[k] > 0 or hops[k] == 0]
            order = np.asarray(keep, np.int32)
        return order

    def run_iteration(self, order: np.ndarray | None = None) -> dict:
        it = self.iteration + 1
        if order is None:
            order = self.order_for(it)
        s = self.spec
        pres = np.asarray([self.schedule[it - 1]], np.float32)
        order = np.ascontiguousarray(order, np.int32)
        stats = np.zeros(16, np.int64)
        P = params_from(self.cfg, 1.0)
        P.node_pres = float(np.float32(self.cfg.node_pres_w))  # scaled by pres inside
        P.astar_cp, P.astar_cv, P.astar_gmax = self.astar
        rc = self.lib.orx_oracle_pathfinder(
            s.X, s.Y, s.L, s.hmask, _p(self.base), self.netlist.n, _p(self.src), _p(self.dst), _p(self.box),
            1, _p(order), _p(pres), C.byref(P), _p(self.hist), _p(self.occ), _p(self.node_hist),
            _p(self.node_use), _p(self.via_col), _p(self.arena), _p(self.path_off), _p(self.path_cap),
            _p(self.path_len), _p(self.path_cost), _p(stats), len(order),
            None if self.terms is None else C.byref(self.terms))
        if rc < 0:
            raise RuntimeError(f"oracle pathfinder failed: {rc}")
        self.iteration = it
        keys = ["routed", "failed", "edge_over", "edge_over_count", "node_over", "node_over_count", "via_hops",
                "barrels", "settled", "relaxed", "path_hops"]
        d = {k: int(stats[i]) for i, k in enumerate(keys)}
        d["iteration"] = it
        d["pres_fac"] = float(pres[0])
        self.history.append(d)
        return d

    def run_iteration_waves(self, threads: int = 0) -> dict:
        """The same iteration as run_iteration, with the nets of each
        disjoint-ROI wave routed on `threads` host threads (0 = all); the
        result equals the sequential order's."""
        it = self.iteration + 1
        order = self.order_for(it)
        s = self.spec
        roi = self.netlist.roi_nodes()
        # largest ROI first inside a wave: the host threads start the long searches first
        waves = [w[np.argsort(-roi[w], kind="stable")] for w in disjoint_waves(self.netlist.box, order)]
        sched = np.ascontiguousarray(np.concatenate(waves) if waves else np.zeros(0, np.int32), np.int32)
        wave_ptr = np.zeros(len(waves) + 1, np.int32)
        np.cumsum([len(w) for w in waves], out=wave_ptr[1:])
        stats = np.zeros(16, np.int64)
        P = params_from(self.cfg, 1.0)
        P.node_pres = float(np.float32(self.cfg.node_pres_w))
        P.astar_cp, P.astar_cv, P.astar_gmax = self.astar
        rc = self.lib.orx_oracle_pathfinder_waves(
            s.X, s.Y, s.L, s.hmask, _p(self.base), self.netlist.n, _p(self.src), _p(self.dst), _p(self.box),
            _p(sched), _p(wave_ptr), len(waves), float(self.schedule[it - 1]), C.byref(P), _p(self.hist),
            _p(self.occ), _p(self.node_hist), _p(self.node_use), _p(self.via_col), _p(self.arena),
            _p(self.path_off), _p(self.path_cap), _p(self.path_len), _p(self.path_cost), _p(stats), int(threads),
            None if self.terms is None else C.byref(self.terms))
        if rc < 0:
            raise RuntimeError(f"oracle wave pathfinder failed: {rc}")
        self.iteration = it
        keys = ["routed", "failed", "edge_over", "edge_over_count", "node_over", "node_over_count", "via_hops",
                "barrels", "settled", "relaxed", "path_hops"]
        d = {k: int(stats[i]) for i, k in enumerate(keys)}
        d["iteration"] = it
        d["pres_fac"] = float(self.schedule[it - 1])
        d["waves"] = len(waves)
        self.history.append(d)
        return d

    def paths(self) -> list[np.ndarray]:
        return [self.arena[o:o + l].copy() for o, l in zip(self.path_off, self.path_len)]

    def state(self) -> dict:
        return {"hist": self.hist, "occ": self.occ, "node_hist": self.node_hist, "node_use": self.node_use,
                "via_col": self.via_col}


def apply_path(spec, path, sign: int, occ: np.ndarray, node_use: np.ndarray, via_col: np.ndarray) -> None:
    """Commit (+1) or rip up (-1) one path in place (UP:10034-10046,
    4635-4659, 4500-4512)."""
    p = np.ascontiguousarray(path, np.int32)
    rc = load().orx_oracle_apply_path(spec.X, spec.Y, spec.L, spec.hmask, _p(p), len(p), int(sign), _p(occ),
                                      _p(node_use), _p(via_col))
    if rc != 0:
        raise ValueError(f"path is not a lattice walk ({rc})")


def overuse(spec, occ, node_use, via_col, edge_cap: int = 1, node_cap: int = 1) -> dict:
    """Iteration-end counts over canonical arcs (UP:1137-1190, 4919-4932)."""
    out = np.zeros(6, np.int64)
    load().orx_oracle_overuse(spec.X, spec.Y, spec.L, _p(occ), _p(node_use), _p(via_col), edge_cap, node_cap,
                              _p(out))
    return dict(zip(("edge_over", "edge_over_count", "node_over", "node_over_count", "via_hops", "barrels"),
                    (int(v) for v in out)))


def update_history(base, hist, occ, node_hist, node_use, params: OracleParams) -> None:
    """h = fl(h*decay) + min(gain*over, cap_mult*base), in place (UP:1226-1289)."""
    load().orx_oracle_update_history(len(hist), _p(base), _p(hist), _p(occ), len(node_hist), _p(node_hist),
                                     _p(node_use), C.byref(params))


def lattice_weights(spec, base, hist, occ, node_hist, node_use, params: OracleParams) -> np.ndarray:
    lib = load()
    out = np.empty(4 * spec.N, np.float32)
    lib.orx_oracle_lattice_weights(spec.X, spec.Y, spec.L, spec.hmask, _p(base), _p(hist), _p(occ),
                                   _p(node_hist), _p(node_use), C.byref(params), _p(out))
    return out


def sssp_box(spec, wcost: np.ndarray, box, src: int, dst: int = -1, full: bool = True):
    """One box query on precomputed weights: (dist, parent, hops, settled, relaxed)."""
    lib = load()
    x0, y0, w, h = (int(v) for v in box)
    R = w * h * spec.L
    dist = np.empty(R, np.float32)
    parent = np.empty(R, np.int32)
    hops = np.empty(R, np.int32)
    stats = np.zeros(2, np.int64)
    wc = np.ascontiguousarray(wcost, np.float32)
    rc = lib.orx_oracle_sssp_box(spec.X, spec.Y, spec.L, spec.hmask, _p(wc), x0, y0, w, h, int(src), int(dst),
                                 1 if full else 0, _p(dist), _p(parent), _p(hops), _p(stats))
    if rc < 0:
        raise RuntimeError(f"oracle sssp_box failed: {rc}")
    return dist, parent, hops, int(stats[0]), int(stats[1])


def sssp_csr(N, indptr, indices, costs, in_rank, in_indptr, in_src, roi_nodes=None, g2r=None, pen=None,
             allowed=None, seeds=(), seed_cost=None, targets=(), tg_cost=None, seed_pen=True, counts=False):
    """Generic CSR query: returns (rc, best_target, dist[R], parent_roi[R]),
    and with ``counts`` also (settled, relaxed): the nodes the heap search
    settled and the arcs it relaxed.  seed_pen: seeds pay their own node
    penalty (multi-seed forms); the single-source find_path_roi form passes
    False."""
    lib = load()
    a = lambda x, t: None if x is None else np.ascontiguousarray(x, t)  # noqa: E731
    indptr, indices = a(indptr, np.int32), a(indices, np.int32)
    costs, in_rank = a(costs, np.float32), a(in_rank, np.int32)
    in_indptr, in_src = a(in_indptr, np.int32), a(in_src, np.int32)
    roi_nodes, g2r, pen = a(roi_nodes, np.int32), a(g2r, np.int32), a(pen, np.float32)
    allowed = a(allowed, np.uint32)
    seeds, targets = a(seeds, np.int32), a(targets, np.int32)
    seed_cost, tg_cost = a(seed_cost, np.float32), a(tg_cost, np.float32)
    R = N if roi_nodes is None else len(roi_nodes)
    dist = np.empty(R, np.float32)
    parent = np.empty(R, np.int32)
    best = np.full(1, -1, np.int32)
    stats = np.zeros(2, np.int64)
    rc = lib.orx_oracle_sssp_csr(N, _p(indptr), _p(indices), _p(costs), _p(in_rank), _p(in_indptr), _p(in_src),
                                 R, _p(roi_nodes), _p(g2r), _p(pen), _p(allowed), 1 if seed_pen else 0,
                                 len(seeds), _p(seeds),
                                 _p(seed_cost), len(targets), _p(targets), _p(tg_cost), _p(dist), _p(parent),
                                 _p(best), _p(stats))
    if rc < 0:
        raise RuntimeError(f"oracle sssp_csr failed: {rc}")
    if counts:
        return rc, int(best[0]), dist, parent, (int(stats[0]), int(stats[1]))
    return rc, int(best[0]), dist, parent


def csr_path(best: int, parent_roi: np.ndarray, roi_nodes=None, g2r=None) -> list[int]:
    """Walk ROI-local parents from the chosen target (global ids, root first)."""
    if best < 0:
        return []
    r = int(g2r[best]) if g2r is not None else best
    out = []
    while r >= 0:
        out.append(int(roi_nodes[r]) if roi_nodes is not None else r)
        r = int(parent_roi[r])
    return out[::-1]


# ---------------------------------------------------------------------------
# Pure-Python twin (tiny cases only): heapq Dijkstra with the canonical key
# ---------------------------------------------------------------------------
def py_box_dijkstra(spec, wcost: np.ndarray, box, src: int):
    """Pure-Python canonical Dijkstra over a box (full sweep).  Returns
    {global node: (dist, hops, parent_global or -1)}."""
    x0, y0, w, h = (int(v) for v in box)
    nb = spec.neighbours()

    def inside(g):
        x, y, _ = spec.coords(g)
        return x0 <= x < x0 + w and y0 <= y < y0 + h

    best = {src: (np.float32(0.0), 0, ROOT_SLOT)}
    heap = [(np.float32(0.0), 0, ROOT_SLOT, src)]
    done = set()
    while heap:
        d, hp, sl, u = heapq.heappop(heap)
        if u in done or best[u] != (d, hp, sl):
            continue
        done.add(u)
        for s in range(4):
            v = int(nb[u, s])
            if v < 0 or not inside(v) or v in done:
                continue
            wt = np.float32(wcost[4 * u + s])
            if np.isinf(wt):
                continue
            nd = np.float32(d + wt)
            key = (nd, hp + 1, 3 - s)
            if v not in best or key < best[v]:
                best[v] = key
                heapq.heappush(heap, (nd, hp + 1, 3 - s, v))
    out = {}
    for v, (d, hp, sl) in best.items():
        par = -1 if sl == ROOT_SLOT else int(nb[v, sl])
        out[v] = (float(d), hp, par)
    return out


# ---------------------------------------------------------------------------
# Bidirectional BFS region of interest (ROIExtractor.extract_roi_bfs,
# unified_pathfinder.py:2257-2384), restated over whole BFS levels
# ---------------------------------------------------------------------------
def roi_bfs(indptr, indices, N: int, src: int, dst: int, initial_radius: int = 40, stagnation_bonus: float = 0.0,
            portal_seeds=None, max_nodes: int = 50_000, blocked=None):
    """(roi_nodes, global_to_roi) of extract_roi_bfs: level-synchronous
    waves from src (mark 1) then dst (mark 2) per depth; a node reached by
    the other wave is marked 3 and the waves have met; stop at max_depth =
    min(int(radius + 2*bonus), 80), at meeting depth + halo (halo =
    min(16, 2 + int(max(0, bonus) / 0.6))), or once more than 1.5*max_nodes
    nodes are seen.  ROI = seen nodes ascending, then the missing must-keep
    nodes; beyond max_nodes, must-keep positions past the cut swap to the
    front positions before the cut."""
    indptr = np.asarray(indptr, np.int64)
    indices = np.asarray(indices, np.int64)
    seen = np.zeros(N, np.uint8)
    seen[src] = 1
    seen[dst] = 2
    fr = {1: np.asarray([src], np.int64), 2: np.asarray([dst], np.int64)}
    max_depth = min(int(initial_radius + stagnation_bonus * 2.0), 80)
    halo = min(16, 2 + int(max(0.0, stagnation_bonus) / 0.6))
    depth, meet, met = 0, None, False
    while depth < max_depth and (len(fr[1]) or len(fr[2])):
        for mark in (1, 2):
            u = fr[mark]
            if len(u) == 0:
                continue
            starts, ends = indptr[u], indptr[u + 1]
            arcs = np.concatenate([np.arange(a, b) for a, b in zip(starts, ends)]) if len(u) else np.zeros(0, np.int64)
            if blocked is not None and len(arcs):
                arcs = arcs[~np.asarray(blocked, bool)[arcs]]
            v = np.unique(indices[arcs])
            old = seen[v]
            new = v[old == 0]
            if np.any((old != 0) & (old != mark)):
                met = True
            seen[v[(old != 0) & (old != mark)]] = 3
            seen[new] = mark
            fr[mark] = new
        depth += 1
        if met and meet is None:
            meet = depth
        if meet is not None and depth >= meet + halo:
            break
        if int((seen > 0).sum()) > max_nodes * 1.5:
            break
    roi = np.flatnonzero(seen > 0).astype(np.int64)
    keep = [src, dst]
    for node, _ in (portal_seeds or []):
        if node not in keep:
            keep.append(node)
    inset = set(roi.tolist())
    missing = [n for n in keep if n not in inset]
    if missing:
        roi = np.append(roi, missing)
    if roi.size > max_nodes:
        kset = set(keep)
        kept = [i for i, n in enumerate(roi.tolist()) if n in kset]
        for swap_pos, k in enumerate(kept):
            if k >= max_nodes:
                roi[k], roi[swap_pos] = roi[swap_pos], roi[k]
        roi = roi[:max_nodes]
    g2r = np.full(N, -1, np.int32)
    g2r[roi] = np.arange(len(roi), dtype=np.int32)
    return roi.astype(np.int32), g2r, {"depth": depth, "meeting_depth": meet, "seen": int((seen > 0).sum())}
