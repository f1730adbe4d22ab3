"""Runs PathFinder iterations of the kernel sources through the host-thread
ASan build (route_driver) -- the same host logic as
orthoroute_amd.router.DevicePathFinder (order, planner, per-net workspace),
with every array on the host.  Test infrastructure."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

import numpy as np

from orthoroute_amd.waves import check_schedule, net_distance, net_order, plan_iteration

HERE = Path(__file__).resolve().parent
DRIVER = HERE / "_out" / "route_driver"


def build_driver() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return DRIVER


class CpuBuildPathFinder:
    def __init__(self, netlist, cfg, workdir: Path, mode: int = 0, max_groups: int = 0, base=None,
                 extra_meta: dict | None = None):
        """mode 0: dependency-driven launch; 1: one launch per net; 2: the
        wave sequence of orx_route_waves; 3: the same waves on the tile
        engine (orx_route_tiles).  extra_meta: driver options
        (tile_steps_cap, bad_slot)."""
        self.extra_meta = dict(extra_meta or {})
        self.nl = netlist
        self.cfg = cfg
        self.spec = spec = netlist.spec
        self.dir = Path(workdir)
        self.dir.mkdir(parents=True, exist_ok=True)
        self.mode = mode
        self.max_groups = max_groups
        N, n = spec.N, netlist.n
        self.base = spec.base_costs() if base is None else np.asarray(base, np.float32).copy()
        self.hist = np.zeros(4 * N, np.float32)
        self.occ = np.zeros(4 * N, np.int32)
        self.wcost = np.zeros(4 * N, np.float32)
        self.node_hist = np.zeros(N, np.float32)
        self.node_use = np.zeros(N, np.int32)
        self.via_col = np.zeros(spec.XY, np.int32)
        R = netlist.roi_nodes().astype(np.int64)
        self.R = R
        self.path_cap = R.astype(np.int32)
        self.path_off = np.zeros(n, np.int64)
        self.path_off[1:] = np.cumsum(R)[:-1]
        self.arena = np.zeros(int(R.sum()), np.int32)
        self.path_len = np.zeros(n, np.int32)
        self.path_cost = np.full(n, np.inf, np.float32)
        self.status = np.zeros(n, np.int32)
        r64 = (R + 63) // 64 * 64
        self.ws_off = np.zeros(n, np.int64)
        self.ws_off[1:] = np.cumsum(r64)[:-1]
        self.ws_nodes = int(r64.sum())
        self.done = np.full(n, -1, np.int32)
        self.distance = net_distance(spec, netlist.src, netlist.dst)
        self.schedule = cfg.schedule(spec.L)
        self.astar = cfg.astar_bounds(self.base)
        self.iteration = 0
        self.nb = spec.neighbours()

    def _net_overuse(self):
        n = self.nl.n
        over = np.zeros(n, np.int64)
        hops = np.maximum(self.path_len - 1, 0).astype(np.int32)
        for k in range(n):
            p = self.arena[self.path_off[k]:self.path_off[k] + self.path_len[k]].astype(np.int64)
            if len(p) < 2:
                continue
            s = np.argmax(self.nb[p[:-1]] == p[1:, None], axis=1)
            o = self.occ[4 * p[:-1] + s] - self.cfg.edge_cap
            over[k] = int(np.maximum(o, 0).sum())
        return over, hops

    def run_iteration(self, timeout: float = 600.0) -> dict:
        it = self.iteration + 1
        n = self.nl.n
        if it > 1:
            over, hops = self._net_overuse()
        else:
            over, hops = np.zeros(n, np.int64), np.zeros(n, np.int32)
        order = net_order(it, self.distance, over, hops)
        wave, sched, dep_ptr, dep_ids = plan_iteration(self.nl.box, order, self.spec.X, self.spec.Y, self.R)
        check_schedule(sched, dep_ptr, dep_ids)
        self.last_wave = wave
        if self.mode == 1:
            sched = order.astype(np.int32)
        p = np.float32(self.schedule[it - 1])
        c = self.cfg
        meta = {
            "X": self.spec.X, "Y": self.spec.Y, "L": self.spec.L, "hmask": self.spec.hmask, "n": n,
            "arena": len(self.arena), "n_deps": len(dep_ids), "ws_nodes": self.ws_nodes,
            "max_nodes": int(self.R.max()), "mode": self.mode, "epoch": it, "max_groups": self.max_groups,
            "pres_fac": float(p), "base_w": c.base_w, "hist_w": c.hist_w,
            "node_pres": float(np.float32(c.node_pres_w) * p), "node_hist_w": c.node_hist_w,
            "edge_cap": c.edge_cap, "node_cap": c.node_cap, "hist_gain": c.hist_gain, "hist_decay": c.hist_decay,
            "hist_cap_mult": c.hist_cap_mult, "node_hist_gain": c.node_hist_gain,
            "astar_cp": self.astar[0], "astar_cv": self.astar[1], "astar_gmax": self.astar[2],
            "delta": c.delta, "block": c.block, "lds_nodes": c.lds_nodes, "flags": 1,
        }
        T = self.nl.terms
        if T is not None:
            meta.update({"terms": 1, "n_seed": len(T.seed_node), "n_tgt": len(T.tgt_node)})
        meta.update(self.extra_meta)
        with open(self.dir / "meta.txt", "w") as f:
            for k, v in meta.items():
                f.write(f"{k} {float(v)!r}\n")
        arrays = {
            "base": self.base, "hist": self.hist, "occ": self.occ, "wcost": self.wcost,
            "node_hist": self.node_hist, "node_use": self.node_use, "via_col": self.via_col,
            "box": np.ascontiguousarray(self.nl.box.reshape(-1), np.int32),
            "src": self.nl.src.astype(np.int32), "dst": self.nl.dst.astype(np.int32),
            "path_arena": self.arena, "path_off": self.path_off, "path_cap": self.path_cap,
            "path_len": self.path_len, "path_cost": self.path_cost, "status": self.status,
            "sched": sched.astype(np.int32), "wave": wave.astype(np.int32), "dep_ptr": dep_ptr.astype(np.int32),
            "dep_ids": dep_ids.astype(np.int32) if len(dep_ids) else np.zeros(1, np.int32),
            "ws_off": self.ws_off, "done": self.done,
        }
        if T is not None:
            arrays.update({"seed_ptr": T.seed_ptr.astype(np.int32), "seed_node": T.seed_node.astype(np.int32),
                           "seed_cost": T.seed_cost.astype(np.float32), "tgt_ptr": T.tgt_ptr.astype(np.int32),
                           "tgt_node": T.tgt_node.astype(np.int32), "tgt_cost": T.tgt_cost.astype(np.float32)})
        for k, a in arrays.items():
            np.ascontiguousarray(a).tofile(self.dir / k)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")
        res = subprocess.run([str(DRIVER), str(self.dir)], capture_output=True, text=True, timeout=timeout, env=env)
        if res.returncode != 0:
            raise RuntimeError(f"route_driver failed ({res.returncode}):\n{res.stderr[-6000:]}")

        def rd(name, like):
            return np.fromfile(self.dir / f"{name}.out", dtype=like.dtype, count=like.size)

        self.hist, self.occ, self.wcost = rd("hist", self.hist), rd("occ", self.occ), rd("wcost", self.wcost)
        self.node_hist, self.node_use = rd("node_hist", self.node_hist), rd("node_use", self.node_use)
        self.via_col, self.arena = rd("via_col", self.via_col), rd("path_arena", self.arena)
        self.path_len, self.path_cost = rd("path_len", self.path_len), rd("path_cost", self.path_cost)
        self.status, self.done = rd("status", self.status), rd("done", self.done)
        stats = np.fromfile(self.dir / "stats.out", np.int64)
        out = np.fromfile(self.dir / "iter_out.out", np.int64)
        self.iteration = it
        return {"routed": int(stats[5]), "failed": int(stats[6]), "edge_over": int(out[0]),
                "edge_over_count": int(out[1]), "node_over": int(out[2]), "node_over_count": int(out[3]),
                "via_hops": int(out[4]), "barrels": int(out[5]), "settled": int(stats[0]), "relaxed": int(stats[1]),
                "path_hops": int(stats[7]), "rounds": int(stats[3]), "phases": int(stats[4]),
                "status": self.status.copy()}

    def paths(self):
        return [self.arena[o:o + l].copy() for o, l in zip(self.path_off, self.path_len)]


SSSP_DRIVER = HERE / "_out" / "sssp_driver"


def sssp_box_host_threads(spec, wcost: np.ndarray, box, src: int, dst: int, workdir: Path, grid: bool = True,
                          delta: float = 2.0, block: int = 256, timeout: float = 600.0, flags: int = 0):
    """One orx_sssp_box query through the host-thread build: (dist, parent,
    stats).  flags: extra ORX_F_* bits (ORX_F_TILES, ORX_F_ASYNC)."""
    d = Path(workdir)
    d.mkdir(parents=True, exist_ok=True)
    x0, y0, w, h = (int(v) for v in box)
    meta = {"X": spec.X, "Y": spec.Y, "L": spec.L, "hmask": spec.hmask, "x0": x0, "y0": y0, "w": w, "h": h,
            "src": src, "dst": dst, "flags": (8 if grid else 0) | 1 | (2 if dst < 0 else 0) | flags, "block": block,
            "delta": delta}
    with open(d / "meta.txt", "w") as f:
        for k, v in meta.items():
            f.write(f"{k} {float(v)!r}\n")
    np.ascontiguousarray(wcost, np.float32).tofile(d / "wcost")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    res = subprocess.run([str(SSSP_DRIVER), str(d)], capture_output=True, text=True, timeout=timeout, env=env)
    if res.returncode != 0:
        raise RuntimeError(f"sssp_driver failed ({res.returncode}):\n{res.stderr[-6000:]}")
    R = w * h * spec.L
    return (np.fromfile(d / "dist.out", np.float32, R), np.fromfile(d / "parent.out", np.int32, R),
            np.fromfile(d / "stats.out", np.int64, 8))


ROI_DRIVER = HERE / "_out" / "roi_driver"


def roi_bfs_host_threads(indptr, indices, N: int, src: int, dst: int, workdir: Path, initial_radius: int = 40,
                         stagnation_bonus: float = 0.0, portal_seeds=None, max_nodes: int = 50_000, blocked=None,
                         timeout: float = 600.0):
    """One orx_roi_bfs + orx_roi_release through the host-thread build:
    (roi, g2r, out[8], clean[2])."""
    d = Path(workdir)
    d.mkdir(parents=True, exist_ok=True)
    keep = [src, dst]
    for node, _ in (portal_seeds or []):
        if node not in keep:
            keep.append(node)
    meta = {"N": N, "E": len(indices), "src": src, "dst": dst,
            "max_depth": min(int(initial_radius + stagnation_bonus * 2.0), 80),
            "halo": min(16, 2 + int(max(0.0, stagnation_bonus) / 0.6)), "max_nodes": max_nodes,
            "n_keep": len(keep), "blocked": 0 if blocked is None else 1}
    with open(d / "meta.txt", "w") as f:
        for k, v in meta.items():
            f.write(f"{k} {float(v)!r}\n")
    np.ascontiguousarray(indptr, np.int32).tofile(d / "indptr")
    ix = np.ascontiguousarray(indices, np.int32)
    (ix if len(ix) else np.zeros(1, np.int32)).tofile(d / "indices")
    np.asarray(keep, np.int32).tofile(d / "keep")
    if blocked is not None:
        np.asarray(blocked, np.uint8).tofile(d / "blocked")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    res = subprocess.run([str(ROI_DRIVER), str(d)], capture_output=True, text=True, timeout=timeout, env=env)
    if res.returncode != 0:
        raise RuntimeError(f"roi_driver failed ({res.returncode}):\n{res.stderr[-6000:]}")
    out = np.fromfile(d / "out.out", np.int64)
    return (np.fromfile(d / "roi.out", np.int32, int(out[0])), np.fromfile(d / "g2r.out", np.int32, N), out,
            np.fromfile(d / "clean.out", np.int64))


CSR_DRIVER = HERE / "_out" / "csr_driver"


def csr_search_host_threads(N, indptr, indices, in_rank, in_indptr, in_src, costs, seeds, targets, workdir: Path,
                            seed_cost=None, tg_cost=None, roi_nodes=None, g2r=None, pen=None, grid: bool = True,
                            walk: bool = True, seed_pen: bool = True, delta: float = 2.0, block: int = 256,
                            timeout: float = 600.0):
    """One orx_sssp_csr query and its parent walk through the host-thread
    build: (keys[R] u64, best, path list)."""
    d = Path(workdir)
    d.mkdir(parents=True, exist_ok=True)
    R = N if roi_nodes is None else len(roi_nodes)
    meta = {"N": N, "E": len(indices), "R": R, "n_seeds": len(seeds), "n_tg": len(targets),
            "flags": (8 if grid else 0) | (4 if seed_pen else 0), "block": block, "delta": delta,
            "roi": int(roi_nodes is not None), "pen": int(pen is not None), "seed_cost": int(seed_cost is not None),
            "tg_cost": int(tg_cost is not None), "walk": int(walk)}
    with open(d / "meta.txt", "w") as f:
        for k, v in meta.items():
            f.write(f"{k} {float(v)!r}\n")
    arrays = {"indptr": (indptr, np.int32), "indices": (indices, np.int32), "in_rank": (in_rank, np.int32),
              "in_indptr": (in_indptr, np.int32), "in_src": (in_src, np.int32), "costs": (costs, np.float32),
              "seeds": (seeds, np.int32), "targets": (targets, np.int32), "seed_cost": (seed_cost, np.float32),
              "tg_cost": (tg_cost, np.float32), "roi_nodes": (roi_nodes, np.int32), "g2r": (g2r, np.int32),
              "pen": (pen, np.float32)}
    for k, (a, t) in arrays.items():
        if a is not None:
            np.ascontiguousarray(a, t).tofile(d / k)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    res = subprocess.run([str(CSR_DRIVER), str(d)], capture_output=True, text=True, timeout=timeout, env=env)
    if res.returncode != 0:
        raise RuntimeError(f"csr_driver failed ({res.returncode}):\n{res.stderr[-6000:]}")
    keys = np.fromfile(d / "keys.out", np.uint64, R)
    best = int(np.fromfile(d / "best.out", np.int32, 1)[0])
    n = int(np.fromfile(d / "len.out", np.int32, 1)[0])
    path = np.fromfile(d / "path.out", np.int32, R)[:max(n, 0)].tolist() if n > 1 else []
    return keys, best, n, path
