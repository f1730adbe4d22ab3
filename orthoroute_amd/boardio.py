"""Job ingestion and result export of the routing service: OrthoRoute board
files (.ORP) in, solution files (.ORS) out.  Both are JSON, gzip-compressed
or plain (docs/ORP_ORS_file_formats.md of the reference).

Board -> lattice follows the reference engine:
  * routing bounds = bounding box of the pads of routable nets (2+ pads)
    +- 3 mm (_calc_bounds, unified_pathfinder.py:3266-3306);
  * grid aligned to the pitch: grid_min = round(min / pitch) * pitch,
    steps = int((grid_max - grid_min) / pitch) + 1, lattice index
    round((w - grid_min) / pitch) (KiCadGeometry, kicad_geometry.py:24-57);
  * layer 0 (F.Cu) vertical, odd layers horizontal (kicad_geometry.py:36-43);
  * a net routes its first two pads (_parse_requests,
    unified_pathfinder.py:4397-4428); pads on one cell make a trivial net
    (UP:4431-4435).

Solution export mirrors export_solution_to_ors (ors_exporter.py:18-98):
straight same-layer runs of a path become tracks, vertical runs vias.
"""
from __future__ import annotations

import gzip
import json
import math
from dataclasses import dataclass, field
from datetime import datetime, timezone
from pathlib import Path

import numpy as np

from .lattice import LatticeSpec
from .netgen import Netlist, roi_boxes, with_portals

ORP_SECTIONS = ("metadata", "board", "pads", "nets", "layers", "drc_rules", "grid_parameters")
ORS_SECTIONS = ("metadata", "geometry", "iteration_metrics", "statistics")
ROUTING_MARGIN_MM = 3.0


def _read_json(path: Path, kind: str) -> dict:
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"{kind} file not found: {path}")
    try:
        with gzip.open(path, "rt", encoding="utf-8") as f:
            data = json.load(f)
    except (gzip.BadGzipFile, OSError):
        with open(path, "r", encoding="utf-8") as f:
            data = json.load(f)
    return data


def _check(data: dict, kind: str, sections) -> dict:
    """Version and section checks of import_board_from_orp /
    import_solution_from_orp (orp_exporter.py:575-588, ors_exporter.py:380-392)."""
    version = data.get("format_version")
    if not version:
        raise ValueError(f"Missing format_version in {kind} file")
    if version != "1.0":
        raise ValueError(f"Unsupported {kind} format version: {version}")
    missing = [s for s in sections if s not in data]
    if missing:
        raise ValueError(f"Missing required sections: {', '.join(missing)}")
    return data


def load_orp(path) -> dict:
    return _check(_read_json(path, "ORP"), "ORP", ORP_SECTIONS)


def load_ors(path) -> dict:
    return _check(_read_json(path, "ORS"), "ORS", ORS_SECTIONS)


def _write_json(data: dict, path, compress: bool) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    opener = gzip.open if compress else open
    with opener(path, "wt", encoding="utf-8") as f:
        json.dump(data, f, indent=2, ensure_ascii=False)


def layer_names(count: int) -> list[str]:
    if count == 2:
        return ["F.Cu", "B.Cu"]
    return ["F.Cu"] + [f"In{i}.Cu" for i in range(1, count - 1)] + ["B.Cu"]


@dataclass
class BoardGrid:
    """Pitch-aligned routing grid (kicad_geometry.py:24-57)."""
    pitch: float
    grid_min_x: float
    grid_min_y: float
    x_steps: int
    y_steps: int

    @classmethod
    def from_bounds(cls, bounds, pitch: float) -> "BoardGrid":
        min_x, min_y, max_x, max_y = bounds
        gx0, gy0 = round(min_x / pitch) * pitch, round(min_y / pitch) * pitch
        gx1, gy1 = round(max_x / pitch) * pitch, round(max_y / pitch) * pitch
        return cls(pitch, gx0, gy0, int((gx1 - gx0) / pitch) + 1, int((gy1 - gy0) / pitch) + 1)

    def world_to_lattice(self, x: float, y: float) -> tuple[int, int]:
        return round((x - self.grid_min_x) / self.pitch), round((y - self.grid_min_y) / self.pitch)

    def lattice_to_world(self, ix: int, iy: int) -> tuple[float, float]:
        return self.grid_min_x + ix * self.pitch, self.grid_min_y + iy * self.pitch


@dataclass
class Board:
    """A board ready for the router: its lattice netlist and what maps the
    nets back to the board."""
    name: str
    grid: BoardGrid
    layers: list[str]
    netlist: Netlist
    net_names: list[str]             # name of each netlist net
    pad_ids: list[tuple[str, str]]   # (source pad, sink pad) of each netlist net
    trivial: list[str] = field(default_factory=list)   # both pads on one cell
    drc: dict = field(default_factory=dict)


def orp_to_board(orp: dict, roi_margin: int = 4, portals: int = 0) -> Board:
    """Netlist of an ORP board.  ``portals`` > 0 gives every pad that many
    escape cells with a terminal on each inner layer (netgen.with_portals);
    0 keeps the pads as two-pin sources and sinks."""
    meta = orp.get("metadata", {})
    board = orp.get("board", {})
    count = int(board.get("layer_count", 2))
    names = [l.get("name", f"Layer{i}") for i, l in enumerate(orp.get("layers", []))] or layer_names(count)
    count = len(names)
    if count < 2:
        raise ValueError("a board needs at least two copper layers")
    pitch = float(orp.get("grid_parameters", {}).get("grid_pitch", 0.4))
    pads = orp.get("pads", [])
    by_net: dict[str, list[dict]] = {}
    for p in pads:
        by_net.setdefault(p.get("net_name") or "", []).append(p)
    routable = []
    for net in orp.get("nets", []):
        name = net.get("name", "")
        ps = by_net.get(name, []) if name else []
        if len(ps) >= 2:
            routable.append((name, ps))
    if not routable:
        raise ValueError("the board has no routable net (a net needs two pads)")
    xs = [float(p["position"]["x"]) for _, ps in routable for p in ps]
    ys = [float(p["position"]["y"]) for _, ps in routable for p in ps]
    m = ROUTING_MARGIN_MM
    grid = BoardGrid.from_bounds((min(xs) - m, min(ys) - m, max(xs) + m, max(ys) + m), pitch)
    spec = LatticeSpec(grid.x_steps, grid.y_steps, count)
    layer_of = {n: i for i, n in enumerate(names)}

    def node(p):
        ix, iy = grid.world_to_lattice(float(p["position"]["x"]), float(p["position"]["y"]))
        z = layer_of.get(p.get("layer", "F.Cu"), 0) if not p.get("drill_size") else 0
        return int(spec.node(ix, iy, z)), (ix, iy)

    src, dst, net_names, pad_ids, trivial = [], [], [], [], []
    for name, ps in routable:
        (a, ca), (b, cb) = node(ps[0]), node(ps[1])
        if ca == cb:
            trivial.append(name)
            continue
        src.append(a)
        dst.append(b)
        net_names.append(name)
        pad_ids.append((ps[0].get("id", ""), ps[1].get("id", "")))
    src, dst = np.asarray(src, np.int64), np.asarray(dst, np.int64)
    nl = Netlist(spec, src.astype(np.int32), dst.astype(np.int32), roi_boxes(spec, src, dst, roi_margin),
                 meta.get("board_name", "board"), {"margin": roi_margin, "pitch": pitch})
    if portals > 0:
        nl = with_portals(nl, candidates=portals)
    return Board(meta.get("board_name", "board"), grid, names, nl, net_names, pad_ids, trivial,
                 orp.get("drc_rules", {}))


def path_geometry(board: Board, path: np.ndarray):
    """Tracks (straight same-layer runs: layer, x0, y0, x1, y1) and vias
    (vertical runs: x, y, from layer, to layer) of one routed path, in
    lattice cells and layer indices."""
    spec = board.netlist.spec
    x, y, z = spec.coords(np.asarray(path, np.int64))
    tracks, vias = [], []
    i, n = 0, len(path)
    while i + 1 < n:
        j = i
        if z[i + 1] != z[i]:
            while j + 1 < n and x[j + 1] == x[i] and y[j + 1] == y[i] and z[j + 1] != z[j]:
                j += 1
            vias.append((int(x[i]), int(y[i]), int(z[i]), int(z[j])))
        else:
            dx, dy = int(x[i + 1] - x[i]), int(y[i + 1] - y[i])
            while j + 1 < n and z[j + 1] == z[i] and int(x[j + 1] - x[j]) == dx and int(y[j + 1] - y[j]) == dy:
                j += 1
            tracks.append((int(z[i]), int(x[i]), int(y[i]), int(x[j]), int(y[j])))
        i = j
    return tracks, vias


def geometry_to_path(board: Board, tracks: list[dict], vias: list[dict], start: tuple[int, int, int]):
    """Lattice node path of one net from its ORS tracks and vias in path
    order, starting at lattice cell ``start`` (x, y, z)."""
    spec = board.netlist.spec
    g = board.grid
    layer_of = {n: i for i, n in enumerate(board.layers)}

    def cell(p):
        return g.world_to_lattice(p["x"], p["y"])

    items = [("t", t) for t in tracks] + [("v", v) for v in vias]
    x, y, z = start
    out = [int(spec.node(x, y, z))]
    pending = list(items)
    while pending:
        for k, (kind, it) in enumerate(pending):
            if kind == "t":
                sx, sy = cell(it["start"])
                if (sx, sy, layer_of[it["layer"]]) != (x, y, z):
                    continue
                ex, ey = cell(it["end"])
                stepx, stepy = int(np.sign(ex - sx)), int(np.sign(ey - sy))
                while (x, y) != (ex, ey):
                    x, y = x + stepx, y + stepy
                    out.append(int(spec.node(x, y, z)))
            else:
                px, py = cell(it["position"])
                if (px, py, layer_of[it["from_layer"]]) != (x, y, z):
                    continue
                z1 = layer_of[it["to_layer"]]
                step = 1 if z1 > z else -1
                while z != z1:
                    z += step
                    out.append(int(spec.node(x, y, z)))
            del pending[k]
            break
        else:
            raise ValueError("tracks and vias do not form one path from the start cell")
    return np.asarray(out, np.int64)


def solution_to_ors(board: Board, paths, history, total_time: float, converged: bool | None = None,
                    notes: str = "") -> dict:
    """ORS dictionary of a routed board (export_solution_to_ors,
    ors_exporter.py:18-98): paths[k] is netlist net k's node path (empty if
    unrouted); history the router's IterationStats."""
    g = board.grid
    d = board.drc.get("default", {})
    width = float(d.get("track_width", 0.25))
    via_d, via_drill = float(d.get("via_diameter", 0.6)), float(d.get("via_drill", 0.3))
    by_net, all_tracks, all_vias, usage = {}, [], [], {}
    wirelength = 0.0
    for k, path in enumerate(paths):
        if len(path) < 2:
            continue
        name = board.net_names[k]
        tr, vi = path_geometry(board, path)
        net = by_net.setdefault(name, {"net_id": name, "tracks": [], "vias": []})
        for z, x0, y0, x1, y1 in tr:
            (ax, ay), (bx, by) = g.lattice_to_world(x0, y0), g.lattice_to_world(x1, y1)
            t = {"layer": board.layers[z], "start": {"x": ax, "y": ay}, "end": {"x": bx, "y": by}, "width": width}
            net["tracks"].append(t)
            all_tracks.append(dict(t, net_id=name))
            usage[board.layers[z]] = usage.get(board.layers[z], 0) + 1
            wirelength += math.hypot(bx - ax, by - ay)
        for x, y, z0, z1 in vi:
            px, py = g.lattice_to_world(x, y)
            v = {"position": {"x": px, "y": py}, "from_layer": board.layers[z0], "to_layer": board.layers[z1],
                 "diameter": via_d, "drill": via_drill}
            net["vias"].append(v)
            all_vias.append(dict(v, net_id=name))
    metrics = [{"iteration": s.iteration, "overuse_count": s.overuse, "nets_routed": s.routed,
                "overflow_cost": float(s.edge_over + s.node_over),
                "wirelength": (s.path_hops - s.via_hops) * g.pitch, "via_count": s.barrels,
                "iteration_time_seconds": s.seconds} for s in history]
    final_over = history[-1].overuse if history else 0
    if converged is None:
        converged = final_over == 0
    return {
        "format_version": "1.0",
        "metadata": {"export_timestamp": datetime.now(timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
                     "orthoroute_version": "0.1.0", "board_name": board.name, "total_iterations": len(history),
                     "converged": bool(converged), "total_time_seconds": float(total_time), "notes": notes},
        "geometry": {"by_net": by_net, "all_tracks": all_tracks, "all_vias": all_vias, "layer_usage": usage},
        "iteration_metrics": metrics,
        "statistics": {"total_tracks": len(all_tracks), "total_vias": len(all_vias),
                       "total_wirelength_mm": wirelength, "nets_routed": len(by_net),
                       "final_overuse_count": int(final_over),
                       "final_overflow_cost": float(metrics[-1]["overflow_cost"]) if metrics else 0.0,
                       "converged": bool(converged), "iterations_completed": len(history)},
    }


def save_ors(data: dict, path, compress: bool = True) -> None:
    _write_json(data, path, compress)


def save_orp(data: dict, path, compress: bool = True) -> None:
    _write_json(_check(data, "ORP", ORP_SECTIONS), path, compress)


def netlist_to_orp(nl: Netlist, pitch: float = 0.4, origin=(100.0, 50.0), name: str | None = None) -> dict:
    """ORP dictionary of a synthetic board: every net's two pins become SMD
    pads on F.Cu at their lattice cell (origin + cell * pitch), nets named
    N<k>, pads U<k>.1 / U<k>.2 (the layout export_board_to_orp writes,
    orp_exporter.py:39-92)."""
    spec = nl.spec
    names = layer_names(spec.L)
    pads, nets = [], []
    xs, ys = [], []
    for k in range(nl.n):
        terms = []
        for j, g in enumerate((int(nl.src[k]), int(nl.dst[k]))):
            x, y, _ = spec.coords(np.asarray([g]))
            px, py = origin[0] + int(x[0]) * pitch, origin[1] + int(y[0]) * pitch
            xs.append(px)
            ys.append(py)
            pid = f"U{k}.{j + 1}"
            pads.append({"id": pid, "component_id": f"U{k}", "component_ref": f"U{k}", "net_id": f"N{k}",
                         "net_name": f"N{k}", "position": {"x": px, "y": py},
                         "size": {"width": 0.3, "height": 0.3}, "drill_size": None, "layer": "F.Cu",
                         "shape": "rect", "angle": 0.0})
            terms.append({"pad_id": pid, "position": {"x": px, "y": py}, "layer": "F.Cu"})
        tx, ty = [t["position"]["x"] for t in terms], [t["position"]["y"] for t in terms]
        nets.append({"id": f"N{k}", "name": f"N{k}", "netclass": "Default", "pad_count": 2, "is_routable": True,
                     "terminals": terms,
                     "bounds": {"x_min": min(tx), "y_min": min(ty), "x_max": max(tx), "y_max": max(ty)}})
    x0, y0, x1, y1 = min(xs), min(ys), max(xs), max(ys)
    return {
        "format_version": "1.0",
        "metadata": {"filename": f"{name or nl.name}.ORP", "board_name": name or nl.name, "board_id": name or nl.name,
                     "export_timestamp": datetime.now(timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
                     "orthoroute_version": "0.1.0"},
        "board": {"bounds": {"x_min": x0, "y_min": y0, "x_max": x1, "y_max": y1, "width": x1 - x0, "height": y1 - y0},
                  "layer_count": spec.L, "thickness": 1.6},
        "pads": pads,
        "nets": nets,
        "layers": [{"name": n, "type": "signal", "stackup_position": i, "thickness": 0.035, "material": "copper",
                    "is_routing_layer": True} for i, n in enumerate(names)],
        "drc_rules": {"default": {"track_width": 0.25, "clearance": 0.2, "via_diameter": 0.6, "via_drill": 0.3,
                                  "netclass": "Default"},
                      "netclasses": {},
                      "min_values": {"track_width": 0.1, "clearance": 0.1, "via_diameter": 0.2, "via_drill": 0.1}},
        "grid_parameters": {"grid_pitch": pitch, "expansion_margin": ROUTING_MARGIN_MM,
                            "coordinate_system": "PCB coordinates in millimeters, origin at board bounds minimum",
                            "board_bounds_used": {"x_min": x0, "y_min": y0, "x_max": x1, "y_max": y1}},
    }


def ors_writer(board: Board, path, compress: bool = True):
    """on_finish callback of a serving job (serve.RouteJob): writes the
    job's solution to ``path`` as .ORS when it retires."""
    def write(job, runner) -> None:
        total = job.finished_at - job.admitted_at if job.admitted_at is not None else 0.0
        save_ors(solution_to_ors(board, runner.paths(), runner.history, total, converged=job.converged), path,
                 compress)
    return write
