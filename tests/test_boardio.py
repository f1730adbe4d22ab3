"""Job ingestion (.ORP) and solution export (.ORS): the reference's file
checks, its board-to-lattice mapping on a hand-computed board, and a
round trip of routed paths through the ORS geometry."""
import gzip
import json

import numpy as np
import pytest

from oracle.oracle import OraclePathFinder
from orthoroute_amd.boardio import (BoardGrid, geometry_to_path, load_orp, load_ors, netlist_to_orp,
                                    orp_to_board, path_geometry, save_orp, save_ors, solution_to_ors)
from orthoroute_amd.netgen import uneven
from orthoroute_amd.router import RouterConfig


def _pad(pid, net, x, y, layer="F.Cu", drill=None):
    return {"id": pid, "component_ref": pid.split(".")[0], "net_name": net, "position": {"x": x, "y": y},
            "size": {"width": 0.3, "height": 0.3}, "drill_size": drill, "layer": layer}


def _toy():
    """Four layers, pitch 0.5: net A (3 pads: the first two are routed), B
    (pads on one cell: trivial), C (one pad: not routable), D (a B.Cu pad)."""
    pads = [_pad("U1.1", "A", 10.0, 20.0), _pad("U2.1", "A", 14.2, 21.0), _pad("U3.1", "A", 30.0, 30.0),
            _pad("U4.1", "B", 12.0, 25.0), _pad("U5.1", "B", 12.1, 25.1), _pad("U6.1", "C", 40.0, 40.0),
            _pad("U7.1", "D", 11.0, 22.0), _pad("U8.1", "D", 13.0, 23.6, layer="B.Cu")]
    nets = [{"id": n, "name": n} for n in ("A", "B", "C", "D")]
    names = ["F.Cu", "In1.Cu", "In2.Cu", "B.Cu"]
    return {"format_version": "1.0", "metadata": {"board_name": "toy"},
            "board": {"bounds": {"x_min": 0, "y_min": 0, "x_max": 50, "y_max": 50}, "layer_count": 4},
            "pads": pads, "nets": nets, "layers": [{"name": n} for n in names],
            "drc_rules": {"default": {"track_width": 0.2, "via_diameter": 0.5, "via_drill": 0.25}},
            "grid_parameters": {"grid_pitch": 0.5}}


def test_file_checks(tmp_path):
    orp = _toy()
    save_orp(orp, tmp_path / "a.orp")
    (tmp_path / "b.orp").write_text(json.dumps(orp))            # plain JSON is accepted too
    assert load_orp(tmp_path / "a.orp") == load_orp(tmp_path / "b.orp") == orp
    with gzip.open(tmp_path / "a.orp", "rt") as f:
        assert json.load(f)["format_version"] == "1.0"
    with pytest.raises(FileNotFoundError):
        load_orp(tmp_path / "none.orp")
    bad = dict(orp, format_version="2.0")
    (tmp_path / "c.orp").write_text(json.dumps(bad))
    with pytest.raises(ValueError, match="Unsupported ORP format version"):
        load_orp(tmp_path / "c.orp")
    del bad["format_version"]
    (tmp_path / "d.orp").write_text(json.dumps(bad))
    with pytest.raises(ValueError, match="Missing format_version"):
        load_orp(tmp_path / "d.orp")
    part = {k: v for k, v in orp.items() if k != "layers"}
    (tmp_path / "e.orp").write_text(json.dumps(part))
    with pytest.raises(ValueError, match="Missing required sections: layers"):
        load_orp(tmp_path / "e.orp")


def test_board_to_lattice_by_hand():
    """Routable pads span x 10..30, y 20..30 (net C has one pad and does not
    count); +-3 mm gives 7..33 x 17..33.  Pitch 0.5: grid_min (7.0, 17.0),
    steps (33-7)/0.5+1 = 53 and (33-17)/0.5+1 = 33.  Net A routes U1.1 (6,
    6) -> U2.1 (round(14.4)=14, round(8.0)=8); B's pads share cell (10, 16)
    and are trivial; D goes from (8, 10) on F.Cu to (12, 13) on B.Cu
    (round(13.2)=13)."""
    b = orp_to_board(_toy())
    assert (b.grid.grid_min_x, b.grid.grid_min_y, b.grid.x_steps, b.grid.y_steps) == (7.0, 17.0, 53, 33)
    spec = b.netlist.spec
    assert (spec.X, spec.Y, spec.L) == (53, 33, 4)
    assert b.net_names == ["A", "D"] and b.trivial == ["B"]
    assert b.pad_ids == [("U1.1", "U2.1"), ("U7.1", "U8.1")]
    assert list(b.netlist.src) == [spec.node(6, 6, 0), spec.node(8, 10, 0)]
    assert list(b.netlist.dst) == [spec.node(14, 8, 0), spec.node(12, 13, 3)]
    # layer 0 vertical, odd layers horizontal (kicad_geometry.py:36-43)
    assert spec.hmask == 0b1010


def test_grid_rounds_half_to_even():
    g = BoardGrid.from_bounds((97.0, 47.0, 110.0, 60.0), 0.4)   # 242.5 -> 242, 117.5 -> 118
    assert abs(g.grid_min_x - 96.8) < 1e-9 and abs(g.grid_min_y - 47.2) < 1e-9


def test_routed_paths_survive_the_ors_round_trip(tmp_path):
    nl = uneven(48, 40, 6, 60, seed=8)
    b = orp_to_board(netlist_to_orp(nl, pitch=0.4))
    assert b.netlist.n == nl.n and not b.trivial
    cfg = RouterConfig(iterations=4)
    o = OraclePathFinder(b.netlist, cfg)
    hist = [o.run_iteration() for _ in range(4)]

    class S:   # the fields solution_to_ors reads from the router's stats
        def __init__(self, d):
            self.__dict__.update(d)
            self.overuse = d["edge_over"] + d["node_over"]
            self.seconds = 0.0

    paths = o.paths()
    ors = solution_to_ors(b, paths, [S(d) for d in hist], total_time=1.0)
    save_ors(ors, tmp_path / "s.ors")
    back = load_ors(tmp_path / "s.ors")
    spec = b.netlist.spec
    planar = vias = 0
    for k, p in enumerate(paths):
        assert len(p) >= 2
        x, y, z = spec.coords(np.asarray(p[:1], np.int64))
        geo = back["geometry"]["by_net"][b.net_names[k]]
        q = geometry_to_path(b, geo["tracks"], geo["vias"], (int(x[0]), int(y[0]), int(z[0])))
        assert np.array_equal(q, p), k
        px, py, pz = spec.coords(np.asarray(p, np.int64))
        planar += int(np.sum(pz[1:] == pz[:-1]))
        vias += len(path_geometry(b, p)[1])
    st = back["statistics"]
    assert st["nets_routed"] == nl.n and st["total_vias"] == vias
    assert abs(st["total_wirelength_mm"] - planar * 0.4) < 1e-6 * planar
    assert st["iterations_completed"] == 4 and st["final_overuse_count"] == hist[-1]["edge_over"] + hist[-1]["node_over"]
    assert [m["nets_routed"] for m in back["iteration_metrics"]] == [d["routed"] for d in hist]


def test_portal_form_of_a_board():
    b = orp_to_board(_toy(), portals=2)
    ns, nt = b.netlist.terms.counts()
    assert list(ns) == [4, 4] and list(nt) == [4, 4]     # 2 escape cells x 2 inner layers


def test_board_file_served_and_exported(tmp_path):
    """A board file through the serving stream: ORP in, an oracle-backed
    router in the server slot, the job's on_finish hook writes the ORS, whose
    geometry equals that of the same board routed directly."""
    from types import SimpleNamespace

    from orthoroute_amd.boardio import ors_writer
    from orthoroute_amd.serve import RouteJob, RoutingServer

    class OracleRunner:
        def __init__(self, netlist, cfg, stream):
            self.o = OraclePathFinder(netlist, cfg)
            self.history, self._d = [], None

        def launch_iteration(self):
            self._d = self.o.run_iteration()

        def ready(self):
            return True

        def finish_iteration(self):
            d = self._d
            s = SimpleNamespace(**d, overuse=d["edge_over"] + d["node_over"], gpu_ms=0.0, seconds=0.0)
            self.history.append(s)
            return s

        def digest(self):
            return str(len(self.history))

        def paths(self):
            return self.o.paths()

    save_orp(netlist_to_orp(uneven(40, 40, 6, 50, seed=3), name="served"), tmp_path / "in.orp")
    board = orp_to_board(load_orp(tmp_path / "in.orp"))
    job = RouteJob(0, board.netlist, 0.0, max_iterations=6, on_finish=ors_writer(board, tmp_path / "out.ors"))
    srv = RoutingServer(RouterConfig(), hbm_budget=10 ** 12, runner=OracleRunner, streams=["s0"], idle_sleep=0.0)
    srv.run([job], time_limit=60)
    served = load_ors(tmp_path / "out.ors")
    direct = OraclePathFinder(board.netlist, RouterConfig())
    for _ in range(job.iterations):
        direct.run_iteration()
    assert served["geometry"]["by_net"] == solution_to_ors(board, direct.paths(), [], 0.0)["geometry"]["by_net"]
    assert served["statistics"]["iterations_completed"] == job.iterations
    assert served["metadata"]["board_name"] == "served"
