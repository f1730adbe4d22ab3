"""Launch-parameter autotuner: bucket width delta, workgroup size and the
ROI size up to which search keys live in LDS.

``tune`` times candidate configurations on a sample of one iteration's waves
(the first ``sample_waves`` waves, which hold the largest ROIs, plus an even
stride over the rest) and keeps the fastest.  Results persist in
``tuned.json`` beside this file, keyed by device architecture and workload
shape.  Results never depend on these parameters (every configuration
computes the same canonical keys); only time does.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
from pathlib import Path

import numpy as np

TUNED_PATH = Path(__file__).resolve().parent / "tuned.json"
DEFAULTS = {"delta": 4.0, "block": 1024, "lds_nodes": 12288}


def _key(nl, arch: str = "gfx950") -> str:
    s = nl.spec
    return f"{arch}:{s.X}x{s.Y}x{s.L}:{nl.n}"


def load_tuned(nl) -> dict:
    try:
        table = json.loads(TUNED_PATH.read_text())
    except (OSError, ValueError):
        return dict(DEFAULTS)
    return dict(table.get(_key(nl), DEFAULTS))


def save_tuned(nl, params: dict) -> None:
    try:
        table = json.loads(TUNED_PATH.read_text())
    except (OSError, ValueError):
        table = {}
    table[_key(nl)] = params
    TUNED_PATH.write_text(json.dumps(table, indent=1, sort_keys=True))


def tune(nl, deltas=(1.0, 2.0, 4.0, 8.0, float("inf")), blocks=(256, 512, 1024),
         lds_nodes=(0, 8192, 12288), sample_waves: int = 12, stride: int = 8, save: bool = True,
         log=print) -> dict:
    """Time every configuration on the same sample of waves of iteration 1
    (each candidate starts from the same fresh state)."""
    import torch

    from . import _lib
    from .router import DevicePathFinder, RouterConfig, _wave_offsets

    results = []
    for delta, block, lds in itertools.product(deltas, blocks, lds_nodes):
        cfg = RouterConfig(iterations=1, delta=delta, block=block, lds_nodes=lds, collect_stats=False)
        d = DevicePathFinder(nl, cfg)
        order = d.order_for(1)
        waves = d.plan(order, "batched")
        pick = list(range(min(sample_waves, len(waves)))) + list(range(sample_waves, len(waves), stride))
        params = cfg.params(float(d.schedule[0]), d.astar)
        d.lat.reprice(params)
        rcfg = d.route_cfg()
        s = torch.cuda.current_stream()
        total = 0.0
        for wi in pick:
            w = waves[wi]
            off, need = _wave_offsets(d.nets.roi_nodes, w)
            d.ws.ensure(need)
            dn = torch.from_numpy(w.astype(np.int32)).to(d.device)
            do = torch.from_numpy(off).to(d.device)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
            _lib.call("orx_route_wave", C.byref(d.lat.struct), C.byref(params), C.byref(d.nets.struct),
                      dn.data_ptr(), do.data_ptr(), len(w), int(d.nets.roi_nodes[w].max()), C.byref(d.ws.struct),
                      C.byref(rcfg), s.cuda_stream)
            e1.record(s)
            e1.synchronize()
            total += e0.elapsed_time(e1)
        results.append((total, {"delta": delta, "block": block, "lds_nodes": lds}))
        log(f"[autotune] delta={delta} block={block} lds_nodes={lds}: {total:.2f} ms on {len(pick)} waves")
        del d
        torch.cuda.empty_cache()
    best = min(results, key=lambda r: r[0])[1]
    if save:
        save_tuned(nl, best)
    return best
