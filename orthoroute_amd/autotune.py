"""Launch-parameter autotuner: bucket width delta, workgroup size and the
ROI size up to which search keys live in LDS.

``tune`` times every candidate configuration on the same full PathFinder
iteration of the workload and keeps the fastest.  Results persist in
``tuned.json`` beside this file, keyed by device architecture and workload
shape.  Results never depend on these parameters (every configuration
computes the same canonical keys); only time does, and the tuner checks it.

    python -m orthoroute_amd.autotune uneven_800
    python -m orthoroute_amd.autotune backplane_400 --deltas 16,256,1024,4096 --blocks 1024 --lds-nodes 16384
"""
from __future__ import annotations

import itertools
import json
import sys
from pathlib import Path

TUNED_PATH = Path(__file__).resolve().parent / "tuned.json"
DEFAULTS = {"delta": 16.0, "block": 1024, "lds_nodes": 16384}


def _key(nl, arch: str = "gfx950") -> str:
    s = nl.spec
    return f"{arch}:{s.X}x{s.Y}x{s.L}:{nl.n}"


def load_tuned(nl) -> dict:
    try:
        table = json.loads(TUNED_PATH.read_text())
    except (OSError, ValueError):
        return dict(DEFAULTS)
    return {**DEFAULTS, **table.get(_key(nl), {})}


def save_tuned(nl, params: dict) -> None:
    try:
        table = json.loads(TUNED_PATH.read_text())
    except (OSError, ValueError):
        table = {}
    table[_key(nl)] = params
    TUNED_PATH.write_text(json.dumps(table, indent=1, sort_keys=True))


def tune(nl, deltas=(4.0, 8.0, 16.0, 32.0, 64.0, 256.0, 1024.0), blocks=(256, 512, 1024), lds_nodes=(8192, 16384),
         iteration: int = 2, save: bool = True, log=print) -> dict:
    """Time every configuration on the same full iteration (``iteration``,
    default 2: the first congested one) of ``nl``.

    One job is routed up to the iteration before, its device state is
    snapshotted, and each candidate restores the snapshot and routes the
    iteration.  Every configuration must produce the same result (checked by
    digest); only the route-kernel time, from HIP events, differs."""
    import dataclasses

    import torch

    from .router import DevicePathFinder, RouterConfig

    base_cfg = RouterConfig(iterations=iteration, collect_stats=False, adapt_delta=False, **DEFAULTS)
    job = DevicePathFinder(nl, base_cfg)
    for _ in range(iteration - 1):
        job.run_iteration()
    snap = job.snapshot()
    results = []
    ref_digest = None
    for delta, block, lds in itertools.product(deltas, blocks, lds_nodes):
        job.restore(snap)
        job.cfg = dataclasses.replace(base_cfg, delta=delta, block=block, lds_nodes=lds)
        st = job.run_iteration()
        dg = job.digest()
        if ref_digest is None:
            ref_digest = dg
        elif dg != ref_digest:
            raise RuntimeError(f"delta={delta} block={block} lds_nodes={lds} changed the result")
        results.append((st.route_ms, {"delta": delta, "block": block, "lds_nodes": lds}))
        log(f"[autotune] delta={delta} block={block} lds_nodes={lds}: route {st.route_ms:.2f} ms "
            f"({st.waves} launches)")
    del job
    torch.cuda.empty_cache()
    best = min(results, key=lambda r: r[0])[1]
    if save:
        save_tuned(nl, best)
    return best


if __name__ == "__main__":
    import argparse

    from . import netgen
    ap = argparse.ArgumentParser(description="time launch parameters on one PathFinder iteration of a workload")
    ap.add_argument("workload", nargs="?", default="uneven_800")
    ap.add_argument("--deltas", default=None, help="comma-separated bucket widths")
    ap.add_argument("--blocks", default=None, help="comma-separated workgroup sizes")
    ap.add_argument("--lds-nodes", default=None, help="comma-separated LDS key capacities")
    ap.add_argument("--no-save", action="store_true")
    a = ap.parse_args()
    kw = {}
    if a.deltas:
        kw["deltas"] = tuple(float(v) for v in a.deltas.split(","))
    if a.blocks:
        kw["blocks"] = tuple(int(v) for v in a.blocks.split(","))
    if a.lds_nodes:
        kw["lds_nodes"] = tuple(int(v) for v in a.lds_nodes.split(","))
    print(json.dumps(tune(getattr(netgen, a.workload)(), save=not a.no_save, log=lambda m: print(m, flush=True),
                          **kw)))
