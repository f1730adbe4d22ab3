"""Launch-parameter autotuner: the per-net engine's bucket width delta,
workgroup size and the ROI size up to which its search keys live in LDS,
and the largest ROI of a wave that still goes to the per-net engine rather
than the tile engine (``small_wave_nodes``).  The tile engine has no bucket
width (it relaxes each tile to its fixpoint); the near-queue ring of the
per-net engine (2,048 entries per half) is a compile-time constant.

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
DEFAULTS = {"delta": 16.0, "block": 1024, "lds_nodes": 16384, "small_wave_nodes": 16384}


def _key(nl, arch: str = "gfx950") -> str:
    s = nl.spec
    return f"{arch}:{s.X}x{s.Y}x{s.L}:{nl.n}"


def load_tuned(nl, path: Path | None = None) -> dict:
    try:
        table = json.loads((path or TUNED_PATH).read_text())
    except (OSError, ValueError):
        return dict(DEFAULTS)
    return {**DEFAULTS, **table.get(_key(nl), {})}


def save_tuned(nl, params: dict, path: Path | None = None) -> None:
    path = path or TUNED_PATH
    try:
        table = json.loads(path.read_text())
    except (OSError, ValueError):
        table = {}
    table[_key(nl)] = params
    path.write_text(json.dumps(table, indent=1, sort_keys=True))


def tune(nl, deltas=(4.0, 8.0, 16.0, 32.0, 64.0, 256.0, 1024.0), blocks=(256, 512, 1024), lds_nodes=(8192, 16384),
         small_wave_nodes=(4096, 16384, 65536), iteration: int = 2, save: bool = True, log=print,
         runner=None, path: Path | None = None) -> dict:
    """Time every configuration on the same full iteration (``iteration``,
    default 2: the first congested one) of ``nl``.

    One job is routed up to the iteration before, its device state is
    snapshotted, and each candidate restores the snapshot and routes the
    iteration.  Every configuration must produce the same result (checked by
    digest; a candidate that changes it raises); only the route time, from
    HIP events, differs.  ``runner(nl, cfg)`` builds the job (default: a
    DevicePathFinder); it needs run_iteration() -> stats with route_ms and
    waves, snapshot(), restore(), digest() and a settable cfg."""
    import dataclasses

    from .router import RouterConfig

    if runner is None:
        from .router import DevicePathFinder
        runner = DevicePathFinder
    base_cfg = RouterConfig(iterations=iteration, collect_stats=False, adapt_delta=False, **DEFAULTS)
    job = runner(nl, base_cfg)
    for _ in range(iteration - 1):
        job.run_iteration()
    snap = job.snapshot()
    results = []
    ref_digest = None
    for delta, block, lds, sw in itertools.product(deltas, blocks, lds_nodes, small_wave_nodes):
        job.restore(snap)
        cand = {"delta": delta, "block": block, "lds_nodes": lds, "small_wave_nodes": sw}
        job.cfg = dataclasses.replace(base_cfg, **cand)
        st = job.run_iteration()
        dg = job.digest()
        if ref_digest is None:
            ref_digest = dg
        elif dg != ref_digest:
            raise RuntimeError(f"{cand} changed the result")
        results.append((st.route_ms, cand))
        log(f"[autotune] {cand}: route {st.route_ms:.2f} ms ({st.waves} waves)")
    del job
    best = min(results, key=lambda r: r[0])[1]
    if save:
        save_tuned(nl, best, path)
    return best


if __name__ == "__main__":
    import argparse

    from . import netgen
    ap = argparse.ArgumentParser(description="time launch parameters on one PathFinder iteration of a workload")
    ap.add_argument("workload", nargs="?", default="uneven_800")
    ap.add_argument("--deltas", default=None, help="comma-separated bucket widths")
    ap.add_argument("--blocks", default=None, help="comma-separated workgroup sizes")
    ap.add_argument("--lds-nodes", default=None, help="comma-separated LDS key capacities")
    ap.add_argument("--small-wave-nodes", default=None, help="comma-separated per-net engine thresholds")
    ap.add_argument("--no-save", action="store_true")
    a = ap.parse_args()
    kw = {}
    if a.deltas:
        kw["deltas"] = tuple(float(v) for v in a.deltas.split(","))
    if a.blocks:
        kw["blocks"] = tuple(int(v) for v in a.blocks.split(","))
    if a.lds_nodes:
        kw["lds_nodes"] = tuple(int(v) for v in a.lds_nodes.split(","))
    if a.small_wave_nodes:
        kw["small_wave_nodes"] = tuple(int(v) for v in a.small_wave_nodes.split(","))
    print(json.dumps(tune(getattr(netgen, a.workload)(), save=not a.no_save, log=lambda m: print(m, flush=True),
                          **kw)))
