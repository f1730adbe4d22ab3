"""Benchmark of the PathFinder hot path on one MI355X.

A step is one PathFinder iteration (rip-up, route and commit every net, then
the overuse / history update) over the 800x800x20 lattice with 8000
lognormal-length two-pin nets (seed 3), batched as disjoint-ROI waves.
W warm-up iterations run first, then K timed iterations; the timed region
starts with every array already resident in HBM.

Printed (one JSON line): nets routed/s of the batched form (``value``),
GTEPS, the plain one-net-per-launch form's iteration time on the same GPU
and the speed-up, the roofline of the route kernel (algorithmic bytes of the
label-setting search: 12 B per relaxed arc + 16 B per settled node, over the
kernel time measured with HIP events on its stream), and a CPU baseline (the
oracle, one core, on a bounded sample of the same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))

HBM_PEAK_GBS = 8000.0


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(nl, cfg) -> dict:
    """Oracle (C restatement of the reference CPU path), 1 core, on iteration 1
    of the same board restricted to every third net."""
    from oracle.oracle import OraclePathFinder
    sub_idx = list(range(0, nl.n, 3))
    from orthoroute_amd.netgen import Netlist
    sub = Netlist(nl.spec, nl.src[sub_idx], nl.dst[sub_idx], nl.box[sub_idx], nl.name + "[::3]")
    o = OraclePathFinder(sub, cfg)
    t0 = time.perf_counter()
    d = o.run_iteration()
    dt = time.perf_counter() - t0
    return {"value": d["routed"] / dt, "unit": "nets/s", "cores": 1, "kind": "port",
            "sample": f"iteration 1 of {sub.name} ({sub.n} nets, every third net of the workload), "
                      f"{dt:.1f} s on {_cpu_model()} (nproc {os.cpu_count()})",
            "gteps": d["relaxed"] / dt / 1e9}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uneven_800")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--delta", type=float, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--lds-nodes", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              f"(the path is single-GPU; sharding is described in DESIGN.md)", file=sys.stderr)
        return 2

    import torch
    from orthoroute_amd import _lib, netgen
    from orthoroute_amd.router import DevicePathFinder, RouterConfig
    from orthoroute_amd.autotune import load_tuned

    _lib.init_device()
    nl = getattr(netgen, args.workload)()
    tuned = load_tuned(nl)
    cfg = RouterConfig(iterations=args.warmup + args.steps,
                       delta=args.delta if args.delta is not None else tuned["delta"],
                       block=args.block if args.block is not None else tuned["block"],
                       lds_nodes=args.lds_nodes if args.lds_nodes is not None else tuned["lds_nodes"])

    dev = DevicePathFinder(nl, cfg)
    first = None
    for _ in range(args.warmup):
        st = dev.run_iteration("batched")
        first = first or st
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed = [dev.run_iteration("batched") for _ in range(args.steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    first = first or timed[0]
    wall = t1 - t0
    routed = sum(s.routed for s in timed)
    relaxed = sum(s.relaxed for s in timed)
    settled = sum(s.settled for s in timed)
    route_s = sum(s.route_ms for s in timed) / 1e3
    waves = sum(s.waves for s in timed)
    alg_bytes = 12 * relaxed + 16 * settled
    achieved = alg_bytes / route_s / 1e9
    footprint = dev.footprint_bytes()
    del dev
    torch.cuda.empty_cache()

    plain = None
    if not args.no_plain:
        pdev = DevicePathFinder(nl, RouterConfig(iterations=1, delta=cfg.delta, block=cfg.block,
                                                 lds_nodes=cfg.lds_nodes))
        pst = pdev.run_iteration("plain")
        plain = {"iteration_ms": pst.seconds * 1e3, "launches": pst.waves,
                 "batched_iteration_ms": first.seconds * 1e3,
                 "speedup": pst.seconds / first.seconds,
                 "bit_exact_iteration1": bool(pst.overuse == first.overuse and pst.routed == first.routed)}
        del pdev
        torch.cuda.empty_cache()

    cpu = None if args.no_cpu else cpu_baseline(nl, cfg)

    out = {
        "metric": "nets routed/s (batched disjoint-ROI PathFinder iteration)",
        "value": routed / wall,
        "unit": "nets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": nl.name, "lattice": [nl.spec.X, nl.spec.Y, nl.spec.L], "nets": nl.n,
                   "roi_margin": 4, "iterations_timed": [timed[0].iteration, timed[-1].iteration],
                   "delta": cfg.delta, "block": cfg.block, "lds_nodes": cfg.lds_nodes,
                   "parallelism": "1 GPU, disjoint-ROI waves"},
        "gteps": relaxed / route_s / 1e9,
        "waves_per_iteration": waves / args.steps,
        "overuse_last": timed[-1].overuse,
        "hbm_footprint_gb": footprint / 1e9,
        "plain_vs_batched": plain,
        "p99_job_latency_s": None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "route_wave_kernel", "algorithmic_bytes_per_iteration": alg_bytes / args.steps,
                     "kernel_ms_per_iteration": route_s * 1e3 / args.steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
