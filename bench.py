"""Benchmark of the PathFinder hot path on one MI355X.

A step is one PathFinder iteration over the 800x800x20 lattice with 8000
lognormal-length two-pin nets (seed 3), batched as disjoint-ROI waves:
- rip-up, route and commit every net;
- then the overuse / history update.

W warm-up iterations run first, then K timed iterations. The timed region
starts with every array already resident in HBM.

Printed: one JSON line.
- ``value``: nets routed/s of the batched form.
- GTEPS.
- The plain one-net-per-launch form on the same GPU (iteration 1, bit-exact
  check of paths) and the speed-up.
- ``roofline`` of the route launches: algorithmic bytes of the search
  (12 B per relaxed arc + 16 B per settled node, the label-setting count)
  over the route time, measured with HIP events on the stream the waves run
  on.  With the tile engine one launch is one iteration's route kernels
  (orx_route_tiles for waves of large ROIs, orx_route_waves for the waves of
  small ones); ``traffic`` is the calibrated FETCH_SIZE + WRITE_SIZE of the
  same launch from the committed rocprofv3 summary.
- ``graph``: the timed iterations repeated from the same state with each
  iteration's launches captured into one hipGraph (capture included).
- ``cpu_baseline``: the oracle on one host core, on iteration 1 of the same
  workload, beside the device's own iteration-1 rate.
- ``sssp_microbench``: 64 full queries on the 1000x1000x12 lattice, one query
  per launch sequence; the oracle's heap Dijkstra on the first query is its
  CPU baseline and must give the same dist bits and parents.
- ``backplane_plain_vs_batched``: the 400x400x16 backplane, plain one net per
  launch against batched waves, iterations 1-2 in lockstep.
- ``seam``: the drop-in seam's find_path_fullgraph_gpu_seeds and
  find_path_roi_gpu on the CSR of a 400x400x16 lattice, with the oracle's
  heap search as CPU baseline and path check.
- ``serving``: a Poisson stream of routing jobs through the
  continuous-batching server, with throughput and p50/p99 job latency.

``--copy-probe`` only runs a plain device copy of known size. It is the
calibration run for rocprofv3 FETCH_SIZE / WRITE_SIZE counters.
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
# The serving leg keeps up to 8 jobs in flight, one HIP stream each; with
# fewer hardware queues than streams, jobs sharing a queue run in lockstep.
# Must be set before the HIP runtime initialises.
try:
    _queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
except ValueError:
    _queues = 0
if _queues < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

HBM_PEAK_GBS = 8000.0
SSSP_TRAFFIC_FILE = REPO / "profiles" / "sssp_tiles_traffic_1000x1000x12.json"
TRAFFIC_FILES = {"tiles": REPO / "profiles" / "route_traffic_uneven_800x800x20.json",
                 "wave": REPO / "profiles" / "route_wave_traffic_uneven_800x800x20.json"}
ROUTE_KERNELS = {"tiles": "the route kernels of one iteration: orx_route_tiles (tile_fill, tile_prep, tile_step, "
                          "tile_drain, tile_post) for waves whose largest ROI exceeds 16384 nodes, orx_route_waves "
                          "(ws_fill, route_wave_kernel) for the others",
                 "wave": "the route kernels of one iteration: orx_route_waves (ws_fill, route_wave_kernel) for "
                         "every wave"}


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(nl, cfg, stride: int) -> dict:
    """Oracle (C restatement of the reference CPU path), 1 core, iteration 1
    on every ``stride``-th net of the workload."""
    from oracle.oracle import OraclePathFinder
    sub = nl.select(list(range(0, nl.n, stride)), f"::{stride}") if stride > 1 else nl
    o = OraclePathFinder(sub, cfg)
    t0 = time.perf_counter()
    d = o.run_iteration()
    dt = time.perf_counter() - t0
    what = "all nets" if stride == 1 else f"every {stride}th net"
    return {"value": d["routed"] / dt, "unit": "nets/s", "cores": 1, "kind": "port",
            "sample": f"PathFinder iteration 1 of {nl.name} on {sub.n} nets ({what}), {dt:.1f} s "
                      f"on {_cpu_model()} (nproc {os.cpu_count()})",
            "gteps": d["relaxed"] / dt / 1e9, "seconds": dt}


def cpu_baseline_all_cores(nl, cfg) -> dict:
    """The oracle with the nets of each disjoint-ROI wave on all host cores
    (OpenMP), iteration 1 of the whole workload: the separately labelled
    all-core CPU figure beside the one-core baseline."""
    from oracle.oracle import OraclePathFinder
    threads = max(1, min(16, os.cpu_count() or 1))
    o = OraclePathFinder(nl, cfg)
    t0 = time.perf_counter()
    d = o.run_iteration_waves(threads)
    dt = time.perf_counter() - t0
    return {"value": d["routed"] / dt, "unit": "nets/s", "cores": threads, "kind": "port",
            "sample": f"PathFinder iteration 1 of {nl.name}, all {nl.n} nets, disjoint-ROI waves on {threads} "
                      f"threads, {dt:.1f} s on {_cpu_model()}", "seconds": dt}


def verify_lockstep(nl, cfg, iterations: int) -> dict:
    """Batched waves against the sequential per-net order (plain one net per
    launch) for `iterations` iterations: per-iteration counts and the digest
    of paths, fp32 costs, occupancy and node use must agree."""
    import dataclasses

    from orthoroute_amd.router import DevicePathFinder
    c = dataclasses.replace(cfg, iterations=iterations, collect_stats=False, stop_at_zero_overuse=False)
    seq, bat = DevicePathFinder(nl, c), DevicePathFinder(nl, c)
    rows = []
    for it in range(iterations):
        ss, sb = seq.run_iteration("plain"), bat.run_iteration("batched")
        same = (ss.overuse == sb.overuse and ss.routed == sb.routed and ss.via_hops == sb.via_hops
                and ss.path_hops == sb.path_hops and seq.digest() == bat.digest())
        rows.append({"iteration": it + 1, "overuse": sb.overuse, "plain_ms": round(ss.seconds * 1e3, 1),
                     "batched_ms": round(sb.seconds * 1e3, 1), "bit_exact": same})
        _log(f"lockstep {nl.name} iteration {it + 1}: overuse {sb.overuse}, plain {ss.seconds * 1e3:.1f} ms, "
             f"batched {sb.seconds * 1e3:.1f} ms, bit-exact {same}")
        if not same:
            break
    return {"workload": nl.name, "iterations": len(rows), "all_bit_exact": all(r["bit_exact"] for r in rows),
            "rows": rows}


def copy_probe(mib: int = 256, reps: int = 5) -> None:
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "reps": reps}))


def sssp_microbench(n_queries: int, delta: float, block: int, seed: int = 1, tiles: bool = True,
                    with_cpu: bool = True) -> dict:
    """Full queries on 1000x1000x12 (costs U[1,4) per physical edge, a seeded
    5% x10), endpoints uniform on layer 0, one query per launch sequence.
    With ``with_cpu``, the oracle (C heap Dijkstra, one core) runs the first
    query on the same weights: its time is the CPU baseline, and its dist
    bits and parents must equal the device's."""
    import numpy as np
    import torch
    from orthoroute_amd.lattice import LatticeSpec
    from orthoroute_amd.router import LatticeState, RouterConfig, Workspace
    from orthoroute_amd.sssp import sssp_box

    spec = LatticeSpec(1000, 1000, 12)
    base = spec.random_edge_costs(seed)
    lat = LatticeState(spec, base=base)
    lat.reprice(RouterConfig().params(0.5))
    ws = Workspace()
    ws.ensure(spec.N)
    rng = np.random.default_rng(seed + 100)
    pairs = rng.integers(0, spec.XY, size=(n_queries, 2))
    stats = torch.zeros(8, dtype=torch.int64, device="cuda")
    box = (0, 0, spec.X, spec.Y)
    # warm-up query (not timed)
    sssp_box(lat, box, int(pairs[0, 0]), int(pairs[0, 1]), full=True, delta=delta, block=block, ws=ws,
             stats=stats, grid=not tiles, tiles=tiles)
    stats.zero_()
    torch.cuda.synchronize()
    times = []
    check = None
    first = None
    for q in range(n_queries):
        t = time.perf_counter()
        dist, parent, rc = sssp_box(lat, box, int(pairs[q, 0]), int(pairs[q, 1]), full=True, delta=delta,
                                    block=block, ws=ws, stats=stats, grid=not tiles, tiles=tiles)
        times.append(time.perf_counter() - t)
        if q == 0:
            check = _check_path_cost(lat, dist, parent, int(pairs[q, 0]), int(pairs[q, 1]))
            first = (dist.cpu().numpy(), parent.cpu().numpy())
    st = stats.cpu().numpy()
    total = float(sum(times))
    relaxed, settled = int(st[1]), int(st[0])
    alg = 12 * relaxed + 16 * settled
    engine = ("tile-step search (orx_sssp_box with ORX_F_TILES: tile_fill, tile_step, tile_drain kernels)"
              if tiles else "grid-wide state machine (orx_sssp_box with ORX_F_GRID: grid_step_kernel)")
    out = {
        "workload": "sssp_1000x1000x12_full_query", "queries": n_queries, "delta": delta, "block": block,
        "engine": engine,
        "ms_per_query": total / n_queries * 1e3, "p50_ms": float(np.percentile(times, 50)) * 1e3,
        "gteps": relaxed / total / 1e9, "relaxed_per_query": relaxed / n_queries,
        "settled_per_query": settled / n_queries, "rounds_per_query": int(st[3]) / n_queries,
        "path_cost_check": check,
        "roofline": {"bound": "hbm", "achieved": alg / total / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / total / 1e9 / HBM_PEAK_GBS, "kernel": engine,
                     "algorithmic_bytes_per_query": alg / n_queries,
                     "timing": "host clock around each synchronous query (all its launches)",
                     "traffic": None, "traffic_source": None},
    }
    if tiles:
        try:
            t = json.loads(SSSP_TRAFFIC_FILE.read_text())
            out["roofline"]["traffic"] = float(t["hbm_bytes_per_query"])
            out["roofline"]["traffic_source"] = str(SSSP_TRAFFIC_FILE.relative_to(REPO))
        except (OSError, ValueError, KeyError):
            pass
    if with_cpu:
        from oracle.oracle import sssp_box as oracle_sssp_box
        w = lat.wcost.cpu().numpy()
        t = time.perf_counter()
        od, op, _, o_settled, o_relaxed = oracle_sssp_box(spec, w, box, int(pairs[0, 0]), int(pairs[0, 1]),
                                                          full=True)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": dt * 1e3, "unit": "ms/query", "cores": 1, "kind": "port",
                               "gteps": o_relaxed / dt / 1e9,
                               "sample": f"the first query (full 12 M-node box), oracle heap Dijkstra, {dt:.1f} s "
                                         f"on {_cpu_model()}"}
        out["same_as_oracle"] = bool(np.array_equal(od.view(np.uint32), first[0].view(np.uint32)) and
                                     np.array_equal(op, first[1]))
        out["speedup_vs_cpu"] = dt * 1e3 / out["ms_per_query"]
    return out


def seam_bench(n_queries: int = 8, seed: int = 11, with_cpu: bool = True) -> dict:
    """The drop-in seam (HipDijkstra) on the CSR of a 400x400x16 lattice
    (2.56 M nodes, 9.9 M arcs; U[1,4) costs, 5% x10, node penalties in
    [0, 0.5)), costs and penalties resident on the device:
    - find_path_fullgraph_gpu_seeds: 6 priced seeds within 40 columns of the
      left edge, 6 priced targets within 40 of the right edge, the whole-chip
      form and the parallel parent walk;
    - find_path_roi_gpu: a 150x150-cell box ROI (360 K nodes) between two
      corners, src and dst inside.
    Each query is synchronous (it returns the host path).  The oracle's heap
    search on the first full-graph query (one core) gives the CPU baseline,
    the label-setting counts for the algorithmic bytes (12 B per relaxed arc
    + 4 B of its penalty + 16 B per settled node), and the path to compare."""
    import numpy as np
    import torch
    from orthoroute_amd.lattice import LatticeSpec, in_rank_and_transpose
    from orthoroute_amd.solver import HipDijkstra

    spec = LatticeSpec(400, 400, 16)
    indptr, indices, _, _ = spec.to_csr()
    rng = np.random.default_rng(seed)
    E = len(indices)
    costs = rng.uniform(1.0, 4.0, E).astype(np.float32)
    costs[rng.random(E) < 0.05] *= np.float32(10.0)
    pen = rng.uniform(0.0, 0.5, spec.N).astype(np.float32)
    hip = HipDijkstra(indptr=indptr, indices=indices)
    d_costs = torch.from_numpy(costs).cuda()
    d_pen = torch.from_numpy(pen).cuda()

    def side(lo, hi, k):
        return np.array([spec.node(int(x), int(y), int(z)) for x, y, z in
                         zip(rng.integers(lo, hi, k), rng.integers(0, spec.Y, k), rng.integers(0, spec.L, k))],
                        np.int32)

    queries = [(side(5, 40, 6), side(360, 395, 6), rng.uniform(0, 2, 6).astype(np.float32),
                rng.uniform(0, 2, 6).astype(np.float32)) for _ in range(n_queries + 1)]

    def full(q):
        s, tg, sc, tc = q
        return hip.find_path_fullgraph_gpu_seeds(costs=d_costs, src_seeds=s, dst_targets=tg, src_seed_costs=sc,
                                                 dst_target_costs=tc, node_penalty=d_pen)

    full(queries[0])                      # warm-up (buffers, graph set-up)
    torch.cuda.synchronize()
    times, paths = [], []
    for q in queries[1:]:
        t0 = time.perf_counter()
        paths.append(full(q))
        times.append(time.perf_counter() - t0)
    out = {"workload": "lattice_csr_400x400x16", "nodes": spec.N, "arcs": E,
           "fullgraph_seeds": {"queries": n_queries, "ms_per_query": float(np.mean(times)) * 1e3,
                               "p50_ms": float(np.median(times)) * 1e3,
                               "mean_path_nodes": float(np.mean([len(p) for p in paths if p]))}}
    # ROI form: a 150 x 150 box (all layers) between two corners
    x0, y0, w = 120, 130, 150
    xs, ys, zs = np.meshgrid(np.arange(x0, x0 + w), np.arange(y0, y0 + w), np.arange(spec.L), indexing="ij")
    roi_nodes = np.sort(spec.node(xs.ravel(), ys.ravel(), zs.ravel()).astype(np.int32))
    g2r = np.full(spec.N, -1, np.int32)
    g2r[roi_nodes] = np.arange(len(roi_nodes), dtype=np.int32)
    d_roi, d_g2r = torch.from_numpy(roi_nodes).cuda(), torch.from_numpy(g2r).cuda()
    d_roi_pen = torch.from_numpy(pen[roi_nodes]).cuda()
    src, dst = int(spec.node(x0 + 2, y0 + 3, 0)), int(spec.node(x0 + w - 3, y0 + w - 2, 0))
    hip.find_path_roi_gpu(src, dst, d_costs, d_roi, d_g2r, node_penalty=d_roi_pen)
    torch.cuda.synchronize()
    rt = []
    for _ in range(n_queries):
        t0 = time.perf_counter()
        rp = hip.find_path_roi_gpu(src, dst, d_costs, d_roi, d_g2r, node_penalty=d_roi_pen)
        rt.append(time.perf_counter() - t0)
    out["roi"] = {"roi_nodes": int(len(roi_nodes)), "queries": n_queries,
                  "ms_per_query": float(np.mean(rt)) * 1e3, "path_nodes": len(rp) if rp else 0}
    if with_cpu:
        from oracle.oracle import csr_path, sssp_csr
        rank, in_indptr, in_src = in_rank_and_transpose(spec.N, indptr, indices)
        s, tg, sc, tc = queries[1]
        t0 = time.perf_counter()
        rc, best, _, parent, (settled, relaxed) = sssp_csr(spec.N, indptr, indices, costs, rank, in_indptr, in_src,
                                                            pen=pen, seeds=s, seed_cost=sc, targets=tg, tg_cost=tc,
                                                            counts=True)
        dt = time.perf_counter() - t0
        alg = 16 * relaxed + 16 * settled
        dev_s = times[0]
        out["fullgraph_seeds"].update({
            "same_as_oracle": csr_path(best, parent) == paths[0],
            "algorithmic_bytes_first_query": alg, "settled_first_query": settled, "relaxed_first_query": relaxed,
            "gteps_first_query": relaxed / dev_s / 1e9,
            "roofline_first_query": {"bound": "hbm", "achieved": alg / dev_s / 1e9, "peak": HBM_PEAK_GBS,
                                     "unit": "GB/s", "frac": alg / dev_s / 1e9 / HBM_PEAK_GBS,
                                     "timing": "host clock around the synchronous query (search, walk, path copy)"},
            "cpu_baseline": {"value": dt * 1e3, "unit": "ms/query", "cores": 1, "kind": "port",
                             "sample": f"the first timed query, oracle heap search on one core, {dt:.2f} s on "
                                       f"{_cpu_model()}"}})
    return out


def _check_path_cost(lat, dist, parent, src: int, dst: int) -> bool:
    """The sink's fp32 distance equals the fp32 sum of entering weights along
    its parent chain (a size-independent property of the result)."""
    import numpy as np
    spec = lat.spec
    d = dist.cpu().numpy()
    p = parent.cpu().numpy()
    w = lat.wcost.cpu().numpy().reshape(-1, 4)
    nb = spec.neighbours()
    path = [dst]
    while path[-1] != src:
        u = int(p[path[-1]])
        if u < 0 or len(path) > spec.N:
            return False
        path.append(u)
    path.reverse()
    acc = np.float32(0.0)
    for u, v in zip(path[:-1], path[1:]):
        s = int(np.flatnonzero(nb[u] == v)[0])
        acc = np.float32(acc + w[u, s])
    return bool(acc == d[dst])


def serving_bench(n_jobs: int, rate: float, seed: int, max_iterations: int, tuned: dict,
                  time_limit: float) -> dict:
    from orthoroute_amd.router import RouterConfig
    from orthoroute_amd.serve import RoutingServer, poisson_jobs

    t0 = time.perf_counter()
    jobs = poisson_jobs(n_jobs, rate, seed=seed, max_iterations=max_iterations)
    gen_s = time.perf_counter() - t0
    cfg = RouterConfig(iterations=max_iterations, stop_at_zero_overuse=True, delta=tuned["delta"],
                       block=tuned["block"], lds_nodes=tuned["lds_nodes"])
    srv = RoutingServer(cfg, max_slots=8, progress=_log)
    summ = srv.run(jobs, time_limit=time_limit)
    summ.update({
        "arrival_rate_jobs_per_s": rate, "seed": seed, "max_iterations": max_iterations,
        "job_generation_s": gen_s,
        "jobs_detail": [{"board": j.netlist.name, "arrival_s": round(j.arrival, 3),
                         "latency_s": round(j.latency, 3), "iterations": j.iterations,
                         "final_overuse": j.overuse_history[-1] if j.overuse_history else None}
                        for j in sorted(srv.finished, key=lambda j: j.job_id)],
    })
    return summ


def portal_bench(nl, cfg, candidates: int, steps: int, cpu_stride: int, with_cpu: bool) -> dict:
    """The workload in portal form (every pad a seed or target on each inner
    layer of ``candidates`` escape cells): one warm-up and ``steps`` timed
    batched iterations, and the oracle on one core on every
    ``cpu_stride``-th net of iteration 1."""
    import dataclasses

    import torch
    from orthoroute_amd.netgen import with_portals
    from orthoroute_amd.router import DevicePathFinder
    pnl = with_portals(nl, candidates=candidates, seed=0)
    ns, nt = pnl.terms.counts()
    c = dataclasses.replace(cfg, iterations=1 + steps, collect_stats=False)
    dev = DevicePathFinder(pnl, c)
    warm = dev.run_iteration("batched")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed = [dev.run_iteration("batched") for _ in range(steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    out = {"workload": pnl.name, "seeds_per_net": float(ns.mean()), "targets_per_net": float(nt.mean()),
           "value": sum(s.routed for s in timed) / wall, "unit": "nets/s", "ms_per_iteration": wall / steps * 1e3,
           "iteration1_ms": warm.seconds * 1e3, "overuse_last": timed[-1].overuse,
           "waves_per_iteration": sum(s.waves for s in timed) / steps}
    del dev
    torch.cuda.empty_cache()
    if with_cpu:
        from oracle.oracle import OraclePathFinder
        sub = pnl.select(list(range(0, pnl.n, cpu_stride)), f"::{cpu_stride}")
        o = OraclePathFinder(sub, dataclasses.replace(cfg, iterations=1))
        t0 = time.perf_counter()
        d = o.run_iteration()
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": d["routed"] / dt, "unit": "nets/s", "cores": 1, "kind": "port",
                               "sample": f"iteration 1 of every {cpu_stride}th net ({sub.n} nets), {dt:.1f} s"}
    return out


def convergence_bench(nl, cfg, max_iterations: int) -> dict:
    """Iterations and wall time from scratch to zero overuse, rerouting every
    net (the plain schedule) and in hotset mode (offenders only)."""
    import dataclasses

    import torch
    from orthoroute_amd.router import DevicePathFinder
    out = {}
    for name, hot in (("reroute_all", False), ("hotset", True)):
        c = dataclasses.replace(cfg, iterations=max_iterations, collect_stats=False, hotset=hot)
        dev = DevicePathFinder(nl, c)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        routed = 0
        for _ in range(max_iterations):
            st = dev.run_iteration("batched")
            routed += st.routed
            if st.overuse == 0:
                break
        torch.cuda.synchronize()
        out[name] = {"iterations": dev.iteration, "seconds": time.perf_counter() - t0, "nets_rerouted": routed,
                     "overuse_last": st.overuse, "converged": st.overuse == 0}
        del dev
        torch.cuda.empty_cache()
    return out


def roi_bench(n: int = 32, samples: int = 4, with_cpu: bool = True) -> dict:
    """Device BFS regions of interest (extract_roi_bfs) on the 400x400x16
    lattice: pins on layer 0 at offset (30, 20), radius 40; ms per
    extraction including the host read of its size, and the oracle (numpy
    restatement, one core) on the first ``samples`` pairs, which must give
    the same ROIs."""
    import numpy as np
    import torch
    from orthoroute_amd.lattice import LatticeSpec
    from orthoroute_amd.roi import DeviceRoiBfs
    from orthoroute_amd.solver import HipDijkstra
    spec = LatticeSpec(400, 400, 16)
    ip, ix, _, _ = spec.to_csr()
    solver = HipDijkstra(indptr=ip, indices=ix)
    r = DeviceRoiBfs(solver)
    rng = np.random.default_rng(2)
    pairs = []
    for _ in range(n):
        x, y = int(rng.integers(0, 370)), int(rng.integers(0, 380))
        pairs.append((int(spec.node(x, y, 0)), int(spec.node(x + 30, y + 20, 0))))
    r.extract_roi_bfs(*pairs[0], initial_radius=40)          # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sizes = []
    for s_, d_ in pairs:
        roi, _ = r.extract_roi_bfs(s_, d_, initial_radius=40)
        sizes.append(int(roi.numel()))
    dt = (time.perf_counter() - t0) / n
    out = {"lattice": [400, 400, 16], "extractions": n, "ms_per_extraction": dt * 1e3,
           "mean_roi_nodes": float(np.mean(sizes)), "depth_last": r.last["depth"]}
    if with_cpu:
        from oracle.oracle import roi_bfs
        same = True
        cpu = 0.0
        for s_, d_ in pairs[:samples]:
            t0 = time.perf_counter()
            oroi, _, _ = roi_bfs(ip, ix, spec.N, s_, d_, initial_radius=40)
            cpu += time.perf_counter() - t0
            roi, _ = r.extract_roi_bfs(s_, d_, initial_radius=40)
            same &= bool(np.array_equal(roi.cpu().numpy(), oroi))
        cpu /= samples
        out["cpu_baseline"] = {"ms_per_extraction": cpu * 1e3, "cores": 1, "kind": "port",
                               "sample": f"the first {samples} pairs, numpy restatement (the oracle's time only)"}
        out["same_as_oracle"] = same
    del r, solver
    torch.cuda.empty_cache()
    return out


def _traffic(engine: str):
    """HBM bytes per route launch (the tile engine: one iteration's route
    kernels; the wave engine: one route_wave_kernel dispatch) from the
    committed rocprofv3 PMC summary (FETCH_SIZE +
    WRITE_SIZE, calibrated by a copy of known size), or None."""
    try:
        t = json.loads(TRAFFIC_FILES[engine].read_text())
        return float(t["hbm_bytes_per_launch"]), t
    except (OSError, ValueError, KeyError):
        return None, None


def backplane_bench(cfg, iterations: int) -> dict:
    """The 400x400x16 backplane (1600 long nets + 400 short ones), plain one
    net per launch against the batched waves, lockstep for ``iterations``
    iterations with a bit-exact check of every iteration."""
    import dataclasses

    from orthoroute_amd.netgen import backplane_400
    nl = backplane_400()
    res = verify_lockstep(nl, dataclasses.replace(cfg, iterations=iterations), iterations)
    for r in res["rows"]:
        r["speedup"] = r["plain_ms"] / r["batched_ms"] if r["batched_ms"] else None
    return res


def graph_leg(dev, snap, steps: int, digest: str) -> dict:
    """The timed iterations again from the same state, each iteration's
    route launches captured into one hipGraph and replayed (capture and
    instantiation inside the timed region, since the wave plan changes every
    iteration); the result must be the direct launches'."""
    import torch
    dev.restore(snap)
    dev.cfg.collect_stats = False
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = [dev.run_iteration("batched", graph=True) for _ in range(steps)]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return {"ms_per_step": wall / steps * 1e3, "value": sum(x.routed for x in st) / wall, "unit": "nets/s",
            "route_ms_per_step": sum(x.route_ms for x in st) / steps,
            "same_result_as_direct": dev.digest() == digest}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="uneven_800")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-stride", type=int, default=1)
    ap.add_argument("--no-sssp", action="store_true")
    ap.add_argument("--sssp-queries", type=int, default=64)
    ap.add_argument("--sssp-delta", type=float, default=16.0)
    ap.add_argument("--sssp-only", action="store_true",
                    help="only run the 1000x1000x12 query microbenchmark and print its JSON (profiling runs)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=6)
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--serve-iterations", type=int, default=50)
    ap.add_argument("--serve-seed", type=int, default=4)
    ap.add_argument("--serve-time-limit", type=float, default=300.0)
    ap.add_argument("--no-roi", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--converge-iterations", type=int, default=30)
    ap.add_argument("--no-portal", action="store_true")
    ap.add_argument("--portal-candidates", type=int, default=1)
    ap.add_argument("--portal-steps", type=int, default=2)
    ap.add_argument("--portal-cpu-stride", type=int, default=4)
    ap.add_argument("--graph", action="store_true", help="capture each iteration's wave launches in a hipGraph")
    ap.add_argument("--no-graph-leg", action="store_true",
                    help="skip the timed iterations repeated with hipGraph capture")
    ap.add_argument("--no-backplane", action="store_true")
    ap.add_argument("--no-seam", action="store_true")
    ap.add_argument("--backplane-iterations", type=int, default=2)
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--no-count", action="store_true",
                    help="skip the counting replay (no algorithmic bytes): for counter passes over the timed kernel")
    ap.add_argument("--verify-lockstep", type=int, default=0, metavar="N",
                    help="only run N iterations of batched vs sequential order on --workload and print the result")
    ap.add_argument("--engine", default="tiles", choices=("tiles", "wave"),
                    help="search engine: tile-step (default) or one workgroup per net")
    ap.add_argument("--delta", type=float, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--lds-nodes", type=int, default=None)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: the route path runs on a "
              f"single GPU (sharding it across GPUs is designed, not implemented)", file=sys.stderr)
        return 2
    if args.copy_probe:
        copy_probe()
        return 0

    import torch
    from orthoroute_amd import _lib, netgen
    from orthoroute_amd.autotune import load_tuned
    from orthoroute_amd.router import DevicePathFinder, RouterConfig

    _lib.init_device()
    if args.sssp_only:
        print(json.dumps(sssp_microbench(args.sssp_queries, args.sssp_delta, 256, tiles=args.engine == "tiles",
                                         with_cpu=not args.no_cpu)))
        return 0
    nl = getattr(netgen, args.workload)()
    tuned = load_tuned(nl)
    if args.verify_lockstep:
        res = verify_lockstep(nl, RouterConfig(delta=tuned["delta"], block=tuned["block"],
                                               lds_nodes=tuned["lds_nodes"]), args.verify_lockstep)
        print(json.dumps(res))
        return 0 if res["all_bit_exact"] else 1
    delta = args.delta if args.delta is not None else tuned["delta"]
    block = args.block if args.block is not None else tuned["block"]
    lds_nodes = args.lds_nodes if args.lds_nodes is not None else tuned["lds_nodes"]
    cfg = RouterConfig(iterations=args.warmup + args.steps, delta=delta, block=block, lds_nodes=lds_nodes,
                       graph=args.graph, engine=args.engine, small_wave_nodes=int(tuned["small_wave_nodes"]))

    _log(f"{nl.name}: {args.warmup} warm-up + {args.steps} timed iterations (delta {delta}, block {block})")
    dev = DevicePathFinder(nl, cfg)
    dev.cfg.collect_stats = False
    first = None
    first_paths = None
    for _ in range(args.warmup):
        st = dev.run_iteration("batched")
        if first is None:
            first, first_paths = st, dev.paths()
    # The timed iterations run without the settled / relaxed counting pass;
    # the counts come from a replay of the same iterations with it on (every
    # configuration computes the same result, so the counts are those of the
    # timed work).
    snap = None if args.no_count else dev.snapshot()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed = [dev.run_iteration("batched") for _ in range(args.steps)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    digest_timed = dev.digest()
    counted = None
    if snap is not None:
        dev.restore(snap)
        dev.cfg.collect_stats = True
        counted = [dev.run_iteration("batched") for _ in range(args.steps)]
        if dev.digest() != digest_timed:
            raise RuntimeError("the counting replay did not reproduce the timed iterations")
    graph = None
    if snap is not None and not args.no_graph_leg and not args.graph:
        _log("the timed iterations again with each iteration's launches captured in a hipGraph")
        graph = graph_leg(dev, snap, args.steps, digest_timed)
        _log(f"graph: {graph['ms_per_step']:.1f} ms/iteration, same result {graph['same_result_as_direct']}")
    if first is None:
        first = timed[0]
    wall = t1 - t0
    routed = sum(s.routed for s in timed)
    route_s = sum(s.route_ms for s in timed) / 1e3
    waves = sum(s.waves for s in timed)
    relaxed = settled = alg_bytes = achieved = None
    if counted is not None:
        relaxed = sum(s.relaxed for s in counted)
        settled = sum(s.settled for s in counted)
        alg_bytes = 12 * relaxed + 16 * settled
        achieved = alg_bytes / route_s / 1e9
    footprint = dev.footprint_bytes()
    _log(f"batched: {wall / args.steps * 1e3:.1f} ms/iteration, overuse {timed[-1].overuse}")
    del dev
    torch.cuda.empty_cache()

    plain = None
    if not args.no_plain:
        _log("plain one-net-per-launch iteration 1")
        pdev = DevicePathFinder(nl, RouterConfig(iterations=1, delta=delta, block=block, lds_nodes=lds_nodes,
                                                 engine=args.engine,
                                                 collect_stats=False))
        pst = pdev.run_iteration("plain")
        same = None
        if first_paths is not None:
            pp = pdev.paths()
            same = bool(len(pp) == len(first_paths) and all((a.shape == b.shape) and (a == b).all()
                                                            for a, b in zip(pp, first_paths)))
        plain = {"iteration_ms": pst.seconds * 1e3, "launches": pst.waves,
                 "batched_iteration_ms": first.seconds * 1e3, "speedup": pst.seconds / first.seconds,
                 "route_kernel_ms_plain": pst.route_ms, "route_kernel_ms_batched": first.route_ms,
                 "bit_exact_iteration1": bool(same and pst.overuse == first.overuse and pst.routed == first.routed)}
        del pdev
        torch.cuda.empty_cache()

    sssp = None
    if not args.no_sssp:
        _log(f"SSSP microbench: {args.sssp_queries} queries on 1000x1000x12")
        sssp = sssp_microbench(args.sssp_queries, args.sssp_delta, 256, tiles=args.engine == "tiles",
                               with_cpu=not args.no_cpu)
        torch.cuda.empty_cache()

    backplane = None
    if not args.no_backplane and args.backplane_iterations > 0:
        _log(f"backplane 400x400x16: plain vs batched, {args.backplane_iterations} iterations lockstep")
        backplane = backplane_bench(RouterConfig(delta=delta, block=block, lds_nodes=lds_nodes, engine=args.engine),
                                    args.backplane_iterations)
        torch.cuda.empty_cache()

    seam = None
    if not args.no_seam:
        _log("seam: find_path_fullgraph_gpu_seeds / find_path_roi_gpu on a 400x400x16 lattice CSR")
        seam = seam_bench(with_cpu=not args.no_cpu)
        _log(f"seam: {seam['fullgraph_seeds']['ms_per_query']:.2f} ms per full-graph query, "
             f"{seam['roi']['ms_per_query']:.2f} ms per ROI query")
        torch.cuda.empty_cache()

    serving = None
    if not args.no_serve:
        _log(f"serving: {args.serve_jobs} jobs at {args.serve_rate} jobs/s")
        serving = serving_bench(args.serve_jobs, args.serve_rate, args.serve_seed, args.serve_iterations, tuned,
                                time_limit=args.serve_time_limit)
        torch.cuda.empty_cache()

    roi = None
    if not args.no_roi:
        _log("device BFS regions of interest on 400x400x16")
        roi = roi_bench(with_cpu=not args.no_cpu)
        _log(f"roi: {roi['ms_per_extraction']:.2f} ms per extraction, {roi['mean_roi_nodes']:.0f} nodes")

    converge = None
    if not args.no_converge:
        _log(f"convergence: up to {args.converge_iterations} iterations, reroute-all and hotset")
        converge = convergence_bench(nl, cfg, args.converge_iterations)
        _log("convergence: " + ", ".join(f"{k} {v['iterations']} it {v['seconds']:.2f} s" for k, v in converge.items()))

    portal = None
    if not args.no_portal and nl.spec.L >= 3:
        _log(f"portal form: {args.portal_candidates} escape cell(s) per pad, {args.portal_steps} timed iterations")
        portal = portal_bench(nl, cfg, args.portal_candidates, args.portal_steps, args.portal_cpu_stride,
                              not args.no_cpu)
        _log(f"portal form: {portal['ms_per_iteration']:.1f} ms/iteration")

    cpu = cpu_all = None
    if not args.no_cpu:
        _log("CPU baseline (oracle, 1 core)")
        cpu = cpu_baseline(nl, cfg, args.cpu_stride)
        cpu["device_iteration1"] = {"value": first.routed / first.seconds, "unit": "nets/s",
                                    "ms": first.seconds * 1e3,
                                    "note": "the device on the same iteration as the CPU sample (iteration 1, "
                                            "warm-up, host clock)"}
        _log("CPU baseline (oracle, all cores over disjoint-ROI waves)")
        cpu_all = cpu_baseline_all_cores(nl, cfg)

    traffic, traffic_src = _traffic(args.engine)
    # one launch: one iteration's route kernels (either engine)
    n_dispatch = args.steps
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
                   "delta": delta, "block": block, "lds_nodes": lds_nodes, "graph": args.graph,
                   "engine": args.engine,
                   "parallelism": "1 GPU, disjoint-ROI waves"},
        "gteps": None if relaxed is None else relaxed / route_s / 1e9,
        "waves_per_iteration": waves / args.steps,
        "overuse_last": timed[-1].overuse,
        "hbm_footprint_gb": footprint / 1e9,
        "plain_vs_batched": plain,
        "graph": graph,
        "backplane_plain_vs_batched": backplane,
        "p99_job_latency_s": serving["p99_latency_s"] if serving else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": None if achieved is None else achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": ROUTE_KERNELS[args.engine],
                     "algorithmic_bytes_per_launch": None if alg_bytes is None else alg_bytes / n_dispatch,
                     "kernel_ms_per_launch": route_s * 1e3 / n_dispatch,
                     "launches_per_iteration": n_dispatch / args.steps,
                     "timing": "HIP events on the job's stream around the route launches of each iteration",
                     "traffic_source": str(TRAFFIC_FILES[args.engine].relative_to(REPO)) if traffic_src else None},
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "sssp_microbench": sssp,
        "serving": serving,
        "seam": seam,
        "portal_nets": portal,
        "convergence": converge,
        "roi_bfs": roi,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
