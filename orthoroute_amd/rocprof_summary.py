"""Per-unit kernel summaries from rocprofv3 output (kernel trace + PMC passes).

Turns the CSV directories of one kernel-trace pass and the FETCH_SIZE /
WRITE_SIZE (and optional CU-busy) counter passes of the same command into a
JSON record and a Markdown table per unit of work:

* ``route``: one unit = one PathFinder iteration (one end_iteration_kernel
  dispatch); the route kernels are the tile engine's (tile_*) and the
  one-workgroup-per-net engine's (route_wave_kernel, ws_fill_kernel), i.e.
  everything a DevicePathFinder enqueues between the re-price and the
  end-of-iteration kernel;
* ``query``: one unit = one whole-box query (one tile_finish_kernel
  dispatch); the kernels are the tile engine's.

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what the same
counters read for a plain copy of known size (``bench.py --copy-probe``:
256 MiB read and written per copy), so calibrated bytes = FETCH x f + WRITE
x w.  The CU-busy pass (GRBM_GUI_ACTIVE, SQ_BUSY_CU_CYCLES, SQ_WAVE_CYCLES,
SQ_WAVES) gives waves per dispatch, the time a wave lives (SQ_WAVE_CYCLES is
in quad-cycles) and the share of CU cycles that are busy over the kernel's
traced duration (SQ_BUSY_CU_CYCLES summed over CUs, as rocprofv3's own
CU-utilisation metric divides it by GRBM cycles and CU count).

    python -m orthoroute_amd.rocprof_summary route RP_DIR OUT_STEM "title" \\
        [--algorithmic-bytes B] [--bench-ms T] [--commands FILE]

RP_DIR holds trace_*/fetch_*/write_*/busy_* (suffix c3 for route, c1 for
query) and copy_fetch/copy_write, as rocprofv3 ``-d`` wrote them.
"""
from __future__ import annotations

import argparse
import csv
import json
from collections import defaultdict
from pathlib import Path

COPY_KIB = 256 * 1024          # the copy probe moves 256 MiB each way per copy
ENGINE_MHZ = 2400.0            # MI355X peak engine clock (agent info Max_Engine_Clk_Fcompute)
UNIT_KERNEL = {"route": "end_iteration_kernel", "query": "tile_finish_kernel"}
PASS_SUFFIX = {"route": "c3", "query": "c1"}


def short_name(name: str) -> str:
    n = name.replace("void ", "").replace("(anonymous namespace)::", "")
    return n.split("(")[0].strip()


def is_route_kernel(kind: str, name: str) -> bool:
    base = name.split("<")[0]
    if kind == "query":
        return base.startswith("tile_") and base not in ("tile_finish_kernel", "tile_single_done_kernel")
    return base.startswith("tile_") or base in ("route_wave_kernel", "ws_fill_kernel")


def _csv(path: Path) -> list[dict]:
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def trace_stats(d: Path):
    rows = _csv(d / "run_kernel_trace.csv")
    calls, ns = defaultdict(int), defaultdict(float)
    for r in rows:
        n = short_name(r["Kernel_Name"])
        calls[n] += 1
        ns[n] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return calls, ns


def counter_sums(d: Path):
    """{kernel: {counter: sum over dispatches}}, {kernel: dispatches}."""
    rows = _csv(d / "run_counter_collection.csv")
    vals = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    for r in rows:
        n = short_name(r["Kernel_Name"])
        vals[n][r["Counter_Name"]] += float(r["Counter_Value"])
        disp[n].add(r["Dispatch_Id"])
    return vals, {k: len(v) for k, v in disp.items()}


def copy_factor(d: Path, counter: str) -> float:
    """KiB the copy of COPY_KIB really moved / KiB the counter read, from
    the kernel with the largest reading (the copy; the fill only writes)."""
    vals, disp = counter_sums(d)
    best = max(vals, key=lambda k: vals[k].get(counter, 0.0) / max(1, disp[k]))
    per = vals[best][counter] / disp[best]
    return COPY_KIB / per


def summarize(kind: str, rp: Path, title: str, algorithmic_bytes: float | None = None,
              bench_ms: float | None = None, commands: list[str] | None = None) -> dict:
    sfx = PASS_SUFFIX[kind]
    calls, ns = trace_stats(rp / f"trace_{sfx}")
    units = calls.get(UNIT_KERNEL[kind], 0)
    if units == 0:
        raise ValueError(f"no {UNIT_KERNEL[kind]} dispatch in the trace")
    fetch, fdisp = counter_sums(rp / f"fetch_{sfx}")
    write, wdisp = counter_sums(rp / f"write_{sfx}")
    f_units = fdisp.get(UNIT_KERNEL[kind], units)
    w_units = wdisp.get(UNIT_KERNEL[kind], units)
    busy, bdisp = ({}, {})
    if (rp / f"busy_{sfx}").exists():
        busy, bdisp = counter_sums(rp / f"busy_{sfx}")
    ff = copy_factor(rp / "copy_fetch", "FETCH_SIZE")
    wf = copy_factor(rp / "copy_write", "WRITE_SIZE")
    per = {}
    for n in sorted(calls):
        if not is_route_kernel(kind, n):
            continue
        c = calls[n] / units
        ms = ns[n] / 1e6 / units
        fk = fetch.get(n, {}).get("FETCH_SIZE", 0.0) / f_units
        wk = write.get(n, {}).get("WRITE_SIZE", 0.0) / w_units
        rec = {"calls_per_unit": c, "avg_us": ns[n] / 1e3 / calls[n], "ms_per_unit": ms,
               "fetch_kib_per_unit_raw": fk, "write_kib_per_unit_raw": wk,
               "hbm_bytes_per_unit": (fk * ff + wk * wf) * 1024.0}
        if n in busy and bdisp.get(n):
            b = busy[n]
            dsp = bdisp[n]
            avg_ns = ns[n] / calls[n]
            cu_cycles = avg_ns * 1e-9 * ENGINE_MHZ * 1e6 * 256
            rec["waves_per_dispatch"] = b.get("SQ_WAVES", 0.0) / dsp
            rec["cu_busy_frac"] = b.get("SQ_BUSY_CU_CYCLES", 0.0) / dsp / cu_cycles if cu_cycles else None
            rec["wave_us_per_wave"] = (4.0 * b.get("SQ_WAVE_CYCLES", 0.0) / max(1.0, b.get("SQ_WAVES", 0.0))
                                       / (ENGINE_MHZ * 1e6) * 1e6)
        per[n] = rec
    total_bytes = sum(r["hbm_bytes_per_unit"] for r in per.values())
    total_ms = sum(r["ms_per_unit"] for r in per.values())
    out = {"title": title, "kind": kind, "units_per_pass": units, "kernel_ms_per_unit_summed": total_ms,
           "hbm_bytes_per_unit": total_bytes, "fetch_calibration": ff, "write_calibration": wf,
           "per_kernel": per, "commands": commands or []}
    # the names bench.py reads
    out["hbm_bytes_per_launch" if kind == "route" else "hbm_bytes_per_query"] = total_bytes
    if algorithmic_bytes:
        out["algorithmic_bytes_per_unit"] = algorithmic_bytes
        out["traffic_over_algorithmic"] = total_bytes / algorithmic_bytes
    if bench_ms:
        out["bench_ms_per_unit"] = bench_ms
    return out


def markdown(rec: dict) -> str:
    unit = "iteration" if rec["kind"] == "route" else "query"
    lines = [f"# {rec['title']}", "",
             f"One unit = one {unit}; per-unit figures average the {rec['units_per_pass']} {unit}s of each pass.",
             "", "Commands (`$REPO` = repository root, each counter set in its own pass):", ""]
    lines += [f"    {c}" for c in rec["commands"]]
    lines += ["", f"Calibration: the copy probe moves 256 MiB each way; FETCH_SIZE x {rec['fetch_calibration']:.3f} + "
              f"WRITE_SIZE x {rec['write_calibration']:.3f} = calibrated HBM bytes.", "",
              f"| kernel | calls / {unit} | avg us | ms / {unit} | FETCH KiB raw | WRITE KiB raw | calibrated MB |"
              " waves / dispatch | CU busy | us per wave |",
              "|---|---|---|---|---|---|---|---|---|---|"]
    for n, r in rec["per_kernel"].items():
        cb = r.get("cu_busy_frac")
        lines.append(f"| {n} | {r['calls_per_unit']:.1f} | {r['avg_us']:.2f} | {r['ms_per_unit']:.2f} | "
                     f"{r['fetch_kib_per_unit_raw']:.0f} | {r['write_kib_per_unit_raw']:.0f} | "
                     f"{r['hbm_bytes_per_unit'] / 1e6:.1f} | {r.get('waves_per_dispatch', float('nan')):.0f} | "
                     f"{'' if cb is None else f'{100 * cb:.0f}%'} | {r.get('wave_us_per_wave', float('nan')):.1f} |")
    lines.append(f"| **sum** | | | **{rec['kernel_ms_per_unit_summed']:.2f}** | | | "
                 f"**{rec['hbm_bytes_per_unit'] / 1e6:.1f}** | | | |")
    if "algorithmic_bytes_per_unit" in rec:
        lines += ["", f"Algorithmic bytes per {unit} (12 B per relaxed arc + 16 B per settled node of the "
                  f"label-setting count): {rec['algorithmic_bytes_per_unit'] / 1e6:.1f} MB; calibrated traffic "
                  f"{rec['hbm_bytes_per_unit'] / 1e6:.1f} MB ({rec['traffic_over_algorithmic']:.1f}x)."]
    if "bench_ms_per_unit" in rec:
        lines += ["", f"Agreement with the bench: {rec['bench_ms_per_unit']:.2f} ms per {unit} measured by the "
                  f"bench against {rec['kernel_ms_per_unit_summed']:.2f} ms of summed kernel time here."]
    return "\n".join(lines) + "\n"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=("route", "query"))
    ap.add_argument("rp_dir")
    ap.add_argument("out_stem")
    ap.add_argument("title")
    ap.add_argument("--algorithmic-bytes", type=float, default=None)
    ap.add_argument("--bench-ms", type=float, default=None)
    ap.add_argument("--commands", default=None, help="text file, one command per line")
    a = ap.parse_args()
    cmds = Path(a.commands).read_text().strip().splitlines() if a.commands else []
    rec = summarize(a.kind, Path(a.rp_dir), a.title, a.algorithmic_bytes, a.bench_ms, cmds)
    Path(a.out_stem + ".json").write_text(json.dumps(rec, indent=1))
    Path(a.out_stem + ".md").write_text(markdown(rec))
    print(markdown(rec))


if __name__ == "__main__":
    main()
