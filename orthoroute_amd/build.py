"""Build liborx_hip.so (the gfx950 kernels + C ABI) and the CPU oracle.

The library is compiled in-tree with hipcc for gfx950 only.  Flags that
matter for parity:
  -ffp-contract=off      every fp32 add/mul rounds once, in source order, so the
                         cost formulas and path sums match the CPU oracle bit
                         for bit;
  -target-feature -packed-fp32-ops
                         no packed fp32 instructions are generated.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
BUILD_DIR = REPO_DIR / "_build"
LIB_PATH = PKG_DIR / "liborx_hip.so"
ORACLE_DIR = REPO_DIR / "oracle"

SOURCES = ["orx_capi.hip", "orx_lattice.hip", "orx_tiles.hip", "orx_csr.hip", "orx_roi.hip", "orx_accountant.hip",
           "orx_plan.cpp"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(cmd)}")


def build_hip(verbose: bool = False) -> Path:
    """Compile every csrc/*.hip for gfx950 and link liborx_hip.so in-tree."""
    hipcc = _hipcc()
    BUILD_DIR.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    missing = [str(s) for s in srcs if not s.exists()]
    if missing:
        raise FileNotFoundError(f"HIP sources missing: {missing}")
    objs = [BUILD_DIR / (s.stem + ".o") for s in srcs]

    def compile_one(pair):
        src, obj = pair
        if src.suffix == ".cpp":  # host-only planner
            cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
        else:
            cmd = [hipcc, *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    with ThreadPoolExecutor(max_workers=min(4, len(srcs))) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    _run(["make", "-s", "-C", str(ORACLE_DIR)])
    return ORACLE_DIR / "liborx_oracle.so"


def main() -> None:
    build_hip(verbose="-v" in sys.argv)
    build_oracle()
    print(f"built {LIB_PATH}")


if __name__ == "__main__":
    main()
