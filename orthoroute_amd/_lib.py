"""ctypes binding of liborx_hip.so (declarations in include/orthoroute_hip.h).

The library is loaded from the package directory.  There is no fallback: if
the shared object is missing or fails to load, every entry point raises
``HipUnavailable`` so a GPU run can never silently use another code path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "liborx_hip.so"

ORX_OK = 0
ORX_NO_PATH = 1
ORX_E_ARG = -1
ORX_E_HIP = -2
ORX_E_OVERFLOW = -3
ORX_E_CYCLE = -4
ORX_E_NODEVICE = -5

ORX_F_STATS = 1
ORX_F_FULL = 2
ORX_F_SEED_PEN = 4
ORX_F_GRID = 8
ORX_F_TILES = 16
ORX_F_ASYNC = 32
ORX_NOALLOC_MAX_NODES = 65536

STATUS_NAMES = {
    ORX_OK: "ok",
    ORX_NO_PATH: "no path",
    ORX_E_ARG: "bad argument",
    ORX_E_HIP: "HIP error",
    ORX_E_OVERFLOW: "overflow",
    ORX_E_CYCLE: "parent cycle",
    ORX_E_NODEVICE: "no gfx950 device",
}

P = C.c_void_p


class OrxLattice(C.Structure):
    _fields_ = [
        ("X", C.c_int32), ("Y", C.c_int32), ("L", C.c_int32), ("hmask", C.c_uint32),
        ("base", P), ("hist", P), ("occ", P), ("wcost", P),
        ("node_hist", P), ("node_use", P), ("via_col", P),
    ]


class OrxParams(C.Structure):
    _fields_ = [
        ("pres_fac", C.c_float), ("base_w", C.c_float), ("hist_w", C.c_float),
        ("node_pres", C.c_float), ("node_hist_w", C.c_float),
        ("edge_cap", C.c_int32), ("node_cap", C.c_int32),
        ("hist_gain", C.c_float), ("hist_decay", C.c_float),
        ("hist_cap_mult", C.c_float), ("node_hist_gain", C.c_float),
        ("astar_cp", C.c_float), ("astar_cv", C.c_float), ("astar_gmax", C.c_float),
    ]


class OrxNets(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("box", P), ("src", P), ("dst", P), ("path_arena", P),
        ("path_off", P), ("path_cap", P), ("path_len", P), ("path_cost", P), ("status", P),
        ("seed_ptr", P), ("seed_node", P), ("seed_cost", P), ("tgt_ptr", P), ("tgt_node", P), ("tgt_cost", P),
    ]


MAX_TARGETS = 128   # ORX_MAX_TARGETS


class OrxWorkspace(C.Structure):
    _fields_ = [("keys", P), ("queues", P), ("bitmaps", P), ("ws_nodes", C.c_int64)]


class OrxRouteCfg(C.Structure):
    _fields_ = [
        ("delta", C.c_float), ("block", C.c_int32), ("lds_nodes", C.c_int32),
        ("flags", C.c_int32), ("stats", P), ("max_groups", C.c_int32),
    ]


class OrxAcct(C.Structure):
    _fields_ = [
        ("E", C.c_int64), ("present", P), ("present_ema", P), ("history", P), ("capacity", P),
        ("total_cost", P), ("live_static", P), ("canon", P),
    ]


class OrxCsr(C.Structure):
    _fields_ = [
        ("N", C.c_int64), ("E", C.c_int64), ("indptr", P), ("indices", P),
        ("in_rank", P), ("in_indptr", P), ("in_src", P),
    ]


# (name, restype, argtypes) for every symbol the header declares.
SIGNATURES = [
    ("orx_init", C.c_int, []),
    ("orx_last_error", C.c_char_p, []),
    ("orx_version", C.c_int, []),
    ("orx_lattice_reprice", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxParams), P]),
    ("orx_lattice_end_iteration", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxParams), P, P]),
    ("orx_route_wave", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxParams), C.POINTER(OrxNets), P, P,
                                 C.c_int32, C.c_int64, C.POINTER(OrxWorkspace), C.POINTER(OrxRouteCfg), P]),
    ("orx_route_waves", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxParams), C.POINTER(OrxNets), P, P,
                                  C.c_int32, P, P, P, C.POINTER(OrxWorkspace), C.POINTER(OrxRouteCfg), P]),
    ("orx_route_tiles", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxParams), C.POINTER(OrxNets), P, P,
                                  C.c_int32, P, P, P, C.POINTER(OrxWorkspace), C.POINTER(OrxRouteCfg), P]),
    ("orx_graph_begin", C.c_int, [C.POINTER(C.c_void_p)]),
    ("orx_graph_end", C.c_int, [P, C.POINTER(C.c_void_p)]),
    ("orx_graph_launch", C.c_int, [P, P]),
    ("orx_graph_destroy", C.c_int, [P]),
    ("orx_route_dag", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxParams), C.POINTER(OrxNets), P, C.c_int32, P, P,
                                P, C.c_int32, P, P, C.c_int64, C.POINTER(OrxWorkspace), C.POINTER(OrxRouteCfg),
                                C.c_int32, P]),
    ("orx_plan_iteration", C.c_int64, [P, P, C.c_int32, C.c_int32, C.c_int32, P, P, P, P, P, C.c_int64]),
    ("orx_roi_bfs", C.c_int, [C.POINTER(OrxCsr), P, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_int64, P,
                              C.c_int32, P, P, C.c_int64, P, P, C.c_int64, P, P, P]),
    ("orx_roi_release", C.c_int, [P, P, C.c_int64, P, P]),
    ("orx_plan_waves", C.c_int64, [P, C.c_int32, P, C.c_int32, C.c_int32, C.c_int32, P, P, P]),
    ("orx_net_overuse", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxNets), C.c_int32, C.c_int32, P, P, P, P]),
    ("orx_apply_paths", C.c_int, [C.POINTER(OrxLattice), C.POINTER(OrxParams), P, P, P, C.c_int32,
                                  C.c_int32, P]),
    ("orx_sssp_box", C.c_int, [C.POINTER(OrxLattice), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64,
                               C.c_int64, P, P, C.POINTER(OrxWorkspace), C.POINTER(OrxRouteCfg), P]),
    ("orx_sssp_box_noalloc", C.c_int, [C.POINTER(OrxLattice), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int64, C.c_int64, P, P, C.POINTER(OrxRouteCfg), P]),
    ("orx_sssp_csr", C.c_int, [C.POINTER(OrxCsr), P, C.c_int64, P, P, P, P, C.c_int32, P, P, C.c_int32, P, P,
                               P, P, C.c_int64, P, C.POINTER(OrxRouteCfg), P]),
    ("orx_backtrace_csr", C.c_int, [C.POINTER(OrxCsr), P, P, P, P, C.c_int32, P, P]),
    ("orx_backtrace_csr_roi", C.c_int, [C.POINTER(OrxCsr), P, C.c_int64, P, P, P, P, C.c_int32, P, P, C.c_int64,
                                        P]),
    ("orx_acct_apply", C.c_int, [C.POINTER(OrxAcct), P, C.c_int64, C.c_int32, C.c_int32, P, P]),
    ("orx_acct_overuse", C.c_int, [C.POINTER(OrxAcct), P, P, P, P]),
    ("orx_acct_sum", C.c_int, [C.POINTER(OrxAcct), C.c_int32, P, P]),
    ("orx_acct_present_ema", C.c_int, [C.POINTER(OrxAcct), C.c_float, C.c_float, P]),
    ("orx_acct_history", C.c_int, [C.POINTER(OrxAcct), C.c_float, P, C.c_float, C.c_float, C.c_int32, P]),
    ("orx_acct_costs", C.c_int, [C.POINTER(OrxAcct), P, C.c_float, C.c_float, C.c_float, C.c_float, P, P,
                                 C.c_int32, P, P, C.c_int32, P, P]),
    ("orx_acct_begin_live", C.c_int, [C.POINTER(OrxAcct), C.c_float, P, P]),
    ("orx_acct_refresh_live", C.c_int, [C.POINTER(OrxAcct), P, C.c_int64, C.c_float, P, P, P]),
    ("orx_acct_from_canonical", C.c_int, [C.POINTER(OrxAcct), P]),
    ("orx_via_hard_block", C.c_int, [P, P, P, P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P, P, P, P,
                                     C.c_int64, P, P]),
    ("orx_via_penalties", C.c_int, [P, P, P, P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P, C.c_float,
                                    C.c_float, P, C.c_int64, P, P]),
]


class HipUnavailable(RuntimeError):
    """liborx_hip.so is not built / not loadable, or no gfx950 device."""


class HipError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed: {STATUS_NAMES.get(code, code)} ({code}): {msg}")
        self.code = code


_lib = None


def load(path: Path | None = None) -> C.CDLL:
    """Load liborx_hip.so and bind every declared symbol (no device needed)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise HipUnavailable(f"{p} is not built; run __graft_entry__.build() or python -m orthoroute_amd.build")
    try:
        lib = C.CDLL(str(p))
    except OSError as e:  # pragma: no cover - depends on the host
        raise HipUnavailable(f"cannot load {p}: {e}") from e
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def call(name: str, *args, allow=(ORX_OK,)) -> int:
    """Call an entry point and raise HipError on a status outside ``allow``."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc not in allow:
        msg = lib.orx_last_error()
        raise HipError(name, rc, msg.decode() if msg else "")
    return rc


_initialised = False


def init_device() -> None:
    """orx_init(): require a visible gfx950 device."""
    global _initialised
    if _initialised:
        return
    lib = load()
    rc = lib.orx_init()
    if rc != ORX_OK:
        raise HipUnavailable(lib.orx_last_error().decode())
    _initialised = True


def ptr(t) -> int:
    """Raw device (or host) address of a torch tensor / numpy array."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data
