"""Per-arc accountant on the reference's CSR arc order, resident in HBM.

``DeviceEdgeAccountant`` is EdgeAccountant (unified_pathfinder.py:1039-1422)
with the same attributes and methods, its arrays held as float32 device
tensors and every array operation run by liborx_hip (orx_acct_* in
include/orthoroute_hip.h).  ``ViaKernelManager`` carries the two via capacity
kernels that act on its cost array (pathfinder/via_kernels.py:186-325).
A router that owns a CSR graph can use both in place of the reference's
classes: arc ids are the CSR arc ids the reference passes.

Results equal the reference's float32 NumPy evaluation bit for bit (the
kernels apply the same float32 operations in the same order); the overuse
and balance sums are formed in float64, which is exact for integer counts
and capacities.  Deliberate differences:
  * arc ids outside [0, E) raise IndexError, negative ones included (NumPy
    would wrap a negative id); ids given as device tensors are checked with
    one device reduction;
  * the via halo of commit_path (``_via_halo``, a via pooling feature, not
    part of this path) is not supported: setting it raises.
"""
from __future__ import annotations

import ctypes as C
import logging

import numpy as np
import torch

from . import _lib

log = logging.getLogger(__name__)

_F32 = torch.float32


def _f32(x) -> float:
    """A Python-float coefficient rounded as NumPy 2 applies it to float32."""
    return float(np.float32(x))


def _stream(stream) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


class DeviceEdgeAccountant:
    """EdgeAccountant (UP:1039-1052) on the device."""

    _ARRAYS = ("present", "present_ema", "history", "capacity")

    def __init__(self, num_edges: int, use_gpu: bool = True, device="cuda", stream=None):
        if not use_gpu:
            raise ValueError("DeviceEdgeAccountant runs on the GPU only")
        _lib.init_device()
        self.E = int(num_edges)
        if self.E < 0:
            raise ValueError("num_edges must be >= 0")
        self.use_gpu = True
        self.device = torch.device(device)
        self.stream = stream
        n = max(self.E, 1)
        dev = self.device
        self._a = {
            "present": torch.zeros(n, dtype=_F32, device=dev),
            "present_ema": torch.zeros(n, dtype=_F32, device=dev),
            "history": torch.zeros(n, dtype=_F32, device=dev),
            "capacity": torch.ones(n, dtype=_F32, device=dev),
        }
        self.canon = torch.zeros(n, dtype=torch.int32, device=dev)
        self._total = None
        self._live_static = None
        self._present_cost_scale = 1.0
        self._blocked_edges = None
        self._inf_edges = 0
        self._sums = torch.zeros(4, dtype=torch.float64, device=dev)
        self._counts = torch.zeros(4, dtype=torch.int64, device=dev)

    # -- attributes of the reference -------------------------------------------
    def _get(self, name):
        return self._a[name][: self.E]

    def _set(self, name, value):
        t = torch.as_tensor(np.asarray(value, np.float32) if not torch.is_tensor(value) else value)
        if t.shape != (self.E,):
            raise ValueError(f"{name} must have {self.E} entries")
        self._a[name][: self.E].copy_(t.to(self.device, _F32))

    present = property(lambda s: s._get("present"), lambda s, v: s._set("present", v))
    present_ema = property(lambda s: s._get("present_ema"), lambda s, v: s._set("present_ema", v))
    history = property(lambda s: s._get("history"), lambda s, v: s._set("history", v))
    capacity = property(lambda s: s._get("capacity"), lambda s, v: s._set("capacity", v))

    @property
    def total_cost(self):
        return None if self._total is None else self._total[: self.E]

    @property
    def canonical(self) -> dict:
        """Canonical use counts as the reference's dict {arc: count} (host copy)."""
        c = self.canon[: self.E].cpu().numpy()
        nz = np.flatnonzero(c)
        return {int(i): int(c[i]) for i in nz}

    @property
    def edge_usage(self) -> dict:
        return self.canonical

    @property
    def _via_halo(self):
        return None

    @_via_halo.setter
    def _via_halo(self, value):
        if value is not None:
            raise NotImplementedError("via halo expansion (via pooling) is not supported")

    def _struct(self) -> _lib.OrxAcct:
        return _lib.OrxAcct(self.E, self._a["present"].data_ptr(), self._a["present_ema"].data_ptr(),
                            self._a["history"].data_ptr(), self._a["capacity"].data_ptr(),
                            None if self._total is None else self._total.data_ptr(),
                            None if self._live_static is None else self._live_static.data_ptr(),
                            self.canon.data_ptr())

    def _ids(self, edge_indices) -> torch.Tensor:
        """Arc ids as an int64 device tensor, range-checked."""
        if torch.is_tensor(edge_indices):
            t = edge_indices.reshape(-1)
            if t.numel() and t.dtype in (torch.float16, torch.float32, torch.float64, torch.bfloat16):
                raise TypeError("arc ids must be integers")
            t = t.to(self.device, torch.int64)
            if t.numel():
                lo, hi = (int(v) for v in torch.aminmax(t))
                if lo < 0 or hi >= self.E:
                    raise IndexError(f"arc id out of range [0, {self.E})")
            return t
        a = np.asarray(edge_indices, dtype=np.int64).reshape(-1)
        if a.size and (a.min() < 0 or a.max() >= self.E):
            raise IndexError(f"arc id out of range [0, {self.E})")
        return torch.from_numpy(a).to(self.device)

    def _device_f32(self, x, n: int, what: str) -> torch.Tensor:
        t = x if torch.is_tensor(x) else torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32)))
        t = t.reshape(-1).to(self.device, _F32).contiguous()
        if t.numel() != n:
            raise ValueError(f"{what} must have {n} entries")
        return t

    # -- usage (UP:1073-1224) --------------------------------------------------
    def refresh_from_canonical(self):
        _lib.call("orx_acct_from_canonical", C.byref(self._struct()), _stream(self.stream))

    def commit_path(self, edge_indices):
        """UP:1099-1120.  A concatenation of paths commits them all."""
        if edge_indices is None or len(edge_indices) == 0:
            return
        ids = self._ids(edge_indices)
        _lib.call("orx_acct_apply", C.byref(self._struct()), ids.data_ptr(), ids.numel(), 1, 0, None,
                  _stream(self.stream))

    def clear_path(self, edge_indices):
        """UP:1122-1135, including the whole-array floor of present at 0."""
        if edge_indices is None or len(edge_indices) == 0:
            return
        ids = self._ids(edge_indices)
        _lib.call("orx_acct_apply", C.byref(self._struct()), ids.data_ptr(), ids.numel(), -1, 1, None,
                  _stream(self.stream))

    def _overuse_arcs(self, mask) -> tuple[int, int]:
        self._sums.zero_()
        self._counts.zero_()
        m = None
        if mask is not None:
            m = torch.as_tensor(np.asarray(mask, bool) if not torch.is_tensor(mask) else mask)
            m = m.reshape(-1).to(self.device, torch.uint8).contiguous()
            if m.numel() != self.E:
                raise ValueError(f"mask must have {self.E} entries")
        _lib.call("orx_acct_overuse", C.byref(self._struct()), None if m is None else m.data_ptr(),
                  self._sums.data_ptr(), self._counts.data_ptr(), _stream(self.stream))
        return int(self._sums[0].item()), int(self._counts[0].item())

    def compute_overuse(self, router_instance=None, canonical_mask=None) -> tuple[int, int]:
        """UP:1137-1190: (total overuse, overused arc count).  The arc mask
        comes from router_instance._canonical_edge_resource_mask() when the
        router has one (or from canonical_mask); the via column / segment
        pools of the router (host arrays) add to the total as in the
        reference."""
        mask = canonical_mask
        if router_instance is not None and hasattr(router_instance, "_canonical_edge_resource_mask"):
            mask = router_instance._canonical_edge_resource_mask()
        arc_sum, arc_count = self._overuse_arcs(mask)
        col_sum = seg_sum = 0
        if router_instance is not None:
            if hasattr(router_instance, "via_col_use") and hasattr(router_instance, "via_col_cap"):
                col_sum = int(np.maximum(0, np.asarray(router_instance.via_col_use) -
                                         np.asarray(router_instance.via_col_cap)).sum())
            if hasattr(router_instance, "via_seg_use") and hasattr(router_instance, "via_seg_cap"):
                seg_sum = int(np.maximum(0, np.asarray(router_instance.via_seg_use) -
                                         np.asarray(router_instance.via_seg_cap)).sum())
        return arc_sum + col_sum + seg_sum, arc_count

    def _sum(self, which: int) -> float:
        self._sums.zero_()
        _lib.call("orx_acct_sum", C.byref(self._struct()), which, self._sums.data_ptr(), _stream(self.stream))
        return float(self._sums[0].item())

    def verify_present_matches_canonical(self) -> bool:
        """UP:1192-1224."""
        mismatch = self._sum(2)
        if mismatch > 0.01:
            log.error(f"[ACCOUNTING] Present/canonical mismatch: {mismatch:.2f}")
            return False
        return True

    def cost_balance_ratio(self, history_weight: float, present_factor: float) -> float:
        """UP:1054-1067."""
        h = float(np.float32(history_weight) * np.float32(self._sum(0)))
        p = float(np.float32(present_factor) * np.float32(self._sum(1)))
        return h / (p + 1e-9)

    # -- costs (UP:1226-1422) --------------------------------------------------
    def update_history(self, gain: float, base_costs=None, history_cap_multiplier=10.0, decay_factor=0.98,
                       use_raw_present=False):
        base = None if base_costs is None else self._device_f32(base_costs, self.E, "base_costs")
        _lib.call("orx_acct_history", C.byref(self._struct()), _f32(gain),
                  None if base is None else base.data_ptr(), _f32(history_cap_multiplier), _f32(decay_factor),
                  1 if use_raw_present else 0, _stream(self.stream))

    def update_present_ema(self, beta: float = 0.60):
        _lib.call("orx_acct_present_ema", C.byref(self._struct()), _f32(beta), _f32(1.0 - beta),
                  _stream(self.stream))

    def update_costs(self, base_costs, pres_fac: float, hist_weight: float = 1.0, add_jitter: bool = True,
                     via_cost_multiplier: float = 1.0, base_cost_weight: float = 0.01, *, edge_layer=None,
                     layer_bias_per_layer=None, edge_kind=None):
        base = self._device_f32(base_costs, self.E, "base_costs")
        if self._total is None:
            self._total = torch.empty(max(self.E, 1), dtype=_F32, device=self.device)
        layered = edge_layer is not None and layer_bias_per_layer is not None and edge_kind is not None
        layer = bias = kind = scale = None
        n_layers = 0
        if layered:
            bias = self._device_f32(layer_bias_per_layer, int(np.size(layer_bias_per_layer))
                                    if not torch.is_tensor(layer_bias_per_layer) else
                                    layer_bias_per_layer.numel(), "layer_bias_per_layer")
            n_layers = bias.numel()
            layer = torch.as_tensor(edge_layer if torch.is_tensor(edge_layer) else np.asarray(edge_layer))
            layer = layer.reshape(-1).to(self.device, torch.int32).contiguous()
            kind = torch.as_tensor(edge_kind if torch.is_tensor(edge_kind) else np.asarray(edge_kind))
            kind = kind.reshape(-1).to(self.device, torch.int32).contiguous()
            if layer.numel() != self.E or kind.numel() != self.E:
                raise ValueError(f"edge_layer / edge_kind must have {self.E} entries")
            if self.E:
                lo, hi = (int(v) for v in torch.aminmax(layer))
                if lo < 0 or hi >= n_layers:
                    raise IndexError("edge_layer out of range of layer_bias_per_layer")
            scale = torch.empty(max(self.E, 1), dtype=_F32, device=self.device)
        blocked = None
        self._inf_edges = 0
        if self._blocked_edges is not None:
            b = torch.as_tensor(self._blocked_edges if torch.is_tensor(self._blocked_edges)
                                else np.asarray(self._blocked_edges))
            if tuple(b.shape) == (self.E,):
                blocked = b.to(self.device, torch.uint8).contiguous()
                self._inf_edges = int(blocked.sum().item())
            else:
                log.error(f"[EXISTING-COPPER] blocked-edge mask {tuple(b.shape)} != total_cost ({self.E},); "
                          f"NOT applied")
        _lib.call("orx_acct_costs", C.byref(self._struct()), base.data_ptr(), _f32(pres_fac), _f32(hist_weight),
                  _f32(via_cost_multiplier), _f32(base_cost_weight),
                  None if layer is None else layer.data_ptr(), None if bias is None else bias.data_ptr(), n_layers,
                  None if kind is None else kind.data_ptr(), None if blocked is None else blocked.data_ptr(),
                  1 if add_jitter else 0, None if scale is None else scale.data_ptr(), _stream(self.stream))
        self._present_cost_scale = scale[: self.E] if layered else float(pres_fac)

    def _scale_args(self):
        s = self._present_cost_scale
        if torch.is_tensor(s):
            return 0.0, s.data_ptr()
        return _f32(s), None

    def begin_live_present_costs(self):
        """UP:1386-1401."""
        if self._total is None:
            raise TypeError("update_costs must run before live present costs")
        if self._live_static is None:
            self._live_static = torch.empty(max(self.E, 1), dtype=_F32, device=self.device)
        sc, sp = self._scale_args()
        _lib.call("orx_acct_begin_live", C.byref(self._struct()), sc, sp, _stream(self.stream))

    def refresh_live_present_costs(self, edge_indices):
        """UP:1403-1422."""
        if edge_indices is None or len(edge_indices) == 0 or self._live_static is None:
            return
        ids = self._ids(edge_indices)
        sc, sp = self._scale_args()
        _lib.call("orx_acct_refresh_live", C.byref(self._struct()), ids.data_ptr(), ids.numel(), sc, sp, None,
                  _stream(self.stream))


def via_edge_metadata(indptr, indices, edge_kind, X: int, Y: int, L: int, device="cuda") -> dict:
    """Via arcs of a CSR lattice graph as device arrays (the GPU branch of
    _build_via_edge_metadata, UP:9885-9928): CSR ids of arcs with
    edge_kind != 0, their source column (x, y), and the layer span
    [min(zu, zv), max(zu, zv)] clipped to [1, L - 2]."""
    dev = torch.device(device)
    ip = torch.as_tensor(np.asarray(indptr, np.int64)).to(dev)
    ix = torch.as_tensor(np.asarray(indices, np.int64)).to(dev)
    k = torch.as_tensor(np.asarray(edge_kind)).to(dev)
    ids = torch.nonzero(k != 0).reshape(-1)
    u = torch.searchsorted(ip, ids, right=True) - 1
    v = ix[ids]
    plane = X * Y
    xu, yu = (u % plane) % X, (u % plane) // X
    zu, zv = u // plane, v // plane
    return {
        "indices": ids.to(torch.int32),
        "xy_coords": torch.stack([xu, yu], dim=1).to(torch.int32).contiguous(),
        "z_lo": torch.clamp(torch.minimum(zu, zv), 1, L - 2).to(torch.int32),
        "z_hi": torch.clamp(torch.maximum(zu, zv), 1, L - 2).to(torch.int32),
    }


class ViaKernelManager:
    """hard_block_via_edges / apply_via_penalties of the reference's
    ViaKernelManager (pathfinder/via_kernels.py:186-325) on device arrays.
    Column arrays are [Nx, Ny] (column index x * Ny + y), segment arrays
    [Nx, Ny, segZ]; total_cost is a float32 device tensor changed in place."""

    def __init__(self, use_gpu: bool = True, device="cuda", stream=None):
        if not use_gpu:
            raise ValueError("ViaKernelManager runs on the GPU only")
        _lib.init_device()
        self.use_gpu = True
        self.device = torch.device(device)
        self.stream = stream
        self._count = torch.zeros(2, dtype=torch.int64, device=self.device)

    def _dev(self, x, dtype):
        if x is None:
            return None
        t = x if torch.is_tensor(x) else torch.from_numpy(np.ascontiguousarray(x))
        return t.to(self.device, dtype).contiguous()

    def _meta(self, via_metadata):
        edge = self._dev(via_metadata["indices"], torch.int32)
        xy = self._dev(via_metadata["xy_coords"], torch.int32)
        zlo = self._dev(via_metadata["z_lo"], torch.int32)
        zhi = self._dev(via_metadata["z_hi"], torch.int32)
        n = edge.numel()
        if xy.numel() != 2 * n or zlo.numel() != n or zhi.numel() != n:
            raise ValueError("via metadata arrays differ in length")
        return edge, xy, zlo, zhi, n

    @staticmethod
    def _check_total(total_cost):
        if not (torch.is_tensor(total_cost) and total_cost.is_cuda and total_cost.dtype == _F32 and
                total_cost.is_contiguous()):
            raise TypeError("total_cost must be a contiguous float32 device tensor (changed in place)")

    def _finish(self) -> int:
        bad = int(self._count[1].item())
        if bad:
            raise IndexError(f"{bad} via entries have a column or arc id out of range")
        return int(self._count[0].item())

    def hard_block_via_edges(self, via_metadata, via_col_use, via_col_cap, via_seg_use, via_seg_cap, total_cost,
                             Ny: int, segZ: int) -> int:
        """Number of via arcs set to +inf (via_kernels.py:186-260)."""
        self._check_total(total_cost)
        edge, xy, zlo, zhi, n = self._meta(via_metadata)
        if n == 0:
            return 0
        cu, cc = self._dev(via_col_use, torch.int16), self._dev(via_col_cap, torch.int16)
        su, sc = self._dev(via_seg_use, torch.int8), self._dev(via_seg_cap, torch.int8)
        Nx = cu.shape[0]
        if tuple(cu.shape) != (Nx, Ny) or tuple(cc.shape) != (Nx, Ny):
            raise ValueError("column arrays must be [Nx, Ny]")
        if su is not None and (tuple(su.shape) != (Nx, Ny, segZ) or tuple(sc.shape) != (Nx, Ny, segZ)):
            raise ValueError("segment arrays must be [Nx, Ny, segZ]")
        self._count.zero_()
        _lib.call("orx_via_hard_block", edge.data_ptr(), xy.data_ptr(), zlo.data_ptr(), zhi.data_ptr(), n, Nx,
                  int(Ny), int(segZ), cu.data_ptr(), cc.data_ptr(), None if su is None else su.data_ptr(),
                  None if sc is None else sc.data_ptr(), total_cost.data_ptr(), total_cost.numel(),
                  self._count.data_ptr(), _stream(self.stream))
        return self._finish()

    def apply_via_penalties(self, via_metadata, via_col_pres, via_seg_pres, col_weight: float, seg_weight: float,
                            total_cost, Ny: int, segZ: int) -> int:
        """Number of via arcs given a positive penalty (via_kernels.py:262-325).
        The listed arcs must be distinct for a bit-exact result (the
        additions into total_cost are atomic)."""
        self._check_total(total_cost)
        edge, xy, zlo, zhi, n = self._meta(via_metadata)
        if n == 0:
            return 0
        cp_ = self._dev(via_col_pres, _F32)
        sp = self._dev(via_seg_pres, _F32)
        Nx = cp_.shape[0]
        if tuple(cp_.shape) != (Nx, Ny):
            raise ValueError("column pressure must be [Nx, Ny]")
        if sp is not None and tuple(sp.shape) != (Nx, Ny, segZ):
            raise ValueError("segment pressure must be [Nx, Ny, segZ]")
        self._count.zero_()
        _lib.call("orx_via_penalties", edge.data_ptr(), xy.data_ptr(), zlo.data_ptr(), zhi.data_ptr(), n, Nx,
                  int(Ny), int(segZ), cp_.data_ptr(), None if sp is None else sp.data_ptr(), _f32(col_weight),
                  _f32(seg_weight), total_cost.data_ptr(), total_cost.numel(), self._count.data_ptr(),
                  _stream(self.stream))
        return self._finish()
