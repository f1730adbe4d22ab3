"""CPU restatement of the reference's per-arc accountant and via capacity
kernels (TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
orthoroute_amd.accountant, never by the product path).

  OracleEdgeAccountant  -- EdgeAccountant, unified_pathfinder.py:1039-1422
  oracle_hard_block     -- hard_block_via_edges_cpu, pathfinder/via_kernels.py:567-608
  oracle_via_penalties  -- apply_via_penalties_cpu, pathfinder/via_kernels.py:611-651
  oracle_via_metadata   -- _build_via_edge_metadata, unified_pathfinder.py:9875-9928

Every array operation is a float32 NumPy operation applied in the reference's
order, with the reference's Python-float coefficients rounded to float32 the
way NumPy 2 promotes them (requirements pin numpy>=2.2), so the results are
the reference's bit for bit.  The canonical dict is kept as a dict, like the
reference's, and the arrays are rebuilt from it only where the reference
does.

Parity status: pinned by the reference's known answers
(tests/unit/test_edge_accountant.py: commit 1/1/1 and repeated 2/1/1
(86-101), EMA 6.0 and steady state 5.0 (110-171), history 1.0 / 7.0 / cap 10 /
EMA 4.0 / raw 9.0 / decay 98.0 and 0.98^10 (181-298), cost 3.51 / base 0.01 /
via 0.02 / layer bias 0.005-0.01-0.015 / uniform vias / jitter range
(305-440), overuse (3, 3) / (0, 0) (447-475)), by closed forms for the
jitter ((e mod 9973) * 1e-6 in float64, rounded) and the via kernels
(hand-computed columns and segments), and by order-independence properties
(a batch of clears equals the clears one by one); see
tests/test_oracle_edge_accountant.py.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32


def _f(x) -> np.float32:
    """A Python float coefficient as NumPy 2 applies it to a float32 array."""
    return F32(x)


def _max0(a: np.ndarray) -> np.ndarray:
    return np.maximum(F32(0), a)


class OracleEdgeAccountant:
    """Arrays and dict of EdgeAccountant (UP:1042-1052) on the CPU."""

    def __init__(self, num_edges: int):
        self.E = int(num_edges)
        self.canonical: dict[int, int] = {}
        self.present = np.zeros(self.E, F32)
        self.present_ema = np.zeros(self.E, F32)
        self.history = np.zeros(self.E, F32)
        self.capacity = np.ones(self.E, F32)
        self.total_cost = None
        self.live_static = None
        self.scale = 1.0              # _present_cost_scale: float or float32 array
        self.blocked = None           # _blocked_edges

    # -- usage ---------------------------------------------------------------
    def commit_path(self, arcs) -> None:
        """UP:1099-1120: dict count +1 per listing, present scatter +1."""
        arcs = [int(a) for a in arcs]
        if not arcs:
            return
        for a in arcs:
            self.canonical[a] = self.canonical.get(a, 0) + 1
        np.add.at(self.present, np.asarray(arcs, np.int64), F32(1))

    def clear_path(self, arcs) -> None:
        """UP:1122-1135: dict count -1 while present (deleted at <= 0),
        present scatter -1, then the whole array floored at 0."""
        arcs = [int(a) for a in arcs]
        if not arcs:
            return
        for a in arcs:
            c = self.canonical.get(a)
            if c is not None:
                if c - 1 <= 0:
                    del self.canonical[a]
                else:
                    self.canonical[a] = c - 1
        np.add.at(self.present, np.asarray(arcs, np.int64), F32(-1))
        np.maximum(self.present, F32(0), out=self.present)

    def refresh_from_canonical(self) -> None:
        """UP:1073-1097."""
        self.present[:] = 0
        for a, c in self.canonical.items():
            if 0 <= a < self.E:
                self.present[a] = F32(c)

    def canonical_counts(self) -> np.ndarray:
        out = np.zeros(self.E, np.int32)
        for a, c in self.canonical.items():
            out[a] = c
        return out

    def compute_overuse(self, mask=None) -> tuple[int, int]:
        """Arc term of UP:1137-1167 (via pools enter through the caller)."""
        over = _max0(self.present - self.capacity)
        if mask is not None:
            over = over[np.asarray(mask, bool)]
        return int(over.sum()), int(np.count_nonzero(over))

    def verify_present_matches_canonical(self) -> bool:
        """UP:1192-1224."""
        rebuilt = np.zeros(self.E, F32)
        for a, c in self.canonical.items():
            if 0 <= a < self.E:
                rebuilt[a] = F32(c)
        return not (np.sum(np.abs(self.present - rebuilt)) > 0.01)

    def cost_balance_ratio(self, history_weight: float, present_factor: float) -> float:
        """UP:1054-1067."""
        h = float(history_weight * self.history.sum())
        p = float(present_factor * _max0(self.present_ema - self.capacity).sum())
        return h / (p + 1e-9)

    # -- costs ---------------------------------------------------------------
    def update_present_ema(self, beta: float = 0.60) -> None:
        """UP:1291-1299."""
        self.present_ema = _f(beta) * self.present + _f(1.0 - beta) * self.present_ema

    def update_history(self, gain: float, base_costs=None, history_cap_multiplier: float = 10.0,
                       decay_factor: float = 0.98, use_raw_present: bool = False) -> None:
        """UP:1226-1289: decay first, then the (capped) overuse increment."""
        self.history *= _f(decay_factor)
        p = self.present if use_raw_present else self.present_ema
        inc = _f(gain) * _max0(p - self.capacity)
        if base_costs is not None:
            inc = np.minimum(inc, _f(history_cap_multiplier) * np.asarray(base_costs, F32))
        self.history += inc

    def update_costs(self, base_costs, pres_fac: float, hist_weight: float = 1.0, add_jitter: bool = True,
                     via_cost_multiplier: float = 1.0, base_cost_weight: float = 0.01, *, edge_layer=None,
                     layer_bias_per_layer=None, edge_kind=None) -> None:
        """UP:1301-1384."""
        base = np.asarray(base_costs, F32)
        over = _max0(self.present_ema - self.capacity)
        layered = edge_layer is not None and layer_bias_per_layer is not None and edge_kind is not None
        if layered:
            per_arc = np.asarray(layer_bias_per_layer, F32)[np.asarray(edge_layer)]
            bias = np.where(np.asarray(edge_kind) == 0, per_arc, F32(1))
            adjusted = base * _f(via_cost_multiplier) * _f(base_cost_weight) * bias
            self.scale = _f(pres_fac) * bias
            present_term = self.scale * over
        else:
            adjusted = base * _f(via_cost_multiplier) * _f(base_cost_weight)
            self.scale = float(pres_fac)
            present_term = _f(pres_fac) * over
        self.total_cost = adjusted + present_term + _f(hist_weight) * self.history
        if self.blocked is not None and np.shape(self.blocked) == self.total_cost.shape:
            self.total_cost[np.asarray(self.blocked, bool)] = F32(np.inf)
        if add_jitter:
            ramp = np.arange(self.E, dtype=F32) % F32(9973)
            self.total_cost += ramp * _f(1e-6)

    def _scale_of(self, idx=None):
        if np.isscalar(self.scale):
            return _f(self.scale)
        return self.scale if idx is None else self.scale[idx]

    def begin_live_present_costs(self) -> None:
        """UP:1386-1401."""
        s = self._scale_of()
        self.live_static = self.total_cost - s * _max0(self.present_ema - self.capacity)
        self.total_cost[:] = self.live_static + s * _max0(self.present + F32(1) - self.capacity)

    def refresh_live_present_costs(self, arcs) -> None:
        """UP:1403-1422."""
        arcs = np.asarray([int(a) for a in arcs], np.int64)
        if arcs.size == 0 or self.live_static is None:
            return
        s = self._scale_of(arcs)
        self.total_cost[arcs] = self.live_static[arcs] + s * _max0(self.present[arcs] + F32(1) -
                                                                   self.capacity[arcs])


# -- via capacity kernels (CPU forms) -----------------------------------------
def oracle_hard_block(meta: dict, col_use, col_cap, seg_use, seg_cap, total_cost: np.ndarray, segZ: int) -> int:
    """hard_block_via_edges_cpu (via_kernels.py:567-608): +inf on via arcs
    whose column, or (column free) one spanned segment, is at capacity."""
    blocked = 0
    for i, e in enumerate(np.asarray(meta["indices"])):
        x, y = (int(v) for v in meta["xy_coords"][i])
        hit = bool(col_use[x, y] >= col_cap[x, y])
        if not hit and seg_use is not None:
            for z in range(int(meta["z_lo"][i]), int(meta["z_hi"][i])):
                s = z - 1
                if 0 <= s < segZ and seg_use[x, y, s] >= seg_cap[x, y, s]:
                    hit = True
                    break
        if hit:
            total_cost[int(e)] = F32(np.inf)
            blocked += 1
    return blocked


def oracle_via_penalties(meta: dict, col_pres, seg_pres, col_weight: float, seg_weight: float,
                         total_cost: np.ndarray, segZ: int) -> int:
    """apply_via_penalties_cpu (via_kernels.py:611-651): column pressure plus
    the spanned segments' pressure, each weighted, added where positive."""
    n = 0
    cw, sw = _f(col_weight), _f(seg_weight)
    for i, e in enumerate(np.asarray(meta["indices"])):
        x, y = (int(v) for v in meta["xy_coords"][i])
        pen = F32(col_pres[x, y]) * cw
        if seg_pres is not None:
            for z in range(int(meta["z_lo"][i]), int(meta["z_hi"][i])):
                s = z - 1
                if 0 <= s < segZ:
                    pen = pen + F32(seg_pres[x, y, s]) * sw
        if pen > 0:
            total_cost[int(e)] = total_cost[int(e)] + pen
            n += 1
    return n


def oracle_via_metadata(indptr, indices, edge_kind, X: int, Y: int, L: int) -> dict:
    """Via arcs of a CSR lattice graph (UP:9885-9928): CSR ids of the arcs
    with edge_kind != 0, their source column (x, y) and the clipped layer
    span [min(zu, zv), max(zu, zv)] within [1, L - 2]."""
    indptr = np.asarray(indptr, np.int64)
    ids = np.flatnonzero(np.asarray(edge_kind)).astype(np.int32)
    u = np.searchsorted(indptr, ids, side="right") - 1
    v = np.asarray(indices, np.int64)[ids]
    plane = X * Y
    xu, yu = (u % plane) % X, (u % plane) // X
    zu, zv = u // plane, v // plane
    return {
        "indices": ids,
        "xy_coords": np.stack([xu, yu], axis=1).astype(np.int32),
        "z_lo": np.clip(np.minimum(zu, zv), 1, L - 2).astype(np.int32),
        "z_hi": np.clip(np.maximum(zu, zv), 1, L - 2).astype(np.int32),
    }
