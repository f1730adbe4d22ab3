"""3D Manhattan routing lattice: geometry, arc layout and base costs.

Node id ``z*X*Y + y*X + x`` (unified_pathfinder.py:118; kicad_geometry.py:59-62).
Layer directions follow Lattice3D._assign_directions
(unified_pathfinder.py:1487-1495): layer 0 vertical, odd layers horizontal,
even layers vertical.  Every layer carries planar arcs along its axis (the
reference's ORTHO_LATERAL_ALL=1 form, unified_pathfinder.py:1691-1701) and
vias join vertically adjacent layers in both directions.

Arc layout (device and host): four slots per node ordered by neighbour id,
``0`` via down, ``1`` planar minus, ``2`` planar plus, ``3`` via up; arc id
``4*u + s``; the reverse arc of slot ``s`` is slot ``3 - s``; absent arcs have
base cost ``+inf``.  ``to_csr`` produces the equivalent CSR sorted by
(src, dst) for the generic seam, with the arc-id map between the two.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

PLANAR_COST = np.float32(1.0)
VIA_COST = np.float32(3.0)


def default_hmask(L: int) -> int:
    """Bit z set = layer z is horizontal: odd layers (unified_pathfinder.py:1487-1495)."""
    m = 0
    for z in range(L):
        if z % 2 == 1:
            m |= 1 << z
    return m


@dataclass(frozen=True)
class LatticeSpec:
    X: int
    Y: int
    L: int
    hmask: int = -1

    def __post_init__(self):
        if self.hmask < 0:
            object.__setattr__(self, "hmask", default_hmask(self.L))
        if not (1 <= self.L <= 32):
            raise ValueError("1 <= L <= 32 layers")
        if self.N >= 2**31:
            raise ValueError("lattice too large for int32 node ids")

    @property
    def N(self) -> int:
        return self.X * self.Y * self.L

    @property
    def XY(self) -> int:
        return self.X * self.Y

    def node(self, x, y, z):
        return (np.asarray(z, np.int64) * self.XY + np.asarray(y, np.int64) * self.X + np.asarray(x, np.int64))

    def coords(self, ids):
        ids = np.asarray(ids, np.int64)
        z = ids // self.XY
        rem = ids - z * self.XY
        y = rem // self.X
        x = rem - y * self.X
        return x, y, z

    def horizontal(self, z) -> np.ndarray:
        return ((self.hmask >> np.asarray(z, np.int64)) & 1).astype(bool)

    def neighbours(self) -> np.ndarray:
        """[N, 4] int64 neighbour ids per slot, -1 when absent."""
        u = np.arange(self.N, dtype=np.int64)
        x, y, z = self.coords(u)
        h = self.horizontal(z)
        nb = np.full((self.N, 4), -1, dtype=np.int64)
        nb[:, 0] = np.where(z > 0, u - self.XY, -1)
        nb[:, 3] = np.where(z < self.L - 1, u + self.XY, -1)
        minus_h = np.where(x > 0, u - 1, -1)
        minus_v = np.where(y > 0, u - self.X, -1)
        plus_h = np.where(x < self.X - 1, u + 1, -1)
        plus_v = np.where(y < self.Y - 1, u + self.X, -1)
        nb[:, 1] = np.where(h, minus_h, minus_v)
        nb[:, 2] = np.where(h, plus_h, plus_v)
        return nb

    def base_costs(self, planar=PLANAR_COST, via=VIA_COST) -> np.ndarray:
        """[4N] fp32 base arc costs: planar 1.0, via 3.0, absent +inf."""
        nb = self.neighbours()
        base = np.empty((self.N, 4), dtype=np.float32)
        base[:, [0, 3]] = np.float32(via)
        base[:, [1, 2]] = np.float32(planar)
        base[nb < 0] = np.inf
        return base.reshape(-1)

    def base_costs_device(self, device, planar=PLANAR_COST, via=VIA_COST):
        """The same [4N] array as base_costs, built on the device (no host
        array, no host-to-device copy) with torch ops on the current stream."""
        import torch
        dev = torch.device(device)
        u = torch.arange(self.N, dtype=torch.int64, device=dev)
        z = u // self.XY
        rem = u - z * self.XY
        y = rem // self.X
        x = rem - y * self.X
        horiz = ((self.hmask >> z) & 1).bool()
        inf = torch.tensor(float("inf"), dtype=torch.float32, device=dev)
        pv = torch.tensor(float(np.float32(planar)), dtype=torch.float32, device=dev)
        vv = torch.tensor(float(np.float32(via)), dtype=torch.float32, device=dev)
        base = torch.empty((self.N, 4), dtype=torch.float32, device=dev)
        base[:, 0] = torch.where(z > 0, vv, inf)
        base[:, 3] = torch.where(z < self.L - 1, vv, inf)
        base[:, 1] = torch.where(torch.where(horiz, x > 0, y > 0), pv, inf)
        base[:, 2] = torch.where(torch.where(horiz, x < self.X - 1, y < self.Y - 1), pv, inf)
        return base.reshape(-1)

    def base_cost_minima(self, planar=PLANAR_COST, via=VIA_COST) -> np.ndarray:
        """A tiny array whose planar / via minima equal those of base_costs
        (for the A* bounds without building the whole array)."""
        b = np.full((1, 4), np.inf, np.float32)
        b[0, 1] = planar
        if self.L > 1:
            b[0, 3] = via
        return b.reshape(-1)

    def random_edge_costs(self, seed: int, lo=1.0, hi=4.0, hot_frac=0.05, hot_mult=10.0) -> np.ndarray:
        """[4N] costs uniform in [lo, hi) per physical edge (shared by both
        arcs) with a seeded fraction multiplied by hot_mult (the 1000x1000x12
        SSSP microbenchmark workload)."""
        nb = self.neighbours()
        rng = np.random.default_rng(seed)
        # one draw per canonical arc (slots 2, 3: neighbour id above u)
        base = np.full((self.N, 4), np.inf, dtype=np.float32)
        for s in (2, 3):
            ok = nb[:, s] >= 0
            cnt = int(ok.sum())
            c = rng.uniform(lo, hi, cnt).astype(np.float32)
            hot = rng.random(cnt) < hot_frac
            c[hot] = (c[hot] * np.float32(hot_mult)).astype(np.float32)
            base[ok, s] = c
            base[nb[ok, s], 3 - s] = c
        return base.reshape(-1)

    def arc_count(self) -> int:
        return int((self.neighbours() >= 0).sum())

    def out_degree(self) -> np.ndarray:
        return (self.neighbours() >= 0).sum(axis=1)

    def to_csr(self, arc_values: np.ndarray | None = None):
        """CSR sorted by (src, dst): returns indptr, indices, csr_to_arc
        (lattice arc id of each CSR arc) and values (arc_values permuted)."""
        nb = self.neighbours()
        ok = nb >= 0
        deg = ok.sum(axis=1)
        indptr = np.zeros(self.N + 1, dtype=np.int64)
        np.cumsum(deg, out=indptr[1:])
        arc_ids = np.nonzero(ok.reshape(-1))[0]  # row-major: (u, slot) ascending == (src, dst)
        indices = nb.reshape(-1)[arc_ids]
        vals = None if arc_values is None else np.asarray(arc_values).reshape(-1)[arc_ids]
        return indptr.astype(np.int32), indices.astype(np.int32), arc_ids.astype(np.int64), vals


def in_rank_and_transpose(N: int, indptr: np.ndarray, indices: np.ndarray):
    """For a CSR graph: in_rank[e] = rank of src(e) among the in-neighbours of
    dst(e) ordered by (src, arc id), plus the transposed CSR (in_indptr,
    in_src) in the same order.  Used by the generic search to recover parents
    from the key slot."""
    indptr = np.asarray(indptr, np.int64)
    indices = np.asarray(indices, np.int64)
    E = len(indices)
    src = np.repeat(np.arange(N, dtype=np.int64), np.diff(indptr))
    order = np.lexsort((np.arange(E), src, indices))  # by dst, then src, then arc id
    in_deg = np.bincount(indices, minlength=N)
    in_indptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(in_deg, out=in_indptr[1:])
    in_src = src[order]
    rank = np.empty(E, dtype=np.int64)
    rank[order] = np.arange(E, dtype=np.int64) - in_indptr[indices[order]]
    if E and rank.max() > 254:
        raise ValueError("in-degree above 255 is not supported by the packed key")
    return rank.astype(np.int32), in_indptr.astype(np.int32), in_src.astype(np.int32)
