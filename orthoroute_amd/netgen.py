"""Synthetic routing workloads on the 3D lattice.

Each generator is seeded and deterministic (NumPy PCG64).  Pins sit on layer
0 (the pin/escape layer) and are pairwise distinct.  Every net gets an ROI
box: its pins' bounding box plus a margin, clipped to the board, spanning all
layers.  Workloads:

* ``plain_64``  -- 64x64x6, 200 two-pin nets, pins uniform without
  replacement on layer 0, seed 0;
* ``backplane_400`` -- 400x400x16, two 40x40 pin arrays at opposite ends
  joined by a seeded permutation (1600 long nets) plus 400 short nets with
  Manhattan pin spacing uniform in 2..20;
* ``uneven_800`` -- 800x800x20, 8000 two-pin nets, Manhattan length
  lognormal (median 25, sigma 1, clipped to 2..700), sources uniform,
  directions random, seed 3, ROI margin 4;
* ``serving_job`` -- one synthetic board of the serving stream;
* ``with_portals`` -- any of the above with portal terminals: every pad
  becomes a seed (source pad) or target (sink pad) on each inner layer of
  one or more escape cells, priced by via depth.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .lattice import LatticeSpec


MAX_TARGETS = 128   # targets per net the route kernel accepts (ORX_MAX_TARGETS)


@dataclass
class Terminals:
    """Seeds and targets of multi-terminal nets, CSR by net: net k's seeds
    are seed_node[seed_ptr[k]:seed_ptr[k+1]] with costs seed_cost[...], its
    targets likewise.  A net with no seeds routes src -> dst.  The search
    semantics are those of find_path_multisource_multisink
    (unified_pathfinder.py:2550-2659): each seed starts at fl(seed_cost +
    its node term), targets are sinks, the best target minimises
    (fl(dist + tgt_cost), node id).  Seeds and targets of a pad on every
    legal layer are what _get_pad_portal_seeds (UP:4054-4180) lists."""
    seed_ptr: np.ndarray   # int32[n+1]
    seed_node: np.ndarray  # int32 global ids
    seed_cost: np.ndarray  # float32
    tgt_ptr: np.ndarray    # int32[n+1]
    tgt_node: np.ndarray
    tgt_cost: np.ndarray

    def select(self, idx) -> "Terminals":
        idx = np.asarray(idx, np.int64)

        def take(ptr, *arrs):
            lo, hi = ptr[idx], ptr[idx + 1]
            cnt = (hi - lo).astype(np.int64)
            nptr = np.zeros(len(idx) + 1, np.int32)
            np.cumsum(cnt, out=nptr[1:])
            sel = np.concatenate([np.arange(a, b) for a, b in zip(lo, hi)]) if len(idx) else np.zeros(0, np.int64)
            sel = sel.astype(np.int64)
            return (nptr, *(a[sel].copy() for a in arrs))

        sp, sn, sc = take(self.seed_ptr, self.seed_node, self.seed_cost)
        tp, tn, tc = take(self.tgt_ptr, self.tgt_node, self.tgt_cost)
        return Terminals(sp, sn, sc, tp, tn, tc)

    def counts(self) -> tuple[np.ndarray, np.ndarray]:
        return np.diff(self.seed_ptr), np.diff(self.tgt_ptr)

    def validate(self, n: int, N: int) -> None:
        for name, ptr, node in (("seed", self.seed_ptr, self.seed_node), ("target", self.tgt_ptr, self.tgt_node)):
            if len(ptr) != n + 1 or ptr[0] != 0 or np.any(np.diff(ptr) < 0) or ptr[-1] != len(node):
                raise ValueError(f"{name}_ptr is not a CSR index over {n} nets")
            if len(node) and (node.min() < 0 or node.max() >= N):
                raise ValueError(f"{name} node outside the lattice")
        ns, nt = self.counts()
        multi = ns > 0
        if np.any(multi & ((nt < 1) | (nt > MAX_TARGETS))):
            raise ValueError(f"a multi-terminal net needs 1..{MAX_TARGETS} targets")
        if len(self.seed_cost) != len(self.seed_node) or len(self.tgt_cost) != len(self.tgt_node):
            raise ValueError("terminal costs and nodes differ in length")
        # packed keys order distances by their fp32 bits: costs must be >= 0
        if np.any(~np.isfinite(self.seed_cost)) or np.any(self.seed_cost < 0):
            raise ValueError("seed costs must be finite and >= 0")
        if np.any(np.isnan(self.tgt_cost)) or np.any(self.tgt_cost < 0):
            raise ValueError("target costs must be >= 0")


@dataclass
class Netlist:
    spec: LatticeSpec
    src: np.ndarray   # int32 global node ids (layer 0)
    dst: np.ndarray
    box: np.ndarray   # int32 [n, 4]: x0, y0, w, h
    name: str = ""
    meta: dict = field(default_factory=dict)
    terms: Terminals | None = None

    @property
    def n(self) -> int:
        return len(self.src)

    def roi_nodes(self) -> np.ndarray:
        return self.box[:, 2].astype(np.int64) * self.box[:, 3] * self.spec.L

    def subset(self, k: int) -> "Netlist":
        return self.select(np.arange(min(k, self.n)), f":{k}")

    def select(self, idx, tag: str = "") -> "Netlist":
        """The nets ``idx`` (in that order) on the same board."""
        idx = np.asarray(idx, np.int64)
        return Netlist(self.spec, self.src[idx].copy(), self.dst[idx].copy(), self.box[idx].copy(),
                       f"{self.name}[{tag or len(idx)}]", dict(self.meta),
                       None if self.terms is None else self.terms.select(idx))


def roi_boxes(spec: LatticeSpec, src: np.ndarray, dst: np.ndarray, margin: int) -> np.ndarray:
    """Bounding box of the two pins plus ``margin`` cells, clipped to the board."""
    sx, sy, _ = spec.coords(src)
    tx, ty, _ = spec.coords(dst)
    x0 = np.maximum(np.minimum(sx, tx) - margin, 0)
    y0 = np.maximum(np.minimum(sy, ty) - margin, 0)
    x1 = np.minimum(np.maximum(sx, tx) + margin, spec.X - 1)
    y1 = np.minimum(np.maximum(sy, ty) + margin, spec.Y - 1)
    return np.stack([x0, y0, x1 - x0 + 1, y1 - y0 + 1], axis=1).astype(np.int32)


def long_nets_full_board(nl: Netlist, steps: int = 125) -> Netlist:
    """The long-net half of the reference engine's ROI rule: a net whose
    pins are at least ``steps`` apart (|dx| + |dy| in cells, UP:7980-7982)
    routes on the full graph (UP:8219-8262), here the whole board as its
    box; shorter nets keep their box.  Multi-terminal nets keep theirs."""
    sx, sy, _ = nl.spec.coords(nl.src)
    tx, ty, _ = nl.spec.coords(nl.dst)
    d = np.abs(tx.astype(np.int64) - sx) + np.abs(ty.astype(np.int64) - sy)
    box = nl.box.copy()
    long = d >= steps
    if nl.terms is not None:
        ns, _ = nl.terms.counts()
        long &= ns == 0
    box[long] = (0, 0, nl.spec.X, nl.spec.Y)
    meta = dict(nl.meta, full_board_steps=steps, full_board_nets=int(long.sum()))
    return Netlist(nl.spec, nl.src.copy(), nl.dst.copy(), box, f"{nl.name}_full{steps}", meta,
                   nl.terms)


def _place_two_pin(rng, spec: LatticeSpec, lengths: np.ndarray, used: np.ndarray,
                   forbid: np.ndarray | None = None, max_tries: int = 64):
    """For each net: a random Manhattan split (|dx|, |dy|) of its length and
    random signs, then a source uniform over the layer-0 cells from which that
    offset stays on the board.  Resamples when a pin cell is already used;
    after ``max_tries`` failures (a length near the board diagonal leaves only
    corner placements) the length shrinks by 10% and sampling continues."""
    X, Y = spec.X, spec.Y
    n = len(lengths)
    src = np.empty(n, np.int64)
    dst = np.empty(n, np.int64)
    for i in range(n):
        ell = int(min(max(lengths[i], 1), (X - 1) + (Y - 1)))
        for attempt in range(16 * max_tries):
            if attempt and attempt % max_tries == 0:
                ell = max(1, int(ell * 0.9))
            ax = min(int(round(ell * float(rng.random()))), X - 1)
            ay = ell - ax
            if ay > Y - 1:
                ay = Y - 1
                ax = ell - ay
            gx = 1 if rng.random() < 0.5 else -1
            gy = 1 if rng.random() < 0.5 else -1
            sx = int(rng.integers(0, X - ax)) + (ax if gx < 0 else 0)
            sy = int(rng.integers(0, Y - ay)) + (ay if gy < 0 else 0)
            tx, ty = sx + gx * ax, sy + gy * ay
            s, t = sy * X + sx, ty * X + tx
            if s == t or used[s] or used[t]:
                continue
            if forbid is not None and (forbid[s] or forbid[t]):
                continue
            used[s] = True
            used[t] = True
            src[i], dst[i] = s, t
            break
        else:
            raise RuntimeError("could not place a net; board too crowded")
    return src, dst


def plain_64(seed: int = 0, margin: int = 4) -> Netlist:
    """64x64x6, 200 two-pin nets with pins uniform without replacement on layer 0."""
    spec = LatticeSpec(64, 64, 6)
    rng = np.random.default_rng(seed)
    cells = rng.choice(spec.XY, size=400, replace=False).astype(np.int64)
    src, dst = cells[0::2], cells[1::2]
    return Netlist(spec, src.astype(np.int32), dst.astype(np.int32), roi_boxes(spec, src, dst, margin),
                   "plain_64x64x6_200", {"seed": seed, "margin": margin})


def lognormal_lengths(rng, n: int, median: float = 25.0, sigma: float = 1.0, lo: int = 2, hi: int = 700):
    ell = rng.lognormal(mean=np.log(median), sigma=sigma, size=n)
    return np.clip(np.rint(ell), lo, hi).astype(np.int64)


def uneven(X: int, Y: int, L: int, n_nets: int, seed: int, margin: int = 4, name: str = "") -> Netlist:
    spec = LatticeSpec(X, Y, L)
    rng = np.random.default_rng(seed)
    lengths = lognormal_lengths(rng, n_nets)
    used = np.zeros(spec.XY, bool)
    src, dst = _place_two_pin(rng, spec, lengths, used)
    return Netlist(spec, src.astype(np.int32), dst.astype(np.int32), roi_boxes(spec, src, dst, margin),
                   name or f"uneven_{X}x{Y}x{L}_{n_nets}", {"seed": seed, "margin": margin})


def uneven_800(seed: int = 3, n_nets: int = 8000, margin: int = 4) -> Netlist:
    """800x800x20 with 8000 lognormal-length two-pin nets (seed 3)."""
    return uneven(800, 800, 20, n_nets, seed, margin, "uneven_800x800x20_8000")


def backplane_400(seed: int = 2, margin: int = 4) -> Netlist:
    """400x400x16: two 40x40 pin arrays at opposite ends joined by a seeded
    permutation (1600 long nets) plus 400 short nets, spacing U{2..20}."""
    spec = LatticeSpec(400, 400, 16)
    rng = np.random.default_rng(seed)
    ax0, bx0, y0 = 20, 340, 180
    gx, gy = np.meshgrid(np.arange(40), np.arange(40), indexing="xy")
    a = ((y0 + gy) * spec.X + (ax0 + gx)).reshape(-1)
    b = ((y0 + gy) * spec.X + (bx0 + gx)).reshape(-1)
    perm = rng.permutation(1600)
    long_src, long_dst = a.astype(np.int64), b[perm].astype(np.int64)
    used = np.zeros(spec.XY, bool)
    used[a] = True
    used[b] = True
    spacing = rng.integers(2, 21, size=400)
    s_src, s_dst = _place_two_pin(rng, spec, spacing, used)
    src = np.concatenate([long_src, s_src])
    dst = np.concatenate([long_dst, s_dst])
    return Netlist(spec, src.astype(np.int32), dst.astype(np.int32), roi_boxes(spec, src, dst, margin),
                   "backplane_400x400x16_2000", {"seed": seed, "margin": margin})


SERVING_SIDES = (128, 256, 512, 1024)
SERVING_LAYERS = (8, 16, 24, 32)


def serving_job(rng, n_min: int = 500, n_max: int = 10000, margin: int = 4) -> Netlist:
    """One board of the serving stream: side in {128,256,512,1024}, layers in
    {8,16,24,32}, 500..10000 nets generated like uneven_800."""
    side = int(rng.choice(SERVING_SIDES))
    L = int(rng.choice(SERVING_LAYERS))
    n = int(rng.integers(n_min, n_max + 1))
    n = min(n, (side * side) // 4)
    seed = int(rng.integers(0, 2**31 - 1))
    return uneven(side, side, L, n, seed, margin, f"job_{side}x{side}x{L}_{n}")


def with_portals(nl: Netlist, candidates: int = 1, via_cost: float = 3.0, discount: float = 0.15,
                 candidate_penalty: float = 0.5, fraction: float = 1.0, seed: int = 0) -> Netlist:
    """Portal form of a netlist (the pads of _get_pad_portal_seeds,
    unified_pathfinder.py:3510-3565 and 4054-4180): each pad at (x, y) on
    layer 0 gets ``candidates`` escape cells (the pad cell, then seeded cells
    within 2 of it), and each escape cell one terminal per inner layer
    1..L-2 priced fp32(via_cost * discount * layer + candidate_penalty * c)
    for the c-th cell.  Source pads give seeds, sink pads targets.  A seeded
    ``fraction`` of the nets gets portals; the rest stay two-pin.  Boxes grow
    to hold every terminal."""
    spec = nl.spec
    if spec.L < 3:
        raise ValueError("portals need at least three layers")
    rng = np.random.default_rng(seed)
    layers = np.arange(1, spec.L - 1)
    pick = rng.random(nl.n) < fraction
    sp, sn, sc, tp, tn, tc = [0], [], [], [0], [], []

    def cells(pad):
        x, y, _ = spec.coords(np.asarray([pad]))
        out = [(int(x[0]), int(y[0]))]
        while len(out) < candidates:
            cx = int(np.clip(out[0][0] + rng.integers(-2, 3), 0, spec.X - 1))
            cy = int(np.clip(out[0][1] + rng.integers(-2, 3), 0, spec.Y - 1))
            if (cx, cy) not in out:
                out.append((cx, cy))
        return out

    def terms(pad):
        nodes, costs = [], []
        for c, (x, y) in enumerate(cells(pad)):
            for z in layers:
                nodes.append(int(z) * spec.XY + y * spec.X + x)
                costs.append(np.float32(via_cost * discount * float(z) + candidate_penalty * c))
        return nodes, costs

    box = nl.box.copy()
    margin = int(nl.meta.get("margin", 4))
    for k in range(nl.n):
        if pick[k]:
            a, ac = terms(int(nl.src[k]))
            b, bc = terms(int(nl.dst[k]))
            if len(b) > MAX_TARGETS:
                raise ValueError("too many targets per net")
            sn += a; sc += ac; tn += b; tc += bc
            xs, ys, _ = spec.coords(np.asarray(a + b, np.int64))
            x0 = max(0, min(int(box[k, 0]), int(xs.min()) - margin))
            y0 = max(0, min(int(box[k, 1]), int(ys.min()) - margin))
            x1 = min(spec.X - 1, max(int(box[k, 0] + box[k, 2] - 1), int(xs.max()) + margin))
            y1 = min(spec.Y - 1, max(int(box[k, 1] + box[k, 3] - 1), int(ys.max()) + margin))
            box[k] = (x0, y0, x1 - x0 + 1, y1 - y0 + 1)
        sp.append(len(sn))
        tp.append(len(tn))
    T = Terminals(np.asarray(sp, np.int32), np.asarray(sn, np.int32), np.asarray(sc, np.float32).reshape(-1),
                  np.asarray(tp, np.int32), np.asarray(tn, np.int32), np.asarray(tc, np.float32).reshape(-1))
    T.validate(nl.n, spec.N)
    meta = dict(nl.meta, portals={"candidates": candidates, "fraction": fraction, "seed": seed})
    return Netlist(spec, nl.src.copy(), nl.dst.copy(), box.astype(np.int32), f"{nl.name}_portals{candidates}", meta,
                   T)
