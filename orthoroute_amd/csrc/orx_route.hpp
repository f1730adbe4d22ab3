// orx_route.hpp -- device and host pieces shared by the two search engines of
// liborx_hip (the per-net workgroup search in orx_lattice.hip and the tile
// engine in orx_tiles.hip): per-net tables, ROI box indexing, the
// accountant pieces a workgroup runs over one path (rip-up, commit, live
// re-price), the settled-set statistics, the parent walk, the terminal
// table of multi-terminal nets and the host-side argument plumbing.
// Everything is in an anonymous namespace: each translation unit gets its
// own copy.
#pragma once
#include "orx_common.hpp"

#include <algorithm>
#include <atomic>
#include "../../include/orthoroute_hip.h"

namespace {

using namespace orx;

struct NetsDev {
    int32_t n;
    const int32_t* box;
    const int32_t* src;
    const int32_t* dst;
    int32_t* path_arena;
    const int64_t* path_off;
    const int32_t* path_cap;
    int32_t* path_len;
    float* path_cost;
    int32_t* status;
    const int32_t* seed_ptr;
    const int32_t* seed_node;
    const float* seed_cost;
    const int32_t* tgt_ptr;
    const int32_t* tgt_node;
    const float* tgt_cost;
};

struct Ctl {
    uint32_t qcnt[3];
    uint32_t fcnt;
    uint32_t fcnt2;
    unsigned long long minf;
    int32_t err;
    int32_t path_len;
    uint32_t rounds, phases, items;
    uint32_t fhist[32];   // far-pile histogram of log2(1 + (f - T) / delta)
    unsigned long long tbest[2];   // best certified (total bits << 32 | node id), alternate checks
    uint32_t tlb[2];               // bound of the uncertified targets' totals (fp32 bits)
    uint32_t best_r;
    float best_cost;
    unsigned long long settled, relaxed;
};

struct Box {
    int32_t x0, y0, w, h, L;
    uint32_t wh, R;
    FastDiv dw, dwh;
};

// Terminal table of a multi-terminal net (find_path_multisource_multisink,
// unified_pathfinder.py:2550-2659): the ROI indices of its targets inside the
// box, ascending, and their ROI-local bounding box.  Targets are sinks: the
// search reaches them but does not expand them (UP:2611-2618).
constexpr int32_t kMaxTargets = ORX_MAX_TARGETS;
struct TermTab {
    uint32_t n;
    uint32_t lx0, lx1, ly0, ly1, z0, z1;
    uint32_t seeds_in;      // seeds inside the box
    uint32_t roi[kMaxTargets];
};

// Seeds of a multi-terminal net: key fl(cost + q_seed), the seed paying its
// own node term (cuda_dijkstra.py:5589-5606, UP:2582-2583).
struct Multi {
    const TermTab* tt;
    const int32_t* seed_node;
    const float* seed_cost;
    int32_t ns;
    const int32_t* tgt_node;
    const float* tgt_cost;   // nullptr: all 0
    int32_t nt;
    Params P;
};

__device__ __forceinline__ bool is_target(const TermTab* T, uint32_t r, uint32_t lx, uint32_t ly, uint32_t z) {
    if (lx < T->lx0 || lx > T->lx1 || ly < T->ly0 || ly > T->ly1 || z < T->z0 || z > T->z1) return false;
    uint32_t lo = 0, hi = T->n;
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (T->roi[m] < r) lo = m + 1;
        else hi = m;
    }
    return lo < T->n && T->roi[lo] == r;
}

// Atomic helpers (workgroup scope: every access to an ROI's search state
// comes from the one workgroup that owns the ROI).
__device__ __forceinline__ uint64_t key_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t key_min(uint64_t* p, uint64_t v) {
    return __hip_atomic_fetch_min((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t bit_set(uint32_t* bm, uint32_t i) {
    uint32_t b = 1u << (i & 31u);
    return __hip_atomic_fetch_or(bm + (i >> 5), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & b;
}
__device__ __forceinline__ void bit_clear(uint32_t* bm, uint32_t i) {
    __hip_atomic_fetch_and(bm + (i >> 5), ~(1u << (i & 31u)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Decode an ROI-local index.
__device__ __forceinline__ void roi_decode(const Box& B, uint32_t r, uint32_t& lx, uint32_t& ly, uint32_t& z) {
    z = B.dwh.div(r);
    uint32_t rem = r - z * B.wh;
    ly = B.dw.div(rem);
    lx = rem - ly * (uint32_t)B.w;
}

// ROI-local neighbour of r in slot s (UINT32_MAX when outside the box).
__device__ __forceinline__ uint32_t roi_nbr(const Box& B, bool horiz, uint32_t r, uint32_t lx, uint32_t ly,
                                            uint32_t z, int s) {
    switch (s) {
    case 0: return z > 0 ? r - B.wh : UINT32_MAX;
    case 1: return horiz ? (lx > 0 ? r - 1 : UINT32_MAX) : (ly > 0 ? r - (uint32_t)B.w : UINT32_MAX);
    case 2: return horiz ? (lx + 1 < (uint32_t)B.w ? r + 1 : UINT32_MAX)
                         : (ly + 1 < (uint32_t)B.h ? r + (uint32_t)B.w : UINT32_MAX);
    default: return z + 1 < (uint32_t)B.L ? r + B.wh : UINT32_MAX;
    }
}

__device__ __forceinline__ int64_t roi_global(const Lat& g, const Box& B, uint32_t lx, uint32_t ly, uint32_t z) {
    return (int64_t)z * g.XY + (int64_t)(B.y0 + (int32_t)ly) * g.X + (B.x0 + (int32_t)lx);
}

// ROI-local index of a global node, UINT32_MAX outside the box.
__device__ __forceinline__ uint32_t global_to_roi(const Lat& g, const Box& B, int64_t gid) {
    const int32_t z = (int32_t)(gid / g.XY);
    const int64_t rem = gid - (int64_t)z * g.XY;
    const int32_t y = (int32_t)(rem / g.X);
    const int32_t x = (int32_t)(rem - (int64_t)y * g.X);
    if (x < B.x0 || x >= B.x0 + B.w || y < B.y0 || y >= B.y0 + B.h) return UINT32_MAX;
    return ((uint32_t)z * (uint32_t)B.h + (uint32_t)(y - B.y0)) * (uint32_t)B.w + (uint32_t)(x - B.x0);
}

// ---------------------------------------------------------------------------
// Accountant pieces run by one workgroup over one path.
// ---------------------------------------------------------------------------
template <int BLOCK>
__device__ void apply_path_block(const Lat& g, const int32_t* path, int32_t len, int32_t sign) {
    for (int32_t i = threadIdx.x; i < len; i += BLOCK) atomicAdd(g.node_use + path[i], sign);
    for (int32_t i = threadIdx.x; i + 1 < len; i += BLOCK) {
        int64_t u = path[i], v = path[i + 1];
        int s = lat_slot(g, u, v);
        atomicAdd(g.occ + 4 * u + s, sign);
        atomicAdd(g.occ + 4 * v + (3 - s), sign);
        bool vert = (s == 0 || s == 3);
        bool prev_vert = false;
        if (i > 0) {
            int ps = lat_slot(g, path[i - 1], u);
            prev_vert = (ps == 0 || ps == 3);
        }
        if (vert && !prev_vert) atomicAdd(g.via_col + (u % g.XY), sign);
    }
}

// Re-price every in-arc of every path node: covers both arcs of every hop
// (each hop arc's head is a path node) and the node-term change.
__device__ __forceinline__ void reprice_in_arcs(const Lat& g, const Params& P, int64_t v) {
    float q = node_cost(g.node_hist[v], __hip_atomic_load(g.node_use + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), P);
    for (int t = 0; t < 4; ++t) {
        int64_t n = lat_nbr(g, v, t);
        if (n < 0) continue;
        int64_t a = 4 * n + (3 - t);
        // the arc's three loads go out together (no branch between them)
        const float b = g.base[a], hst = g.hist[a];
        const int32_t o = __hip_atomic_load(g.occ + a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        g.wcost[a] = isinf(b) ? INFINITY : __fadd_rn(arc_cost(b, hst, o, P), q);
    }
}

template <int BLOCK>
__device__ void reprice_path_block(const Lat& g, const Params& P, const int32_t* path, int32_t len) {
    for (int32_t i = threadIdx.x; i < len; i += BLOCK) reprice_in_arcs(g, P, path[i]);
}

// A* heuristic toward the sink box [tx0, tx1] x [ty0, ty1] x [tz0, tz1]
// (global coords; one node for a two-pin net, the targets' bounding box for
// a multi-terminal net):
//   h(v) = cp*(dx + dy) + cv*dz, d* = distance to the box along that axis,
// evaluated exactly in double.  Every sink has h = 0.
// cp / cv are per-move lower bounds of the entering weight reduced by a
// margin of at least half an fp32 ulp below gmax, so along every tight arc
// u->v with g(v) < gmax, f(u) = g(u) + h(u) <= f(v): bucketing by f settles
// exactly the canonical keys of the sinks and their parent chains.  A result
// with a sink at g >= gmax is recomputed without the heuristic by the caller.
struct Heur {
    bool on;
    double cp, cv;
    int32_t tx0, tx1, ty0, ty1, tz0, tz1;
    __device__ __forceinline__ double h(int32_t x, int32_t y, int32_t z) const {
        if (!on) return 0.0;
        const int32_t dx = x < tx0 ? tx0 - x : (x > tx1 ? x - tx1 : 0);
        const int32_t dy = y < ty0 ? ty0 - y : (y > ty1 ? y - ty1 : 0);
        const int32_t dz = z < tz0 ? tz0 - z : (z > tz1 ? z - tz1 : 0);
        return cp * (double)(dx + dy) + cv * (double)dz;
    }
};

__device__ __forceinline__ double box_h(const Heur& H, const Box& B, uint32_t lx, uint32_t ly, uint32_t z) {
    return H.h(B.x0 + (int32_t)lx, B.y0 + (int32_t)ly, (int32_t)z);
}

// Heuristic of the neighbour of (lx, ly, z) in slot s.
__device__ __forceinline__ double nbr_h(const Heur& H, const Box& B, bool horiz, uint32_t lx, uint32_t ly, uint32_t z,
                                       int s) {
    int32_t x = B.x0 + (int32_t)lx, y = B.y0 + (int32_t)ly, zz = (int32_t)z;
    switch (s) {
    case 0: zz -= 1; break;
    case 1: if (horiz) x -= 1; else y -= 1; break;
    case 2: if (horiz) x += 1; else y += 1; break;
    default: zz += 1; break;
    }
    return H.h(x, y, zz);
}

__device__ __forceinline__ unsigned long long dbits(double d) { return (unsigned long long)__double_as_longlong(d); }

// Settled-set statistics of the finished search: the nodes a label-setting
// search popping in (f, key, node id) order settles up to and including the
// sink (every reached node without a sink), and the in-ROI out-degree of
// those popped before the sink.  ROI-local indices order like global ids.
// Every node ordered before the sink has f <= f(dst) < T, so its key is final.
// With a terminal table, rdst is the chosen target (full: none chosen) and
// no target contributes out-arcs (targets are sinks).
template <int BLOCK>
__device__ void box_stats(const Lat& g, const Box& B, const Heur& H, const uint64_t* keys, uint32_t rdst, bool full,
                          Ctl* ctl, const TermTab* TT = nullptr) {
    const uint64_t kd = full ? ORX_KEY_INF : key_load(keys + rdst);
    const double fd = full ? (double)INFINITY : (double)kdist(kd);
    unsigned long long settled = 0, relaxed = 0;
    for (uint32_t r = threadIdx.x; r < B.R; r += BLOCK) {
        const uint64_t k = key_load(keys + r);
        if (k == ORX_KEY_INF) continue;
        uint32_t lx, ly, z;
        roi_decode(B, r, lx, ly, z);
        if (!full) {
            const double f = (double)kdist(k) + box_h(H, B, lx, ly, z);
            if (f > fd || (f == fd && (k > kd || (k == kd && r > rdst)))) continue;
        }
        settled += 1;
        if (TT ? is_target(TT, r, lx, ly, z) : (!full && r == rdst)) continue;
        const bool horiz = (g.hmask >> z) & 1u;
        for (int s = 0; s < 4; ++s) relaxed += roi_nbr(B, horiz, r, lx, ly, z, s) != UINT32_MAX;
    }
    settled = wave_sum(settled);
    relaxed = wave_sum(relaxed);
    if (lane_id() == 0) {
        atomicAdd(&ctl->settled, settled);
        atomicAdd(&ctl->relaxed, relaxed);
    }
}


// Parent walk by one wavefront, a straight run per memory round trip.  A
// node's parent lies one move away in the direction its in-slot names, and a
// shortest path on the layered lattice is a few long straight runs (a turn
// costs a via).  So from the current node the 64 lanes load the keys of the
// next 64 nodes in its parent direction at once; the run continues through
// every node whose in-slot is the same (same slot, same direction: planar
// moves keep their layer), and the first node whose slot differs (a turn or
// the root) ends the trip and starts the next one.  Lane j writes the path
// entry of the j-th node of the run, root first.  The walk must end at rsrc,
// or at any root (a seed) when rsrc is UINT32_MAX.  Returns the path length,
// 0 when rdst was not reached, or a negative status.
// Call with all 64 lanes of one wave; every lane returns the same value.
__device__ int32_t box_backtrace_wave(const Lat& g, const Box& B, const uint64_t* keys, uint32_t rsrc,
                                      uint32_t rdst, int32_t* path, int32_t cap) {
    const int32_t lane = (int32_t)lane_id();
    const uint64_t kd = key_load(keys + rdst);
    if (kd == ORX_KEY_INF) return 0;
    const uint32_t hops = khops(kd);
    if ((int64_t)hops + 1 > cap) return ORX_E_OVERFLOW;
    uint32_t ux, uy, uz;
    roi_decode(B, rdst, ux, uy, uz);
    int32_t cx = (int32_t)ux, cy = (int32_t)uy, cz = (int32_t)uz;
    int32_t i = (int32_t)hops;          // path index of the current node
    uint32_t sl = kslot(kd);            // its in-slot
    if (lane == 0) path[i] = (int32_t)roi_global(g, B, (uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
    for (;;) {
        if (i == 0) {
            const uint32_t v = ((uint32_t)cz * (uint32_t)B.h + (uint32_t)cy) * (uint32_t)B.w + (uint32_t)cx;
            return (sl == ORX_ROOT_SLOT && (rsrc == UINT32_MAX || v == rsrc)) ? (int32_t)hops + 1 : ORX_E_CYCLE;
        }
        if (sl > 3) return ORX_E_CYCLE;
        int32_t mx = 0, my = 0, mz = 0;
        if (sl == 0) {
            mz = -1;
        } else if (sl == 3) {
            mz = 1;
        } else {
            const int32_t d = sl == 1 ? -1 : 1;
            if ((g.hmask >> cz) & 1u) mx = d;
            else my = d;
        }
        // lane j: the node j + 1 moves ahead
        const int32_t nx = cx + (lane + 1) * mx, ny = cy + (lane + 1) * my, nz = cz + (lane + 1) * mz;
        const bool inside = nx >= 0 && nx < B.w && ny >= 0 && ny < B.h && nz >= 0 && nz < B.L;
        uint64_t k = ORX_KEY_INF;
        if (inside) k = key_load(keys + ((uint32_t)nz * (uint32_t)B.h + (uint32_t)ny) * (uint32_t)B.w + (uint32_t)nx);
        // the run ends at the first lane that is outside the box or turns
        const uint64_t brk = __ballot(!inside || kslot(k) != sl);
        const int32_t b = brk ? (int32_t)__ffsll((long long)brk) - 1 : 64;
        if (b < 64 && !((__ballot(inside) >> b) & 1ull)) return ORX_E_CYCLE;   // a parent outside the box
        const int32_t take = b < 64 ? b + 1 : 64;   // nodes of the run, the turning node included
        if (take > i) return ORX_E_CYCLE;          // longer than the target's hop count
        if (lane < take) path[i - 1 - lane] = (int32_t)roi_global(g, B, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz);
        sl = kslot(__shfl(k, take - 1));
        cx += take * mx; cy += take * my; cz += take * mz;
        i -= take;
    }
}

__device__ __forceinline__ Heur make_heur(const Lat& g, const Params& P, int64_t dst, bool allowed) {
    Heur H;
    H.on = allowed && P.astar_cp > 0.0f && P.astar_cv > 0.0f && P.astar_gmax > 0.0f;
    H.cp = P.astar_cp;
    H.cv = P.astar_cv;
    H.tz0 = H.tz1 = (int32_t)(dst / g.XY);
    const int64_t rem = dst - (int64_t)H.tz0 * g.XY;
    H.ty0 = H.ty1 = (int32_t)(rem / g.X);
    H.tx0 = H.tx1 = (int32_t)(rem - (int64_t)H.ty0 * g.X);
    return H;
}

// Heuristic toward the bounding box of a terminal table's targets.
__device__ __forceinline__ Heur make_heur_box(const Params& P, const Box& B, const TermTab& T, bool allowed) {
    Heur H;
    H.on = allowed && P.astar_cp > 0.0f && P.astar_cv > 0.0f && P.astar_gmax > 0.0f;
    H.cp = P.astar_cp;
    H.cv = P.astar_cv;
    H.tx0 = B.x0 + (int32_t)T.lx0; H.tx1 = B.x0 + (int32_t)T.lx1;
    H.ty0 = B.y0 + (int32_t)T.ly0; H.ty1 = B.y0 + (int32_t)T.ly1;
    H.tz0 = (int32_t)T.z0; H.tz1 = (int32_t)T.z1;
    return H;
}

__device__ __forceinline__ Box make_box(const int32_t* b, int32_t L) {
    Box B;
    B.x0 = b[0]; B.y0 = b[1]; B.w = b[2]; B.h = b[3]; B.L = L;
    B.wh = (uint32_t)B.w * (uint32_t)B.h;
    B.R = B.wh * (uint32_t)L;
    B.dw = FastDiv((uint32_t)B.w);
    B.dwh = FastDiv(B.wh);
    return B;
}

// Terminal table of one net: its targets inside the box (sorted ROI indices
// and their bounding box) and the number of its seeds inside the box.
template <int BLOCK>
__device__ void build_terms(const Lat& g, const Box& B, const int32_t* tn, int32_t nt, const int32_t* sn, int32_t ns,
                            TermTab& T) {
    static_assert(kMaxTargets <= BLOCK, "one lane per target");
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        T.n = 0; T.seeds_in = 0;
        T.lx0 = T.ly0 = T.z0 = UINT32_MAX;
        T.lx1 = T.ly1 = T.z1 = 0;
    }
    __syncthreads();
    if ((int32_t)tid < nt) {
        const uint32_t r = global_to_roi(g, B, tn[tid]);
        if (r != UINT32_MAX) {
            uint32_t lx, ly, z;
            roi_decode(B, r, lx, ly, z);
            T.roi[atomicAdd(&T.n, 1u)] = r;
            atomicMin(&T.lx0, lx); atomicMax(&T.lx1, lx);
            atomicMin(&T.ly0, ly); atomicMax(&T.ly1, ly);
            atomicMin(&T.z0, z); atomicMax(&T.z1, z);
        }
    }
    uint32_t sin = 0;
    for (int32_t i = (int32_t)tid; i < ns; i += BLOCK) sin += global_to_roi(g, B, sn[i]) != UINT32_MAX ? 1u : 0u;
    sin = wave_sum(sin);
    if (lane_id() == 0 && sin) atomicAdd(&T.seeds_in, sin);
    __syncthreads();
    // ascending order: each entry's rank (ties by position)
    const uint32_t n = T.n;
    uint32_t v = 0, rank = 0;
    if (tid < n) {
        v = T.roi[tid];
        for (uint32_t j = 0; j < n; ++j) {
            const uint32_t o = T.roi[j];
            rank += (o < v || (o == v && j < tid)) ? 1u : 0u;
        }
    }
    __syncthreads();
    if (tid < n) T.roi[rank] = v;
    __syncthreads();
}

// Best target of a finished multi-terminal search, by one wavefront: the
// reached target with the smallest (fl(dist + target cost), node id) and a
// finite total (UP:2605-2618 keeps the first strict improvement in pop
// order).  Targets the search did not certify have totals above the
// certified best, so they never win.
__device__ void pick_target(const Lat& g, const Box& B, const uint64_t* keys, const int32_t* tn, const float* tc,
                            int32_t nt, Ctl& ctl) {
    const uint32_t lane = lane_id();
    float bt = INFINITY;
    int64_t bg = INT64_MAX;
    for (int32_t i = (int32_t)lane; i < nt; i += 64) {
        const int64_t gid = tn[i];
        const uint32_t r = global_to_roi(g, B, gid);
        if (r == UINT32_MAX) continue;
        const uint64_t k = key_load(keys + r);
        if (k == ORX_KEY_INF) continue;
        const float t = __fadd_rn(kdist(k), tc ? tc[i] : 0.0f);
        if (!(t < INFINITY)) continue;
        if (t < bt || (t == bt && gid < bg)) { bt = t; bg = gid; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float ot = __shfl_xor(bt, o);
        const int64_t og = __shfl_xor(bg, o);
        if (ot < bt || (ot == bt && og < bg)) { bt = ot; bg = og; }
    }
    if (lane == 0) {
        ctl.best_r = bg == INT64_MAX ? UINT32_MAX : global_to_roi(g, B, bg);
        ctl.best_cost = bt;
    }
}


// ---------------------------------------------------------------------------
// Host-side plumbing shared by the entry points.
// ---------------------------------------------------------------------------
Lat make_lat(const orx_lattice* l) {
    Lat g;
    g.X = l->X; g.Y = l->Y; g.L = l->L; g.hmask = l->hmask;
    g.XY = (int64_t)l->X * l->Y;
    g.base = l->base; g.hist = l->hist; g.occ = l->occ; g.wcost = l->wcost;
    g.node_hist = l->node_hist; g.node_use = l->node_use; g.via_col = l->via_col;
    return g;
}
Params make_params(const orx_params* p) {
    Params P;
    P.pres_fac = p->pres_fac; P.base_w = p->base_w; P.hist_w = p->hist_w; P.node_pres = p->node_pres;
    P.node_hist_w = p->node_hist_w; P.edge_cap = p->edge_cap; P.node_cap = p->node_cap;
    P.hist_gain = p->hist_gain; P.hist_decay = p->hist_decay; P.hist_cap_mult = p->hist_cap_mult;
    P.node_hist_gain = p->node_hist_gain;
    P.astar_cp = p->astar_cp; P.astar_cv = p->astar_cv; P.astar_gmax = p->astar_gmax;
    return P;
}

int check_lattice(const orx_lattice* l) {
    if (!l) return orx_host::fail(ORX_E_ARG, "lattice is NULL");
    if (l->X < 1 || l->Y < 1 || l->L < 1 || l->L > 32)
        return orx_host::fail(ORX_E_ARG, "lattice dims must be X,Y >= 1 and 1 <= L <= 32");
    if ((int64_t)l->X * l->Y * l->L >= (1ll << 31))
        return orx_host::fail(ORX_E_ARG, "lattice too large for int32 node ids");
    if (!l->base || !l->hist || !l->occ || !l->wcost || !l->node_hist || !l->node_use || !l->via_col)
        return orx_host::fail(ORX_E_ARG, "lattice has a NULL array");
    if (((uintptr_t)l->base | (uintptr_t)l->hist | (uintptr_t)l->occ | (uintptr_t)l->wcost) & 15u)
        return orx_host::fail(ORX_E_ARG, "arc arrays must be 16-byte aligned");
    return ORX_OK;
}

// Largest dynamic LDS a launch may ask for: 160 KiB per CU less a reserve
// for the kernels' static LDS (the search control block, a few hundred
// bytes).
constexpr size_t kMaxDynLds = 160 * 1024 - 1024;

// Raise a kernel's dynamic-LDS limit once (orx_init sets every kernel to
// kMaxDynLds up front, so launches -- also those recorded into a captured
// graph -- make no attribute calls).
template <auto KERNEL>
int set_lds(size_t bytes) {
    static std::atomic<size_t> configured{0};
    if (bytes <= configured.load(std::memory_order_relaxed)) return ORX_OK;
    ORX_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(KERNEL),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    size_t cur = configured.load(std::memory_order_relaxed);
    while (cur < bytes && !configured.compare_exchange_weak(cur, bytes)) {
    }
    return ORX_OK;
}


int device_cus() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            cus = n;
        else
            return 1;
    }
    return cus;
}

NetsDev make_nets(const orx_nets* nets) {
    return NetsDev{nets->n, nets->box, nets->src, nets->dst, nets->path_arena, nets->path_off,
                   nets->path_cap, nets->path_len, nets->path_cost, nets->status, nets->seed_ptr,
                   nets->seed_node, nets->seed_cost, nets->tgt_ptr, nets->tgt_node, nets->tgt_cost};
}

}  // namespace
