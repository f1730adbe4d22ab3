/*
 * orx_oracle.c -- CPU restatement of OrthoRoute's PathFinder hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests, the smoke
 * check and the CPU-baseline leg of the benchmark link it as the checker.  The
 * product path (orthoroute_amd) never calls it.
 *
 * What it restates (all citations are files of bbenchoff/OrthoRoute):
 *   - heap Dijkstra on an ROI, early exit at the sink, parent walk
 *       orthoroute/algorithms/manhattan/unified_pathfinder.py:2451-2548
 *   - multi-seed / multi-target Dijkstra (targets are not expanded)
 *       unified_pathfinder.py:2550-2659, seed pricing cuda_dijkstra.py:5589-5606
 *   - node penalty folded into the entered arc's weight
 *       cuda_dijkstra.py:5079-5081 (tests/test_gpu_fallback.py:104-117)
 *   - EdgeAccountant: commit/clear scatter-add of both arcs of a physical edge,
 *       live prospective present cost, history update, overuse
 *       unified_pathfinder.py:1074-1135, 1226-1289, 1301-1422, 1137-1190,
 *       _path_to_edges 9981-10046, canonical arc mask 4919-4932
 *   - path-node use and via-column (barrel) accounting
 *       unified_pathfinder.py:4470-4546, 4623-4659
 *   - per-net rip-up -> route -> commit -> re-price loop and the iteration
 *       order refresh -> cost -> route -> overuse -> history -> escalate
 *       unified_pathfinder.py:7903-8449, 6163-7391
 *
 * Canonical semantics (both the oracle and the HIP path implement exactly this):
 *   key(v) = (dist f32, hops, in-slot) compared lexicographically, packed as
 *   u64 = dist_bits<<32 | hops<<8 | slot.  dist is the left fold of fp32 adds
 *   along the path, d' = fl(d + w), w = fl(c_arc + q_head).  The in-slot is
 *   the rank of the predecessor among v's in-neighbours sorted by node id, so
 *   the minimum key picks minimal distance, then minimal hops, then the lowest
 *   predecessor id.  The reference's heapq pops by (dist, id) and keeps the
 *   first strict improvement; this fixed tie-break replaces that order
 *   dependence and makes parents acyclic under fp32 absorption
 *   (tests/test_gpu_fallback.py:148-173).
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off; no FMA contraction so
 * every fp32 operation rounds exactly once, in the written order).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include <math.h>

#define KEY_INF 0xFFFFFFFFFFFFFFFFull
#define ROOT_SLOT 0xFFu
#define HOPS_MAX 0xFFFFFFu

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint64_t kpack(float d, uint32_t h, uint32_t s) {
    return ((uint64_t)f2u(d) << 32) | ((uint64_t)h << 8) | (uint64_t)s;
}
static inline float kdist(uint64_t k) { return u2f((uint32_t)(k >> 32)); }
static inline uint32_t khops(uint64_t k) { return (uint32_t)(k >> 8) & HOPS_MAX; }
static inline uint32_t kslot(uint64_t k) { return (uint32_t)k & 0xFFu; }

/* ------------------------------------------------------------------------ */
/* Cost parameters (mirrors orx_cost_params in include/orthoroute_hip.h)     */
/* ------------------------------------------------------------------------ */
typedef struct {
    float pres_fac;        /* present factor of this iteration              */
    float base_w;          /* base-cost weight (reference base_cost_weight)  */
    float hist_w;          /* history weight                                 */
    float node_pres;       /* present factor applied to node overuse         */
    float node_hist_w;     /* node-history weight                            */
    int32_t edge_cap;      /* capacity of one physical edge                  */
    int32_t node_cap;      /* capacity of one node                           */
    float hist_gain;       /* history increment per unit of overuse          */
    float hist_decay;      /* history *= decay before the increment          */
    float hist_cap_mult;   /* increment <= mult*base when > 0                 */
    float node_hist_gain;  /* node-history increment per unit of overuse     */
    float astar_cp;        /* A* planar step bound (0 = plain Dijkstra)      */
    float astar_cv;        /* A* via step bound                              */
    float astar_gmax;      /* A* result trusted only below this distance     */
} orx_params;

/* c_arc = fl(fl(fl(b*bw) + fl(hw*h)) + fl(p*max(0, occ+1-cap)))
 * unified_pathfinder.py:1348-1361 (static part) and 1403-1422 (prospective
 * present term re-priced live per net). */
static inline float arc_cost(float b, float h, int32_t occ, const orx_params* P) {
    int32_t over = occ + 1 - P->edge_cap;
    if (over < 0) over = 0;
    float t1 = b * P->base_w;
    float t2 = P->hist_w * h;
    float t3 = t1 + t2;
    float t4 = P->pres_fac * (float)over;
    return t3 + t4;
}
/* q_v = fl(fl(nhw*nh) + fl(np*max(0, use+1-cap)))  -- ownership-as-cost node
 * term, unified_pathfinder.py:5308-5383 (priced when the node is entered). */
static inline float node_cost(float nh, int32_t use, const orx_params* P) {
    int32_t over = use + 1 - P->node_cap;
    if (over < 0) over = 0;
    float t1 = P->node_hist_w * nh;
    float t2 = P->node_pres * (float)over;
    return t1 + t2;
}

/* ------------------------------------------------------------------------ */
/* Lattice: X*Y*L, node id z*X*Y + y*X + x (unified_pathfinder.py:118,       */
/* kicad_geometry.py:59-62).  4 arc slots per node, ordered by neighbour id: */
/* 0 = via down (z-1), 1 = planar minus, 2 = planar plus, 3 = via up (z+1).  */
/* Planar axis per layer: bit z of hmask set = horizontal (x), else vertical */
/* (y); the reference assigns layer 0 = V, odd = H, even = V                 */
/* (unified_pathfinder.py:1487-1495).  Reverse of slot s is slot 3-s.        */
/* ------------------------------------------------------------------------ */
typedef struct { int32_t X, Y, L; uint32_t hmask; } lat_t;

static inline int64_t lat_nbr(const lat_t* g, int64_t u, int s) {
    int64_t XY = (int64_t)g->X * g->Y;
    int32_t z = (int32_t)(u / XY);
    int64_t rem = u - (int64_t)z * XY;
    int32_t y = (int32_t)(rem / g->X);
    int32_t x = (int32_t)(rem - (int64_t)y * g->X);
    int horiz = (g->hmask >> z) & 1u;
    switch (s) {
    case 0: return z > 0 ? u - XY : -1;
    case 1: if (horiz) return x > 0 ? u - 1 : -1; return y > 0 ? u - g->X : -1;
    case 2: if (horiz) return x < g->X - 1 ? u + 1 : -1; return y < g->Y - 1 ? u + g->X : -1;
    default: return z < g->L - 1 ? u + XY : -1;
    }
}

/* slot of the arc u->v (v must be a lattice neighbour of u), -1 otherwise */
static inline int lat_slot(const lat_t* g, int64_t u, int64_t v) {
    for (int s = 0; s < 4; ++s) if (lat_nbr(g, u, s) == v) return s;
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Binary min-heap of (f, key, node) with lazy deletion                     */
/* ------------------------------------------------------------------------ */
typedef struct { double f; uint64_t k; int32_t v; } hent;
typedef struct { hent* a; int64_t n, cap; } heap_t;

/* Pop order (f, key, node), f = g + heuristic (f = g for plain Dijkstra):
 * equal packed keys pop by node id, so "the nodes settled up to the sink"
 * is canonical. */
static inline int hless(double fa, uint64_t ka, int32_t va, double fb, uint64_t kb, int32_t vb) {
    return fa < fb || (fa == fb && (ka < kb || (ka == kb && va < vb)));
}

static int heap_push(heap_t* h, double f, uint64_t k, int32_t v) {
    if (h->n == h->cap) {
        int64_t nc = h->cap ? h->cap * 2 : 1024;
        hent* na = (hent*)realloc(h->a, (size_t)nc * sizeof(hent));
        if (!na) return -1;
        h->a = na; h->cap = nc;
    }
    int64_t i = h->n++;
    while (i > 0) {
        int64_t p = (i - 1) >> 1;
        if (!hless(f, k, v, h->a[p].f, h->a[p].k, h->a[p].v)) break;
        h->a[i] = h->a[p];
        i = p;
    }
    h->a[i].f = f; h->a[i].k = k; h->a[i].v = v;
    return 0;
}
static hent heap_pop(heap_t* h) {
    hent top = h->a[0];
    hent last = h->a[--h->n];
    int64_t i = 0, n = h->n;
    for (;;) {
        int64_t l = 2 * i + 1;
        if (l >= n) break;
        int64_t c = (l + 1 < n && hless(h->a[l + 1].f, h->a[l + 1].k, h->a[l + 1].v,
                                        h->a[l].f, h->a[l].k, h->a[l].v)) ? l + 1 : l;
        if (!hless(h->a[c].f, h->a[c].k, h->a[c].v, last.f, last.k, last.v)) break;
        h->a[i] = h->a[c];
        i = c;
    }
    if (n > 0) h->a[i] = last;
    return top;
}

/* ------------------------------------------------------------------------ */
/* Box-ROI lattice Dijkstra.  ROI = [x0,x0+w) x [y0,y0+h) x [0,L).           */
/* ROI-local index r = (z*h + ly)*w + lx.                                    */
/* The arc weight is computed from the accountant state exactly as the       */
/* device keeps it: w = fl(arc_cost + node_cost(head)).                      */
/* ------------------------------------------------------------------------ */
typedef struct {
    const float* base; const float* hist; const int32_t* occ;
    const float* node_hist; const int32_t* node_use;
    const float* wcost;     /* optional precomputed [4N] weights (SSSP bench) */
} lat_state;

static inline float arc_weight(const lat_t* g, const lat_state* S, const orx_params* P,
                               int64_t u, int s, int64_t v) {
    (void)g;
    if (S->wcost) return S->wcost[4 * u + s];
    float b = S->base[4 * u + s];
    if (isinf(b)) return INFINITY;
    float c = arc_cost(b, S->hist[4 * u + s], S->occ[4 * u + s], P);
    float q = node_cost(S->node_hist[v], S->node_use[v], P);
    return c + q;
}

typedef struct {
    uint64_t* key; uint8_t* done; uint8_t* tmark; heap_t heap; int64_t cap;
} sssp_ws;

/* ROI-local index of global node gid in box (x0,y0,w,h), -1 when outside */
static inline int64_t box_index(const lat_t* g, int32_t x0, int32_t y0, int32_t w, int32_t h, int64_t gid) {
    int64_t XY = (int64_t)g->X * g->Y;
    int32_t z = (int32_t)(gid / XY);
    int64_t rem = gid - (int64_t)z * XY;
    int32_t y = (int32_t)(rem / g->X), x = (int32_t)(rem - (int64_t)y * g->X);
    if (x < x0 || x >= x0 + w || y < y0 || y >= y0 + h) return -1;
    return ((int64_t)z * h + (y - y0)) * w + (x - x0);
}

static int ws_reserve(sssp_ws* W, int64_t R) {
    if (R <= W->cap) return 0;
    free(W->key); free(W->done); free(W->tmark);
    W->key = (uint64_t*)malloc((size_t)R * 8);
    W->done = (uint8_t*)malloc((size_t)R);
    W->tmark = (uint8_t*)calloc((size_t)R, 1);
    if (!W->key || !W->done || !W->tmark) return -1;
    W->cap = R;
    return 0;
}
static void ws_free(sssp_ws* W) {
    free(W->key); free(W->done); free(W->tmark); free(W->heap.a);
    memset(W, 0, sizeof(*W));
}

/* A* heuristic toward the sink box [tx0,tx1] x [ty0,ty1] x [tz0,tz1] (the
 * sink itself for a two-pin net, the targets' bounding box for a
 * multi-terminal net):
 *   h(v) = cp*(dx + dy) + cv*dz, d* = distance to the box along that axis,
 * evaluated exactly in double; h = 0 on every sink.
 * cp / cv are per-move lower bounds of the entering weight minus a margin
 * m >= half an fp32 ulp below gmax, so for every tight arc u->v with
 * g(v) < gmax, f(u) = g(u)+h(u) <= g(v)+h(v) = f(v): popping by f settles
 * exactly the canonical keys.  A result with a sink at g >= gmax is
 * recomputed without the heuristic (box_route, box_route_multi). */
typedef struct { int on; double cp, cv; int32_t tx0, tx1, ty0, ty1, tz0, tz1; } astar_t;

static inline int32_t axis_gap(int32_t v, int32_t lo, int32_t hi) {
    return v < lo ? lo - v : (v > hi ? v - hi : 0);
}
static inline double heur(const astar_t* A, const lat_t* g, int64_t gid) {
    if (!A || !A->on) return 0.0;
    int64_t XY = (int64_t)g->X * g->Y;
    int32_t z = (int32_t)(gid / XY);
    int64_t rem = gid - (int64_t)z * XY;
    int32_t y = (int32_t)(rem / g->X), x = (int32_t)(rem - (int64_t)y * g->X);
    int32_t dx = axis_gap(x, A->tx0, A->tx1), dy = axis_gap(y, A->ty0, A->ty1), dz = axis_gap(z, A->tz0, A->tz1);
    return A->cp * (double)(dx + dy) + A->cv * (double)dz;
}

/* Returns 0 when dst was reached (or full sweep), 1 when unreachable, <0 on error.
 * stats[0] += settled nodes, stats[1] += relaxed arcs (out-arcs of settled
 * nodes inside the ROI, each arc once). */
static int box_sssp(const lat_t* g, const lat_state* S, const orx_params* P,
                    int32_t x0, int32_t y0, int32_t w, int32_t h,
                    int64_t src, int64_t dst, int full, sssp_ws* W, int64_t* stats,
                    const astar_t* A) {
    const int64_t XY = (int64_t)g->X * g->Y;
    const int64_t R = (int64_t)w * h * g->L;
    if (ws_reserve(W, R)) return -2;
    for (int64_t r = 0; r < R; ++r) { W->key[r] = KEY_INF; W->done[r] = 0; }
    W->heap.n = 0;
#define G2R(gid) ({ int64_t _g = (gid); int32_t _z = (int32_t)(_g / XY); int64_t _rem = _g - (int64_t)_z * XY; \
        int32_t _y = (int32_t)(_rem / g->X); int32_t _x = (int32_t)(_rem - (int64_t)_y * g->X); \
        (_x < x0 || _x >= x0 + w || _y < y0 || _y >= y0 + h) ? (int64_t)-1 : \
        ((int64_t)_z * h + (_y - y0)) * w + (_x - x0); })
#define R2G(r) ({ int64_t _r = (r); int64_t _wh = (int64_t)w * h; int32_t _z = (int32_t)(_r / _wh); \
        int64_t _rm = _r - _z * _wh; int32_t _ly = (int32_t)(_rm / w); int32_t _lx = (int32_t)(_rm - (int64_t)_ly * w); \
        (int64_t)_z * XY + (int64_t)(y0 + _ly) * g->X + (x0 + _lx); })
    int64_t rs = G2R(src);
    int64_t rd = dst >= 0 ? G2R(dst) : -1;
    if (rs < 0 || (dst >= 0 && rd < 0)) return -3;
    W->key[rs] = kpack(0.0f, 0, ROOT_SLOT);
    if (heap_push(&W->heap, heur(A, g, src), W->key[rs], (int32_t)rs)) return -2;
    while (W->heap.n) {
        hent e = heap_pop(&W->heap);
        if (W->done[e.v] || e.k != W->key[e.v]) continue;
        W->done[e.v] = 1;
        stats[0] += 1;
        if (!full && e.v == rd) break;
        int64_t gu = R2G(e.v);
        float du = kdist(e.k);
        uint32_t hu = khops(e.k);
        for (int s = 0; s < 4; ++s) {
            int64_t gv = lat_nbr(g, gu, s);
            if (gv < 0) continue;
            int64_t rv = G2R(gv);
            if (rv < 0) continue;
            stats[1] += 1;
            if (W->done[rv]) continue;
            float wt = arc_weight(g, S, P, gu, s, gv);
            if (isinf(wt)) continue;
            float nd = du + wt;
            if (hu + 1 > HOPS_MAX) return -4;
            uint64_t nk = kpack(nd, hu + 1, (uint32_t)(3 - s));
            if (nk < W->key[rv]) {
                W->key[rv] = nk;
                if (heap_push(&W->heap, (double)nd + heur(A, g, gv), nk, (int32_t)rv)) return -2;
            }
        }
    }
    if (dst >= 0 && W->key[rd] == KEY_INF) return 1;
    return 0;
#undef G2R
#undef R2G
}

/* Sink-directed search of the router: A* when P enables it, recomputed
 * without the heuristic when the result is not below astar_gmax.  Stats are
 * those of the search whose keys are kept. */
static int box_route(const lat_t* g, const lat_state* S, const orx_params* P,
                     int32_t x0, int32_t y0, int32_t w, int32_t h,
                     int64_t src, int64_t dst, sssp_ws* W, int64_t* stats) {
    int64_t st[2] = {0, 0};
    if (P->astar_cp > 0.0f && P->astar_cv > 0.0f && P->astar_gmax > 0.0f) {
        const int64_t XY = (int64_t)g->X * g->Y;
        astar_t A;
        A.on = 1; A.cp = P->astar_cp; A.cv = P->astar_cv;
        A.tz0 = A.tz1 = (int32_t)(dst / XY);
        int64_t rem = dst - (int64_t)A.tz0 * XY;
        A.ty0 = A.ty1 = (int32_t)(rem / g->X); A.tx0 = A.tx1 = (int32_t)(rem - (int64_t)A.ty0 * g->X);
        int r = box_sssp(g, S, P, x0, y0, w, h, src, dst, 0, W, st, &A);
        if (r < 0 || r == 1) { stats[0] += st[0]; stats[1] += st[1]; return r; }
        uint64_t kd = W->key[box_index(g, x0, y0, w, h, dst)];
        if (kdist(kd) < P->astar_gmax) { stats[0] += st[0]; stats[1] += st[1]; return r; }
        st[0] = st[1] = 0;
    }
    int r = box_sssp(g, S, P, x0, y0, w, h, src, dst, 0, W, st, NULL);
    stats[0] += st[0]; stats[1] += st[1];
    return r;
}

/* Walk parents from dst; writes hops+1 global ids into path (src first).
 * Returns path length, or <0 on error. */
static int64_t box_backtrace(const lat_t* g, int32_t x0, int32_t y0, int32_t w, int32_t h,
                             int64_t dst, const sssp_ws* W, int32_t* path, int64_t cap) {
    const int64_t XY = (int64_t)g->X * g->Y;
    int32_t z = (int32_t)(dst / XY);
    int64_t rem = dst - (int64_t)z * XY;
    int32_t y = (int32_t)(rem / g->X), x = (int32_t)(rem - (int64_t)y * g->X);
    int64_t rd = ((int64_t)z * h + (y - y0)) * w + (x - x0);
    uint64_t k = W->key[rd];
    if (k == KEY_INF) return -1;
    int64_t len = (int64_t)khops(k) + 1;
    if (len > cap) return -5;
    int64_t v = dst;
    for (int64_t i = len - 1; i >= 0; --i) {
        path[i] = (int32_t)v;
        int32_t zz = (int32_t)(v / XY);
        int64_t rr = v - (int64_t)zz * XY;
        int32_t yy = (int32_t)(rr / g->X), xx = (int32_t)(rr - (int64_t)yy * g->X);
        int64_t r = ((int64_t)zz * h + (yy - y0)) * w + (xx - x0);
        uint32_t sl = kslot(W->key[r]);
        if (i == 0) { if (sl != ROOT_SLOT) return -6; break; }
        if (sl > 3) return -6;
        v = lat_nbr(g, v, (int)sl);
        if (v < 0) return -6;
    }
    return len;
}

/* ------------------------------------------------------------------------ */
/* Multi-terminal nets: find_path_multisource_multisink                      */
/* (unified_pathfinder.py:2550-2659) on the net's box.  Every seed inside    */
/* the box starts at fl(seed_cost + q_seed), the seed paying its own node    */
/* term (UP:2582-2583, cuda_dijkstra.py:5589-5606; a seed listed twice keeps */
/* the smaller key); targets are popped but never expanded (UP:2611-2618);   */
/* the search stops once every target in the box is popped.  The chosen     */
/* target minimises (fl(dist + target_cost), node id) over reached targets   */
/* with a finite total (UP:2612-2617 keeps the first strict improvement).    */
/* A one-node result (the best target is a seed) is no path (UP:2647-2648). */
/* ------------------------------------------------------------------------ */
typedef struct {
    const int32_t* seed_ptr;   /* [n+1]; NULL = every net is two-pin        */
    const int32_t* seed_node;
    const float* seed_cost;
    const int32_t* tgt_ptr;    /* [n+1] */
    const int32_t* tgt_node;
    const float* tgt_cost;     /* NULL = all 0 */
} orx_terms;

#define MAX_TARGETS 128

/* Smallest fp32 value >= x. */
static inline float f32_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

/* The search pops in (f, key, node) order and stops before popping a node of
 * f = F once the best certified target cannot be beaten: a target popped
 * with dist < F' = min(F, gmax) is certified (exact), and every other target
 * has a true dist >= F', so its total is at least fl(F'^ + tc) with F'^ the
 * smallest fp32 >= F'; the search stops when that bound exceeds the best
 * certified total for all of them (gmax = +inf without the heuristic).
 * Returns 0 when the search ran (targets reached or not), 1 when no seed or
 * no target lies in the box, <0 on error. */
static int box_sssp_multi(const lat_t* g, const lat_state* S, const orx_params* P,
                          int32_t x0, int32_t y0, int32_t w, int32_t h,
                          int32_t ns, const int32_t* sn, const float* sc, int32_t nt, const int32_t* tn,
                          const float* tc, sssp_ws* W, const astar_t* A, double gmax) {
    const int64_t XY = (int64_t)g->X * g->Y;
    const int64_t R = (int64_t)w * h * g->L;
    if (ws_reserve(W, R)) return -2;
    for (int64_t r = 0; r < R; ++r) { W->key[r] = KEY_INF; W->done[r] = 0; W->tmark[r] = 0; }
    W->heap.n = 0;
    int64_t tr[MAX_TARGETS];
    uint8_t cert[MAX_TARGETS];
    int in_box = 0;
    for (int32_t i = 0; i < nt; ++i) {
        tr[i] = box_index(g, x0, y0, w, h, tn[i]);
        cert[i] = 0;
        if (tr[i] >= 0) { W->tmark[tr[i]] = 1; in_box = 1; }
    }
    int seeded = 0;
    for (int32_t i = 0; i < ns; ++i) {
        int64_t r = box_index(g, x0, y0, w, h, sn[i]);
        if (r < 0) continue;
        float c = sc[i] + node_cost(S->node_hist[sn[i]], S->node_use[sn[i]], P);
        uint64_t k = kpack(c, 0, ROOT_SLOT);
        seeded = 1;
        if (k < W->key[r]) {
            W->key[r] = k;
            if (heap_push(&W->heap, (double)c + heur(A, g, sn[i]), k, (int32_t)r)) return -2;
        }
    }
    if (!seeded || !in_box) return 1;
    float best = INFINITY;     /* best certified total */
    int have = 0;
    float tcmin = INFINITY;    /* smallest cost among uncertified targets */
    for (int32_t i = 0; i < nt; ++i)
        if (tr[i] >= 0) tcmin = fminf(tcmin, tc ? tc[i] : 0.0f);
    const int64_t wh = (int64_t)w * h;
    while (W->heap.n) {
        hent e = heap_pop(&W->heap);
        if (W->done[e.v] || e.k != W->key[e.v]) continue;
        if (have) {
            const float lb = f32_up(e.f < gmax ? e.f : gmax) + tcmin;
            if (lb > best) break;
        }
        W->done[e.v] = 1;
        if (W->tmark[e.v]) {
            const float d = kdist(e.k);
            if ((double)d < gmax) {
                tcmin = INFINITY;
                for (int32_t i = 0; i < nt; ++i) {
                    if (tr[i] < 0) continue;
                    const float c = tc ? tc[i] : 0.0f;
                    if (tr[i] == e.v) {
                        cert[i] = 1;
                        const float t = d + c;
                        if (t < INFINITY && t < best) { best = t; have = 1; }
                    }
                    if (!cert[i]) tcmin = fminf(tcmin, c);
                }
            }
            continue;   /* targets are sinks */
        }
        int32_t z = (int32_t)(e.v / wh);
        int64_t rm = e.v - z * wh;
        int32_t ly = (int32_t)(rm / w), lx = (int32_t)(rm - (int64_t)ly * w);
        int64_t gu = (int64_t)z * XY + (int64_t)(y0 + ly) * g->X + (x0 + lx);
        float du = kdist(e.k);
        uint32_t hu = khops(e.k);
        for (int s = 0; s < 4; ++s) {
            int64_t gv = lat_nbr(g, gu, s);
            if (gv < 0) continue;
            int64_t rv = box_index(g, x0, y0, w, h, gv);
            if (rv < 0 || W->done[rv]) continue;
            float wt = arc_weight(g, S, P, gu, s, gv);
            if (isinf(wt)) continue;
            float nd = du + wt;
            if (hu + 1 > HOPS_MAX) return -4;
            uint64_t nk = kpack(nd, hu + 1, (uint32_t)(3 - s));
            if (nk < W->key[rv]) {
                W->key[rv] = nk;
                if (heap_push(&W->heap, (double)nd + heur(A, g, gv), nk, (int32_t)rv)) return -2;
            }
        }
    }
    return 0;
}

/* Best popped target, (fl(dist + tc), node id) over those with a finite
 * total: its global id (-1 when none) and total. */
static void pick_target(const lat_t* g, int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t nt,
                        const int32_t* tn, const float* tc, const sssp_ws* W, int64_t* best, float* total) {
    *best = -1;
    *total = INFINITY;
    for (int32_t i = 0; i < nt; ++i) {
        int64_t r = box_index(g, x0, y0, w, h, tn[i]);
        if (r < 0 || !W->done[r]) continue;
        float t = kdist(W->key[r]) + (tc ? tc[i] : 0.0f);
        if (!(t < INFINITY)) continue;
        if (t < *total || (t == *total && tn[i] < *best)) { *total = t; *best = tn[i]; }
    }
}

/* Search counts of a finished multi-terminal search, defined by the keys
 * alone: the nodes a label-setting search popping in (f, key, node) order
 * settles up to and including the chosen target (every reached node when
 * there is none), and the in-box out-degree of those that are not targets. */
static void count_upto(const lat_t* g, int32_t x0, int32_t y0, int32_t w, int32_t h, const sssp_ws* W,
                       const astar_t* A, int64_t best, int64_t* stats) {
    const int64_t XY = (int64_t)g->X * g->Y, wh = (int64_t)w * h, R = wh * g->L;
    double fb = INFINITY;
    uint64_t kb = KEY_INF;
    int64_t rb = -1;
    if (best >= 0) {
        rb = box_index(g, x0, y0, w, h, best);
        kb = W->key[rb];
        fb = (double)kdist(kb) + heur(A, g, best);
    }
    for (int64_t r = 0; r < R; ++r) {
        uint64_t k = W->key[r];
        if (k == KEY_INF) continue;
        int32_t z = (int32_t)(r / wh);
        int64_t rm = r - z * wh;
        int32_t ly = (int32_t)(rm / w), lx = (int32_t)(rm - (int64_t)ly * w);
        int64_t gid = (int64_t)z * XY + (int64_t)(y0 + ly) * g->X + (x0 + lx);
        if (best >= 0) {
            double f = (double)kdist(k) + heur(A, g, gid);
            if (f > fb || (f == fb && (k > kb || (k == kb && r > rb)))) continue;
        }
        stats[0] += 1;
        if (W->tmark[r]) continue;
        for (int s = 0; s < 4; ++s) {
            int64_t gv = lat_nbr(g, gid, s);
            if (gv >= 0 && box_index(g, x0, y0, w, h, gv) >= 0) stats[1] += 1;
        }
    }
}

/* Multi-terminal route: A* toward the targets' bounding box when P enables
 * it, redone without the heuristic when the chosen total is >= astar_gmax
 * (below it the chosen target, its total and its parent chain are exact:
 * the true best total is no larger, so the true best target lies below
 * gmax and is certified).  Returns 0 routed (best, total set), 1 no path,
 * <0 error. */
static int box_route_multi(const lat_t* g, const lat_state* S, const orx_params* P,
                           int32_t x0, int32_t y0, int32_t w, int32_t h,
                           int32_t ns, const int32_t* sn, const float* sc, int32_t nt, const int32_t* tn,
                           const float* tc, sssp_ws* W, int64_t* stats, int64_t* best, float* total) {
    const int64_t XY = (int64_t)g->X * g->Y;
    if (P->astar_cp > 0.0f && P->astar_cv > 0.0f && P->astar_gmax > 0.0f) {
        astar_t A;
        A.on = 1; A.cp = P->astar_cp; A.cv = P->astar_cv;
        A.tx0 = A.ty0 = A.tz0 = INT32_MAX;
        A.tx1 = A.ty1 = A.tz1 = INT32_MIN;
        for (int32_t i = 0; i < nt; ++i) {
            if (box_index(g, x0, y0, w, h, tn[i]) < 0) continue;
            int32_t z = (int32_t)(tn[i] / XY);
            int64_t rem = tn[i] - (int64_t)z * XY;
            int32_t y = (int32_t)(rem / g->X), x = (int32_t)(rem - (int64_t)y * g->X);
            if (x < A.tx0) A.tx0 = x;
            if (x > A.tx1) A.tx1 = x;
            if (y < A.ty0) A.ty0 = y;
            if (y > A.ty1) A.ty1 = y;
            if (z < A.tz0) A.tz0 = z;
            if (z > A.tz1) A.tz1 = z;
        }
        int r = box_sssp_multi(g, S, P, x0, y0, w, h, ns, sn, sc, nt, tn, tc, W, &A, (double)P->astar_gmax);
        if (r != 0) return r;
        pick_target(g, x0, y0, w, h, nt, tn, tc, W, best, total);
        if (*best < 0 || *total < P->astar_gmax) {
            count_upto(g, x0, y0, w, h, W, &A, *best, stats);
            return *best < 0 ? 1 : 0;
        }
    }
    int r = box_sssp_multi(g, S, P, x0, y0, w, h, ns, sn, sc, nt, tn, tc, W, NULL, INFINITY);
    if (r != 0) return r;
    pick_target(g, x0, y0, w, h, nt, tn, tc, W, best, total);
    count_upto(g, x0, y0, w, h, W, NULL, *best, stats);
    return *best < 0 ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Accountant: commit (+1) / rip-up (-1) of one path                        */
/* Both arcs of each hop (unified_pathfinder.py:10034-10046), every unique  */
/* path node (4635-4659), one via-column use per contiguous vertical run    */
/* (4500-4512).                                                             */
/* ------------------------------------------------------------------------ */
static int apply_path(const lat_t* g, const int32_t* path, int64_t len, int sign,
                      int32_t* occ, int32_t* node_use, int32_t* via_col) {
    const int64_t XY = (int64_t)g->X * g->Y;
    int prev_vert = 0;
    for (int64_t i = 0; i < len; ++i) node_use[path[i]] += sign;
    for (int64_t i = 0; i + 1 < len; ++i) {
        int64_t u = path[i], v = path[i + 1];
        int s = lat_slot(g, u, v);
        if (s < 0) return -7;
        occ[4 * u + s] += sign;
        occ[4 * v + (3 - s)] += sign;
        int vert = (s == 0 || s == 3);
        if (vert && !prev_vert) via_col[u % XY] += sign;
        prev_vert = vert;
    }
    return 0;
}

/* One commit (+1) or rip-up (-1) of one path, as the driver applies it
 * (exported for the accountant's known-answer checks). */
int orx_oracle_apply_path(int32_t X, int32_t Y, int32_t L, uint32_t hmask, const int32_t* path, int64_t len,
                          int32_t sign, int32_t* occ, int32_t* node_use, int32_t* via_col) {
    lat_t g = { X, Y, L, hmask };
    return apply_path(&g, path, len, sign, occ, node_use, via_col);
}

/* ------------------------------------------------------------------------ */
/* Iteration-end accounting: overuse (canonical arcs only, slots 2 and 3,    */
/* unified_pathfinder.py:4919-4932), then history                           */
/* h = fl(fl(h*decay) + min(fl(gain*over), fl(capm*b)))  (1226-1269, raw      */
/* present).  out[0..5] = edge_over, edge_over_count, node_over,             */
/* node_over_count, via_hops, barrels.                                       */
/* ------------------------------------------------------------------------ */
void orx_oracle_overuse(int32_t X, int32_t Y, int32_t L, const int32_t* occ, const int32_t* node_use,
                        const int32_t* via_col, int32_t edge_cap, int32_t node_cap, int64_t* out) {
    int64_t N = (int64_t)X * Y * L, eo = 0, ec = 0, no = 0, nc = 0, vh = 0, br = 0;
    for (int64_t u = 0; u < N; ++u) {
        for (int s = 2; s < 4; ++s) {
            int32_t o = occ[4 * u + s] - edge_cap;
            if (o > 0) { eo += o; ec += 1; }
        }
        vh += occ[4 * u + 3];
        int32_t on = node_use[u] - node_cap;
        if (on > 0) { no += on; nc += 1; }
    }
    for (int64_t c = 0; c < (int64_t)X * Y; ++c) br += via_col[c];
    out[0] = eo; out[1] = ec; out[2] = no; out[3] = nc; out[4] = vh; out[5] = br;
}

void orx_oracle_update_history(int64_t n_arcs, const float* base, float* hist, const int32_t* occ,
                               int64_t n_nodes, float* node_hist, const int32_t* node_use,
                               const orx_params* P) {
    for (int64_t a = 0; a < n_arcs; ++a) {
        int32_t o = occ[a] - P->edge_cap;
        if (o < 0) o = 0;
        float inc = P->hist_gain * (float)o;
        if (P->hist_cap_mult > 0.0f) {
            float capv = P->hist_cap_mult * base[a];
            if (capv < inc) inc = capv;
        }
        float hd = hist[a] * P->hist_decay;
        hist[a] = hd + inc;
    }
    for (int64_t v = 0; v < n_nodes; ++v) {
        int32_t o = node_use[v] - P->node_cap;
        if (o < 0) o = 0;
        float inc = P->node_hist_gain * (float)o;
        float hd = node_hist[v] * P->hist_decay;
        node_hist[v] = hd + inc;
    }
}

/* Entering-arc weights of the whole lattice: w = fl(arc_cost + node_cost(head)). */
void orx_oracle_lattice_weights(int32_t X, int32_t Y, int32_t L, uint32_t hmask, const float* base,
                                const float* hist, const int32_t* occ, const float* node_hist,
                                const int32_t* node_use, const orx_params* P, float* wout) {
    lat_t g = { X, Y, L, hmask };
    lat_state S = { base, hist, occ, node_hist, node_use, NULL };
    int64_t N = (int64_t)X * Y * L;
    for (int64_t u = 0; u < N; ++u)
        for (int s = 0; s < 4; ++s) {
            int64_t v = lat_nbr(&g, u, s);
            wout[4 * u + s] = v < 0 ? INFINITY : arc_weight(&g, &S, P, u, s, v);
        }
}

/* Per-net forward-arc overuse of the committed path and its hop count (the
 * congestion term of _order_nets_by_difficulty, unified_pathfinder.py:7858-7863). */
void orx_oracle_net_overuse(int32_t X, int32_t Y, int32_t L, uint32_t hmask, const int32_t* occ, int32_t edge_cap,
                            int32_t n_nets, const int32_t* path_arena, const int64_t* path_off,
                            const int32_t* path_len, int64_t* over_out, int32_t* hops_out) {
    lat_t g = { X, Y, L, hmask };
    for (int32_t k = 0; k < n_nets; ++k) {
        const int32_t* p = path_arena + path_off[k];
        int64_t over = 0;
        for (int32_t i = 0; i + 1 < path_len[k]; ++i) {
            int s = lat_slot(&g, p[i], p[i + 1]);
            int32_t o = occ[4 * (int64_t)p[i] + s] - edge_cap;
            if (o > 0) over += o;
        }
        over_out[k] = over;
        hops_out[k] = path_len[k] > 0 ? path_len[k] - 1 : 0;
    }
}

/* ------------------------------------------------------------------------ */
/* Plain PathFinder driver over box ROIs.                                    */
/* For each iteration it: for each net in order[it]: rip up its old path,    */
/* route it in its ROI with live costs, commit.  Then overuse stats and the  */
/* history update.  pres[it] is the present factor of iteration it.          */
/* n_order: entries of order[] per iteration, the nets routed in it (a     */
/* hotset routes a subset); < 0 means n_nets.                                */
/* stats per iteration (16 int64): 0 routed, 1 failed, 2 edge_over,          */
/* 3 edge_over_count, 4 node_over, 5 node_over_count, 6 via_hops, 7 barrels, */
/* 8 settled nodes, 9 relaxed arcs, 10 total path hops.                      */
/* dist_out[n] receives each net's final path cost (fp32) of the last        */
/* iteration it was routed in.                                               */
/* ------------------------------------------------------------------------ */
/* The per-net body of _route_all (UP:7949-8430): rip up the old path,
 * search, backtrace, commit.  Touches only arcs and nodes inside the net's
 * box.  st[0] routed, st[1] failed, st[8..9] settled / relaxed, st[10] path
 * hops.  Returns 0 or a negative status. */
static int route_one(const lat_t* g, const lat_state* S, const orx_params* P, int32_t k,
                     const int32_t* src, const int32_t* dst, const int32_t* box, const orx_terms* T,
                     int32_t* occ, int32_t* node_use, int32_t* via_col, int32_t* path_arena,
                     const int64_t* path_off, const int64_t* path_cap, int32_t* path_len, float* dist_out,
                     sssp_ws* W, int64_t* st) {
    int32_t* pk = path_arena + path_off[k];
    if (path_len[k] > 0) {
        if (apply_path(g, pk, path_len[k], -1, occ, node_use, via_col)) return -7;
        path_len[k] = 0;
    }
    const int32_t* b = box + 4 * (int64_t)k;
    if (T && T->seed_ptr && T->seed_ptr[k + 1] > T->seed_ptr[k]) {
        const int32_t s0 = T->seed_ptr[k], ns = T->seed_ptr[k + 1] - s0;
        const int32_t t0 = T->tgt_ptr ? T->tgt_ptr[k] : 0;
        const int32_t nt = T->tgt_ptr ? T->tgt_ptr[k + 1] - t0 : 0;
        if (nt < 1 || nt > MAX_TARGETS) return -3;
        int64_t best;
        float total;
        int r = box_route_multi(g, S, P, b[0], b[1], b[2], b[3], ns, T->seed_node + s0, T->seed_cost + s0, nt,
                                T->tgt_node + t0, T->tgt_cost ? T->tgt_cost + t0 : NULL, W, st + 8, &best,
                                &total);
        if (r < 0) return r;
        int64_t len = r == 1 ? 0 : box_backtrace(g, b[0], b[1], b[2], b[3], best, W, pk, path_cap[k]);
        if (len < 0) return (int)len;
        if (len < 2) { st[1] += 1; return 0; }
        path_len[k] = (int32_t)len;
        if (dist_out) dist_out[k] = total;
        if (apply_path(g, pk, len, +1, occ, node_use, via_col)) return -7;
        st[0] += 1;
        st[10] += len - 1;
        return 0;
    }
    int r = box_route(g, S, P, b[0], b[1], b[2], b[3], src[k], dst[k], W, st + 8);
    if (r < 0) return r;
    if (r == 1) { st[1] += 1; return 0; }
    int64_t len = box_backtrace(g, b[0], b[1], b[2], b[3], dst[k], W, pk, path_cap[k]);
    if (len < 0) return (int)len;
    path_len[k] = (int32_t)len;
    if (dist_out) dist_out[k] = kdist(W->key[box_index(g, b[0], b[1], b[2], b[3], dst[k])]);
    if (apply_path(g, pk, len, +1, occ, node_use, via_col)) return -7;
    st[0] += 1;
    st[10] += len - 1;
    return 0;
}

int orx_oracle_pathfinder(int32_t X, int32_t Y, int32_t L, uint32_t hmask, const float* base,
                          int32_t n_nets, const int32_t* src, const int32_t* dst, const int32_t* box,
                          int32_t iters, const int32_t* order, const float* pres,
                          const orx_params* P0,
                          float* hist, int32_t* occ, float* node_hist, int32_t* node_use,
                          int32_t* via_col, int32_t* path_arena, const int64_t* path_off,
                          const int64_t* path_cap, int32_t* path_len, float* dist_out,
                          int64_t* stats, int64_t n_order, const orx_terms* terms) {
    lat_t g = { X, Y, L, hmask };
    int64_t N = (int64_t)X * Y * L;
    sssp_ws W; memset(&W, 0, sizeof(W));
    orx_params P = *P0;
    int rc = 0;
    const int64_t per = (n_order < 0 || n_order > n_nets) ? n_nets : n_order;
    for (int32_t it = 0; it < iters; ++it) {
        int64_t* st = stats + (int64_t)it * 16;
        memset(st, 0, 16 * sizeof(int64_t));
        P.pres_fac = pres[it];
        P.node_pres = P0->node_pres * pres[it];
        lat_state S = { base, hist, occ, node_hist, node_use, NULL };
        for (int64_t i = 0; i < per; ++i) {
            int32_t k = order[(int64_t)it * per + i];
            int r = route_one(&g, &S, &P, k, src, dst, box, terms, occ, node_use, via_col, path_arena, path_off,
                              path_cap, path_len, dist_out, &W, st);
            if (r < 0) { rc = r; goto done; }
        }
        {
            int64_t ov[6];
            orx_oracle_overuse(X, Y, L, occ, node_use, via_col, P.edge_cap, P.node_cap, ov);
            for (int j = 0; j < 6; ++j) st[2 + j] = ov[j];
        }
        orx_oracle_update_history(4 * N, base, hist, occ, N, node_hist, node_use, &P);
    }
done:
    ws_free(&W);
    return rc;
}

/* Single box-ROI query on precomputed weights wcost[4N] (SSSP microbench).
 * full != 0 sweeps the whole ROI (no early exit).  dist_out / parent_out are
 * ROI-local arrays of w*h*L (parent as a global id, -1 for unreached and the
 * source).  hops_out optional.  stats[0] settled, stats[1] relaxed arcs. */
int orx_oracle_sssp_box(int32_t X, int32_t Y, int32_t L, uint32_t hmask, const float* wcost,
                        int32_t x0, int32_t y0, int32_t w, int32_t h, int64_t src, int64_t dst,
                        int32_t full, float* dist_out, int32_t* parent_out, int32_t* hops_out,
                        int64_t* stats) {
    lat_t g = { X, Y, L, hmask };
    lat_state S = { NULL, NULL, NULL, NULL, NULL, wcost };
    orx_params P; memset(&P, 0, sizeof(P));
    sssp_ws W; memset(&W, 0, sizeof(W));
    stats[0] = stats[1] = 0;
    int r = box_sssp(&g, &S, &P, x0, y0, w, h, src, full ? -1 : dst, full, &W, stats, NULL);
    if (r >= 0) {
        int64_t R = (int64_t)w * h * L, XY = (int64_t)X * Y, wh = (int64_t)w * h;
        for (int64_t i = 0; i < R; ++i) {
            uint64_t k = W.key[i];
            if (dist_out) dist_out[i] = k == KEY_INF ? INFINITY : kdist(k);
            if (hops_out) hops_out[i] = k == KEY_INF ? -1 : (int32_t)khops(k);
            if (parent_out) {
                if (k == KEY_INF || kslot(k) == ROOT_SLOT) parent_out[i] = -1;
                else {
                    int32_t z = (int32_t)(i / wh);
                    int64_t rm = i - z * wh;
                    int32_t ly = (int32_t)(rm / w), lx = (int32_t)(rm - (int64_t)ly * w);
                    int64_t gid = (int64_t)z * XY + (int64_t)(y0 + ly) * X + (x0 + lx);
                    parent_out[i] = (int32_t)lat_nbr(&g, gid, (int)kslot(k));
                }
            }
        }
        if (dst >= 0 && full) {
            /* report reachability of dst in full mode */
        }
    }
    ws_free(&W);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Generic CSR query (the drop-in seam find_path_roi_gpu /                   */
/* find_path_fullgraph_gpu_seeds, cuda_dijkstra.py:4951-5040, 5516-5850).    */
/* roi: g2r[N] maps a global id to an ROI index or -1 (NULL = whole graph,   */
/* r = global id).  pen: node penalty indexed by ROI index (NULL = none).    */
/* in_rank[e]: rank of the arc's source among the in-neighbours of its       */
/* target sorted by (source id, arc id).  Seeds start at seed_cost, plus    */
/* pen(seed) when seed_pen != 0 (multi-seed forms, unified_pathfinder.py:    */
/* 2582-2583, cuda_dijkstra.py:5589-5606; the single-source find_path_roi   */
/* starts at 0, UP:2497); targets are not expanded; the chosen target       */
/* minimises (fl(d + target_cost), node id).  allowed (bitmap over global    */
/* ids, NULL = all).  parent_out[R]: predecessor ROI index of each settled   */
/* node (-1 root/unreached); dist_out[R]; best_out[0] = chosen target global */
/* id, -1 if none.  Returns 0 found, 1 none, <0 error.                       */
/* ------------------------------------------------------------------------ */
int orx_oracle_sssp_csr(int64_t N, const int32_t* indptr, const int32_t* indices, const float* costs,
                        const int32_t* in_rank, const int32_t* in_indptr, const int32_t* in_src,
                        int64_t R, const int32_t* roi_nodes, const int32_t* g2r, const float* pen,
                        const uint32_t* allowed, int32_t seed_pen,
                        int32_t n_seeds, const int32_t* seeds, const float* seed_cost,
                        int32_t n_tg, const int32_t* targets, const float* tg_cost,
                        float* dist_out, int32_t* parent_out, int32_t* best_out, int64_t* stats) {
    (void)N;
    uint64_t* key = (uint64_t*)malloc((size_t)R * 8);
    uint8_t* done = (uint8_t*)calloc((size_t)R, 1);
    uint8_t* is_tg = (uint8_t*)calloc((size_t)R, 1);
    heap_t H; memset(&H, 0, sizeof(H));
    int rc = 0;
    if (!key || !done || !is_tg) { rc = -2; goto out; }
    for (int64_t r = 0; r < R; ++r) key[r] = KEY_INF;
#define MAPR(gid) (g2r ? (int64_t)g2r[(gid)] : (int64_t)(gid))
#define ALLOWED(gid) (!allowed || ((allowed[(gid) >> 5] >> ((gid) & 31)) & 1u))
    for (int32_t i = 0; i < n_tg; ++i) {
        int64_t r = MAPR(targets[i]);
        if (r >= 0) is_tg[r] = 1;
    }
    for (int32_t i = 0; i < n_seeds; ++i) {
        int64_t r = MAPR(seeds[i]);
        if (r < 0 || !ALLOWED(seeds[i])) continue;
        float c = seed_cost ? seed_cost[i] : 0.0f;
        if (pen && seed_pen) c = c + pen[r];
        uint64_t k = kpack(c, 0, ROOT_SLOT);
        if (k < key[r]) { key[r] = k; if (heap_push(&H, (double)c, k, (int32_t)r)) { rc = -2; goto out; } }
    }
    int64_t remaining = 0;
    for (int64_t r = 0; r < R; ++r) remaining += is_tg[r];
    while (H.n) {
        hent e = heap_pop(&H);
        if (done[e.v] || e.k != key[e.v]) continue;
        done[e.v] = 1;
        stats[0] += 1;
        if (is_tg[e.v]) {
            if (--remaining == 0) break;
            continue;
        }
        int64_t gu = roi_nodes ? roi_nodes[e.v] : e.v;
        float du = kdist(e.k);
        uint32_t hu = khops(e.k);
        for (int32_t a = indptr[gu]; a < indptr[gu + 1]; ++a) {
            int64_t gv = indices[a];
            int64_t rv = MAPR(gv);
            if (rv < 0 || !ALLOWED(gv)) continue;
            stats[1] += 1;
            if (done[rv]) continue;
            float wt = costs[a];
            if (pen) wt = wt + pen[rv];
            if (isinf(wt)) continue;
            float nd = du + wt;
            if (hu + 1 > HOPS_MAX || in_rank[a] > 254) { rc = -4; goto out; }
            uint64_t nk = kpack(nd, hu + 1, (uint32_t)in_rank[a]);
            if (nk < key[rv]) { key[rv] = nk; if (heap_push(&H, (double)nd, nk, (int32_t)rv)) { rc = -2; goto out; } }
        }
    }
    {
        int64_t best = -1; float bestc = INFINITY; int64_t bestg = -1;
        for (int32_t i = 0; i < n_tg; ++i) {
            int64_t r = MAPR(targets[i]);
            if (r < 0 || key[r] == KEY_INF) continue;
            float t = kdist(key[r]) + (tg_cost ? tg_cost[i] : 0.0f);
            if (t < bestc || (t == bestc && targets[i] < bestg)) { bestc = t; best = r; bestg = targets[i]; }
        }
        best_out[0] = (int32_t)bestg;
        if (best < 0) rc = 1;
    }
    for (int64_t r = 0; r < R; ++r) {
        uint64_t k = key[r];
        if (dist_out) dist_out[r] = k == KEY_INF ? INFINITY : kdist(k);
        if (parent_out) {
            if (k == KEY_INF || kslot(k) == ROOT_SLOT) parent_out[r] = -1;
            else {
                int64_t gv = roi_nodes ? roi_nodes[r] : r;
                int64_t gp = in_src[in_indptr[gv] + (int32_t)kslot(k)];
                parent_out[r] = (int32_t)MAPR(gp);
            }
        }
    }
#undef MAPR
#undef ALLOWED
out:
    free(key); free(done); free(is_tg); free(H.a);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* One PathFinder iteration with the nets of each disjoint-ROI wave routed   */
/* on host threads (OpenMP): the all-core CPU baseline.  sched[n] lists the  */
/* nets wave by wave, wave_ptr[n_waves+1] the wave boundaries.  Nets of one  */
/* wave touch disjoint arcs and nodes, so the result equals the sequential   */
/* order's.  stats[16] as one iteration of orx_oracle_pathfinder.            */
/* ------------------------------------------------------------------------ */
int orx_oracle_pathfinder_waves(int32_t X, int32_t Y, int32_t L, uint32_t hmask, const float* base,
                                int32_t n_nets, const int32_t* src, const int32_t* dst, const int32_t* box,
                                const int32_t* sched, const int32_t* wave_ptr, int32_t n_waves, float pres_it,
                                const orx_params* P0, float* hist, int32_t* occ, float* node_hist,
                                int32_t* node_use, int32_t* via_col, int32_t* path_arena,
                                const int64_t* path_off, const int64_t* path_cap, int32_t* path_len,
                                float* dist_out, int64_t* stats, int32_t threads, const orx_terms* terms) {
    lat_t g = { X, Y, L, hmask };
    int64_t N = (int64_t)X * Y * L;
    orx_params P = *P0;
    P.pres_fac = pres_it;
    P.node_pres = P0->node_pres * pres_it;
    lat_state S = { base, hist, occ, node_hist, node_use, NULL };
    int T = threads > 0 ? threads : omp_get_max_threads();
    sssp_ws* W = (sssp_ws*)calloc((size_t)T, sizeof(sssp_ws));
    int64_t* tst = (int64_t*)calloc((size_t)T * 16, sizeof(int64_t));
    if (!W || !tst) { free(W); free(tst); return -8; }
    int rc = 0;
    (void)n_nets;
    for (int32_t w = 0; w < n_waves && rc == 0; ++w) {
        const int32_t a = wave_ptr[w], b = wave_ptr[w + 1];
#pragma omp parallel for schedule(dynamic, 1) num_threads(T)
        for (int32_t i = a; i < b; ++i) {
            const int t = omp_get_thread_num();
            int r = route_one(&g, &S, &P, sched[i], src, dst, box, terms, occ, node_use, via_col, path_arena,
                              path_off, path_cap, path_len, dist_out, &W[t], tst + 16 * (int64_t)t);
            if (r < 0) {
#pragma omp atomic write
                rc = r;
            }
        }
    }
    memset(stats, 0, 16 * sizeof(int64_t));
    for (int t = 0; t < T; ++t) {
        for (int j = 0; j < 16; ++j) stats[j] += tst[16 * (int64_t)t + j];
        ws_free(&W[t]);
    }
    free(W);
    free(tst);
    if (rc < 0) return rc;
    int64_t ov[6];
    orx_oracle_overuse(X, Y, L, occ, node_use, via_col, P.edge_cap, P.node_cap, ov);
    for (int j = 0; j < 6; ++j) stats[2 + j] = ov[j];
    orx_oracle_update_history(4 * N, base, hist, occ, N, node_hist, node_use, &P);
    return 0;
}
