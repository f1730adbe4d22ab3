// orx_lattice.hip -- lattice PathFinder kernels for gfx950.
//
// route_wave_kernel: one workgroup per net of a disjoint-ROI wave.  Each
// workgroup runs the whole per-net body of the reference's _route_all
// (unified_pathfinder.py:7949-8430) on its ROI box:
//   rip-up of the old path (EdgeAccountant.clear_path UP:1118-1135 and
//   _clear_path_node_use UP:4648-4659), live re-price of the touched arcs
//   (refresh_live_present_costs UP:1403-1422), delta-stepping SSSP over the
//   ROI (replaces the wavefront search CD:2800-3533 and the heap Dijkstra
//   UP:2451-2548), parent backtrace (CD:1845-1950), commit (UP:1091-1116,
//   4635-4646, via columns UP:4470-4546) and re-price again.
// Nets of one wave have pairwise disjoint ROI boxes, so their commits touch
// disjoint arcs and the wave is equivalent to routing them sequentially.
//
// SSSP: near-far delta-stepping.  Keys live in LDS when the ROI fits
// (lds_nodes), else in the global workspace.  The near frontier is
// deduplicated with a per-round bitmap and appended with wave ballots + a
// prefix popcount (one LDS atomic per wave); the far pile is deduplicated
// with a third bitmap and re-split when the near bucket [.., T) drains:
// T <- min(far) + delta.  With early exit the search stops once the sink's
// distance is below T (all keys below T are final).
#include "orx_route.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include "../../include/orthoroute_hip.h"

using namespace orx;

namespace {

// ---------------------------------------------------------------------------
// Near-far delta-stepping over one ROI box, one workgroup, bucketed by
// f = g + h (h = 0 without a heuristic).
// ---------------------------------------------------------------------------
// Entries of the LDS ring that holds the first part of a round's near queue.
constexpr uint32_t kRing = 2048;
// Internal route flag (above the public ORX_F_* bits): the launch's global
// workspace range was cleared by ws_fill_kernel just before it.
constexpr int32_t kPrefilled = 1 << 16;

__device__ __forceinline__ uint32_t ring_get(const uint32_t* lq, const uint32_t* gq, uint32_t i) {
    return i < kRing ? lq[i] : gq[i];
}
__device__ __forceinline__ void ring_put(uint32_t* lq, uint32_t* gq, uint32_t i, uint32_t v) {
    if (i < kRing) lq[i] = v;
    else gq[i] = v;
}

// Seeds of a multi-terminal net: every seed inside the box gets the key
// (fl(cost + q_seed), 0 hops, root), duplicates keeping the minimum; the
// first threshold is delta above the smallest seed f, and the seeds below
// it form round 0's near queue, the others the far pile.
template <int BLOCK>
__device__ double seed_init(const Lat& g, const Box& B, const Multi& M, float delta, const Heur& H, uint64_t* keys,
                            uint32_t* bmn0, uint32_t* bmf, uint32_t* qA, uint32_t* lqA, uint32_t* fA, Ctl* ctl) {
    const uint32_t tid = threadIdx.x;
    for (int32_t i = (int32_t)tid; i < M.ns; i += BLOCK) {
        const int64_t gs = M.seed_node[i];
        const uint32_t rs = global_to_roi(g, B, gs);
        if (rs == UINT32_MAX) continue;
        const float q = node_cost(g.node_hist[gs],
                                  __hip_atomic_load(g.node_use + gs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), M.P);
        key_min(keys + rs, kpack(__fadd_rn(M.seed_cost[i], q), 0u, ORX_ROOT_SLOT));
    }
    __syncthreads();
    double lmin = INFINITY;
    for (int32_t i = (int32_t)tid; i < M.ns; i += BLOCK) {
        const uint32_t rs = global_to_roi(g, B, M.seed_node[i]);
        if (rs == UINT32_MAX) continue;
        uint32_t lx, ly, z;
        roi_decode(B, rs, lx, ly, z);
        lmin = fmin(lmin, (double)kdist(key_load(keys + rs)) + box_h(H, B, lx, ly, z));
    }
    for (int o = 32; o > 0; o >>= 1) lmin = fmin(lmin, __shfl_xor(lmin, o));
    if (lane_id() == 0 && lmin < (double)INFINITY) atomicMin(&ctl->minf, dbits(lmin));
    __syncthreads();
    const double f0 = __longlong_as_double((long long)ctl->minf);
    double T = f0 + (double)delta;
    if (!(T > f0)) T = nextafter(f0, (double)INFINITY);
    for (int32_t base = 0; base < M.ns; base += BLOCK) {
        const int32_t i = base + (int32_t)tid;
        bool near = false, far = false;
        uint32_t rs = UINT32_MAX;
        if (i < M.ns) rs = global_to_roi(g, B, M.seed_node[i]);
        if (rs != UINT32_MAX) {
            uint32_t lx, ly, z;
            roi_decode(B, rs, lx, ly, z);
            const double f = (double)kdist(key_load(keys + rs)) + box_h(H, B, lx, ly, z);
            if (f < T) near = bit_set(bmn0, rs) == 0;
            else far = bit_set(bmf, rs) == 0;
        }
        int at = wave_append(near, &ctl->qcnt[0]);
        if (near) ring_put(lqA, qA, (uint32_t)at, rs);
        at = wave_append(far, &ctl->fcnt);
        if (far) fA[at] = rs;
    }
    __syncthreads();
    return T;
}

// skip_keys / skip_bitmaps: those arrays were cleared before the launch
// (global workspace ranges filled chip-wide by ws_fill_kernel).
// multi == false: one source rsrc at key 0 and one sink rdst.  Otherwise the
// seeds and targets of M (rsrc / rdst unused); the search then stops once
// the best certified target cannot be beaten (see the stop test below).
template <int BLOCK>
__device__ void box_sssp(const Lat& g, const Box& B, uint32_t rsrc, uint32_t rdst, bool full, float delta,
                         const Heur& H, uint64_t* keys, uint32_t* bmn0, uint32_t* bmn1, uint32_t* bmf, uint32_t* qA,
                         uint32_t* qB, uint32_t* lqA, uint32_t* lqB, uint32_t* fA, uint32_t* fB, bool skip_keys,
                         bool skip_bitmaps, Ctl* ctl, bool multi, const Multi& M) {
    const uint32_t tid = threadIdx.x;
    const uint32_t R = B.R;
    const uint32_t nwords = (R + 31) / 32;
    if (!skip_keys)
        for (uint32_t r = tid; r < R; r += BLOCK) keys[r] = ORX_KEY_INF;
    if (!skip_bitmaps)
        for (uint32_t i = tid; i < nwords; i += BLOCK) { bmn0[i] = 0; bmn1[i] = 0; bmf[i] = 0; }
    if (tid == 0) {
        ctl->qcnt[0] = multi ? 0u : 1u; ctl->qcnt[1] = 0; ctl->qcnt[2] = 0;
        ctl->fcnt = 0; ctl->fcnt2 = 0; ctl->err = 0;
        ctl->tbest[0] = ctl->tbest[1] = ~0ull;
        ctl->tlb[0] = ctl->tlb[1] = 0x7F800000u;
        ctl->minf = dbits((double)INFINITY);
    }
    __syncthreads();
    double T;
    if (multi) {
        T = seed_init<BLOCK>(g, B, M, delta, H, keys, bmn0, bmf, qA, lqA, fA, ctl);
    } else {
        uint32_t lx, ly, z;
        roi_decode(B, rsrc, lx, ly, z);
        const double f0 = box_h(H, B, lx, ly, z);
        T = f0 + (double)delta;
        if (!(T > f0)) T = nextafter(f0, (double)INFINITY);
        if (tid == 0) {
            keys[rsrc] = kpack(0.0f, 0u, ORX_ROOT_SLOT);
            lqA[0] = rsrc;
            bit_set(bmn0, rsrc);
        }
        __syncthreads();
    }

    uint32_t r = 0;     // round counter
    uint32_t chk = 0;   // target checks (multi-terminal stop test)
    for (;;) {
        // ---------------- near phase: relax until the bucket [.., T) drains
        for (;;) {
            const uint32_t n = ctl->qcnt[r % 3];
            if (n == 0) break;
            const uint32_t* cur = (r & 1) ? qB : qA;
            uint32_t* nxt = (r & 1) ? qA : qB;
            const uint32_t* lcur = (r & 1) ? lqB : lqA;
            uint32_t* lnxt = (r & 1) ? lqA : lqB;
            uint32_t* bmc = (r & 1) ? bmn1 : bmn0;
            uint32_t* bmx = (r & 1) ? bmn0 : bmn1;
            uint32_t* cnx = &ctl->qcnt[(r + 1) % 3];
            if (tid == 0) { ctl->qcnt[(r + 2) % 3] = 0; ctl->rounds++; ctl->items += n; }
            for (uint32_t base = 0; base < n; base += BLOCK) {
                const uint32_t i = base + tid;
                const bool active = i < n;
                // The memory operations of a lane's four arcs are issued as
                // four independent batches (loads, then key atomics, then
                // bitmap atomics) so each step waits for memory once, not
                // once per arc.
                uint32_t u = 0, lx = 0, ly = 0, z = 0;
                bool horiz = false;
                uint64_t ku = ORX_KEY_INF;
                float w4[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
                uint32_t nb[4] = {UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX};
                if (active) {
                    u = ring_get(lcur, cur, i);
                    bit_clear(bmc, u);
                    ku = key_load(keys + u);
                    roi_decode(B, u, lx, ly, z);
                    horiz = (g.hmask >> z) & 1u;
                    if (!(multi && is_target(M.tt, u, lx, ly, z))) {   // targets are sinks
                        const int64_t gu = roi_global(g, B, lx, ly, z);
                        const float4 wv = reinterpret_cast<const float4*>(g.wcost)[gu];
                        w4[0] = wv.x; w4[1] = wv.y; w4[2] = wv.z; w4[3] = wv.w;
#pragma unroll
                        for (int s = 0; s < 4; ++s) nb[s] = roi_nbr(B, horiz, u, lx, ly, z, s);
                    }
                }
                const float du = kdist(ku);
                const uint32_t hu = khops(ku);
                float nd[4];
                uint64_t nk[4], old[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    nd[s] = __fadd_rn(du, w4[s]);
                    nk[s] = kpack(nd[s], hu + 1, 3u - (uint32_t)s);
                    old[s] = 0;
                    if (nb[s] != UINT32_MAX && !isinf(w4[s])) old[s] = key_min(keys + nb[s], nk[s]);
                }
                bool to_near[4], to_far[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    to_near[s] = false;
                    to_far[s] = false;
                    if (nk[s] < old[s]) {
                        const double nf = (double)nd[s] + nbr_h(H, B, horiz, lx, ly, z, s);
                        if (nf < T) to_near[s] = true;
                        else to_far[s] = true;
                    }
                }
                uint32_t was[4];
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    was[s] = 1;
                    if (to_near[s]) was[s] = bit_set(bmx, nb[s]);
                    else if (to_far[s]) was[s] = bit_set(bmf, nb[s]);
                }
                uint32_t cn = 0, cf = 0;
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    cn += (to_near[s] && was[s] == 0) ? 1u : 0u;
                    cf += (to_far[s] && was[s] == 0) ? 1u : 0u;
                }
                uint32_t an = wave_append_n(cn, cnx);
                uint32_t af = wave_append_n(cf, &ctl->fcnt);
#pragma unroll
                for (int s = 0; s < 4; ++s) {
                    if (to_near[s] && was[s] == 0) ring_put(lnxt, nxt, an++, nb[s]);
                    if (to_far[s] && was[s] == 0) fA[af++] = nb[s];
                }
            }
            ++r;
            __syncthreads();
        }
        // ---------------- every key with f below T is final
        if (!full && multi) {
            // A target with dist < T' = min(T, gmax) (T alone without the
            // heuristic) is certified: final and exact (h = 0 on targets).
            // Every other target has a true dist >= T', so its total is at
            // least fl(T'^ + cost), T'^ the smallest fp32 >= T'.  Stop when
            // that bound exceeds the best certified total for all of them.
            // Checks alternate between two slot pairs: the pair read now is
            // reset after the next check's barrier, when every lane has read it.
            const uint32_t par = chk++ & 1u;
            double Tc = T;
            if (H.on && Tc > (double)M.P.astar_gmax) Tc = (double)M.P.astar_gmax;
            float Tu = (float)Tc;
            if ((double)Tu < Tc) Tu = nextafterf(Tu, INFINITY);
            if ((int32_t)tid < M.nt) {
                const int64_t gid = M.tgt_node[tid];
                const uint32_t rt = global_to_roi(g, B, gid);
                if (rt != UINT32_MAX) {
                    const uint64_t kt = key_load(keys + rt);
                    const float c = M.tgt_cost ? M.tgt_cost[tid] : 0.0f;
                    if (kt != ORX_KEY_INF && (double)kdist(kt) < Tc) {
                        const float tot = __fadd_rn(kdist(kt), c);
                        if (tot < INFINITY)
                            atomicMin(&ctl->tbest[par],
                                      ((unsigned long long)__float_as_uint(tot) << 32) | (uint32_t)gid);
                    } else {
                        atomicMin(&ctl->tlb[par], __float_as_uint(__fadd_rn(Tu, c)));
                    }
                }
            }
            __syncthreads();
            const unsigned long long bb = ctl->tbest[par];
            const uint32_t lb = ctl->tlb[par];
            if (tid == 0) { ctl->tbest[par ^ 1u] = ~0ull; ctl->tlb[par ^ 1u] = 0x7F800000u; }
            if (bb != ~0ull && lb > (uint32_t)(bb >> 32)) break;
        } else if (!full) {
            const uint64_t kd = key_load(keys + rdst);
            if (kd != ORX_KEY_INF && (double)kdist(kd) < T) break;  // h(dst) = 0
        }
        const uint32_t nf = ctl->fcnt;
        if (nf == 0) break;
        // split pass 1: drop settled far entries, compact live ones, find
        // min f and a log2 histogram of how far above T the live ones are
        if (tid == 0) { ctl->fcnt2 = 0; ctl->minf = dbits((double)INFINITY); ctl->phases++; }
        if (tid < 32) ctl->fhist[tid] = 0;
        __syncthreads();
        double lmin = INFINITY;
        for (uint32_t base = 0; base < nf; base += BLOCK) {
            const uint32_t i = base + tid;
            const bool active = i < nf;
            bool keep = false;
            uint32_t v = 0;
            if (active) {
                v = fA[i];
                uint32_t lx, ly, z;
                roi_decode(B, v, lx, ly, z);
                const double f = (double)kdist(key_load(keys + v)) + box_h(H, B, lx, ly, z);
                if (f < T) bit_clear(bmf, v);
                else {
                    keep = true;
                    lmin = fmin(lmin, f);
                    const double y = 1.0 + (f - T) / (double)delta;
                    int b = (y < 2.0) ? 0 : ilogb(y);
                    b = b > 31 ? 31 : b;
                    atomicAdd(&ctl->fhist[b], 1u);
                }
            }
            const int at = wave_append(keep, &ctl->fcnt2);
            if (keep) fB[at] = v;
        }
        for (int o = 32; o > 0; o >>= 1) lmin = fmin(lmin, __shfl_xor(lmin, o));
        if (lane_id() == 0 && lmin < (double)INFINITY) atomicMin(&ctl->minf, dbits(lmin));
        __syncthreads();
        const uint32_t nf2 = ctl->fcnt2;
        if (nf2 == 0) break;
        const double mn = __longlong_as_double((long long)ctl->minf);
        // Next threshold: at least delta above the minimum, and wide enough
        // that the next near phase gets about two workgroups' worth of far
        // entries.  With costs spread far beyond delta (a congested board
        // late in the negotiation) a fixed delta would run a phase per few
        // nodes and rescan the whole far pile each time.  Any threshold
        // above the minimum reaches the same fixpoint.
        double Tn = mn + (double)delta;
        if (!isinf(delta)) {
            const uint32_t target = nf2 < 2u * BLOCK ? nf2 : 2u * BLOCK;
            uint32_t c = 0;
            int b = 31;
            for (int i = 0; i < 32; ++i) {
                c += ctl->fhist[i];
                if (c >= target) { b = i; break; }
            }
            const double edge = T + (double)delta * (ldexp(1.0, b + 1) - 1.0);
            if (edge > Tn) Tn = edge;
        }
        if (!(Tn > mn)) Tn = nextafter(mn, (double)INFINITY);
        if (tid == 0) ctl->fcnt = 0;
        __syncthreads();
        // split pass 2: entries below the new threshold form round r's queue
        uint32_t* qr = (r & 1) ? qB : qA;
        uint32_t* lqr = (r & 1) ? lqB : lqA;
        uint32_t* bmr = (r & 1) ? bmn1 : bmn0;
        for (uint32_t base = 0; base < nf2; base += BLOCK) {
            const uint32_t i = base + tid;
            const bool active = i < nf2;
            bool to_near = false, stay = false;
            uint32_t v = 0;
            if (active) {
                v = fB[i];
                uint32_t lx, ly, z;
                roi_decode(B, v, lx, ly, z);
                const double f = (double)kdist(key_load(keys + v)) + box_h(H, B, lx, ly, z);
                if (f < Tn) { to_near = true; bit_clear(bmf, v); bit_set(bmr, v); }
                else stay = true;
            }
            int at = wave_append(to_near, &ctl->qcnt[r % 3]);
            if (to_near) ring_put(lqr, qr, (uint32_t)at, v);
            at = wave_append(stay, &ctl->fcnt);
            if (stay) fA[at] = v;
        }
        T = Tn;
        __syncthreads();
    }
    __syncthreads();
}

// Search-state pointers of one wave slot.  The dynamic LDS of a search
// workgroup is laid out as
//   [ near ring A: kRing u32 ][ near ring B: kRing u32 ][ region ]
// The near queue of a round lives in the ring while it has at most kRing
// entries and spills to the global queue beyond that.  The region holds the
// keys and the three bitmaps when the ROI has at most lds_nodes nodes and
// they fit, else only the bitmaps when those fit; the rest stays in the
// global workspace.
struct SlotState {
    uint64_t* keys;
    uint32_t *bmn0, *bmn1, *bmf, *qA, *qB, *fA, *fB;
    uint32_t *lqA, *lqB;
    bool keys_global, bm_global;   // kept in the global workspace
};

__device__ __forceinline__ SlotState slot_state(const orx_workspace& ws, int64_t off, uint32_t R, uint32_t* lds,
                                                int32_t region_bytes, int32_t lds_nodes) {
    SlotState S;
    S.qA = ws.queues + off;
    S.qB = ws.queues + ws.ws_nodes + off;
    S.fA = ws.queues + 2 * ws.ws_nodes + off;
    S.fB = ws.queues + 3 * ws.ws_nodes + off;
    S.lqA = lds;
    S.lqB = lds + kRing;
    S.keys_global = false;
    S.bm_global = false;
    uint32_t* region = lds + 2 * kRing;
    const uint32_t words = (R + 31) / 32;
    const int64_t bm_bytes = 3ll * words * 4;
    const int64_t key_bytes = 8ll * R;
    if ((int64_t)R <= (int64_t)lds_nodes && key_bytes + bm_bytes <= (int64_t)region_bytes) {
        S.keys = reinterpret_cast<uint64_t*>(region);
        uint32_t* bm = reinterpret_cast<uint32_t*>(S.keys + R);
        S.bmn0 = bm; S.bmn1 = bm + words; S.bmf = bm + 2 * words;
    } else {
        S.keys = ws.keys + off;
        S.keys_global = true;
        if (bm_bytes <= (int64_t)region_bytes) {
            S.bmn0 = region; S.bmn1 = region + words; S.bmf = region + 2 * words;
        } else {
            S.bm_global = true;
            S.bmn0 = ws.bitmaps + off / 32;
            S.bmn1 = ws.bitmaps + ws.ws_nodes / 32 + off / 32;
            S.bmf = ws.bitmaps + 2 * (ws.ws_nodes / 32) + off / 32;
        }
    }
    return S;
}

// The per-net body of _route_all (unified_pathfinder.py:7949-8430), run by
// one workgroup: rip-up, search, backtrace, commit.  Ends with a barrier.
//
// A net with seeds in nets.seed_ptr is a multi-terminal net
// (find_path_multisource_multisink, UP:2550-2659, on the net's box): seeds
// priced with their cost plus their node term, targets not expanded, the
// best target by (fl(dist + target cost), node id); no path when no seed or
// target lies in the box, nothing is reached, or the best target is itself a
// seed (a one-node path, UP:2647-2648).  Other nets route src -> dst.
// Both forms share one search call site (one inlined copy of box_sssp).
template <int BLOCK>
__device__ void route_net(const Lat& g, const Params& P, const NetsDev& nets, int32_t k, const orx_workspace& ws,
                          int64_t ws_off, float delta, int32_t lds_nodes, int32_t lds_region, int32_t flags,
                          unsigned long long* stats, uint32_t* lds, Ctl& ctl, TermTab& TT) {
    const Box B = make_box(nets.box + 4 * (int64_t)k, g.L);
    const SlotState S = slot_state(ws, ws_off, B.R, lds, lds_region, lds_nodes);
    int32_t* path = nets.path_arena + nets.path_off[k];
    const int32_t cap = nets.path_cap[k];

    if (threadIdx.x == 0) { ctl.settled = 0; ctl.relaxed = 0; ctl.rounds = 0; ctl.phases = 0; ctl.items = 0; }
    // 1. rip up the previous path and re-price what it held
    const int32_t old_len = nets.path_len[k];
    if (old_len > 0) {
        apply_path_block<BLOCK>(g, path, old_len, -1);
        __syncthreads();
        reprice_path_block<BLOCK>(g, P, path, old_len);
        __syncthreads();
    }
    // 2. terminals
    const bool full = (flags & ORX_F_FULL) != 0;
    const bool multi = nets.seed_ptr && nets.seed_ptr[k + 1] > nets.seed_ptr[k];
    int32_t st = ORX_OK;
    uint32_t rsrc = 0, rdst = 0;
    const int32_t* tn = nullptr;
    const float* tc = nullptr;
    int32_t nt = 0;
    Multi M;
    Heur H;
    if (multi) {
        const int32_t s0 = nets.seed_ptr[k];
        const int32_t t0 = nets.tgt_ptr ? nets.tgt_ptr[k] : 0;
        nt = nets.tgt_ptr ? nets.tgt_ptr[k + 1] - t0 : 0;
        if (nt < 1 || nt > kMaxTargets || !nets.seed_node || !nets.seed_cost || !nets.tgt_node) {
            st = ORX_E_ARG;
        } else {
            tn = nets.tgt_node + t0;
            tc = nets.tgt_cost ? nets.tgt_cost + t0 : nullptr;
            M.tt = &TT; M.seed_node = nets.seed_node + s0; M.seed_cost = nets.seed_cost + s0;
            M.ns = nets.seed_ptr[k + 1] - s0; M.P = P;
            M.tgt_node = tn; M.tgt_cost = tc; M.nt = nt;
            build_terms<BLOCK>(g, B, tn, nt, M.seed_node, M.ns, TT);
            if (TT.n == 0 || TT.seeds_in == 0) st = ORX_NO_PATH;
            H = make_heur_box(P, B, TT, !full);
        }
        rsrc = UINT32_MAX;   // the walk may end at any seed
    } else {
        rsrc = global_to_roi(g, B, nets.src[k]);
        rdst = global_to_roi(g, B, nets.dst[k]);
        if (rsrc == UINT32_MAX || rdst == UINT32_MAX) st = ORX_E_ARG;
        H = make_heur(g, P, nets.dst[k], !full);
    }
    int32_t len = 0;
    float cost = INFINITY;
    if (st == ORX_OK) {
        // 3. search; redone without the heuristic when the sink's distance
        // (a multi-terminal net: the chosen total) is >= gmax, where the A*
        // margin no longer covers the rounding.  Below it the result is
        // exact: the true best total is no larger, so the true best target
        // lies below gmax and is certified.
        bool pre = (flags & kPrefilled) != 0;
        for (;;) {
            box_sssp<BLOCK>(g, B, multi ? 0u : rsrc, rdst, full, delta, H, S.keys, S.bmn0, S.bmn1, S.bmf, S.qA,
                            S.qB, S.lqA, S.lqB, S.fA, S.fB, pre && S.keys_global, pre && S.bm_global, &ctl,
                            multi, M);
            bool big;
            if (multi) {
                if (threadIdx.x < 64) pick_target(g, B, S.keys, tn, tc, nt, ctl);
                __syncthreads();
                big = ctl.best_r != UINT32_MAX && !(ctl.best_cost < P.astar_gmax);
            } else {
                const uint64_t kd = key_load(S.keys + rdst);
                big = kd != ORX_KEY_INF && !(kdist(kd) < P.astar_gmax);
            }
            if (!H.on || !big) break;
            H.on = false;
            pre = false;
        }
        bool all = full;
        if (multi) {
            rdst = ctl.best_r;
            all = full || rdst == UINT32_MAX;
        }
        if (flags & ORX_F_STATS) box_stats<BLOCK>(g, B, H, S.keys, rdst, all, &ctl, multi ? &TT : nullptr);
        // 4. backtrace
        if (rdst != UINT32_MAX && threadIdx.x < 64) {
            const int32_t l = box_backtrace_wave(g, B, S.keys, rsrc, rdst, path, cap);
            if (threadIdx.x == 0) ctl.path_len = l;
        }
        __syncthreads();
        const int32_t l = rdst == UINT32_MAX ? 0 : ctl.path_len;
        if (l < 0) st = l;
        else if (l == 0 || (multi && l == 1)) st = ORX_NO_PATH;
        else {
            len = l;
            cost = multi ? ctl.best_cost : kdist(key_load(S.keys + rdst));
        }
    }
    if (threadIdx.x == 0) {
        nets.status[k] = st;
        nets.path_len[k] = len;
        nets.path_cost[k] = cost;
    }
    // 5. commit and re-price
    if (len > 0) {
        apply_path_block<BLOCK>(g, path, len, +1);
        __syncthreads();
        reprice_path_block<BLOCK>(g, P, path, len);
    }
    if (stats && threadIdx.x == 0) {
        atomicAdd(stats + 0, ctl.settled);
        atomicAdd(stats + 1, ctl.relaxed);
        atomicAdd(stats + 2, (unsigned long long)ctl.items);
        atomicAdd(stats + 3, (unsigned long long)ctl.rounds);
        atomicAdd(stats + 4, (unsigned long long)ctl.phases);
        atomicAdd(stats + (len > 0 ? 5 : 6), 1ull);
        if (len > 0) atomicAdd(stats + 7, (unsigned long long)(len - 1));
    }
    __syncthreads();
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void route_wave_kernel(Lat g, Params P, NetsDev nets, const int32_t* wave_nets,
                                                           const int64_t* ws_off, orx_workspace ws, float delta,
                                                           int32_t lds_nodes, int32_t lds_region, int32_t flags,
                                                           unsigned long long* stats, int64_t ws_limit,
                                                           int64_t max_nodes) {
    ORX_DYNAMIC_LDS(uint64_t, lds_mem);
    ORX_SHARED(Ctl, ctl);
    ORX_SHARED(TermTab, tt);
    // The plan's net id and workspace slice must fit: a net id below n, the
    // ROI no larger than the launch's largest (which sized the LDS plan and
    // the prefilled range), its slice 64-aligned inside [0, ws_limit).  A
    // slot that does not is refused (status ORX_E_ARG, counted as failed).
    const int32_t k = wave_nets[blockIdx.x];
    const int64_t off = ws_off[blockIdx.x];
    bool ok = k >= 0 && k < nets.n;
    if (ok) {
        const int64_t R = (int64_t)nets.box[4 * (int64_t)k + 2] * nets.box[4 * (int64_t)k + 3] * g.L;
        ok = R >= 1 && R <= max_nodes && off >= 0 && (off & 63) == 0 && off + R <= ws_limit;
    }
    if (!ok) {
        if (threadIdx.x == 0) {
            if (k >= 0 && k < nets.n) {
                nets.status[k] = ORX_E_ARG;
                nets.path_len[k] = 0;
            }
            if (stats) atomicAdd(stats + 6, 1ull);
        }
        return;
    }
    route_net<BLOCK>(g, P, nets, k, ws, off, delta, lds_nodes, lds_region, flags, stats,
                     reinterpret_cast<uint32_t*>(lds_mem), ctl, tt);
}

// Persistent dependency-driven form: workgroups claim nets in schedule order
// (a topological order of the dependency lists) from a global counter, wait
// until every dependency of the claimed net is published for this epoch,
// route it and publish it.  Every awaited net was claimed earlier by a
// running workgroup, so the wait always ends.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void route_dag_kernel(Lat g, Params P, NetsDev nets, const int32_t* sched,
                                                          int32_t n, const int32_t* dep_ptr, const int32_t* dep_ids,
                                                          int32_t* done, int32_t epoch, uint32_t* next,
                                                          const int64_t* ws_off, orx_workspace ws, float delta,
                                                          int32_t lds_nodes, int32_t lds_region, int32_t flags,
                                                          unsigned long long* stats) {
    ORX_DYNAMIC_LDS(uint64_t, lds_mem);
    ORX_SHARED(Ctl, ctl);
    ORX_SHARED(TermTab, tt);
    ORX_SHARED(int32_t, s_pick);
    for (;;) {
        if (threadIdx.x == 0) s_pick = (int32_t)atomicAdd(next, 1u);
        __syncthreads();
        const int32_t i = s_pick;
        if (i >= n) break;
        const int32_t k = sched[i];
        if (threadIdx.x == 0) {
            for (int32_t d = dep_ptr[k]; d < dep_ptr[k + 1]; ++d) {
                const int32_t p = dep_ids[d];
                while (__hip_atomic_load(done + p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != epoch)
                    __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();
        route_net<BLOCK>(g, P, nets, k, ws, ws_off[k], delta, lds_nodes, lds_region, flags, stats,
                         reinterpret_cast<uint32_t*>(lds_mem), ctl, tt);
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(done + k, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Clears the first n_keys keys (all-ones) and the first n_words words of
// each of the three bitmap arrays of a workspace, chip-wide and coalesced:
// the search state of the largest ROIs of a wave, which one workgroup
// would otherwise clear alone.
__global__ __launch_bounds__(256) void ws_fill_kernel(orx_workspace ws, int64_t n_keys, int64_t n_words) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    ulonglong2* k2 = reinterpret_cast<ulonglong2*>(ws.keys);
    const ulonglong2 inf2 = make_ulonglong2(ORX_KEY_INF, ORX_KEY_INF);
    for (int64_t i = t; i < n_keys / 2; i += stride) k2[i] = inf2;
    if (t == 0 && (n_keys & 1)) ws.keys[n_keys - 1] = ORX_KEY_INF;
    const int64_t wn = ws.ws_nodes / 32;
    for (int64_t i = t; i < n_words; i += stride) {
        ws.bitmaps[i] = 0u;
        ws.bitmaps[wn + i] = 0u;
        ws.bitmaps[2 * wn + i] = 0u;
    }
}

__global__ void zero_u32_kernel(uint32_t* p) {
    if (threadIdx.x == 0) p[0] = 0;
}

// Single box query writing ROI-local dist / parent arrays.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void sssp_box_kernel(Lat g, orx_workspace ws, int32_t x0, int32_t y0, int32_t w,
                                                         int32_t h, int64_t src, int64_t dst, float delta,
                                                         int32_t lds_nodes, int32_t lds_region, int32_t flags,
                                                         float* dist,
                                                         int32_t* parent, unsigned long long* stats,
                                                         int32_t* status) {
    ORX_DYNAMIC_LDS(uint64_t, lds_mem);
    ORX_SHARED(Ctl, ctl);
    const int32_t bx[4] = {x0, y0, w, h};
    const Box B = make_box(bx, g.L);
    const SlotState S = slot_state(ws, 0, B.R, reinterpret_cast<uint32_t*>(lds_mem), lds_region, lds_nodes);
    if (threadIdx.x == 0) { ctl.settled = 0; ctl.relaxed = 0; ctl.rounds = 0; ctl.phases = 0; ctl.items = 0; }
    const uint32_t rsrc = global_to_roi(g, B, src);
    const bool full = (flags & ORX_F_FULL) != 0 || dst < 0;
    const uint32_t rdst = dst >= 0 ? global_to_roi(g, B, dst) : 0;
    if (rsrc == UINT32_MAX || rdst == UINT32_MAX) {
        if (threadIdx.x == 0) *status = ORX_E_ARG;
        return;
    }
    Heur H;
    H.on = false;
    const Multi none{};
    box_sssp<BLOCK>(g, B, rsrc, rdst, full, delta, H, S.keys, S.bmn0, S.bmn1, S.bmf, S.qA, S.qB, S.lqA, S.lqB, S.fA,
                    S.fB, false, false, &ctl, false, none);
    if (flags & ORX_F_STATS) box_stats<BLOCK>(g, B, H, S.keys, rdst, full, &ctl);
    for (uint32_t r = threadIdx.x; r < B.R; r += BLOCK) {
        const uint64_t k = key_load(S.keys + r);
        if (dist) dist[r] = k == ORX_KEY_INF ? INFINITY : kdist(k);
        if (parent) {
            int32_t p = -1;
            if (k != ORX_KEY_INF && kslot(k) != ORX_ROOT_SLOT) {
                uint32_t lx, ly, z;
                roi_decode(B, r, lx, ly, z);
                const bool horiz = (g.hmask >> z) & 1u;
                const uint32_t pr = roi_nbr(B, horiz, r, lx, ly, z, (int)kslot(k));
                if (pr != UINT32_MAX) {
                    uint32_t plx, ply, pz;
                    roi_decode(B, pr, plx, ply, pz);
                    p = (int32_t)roi_global(g, B, plx, ply, pz);
                }
            }
            parent[r] = p;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint64_t kd = key_load(S.keys + rdst);
        *status = (!full && kd == ORX_KEY_INF) ? ORX_NO_PATH : ORX_OK;
        if (stats) {
            atomicAdd(stats + 0, ctl.settled);
            atomicAdd(stats + 1, ctl.relaxed);
            atomicAdd(stats + 2, (unsigned long long)ctl.items);
            atomicAdd(stats + 3, (unsigned long long)ctl.rounds);
            atomicAdd(stats + 4, (unsigned long long)ctl.phases);
        }
    }
}

// ---------------------------------------------------------------------------
// Whole-lattice streaming kernels (HBM-bound)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reprice_all_kernel(Lat g, Params P, int64_t N) {
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < N; u += (int64_t)gridDim.x * blockDim.x) {
        const float4 b4 = reinterpret_cast<const float4*>(g.base)[u];
        const float4 h4 = reinterpret_cast<const float4*>(g.hist)[u];
        const int4 o4 = reinterpret_cast<const int4*>(g.occ)[u];
        const float bb[4] = {b4.x, b4.y, b4.z, b4.w};
        const float hh[4] = {h4.x, h4.y, h4.z, h4.w};
        const int32_t oo[4] = {o4.x, o4.y, o4.z, o4.w};
        float ww[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            ww[s] = INFINITY;
            if (isinf(bb[s])) continue;
            const int64_t v = lat_nbr(g, u, s);
            if (v < 0) continue;
            const float q = node_cost(g.node_hist[v], g.node_use[v], P);
            ww[s] = __fadd_rn(arc_cost(bb[s], hh[s], oo[s], P), q);
        }
        reinterpret_cast<float4*>(g.wcost)[u] = make_float4(ww[0], ww[1], ww[2], ww[3]);
    }
}

// Overuse + history, fused.  out: edge_over, edge_over_count, node_over,
// node_over_count, via_hops, barrels.
__global__ __launch_bounds__(256) void end_iteration_kernel(Lat g, Params P, int64_t N, unsigned long long* out) {
    unsigned long long eo = 0, ec = 0, no = 0, nc = 0, vh = 0, br = 0;
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < N; u += (int64_t)gridDim.x * blockDim.x) {
        const float4 b4 = reinterpret_cast<const float4*>(g.base)[u];
        float4 h4 = reinterpret_cast<const float4*>(g.hist)[u];
        const int4 o4 = reinterpret_cast<const int4*>(g.occ)[u];
        const float bb[4] = {b4.x, b4.y, b4.z, b4.w};
        float hh[4] = {h4.x, h4.y, h4.z, h4.w};
        const int32_t oo[4] = {o4.x, o4.y, o4.z, o4.w};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            int32_t over = oo[s] - P.edge_cap;
            over = over < 0 ? 0 : over;
            if (s >= 2 && over > 0) { eo += over; ec += 1; }
            float inc = __fmul_rn(P.hist_gain, (float)over);
            if (P.hist_cap_mult > 0.0f) {
                const float capv = __fmul_rn(P.hist_cap_mult, bb[s]);
                if (capv < inc) inc = capv;
            }
            hh[s] = __fadd_rn(__fmul_rn(hh[s], P.hist_decay), inc);
        }
        vh += (unsigned long long)oo[3];
        reinterpret_cast<float4*>(g.hist)[u] = make_float4(hh[0], hh[1], hh[2], hh[3]);
        int32_t on = g.node_use[u] - P.node_cap;
        on = on < 0 ? 0 : on;
        if (on > 0) { no += on; nc += 1; }
        g.node_hist[u] = __fadd_rn(__fmul_rn(g.node_hist[u], P.hist_decay), __fmul_rn(P.node_hist_gain, (float)on));
        if (u < g.XY) br += (unsigned long long)g.via_col[u];
    }
    eo = wave_sum(eo); ec = wave_sum(ec); no = wave_sum(no); nc = wave_sum(nc); vh = wave_sum(vh); br = wave_sum(br);
    if (lane_id() == 0) {
        if (eo) atomicAdd(out + 0, eo);
        if (ec) atomicAdd(out + 1, ec);
        if (no) atomicAdd(out + 2, no);
        if (nc) atomicAdd(out + 3, nc);
        if (vh) atomicAdd(out + 4, vh);
        if (br) atomicAdd(out + 5, br);
    }
}

__global__ __launch_bounds__(256) void zero_i64_kernel(unsigned long long* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void apply_paths_kernel(Lat g, const int32_t* paths, const int64_t* off,
                                                            const int32_t* len, int32_t sign) {
    apply_path_block<BLOCK>(g, paths + off[blockIdx.x], len[blockIdx.x], sign);
}
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void reprice_paths_kernel(Lat g, Params P, const int32_t* paths,
                                                              const int64_t* off, const int32_t* len) {
    reprice_path_block<BLOCK>(g, P, paths + off[blockIdx.x], len[blockIdx.x]);
}

// Per-net overuse over the forward arcs of its committed path: the
// numerator of _order_nets_by_difficulty's congestion term
// (unified_pathfinder.py:7858-7863; both arcs of a hop hold equal occupancy).
// One wavefront per net, its lanes striding over the path's hops, so a long
// path costs len / 64 dependent loads instead of len.
__global__ __launch_bounds__(256) void net_overuse_kernel(Lat g, NetsDev nets, int32_t edge_cap, int32_t node_cap,
                                                          int64_t* out_over, int64_t* out_node_over,
                                                          int32_t* out_hops) {
    const int32_t k = (int32_t)(blockIdx.x * 4u + (threadIdx.x >> 6));
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    if (k >= nets.n) return;   // uniform per wavefront
    const int32_t len = nets.path_len[k];
    const int32_t* path = nets.path_arena + nets.path_off[k];
    int64_t over = 0, node_over = 0;
    for (int32_t i = lane; i < len; i += 64) {
        const int64_t u = path[i];
        if (out_node_over) {
            const int32_t o = g.node_use[u] - node_cap;
            node_over += o > 0 ? o : 0;
        }
        if (i + 1 < len) {
            const int s = lat_slot(g, u, path[i + 1]);
            const int32_t o = g.occ[4 * u + s] - edge_cap;
            over += o > 0 ? o : 0;
        }
    }
    for (int m = 32; m >= 1; m >>= 1) {
        over += __shfl_xor(over, m);
        node_over += __shfl_xor(node_over, m);
    }
    if (lane == 0) {
        out_over[k] = over;
        if (out_node_over) out_node_over[k] = node_over;
        out_hops[k] = len > 0 ? len - 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// Grid-wide box search: one query over an ROI too large for one workgroup
// (the 1000x1000x12 single-query microbenchmark, array-reporting form of
// find_path_roi_gpu).  A device-side state machine advanced by repeated
// launches of one step kernel; no workgroup ever waits on another:
//   RELAX  : every workgroup relaxes a slice of the round's near queue;
//   SPLIT1 : drop settled far entries, compact the live ones, min f;
//   SPLIT2 : entries below the new threshold form the next near queue.
// Each launch learns its step from the one before it, with no end-of-step
// ticket: launch k reads the record rec[k & 1] that launch k - 1 wrote and
// the counters launch k - 1 left (final once it ended), settles what the
// record left open (a RELAX whose queue came out empty is a SPLIT1, or DONE),
// and its workgroup 0 writes the record of launch k + 1 into rec[(k + 1) & 1],
// which no workgroup of launch k reads.  Every counter a launch appends to
// was cleared by an earlier launch that did not read it.  Launch sequences
// are replayed as a hipGraph by the host (an even number of launches, so the
// record parity carries over).  Same keys, tie-break and near-far rule as
// box_sssp.
// ---------------------------------------------------------------------------
enum : uint32_t { GS_RELAX = 0, GS_SPLIT1 = 1, GS_SPLIT2 = 2, GS_DONE = 3 };

// The step one launch executes, as far as the launch before it can tell.
struct GridRec {
    uint32_t state;     // RELAX / SPLIT2 / DONE (a RELAX may turn out a SPLIT1)
    uint32_t r;         // near-far round: queue and bitmap parity, count slot
    uint32_t phase;     // far-pile counter of the phase: fcnt[phase & 1]
    uint32_t pad;
    double T;           // threshold of the phase
};

struct GridCtl {
    GridRec rec[2];
    uint32_t qcnt[3];   // near-queue sizes, by round mod 3
    uint32_t fcnt[2];   // far-pile size, by phase parity
    uint32_t fcnt2;     // live far entries compacted by SPLIT1
    uint32_t done, err;
    unsigned long long minf;
    unsigned long long rounds, phases, items, settled, relaxed;
};

__device__ __forceinline__ uint64_t gkey_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t gkey_min(uint64_t* p, uint64_t v) {
    return __hip_atomic_fetch_min((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gbit_set(uint32_t* bm, uint32_t i) {
    const uint32_t b = 1u << (i & 31u);
    return __hip_atomic_fetch_or(bm + (i >> 5), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & b;
}
__device__ __forceinline__ void gbit_clear(uint32_t* bm, uint32_t i) {
    __hip_atomic_fetch_and(bm + (i >> 5), ~(1u << (i & 31u)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gload_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct GridArgs {
    Lat g;
    Box B;
    uint32_t rsrc, rdst;
    int32_t full;
    float delta;
    uint32_t cap;       // entries of each queue array
    uint64_t* keys;
    uint32_t *bmn0, *bmn1, *bmf, *qA, *qB, *fA, *fB;
    GridCtl* ctl;
};

__global__ __launch_bounds__(256) void grid_init_kernel(GridArgs A) {
    const uint32_t R = A.B.R, words = (R + 31) / 32;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < R; i += gridDim.x * blockDim.x)
        A.keys[i] = ORX_KEY_INF;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) {
        A.bmn0[i] = 0; A.bmn1[i] = 0; A.bmf[i] = 0;
    }
}

__global__ void grid_seed_kernel(GridArgs A) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    GridCtl* c = A.ctl;
    double T = (double)A.delta;
    if (!(T > 0.0)) T = nextafter(0.0, (double)INFINITY);
    c->rec[0].state = GS_RELAX; c->rec[0].r = 0; c->rec[0].phase = 0; c->rec[0].pad = 0; c->rec[0].T = T;
    c->rec[1] = c->rec[0];
    c->qcnt[0] = 1; c->qcnt[1] = 0; c->qcnt[2] = 0;
    c->fcnt[0] = 0; c->fcnt[1] = 0; c->fcnt2 = 0; c->done = 0; c->err = 0;
    c->minf = dbits((double)INFINITY);
    c->rounds = 0; c->phases = 0; c->items = 0; c->settled = 0; c->relaxed = 0;
    A.keys[A.rsrc] = kpack(0.0f, 0u, ORX_ROOT_SLOT);
    A.qA[0] = A.rsrc;
    A.bmn0[A.rsrc >> 5] |= 1u << (A.rsrc & 31u);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void grid_step_kernel(GridArgs A, uint32_t par) {
    GridCtl* ctl = A.ctl;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    // Every control field a step reads is loaded at once (one memory wait,
    // not one per field; counts are picked after).
    const GridRec rc = ctl->rec[par];
    const uint32_t qc0 = gload_u32(&ctl->qcnt[0]), qc1 = gload_u32(&ctl->qcnt[1]), qc2 = gload_u32(&ctl->qcnt[2]);
    const uint32_t fc0 = gload_u32(&ctl->fcnt[0]), fc1 = gload_u32(&ctl->fcnt[1]);
    const uint32_t nf2 = gload_u32(&ctl->fcnt2);
    const uint64_t kd = A.full ? ORX_KEY_INF : gkey_load(A.keys + A.rdst);
    const unsigned long long mnb = __hip_atomic_load(&ctl->minf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t r = rc.r, j = rc.phase;
    const double T = rc.T;
    const uint32_t rq = r % 3;
    const uint32_t n = rq == 0 ? qc0 : (rq == 1 ? qc1 : qc2);
    const uint32_t nf = (j & 1) ? fc1 : fc0;
    uint32_t state = rc.state;
    if (state == GS_RELAX && n == 0) {
        // the phase's bucket drained: stop (sink final, or nothing left
        // far), else split the far pile
        const bool sink = kd != ORX_KEY_INF && (double)kdist(kd) < T;
        state = (sink || nf == 0) ? GS_DONE : GS_SPLIT1;
    } else if (state == GS_SPLIT2 && nf2 == 0) {
        state = GS_DONE;
    }
    // counts beyond the queue arrays would be a broken state: stop with an error
    bool bad = false;
    if ((state == GS_RELAX && n > A.cap) || (state == GS_SPLIT1 && nf > A.cap) || (state == GS_SPLIT2 && nf2 > A.cap)) {
        bad = true;
        state = GS_DONE;
    }
    double Tn = 0.0;
    if (state == GS_SPLIT2) {
        const double mn = __longlong_as_double((long long)mnb);
        Tn = mn + (double)A.delta;
        if (!(Tn > mn)) Tn = nextafter(mn, (double)INFINITY);
    }
    if (lead) {
        // the next launch's record, and the counters it may append to
        GridRec nx = rc;
        if (state == GS_DONE) {
            nx.state = GS_DONE;
            ctl->done = 1;
            if (bad) ctl->err = 1;
        } else if (state == GS_RELAX) {
            nx.r = r + 1;
            ctl->qcnt[(r + 2) % 3] = 0;
            ctl->fcnt2 = 0;
            ctl->minf = dbits((double)INFINITY);
            ctl->rounds += 1;
            ctl->items += n;
        } else if (state == GS_SPLIT1) {
            nx.state = GS_SPLIT2;
            ctl->phases += 1;
        } else {  // SPLIT2: the next phase starts at the new threshold
            nx.state = GS_RELAX;
            nx.phase = j + 1;
            nx.T = Tn;
            ctl->fcnt[j & 1] = 0;
        }
        ctl->rec[par ^ 1u] = nx;
    }
    if (state == GS_DONE) return;
    const Box& B = A.B;
    const Lat& g = A.g;
    const uint32_t stride = gridDim.x * BLOCK;
    const uint32_t first = blockIdx.x * BLOCK;
    if (state == GS_RELAX) {
        uint32_t* cur = (r & 1) ? A.qB : A.qA;
        uint32_t* nxt = (r & 1) ? A.qA : A.qB;
        uint32_t* bmc = (r & 1) ? A.bmn1 : A.bmn0;
        uint32_t* bmx = (r & 1) ? A.bmn0 : A.bmn1;
        uint32_t* cnx = &ctl->qcnt[(r + 1) % 3];
        uint32_t* cfar = &ctl->fcnt[j & 1];
        for (uint32_t base = first; base < n; base += stride) {
            const uint32_t i = base + threadIdx.x;
            const bool active = i < n;
            uint32_t u = 0, lx = 0, ly = 0, z = 0;
            bool horiz = false;
            uint64_t ku = ORX_KEY_INF;
            float w4[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
            uint32_t nb[4] = {UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX};
            if (active) {
                u = cur[i];
                gbit_clear(bmc, u);
                ku = gkey_load(A.keys + u);
                roi_decode(B, u, lx, ly, z);
                horiz = (g.hmask >> z) & 1u;
                const float4 wv = reinterpret_cast<const float4*>(g.wcost)[roi_global(g, B, lx, ly, z)];
                w4[0] = wv.x; w4[1] = wv.y; w4[2] = wv.z; w4[3] = wv.w;
#pragma unroll
                for (int s = 0; s < 4; ++s) nb[s] = roi_nbr(B, horiz, u, lx, ly, z, s);
            }
            const float du = kdist(ku);
            const uint32_t hu = khops(ku);
            // The four arcs' atomics go out as batches (key minima, then
            // queue bits, then one reservation per queue and wavefront), so
            // a lane waits for device memory three times, not once per arc
            // and operation.
            float nd[4];
            uint64_t nk[4], old[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                nd[s] = __fadd_rn(du, w4[s]);
                nk[s] = kpack(nd[s], hu + 1, 3u - (uint32_t)s);
                old[s] = 0;
                if (nb[s] != UINT32_MAX && !isinf(w4[s])) old[s] = gkey_min(A.keys + nb[s], nk[s]);
            }
            bool to_near[4], to_far[4];
            uint32_t was[4];
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                to_near[s] = nk[s] < old[s] && (double)nd[s] < T;
                to_far[s] = nk[s] < old[s] && !to_near[s];
                was[s] = 1;
                if (to_near[s]) was[s] = gbit_set(bmx, nb[s]);
                else if (to_far[s]) was[s] = gbit_set(A.bmf, nb[s]);
            }
            uint32_t cn = 0, cf = 0;
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                cn += (to_near[s] && was[s] == 0) ? 1u : 0u;
                cf += (to_far[s] && was[s] == 0) ? 1u : 0u;
            }
            uint32_t an = wave_append_n(cn, cnx);
            uint32_t af = wave_append_n(cf, cfar);
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                if (to_near[s] && was[s] == 0) nxt[an++] = nb[s];
                if (to_far[s] && was[s] == 0) A.fA[af++] = nb[s];
            }
        }
    } else if (state == GS_SPLIT1) {
        double lmin = INFINITY;
        for (uint32_t base = first; base < nf; base += stride) {
            const uint32_t i = base + threadIdx.x;
            bool keep = false;
            uint32_t v = 0;
            if (i < nf) {
                v = A.fA[i];
                const double f = (double)kdist(gkey_load(A.keys + v));
                if (f < T) gbit_clear(A.bmf, v);
                else { keep = true; lmin = fmin(lmin, f); }
            }
            const int at = wave_append(keep, &ctl->fcnt2);
            if (keep) A.fB[at] = v;
        }
        for (int o = 32; o > 0; o >>= 1) lmin = fmin(lmin, __shfl_xor(lmin, o));
        if (lane_id() == 0 && lmin < (double)INFINITY)
            __hip_atomic_fetch_min(&ctl->minf, dbits(lmin), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {  // GS_SPLIT2
        uint32_t* qr = (r & 1) ? A.qB : A.qA;
        uint32_t* bmr = (r & 1) ? A.bmn1 : A.bmn0;
        uint32_t* cstay = &ctl->fcnt[(j + 1) & 1];
        for (uint32_t base = first; base < nf2; base += stride) {
            const uint32_t i = base + threadIdx.x;
            bool to_near = false, stay = false;
            uint32_t v = 0;
            if (i < nf2) {
                v = A.fB[i];
                const double f = (double)kdist(gkey_load(A.keys + v));
                if (f < Tn) { to_near = true; gbit_clear(A.bmf, v); gbit_set(bmr, v); }
                else stay = true;
            }
            int at = wave_append(to_near, &ctl->qcnt[r % 3]);
            if (to_near) qr[at] = v;
            at = wave_append(stay, cstay);
            if (stay) A.fA[at] = v;
        }
    }
}

// dist / parent arrays and settled-set statistics of a finished grid search.
__global__ __launch_bounds__(256) void grid_finish_kernel(GridArgs A, float* dist, int32_t* parent) {
    const Box& B = A.B;
    const Lat& g = A.g;
    const uint64_t kd = A.full ? ORX_KEY_INF : A.keys[A.rdst];
    unsigned long long settled = 0, relaxed = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < B.R; r += gridDim.x * blockDim.x) {
        const uint64_t k = A.keys[r];
        uint32_t lx, ly, z;
        roi_decode(B, r, lx, ly, z);
        const bool horiz = (g.hmask >> z) & 1u;
        if (dist) dist[r] = k == ORX_KEY_INF ? INFINITY : kdist(k);
        if (parent) {
            int32_t p = -1;
            if (k != ORX_KEY_INF && kslot(k) != ORX_ROOT_SLOT) {
                const uint32_t pr = roi_nbr(B, horiz, r, lx, ly, z, (int)kslot(k));
                if (pr != UINT32_MAX) {
                    uint32_t plx, ply, pz;
                    roi_decode(B, pr, plx, ply, pz);
                    p = (int32_t)roi_global(g, B, plx, ply, pz);
                }
            }
            parent[r] = p;
        }
        if (k == ORX_KEY_INF) continue;
        if (!A.full && (k > kd || (k == kd && r > A.rdst))) continue;
        settled += 1;
        if (!A.full && r == A.rdst) continue;
        for (int s = 0; s < 4; ++s) relaxed += roi_nbr(B, horiz, r, lx, ly, z, s) != UINT32_MAX;
    }
    settled = wave_sum(settled);
    relaxed = wave_sum(relaxed);
    if (lane_id() == 0) {
        if (settled) atomicAdd(&A.ctl->settled, settled);
        if (relaxed) atomicAdd(&A.ctl->relaxed, relaxed);
    }
}

// Dynamic LDS of a search launch whose largest ROI has max_nodes nodes: the
// two near-queue rings plus a region big enough for the keys and bitmaps of
// an ROI of min(lds_nodes, max_nodes) nodes or the bitmaps of the largest
// ROI, whichever is larger, capped by the LDS of a CU (each ROI then keeps
// in LDS what fits; see slot_state).
struct LdsPlan {
    size_t bytes;
    int32_t region;
    int32_t lds_nodes;
};
LdsPlan lds_plan(int64_t max_nodes, int32_t lds_nodes) {
    if (lds_nodes < 0) lds_nodes = 0;
    const int64_t ring = 2ll * kRing * 4;
    const int64_t key_n = std::min<int64_t>(lds_nodes, max_nodes);
    const int64_t key_need = key_n > 0 ? key_n * 8 + 3 * ((key_n + 31) / 32) * 4 : 0;
    const int64_t bm_need = 3 * ((max_nodes + 31) / 32) * 4;
    int64_t region = std::max(key_need, bm_need);
    region = std::min<int64_t>(region, (int64_t)kMaxDynLds - ring) / 16 * 16;
    return LdsPlan{(size_t)(ring + region), (int32_t)region, lds_nodes};
}

int32_t L_of(const orx_lattice* l) { return l->L; }

__global__ void grid_stats_kernel(const GridCtl* c, unsigned long long* stats, int32_t* status, int32_t full,
                                  const uint64_t* keys, uint32_t rdst) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (stats) {
        stats[0] += c->settled; stats[1] += c->relaxed; stats[2] += c->items;
        stats[3] += c->rounds; stats[4] += c->phases;
    }
    *status = c->err ? ORX_E_OVERFLOW : ((!full && keys[rdst] == ORX_KEY_INF) ? ORX_NO_PATH : ORX_OK);
}

// Host driver of the grid state machine: seed, replay a captured graph of
// kGridSteps step launches until the device reports done, then write the
// arrays.  Runs on a private stream after the caller's stream drained.
constexpr int kGridSteps = 64;

int sssp_box_grid(const Lat& g, int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t L, uint32_t rsrc, uint32_t rdst,
                  bool full, float* d_dist, int32_t* d_parent, const orx_workspace* ws, const orx_route_cfg* cfg,
                  hipStream_t caller) {
    GridArgs A;
    A.g = g;
    A.B.x0 = x0; A.B.y0 = y0; A.B.w = w; A.B.h = h; A.B.L = L;
    A.B.wh = (uint32_t)w * (uint32_t)h; A.B.R = A.B.wh * (uint32_t)L;
    A.B.dw = FastDiv((uint32_t)w); A.B.dwh = FastDiv(A.B.wh);
    A.rsrc = rsrc; A.rdst = rdst; A.full = full ? 1 : 0; A.delta = cfg->delta;
    const int64_t wn = ws->ws_nodes;
    A.keys = ws->keys;
    A.qA = ws->queues; A.qB = ws->queues + wn; A.fA = ws->queues + 2 * wn; A.fB = ws->queues + 3 * wn;
    A.bmn0 = ws->bitmaps; A.bmn1 = ws->bitmaps + wn / 32; A.bmf = ws->bitmaps + 2 * (wn / 32);
    A.cap = (uint32_t)std::min<int64_t>(wn, (int64_t)UINT32_MAX);
    int dev = 0, cus = 1;
    ORX_HIP_CHECK(hipGetDevice(&dev));
    ORX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    ORX_HIP_CHECK(hipStreamSynchronize(caller));
    hipStream_t ps = nullptr;
    ORX_HIP_CHECK(hipStreamCreateWithFlags(&ps, hipStreamNonBlocking));
    GridCtl* ctl = nullptr;
    int32_t* d_status = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = ORX_OK;
    hipError_t e = hipMallocAsync((void**)&ctl, sizeof(GridCtl), ps);
    if (e == hipSuccess) e = hipMallocAsync((void**)&d_status, sizeof(int32_t), ps);
    if (e == hipSuccess) {
        A.ctl = ctl;
        const int fill_grid = std::max(1, std::min(cus * 8, (int)((A.B.R + 255) / 256)));
        hipLaunchKernelGGL(grid_init_kernel, dim3(fill_grid), dim3(256), 0, ps, A);
        hipLaunchKernelGGL(grid_seed_kernel, dim3(1), dim3(64), 0, ps, A);
        e = hipGetLastError();
    }
    const int block = cfg->block;
    // Half a workgroup per CU: a step is short (a few thousand frontier
    // nodes), so fewer workgroups mean a cheaper dispatch (1000x1000x12, 256
    // lanes: 64 or 128 groups 42 ms per query, 512 groups 55 ms, 1024 groups
    // 76 ms, measured with an end-of-step ticket).
    const int grid = std::max(1, cus / 2);
    if (e == hipSuccess) e = hipStreamBeginCapture(ps, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        // kGridSteps is even: launch i reads control record i & 1, and every
        // replay starts again at record 0
        static_assert(kGridSteps % 2 == 0, "record parity must carry over between replays");
        for (int i = 0; i < kGridSteps; ++i) {
            const uint32_t par = (uint32_t)(i & 1);
            switch (block) {
            case 512: hipLaunchKernelGGL(grid_step_kernel<512>, dim3(grid), dim3(512), 0, ps, A, par); break;
            case 1024: hipLaunchKernelGGL(grid_step_kernel<1024>, dim3(grid), dim3(1024), 0, ps, A, par); break;
            default: hipLaunchKernelGGL(grid_step_kernel<256>, dim3(grid), dim3(256), 0, ps, A, par); break;
            }
        }
        hipError_t e2 = hipStreamEndCapture(ps, &graph);
        e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    // Replays run two deep: replay k + 1 is queued before the host waits for
    // replay k's done flag, so the device does not idle through the host's
    // round trip (the replay queued after the end runs only no-op launches).
    uint32_t* h_done = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (e == hipSuccess) e = hipHostMalloc((void**)&h_done, 2 * sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    auto replay = [&](int slot) -> hipError_t {
        hipError_t x = hipGraphLaunch(exec, ps);
        if (x == hipSuccess) x = hipMemcpyAsync(h_done + slot, &ctl->done, sizeof(uint32_t), hipMemcpyDeviceToHost, ps);
        if (x == hipSuccess) x = hipEventRecord(ev[slot], ps);
        return x;
    };
    if (e == hipSuccess) e = replay(0);
    if (e == hipSuccess) {
        // bounded wait: a search that has not finished after the time limit
        // reports an error instead of looping forever
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0;; ++k) {
            const int cur = k & 1;
            e = replay(cur ^ 1);
            if (e == hipSuccess) e = hipEventSynchronize(ev[cur]);
            if (e != hipSuccess || h_done[cur]) break;
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300)) {
                rc = orx_host::fail(ORX_E_OVERFLOW, "grid search did not finish within 300 s");
                break;
            }
        }
    }
    if (e == hipSuccess && rc == ORX_OK) {
        const int fin_grid = std::max(1, std::min(cus * 8, (int)((A.B.R + 255) / 256)));
        hipLaunchKernelGGL(grid_finish_kernel, dim3(fin_grid), dim3(256), 0, ps, A, d_dist, d_parent);
        hipLaunchKernelGGL(grid_stats_kernel, dim3(1), dim3(64), 0, ps, (const GridCtl*)ctl,
                           (unsigned long long*)cfg->stats, d_status, A.full, (const uint64_t*)A.keys, rdst);
        int32_t h_status = 0;
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&h_status, d_status, sizeof(int32_t), hipMemcpyDeviceToHost, ps);
        if (e == hipSuccess) e = hipStreamSynchronize(ps);
        if (e == hipSuccess) rc = h_status == ORX_E_OVERFLOW
                                      ? orx_host::fail(ORX_E_OVERFLOW, "grid search: a queue count exceeded its array")
                                      : h_status;
    }
    if (e != hipSuccess && rc == ORX_OK) rc = orx_host::check_hip(e, "orx_sssp_box (grid)");
    (void)hipStreamSynchronize(ps);   // no replay or copy in flight past this point
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    if (h_done) (void)hipHostFree(h_done);
    if (ctl) (void)hipFreeAsync(ctl, ps);
    if (d_status) (void)hipFreeAsync(d_status, ps);
    (void)hipStreamSynchronize(ps);
    (void)hipStreamDestroy(ps);
    return rc;
}

// One route_wave_kernel launch (K nets, one workgroup each) after the
// argument checks shared by orx_route_wave and orx_route_waves.
// ws_extent > 0: the wave's nets own workspace nodes inside [0, ws_extent).
// If the largest ROI keeps its keys in the global workspace, that range is
// cleared chip-wide first and the searches skip their own clear.
int launch_route_wave(const Lat& g, const Params& P, const NetsDev& nd, const int32_t* d_wave_nets,
                      const int64_t* d_ws_off, int32_t K, int64_t max_nodes, int64_t ws_extent,
                      const orx_workspace* ws, const orx_route_cfg* cfg, hipStream_t s) {
    if (K <= 0) return ORX_OK;
    if (max_nodes < 1 || max_nodes >= (1ll << 24)) return orx_host::fail(ORX_E_ARG, "ROI size outside [1, 2^24)");
    const LdsPlan lp = lds_plan(max_nodes, cfg->lds_nodes);
    const size_t lds = lp.bytes;
    unsigned long long* st = (unsigned long long*)cfg->stats;
    int rc = ORX_OK;
    int32_t flags = cfg->flags & ~kPrefilled;
    const int64_t limit = ws_extent > 0 ? std::min<int64_t>(ws_extent, ws->ws_nodes) : ws->ws_nodes;
    if (ws_extent > 0) {
        // whether a net keeps keys / bitmaps in the workspace grows with its
        // ROI, so the largest ROI of the wave decides
        const int64_t words = (max_nodes + 31) / 32;
        const bool keys_global = !(max_nodes <= lp.lds_nodes && 8 * max_nodes + 12 * words <= lp.region);
        const bool bm_global = 12 * words > lp.region;
        if (keys_global) {
            const int64_t ext = std::min<int64_t>(ws_extent, ws->ws_nodes);
            const int64_t per = std::max<int64_t>(ext / 2, (ext + 31) / 32);
            const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(device_cus() * 8, (per + 255) / 256));
            hipLaunchKernelGGL(ws_fill_kernel, dim3(grid), dim3(256), 0, s, *ws, ext,
                               bm_global ? (ext + 31) / 32 : (int64_t)0);
            flags |= kPrefilled;
        }
    }
#define ORX_LAUNCH_ROUTE(BS)                                                                              \
    do {                                                                                                  \
        rc = set_lds<route_wave_kernel<BS>>(lds);                                                         \
        if (rc) return rc;                                                                                \
        hipLaunchKernelGGL(route_wave_kernel<BS>, dim3(K), dim3(BS), lds, s, g, P, nd, d_wave_nets,       \
                           d_ws_off, *ws, cfg->delta, lp.lds_nodes, lp.region, flags, st, limit, max_nodes); \
    } while (0)
    switch (cfg->block) {
    case 256: ORX_LAUNCH_ROUTE(256); break;
    case 512: ORX_LAUNCH_ROUTE(512); break;
    case 1024: ORX_LAUNCH_ROUTE(1024); break;
    default: return orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    }
#undef ORX_LAUNCH_ROUTE
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Search scratch of orx_sssp_box_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends, so the entry point allocates
// nothing.  Layout for up to ORX_NOALLOC_MAX_NODES ROI nodes: keys (u64),
// four queues (u32), three bitmaps, the status word.  Host code never uses
// the array's name as an address: each device's address comes from
// hipGetSymbolAddress, once, with that device current.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocNodes = ORX_NOALLOC_MAX_NODES;
constexpr size_t kScratchKeys = 8 * kNoAllocNodes;
constexpr size_t kScratchQueues = 16 * kNoAllocNodes;
constexpr size_t kScratchBitmaps = 3 * kNoAllocNodes / 8;
constexpr size_t kScratchBytes = kScratchKeys + kScratchQueues + kScratchBitmaps + 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// This device's copy of the scratch array (looked up on first use).
int scratch_for_current_device(unsigned char** out) {
    int dev = 0;
    ORX_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) return orx_host::fail(ORX_E_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        ORX_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return orx_host::fail(ORX_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return ORX_OK;
}

namespace orx_host {

int prepare_lattice_kernels() {
    int rc = ORX_OK;
    if (!rc) rc = set_lds<route_wave_kernel<256>>(kMaxDynLds);
    if (!rc) rc = set_lds<route_wave_kernel<512>>(kMaxDynLds);
    if (!rc) rc = set_lds<route_wave_kernel<1024>>(kMaxDynLds);
    if (!rc) rc = set_lds<route_dag_kernel<256>>(kMaxDynLds);
    if (!rc) rc = set_lds<route_dag_kernel<512>>(kMaxDynLds);
    if (!rc) rc = set_lds<route_dag_kernel<1024>>(kMaxDynLds);
    if (!rc) rc = set_lds<sssp_box_kernel<256>>(kMaxDynLds);
    if (!rc) rc = set_lds<sssp_box_kernel<512>>(kMaxDynLds);
    if (!rc) rc = set_lds<sssp_box_kernel<1024>>(kMaxDynLds);
    unsigned char* sp = nullptr;
    if (!rc) rc = scratch_for_current_device(&sp);   // the allocation-free entry's scratch, this device
    return rc;
}

}  // namespace orx_host

extern "C" {

int orx_lattice_reprice(const orx_lattice* lat, const orx_params* p, orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!p) return orx_host::fail(ORX_E_ARG, "params is NULL");
    const int64_t N = (int64_t)lat->X * lat->Y * lat->L;
    const int grid = (int)std::min<int64_t>((N + 255) / 256, 65536);
    hipLaunchKernelGGL(reprice_all_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, make_lat(lat),
                       make_params(p), N);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_lattice_end_iteration(const orx_lattice* lat, const orx_params* p, int64_t* d_out, orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!p || !d_out) return orx_host::fail(ORX_E_ARG, "params / out is NULL");
    const int64_t N = (int64_t)lat->X * lat->Y * lat->L;
    hipLaunchKernelGGL(zero_i64_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, (unsigned long long*)d_out, 8);
    // grid-stride over 32 workgroups per CU at most (8192 on MI355X)
    const int grid = (int)std::min<int64_t>((N + 255) / 256, 32ll * device_cus());
    hipLaunchKernelGGL(end_iteration_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, make_lat(lat),
                       make_params(p), N, (unsigned long long*)d_out);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_route_wave(const orx_lattice* lat, const orx_params* p, const orx_nets* nets, const int32_t* d_wave_nets,
                   const int64_t* d_ws_off, int32_t K, int64_t max_nodes, const orx_workspace* ws,
                   const orx_route_cfg* cfg, orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!p || !nets || !ws || !cfg || !d_wave_nets || !d_ws_off) return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (K <= 0) return ORX_OK;
    if (ws->ws_nodes % 64) return orx_host::fail(ORX_E_ARG, "ws_nodes must be a multiple of 64");
    return launch_route_wave(make_lat(lat), make_params(p), make_nets(nets), d_wave_nets, d_ws_off, K, max_nodes, 0,
                             ws, cfg, (hipStream_t)stream);
}

int orx_route_waves(const orx_lattice* lat, const orx_params* p, const orx_nets* nets, const int32_t* d_nets,
                    const int64_t* d_ws_off, int32_t n_waves, const int32_t* h_wave_ptr, const int64_t* h_wave_max,
                    const int64_t* h_wave_ws, const orx_workspace* ws, const orx_route_cfg* cfg,
                    orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!p || !nets || !ws || !cfg || !h_wave_ptr || !h_wave_max || !h_wave_ws)
        return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (n_waves <= 0) return ORX_OK;
    if (!d_nets || !d_ws_off) return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (ws->ws_nodes % 64) return orx_host::fail(ORX_E_ARG, "ws_nodes must be a multiple of 64");
    if (h_wave_ptr[0] != 0) return orx_host::fail(ORX_E_ARG, "wave_ptr[0] must be 0");
    for (int32_t i = 0; i < n_waves; ++i) {
        if (h_wave_ptr[i + 1] < h_wave_ptr[i] || h_wave_ptr[i + 1] > nets->n)
            return orx_host::fail(ORX_E_ARG, "wave_ptr must be non-decreasing and at most n");
        if (h_wave_ws[i] < 0 || h_wave_ws[i] > ws->ws_nodes)
            return orx_host::fail(ORX_E_ARG, "wave workspace extent outside [0, ws_nodes]");
    }
    const Lat g = make_lat(lat);
    const Params P = make_params(p);
    const NetsDev nd = make_nets(nets);
    const hipStream_t s = (hipStream_t)stream;
    for (int32_t i = 0; i < n_waves; ++i) {
        const int32_t a = h_wave_ptr[i], K = h_wave_ptr[i + 1] - a;
        rc = launch_route_wave(g, P, nd, d_nets + a, d_ws_off + a, K, h_wave_max[i], h_wave_ws[i], ws, cfg, s);
        if (rc) return rc;
    }
    return ORX_OK;
}

int orx_route_dag(const orx_lattice* lat, const orx_params* p, const orx_nets* nets, const int32_t* d_sched,
                  int32_t n, const int32_t* d_dep_ptr, const int32_t* d_dep_ids, int32_t* d_done, int32_t epoch,
                  uint32_t* d_counter, const int64_t* d_ws_off, int64_t max_nodes, const orx_workspace* ws,
                  const orx_route_cfg* cfg, int32_t max_groups, orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!p || !nets || !ws || !cfg || !d_sched || !d_dep_ptr || !d_done || !d_counter || !d_ws_off)
        return orx_host::fail(ORX_E_ARG, "NULL argument");
    // Switched off on the device: its first gfx950 run ended in an illegal
    // address fault whose cause has not been found (the host-thread build of
    // this kernel, with sequential and concurrent workgroups under
    // AddressSanitizer, is clean and bit-exact).  Only that build, which
    // defines ORX_HOST_THREAD_BUILD, may run it.
#ifndef ORX_HOST_THREAD_BUILD
    return orx_host::fail(ORX_E_ARG, "orx_route_dag is switched off on the device (unexplained fault)");
#endif
    if (n <= 0) return ORX_OK;
    if (max_nodes >= (1ll << 24)) return orx_host::fail(ORX_E_ARG, "ROI larger than 2^24 nodes");
    if (ws->ws_nodes % 64) return orx_host::fail(ORX_E_ARG, "ws_nodes must be a multiple of 64");
    const LdsPlan lp = lds_plan(max_nodes, cfg->lds_nodes);
    const size_t lds = lp.bytes;
    const NetsDev nd = make_nets(nets);
    unsigned long long* st = (unsigned long long*)cfg->stats;
    const hipStream_t s = (hipStream_t)stream;
    const Lat g = make_lat(lat);
    const Params P = make_params(p);
    int dev = 0, cus = 0;
    ORX_HIP_CHECK(hipGetDevice(&dev));
    ORX_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, d_counter);
#define ORX_LAUNCH_DAG(BS)                                                                                \
    do {                                                                                                  \
        rc = set_lds<route_dag_kernel<BS>>(lds);                                                          \
        if (rc) return rc;                                                                                \
        int per_cu = 0;                                                                                   \
        ORX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, route_dag_kernel<BS>, BS, lds)); \
        int grid = std::max(1, per_cu) * std::max(1, cus);                                                \
        if (max_groups > 0) grid = std::min(grid, max_groups);                                            \
        grid = std::min(grid, n);                                                                         \
        hipLaunchKernelGGL(route_dag_kernel<BS>, dim3(grid), dim3(BS), lds, s, g, P, nd, d_sched, n,      \
                           d_dep_ptr, d_dep_ids, d_done, epoch, d_counter, d_ws_off, *ws, cfg->delta,     \
                           lp.lds_nodes, lp.region, cfg->flags, st);                                      \
    } while (0)
    switch (cfg->block) {
    case 256: ORX_LAUNCH_DAG(256); break;
    case 512: ORX_LAUNCH_DAG(512); break;
    case 1024: ORX_LAUNCH_DAG(1024); break;
    default: return orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    }
#undef ORX_LAUNCH_DAG
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_net_overuse(const orx_lattice* lat, const orx_nets* nets, int32_t edge_cap, int32_t node_cap,
                    int64_t* d_over, int64_t* d_node_over, int32_t* d_hops, orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!nets || !d_over || !d_hops) return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (nets->n <= 0) return ORX_OK;
    const NetsDev nd = make_nets(nets);
    hipLaunchKernelGGL(net_overuse_kernel, dim3((nets->n + 3) / 4), dim3(256), 0, (hipStream_t)stream,
                       make_lat(lat), nd, edge_cap, node_cap, d_over, d_node_over, d_hops);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_apply_paths(const orx_lattice* lat, const orx_params* p, const int32_t* d_paths, const int64_t* d_off,
                    const int32_t* d_len, int32_t K, int32_t sign, orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!p || !d_paths || !d_off || !d_len) return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (sign != 1 && sign != -1) return orx_host::fail(ORX_E_ARG, "sign must be +1 or -1");
    if (K <= 0) return ORX_OK;
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(apply_paths_kernel<256>, dim3(K), dim3(256), 0, s, make_lat(lat), d_paths, d_off, d_len, sign);
    hipLaunchKernelGGL(reprice_paths_kernel<256>, dim3(K), dim3(256), 0, s, make_lat(lat), make_params(p), d_paths,
                       d_off, d_len);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_sssp_box(const orx_lattice* lat, int32_t x0, int32_t y0, int32_t w, int32_t h, int64_t src, int64_t dst,
                 float* d_dist, int32_t* d_parent, const orx_workspace* ws, const orx_route_cfg* cfg,
                 orx_stream_t stream) {
    if (!lat || !lat->wcost) return orx_host::fail(ORX_E_ARG, "lattice / wcost is NULL");
    if (!ws || !cfg) return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (x0 < 0 || y0 < 0 || w < 1 || h < 1 || x0 + w > lat->X || y0 + h > lat->Y)
        return orx_host::fail(ORX_E_ARG, "box outside the lattice");
    const int64_t R = (int64_t)w * h * lat->L;
    if (R >= (1ll << 24)) return orx_host::fail(ORX_E_ARG, "ROI larger than 2^24 nodes");
    if (ws->ws_nodes < R || ws->ws_nodes % 64) return orx_host::fail(ORX_E_ARG, "workspace smaller than the ROI");
    const int64_t XY = (int64_t)lat->X * lat->Y;
    auto to_roi = [&](int64_t gid) -> int64_t {
        if (gid < 0 || gid >= XY * lat->L) return -1;
        const int64_t z = gid / XY, rem = gid - z * XY, y = rem / lat->X, x = rem - y * lat->X;
        if (x < x0 || x >= x0 + w || y < y0 || y >= y0 + h) return -1;
        return (z * h + (y - y0)) * w + (x - x0);
    };
    if (to_roi(src) < 0 || (dst >= 0 && to_roi(dst) < 0)) return orx_host::fail(ORX_E_ARG, "src / dst outside the box");
    const hipStream_t s = (hipStream_t)stream;
    Lat g;
    g.X = lat->X; g.Y = lat->Y; g.L = lat->L; g.hmask = lat->hmask; g.XY = XY;
    g.base = lat->base; g.hist = lat->hist; g.occ = lat->occ; g.wcost = lat->wcost;
    g.node_hist = lat->node_hist; g.node_use = lat->node_use; g.via_col = lat->via_col;
    if (cfg->flags & ORX_F_TILES)
        return orx_host::sssp_box_tiles(lat, x0, y0, w, h, src, dst, d_dist, d_parent, ws, cfg, s);
    const bool use_grid = (cfg->flags & ORX_F_GRID) != 0 || R > (1 << 18);
    if (use_grid) return sssp_box_grid(g, x0, y0, w, h, L_of(lat), (uint32_t)to_roi(src),
                                       dst >= 0 ? (uint32_t)to_roi(dst) : 0u, dst < 0 || (cfg->flags & ORX_F_FULL),
                                       d_dist, d_parent, ws, cfg, s);
    const LdsPlan lp = lds_plan(R, cfg->lds_nodes);
    const size_t lds = lp.bytes;
    int32_t* d_status = nullptr;
    int32_t h_status = 0;
    ORX_HIP_CHECK(hipMallocAsync((void**)&d_status, sizeof(int32_t), s));
    int rc = ORX_OK;
#define ORX_LAUNCH_BOX(BS)                                                                                 \
    do {                                                                                                   \
        rc = set_lds<sssp_box_kernel<BS>>(lds);                                                            \
        if (rc) break;                                                                                     \
        hipLaunchKernelGGL(sssp_box_kernel<BS>, dim3(1), dim3(BS), lds, s, g, *ws, x0, y0, w, h, src, dst, \
                           cfg->delta, lp.lds_nodes, lp.region, cfg->flags, d_dist, d_parent,              \
                           (unsigned long long*)cfg->stats, d_status);                                     \
    } while (0)
    switch (cfg->block) {
    case 256: ORX_LAUNCH_BOX(256); break;
    case 512: ORX_LAUNCH_BOX(512); break;
    case 1024: ORX_LAUNCH_BOX(1024); break;
    default: rc = orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    }
#undef ORX_LAUNCH_BOX
    if (rc == ORX_OK) {
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&h_status, d_status, sizeof(int32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = orx_host::check_hip(e, "orx_sssp_box");
        else rc = h_status;
    }
    (void)hipFreeAsync(d_status, s);
    return rc;
}

int orx_sssp_box_noalloc(const orx_lattice* lat, int32_t x0, int32_t y0, int32_t w, int32_t h, int64_t src,
                         int64_t dst, float* d_dist, int32_t* d_parent, const orx_route_cfg* cfg,
                         orx_stream_t stream) {
    if (!lat || !lat->wcost) return orx_host::fail(ORX_E_ARG, "lattice / wcost is NULL");
    if (!cfg) return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (x0 < 0 || y0 < 0 || w < 1 || h < 1 || x0 + w > lat->X || y0 + h > lat->Y)
        return orx_host::fail(ORX_E_ARG, "box outside the lattice");
    const int64_t R = (int64_t)w * h * lat->L;
    if (R > kNoAllocNodes)
        return orx_host::fail(ORX_E_ARG, "ROI exceeds the allocation-free scratch (ORX_NOALLOC_MAX_NODES nodes)");
    if (cfg->flags & (ORX_F_GRID | ORX_F_TILES))
        return orx_host::fail(ORX_E_ARG, "the allocation-free query runs in one workgroup (no grid / tile flags)");
    const int64_t XY = (int64_t)lat->X * lat->Y;
    auto to_roi = [&](int64_t gid) -> int64_t {
        if (gid < 0 || gid >= XY * lat->L) return -1;
        const int64_t z = gid / XY, rem = gid - z * XY, y = rem / lat->X, x = rem - y * lat->X;
        if (x < x0 || x >= x0 + w || y < y0 || y >= y0 + h) return -1;
        return (z * h + (y - y0)) * w + (x - x0);
    };
    if (to_roi(src) < 0 || (dst >= 0 && to_roi(dst) < 0)) return orx_host::fail(ORX_E_ARG, "src / dst outside the box");
    unsigned char* sp = nullptr;
    int rc = scratch_for_current_device(&sp);
    if (rc) return rc;
    // the workspace slice of the scratch, sized to this ROI's node count
    // rounded to 64 (every array of the one-workgroup search is cleared by
    // the kernel itself; the status word is written by it)
    orx_workspace ws;
    ws.ws_nodes = kNoAllocNodes;
    ws.keys = reinterpret_cast<uint64_t*>(sp);
    ws.queues = reinterpret_cast<uint32_t*>(sp + kScratchKeys);
    ws.bitmaps = reinterpret_cast<uint32_t*>(sp + kScratchKeys + kScratchQueues);
    int32_t* d_status = reinterpret_cast<int32_t*>(sp + kScratchKeys + kScratchQueues + kScratchBitmaps);
    const Lat g = make_lat(lat);
    const hipStream_t s = (hipStream_t)stream;
    const LdsPlan lp = lds_plan(R, cfg->lds_nodes);
    const size_t lds = lp.bytes;
#define ORX_LAUNCH_BOX(BS)                                                                                  \
    do {                                                                                                    \
        rc = set_lds<sssp_box_kernel<BS>>(lds);                                                             \
        if (rc) break;                                                                                      \
        hipLaunchKernelGGL(sssp_box_kernel<BS>, dim3(1), dim3(BS), lds, s, g, ws, x0, y0, w, h, src, dst,   \
                           cfg->delta, lp.lds_nodes, lp.region, cfg->flags, d_dist, d_parent,               \
                           (unsigned long long*)cfg->stats, d_status);                                      \
    } while (0)
    switch (cfg->block) {
    case 256: ORX_LAUNCH_BOX(256); break;
    case 512: ORX_LAUNCH_BOX(512); break;
    case 1024: ORX_LAUNCH_BOX(1024); break;
    default: rc = orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    }
#undef ORX_LAUNCH_BOX
    if (rc) return rc;
    int32_t h_status = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(&h_status, d_status, sizeof(int32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return orx_host::check_hip(e, "orx_sssp_box_noalloc");
    return h_status;
}

}  // extern "C"
