// orx_csr.hip -- generic CSR search behind the gpu_solver seam.
//
// Serves find_path_roi_gpu (cuda_dijkstra.py:4951-5040: ROI node list +
// global_to_roi map, ROI-local node penalty folded into the entered arc,
// cuda_dijkstra.py:5079-5081) and find_path_fullgraph_gpu_seeds
// (cuda_dijkstra.py:5516-5850: seeds priced with their cost plus the seed's
// node penalty, cuda_dijkstra.py:5589-5606; targets with target costs;
// optional allowed bitmap).  Same near-far delta-stepping and packed
// (dist, hops, in-rank) keys as the lattice path; the in-rank of an arc is
// precomputed on the host from the transposed graph.  Targets are not
// expanded (unified_pathfinder.py:2611-2618).
#include "orx_common.hpp"

#include <algorithm>
#include <chrono>
#include "../../include/orthoroute_hip.h"

using namespace orx;

namespace {

struct CsrCtl {
    uint32_t qcnt[3];
    uint32_t fcnt, fcnt2, minbits;
    uint32_t tg_left;
    int32_t err;
};

__device__ __forceinline__ uint64_t gkey_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint64_t gkey_min(uint64_t* p, uint64_t v) {
    return __hip_atomic_fetch_min((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t gbit_set(uint32_t* bm, uint32_t i) {
    uint32_t b = 1u << (i & 31u);
    return __hip_atomic_fetch_or(bm + (i >> 5), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & b;
}
__device__ __forceinline__ void gbit_clear(uint32_t* bm, uint32_t i) {
    __hip_atomic_fetch_and(bm + (i >> 5), ~(1u << (i & 31u)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ bool gbit_test(const uint32_t* bm, uint32_t i) {
    return (__hip_atomic_load(bm + (i >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> (i & 31u)) & 1u;
}
__device__ __forceinline__ float wmin(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}

struct CsrArgs {
    orx_csr g;
    const float* costs;
    int64_t R;
    const int32_t* roi_nodes;
    const int32_t* g2r;
    const float* pen;
    const uint32_t* allowed;
    int32_t n_seeds;
    const int32_t* seeds;
    const float* seed_cost;
    int32_t n_tg;
    const int32_t* targets;
    const float* tg_cost;
    uint64_t* keys;
    uint32_t* scratch;  // 4R queue words + 4*ceil(R/32) bitmap words
    int32_t* best;
    float delta;
    int32_t seed_pen;
};

__device__ __forceinline__ int64_t to_roi(const CsrArgs& A, int64_t gid) {
    return A.g2r ? (int64_t)A.g2r[gid] : gid;
}
__device__ __forceinline__ bool allowed_node(const CsrArgs& A, int64_t gid) {
    return !A.allowed || ((A.allowed[gid >> 5] >> (gid & 31)) & 1u);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void csr_sssp_kernel(CsrArgs A) {
    ORX_SHARED(CsrCtl, ctl);
    const uint32_t tid = threadIdx.x;
    const uint32_t R = (uint32_t)A.R;
    const uint32_t words = (R + 31) / 32;
    uint32_t* qA = A.scratch;
    uint32_t* qB = qA + R;
    uint32_t* fA = qB + R;
    uint32_t* fB = fA + R;
    uint32_t* bmn0 = fB + R;
    uint32_t* bmn1 = bmn0 + words;
    uint32_t* bmf = bmn1 + words;
    uint32_t* bmt = bmf + words;  // targets
    uint64_t* keys = A.keys;

    for (uint32_t r = tid; r < R; r += BLOCK) keys[r] = ORX_KEY_INF;
    for (uint32_t i = tid; i < words; i += BLOCK) { bmn0[i] = 0; bmn1[i] = 0; bmf[i] = 0; bmt[i] = 0; }
    if (tid == 0) {
        ctl.qcnt[0] = 0; ctl.qcnt[1] = 0; ctl.qcnt[2] = 0;
        ctl.fcnt = 0; ctl.fcnt2 = 0; ctl.tg_left = 0; ctl.err = 0;
    }
    __syncthreads();
    // targets
    for (int32_t i = tid; i < A.n_tg; i += BLOCK) {
        const int64_t r = to_roi(A, A.targets[i]);
        if (r >= 0 && r < (int64_t)R) gbit_set(bmt, (uint32_t)r);
    }
    // seeds: key = fl(seed_cost + pen(seed)), duplicates keep the minimum
    for (int32_t i = tid; i < A.n_seeds; i += BLOCK) {
        const int64_t gs = A.seeds[i];
        const int64_t r = to_roi(A, gs);
        if (r < 0 || r >= (int64_t)R || !allowed_node(A, gs)) continue;
        float c = A.seed_cost ? A.seed_cost[i] : 0.0f;
        if (A.pen && A.seed_pen) c = __fadd_rn(c, A.pen[r]);
        gkey_min(keys + r, kpack(c, 0u, ORX_ROOT_SLOT));
    }
    __syncthreads();
    // distinct targets still to settle
    {
        uint32_t cnt = 0;
        for (uint32_t i = tid; i < words; i += BLOCK) cnt += __popc(bmt[i]);
        for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
        if (lane_id() == 0 && cnt) atomicAdd(&ctl.tg_left, cnt);
    }
    // initial threshold from the smallest seed
    float lmin = INFINITY;
    for (int32_t i = tid; i < A.n_seeds; i += BLOCK) {
        const int64_t r = to_roi(A, A.seeds[i]);
        if (r < 0 || r >= (int64_t)R || !allowed_node(A, A.seeds[i])) continue;
        lmin = fminf(lmin, kdist(gkey_load(keys + r)));
    }
    lmin = wmin(lmin);
    if (tid == 0) ctl.minbits = 0x7F800000u;
    __syncthreads();
    if (lane_id() == 0 && lmin < INFINITY) atomicMin(&ctl.minbits, __float_as_uint(lmin));
    __syncthreads();
    if (ctl.minbits == 0x7F800000u) {
        if (tid == 0) A.best[0] = -1;
        return;
    }
    float T;
    {
        const float mn = __uint_as_float(ctl.minbits);
        T = __fadd_rn(mn, A.delta);
        if (!(T > mn)) T = nextafterf(mn, INFINITY);
    }
    // seed queues: below T -> round 0 near queue, else far (dedup by bitmaps)
    for (int32_t base = 0; base < A.n_seeds; base += BLOCK) {
        const int32_t i = base + tid;
        bool to_near = false, to_far = false;
        uint32_t r = 0;
        if (i < A.n_seeds) {
            const int64_t rr = to_roi(A, A.seeds[i]);
            if (rr >= 0 && rr < (int64_t)R && allowed_node(A, A.seeds[i])) {
                r = (uint32_t)rr;
                const float d = kdist(gkey_load(keys + r));
                if (d < T) to_near = gbit_set(bmn0, r) == 0;
                else to_far = gbit_set(bmf, r) == 0;
            }
        }
        int at = wave_append(to_near, &ctl.qcnt[0]);
        if (to_near) qA[at] = r;
        at = wave_append(to_far, &ctl.fcnt);
        if (to_far) fA[at] = r;
    }
    __syncthreads();

    uint32_t rnd = 0;
    for (;;) {
        for (;;) {
            const uint32_t n = ctl.qcnt[rnd % 3];
            if (n == 0) break;
            uint32_t* cur = (rnd & 1) ? qB : qA;
            uint32_t* nxt = (rnd & 1) ? qA : qB;
            uint32_t* bmc = (rnd & 1) ? bmn1 : bmn0;
            uint32_t* bmx = (rnd & 1) ? bmn0 : bmn1;
            uint32_t* cnx = &ctl.qcnt[(rnd + 1) % 3];
            if (tid == 0) ctl.qcnt[(rnd + 2) % 3] = 0;
            for (uint32_t base = 0; base < n; base += BLOCK) {
                const uint32_t i = base + tid;
                bool active = i < n;
                uint32_t u = 0;
                uint64_t ku = ORX_KEY_INF;
                int32_t e0 = 0, e1 = 0;
                if (active) {
                    u = cur[i];
                    gbit_clear(bmc, u);
                    if (gbit_test(bmt, u)) active = false;  // targets are sinks
                    else {
                        ku = gkey_load(keys + u);
                        const int64_t gu = A.roi_nodes ? A.roi_nodes[u] : u;
                        e0 = A.g.indptr[gu];
                        e1 = A.g.indptr[gu + 1];
                    }
                }
                const float du = kdist(ku);
                const uint32_t hu = khops(ku);
                // uniform trip count across the wave for the ballots
                int32_t deg = active ? e1 - e0 : 0;
                int32_t maxdeg = deg;
                for (int o = 32; o > 0; o >>= 1) maxdeg = max(maxdeg, __shfl_xor(maxdeg, o));
                for (int32_t j = 0; j < maxdeg; ++j) {
                    bool push_near = false, push_far = false;
                    uint32_t v = 0;
                    if (j < deg) {
                        const int32_t e = e0 + j;
                        const int64_t gv = A.g.indices[e];
                        const int64_t rv = to_roi(A, gv);
                        if (rv >= 0 && rv < (int64_t)R && allowed_node(A, gv)) {
                            float w = A.costs[e];
                            if (A.pen) w = __fadd_rn(w, A.pen[rv]);
                            if (!isinf(w)) {
                                v = (uint32_t)rv;
                                const float nd = __fadd_rn(du, w);
                                const uint64_t nk = kpack(nd, hu + 1, (uint32_t)A.g.in_rank[e]);
                                if (nk < gkey_load(keys + v)) {
                                    const uint64_t old = gkey_min(keys + v, nk);
                                    if (nk < old) {
                                        if (nd < T) push_near = gbit_set(bmx, v) == 0;
                                        else push_far = gbit_set(bmf, v) == 0;
                                    }
                                }
                            }
                        }
                    }
                    int at = wave_append(push_near, cnx);
                    if (push_near) nxt[at] = v;
                    at = wave_append(push_far, &ctl.fcnt);
                    if (push_far) fA[at] = v;
                }
            }
            ++rnd;
            __syncthreads();
        }
        // every target below T is final: stop when all targets are
        uint32_t left = 0;
        for (int32_t i = tid; i < A.n_tg; i += BLOCK) {
            const int64_t r = to_roi(A, A.targets[i]);
            if (r < 0 || r >= (int64_t)R) continue;
            const uint64_t k = gkey_load(keys + r);
            left += (k == ORX_KEY_INF || !(kdist(k) < T)) ? 1u : 0u;
        }
        for (int o = 32; o > 0; o >>= 1) left += __shfl_xor(left, o);
        if (tid == 0) ctl.tg_left = 0;
        __syncthreads();
        if (lane_id() == 0 && left) atomicAdd(&ctl.tg_left, left);
        __syncthreads();
        if (A.n_tg > 0 && ctl.tg_left == 0) break;
        const uint32_t nf = ctl.fcnt;
        if (nf == 0) break;
        if (tid == 0) { ctl.fcnt2 = 0; ctl.minbits = 0x7F800000u; }
        __syncthreads();
        float fm = INFINITY;
        for (uint32_t base = 0; base < nf; base += BLOCK) {
            const uint32_t i = base + tid;
            bool keep = false;
            uint32_t v = 0;
            if (i < nf) {
                v = fA[i];
                const float d = kdist(gkey_load(keys + v));
                if (d < T) gbit_clear(bmf, v);
                else { keep = true; fm = fminf(fm, d); }
            }
            const int at = wave_append(keep, &ctl.fcnt2);
            if (keep) fB[at] = v;
        }
        fm = wmin(fm);
        if (lane_id() == 0 && fm < INFINITY) atomicMin(&ctl.minbits, __float_as_uint(fm));
        __syncthreads();
        const uint32_t nf2 = ctl.fcnt2;
        if (nf2 == 0) break;
        const float mn = __uint_as_float(ctl.minbits);
        float Tn = __fadd_rn(mn, A.delta);
        if (!(Tn > mn)) Tn = nextafterf(mn, INFINITY);
        if (tid == 0) ctl.fcnt = 0;
        __syncthreads();
        uint32_t* qr = (rnd & 1) ? qB : qA;
        uint32_t* bmr = (rnd & 1) ? bmn1 : bmn0;
        for (uint32_t base = 0; base < nf2; base += BLOCK) {
            const uint32_t i = base + tid;
            bool to_near = false, stay = false;
            uint32_t v = 0;
            if (i < nf2) {
                v = fB[i];
                const float d = kdist(gkey_load(keys + v));
                if (d < Tn) { to_near = true; gbit_clear(bmf, v); gbit_set(bmr, v); }
                else stay = true;
            }
            int at = wave_append(to_near, &ctl.qcnt[rnd % 3]);
            if (to_near) qr[at] = v;
            at = wave_append(stay, &ctl.fcnt);
            if (stay) fA[at] = v;
        }
        T = Tn;
        __syncthreads();
    }
    __syncthreads();
    // best target: min (fl(d + target_cost), node id)
    if (tid == 0) {
        float bc = INFINITY;
        int32_t bg = -1;
        for (int32_t i = 0; i < A.n_tg; ++i) {
            const int64_t r = to_roi(A, A.targets[i]);
            if (r < 0 || r >= (int64_t)R) continue;
            const uint64_t k = gkey_load(keys + r);
            if (k == ORX_KEY_INF) continue;
            const float t = __fadd_rn(kdist(k), A.tg_cost ? A.tg_cost[i] : 0.0f);
            if (t < bc || (t == bc && A.targets[i] < bg)) { bc = t; bg = A.targets[i]; }
        }
        A.best[0] = bg;
    }
}

// Single-lane parent walk over CSR keys.
__global__ void csr_backtrace_kernel(orx_csr g, const uint64_t* keys, const int32_t* g2r, const int32_t* d_target,
                                     int32_t* path, int32_t cap, int32_t* out_len) {
    if (threadIdx.x != 0) return;
    const int32_t target = d_target[0];
    if (target < 0) { *out_len = ORX_NO_PATH; return; }
    const int64_t rt = g2r ? (int64_t)g2r[target] : (int64_t)target;
    if (rt < 0) { *out_len = ORX_E_ARG; return; }
    const uint64_t kt = keys[rt];
    if (kt == ORX_KEY_INF) { *out_len = ORX_NO_PATH; return; }
    const uint32_t hops = khops(kt);
    if ((int64_t)hops + 1 > cap) { *out_len = ORX_E_OVERFLOW; return; }
    int64_t gv = target;
    int64_t rv = rt;
    for (int32_t i = (int32_t)hops; i >= 0; --i) {
        path[i] = (int32_t)gv;
        const uint32_t sl = kslot(keys[rv]);
        if (i == 0) { *out_len = sl == ORX_ROOT_SLOT ? (int32_t)hops + 1 : ORX_E_CYCLE; return; }
        if (sl == ORX_ROOT_SLOT) { *out_len = ORX_E_CYCLE; return; }
        const int32_t lo = g.in_indptr[gv], hi = g.in_indptr[gv + 1];
        if ((int32_t)sl >= hi - lo) { *out_len = ORX_E_CYCLE; return; }
        gv = g.in_src[lo + (int32_t)sl];
        rv = g2r ? (int64_t)g2r[gv] : gv;
        if (rv < 0) { *out_len = ORX_E_CYCLE; return; }
    }
    *out_len = ORX_E_CYCLE;
}

// ---------------------------------------------------------------------------
// Whole-chip form of the CSR query (ORX_F_GRID, automatic from
// kCsrGridMinR ROI nodes): the same near-far rounds over the same keys, as a
// device-side state machine of grid-wide step launches (the lattice grid
// search's discipline).  Each launch learns its step from the control record
// the launch before it wrote (RELAX, SPLIT1, SPLIT2 or DONE), appends only to
// counters that launch does not read to choose its step, and no workgroup
// ever waits for another.  A drained near queue ends the search when the far
// pile is empty or, with targets, when every target's key is below the
// threshold (final); targets are sinks throughout.
// ---------------------------------------------------------------------------
constexpr int64_t kCsrGridMinR = 65536;
// step launches of one captured graph (even: record parity); the host-thread
// test build, where every launch starts its workgroups' threads, uses fewer
#ifndef ORX_CSR_GRID_STEPS
#define ORX_CSR_GRID_STEPS 64
#endif
constexpr int kCsrGridSteps = ORX_CSR_GRID_STEPS;
enum : uint32_t { CG_RELAX = 0, CG_SPLIT1 = 1, CG_SPLIT2 = 2, CG_DONE = 3 };

struct CgRec {
    uint32_t state;   // RELAX / SPLIT2 / DONE (a RELAX may turn out a SPLIT1 or DONE)
    uint32_t r;       // near-far round: queue and bitmap parity, count slot
    uint32_t phase;   // far-pile counter of the phase: fcnt[phase & 1]
    float T;          // threshold of the phase
};

struct CgCtl {
    CgRec rec[2];
    uint32_t qcnt[3];     // near-queue sizes, by round mod 3
    uint32_t fcnt[2];     // far-pile size, by phase parity
    uint32_t fcnt2;       // live far entries compacted by SPLIT1
    uint32_t done, err;
    uint32_t minf;        // float bits of the far pile's smallest key (distances are >= 0)
    uint32_t pad;
    unsigned long long best;    // (float bits of fl(dist + target cost)) << 32 | target global id
    unsigned long long rounds, phases, items;
};
// control words behind the search scratch (8-byte aligned)
constexpr int64_t kCgCtlWords = (int64_t)((sizeof(CgCtl) + 7) / 8 * 2);

__host__ __device__ inline int64_t cg_ctl_offset(int64_t R) {
    const int64_t words = (R + 31) / 32;
    return (4 * R + 4 * words + 1) / 2 * 2;
}

struct CgArgs {
    CsrArgs a;
    uint32_t *qA, *qB, *fA, *fB, *bmn0, *bmn1, *bmf, *bmt;
    CgCtl* ctl;
};

__device__ __forceinline__ uint64_t agent_key(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t agent_key_min(uint64_t* p, uint64_t v) {
    return __hip_atomic_fetch_min((unsigned long long*)p, (unsigned long long)v, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t agent_bit_set(uint32_t* bm, uint32_t i) {
    const uint32_t b = 1u << (i & 31u);
    return __hip_atomic_fetch_or(bm + (i >> 5), b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & b;
}
__device__ __forceinline__ void agent_bit_clear(uint32_t* bm, uint32_t i) {
    __hip_atomic_fetch_and(bm + (i >> 5), ~(1u << (i & 31u)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool agent_bit_test(const uint32_t* bm, uint32_t i) {
    return (__hip_atomic_load(bm + (i >> 5), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> (i & 31u)) & 1u;
}
__device__ __forceinline__ uint32_t agent_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float thresh_after(float mn, float delta) {
    float T = __fadd_rn(mn, delta);
    return (T > mn) ? T : nextafterf(mn, INFINITY);
}

__global__ __launch_bounds__(256) void cg_init_kernel(CgArgs A) {
    const uint32_t R = (uint32_t)A.a.R, words = (R + 31) / 32;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < R; i += stride) A.a.keys[i] = ORX_KEY_INF;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) {
        A.bmn0[i] = 0u; A.bmn1[i] = 0u; A.bmf[i] = 0u; A.bmt[i] = 0u;
    }
}

// One workgroup: target bits, seed keys fl(seed cost + the seed's node
// penalty) (duplicates keep the minimum), the first threshold from the
// smallest seed, the seeds queued near or far, the control block.
__global__ __launch_bounds__(256) void cg_seed_kernel(CgArgs G) {
    ORX_SHARED(uint32_t, s_min);
    ORX_SHARED(uint32_t, s_near);
    ORX_SHARED(uint32_t, s_far);
    const CsrArgs& A = G.a;
    const uint32_t tid = threadIdx.x;
    const uint32_t R = (uint32_t)A.R;
    for (int32_t i = tid; i < A.n_tg; i += 256) {
        const int64_t r = to_roi(A, A.targets[i]);
        if (r >= 0 && r < (int64_t)R) agent_bit_set(G.bmt, (uint32_t)r);
    }
    for (int32_t i = tid; i < A.n_seeds; i += 256) {
        const int64_t gs = A.seeds[i];
        const int64_t r = to_roi(A, gs);
        if (r < 0 || r >= (int64_t)R || !allowed_node(A, gs)) continue;
        float c = A.seed_cost ? A.seed_cost[i] : 0.0f;
        if (A.pen && A.seed_pen) c = __fadd_rn(c, A.pen[r]);
        agent_key_min(A.keys + r, kpack(c, 0u, ORX_ROOT_SLOT));
    }
    if (tid == 0) { s_min = 0x7F800000u; s_near = 0u; s_far = 0u; }
    __syncthreads();
    float lmin = INFINITY;
    for (int32_t i = tid; i < A.n_seeds; i += 256) {
        const int64_t r = to_roi(A, A.seeds[i]);
        if (r < 0 || r >= (int64_t)R || !allowed_node(A, A.seeds[i])) continue;
        lmin = fminf(lmin, kdist(agent_key(A.keys + r)));
    }
    lmin = wmin(lmin);
    if (lane_id() == 0 && lmin < INFINITY) atomicMin(&s_min, __float_as_uint(lmin));
    __syncthreads();
    CgCtl* c = G.ctl;
    const bool none = s_min == 0x7F800000u;
    const float T = none ? 0.0f : thresh_after(__uint_as_float(s_min), A.delta);
    for (int32_t base = 0; base < A.n_seeds && !none; base += 256) {
        const int32_t i = base + (int32_t)tid;
        bool to_near = false, to_far = false;
        uint32_t r = 0;
        if (i < A.n_seeds) {
            const int64_t rr = to_roi(A, A.seeds[i]);
            if (rr >= 0 && rr < (int64_t)R && allowed_node(A, A.seeds[i])) {
                r = (uint32_t)rr;
                const float d = kdist(agent_key(A.keys + r));
                if (d < T) to_near = agent_bit_set(G.bmn0, r) == 0;
                else to_far = agent_bit_set(G.bmf, r) == 0;
            }
        }
        int at = wave_append(to_near, &s_near);
        if (to_near) G.qA[at] = r;
        at = wave_append(to_far, &s_far);
        if (to_far) G.fA[at] = r;
    }
    __syncthreads();
    if (tid == 0) {
        CgRec r0;
        r0.state = none ? CG_DONE : CG_RELAX;
        r0.r = 0;
        r0.phase = 0;
        r0.T = T;
        c->rec[0] = r0;
        c->rec[1] = r0;
        c->qcnt[0] = s_near; c->qcnt[1] = 0; c->qcnt[2] = 0;
        c->fcnt[0] = s_far; c->fcnt[1] = 0; c->fcnt2 = 0;
        c->done = none ? 1u : 0u;
        c->err = 0;
        c->minf = 0x7F800000u;
        c->pad = 0;
        c->best = ~0ull;
        c->rounds = 0; c->phases = 0; c->items = 0;
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void cg_step_kernel(CgArgs G, uint32_t par) {
    ORX_SHARED(uint32_t, s_left);
    const CsrArgs& A = G.a;
    CgCtl* ctl = G.ctl;
    const uint32_t tid = threadIdx.x;
    const bool lead = blockIdx.x == 0 && tid == 0;
    const uint32_t R = (uint32_t)A.R;
    const CgRec rc = ctl->rec[par];
    const uint32_t qc0 = agent_u32(&ctl->qcnt[0]), qc1 = agent_u32(&ctl->qcnt[1]), qc2 = agent_u32(&ctl->qcnt[2]);
    const uint32_t fc0 = agent_u32(&ctl->fcnt[0]), fc1 = agent_u32(&ctl->fcnt[1]);
    const uint32_t nf2 = agent_u32(&ctl->fcnt2);
    const uint32_t mnb = agent_u32(&ctl->minf);
    const uint32_t r = rc.r, j = rc.phase;
    const float T = rc.T;
    const uint32_t rq = r % 3;
    const uint32_t n = rq == 0 ? qc0 : (rq == 1 ? qc1 : qc2);
    const uint32_t nf = (j & 1) ? fc1 : fc0;
    uint32_t state = rc.state;
    if (state == CG_RELAX && n == 0) {
        // the bucket drained: every key below T is final; stop once every
        // target is (or nothing is left far), else split the far pile.  No
        // key changes during this launch, so every workgroup counts alike.
        bool targets_final = false;
        if (A.n_tg > 0) {
            if (tid == 0) s_left = 0u;
            __syncthreads();
            uint32_t left = 0;
            for (int32_t i = tid; i < A.n_tg; i += BLOCK) {
                const int64_t rt = to_roi(A, A.targets[i]);
                if (rt < 0 || rt >= (int64_t)R) continue;
                const uint64_t k = agent_key(A.keys + rt);
                left += (k == ORX_KEY_INF || !(kdist(k) < T)) ? 1u : 0u;
            }
            for (int o = 32; o > 0; o >>= 1) left += __shfl_xor(left, o);
            if (lane_id() == 0 && left) atomicAdd(&s_left, left);
            __syncthreads();
            targets_final = s_left == 0u;
        }
        state = (targets_final || nf == 0) ? CG_DONE : CG_SPLIT1;
    } else if (state == CG_SPLIT2 && nf2 == 0) {
        state = CG_DONE;
    }
    bool bad = false;
    if ((state == CG_RELAX && n > R) || (state == CG_SPLIT1 && nf > R) || (state == CG_SPLIT2 && nf2 > R)) {
        bad = true;
        state = CG_DONE;
    }
    const float Tn = state == CG_SPLIT2 ? thresh_after(__uint_as_float(mnb), A.delta) : 0.0f;
    if (lead) {
        CgRec nx = rc;
        if (state == CG_DONE) {
            nx.state = CG_DONE;
            ctl->done = 1u;
            if (bad) ctl->err = 1u;
        } else if (state == CG_RELAX) {
            nx.r = r + 1;
            ctl->qcnt[(r + 2) % 3] = 0;
            ctl->fcnt2 = 0;
            ctl->minf = 0x7F800000u;
            ctl->rounds += 1;
            ctl->items += n;
        } else if (state == CG_SPLIT1) {
            nx.state = CG_SPLIT2;
            ctl->phases += 1;
        } else {   // SPLIT2: the next phase starts at the new threshold
            nx.state = CG_RELAX;
            nx.phase = j + 1;
            nx.T = Tn;
            ctl->fcnt[j & 1] = 0;
        }
        ctl->rec[par ^ 1u] = nx;
    }
    if (state == CG_DONE) return;
    const uint32_t stride = gridDim.x * BLOCK;
    const uint32_t first = blockIdx.x * BLOCK;
    if (state == CG_RELAX) {
        uint32_t* cur = (r & 1) ? G.qB : G.qA;
        uint32_t* nxt = (r & 1) ? G.qA : G.qB;
        uint32_t* bmc = (r & 1) ? G.bmn1 : G.bmn0;
        uint32_t* bmx = (r & 1) ? G.bmn0 : G.bmn1;
        uint32_t* cnx = &ctl->qcnt[(r + 1) % 3];
        uint32_t* cfar = &ctl->fcnt[j & 1];
        for (uint32_t base = first; base < n; base += stride) {
            const uint32_t i = base + tid;
            bool active = i < n;
            uint32_t u = 0;
            uint64_t ku = ORX_KEY_INF;
            int32_t e0 = 0, e1 = 0;
            if (active) {
                u = cur[i];
                agent_bit_clear(bmc, u);
                if (agent_bit_test(G.bmt, u)) active = false;   // targets are sinks
                else {
                    ku = agent_key(A.keys + u);
                    const int64_t gu = A.roi_nodes ? A.roi_nodes[u] : u;
                    e0 = A.g.indptr[gu];
                    e1 = A.g.indptr[gu + 1];
                }
            }
            const float du = kdist(ku);
            const uint32_t hu = khops(ku);
            // a wave-uniform trip count for the wave-aggregated appends
            const int32_t deg = active ? e1 - e0 : 0;
            int32_t maxdeg = deg;
            for (int o = 32; o > 0; o >>= 1) maxdeg = max(maxdeg, __shfl_xor(maxdeg, o));
            for (int32_t k = 0; k < maxdeg; ++k) {
                bool push_near = false, push_far = false;
                uint32_t v = 0;
                if (k < deg) {
                    const int32_t e = e0 + k;
                    const int64_t gv = A.g.indices[e];
                    const int64_t rv = to_roi(A, gv);
                    if (rv >= 0 && rv < (int64_t)R && allowed_node(A, gv)) {
                        float w = A.costs[e];
                        if (A.pen) w = __fadd_rn(w, A.pen[rv]);
                        if (!isinf(w)) {
                            v = (uint32_t)rv;
                            const float nd = __fadd_rn(du, w);
                            const uint64_t nk = kpack(nd, hu + 1, (uint32_t)A.g.in_rank[e]);
                            if (nk < agent_key(A.keys + v)) {
                                const uint64_t old = agent_key_min(A.keys + v, nk);
                                if (nk < old) {
                                    if (nd < T) push_near = agent_bit_set(bmx, v) == 0;
                                    else push_far = agent_bit_set(G.bmf, v) == 0;
                                }
                            }
                        }
                    }
                }
                int at = wave_append(push_near, cnx);
                if (push_near) nxt[at] = v;
                at = wave_append(push_far, cfar);
                if (push_far) G.fA[at] = v;
            }
        }
    } else if (state == CG_SPLIT1) {
        float lmin = INFINITY;
        for (uint32_t base = first; base < nf; base += stride) {
            const uint32_t i = base + tid;
            bool keep = false;
            uint32_t v = 0;
            if (i < nf) {
                v = G.fA[i];
                const float d = kdist(agent_key(A.keys + v));
                if (d < T) agent_bit_clear(G.bmf, v);
                else { keep = true; lmin = fminf(lmin, d); }
            }
            const int at = wave_append(keep, &ctl->fcnt2);
            if (keep) G.fB[at] = v;
        }
        lmin = wmin(lmin);
        if (lane_id() == 0 && lmin < INFINITY)
            __hip_atomic_fetch_min(&ctl->minf, __float_as_uint(lmin), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {   // CG_SPLIT2
        uint32_t* qr = (r & 1) ? G.qB : G.qA;
        uint32_t* bmr = (r & 1) ? G.bmn1 : G.bmn0;
        uint32_t* cstay = &ctl->fcnt[(j + 1) & 1];
        for (uint32_t base = first; base < nf2; base += stride) {
            const uint32_t i = base + tid;
            bool to_near = false, stay = false;
            uint32_t v = 0;
            if (i < nf2) {
                v = G.fB[i];
                const float d = kdist(agent_key(A.keys + v));
                if (d < Tn) { to_near = true; agent_bit_clear(G.bmf, v); agent_bit_set(bmr, v); }
                else stay = true;
            }
            int at = wave_append(to_near, &ctl->qcnt[r % 3]);
            if (to_near) qr[at] = v;
            at = wave_append(stay, cstay);
            if (stay) G.fA[at] = v;
        }
    }
}

// The chosen target: the smallest (fl(dist + target cost), node id) over
// reached targets (UP:2605-2618), or -1.
__global__ __launch_bounds__(256) void cg_best_kernel(CgArgs G) {
    const CsrArgs& A = G.a;
    const uint32_t R = (uint32_t)A.R;
    unsigned long long b = ~0ull;
    for (int32_t i = threadIdx.x; i < A.n_tg; i += 256) {
        const int64_t rt = to_roi(A, A.targets[i]);
        if (rt < 0 || rt >= (int64_t)R) continue;
        const uint64_t k = agent_key(A.keys + rt);
        if (k == ORX_KEY_INF) continue;
        const float t = __fadd_rn(kdist(k), A.tg_cost ? A.tg_cost[i] : 0.0f);
        if (isinf(t)) continue;
        const unsigned long long c = ((unsigned long long)__float_as_uint(t) << 32) | (uint32_t)A.targets[i];
        b = c < b ? c : b;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long x = __shfl_xor(b, o);
        b = x < b ? x : b;
    }
    if (lane_id() == 0 && b != ~0ull) atomicMin(&G.ctl->best, b);
}

__global__ void cg_best_write_kernel(const CgCtl* c, int32_t* best) {
    if (threadIdx.x == 0 && blockIdx.x == 0) best[0] = c->best == ~0ull ? -1 : (int32_t)(uint32_t)c->best;
}

// ---------------------------------------------------------------------------
// Parallel parent walk of a CSR result: the parent of every ROI node from
// its key slot (-1 root, -2 a broken slot, -3 unreached), 32-hop jumps by
// five rounds of pointer doubling, then one lane finds the nodes 32 hops
// apart along the chosen target's chain and 256 lanes fill the segments
// between them.  A path of h hops costs about h / 32 + 32 dependent loads
// instead of 3 h.
// ---------------------------------------------------------------------------
constexpr int32_t kJumpLog = 5;
constexpr int32_t kJump = 1 << kJumpLog;

__global__ __launch_bounds__(256) void cg_parent_kernel(orx_csr g, const uint64_t* keys, int64_t R,
                                                        const int32_t* roi_nodes, const int32_t* g2r,
                                                        int32_t* P0, int32_t* J0) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = keys[r];
        int32_t p = -3;
        if (k != ORX_KEY_INF) {
            const uint32_t sl = kslot(k);
            if (sl == ORX_ROOT_SLOT) {
                p = -1;
            } else {
                const int64_t gv = roi_nodes ? roi_nodes[r] : r;
                const int32_t lo = g.in_indptr[gv], hi = g.in_indptr[gv + 1];
                p = -2;
                if ((int32_t)sl < hi - lo) {
                    const int64_t gp = g.in_src[lo + (int32_t)sl];
                    const int64_t rp = g2r ? (int64_t)g2r[gp] : gp;
                    if (rp >= 0 && rp < R) p = (int32_t)rp;
                }
            }
        }
        P0[r] = p;
        J0[r] = p;
    }
}

__global__ __launch_bounds__(256) void cg_jump_kernel(const int32_t* Jin, int32_t* Jout, int64_t R) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t a = Jin[r];
        Jout[r] = a < 0 ? a : Jin[a];
    }
}

__global__ __launch_bounds__(256) void cg_path_kernel(const uint64_t* keys, int64_t R, const int32_t* roi_nodes,
                                                      const int32_t* g2r, const int32_t* d_target,
                                                      const int32_t* P0, const int32_t* J, int32_t* path,
                                                      int32_t cap, int32_t* out_len) {
    struct Anchors { int32_t v[256]; };
    ORX_SHARED(int32_t, s_err);
    ORX_SHARED(int32_t, s_next);
    ORX_SHARED(Anchors, an);
    int32_t* anch = an.v;
    const int32_t tid = (int32_t)threadIdx.x;
    const int32_t target = d_target[0];
    int32_t status = 0;
    int64_t rt = -1;
    uint32_t H = 0;
    if (target < 0) status = ORX_NO_PATH;
    else {
        rt = g2r ? (int64_t)g2r[target] : (int64_t)target;
        if (rt < 0 || rt >= R) status = ORX_E_ARG;
        else if (keys[rt] == ORX_KEY_INF) status = ORX_NO_PATH;
        else {
            H = khops(keys[rt]);
            if ((int64_t)H + 1 > cap) status = ORX_E_OVERFLOW;
        }
    }
    if (status != 0) {
        if (tid == 0) *out_len = status;
        return;
    }
    if (tid == 0) { s_err = 0; s_next = (int32_t)rt; }
    __syncthreads();
    const int64_t nanch = (int64_t)H / kJump + 1;   // anchors at depths H, H - 32, ... >= 0
    for (int64_t c0 = 0; c0 < nanch; c0 += 256) {
        const int32_t cnt = (int32_t)(nanch - c0 < 256 ? nanch - c0 : 256);
        if (tid == 0) {
            int32_t a = s_next;
            for (int32_t k = 0; k < cnt; ++k) {
                anch[k] = a;
                if (k + 1 < cnt || c0 + cnt < nanch) {
                    a = a >= 0 ? J[a] : -2;
                    if (a < 0) { s_err = ORX_E_CYCLE; break; }
                }
            }
            s_next = a;
        }
        __syncthreads();
        if (s_err == 0 && tid < cnt) {
            int64_t depth = (int64_t)H - (int64_t)kJump * (c0 + tid);
            int32_t v = anch[tid];
            for (int32_t s = 0; s < kJump && depth >= 0; ++s, --depth) {
                if (v < 0 || v >= R) { s_err = ORX_E_CYCLE; break; }
                path[depth] = roi_nodes ? roi_nodes[v] : v;
                const int32_t pv = P0[v];
                if (depth == 0) {
                    if (pv != -1) s_err = ORX_E_CYCLE;   // the chain must end at a root
                } else if (pv < 0) {
                    s_err = ORX_E_CYCLE;
                    break;
                }
                v = pv;
            }
        }
        __syncthreads();
        if (s_err != 0) break;
    }
    if (tid == 0) *out_len = s_err != 0 ? s_err : (int32_t)H + 1;
}

}  // namespace

namespace {

int csr_cus() {
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

// Whole-chip query: init, seed, replays of a captured graph of
// kCsrGridSteps step launches until the device reports DONE (two replays
// queued, so the device does not wait for the host's round trip), then the
// chosen target.  Runs on a private stream after the caller's drained.
int sssp_csr_grid(const CsrArgs& base, int64_t R, uint32_t* scratch, int32_t* d_best, int32_t block,
                  hipStream_t caller) {
    CgArgs G;
    G.a = base;
    const int64_t words = (R + 31) / 32;
    G.qA = scratch; G.qB = G.qA + R; G.fA = G.qB + R; G.fB = G.fA + R;
    G.bmn0 = G.fB + R; G.bmn1 = G.bmn0 + words; G.bmf = G.bmn1 + words; G.bmt = G.bmf + words;
    G.ctl = reinterpret_cast<CgCtl*>(scratch + cg_ctl_offset(R));
    ORX_HIP_CHECK(hipStreamSynchronize(caller));
    hipStream_t ps = nullptr;
    ORX_HIP_CHECK(hipStreamCreateWithFlags(&ps, hipStreamNonBlocking));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* h_done = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = ORX_OK;
    const int fill_grid = (int)std::max<int64_t>(1, std::min<int64_t>(csr_cus() * 8, (R + 255) / 256));
    const int grid = std::max(1, csr_cus()) * (block == 256 ? 2 : 1);
    hipLaunchKernelGGL(cg_init_kernel, dim3(fill_grid), dim3(256), 0, ps, G);
    hipLaunchKernelGGL(cg_seed_kernel, dim3(1), dim3(256), 0, ps, G);
    hipError_t e = hipGetLastError();
    static_assert(kCsrGridSteps % 2 == 0, "record parity must carry over between replays");
    if (e == hipSuccess) e = hipStreamBeginCapture(ps, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        for (int i = 0; i < kCsrGridSteps; ++i) {
            const uint32_t par = (uint32_t)(i & 1);
            switch (block) {
            case 512: hipLaunchKernelGGL(cg_step_kernel<512>, dim3(grid), dim3(512), 0, ps, G, par); break;
            case 1024: hipLaunchKernelGGL(cg_step_kernel<1024>, dim3(grid), dim3(1024), 0, ps, G, par); break;
            default: hipLaunchKernelGGL(cg_step_kernel<256>, dim3(grid), dim3(256), 0, ps, G, par); break;
            }
        }
        e = hipStreamEndCapture(ps, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipHostMalloc((void**)&h_done, 2 * sizeof(uint32_t), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    auto replay = [&](int slot) -> hipError_t {
        hipError_t x = hipGraphLaunch(exec, ps);
        if (x == hipSuccess) x = hipMemcpyAsync(h_done + slot, &G.ctl->done, sizeof(uint32_t), hipMemcpyDeviceToHost, ps);
        if (x == hipSuccess) x = hipEventRecord(ev[slot], ps);
        return x;
    };
    if (e == hipSuccess) e = replay(0);
    if (e == hipSuccess) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int it = 0;; ++it) {
            const int cur = it & 1;
            e = replay(cur ^ 1);
            if (e == hipSuccess) e = hipEventSynchronize(ev[cur]);
            if (e != hipSuccess || h_done[cur] != 0) break;
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300)) {
                rc = orx_host::fail(ORX_E_OVERFLOW, "CSR grid search did not finish within 300 s");
                break;
            }
        }
    }
    uint32_t h_err = 0;
    if (e == hipSuccess && rc == ORX_OK) {
        hipLaunchKernelGGL(cg_best_kernel, dim3(1), dim3(256), 0, ps, G);
        hipLaunchKernelGGL(cg_best_write_kernel, dim3(1), dim3(64), 0, ps, (const CgCtl*)G.ctl, d_best);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&h_err, &G.ctl->err, sizeof(uint32_t), hipMemcpyDeviceToHost, ps);
        if (e == hipSuccess) e = hipStreamSynchronize(ps);
        if (e == hipSuccess && h_err) rc = orx_host::fail(ORX_E_OVERFLOW, "CSR grid search: a queue count overflowed");
    }
    if (e != hipSuccess && rc == ORX_OK) rc = orx_host::check_hip(e, "orx_sssp_csr (grid)");
    (void)hipStreamSynchronize(ps);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    if (h_done) (void)hipHostFree(h_done);
    (void)hipStreamDestroy(ps);
    return rc;
}

}  // namespace

extern "C" {

int orx_sssp_csr(const orx_csr* g, const float* d_costs, int64_t R, const int32_t* d_roi_nodes, const int32_t* d_g2r,
                 const float* d_pen, const uint32_t* d_allowed, int32_t n_seeds, const int32_t* d_seeds,
                 const float* d_seed_cost, int32_t n_targets, const int32_t* d_targets, const float* d_target_cost,
                 uint64_t* d_key, uint32_t* d_scratch, int64_t scratch_words, int32_t* d_best,
                 const orx_route_cfg* cfg, orx_stream_t stream) {
    if (!g || !g->indptr || !g->indices || !g->in_rank || !d_costs || !d_key || !d_scratch || !d_best || !cfg)
        return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (R <= 0 || R >= (1ll << 24)) return orx_host::fail(ORX_E_ARG, "R must be in [1, 2^24)");
    if ((d_roi_nodes == nullptr) != (d_g2r == nullptr))
        return orx_host::fail(ORX_E_ARG, "roi_nodes and g2r must both be given or both NULL");
    if (!d_roi_nodes && R != g->N) return orx_host::fail(ORX_E_ARG, "whole-graph search needs R == N");
    const int64_t words = (R + 31) / 32;
    if (scratch_words < 4 * R + 4 * words) return orx_host::fail(ORX_E_ARG, "scratch too small (4R + 4*ceil(R/32))");
    const bool grid = (cfg->flags & ORX_F_GRID) != 0;
    if (grid && scratch_words < cg_ctl_offset(R) + kCgCtlWords)
        return orx_host::fail(ORX_E_ARG, "the whole-chip form needs scratch of 4R + 4*ceil(R/32) + 64 words");
    if (n_seeds < 1 || !d_seeds) return orx_host::fail(ORX_E_ARG, "no seeds");
    CsrArgs A;
    A.g = *g; A.costs = d_costs; A.R = R; A.roi_nodes = d_roi_nodes; A.g2r = d_g2r; A.pen = d_pen;
    A.allowed = d_allowed; A.n_seeds = n_seeds; A.seeds = d_seeds; A.seed_cost = d_seed_cost;
    A.n_tg = n_targets; A.targets = d_targets; A.tg_cost = d_target_cost; A.keys = d_key;
    A.scratch = d_scratch; A.best = d_best; A.delta = cfg->delta;
    A.seed_pen = (cfg->flags & ORX_F_SEED_PEN) ? 1 : 0;
    const hipStream_t s = (hipStream_t)stream;
    if (cfg->block != 256 && cfg->block != 512 && cfg->block != 1024)
        return orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    if (grid) return sssp_csr_grid(A, R, d_scratch, d_best, cfg->block, s);
    switch (cfg->block) {
    case 256: hipLaunchKernelGGL(csr_sssp_kernel<256>, dim3(1), dim3(256), 0, s, A); break;
    case 512: hipLaunchKernelGGL(csr_sssp_kernel<512>, dim3(1), dim3(512), 0, s, A); break;
    case 1024: hipLaunchKernelGGL(csr_sssp_kernel<1024>, dim3(1), dim3(1024), 0, s, A); break;
    default: return orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    }
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_backtrace_csr_roi(const orx_csr* g, const uint64_t* d_key, int64_t R, const int32_t* d_roi_nodes,
                          const int32_t* d_g2r, const int32_t* d_target, int32_t* d_path, int32_t cap,
                          int32_t* d_len, uint32_t* d_scratch, int64_t scratch_words, orx_stream_t stream) {
    if (!g || !g->in_indptr || !g->in_src || !d_key || !d_target || !d_path || !d_len || !d_scratch || cap < 1)
        return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (R <= 0 || R >= (1ll << 24)) return orx_host::fail(ORX_E_ARG, "R must be in [1, 2^24)");
    if ((d_roi_nodes == nullptr) != (d_g2r == nullptr))
        return orx_host::fail(ORX_E_ARG, "roi_nodes and g2r must both be given or both NULL");
    if (!d_roi_nodes && R != g->N) return orx_host::fail(ORX_E_ARG, "whole-graph walk needs R == N");
    if (scratch_words < 3 * R) return orx_host::fail(ORX_E_ARG, "scratch too small (3R words)");
    const hipStream_t s = (hipStream_t)stream;
    int32_t* P0 = reinterpret_cast<int32_t*>(d_scratch);
    int32_t* Ja = P0 + R;
    int32_t* Jb = Ja + R;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(csr_cus() * 8, (R + 255) / 256));
    hipLaunchKernelGGL(cg_parent_kernel, dim3(grid), dim3(256), 0, s, *g, d_key, R, d_roi_nodes, d_g2r, P0, Ja);
    for (int k = 0; k < kJumpLog; ++k) {
        hipLaunchKernelGGL(cg_jump_kernel, dim3(grid), dim3(256), 0, s, (const int32_t*)Ja, Jb, R);
        std::swap(Ja, Jb);
    }
    hipLaunchKernelGGL(cg_path_kernel, dim3(1), dim3(256), 0, s, d_key, R, d_roi_nodes, d_g2r, d_target,
                       (const int32_t*)P0, (const int32_t*)Ja, d_path, cap, d_len);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_backtrace_csr(const orx_csr* g, const uint64_t* d_key, const int32_t* d_g2r, const int32_t* d_target,
                      int32_t* d_path, int32_t cap, int32_t* d_len, orx_stream_t stream) {
    if (!g || !g->in_indptr || !g->in_src || !d_key || !d_target || !d_path || !d_len || cap < 1)
        return orx_host::fail(ORX_E_ARG, "NULL argument");
    hipLaunchKernelGGL(csr_backtrace_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, *g, d_key, d_g2r,
                       d_target, d_path, cap, d_len);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

}  // extern "C"
