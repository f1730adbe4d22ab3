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
    if (n_seeds < 1 || !d_seeds) return orx_host::fail(ORX_E_ARG, "no seeds");
    CsrArgs A;
    A.g = *g; A.costs = d_costs; A.R = R; A.roi_nodes = d_roi_nodes; A.g2r = d_g2r; A.pen = d_pen;
    A.allowed = d_allowed; A.n_seeds = n_seeds; A.seeds = d_seeds; A.seed_cost = d_seed_cost;
    A.n_tg = n_targets; A.targets = d_targets; A.tg_cost = d_target_cost; A.keys = d_key;
    A.scratch = d_scratch; A.best = d_best; A.delta = cfg->delta;
    A.seed_pen = (cfg->flags & ORX_F_SEED_PEN) ? 1 : 0;
    const hipStream_t s = (hipStream_t)stream;
    switch (cfg->block) {
    case 256: hipLaunchKernelGGL(csr_sssp_kernel<256>, dim3(1), dim3(256), 0, s, A); break;
    case 512: hipLaunchKernelGGL(csr_sssp_kernel<512>, dim3(1), dim3(512), 0, s, A); break;
    case 1024: hipLaunchKernelGGL(csr_sssp_kernel<1024>, dim3(1), dim3(1024), 0, s, A); break;
    default: return orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    }
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
