// orx_accountant.hip -- per-arc congestion accounting on the caller's CSR arc
// order: the device form of EdgeAccountant (unified_pathfinder.py:1039-1422)
// and of the via capacity kernels that act on its cost array
// (ViaKernelManager.hard_block_via_edges / apply_via_penalties,
// pathfinder/via_kernels.py:186-344, with the arithmetic of their CPU forms
// hard_block_via_edges_cpu / apply_via_penalties_cpu, via_kernels.py:567-651).
//
// Every array holds E float32 (or int32 count) entries indexed by CSR arc id.
// Each entry point is one or two elementwise or reduction launches on the
// caller's stream; all are HBM-bound streams over E entries (or over the
// index list they are given), coalesced, grid-stride, 256 lanes a workgroup.
//
// The arithmetic repeats the reference's NumPy evaluation exactly: every
// product and sum is one IEEE float32 operation in the reference's order (the
// library is built without contraction), Python-float coefficients enter as
// their float32 roundings (NumPy 2 promotion), and NumPy's maximum / minimum
// NaN rules are kept.
#include "orx_common.hpp"

#include <algorithm>

namespace {

constexpr int kAcctBlock = 256;

// np.maximum(0, x): +0 for x <= 0, NaN stays NaN (first operand kept on ties).
__device__ __forceinline__ float np_max0(float x) { return 0.0f >= x ? 0.0f : x; }
// np.maximum(x, 0) as the in-place floor of clear_path.
__device__ __forceinline__ float np_floor0(float x) { return (x >= 0.0f || isnan(x)) ? x : 0.0f; }
// np.minimum(a, b): NaN in either operand propagates.
__device__ __forceinline__ float np_min(float a, float b) { return (a <= b || isnan(a)) ? a : b; }

__device__ __forceinline__ int64_t gtid() { return (int64_t)blockIdx.x * kAcctBlock + threadIdx.x; }
__device__ __forceinline__ int64_t gstride() { return (int64_t)gridDim.x * kAcctBlock; }

// commit_path / clear_path (UP:1099-1135): present[e] += sign for every listed
// arc (repeats count once each); the canonical count of an arc rises by one
// per listing, or drops by one per listing while it is positive (the dict
// entry is deleted at zero).  Out-of-range ids are skipped and counted.
__global__ __launch_bounds__(kAcctBlock) void acct_apply_kernel(float* present, int32_t* canon, int64_t E,
                                                               const int64_t* idx, int64_t n, int32_t sign,
                                                               unsigned long long* bad) {
    const float d = sign > 0 ? 1.0f : -1.0f;
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const int64_t e = idx[i];
        if (e < 0 || e >= E) {
            if (bad) atomicAdd(bad, 1ull);
            continue;
        }
        atomicAdd(present + e, d);
        if (sign > 0) {
            atomicAdd(canon + e, 1);
        } else {
            int32_t c = __hip_atomic_load(canon + e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            while (c > 0) {
                const int32_t prev = atomicCAS(canon + e, c, c - 1);
                if (prev == c) break;
                c = prev;
            }
        }
    }
}

// np.maximum(present, 0, out=present) over the whole array (UP:1134-1135).
__global__ __launch_bounds__(kAcctBlock) void acct_floor_kernel(float* present, int64_t E) {
    for (int64_t e = gtid(); e < E; e += gstride()) {
        const float p = present[e];
        const float f = np_floor0(p);
        if (__float_as_uint(f) != __float_as_uint(p)) present[e] = f;
    }
}

// Workgroup sum of a double and an unsigned count, added once per workgroup
// into out (float64 atomics).
__device__ __forceinline__ void block_add(double v, unsigned long long c, double* out_sum,
                                          unsigned long long* out_cnt) {
    ORX_SHARED(double, sv[kAcctBlock]);
    ORX_SHARED(unsigned long long, sc[kAcctBlock]);
    sv[threadIdx.x] = v;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int w = kAcctBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sv[threadIdx.x] += sv[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (out_sum && sv[0] != 0.0) atomicAdd(out_sum, sv[0]);
        if (out_cnt && sc[0]) atomicAdd(out_cnt, sc[0]);
    }
}

// compute_overuse's edge term (UP:1137-1167): over = max(0, present - cap),
// restricted to mask[e] != 0 when a mask is given; out_sum += sum(over),
// out_cnt += count_nonzero(over).
__global__ __launch_bounds__(kAcctBlock) void acct_overuse_kernel(const float* present, const float* cap,
                                                                 const uint8_t* mask, int64_t E, double* out_sum,
                                                                 unsigned long long* out_cnt) {
    double s = 0.0;
    unsigned long long c = 0;
    for (int64_t e = gtid(); e < E; e += gstride()) {
        if (mask && !mask[e]) continue;
        const float o = np_max0(present[e] - cap[e]);
        s += (double)o;
        c += (o != 0.0f) ? 1ull : 0ull;
    }
    block_add(s, c, out_sum, out_cnt);
}

// Sums behind cost_balance_ratio (UP:1054-1067) and
// verify_present_matches_canonical (UP:1192-1224): which = 0: history;
// 1: max(0, present_ema - capacity); 2: |present - canonical count|.
__global__ __launch_bounds__(kAcctBlock) void acct_sum_kernel(const float* a, const float* b, const int32_t* canon,
                                                             int64_t E, int32_t which, double* out) {
    double s = 0.0;
    for (int64_t e = gtid(); e < E; e += gstride()) {
        float v;
        if (which == 0) v = a[e];
        else if (which == 1) v = np_max0(a[e] - b[e]);
        else v = fabsf(a[e] - (float)canon[e]);
        s += (double)v;
    }
    block_add(s, 0ull, out, nullptr);
}

// update_present_ema (UP:1291-1299): ema = beta * present + (1 - beta) * ema.
__global__ __launch_bounds__(kAcctBlock) void acct_ema_kernel(const float* present, float* ema, int64_t E, float beta,
                                                             float one_minus_beta) {
    for (int64_t e = gtid(); e < E; e += gstride())
        ema[e] = __fadd_rn(__fmul_rn(beta, present[e]), __fmul_rn(one_minus_beta, ema[e]));
}

// update_history (UP:1226-1289): history = history * decay +
// min(gain * max(0, p - cap), cap_mult * base), p = present or present_ema;
// without base the increment is not capped.
__global__ __launch_bounds__(kAcctBlock) void acct_history_kernel(const float* p, const float* cap,
                                                                 const float* base, float* hist, int64_t E,
                                                                 float gain, float cap_mult, float decay) {
    for (int64_t e = gtid(); e < E; e += gstride()) {
        const float h = __fmul_rn(hist[e], decay);
        float inc = __fmul_rn(gain, np_max0(p[e] - cap[e]));
        if (base) inc = np_min(inc, __fmul_rn(cap_mult, base[e]));
        hist[e] = __fadd_rn(h, inc);
    }
}

struct CostArgs {
    const float* base;
    const float* ema;
    const float* cap;
    const float* hist;
    float* total;
    float* scale_out;            // per-arc present-cost scale, layer-biased form only
    const int32_t* edge_layer;   // NULL: no layer bias
    const float* layer_bias;
    const int32_t* edge_kind;
    const uint8_t* blocked;      // NULL: no hard-blocked arcs
    int64_t E;
    int32_t n_layers;
    float pres_fac, hist_w, via_mult, base_w;
    int32_t jitter;
};

// update_costs (UP:1301-1384):
//   bias  = kind == 0 ? layer_bias[layer] : 1      (1 without layer arrays)
//   total = ((base * via_mult) * base_w) * bias + (pres_fac * bias) * over
//           + hist_w * history,   over = max(0, present_ema - capacity)
// blocked arcs then become +inf, and the jitter (e mod 9973) * 1e-6 is added,
// e taken as NumPy's float32 arange does (the float32 rounding of e).
__global__ __launch_bounds__(kAcctBlock) void acct_costs_kernel(CostArgs A) {
    for (int64_t e = gtid(); e < A.E; e += gstride()) {
        float bias = 1.0f;
        if (A.edge_layer && A.edge_kind[e] == 0) {
            const int32_t l = A.edge_layer[e];
            bias = (l >= 0 && l < A.n_layers) ? A.layer_bias[l] : __uint_as_float(0x7FC00000u);
        }
        const float over = np_max0(A.ema[e] - A.cap[e]);
        const float adj = __fmul_rn(__fmul_rn(__fmul_rn(A.base[e], A.via_mult), A.base_w), bias);
        const float sc = __fmul_rn(A.pres_fac, bias);
        float t = __fadd_rn(__fadd_rn(adj, __fmul_rn(sc, over)), __fmul_rn(A.hist_w, A.hist[e]));
        if (A.blocked && A.blocked[e]) t = __uint_as_float(0x7F800000u);
        if (A.jitter) {
            // fmod of the integer-valued float32 (float)e by 9973, taken exactly
            // in integers.
            const int64_t fe = (int64_t)(float)e;
            t = __fadd_rn(t, __fmul_rn((float)(fe % 9973), 1e-6f));
        }
        A.total[e] = t;
        if (A.scale_out) A.scale_out[e] = sc;
    }
}

// begin_live_present_costs (UP:1386-1401): static = total - s * max(0, ema - cap);
// total = static + s * max(0, present + 1 - cap); s scalar or per arc.
__global__ __launch_bounds__(kAcctBlock) void acct_begin_live_kernel(const float* present, const float* ema,
                                                                    const float* cap, float* total, float* live,
                                                                    int64_t E, float scale, const float* scale_arr) {
    for (int64_t e = gtid(); e < E; e += gstride()) {
        const float s = scale_arr ? scale_arr[e] : scale;
        const float st = __fsub_rn(total[e], __fmul_rn(s, np_max0(ema[e] - cap[e])));
        live[e] = st;
        total[e] = __fadd_rn(st, __fmul_rn(s, np_max0(__fadd_rn(present[e], 1.0f) - cap[e])));
    }
}

// refresh_live_present_costs (UP:1403-1422) on the listed arcs.
__global__ __launch_bounds__(kAcctBlock) void acct_refresh_live_kernel(const float* present, const float* cap,
                                                                      float* total, const float* live, int64_t E,
                                                                      const int64_t* idx, int64_t n, float scale,
                                                                      const float* scale_arr,
                                                                      unsigned long long* bad) {
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const int64_t e = idx[i];
        if (e < 0 || e >= E) {
            if (bad) atomicAdd(bad, 1ull);
            continue;
        }
        const float s = scale_arr ? scale_arr[e] : scale;
        total[e] = __fadd_rn(live[e], __fmul_rn(s, np_max0(__fadd_rn(present[e], 1.0f) - cap[e])));
    }
}

// refresh_from_canonical (UP:1073-1097): present = canonical counts.
__global__ __launch_bounds__(kAcctBlock) void acct_from_canon_kernel(float* present, const int32_t* canon,
                                                                    int64_t E) {
    for (int64_t e = gtid(); e < E; e += gstride()) present[e] = (float)canon[e];
}

struct ViaArgs {
    const int32_t* edge;   // [n] CSR arc ids of via arcs
    const int32_t* xy;     // [n, 2] (x, y) of each arc's column
    const int32_t* z_lo;   // [n]
    const int32_t* z_hi;   // [n]
    int32_t n, Nx, Ny, segZ;
    int64_t E;
    float* total;
    unsigned long long* count;  // [0]: arcs blocked / penalised, [1]: entries out of range
};

__device__ __forceinline__ bool via_entry(const ViaArgs& V, int32_t i, int64_t& col, int64_t& e) {
    const int32_t x = V.xy[2 * i], y = V.xy[2 * i + 1];
    e = V.edge[i];
    if (x < 0 || x >= V.Nx || y < 0 || y >= V.Ny || e < 0 || e >= V.E) {
        atomicAdd(V.count + 1, 1ull);
        return false;
    }
    col = (int64_t)x * V.Ny + y;
    return true;
}

// hard_block_via_edges (via_kernels.py:186-260, 567-608): an arc whose
// column is at capacity, or (column free) one of whose segments z-1,
// z in [z_lo, z_hi), is at capacity, gets cost +inf.
__global__ __launch_bounds__(kAcctBlock) void via_hard_block_kernel(ViaArgs V, const int16_t* col_use,
                                                                   const int16_t* col_cap, const int8_t* seg_use,
                                                                   const int8_t* seg_cap) {
    const int64_t i = gtid();
    if (i >= V.n) return;
    int64_t col, e;
    if (!via_entry(V, (int32_t)i, col, e)) return;
    bool blocked = col_use[col] >= col_cap[col];
    if (!blocked && seg_use) {
        for (int32_t z = V.z_lo[i]; z < V.z_hi[i]; ++z) {
            const int32_t s = z - 1;
            if (s >= 0 && s < V.segZ && seg_use[col * V.segZ + s] >= seg_cap[col * V.segZ + s]) {
                blocked = true;
                break;
            }
        }
    }
    if (blocked) {
        V.total[e] = __uint_as_float(0x7F800000u);
        atomicAdd(V.count, 1ull);
    }
}

// apply_via_penalties (via_kernels.py:262-325, 611-651): penalty =
// col_pres[col] * col_w, plus seg_pres[col, z-1] * seg_w over the spanned
// segments (each product and sum one float32 rounding, in order); a
// positive penalty is added to the arc's cost.
__global__ __launch_bounds__(kAcctBlock) void via_penalty_kernel(ViaArgs V, const float* col_pres,
                                                                const float* seg_pres, float col_w, float seg_w) {
    const int64_t i = gtid();
    if (i >= V.n) return;
    int64_t col, e;
    if (!via_entry(V, (int32_t)i, col, e)) return;
    float pen = __fmul_rn(col_pres[col], col_w);
    if (seg_pres) {
        for (int32_t z = V.z_lo[i]; z < V.z_hi[i]; ++z) {
            const int32_t s = z - 1;
            if (s >= 0 && s < V.segZ) pen = __fadd_rn(pen, __fmul_rn(seg_pres[col * V.segZ + s], seg_w));
        }
    }
    if (pen > 0.0f) {
        atomicAdd(V.total + e, pen);
        atomicAdd(V.count, 1ull);
    }
}

unsigned grid_for(int64_t n) {
    const int64_t g = (n + kAcctBlock - 1) / kAcctBlock;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, 8192));
}

int check_acct(const orx_acct* a) {
    if (!a) return orx_host::fail(ORX_E_ARG, "accountant is NULL");
    if (a->E < 0) return orx_host::fail(ORX_E_ARG, "negative arc count");
    if (!a->present || !a->present_ema || !a->history || !a->capacity || !a->canon)
        return orx_host::fail(ORX_E_ARG, "accountant array is NULL");
    return ORX_OK;
}

}  // namespace

extern "C" {

int orx_acct_apply(const orx_acct* a, const int64_t* d_idx, int64_t n, int32_t sign, int32_t floor_present,
                   unsigned long long* d_bad, orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (sign != 1 && sign != -1) return orx_host::fail(ORX_E_ARG, "sign must be +1 or -1");
    if (n < 0 || (n > 0 && !d_idx)) return orx_host::fail(ORX_E_ARG, "bad index list");
    const hipStream_t s = (hipStream_t)stream;
    if (n > 0)
        hipLaunchKernelGGL(acct_apply_kernel, dim3(grid_for(n)), dim3(kAcctBlock), 0, s, a->present, a->canon, a->E,
                           d_idx, n, sign, d_bad);
    if (floor_present && a->E > 0)
        hipLaunchKernelGGL(acct_floor_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, s, a->present, a->E);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_overuse(const orx_acct* a, const uint8_t* d_mask, double* d_sum, unsigned long long* d_count,
                     orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (!d_sum || !d_count) return orx_host::fail(ORX_E_ARG, "NULL output");
    if (a->E > 0)
        hipLaunchKernelGGL(acct_overuse_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, (hipStream_t)stream,
                           a->present, a->capacity, d_mask, a->E, d_sum, d_count);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_sum(const orx_acct* a, int32_t which, double* d_out, orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (!d_out || which < 0 || which > 2) return orx_host::fail(ORX_E_ARG, "bad sum request");
    const float* x = which == 0 ? a->history : (which == 1 ? a->present_ema : a->present);
    if (a->E > 0)
        hipLaunchKernelGGL(acct_sum_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, (hipStream_t)stream, x,
                           a->capacity, a->canon, a->E, which, d_out);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_present_ema(const orx_acct* a, float beta, float one_minus_beta, orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (a->E > 0)
        hipLaunchKernelGGL(acct_ema_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, (hipStream_t)stream,
                           a->present, a->present_ema, a->E, beta, one_minus_beta);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_history(const orx_acct* a, float gain, const float* d_base, float cap_mult, float decay,
                     int32_t use_raw_present, orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (a->E > 0)
        hipLaunchKernelGGL(acct_history_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, (hipStream_t)stream,
                           use_raw_present ? a->present : a->present_ema, a->capacity, d_base, a->history, a->E,
                           gain, cap_mult, decay);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_costs(const orx_acct* a, const float* d_base, float pres_fac, float hist_w, float via_mult,
                   float base_w, const int32_t* d_edge_layer, const float* d_layer_bias, int32_t n_layers,
                   const int32_t* d_edge_kind, const uint8_t* d_blocked, int32_t jitter, float* d_scale_out,
                   orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (!a->total_cost || !d_base) return orx_host::fail(ORX_E_ARG, "total_cost / base is NULL");
    const bool layered = d_edge_layer || d_layer_bias || d_edge_kind;
    if (layered && (!d_edge_layer || !d_layer_bias || !d_edge_kind || n_layers < 1))
        return orx_host::fail(ORX_E_ARG, "layer bias needs edge_layer, layer_bias and edge_kind");
    if (layered && !d_scale_out) return orx_host::fail(ORX_E_ARG, "layer bias needs a per-arc scale output");
    CostArgs A;
    A.base = d_base; A.ema = a->present_ema; A.cap = a->capacity; A.hist = a->history; A.total = a->total_cost;
    A.scale_out = layered ? d_scale_out : nullptr;
    A.edge_layer = layered ? d_edge_layer : nullptr;
    A.layer_bias = d_layer_bias;
    A.edge_kind = d_edge_kind;
    A.blocked = d_blocked;
    A.E = a->E;
    A.n_layers = n_layers;
    A.pres_fac = pres_fac; A.hist_w = hist_w; A.via_mult = via_mult; A.base_w = base_w;
    A.jitter = jitter;
    if (a->E > 0)
        hipLaunchKernelGGL(acct_costs_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, (hipStream_t)stream, A);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_begin_live(const orx_acct* a, float scale, const float* d_scale, orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (!a->total_cost || !a->live_static) return orx_host::fail(ORX_E_ARG, "total_cost / live_static is NULL");
    if (a->E > 0)
        hipLaunchKernelGGL(acct_begin_live_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, (hipStream_t)stream,
                           a->present, a->present_ema, a->capacity, a->total_cost, a->live_static, a->E, scale,
                           d_scale);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_refresh_live(const orx_acct* a, const int64_t* d_idx, int64_t n, float scale, const float* d_scale,
                          unsigned long long* d_bad, orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (!a->total_cost || !a->live_static) return orx_host::fail(ORX_E_ARG, "total_cost / live_static is NULL");
    if (n < 0 || (n > 0 && !d_idx)) return orx_host::fail(ORX_E_ARG, "bad index list");
    if (n > 0)
        hipLaunchKernelGGL(acct_refresh_live_kernel, dim3(grid_for(n)), dim3(kAcctBlock), 0, (hipStream_t)stream,
                           a->present, a->capacity, a->total_cost, a->live_static, a->E, d_idx, n, scale, d_scale,
                           d_bad);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_acct_from_canonical(const orx_acct* a, orx_stream_t stream) {
    int rc = check_acct(a);
    if (rc) return rc;
    if (a->E > 0)
        hipLaunchKernelGGL(acct_from_canon_kernel, dim3(grid_for(a->E)), dim3(kAcctBlock), 0, (hipStream_t)stream,
                           a->present, a->canon, a->E);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

static int via_args(ViaArgs& V, const int32_t* d_edge, const int32_t* d_xy, const int32_t* d_z_lo,
                    const int32_t* d_z_hi, int32_t n, int32_t Nx, int32_t Ny, int32_t segZ, float* d_total,
                    int64_t E, unsigned long long* d_count) {
    if (n < 0 || Nx < 1 || Ny < 1 || segZ < 0 || E < 0) return orx_host::fail(ORX_E_ARG, "bad via shape");
    if (n > 0 && (!d_edge || !d_xy || !d_z_lo || !d_z_hi || !d_total))
        return orx_host::fail(ORX_E_ARG, "via metadata / total_cost is NULL");
    if (!d_count) return orx_host::fail(ORX_E_ARG, "NULL count output");
    V.edge = d_edge; V.xy = d_xy; V.z_lo = d_z_lo; V.z_hi = d_z_hi;
    V.n = n; V.Nx = Nx; V.Ny = Ny; V.segZ = segZ; V.E = E;
    V.total = d_total; V.count = d_count;
    return ORX_OK;
}

int orx_via_hard_block(const int32_t* d_edge, const int32_t* d_xy, const int32_t* d_z_lo, const int32_t* d_z_hi,
                       int32_t n, int32_t Nx, int32_t Ny, int32_t segZ, const int16_t* d_col_use,
                       const int16_t* d_col_cap, const int8_t* d_seg_use, const int8_t* d_seg_cap, float* d_total,
                       int64_t E, unsigned long long* d_count, orx_stream_t stream) {
    ViaArgs V;
    int rc = via_args(V, d_edge, d_xy, d_z_lo, d_z_hi, n, Nx, Ny, segZ, d_total, E, d_count);
    if (rc) return rc;
    if (n > 0 && (!d_col_use || !d_col_cap)) return orx_host::fail(ORX_E_ARG, "column arrays are NULL");
    if ((d_seg_use == nullptr) != (d_seg_cap == nullptr))
        return orx_host::fail(ORX_E_ARG, "segment use and capacity go together");
    if (n > 0)
        hipLaunchKernelGGL(via_hard_block_kernel, dim3((n + kAcctBlock - 1) / kAcctBlock), dim3(kAcctBlock), 0,
                           (hipStream_t)stream, V, d_col_use, d_col_cap, d_seg_use, d_seg_cap);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_via_penalties(const int32_t* d_edge, const int32_t* d_xy, const int32_t* d_z_lo, const int32_t* d_z_hi,
                      int32_t n, int32_t Nx, int32_t Ny, int32_t segZ, const float* d_col_pres,
                      const float* d_seg_pres, float col_w, float seg_w, float* d_total, int64_t E,
                      unsigned long long* d_count, orx_stream_t stream) {
    ViaArgs V;
    int rc = via_args(V, d_edge, d_xy, d_z_lo, d_z_hi, n, Nx, Ny, segZ, d_total, E, d_count);
    if (rc) return rc;
    if (n > 0 && !d_col_pres) return orx_host::fail(ORX_E_ARG, "column pressure is NULL");
    if (n > 0)
        hipLaunchKernelGGL(via_penalty_kernel, dim3((n + kAcctBlock - 1) / kAcctBlock), dim3(kAcctBlock), 0,
                           (hipStream_t)stream, V, d_col_pres, d_seg_pres, col_w, seg_w);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

}  // extern "C"
