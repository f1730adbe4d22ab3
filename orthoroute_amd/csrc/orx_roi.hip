// orx_roi.hip -- bidirectional BFS region of interest on the device.
//
// ROIExtractor.extract_roi_bfs (unified_pathfinder.py:2257-2384) for a CSR
// graph: level-synchronous BFS waves from src (mark 1) and dst (mark 2); a
// node reached by both waves is marked 3 and the waves have met.  Levels
// stop at max_depth, at the meeting level + halo, or once more than
// 1.5 * max_nodes nodes are seen.  The ROI is the seen nodes in ascending
// id order, then the must-keep nodes (src, dst, portal seeds) not among
// them; beyond max_nodes, must-keep nodes past the cut are swapped to the
// front positions and the list is cut (UP:2357-2379).  global_to_roi is
// written for the ROI's nodes only.
//
// Marks are 2-bit fields (16 nodes per u32).  A step ORs its mark into the
// field: 0 -> mark is a newly seen node (the next frontier), a field holding
// the other wave's mark (or 3) records a meeting -- exactly the per-node
// transitions of the sequential step, in any order.  The frontiers of a
// level are processed by one workgroup; the sorted node list comes from
// chip-wide count / scan / scatter kernels, which also clear the marks.
#include "orx_common.hpp"
#include "../../include/orthoroute_hip.h"

using namespace orx;

namespace {

constexpr int kBfsBlock = 1024;

struct RoiArgs {
    orx_csr g;
    const uint8_t* blocked;
    int64_t src, dst;
    int32_t max_depth, halo;
    int64_t max_nodes;
    uint32_t* marks;
    int32_t* q;        // [4 * qcap]: src frontier A / B, dst frontier A / B
    int64_t qcap;
    int64_t* out;      // [8]
};

struct WaveSums16 { uint32_t v[16]; };
struct WaveSums4 { uint32_t v[4]; };

__device__ __forceinline__ uint32_t mark_or(uint32_t* marks, int64_t v, uint32_t mark) {
    const uint32_t sh = 2u * (uint32_t)(v & 15);
    const uint32_t old = atomicOr(marks + (v >> 4), mark << sh);
    return (old >> sh) & 3u;
}

__global__ __launch_bounds__(kBfsBlock) void roi_bfs_kernel(RoiArgs A) {
    ORX_SHARED(uint32_t, s_n0);
    ORX_SHARED(uint32_t, s_n1);
    ORX_SHARED(uint32_t, s_next);
    ORX_SHARED(int32_t, s_met);
    ORX_SHARED(int32_t, s_err);
    const uint32_t tid = threadIdx.x;
    const int64_t qc = A.qcap;
    if (tid == 0) {
        // seen[src] = 1, then seen[dst] = 2 (an assignment, as in the reference)
        uint32_t* ws = A.marks + (A.src >> 4);
        *ws = (*ws & ~(3u << (2 * (A.src & 15)))) | (1u << (2 * (A.src & 15)));
        uint32_t* wd = A.marks + (A.dst >> 4);
        *wd = (*wd & ~(3u << (2 * (A.dst & 15)))) | (2u << (2 * (A.dst & 15)));
        A.q[0] = (int32_t)A.src;
        A.q[2 * qc] = (int32_t)A.dst;
        s_n0 = 1;
        s_n1 = 1;
        s_met = 0;
        s_err = 0;
    }
    __threadfence_block();
    __syncthreads();
    int64_t seen = A.src == A.dst ? 1 : 2;
    int32_t depth = 0, meet = -1;
    uint32_t par0 = 0, par1 = 0;   // which half of each wave's queue pair is current
    for (;;) {
        const uint32_t n0 = s_n0, n1 = s_n1;
        if (!(depth < A.max_depth && (n0 > 0 || n1 > 0)) || s_err) break;
        for (int w = 0; w < 2; ++w) {
            const uint32_t mark = w == 0 ? 1u : 2u;
            const uint32_t n = w == 0 ? n0 : n1;
            const uint32_t par = w == 0 ? par0 : par1;
            const int32_t* cur = A.q + (int64_t)(2 * w + par) * qc;
            int32_t* nxt = A.q + (int64_t)(2 * w + (par ^ 1u)) * qc;
            if (tid == 0) s_next = 0;
            __syncthreads();
            for (uint32_t i = tid; i < n; i += kBfsBlock) {
                const int64_t u = cur[i];
                const int32_t s = A.g.indptr[u], e = A.g.indptr[u + 1];
                for (int32_t ei = s; ei < e; ++ei) {
                    if (A.blocked && A.blocked[ei]) continue;
                    const int64_t v = A.g.indices[ei];
                    const uint32_t f = mark_or(A.marks, v, mark);
                    if (f == 0u) {
                        const uint32_t at = atomicAdd(&s_next, 1u);
                        if ((int64_t)at < qc) nxt[at] = (int32_t)v;
                        else s_err = 1;
                    } else if (f != mark) {
                        s_met = 1;
                    }
                }
            }
            __threadfence_block();
            __syncthreads();
            const uint32_t nn = s_next;
            seen += nn;
            if (tid == 0) { if (w == 0) s_n0 = nn; else s_n1 = nn; }
            if (w == 0) par0 ^= 1u; else par1 ^= 1u;
            __syncthreads();
        }
        ++depth;
        if (s_met && meet < 0) meet = depth;
        if (meet >= 0 && depth >= meet + A.halo) break;
        if ((double)seen > (double)A.max_nodes * 1.5) break;
    }
    if (tid == 0) {
        A.out[1] = depth;
        A.out[2] = meet;
        A.out[3] = seen;
        A.out[4] = s_err;
    }
}

// Per-block count of seen nodes (nonzero 2-bit fields), one u32 word per lane.
__global__ __launch_bounds__(256) void roi_count_kernel(const uint32_t* marks, int64_t words, uint32_t* counts) {
    ORX_SHARED(uint32_t, s_sum);
    const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t c = 0;
    if (w < words) {
        const uint32_t m = marks[w];
        c = (uint32_t)__popc((m | (m >> 1)) & 0x55555555u);
    }
    if (threadIdx.x == 0) s_sum = 0;
    __syncthreads();
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane_id() == 0 && c) atomicAdd(&s_sum, c);
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = s_sum;
}

// Exclusive scan of the block counts by one workgroup; out[5] = total.
__global__ __launch_bounds__(1024) void roi_scan_kernel(uint32_t* counts, int64_t blocks, int64_t* out) {
    ORX_SHARED(unsigned long long, s_carry);
    ORX_SHARED(WaveSums16, s_wave);
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int64_t base = 0; base < blocks; base += 1024) {
        const int64_t i = base + threadIdx.x;
        const uint32_t c = i < blocks ? counts[i] : 0u;
        uint32_t incl = c;   // inclusive scan within the wave
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if ((int)lane_id() >= o) incl += t;
        }
        const uint32_t wv = threadIdx.x >> 6;
        if (lane_id() == 63) s_wave.v[wv] = incl;
        __syncthreads();
        uint32_t before = 0;
        for (uint32_t k = 0; k < wv; ++k) before += s_wave.v[k];
        const unsigned long long carry = s_carry;
        if (i < blocks) counts[i] = (uint32_t)(carry + before + incl - c);
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = carry + before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) out[5] = (int64_t)s_carry;
}

// Writes the seen nodes in ascending id order (block base from the scan,
// a block-level prefix of the per-lane counts) and clears the marks.
__global__ __launch_bounds__(256) void roi_scatter_kernel(uint32_t* marks, int64_t words, const uint32_t* base,
                                                          int32_t* roi, int64_t cap) {
    ORX_SHARED(WaveSums4, s_wave);
    const int64_t w = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t m = 0;
    if (w < words) m = marks[w];
    const uint32_t nz = (m | (m >> 1)) & 0x55555555u;
    const uint32_t c = (uint32_t)__popc(nz);
    uint32_t incl = c;
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if ((int)lane_id() >= o) incl += t;
    }
    const uint32_t wv = threadIdx.x >> 6;
    if (lane_id() == 63) s_wave.v[wv] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t k = 0; k < wv; ++k) before += s_wave.v[k];
    int64_t at = (int64_t)base[blockIdx.x] + before + incl - c;
    for (int f = 0; f < 16 && w < words; ++f) {
        if ((nz >> (2 * f)) & 1u) {
            if (at < cap) roi[at] = (int32_t)(w * 16 + f);
            ++at;
        }
    }
    if (w < words && m) marks[w] = 0u;
}

// Must-keep nodes and the cut (one lane).  keep[] = src, dst, then the
// portal seeds not already listed; those missing from the sorted list are
// appended; beyond max_nodes the must-keep positions past the cut are
// swapped to the front positions (UP:2357-2379).  out[0] = ROI size.
__global__ void roi_finalize_kernel(int32_t* roi, int64_t cap, const int32_t* keep, int32_t n_keep,
                                    int64_t max_nodes, int64_t* out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t S = out[5];
    if (S > cap) { out[4] = 2; out[0] = 0; return; }
    auto find = [&](int32_t v, int64_t n) -> int64_t {   // position in the sorted prefix, -1 if absent
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (roi[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        return (lo < n && roi[lo] == v) ? lo : -1;
    };
    int64_t L = S;
    for (int32_t i = 0; i < n_keep; ++i) {
        const int32_t v = keep[i];
        bool present = find(v, S) >= 0;
        for (int64_t j = S; j < L && !present; ++j) present = roi[j] == v;
        if (!present) {
            if (L >= cap) { out[4] = 2; out[0] = 0; return; }
            roi[L++] = v;
        }
    }
    if (L > max_nodes) {
        // positions of must-keep nodes, ascending: sorted prefix first, then the appended tail
        int64_t swap_pos = 0;
        int64_t pos_list[ORX_MAX_TARGETS + 2];
        int32_t np = 0;
        for (int32_t i = 0; i < n_keep && np < ORX_MAX_TARGETS + 2; ++i) {
            const int64_t p = find(keep[i], S);
            if (p < 0) continue;
            bool dup = false;
            for (int32_t j = 0; j < np; ++j) dup |= pos_list[j] == p;
            if (!dup) pos_list[np++] = p;
        }
        for (int32_t a = 1; a < np; ++a)          // insertion sort of the prefix positions
            for (int32_t b = a; b > 0 && pos_list[b - 1] > pos_list[b]; --b) {
                const int64_t t = pos_list[b]; pos_list[b] = pos_list[b - 1]; pos_list[b - 1] = t;
            }
        for (int64_t j = S; j < L && np < ORX_MAX_TARGETS + 2; ++j) pos_list[np++] = j;
        for (int32_t k = 0; k < np; ++k, ++swap_pos) {
            const int64_t kept = pos_list[k];
            if (kept >= max_nodes) {
                const int32_t t = roi[kept]; roi[kept] = roi[swap_pos]; roi[swap_pos] = t;
            }
        }
        L = max_nodes;
    }
    out[0] = L;
}

__global__ __launch_bounds__(256) void roi_g2r_kernel(const int32_t* roi, const int64_t* out, int32_t* g2r,
                                                      int32_t set) {
    const int64_t n = out[0];
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        g2r[roi[i]] = set ? (int32_t)i : -1;
}

__global__ void roi_zero_kernel(int64_t* out) {
    if (threadIdx.x < 8) out[threadIdx.x] = threadIdx.x == 2 ? -1 : 0;
}

}  // namespace

extern "C" {

int orx_roi_bfs(const orx_csr* g, const uint8_t* d_blocked, int64_t src, int64_t dst, int32_t max_depth,
                int32_t halo, int64_t max_nodes, const int32_t* d_keep, int32_t n_keep, uint32_t* d_marks,
                int32_t* d_queue, int64_t qcap, uint32_t* d_blocks, int32_t* d_roi, int64_t cap, int32_t* d_g2r,
                int64_t* d_out, orx_stream_t stream) {
    if (!g || !g->indptr || !g->indices || !d_marks || !d_queue || !d_blocks || !d_roi || !d_g2r || !d_out)
        return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (src < 0 || src >= g->N || dst < 0 || dst >= g->N) return orx_host::fail(ORX_E_ARG, "src / dst outside");
    if (n_keep < 2 || n_keep > ORX_MAX_TARGETS + 2 || !d_keep)
        return orx_host::fail(ORX_E_ARG, "keep lists src, dst and at most ORX_MAX_TARGETS seeds");
    if (qcap < 1 || cap < 1 || max_nodes < 1 || g->N >= (1ll << 31))
        return orx_host::fail(ORX_E_ARG, "bad capacity");
    const hipStream_t s = (hipStream_t)stream;
    const int64_t words = (g->N + 15) / 16;
    const int64_t blocks = (words + 255) / 256;
    RoiArgs A;
    A.g = *g; A.blocked = d_blocked; A.src = src; A.dst = dst; A.max_depth = max_depth; A.halo = halo;
    A.max_nodes = max_nodes; A.marks = d_marks; A.q = d_queue; A.qcap = qcap; A.out = d_out;
    hipLaunchKernelGGL(roi_zero_kernel, dim3(1), dim3(64), 0, s, d_out);
    hipLaunchKernelGGL(roi_bfs_kernel, dim3(1), dim3(kBfsBlock), 0, s, A);
    hipLaunchKernelGGL(roi_count_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_marks, words, d_blocks);
    hipLaunchKernelGGL(roi_scan_kernel, dim3(1), dim3(1024), 0, s, d_blocks, blocks, d_out);
    hipLaunchKernelGGL(roi_scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_marks, words, d_blocks, d_roi,
                       cap);
    hipLaunchKernelGGL(roi_finalize_kernel, dim3(1), dim3(64), 0, s, d_roi, cap, d_keep, n_keep, max_nodes, d_out);
    const unsigned gb = (unsigned)std::min<int64_t>((cap + 255) / 256, 4096);
    hipLaunchKernelGGL(roi_g2r_kernel, dim3(gb), dim3(256), 0, s, d_roi, d_out, d_g2r, 1);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

int orx_roi_release(const int32_t* d_roi, const int64_t* d_out, int64_t cap, int32_t* d_g2r, orx_stream_t stream) {
    if (!d_roi || !d_out || !d_g2r || cap < 1) return orx_host::fail(ORX_E_ARG, "NULL argument");
    const unsigned gb = (unsigned)std::min<int64_t>((cap + 255) / 256, 4096);
    hipLaunchKernelGGL(roi_g2r_kernel, dim3(gb), dim3(256), 0, (hipStream_t)stream, d_roi, d_out, d_g2r, 0);
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

}  // extern "C"
