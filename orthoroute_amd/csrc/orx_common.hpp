// orx_common.hpp -- shared device helpers of liborx_hip (gfx950, wave64).
//
// Canonical search key (identical bit layout to the CPU oracle):
//   key = dist_bits(f32) << 32 | hops << 8 | in_slot
// Non-negative IEEE floats order like their bit patterns, so one unsigned
// 64-bit atomicMin resolves distance, then hop count, then the lowest
// predecessor id (in_slot ranks the predecessor among the node's
// in-neighbours by id).  The parent is recovered from the slot, so there is
// no separate (racy) parent write as in cuda_dijkstra.py:676-681.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/orthoroute_hip.h"

#pragma clang fp contract(off)

// Dynamic LDS of a kernel.  The host-thread test build of the kernels
// (tests/cpu_build) predefines it to per-workgroup host memory.
#ifndef ORX_DYNAMIC_LDS
#define ORX_DYNAMIC_LDS(T, name) extern __shared__ T name[]
#endif
// One workgroup-shared variable of a kernel (same test build: per-workgroup
// host storage, so workgroups can run concurrently there).
#ifndef ORX_SHARED
#define ORX_SHARED(T, name) __shared__ T name
#endif

#define ORX_KEY_INF 0xFFFFFFFFFFFFFFFFull
#define ORX_ROOT_SLOT 0xFFu
#define ORX_HOPS_MAX 0xFFFFFFu

namespace orx {

__device__ __forceinline__ uint64_t kpack(float d, uint32_t h, uint32_t s) {
    return ((uint64_t)__float_as_uint(d) << 32) | ((uint64_t)h << 8) | (uint64_t)s;
}
__device__ __forceinline__ float kdist(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint32_t khops(uint64_t k) { return (uint32_t)(k >> 8) & ORX_HOPS_MAX; }
__device__ __forceinline__ uint32_t kslot(uint64_t k) { return (uint32_t)k & 0xFFu; }

// Unsigned division by a runtime constant (Granlund-Montgomery, 32-bit).
struct FastDiv {
    uint32_t d, m, s;
    __host__ __device__ FastDiv() : d(1), m(0), s(0) {}
    __host__ __device__ explicit FastDiv(uint32_t dv) : d(dv) {
        uint32_t l = 0;
        while (l < 32 && (1ull << l) < dv) ++l;
        s = l;
        m = (uint32_t)((((1ull << 32) * ((1ull << l) - dv)) / dv) + 1);
        if (dv == 1) { m = 0; s = 0; }
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const {
        if (d == 1) return n;
        uint32_t t = __umulhi(n, m);
        return (t + ((n - t) >> 1)) >> (s - 1);
    }
};

// Wave64 lane helpers.
__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Direct global -> LDS load of 16 bytes per lane (gfx950 global_load_lds_dwordx4):
// lane l's bytes land at lds_wave_base + 16 * l, with no register staging.
// lds_wave_base must be the same on every lane of the wave.  Completion is
// awaited with lds_load_wait() before the barrier that publishes the data.
// The host-thread test build provides its own versions of both.
#ifndef ORX_HOST_THREAD_BUILD
__device__ __forceinline__ void lds_load16(const void* g, void* lds_wave_base) {
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)g,
                                     (__attribute__((address_space(3))) void*)lds_wave_base, 16, 0, 0);
}
__device__ __forceinline__ void lds_load_wait() { __builtin_amdgcn_s_waitcnt(0); }
#endif

// Wavefront-aggregated append: every lane of the wave must call this (uniform
// control flow); lanes with want=true get consecutive slots.  One LDS atomic
// per wave, ballot + prefix popcount for the lane offsets.
__device__ __forceinline__ int wave_append(bool want, uint32_t* counter) {
    uint64_t m = __ballot(want);
    if (m == 0) return -1;
    uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return want ? (int)(base + lanes_below(m)) : -1;
}

// Wavefront-aggregated append of a variable number of entries (0..7 per
// lane): the exclusive prefix of the counts comes from three ballot
// bit-slices and mbcnt, and the wave takes its range with one LDS atomic.
// Uniform control flow required.  Returns the lane's first slot (only
// meaningful when cnt > 0).
__device__ __forceinline__ uint32_t wave_append_n(uint32_t cnt, uint32_t* counter) {
    const uint64_t m0 = __ballot(cnt & 1u), m1 = __ballot(cnt & 2u), m2 = __ballot(cnt & 4u);
    const uint32_t total = (uint32_t)__popcll(m0) + 2u * (uint32_t)__popcll(m1) + 4u * (uint32_t)__popcll(m2);
    if (total == 0) return 0;
    const uint32_t before = lanes_below(m0) + 2u * lanes_below(m1) + 4u * lanes_below(m2);
    const uint64_t any = m0 | m1 | m2;
    const uint32_t leader = (uint32_t)__ffsll((long long)any) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, total);
    base = __shfl(base, leader);
    return base + before;
}

// Cost formulas: must match the oracle bit for bit (contract(off) above, and
// the library is built with -ffp-contract=off).
struct Params {
    float pres_fac, base_w, hist_w, node_pres, node_hist_w;
    int32_t edge_cap, node_cap;
    float hist_gain, hist_decay, hist_cap_mult, node_hist_gain;
    float astar_cp, astar_cv, astar_gmax;
};

__device__ __forceinline__ float arc_cost(float b, float h, int32_t occ, const Params& P) {
    int32_t over = occ + 1 - P.edge_cap;
    over = over < 0 ? 0 : over;
    float t1 = __fmul_rn(b, P.base_w);
    float t2 = __fmul_rn(P.hist_w, h);
    float t3 = __fadd_rn(t1, t2);
    float t4 = __fmul_rn(P.pres_fac, (float)over);
    return __fadd_rn(t3, t4);
}
__device__ __forceinline__ float node_cost(float nh, int32_t use, const Params& P) {
    int32_t over = use + 1 - P.node_cap;
    over = over < 0 ? 0 : over;
    float t1 = __fmul_rn(P.node_hist_w, nh);
    float t2 = __fmul_rn(P.node_pres, (float)over);
    return __fadd_rn(t1, t2);
}

struct Lat {
    int32_t X, Y, L;
    uint32_t hmask;
    int64_t XY;
    const float* base;
    float* hist;
    int32_t* occ;
    float* wcost;
    float* node_hist;
    int32_t* node_use;
    int32_t* via_col;
};

// Global lattice neighbour of node u in slot s, -1 if absent.
__device__ __forceinline__ int64_t lat_nbr(const Lat& g, int64_t u, int s) {
    int32_t z = (int32_t)(u / g.XY);
    int64_t rem = u - (int64_t)z * g.XY;
    int32_t y = (int32_t)(rem / g.X);
    int32_t x = (int32_t)(rem - (int64_t)y * g.X);
    bool horiz = (g.hmask >> z) & 1u;
    switch (s) {
    case 0: return z > 0 ? u - g.XY : -1;
    case 1: return horiz ? (x > 0 ? u - 1 : -1) : (y > 0 ? u - g.X : -1);
    case 2: return horiz ? (x < g.X - 1 ? u + 1 : -1) : (y < g.Y - 1 ? u + g.X : -1);
    default: return z < g.L - 1 ? u + g.XY : -1;
    }
}
// Slot of arc u->v for lattice neighbours u, v.
__device__ __forceinline__ int lat_slot(const Lat& g, int64_t u, int64_t v) {
    int64_t zu = u / g.XY, zv = v / g.XY;
    if (zv < zu) return 0;
    if (zv > zu) return 3;
    return v < u ? 1 : 2;
}

}  // namespace orx

// ---------------------------------------------------------------------------
// Host-side error plumbing shared by the C entry points.
// ---------------------------------------------------------------------------
namespace orx_host {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);
int prepare_lattice_kernels();  // raise the dynamic-LDS limit of the search kernels once
int prepare_tile_kernels();     // same for the tile engine
// whole-box query on the tile engine (orx_sssp_box with ORX_F_TILES)
int sssp_box_tiles(const orx_lattice* lat, int32_t x0, int32_t y0, int32_t w, int32_t h, int64_t src, int64_t dst,
                   float* d_dist, int32_t* d_parent, const orx_workspace* ws, const orx_route_cfg* cfg,
                   hipStream_t caller);
}  // namespace orx_host

#define ORX_HIP_CHECK(expr)                                              \
    do {                                                                 \
        hipError_t _e = (expr);                                          \
        if (_e != hipSuccess) return orx_host::check_hip(_e, #expr);     \
    } while (0)
