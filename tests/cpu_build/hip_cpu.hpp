// hip_cpu.hpp -- runs the liborx_hip kernel sources on host threads.
//
// Test infrastructure: a CPU build of the device code for bounds checking
// (AddressSanitizer) and for debugging without a GPU.  Execution model:
//   * a launch runs its workgroups one after another;
//   * a workgroup runs blockDim.x OS threads; __syncthreads is a barrier;
//   * wave64 operations (__ballot, __shfl, __shfl_xor, mbcnt) exchange values
//     through a per-wave array between two barriers of that wave, so they
//     require uniform control flow across the 64 lanes, as on the GPU;
//   * workgroup-shared variables (ORX_SHARED) and the dynamic LDS block live
//     in the workgroup's own heap storage; the dynamic block has exactly the
//     launch's size;
//   * workgroups run one after another, or up to HIPCPU_CONCURRENT_WG at a
//     time (environment variable) to expose inter-workgroup races;
//   * device memory is host memory; atomics are std::atomic_ref.
#pragma once
#define ORX_HOST_THREAD_BUILD 1
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <barrier>
#include <cmath>
#include <cstdlib>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __constant__
#define __align__(n) alignas(n)
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(x)
#define __shared__ static
#include <map>

struct alignas(16) float4 { float x, y, z, w; };
struct alignas(8) float2 { float x, y; };
inline float2 make_float2(float a, float b) { return float2{a, b}; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
struct alignas(16) int4 { int x, y, z, w; };
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
struct alignas(16) ulonglong2 { unsigned long long x, y; };
inline ulonglong2 make_ulonglong2(unsigned long long a, unsigned long long b) { return ulonglong2{a, b}; }

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct hipcpu_idx { unsigned x, y, z; };

namespace hipcpu {
inline thread_local hipcpu_idx t_thread{0, 0, 0};
inline hipcpu_idx g_bdim{1, 1, 1};
inline hipcpu_idx g_gdim{1, 1, 1};

struct Block {
    hipcpu_idx idx{0, 0, 0};
    std::unique_ptr<unsigned char[]> dyn;
    std::unique_ptr<std::barrier<>> bar;
    std::vector<std::unique_ptr<std::barrier<>>> wave_bar;
    std::vector<uint64_t> exch;  // 64 slots per wave
    std::mutex shm_mu;
    std::map<const void*, std::unique_ptr<unsigned char[]>> shm;
};
inline thread_local Block* t_blk = nullptr;
#define g_blk t_blk

template <typename T>
T& wg_shared(const void* tag) {
    std::lock_guard<std::mutex> lk(t_blk->shm_mu);
    auto& slot = t_blk->shm[tag];
    if (!slot) {
        slot.reset(new unsigned char[sizeof(T)]);
        memset(slot.get(), 0xA5, sizeof(T));
    }
    return *reinterpret_cast<T*>(slot.get());
}

inline unsigned lane() { return t_thread.x & 63u; }
inline unsigned wave() { return t_thread.x >> 6; }
inline void wave_sync() { g_blk->wave_bar[wave()]->arrive_and_wait(); }
inline unsigned wave_lanes() {
    unsigned n = g_bdim.x - (wave() << 6);
    return n > 64 ? 64 : n;
}
template <typename T>
inline uint64_t to_bits(T v) {
    uint64_t b = 0;
    memcpy(&b, &v, sizeof(T));
    return b;
}
template <typename T>
inline T from_bits(uint64_t b) {
    T v;
    memcpy(&v, &b, sizeof(T));
    return v;
}
template <typename T>
T shfl(T v, unsigned src) {
    uint64_t* ex = g_blk->exch.data() + 64 * wave();
    ex[lane()] = to_bits(v);
    wave_sync();
    T r = from_bits<T>(ex[src & 63u]);
    wave_sync();
    return r;
}
inline uint64_t ballot(bool p) {
    uint64_t* ex = g_blk->exch.data() + 64 * wave();
    ex[lane()] = p ? 1 : 0;
    wave_sync();
    uint64_t m = 0;
    for (unsigned i = 0; i < wave_lanes(); ++i) m |= (ex[i] ? 1ull : 0ull) << i;
    wave_sync();
    return m;
}
inline unsigned mbcnt_lo(unsigned mask, unsigned acc) {
    unsigned l = lane() < 32 ? lane() : 32;
    unsigned m = l == 32 ? mask : (mask & ((1u << l) - 1u));
    return acc + (unsigned)__builtin_popcount(m);
}
inline unsigned mbcnt_hi(unsigned mask, unsigned acc) {
    unsigned l = lane() < 32 ? 0 : lane() - 32;
    unsigned m = mask & ((1u << l) - 1u);
    if (l == 0) m = 0;
    return acc + (unsigned)__builtin_popcount(m);
}

// Worker threads kept across launches (a launch of a 512-lane workgroup
// would otherwise start and join 512 OS threads).  run(n, f) calls f(i) on
// threads 0..n-1 and returns when every call has returned; launches come
// from one host thread, one at a time.  The pool is never destroyed: its
// idle threads end with the process.
class Pool {
public:
    void run(unsigned n, const std::function<void(unsigned)>& f) {
        std::unique_lock<std::mutex> lk(m_);
        while (th_.size() < n) {
            const unsigned i = (unsigned)th_.size();
            th_.emplace_back([this, i]() { work(i); });
            th_.back().detach();
        }
        job_ = &f;
        active_ = n;
        left_ = n;
        ++gen_;
        go_.notify_all();
        done_.wait(lk, [this]() { return left_ == 0; });
        job_ = nullptr;
    }

private:
    void work(unsigned i) {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            go_.wait(lk, [&]() { return gen_ != seen; });
            seen = gen_;
            if (i >= active_) continue;
            const std::function<void(unsigned)>* f = job_;
            lk.unlock();
            (*f)(i);
            lk.lock();
            if (--left_ == 0) done_.notify_one();
        }
    }
    std::mutex m_;
    std::condition_variable go_, done_;
    std::vector<std::thread> th_;
    const std::function<void(unsigned)>* job_ = nullptr;
    uint64_t gen_ = 0;
    unsigned active_ = 0, left_ = 0;
};
inline Pool& pool() {
    static Pool* p = new Pool;
    return *p;
}

template <typename K, typename... A>
void launch(K kernel, dim3 grid, dim3 block, size_t lds, A... args) {
    g_bdim = {block.x, 1, 1};
    g_gdim = {grid.x, 1, 1};
    unsigned conc = 1;
    if (const char* e = getenv("HIPCPU_CONCURRENT_WG")) conc = (unsigned)atoi(e);
    if (conc < 1) conc = 1;
    if (grid.x == 0) return;
    // One pool of nb * blockDim threads runs the launch in steps of nb
    // workgroups; thread (0, 0) builds each step's workgroup state between
    // two pool-wide barriers, so every step starts after the previous one
    // has ended, as with a fresh set of threads per step.
    const unsigned nb = std::min(conc, grid.x);
    std::vector<std::unique_ptr<Block>> blks(nb);
    std::barrier<> step((ptrdiff_t)nb * block.x);
    auto build = [&](unsigned b0) {
        for (unsigned j = 0; j < nb; ++j) {
            blks[j].reset();
            if (b0 + j >= grid.x) continue;
            auto blk = std::make_unique<Block>();
            blk->idx = {b0 + j, 0, 0};
            blk->dyn.reset(new unsigned char[lds ? lds : 1]);
            memset(blk->dyn.get(), 0xA5, lds ? lds : 1);
            blk->bar = std::make_unique<std::barrier<>>((ptrdiff_t)block.x);
            unsigned nw = (block.x + 63) / 64;
            for (unsigned w = 0; w < nw; ++w) {
                unsigned n = block.x - w * 64;
                blk->wave_bar.push_back(std::make_unique<std::barrier<>>((ptrdiff_t)(n > 64 ? 64 : n)));
            }
            blk->exch.assign(64 * nw, 0);
            blks[j] = std::move(blk);
        }
    };
    const std::function<void(unsigned)> body = [&](unsigned i) {
        const unsigned j = i / block.x, t = i % block.x;
        for (unsigned b0 = 0; b0 < grid.x; b0 += nb) {
            if (j == 0 && t == 0) build(b0);
            step.arrive_and_wait();
            if (blks[j]) {
                t_thread = {t, 0, 0};
                t_blk = blks[j].get();
                kernel(args...);
            }
            step.arrive_and_wait();
        }
        t_blk = nullptr;
    };
    pool().run(nb * block.x, body);
}
}  // namespace hipcpu

#define threadIdx (hipcpu::t_thread)
#define blockIdx (hipcpu::t_blk->idx)
#define blockDim (hipcpu::g_bdim)
#define gridDim (hipcpu::g_gdim)
#define ORX_DYNAMIC_LDS(T, name) T* name = reinterpret_cast<T*>(hipcpu::t_blk->dyn.get())
#define ORX_SHARED(T, name) static const char hipcpu_tag_##name = 0; T& name = hipcpu::wg_shared<T>(&hipcpu_tag_##name)

inline void __syncthreads() { hipcpu::g_blk->bar->arrive_and_wait(); }
inline void __threadfence() { std::atomic_thread_fence(std::memory_order_seq_cst); }
inline void __threadfence_block() { std::atomic_thread_fence(std::memory_order_seq_cst); }
template <typename T>
T __shfl(T v, int src) { return hipcpu::shfl(v, (unsigned)src); }
template <typename T>
T __shfl_up(T v, unsigned d) {
    uint64_t* ex = hipcpu::g_blk->exch.data() + 64 * hipcpu::wave();
    ex[hipcpu::lane()] = hipcpu::to_bits(v);
    hipcpu::wave_sync();
    T r = hipcpu::lane() >= d ? hipcpu::from_bits<T>(ex[hipcpu::lane() - d]) : v;
    hipcpu::wave_sync();
    return r;
}
template <typename T>
T __shfl_xor(T v, int m) {
    uint64_t* ex = hipcpu::g_blk->exch.data() + 64 * hipcpu::wave();
    ex[hipcpu::lane()] = hipcpu::to_bits(v);
    hipcpu::wave_sync();
    T r = hipcpu::from_bits<T>(ex[(hipcpu::lane() ^ (unsigned)m) & 63u]);
    hipcpu::wave_sync();
    return r;
}
inline uint64_t __ballot(bool p) { return hipcpu::ballot(p); }
// direct global -> LDS load: lane l writes 16 bytes at lds_wave_base + 16 l
inline void lds_load16(const void* g, void* lds_wave_base) {
    memcpy(static_cast<unsigned char*>(lds_wave_base) + 16 * hipcpu::lane(), g, 16);
}
inline void lds_load_wait() {}
#define __builtin_amdgcn_mbcnt_lo(m, a) hipcpu::mbcnt_lo((m), (a))
#define __builtin_amdgcn_mbcnt_hi(m, a) hipcpu::mbcnt_hi((m), (a))
#define __builtin_amdgcn_s_sleep(n) std::this_thread::yield()
#define __builtin_amdgcn_sched_barrier(n) ((void)0)
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline int __ffs(unsigned v) { return __builtin_ffs((int)v); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline unsigned __umulhi(unsigned a, unsigned b) { return (unsigned)(((uint64_t)a * b) >> 32); }
inline unsigned __float_as_uint(float f) { unsigned u; memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(unsigned u) { float f; memcpy(&f, &u, 4); return f; }
inline long long __double_as_longlong(double d) { long long u; memcpy(&u, &d, 8); return u; }
inline double __longlong_as_double(long long u) { double d; memcpy(&d, &u, 8); return d; }
inline float __fadd_rn(float a, float b) { return a + b; }
inline float __fmul_rn(float a, float b) { return a * b; }
using std::isinf;

#define __ATOMIC_RELAXED_HIP 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 0
#define __HIP_MEMORY_SCOPE_AGENT 0
template <typename T>
T __hip_atomic_load(const T* p, int, int) { return std::atomic_ref<T>(*const_cast<T*>(p)).load(); }
template <typename T>
void __hip_atomic_store(T* p, T v, int, int) { std::atomic_ref<T>(*p).store(v); }
template <typename T>
T __hip_atomic_fetch_min(T* p, T v, int, int) {
    std::atomic_ref<T> a(*p);
    T cur = a.load();
    while (v < cur && !a.compare_exchange_weak(cur, v)) {}
    return cur;
}
template <typename T>
T __hip_atomic_fetch_or(T* p, T v, int, int) { return std::atomic_ref<T>(*p).fetch_or(v); }
template <typename T>
T __hip_atomic_fetch_add(T* p, T v, int, int) { return std::atomic_ref<T>(*p).fetch_add(v); }
template <typename T>
T __hip_atomic_fetch_sub(T* p, T v, int, int) { return std::atomic_ref<T>(*p).fetch_sub(v); }
template <typename T>
T __hip_atomic_exchange(T* p, T v, int, int) { return std::atomic_ref<T>(*p).exchange(v); }
template <typename T>
bool __hip_atomic_compare_exchange_strong(T* p, T* expected, T desired, int, int, int) {
    return std::atomic_ref<T>(*p).compare_exchange_strong(*expected, desired);
}
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)
// the device wall clock: 100 MHz, as hipDeviceAttributeWallClockRate reports
inline uint64_t wall_clock64() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch()).count() / 10u;
}
template <typename T>
T __hip_atomic_fetch_and(T* p, T v, int, int) { return std::atomic_ref<T>(*p).fetch_and(v); }
template <typename T, typename U>
T atomicAdd(T* p, U v) { return std::atomic_ref<T>(*p).fetch_add((T)v); }
template <typename T, typename U>
T atomicMin(T* p, U v) { return __hip_atomic_fetch_min(p, (T)v, 0, 0); }
template <typename T, typename U>
T atomicOr(T* p, U v) { return std::atomic_ref<T>(*p).fetch_or((T)v); }
template <typename T, typename U>
T atomicAnd(T* p, U v) { return std::atomic_ref<T>(*p).fetch_and((T)v); }
template <typename T, typename U, typename V>
T atomicCAS(T* p, U expected, V desired) {
    T cur = (T)expected;
    std::atomic_ref<T>(*p).compare_exchange_strong(cur, (T)desired);
    return cur;
}
template <typename T, typename U>
T atomicMax(T* p, U v) {
    std::atomic_ref<T> a(*p);
    T cur = a.load();
    while ((T)v > cur && !a.compare_exchange_weak(cur, (T)v)) {}
    return cur;
}

// runtime API
typedef int hipError_t;
typedef void* hipStream_t;
#define hipSuccess 0
enum { hipFuncAttributeMaxDynamicSharedMemorySize = 8, hipDeviceAttributeMultiprocessorCount = 16,
       hipDeviceAttributeWallClockRate = 17,
       hipMemcpyDeviceToHost = 2 };
struct hipDeviceProp_t { char gcnArchName[256]; };
inline const char* hipGetErrorString(hipError_t) { return "hipcpu error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) { strcpy(p->gcnArchName, "gfx950:cpu"); return 0; }
inline hipError_t hipDeviceGetAttribute(int* v, int a, int) {
    *v = a == hipDeviceAttributeWallClockRate ? 100000 : 4;
    return hipSuccess;
}
inline hipError_t hipFuncSetAttribute(const void*, int, int) { return hipSuccess; }
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
template <typename K>
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) { *n = 1; return hipSuccess; }
inline hipError_t hipMallocAsync(void** p, size_t b, hipStream_t) { *p = malloc(b); return hipSuccess; }
inline hipError_t hipFreeAsync(void* p, hipStream_t) { free(p); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t b, int, hipStream_t) { memcpy(d, s, b); return 0; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t b, hipStream_t) { memset(d, v, b); return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
// streams run synchronously here, so events are always complete
typedef void* hipEvent_t;
constexpr unsigned hipEventDisableTiming = 2;
constexpr unsigned hipHostMallocDefault = 0;
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (hipEvent_t)0x1; return 0; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return 0; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return 0; }
inline hipError_t hipEventDestroy(hipEvent_t) { return 0; }
inline hipError_t hipHostMalloc(void** p, size_t b, unsigned) { *p = malloc(b); return *p ? 0 : 1; }
inline hipError_t hipHostFree(void* p) { free(p); return 0; }
// stream capture: launches on a capturing stream are recorded, and replayed
// in order by hipGraphLaunch
namespace hipcpu {
inline bool g_capturing = false;
inline std::vector<std::function<void()>>* g_capture = nullptr;
template <typename K, typename... A>
void launch_or_record(K kernel, dim3 grid, dim3 block, size_t lds, A... args) {
    if (g_capturing) {
        g_capture->push_back([=]() { launch(kernel, grid, block, lds, args...); });
        return;
    }
    launch(kernel, grid, block, lds, args...);
}
}  // namespace hipcpu
struct hipcpu_graph { std::vector<std::function<void()>> ops; };
typedef hipcpu_graph* hipGraph_t;
typedef hipcpu_graph* hipGraphExec_t;
enum { hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (hipStream_t)0x1; return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
inline hipError_t hipStreamBeginCapture(hipStream_t, int) {
    hipcpu::g_capturing = true;
    hipcpu::g_capture = new std::vector<std::function<void()>>();
    return 0;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    hipcpu::g_capturing = false;
    *g = new hipcpu_graph{std::move(*hipcpu::g_capture)};
    delete hipcpu::g_capture;
    hipcpu::g_capture = nullptr;
    return 0;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) {
    *e = new hipcpu_graph{g->ops};
    return 0;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto& f : e->ops) f();
    return 0;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return 0; }
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return 0; }
#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
    hipcpu::launch_or_record(kernel, dim3(grid), dim3(block), (size_t)(lds), __VA_ARGS__)
