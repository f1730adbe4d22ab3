// orx_tiles.hip -- tile-step search engine for gfx950.
//
// A search over an ROI box (a net of a disjoint-ROI wave, or one whole-
// lattice query) is split into tiles of kTX x ty cells spanning every layer.
// One step launch gives each active tile a workgroup, which
//   1. loads the tile's keys and entering weights into LDS, plus the keys of
//      the nodes just outside it (its halo) relaxed across the boundary;
//   2. relaxes the tile to a local fixpoint with line scans: a planar phase
//      scans every planar line of the tile forward and backward (one lane per
//      line, the line in registers), a via phase scans every via column up
//      and down; the phases alternate until one changes nothing (on routing
//      waves at most kRoutePhaseCap phases: a visit still changing keys then
//      stops and its tile is visited again, from scratch, in the next step);
//   3. writes the keys that improved back and activates the neighbour tiles
//      whose halo it changed, for the next step.
// Steps repeat until no tile is active.  Every relaxation is one arc of the
// lattice, fl(d(u) + w(u,v)) with the canonical (distance, hops, in-slot)
// key, so the keys reach the same unique fixpoint as the heap Dijkstra of
// unified_pathfinder.py:2451-2548 (and the per-net search of
// orx_lattice.hip): every path, cost and parent slot is bit-identical.  A
// path of many hops costs about one step per tile it crosses instead of one
// dependent round per hop.
//
// Two launch forms drive the visits.
//
// Step form: no workgroup ever waits for another.  Step k reads the item
// list k & 1 and its size cnt[k % 3] (final once step k - 1 ended), appends
// the next step's items to list (k + 1) & 1 and cnt[(k + 1) % 3], and clears
// cnt[(k + 2) % 3], which no launch in flight reads.  A tile is queued at
// most once per list (a bit per tile and list parity).  Two tiles of one
// step may read each other's boundary keys while they change: any value read
// is an upper bound of the final key, and a tile that changed its boundary
// has queued the neighbour again, so the last visit of every tile sees final
// halo keys.
//
// Queue form (ORX_F_ASYNC): one persistent launch per wave.  Its workgroups
// take tiles from a device ring in ticket order and visit each as soon as it
// is queued, with no step boundary and no per-step launch.  A state word per
// tile (dirty / running / seen) keeps every tile queued at most once and
// never visited by two workgroups at once: a neighbour that changes a halo
// sets dirty; a tile that is idle is pushed by whoever set dirty; a running
// one is pushed again by its own visitor when the visit ends.  Key stores
// are published with a release before the activation that follows them, and
// a visit starts after an acquire of its ring entry, so a visit sees every
// halo key that was written before its tile was queued.  A pending counter
// (tiles queued or running) reaches zero only when the search is done; idle
// workgroups wait for a ring entry or for that, and every wait is bounded by
// a wall-clock deadline after which the launch reports an error.
//
// The per-net body around the search (rip-up, commit, live re-price;
// _route_all, unified_pathfinder.py:7949-8430) runs in one prep and one post
// launch per wave, one workgroup per net; nets of a wave have disjoint boxes,
// so the wave equals the sequential order.
#include "orx_route.hpp"

#include <chrono>
#include <cstdlib>

namespace {

constexpr int32_t kTX = 16;            // tile cells along x
constexpr int32_t kPitch = kTX + 1;    // LDS row pitch in nodes (odd: lanes scanning different rows spread over banks)
constexpr int32_t kLines = 16;         // planar lines per layer (max of tile width and height)
constexpr int32_t kMaxL = 32;
constexpr int32_t kTileBlock = 512;    // lanes of a step workgroup
constexpr uint32_t kDrainVisits = 1u << 24;   // bound of the single-workgroup drain
constexpr int32_t kTileSteps = 48;     // steps of one captured graph (multiple of 6: list and counter parity)
// A routing wave's step lasts as long as its slowest visit, and a visit of
// the search front can run 6-10 phases.  Capped at this many phases, a visit
// that still changes keys stops and its tile is queued for the next step
// with its visited bit cleared (so that visit scans every line and column
// holding a finite key, restoring the local fixpoint).  Same keys, shorter
// steps: 800x800x20 route 49.8 -> 44.3 ms per iteration at 4 phases.
constexpr int32_t kRoutePhaseCap = 4;

// Queue form: tile state bits, ring entry fields (epoch 24 bits | slot 20 |
// tile 20), and the wall-clock bound of one persistent launch.
constexpr uint32_t kTsDirty = 1u, kTsRun = 2u, kTsSeen = 4u;
constexpr uint32_t kRingField = 20;
constexpr uint32_t kRingMax = 1u << kRingField;   // slots and tiles per slot of one launch
constexpr int64_t kRunSeconds = 20;

// Layer class of a launch: the step kernels are instantiated for the layer
// counts of the workloads (6, 8, 12, 16, 20, 24, 32), where a via column has
// a compile-time length, and once for any other count up to 32 (class 0).
// The tile height is a constant of the class, so every LDS offset of a scan
// is an immediate: 16 rows up to 20 layers (16 x 16 x 20 = 5120 nodes), 10
// rows above and in class 0.
__host__ __device__ constexpr int32_t layer_class(int32_t L) {
    return (L == 6 || L == 8 || L == 12 || L == 16 || L == 20 || L == 24 || L == 32) ? L : 0;
}
__host__ __device__ constexpr int32_t class_ty(int32_t LC) { return (LC != 0 && LC <= 20) ? 16 : 10; }
__host__ __device__ inline int32_t tile_ty(int32_t L) { return class_ty(layer_class(L)); }

// Dynamic LDS of a step workgroup for L layers, on the padded node index
// (z * ty + ly) * kPitch + lx:
//   keys  u64;
//   w4    the float4 entering weights (slots 0..3), filled by direct
//         global->LDS loads (the padding column is not loaded);
//   cm/cp the two halo carries of every planar line (z * kLines + line);
//   tg    a target byte per node;
//   mk    the scan marks, bit masks (kMk* offsets): a planar line or via
//         column is marked when a node of it changed since its last scan.
struct TileLds {
    int32_t ty, cells, keys, w4, cm, cp, tg, mk, bytes;
};
// Mark words, in two copies each (buffer b at + b * kMkSet): a row ly of the
// H layers is bit z of mk[kMkHRow + ly], a column lx of the V layers bit z
// of mk[kMkVCol + lx]; via column (lx, ly) is bit lx of mk[kMkViaRow + ly]
// (marked by an H-row scan) or bit ly of mk[kMkViaCol + lx] (by a V-column
// scan).  Phase j reads one copy and sets bits in a copy of the other kind;
// the copy it reads is zeroed in phase j + 1, when no phase reads or sets it.
constexpr int32_t kMkHRow = 0, kMkVCol = 16, kMkViaRow = 64, kMkViaCol = 80, kMkSet = 32, kMkWords = 128;
__host__ __device__ inline TileLds tile_lds(int32_t L) {
    TileLds t;
    t.ty = tile_ty(L);
    t.cells = L * t.ty * kPitch;
    t.keys = 0;
    t.w4 = (t.keys + 8 * t.cells + 15) / 16 * 16;
    t.cm = t.w4 + 16 * t.cells;
    t.cp = t.cm + 8 * L * kLines;
    t.tg = t.cp + 8 * L * kLines;
    t.mk = (t.tg + t.cells + 15) / 16 * 16;
    t.bytes = (t.mk + 4 * kMkWords + 15) / 16 * 16;
    return t;
}

// What a tile visit reads: the lattice geometry and weights, the slot table
// (box and terminals of each net), the workspace.  Kept small: the step
// kernels hold it in scalar registers.
struct TileStep {
    int32_t X, L;
    uint32_t hmask;
    int64_t XY;
    const float* wcost;
    const int32_t* box;        // nets' boxes [4n]
    const int32_t* seed_ptr;   // nets' terminal tables (multi-terminal nets), or nullptr
    const int32_t* tgt_ptr;
    const int32_t* tgt_node;
    const int32_t* slot_net;   // [K] net of each slot; nullptr: one search given by sbox
    const int64_t* slot_off;   // [K] workspace offset of each slot (multiple of 64); nullptr: 0
    int32_t sbox[4];
    uint64_t* keys;            // workspace keys
    uint32_t* item_s[2];       // item lists by parity: slot ...
    uint32_t* item_t[2];       // ... and tile
    uint32_t* marks[2];        // queued bit per tile and list parity; slot s at word slot_off[s] / 32
    uint32_t* seen;            // visited bit per tile, same indexing
    uint32_t* cnt;             // [3] list sizes by step mod 3
    unsigned long long* info;  // [4] tile visits, phases, drain visits, error flag
    int64_t cap;               // entries of each item list
    int32_t flags;
    int32_t phase_cap;         // step form: phases after which a visit stops and queues its tile again (0: none)
};

// Everything the prep / post launches of a wave need besides.
struct TileArgs {
    TileStep S;
    Lat g;
    Params P;
    NetsDev nets;
    int32_t K;
    int64_t ssrc;
    int64_t extent;            // the searches own workspace nodes inside [0, extent)
    unsigned long long* stats;
};

struct SlotGeo {
    int32_t k;
    Box B;
    int64_t off;
    int32_t ntx, nty, ty;
};

__device__ __forceinline__ SlotGeo slot_geo(const TileStep& S, uint32_t s) {
    SlotGeo G;
    G.k = S.slot_net ? S.slot_net[s] : 0;
    G.B = make_box(S.slot_net ? S.box + 4 * (int64_t)G.k : S.sbox, S.L);
    G.off = S.slot_off ? S.slot_off[s] : 0;
    G.ty = tile_ty(S.L);
    G.ntx = (G.B.w + kTX - 1) / kTX;
    G.nty = (G.B.h + G.ty - 1) / G.ty;
    return G;
}

// The key a relaxation over an arc of weight w offers (+inf from an
// unreached node or over an absent arc).  Written with selects, no branches:
// the scans run it once per node and direction.
__device__ __forceinline__ uint64_t relax_key(uint64_t ku, float w, uint32_t slot) {
    const uint32_t hi = (uint32_t)(ku >> 32), lo = (uint32_t)ku;
    const uint32_t bad = (hi == 0xFFFFFFFFu || isinf(w)) ? 0xFFFFFFFFu : 0u;   // unreached, or no arc
    const uint32_t nhi = __float_as_uint(__fadd_rn(__uint_as_float(hi), w)) | bad;
    const uint32_t nlo = (((lo & 0xFFFFFF00u) + 0x100u) | slot) | bad;
    return ((uint64_t)nhi << 32) | nlo;
}

__device__ __forceinline__ uint64_t gload_key(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gstore_key(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t gload_u32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Queue tile t of slot s on list par (at most once per list parity).
__device__ __forceinline__ void queue_tile(const TileStep& S, uint32_t s, int64_t off, uint32_t t, uint32_t par,
                                           uint32_t* cnt) {
    uint32_t* m = S.marks[par] + off / 32 + (t >> 5);
    const uint32_t b = 1u << (t & 31u);
    if (__hip_atomic_fetch_or(m, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & b) return;
    const uint32_t at = atomicAdd(cnt, 1u);
    if ((int64_t)at >= S.cap) {
        atomicOr(S.info + 3, 1ull);   // list overflow: a broken plan; the tile is dropped
        return;
    }
    S.item_s[par][at] = s;
    S.item_t[par][at] = t;
}

// Queue form.  The ring is the first two queue arrays read as u64 entries
// (S.cap of them); the state word of tile t of the slot at workspace offset
// off is item_s[1][off + t]; cnt[0] / cnt[1] / cnt[2] are the ring's head
// and tail tickets and the pending count, cnt[3] the launch epoch.
__device__ __forceinline__ uint64_t* ring_of(const TileStep& S) { return reinterpret_cast<uint64_t*>(S.item_s[0]); }
__device__ __forceinline__ uint32_t* tstate_of(const TileStep& S, int64_t off, uint32_t t) {
    return S.item_s[1] + off + t;
}

// Append tile t of slot s to the ring (one more pending tile).  A ring
// entry of this epoch that was never taken means more live entries than the
// ring holds: reported as an overflow.
__device__ __forceinline__ void ring_push(const TileStep& S, uint32_t epoch, uint32_t s, uint32_t t) {
    __hip_atomic_fetch_add(S.cnt + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t at = __hip_atomic_fetch_add(S.cnt + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t* e = ring_of(S) + (int64_t)(at % (uint64_t)S.cap);
    if ((__hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 40) == epoch) atomicOr(S.info + 3, 64ull);
    __hip_atomic_store(e, ((uint64_t)epoch << 40) | ((uint64_t)s << kRingField) | t, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Ask for a visit of tile t of slot s (its halo changed, or it holds a
// seed).  The release publishes the caller's earlier key stores to the
// visit this request leads to.
__device__ __forceinline__ void ring_activate(const TileStep& S, uint32_t epoch, uint32_t s, int64_t off, uint32_t t) {
    const uint32_t old =
        __hip_atomic_fetch_or(tstate_of(S, off, t), kTsDirty, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if ((old & (kTsDirty | kTsRun)) == 0) ring_push(S, epoch, s, t);
}

// Take the next ticket and wait for its entry; 0 when the search is done (no
// tile queued or running) or the deadline passed (reported as an error).
// The wait polls with relaxed loads and ends with one acquire fence.
__device__ uint64_t ring_pop(const TileStep& S, uint32_t epoch, uint64_t deadline) {
    const uint32_t at = __hip_atomic_fetch_add(S.cnt + 0, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint64_t* e = ring_of(S) + (int64_t)(at % (uint64_t)S.cap);
    for (uint32_t spin = 0;; ++spin) {
        const uint64_t v = __hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((v >> 40) == epoch) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            __hip_atomic_store(e, (uint64_t)0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return v;
        }
        if (__hip_atomic_load(S.cnt + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) return 0;
        if ((spin & 63u) == 63u && wall_clock64() > deadline) {
            atomicOr(S.info + 3, 32ull);
            return 0;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

struct TileShared {
    uint32_t changed[3];
    uint32_t act[4];   // halo of the left / right / lower / upper neighbour changed
    uint32_t first;    // first visit of the tile in this search (its finite keys are unrelaxed seeds)
    unsigned char zl[64];   // the H layers in order, then (from 32) the V layers
};

// The layer lists of TileShared, filled once per launch (they depend only on
// the layer count and the H mask): lane z < L writes layer z at its rank
// among the layers of its kind.  The first barrier of a visit orders the
// writes before any read.
__device__ __forceinline__ void tile_layer_lists(const TileStep& S, int32_t L, TileShared& sh) {
    const int32_t z = (int32_t)threadIdx.x;
    if (z < L) {
        const uint32_t h = (S.hmask >> z) & 1u;
        const uint32_t below = (1u << z) - 1u;
        const int32_t rank = __popc((h ? S.hmask : ~S.hmask) & below);
        sh.zl[(h ? 0 : 32) + rank] = (unsigned char)z;
    }
}

// One relaxation step of a scan: node key k receives the carry c (kept if
// smaller), then offers its own relaxation over weight w to the next node.
//
// The carry's own infinity is tracked in cinf, so whether the kept key is
// +inf ("no path yet": both the carry and the node's key are +inf) comes
// from flags known before the compare instead of a test of the kept key's
// bits after it; the offered key is relax_key's, bit for bit.  The hop
// count steps as (lo | 0xff) + 1: the low byte (the in-slot) carries into
// the hop field, as (lo & ~0xff) + 0x100 does.
__device__ __forceinline__ uint64_t scan_step(uint64_t& k, uint64_t& c, bool& cinf, float w, uint32_t slot) {
    const bool kinf = k == ORX_KEY_INF;
    const bool lt = c < k;
    k = lt ? c : k;
    const bool bad = (cinf && kinf) || isinf(w);
    const uint32_t hi = (uint32_t)(k >> 32), lo = (uint32_t)k;
    const uint32_t nhi = bad ? 0xFFFFFFFFu : __float_as_uint(__fadd_rn(__uint_as_float(hi), w));
    const uint32_t nlo = bad ? 0xFFFFFFFFu : (lo | 0xFFu) + (1u + slot);
    c = ((uint64_t)nhi << 32) | nlo;
    cinf = bad;
    return lt;
}

// One scan direction over a line of compile-time length N at node indices
// b0 + i * ST (i = 0 .. N-1 forward, N-1 .. 0 backward), weights at float
// WO of each node's float4, in-slot SLOT; the carry c enters the first node.
// Chunks of 4 nodes: a chunk's LDS reads go out together (immediate
// offsets), its relaxations run on registers, its keys are written back
// (an unchanged key rewrites itself: cheaper than a branch per node; a node
// belongs to one line or column, so no other lane touches it in the phase);
// the next chunk's reads are issued before the current chunk is relaxed,
// and scheduling barriers keep the chunks from piling up registers.  MT:
// targets (tg bits) have no out-arcs.  Returns the changed mask (bit i =
// node i).
template <int N, int ST, int WO, uint32_t SLOT, bool FWD, bool MT>
__device__ __forceinline__ uint32_t scan_fixed(uint64_t* K, const float* wf, int32_t b0, uint32_t tg, uint64_t c) {
    uint32_t cm = 0;
    bool cinf = c == ORX_KEY_INF;
    constexpr int NC = (N + 3) / 4;
    uint64_t k[2][4];
    float w[2][4];
    // chunk c's reads are issued before chunk c - 1 is relaxed (two register
    // sets), so their LDS latency overlaps that chunk's dependent chain
    auto fetch = [&](int c0, uint64_t* kk, float* ww) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = FWD ? c0 + q : N - 1 - c0 - q;
            if (c0 + q < N) {
                kk[q] = K[b0 + i * ST];
                ww[q] = wf[4 * i * ST + WO];
                if (MT && ((tg >> i) & 1u)) ww[q] = INFINITY;
            }
        }
    };
    fetch(0, k[0], w[0]);
#pragma unroll
    for (int ci = 0; ci < NC; ++ci) {
        const int c0 = 4 * ci;
        uint64_t* kk = k[ci & 1];
        float* ww = w[ci & 1];
        if (ci + 1 < NC) fetch(c0 + 4, k[(ci + 1) & 1], w[(ci + 1) & 1]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = FWD ? c0 + q : N - 1 - c0 - q;
            if (c0 + q < N && scan_step(kk[q], c, cinf, ww[q], SLOT)) cm |= 1u << i;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int i = FWD ? c0 + q : N - 1 - c0 - q;
            if (c0 + q < N) K[b0 + i * ST] = kk[q];
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    return cm;
}

// The same over a line of run-time length m <= 32 (edge tiles, class 0
// columns), one chunk of 4 at a time.
template <int ST, int WO, uint32_t SLOT, bool FWD, bool MT>
__device__ __forceinline__ uint32_t scan_var(uint64_t* K, const float* wf, int32_t b0, int32_t m, uint32_t tg,
                                             uint64_t c) {
    uint32_t cm = 0;
    bool cinf = c == ORX_KEY_INF;
#pragma unroll 1
    for (int32_t c0 = 0; c0 < m; c0 += 4) {
        uint64_t k[4];
        float w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t sq = c0 + q;
            const int32_t i = FWD ? sq : m - 1 - sq;
            const int32_t ic = sq < m ? i : (FWD ? m - 1 : 0);   // clamped: always a node of the line
            k[q] = K[b0 + ic * ST];
            w[q] = wf[4 * ic * ST + WO];
            if (MT && ((tg >> ic) & 1u)) w[q] = INFINITY;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t sq = c0 + q;
            const int32_t i = FWD ? sq : m - 1 - sq;
            if (sq < m && scan_step(k[q], c, cinf, w[q], SLOT)) cm |= 1u << i;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t sq = c0 + q;
            const int32_t i = FWD ? sq : m - 1 - sq;
            if (sq < m) K[b0 + i * ST] = k[q];
        }
    }
    return cm;
}

// Planar line of n nodes (node i at b0 + i * ST: ST = 1 for a row of an H
// layer, kPitch for a column of a V layer), forward from the minus halo
// (slot 2 weights, in-slot 1), then backward from the plus halo (slot 1,
// in-slot 2).  A scan leaves every arc of its line satisfied: a key lowered
// by the backward scan came through its plus neighbour, and relaxing it
// forward again adds two hops.  Full lines (n = NF) take the fixed path.
template <int ST, int NF, bool MT>
__device__ __forceinline__ uint32_t scan_line(uint64_t* K, const float4* W4, int32_t b0, int32_t n, uint32_t tg,
                                              uint64_t cm_in, uint64_t cp_in) {
    const float* wf = reinterpret_cast<const float*>(W4 + b0);
    if (n == NF)
        return scan_fixed<NF, ST, 2, 1u, true, MT>(K, wf, b0, tg, cm_in) |
               scan_fixed<NF, ST, 1, 2u, false, MT>(K, wf, b0, tg, cp_in);
    return scan_var<ST, 2, 1u, true, MT>(K, wf, b0, n, tg, cm_in) |
           scan_var<ST, 1, 2u, false, MT>(K, wf, b0, n, tg, cp_in);
}

// Via column at node index b0 (layer stride ZS), up (slot 3, in-slot 0) then
// down (slot 0, in-slot 3), over the L layers (LC = L, or 0: run-time L).
template <int LC, int ZS, bool MT>
__device__ __forceinline__ uint32_t scan_column(uint64_t* K, const float4* W4, int32_t b0, int32_t L, uint32_t tg) {
    const float* wf = reinterpret_cast<const float*>(W4 + b0);
    if constexpr (LC != 0)
        return scan_fixed<LC, ZS, 3, 0u, true, MT>(K, wf, b0, tg, ORX_KEY_INF) |
               scan_fixed<LC, ZS, 0, 3u, false, MT>(K, wf, b0, tg, ORX_KEY_INF);
    else
        return scan_var<ZS, 3, 0u, true, MT>(K, wf, b0, L, tg, ORX_KEY_INF) |
               scan_var<ZS, 0, 3u, false, MT>(K, wf, b0, L, tg, ORX_KEY_INF);
}

// Key-load and write-back loop trip count per lane.
constexpr int32_t kPerLane = (16 * 16 * 20 + kTileBlock - 1) / kTileBlock;

// The fixpoint phases of a visit.  Planar phase: the lanes of the rows of
// the H layers first, then (from a multiple of 64, so every wave scans one
// orientation) the lanes of the columns of the V layers.  A line is scanned
// when its mark is set (a node of it changed in the via phase before, or it
// holds a seed on the tile's first visit) or, in the first phase, when its
// halo carry improves its end node; the via phase likewise scans the marked
// columns.  A scan that changes nodes marks, with one LDS atomic, the lines
// of the other phase through them, and the neighbour whose halo it is.  The
// marks are double-buffered (h = (j / 2) & 1): a planar phase reads planar
// copy h and sets via copy h, a via phase reads via copy h and sets planar
// copy h ^ 1; each phase zeroes the copy the phase before it read, so no
// mark is cleared bit by bit.  ZL: the H layers in order, then (from 32) the
// V layers.
// Returns whether a key changed.
template <int NT, int LC, bool MT>
__device__ __forceinline__ bool tile_phase(const TileStep& S, int32_t j, int32_t L, int32_t tw,
                                           int32_t th, uint64_t* K, const float4* W4, const uint64_t* CM,
                                           const uint64_t* CP, const unsigned char* TG, uint32_t* MK,
                                           const unsigned char* ZL, TileShared& sh) {
    constexpr int32_t TY = class_ty(LC);
    constexpr int32_t ZS = TY * kPitch;
    const uint32_t tid = threadIdx.x;
    const uint32_t hm = S.hmask & (L < 32 ? (1u << L) - 1u : ~0u);
    uint32_t chg = 0;
    const int32_t h = (j >> 1) & 1;
    if ((j & 1) == 0) {
        // the via copy the via phase before read: free until planar phase j + 2
        if (j >= 2 && (int32_t)tid < kMkSet) MK[kMkViaRow + (h ^ 1) * kMkSet + tid] = 0u;
        uint32_t* const PR = MK + h * kMkSet;         // planar marks read
        uint32_t* const VW = MK + h * kMkSet;         // via marks set (offset by kMkViaRow below)
        const int32_t nh = __popc(hm);
        const int32_t hl = (nh * kLines + 63) / 64 * 64, pl = hl + (L - nh) * kLines;
#pragma unroll 1
        for (int32_t p = (int32_t)tid; p < pl; p += NT) {
            const bool horiz = p < hl;
            const int32_t q = horiz ? p : p - hl, k = q >> 4, l = q & 15;
            const int32_t n = horiz ? tw : th;
            if (k >= (horiz ? nh : L - nh) || l >= (horiz ? th : tw)) continue;
            const int32_t z = ZL[(horiz ? 0 : 32) + k];
            const int32_t li = z * kLines + l;
            const int32_t b0 = horiz ? (z * TY + l) * kPitch : z * ZS + l;
            const int32_t st = horiz ? 1 : kPitch;
            const uint32_t bit = 1u << z;
            const uint64_t cmi = CM[li], cpi = CP[li];
            const bool marked = (PR[(horiz ? kMkHRow : kMkVCol) + l] & bit) != 0;
            const bool need = marked || (j == 0 && (cmi < K[b0] || cpi < K[b0 + (n - 1) * st]));
            if (!need) continue;
            uint32_t tg = 0;
            if (MT)
                for (int32_t i = 0; i < n; ++i) tg |= (uint32_t)TG[b0 + i * st] << i;
            const uint32_t cm = horiz ? scan_line<1, kTX, MT>(K, W4, b0, n, tg, cmi, cpi)
                                      : scan_line<kPitch, TY, MT>(K, W4, b0, n, tg, cmi, cpi);
            if (cm) {
                if (cm & 1u) sh.act[horiz ? 0 : 2] = 1u;
                if ((cm >> (n - 1)) & 1u) sh.act[horiz ? 1 : 3] = 1u;
                // H row l: via columns (lx in cm, l); V column l: (l, ly in cm)
                atomicOr(VW + (horiz ? kMkViaRow : kMkViaCol) + l, cm);
                chg |= cm;
            }
        }
    } else {
        // the planar copy the planar phase before read: free until via phase j + 2
        if ((int32_t)tid < kMkSet) MK[h * kMkSet + tid] = 0u;
        const uint32_t* const VR = MK + h * kMkSet;       // via marks read (offset by kMkViaRow below)
        uint32_t* const PW = MK + (h ^ 1) * kMkSet;       // planar marks set
        const int32_t lx = (int32_t)tid & (kTX - 1), ly = (int32_t)tid >> 4;
        if ((int32_t)tid < kTX * TY && lx < tw && ly < th &&
            (((VR[kMkViaRow + ly] >> lx) | (VR[kMkViaCol + lx] >> ly)) & 1u)) {
            const int32_t b0 = ly * kPitch + lx;
            uint32_t tg = 0;
            if (MT)
                for (int32_t z = 0; z < L; ++z) tg |= (uint32_t)TG[b0 + z * ZS] << z;
            const uint32_t cm = scan_column<LC, ZS, MT>(K, W4, b0, L, tg);
            if (cm) {
                if (lx == 0 && (cm & hm)) sh.act[0] = 1u;
                if (lx == tw - 1 && (cm & hm)) sh.act[1] = 1u;
                if (ly == 0 && (cm & ~hm)) sh.act[2] = 1u;
                if (ly == th - 1 && (cm & ~hm)) sh.act[3] = 1u;
                // changed nodes of H layers lie on row ly, of V layers on column lx
                if (cm & hm) atomicOr(PW + kMkHRow + ly, cm & hm);
                if (cm & ~hm) atomicOr(PW + kMkVCol + lx, cm & ~hm);
                chg = cm;
            }
        }
    }
    return chg != 0;
}

// One visit of tile t of slot s: load, local fixpoint, write-back, activation.
// Every lane of the workgroup calls it; it ends with a barrier.  The global
// loads are issued together (one memory wait per visit): the weights go
// straight into LDS, the keys through registers.  AQ: the queue form (pin,
// pout and cnt_out unused; epoch tags the activations).
template <int NT, int LC, bool AQ>
__device__ void tile_visit(const TileStep& S, uint32_t s, uint32_t t, uint32_t pin, uint32_t pout, uint32_t* cnt_out,
                           uint32_t epoch, unsigned char* lds, TileShared& sh) {
    constexpr int32_t TY = class_ty(LC);
    constexpr int32_t ZS = TY * kPitch;   // node-index stride of one layer
    static_assert(NT * kPerLane >= kTX * 16 * 20 && NT * kPerLane >= kTX * 10 * 32, "the key loop covers a tile");
    static_assert(NT >= kTX * TY, "one lane per via column");
    const uint32_t tid = threadIdx.x;
    const SlotGeo G = slot_geo(S, s);
    const Box& B = G.B;
    const int32_t L = LC ? LC : S.L;
    const int32_t tx = (int32_t)(t % (uint32_t)G.ntx), tyi = (int32_t)(t / (uint32_t)G.ntx);
    const int32_t ax = tx * kTX, ay = tyi * TY;
    const int32_t tw = min(kTX, B.w - ax), th = min(TY, B.h - ay);
    uint64_t* keys = S.keys + G.off;
    const TileLds TL = tile_lds(L);
    uint64_t* K = reinterpret_cast<uint64_t*>(lds + TL.keys);
    float4* W4 = reinterpret_cast<float4*>(lds + TL.w4);
    uint64_t* CM = reinterpret_cast<uint64_t*>(lds + TL.cm);
    uint64_t* CP = reinterpret_cast<uint64_t*>(lds + TL.cp);
    unsigned char* TG = lds + TL.tg;
    uint32_t* MK = reinterpret_cast<uint32_t*>(lds + TL.mk);
    const float4* w4 = reinterpret_cast<const float4*>(S.wcost);
    const int64_t XY = S.XY;
    const int32_t span = kTX * TY * L;
    const bool multi = S.slot_net && S.seed_ptr && S.seed_ptr[G.k + 1] > S.seed_ptr[G.k];
    uint32_t* seen = S.seen + G.off / 32 + (t >> 5);
    const uint32_t tbit = 1u << (t & 31u);

    uint32_t seen_w = 0;
    if (tid == 0) {
        if constexpr (AQ) {
            // taken: running and seen, dirty cleared (a halo change from now
            // on asks for another visit)
            seen_w = (__hip_atomic_exchange(tstate_of(S, G.off, t), kTsRun | kTsSeen, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_AGENT) & kTsSeen)
                         ? tbit
                         : 0u;
        } else {
            __hip_atomic_fetch_and(S.marks[pin] + G.off / 32 + (t >> 5), ~tbit, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            seen_w = __hip_atomic_fetch_or(seen, tbit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        sh.changed[0] = sh.changed[1] = sh.changed[2] = 0;
        sh.act[0] = sh.act[1] = sh.act[2] = sh.act[3] = 0;
    }
    // 1. every global load of the visit at once: the weights straight into
    // LDS (16 bytes per lane, each wave filling 64 consecutive node indices;
    // the padding column is skipped), the keys (lanes walk x fastest:
    // coalesced rows) and the halo of each planar line -- the key of the
    // node beyond each end and the weight of its arc into the line -- into
    // registers
    {
        const int32_t units = L * ZS;
        const int32_t wave = (int32_t)(tid >> 6), lane = (int32_t)(tid & 63);
        for (int32_t u0 = wave * 64; u0 < units; u0 += NT) {
            const int32_t m = u0 + lane, r = m / kPitch, lx = m - r * kPitch, z = r / TY, ly = r - z * TY;
            if (m < units && lx < tw && ly < th)
                lds_load16(w4 + (uint32_t)z * (uint32_t)XY + (uint32_t)((B.y0 + ay + ly) * S.X + (B.x0 + ax + lx)),
                           W4 + u0);
        }
    }
    uint64_t lk[kPerLane];
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
        const int32_t i = (int32_t)tid + u * NT;
        lk[u] = ORX_KEY_INF;
        if (i < span) {
            const int32_t lx = i & (kTX - 1), q = i >> 4, z = q / TY, ly = q - z * TY;
            // ROI indices are below 2^24
            if (lx < tw && ly < th) lk[u] = gload_key(keys + (uint32_t)((z * B.h + (ay + ly)) * B.w + (ax + lx)));
        }
    }
    uint64_t hk[2] = {ORX_KEY_INF, ORX_KEY_INF};
    float hw[2] = {INFINITY, INFINITY};
    if ((int32_t)tid < L * kLines) {
        const int32_t z = (int32_t)tid >> 4, l = (int32_t)tid & 15;
        const bool horiz = (S.hmask >> z) & 1u;
        if (horiz && l < th) {
            const int32_t y = ay + l;
            const int64_t row = ((int64_t)z * B.h + y) * B.w, grow = (int64_t)z * XY + (int64_t)(B.y0 + y) * S.X + B.x0;
            if (ax > 0) {
                hk[0] = gload_key(keys + row + ax - 1);
                hw[0] = S.wcost[4 * (grow + ax - 1) + 2];
            }
            if (ax + tw < B.w) {
                hk[1] = gload_key(keys + row + ax + tw);
                hw[1] = S.wcost[4 * (grow + ax + tw) + 1];
            }
        } else if (!horiz && l < tw) {
            const int32_t x = ax + l;
            const int64_t lay = (int64_t)z * B.h, glay = (int64_t)z * XY + B.x0 + x;
            if (ay > 0) {
                hk[0] = gload_key(keys + (lay + ay - 1) * B.w + x);
                hw[0] = S.wcost[4 * (glay + (int64_t)(B.y0 + ay - 1) * S.X) + 2];
            }
            if (ay + th < B.h) {
                hk[1] = gload_key(keys + (lay + ay + th) * B.w + x);
                hw[1] = S.wcost[4 * (glay + (int64_t)(B.y0 + ay + th) * S.X) + 1];
            }
        }
    }
    if ((int32_t)tid < kMkWords) MK[tid] = 0u;
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
        const int32_t i = (int32_t)tid + u * NT;
        if (i < span) {
            const int32_t li = (i >> 4) * kPitch + (i & (kTX - 1));
            if (multi) TG[li] = 0;
            K[li] = lk[u];
        }
    }
    if ((int32_t)tid < L * kLines) {
        CM[tid] = relax_key(hk[0], hw[0], 1u);
        CP[tid] = relax_key(hk[1], hw[1], 2u);
    }
    if (tid == 0) sh.first = (seen_w & tbit) == 0;
    lds_load_wait();
    __syncthreads();
    // The tile's keys are a fixpoint for the halo of its last visit (no one
    // else writes them), so only lines whose carry improves their end node
    // need a scan -- except on its first visit, when finite keys are seeds
    // not yet relaxed: their line and column are marked.
    if (sh.first) {
#pragma unroll
        for (int u = 0; u < kPerLane; ++u) {
            const int32_t i = (int32_t)tid + u * NT;
            if (i < span && lk[u] != ORX_KEY_INF) {
                const int32_t lx = i & (kTX - 1), q = i >> 4, z = q / TY, ly = q - z * TY;
                if ((S.hmask >> z) & 1u)
                    atomicOr(MK + kMkHRow + ly, 1u << z);
                else
                    atomicOr(MK + kMkVCol + lx, 1u << z);
                atomicOr(MK + kMkViaRow + ly, 1u << lx);
            }
        }
        __syncthreads();
    }
    // 2. targets of a multi-terminal net are sinks: reached, never expanded
    // (UP:2611-2618).  Mark those in the tile; a target on the halo gives
    // its line no carry.
    if (multi) {
        const int32_t t0 = S.tgt_ptr[G.k], nt = S.tgt_ptr[G.k + 1] - t0;
        for (int32_t i = (int32_t)tid; i < nt; i += NT) {
            const int64_t gid = S.tgt_node[t0 + i];
            const int32_t z = (int32_t)(gid / XY);
            const int64_t rem = gid - (int64_t)z * XY;
            const int32_t y = (int32_t)(rem / S.X), x = (int32_t)(rem - (int64_t)y * S.X);
            const int32_t lx = x - B.x0 - ax, ly = y - B.y0 - ay;
            const bool horiz = (S.hmask >> z) & 1u;
            if (lx >= 0 && lx < tw && ly >= 0 && ly < th) {
                TG[(z * TY + ly) * kPitch + lx] = 1;
            } else if (horiz && ly >= 0 && ly < th) {
                if (lx == -1) CM[z * kLines + ly] = ORX_KEY_INF;
                if (lx == tw) CP[z * kLines + ly] = ORX_KEY_INF;
            } else if (!horiz && lx >= 0 && lx < tw) {
                if (ly == -1) CM[z * kLines + lx] = ORX_KEY_INF;
                if (ly == th) CP[z * kLines + lx] = ORX_KEY_INF;
            }
        }
        __syncthreads();
    }
    // 3. local fixpoint: planar and via phases alternate; stop once a phase
    // changes nothing (after both ran): every arc of the tile then holds
    // (the tile's keys were a fixpoint for the old halo after its last
    // visit, and only the carries are new).
    int32_t j = 0;
    bool capped = false;
    for (;; ++j) {
        const bool chg = multi ? tile_phase<NT, LC, true>(S, j, L, tw, th, K, W4, CM, CP, TG, MK, sh.zl, sh)
                               : tile_phase<NT, LC, false>(S, j, L, tw, th, K, W4, CM, CP, TG, MK, sh.zl, sh);
        if (chg) sh.changed[j % 3] = 1u;
        __syncthreads();
        const bool any = sh.changed[j % 3] != 0;
        if (tid == 0) sh.changed[(j + 2) % 3] = 0;
        if (!any && j >= 1) break;
        if (!AQ && S.phase_cap > 0 && j + 1 >= S.phase_cap) {
            capped = true;
            break;
        }
        if (j >= 4 * (kLines + kMaxL) * kLines) {   // far beyond any real tile: a broken state
            if (tid == 0) atomicOr(S.info + 3, 2ull);
            break;
        }
    }
    // 4. activation of the neighbours whose halo changed (lanes 0-3), while
    // every lane writes back what improved (step form); the queue form
    // writes back first and activates after a release of every lane's stores
    auto activate = [&]() {
        if (tid < 4 && sh.act[tid]) {
            const int32_t nx = tx + (tid == 0 ? -1 : (tid == 1 ? 1 : 0));
            const int32_t ny = tyi + (tid == 2 ? -1 : (tid == 3 ? 1 : 0));
            if (nx >= 0 && nx < G.ntx && ny >= 0 && ny < G.nty) {
                if constexpr (AQ)
                    ring_activate(S, epoch, s, G.off, (uint32_t)(ny * G.ntx + nx));
                else
                    queue_tile(S, s, G.off, (uint32_t)(ny * G.ntx + nx), pout, cnt_out);
            }
        }
    };
    if constexpr (!AQ) {
        activate();
        // stopped before its fixpoint: visited again next step as if for the
        // first time (every line and column holding a finite key is scanned)
        if (capped && tid == 0) {
            __hip_atomic_fetch_and(seen, ~tbit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            queue_tile(S, s, G.off, t, pout, cnt_out);
        }
    }
    // write-back of the keys that changed (against the loaded ones, still in
    // registers); the indices are recomputed from an opaque copy of the lane
    // id, so the compiler does not keep the load loop's indices alive
    // through the fixpoint
    uint32_t wtid = tid;
    asm volatile("" : "+v"(wtid));
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
        const int32_t i = (int32_t)wtid + u * NT;
        if (i < span) {
            const int32_t lx = i & (kTX - 1), q = i >> 4;
            const uint64_t kn = K[q * kPitch + lx];
            if (kn != lk[u]) {
                const int32_t z = q / TY, ly = q - z * TY;
                gstore_key(keys + (uint32_t)((z * B.h + (ay + ly)) * B.w + (ax + lx)), kn);
            }
        }
    }
    if constexpr (AQ) {
        // the barrier orders every lane's key stores before the activations'
        // release (agent scope) in lanes 0-3
        __syncthreads();
        activate();
        // every push of lanes 0-3 (pending raised) before lane 0 drops this
        // run's count: pending may not touch zero while a push is under way
        __syncthreads();
        if (tid == 0) {
            // the visit ends: idle again, unless a halo changed meanwhile --
            // then this tile goes back on the ring (pending is raised by the
            // push before this run's own count is dropped)
            uint32_t* st = tstate_of(S, G.off, t);
            uint32_t expect = kTsRun | kTsSeen;
            if (!__hip_atomic_compare_exchange_strong(st, &expect, kTsSeen, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_exchange(st, kTsDirty | kTsSeen, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                ring_push(S, epoch, s, t);
            }
            __hip_atomic_fetch_sub(S.cnt + 2, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (tid == 0 && (S.flags & ORX_F_STATS)) {
        atomicAdd(S.info + 0, 1ull);
        atomicAdd(S.info + 1, (unsigned long long)(j + 1));
    }
    __syncthreads();
}

// Keys of [0, n) to +inf, the step counters / info cleared and the launch
// epoch of the queue form advanced (1 .. 2^24 - 1).  Queue form: the ring's
// first n entries (the ring of a launch over n workspace nodes) cleared too,
// so no entry left by another engine's use of the queue arrays can pass for
// one of this launch.
__global__ __launch_bounds__(256) void tile_fill_kernel(TileStep S, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    ulonglong2* k2 = reinterpret_cast<ulonglong2*>(S.keys);
    const ulonglong2 inf2 = make_ulonglong2(ORX_KEY_INF, ORX_KEY_INF);
    for (int64_t i = t; i < n / 2; i += stride) k2[i] = inf2;
    if (S.flags & ORX_F_ASYNC) {
        ulonglong2* r2 = reinterpret_cast<ulonglong2*>(ring_of(S));
        const ulonglong2 zero2 = make_ulonglong2(0ull, 0ull);
        for (int64_t i = t; i < n / 2; i += stride) r2[i] = zero2;
        if (t == 0 && (n & 1)) ring_of(S)[n - 1] = 0ull;
    }
    if (t == 0) {
        if (n & 1) S.keys[n - 1] = ORX_KEY_INF;
        S.cnt[0] = S.cnt[1] = S.cnt[2] = 0;
        S.cnt[3] = S.cnt[3] % 0xFFFFFFu + 1u;
        S.info[0] = S.info[1] = S.info[2] = S.info[3] = 0;
    }
}

// The head of the per-net body, one workgroup per slot: rip up the old path
// and re-price what it held (EdgeAccountant.clear_path UP:1118-1135,
// _clear_path_node_use UP:4648-4659, refresh_live_present_costs
// UP:1403-1422), check the terminals, clear the slot's tile bits, seed the
// search and queue the seeds' tiles on list 0.  Status: ORX_OK, ORX_NO_PATH
// (no seed or target in the box) or ORX_E_ARG (bad terminals or workspace
// offset); nets not ORX_OK are not searched.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void tile_prep_kernel(TileArgs A, int32_t rip) {
    ORX_SHARED(uint32_t, s_ok);
    const TileStep& S = A.S;
    const uint32_t s = blockIdx.x, tid = threadIdx.x;
    const Lat& g = A.g;
    const NetsDev& nets = A.nets;
    const bool wave = S.slot_net != nullptr;
    if (wave && (S.slot_net[s] < 0 || S.slot_net[s] >= nets.n)) {   // a broken plan: refused
        if (tid == 0) atomicOr(S.info + 3, 16ull);
        return;
    }
    const SlotGeo G = slot_geo(S, s);
    const Box& B = G.B;
    const int32_t k = G.k;
    if (rip && wave) {
        int32_t* path = nets.path_arena + nets.path_off[k];
        const int32_t old_len = nets.path_len[k];
        if (old_len > 0) {
            apply_path_block<BLOCK>(g, path, old_len, -1);
            __syncthreads();
            reprice_path_block<BLOCK>(g, A.P, path, old_len);
        }
    }
    const bool multi = wave && nets.seed_ptr && nets.seed_ptr[k + 1] > nets.seed_ptr[k];
    int32_t st = ORX_OK;
    if (G.off < 0 || (G.off & 63) || G.off + (int64_t)B.R > A.extent) st = ORX_E_ARG;
    uint32_t rsrc = UINT32_MAX;
    int32_t s0 = 0, ns = 0;
    if (multi) {
        s0 = nets.seed_ptr[k];
        ns = nets.seed_ptr[k + 1] - s0;
        const int32_t t0 = nets.tgt_ptr ? nets.tgt_ptr[k] : 0;
        const int32_t nt = nets.tgt_ptr ? nets.tgt_ptr[k + 1] - t0 : 0;
        if (nt < 1 || nt > kMaxTargets || !nets.seed_node || !nets.seed_cost || !nets.tgt_node) st = ORX_E_ARG;
        if (tid == 0) s_ok = 0;
        __syncthreads();
        if (st == ORX_OK) {
            // a search with no seed or no target in the box finds nothing
            uint32_t seeds_in = 0, tg_in = 0;
            for (int32_t i = (int32_t)tid; i < ns; i += BLOCK)
                seeds_in += global_to_roi(g, B, nets.seed_node[s0 + i]) != UINT32_MAX;
            for (int32_t i = (int32_t)tid; i < nt; i += BLOCK)
                tg_in += global_to_roi(g, B, nets.tgt_node[t0 + i]) != UINT32_MAX;
            if (seeds_in) atomicOr(&s_ok, 1u);
            if (tg_in) atomicOr(&s_ok, 2u);
        }
        __syncthreads();
        if (st == ORX_OK && s_ok != 3u) st = ORX_NO_PATH;
    } else if (st == ORX_OK) {
        rsrc = global_to_roi(g, B, wave ? (int64_t)nets.src[k] : A.ssrc);
        if (rsrc == UINT32_MAX) st = ORX_E_ARG;
        if (wave && global_to_roi(g, B, nets.dst[k]) == UINT32_MAX) st = ORX_E_ARG;
    }
    const int32_t ntiles = G.ntx * G.nty;
    const bool async = (S.flags & ORX_F_ASYNC) != 0;
    if (st == ORX_OK && async && (uint32_t)ntiles > kRingMax) st = ORX_E_ARG;
    if (tid == 0 && wave) nets.status[k] = st;
    if (st != ORX_OK) return;
    const uint32_t epoch = gload_u32(S.cnt + 3);
    if (async) {
        for (int32_t i = (int32_t)tid; i < ntiles; i += BLOCK) *tstate_of(S, G.off, (uint32_t)i) = 0u;
    } else {
        for (int32_t i = (int32_t)tid; i < (ntiles + 31) / 32; i += BLOCK) {
            S.marks[0][G.off / 32 + i] = 0u;
            S.marks[1][G.off / 32 + i] = 0u;
            S.seen[G.off / 32 + i] = 0u;
        }
    }
    __syncthreads();
    uint64_t* keys = S.keys + G.off;
    auto tile_of = [&](uint32_t r) -> uint32_t {
        uint32_t lx, ly, z;
        roi_decode(B, r, lx, ly, z);
        return (ly / (uint32_t)G.ty) * (uint32_t)G.ntx + lx / (uint32_t)kTX;
    };
    if (!multi) {
        if (tid == 0) {
            keys[rsrc] = kpack(0.0f, 0u, ORX_ROOT_SLOT);
            if (async)
                ring_activate(S, epoch, s, G.off, tile_of(rsrc));
            else
                queue_tile(S, s, G.off, tile_of(rsrc), 0u, &S.cnt[0]);
        }
        return;
    }
    // seeds: fl(seed cost + the seed's node term), duplicates keep the
    // minimum (cuda_dijkstra.py:5589-5606, UP:2582-2583)
    for (int32_t i = (int32_t)tid; i < ns; i += BLOCK) {
        const int64_t gs = nets.seed_node[s0 + i];
        const uint32_t rs = global_to_roi(g, B, gs);
        if (rs == UINT32_MAX) continue;
        const float q = node_cost(g.node_hist[gs],
                                  __hip_atomic_load(g.node_use + gs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), A.P);
        atomicMin((unsigned long long*)(keys + rs),
                  (unsigned long long)kpack(__fadd_rn(nets.seed_cost[s0 + i], q), 0u, ORX_ROOT_SLOT));
    }
    __syncthreads();
    for (int32_t i = (int32_t)tid; i < ns; i += BLOCK) {
        const uint32_t rs = global_to_roi(g, B, nets.seed_node[s0 + i]);
        if (rs == UINT32_MAX) continue;
        if (async)
            ring_activate(S, epoch, s, G.off, tile_of(rs));
        else
            queue_tile(S, s, G.off, tile_of(rs), 0u, &S.cnt[0]);
    }
}

// One step: every workgroup visits items of list k & 1.
template <int LC>
__global__ __launch_bounds__(kTileBlock) void tile_step_kernel(TileStep S, uint32_t k) {
    ORX_DYNAMIC_LDS(uint64_t, lds_mem);
    ORX_SHARED(TileShared, sh);
    const uint32_t n = gload_u32(S.cnt + k % 3);
    if (blockIdx.x == 0 && threadIdx.x == 0) S.cnt[(k + 2) % 3] = 0;
    if ((int64_t)n > S.cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(S.info + 3, 4ull);
        return;
    }
    const uint32_t pin = k & 1u, pout = pin ^ 1u;
    tile_layer_lists(S, LC ? LC : S.L, sh);
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x)
        tile_visit<kTileBlock, LC, false>(S, S.item_s[pin][i], S.item_t[pin][i], pin, pout, S.cnt + (k + 1) % 3, 0u,
                                          reinterpret_cast<unsigned char*>(lds_mem), sh);
}

// The tail of a step budget: one workgroup runs the remaining steps until no
// tile is active (bounded; a search still running past the bound is an
// error).  With one workgroup, nothing runs beside it.
template <int LC>
__global__ __launch_bounds__(kTileBlock) void tile_drain_kernel(TileStep S, uint32_t k0) {
    ORX_DYNAMIC_LDS(uint64_t, lds_mem);
    ORX_SHARED(TileShared, sh);
    uint32_t visits = 0;
    tile_layer_lists(S, LC ? LC : S.L, sh);
    for (uint32_t k = k0;; ++k) {
        const uint32_t n = gload_u32(S.cnt + k % 3);
        __syncthreads();   // every lane has read n before lane 0 clears a counter
        if (n == 0) break;
        if ((int64_t)n > S.cap || visits > kDrainVisits) {
            if (threadIdx.x == 0) atomicOr(S.info + 3, 8ull);
            break;
        }
        if (threadIdx.x == 0) S.cnt[(k + 2) % 3] = 0;
        const uint32_t pin = k & 1u, pout = pin ^ 1u;
        for (uint32_t i = 0; i < n; ++i)
            tile_visit<kTileBlock, LC, false>(S, S.item_s[pin][i], S.item_t[pin][i], pin, pout,
                                              S.cnt + (k + 1) % 3, 0u, reinterpret_cast<unsigned char*>(lds_mem),
                                              sh);
        visits += n;
        __threadfence();
        __syncthreads();
    }
    if (threadIdx.x == 0 && visits) atomicAdd(S.info + 2, (unsigned long long)visits);
}

// Queue form: the persistent launch of one wave (or one search).  Lane 0
// takes a ticket and waits for its ring entry; the workgroup visits that
// tile; until no tile is queued or running, or the deadline.
template <int LC>
__global__ __launch_bounds__(kTileBlock) void tile_run_kernel(TileStep S, uint64_t deadline_ticks) {
    ORX_DYNAMIC_LDS(uint64_t, lds_mem);
    ORX_SHARED(TileShared, sh);
    ORX_SHARED(uint64_t, s_item);
    const uint32_t epoch = gload_u32(S.cnt + 3);
    const uint64_t deadline = wall_clock64() + deadline_ticks;
    tile_layer_lists(S, LC ? LC : S.L, sh);
    for (;;) {
        if (threadIdx.x == 0) s_item = ring_pop(S, epoch, deadline);
        __syncthreads();
        const uint64_t v = s_item;
        if (v == 0) break;
        tile_visit<kTileBlock, LC, true>(S, (uint32_t)(v >> kRingField) & (kRingMax - 1u), (uint32_t)v & (kRingMax - 1u),
                                         0u, 0u, nullptr, epoch, reinterpret_cast<unsigned char*>(lds_mem), sh);
    }
}

// The tail of the per-net body, one workgroup per slot: the chosen target
// of a multi-terminal net (UP:2605-2618), the settled-set statistics, the
// parent walk, then commit and re-price (UP:1091-1116, 4635-4646, via
// columns UP:4470-4546, UP:1403-1422).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void tile_post_kernel(TileArgs A) {
    ORX_SHARED(Ctl, ctl);
    ORX_SHARED(TermTab, tt);
    const TileStep& S = A.S;
    const uint32_t s = blockIdx.x, tid = threadIdx.x;
    const Lat& g = A.g;
    const Params& P = A.P;
    const NetsDev& nets = A.nets;
    if (S.slot_net[s] < 0 || S.slot_net[s] >= nets.n) {   // refused by the prep
        if (tid == 0 && A.stats) atomicAdd(A.stats + 6, 1ull);
        return;
    }
    const SlotGeo G = slot_geo(S, s);
    const Box& B = G.B;
    const int32_t k = G.k;
    const uint64_t* keys = S.keys + G.off;
    const bool full = (S.flags & ORX_F_FULL) != 0;
    const bool multi = nets.seed_ptr && nets.seed_ptr[k + 1] > nets.seed_ptr[k];
    int32_t st = nets.status[k];
    // a list overflow, broken state or missed deadline of this wave's search
    // leaves its keys incomplete: no net of the wave commits
    if (st == ORX_OK && __hip_atomic_load(S.info + 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        st = ORX_E_OVERFLOW;
    if (tid == 0) { ctl.settled = 0; ctl.relaxed = 0; ctl.path_len = 0; }
    int32_t* path = nets.path_arena + nets.path_off[k];
    const int32_t cap = nets.path_cap[k];
    int32_t len = 0;
    float cost = INFINITY;
    __syncthreads();
    if (st == ORX_OK) {
        uint32_t rsrc = UINT32_MAX, rdst = UINT32_MAX;
        Heur H;
        bool big = false;
        if (multi) {
            const int32_t s0 = nets.seed_ptr[k], t0 = nets.tgt_ptr[k];
            const int32_t nt = nets.tgt_ptr[k + 1] - t0;
            const int32_t* tn = nets.tgt_node + t0;
            const float* tc = nets.tgt_cost ? nets.tgt_cost + t0 : nullptr;
            build_terms<BLOCK>(g, B, tn, nt, nets.seed_node + s0, nets.seed_ptr[k + 1] - s0, tt);
            if (tid < 64) pick_target(g, B, keys, tn, tc, nt, ctl);
            __syncthreads();
            rdst = ctl.best_r;
            big = rdst != UINT32_MAX && !(ctl.best_cost < P.astar_gmax);
            H = make_heur_box(P, B, tt, !full && !big);
        } else {
            rsrc = global_to_roi(g, B, nets.src[k]);
            rdst = global_to_roi(g, B, nets.dst[k]);
            const uint64_t kd = key_load(keys + rdst);
            big = kd != ORX_KEY_INF && !(kdist(kd) < P.astar_gmax);
            H = make_heur(g, P, nets.dst[k], !full && !big);
        }
        if (S.flags & ORX_F_STATS) {
            const bool all = full || rdst == UINT32_MAX;
            box_stats<BLOCK>(g, B, H, keys, rdst, all, &ctl, multi ? &tt : nullptr);
        }
        if (rdst != UINT32_MAX && tid < 64) {
            const int32_t l = box_backtrace_wave(g, B, keys, rsrc, rdst, path, cap);
            if (tid == 0) ctl.path_len = l;
        }
        __syncthreads();
        const int32_t l = rdst == UINT32_MAX ? 0 : ctl.path_len;
        if (l < 0) st = l;
        else if (l == 0 || (multi && l == 1)) st = ORX_NO_PATH;
        else {
            len = l;
            cost = multi ? ctl.best_cost : kdist(key_load(keys + rdst));
        }
    }
    if (tid == 0) {
        nets.status[k] = st;
        nets.path_len[k] = len;
        nets.path_cost[k] = cost;
    }
    if (len > 0) {
        apply_path_block<BLOCK>(g, path, len, +1);
        __syncthreads();
        reprice_path_block<BLOCK>(g, P, path, len);
    }
    unsigned long long* stats = A.stats;
    if (stats && tid == 0) {
        atomicAdd(stats + 0, ctl.settled);
        atomicAdd(stats + 1, ctl.relaxed);
        atomicAdd(stats + (len > 0 ? 5 : 6), 1ull);
        if (len > 0) atomicAdd(stats + 7, (unsigned long long)(len - 1));
        if (s == 0) {
            atomicAdd(stats + 2, S.info[0]);
            atomicAdd(stats + 3, S.info[1]);
            atomicAdd(stats + 4, S.info[2]);
        }
    }
}

// dist / parent arrays and statistics of a finished single search (the
// array-reporting form of find_path_roi_gpu).
__global__ __launch_bounds__(256) void tile_finish_kernel(TileArgs A, uint32_t rdst, int32_t full, float* dist,
                                                          int32_t* parent, unsigned long long* acc) {
    const TileStep& S = A.S;
    const Box B = make_box(S.sbox, S.L);
    const Lat& g = A.g;
    const uint64_t kd = full ? ORX_KEY_INF : S.keys[rdst];
    unsigned long long settled = 0, relaxed = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < B.R; r += gridDim.x * blockDim.x) {
        const uint64_t k = S.keys[r];
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
        if (!full && (k > kd || (k == kd && r > rdst))) continue;
        settled += 1;
        if (!full && r == rdst) continue;
        for (int s = 0; s < 4; ++s) relaxed += roi_nbr(B, horiz, r, lx, ly, z, s) != UINT32_MAX;
    }
    settled = wave_sum(settled);
    relaxed = wave_sum(relaxed);
    if (lane_id() == 0) {
        if (settled) atomicAdd(acc + 0, settled);
        if (relaxed) atomicAdd(acc + 1, relaxed);
    }
}

__global__ void tile_single_done_kernel(TileArgs A, uint32_t rdst, int32_t full, unsigned long long* acc,
                                        int32_t* status) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const TileStep& S = A.S;
    if (A.stats) {
        atomicAdd(A.stats + 0, acc[0]);
        atomicAdd(A.stats + 1, acc[1]);
        atomicAdd(A.stats + 2, S.info[0]);
        atomicAdd(A.stats + 3, S.info[1]);
        atomicAdd(A.stats + 4, S.info[2]);
    }
    *status = S.info[3] ? ORX_E_OVERFLOW : ((!full && S.keys[rdst] == ORX_KEY_INF) ? ORX_NO_PATH : ORX_OK);
}

// Workspace layout of the tile engine: keys as the per-net engine; the four
// queue arrays hold the two item lists (slot, tile); the three bitmap arrays
// the queued bits (two list parities) and the visited bits, at word
// slot_off / 32 of each; the last kTileCtlWords words of the third array the
// step counters and the info block, so the searches own at most the first
// ws_nodes - kTileReserve workspace nodes.
constexpr int64_t kTileCtlWords = 16;
constexpr int64_t kTileReserve = 32 * kTileCtlWords;

// Phase cap of a routing wave's tile visits (step form): ORX_TILE_PHASE_CAP
// if set (0: none, else at least 3), otherwise kRoutePhaseCap.  The
// whole-box query runs uncapped.
int32_t tile_phase_cap();

TileArgs tile_args(const orx_lattice* lat, const orx_params* p, const orx_workspace* ws, int32_t flags) {
    TileArgs A;
    A.g = make_lat(lat);
    A.P = p ? make_params(p) : Params{};
    A.nets = NetsDev{};
    A.K = 0;
    A.ssrc = 0;
    A.stats = nullptr;
    TileStep& S = A.S;
    S.X = lat->X;
    S.L = lat->L;
    S.hmask = lat->hmask;
    S.XY = (int64_t)lat->X * lat->Y;
    S.wcost = lat->wcost;
    S.box = nullptr;
    S.seed_ptr = S.tgt_ptr = S.tgt_node = nullptr;
    S.slot_net = nullptr;
    S.slot_off = nullptr;
    S.sbox[0] = S.sbox[1] = S.sbox[2] = S.sbox[3] = 0;
    const int64_t wn = ws->ws_nodes;
    S.keys = ws->keys;
    S.item_s[0] = ws->queues;
    S.item_t[0] = ws->queues + wn;
    S.item_s[1] = ws->queues + 2 * wn;
    S.item_t[1] = ws->queues + 3 * wn;
    S.marks[0] = ws->bitmaps;
    S.marks[1] = ws->bitmaps + wn / 32;
    S.seen = ws->bitmaps + 2 * (wn / 32);
    uint32_t* ctl = ws->bitmaps + 3 * (wn / 32) - kTileCtlWords;
    S.cnt = ctl;
    S.info = reinterpret_cast<unsigned long long*>(ctl + 4);
    S.cap = wn;
    S.flags = flags;
    S.phase_cap = 0;
    A.extent = wn - kTileReserve;
    return A;
}

int check_tile_ws(const orx_workspace* ws) {
    if (!ws || !ws->keys || !ws->queues || !ws->bitmaps) return orx_host::fail(ORX_E_ARG, "workspace has a NULL array");
    if (ws->ws_nodes % 64) return orx_host::fail(ORX_E_ARG, "ws_nodes must be a multiple of 64");
    if (ws->ws_nodes < 2 * kTileReserve)
        return orx_host::fail(ORX_E_ARG, "the tile engine needs a workspace of at least 1024 nodes");
    if ((uintptr_t)ws->keys & 15u) return orx_host::fail(ORX_E_ARG, "workspace keys must be 16-byte aligned");
    if ((uintptr_t)ws->bitmaps & 7u) return orx_host::fail(ORX_E_ARG, "workspace bitmaps must be 8-byte aligned");
    if ((uintptr_t)ws->queues & 7u) return orx_host::fail(ORX_E_ARG, "workspace queues must be 8-byte aligned");
    return ORX_OK;
}

template <int LC>
int tile_lds_ready_lc(size_t b) {
    int rc = set_lds<tile_step_kernel<LC>>(b);
    if (!rc) rc = set_lds<tile_drain_kernel<LC>>(b);
    if (!rc) rc = set_lds<tile_run_kernel<LC>>(b);
    return rc;
}

#define ORX_TILE_CLASSES(X) X(0) X(6) X(8) X(12) X(16) X(20) X(24) X(32)

int tile_lds_ready(int32_t L) {
    const size_t b = (size_t)tile_lds(L).bytes;
    switch (layer_class(L)) {
#define ORX_CASE(LC) case LC: return tile_lds_ready_lc<LC>(b);
        ORX_TILE_CLASSES(ORX_CASE)
#undef ORX_CASE
    }
    return orx_host::fail(ORX_E_ARG, "no tile kernel for this layer count");
}

// Step workgroups per launch: as many as the LDS of the chip holds, or the
// caller's cap (max_groups > 0).
int tile_grid(int32_t L, int32_t max_groups) {
    const int per_cu = std::max(1, (int)(kMaxDynLds / (size_t)tile_lds(L).bytes));
    const int g = device_cus() * per_cu;
    return max_groups > 0 ? std::min(g, (int)max_groups) : g;
}

void launch_step(const TileStep& S, uint32_t k, int grid, size_t lds, hipStream_t s) {
    switch (layer_class(S.L)) {
#define ORX_CASE(LC) \
    case LC: hipLaunchKernelGGL(tile_step_kernel<LC>, dim3(grid), dim3(kTileBlock), lds, s, S, k); break;
        ORX_TILE_CLASSES(ORX_CASE)
#undef ORX_CASE
    }
}

// Wall-clock ticks of the persistent launch's deadline: kRunSeconds, or
// ORX_TILE_RUN_SECONDS from the environment (read once; test builds use it to
// stretch the bound for slow emulation or to force the error path).
int32_t tile_phase_cap() {
    static int32_t cap = -1;
    if (cap < 0) {
        cap = kRoutePhaseCap;
        if (const char* e = std::getenv("ORX_TILE_PHASE_CAP")) {
            const int v = std::atoi(e);
            if (v == 0 || (v >= 3 && v < 1000)) cap = v;
        }
    }
    return cap;
}

uint64_t run_deadline_ticks() {
    static uint64_t ticks = 0;
    if (ticks == 0) {
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;
        double secs = (double)kRunSeconds;
        if (const char* e = std::getenv("ORX_TILE_RUN_SECONDS")) {
            const double v = std::atof(e);
            if (v > 0.0 && v < 1e6) secs = v;
        }
        ticks = std::max<uint64_t>(1, (uint64_t)((double)khz * 1000.0 * secs));
    }
    return ticks;
}

// Workgroups of a persistent launch: the chip's, or fewer when the wave has
// fewer tiles (every slot's tiles, bounded by its share of the extent).
int run_grid(int32_t L, int32_t K, int64_t extent, int32_t max_groups) {
    const int64_t tile_nodes = (int64_t)kTX * tile_ty(L) * L;
    const int64_t tiles = (int64_t)K + 2 * ((extent + tile_nodes - 1) / tile_nodes);
    return (int)std::max<int64_t>(1, std::min<int64_t>(tile_grid(L, max_groups), tiles));
}

void launch_run(const TileStep& S, int grid, size_t lds, hipStream_t s) {
    switch (layer_class(S.L)) {
#define ORX_CASE(LC)                                                                                 \
    case LC:                                                                                         \
        hipLaunchKernelGGL(tile_run_kernel<LC>, dim3(grid), dim3(kTileBlock), lds, s, S, run_deadline_ticks()); \
        break;
        ORX_TILE_CLASSES(ORX_CASE)
#undef ORX_CASE
    }
}

void launch_drain(const TileStep& S, uint32_t k0, size_t lds, hipStream_t s) {
    switch (layer_class(S.L)) {
#define ORX_CASE(LC) \
    case LC: hipLaunchKernelGGL(tile_drain_kernel<LC>, dim3(1), dim3(kTileBlock), lds, s, S, k0); break;
        ORX_TILE_CLASSES(ORX_CASE)
#undef ORX_CASE
    }
}

// One wave: fill, prep, `steps` step launches, drain, post.
int launch_tile_wave(TileArgs A, const int32_t* d_wave_nets, const int64_t* d_ws_off, int32_t K, int64_t ws_extent,
                     int32_t steps, int32_t block, int32_t max_groups, hipStream_t s) {
    if (K <= 0) return ORX_OK;
    A.S.slot_net = d_wave_nets;
    A.S.slot_off = d_ws_off;
    A.K = K;
    A.extent = std::min<int64_t>(ws_extent, A.S.cap - kTileReserve);
    // queue form: the ring holds as many entries as the wave has workspace
    // nodes (at least one per tile), cleared by the fill
    if (A.S.flags & ORX_F_ASYNC) A.S.cap = std::max<int64_t>(A.extent, 2);
    const size_t lds = (size_t)tile_lds(A.S.L).bytes;
    const int fill_grid = (int)std::max<int64_t>(1, std::min<int64_t>(device_cus() * 8, (A.extent / 2 + 255) / 256));
    hipLaunchKernelGGL(tile_fill_kernel, dim3(fill_grid), dim3(256), 0, s, A.S, A.extent);
    hipLaunchKernelGGL(tile_prep_kernel<256>, dim3(K), dim3(256), 0, s, A, 1);
    if (A.S.flags & ORX_F_ASYNC) {
        launch_run(A.S, run_grid(A.S.L, K, A.extent, max_groups), lds, s);
    } else {
        const int grid = tile_grid(A.S.L, max_groups);
        for (int32_t k = 0; k < steps; ++k) launch_step(A.S, (uint32_t)k, grid, lds, s);
        launch_drain(A.S, (uint32_t)steps, lds, s);
    }
    switch (block) {
    case 256: hipLaunchKernelGGL(tile_post_kernel<256>, dim3(K), dim3(256), 0, s, A); break;
    case 512: hipLaunchKernelGGL(tile_post_kernel<512>, dim3(K), dim3(512), 0, s, A); break;
    case 1024: hipLaunchKernelGGL(tile_post_kernel<1024>, dim3(K), dim3(1024), 0, s, A); break;
    default: return orx_host::fail(ORX_E_ARG, "block must be 256, 512 or 1024");
    }
    ORX_HIP_CHECK(hipGetLastError());
    return ORX_OK;
}

}  // namespace

namespace orx_host {
int prepare_tile_kernels() {
    int rc = ORX_OK;
#define ORX_CASE(LC) if (!rc) rc = tile_lds_ready_lc<LC>(kMaxDynLds);
    ORX_TILE_CLASSES(ORX_CASE)
#undef ORX_CASE
    return rc;
}

// Whole-box query on the tile engine (orx_sssp_box with ORX_F_TILES): fill,
// seed, replays of a captured graph of kTileSteps steps until no tile is
// active, then the dist / parent arrays.  Runs on a private stream after the
// caller's stream drained.
int sssp_box_tiles(const orx_lattice* lat, int32_t x0, int32_t y0, int32_t w, int32_t h, int64_t src, int64_t dst,
                   float* d_dist, int32_t* d_parent, const orx_workspace* ws, const orx_route_cfg* cfg,
                   hipStream_t caller) {
    int rc = check_tile_ws(ws);
    if (rc) return rc;
    if (lat->L > kMaxL) return orx_host::fail(ORX_E_ARG, "the tile engine supports at most 32 layers");
    TileArgs A = tile_args(lat, nullptr, ws, cfg->flags);
    A.stats = (unsigned long long*)cfg->stats;
    A.S.sbox[0] = x0; A.S.sbox[1] = y0; A.S.sbox[2] = w; A.S.sbox[3] = h;
    A.ssrc = src;
    A.K = 1;
    const int64_t R = (int64_t)w * h * lat->L;
    if (R > A.extent) return orx_host::fail(ORX_E_ARG, "workspace smaller than the ROI plus the tile engine's reserve");
    A.extent = R;
    if (cfg->flags & ORX_F_ASYNC) A.S.cap = std::max<int64_t>(R, 2);
    const bool full = dst < 0 || (cfg->flags & ORX_F_FULL);
    uint32_t rdst = 0;
    if (dst >= 0) {
        const int64_t XY = (int64_t)lat->X * lat->Y, z = dst / XY, rem = dst - z * XY, y = rem / lat->X,
                      x = rem - y * lat->X;
        rdst = (uint32_t)((z * h + (y - y0)) * w + (x - x0));
    }
    rc = tile_lds_ready(lat->L);
    if (rc) return rc;
    ORX_HIP_CHECK(hipStreamSynchronize(caller));
    hipStream_t ps = nullptr;
    ORX_HIP_CHECK(hipStreamCreateWithFlags(&ps, hipStreamNonBlocking));
    int32_t* d_status = nullptr;
    unsigned long long* d_acc = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* h_cnt = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipError_t e = hipMallocAsync((void**)&d_status, sizeof(int32_t), ps);
    if (e == hipSuccess) e = hipMallocAsync((void**)&d_acc, 2 * sizeof(unsigned long long), ps);
    if (e == hipSuccess) e = hipMemsetAsync(d_acc, 0, 2 * sizeof(unsigned long long), ps);
    const size_t lds = (size_t)tile_lds(lat->L).bytes;
    const int grid = tile_grid(lat->L, cfg->max_groups);
    if (e == hipSuccess) {
        const int fill_grid = (int)std::max<int64_t>(1, std::min<int64_t>(device_cus() * 8, (R / 2 + 255) / 256));
        hipLaunchKernelGGL(tile_fill_kernel, dim3(fill_grid), dim3(256), 0, ps, A.S, R);
        hipLaunchKernelGGL(tile_prep_kernel<256>, dim3(1), dim3(256), 0, ps, A, 0);
        e = hipGetLastError();
    }
    static_assert(kTileSteps % 6 == 0, "list (mod 2) and counter (mod 3) parity must carry over between replays");
    const bool async = (cfg->flags & ORX_F_ASYNC) != 0;
    if (async && e == hipSuccess) {
        // queue form: one persistent launch runs the search to its end
        launch_run(A.S, run_grid(lat->L, 1, R, cfg->max_groups), lds, ps);
        e = hipGetLastError();
    }
    if (!async && e == hipSuccess) e = hipStreamBeginCapture(ps, hipStreamCaptureModeThreadLocal);
    if (!async && e == hipSuccess) {
        for (int i = 0; i < kTileSteps; ++i) launch_step(A.S, (uint32_t)i, grid, lds, ps);
        e = hipStreamEndCapture(ps, &graph);
    }
    if (!async && e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (!async && e == hipSuccess) e = hipHostMalloc((void**)&h_cnt, 2 * sizeof(uint32_t), hipHostMallocDefault);
    if (!async && e == hipSuccess) e = hipEventCreateWithFlags(&ev[0], hipEventDisableTiming);
    if (!async && e == hipSuccess) e = hipEventCreateWithFlags(&ev[1], hipEventDisableTiming);
    // after each replay the next step's list size is cnt[0] (kTileSteps % 3
    // == 0); replays run two deep so the device does not idle through the
    // host's round trip
    auto replay = [&](int slot) -> hipError_t {
        hipError_t x = hipGraphLaunch(exec, ps);
        if (x == hipSuccess) x = hipMemcpyAsync(h_cnt + slot, A.S.cnt, sizeof(uint32_t), hipMemcpyDeviceToHost, ps);
        if (x == hipSuccess) x = hipEventRecord(ev[slot], ps);
        return x;
    };
    if (!async && e == hipSuccess) e = replay(0);
    if (!async && e == hipSuccess) {
        const auto t0 = std::chrono::steady_clock::now();
        for (int it = 0;; ++it) {
            const int cur = it & 1;
            e = replay(cur ^ 1);
            if (e == hipSuccess) e = hipEventSynchronize(ev[cur]);
            if (e != hipSuccess || h_cnt[cur] == 0) break;
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(300)) {
                rc = orx_host::fail(ORX_E_OVERFLOW, "tile search did not finish within 300 s");
                break;
            }
        }
    }
    if (e == hipSuccess && rc == ORX_OK) {
        const int fin_grid = (int)std::max<int64_t>(1, std::min<int64_t>(device_cus() * 8, (R + 255) / 256));
        hipLaunchKernelGGL(tile_finish_kernel, dim3(fin_grid), dim3(256), 0, ps, A, rdst, full ? 1 : 0, d_dist,
                           d_parent, d_acc);
        hipLaunchKernelGGL(tile_single_done_kernel, dim3(1), dim3(64), 0, ps, A, rdst, full ? 1 : 0, d_acc,
                           d_status);
        int32_t h_status = 0;
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&h_status, d_status, sizeof(int32_t), hipMemcpyDeviceToHost, ps);
        if (e == hipSuccess) e = hipStreamSynchronize(ps);
        if (e == hipSuccess)
            rc = h_status == ORX_E_OVERFLOW ? orx_host::fail(ORX_E_OVERFLOW, "tile search: a list or bound overflowed")
                                            : h_status;
    }
    if (e != hipSuccess && rc == ORX_OK) rc = orx_host::check_hip(e, "orx_sssp_box (tiles)");
    (void)hipStreamSynchronize(ps);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev[0]) (void)hipEventDestroy(ev[0]);
    if (ev[1]) (void)hipEventDestroy(ev[1]);
    if (h_cnt) (void)hipHostFree(h_cnt);
    if (d_status) (void)hipFreeAsync(d_status, ps);
    if (d_acc) (void)hipFreeAsync(d_acc, ps);
    (void)hipStreamSynchronize(ps);
    (void)hipStreamDestroy(ps);
    return rc;
}
}  // namespace orx_host

extern "C" {

int orx_route_tiles(const orx_lattice* lat, const orx_params* p, const orx_nets* nets, const int32_t* d_nets,
                    const int64_t* d_ws_off, int32_t n_waves, const int32_t* h_wave_ptr, const int64_t* h_wave_ws,
                    const int32_t* h_wave_steps, const orx_workspace* ws, const orx_route_cfg* cfg,
                    orx_stream_t stream) {
    int rc = check_lattice(lat);
    if (rc) return rc;
    if (!p || !nets || !ws || !cfg || !h_wave_ptr || !h_wave_ws || !h_wave_steps)
        return orx_host::fail(ORX_E_ARG, "NULL argument");
    if (n_waves <= 0) return ORX_OK;
    if (!d_nets || !d_ws_off || !nets->box) return orx_host::fail(ORX_E_ARG, "NULL argument");
    rc = check_tile_ws(ws);
    if (rc) return rc;
    if (lat->L > kMaxL) return orx_host::fail(ORX_E_ARG, "the tile engine supports at most 32 layers");
    if (h_wave_ptr[0] != 0) return orx_host::fail(ORX_E_ARG, "wave_ptr[0] must be 0");
    for (int32_t i = 0; i < n_waves; ++i) {
        if (h_wave_ptr[i + 1] < h_wave_ptr[i] || h_wave_ptr[i + 1] > nets->n)
            return orx_host::fail(ORX_E_ARG, "wave_ptr must be non-decreasing and at most n");
        if (h_wave_ws[i] < 0 || h_wave_ws[i] > ws->ws_nodes - kTileReserve)
            return orx_host::fail(ORX_E_ARG, "wave workspace extent outside [0, ws_nodes - 512]");
        if (h_wave_steps[i] < 0 || h_wave_steps[i] > 65536)
            return orx_host::fail(ORX_E_ARG, "wave step budget outside [0, 65536]");
        if ((cfg->flags & ORX_F_ASYNC) && h_wave_ptr[i + 1] - h_wave_ptr[i] > (int32_t)kRingMax)
            return orx_host::fail(ORX_E_ARG, "the queue form takes at most 2^20 nets per wave");
    }
    rc = tile_lds_ready(lat->L);
    if (rc) return rc;
    TileArgs A = tile_args(lat, p, ws, cfg->flags);
    A.S.phase_cap = tile_phase_cap();
    A.nets = make_nets(nets);
    A.S.box = nets->box;
    A.S.seed_ptr = nets->seed_ptr;
    A.S.tgt_ptr = nets->tgt_ptr;
    A.S.tgt_node = nets->tgt_node;
    A.stats = (unsigned long long*)cfg->stats;
    const hipStream_t s = (hipStream_t)stream;
    for (int32_t i = 0; i < n_waves; ++i) {
        const int32_t a = h_wave_ptr[i], K = h_wave_ptr[i + 1] - a;
        rc = launch_tile_wave(A, d_nets + a, d_ws_off + a, K, h_wave_ws[i], h_wave_steps[i], cfg->block,
                              cfg->max_groups, s);
        if (rc) return rc;
    }
    return ORX_OK;
}

}  // extern "C"
