/*
 * orthoroute_hip.h -- C ABI of the MI355X (gfx950) PathFinder / SSSP library
 * (liborx_hip.so).
 *
 * Plain C: every buffer is a raw device pointer the caller owns (the Python
 * host keeps them in PyTorch-ROCm tensors); the library never frees caller
 * memory and never throws across the ABI.  Every entry point takes the HIP
 * stream to enqueue on and returns an int status:
 *     ORX_OK (0), ORX_NO_PATH (1), or a negative ORX_E_* error;
 * orx_last_error() returns the message of the last failure on this thread.
 *
 * Each declaration names the interface of bbenchoff/OrthoRoute it replaces
 * (file:line; UP = orthoroute/algorithms/manhattan/unified_pathfinder.py,
 * CD = orthoroute/algorithms/manhattan/pathfinder/cuda_dijkstra.py,
 * PK = orthoroute/algorithms/manhattan/pathfinder/persistent_kernel.py).
 */
#ifndef ORTHOROUTE_HIP_H
#define ORTHOROUTE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* orx_stream_t; /* a hipStream_t; NULL = default stream */

enum {
    ORX_OK = 0,
    ORX_NO_PATH = 1,
    ORX_E_ARG = -1,        /* bad argument / shape                          */
    ORX_E_HIP = -2,        /* HIP runtime error (launch, attribute)         */
    ORX_E_OVERFLOW = -3,   /* a queue, path or hop counter exceeded its cap */
    ORX_E_CYCLE = -4,      /* parent walk did not reach a root              */
    ORX_E_NODEVICE = -5
};

/* ---------------------------------------------------------------------- */
/* Lattice state resident in HBM.  N = X*Y*L nodes, node id z*X*Y+y*X+x   */
/* (UP:118, kicad_geometry.py:59-62).  Four arc slots per node, ordered by */
/* neighbour id: 0 via down, 1 planar minus, 2 planar plus, 3 via up; arc  */
/* id 4*u+s; the reverse of slot s is 3-s.  Bit z of hmask set = layer z   */
/* routes along x (H), else along y (V) (UP:1487-1495).  Absent arcs have  */
/* base = +inf.  Replaces CSRGraph + EdgeAccountant arrays (UP:851-1002,   */
/* UP:1039-1052) and the via / path-node pools (UP:3019-3048, 4623-4659).   */
/* ---------------------------------------------------------------------- */
typedef struct orx_lattice {
    int32_t X, Y, L;
    uint32_t hmask;
    const float* base;   /* [4N] base arc cost                               */
    float* hist;         /* [4N] arc history                                 */
    int32_t* occ;        /* [4N] arc occupancy (both arcs of an edge equal)  */
    float* wcost;        /* [4N] live entering weight fl(c_arc + q_head)     */
    float* node_hist;    /* [N]                                              */
    int32_t* node_use;   /* [N]  path-node use                               */
    int32_t* via_col;    /* [X*Y] via barrels per column                     */
} orx_lattice;

/* Cost / history parameters (UP:1301-1384 update_costs, 1226-1289
 * update_history, 5308-5383 node penalty).  The arc cost is
 *   fl(fl(fl(b*base_w) + fl(hist_w*h)) + fl(pres_fac*max(0, occ+1-edge_cap)))
 * the node term fl(fl(node_hist_w*nh) + fl(node_pres*max(0, use+1-node_cap))),
 * and the entering weight their fp32 sum. */
typedef struct orx_params {
    float pres_fac;
    float base_w;
    float hist_w;
    float node_pres;
    float node_hist_w;
    int32_t edge_cap;
    int32_t node_cap;
    float hist_gain;
    float hist_decay;
    float hist_cap_mult;   /* > 0: increment <= mult*base (UP:1255-1262)   */
    float node_hist_gain;
    /* A* toward the sink in orx_route_wave (0 = plain Dijkstra order):
     * per-move lower bounds of the entering weight, planar / via, each at
     * least half an fp32 ulp of astar_gmax below the true minimum; results
     * with a sink distance (multi-terminal nets: a chosen total) >=
     * astar_gmax are recomputed without it. */
    float astar_cp;
    float astar_cv;
    float astar_gmax;
} orx_params;

/* Per-net tables of one routing job (resident in HBM). */
typedef struct orx_nets {
    int32_t n;
    const int32_t* box;       /* [4n] ROI box x0, y0, w, h (all layers)      */
    const int32_t* src;       /* [n] global source node                     */
    const int32_t* dst;       /* [n] global sink node                       */
    int32_t* path_arena;      /* concatenated node paths, src first         */
    const int64_t* path_off;  /* [n] offset of each net's path              */
    const int32_t* path_cap;  /* [n] capacity (>= w*h*L)                    */
    int32_t* path_len;        /* [n] 0 = unrouted                           */
    float* path_cost;         /* [n] fp32 distance of the sink              */
    int32_t* status;          /* [n] last status (ORX_OK / ORX_NO_PATH / <0) */
    /* Multi-terminal nets (portal seeds and targets,
     * find_path_multisource_multisink UP:2550-2659 with the seeds of
     * _get_pad_portal_seeds UP:4054-4180), or NULL seed_ptr for none.  Net k
     * with seed_ptr[k+1] > seed_ptr[k] ignores src/dst and routes from the
     * seeds seed_node[seed_ptr[k] .. seed_ptr[k+1]), each starting at
     * fl(seed_cost + the seed's node term) (CD:5589-5606; duplicates keep the
     * minimum), to the reached target of tgt_node[tgt_ptr[k] ..
     * tgt_ptr[k+1]) with the smallest (fl(dist + tgt_cost), node id).
     * Targets are not expanded (UP:2611-2618); 1..ORX_MAX_TARGETS per net;
     * terminals outside the net's box are ignored.  path_cost is then that
     * total.  A best target that is itself a seed gives ORX_NO_PATH
     * (a one-node path, UP:2647-2648).  Costs are >= 0 (seed costs
     * finite); tgt_cost may be NULL (all 0). */
    const int32_t* seed_ptr;  /* [n+1] */
    const int32_t* seed_node; /* global node ids                             */
    const float* seed_cost;
    const int32_t* tgt_ptr;   /* [n+1] */
    const int32_t* tgt_node;
    const float* tgt_cost;
} orx_nets;

#define ORX_MAX_TARGETS 128

/* Scratch for one wave: every slot k of the wave owns nodes
 * [ws_off[k], ws_off[k] + R_k) of keys (u64), 4 queues (u32, stride
 * ws_nodes) and 3 bitmaps (u32 words, offset ws_off[k]/32).  ws_off must be
 * multiples of 64.  Sizes: keys 8*ws_nodes B, queues 16*ws_nodes B,
 * bitmaps 3*ws_nodes/8 B. */
typedef struct orx_workspace {
    uint64_t* keys;
    uint32_t* queues;
    uint32_t* bitmaps;
    int64_t ws_nodes;
} orx_workspace;

typedef struct orx_route_cfg {
    float delta;            /* bucket width; +inf = pure frontier relaxation */
    int32_t block;          /* workgroup size: 256, 512 or 1024              */
    int32_t lds_nodes;      /* ROIs up to this many nodes keep keys in LDS   */
    int32_t flags;          /* ORX_F_* */
    int64_t* stats;         /* device [8] accumulators or NULL               */
    int32_t max_groups;     /* tile engine: cap on the workgroups of a step  */
                            /* or persistent launch (0: the whole chip); a   */
                            /* server running several jobs at once gives     */
                            /* each its share of the CUs                     */
} orx_route_cfg;

#define ORX_F_STATS 1       /* count settled nodes / relaxed arcs per query  */
#define ORX_F_FULL 2        /* no early exit: settle the whole ROI           */
#define ORX_F_GRID 8        /* orx_sssp_box: grid-wide search (automatic     */
                            /* above 2^18 ROI nodes)                         */
#define ORX_F_SEED_PEN 4    /* CSR seeds pay their own node penalty (multi-  */
                            /* seed forms UP:2582-2583, CD:5589-5606)        */
#define ORX_F_TILES 16      /* orx_sssp_box: tile-step engine (the whole box  */
                            /* relaxed to its fixpoint, see orx_route_tiles)  */
#define ORX_F_ASYNC 32      /* tile engine, queue form: one persistent launch */
                            /* per wave visits tiles as they are queued (no   */
                            /* step launches; h_wave_steps unused)            */
                            /* (bits 16 and up are reserved)                 */

/* stats[]: 0 settled nodes (label-setting count up to the sink; for a
 * multi-terminal net up to the chosen target), 1 relaxed arcs (their in-ROI
 * out-degree, targets excluded), 2 queue items processed, 3 near rounds,
 * 4 threshold advances, 5 nets routed, 6 nets failed, 7 path hops. */

/* Library init: checks for a gfx950 device.  Replaces the CuPy
 * RawKernel JIT set-up of CD:98-150 / PK:519-529. */
int orx_init(void);
const char* orx_last_error(void);
int orx_version(void);

/* Recompute wcost for every arc from base/hist/occ/node arrays (the
 * per-iteration cost refresh: UP:1301-1384 update_costs followed by
 * UP:1386-1401 begin_live_present_costs). */
int orx_lattice_reprice(const orx_lattice* lat, const orx_params* p, orx_stream_t stream);

/* End-of-iteration accounting, fused: overuse over canonical arcs
 * (UP:1137-1190, 4919-4932), node overuse (UP:10718-10726), via hops and
 * barrels, then the history update (UP:1226-1289 with raw present, node
 * history UP:10048-10081).  d_out (device int64[8]) receives edge_over,
 * edge_over_count, node_over, node_over_count, via_hops, barrels; it is
 * zeroed by the call. */
int orx_lattice_end_iteration(const orx_lattice* lat, const orx_params* p, int64_t* d_out,
                              orx_stream_t stream);

/* Route one wave of nets whose ROI boxes are pairwise disjoint, one
 * workgroup per net: rip up the old path, delta-stepping SSSP over the ROI
 * box, parent backtrace, commit, live re-price of touched arcs.  Equivalent
 * to running the per-net body of _route_all (UP:7949-8430) for these nets in
 * any order.  d_wave_nets[K] are net ids, d_ws_off[K] workspace offsets;
 * max_nodes = max ROI size in the wave. */
int orx_route_wave(const orx_lattice* lat, const orx_params* p, const orx_nets* nets,
                   const int32_t* d_wave_nets, const int64_t* d_ws_off, int32_t K,
                   int64_t max_nodes, const orx_workspace* ws, const orx_route_cfg* cfg,
                   orx_stream_t stream);

/* The whole routing step of one iteration as a launch sequence: wave i is
 * the nets d_nets[h_wave_ptr[i] .. h_wave_ptr[i+1]) with workspace offsets
 * d_ws_off[same range], routed as by orx_route_wave, waves in order on one
 * stream (the per-net loop of _route_all, UP:7949-8430, batched).
 * h_wave_ptr[n_waves+1], h_wave_max[n_waves] (largest ROI of each wave)
 * and h_wave_ws[n_waves] (the wave's nets own workspace nodes inside
 * [0, h_wave_ws[i]); the library clears that range chip-wide before the
 * wave when its largest ROI keeps search state in the workspace) are HOST
 * arrays. */
int orx_route_waves(const orx_lattice* lat, const orx_params* p, const orx_nets* nets,
                    const int32_t* d_nets, const int64_t* d_ws_off, int32_t n_waves,
                    const int32_t* h_wave_ptr, const int64_t* h_wave_max, const int64_t* h_wave_ws,
                    const orx_workspace* ws, const orx_route_cfg* cfg, orx_stream_t stream);

/* The routing step of one iteration on the tile-step engine: the same
 * waves, offsets and results as orx_route_waves (the per-net loop of
 * _route_all, UP:7949-8430, batched over disjoint-ROI waves), with each
 * net's search (heap Dijkstra UP:2451-2548 / the wavefront and persistent
 * searches CD:2800-3533, PK:260-516) run as tiles of 16 x ty cells x all
 * layers (ty = min(16, 5120 / (16 L))), relaxed to a local fixpoint in LDS
 * by line scans and exchanging boundary keys between step launches.  Every
 * net's box is searched to its fixpoint (no early exit), so paths, costs
 * and parents are the canonical ones.  Per wave: one fill, one prep (rip-up,
 * seeds), h_wave_steps[i] step launches, one single-workgroup drain of
 * whatever is still active, one post (backtrace, commit, re-price).
 * h_wave_steps ~ the largest net's tile count along x plus along y.  The
 * workspace needs ws_nodes >= 512 and >= every h_wave_ws[i]; its queue and
 * bitmap arrays hold the tile lists, bits and step counters.  L <= 32. */
int orx_route_tiles(const orx_lattice* lat, const orx_params* p, const orx_nets* nets, const int32_t* d_nets,
                    const int64_t* d_ws_off, int32_t n_waves, const int32_t* h_wave_ptr, const int64_t* h_wave_ws,
                    const int32_t* h_wave_steps, const orx_workspace* ws, const orx_route_cfg* cfg,
                    orx_stream_t stream);

/* Launch-sequence capture into a hipGraph (no reference counterpart: the
 * reference launches each CuPy kernel from Python, CD:2968-3043).
 * orx_graph_begin returns a private capture stream; orx_* calls enqueued on
 * it are recorded, not run; orx_graph_end instantiates them;
 * orx_graph_launch replays the graph on a stream; orx_graph_destroy frees it
 * once no launch of it is pending. */
typedef void* orx_graph_t;
int orx_graph_begin(orx_stream_t* capture_stream);
int orx_graph_end(orx_stream_t capture_stream, orx_graph_t* graph);
int orx_graph_launch(orx_graph_t graph, orx_stream_t stream);
int orx_graph_destroy(orx_graph_t graph);

/* Route every net of one iteration in ONE persistent launch.  d_sched[n]
 * is a topological order of the nets; net k may start once every net in
 * d_dep_ids[d_dep_ptr[k] .. d_dep_ptr[k+1]) has published d_done[id] == epoch
 * (the last earlier-ordered nets covering a cell of its ROI box, from
 * orx_plan_iteration).  This reproduces the sequential per-net loop of
 * _route_all (UP:7903-8449) exactly, with disjoint-ROI nets in flight
 * together.  d_ws_off[k]: per-net workspace offset (multiple of 64);
 * d_counter: one u32 the call zeroes; max_groups <= 0: one group per CU slot
 * the occupancy allows. */
int orx_route_dag(const orx_lattice* lat, const orx_params* p, const orx_nets* nets,
                  const int32_t* d_sched, int32_t n, const int32_t* d_dep_ptr,
                  const int32_t* d_dep_ids, int32_t* d_done, int32_t epoch, uint32_t* d_counter,
                  const int64_t* d_ws_off, int64_t max_nodes, const orx_workspace* ws,
                  const orx_route_cfg* cfg, int32_t max_groups, orx_stream_t stream);

/* Host planner of one iteration (no device work): disjoint-ROI wave of every
 * net, a schedule (by wave, larger ROI first) and the dependency lists of
 * orx_route_dag.  Inputs: box[4n], order[n] (routing order), board X x Y,
 * roi_nodes[n].  Outputs: wave[n], sched[n], dep_ptr[n+1], dep_ids[dep_cap].
 * Returns the number of dependencies (> dep_cap: nothing written to dep_ids,
 * call again with a larger buffer) or < 0 on bad input. */
int64_t orx_plan_iteration(const int32_t* box, const int32_t* order, int32_t n, int32_t X,
                           int32_t Y, const int64_t* roi_nodes, int32_t* wave, int32_t* sched,
                           int32_t* dep_ptr, int32_t* dep_ids, int64_t dep_cap);

/* Host planner, waves only: the same waves as orx_plan_iteration and the
 * schedule sched[n] (by wave, larger ROI first, then routing order),
 * without dependency lists, on a tiled grid (8x8-cell tiles) so large boxes
 * cost their tile count plus their border cells.  order[n] lists distinct
 * net ids below n_nets (all nets, or a hotset); box, roi_nodes and wave are
 * indexed by net id (n_nets entries; only the ordered nets' wave entries are
 * written).  Returns the number of waves, or < 0 on bad input.  No device
 * work. */
int64_t orx_plan_waves(const int32_t* box, int32_t n_nets, const int32_t* order, int32_t n, int32_t X, int32_t Y,
                       const int64_t* roi_nodes, int32_t* wave, int32_t* sched);

/* Per-net overuse of its committed path, summed over the forward arcs
 * (d_over[n]) and its hop count (d_hops[n]): the congestion input of
 * _order_nets_by_difficulty (UP:7840-7901).  d_node_over[n] (NULL to skip)
 * receives the overuse max(0, use - node_cap) summed over the path's nodes:
 * with d_over it marks the nets that hold an over-capacity resource, the
 * offenders a hotset reroutes (_build_hotset UP:9143-9171). */
int orx_net_overuse(const orx_lattice* lat, const orx_nets* nets, int32_t edge_cap, int32_t node_cap,
                    int64_t* d_over, int64_t* d_node_over, int32_t* d_hops, orx_stream_t stream);

/* Commit (sign=+1) or rip up (sign=-1) K device paths into the occupancy
 * arrays and re-price the touched arcs (EdgeAccountant.commit_path /
 * clear_path + refresh_live_present_costs, UP:1074-1135, 1403-1422;
 * _mark/_clear_path_node_use UP:4635-4659; via columns UP:4470-4546).
 * Paths of one batch may share arcs and nodes (the updates are atomic and
 * the re-pricing reads the final counts). */
int orx_apply_paths(const orx_lattice* lat, const orx_params* p, const int32_t* d_paths,
                    const int64_t* d_off, const int32_t* d_len, int32_t K, int32_t sign,
                    orx_stream_t stream);

/* One box-ROI query on the current wcost (the SSSP microbenchmark and the
 * array-reporting form of find_path_roi_gpu, CD:4951-5040).  Writes
 * ROI-local dist (fp32) and parent (global id, -1 root / unreached) arrays.
 * Small ROIs run in one workgroup; larger ones (or ORX_F_GRID) run grid-wide
 * as a device-side state machine replayed from a captured hipGraph (the
 * persistent kernel of PK:260-516 without inter-workgroup waits).
 * Synchronous: returns after the arrays are written. */
int orx_sssp_box(const orx_lattice* lat, int32_t x0, int32_t y0, int32_t w, int32_t h,
                 int64_t src, int64_t dst, float* d_dist, int32_t* d_parent,
                 const orx_workspace* ws, const orx_route_cfg* cfg, orx_stream_t stream);

/* orx_sssp_box without any allocation (no counterpart in the reference):
 * for callers that cannot call hipMalloc -- code under hipGraph capture,
 * tight serving loops, harnesses that forbid allocation.  The same query,
 * arrays and status as orx_sssp_box on its one-workgroup path, with the
 * search scratch (keys, queues, bitmaps, status word) in a static device
 * array of the library, one per GPU, held until the process ends.  ROIs up
 * to ORX_NOALLOC_MAX_NODES nodes; larger ones, and ORX_F_GRID /
 * ORX_F_TILES, are refused with ORX_E_ARG.  Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU (other streams, other
 * threads) race.  Synchronous, like orx_sssp_box. */
#define ORX_NOALLOC_MAX_NODES 65536
int orx_sssp_box_noalloc(const orx_lattice* lat, int32_t x0, int32_t y0, int32_t w, int32_t h, int64_t src,
                         int64_t dst, float* d_dist, int32_t* d_parent, const orx_route_cfg* cfg,
                         orx_stream_t stream);

/* Generic CSR query behind the gpu_solver seam:
 *   find_path_roi_gpu(src, dst, costs, roi_nodes, global_to_roi, node_penalty)
 *     (CD:4951-5040; UP:2465-2480), and
 *   find_path_fullgraph_gpu_seeds(costs, src_seeds, dst_targets, ...)
 *     (CD:5516-5850, PK:260-516).
 * The graph must carry in_rank[e] (rank of the arc's source among its
 * target's in-neighbours ordered by (source, arc)) and the transposed
 * in_indptr / in_src.  roi_nodes / g2r NULL = whole graph; pen indexed by
 * ROI index (NULL = none); allowed bitmap over global ids (NULL = all).
 * Enqueues the search; d_key[R] receives the packed dist/hops/slot keys and
 * d_best[0] the chosen target (global id, -1 none).  d_scratch needs
 * 4*R + 4*ceil(R/32) words.
 * cfg->flags ORX_F_GRID: the whole-chip form (the persistent full-graph
 * kernel of PK:260-516, driven from CD:5516-5850): grid-wide near-far step
 * launches as a device-side state machine, replayed from a captured graph
 * until done, with the same keys and chosen target as the one-workgroup
 * form; it synchronises with the caller's stream before it starts and
 * returns when the result is complete; d_scratch then needs
 * 4*R + 4*ceil(R/32) + 64 words. */
typedef struct orx_csr {
    int64_t N, E;
    const int32_t* indptr;
    const int32_t* indices;
    const int32_t* in_rank;
    const int32_t* in_indptr;
    const int32_t* in_src;
} orx_csr;

int orx_sssp_csr(const orx_csr* g, const float* d_costs, int64_t R, const int32_t* d_roi_nodes,
                 const int32_t* d_g2r, const float* d_pen, const uint32_t* d_allowed,
                 int32_t n_seeds, const int32_t* d_seeds, const float* d_seed_cost,
                 int32_t n_targets, const int32_t* d_targets, const float* d_target_cost,
                 uint64_t* d_key, uint32_t* d_scratch, int64_t scratch_words, int32_t* d_best,
                 const orx_route_cfg* cfg, orx_stream_t stream);

/* Bidirectional BFS region of interest on the device
 * (ROIExtractor.extract_roi_bfs UP:2257-2384; replaces its Python BFS and
 * the per-net N-sized global_to_roi allocation, UP:2381-2382).  The caller
 * passes the reference's derived limits: max_depth = min(initial_radius +
 * 2*stagnation_bonus, 80), halo = min(16, 2 + floor(max(0, bonus)/0.6)),
 * max_nodes (max_roi_nodes, 50 000), and d_keep[n_keep] = src, dst, then the
 * portal seeds not yet listed.  d_blocked[E] (or NULL) skips arcs.
 * Buffers: d_marks[ceil(N/16)] zero on entry (left zero), d_queue[4*qcap],
 * d_blocks[ceil(N/4096)], d_roi[cap], d_g2r[N] -1 outside earlier ROIs.
 * On return (stream order) d_roi[0 .. d_out[0]) is the ROI, ascending node
 * ids then the appended must-keep nodes, cut to max_nodes as the reference
 * does; d_g2r maps those nodes; d_out[1] depth, [2] meeting depth (-1),
 * [3] seen nodes, [4] 0 or an overflow flag (queue 1, list 2), [5] seen
 * nodes counted by the compaction.  orx_roi_release resets their d_g2r
 * entries to -1; call it before the next orx_roi_bfs on the same buffers. */
int orx_roi_bfs(const orx_csr* g, const uint8_t* d_blocked, int64_t src, int64_t dst, int32_t max_depth,
                int32_t halo, int64_t max_nodes, const int32_t* d_keep, int32_t n_keep, uint32_t* d_marks,
                int32_t* d_queue, int64_t qcap, uint32_t* d_blocks, int32_t* d_roi, int64_t cap, int32_t* d_g2r,
                int64_t* d_out, orx_stream_t stream);
int orx_roi_release(const int32_t* d_roi, const int64_t* d_out, int64_t cap, int32_t* d_g2r, orx_stream_t stream);

/* Parent walk of a CSR query result (backtrace_parent CD:5857-5900,
 * backtrace_paths CD:1845-1950) from the target in d_target[0] (device, e.g.
 * the d_best of orx_sssp_csr): writes the global path, root first, into
 * d_path[cap]; d_len[0] = length, ORX_NO_PATH or a negative status. */
int orx_backtrace_csr(const orx_csr* g, const uint64_t* d_key, const int32_t* d_g2r,
                      const int32_t* d_target, int32_t* d_path, int32_t cap, int32_t* d_len,
                      orx_stream_t stream);

/* The same walk on the whole chip (backtrace_parent CD:5857-5900 walks one
 * lane per net): the parent of every ROI node from its key slot, 32-hop
 * jumps by pointer doubling, then the segments between nodes 32 hops apart
 * filled in parallel; a path of h hops costs ~h/32 + 32 dependent loads.
 * roi_nodes / g2r as for orx_sssp_csr (NULL = whole graph, R == N);
 * d_scratch needs 3*R words (the search scratch may be reused).  Same
 * outputs as orx_backtrace_csr. */
int orx_backtrace_csr_roi(const orx_csr* g, const uint64_t* d_key, int64_t R, const int32_t* d_roi_nodes,
                          const int32_t* d_g2r, const int32_t* d_target, int32_t* d_path, int32_t cap,
                          int32_t* d_len, uint32_t* d_scratch, int64_t scratch_words, orx_stream_t stream);

/* ---- Per-arc accountant on the CSR arc order (EdgeAccountant, UP:1039-1422).
 * Device arrays of E entries indexed by CSR arc id.  The arithmetic is the
 * reference's float32 NumPy evaluation, operation for operation; coefficients
 * are passed as the float32 roundings of the reference's Python floats.
 * Every call is asynchronous on `stream`; outputs marked d_ are device memory
 * the caller clears when they accumulate. */
typedef struct orx_acct {
    int64_t E;
    float* present;      /* committed uses per arc (EdgeAccountant.present)      */
    float* present_ema;  /* smoothed present                                    */
    float* history;
    float* capacity;
    float* total_cost;   /* may be NULL until orx_acct_costs is used            */
    float* live_static;  /* static part under live re-pricing (NULL if unused)  */
    int32_t* canon;      /* canonical use count per arc (EdgeAccountant.canonical) */
} orx_acct;

/* commit_path (sign=+1, UP:1099-1120) / clear_path (sign=-1, UP:1122-1135)
 * of the n arc ids d_idx (a batch of paths may be concatenated: the result
 * equals committing / clearing them one by one).  floor_present applies the
 * reference's whole-array floor np.maximum(present, 0) after a clear.  Ids
 * outside [0, E) are skipped and counted into *d_bad (may be NULL). */
int orx_acct_apply(const orx_acct* a, const int64_t* d_idx, int64_t n, int32_t sign, int32_t floor_present,
                   unsigned long long* d_bad, orx_stream_t stream);
/* compute_overuse's arc term (UP:1137-1167): *d_sum += sum of
 * max(0, present - capacity) in float64, *d_count += arcs with nonzero
 * overuse; d_mask (NULL = all) restricts both to the canonical arcs
 * (_canonical_edge_resource_mask). */
int orx_acct_overuse(const orx_acct* a, const uint8_t* d_mask, double* d_sum, unsigned long long* d_count,
                     orx_stream_t stream);
/* *d_out += float64 sum of: which=0 history; 1 max(0, present_ema - capacity)
 * (cost_balance_ratio, UP:1054-1067); 2 |present - canon|
 * (verify_present_matches_canonical, UP:1192-1224). */
int orx_acct_sum(const orx_acct* a, int32_t which, double* d_out, orx_stream_t stream);
/* update_present_ema (UP:1291-1299); one_minus_beta = float32(1 - beta). */
int orx_acct_present_ema(const orx_acct* a, float beta, float one_minus_beta, orx_stream_t stream);
/* update_history (UP:1226-1289); d_base NULL = no increment cap. */
int orx_acct_history(const orx_acct* a, float gain, const float* d_base, float cap_mult, float decay,
                     int32_t use_raw_present, orx_stream_t stream);
/* update_costs (UP:1301-1384) into total_cost.  Layer bias applies when
 * d_edge_layer, d_layer_bias[n_layers] and d_edge_kind are all given; then
 * d_scale_out receives the per-arc present-cost scale pres_fac * bias.
 * d_blocked (NULL = none) marks hard-blocked arcs (+inf, the _blocked_edges
 * mask); jitter != 0 adds (arc mod 9973) * 1e-6. */
int orx_acct_costs(const orx_acct* a, const float* d_base, float pres_fac, float hist_w, float via_mult,
                   float base_w, const int32_t* d_edge_layer, const float* d_layer_bias, int32_t n_layers,
                   const int32_t* d_edge_kind, const uint8_t* d_blocked, int32_t jitter, float* d_scale_out,
                   orx_stream_t stream);
/* begin_live_present_costs (UP:1386-1401) / refresh_live_present_costs
 * (UP:1403-1422) with the present-cost scale `scale`, or per arc d_scale. */
int orx_acct_begin_live(const orx_acct* a, float scale, const float* d_scale, orx_stream_t stream);
int orx_acct_refresh_live(const orx_acct* a, const int64_t* d_idx, int64_t n, float scale, const float* d_scale,
                          unsigned long long* d_bad, orx_stream_t stream);
/* refresh_from_canonical (UP:1073-1097): present = canon. */
int orx_acct_from_canonical(const orx_acct* a, orx_stream_t stream);

/* Via capacity kernels on a CSR cost array (ViaKernelManager,
 * pathfinder/via_kernels.py:186-325; arithmetic of the CPU forms
 * via_kernels.py:567-651).  Via arc i: CSR id d_edge[i], column
 * d_xy[2i], d_xy[2i+1] (column index x * Ny + y), segments z - 1 for z in
 * [d_z_lo[i], d_z_hi[i]).  d_count[0] += arcs blocked / penalised,
 * d_count[1] += entries with a column or arc id out of range (skipped).
 * Segment arrays may be NULL (no segment term). */
int orx_via_hard_block(const int32_t* d_edge, const int32_t* d_xy, const int32_t* d_z_lo, const int32_t* d_z_hi,
                       int32_t n, int32_t Nx, int32_t Ny, int32_t segZ, const int16_t* d_col_use,
                       const int16_t* d_col_cap, const int8_t* d_seg_use, const int8_t* d_seg_cap, float* d_total,
                       int64_t E, unsigned long long* d_count, orx_stream_t stream);
int orx_via_penalties(const int32_t* d_edge, const int32_t* d_xy, const int32_t* d_z_lo, const int32_t* d_z_hi,
                      int32_t n, int32_t Nx, int32_t Ny, int32_t segZ, const float* d_col_pres,
                      const float* d_seg_pres, float col_w, float seg_w, float* d_total, int64_t E,
                      unsigned long long* d_count, orx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* ORTHOROUTE_HIP_H */
