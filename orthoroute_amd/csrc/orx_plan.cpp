// orx_plan.cpp -- host-side planner of one PathFinder iteration (C ABI).
//
// Given the iteration's routing order and the nets' ROI boxes, computes
//   wave[k]  : 1 + the largest wave of any earlier net whose box overlaps net
//              k's box (0 for none) -- the disjoint-ROI waves;
//   sched[]  : a topological order of the nets (by wave, larger ROI first
//              inside a wave);
//   deps     : for each net, the distinct nets that last covered a cell of its
//              box before it in the routing order (CSR: dep_ptr / dep_ids).
// Waiting for those last writers implies waiting for every earlier
// overlapping net (each of them is a dependency of some last writer), so
// routing every net after its dependencies reproduces the sequential order's
// result exactly: a net reads and writes only arcs and nodes inside its box.
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/orthoroute_hip.h"

extern "C" int64_t orx_plan_iteration(const int32_t* box, const int32_t* order, int32_t n, int32_t X, int32_t Y,
                                      const int64_t* roi_nodes, int32_t* wave, int32_t* sched, int32_t* dep_ptr,
                                      int32_t* dep_ids, int64_t dep_cap) {
    if (n < 0 || X <= 0 || Y <= 0) return -1;
    std::vector<int32_t> last((size_t)X * Y, -1);   // last net covering each cell
    std::vector<int32_t> lastw((size_t)X * Y, -1);  // its wave
    std::vector<int32_t> mark((size_t)n, -1);
    std::vector<std::vector<int32_t>> deps((size_t)n);
    int32_t nw = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t k = order[i];
        const int32_t* b = box + 4 * (int64_t)k;
        const int32_t x0 = b[0], y0 = b[1], w = b[2], h = b[3];
        if (x0 < 0 || y0 < 0 || w < 1 || h < 1 || x0 + w > X || y0 + h > Y) return -2;
        int32_t wk = 0;
        std::vector<int32_t>& dk = deps[(size_t)k];
        for (int32_t y = y0; y < y0 + h; ++y) {
            const int32_t* lrow = last.data() + (size_t)y * X;
            const int32_t* wrow = lastw.data() + (size_t)y * X;
            for (int32_t x = x0; x < x0 + w; ++x) {
                const int32_t p = lrow[x];
                if (p >= 0 && mark[(size_t)p] != k) {
                    mark[(size_t)p] = k;
                    dk.push_back(p);
                    wk = std::max(wk, wrow[x] + 1);
                }
            }
        }
        for (int32_t y = y0; y < y0 + h; ++y) {
            std::fill(last.begin() + (size_t)y * X + x0, last.begin() + (size_t)y * X + x0 + w, k);
            std::fill(lastw.begin() + (size_t)y * X + x0, lastw.begin() + (size_t)y * X + x0 + w, wk);
        }
        wave[k] = wk;
        nw = std::max(nw, wk + 1);
    }
    // schedule: by wave, larger ROI first, then routing order
    std::vector<int32_t> pos((size_t)n);
    for (int32_t i = 0; i < n; ++i) pos[(size_t)order[i]] = i;
    std::vector<int32_t> s(order, order + n);
    std::stable_sort(s.begin(), s.end(), [&](int32_t a, int32_t c) {
        if (wave[a] != wave[c]) return wave[a] < wave[c];
        if (roi_nodes[a] != roi_nodes[c]) return roi_nodes[a] > roi_nodes[c];
        return pos[(size_t)a] < pos[(size_t)c];
    });
    std::memcpy(sched, s.data(), sizeof(int32_t) * (size_t)n);
    int64_t total = 0;
    for (int32_t k = 0; k < n; ++k) total += (int64_t)deps[(size_t)k].size();
    if (total > dep_cap) return total;  // caller grows dep_ids and retries
    int64_t at = 0;
    for (int32_t k = 0; k < n; ++k) {
        dep_ptr[k] = (int32_t)at;
        for (int32_t p : deps[(size_t)k]) dep_ids[at++] = p;
    }
    dep_ptr[n] = (int32_t)at;
    (void)nw;
    return total;
}

// Waves only (no dependency lists), for the per-iteration launch plan.
// wave[k] = 1 + the largest wave of any earlier net (in `order`) whose box
// overlaps net k's box -- the same waves as orx_plan_iteration -- computed on
// 8x8-cell tiles.  Each tile keeps the wave of the last box covering all of
// it (tfull), the largest wave anywhere in it (tmax), and the boxes that
// covered part of it since as up to K stamps (a rectangle and its wave);
// when the stamps are full they are written into the tile's cells.  A box
// visits each tile it touches once to read (skipped when tmax is no larger
// than what it has found) and once to write (one stamp, or tfull).  The
// wave of a cell is the max of tfull, the stamps covering it and the cell.
// Every box's wave exceeds every value inside it, so values only rise.
extern "C" int64_t orx_plan_waves(const int32_t* box, int32_t n_nets, const int32_t* order, int32_t n, int32_t X,
                                  int32_t Y, const int64_t* roi_nodes, int32_t* wave, int32_t* sched) {
    if (n < 0 || n > n_nets || X <= 0 || Y <= 0) return -1;
    for (int32_t i = 0; i < n; ++i)
        if (order[i] < 0 || order[i] >= n_nets) return -3;
    constexpr int32_t T = 8, K = 8;   // tile side in cells, stamps per tile
    struct Stamp { uint8_t x0, x1, y0, y1; int32_t w; };
    const int32_t TX = (X + T - 1) / T, TY = (Y + T - 1) / T;
    const size_t NT = (size_t)TX * TY;
    std::vector<int32_t> cell(NT * T * T, -1), tfull(NT, -1), tmax(NT, -1);
    std::vector<Stamp> stamps(NT * K);
    std::vector<uint8_t> ns(NT, 0);
    int32_t nw = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t k = order[i];
        const int32_t* b = box + 4 * (int64_t)k;
        const int32_t x0 = b[0], y0 = b[1], x1 = b[0] + b[2], y1 = b[1] + b[3];
        if (x0 < 0 || y0 < 0 || b[2] < 1 || b[3] < 1 || x1 > X || y1 > Y) return -2;
        const int32_t tx0 = x0 / T, tx1 = (x1 - 1) / T, ty0 = y0 / T, ty1 = (y1 - 1) / T;
        int32_t m = -1;
        for (int32_t ty = ty0; ty <= ty1; ++ty) {
            const int32_t cy0 = std::max(y0, ty * T) - ty * T, cy1 = std::min(y1, ty * T + T) - ty * T;
            const bool full_y = cy0 == 0 && cy1 == std::min(T, Y - ty * T);
            for (int32_t tx = tx0; tx <= tx1; ++tx) {
                const size_t t = (size_t)ty * TX + tx;
                if (tmax[t] <= m) continue;
                const int32_t cx0 = std::max(x0, tx * T) - tx * T, cx1 = std::min(x1, tx * T + T) - tx * T;
                if (full_y && cx0 == 0 && cx1 == std::min(T, X - tx * T)) {
                    m = tmax[t];
                    continue;
                }
                int32_t cm = tfull[t];
                const Stamp* st = stamps.data() + t * K;
                for (int32_t q = 0; q < ns[t]; ++q)
                    if (st[q].w > cm && st[q].x0 < cx1 && cx0 < st[q].x1 && st[q].y0 < cy1 && cy0 < st[q].y1) cm = st[q].w;
                const int32_t* c = cell.data() + t * (T * T);
                for (int32_t y = cy0; y < cy1; ++y)
                    for (int32_t x = cx0; x < cx1; ++x) cm = std::max(cm, c[y * T + x]);
                m = std::max(m, cm);
            }
        }
        const int32_t wk = m + 1;
        wave[k] = wk;
        nw = std::max(nw, wk + 1);
        for (int32_t ty = ty0; ty <= ty1; ++ty) {
            const int32_t cy0 = std::max(y0, ty * T) - ty * T, cy1 = std::min(y1, ty * T + T) - ty * T;
            const bool full_y = cy0 == 0 && cy1 == std::min(T, Y - ty * T);
            for (int32_t tx = tx0; tx <= tx1; ++tx) {
                const size_t t = (size_t)ty * TX + tx;
                const int32_t cx0 = std::max(x0, tx * T) - tx * T, cx1 = std::min(x1, tx * T + T) - tx * T;
                tmax[t] = std::max(tmax[t], wk);
                if (full_y && cx0 == 0 && cx1 == std::min(T, X - tx * T)) {
                    tfull[t] = wk;
                    ns[t] = 0;
                    continue;
                }
                Stamp* st = stamps.data() + t * K;
                if (ns[t] == K) {
                    int32_t* c = cell.data() + t * (T * T);
                    for (int32_t q = 0; q < K; ++q)
                        for (int32_t y = st[q].y0; y < st[q].y1; ++y)
                            for (int32_t x = st[q].x0; x < st[q].x1; ++x) c[y * T + x] = std::max(c[y * T + x], st[q].w);
                    ns[t] = 0;
                }
                st[ns[t]++] = Stamp{(uint8_t)cx0, (uint8_t)cx1, (uint8_t)cy0, (uint8_t)cy1, wk};
            }
        }
    }
    // schedule: by wave, larger ROI first, then routing order (counting sort
    // on the wave, a stable sort inside each wave)
    std::vector<int32_t> start((size_t)nw + 1, 0);
    for (int32_t i = 0; i < n; ++i) start[(size_t)wave[order[i]] + 1]++;
    for (int32_t w = 0; w < nw; ++w) start[(size_t)w + 1] += start[(size_t)w];
    std::vector<int32_t> fill(start.begin(), start.end() - 1);
    for (int32_t i = 0; i < n; ++i) sched[fill[(size_t)wave[order[i]]]++] = order[i];
    for (int32_t w = 0; w < nw; ++w)
        std::stable_sort(sched + start[(size_t)w], sched + start[(size_t)w + 1],
                         [&](int32_t a, int32_t c) { return roi_nodes[a] > roi_nodes[c]; });
    return nw;
}
