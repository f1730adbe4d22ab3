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
