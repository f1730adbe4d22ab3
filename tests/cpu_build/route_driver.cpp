// route_driver.cpp -- runs one PathFinder iteration of the liborx_hip kernel
// sources on host threads (hip_cpu.hpp) through the C ABI, on buffers of
// exactly their documented sizes, so AddressSanitizer reports any access
// outside them.  Test infrastructure.
//
// usage: route_driver <dir> ; <dir> holds meta.txt (key value lines) and the
// raw little-endian arrays named below; updated state arrays are written
// back as <name>.out.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/orthoroute_hip.h"

static std::string g_dir;
static std::map<std::string, double> g_meta;

template <typename T>
static std::vector<T> load(const char* name, size_t n) {
    std::vector<T> v(n);
    std::string p = g_dir + "/" + name;
    FILE* f = fopen(p.c_str(), "rb");
    if (!f) { fprintf(stderr, "missing %s\n", p.c_str()); exit(2); }
    size_t got = fread(v.data(), sizeof(T), n, f);
    fclose(f);
    if (got != n) { fprintf(stderr, "%s: %zu of %zu\n", name, got, n); exit(2); }
    return v;
}
template <typename T>
static void save(const char* name, const std::vector<T>& v) {
    std::string p = g_dir + "/" + name + std::string(".out");
    FILE* f = fopen(p.c_str(), "wb");
    fwrite(v.data(), sizeof(T), v.size(), f);
    fclose(f);
}
static int64_t M(const char* k) {
    auto it = g_meta.find(k);
    if (it == g_meta.end()) { fprintf(stderr, "meta %s missing\n", k); exit(2); }
    return (int64_t)it->second;
}
static double Mf(const char* k) { return g_meta.at(k); }

#define CHECK(call)                                                                 \
    do {                                                                            \
        int _rc = (call);                                                           \
        if (_rc != ORX_OK) {                                                        \
            fprintf(stderr, "%s -> %d: %s\n", #call, _rc, orx_last_error());       \
            return 3;                                                               \
        }                                                                           \
    } while (0)

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    g_dir = argv[1];
    {
        FILE* f = fopen((g_dir + "/meta.txt").c_str(), "r");
        if (!f) return 2;
        char key[128];
        double val;
        while (fscanf(f, "%127s %lf", key, &val) == 2) g_meta[key] = val;
        fclose(f);
    }
    const int32_t X = (int32_t)M("X"), Y = (int32_t)M("Y"), L = (int32_t)M("L");
    const int64_t N = (int64_t)X * Y * L;
    const int32_t n = (int32_t)M("n");
    auto base = load<float>("base", 4 * N);
    auto hist = load<float>("hist", 4 * N);
    auto occ = load<int32_t>("occ", 4 * N);
    auto wcost = load<float>("wcost", 4 * N);
    auto node_hist = load<float>("node_hist", N);
    auto node_use = load<int32_t>("node_use", N);
    auto via_col = load<int32_t>("via_col", (size_t)X * Y);
    auto box = load<int32_t>("box", 4 * (size_t)n);
    auto src = load<int32_t>("src", n);
    auto dst = load<int32_t>("dst", n);
    auto arena = load<int32_t>("path_arena", (size_t)M("arena"));
    auto path_off = load<int64_t>("path_off", n);
    auto path_cap = load<int32_t>("path_cap", n);
    auto path_len = load<int32_t>("path_len", n);
    auto path_cost = load<float>("path_cost", n);
    auto status = load<int32_t>("status", n);
    auto sched = load<int32_t>("sched", n);
    auto dep_ptr = load<int32_t>("dep_ptr", (size_t)n + 1);
    auto dep_ids = load<int32_t>("dep_ids", (size_t)std::max<int64_t>(1, M("n_deps")));
    auto ws_off = load<int64_t>("ws_off", n);
    auto done = load<int32_t>("done", n);
    const int64_t wsn = std::max<int64_t>(M("ws_nodes") + 512, 1024);
    std::vector<uint64_t> keys((size_t)wsn);
    std::vector<uint32_t> queues((size_t)(4 * wsn));
    std::vector<uint32_t> bitmaps((size_t)(3 * wsn / 32));
    std::vector<uint32_t> counter(1);
    std::vector<int64_t> stats(8, 0), iter_out(8, 0);

    orx_lattice lat{X, Y, L, (uint32_t)M("hmask"), base.data(), hist.data(), occ.data(), wcost.data(),
                    node_hist.data(), node_use.data(), via_col.data()};
    orx_params P{(float)Mf("pres_fac"), (float)Mf("base_w"), (float)Mf("hist_w"), (float)Mf("node_pres"),
                 (float)Mf("node_hist_w"), (int32_t)M("edge_cap"), (int32_t)M("node_cap"), (float)Mf("hist_gain"),
                 (float)Mf("hist_decay"), (float)Mf("hist_cap_mult"), (float)Mf("node_hist_gain"),
                 (float)Mf("astar_cp"), (float)Mf("astar_cv"), (float)Mf("astar_gmax")};
    orx_nets nets{n, box.data(), src.data(), dst.data(), arena.data(), path_off.data(), path_cap.data(),
                  path_len.data(), path_cost.data(), status.data()};
    // multi-terminal tables, each of exactly its size
    std::vector<int32_t> seed_ptr, seed_node, tgt_ptr, tgt_node;
    std::vector<float> seed_cost, tgt_cost;
    if (g_meta.count("terms") && M("terms")) {
        seed_ptr = load<int32_t>("seed_ptr", (size_t)n + 1);
        tgt_ptr = load<int32_t>("tgt_ptr", (size_t)n + 1);
        seed_node = load<int32_t>("seed_node", (size_t)M("n_seed"));
        seed_cost = load<float>("seed_cost", (size_t)M("n_seed"));
        tgt_node = load<int32_t>("tgt_node", (size_t)M("n_tgt"));
        tgt_cost = load<float>("tgt_cost", (size_t)M("n_tgt"));
        nets.seed_ptr = seed_ptr.data(); nets.seed_node = seed_node.data(); nets.seed_cost = seed_cost.data();
        nets.tgt_ptr = tgt_ptr.data(); nets.tgt_node = tgt_node.data(); nets.tgt_cost = tgt_cost.data();
    }
    orx_workspace ws{keys.data(), queues.data(), bitmaps.data(), wsn};
    orx_route_cfg cfg{(float)Mf("delta"), (int32_t)M("block"), (int32_t)M("lds_nodes"), (int32_t)M("flags"),
                      stats.data()};
    CHECK(orx_init());
    CHECK(orx_lattice_reprice(&lat, &P, nullptr));
    // 0 = dependency-driven launch, 1 = one launch per net in sched order,
    // 2 = the wave sequence through orx_route_waves (sched is ordered by wave)
    const int64_t mode = M("mode");
    if (mode == 2 || mode == 3) {
        auto wave = load<int32_t>("wave", n);
        std::vector<int32_t> ptr{0};
        std::vector<int64_t> wmax, wws, offs((size_t)n);
        int64_t acc = 0, mx = 0;
        for (int32_t i = 0; i < n; ++i) {
            const int32_t k = sched[i];
            if (i > 0 && wave[k] != wave[sched[i - 1]]) {
                ptr.push_back(i); wmax.push_back(mx); wws.push_back(acc);
                acc = 0; mx = 0;
            }
            const int64_t R = (int64_t)box[4 * (size_t)k + 2] * box[4 * (size_t)k + 3] * L;
            offs[(size_t)i] = acc;
            acc += (R + 63) / 64 * 64;
            mx = std::max(mx, R);
        }
        ptr.push_back(n); wmax.push_back(mx); wws.push_back(acc);
        // a broken plan for the refusal tests: one slot's workspace offset
        // off its 64-node alignment
        if (g_meta.count("bad_slot")) offs[(size_t)M("bad_slot")] += 1;
        if (mode == 2) {
            CHECK(orx_route_waves(&lat, &P, &nets, sched.data(), offs.data(), (int32_t)wmax.size(), ptr.data(),
                                  wmax.data(), wws.data(), &ws, &cfg, nullptr));
        } else {
            // tile engine: step budget per wave = the largest tile count along
            // x plus along y, capped by tile_steps_cap (< 0: no cap), so
            // the single-workgroup drain runs the rest
            const int64_t capk = g_meta.count("tile_steps_cap") ? M("tile_steps_cap") : -1;
            const int32_t ty = (L == 6 || L == 8 || L == 12 || L == 16 || L == 20) ? 16 : 10;
            std::vector<int32_t> steps(wmax.size(), 0);
            for (size_t w = 0; w < wmax.size(); ++w) {
                for (int32_t i = ptr[w]; i < ptr[w + 1]; ++i) {
                    const int32_t k = sched[i];
                    const int32_t bw = box[4 * (size_t)k + 2], bh = box[4 * (size_t)k + 3];
                    steps[w] = std::max(steps[w], (bw + 15) / 16 + (bh + ty - 1) / ty);
                }
                if (capk >= 0) steps[w] = std::min<int32_t>(steps[w], (int32_t)capk);
            }
            CHECK(orx_route_tiles(&lat, &P, &nets, sched.data(), offs.data(), (int32_t)wmax.size(), ptr.data(),
                                  wws.data(), steps.data(), &ws, &cfg, nullptr));
        }
    } else if (mode == 0) {
        CHECK(orx_route_dag(&lat, &P, &nets, sched.data(), n, dep_ptr.data(), dep_ids.data(), done.data(),
                            (int32_t)M("epoch"), counter.data(), ws_off.data(), M("max_nodes"), &ws, &cfg,
                            (int32_t)M("max_groups"), nullptr));
    } else {
        for (int32_t i = 0; i < n; ++i) {
            std::vector<int32_t> one(1, sched[i]);
            std::vector<int64_t> off(1, ws_off[sched[i]]);
            CHECK(orx_route_wave(&lat, &P, &nets, one.data(), off.data(), 1, M("max_nodes"), &ws, &cfg, nullptr));
        }
    }
    CHECK(orx_lattice_end_iteration(&lat, &P, iter_out.data(), nullptr));
    save("hist", hist); save("occ", occ); save("wcost", wcost); save("node_hist", node_hist);
    save("node_use", node_use); save("via_col", via_col); save("path_arena", arena); save("path_len", path_len);
    save("path_cost", path_cost); save("status", status); save("done", done); save("stats", stats);
    save("iter_out", iter_out);
    return 0;
}
