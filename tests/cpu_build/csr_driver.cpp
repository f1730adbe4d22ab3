// csr_driver.cpp -- one orx_sssp_csr query plus its parent walk, of the
// liborx_hip kernel sources on host threads (hip_cpu.hpp), every buffer of
// exact size (AddressSanitizer).
// usage: csr_driver <dir>; <dir>/meta.txt: N E R n_seeds n_tg flags block
// delta roi (0/1) pen (0/1) seed_cost (0/1) tg_cost (0/1) walk (0: one lane,
// 1: parallel); arrays indptr indices in_rank in_indptr in_src costs seeds
// targets [seed_cost tg_cost roi_nodes g2r pen] as raw files.  Writes
// keys.out (u64[R]), best.out (i32), path.out (i32[R]), len.out (i32).
// Test infrastructure.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/orthoroute_hip.h"

static std::string g_dir;

template <typename T>
static bool load(const char* name, std::vector<T>& v, size_t n) {
    v.assign(n ? n : 1, T{});
    if (n == 0) return true;
    FILE* f = fopen((g_dir + "/" + name).c_str(), "rb");
    if (!f) return false;
    const bool ok = fread(v.data(), sizeof(T), n, f) == n;
    fclose(f);
    v.resize(n);
    return ok;
}

template <typename T>
static void save(const char* name, const T* p, size_t n) {
    FILE* f = fopen((g_dir + "/" + name).c_str(), "wb");
    fwrite(p, sizeof(T), n, f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    g_dir = argv[1];
    std::map<std::string, double> m;
    FILE* f = fopen((g_dir + "/meta.txt").c_str(), "r");
    if (!f) return 2;
    char k[128];
    double v;
    while (fscanf(f, "%127s %lf", k, &v) == 2) m[k] = v;
    fclose(f);
    const int64_t N = (int64_t)m["N"], E = (int64_t)m["E"], R = (int64_t)m["R"];
    const int32_t ns = (int32_t)m["n_seeds"], nt = (int32_t)m["n_tg"];
    std::vector<int32_t> indptr, indices, in_rank, in_indptr, in_src, seeds, targets, roi, g2r;
    std::vector<float> costs, seed_cost, tg_cost, pen;
    if (!load("indptr", indptr, N + 1) || !load("indices", indices, E) || !load("in_rank", in_rank, E) ||
        !load("in_indptr", in_indptr, N + 1) || !load("in_src", in_src, E) || !load("costs", costs, E) ||
        !load("seeds", seeds, ns) || !load("targets", targets, nt))
        return 2;
    const bool has_roi = m["roi"] != 0.0, has_pen = m["pen"] != 0.0;
    const bool has_sc = m["seed_cost"] != 0.0, has_tc = m["tg_cost"] != 0.0;
    if (has_roi && (!load("roi_nodes", roi, R) || !load("g2r", g2r, N))) return 2;
    if (has_pen && !load("pen", pen, R)) return 2;
    if (has_sc && !load("seed_cost", seed_cost, ns)) return 2;
    if (has_tc && !load("tg_cost", tg_cost, nt)) return 2;
    orx_csr g{N, E, indptr.data(), indices.data(), in_rank.data(), in_indptr.data(), in_src.data()};
    const int64_t words = (R + 31) / 32;
    const int64_t sw = (4 * R + 4 * words + 1) / 2 * 2 + 64;
    std::vector<uint32_t> scratch((size_t)sw);
    std::vector<uint64_t> keys((size_t)R);
    std::vector<int32_t> path((size_t)R);
    int32_t best = -7, len = -7;
    orx_route_cfg cfg{(float)m["delta"], (int32_t)m["block"], 0, (int32_t)m["flags"], nullptr};
    if (orx_init() != ORX_OK) return 3;
    int rc = orx_sssp_csr(&g, costs.data(), R, has_roi ? roi.data() : nullptr, has_roi ? g2r.data() : nullptr,
                          has_pen ? pen.data() : nullptr, nullptr, ns, seeds.data(), has_sc ? seed_cost.data() : nullptr,
                          nt, targets.data(), has_tc ? tg_cost.data() : nullptr, keys.data(), scratch.data(), sw, &best,
                          &cfg, nullptr);
    if (rc < 0) { fprintf(stderr, "orx_sssp_csr -> %d: %s\n", rc, orx_last_error()); return 4; }
    if (m["walk"] != 0.0)
        rc = orx_backtrace_csr_roi(&g, keys.data(), R, has_roi ? roi.data() : nullptr, has_roi ? g2r.data() : nullptr,
                                   &best, path.data(), (int32_t)R, &len, scratch.data(), sw, nullptr);
    else
        rc = orx_backtrace_csr(&g, keys.data(), has_roi ? g2r.data() : nullptr, &best, path.data(), (int32_t)R, &len,
                               nullptr);
    if (rc < 0) { fprintf(stderr, "backtrace -> %d: %s\n", rc, orx_last_error()); return 5; }
    save("keys.out", keys.data(), keys.size());
    save("best.out", &best, 1);
    save("path.out", path.data(), path.size());
    save("len.out", &len, 1);
    return 0;
}
