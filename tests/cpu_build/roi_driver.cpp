// roi_driver.cpp -- one orx_roi_bfs call of the liborx_hip kernel sources on
// host threads (hip_cpu.hpp), every buffer of exactly its documented size
// (AddressSanitizer), then orx_roi_release.  Test infrastructure.
// usage: roi_driver <dir>; <dir>/meta.txt: N E src dst max_depth halo
// max_nodes n_keep blocked; arrays indptr (i32[N+1]), indices (i32[E]),
// keep (i32[n_keep]), blocked (u8[E], when blocked = 1).  Writes roi.out
// (the ROI), g2r.out (i32[N] before release), out.out (i64[8]) and clean.out
// (i64[2]: marks words and g2r entries not back at their idle value).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/orthoroute_hip.h"

template <typename T>
static std::vector<T> load(const std::string& path, size_t n) {
    std::vector<T> v(n);
    FILE* f = fopen(path.c_str(), "rb");
    if (!f || fread(v.data(), sizeof(T), n, f) != n) { fprintf(stderr, "bad %s\n", path.c_str()); exit(2); }
    fclose(f);
    return v;
}
template <typename T>
static void save(const std::string& path, const T* p, size_t n) {
    FILE* f = fopen(path.c_str(), "wb");
    fwrite(p, sizeof(T), n, f);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) return 1;
    std::string dir = argv[1];
    std::map<std::string, double> m;
    FILE* f = fopen((dir + "/meta.txt").c_str(), "r");
    if (!f) return 2;
    char k[128];
    double v;
    while (fscanf(f, "%127s %lf", k, &v) == 2) m[k] = v;
    fclose(f);
    const int64_t N = (int64_t)m["N"], E = (int64_t)m["E"];
    const int32_t n_keep = (int32_t)m["n_keep"];
    auto indptr = load<int32_t>(dir + "/indptr", (size_t)N + 1);
    auto indices = load<int32_t>(dir + "/indices", (size_t)std::max<int64_t>(E, 1));
    auto keep = load<int32_t>(dir + "/keep", (size_t)n_keep);
    std::vector<uint8_t> blocked;
    if (m["blocked"] != 0) blocked = load<uint8_t>(dir + "/blocked", (size_t)E);
    const int64_t qcap = N, cap = N + ORX_MAX_TARGETS + 2;
    std::vector<uint32_t> marks((size_t)((N + 15) / 16), 0u);
    std::vector<int32_t> queue((size_t)(4 * qcap));
    std::vector<uint32_t> blocks((size_t)std::max<int64_t>(1, (N + 4095) / 4096));
    std::vector<int32_t> roi((size_t)cap), g2r((size_t)N, -1);
    std::vector<int64_t> out(8, 0);
    orx_csr g{N, E, indptr.data(), indices.data(), nullptr, nullptr, nullptr};
    if (orx_init() != ORX_OK) return 3;
    int rc = orx_roi_bfs(&g, blocked.empty() ? nullptr : blocked.data(), (int64_t)m["src"], (int64_t)m["dst"],
                         (int32_t)m["max_depth"], (int32_t)m["halo"], (int64_t)m["max_nodes"], keep.data(), n_keep,
                         marks.data(), queue.data(), qcap, blocks.data(), roi.data(), cap, g2r.data(), out.data(),
                         nullptr);
    if (rc < 0) { fprintf(stderr, "orx_roi_bfs -> %d: %s\n", rc, orx_last_error()); return 4; }
    save(dir + "/roi.out", roi.data(), (size_t)out[0]);
    save(dir + "/g2r.out", g2r.data(), g2r.size());
    save(dir + "/out.out", out.data(), out.size());
    rc = orx_roi_release(roi.data(), out.data(), cap, g2r.data(), nullptr);
    if (rc < 0) { fprintf(stderr, "orx_roi_release -> %d: %s\n", rc, orx_last_error()); return 5; }
    int64_t clean[2] = {0, 0};
    for (uint32_t w : marks) clean[0] += w != 0u;
    for (int32_t x : g2r) clean[1] += x != -1;
    save(dir + "/clean.out", clean, 2);
    return 0;
}
