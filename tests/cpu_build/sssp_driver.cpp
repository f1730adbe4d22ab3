// sssp_driver.cpp -- one orx_sssp_box query of the liborx_hip kernel sources
// on host threads (hip_cpu.hpp), buffers of exact size (AddressSanitizer).
// usage: sssp_driver <dir>; <dir>/meta.txt: X Y L hmask x0 y0 w h src dst
// flags block delta; <dir>/wcost (f32[4N]).  Writes dist.out, parent.out,
// stats.out.  Test infrastructure.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/orthoroute_hip.h"

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
    const int32_t X = (int32_t)m["X"], Y = (int32_t)m["Y"], L = (int32_t)m["L"];
    const int64_t N = (int64_t)X * Y * L;
    std::vector<float> wcost((size_t)(4 * N));
    f = fopen((dir + "/wcost").c_str(), "rb");
    if (!f || fread(wcost.data(), 4, wcost.size(), f) != wcost.size()) return 2;
    fclose(f);
    const int32_t w = (int32_t)m["w"], h = (int32_t)m["h"];
    const int64_t R = (int64_t)w * h * L;
    // room for the ROI plus the tile engine's 512 reserved nodes
    const int64_t wn = (R + 512 + 63) / 64 * 64;
    std::vector<uint64_t> keys((size_t)wn);
    std::vector<uint32_t> queues((size_t)(4 * wn)), bitmaps((size_t)(3 * wn / 32));
    std::vector<float> dist((size_t)R);
    std::vector<int32_t> parent((size_t)R);
    std::vector<int64_t> stats(8, 0);
    orx_lattice lat{X, Y, L, (uint32_t)m["hmask"], nullptr, nullptr, nullptr, wcost.data(), nullptr, nullptr,
                    nullptr};
    orx_workspace ws{keys.data(), queues.data(), bitmaps.data(), wn};
    orx_route_cfg cfg{(float)m["delta"], (int32_t)m["block"], 0, (int32_t)m["flags"], stats.data()};
    if (orx_init() != ORX_OK) return 3;
    int rc = orx_sssp_box(&lat, (int32_t)m["x0"], (int32_t)m["y0"], w, h, (int64_t)m["src"], (int64_t)m["dst"],
                          dist.data(), parent.data(), &ws, &cfg, nullptr);
    if (rc < 0) { fprintf(stderr, "orx_sssp_box -> %d: %s\n", rc, orx_last_error()); return 4; }
    f = fopen((dir + "/dist.out").c_str(), "wb"); fwrite(dist.data(), 4, dist.size(), f); fclose(f);
    f = fopen((dir + "/parent.out").c_str(), "wb"); fwrite(parent.data(), 4, parent.size(), f); fclose(f);
    f = fopen((dir + "/stats.out").c_str(), "wb"); fwrite(stats.data(), 8, stats.size(), f); fclose(f);
    return 0;
}
