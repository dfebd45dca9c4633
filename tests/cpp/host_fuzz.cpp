// Fuzz of the library's host-side parsers and validators, built with
// AddressSanitizer and UBSan (CPU only, no device is touched):
//   rt_obj_load      on mutated OBJ text (odd tokens, huge and negative
//                    indices, truncation, binary bytes);
//   rt_scene_check   on random and corrupted trees (bad child and object
//                    indices, cycles, NaN and infinite bounds);
//   rt_scene_lights  on out-of-range material indices;
//   rt_bvh_build_seeded on degenerate triangles (NaN, infinite, coincident);
//   rt_image_write   on small images.
// A sanitizer report or a crash fails the run; every call may refuse its
// input, none may read or write out of bounds.
// usage: host_fuzz SEED ITERATIONS DIR
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "rtamd.h"
#include "rtamd_workloads.h"

static std::mt19937_64 rng;
static uint64_t rnd(uint64_t n) { return n ? rng() % n : 0; }

static const char* kTokens[] = {"v", "f", "vt", "vn", "#", "-", "/", "//", "1e309", "nan", "inf", "-inf",
                                "999999999999", "4294967296", "2147483648", "-2147483649", "0", "-0", "1", "-1",
                                "3", " ", "\n", "\t", "\r\n", "o", "g", "s", "usemtl", "mtllib", "0.5", "1e-45",
                                "v 1 2 3", "f 1 2 3", "f 1/1/1 2/2/2 3/3/3", "f 1 2 3 4 5", "f 0 0 0", "f -9 1 2",
                                "\\", "\0", "\xff", "vn 0 0 1", "vt 0.5 0.5"};

static std::string mutate(const std::string& base) {
    std::string s = base;
    int ops = 1 + (int)rnd(8);
    for (int k = 0; k < ops; ++k) {
        switch (rnd(6)) {
            case 0: if (!s.empty()) s[rnd(s.size())] = (char)rnd(256); break;
            case 1: {
                const char* t = kTokens[rnd(sizeof kTokens / sizeof *kTokens)];
                s.insert(rnd(s.size() + 1), t);
                break;
            }
            case 2: if (!s.empty()) s.erase(rnd(s.size()), rnd(16)); break;
            case 3: s.resize(rnd(s.size() + 1)); break;
            case 4: s += s.substr(rnd(s.size() + 1)); break;
            case 5: s.insert(rnd(s.size() + 1), std::string(rnd(300), (char)('0' + rnd(10)))); break;
        }
    }
    return s;
}

static float rf() {
    switch (rnd(20)) {
        case 0: return NAN;
        case 1: return INFINITY;
        case 2: return -INFINITY;
        case 3: return 1e30f;
        case 4: return 0.0f;
        default: return (float)((double)rnd(2000001) / 1000000.0 - 1.0);
    }
}

int main(int argc, char** argv) {
    const uint64_t seed = argc > 1 ? strtoull(argv[1], nullptr, 10) : 1;
    const int iters = argc > 2 ? atoi(argv[2]) : 2000;
    const std::string dir = argc > 3 ? argv[3] : ".";
    rng.seed(seed);
    const std::string base =
        "# cube corner\nv 0 0 0\nv 1 0 0\nv 0 1 0\nv 1 1 0\nvt 0 0\nvn 0 0 1\nf 1 2 3\nf 2/1/1 4/1/1 3/1/1\n"
        "f -1 -2 -3\nf 1//1 2//1 4//1\n";
    const std::string path = dir + "/fuzz.obj";
    long ok = 0, bad = 0;
    for (int it = 0; it < iters; ++it) {
        // 1. OBJ reader on mutated text
        {
            std::string s = mutate(base);
            FILE* f = std::fopen(path.c_str(), "wb");
            std::fwrite(s.data(), 1, s.size(), f);
            std::fclose(f);
            rt_triangle* tris = nullptr;
            uint32_t n = 0;
            int rc = rt_obj_load(path.c_str(), (uint32_t)rnd(4), &tris, &n);
            (rc == RT_OK ? ok : bad)++;
            if (rc == RT_OK && n) {
                volatile float sink = 0;
                for (uint32_t i = 0; i < n; ++i) sink = sink + tris[i].v0[0] + tris[i].v2[2];
            }
            rt_free(tris);
        }
        // 2. scene validation on random arrays
        {
            const uint32_t nt = 1 + (uint32_t)rnd(40), nn = (uint32_t)rnd(2 * nt + 4), nm = (uint32_t)rnd(6),
                           nl = (uint32_t)rnd(6), ns = (uint32_t)rnd(20);
            std::vector<rt_triangle> t(nt);
            for (auto& x : t) {
                for (int k = 0; k < 3; ++k) x.v0[k] = rf(), x.v1[k] = rf(), x.v2[k] = rf();
                x.materialIndex = (uint32_t)rnd(nm + 2);
            }
            std::vector<rt_material> m(nm);
            for (auto& x : m) { x.type = (uint32_t)rnd(5); for (int k = 0; k < 3; ++k) x.albedo[k] = rf(); }
            std::vector<rt_light> l(nl);
            for (auto& x : l) { x.triangleIndex = (uint32_t)rnd(nt + 2); x.area = rf(); }
            std::vector<rt_sphere> sp(ns);
            for (auto& x : sp) { for (int k = 0; k < 4; ++k) x.s[k] = rf(); x.materialIndex = (uint32_t)rnd(nm + 2); }
            std::vector<rt_bvh_node> nodes(nn);
            if (nn && rnd(2)) {
                // a valid tree from the host builder, then a few corruptions
                std::vector<rt_bvh_node> built(2 * (size_t)nt - 1);
                if (rt_bvh_build_seeded(t.data(), nt, 1, 1, built.data()) == RT_OK) nodes = built;
                for (int k = (int)rnd(3); k > 0 && !nodes.empty(); --k) {
                    rt_bvh_node& x = nodes[rnd(nodes.size())];
                    switch (rnd(4)) {
                        case 0: x.leftNodeIndex = (int32_t)rnd(nodes.size() + 3) - 2; break;
                        case 1: x.rightNodeIndex = (int32_t)rnd(nodes.size() + 3) - 2; break;
                        case 2: x.objectIndex = (int32_t)rnd(nt + 3) - 2; break;
                        case 3: x.min[rnd(3)] = rf(); break;
                    }
                }
            } else {
                for (auto& x : nodes) {
                    for (int k = 0; k < 3; ++k) x.min[k] = rf(), x.max[k] = rf();
                    x.leftNodeIndex = (int32_t)rnd(nn + 3) - 2;
                    x.rightNodeIndex = (int32_t)rnd(nn + 3) - 2;
                    x.objectIndex = (int32_t)rnd(nt + 3) - 2;
                }
            }
            rt_scene_desc d{};
            d.triangles = t.data(); d.num_triangles = nt;
            d.materials = m.data(); d.num_materials = nm;
            d.nodes = nodes.data(); d.num_nodes = (uint32_t)nodes.size();
            d.lights = l.data(); d.num_lights = nl;
            d.spheres = sp.data(); d.num_spheres = ns;
            rt_scene_info info{};
            char err[256];
            int rc = rt_scene_check(&d, &info, err, sizeof err);
            (rc == RT_OK ? ok : bad)++;
            // the light rule over the same arrays
            uint32_t cnt = 0;
            std::vector<rt_light> lo(nt);
            rc = rt_scene_lights(t.data(), nt, m.data(), nm, lo.data(), nt, &cnt);
            (rc == RT_OK ? ok : bad)++;
        }
        // 3. host builder on degenerate triangles
        {
            const uint32_t nt = 1 + (uint32_t)rnd(300);
            std::vector<rt_triangle> t(nt);
            const bool same = rnd(4) == 0;
            for (auto& x : t) {
                for (int k = 0; k < 3; ++k) {
                    x.v0[k] = same ? 0.5f : rf();
                    x.v1[k] = same ? 0.5f : rf();
                    x.v2[k] = same ? 0.5f : rf();
                }
                x.materialIndex = 0;
            }
            std::vector<rt_bvh_node> nodes(2 * (size_t)nt - 1);
            int rc = rt_bvh_build_seeded(t.data(), nt, (uint32_t)rnd(5), 1 + (uint32_t)rnd(4), nodes.data());
            (rc == RT_OK ? ok : bad)++;
        }
        // 4. image writer on small images
        if (it % 50 == 0) {
            const uint32_t w = (uint32_t)rnd(70), h = (uint32_t)rnd(70);
            std::vector<uint8_t> img((size_t)w * h * 4 + 1);
            for (auto& b : img) b = (uint8_t)rnd(256);
            const std::string out = dir + (rnd(2) ? "/fuzz.png" : "/fuzz.ppm");
            int rc = rt_image_write(out.c_str(), img.data(), w, h);
            (rc == RT_OK ? ok : bad)++;
        }
    }
    std::printf("fuzz seed %llu: %d iterations, %ld accepted, %ld refused\n", (unsigned long long)seed, iters, ok, bad);
    return 0;
}
