// Host-only checks of include/rtamd.hpp (no device needed):
//   1. GpuScene::buildBvhAndLights gives the nodes and lights of the
//      library's own scene generator (Bvh::createBvh's rule, seed 1, and the
//      light rule of RtScene.h:87-96), byte for byte;
//   2. a failing call throws rtamd::Error carrying the status and message;
//   3. with --no-device, rtamd::Context(0) throws (the library fails loudly
//      without a GPU instead of falling back to the CPU).
// Prints "hpp ok" and exits 0 when every check holds.
#include <cstdio>
#include <cstring>
#include <string>

#include "rtamd.hpp"

#define EXPECT(cond)                                                          \
    do {                                                                      \
        if (!(cond)) {                                                        \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            return 1;                                                         \
        }                                                                     \
    } while (0)

template <class T>
static bool same_bytes(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

int main(int argc, char** argv) {
    const bool no_device = argc > 1 && std::string(argv[1]) == "--no-device";

    rtw_scene w{};
    rtamd::check(rtw_scene_soup(777, 9, 4, &w), nullptr, "rtw_scene_soup");
    rtamd::GpuScene g = rtamd::GpuScene::adopt(w);
    EXPECT(g.triangles.size() == 777 && g.bvhNodes.size() == 2 * 777 - 1 && !g.lights.empty());
    const std::vector<rt_bvh_node> nodes = g.bvhNodes;
    const std::vector<rt_light> lights = g.lights;
    g.bvhNodes.clear();
    g.lights.clear();
    g.buildBvhAndLights(3);
    EXPECT(same_bytes(g.bvhNodes, nodes));
    EXPECT(same_bytes(g.lights, lights));

    const rt_scene_desc d = g.desc();
    EXPECT(d.num_triangles == 777 && d.num_nodes == 2 * 777 - 1 && d.triangles == g.triangles.data());

    bool threw = false;
    try {
        rtamd::check(rt_bvh_build_seeded(nullptr, 5, 1, 1, nullptr), nullptr, "rt_bvh_build_seeded");
    } catch (const rtamd::Error& e) {
        threw = e.status() == RT_ERR_INVALID && std::string(e.what()).find("rt_bvh_build_seeded") == 0;
    }
    EXPECT(threw);

    if (no_device) {
        threw = false;
        try {
            rtamd::Context ctx(0);
        } catch (const rtamd::Error& e) {
            threw = e.status() != RT_OK && std::string(e.what()).find("rt_create") == 0;
            std::printf("no device: %s\n", e.what());
        }
        EXPECT(threw);
    }
    std::printf("hpp ok\n");
    return 0;
}
