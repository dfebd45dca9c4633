// rtamd_frame_loop: the reference application's frame loop as a C++ host
// program over rtamd.hpp, without a window.  It follows src/main.cpp:
//   scene      RtScene.h:44-101 (the doge scene from its OBJ directory) or a
//              generated triangle soup, BVH from Bvh::createBvh's rule;
//   camera     Camera camera(glm::vec3(1.8f, 8.6f, 1.1f)) (main.cpp:37);
//   per frame  updateScene (main.cpp:166-183): the UBO with camera, time and
//              currentSample; the compute dispatch and the target->accumulation
//              copy (main.cpp:228-261); the post-process pass on the last frame.
// The display image is written as PNG or PPM (by extension); --raw also
// writes the target's RGBA8 bytes, row-major.
//
// usage: rtamd_frame_loop W H FRAMES OUT.png|OUT.ppm [--doge DIR] [--soup N SEED]
//                         [--cam X Y Z] [--vfov DEG] [--bounces B] [--raw PATH]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>

#include "rtamd.hpp"

namespace {

struct Options {
    uint32_t w = 0, h = 0, frames = 0;
    std::string out, raw, doge;
    uint32_t soup_n = 1024, soup_seed = 1;
    float cam[3] = {1.8f, 8.6f, 1.1f};
    float vfov = 30.0f;     // ray-trace-compute.comp:358
    uint32_t bounces = 2;   // NUM_BOUNCES, ray-trace-compute.comp:313
};

[[noreturn]] void usage(const char* msg) {
    std::fprintf(stderr,
                 "rtamd_frame_loop: %s\n"
                 "usage: rtamd_frame_loop W H FRAMES OUT.png|OUT.ppm [--doge DIR] [--soup N SEED]\n"
                 "                        [--cam X Y Z] [--vfov DEG] [--bounces B] [--raw PATH]\n",
                 msg);
    std::exit(2);
}

uint32_t to_u32(const char* s, const char* what) {
    char* end = nullptr;
    unsigned long v = std::strtoul(s, &end, 10);
    if (!*s || *end || v > 0xffffffffUL) usage((std::string("bad ") + what + ": " + s).c_str());
    return (uint32_t)v;
}

float to_f32(const char* s, const char* what) {
    char* end = nullptr;
    float v = std::strtof(s, &end);
    if (!*s || *end) usage((std::string("bad ") + what + ": " + s).c_str());
    return v;
}

Options parse(int argc, char** argv) {
    if (argc < 5) usage("missing arguments");
    Options o;
    o.w = to_u32(argv[1], "width");
    o.h = to_u32(argv[2], "height");
    o.frames = to_u32(argv[3], "frames");
    o.out = argv[4];
    for (int i = 5; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](int k) {
            if (i + k >= argc) usage(("missing value for " + a).c_str());
        };
        if (a == "--doge") {
            need(1);
            o.doge = argv[++i];
        } else if (a == "--soup") {
            need(2);
            o.soup_n = to_u32(argv[i + 1], "soup size");
            o.soup_seed = to_u32(argv[i + 2], "soup seed");
            i += 2;
        } else if (a == "--cam") {
            need(3);
            for (int k = 0; k < 3; ++k) o.cam[k] = to_f32(argv[i + 1 + k], "camera");
            i += 3;
        } else if (a == "--vfov") {
            need(1);
            o.vfov = to_f32(argv[++i], "vfov");
        } else if (a == "--bounces") {
            need(1);
            o.bounces = to_u32(argv[++i], "bounces");
        } else if (a == "--raw") {
            need(1);
            o.raw = argv[++i];
        } else {
            usage(("unknown option " + a).c_str());
        }
    }
    if (o.frames == 0) usage("FRAMES must be at least 1");
    return o;
}

rtamd::GpuScene make_scene(const Options& o) {
    rtw_scene w{};
    if (!o.doge.empty()) {
        const int rc = rtw_scene_doge(o.doge.c_str(), 8, &w);
        if (rc != RT_OK) {
            const char* detail = rt_obj_last_error();
            throw rtamd::Error(rc, "rtw_scene_doge(" + o.doge + "): " +
                                       (detail && *detail ? detail : rt_status_string(rc)));
        }
    } else {
        rtamd::check(rtw_scene_soup(o.soup_n, o.soup_seed, 8, &w), nullptr, "rtw_scene_soup");
    }
    return rtamd::GpuScene::adopt(w);
}

}  // namespace

int main(int argc, char** argv) {
    const Options o = parse(argc, argv);
    try {
        const rtamd::GpuScene host = make_scene(o);
        rtamd::Context ctx(0);
        const rtamd::Scene scene(ctx, host);
        rt_render_params prm{};
        prm.vfov_deg = o.vfov;
        prm.bounces = o.bounces;
        rtamd::ComputeModel model(ctx, scene, prm, o.w, o.h);
        for (uint32_t k = 0; k < o.frames; ++k) model.computeCommand(o.cam, 0.0f, false);
        model.postProcess();
        const std::vector<uint8_t> display = model.read(2, o.w, o.h);
        rtamd::writeImage(o.out, display, o.w, o.h);
        if (!o.raw.empty()) {
            const std::vector<uint8_t> target = model.read(0, o.w, o.h);
            std::ofstream f(o.raw, std::ios::binary);
            f.write(reinterpret_cast<const char*>(target.data()), (std::streamsize)target.size());
            if (!f) throw rtamd::Error(RT_ERR_INVALID, "cannot write " + o.raw);
        }
        std::printf("rtamd_frame_loop: %u triangles, %u nodes, %u lights, %u spheres; %u frames %ux%u; "
                    "currentSample %u; wrote %s\n",
                    scene.info().num_triangles, scene.info().num_nodes, scene.info().num_lights,
                    scene.info().num_spheres, o.frames, o.w, o.h, model.currentSample(), o.out.c_str());
    } catch (const rtamd::Error& e) {
        std::fprintf(stderr, "rtamd_frame_loop: %s (status %d)\n", e.what(), e.status());
        return 1;
    }
    return 0;
}
