/*
 * rtamd.hpp -- C++ host layer over the C-ABI of include/rtamd.h, for the
 * reference's own host code (C++17).  Header only: RAII owners and
 * exceptions over the rt_* entry points, shaped after the reference's host
 * objects:
 *   rtamd::GpuScene      GpuModel::Scene (src/ray-tracing/RtScene.h:35-102):
 *                        triangles, spheres, materials, lights, bvhNodes
 *   rtamd::Scene         the uploaded scene (the storage buffers b3-b7 that
 *                        main.cpp:84-153 binds to ComputeMaterial)
 *   rtamd::ComputeModel  ComputeModel::computeCommand (src/scene/
 *                        ComputeModel.cpp:21-25) together with the
 *                        target->accumulation copy main.cpp:228-261 records
 *                        after it and the UBO updateScene fills
 *                        (main.cpp:166-183)
 * Every failing call throws rtamd::Error with the library's status and
 * message, where the reference throws std::runtime_error on a failed
 * VkResult.
 */
#ifndef RTAMD_HPP
#define RTAMD_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rtamd.h"
#include "rtamd_workloads.h"

namespace rtamd {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int rc, const rt_ctx* ctx, const char* what) {
    if (rc == RT_OK) return;
    std::string msg = std::string(what) + ": ";
    const char* detail = ctx ? rt_last_error(ctx) : nullptr;
    msg += detail && *detail ? detail : rt_status_string(rc);
    throw Error(rc, msg);
}

/* The host arrays of GpuModel::Scene, in the same record layouts. */
struct GpuScene {
    std::vector<rt_triangle> triangles;
    std::vector<rt_sphere> spheres;
    std::vector<rt_material> materials;
    std::vector<rt_light> lights;
    std::vector<rt_bvh_node> bvhNodes;

    rt_scene_desc desc() const {
        rt_scene_desc d{};
        d.triangles = triangles.data();
        d.num_triangles = (uint32_t)triangles.size();
        d.materials = materials.data();
        d.num_materials = (uint32_t)materials.size();
        d.nodes = bvhNodes.data();
        d.num_nodes = (uint32_t)bvhNodes.size();
        d.lights = lights.data();
        d.num_lights = (uint32_t)lights.size();
        d.spheres = spheres.data();
        d.num_spheres = (uint32_t)spheres.size();
        return d;
    }

    /* Takes ownership of a library-generated scene (rtw_scene_*). */
    static GpuScene adopt(rtw_scene& w) {
        GpuScene g;
        g.triangles.assign(w.triangles, w.triangles + w.num_triangles);
        g.materials.assign(w.materials, w.materials + w.num_materials);
        g.bvhNodes.assign(w.nodes, w.nodes + w.num_nodes);
        g.lights.assign(w.lights, w.lights + w.num_lights);
        g.spheres.assign(w.spheres, w.spheres + w.num_spheres);
        rtw_scene_free(&w);
        return g;
    }

    /* Bvh::createBvh over the triangles (reference builder, seed 1) and the
     * light rule of RtScene.h:87-96. */
    void buildBvhAndLights(uint32_t threads = 1) {
        bvhNodes.resize(triangles.empty() ? 0 : 2 * triangles.size() - 1);
        check(rt_bvh_build_seeded(triangles.data(), (uint32_t)triangles.size(), 1, threads, bvhNodes.data()),
              nullptr, "rt_bvh_build_seeded");
        uint32_t n = 0;
        check(rt_scene_lights(triangles.data(), (uint32_t)triangles.size(), materials.data(),
                              (uint32_t)materials.size(), nullptr, 0, &n),
              nullptr, "rt_scene_lights");
        lights.resize(n);
        check(rt_scene_lights(triangles.data(), (uint32_t)triangles.size(), materials.data(),
                              (uint32_t)materials.size(), lights.data(), n, &n),
              nullptr, "rt_scene_lights");
    }
};

class Context {
public:
    explicit Context(int device = 0) {
        rt_ctx* c = nullptr;
        check(rt_create(&c, device), nullptr, "rt_create");
        ctx_.reset(c);
    }
    rt_ctx* get() const { return ctx_.get(); }
    void setTuning(const rt_tuning& t) { check(rt_tuning_set(get(), &t), get(), "rt_tuning_set"); }

private:
    struct Del {
        void operator()(rt_ctx* c) const { rt_destroy(c); }
    };
    std::unique_ptr<rt_ctx, Del> ctx_;
};

class Scene {
public:
    Scene(Context& ctx, const GpuScene& g) : ctx_(&ctx) {
        const rt_scene_desc d = g.desc();
        rt_scene* s = nullptr;
        check(rt_scene_upload(ctx.get(), &d, &s), ctx.get(), "rt_scene_upload");
        scene_ = s;
        check(rt_scene_info_get(s, &info_), ctx.get(), "rt_scene_info_get");
    }
    ~Scene() {
        if (scene_) rt_scene_release(ctx_->get(), scene_);
    }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    const rt_scene* get() const { return scene_; }
    const rt_scene_info& info() const { return info_; }

    /* hit_bvh over device rays (ray-trace-compute.comp:263-311). */
    void closestHit(const rt_ray* rays_dev, uint32_t n, rt_hit* hits_dev, const rt_query_params* params = nullptr,
                    void* stream = nullptr) const {
        check(rt_closest_hit(ctx_->get(), scene_, rays_dev, n, hits_dev, params, stream), ctx_->get(),
              "rt_closest_hit");
    }

private:
    Context* ctx_;
    rt_scene* scene_ = nullptr;
    rt_scene_info info_{};
};

/* The compute pass of the reference's frame: main() over the image, then
 * the copy of the target into the accumulation image. */
class ComputeModel {
public:
    ComputeModel(Context& ctx, const Scene& scene, const rt_render_params& params, uint32_t width, uint32_t height)
        : ctx_(&ctx) {
        rt_frame_loop* f = nullptr;
        check(rt_frame_loop_create(ctx.get(), scene.get(), &params, width, height, &f), ctx.get(),
              "rt_frame_loop_create");
        loop_.reset(f);
    }
    /* One frame with the UBO main.cpp:166-183 fills: camera position and
     * time; currentSample restarts at 0 when `moved`, else counts up. */
    void computeCommand(const float cameraPos[3], float time, bool moved, void* stream = nullptr) {
        check(rt_frame_loop_step(loop_.get(), cameraPos, time, moved ? 1 : 0, stream), ctx_->get(),
              "rt_frame_loop_step");
    }
    /* The post-process pass (post-process-shader.frag:62-68) into the display image. */
    void postProcess(void* stream = nullptr) {
        check(rt_frame_loop_display(loop_.get(), stream), ctx_->get(), "rt_frame_loop_display");
    }
    /* 0 target, 1 accumulation, 2 display: w*h*4 bytes copied to the host. */
    std::vector<uint8_t> read(int which, uint32_t width, uint32_t height, void* stream = nullptr) const {
        std::vector<uint8_t> out((size_t)width * height * 4);
        check(rt_frame_loop_read(loop_.get(), which, out.data(), stream), ctx_->get(), "rt_frame_loop_read");
        return out;
    }
    uint32_t currentSample() const {
        uint32_t cs = 0;
        check(rt_frame_loop_images(loop_.get(), nullptr, nullptr, nullptr, &cs), ctx_->get(), "rt_frame_loop_images");
        return cs;
    }

private:
    struct Del {
        void operator()(rt_frame_loop* f) const { rt_frame_loop_destroy(f); }
    };
    Context* ctx_;
    std::unique_ptr<rt_frame_loop, Del> loop_;
};

inline void writeImage(const std::string& path, const std::vector<uint8_t>& rgba, uint32_t width, uint32_t height) {
    check(rt_image_write(path.c_str(), rgba.data(), width, height), nullptr, "rt_image_write");
}

}  // namespace rtamd

#endif /* RTAMD_HPP */
