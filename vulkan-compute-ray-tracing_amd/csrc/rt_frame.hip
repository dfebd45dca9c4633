// rt_frame.hip -- the reference's frame loop around the traced image and
// its display pass.
//
// Frame loop (src/main.cpp:166-183 updateScene, :228-261 command buffer):
// a camera move resets currentSample to 0; the UBO takes the camera
// position, the time, currentSample and the scene's counts; the compute
// pass renders into the target reading the accumulation image; the target
// is copied into the accumulation image; currentSample is incremented.
//
// Display (post-process-shader.frag:62-68): the target is sampled with a
// linear filter at texel centres (the quad covers the target one to one,
// so the sample is the texel, c/255), raised to 1/gamma with gamma = 2.2,
// alpha 1, and stored to an 8-bit UNORM image.  The per-channel map has
// 256 inputs, so it is a table: out = rne(pow(fl(c/255), fl(1/2.2)) * 255)
// with pow correctly rounded to binary32 (pow in double precision, then
// rounded).  The kernel reads 4 bytes and writes 4 bytes per pixel.
#include <cmath>
#include <cstring>

#include "rt_common.hpp"

namespace rtamd {
namespace {

struct GammaLut {
    uint8_t v[256];
};

__global__ __launch_bounds__(256) void k_gamma(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                               uint32_t count, GammaLut lut) {
    __shared__ uint8_t t[256];
    t[threadIdx.x] = lut.v[threadIdx.x];
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t c = src[i];
        dst[i] = (uint32_t)t[c & 0xffu] | ((uint32_t)t[(c >> 8) & 0xffu] << 8) | ((uint32_t)t[(c >> 16) & 0xffu] << 16) |
                 0xff000000u;
    }
}

// UNORM8 store of the shader's output: clamp to [0, 1], round half to even.
uint8_t unorm8(float x) {
    if (!(x > 0.0f)) return 0;
    if (x >= 1.0f) return 255;
    return (uint8_t)std::nearbyint(x * 255.0f);
}

GammaLut make_gamma_lut() {
    GammaLut l;
    const float gamma = 2.2f;
    const float e = 1.0f / gamma;
    for (int c = 0; c < 256; ++c) {
        const float x = (float)c / 255.0f;
        const float y = (float)std::pow((double)x, (double)e);
        l.v[c] = unorm8(y);
    }
    return l;
}

}  // namespace
}  // namespace rtamd

using namespace rtamd;

extern "C" int rt_postprocess_gamma(rt_ctx* ctx, const uint8_t* src_dev, uint8_t* dst_dev, uint32_t w, uint32_t h,
                                    void* stream) {
    if (!ctx) return RT_ERR_INVALID;
    const uint64_t count = (uint64_t)w * h;
    if (count == 0) return RT_OK;
    if (!src_dev || !dst_dev) return set_error(ctx, RT_ERR_INVALID, "rt_postprocess_gamma: null image");
    if (((uintptr_t)src_dev & 3u) || ((uintptr_t)dst_dev & 3u))
        return set_error(ctx, RT_ERR_INVALID, "rt_postprocess_gamma: images must be 4-byte aligned");
    // the pass's 32-bit grid-stride index stays below 2^32 for fewer than 2^31 pixels
    if (count > 0x7fffffffull) return set_error(ctx, RT_ERR_INVALID, "rt_postprocess_gamma: image too large");
    static const GammaLut lut = make_gamma_lut();
    RT_HIP(ctx, hipSetDevice(ctx->device));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((count + 255) / 256, (uint64_t)ctx->num_cus * 16u);
    hipLaunchKernelGGL(k_gamma, dim3(grid), dim3(256), 0, (hipStream_t)stream, (const uint32_t*)src_dev,
                       (uint32_t*)dst_dev, (uint32_t)count, lut);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

extern "C" int rt_gamma_table(uint8_t out[256]) {
    if (!out) return RT_ERR_INVALID;
    static const GammaLut lut = make_gamma_lut();
    std::memcpy(out, lut.v, 256);
    return RT_OK;
}

struct rt_frame_loop {
    rt_ctx* ctx = nullptr;
    const rt_scene* scene = nullptr;
    rt_render_params params{};
    uint32_t w = 0, h = 0;
    uint8_t* target = nullptr;
    uint8_t* accum = nullptr;
    uint8_t* display = nullptr;
    uint32_t current_sample = 0;
};

extern "C" int rt_frame_loop_create(rt_ctx* ctx, const rt_scene* scene, const rt_render_params* params, uint32_t w,
                                    uint32_t h, rt_frame_loop** out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    *out = nullptr;
    if (!scene_usable(ctx, scene)) return set_error(ctx, RT_ERR_INVALID, "rt_frame_loop_create: scene not on this device");
    if (!params || w == 0 || h == 0) return set_error(ctx, RT_ERR_INVALID, "rt_frame_loop_create: empty image or null params");
    if ((uint64_t)w * h > 0x7fffffffull) return set_error(ctx, RT_ERR_INVALID, "rt_frame_loop_create: image too large");
    RT_HIP(ctx, hipSetDevice(ctx->device));
    rt_frame_loop* f = new (std::nothrow) rt_frame_loop();
    if (!f) return RT_ERR_NOMEM;
    f->ctx = ctx;
    f->scene = scene;
    f->params = *params;
    f->w = w;
    f->h = h;
    const size_t bytes = (size_t)w * h * 4;
    hipError_t e = hipMalloc((void**)&f->target, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&f->accum, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&f->display, bytes);
    if (e == hipSuccess) e = hipMemset(f->target, 0, bytes);
    if (e == hipSuccess) e = hipMemset(f->accum, 0, bytes);
    if (e == hipSuccess) e = hipMemset(f->display, 0, bytes);
    if (e != hipSuccess) {
        if (f->target) (void)hipFree(f->target);
        if (f->accum) (void)hipFree(f->accum);
        if (f->display) (void)hipFree(f->display);
        delete f;
        return hip_error(ctx, e, "rt_frame_loop_create");
    }
    *out = f;
    return RT_OK;
}

extern "C" int rt_frame_loop_destroy(rt_frame_loop* f) {
    if (!f) return RT_OK;
    (void)hipSetDevice(f->ctx->device);
    (void)hipDeviceSynchronize();
    (void)hipFree(f->target);
    (void)hipFree(f->accum);
    (void)hipFree(f->display);
    delete f;
    return RT_OK;
}

extern "C" int rt_frame_loop_step(rt_frame_loop* f, const float cam_pos[3], float time, int moved, void* stream) {
    if (!f || !cam_pos) return RT_ERR_INVALID;
    if (moved) f->current_sample = 0;
    rt_ubo ubo{};
    std::memcpy(ubo.camPos, cam_pos, sizeof ubo.camPos);
    ubo.time = time;
    ubo.currentSample = f->current_sample;
    rt_scene_info info{};
    int rc = rt_scene_info_get(f->scene, &info);
    if (rc) return rc;
    ubo.numTriangles = info.num_triangles;
    ubo.numLights = info.num_lights;
    ubo.numSpheres = info.num_spheres;
    rc = rt_render(f->ctx, f->scene, &ubo, &f->params, f->target, f->accum, f->w, f->h, stream);
    if (rc) return rc;
    RT_HIP(f->ctx, hipMemcpyAsync(f->accum, f->target, (size_t)f->w * f->h * 4, hipMemcpyDeviceToDevice,
                                  (hipStream_t)stream));
    f->current_sample++;
    return RT_OK;
}

extern "C" int rt_frame_loop_display(rt_frame_loop* f, void* stream) {
    if (!f) return RT_ERR_INVALID;
    return rt_postprocess_gamma(f->ctx, f->target, f->display, f->w, f->h, stream);
}

extern "C" int rt_frame_loop_images(const rt_frame_loop* f, uint8_t** target_dev, uint8_t** accum_dev,
                                    uint8_t** display_dev, uint32_t* current_sample) {
    if (!f) return RT_ERR_INVALID;
    if (target_dev) *target_dev = f->target;
    if (accum_dev) *accum_dev = f->accum;
    if (display_dev) *display_dev = f->display;
    if (current_sample) *current_sample = f->current_sample;
    return RT_OK;
}

extern "C" int rt_frame_loop_read(const rt_frame_loop* f, int which, uint8_t* host, void* stream) {
    if (!f || !host || which < 0 || which > 2) return RT_ERR_INVALID;
    const uint8_t* src = which == 0 ? f->target : which == 1 ? f->accum : f->display;
    RT_HIP(f->ctx, hipSetDevice(f->ctx->device));
    RT_HIP(f->ctx, hipMemcpyAsync(host, src, (size_t)f->w * f->h * 4, hipMemcpyDeviceToHost, (hipStream_t)stream));
    RT_HIP(f->ctx, hipStreamSynchronize((hipStream_t)stream));
    return RT_OK;
}
