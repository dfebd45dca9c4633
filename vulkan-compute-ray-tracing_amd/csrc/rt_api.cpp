// rt_api.cpp -- context, errors and the query/render entry points of the
// C-ABI declared in include/rtamd.h.
#include <cstdio>
#include <cstring>
#include <new>

#include "rt_common.hpp"

namespace rtamd {

int set_error(rt_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_error(rt_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(ctx, e == hipErrorOutOfMemory ? RT_ERR_NOMEM : RT_ERR_HIP, m);
}

int ensure_workspace(rt_ctx* ctx, int slot, size_t bytes) {
    if (ctx->ws_bytes[slot] >= bytes) return RT_OK;
    if (ctx->ws[slot]) {
        RT_HIP(ctx, hipDeviceSynchronize());
        RT_HIP(ctx, hipFree(ctx->ws[slot]));
        ctx->ws[slot] = nullptr;
        ctx->ws_bytes[slot] = 0;
    }
    size_t grown = bytes + bytes / 8;
    RT_HIP(ctx, hipMalloc(&ctx->ws[slot], grown));
    ctx->ws_bytes[slot] = grown;
    return RT_OK;
}

int render_impl(rt_ctx* ctx, const rt_scene* sc, const rt_ubo& ubo, const rt_render_params& prm, uint32_t frames,
                uint32_t* target, const uint32_t* accum_in, uint32_t* accum_out, uint32_t w, uint32_t h,
                hipStream_t s);

}  // namespace rtamd

using namespace rtamd;

extern "C" const char* rt_status_string(int status) {
    switch (status) {
    case RT_OK: return "ok";
    case RT_ERR_INVALID: return "invalid argument";
    case RT_ERR_HIP: return "HIP runtime error";
    case RT_ERR_NOMEM: return "out of memory";
    case RT_ERR_UNSUPPORTED: return "unsupported";
    case RT_ERR_BUSY: return "busy";
    case RT_ERR_CANCELLED: return "cancelled";
    case RT_ERR_NOT_FOUND: return "not found";
    }
    return "unknown status";
}

extern "C" int rt_create(rt_ctx** out, int device) {
    if (!out) return RT_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return RT_ERR_HIP;
    if (device < 0 || device >= count) return RT_ERR_INVALID;
    rt_ctx* c = new (std::nothrow) rt_ctx();
    if (!c) return RT_ERR_NOMEM;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return RT_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return RT_ERR_HIP; }
    c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    // the allocation-free query's scratch: its device address, once per device
    (void)noalloc_setup(device);
    if (hipMalloc((void**)&c->d_counters, kRenderCounters * sizeof(uint32_t)) != hipSuccess) { delete c; return RT_ERR_NOMEM; }
    if (hipMemset(c->d_counters, 0, kRenderCounters * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc((void**)&c->d_cursors, kCursorSlots * sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(c->d_counters);
        delete c;
        return RT_ERR_HIP;
    }
    *out = c;
    return RT_OK;
}

extern "C" int rt_destroy(rt_ctx* ctx) {
    if (!ctx) return RT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (int i = 0; i < rt_ctx::kWorkspaces; ++i)
        if (ctx->ws[i]) (void)hipFree(ctx->ws[i]);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_cursors) (void)hipFree(ctx->d_cursors);
    if (ctx->d_render_sums) (void)hipFree(ctx->d_render_sums);
    for (hipEvent_t e : ctx->render_ev)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->query_ev)
        if (e) (void)hipEventDestroy(e);
    delete ctx;
    return RT_OK;
}

extern "C" int rt_tuning_set(rt_ctx* ctx, const rt_tuning* t) {
    if (!ctx || !t) return RT_ERR_INVALID;
    if (t->node_layout > RT_LAYOUT_W4) return set_error(ctx, RT_ERR_INVALID, "rt_tuning_set: unknown node layout");
    if (t->refill > 64) return set_error(ctx, RT_ERR_INVALID, "rt_tuning_set: refill must be 0..64");
    if (t->tree > RT_TREE_DEVICE) return set_error(ctx, RT_ERR_INVALID, "rt_tuning_set: unknown tree");
    if (t->ray_order > RT_ORDER_SORTED) return set_error(ctx, RT_ERR_INVALID, "rt_tuning_set: unknown ray order");
    if (t->steal > RT_STEAL_NEVER)
        return set_error(ctx, RT_ERR_INVALID, "rt_tuning_set: steal must be 0..64 or RT_STEAL_NEVER");
    if (t->stack_lds != 0 && t->stack_lds != 8 && t->stack_lds != 16 && t->stack_lds != 32)
        return set_error(ctx, RT_ERR_INVALID, "rt_tuning_set: stack_lds must be 0, 8, 16 or 32");
    ctx->tuning = *t;
    ctx->tuning_gen++;          // batches captured under the old tuning re-capture
    return RT_OK;
}

extern "C" int rt_tuning_get(const rt_ctx* ctx, rt_tuning* t) {
    if (!ctx || !t) return RT_ERR_INVALID;
    *t = ctx->tuning;
    return RT_OK;
}

extern "C" const char* rt_last_error(const rt_ctx* ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

namespace {

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

int query_common(rt_ctx* ctx, const rt_scene* scene, const rt_ray* rays, uint32_t n, rt_hit* hits,
                 const rt_query_params* params, uint32_t* nv, uint32_t* tt, void* stream) {
    if (!ctx) return RT_ERR_INVALID;
    if (!scene_usable(ctx, scene)) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit: scene not on this context's device");
    if (n == 0) return RT_OK;
    // the kernels' 32-bit grid-stride indices stay below 2^32 for n < 2^31
    if (n > 0x7fffffffu) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit: more than 2^31 - 1 rays in one call");
    if (!rays || !hits) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit: null ray or hit buffer");
    if (!aligned16(rays) || !aligned16(hits)) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit: buffers must be 16-byte aligned");
    rt_query_params p{};
    if (params) p = *params;
    if (p.traversal > RT_TRAVERSE_REFERENCE) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit: unknown traversal");
    if (p.spheres > RT_SPHERES_SCENE) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit: unknown sphere rule");
    RT_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    TraceArgs a{scene, rays, n, nullptr, hits, p.spheres, nv, tt};
    bool packed = p.traversal == RT_TRAVERSE_PACKED ||
                  (p.traversal == RT_TRAVERSE_AUTO && scene->fast && p.max_stack == 0);
    if (nv || tt) packed = true;
    if (packed && p.max_stack)
        return set_error(ctx, RT_ERR_UNSUPPORTED, "rt_closest_hit: max_stack applies to reference traversal");
    // stage timing: ev[0] start, ev[1] order ready, ev[2] traversal done, ev[3] spheres done
    hipEvent_t* ev = ctx->time_render ? ctx->query_ev : nullptr;
    ctx->last_query_timed = false;
    ctx->last_query_sorted = false;
    if (ev) {
        for (int i = 0; i < 4; ++i)
            if (!ev[i]) RT_HIP(ctx, hipEventCreate(&ev[i]));
        RT_HIP(ctx, hipEventRecord(ev[0], s));
    }
    int rc;
    if (packed) {
        rc = ray_order(ctx, a, s, &a.order);
        if (rc) return rc;
        ctx->last_query_sorted = a.order != nullptr;
        if (ev) RT_HIP(ctx, hipEventRecord(ev[1], s));
        rc = launch_trace_packed(ctx, a, s);
    } else {
        if (ev) RT_HIP(ctx, hipEventRecord(ev[1], s));
        rc = launch_trace_reference(ctx, a, p.max_stack, s);
    }
    if (rc) return rc;
    if (ev) RT_HIP(ctx, hipEventRecord(ev[2], s));
    rc = launch_sphere_pass(ctx, a, s);
    if (rc) return rc;
    if (ev) {
        RT_HIP(ctx, hipEventRecord(ev[3], s));
        ctx->last_query_timed = true;
    }
    return RT_OK;
}

}  // namespace

extern "C" int rt_closest_hit(rt_ctx* ctx, const rt_scene* scene, const rt_ray* rays_dev, uint32_t n,
                              rt_hit* hits_dev, const rt_query_params* params, void* stream) {
    return query_common(ctx, scene, rays_dev, n, hits_dev, params, nullptr, nullptr, stream);
}

extern "C" int rt_closest_hit_stats(rt_ctx* ctx, const rt_scene* scene, const rt_ray* rays_dev, uint32_t n,
                                    rt_hit* hits_dev, uint32_t* node_visits_dev, uint32_t* tri_tests_dev,
                                    void* stream) {
    if (ctx && (!node_visits_dev || !tri_tests_dev))
        return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit_stats: null counter buffer");
    return query_common(ctx, scene, rays_dev, n, hits_dev, nullptr, node_visits_dev, tri_tests_dev, stream);
}

namespace {

int render_common(rt_ctx* ctx, const rt_scene* scene, const rt_ubo* ubo, const rt_render_params* params,
                  uint32_t frames, uint8_t* target, const uint8_t* accum_in, uint8_t* accum_out, uint32_t w,
                  uint32_t h, void* stream) {
    if (!ctx) return RT_ERR_INVALID;
    if (!scene_usable(ctx, scene)) return set_error(ctx, RT_ERR_INVALID, "render: scene not on this context's device");
    if (!ubo || !params) return set_error(ctx, RT_ERR_INVALID, "render: null ubo or params");
    if (w == 0 || h == 0 || frames == 0) return RT_OK;
    if (!target || !accum_in) return set_error(ctx, RT_ERR_INVALID, "render: null image");
    if (((uintptr_t)target & 3u) || ((uintptr_t)accum_in & 3u))
        return set_error(ctx, RT_ERR_INVALID, "render: images must be 4-byte aligned");
    if (params->traversal > RT_TRAVERSE_REFERENCE) return set_error(ctx, RT_ERR_INVALID, "render: unknown traversal");
    if (params->spheres > RT_SPHERES_SCENE) return set_error(ctx, RT_ERR_INVALID, "render: unknown sphere rule");
    if (!(params->vfov_deg > 0.0f && params->vfov_deg < 180.0f))
        return set_error(ctx, RT_ERR_INVALID, "render: vfov must lie in (0, 180) degrees");
    RT_HIP(ctx, hipSetDevice(ctx->device));
    ctx->last_timed = false;
    ctx->last_counted = false;
    int rc = render_impl(ctx, scene, *ubo, *params, frames, (uint32_t*)target, (const uint32_t*)accum_in,
                         (uint32_t*)accum_out, w, h, (hipStream_t)stream);
    if (rc) return rc;
    ctx->last_primary = (uint64_t)w * h * frames;
    ctx->last_secondary = UINT64_MAX;   // read lazily from the counters
    ctx->last_bounces = params->bounces;
    ctx->last_timed = ctx->time_render && params->bounces > 0;
    ctx->last_counted = ctx->count_render && ctx->d_render_sums;
    ctx->last_render_stream = (hipStream_t)stream;
    return RT_OK;
}

}  // namespace

extern "C" int rt_render(rt_ctx* ctx, const rt_scene* scene, const rt_ubo* ubo, const rt_render_params* params,
                         uint8_t* target_dev, const uint8_t* accum_dev, uint32_t w, uint32_t h, void* stream) {
    return render_common(ctx, scene, ubo, params, 1, target_dev, accum_dev, nullptr, w, h, stream);
}

extern "C" int rt_render_frames(rt_ctx* ctx, const rt_scene* scene, const rt_ubo* ubo, const rt_render_params* params,
                                uint32_t frames, uint8_t* target_dev, uint8_t* accum_dev, uint32_t w, uint32_t h,
                                void* stream) {
    return render_common(ctx, scene, ubo, params, frames, target_dev, accum_dev, accum_dev, w, h, stream);
}

extern "C" int rt_render_set_timing(rt_ctx* ctx, int enable) {
    if (!ctx) return RT_ERR_INVALID;
    ctx->time_render = enable != 0;
    return RT_OK;
}

extern "C" int rt_render_set_counting(rt_ctx* ctx, int enable) {
    if (!ctx) return RT_ERR_INVALID;
    if (enable && !ctx->d_render_sums) {
        RT_HIP(ctx, hipSetDevice(ctx->device));
        RT_HIP(ctx, hipMalloc((void**)&ctx->d_render_sums, 2 * sizeof(unsigned long long)));
    }
    ctx->count_render = enable != 0;
    return RT_OK;
}

extern "C" int rt_render_last_counts(rt_ctx* ctx, uint64_t* node_records, uint64_t* tri_records) {
    if (!ctx || !node_records || !tri_records) return RT_ERR_INVALID;
    if (!ctx->last_counted) return set_error(ctx, RT_ERR_INVALID, "rt_render_last_counts: the last render was not counted");
    unsigned long long h[2] = {0, 0};
    RT_HIP(ctx, hipSetDevice(ctx->device));
    // the totals are final once the render's stream has run it
    RT_HIP(ctx, hipMemcpyAsync(h, ctx->d_render_sums, sizeof h, hipMemcpyDeviceToHost, ctx->last_render_stream));
    RT_HIP(ctx, hipStreamSynchronize(ctx->last_render_stream));
    *node_records = h[0];
    *tri_records = h[1];
    return RT_OK;
}

extern "C" int rt_render_last_stats(rt_ctx* ctx, rt_render_stats* out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    std::memset(out, 0, sizeof *out);
    int rc = rt_render_last_rays(ctx, &out->primary_rays, &out->secondary_rays);
    if (rc) return rc;
    out->bounces = ctx->last_bounces;
    if (!ctx->last_timed) return RT_OK;
    hipEvent_t* ev = ctx->render_ev;
    const uint32_t B = ctx->last_bounces;
    RT_HIP(ctx, hipEventSynchronize(ev[1 + 2 * B]));
    auto ms = [&](hipEvent_t a, hipEvent_t b) -> float {
        float t = 0.0f;
        return hipEventElapsedTime(&t, a, b) == hipSuccess ? t : 0.0f;
    };
    out->total_ms = ms(ev[0], ev[1 + 2 * B]);
    out->trace_primary_ms = ms(ev[1], ev[2]);
    for (uint32_t b = 1; b < B; ++b) out->trace_secondary_ms += ms(ev[1 + 2 * b], ev[2 + 2 * b]);
    out->shade_ms = out->total_ms - out->trace_primary_ms - out->trace_secondary_ms;
    out->timed = 1;
    return RT_OK;
}

extern "C" int rt_query_last_stats(rt_ctx* ctx, rt_query_stats* out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    std::memset(out, 0, sizeof *out);
    out->sorted = ctx->last_query_sorted ? 1u : 0u;
    if (!ctx->last_query_timed) return RT_OK;
    hipEvent_t* ev = ctx->query_ev;
    RT_HIP(ctx, hipEventSynchronize(ev[3]));
    auto ms = [&](hipEvent_t a, hipEvent_t b) -> float {
        float t = 0.0f;
        return hipEventElapsedTime(&t, a, b) == hipSuccess ? t : 0.0f;
    };
    out->order_ms = ms(ev[0], ev[1]);
    out->trace_ms = ms(ev[1], ev[2]);
    out->spheres_ms = ms(ev[2], ev[3]);
    out->total_ms = ms(ev[0], ev[3]);
    out->timed = 1;
    return RT_OK;
}

extern "C" int rt_render_last_rays(const rt_ctx* cctx, uint64_t* primary, uint64_t* secondary) {
    rt_ctx* ctx = const_cast<rt_ctx*>(cctx);
    if (!ctx || !primary || !secondary) return RT_ERR_INVALID;
    if (ctx->last_secondary == UINT64_MAX) {
        uint32_t c[64];
        RT_HIP(ctx, hipDeviceSynchronize());
        RT_HIP(ctx, hipMemcpy(c, ctx->d_counters, sizeof c, hipMemcpyDeviceToHost));
        uint64_t s = 0;
        for (int i = 1; i < 64; ++i) s += c[i];
        ctx->last_secondary = s;
    }
    *primary = ctx->last_primary;
    *secondary = ctx->last_secondary;
    return RT_OK;
}
