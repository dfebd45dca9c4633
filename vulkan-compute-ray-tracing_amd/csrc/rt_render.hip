// rt_render.hip -- the render path of ray-trace-compute.comp:main as
// wavefront stages on gfx950.
//
// The shader runs one invocation per pixel that loops over bounces
// (ray_color, :314-338).  Here every (pixel, frame) path keeps its state in
// HBM and the bounce loop becomes a sequence of launches:
//
//   k_generate           camera ray per path (main, :354-372; ray_color :319)
//   repeat `bounces` times:
//     trace              closest hit of every live ray (rt_trace.hip)
//     k_shade            hit record + scatter (:118-155, :205-220), colour
//                        product, and compaction of the surviving rays into
//                        the next queue with one atomic per wave
//   k_store              progressive 8-bit average per pixel (:375-379)
//
// Each path's arithmetic and RNG stream depend only on (pixel, sample), so
// the queue order chosen by compaction never changes a result.  The counters
// and ray cursors a stage needs at zero are cleared by the stage before it
// on the same stream (k_generate for the first bounce, k_shade for the next),
// so a frame launches no fill kernels.
#include <cmath>

#include "rt_common.hpp"
#include "rt_math.hpp"

#pragma clang fp contract(off)

namespace rtamd {

namespace {

struct Camera {
    f3 origin, llc, horizontal, vertical;
};

struct HitRec {
    f3 p, normal;
    uint32_t mat;
    int back;
};

// Onb (definitions.glsl:42-49) and onbLocal (:51-53).
__device__ __forceinline__ void make_onb(f3 n, f3& u, f3& v, f3& w) {
    w = normalize(n);
    f3 a = (fabsf(w.x) > 0.9f) ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    v = normalize(cross(w, a));
    u = cross(w, v);
}

// random_cosine_direction (random.glsl:42-52)
__device__ __forceinline__ f3 random_cosine_direction(uint32_t& rng) {
    float r1 = rng_next(rng);
    float r2 = rng_next(rng);
    float z = sqrtf(1.0f - r2);
    float phi = 2.0f * kPi * r1;
    float x = cos_cr(phi) * sqrtf(r2);
    float y = sin_cr(phi) * sqrtf(r2);
    return f3{x, y, z};
}

// sampleLambertian (:91-99); the pdf never reaches the colour.
__device__ __forceinline__ f3 sample_lambertian(f3 normal, uint32_t& rng) {
    f3 u, v, w;
    make_onb(normal, u, v, w);
    f3 rc = random_cosine_direction(rng);
    return normalize((rc.x * u + rc.y * v) + rc.z * w);
}

// sampleGlass (:106-116)
__device__ __forceinline__ f3 sample_glass(f3 I, const HitRec& rec) {
    float ir = 1.5f;
    float ratio = (float)(1 - rec.back) * 1.0f / ir + (float)rec.back * ir;
    f3 i = normalize(I);
    float cos_theta = gmin(dot(-i, rec.normal), 1.0f);
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    float t = floorf(gclamp(ratio * sin_theta, 0.0f, 1.0f));
    return t * reflect(i, rec.normal) + (1.0f - t) * refract(i, rec.normal, ratio);
}

// randomOnATriangle (:67-74) + sampleLight (:76-89).  An index past the end
// of lights[] (a draw of exactly 1.0, or no lights) reads a zero record.
__device__ __forceinline__ f3 sample_light(const rt_triangle* __restrict__ tris, const rt_light* __restrict__ lights,
                                           uint32_t nlights, f3 p, float& light_cosine, uint32_t& rng) {
    int li = (int)floorf((float)(int)nlights * rng_next(rng));
    uint32_t ti = (li >= 0 && (uint32_t)li < nlights) ? lights[li].triangleIndex : 0u;
    float s = rng_next(rng);
    float t = rng_next(rng);
    const rt_triangle& tr = tris[ti];
    f3 v0 = ld3(tr.v0), v1 = ld3(tr.v1), v2 = ld3(tr.v2);
    f3 v01 = -v0 + v1;
    f3 v02 = -v0 + v2;
    f3 on_light = (v0 + s * v01) + t * v02;
    f3 to_light = normalize(on_light - p);
    light_cosine = fabsf(to_light.y);
    return to_light;
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Appends (ray, path id) for the lanes with `keep` set: one atomic per wave,
// slots assigned by a ballot prefix count.
__device__ __forceinline__ void wave_append(bool keep, uint32_t* counter, rt_ray* out_rays, uint32_t* out_pid,
                                            const f3& o, const f3& d, uint32_t pid) {
    uint64_t mask = __ballot(keep);
    if (mask == 0) return;
    uint32_t total = (uint32_t)__popcll(mask);
    uint32_t leader = (uint32_t)__builtin_ctzll(mask);
    uint32_t lane = lane_id();
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, total);
    base = __shfl(base, (int)leader);
    if (keep) {
        uint32_t below = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        uint32_t slot = base + below;
        float4* rp = reinterpret_cast<float4*>(out_rays) + 2 * (size_t)slot;
        rp[0] = make_float4(o.x, o.y, o.z, 0.001f);
        rp[1] = make_float4(d.x, d.y, d.z, 10000.0f);
        out_pid[slot] = pid;
    }
}

__global__ __launch_bounds__(256) void k_generate(uint32_t w, uint32_t h, uint32_t frames, uint32_t sample0, Camera cam,
                                                  rt_ray* __restrict__ rays, uint32_t* __restrict__ pid,
                                                  float4* __restrict__ color, uint32_t* __restrict__ rng,
                                                  uint32_t* __restrict__ counters, uint32_t* __restrict__ cursor,
                                                  unsigned long long* __restrict__ sums) {
    // the render's bounce counters, record totals and first ray cursors
    // start at zero (written here instead of by fill launches)
    if (blockIdx.x == 0) {
        const uint32_t t = threadIdx.x;
        if (t < kRenderCounters) counters[t] = 0u;
        if (cursor && t < kSegments) cursor[t * kCursorStride] = 0u;
        if (sums && t < 2) sums[t] = 0ull;
    }
    const uint32_t pix_count = w * h;
    const uint32_t total = pix_count * frames;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < total; p += gridDim.x * blockDim.x) {
        uint32_t f = p / pix_count, pix = p - f * pix_count;
        uint32_t y = pix / w, x = pix - y * w;
        float u = (float)x / (float)w, v = (float)y / (float)h;
        f3 dir = ((cam.llc + u * cam.horizontal) + v * cam.vertical) - cam.origin;
        dir = normalize(dir);
        float4* rp = reinterpret_cast<float4*>(rays) + 2 * (size_t)p;
        rp[0] = make_float4(cam.origin.x, cam.origin.y, cam.origin.z, 0.001f);
        rp[1] = make_float4(dir.x, dir.y, dir.z, 10000.0f);
        pid[p] = p;
        color[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        rng[p] = rng_seed(x, y, sample0 + f);
    }
}

// One bounce of ray_color's loop for every live ray of the queue.
__global__ __launch_bounds__(256) void k_shade(
    const rt_ray* __restrict__ rays, const uint32_t* __restrict__ pid_in, const uint32_t* __restrict__ n_in,
    uint32_t n_host, const rt_hit* __restrict__ hits, rt_ray* __restrict__ out_rays, uint32_t* __restrict__ out_pid,
    uint32_t* __restrict__ n_out, float4* __restrict__ color, uint32_t* __restrict__ rngs,
    const rt_triangle* __restrict__ tris, const rt_material* __restrict__ mats, const rt_light* __restrict__ lights,
    uint32_t nlights, const rt_sphere* __restrict__ spheres, uint32_t ntri, uint32_t last,
    uint32_t* __restrict__ next_cursor) {
    // the next bounce's traversal finds its ray cursors zeroed (no fill launch)
    if (next_cursor && blockIdx.x == 0 && threadIdx.x < kSegments) next_cursor[threadIdx.x * kCursorStride] = 0u;
    const uint32_t n = n_in ? *n_in : n_host;
    const uint32_t stride = gridDim.x * blockDim.x;
    // every lane of a wave runs the same number of iterations so the
    // ballot in wave_append sees the whole wave
    const uint32_t n_round = (n + 63u) & ~63u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += stride) {
        bool live = i < n;
        bool keep = false;
        f3 o{0, 0, 0}, d{0, 0, 0};
        uint32_t pid = 0;
        if (live) {
            pid = pid_in[i];
            const float4* rp = reinterpret_cast<const float4*>(rays) + 2 * (size_t)i;
            float4 r0 = rp[0], r1 = rp[1];
            f3 ro{r0.x, r0.y, r0.z}, rd{r1.x, r1.y, r1.z};
            float4 hh = reinterpret_cast<const float4*>(hits)[i];
            int32_t id = __float_as_int(hh.x);
            float t = hh.y;
            float4 c4 = color[pid];
            f3 c{c4.x, c4.y, c4.z};
            if (id < 0) {
                c = c * 0.0f;                       // miss: final_color *= 0.0
            } else {
                HitRec rec;
                if ((uint32_t)id < ntri) {          // hit_triangle (:205-220)
                    const rt_triangle& tr = tris[id];
                    f3 v0 = ld3(tr.v0), v1 = ld3(tr.v1), v2 = ld3(tr.v2);
                    f3 a = v0 - v1;
                    f3 b = v2 - v0;
                    f3 nn = cross(b, a);
                    rec.p = ro + t * rd;
                    rec.normal = normalize(nn);
                    rec.back = dot(rd, rec.normal) > 0.0f ? 1 : 0;
                    rec.normal = rec.normal * (float)(1 - 2 * rec.back);
                    rec.p = rec.p + rec.normal * 0.0001f;
                    rec.mat = tr.materialIndex;
                } else {                            // hit_sphere (:198-201)
                    const rt_sphere& s = spheres[(uint32_t)id - ntri];
                    f3 center{s.s[0], s.s[1], s.s[2]};
                    rec.back = hh.z != 0.0f ? 1 : 0;
                    rec.p = ro + t * rd;
                    rec.normal = ((float)(1 - 2 * rec.back) * (rec.p - center)) / s.s[3];
                    rec.mat = s.materialIndex;
                }
                // scatter (:118-155)
                const rt_material& m = mats[rec.mat];
                f3 albedo = ld3(m.albedo);
                if (m.type == (uint32_t)RT_MAT_GLASS) albedo = f3{1.0f, 1.0f, 1.0f};
                c = c * albedo;
                if (m.type != (uint32_t)RT_MAT_LIGHT && !last) {
                    uint32_t rng = rngs[pid];
                    f3 material_sample{0.0f, 0.0f, 0.0f};
                    if (m.type == (uint32_t)RT_MAT_LAMBERTIAN) material_sample = sample_lambertian(rec.normal, rng);
                    else if (m.type == (uint32_t)RT_MAT_METAL) material_sample = reflect(rd, rec.normal);
                    else if (m.type == (uint32_t)RT_MAT_GLASS) material_sample = sample_glass(rd, rec);
                    f3 final_sample = material_sample;
                    if (rng_next(rng) < 0.5f && m.type == (uint32_t)RT_MAT_LAMBERTIAN) {
                        float light_cosine;
                        final_sample = sample_light(tris, lights, nlights, rec.p, light_cosine, rng);
                        if (fabsf(light_cosine) < 0.001f) final_sample = material_sample;
                    }
                    rngs[pid] = rng;
                    o = rec.p;
                    d = final_sample;
                    keep = true;
                }
            }
            color[pid] = make_float4(c.x, c.y, c.z, 0.0f);
        }
        wave_append(keep, n_out, out_rays, out_pid, o, d, pid);
    }
}

__device__ __forceinline__ float unorm8_load(uint32_t c) { return (float)c / 255.0f; }
__device__ __forceinline__ uint32_t unorm8_store(float x) {
    if (!(x > 0.0f)) x = 0.0f;
    if (x > 1.0f) x = 1.0f;
    return (uint32_t)rintf(x * 255.0f);
}

// main()'s accumulation (:375-379) applied frame after frame per pixel.
__global__ __launch_bounds__(256) void k_store(uint32_t pix_count, uint32_t frames, uint32_t sample0,
                                               const float4* __restrict__ color, const uint32_t* __restrict__ accum_in,
                                               uint32_t* __restrict__ target, uint32_t* __restrict__ accum_out) {
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < pix_count; pix += gridDim.x * blockDim.x) {
        uint32_t acc = accum_in[pix];
        for (uint32_t f = 0; f < frames; ++f) {
            float4 c = color[(size_t)f * pix_count + pix];
            float cs = (float)(sample0 + f);
            float m = gmin(cs, 1.0f);
            float c4[4] = {c.x, c.y, c.z, 1.0f};
            uint32_t out = 0;
            for (int k = 0; k < 4; ++k) {
                float cur = unorm8_load((acc >> (8 * k)) & 0xffu) * m;
                float tw = (c4[k] + cur * cs) / (cs + 1.0f);
                out |= unorm8_store(tw) << (8 * k);
            }
            acc = out;
        }
        target[pix] = acc;
        if (accum_out) accum_out[pix] = acc;
    }
}

// Camera block of main() (:357-369) on the host, same float operations.
Camera make_camera(const rt_ubo& ubo, float vfov, uint32_t w, uint32_t h) {
    Camera c;
    float isx = (float)w, isy = (float)h;
    float theta = vfov * kPi / 180.0f;
    float hh = (float)std::tan((double)(theta / 2.0f));
    float viewport_height = 2.0f * hh;
    float viewport_width = isx / isy * viewport_height;
    float focal_length = 1.0f;
    c.horizontal = f3{viewport_width, 0.0f, 0.0f};
    c.vertical = f3{0.0f, -viewport_height, 0.0f};
    c.origin = f3{ubo.camPos[2], ubo.camPos[0], ubo.camPos[1]} * f3{-1.0f, 1.0f, 1.0f};
    c.llc = ((c.origin - c.horizontal / 2.0f) - c.vertical / 2.0f) - f3{0.0f, 0.0f, focal_length};
    return c;
}

inline uint32_t grid_for(uint64_t n, uint32_t cap) {
    uint64_t b = (n + 255) / 256;
    if (b > cap) b = cap;
    return b ? (uint32_t)b : 1u;
}

}  // namespace

// Renders `frames` frames; accum_out may be null (single-frame call, the
// caller copies the target into the accumulation image as main.cpp does).
int render_impl(rt_ctx* ctx, const rt_scene* sc, const rt_ubo& ubo, const rt_render_params& prm, uint32_t frames,
                uint32_t* target, const uint32_t* accum_in, uint32_t* accum_out, uint32_t w, uint32_t h,
                hipStream_t s) {
    const uint64_t P64 = (uint64_t)w * h * frames;
    if (P64 == 0) return RT_OK;
    if (P64 > 0x7fffffffull) return set_error(ctx, RT_ERR_INVALID, "render: too many paths in one call");
    const uint32_t P = (uint32_t)P64;
    const uint32_t bounces = prm.bounces;
    const uint32_t trav = prm.traversal;
    const bool use_packed = trav == RT_TRAVERSE_PACKED || (trav == RT_TRAVERSE_AUTO && sc->fast && prm.max_stack == 0);
    if (trav == RT_TRAVERSE_PACKED && (!sc->fast || prm.max_stack))
        return set_error(ctx, RT_ERR_UNSUPPORTED, "render: packed traversal needs an enclosing tree and max_stack 0");

    // workspace: 2 ray queues, hits, 2 path-id queues, colour, rng, counters
    const size_t rq = (size_t)P * sizeof(rt_ray);
    const size_t need = 2 * rq + (size_t)P * sizeof(rt_hit) + 2 * (size_t)P * 4 + (size_t)P * 16 + (size_t)P * 4 + 256;
    int rc = ensure_workspace(ctx, kWsRender, need);
    if (rc) return rc;
    char* base = (char*)ctx->ws[kWsRender];
    rt_ray* rays[2] = {(rt_ray*)base, (rt_ray*)(base + rq)};
    base += 2 * rq;
    rt_hit* hits = (rt_hit*)base; base += (size_t)P * sizeof(rt_hit);
    uint32_t* pids[2] = {(uint32_t*)base, (uint32_t*)(base + (size_t)P * 4)};
    base += 2 * (size_t)P * 4;
    float4* color = (float4*)base; base += (size_t)P * 16;
    uint32_t* rng = (uint32_t*)base; base += (size_t)P * 4;
    uint32_t* counters = ctx->d_counters;   // [b] = live rays entering bounce b

    if (bounces + 1 > kRenderCounters) return set_error(ctx, RT_ERR_INVALID, "render: at most 63 bounces");
    unsigned long long* sums = ctx->count_render ? ctx->d_render_sums : nullptr;
    // each bounce's packed traversal takes a cursor block that the kernel
    // before it (generate, then the previous bounce's shading) zeroed
    uint32_t* cursor = use_packed ? next_cursor(ctx) : nullptr;

    // timing mode: ev[0] before generate, ev[1+2b] / ev[2+2b] around bounce
    // b's traversal, ev[1+2*bounces] after store
    hipEvent_t* ev = ctx->time_render ? ctx->render_ev : nullptr;
    if (ev) {
        for (uint32_t i = 0; i < 2 * bounces + 2; ++i)
            if (!ev[i]) RT_HIP(ctx, hipEventCreate(&ev[i]));
        RT_HIP(ctx, hipEventRecord(ev[0], s));
    }

    Camera cam = make_camera(ubo, prm.vfov_deg, w, h);
    const uint32_t cu_cap = (uint32_t)ctx->num_cus * 16u;
    hipLaunchKernelGGL(k_generate, dim3(grid_for(P, cu_cap)), dim3(256), 0, s, w, h, frames, ubo.currentSample, cam,
                       rays[0], pids[0], color, rng, counters, cursor, sums);
    RT_HIP(ctx, hipGetLastError());

    for (uint32_t b = 0; b < bounces; ++b) {
        const uint32_t* n_dev = b == 0 ? nullptr : counters + b;
        TraceArgs ta{sc, rays[b & 1], P, n_dev, hits, prm.spheres, nullptr, nullptr, true};
        if (use_packed) {
            ta.sums = sums;
            ta.cursor = cursor;
            ta.cursor_zeroed = true;
        }
        if (ev) RT_HIP(ctx, hipEventRecord(ev[1 + 2 * b], s));
        // secondary bounces start on surfaces all over the scene: trace them
        // in origin order (camera rays share one origin and are coherent)
        if (b > 0 && use_packed) {
            rc = ray_order(ctx, ta, s, &ta.order);
            if (rc) return rc;
        }
        rc = use_packed ? launch_trace_packed(ctx, ta, s) : launch_trace_reference(ctx, ta, prm.max_stack, s);
        if (rc) return rc;
        rc = launch_sphere_pass(ctx, ta, s);
        if (rc) return rc;
        if (ev) RT_HIP(ctx, hipEventRecord(ev[2 + 2 * b], s));
        cursor = use_packed && b + 1 < bounces ? next_cursor(ctx) : nullptr;
        hipLaunchKernelGGL(k_shade, dim3(grid_for(P, cu_cap)), dim3(256), 0, s, rays[b & 1], pids[b & 1], n_dev, P,
                           hits, rays[(b + 1) & 1], pids[(b + 1) & 1], counters + b + 1, color, rng, sc->d_tris,
                           sc->d_mats, sc->d_lights, sc->info.num_lights, sc->d_spheres, sc->info.num_triangles,
                           (uint32_t)(b + 1 == bounces), cursor);
        RT_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_store, dim3(grid_for((uint64_t)w * h, cu_cap)), dim3(256), 0, s, w * h, frames,
                       ubo.currentSample, color, accum_in, target, accum_out);
    RT_HIP(ctx, hipGetLastError());
    if (ev) RT_HIP(ctx, hipEventRecord(ev[1 + 2 * bounces], s));
    return RT_OK;
}

}  // namespace rtamd
