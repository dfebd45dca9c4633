// rt_workloads.cpp -- synthetic scenes and ray batches (rtamd_workloads.h).
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rt_common.hpp"
#include "rt_math.hpp"
#include "rtamd_workloads.h"

#pragma clang fp contract(off)

namespace rtamd {
namespace {

// SplitMix64 (Steele, Lea & Flood 2014).
struct SplitMix {
    uint64_t s;
    explicit SplitMix(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform() { return (float)(next() >> 40) * 0x1p-24f; }               // [0, 1)
    float uniform(float a, float b) { return a + (b - a) * uniform(); }
    double gaussian() {                                                         // Box-Muller
        double u1 = ((double)(next() >> 11) + 1.0) * 0x1p-53;
        double u2 = (double)(next() >> 11) * 0x1p-53;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    f3 gauss3() { return f3{(float)gaussian(), (float)gaussian(), (float)gaussian()}; }
};

uint64_t mix64(uint64_t a, uint64_t b) {
    SplitMix m(a ^ (b * 0xD1B54A32D192ED03ull));
    return m.next();
}

template <typename T>
T* dup(const std::vector<T>& v) {
    if (v.empty()) return nullptr;
    T* p = (T*)std::malloc(sizeof(T) * v.size());
    if (p) std::memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

rt_material mat(uint32_t type, float r, float g, float b) {
    rt_material m;
    std::memset(&m, 0, sizeof m);
    m.type = type;
    m.albedo[0] = r; m.albedo[1] = g; m.albedo[2] = b;
    return m;
}

rt_triangle tri(f3 a, f3 b, f3 c, uint32_t m) {
    rt_triangle t;
    std::memset(&t, 0, sizeof t);
    t.v0[0] = a.x; t.v0[1] = a.y; t.v0[2] = a.z;
    t.v1[0] = b.x; t.v1[1] = b.y; t.v1[2] = b.z;
    t.v2[0] = c.x; t.v2[1] = c.y; t.v2[2] = c.z;
    t.materialIndex = m;
    return t;
}

std::vector<rt_material> standard_materials() {
    return {mat(RT_MAT_LAMBERTIAN, 0.7f, 0.7f, 0.7f), mat(RT_MAT_METAL, 1.0f, 1.0f, 1.0f),
            mat(RT_MAT_GLASS, 1.0f, 1.0f, 1.0f), mat(RT_MAT_LIGHT, 4.0f, 4.0f, 4.0f)};
}

int finish(std::vector<rt_triangle>& tris, std::vector<rt_material>& mats, std::vector<rt_sphere>& spheres,
           uint32_t threads, rtw_scene* out) {
    const uint32_t n = (uint32_t)tris.size();
    std::vector<rt_bvh_node> nodes(2 * (size_t)n - 1);
    int rc = rt_bvh_build_seeded(tris.data(), n, 1, threads, nodes.data());
    if (rc) return rc;
    uint32_t nl = 0;
    rc = rt_scene_lights(tris.data(), n, mats.data(), (uint32_t)mats.size(), nullptr, 0, &nl);
    if (rc) return rc;
    std::vector<rt_light> lights(nl);
    rc = rt_scene_lights(tris.data(), n, mats.data(), (uint32_t)mats.size(), lights.data(), nl, &nl);
    if (rc) return rc;
    std::memset(out, 0, sizeof *out);
    out->triangles = dup(tris); out->num_triangles = n;
    out->materials = dup(mats); out->num_materials = (uint32_t)mats.size();
    out->nodes = dup(nodes); out->num_nodes = (uint32_t)nodes.size();
    out->lights = dup(lights); out->num_lights = nl;
    out->spheres = dup(spheres); out->num_spheres = (uint32_t)spheres.size();
    if (!out->triangles || !out->materials || !out->nodes || (nl && !out->lights) || (!spheres.empty() && !out->spheres)) {
        rtw_scene_free(out);
        return RT_ERR_NOMEM;
    }
    return RT_OK;
}

// 3-D lattice value noise: a hashed value in [-1, 1] at each integer point,
// blended trilinearly with smoothstep weights.
double lattice(int64_t x, int64_t y, int64_t z, uint64_t seed) {
    uint64_t h = mix64(mix64(mix64(seed, (uint64_t)x), (uint64_t)y), (uint64_t)z);
    return (double)(h >> 11) * 0x1p-53 * 2.0 - 1.0;
}
double value_noise(double x, double y, double z, uint64_t seed) {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
    auto sm = [](double t) { return t * t * (3.0 - 2.0 * t); };
    double tx = sm(x - fx), ty = sm(y - fy), tz = sm(z - fz);
    double acc = 0.0;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                double w = (dx ? tx : 1.0 - tx) * (dy ? ty : 1.0 - ty) * (dz ? tz : 1.0 - tz);
                acc += w * lattice(ix + dx, iy + dy, iz + dz, seed);
            }
    return acc;
}

// Page-locked buffers released together.
struct PinnedArena {
    std::vector<void*> ptrs;
    ~PinnedArena() {
        for (void* q : ptrs) rt_host_free(q);
    }
    void* get(size_t bytes) {
        void* q = nullptr;
        if (rt_host_alloc(bytes, &q) != RT_OK) return nullptr;
        ptrs.push_back(q);
        return q;
    }
};

}  // namespace
}  // namespace rtamd

using namespace rtamd;

extern "C" void rtw_scene_free(rtw_scene* s) {
    if (!s) return;
    std::free(s->triangles);
    std::free(s->materials);
    std::free(s->nodes);
    std::free(s->lights);
    std::free(s->spheres);
    std::memset(s, 0, sizeof *s);
}

extern "C" uint32_t rtw_log_uniform(uint64_t seed, uint32_t lo, uint32_t hi) {
    SplitMix r(seed);
    double u = (double)(r.next() >> 11) * 0x1p-53;
    double v = std::floor(std::exp(u * std::log((double)hi / (double)lo)) * (double)lo);
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return (uint32_t)v;
}

extern "C" int rtw_scene_soup(uint32_t n, uint64_t seed, uint32_t threads, rtw_scene* out) {
    if (!out || n == 0) return RT_ERR_INVALID;
    try {
        SplitMix r(seed);
        std::vector<rt_triangle> tris(n);
        for (uint32_t i = 0; i < n; ++i) {
            f3 c{r.uniform(-1.0f, 1.0f), r.uniform(-1.0f, 1.0f), r.uniform(-1.0f, 1.0f)};
            f3 v[3];
            for (int k = 0; k < 3; ++k)
                v[k] = c + f3{r.uniform(-0.1f, 0.1f), r.uniform(-0.1f, 0.1f), r.uniform(-0.1f, 0.1f)};
            tris[i] = tri(v[0], v[1], v[2], i % 4u);
        }
        std::vector<rt_material> mats = standard_materials();
        std::vector<rt_sphere> spheres(2);
        std::memset(spheres.data(), 0, sizeof(rt_sphere) * 2);
        spheres[0].s[0] = 0.6f; spheres[0].s[3] = 0.3f; spheres[0].materialIndex = 0;
        spheres[1].s[0] = -0.6f; spheres[1].s[3] = 0.3f; spheres[1].materialIndex = 0;
        return finish(tris, mats, spheres, threads, out);
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    }
}

extern "C" int rtw_scene_doge(const char* dir, uint32_t threads, rtw_scene* out) {
    if (!dir || !out) return RT_ERR_INVALID;
    try {
        // materials of RtScene.h:49-62, in that order
        std::vector<rt_material> mats = {mat(RT_MAT_LAMBERTIAN, 0.3f, 0.3f, 0.3f), mat(RT_MAT_LAMBERTIAN, 0.9f, 0.1f, 0.1f),
                                         mat(RT_MAT_LAMBERTIAN, 0.1f, 0.9f, 0.1f), mat(RT_MAT_LIGHT, 2.0f, 2.0f, 2.0f),
                                         mat(RT_MAT_METAL, 1.0f, 1.0f, 1.0f), mat(RT_MAT_GLASS, 1.0f, 1.0f, 1.0f)};
        // models and their materials, concatenated in the order of RtScene.h:81-88
        struct Part {
            const char* file;
            uint32_t material;
        };
        const Part parts[] = {{"buff-doge.obj", 0}, {"cheems.obj", 0}, {"right.obj", 1}, {"left.obj", 2},
                              {"back.obj", 0},      {"ceil.obj", 0},   {"floor.obj", 0}, {"light.obj", 3}};
        std::vector<rt_triangle> tris;
        for (const Part& pt : parts) {
            std::string err;
            const std::string path = std::string(dir) + "/" + pt.file;
            int rc = rtamd::obj_triangles(path.c_str(), pt.material, tris, err);
            if (rc) return rc;
        }
        if (tris.empty()) return RT_ERR_INVALID;
        // the glass sphere of RtScene.h:97: centre (0.6, 1, -1), radius 0.6, material 5
        std::vector<rt_sphere> spheres(1);
        std::memset(spheres.data(), 0, sizeof(rt_sphere));
        spheres[0].s[0] = 0.6f; spheres[0].s[1] = 1.0f; spheres[0].s[2] = -1.0f; spheres[0].s[3] = 0.6f;
        spheres[0].materialIndex = 5;
        return finish(tris, mats, spheres, threads, out);
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    }
}

extern "C" int rtw_scene_icosphere(uint32_t level, uint64_t seed, float amp, float freq, uint32_t threads,
                                   rtw_scene* out) {
    if (!out || level > 10) return RT_ERR_INVALID;
    try {
        const double t = (1.0 + std::sqrt(5.0)) / 2.0;
        std::vector<double> vx = {-1, 1, -1, 1, 0, 0, 0, 0, t, t, -t, -t};
        std::vector<double> vy = {t, t, -t, -t, -1, 1, -1, 1, 0, 0, 0, 0};
        std::vector<double> vz = {0, 0, 0, 0, t, t, -t, -t, -1, 1, -1, 1};
        for (size_t i = 0; i < vx.size(); ++i) {
            double l = std::sqrt(vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
            vx[i] /= l; vy[i] /= l; vz[i] /= l;
        }
        std::vector<uint32_t> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2,
                                   10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5,
                                   2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
        for (uint32_t l = 0; l < level; ++l) {
            std::unordered_map<uint64_t, uint32_t> cache;
            auto midpoint = [&](uint32_t a, uint32_t b) -> uint32_t {
                uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
                auto it = cache.find(key);
                if (it != cache.end()) return it->second;
                double x = (vx[a] + vx[b]) * 0.5, y = (vy[a] + vy[b]) * 0.5, z = (vz[a] + vz[b]) * 0.5;
                double len = std::sqrt(x * x + y * y + z * z);
                vx.push_back(x / len); vy.push_back(y / len); vz.push_back(z / len);
                uint32_t id = (uint32_t)vx.size() - 1;
                cache.emplace(key, id);
                return id;
            };
            std::vector<uint32_t> nf;
            nf.reserve(f.size() * 4);
            for (size_t i = 0; i < f.size(); i += 3) {
                uint32_t a = f[i], b = f[i + 1], c = f[i + 2];
                uint32_t ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
                uint32_t q[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
                nf.insert(nf.end(), q, q + 12);
            }
            f.swap(nf);
        }
        std::vector<f3> pos(vx.size());
        for (size_t i = 0; i < vx.size(); ++i) {
            double s = 1.0 + (double)amp * value_noise(freq * vx[i], freq * vy[i], freq * vz[i], seed);
            pos[i] = f3{(float)(vx[i] * s), (float)(vy[i] * s), (float)(vz[i] * s)};
        }
        std::vector<rt_triangle> tris(f.size() / 3);
        for (size_t i = 0; i < tris.size(); ++i) tris[i] = tri(pos[f[3 * i]], pos[f[3 * i + 1]], pos[f[3 * i + 2]], 0);
        std::vector<rt_material> mats = {mat(RT_MAT_LAMBERTIAN, 0.7f, 0.7f, 0.7f)};
        std::vector<rt_sphere> spheres;
        return finish(tris, mats, spheres, threads, out);
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    }
}

extern "C" int rtw_scene_clusters(uint32_t n, uint64_t seed, uint32_t threads, rtw_scene* out) {
    if (!out || n == 0) return RT_ERR_INVALID;
    try {
        const uint32_t per = 1024;
        const uint32_t nc = (n + per - 1) / per;
        SplitMix r(seed);
        std::vector<f3> centres(nc);
        for (auto& c : centres) c = f3{r.uniform(-4.0f, 4.0f), r.uniform(-4.0f, 4.0f), r.uniform(-4.0f, 4.0f)};
        std::vector<rt_triangle> tris(n);
        auto gen = [&](uint32_t c0, uint32_t c1) {
            for (uint32_t c = c0; c < c1; ++c) {
                SplitMix cr(mix64(seed, c + 1));
                uint32_t lo = c * per, hi = std::min(n, lo + per);
                for (uint32_t i = lo; i < hi; ++i) {
                    f3 g = cr.gauss3();
                    f3 cen = centres[c] + 0.15f * g;
                    float edge = cr.uniform(0.005f, 0.03f);
                    f3 a = normalize(cr.gauss3());
                    f3 b = normalize(cross(a, cr.gauss3()));
                    double th0 = 6.283185307179586 * (double)cr.uniform();
                    float R = edge / std::sqrt(3.0f);
                    f3 v[3];
                    for (int k = 0; k < 3; ++k) {
                        double th = th0 + 2.0943951023931953 * k;
                        float ck = (float)std::cos(th), sk = (float)std::sin(th);
                        v[k] = cen + R * (ck * a + sk * b);
                    }
                    float u = cr.uniform();
                    uint32_t m = u < 0.60f ? 0u : u < 0.85f ? 1u : u < 0.95f ? 2u : 3u;
                    tris[i] = tri(v[0], v[1], v[2], m);
                }
            }
        };
        uint32_t T = std::max(1u, std::min(threads, 64u));
        if (T == 1 || nc < 64) {
            gen(0, nc);
        } else {
            std::vector<std::thread> th;
            for (uint32_t t = 0; t < T; ++t) th.emplace_back(gen, nc * t / T, nc * (t + 1) / T);
            for (auto& x : th) x.join();
        }
        std::vector<rt_material> mats = standard_materials();
        std::vector<rt_sphere> spheres;
        return finish(tris, mats, spheres, threads, out);
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    }
}

extern "C" int rtw_rays_camera_tiles(const float cam_pos[3], float vfov_deg, uint32_t w, uint32_t h, uint32_t tile,
                                     uint32_t first, uint32_t count, rt_ray* out) {
    if (!cam_pos || !out || !tile || w % tile || h % tile) return RT_ERR_INVALID;
    if ((uint64_t)first + count > (uint64_t)w * h) return RT_ERR_INVALID;
    // camera block of ray-trace-compute.comp:357-372 with the given vfov
    float isx = (float)w, isy = (float)h;
    float theta = vfov_deg * kPi / 180.0f;
    float hh = (float)std::tan((double)(theta / 2.0f));
    float vh = 2.0f * hh, vw = isx / isy * vh;
    f3 hor{vw, 0.0f, 0.0f}, ver{0.0f, -vh, 0.0f};
    f3 origin = f3{cam_pos[2], cam_pos[0], cam_pos[1]} * f3{-1.0f, 1.0f, 1.0f};
    f3 llc = ((origin - hor / 2.0f) - ver / 2.0f) - f3{0.0f, 0.0f, 1.0f};
    const uint32_t tiles_x = w / tile, tt = tile * tile;
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t q = first + k, t = q / tt, in = q % tt;
        uint32_t x = (t % tiles_x) * tile + in % tile, y = (t / tiles_x) * tile + in / tile;
        float u = (float)x / isx, v = (float)y / isy;
        f3 d = normalize(((llc + u * hor) + v * ver) - origin);
        rt_ray& r = out[k];
        r.o[0] = origin.x; r.o[1] = origin.y; r.o[2] = origin.z; r.tmin = 0.001f;
        r.d[0] = d.x; r.d[1] = d.y; r.d[2] = d.z; r.tmax = 10000.0f;
    }
    return RT_OK;
}

extern "C" int rtw_rays_shell(uint32_t n, uint64_t seed, float radius, float half, rt_ray* out) {
    if (!out && n) return RT_ERR_INVALID;
    SplitMix r(seed);
    for (uint32_t i = 0; i < n; ++i) {
        f3 o = radius * normalize(r.gauss3());
        f3 target{r.uniform(-half, half), r.uniform(-half, half), r.uniform(-half, half)};
        f3 d = normalize(target - o);
        rt_ray& q = out[i];
        q.o[0] = o.x; q.o[1] = o.y; q.o[2] = o.z; q.tmin = 0.001f;
        q.d[0] = d.x; q.d[1] = d.y; q.d[2] = d.z; q.tmax = 10000.0f;
    }
    return RT_OK;
}

extern "C" int rtw_rays_box(uint32_t n, uint64_t seed, float extent, rt_ray* out) {
    if (!out && n) return RT_ERR_INVALID;
    SplitMix r(seed);
    for (uint32_t i = 0; i < n; ++i) {
        f3 o{r.uniform(-extent, extent), r.uniform(-extent, extent), r.uniform(-extent, extent)};
        f3 d = normalize(r.gauss3());
        rt_ray& q = out[i];
        q.o[0] = o.x; q.o[1] = o.y; q.o[2] = o.z; q.tmin = 0.001f;
        q.d[0] = d.x; q.d[1] = d.y; q.d[2] = d.z; q.tmax = 10000.0f;
    }
    return RT_OK;
}

// ---------------------------------------------------------------------------
// Poisson request stream against a server (rtamd_workloads.h).
extern "C" int rtw_serve_poisson(rt_server* S, const rtw_serve_params* p, rtw_serve_result* res) {
    return rtw_serve_poisson_sampled(S, p, nullptr, res);
}

extern "C" int rtw_serve_poisson_sampled(rt_server* S, const rtw_serve_params* p, rtw_serve_sampling* smp,
                                         rtw_serve_result* res) {
    using Clock = std::chrono::steady_clock;
    if (!S || !p || !res || p->rate_per_s <= 0.0 || p->duration_s <= 0.0 || p->query_min == 0 ||
        p->query_max < p->query_min || p->render_w == 0 || p->render_h == 0)
        return RT_ERR_INVALID;
    if (smp) {
        if (smp->every == 0 || (smp->capacity && !smp->samples) || (smp->swaps_capacity && !smp->swap_ids))
            return RT_ERR_INVALID;
        for (uint32_t i = 0; i < smp->capacity; ++i)
            if (!smp->samples[i].hits || !smp->samples[i].image) return RT_ERR_INVALID;
        smp->taken = 0;
        smp->swaps_taken = 0;
        for (uint32_t i = 0; i < smp->swaps_capacity; ++i) smp->swap_ids[i] = 0;   // 0: that swap failed
    }
    std::memset(res, 0, sizeof *res);
    rt_server_stats_reset(S);
    SplitMix rng(p->seed);
    // request buffers in page-locked memory, as a serving client would use
    const int kRing = 32;
    PinnedArena mem;
    rt_ray* pool = (rt_ray*)mem.get((size_t)p->query_max * sizeof(rt_ray));
    std::vector<rt_hit*> hit_ring(kRing);
    std::vector<uint8_t*> tgt_ring(kRing);
    for (int i = 0; i < kRing; ++i) {
        hit_ring[i] = (rt_hit*)mem.get((size_t)p->query_max * sizeof(rt_hit));
        tgt_ring[i] = (uint8_t*)mem.get((size_t)p->render_w * p->render_h * 4);
        if (!hit_ring[i] || !tgt_ring[i]) return RT_ERR_NOMEM;
    }
    if (!pool) return RT_ERR_NOMEM;
    rtw_rays_box(p->query_max, p->seed ^ 0x5eedull, 4.0f, pool);

    std::mutex ids_mu;
    std::vector<uint64_t> ids;
    auto refresh = [&] {
        uint32_t cnt = 0;
        rt_server_scene_list(S, nullptr, 0, &cnt);
        std::vector<uint64_t> v(cnt);
        rt_server_scene_list(S, v.data(), cnt, &cnt);
        v.resize(std::min<size_t>(v.size(), cnt));
        std::lock_guard<std::mutex> lk(ids_mu);
        ids.swap(v);
    };
    refresh();
    if (ids.empty()) return RT_ERR_NOT_FOUND;

    const auto t_start = Clock::now();
    std::atomic<bool> done{false};
    std::atomic<uint64_t> swaps{0};
    std::thread swapper([&] {
        if (p->swap_interval_s <= 0.0 || p->swap_triangles == 0) return;
        uint32_t k = 0;
        auto next = t_start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(p->swap_interval_s));
        while (!done.load()) {
            while (!done.load() && Clock::now() < next) std::this_thread::sleep_for(std::chrono::milliseconds(2));
            if (done.load()) break;
            rtw_scene ws;
            if (rtw_scene_clusters(p->swap_triangles, p->seed + 1000 + k, std::max(1u, p->threads), &ws) == RT_OK) {
                rt_scene_desc d{ws.triangles, ws.num_triangles, ws.materials, ws.num_materials, ws.nodes, ws.num_nodes,
                                ws.lights, ws.num_lights, ws.spheres, ws.num_spheres};
                uint64_t id = 0;
                if (rt_server_scene_add(S, &d, &id) == RT_OK) {
                    swaps++;
                    if (smp && k < smp->swaps_capacity) {
                        smp->swap_ids[k] = id;
                        smp->swaps_taken = k + 1;
                    }
                }
                rtw_scene_free(&ws);
                refresh();
            }
            ++k;
            next += std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(p->swap_interval_s));
        }
    });

    double t = 0.0;
    uint64_t k = 0;
    const double lq = std::log((double)p->query_max / (double)p->query_min);
    for (;;) {
        double u = ((double)(rng.next() >> 11) + 1.0) * 0x1p-53;
        t += -std::log(u) / p->rate_per_s;
        if (t >= p->duration_s) break;
        std::this_thread::sleep_until(t_start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t)));
        const bool render = (double)(rng.next() >> 11) * 0x1p-53 < p->render_fraction;
        const double us = (double)(rng.next() >> 11) * 0x1p-53;
        const double un = (double)(rng.next() >> 11) * 0x1p-53;
        // a kept arrival writes to its own buffers, which no later request reuses
        rtw_serve_sample* keep =
            (smp && k % smp->every == 0 && smp->taken < smp->capacity) ? &smp->samples[smp->taken] : nullptr;
        rt_hit* hits_out = keep ? keep->hits : hit_ring[k % kRing];
        uint8_t* image_out = keep ? keep->image : tgt_ring[k % kRing];
        for (int attempt = 0; attempt < 2; ++attempt) {
            uint64_t id;
            {
                std::lock_guard<std::mutex> lk(ids_mu);
                if (ids.empty()) break;
                id = ids[std::min(ids.size() - 1, (size_t)(us * (double)ids.size()))];
            }
            int rc;
            uint32_t n = 0;
            if (render) {
                rt_ubo ubo{};
                ubo.camPos[1] = 10.0f;            // eye at (0, 0, 10)
                rt_render_params rp{};
                rp.vfov_deg = 60.0f;
                rp.bounces = p->render_bounces;
                rc = rt_server_submit_render(S, id, &ubo, &rp, p->render_w, p->render_h, image_out, nullptr);
            } else {
                n = (uint32_t)std::floor(std::exp(un * lq) * (double)p->query_min);
                n = std::max(p->query_min, std::min(p->query_max, n));
                rc = rt_server_submit_query(S, id, pool, n, hits_out, nullptr);
            }
            if (rc == RT_ERR_NOT_FOUND) { refresh(); continue; }   // the scene was just evicted
            if (rc == RT_OK) {
                res->requests++;
                (render ? res->renders : res->queries)++;
                if (keep) {
                    keep->scene_id = id;
                    keep->arrival = k;
                    keep->kind = render ? 1u : 0u;
                    keep->count = n;
                    smp->taken++;
                }
            }
            break;
        }
        ++k;
    }
    done = true;
    swapper.join();
    rt_server_drain(S);
    res->wall_s = std::chrono::duration<double>(Clock::now() - t_start).count();
    res->swaps = swaps.load();
    return rt_server_stats_get(S, &res->stats);
}
