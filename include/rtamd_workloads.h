/*
 * rtamd_workloads.h -- synthetic scenes and ray batches of the benchmark
 * workloads (part of librtamd.so).  These have no counterpart in the
 * reference, whose only scene is the hard-coded doge scene of
 * src/ray-tracing/RtScene.h:44-101; they build inputs of the same record
 * layouts so tests, benchmarks and the server can run without assets.
 *
 * Randomness: SplitMix64 (Steele, Lea & Flood, OOPSLA 2014) seeded with the
 * workload seed; uniform floats take the top 24 bits (x >> 40) * 2^-24;
 * Gaussians use the Box-Muller transform in double precision.  BVHs come
 * from rt_bvh_build_seeded(seed 1) -- the reference builder's tree in a
 * process whose rand() is untouched.  Every array is malloc'ed by the
 * library and released with rtw_scene_free / free().
 */
#ifndef RTAMD_WORKLOADS_H
#define RTAMD_WORKLOADS_H

#include "rtamd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtw_scene {
    rt_triangle* triangles; uint32_t num_triangles;
    rt_material* materials; uint32_t num_materials;
    rt_bvh_node* nodes;     uint32_t num_nodes;
    rt_light*    lights;    uint32_t num_lights;
    rt_sphere*   spheres;   uint32_t num_spheres;
} rtw_scene;

/* Random soup: triangle i has centroid U[-1,1]^3 and vertices centroid +
 * U[-0.1,0.1]^3; materials [0] lambertian 0.7, [1] metal 1.0, [2] glass,
 * [3] light 4.0 assigned by i mod 4; spheres (+-0.6, 0, 0, r 0.3) with
 * material 0; lights by the reference rule. */
int rtw_scene_soup(uint32_t n, uint64_t seed, uint32_t threads, rtw_scene* out);

/* Level-`level` subdivided icosahedron (20 * 4^level triangles) on the unit
 * sphere, each vertex moved radially by amp * value_noise(seed, freq * p);
 * one lambertian material, no lights. */
int rtw_scene_icosphere(uint32_t level, uint64_t seed, float amp, float freq, uint32_t threads, rtw_scene* out);

/* Clustered soup: ceil(n/1024) Gaussian clusters (centre U[-4,4]^3, sigma
 * 0.15) of equilateral triangles with edge U[0.005,0.03] and random
 * orientation; materials drawn 60/25/10/5 % lambertian/metal/glass/light. */
int rtw_scene_clusters(uint32_t n, uint64_t seed, uint32_t threads, rtw_scene* out);

/* The reference's own scene (src/ray-tracing/RtScene.h:44-101) from the OBJ
 * files in `dir` (the reference's resources/models/doge_scene): its six
 * materials, the eight models concatenated in its order with its material
 * indices, its light rule, the glass sphere (0.6, 1, -1, r 0.6) and the
 * reference builder's BVH (rt_bvh_build_seeded, seed 1). */
int rtw_scene_doge(const char* dir, uint32_t threads, rtw_scene* out);

void rtw_scene_free(rtw_scene* s);

/* Camera rays of the shader's model (origin camPos.zxy*(-1,1,1), vfov, unit
 * directions, tmin 0.001, tmax 10000) for a w x h image, emitted tile by
 * tile (tile x tile pixels, tiles in row order), pixels row-major inside a
 * tile; rays [first, first+count) of that order are written. */
int rtw_rays_camera_tiles(const float cam_pos[3], float vfov_deg, uint32_t w, uint32_t h, uint32_t tile,
                          uint32_t first, uint32_t count, rt_ray* out);

/* Incoherent rays: origin = radius * normalize(gaussian^3), direction =
 * normalize(U[-half,half]^3 - origin). */
int rtw_rays_shell(uint32_t n, uint64_t seed, float radius, float half, rt_ray* out);

/* Uniform rays: origin U[-extent,extent]^3, direction uniform on the sphere. */
int rtw_rays_box(uint32_t n, uint64_t seed, float extent, rt_ray* out);

/* Poisson request stream against a server (the served workload): arrivals
 * at `rate_per_s` for `duration_s`; each request is a closest-hit query of
 * log-uniform size in [query_min, query_max] rays (origins U[-4,4]^3,
 * uniform directions) or, with probability render_fraction, a render_w x
 * render_h one-sample render with render_bounces bounces, of a uniformly
 * chosen resident scene.  Every swap_interval_s a new clustered-soup scene
 * of swap_triangles is added from a helper thread (the server evicts the
 * least recently used scene when its budget is exceeded).  Waits for every
 * request, then returns the server's statistics. */
typedef struct rtw_serve_params {
    double rate_per_s, duration_s, render_fraction, swap_interval_s;
    uint32_t query_min, query_max;
    uint32_t render_w, render_h, render_bounces;
    uint32_t swap_triangles;
    uint64_t seed;
    uint32_t threads;            /* host threads for scene generation */
    uint32_t reserved;
} rtw_serve_params;

typedef struct rtw_serve_result {
    rt_server_stats stats;
    double wall_s;
    uint64_t requests, queries, renders, swaps;
} rtw_serve_result;

int rtw_serve_poisson(rt_server* server, const rtw_serve_params* params, rtw_serve_result* result);

/* The same run, keeping some requests' results for checking.  A query's
 * rays are the first `count` of the run's ray pool, rtw_rays_box(query_max,
 * seed ^ 0x5eed, 4.0); a render is eye (0, 0, 10) (camPos (0, 10, 0)),
 * vfov 60, render_bounces bounces, currentSample 0.  Arrival k is kept when
 * k % every == 0 while entries are left; its result goes to the entry's own
 * buffers (hits: query_max records, image: render_w * render_h * 4 bytes,
 * both the caller's), not to the run's shared ring.  The k-th scene the
 * swapper adds is rtw_scene_clusters(swap_triangles, seed + 1000 + k) and
 * its server id is written to swap_ids[k]. */
typedef struct rtw_serve_sample {
    rt_hit* hits;                /* in: caller's buffer for a kept query      */
    uint8_t* image;              /* in: caller's buffer for a kept render     */
    uint64_t scene_id;           /* out: the server scene it ran on           */
    uint64_t arrival;            /* out: k                                    */
    uint32_t kind;               /* out: 0 = query, 1 = render                */
    uint32_t count;              /* out: query rays                           */
} rtw_serve_sample;

typedef struct rtw_serve_sampling {
    uint32_t every;              /* keep arrivals k with k % every == 0 (>= 1)  */
    uint32_t capacity;           /* entries in samples                         */
    rtw_serve_sample* samples;
    uint32_t taken;              /* out: entries filled                        */
    uint32_t swaps_capacity;     /* entries in swap_ids                        */
    uint64_t* swap_ids;
    uint32_t swaps_taken;        /* out                                        */
    uint32_t reserved;
} rtw_serve_sampling;

int rtw_serve_poisson_sampled(rt_server* server, const rtw_serve_params* params, rtw_serve_sampling* sampling,
                              rtw_serve_result* result);

/* Log-uniform integer in [lo, hi]: floor(exp(U * ln(hi/lo)) * lo), U from
 * the seed's first draw (used for scene and batch sizes). */
uint32_t rtw_log_uniform(uint64_t seed, uint32_t lo, uint32_t hi);

#ifdef __cplusplus
}
#endif
#endif /* RTAMD_WORKLOADS_H */
