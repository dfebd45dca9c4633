/*
 * rtamd.h -- C-ABI of the MI355X ray-traversal library (librtamd.so).
 *
 * This is the drop-in boundary for the compute ray tracer's hot path: the
 * scene arrays, flat BVH node array and RGBA8 target/accumulation images that
 * the reference binds to descriptor set 0 and dispatches with
 *     ComputeModel::computeCommand  (src/scene/ComputeModel.cpp:21-25)
 * over the shader resources/shaders/source/ray-trace-compute.comp.
 *
 * Conventions
 *   - Every entry point returns an int status (RT_OK == 0).  On failure the
 *     message is kept per context and read with rt_last_error().  No C++
 *     exception crosses this boundary.
 *   - Record structs below are byte-identical to the reference's std430
 *     records (src/ray-tracing/GpuModels.h:26-63, definitions.glsl:1-29).
 *   - Pointers named *_dev are device (HBM) pointers; everything else is host
 *     memory.  Streams are hipStream_t passed as void* (NULL = the context's
 *     own stream).  A context is externally synchronised.
 *   - All device work on one context must be ordered on ONE stream: a
 *     context's workspaces (the sorted trace order, render path state,
 *     traversal stacks) are shared by its launches, and calls return before
 *     their kernels finish.  Two launches of one context on different
 *     streams may overlap on the device and overwrite each other's
 *     workspaces.  Work that must overlap uses one context per stream (the
 *     server gives each of its slots its own); scenes can be shared by
 *     contexts on the same device.
 */
#ifndef RTAMD_H
#define RTAMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
    RT_OK = 0,
    RT_ERR_INVALID = 1,     /* bad argument or malformed scene             */
    RT_ERR_HIP = 2,         /* a HIP runtime call failed                    */
    RT_ERR_NOMEM = 3,       /* allocation failed or HBM budget exceeded     */
    RT_ERR_UNSUPPORTED = 4, /* valid request this build cannot serve        */
    RT_ERR_BUSY = 5,        /* server queue full / slot unavailable         */
    RT_ERR_CANCELLED = 6,   /* request cancelled or deadline expired        */
    RT_ERR_NOT_FOUND = 7    /* unknown scene / request id                   */
};

/* Material type ids: definitions.glsl:37-40, GpuModels.h:18-24. */
enum { RT_MAT_LIGHT = 0, RT_MAT_LAMBERTIAN = 1, RT_MAT_METAL = 2, RT_MAT_GLASS = 3 };

/* ----------------------------------------------------------------- records */
/* GpuModel::Triangle (GpuModels.h:32-38): vec3 slots at 0/16/32, index @44. */
typedef struct rt_triangle {
    float v0[3]; float _pad0;
    float v1[3]; float _pad1;
    float v2[3]; uint32_t materialIndex;
} rt_triangle;                                   /* 48 bytes */

/* GpuModel::Material (GpuModels.h:26-30): type @0, albedo @16. */
typedef struct rt_material {
    uint32_t type; uint32_t _pad0[3];
    float albedo[3]; float _pad1;
} rt_material;                                   /* 32 bytes */

/* GpuModel::BvhNode (GpuModels.h:47-54): min @0, max @16, left @28,
 * right @32, objectIndex @36; -1 marks "none". */
typedef struct rt_bvh_node {
    float min[3]; float _pad0;
    float max[3]; int32_t leftNodeIndex;
    int32_t rightNodeIndex; int32_t objectIndex; int32_t _pad1[2];
} rt_bvh_node;                                   /* 48 bytes */

/* GpuModel::Light (GpuModels.h:57-63). */
typedef struct rt_light {
    uint32_t triangleIndex; float area;
} rt_light;                                      /* 8 bytes */

/* GpuModel::Sphere (GpuModels.h:40-44): (cx,cy,cz,r) @0, index @16. */
typedef struct rt_sphere {
    float s[4]; uint32_t materialIndex; uint32_t _pad[3];
} rt_sphere;                                     /* 32 bytes */

/* UniformBufferObject (main.cpp:39-47 == ray-trace-compute.comp:8-15). */
typedef struct rt_ubo {
    float camPos[3]; float time;
    uint32_t currentSample; uint32_t numTriangles; uint32_t numLights; uint32_t numSpheres;
} rt_ubo;                                        /* 32 bytes */

/* Query ray: origin, tmin; direction, tmax.  The shader's hit_bvh uses
 * tmin = 0.001, tmax = 10000 (ray-trace-compute.comp:264-265). */
typedef struct rt_ray {
    float o[3]; float tmin;
    float d[3]; float tmax;
} rt_ray;                                        /* 32 bytes */

/* Closest hit: id = triangle index (or num_triangles + sphere index when
 * spheres are enabled), -1 = miss.  t, u, v as triIntersect returns them
 * (ray-trace-compute.comp:157-173): u weights v1, v weights v2.  A sphere
 * hit reports u = backFaceInt (1 when the far root was taken), v = 0. */
typedef struct rt_hit {
    int32_t id; float t; float u; float v;
} rt_hit;                                        /* 16 bytes */

/* ------------------------------------------------------------ scene upload */
typedef struct rt_scene_desc {
    const rt_triangle* triangles; uint32_t num_triangles;
    const rt_material* materials; uint32_t num_materials;
    const rt_bvh_node* nodes;     uint32_t num_nodes;
    const rt_light*    lights;    uint32_t num_lights;
    const rt_sphere*   spheres;   uint32_t num_spheres;
} rt_scene_desc;

/* Facts the library derived from a scene at upload. */
typedef struct rt_scene_info {
    uint32_t num_triangles, num_nodes, num_lights, num_spheres, num_materials;
    uint32_t tree_height;     /* max leaf depth below the root (root = 0)     */
    uint32_t reachable_leaves;
    uint32_t enclosing;       /* 1: every box holds its subtree (fast path ok) */
    uint32_t fast_path;       /* 1: packed ordered traversal is used           */
    uint64_t device_bytes;    /* HBM held by the resident copy                 */
} rt_scene_info;

typedef struct rt_ctx rt_ctx;
typedef struct rt_scene rt_scene;

/* ---------------------------------------------------------------- context */
/* Replaces VulkanApplicationContext device/allocator set-up
 * (src/app-context/VulkanApplicationContext.cpp:47-93). */
int rt_create(rt_ctx** out, int device);
int rt_destroy(rt_ctx* ctx);
const char* rt_last_error(const rt_ctx* ctx);
const char* rt_status_string(int status);

/* Copies the arrays into HBM, validates indices and the tree, and repacks
 * nodes/triangles for traversal.  Replaces the six BufferUtils::createBundle
 * uploads (main.cpp:84-106, Buffer.h:83-102) and ComputeMaterial::init
 * (ComputeMaterial.cpp:15-26).  Counts are whole-array counts (the shader's
 * .length()), not the one-element descriptor range of Buffer.h:85.  An empty
 * scene (no triangles or no nodes) is rejected with RT_ERR_INVALID. */
int rt_scene_upload(rt_ctx* ctx, const rt_scene_desc* desc, rt_scene** out);
int rt_scene_info_get(const rt_scene* scene, rt_scene_info* info);
/* The validation rt_scene_upload performs, without a device: fills the
 * tree facts of *info (device_bytes = 0) and, on RT_ERR_INVALID, copies the
 * reason into err[0..errlen). */
int rt_scene_check(const rt_scene_desc* desc, rt_scene_info* info, char* err, size_t errlen);
int rt_scene_release(rt_ctx* ctx, rt_scene* scene);

/* -------------------------------------------------------- closest-hit query */
enum {
    RT_TRAVERSE_AUTO = 0,       /* packed ordered traversal when the tree
                                   encloses its triangles, else reference    */
    RT_TRAVERSE_PACKED = 1,     /* force packed ordered traversal            */
    RT_TRAVERSE_REFERENCE = 2   /* hit_bvh's own visiting order and line test */
};
/* Sphere rules.  The shader's main path ignores spheres (hit_bvh only);
 * its unused hit_scene (ray-trace-compute.comp:222-247) tests them after
 * the triangles with closest_so_far reset to t_max (comp:236), so a sphere
 * hit replaces a nearer triangle hit.  Both rules run as a pass after the
 * triangle traversal; scenes with more than 16 spheres walk a sphere tree
 * built at upload (the reference's open "include spheres in bvh") that
 * returns exactly what the loop over all spheres returns: the smallest
 * accepted root, the highest index between equal roots.  A sphere hit
 * reports id = num_triangles + sphere index, u = backFaceInt, v = 0. */
enum {
    RT_SPHERES_OFF = 0,         /* triangles only (the shader's hit_bvh)       */
    RT_SPHERES_NEAREST = 1,     /* nearest of triangles and spheres            */
    RT_SPHERES_SCENE = 2        /* hit_scene's rule: a sphere hit overrides    */
};
typedef struct rt_query_params {
    uint32_t traversal;   /* RT_TRAVERSE_*                                       */
    uint32_t max_stack;   /* reference mode only: 0 = unlimited, 16 = shader's
                             MAX_STACK_DEPTH truncation (ray-trace-compute.comp:262) */
    uint32_t spheres;     /* RT_SPHERES_*                                        */
    uint32_t reserved;
} rt_query_params;

/* hit_bvh (ray-trace-compute.comp:263-311) over n rays: rays_dev[n] ->
 * hits_dev[n], n < 2^31.  Enqueued on `stream`; returns without waiting. */
int rt_closest_hit(rt_ctx* ctx, const rt_scene* scene, const rt_ray* rays_dev,
                   uint32_t n, rt_hit* hits_dev, const rt_query_params* params,
                   void* stream);

/* rt_closest_hit without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under stream
 * capture, serving loops and harnesses that forbid hipMalloc).  Its
 * scratch (ray cursors, the trace-order sort) is one static device array
 * per device of about 81 MiB, held until the process ends.  Limits: n <=
 * 4,194,304 (RT_ERR_INVALID above); the packed traversal only, whole stack
 * in LDS (stack_lds 0), binary layouts only once their records exist
 * (RT_ERR_UNSUPPORTED otherwise).  Calls on one device share the scratch:
 * calls that overlap on one device race, so order them on one stream.
 * Hits equal rt_closest_hit's byte for byte. */
int rt_closest_hit_noalloc(rt_ctx* ctx, const rt_scene* scene, const rt_ray* rays_dev,
                           uint32_t n, rt_hit* hits_dev, const rt_query_params* params,
                           void* stream);

/* Per-ray traversal counters of the packed kernel (node visits = 64-byte
 * node records fetched, tri tests = 48-byte triangle records fetched). */
int rt_closest_hit_stats(rt_ctx* ctx, const rt_scene* scene, const rt_ray* rays_dev,
                         uint32_t n, rt_hit* hits_dev, uint32_t* node_visits_dev,
                         uint32_t* tri_tests_dev, void* stream);

/* ------------------------------------------------------------------ tuning */
/* Launch parameters of the packed traversal (no counterpart in the
 * reference, whose only knob is the compile-time MAX_STACK_DEPTH).  Zero
 * fields keep the library default.  They never change results, except
 * `tree`, which can change which of two hits at equal t is reported. */
enum { RT_LAYOUT_AUTO = 0, RT_LAYOUT_F32 = 1, RT_LAYOUT_Q16 = 2, RT_LAYOUT_W4 = 3 };
/* Tree of the packed traversal, read by rt_scene_upload:
 * RT_TREE_GIVEN   the caller's nodes (e.g. Bvh::createBvh's)
 * RT_TREE_DEVICE  a linear BVH built on the device at upload
 *                 (rt_bvh_build_device); the shader-order traversal keeps the
 *                 caller's nodes, and a scene given without nodes uses the
 *                 device tree for both */
enum { RT_TREE_GIVEN = 0, RT_TREE_DEVICE = 1 };
/* Trace order of rt_closest_hit's rays (results always land at each ray's
 * own index): RT_ORDER_AUTO sorts batches of 2^17 rays or more by the
 * Morton code of their origins in the scene's box (8 bits per axis), so the
 * lanes of a wave start near each other and share node fetches;
 * RT_ORDER_GIVEN traces them in index order; RT_ORDER_SORTED always sorts. */
enum { RT_ORDER_AUTO = 0, RT_ORDER_GIVEN = 1, RT_ORDER_SORTED = 2 };
typedef struct rt_tuning {
    uint32_t node_layout;   /* RT_LAYOUT_*: binary 64-byte fp32 or 32-byte 16-bit
                               nodes, or 4-wide 64-byte 16-bit nodes            */
    uint32_t refill;        /* idle lanes (1..64) that make a wave fetch rays;
                               0 = by scene: from 2^17 triangles up (long
                               traversals) 28 for queries and 24 for render
                               bounces, 48 below (short ones)                    */
    uint32_t blocks_per_cu; /* persistent blocks per CU (0 = occupancy limit)     */
    uint32_t stack_lds;     /* 4-wide layout: 0 = whole traversal stack in LDS;
                               8, 16 or 32 = that many top entries in LDS, deeper
                               ones spill to HBM (more waves fit per CU)          */
    uint32_t tree;          /* RT_TREE_*                                          */
    uint32_t ray_order;     /* RT_ORDER_*: the order a query's rays are traced in */
    uint32_t steal;         /* 4-wide layout, whole stack in LDS: once a wave's
                               rays are used up, this many idle lanes (1..64)
                               make it hand subtrees from its busy lanes' stacks
                               to the idle ones (hits unchanged); 0 = 32,
                               RT_STEAL_NEVER = never                            */
    uint32_t reserved;
} rt_tuning;
enum { RT_STEAL_NEVER = 65 };
int rt_tuning_set(rt_ctx* ctx, const rt_tuning* t);
int rt_tuning_get(const rt_ctx* ctx, rt_tuning* t);

/* ----------------------------------------------------------------- batched */
/* Many (scene, ray batch) items traced by ONE persistent launch with a
 * global work queue, optionally regrouping each item's rays first (by
 * origin cell, or by direction octant and origin cell), replayed from a
 * hipGraph.  Each item gives
 * the same hits as rt_closest_hit on it alone (triangles only; the scenes
 * must take the packed path).  No counterpart in the reference, which
 * dispatches one image of one scene per frame (main.cpp:228). */
typedef struct rt_batch_item {
    const rt_scene* scene;
    const rt_ray* rays_dev;
    rt_hit* hits_dev;
    uint32_t count;
    uint32_t reserved;
} rt_batch_item;

typedef struct rt_batch_config {
    uint32_t block;          /* threads per work-group: 128 or 256 (0 = 256)       */
    uint32_t stack_lds;      /* 0 = whole stack in LDS; 8, 16 or 32 = top entries
                                (4-wide records only)                             */
    uint32_t regroup;        /* 0 = rays in the caller's order; 1 = bin each item's
                                rays by origin cell (16^3, Morton order); 2 = by
                                direction octant, then origin cell (8^3)          */
    uint32_t refill;         /* idle lanes (1..64) before a wave fetches rays       */
    uint32_t blocks_per_cu;  /* 0 = occupancy limit                                 */
    uint32_t use_graph;      /* 1 = replay the launch sequence from a hipGraph      */
    uint32_t regroup_interval; /* rays binned together: 0 = each whole item, or a
                                power of two >= 65536 = windows of that many
                                consecutive rays of an item                       */
    uint32_t node_layout;    /* node records read: RT_LAYOUT_W4 (or 0), 4-wide
                                64-byte records; RT_LAYOUT_Q16, binary 32-byte
                                records (built for the scenes on first use)       */
} rt_batch_config;

typedef struct rt_batch rt_batch;

/* Builds the device item table and work queue for `n` items (their ray and
 * hit buffers are the caller's) and captures the launch sequence.  NULL
 * config = defaults (256, 0, 0, 32, 0, 1, 0, RT_LAYOUT_W4). */
int rt_batch_create(rt_ctx* ctx, const rt_batch_item* items, uint32_t n, const rt_batch_config* config,
                    rt_batch** out);
int rt_batch_set_config(rt_ctx* ctx, rt_batch* batch, const rt_batch_config* config);
int rt_batch_get_config(const rt_batch* batch, rt_batch_config* config);
/* Enqueues the whole batch on `stream`. */
int rt_batch_launch(rt_ctx* ctx, rt_batch* batch, void* stream);
/* Times candidate configurations on this batch (coordinate descent, two
 * passes, over block, node_layout, stack_lds, regroup, regroup_interval,
 * refill and blocks_per_cu; `trials` timed launches each), applies the
 * fastest and returns it.  With a cache file, a stored
 * result for the same shape (item count, log2 rays, log2 triangles) is reused
 * and new results are appended. */
int rt_batch_autotune(rt_ctx* ctx, rt_batch* batch, uint32_t trials, const char* cache_path,
                      rt_batch_config* best);
int rt_batch_destroy(rt_ctx* ctx, rt_batch* batch);

/* ------------------------------------------------------------------ render */
typedef struct rt_render_params {
    float    vfov_deg;    /* shader hard-codes 30 (ray-trace-compute.comp:358) */
    uint32_t bounces;     /* shader NUM_BOUNCES = 2 (ray-trace-compute.comp:313) */
    uint32_t max_stack;   /* 0 = unlimited; 16 = reference truncation (reference
                             traversal only)                                    */
    uint32_t spheres;     /* RT_SPHERES_*; 0 = the shader's path (hit_bvh)      */
    uint32_t traversal;   /* RT_TRAVERSE_*                                      */
    uint32_t reserved[3];
} rt_render_params;

/* One frame of main() (ray-trace-compute.comp:352-380): reads accum_dev
 * (RGBA8, pitch w*4) and writes target_dev (RGBA8).  The whole w x h image
 * is rendered (no H/32 truncation of main.cpp:228).  ubo->currentSample
 * seeds the RNG and weights the progressive average. */
int rt_render(rt_ctx* ctx, const rt_scene* scene, const rt_ubo* ubo,
              const rt_render_params* params, uint8_t* target_dev,
              const uint8_t* accum_dev, uint32_t w, uint32_t h, void* stream);

/* `frames` consecutive frames, currentSample = ubo->currentSample + k, with
 * the target->accumulation copy of main.cpp:253-261 between frames.  All
 * frames' paths are traced together; the per-pixel 8-bit running average is
 * applied in frame order, so the result equals `frames` calls of rt_render
 * each followed by a copy target->accum.  On return both images hold the
 * last frame. */
int rt_render_frames(rt_ctx* ctx, const rt_scene* scene, const rt_ubo* ubo,
                     const rt_render_params* params, uint32_t frames,
                     uint8_t* target_dev, uint8_t* accum_dev, uint32_t w,
                     uint32_t h, void* stream);

/* Traversal count of the last rt_render / rt_render_frames on this context
 * (one count per hit_bvh call, i.e. per ray). */
int rt_render_last_rays(const rt_ctx* ctx, uint64_t* primary, uint64_t* secondary);

/* Stage timing of renders and queries (no counterpart in the reference,
 * whose frame loop has no timers).  With timing on, each render records HIP
 * events around every bounce's traversal kernel; rt_render_last_stats waits
 * for the last render and splits its time into the bounce-0 traversal, the
 * traversals of bounces 1.. (all secondary, incoherent rays) and the rest
 * (generate, shade, store).  Each rt_closest_hit records events around its
 * ordering (key kernel and sort), its traversal (cursor reset and the
 * traversal kernel) and its sphere pass; rt_query_last_stats waits for the
 * last query and returns those times.  Without timing the *_ms fields are 0
 * and `timed` is 0. */
typedef struct rt_render_stats {
    uint64_t primary_rays, secondary_rays;
    uint32_t bounces, timed;
    float total_ms, trace_primary_ms, trace_secondary_ms, shade_ms;
} rt_render_stats;
int rt_render_set_timing(rt_ctx* ctx, int enable);
int rt_render_last_stats(rt_ctx* ctx, rt_render_stats* out);
/* Record counting of renders (no counterpart in the reference): while on,
 * a render's packed traversals run the counting variant (no drain work
 * sharing, images unchanged) and rt_render_last_counts returns the node
 * and triangle records its rays read in total -- the render's algorithmic
 * bytes.  Waits for the render. */
int rt_render_set_counting(rt_ctx* ctx, int enable);
int rt_render_last_counts(rt_ctx* ctx, uint64_t* node_records, uint64_t* tri_records);
typedef struct rt_query_stats {
    uint32_t timed, sorted;   /* sorted: the rays were traced in origin order */
    float order_ms, trace_ms, spheres_ms, total_ms;
} rt_query_stats;
int rt_query_last_stats(rt_ctx* ctx, rt_query_stats* out);

/* -------------------------------------------------- frame loop and display */
/* The reference's main loop around the compute pass (src/main.cpp:166-183
 * updateScene, :228-261 the dispatch and the target->accumulation copy):
 * rt_frame_loop_step resets currentSample to 0 when `moved`, fills the UBO
 * (camera, time, currentSample, the scene's counts), renders into the
 * loop's target reading its accumulation image, copies target to
 * accumulation and increments currentSample.  The loop owns three device
 * RGBA8 images of w*h*4 bytes: target, accumulation and display. */
typedef struct rt_frame_loop rt_frame_loop;
int rt_frame_loop_create(rt_ctx* ctx, const rt_scene* scene, const rt_render_params* params, uint32_t w,
                         uint32_t h, rt_frame_loop** out);
int rt_frame_loop_destroy(rt_frame_loop* loop);
int rt_frame_loop_step(rt_frame_loop* loop, const float cam_pos[3], float time, int moved, void* stream);
/* The display pass of post-process-shader.frag:62-68 applied to the target:
 * display = pow(target, 1/2.2) per colour channel, alpha 255. */
int rt_frame_loop_display(rt_frame_loop* loop, void* stream);
/* Device pointers of the loop's images and the next frame's currentSample. */
int rt_frame_loop_images(const rt_frame_loop* loop, uint8_t** target_dev, uint8_t** accum_dev,
                         uint8_t** display_dev, uint32_t* current_sample);
/* Copies image `which` (0 target, 1 accumulation, 2 display) to host memory
 * of w*h*4 bytes after the work queued on `stream`; returns when done. */
int rt_frame_loop_read(const rt_frame_loop* loop, int which, uint8_t* host, void* stream);
/* The display pass on any RGBA8 image (device memory, 4-byte aligned):
 * out = rne(pow(fl(c/255), fl(1/2.2)) * 255) per colour channel with pow
 * correctly rounded to binary32, alpha 255.  The shader samples the target
 * through a linear filter at texel centres, which returns the texel. */
int rt_postprocess_gamma(rt_ctx* ctx, const uint8_t* src_dev, uint8_t* dst_dev, uint32_t w, uint32_t h,
                         void* stream);
/* The 256-entry channel map rt_postprocess_gamma applies. */
int rt_gamma_table(uint8_t out[256]);
/* Writes a host RGBA8 image (rows of w*4 bytes) to `path`: ".png" (RGBA,
 * stored deflate blocks) or ".ppm" (binary P6, alpha dropped); other
 * extensions give RT_ERR_UNSUPPORTED.  Not in the reference, which
 * presents through the swapchain. */
int rt_image_write(const char* path, const uint8_t* rgba, uint32_t w, uint32_t h);

/* ------------------------------------------------------------------ served */
/* A request loop on its own thread: scenes stay resident under an HBM
 * budget (least recently used scenes without pending work are evicted to
 * admit a new one), closest-hit and render requests queue on arrival and
 * are launched as device slots free up.  In RT_SERVE_BATCHED mode each slot
 * launch traces every queued query (up to max_batch_rays) in one batched
 * launch, and up to `slots` launches overlap on the device; RT_SERVE_FIFO
 * runs one request at a time with the plain kernel.  Request buffers are
 * host memory owned by the caller until the request completes.  Latency is
 * measured from rt_server_submit_* to the result being in the caller's
 * buffer.  No counterpart in the reference. */
enum { RT_SERVE_BATCHED = 0, RT_SERVE_FIFO = 1 };
typedef struct rt_server_config {
    uint32_t mode;            /* RT_SERVE_*                                     */
    uint32_t slots;           /* concurrent device launches (1..8, 0 = 4)       */
    uint32_t max_batch_rays;  /* rays per batched launch (0 = 1,048,576; <= 2^30) */
    uint32_t max_render_pixels; /* largest render request (0 = 1,048,576; <= 2^30) */
    uint64_t hbm_budget;      /* bytes for resident scenes (0 = no limit)       */
    uint32_t deadline_ms;     /* a request still queued this long after its
                                 submission is dropped with RT_ERR_CANCELLED
                                 (0 = no deadline)                              */
    uint32_t reserved[3];
} rt_server_config;

typedef struct rt_server_stats {
    uint64_t submitted, completed, failed, launches;
    uint64_t scenes_added, scenes_evicted, scenes_refused;
    uint64_t rays;                 /* traversals done (query rays + render rays) */
    uint64_t resident_bytes;
    double p50_ms, p99_ms, mean_ms, max_ms;   /* over completed requests        */
    uint64_t peak_resident_bytes;  /* most scene bytes ever allocated at once    */
    uint64_t cancelled;            /* dropped by rt_server_cancel or a deadline  */
} rt_server_stats;

typedef struct rt_server rt_server;
/* Page-locked host memory for request buffers: rays, hits and targets in
 * such memory are copied by DMA straight to and from the device; pageable
 * buffers go through the slot's staging copy. */
int rt_host_alloc(size_t bytes, void** out);
int rt_host_free(void* p);
int rt_server_create(rt_ctx* ctx, const rt_server_config* config, rt_server** out);
int rt_server_destroy(rt_server* server);
/* Makes a scene resident.  Admission comes before any device allocation:
 * the scene is validated and packed on the host, which fixes its device
 * bytes; least recently used idle scenes are then evicted (and freed) while
 * the budget, or the device's free memory, cannot hold it, and only then is
 * it uploaded -- resident bytes never exceed the budget, not even during a
 * swap.  RT_ERR_NOMEM when it cannot fit. */
int rt_server_scene_add(rt_server* server, const rt_scene_desc* desc, uint64_t* scene_id);
int rt_server_scene_remove(rt_server* server, uint64_t scene_id);
/* Writes up to `cap` resident scene ids; *count receives the total. */
int rt_server_scene_list(rt_server* server, uint64_t* ids, uint32_t cap, uint32_t* count);
int rt_server_submit_query(rt_server* server, uint64_t scene_id, const rt_ray* rays, uint32_t n, rt_hit* hits,
                           uint64_t* request_id);
/* One frame (rt_render) into `target` (w*h*4 bytes, RGBA8) from a zero
 * accumulation image. */
int rt_server_submit_render(rt_server* server, uint64_t scene_id, const rt_ubo* ubo,
                            const rt_render_params* params, uint32_t w, uint32_t h, uint8_t* target,
                            uint64_t* request_id);
/* Waits for a request (timeout_ms < 0: forever); *status gets its result:
 * RT_OK, RT_ERR_CANCELLED (cancelled, or dropped at its deadline), or the
 * launch's error. */
int rt_server_wait(rt_server* server, uint64_t request_id, int timeout_ms, int* status);
/* Cancels a request that is still queued: it completes at once with
 * RT_ERR_CANCELLED and never runs.  RT_ERR_BUSY when it is already on the
 * device (it then completes normally), RT_ERR_NOT_FOUND when it is unknown
 * or finished.  The reference has no requests to cancel: a failure there
 * ends the process (main.cpp:449-457). */
int rt_server_cancel(rt_server* server, uint64_t request_id);
/* Waits until every submitted request has completed. */
int rt_server_drain(rt_server* server);
int rt_server_stats_get(rt_server* server, rt_server_stats* stats);
int rt_server_stats_reset(rt_server* server);
const char* rt_server_last_error(const rt_server* server);

/* ---------------------------------------------------------------- OBJ input */
/* Triangles of one Wavefront OBJ model, as GpuModel::getTriangles makes
 * them (src/ray-tracing/RtScene.h:13-30 over Mesh::Mesh, src/scene/mesh.cpp
 * :96-139): triangle k is the positions of face k's corners in file order
 * (the reference's vertex de-duplication changes only its index buffer),
 * each with `material_index`.  Faces with k > 3 corners become the fan
 * (0, i, i+1); corner indices are 1-based or negative (relative); numbers
 * are parsed as double and rounded to float.  *out is malloc'ed (release
 * with rt_free); RT_ERR_NOT_FOUND for a missing file, RT_ERR_INVALID for a
 * malformed one (rt_obj_last_error gives the line, per thread). */
int rt_obj_load(const char* path, uint32_t material_index, rt_triangle** out, uint32_t* count);
const char* rt_obj_last_error(void);
void rt_free(void* p);

/* --------------------------------------------------------------- host BVH */
/* Bvh::createBvh (src/ray-tracing/Bvh.h:141-209): 2n-1 nodes into out[].
 * Consumes the C library's rand() exactly as the reference does (one call
 * per node, default seed unless the caller seeded it).  Use
 * rt_bvh_build_seeded for a builder independent of the process's rand(). */
int rt_bvh_build(const rt_triangle* tris, uint32_t n, rt_bvh_node* out);
/* Same tree, with the axis stream of glibc random() seeded by `seed`
 * (seed 1 == a fresh process), run on `threads` host threads. */
int rt_bvh_build_seeded(const rt_triangle* tris, uint32_t n, uint32_t seed,
                        uint32_t threads, rt_bvh_node* out);
/* GPU builder (no counterpart in the reference, whose builder runs on the
 * host, Bvh.h:141-209): a linear BVH (Karras, HPG 2012) over the box
 * centres of the device triangles -- 63-bit Morton codes over the centres'
 * bounds, a stable radix sort, a radix tree whose equal codes are ordered
 * by sorted position -- written to nodes_dev in the reference's node
 * layout: 2n-1 nodes, root at node 0, internal node k at index k, the leaf
 * of sorted position k at index n-1+k, leaf boxes the reference's
 * triangle boxes (bounds widened by 0.0001, Bvh.h:96-97), parent boxes the
 * union of their children's; codes use the centres of those boxes.  The tree differs
 * from Bvh::createBvh's; every closest hit is the same.  Both pointers are
 * device memory, 16-byte aligned; runs on `stream`; n <= 2^30. */
int rt_bvh_build_device(rt_ctx* ctx, const rt_triangle* tris_dev, uint32_t n, rt_bvh_node* nodes_dev, void* stream);
/* Light list rule of RtScene.h:87-96: for each triangle whose material is
 * a light, Light{i, 0.5*|v0 x v1|}.  Writes up to `cap` lights and returns
 * the count in *count. */
int rt_scene_lights(const rt_triangle* tris, uint32_t n, const rt_material* mats,
                    uint32_t num_mats, rt_light* out, uint32_t cap, uint32_t* count);

#ifdef __cplusplus
}
#endif
#endif /* RTAMD_H */
