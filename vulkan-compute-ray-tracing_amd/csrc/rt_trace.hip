// rt_trace.hip -- closest-hit traversal kernels for gfx950 (wave64).
//
// Two kernels compute hit_bvh (ray-trace-compute.comp:263-311):
//
//  * k_trace_packed: ordered traversal of the repacked tree (64-byte nodes
//    carrying both children's boxes, 48-byte triangles with the edge vectors
//    and normal precomputed).  Children are tested against the live interval
//    [tmin, closest], the nearer is entered and the farther deferred on a
//    per-lane stack in LDS.  The triangle test repeats triIntersect /
//    hit_triangle operation for operation, so t, u, v and the accept
//    decision are bit-identical to the shader's; the box test is only a
//    conservative filter (trees whose boxes enclose their subtrees, checked
//    at upload), so the closest hit equals the shader's up to exact ties.
//
//  * k_trace_reference: the shader's own visiting order -- a line test with
//    six divisions per box (intersectAABB, :250-258), both children pushed
//    unconditionally, right popped first, optional MAX_STACK_DEPTH guard.
//    Used for trees that do not enclose their subtrees, for the 16-entry
//    truncation mode, and as a tie-exact mode.
#include <rocprim/device/device_radix_sort.hpp>

#include "rt_trace_dev.hpp"

#pragma clang fp contract(off)

namespace rtamd {

namespace {

// Closest-hit queries (rt_closest_hit, rt_closest_hit_stats).
template <int BLOCK, int STACK, int LAYOUT, bool STATS, int RING>
__global__ __launch_bounds__(BLOCK) void k_trace_packed(TraceScene sc, TraceIo io, int32_t* spill) {
    __shared__ int32_t stk[(RING ? RING : STACK) * BLOCK];
    trace_persistent<BLOCK, STACK, LAYOUT, STATS, false, RING>(sc, io, BatchIo{}, stk, spill);
}

// The render's per-bounce traces over the live-ray queue (count on device).
template <int BLOCK, int STACK, int LAYOUT, int RING>
__global__ __launch_bounds__(BLOCK) void k_trace_render(TraceScene sc, TraceIo io, int32_t* spill) {
    __shared__ int32_t stk[(RING ? RING : STACK) * BLOCK];
    trace_persistent<BLOCK, STACK, LAYOUT, false, false, RING>(sc, io, BatchIo{}, stk, spill);
}

// intersectAABB (ray-trace-compute.comp:250-258) with GLSL min/max.
__device__ __forceinline__ void aabb_line(const rt_bvh_node& nd, const f3& o, const f3& d, float& tn, float& tf) {
    f3 tMin = (ld3(nd.min) - o) / d;
    f3 tMax = (ld3(nd.max) - o) / d;
    f3 t1 = gmin(tMin, tMax);
    f3 t2 = gmax(tMin, tMax);
    tn = gmax(gmax(t1.x, t1.y), t1.z);
    tf = gmin(gmin(t2.x, t2.y), t2.z);
}

// hit_triangle on the original record: identical arithmetic to tri_test.
__device__ __forceinline__ bool tri_test_ref(const rt_triangle& tr, const f3& o, const f3& d, float tmin, float tmax,
                                             float& t, float& u, float& v) {
    f3 v0 = ld3(tr.v0), v1 = ld3(tr.v1), v2 = ld3(tr.v2);
    f3 a = v0 - v1;
    f3 b = v2 - v0;
    f3 p = v0 - o;
    f3 n = cross(b, a);
    f3 q = cross(p, d);
    float idet = 1.0f / dot(d, n);
    u = dot(q, b) * idet;
    v = dot(q, a) * idet;
    t = dot(n, p) * idet;
    bool inside = !(u < 0.0f || u > 1.0f || v < 0.0f || (u + v) > 1.0f);
    return inside && t > tmin && t < tmax;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_trace_reference(
    const rt_bvh_node* __restrict__ nodes, const rt_triangle* __restrict__ tris, const rt_ray* __restrict__ rays,
    uint32_t n_host, const uint32_t* __restrict__ n_dev, rt_hit* __restrict__ hits, int32_t* __restrict__ stack,
    uint32_t cap, uint32_t max_stack) {
    const uint32_t n = n_dev ? *n_dev : n_host;
    const uint32_t stride = gridDim.x * BLOCK;
    const uint32_t g = blockIdx.x * BLOCK + threadIdx.x;
    for (uint32_t i = g; i < n; i += stride) {
        RayIn r = load_ray(rays, i);
        float closest = r.tmax, bu = 0.0f, bv = 0.0f;
        int32_t best = -1;
        uint32_t sp = 0;
        stack[(size_t)sp * stride + g] = 0;
        sp++;
        while (sp > 0 && (max_stack == 0u || sp < max_stack)) {
            sp--;
            int32_t cur = stack[(size_t)sp * stride + g];
            if (cur == -1) continue;
            const rt_bvh_node nd = nodes[cur];
            float tn, tf;
            aabb_line(nd, r.o, r.d, tn, tf);
            if (tn > tf) continue;
            if (nd.objectIndex != -1) {
                float t, u, v;
                if (tri_test_ref(tris[nd.objectIndex], r.o, r.d, r.tmin, closest, t, u, v)) {
                    closest = t; bu = u; bv = v; best = nd.objectIndex;
                }
            }
            if (sp + 2 > cap) break;  // unreachable: cap covers the tree height
            stack[(size_t)sp * stride + g] = nd.leftNodeIndex;
            sp++;
            stack[(size_t)sp * stride + g] = nd.rightNodeIndex;
            sp++;
        }
        if (best >= 0) store_hit(hits, i, best, closest, bu, bv);
        else store_hit(hits, i, -1, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = value;
}

constexpr int kBlock = 256;

// Trace order of a query: rays sorted by the Morton code of their origin's
// cell in a 256^3 grid over the scene's node bounds (the 16-bit frame), so
// the lanes of a wave start near each other and fetch the same upper nodes.
// Only the order changes: each ray is still traced alone and its hit lands
// at its own index.
constexpr uint32_t kOrderBits = 8;                 // cells per axis: 2^8
constexpr uint32_t kQueueOrderBits = 7;            // render queues: 2^7 per axis, three radix passes
constexpr uint32_t kOrderMinRays = 1u << 17;
constexpr uint32_t kOrderMinTriangles = 1u << 20;

__device__ __forceinline__ uint32_t spread10(uint32_t x) {   // bit b -> bit 3b, b < 10
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ __forceinline__ uint32_t order_cell(float o, float org, float rcp) {
    // (o - org) * rcp spans [2, 65002] over the bounds; 2^8 cells
    const float q = (o - org) * rcp * (256.0f / 65000.0f);
    return q >= 255.0f ? 255u : (q > 0.0f ? (uint32_t)q : 0u);
}

__global__ __launch_bounds__(256) void k_ray_keys(const rt_ray* __restrict__ rays, uint32_t n, QFrame f,
                                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 o = reinterpret_cast<const float4*>(rays)[2 * (size_t)i];
        keys[i] = (spread10(order_cell(o.x, f.ox, f.rx)) << 2) | (spread10(order_cell(o.y, f.oy, f.ry)) << 1) |
                  spread10(order_cell(o.z, f.oz, f.rz));
        idx[i] = i;
    }
}

// The same keys for a render bounce's queue, whose length lives on the
// device: slots at or past *n_dev get key 2^24, above every cell key, so the
// live rays sort first.
__global__ __launch_bounds__(256) void k_queue_keys(const rt_ray* __restrict__ rays, const uint32_t* __restrict__ n_dev,
                                                    uint32_t cap, QFrame f, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ idx) {
    const uint32_t n = *n_dev;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) {
        uint32_t k = 1u << (3 * kQueueOrderBits);
        if (i < n) {
            const float4 o = reinterpret_cast<const float4*>(rays)[2 * (size_t)i];
            constexpr uint32_t sh = kOrderBits - kQueueOrderBits;
            k = (spread10(order_cell(o.x, f.ox, f.rx) >> sh) << 2) |
                (spread10(order_cell(o.y, f.oy, f.ry) >> sh) << 1) | spread10(order_cell(o.z, f.oz, f.rz) >> sh);
        }
        keys[i] = k;
        idx[i] = i;
    }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

// Ray cursors for persistent launches: a ring of counters so launches on
// different streams never share one.  A launch zeroes its block with a fill
// kernel on its own stream right before use, unless the caller had an
// earlier kernel on that stream do it (TraceArgs::cursor_zeroed).
uint32_t* next_cursor(rt_ctx* ctx) {
    uint32_t k = ctx->cursor_next++ % kCursorBlocks;
    return ctx->d_cursors + (size_t)k * kCursorBlock;
}

namespace {

template <int STACK, int LAYOUT, int RING>
int launch_packed_t(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, bool stats) {
    const rt_scene* sc = a.scene;
    // the counting variant also serves a render whose record totals are asked for
    auto kern = stats      ? k_trace_packed<kBlock, STACK, LAYOUT, true, RING>
                : a.render ? k_trace_render<kBlock, STACK, LAYOUT, RING>
                           : k_trace_packed<kBlock, STACK, LAYOUT, false, RING>;
    int per_cu = 0;
    RT_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0));
    if (per_cu < 1) per_cu = 1;
    if (ctx->tuning.blocks_per_cu && (int)ctx->tuning.blocks_per_cu < per_cu) per_cu = (int)ctx->tuning.blocks_per_cu;
    uint32_t blocks = (uint32_t)ctx->num_cus * (uint32_t)per_cu;
    if (!a.n_dev) {
        // no more waves than rays
        uint32_t need = (a.n + kBlock - 1) / kBlock;
        if (need < blocks) blocks = need;
        if (blocks == 0) return RT_OK;
    }
    int32_t* spill = nullptr;
    if (RING) {
        if (a.no_alloc)
            return set_error(ctx, RT_ERR_UNSUPPORTED, "an LDS stack ring spills to a workspace: set stack_lds to 0");
        int rc = ensure_workspace(ctx, kWsSpill, (size_t)(STACK - RING) * blocks * kBlock * sizeof(int32_t));
        if (rc) return rc;
        spill = (int32_t*)ctx->ws[kWsSpill];
    }
    uint32_t* cursor = a.cursor ? a.cursor : next_cursor(ctx);
    if (!(a.cursor && a.cursor_zeroed)) {
        const int rc = launch_fill_u32(ctx, cursor, 0u, kCursorBlock, s);
        if (rc) return rc;
    }
    // segment the ray range per XCD when there is enough of it
    const uint32_t segs = (a.n_dev || a.n >= (1u << 16)) ? kSegments : 1u;
    // 1..64 idle lanes trigger a refetch (64: only when the whole wave is
    // idle).  The loop is bound by VALU issue, with each step's dependent
    // record fetch close behind, so idle lanes waste issue slots and every
    // refill stalls its wave on three dependent fetches: a large scene's long
    // traversals want lanes refilled early, while a small scene's short ones
    // finish so fast that each refill's fetches are better shared by more
    // rays.  Measured from 2^17 triangles up, with the drain's
    // work sharing taking over the lanes idle at the end: a query is best at
    // 28 (4M-triangle query 6.06-6.08 ms against 6.10-6.11 at 24, 6.12-6.14
    // at 32, 6.27 at 20), a render's bounces at 24 (1080p render 32.7 ms
    // against 32.8 at 28 and 33.4 at 32).  Below 2^17 triangles 48 is best
    // (-1..4% against 32).
    uint32_t refill = ctx->tuning.refill ? ctx->tuning.refill
                                         : (sc->info.num_triangles >= (1u << 17) ? (a.render ? 24u : 28u) : 48u);
    if (refill > 64u) refill = 64u;
    TraceScene ts{sc->d_pnodes, sc->d_qnodes, sc->d_w4nodes, sc->d_ptris, sc->d_ptri_id, sc->frame,
                  LAYOUT == kLayoutW4 || LAYOUT == kLayoutW4Joined ? sc->root4 : sc->root_ref, sc->tri_off};
    TraceIo io{a.rays, a.n_dev, a.hits, a.node_visits, a.tri_tests, cursor, a.n, refill, a.order, segs,
               drain_steal(ctx), a.sums};
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, ts, io, spill);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

template <int STACK>
int launch_packed_layout(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, bool stats) {
    if (ctx->tuning.node_layout == RT_LAYOUT_F32) return launch_packed_t<STACK, kLayoutF32, 0>(ctx, a, s, stats);
    return launch_packed_t<STACK, kLayoutQ16, 0>(ctx, a, s, stats);
}

template <int STACK>
int launch_w4(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, bool stats) {
    const uint32_t ring = ctx->tuning.stack_lds;
    if (ring == 8 && STACK > 8) return launch_packed_t<STACK, kLayoutW4, 8>(ctx, a, s, stats);
    if (ring == 32 && STACK > 32) return launch_packed_t<STACK, kLayoutW4, 32>(ctx, a, s, stats);
    if (ring == 16 && STACK > 16) return launch_packed_t<STACK, kLayoutW4, 16>(ctx, a, s, stats);
    if (a.scene->tri_off != kSeparateArrays) return launch_packed_t<STACK, kLayoutW4Joined, 0>(ctx, a, s, stats);
    return launch_packed_t<STACK, kLayoutW4, 0>(ctx, a, s, stats);
}

}  // namespace

// Idle lanes before a drained wave hands subtrees to them: a hand-over step
// costs about a traversal step, so it waits for many takers.  Measured on the
// 4M-triangle query (ms per launch): 8 6.42, 16 6.28, 24 6.22, 32 6.22,
// 40 6.31, 48 6.44, never 7.63.
uint32_t drain_steal(const rt_ctx* ctx) { return ctx->tuning.steal ? ctx->tuning.steal : 32u; }

namespace {

// Whether a packed query is traced in sorted order, and the sort's key bits.
bool order_wanted(const rt_ctx* ctx, const TraceArgs& a) {
    const uint32_t mode = ctx->tuning.ray_order;
    if (mode == RT_ORDER_GIVEN || !a.scene->fast) return false;
    // by default only where it pays: many rays, and a tree far larger than
    // the caches (a small scene's nodes stay in L2 in any order).  For a
    // render queue `n` is its capacity.
    return !(mode == RT_ORDER_AUTO && (a.n < kOrderMinRays || a.scene->info.num_triangles < kOrderMinTriangles));
}

// a render queue's keys carry a top bit for the dead slots past its live
// count; 7 bits per axis keep them at 22 bits, three 8-bit radix passes
uint32_t order_key_bits(bool queue) { return queue ? 3u * kQueueOrderBits + 1u : 3u * kOrderBits; }

// Workspace of a sort of n rays: keys and ray indices, each double
// buffered, then the radix sort's temporary storage.
int order_bytes(rt_ctx* ctx, uint32_t n, bool queue, size_t* vec, size_t* sort_bytes) {
    *vec = align256(sizeof(uint32_t) * (size_t)n);
    rocprim::double_buffer<uint32_t> k(nullptr, nullptr), v(nullptr, nullptr);
    *sort_bytes = 0;
    RT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, *sort_bytes, k, v, n, 0u, order_key_bits(queue), (hipStream_t)0));
    return RT_OK;
}

}  // namespace

int ray_order_bytes(rt_ctx* ctx, uint32_t n, size_t* bytes) {
    size_t vec = 0, sort_bytes = 0;
    const int rc = order_bytes(ctx, n, false, &vec, &sort_bytes);
    if (rc) return rc;
    *bytes = 4 * vec + align256(sort_bytes);
    return RT_OK;
}

int ray_order_in(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, char* ws, size_t ws_bytes, const uint32_t** order) {
    *order = nullptr;
    if (!order_wanted(ctx, a)) return RT_OK;
    const uint32_t n = a.n;
    const bool queue = a.n_dev != nullptr;
    size_t vec = 0, sort_bytes = 0;
    int rc = order_bytes(ctx, n, queue, &vec, &sort_bytes);
    if (rc) return rc;
    if (!ws) {
        rc = ensure_workspace(ctx, kWsOrder, 4 * vec + align256(sort_bytes));
        if (rc) return rc;
        ws = (char*)ctx->ws[kWsOrder];
    } else if (4 * vec + align256(sort_bytes) > ws_bytes) {
        return set_error(ctx, RT_ERR_UNSUPPORTED, "ray order: the sort of " + std::to_string(n) + " rays needs " +
                                                      std::to_string(4 * vec + align256(sort_bytes)) +
                                                      " bytes of workspace, more than the " +
                                                      std::to_string(ws_bytes) + " given");
    }
    uint32_t* keys = (uint32_t*)ws;
    uint32_t* keys2 = (uint32_t*)(ws + vec);
    uint32_t* idx = (uint32_t*)(ws + 2 * vec);
    uint32_t* idx2 = (uint32_t*)(ws + 3 * vec);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n + 255u) / 256u, (uint64_t)ctx->num_cus * 16u);
    if (queue)
        hipLaunchKernelGGL(k_queue_keys, dim3(grid), dim3(256), 0, s, a.rays, a.n_dev, n, a.scene->frame, keys, idx);
    else
        hipLaunchKernelGGL(k_ray_keys, dim3(grid), dim3(256), 0, s, a.rays, n, a.scene->frame, keys, idx);
    RT_HIP(ctx, hipGetLastError());
    // double buffers: the sort ping-pongs between each pair, and the sorted
    // ray indices end in whichever buffer holds the last pass
    rocprim::double_buffer<uint32_t> kb(keys, keys2), vb(idx, idx2);
    RT_HIP(ctx, rocprim::radix_sort_pairs((void*)(ws + 4 * vec), sort_bytes, kb, vb, n, 0u, order_key_bits(queue), s));
    *order = vb.current();
    return RT_OK;
}

// The sorted order of a query's rays in ctx's order workspace, or null when
// the batch keeps index order.
int ray_order(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, const uint32_t** order) {
    return ray_order_in(ctx, a, s, nullptr, 0, order);
}

int resolved_layout(const rt_ctx* ctx) {
    uint32_t l = ctx->tuning.node_layout;
    return l == RT_LAYOUT_AUTO ? (int)RT_LAYOUT_W4 : (int)l;
}

int launch_trace_packed(rt_ctx* ctx, const TraceArgs& a, hipStream_t s) {
    const rt_scene* sc = a.scene;
    if (!sc->fast) return set_error(ctx, RT_ERR_UNSUPPORTED, "packed traversal needs a tree that encloses its triangles");
    const bool stats = (a.node_visits && a.tri_tests) || a.sums;
    if (resolved_layout(ctx) == RT_LAYOUT_W4) {
        // up to three deferred children per 4-wide node on the path
        const uint32_t need = 3u * sc->height4 + 1u;
        if (need <= 24) return launch_w4<24>(ctx, a, s, stats);
        if (need <= 36) return launch_w4<36>(ctx, a, s, stats);
        if (need <= 48) return launch_w4<48>(ctx, a, s, stats);
        if (need <= 64) return launch_w4<64>(ctx, a, s, stats);
        if (need <= 96) return launch_w4<96>(ctx, a, s, stats);
        return set_error(ctx, RT_ERR_UNSUPPORTED, "4-wide tree deeper than 31 levels");
    }
    // binary records are built on first use (an upload under the 4-wide
    // layout leaves them out); the scene's arrays are immutable otherwise
    const uint32_t layout = (uint32_t)resolved_layout(ctx);
    if ((layout == RT_LAYOUT_F32 && !sc->d_pnodes) || (layout == RT_LAYOUT_Q16 && !sc->d_qnodes)) {
        if (a.no_alloc)
            return set_error(ctx, RT_ERR_UNSUPPORTED, "the scene's binary node records are not built yet");
        const int rc = scene_binary_records(ctx, const_cast<rt_scene*>(sc), layout);
        if (rc) return rc;
    }
    const uint32_t h = sc->packed_height;
    if (h <= 16) return launch_packed_layout<16>(ctx, a, s, stats);
    if (h <= 20) return launch_packed_layout<20>(ctx, a, s, stats);
    if (h <= 22) return launch_packed_layout<22>(ctx, a, s, stats);
    if (h <= 24) return launch_packed_layout<24>(ctx, a, s, stats);
    if (h <= 28) return launch_packed_layout<28>(ctx, a, s, stats);
    if (h <= 32) return launch_packed_layout<32>(ctx, a, s, stats);
    if (h <= 48) return launch_packed_layout<48>(ctx, a, s, stats);
    if (h <= 64) return launch_packed_layout<64>(ctx, a, s, stats);
    return set_error(ctx, RT_ERR_UNSUPPORTED, "packed tree deeper than 64 levels");
}

int launch_trace_reference(rt_ctx* ctx, const TraceArgs& a, uint32_t max_stack, hipStream_t s) {
    const rt_scene* sc = a.scene;
    uint32_t blocks = (uint32_t)ctx->num_cus * 8u;
    if (!a.n_dev) {
        uint32_t need = (a.n + kBlock - 1) / kBlock;
        if (need < blocks) blocks = need;
        if (blocks == 0) return RT_OK;
    }
    // Line-test DFS holds at most (height + 1) entries; the guard variant
    // never exceeds max_stack.
    uint32_t cap = max_stack ? max_stack : sc->ref_height + 3u;
    size_t bytes = (size_t)cap * blocks * kBlock * sizeof(int32_t);
    int rc = ensure_workspace(ctx, kWsStack, bytes);
    if (rc) return rc;
    hipLaunchKernelGGL((k_trace_reference<kBlock>), dim3(blocks), dim3(kBlock), 0, s, sc->d_nodes, sc->d_tris, a.rays,
                       a.n, a.n_dev, a.hits, (int32_t*)ctx->ws[kWsStack], cap, max_stack);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

int launch_fill_u32(rt_ctx* ctx, uint32_t* p, uint32_t value, uint32_t count, hipStream_t s) {
    if (!count) return RT_OK;
    uint32_t blocks = (count + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, p, value, count);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

}  // namespace rtamd
