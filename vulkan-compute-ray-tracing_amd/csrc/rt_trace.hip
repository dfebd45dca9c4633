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
#include "rt_common.hpp"
#include "rt_math.hpp"

#pragma clang fp contract(off)

namespace rtamd {

namespace {

struct RayIn {
    f3 o, d;
    float tmin, tmax;
};

__device__ __forceinline__ RayIn load_ray(const rt_ray* rays, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(rays) + 2 * (size_t)i;
    float4 a = p[0], b = p[1];
    return RayIn{f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, a.w, b.w};
}

__device__ __forceinline__ void store_hit(rt_hit* hits, uint32_t i, int32_t id, float t, float u, float v) {
    float4 h = make_float4(__int_as_float(id), t, u, v);
    reinterpret_cast<float4*>(hits)[i] = h;
}

// Conservative slab test of one box against [tmin, tmax].  Rounding of the
// reciprocal form is absorbed by a relative slack of 2^-16 (far larger than
// the few-ulp error), so a box that holds an accepted hit is never culled.
__device__ __forceinline__ bool slab(float lox, float loy, float loz, float hix, float hiy, float hiz,
                                     const f3& o, const f3& inv, float tmin, float tmax, float& tnear) {
    float tx0 = (lox - o.x) * inv.x, tx1 = (hix - o.x) * inv.x;
    float ty0 = (loy - o.y) * inv.y, ty1 = (hiy - o.y) * inv.y;
    float tz0 = (loz - o.z) * inv.z, tz1 = (hiz - o.z) * inv.z;
    float nearv = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float farv = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    tnear = nearv;
    return nearv <= farv + (fabsf(farv) + fabsf(nearv)) * 0x1p-16f;
}

// triIntersect + hit_triangle's accept test on a packed triangle.
__device__ __forceinline__ bool tri_test(const PTri& tr, const f3& o, const f3& d, float tmin, float tbest,
                                         float& t, float& u, float& v) {
    f3 v0{tr.a.x, tr.a.y, tr.a.z};
    f3 a{tr.a.w, tr.b.x, tr.b.y};
    f3 b{tr.b.z, tr.b.w, tr.c.x};
    f3 n{tr.c.y, tr.c.z, tr.c.w};
    f3 p = v0 - o;
    f3 q = cross(p, d);
    float idet = 1.0f / dot(d, n);
    u = dot(q, b) * idet;
    v = dot(q, a) * idet;
    t = dot(n, p) * idet;
    bool inside = !(u < 0.0f || u > 1.0f || v < 0.0f || (u + v) > 1.0f);
    return inside && t > tmin && t < tbest;
}

// hit_sphere (ray-trace-compute.comp:175-203): returns the accepted root
// and whether it is the far one (backFaceInt).
__device__ __forceinline__ bool sphere_test(const rt_sphere& s, const f3& o, const f3& d, float tmin, float tmax,
                                            float& root, int& back) {
    f3 center{s.s[0], s.s[1], s.s[2]};
    float radius = s.s[3];
    f3 oc = o - center;
    float a = dot(d, d);
    float half_b = dot(oc, d);
    float c = dot(oc, oc) - radius * radius;
    float disc = half_b * half_b - a * c;
    if (disc < 0.0f) return false;
    float sqrtd = sqrtf(disc);
    back = 0;
    root = (-half_b - sqrtd) / a;
    if (root < tmin || tmax < root) {
        root = (-half_b + sqrtd) / a;
        back = 1;
        if (root < tmin || tmax < root) return false;
    }
    return true;
}

// Reciprocal for the box tests: a zero (or tiny) direction component
// becomes +-2^-80 so every slab bound stays finite (an infinite bound would
// make the slack term infinite and keep every box).  The ray this describes
// departs from the true one by < 1e-20 over t <= 1e4, far inside the boxes'
// padding, so the test stays conservative.
__device__ __forceinline__ float safe_inv(float d) {
    const float tiny = 0x1p-80f;
    float a = fabsf(d) < tiny ? copysignf(tiny, d) : d;
    return 1.0f / a;
}

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Scene and launch arguments of the persistent traversal, by value.
struct TraceScene {
    const PNode* nodes;       // kLayoutF32
    const PNodeQ* qnodes;     // kLayoutQ16
    const PNode4* w4nodes;    // kLayoutW4
    const PTri* tris;
    const int32_t* tri_id;
    const rt_sphere* spheres;
    QFrame frame;
    int32_t root;
    uint32_t nspheres;
    uint32_t ntri;
};
struct TraceIo {
    const rt_ray* rays;
    const uint32_t* n_dev;
    rt_hit* hits;
    uint32_t* node_visits;
    uint32_t* tri_tests;
    uint32_t* cursor;
    uint32_t n_host;
    uint32_t refill;
};

__device__ __forceinline__ float lo16(uint32_t w) { return (float)(w & 0xffffu); }

// Compare-exchange of (key, value) pairs: afterwards ka <= kb.
__device__ __forceinline__ void cswap(float& ka, int32_t& va, float& kb, int32_t& vb) {
    bool s = kb < ka;
    float k = s ? kb : ka;
    kb = s ? ka : kb;
    ka = k;
    int32_t v = s ? vb : va;
    vb = s ? va : vb;
    va = v;
}
__device__ __forceinline__ float hi16(uint32_t w) { return (float)(w >> 16); }

// Persistent traversal with per-lane ray fetch.  Each lane owns one ray at a
// time; when at least `refill` lanes of the wave are idle, the wave takes
// that many ray indices from the global cursor with one atomic and hands
// them out by ballot prefix count, so a long ray never holds its wave's
// other lanes idle.  One loop iteration reads one node record or one
// triangle record per busy lane.  With 16-bit nodes the ray is moved once
// into the nodes' fixed-point frame (origin (o - org) / scale, reciprocal
// direction times scale): slab distances t are the same numbers, so tmin,
// tmax and the closest t apply unchanged.
template <int BLOCK, int STACK, int LAYOUT, bool STATS>
__device__ __forceinline__ void trace_persistent(const TraceScene& sc, const TraceIo& io, int32_t* __restrict__ stk) {
    const uint32_t n = io.n_dev ? *io.n_dev : io.n_host;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint64_t lanes_below = (1ull << lane) - 1ull;

    bool busy = false;
    bool drained = false;           // wave-uniform: the cursor passed n
    uint32_t ray_i = 0;
    RayIn r{};
    f3 bo{0.0f, 0.0f, 0.0f};        // ray origin in the box frame
    f3 inv{0.0f, 0.0f, 0.0f};       // reciprocal direction in the box frame
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;
    int32_t best = -1;
    uint32_t nv = 0, nt = 0;
    int sp = 0;
    int32_t cur = 0;

    for (;;) {
        // ---- refill idle lanes
        const uint64_t idle = __ballot(!busy);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!drained && n_idle >= io.refill) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(idle);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(io.cursor, n_idle);
            base = __shfl(base, (int)leader);
            if (base + n_idle >= n) drained = true;
            if (!busy) {
                uint32_t mine = base + (uint32_t)__popcll(idle & lanes_below);
                if (mine < n) {
                    ray_i = mine;
                    r = load_ray(io.rays, mine);
                    inv = f3{safe_inv(r.d.x), safe_inv(r.d.y), safe_inv(r.d.z)};
                    bo = r.o;
                    if (LAYOUT == kLayoutQ16 || LAYOUT == kLayoutW4) {
                        const QFrame& q = sc.frame;
                        bo = f3{(r.o.x - q.ox) * q.rx, (r.o.y - q.oy) * q.ry, (r.o.z - q.oz) * q.rz};
                        inv = f3{inv.x * q.sx, inv.y * q.sy, inv.z * q.sz};
                    }
                    best_t = r.tmax; bu = 0.0f; bv = 0.0f; best = -1;
                    nv = 0; nt = 0; sp = 0;
                    cur = sc.root;
                    busy = true;
                    if (sc.root == kEmptyRoot) cur = 0, sp = -1;   // finishes on the first step
                }
            }
        }
        if (__ballot(busy) == 0) {
            if (drained) break;
            continue;
        }
        // ---- one traversal step per busy lane
        if (busy) {
            bool pop = true;
            if (sp >= 0) {
                if (cur >= 0 && LAYOUT == kLayoutW4) {
                    const uint4* np = reinterpret_cast<const uint4*>(sc.w4nodes + cur);
                    uint4 a = np[0], b = np[1], c = np[2];
                    int4 d = reinterpret_cast<const int4*>(np)[3];
                    if (STATS) nv++;
                    float k0 = __builtin_inff(), k1 = __builtin_inff(), k2 = __builtin_inff(), k3 = __builtin_inff();
                    float tn;
                    int nh = 0;
                    if (d.x != kNoChild && slab(lo16(a.x), hi16(a.x), lo16(a.y), hi16(a.y), lo16(a.z), hi16(a.z), bo, inv,
                                                r.tmin, best_t, tn)) { k0 = tn; nh++; }
                    if (d.y != kNoChild && slab(lo16(a.w), hi16(a.w), lo16(b.x), hi16(b.x), lo16(b.y), hi16(b.y), bo, inv,
                                                r.tmin, best_t, tn)) { k1 = tn; nh++; }
                    if (d.z != kNoChild && slab(lo16(b.z), hi16(b.z), lo16(b.w), hi16(b.w), lo16(c.x), hi16(c.x), bo, inv,
                                                r.tmin, best_t, tn)) { k2 = tn; nh++; }
                    if (d.w != kNoChild && slab(lo16(c.y), hi16(c.y), lo16(c.z), hi16(c.z), lo16(c.w), hi16(c.w), bo, inv,
                                                r.tmin, best_t, tn)) { k3 = tn; nh++; }
                    if (nh > 0) {
                        int32_t v0 = d.x, v1 = d.y, v2 = d.z, v3 = d.w;
                        // sort the four (entry distance, child) pairs; misses carry +inf
                        cswap(k0, v0, k1, v1);
                        cswap(k2, v2, k3, v3);
                        cswap(k0, v0, k2, v2);
                        cswap(k1, v1, k3, v3);
                        cswap(k1, v1, k2, v2);
                        // farthest pushed first so the nearest deferred child pops first
                        if (nh > 3) { stk[sp * BLOCK + tid] = v3; sp++; }
                        if (nh > 2) { stk[sp * BLOCK + tid] = v2; sp++; }
                        if (nh > 1) { stk[sp * BLOCK + tid] = v1; sp++; }
                        cur = v0;
                        pop = false;
                    }
                } else if (cur >= 0) {
                    bool h0, h1;
                    float tn0, tn1;
                    int32_t c0, c1;
                    if (LAYOUT == kLayoutQ16) {
                        const uint4* np = reinterpret_cast<const uint4*>(sc.qnodes + cur);
                        uint4 a = np[0], b = np[1];
                        h0 = slab(lo16(a.x), hi16(a.x), lo16(a.y), hi16(a.y), lo16(a.z), hi16(a.z), bo, inv,
                                  r.tmin, best_t, tn0);
                        h1 = slab(lo16(a.w), hi16(a.w), lo16(b.x), hi16(b.x), lo16(b.y), hi16(b.y), bo, inv,
                                  r.tmin, best_t, tn1);
                        c0 = (int32_t)b.z;
                        c1 = (int32_t)b.w;
                    } else {
                        const float4* np = reinterpret_cast<const float4*>(sc.nodes + cur);
                        float4 a = np[0], b = np[1], c = np[2];
                        int4 dd = reinterpret_cast<const int4*>(np)[3];
                        h0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, bo, inv, r.tmin, best_t, tn0);
                        h1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, bo, inv, r.tmin, best_t, tn1);
                        c0 = dd.x;
                        c1 = dd.y;
                    }
                    if (STATS) nv++;
                    if (h0 && h1) {
                        bool swap = tn1 < tn0;
                        stk[sp * BLOCK + tid] = swap ? c0 : c1;
                        sp++;
                        cur = swap ? c1 : c0;
                        pop = false;
                    } else if (h0 || h1) {
                        cur = h0 ? c0 : c1;
                        pop = false;
                    }
                } else {
                    const int32_t slot = ~cur;
                    const float4* tp = reinterpret_cast<const float4*>(sc.tris + slot);
                    PTri tr{tp[0], tp[1], tp[2]};
                    if (STATS) nt++;
                    float t, u, v;
                    if (tri_test(tr, r.o, r.d, r.tmin, best_t, t, u, v)) {
                        best_t = t; bu = u; bv = v; best = slot;
                    }
                }
            }
            if (pop) {
                if (sp > 0) {
                    sp--;
                    cur = stk[sp * BLOCK + tid];
                } else {
                    // ---- ray finished: spheres, then the result
                    int32_t id = best >= 0 ? sc.tri_id[best] : -1;
                    if (sc.nspheres) {
                        float closest = best >= 0 ? best_t : r.tmax;
                        for (uint32_t j = 0; j < sc.nspheres; ++j) {
                            float root_t; int back;
                            if (sphere_test(sc.spheres[j], r.o, r.d, r.tmin, closest, root_t, back)) {
                                closest = root_t; best_t = root_t; id = (int32_t)(sc.ntri + j);
                                bu = (float)back; bv = 0.0f;
                            }
                        }
                    }
                    if (id >= 0) store_hit(io.hits, ray_i, id, best_t, bu, bv);
                    else store_hit(io.hits, ray_i, -1, 0.0f, 0.0f, 0.0f);
                    if (STATS) { io.node_visits[ray_i] = nv; io.tri_tests[ray_i] = nt; }
                    busy = false;
                }
            }
        }
    }
}

// Closest-hit queries (rt_closest_hit, rt_closest_hit_stats).
template <int BLOCK, int STACK, int LAYOUT, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_trace_packed(TraceScene sc, TraceIo io) {
    __shared__ int32_t stk[STACK * BLOCK];
    trace_persistent<BLOCK, STACK, LAYOUT, STATS>(sc, io, stk);
}

// The render's per-bounce traces over the live-ray queue (count on device).
template <int BLOCK, int STACK, int LAYOUT>
__global__ __launch_bounds__(BLOCK) void k_trace_render(TraceScene sc, TraceIo io) {
    __shared__ int32_t stk[STACK * BLOCK];
    trace_persistent<BLOCK, STACK, LAYOUT, false>(sc, io, stk);
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
    uint32_t cap, uint32_t max_stack, const rt_sphere* __restrict__ spheres, uint32_t nspheres, uint32_t ntri) {
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
        int32_t id = best;
        float best_t = closest;
        if (nspheres) {
            float cl = best >= 0 ? closest : r.tmax;
            for (uint32_t j = 0; j < nspheres; ++j) {
                float root_t; int back;
                if (sphere_test(spheres[j], r.o, r.d, r.tmin, cl, root_t, back)) {
                    cl = root_t; best_t = root_t; id = (int32_t)(ntri + j); bu = (float)back; bv = 0.0f;
                }
            }
        }
        if (id >= 0) store_hit(hits, i, id, best_t, bu, bv);
        else store_hit(hits, i, -1, 0.0f, 0.0f, 0.0f);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t value, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = value;
}

constexpr int kBlock = 256;

// Ray cursors for persistent launches: a ring of counters so launches on
// different streams never share one; each is zeroed by a fill kernel on the
// launch's own stream right before use.
uint32_t* next_cursor(rt_ctx* ctx) {
    uint32_t k = ctx->cursor_next++ % kCursorSlots;
    return ctx->d_cursors + k;
}

template <int STACK, int LAYOUT>
int launch_packed_t(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, bool stats) {
    const rt_scene* sc = a.scene;
    auto kern = a.render ? k_trace_render<kBlock, STACK, LAYOUT>
                : stats    ? k_trace_packed<kBlock, STACK, LAYOUT, true>
                           : k_trace_packed<kBlock, STACK, LAYOUT, false>;
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
    uint32_t* cursor = next_cursor(ctx);
    int rc = launch_fill_u32(ctx, cursor, 0u, 1u, s);
    if (rc) return rc;
    // 1..64 idle lanes trigger a refetch (64: only when the whole wave is idle)
    uint32_t refill = ctx->tuning.refill ? ctx->tuning.refill : 32u;
    if (refill > 64u) refill = 64u;
    TraceScene ts{sc->d_pnodes, sc->d_qnodes, sc->d_w4nodes, sc->d_ptris, sc->d_ptri_id, sc->d_spheres, sc->frame,
                  LAYOUT == kLayoutW4 ? sc->root4 : sc->root_ref, a.spheres ? sc->info.num_spheres : 0u,
                  sc->info.num_triangles};
    TraceIo io{a.rays, a.n_dev, a.hits, a.node_visits, a.tri_tests, cursor, a.n, refill};
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kBlock), 0, s, ts, io);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

template <int STACK>
int launch_packed_layout(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, bool stats) {
    if (ctx->tuning.node_layout == RT_LAYOUT_F32) return launch_packed_t<STACK, kLayoutF32>(ctx, a, s, stats);
    return launch_packed_t<STACK, kLayoutQ16>(ctx, a, s, stats);
}

}  // namespace

int resolved_layout(const rt_ctx* ctx) {
    uint32_t l = ctx->tuning.node_layout;
    return l == RT_LAYOUT_AUTO ? (int)RT_LAYOUT_W4 : (int)l;
}

int launch_trace_packed(rt_ctx* ctx, const TraceArgs& a, hipStream_t s) {
    const rt_scene* sc = a.scene;
    if (!sc->fast) return set_error(ctx, RT_ERR_UNSUPPORTED, "packed traversal needs a tree that encloses its triangles");
    const bool stats = a.node_visits && a.tri_tests;
    if (resolved_layout(ctx) == RT_LAYOUT_W4) {
        // up to three deferred children per 4-wide node on the path
        const uint32_t need = 3u * sc->height4 + 1u;
        if (need <= 24) return launch_packed_t<24, kLayoutW4>(ctx, a, s, stats);
        if (need <= 36) return launch_packed_t<36, kLayoutW4>(ctx, a, s, stats);
        if (need <= 48) return launch_packed_t<48, kLayoutW4>(ctx, a, s, stats);
        if (need <= 64) return launch_packed_t<64, kLayoutW4>(ctx, a, s, stats);
        if (need <= 96) return launch_packed_t<96, kLayoutW4>(ctx, a, s, stats);
        return set_error(ctx, RT_ERR_UNSUPPORTED, "4-wide tree deeper than 31 levels");
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
                       a.n, a.n_dev, a.hits, (int32_t*)ctx->ws[kWsStack], cap, max_stack, sc->d_spheres,
                       a.spheres ? sc->info.num_spheres : 0u, sc->info.num_triangles);
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
