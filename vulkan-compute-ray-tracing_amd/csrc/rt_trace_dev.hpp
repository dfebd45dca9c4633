// rt_trace_dev.hpp -- device-side traversal shared by the plain query,
// render and batched kernels: ray/hit records, the conservative slab test,
// the exact triangle and sphere tests of the shader, and the persistent
// ordered traversal loop.
#pragma once

#include "rt_common.hpp"
#include "rt_math.hpp"

#pragma clang fp contract(off)

namespace rtamd {

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

// Streaming forms for the persistent traversal: each ray is read once and
// each hit written once, so both carry the non-temporal hint and leave the
// caches to the node and triangle records, which every ray re-reads.
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ RayIn load_ray_stream(const rt_ray* rays, uint32_t i) {
    const v4f* p = reinterpret_cast<const v4f*>(rays) + 2 * (size_t)i;
    const v4f a = __builtin_nontemporal_load(p);
    const v4f b = __builtin_nontemporal_load(p + 1);
    return RayIn{f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, a.w, b.w};
}

__device__ __forceinline__ void store_hit_stream(rt_hit* hits, uint32_t i, int32_t id, float t, float u, float v) {
    const v4f h = {__int_as_float(id), t, u, v};
    __builtin_nontemporal_store(h, reinterpret_cast<v4f*>(hits) + i);
}

// Accept test of a slab interval: near - |near| * 2^-15 <= far, one fused
// multiply-add and a compare.  The computed near and far each carry a few ulp
// (< 2^-21 relative) of rounding; where the exact interval is non-empty and
// rounding inverted it, near and far are within that error of each other, so
// 2^-15 |near| covers both.  Where |far| is much larger than |near| and the
// exact interval is non-empty, far is positive and far above near already.
__device__ __forceinline__ bool slab_accept(float nearv, float farv) {
    return __builtin_fmaf(-fabsf(nearv), 0x1p-15f, nearv) <= farv;
}

// Conservative slab test of one box against [tmin, tmax], so a box that
// holds an accepted hit is never culled (slab_accept).
__device__ __forceinline__ bool slab(float lox, float loy, float loz, float hix, float hiy, float hiz,
                                     const f3& o, const f3& inv, float tmin, float tmax, float& tnear) {
    float tx0 = (lox - o.x) * inv.x, tx1 = (hix - o.x) * inv.x;
    float ty0 = (loy - o.y) * inv.y, ty1 = (hiy - o.y) * inv.y;
    float tz0 = (loz - o.z) * inv.z, tz1 = (hiz - o.z) * inv.z;
    float nearv = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    float farv = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    tnear = nearv;
    return slab_accept(nearv, farv);
}

// Slab test of a 16-bit box in the frame's units with each axis' planes in
// the order the ray meets them: (nx, ny, nz) the planes it enters by,
// (fx, fy, fz) those it leaves by (for inv >= 0 the low plane first, for
// inv < 0 the high one), so no min/max pairs are needed.  Planes arrive
// biased (biased_plane): the float 2^23 + q, met at t = (2^23 + q) * inv + c
// with c = -(2^23 + bo) * inv (plane_bias), one fused multiply-add.  The fma
// is exact before its final rounding, so against the exact (q - bo) * inv
// the only absolute error is the rounding of c, (2^23 + |bo|) * 2^-24
// quanta; frame_far keeps |bo| < 2^21, so the planes move by under 0.63
// quanta, inside the one-quantum padding every quantized bound carries.  The
// remaining errors are relative (inv, the final rounding) and slab_accept
// covers them.
__device__ __forceinline__ bool slab_ordered(float nx, float ny, float nz, float fx, float fy, float fz,
                                             const f3& c, const f3& inv, float tmin, float tmax, float& tnear) {
    const float tnx = __builtin_fmaf(nx, inv.x, c.x), tfx = __builtin_fmaf(fx, inv.x, c.x);
    const float tny = __builtin_fmaf(ny, inv.y, c.y), tfy = __builtin_fmaf(fy, inv.y, c.y);
    const float tnz = __builtin_fmaf(nz, inv.z, c.z), tfz = __builtin_fmaf(fz, inv.z, c.z);
    const float nearv = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
    const float farv = fminf(fminf(tfx, tfy), fminf(tfz, tmax));
    tnear = nearv;
    return slab_accept(nearv, farv);
}

// A 16-bit plane of a 4-wide record word as the float 2^23 + q: one byte
// permute puts the half the selector names under the exponent byte 0x4B
// (2^23, whose unit in the last place is 1), so the float holds q exactly
// with no conversion.  kSelLow takes the word's low half, kSelHigh its high half.
constexpr uint32_t kPlaneExponent = 0x4B000000u;
constexpr uint32_t kSelLow = 0x070C0100u;
constexpr uint32_t kSelHigh = 0x070C0302u;
__device__ __forceinline__ float biased_plane(uint32_t w, uint32_t sel) {
    return __uint_as_float(__builtin_amdgcn_perm(kPlaneExponent, w, sel));
}

// c of a biased plane (slab_ordered): -(2^23 + bo) * inv, rounded once.
__device__ __forceinline__ float plane_bias(double bo, float inv) {
    return (float)(-(0x1p23 + bo) * (double)inv);
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

// A ray origin at least this many quanta outside the 16-bit frame (32
// frame extents) is moved along the ray first (frame_far).
constexpr float kFarOrigin = 2097152.0f;

// Places a ray in a scene's 16-bit frame: a plane at q quanta is met at
// t = t_shift + (q - bo) * inv.  Normally t_shift = 0 and bo is the origin's
// position, (o - org) / scale in float.  The float rounding of bo moves every
// plane by |bo| * 2^-23 quanta, which the one-quantum padding of quantized
// bounds absorbs only while |bo| is small; so an origin at least kFarOrigin
// quanta out is advanced, in double precision, to where the ray meets the
// frame's box, and t_shift records how far.  Returns false when the ray
// misses that box within [tmin, tmax]: it can hit no node box (a tree whose
// root is a single triangle has no node box, and the frame bounds nothing).
// bod receives the advanced origin in double, for plane_bias.
__device__ __forceinline__ bool frame_far(const RayIn& r, const QFrame& q, f3& bo, double* bod, float& t_shift) {
    const double o[3] = {((double)r.o.x - (double)q.ox) / (double)q.sx, ((double)r.o.y - (double)q.oy) / (double)q.sy,
                         ((double)r.o.z - (double)q.oz) / (double)q.sz};
    const double d[3] = {(double)r.d.x / (double)q.sx, (double)r.d.y / (double)q.sy, (double)r.d.z / (double)q.sz};
    double t0 = (double)r.tmin, t1 = (double)r.tmax;
    for (int k = 0; k < 3; ++k) {
        // the box [-1, 65536] holds every quantized bound with room to spare
        if (d[k] == 0.0) {
            if (o[k] < -1.0 || o[k] > 65536.0) return false;
            continue;
        }
        const double ta = (-1.0 - o[k]) / d[k], tb = (65536.0 - o[k]) / d[k];
        t0 = fmax(t0, fmin(ta, tb));
        t1 = fmin(t1, fmax(ta, tb));
    }
    if (!(t0 <= t1)) return false;
    t_shift = (float)t0;
    const double ts = (double)t_shift;
    for (int k = 0; k < 3; ++k) bod[k] = o[k] + ts * d[k];
    bo = f3{(float)bod[0], (float)bod[1], (float)bod[2]};
    return true;
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
    QFrame frame;
    int32_t root;
    uint32_t tri_off;         // kLayoutW4Joined: tris = w4nodes + tri_off bytes
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
    const uint32_t* order;    // trace order (ray indices), or null: index order
    uint32_t segs;            // segments of the ray range (1 or kSegments), see below
    uint32_t steal;           // idle lanes that make a drained wave share work (65: never)
    unsigned long long* sums; // counting variant: totals of node and triangle records, or null
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

// Position of the n-th set bit (from 0) of m; n < popcount(m).  A binary
// search on popcounts of the low halves.
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t n) {
    const uint32_t lo = (uint32_t)m;
    const uint32_t c = (uint32_t)__popc(lo);
    uint32_t w = n < c ? lo : (uint32_t)(m >> 32);
    uint32_t pos = n < c ? 0u : 32u;
    n = n < c ? n : n - c;
    for (uint32_t s = 16; s != 0; s >>= 1) {
        const uint32_t cs = (uint32_t)__popc(w & ((1u << s) - 1u));
        const bool up = n >= cs;
        n = up ? n - cs : n;
        w = up ? w >> s : w;
        pos = up ? pos + s : pos;
    }
    return pos;
}

__device__ __forceinline__ float shfl_f(float x, uint32_t src) {
    return __int_as_float(__shfl(__float_as_int(x), (int)src));
}

// Persistent traversal with per-lane ray fetch.  Each lane owns one ray at a
// time; when at least `refill` lanes of the wave are idle, the wave takes
// that many ray indices from the global cursor with one atomic and hands
// them out by ballot prefix count, so a long ray never holds its wave's
// other lanes idle.  One loop iteration reads one node record or one
// triangle record per busy lane.  With 16-bit nodes the ray is moved once
// into the nodes' fixed-point frame (origin (o - org) / scale, reciprocal
// direction times scale): slab distances t are the same numbers, so tmin,
// tmax and the closest t apply unchanged -- except for a ray from far
// outside the frame, whose node tests run on t - t_shift (frame_far).
//
// BATCHED: rays come from many items (BatchIo); each lane carries its own
// scene pointers.  RING > 0: only the top RING stack entries of a lane live
// in LDS (a ring); deeper entries move to `spill` in HBM, one word per
// push/pop beyond RING.  RING = 0 keeps the whole stack (STACK entries) in LDS.
//
// Work sharing in the drain (4-wide layouts, whole stack in LDS, no
// counters; in batched launches the helper also takes the donor's scene):
// once a wave's rays are used up, its idle lanes would wait for the
// longest rays of the wave.  Instead each idle lane takes the bottom entry of
// a busy lane's stack -- the subtree that lane would visit last -- and
// traverses it for the ray's owner (a helper).  The ordered traversal's
// visiting sequence is cut into pieces: a lane holds the piece
// [plo, plo + 2^pk), and a steal halves the donor's piece, the donor keeping
// the first half (its own remaining work) and the helper the second (the
// stolen subtree, which comes after it).  Each piece accepts hits as the
// whole traversal would (strictly closer than its best so far); a helper
// starts from the donor's closest t, and at its end hands (t, piece, hit) to
// the owner, who keeps the smaller t and on equal t the earlier piece.  The
// result is the whole traversal's: the closest t, first in visiting order.
// The owner writes it once its own piece and every helper have finished.
template <int BLOCK, int STACK, int LAYOUT, bool STATS, bool BATCHED, int RING>
__device__ __forceinline__ void trace_persistent(const TraceScene& sc, const TraceIo& io, const BatchIo& bio,
                                                 int32_t* __restrict__ stk, int32_t* __restrict__ spill) {
    static_assert(!BATCHED || LAYOUT == kLayoutW4 || LAYOUT == kLayoutQ16,
                  "batched launches read 4-wide or binary 16-bit records");
    constexpr bool kW4 = LAYOUT == kLayoutW4 || LAYOUT == kLayoutW4Joined;
    constexpr bool STEAL = kW4 && RING == 0 && !STATS;
    static_assert(RING == 0 || (RING & (RING - 1)) == 0, "ring size is a power of two");
    const uint32_t n = io.n_dev ? *io.n_dev : io.n_host;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = lane_id();
    const uint64_t lanes_below = (1ull << lane) - 1ull;
    const size_t gstride = (size_t)gridDim.x * BLOCK;
    const size_t gtid = (size_t)blockIdx.x * BLOCK + tid;

    bool busy = false;
    bool drained = false;           // wave-uniform: every segment is used up
    // Segments: the ray range (all items' rays, for batched launches) is cut into io.segs
    // contiguous pieces, each with its own cursor (io.cursor[k * kCursorStride]).
    // Blocks are dispatched to the 8 XCDs round robin, so each XCD's blocks
    // start on one run of adjacent segments: each XCD's L2 then serves one
    // region of the (sorted) rays, and the cursors' atomics are spread out.
    // A wave whose segment is used up moves to the next one, and it is
    // drained once it has found all of them used up.  Correctness does not
    // depend on the dispatch order, only locality does.
    // block b runs on XCD b % kXcds; with kSegments segments each XCD starts
    // on its own run of kSegments / kXcds adjacent ones
    uint32_t seg = io.segs == kSegments
                       ? (blockIdx.x % kXcds) * (kSegments / kXcds) + (blockIdx.x / kXcds) % (kSegments / kXcds)
                       : (uint32_t)(blockIdx.x % io.segs);
    uint32_t tried = 0;
    // the segment's ray range [seg_lo, seg_hi); io.segs is 1 or kSegments,
    // both powers of two, so the bounds are a product and a shift
    const uint32_t seg_shift = io.segs == kSegments ? kSegmentShift : 0u;
    auto seg_bound = [&](uint32_t k) -> uint32_t { return (uint32_t)(((uint64_t)n * k) >> seg_shift); };
    uint32_t seg_lo = seg_bound(seg), seg_hi = seg_bound(seg + 1);
    uint32_t ray_i = 0;
    RayIn r{};
    f3 bo{0.0f, 0.0f, 0.0f};        // ray origin in the box frame
    f3 inv{0.0f, 0.0f, 0.0f};       // reciprocal direction in the box frame
    // 4-wide records: the byte-permute selectors of each axis' entry and exit
    // planes (the low half first where the ray runs toward +axis)
    uint32_t sn_x = kSelLow, sn_y = kSelLow, sn_z = kSelLow;
    uint32_t sf_x = kSelHigh, sf_y = kSelHigh, sf_z = kSelHigh;
    auto set_planes = [&](float ix, float iy, float iz) {
        const bool nx = __float_as_uint(ix) >> 31, ny = __float_as_uint(iy) >> 31, nz = __float_as_uint(iz) >> 31;
        sn_x = nx ? kSelHigh : kSelLow;
        sf_x = nx ? kSelLow : kSelHigh;
        sn_y = ny ? kSelHigh : kSelLow;
        sf_y = ny ? kSelLow : kSelHigh;
        sn_z = nz ? kSelHigh : kSelLow;
        sf_z = nz ? kSelLow : kSelHigh;
    };
    f3 cq{0.0f, 0.0f, 0.0f};        // plane_bias: 4-wide box planes are biased_plane * inv + cq
    // node tests run on t - t_shift (frame_far); tmin_rel and best_rel are
    // tmin and the closest hit's t on that scale (equal to them when t_shift = 0)
    float t_shift = 0.0f, tmin_rel = 0.0f, best_rel = 0.0f;
    float best_t = 0.0f, bu = 0.0f, bv = 0.0f;
    int32_t best = -1;
    uint32_t nv = 0, nt = 0;
    uint64_t sum_nv = 0, sum_nt = 0;   // STATS: this lane's totals over its rays
    int sp = 0;
    int32_t cur = 0;
    // work sharing (STEAL): the stack holds entries [bot, sp); the lane's
    // piece of its ray's visiting sequence is [plo, plo + 2^pk)
    int bot = 0;
    uint32_t plo = 0, pk = 31;
    int32_t owner = -1;             // a helper's owner lane; -1 for the owner
    uint32_t pending = 0;           // owner: helpers still running
    uint32_t best_lo = 0;           // the piece that found `best`
    // owner whose own piece is done while helpers still run: not busy (it
    // takes no steps), not idle (it keeps its ray); only in a drained wave
    bool waiting = false;
    bool handoff = false;           // helper: done, result not yet handed over
    // the lane's scene (uniform unless BATCHED)
    const PNode4* l_w4 = sc.w4nodes;
    const PNodeQ* l_q = sc.qnodes;
    const PTri* l_tris = sc.tris;
    const int32_t* l_triid = sc.tri_id;
    rt_hit* l_hits = io.hits;

    auto push = [&](int32_t v) {
        if (RING == 0) {
            stk[sp * BLOCK + tid] = v;
        } else {
            const int slot = sp & (RING - 1);
            if (sp >= RING) spill[(size_t)(sp - RING) * gstride + gtid] = stk[slot * BLOCK + tid];
            stk[slot * BLOCK + tid] = v;
        }
        sp++;
    };
    auto pop = [&]() -> int32_t {
        sp--;
        if (RING == 0) return stk[sp * BLOCK + tid];
        const int slot = sp & (RING - 1);
        int32_t v = stk[slot * BLOCK + tid];
        if (sp >= RING) stk[slot * BLOCK + tid] = spill[(size_t)(sp - RING) * gstride + gtid];
        return v;
    };

    // ---- one traversal step per busy lane; SH: with the drain's work sharing
    auto step = [&](auto sh) {
        constexpr bool SH = decltype(sh)::value;
        if (!busy) return;
        bool next = false;          // cur holds the next record to read
        if (sp >= 0 && kW4) {
            // One memory round trip per step whatever the lanes' mix of
            // nodes and triangles: every lane issues its three 16-byte
            // loads (node words 0-2 or the triangle record) and a fourth
            // (node refs, or the triangle's first word again, so it never
            // reads past the array) before any lane branches.
            const bool node = cur >= 0;
            uint4 a, b, c, dw;
            if constexpr (LAYOUT == kLayoutW4Joined) {
                // one allocation: 32-bit byte offsets from its base, so
                // the loads take the base from scalar registers and no
                // lane computes a 64-bit address
                const char* base = reinterpret_cast<const char*>(sc.w4nodes);
                const uint32_t slot = (uint32_t)(~cur);
                const uint32_t off_node = (uint32_t)cur << 6;
                const uint32_t off_tri = sc.tri_off + (slot << 5) + (slot << 4);
                const uint32_t off = node ? off_node : off_tri;
                const uint32_t off_d = node ? off_node + 48u : off_tri;
                a = *reinterpret_cast<const uint4*>(base + off);
                b = *reinterpret_cast<const uint4*>(base + off + 16u);
                c = *reinterpret_cast<const uint4*>(base + off + 32u);
                dw = *reinterpret_cast<const uint4*>(base + off_d);
            } else {
                const uint4* np = node ? reinterpret_cast<const uint4*>(l_w4 + cur)
                                       : reinterpret_cast<const uint4*>(l_tris + ~cur);
                a = np[0];
                b = np[1];
                c = np[2];
                dw = np[node ? 3 : 0];
            }
            // Only the node branch reads dw, so the compiler would sink its
            // load into that branch, after the triangle branch has already
            // waited for the other three: two round trips for a wave that
            // holds both.  These empty statements use all four records here,
            // before the branch, so the loads are issued together, waited for
            // once, and none is still in flight when the next step reuses
            // their registers (which would otherwise cost a wait for every
            // outstanding memory operation, hit stores included).
            asm volatile("" : "+v"(dw.x), "+v"(dw.y), "+v"(dw.z), "+v"(dw.w));
            asm volatile("" : "+v"(a.x), "+v"(a.y), "+v"(a.z), "+v"(a.w), "+v"(b.x), "+v"(b.y), "+v"(b.z),
                         "+v"(b.w), "+v"(c.x), "+v"(c.y), "+v"(c.z), "+v"(c.w));
            if (node) {
                const int4 d = make_int4((int)dw.x, (int)dw.y, (int)dw.z, (int)dw.w);
                if (STATS) nv++;
                float k0 = __builtin_inff(), k1 = __builtin_inff(), k2 = __builtin_inff(), k3 = __builtin_inff();
                float tn;
                int nh = 0;
                // child k's words: entry and exit planes by the ray's selectors
                auto box = [&](uint32_t wx, uint32_t wy, uint32_t wz, float& key) {
                    if (slab_ordered(biased_plane(wx, sn_x), biased_plane(wy, sn_y), biased_plane(wz, sn_z),
                                     biased_plane(wx, sf_x), biased_plane(wy, sf_y), biased_plane(wz, sf_z), cq,
                                     inv, tmin_rel, best_rel, tn)) {
                        key = tn;
                        nh++;
                    }
                };
                if (d.x != kNoChild) box(a.x, a.y, a.z, k0);
                if (d.y != kNoChild) box(a.w, b.x, b.y, k1);
                if (d.z != kNoChild) box(b.z, b.w, c.x, k2);
                if (d.w != kNoChild) box(c.y, c.z, c.w, k3);
                if (nh > 0) {
                    int32_t v0 = d.x, v1 = d.y, v2 = d.z, v3 = d.w;
                    // sort the four (entry distance, child) pairs; misses carry +inf
                    cswap(k0, v0, k1, v1);
                    cswap(k2, v2, k3, v3);
                    cswap(k0, v0, k2, v2);
                    cswap(k1, v1, k3, v3);
                    cswap(k1, v1, k2, v2);
                    // farthest pushed first so the nearest deferred child pops first
                    if (nh > 3) push(v3);
                    if (nh > 2) push(v2);
                    if (nh > 1) push(v1);
                    cur = v0;
                    next = true;
                }
            } else {
                const int32_t slot = ~cur;
                const PTri tr{make_float4(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(a.z),
                                          __uint_as_float(a.w)),
                              make_float4(__uint_as_float(b.x), __uint_as_float(b.y), __uint_as_float(b.z),
                                          __uint_as_float(b.w)),
                              make_float4(__uint_as_float(c.x), __uint_as_float(c.y), __uint_as_float(c.z),
                                          __uint_as_float(c.w))};
                if (STATS) nt++;
                float t, u, v;
                // a best handed over from a later piece yields to an equal t of this one
                float cap = best_t;
                if (SH && best_lo > plo) cap = __uint_as_float(__float_as_uint(best_t) + 1u);
                if (tri_test(tr, r.o, r.d, r.tmin, cap, t, u, v)) {
                    best_t = t; bu = u; bv = v; best = slot;
                    best_rel = t - t_shift;
                    if (SH) best_lo = plo;
                }
            }
        } else if (sp >= 0) {
            if (cur >= 0) {
                bool h0, h1;
                float tn0, tn1;
                int32_t c0, c1;
                if (LAYOUT == kLayoutQ16) {
                    const uint4* np = reinterpret_cast<const uint4*>(l_q + cur);
                    uint4 a = np[0], b = np[1];
                    h0 = slab(lo16(a.x), hi16(a.x), lo16(a.y), hi16(a.y), lo16(a.z), hi16(a.z), bo, inv,
                              tmin_rel, best_rel, tn0);
                    h1 = slab(lo16(a.w), hi16(a.w), lo16(b.x), hi16(b.x), lo16(b.y), hi16(b.y), bo, inv,
                              tmin_rel, best_rel, tn1);
                    c0 = (int32_t)b.z;
                    c1 = (int32_t)b.w;
                } else {
                    const float4* np = reinterpret_cast<const float4*>(sc.nodes + cur);
                    float4 a = np[0], b = np[1], c = np[2];
                    int4 dd = reinterpret_cast<const int4*>(np)[3];
                    h0 = slab(a.x, a.y, a.z, a.w, b.x, b.y, bo, inv, tmin_rel, best_rel, tn0);
                    h1 = slab(b.z, b.w, c.x, c.y, c.z, c.w, bo, inv, tmin_rel, best_rel, tn1);
                    c0 = dd.x;
                    c1 = dd.y;
                }
                if (STATS) nv++;
                if (h0 && h1) {
                    bool swap = tn1 < tn0;
                    push(swap ? c0 : c1);
                    cur = swap ? c1 : c0;
                    next = true;
                } else if (h0 || h1) {
                    cur = h0 ? c0 : c1;
                    next = true;
                }
            } else {
                const int32_t slot = ~cur;
                const float4* tp = reinterpret_cast<const float4*>(l_tris + slot);
                PTri tr{tp[0], tp[1], tp[2]};
                if (STATS) nt++;
                float t, u, v;
                if (tri_test(tr, r.o, r.d, r.tmin, best_t, t, u, v)) {
                    best_t = t; bu = u; bv = v; best = slot;
                    best_rel = t - t_shift;
                }
            }
        }
        if (!next) {
            if (sp > (SH ? bot : 0)) {
                cur = pop();
            } else if (SH && owner >= 0) {
                handoff = true;             // a helper's piece is done
                busy = false;
            } else if (SH && pending != 0) {
                waiting = true;             // the owner's piece is done
                busy = false;
            } else {
                // ---- ray finished: the result (spheres follow in their own pass)
                const int32_t id = best >= 0 ? l_triid[best] : -1;
                if (id >= 0) store_hit_stream(l_hits, ray_i, id, best_t, bu, bv);
                else store_hit_stream(l_hits, ray_i, -1, 0.0f, 0.0f, 0.0f);
                if (STATS) {
                    if (io.node_visits) { io.node_visits[ray_i] = nv; io.tri_tests[ray_i] = nt; }
                    sum_nv += nv;
                    sum_nt += nt;
                }
                busy = false;
            }
        }
    };

    for (;;) {
        // ---- refill idle lanes
        const uint64_t idle = __ballot(!busy);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        if (!drained && n_idle >= io.refill) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(idle);
            const uint32_t lo = seg_lo, hi = seg_hi;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(io.cursor + (size_t)seg * kCursorStride, n_idle);
            base = __shfl(base, (int)leader);
            if ((uint64_t)lo + base + n_idle >= hi) {     // this segment is used up
                if (++tried >= io.segs) {
                    drained = true;
                    // a drained wave holds the launch's tail: it issues
                    // ahead of the SIMD's waves still fetching rays
                    // (4M-triangle query 6.07 -> 6.04 ms)
                    __builtin_amdgcn_s_setprio(3);
                } else {
                    seg = seg + 1 == io.segs ? 0u : seg + 1;
                    seg_lo = seg_bound(seg);
                    seg_hi = seg_bound(seg + 1);
                }
            }
            if (!busy) {
                const uint64_t m = (uint64_t)lo + base + (uint32_t)__popcll(idle & lanes_below);
                const uint32_t mine = m < hi ? (uint32_t)m : n;
                if (mine < n) {
                    QFrame q = sc.frame;
                    int32_t root = sc.root;
                    if (BATCHED) {
                        uint32_t b = bio.chunk_item[mine >> 6];
                        while (mine >= bio.offsets[b + 1]) ++b;
                        const BatchEntry* e = bio.table + b;
                        if (kW4) l_w4 = e->nodes;
                        else l_q = e->qnodes;
                        l_tris = e->tris;
                        l_triid = e->tri_id;
                        l_hits = e->hits;
                        root = kW4 ? e->root : e->root_q;
                        q = e->frame;
                        ray_i = bio.perm ? bio.perm[mine] : mine - e->base;
                        r = load_ray_stream(e->rays, ray_i);
                    } else {
                        ray_i = io.order ? io.order[mine] : mine;
                        r = load_ray_stream(io.rays, ray_i);
                    }
                    inv = f3{safe_inv(r.d.x), safe_inv(r.d.y), safe_inv(r.d.z)};
                    bo = r.o;
                    t_shift = 0.0f;
                    bool reach = true;
                    if (LAYOUT == kLayoutQ16 || kW4) {
                        bo = f3{(r.o.x - q.ox) * q.rx, (r.o.y - q.oy) * q.ry, (r.o.z - q.oz) * q.rz};
                        inv = f3{inv.x * q.sx, inv.y * q.sy, inv.z * q.sz};
                        // the origin in the frame in double, as the quantizer placed the bounds
                        double bod[3] = {0.0, 0.0, 0.0};
                        if (kW4) {
                            bod[0] = ((double)r.o.x - (double)q.ox) / (double)q.sx;
                            bod[1] = ((double)r.o.y - (double)q.oy) / (double)q.sy;
                            bod[2] = ((double)r.o.z - (double)q.oz) / (double)q.sz;
                        }
                        if (!(fmaxf(fmaxf(fabsf(bo.x), fabsf(bo.y)), fabsf(bo.z)) < kFarOrigin))
                            reach = frame_far(r, q, bo, bod, t_shift);
                        if (kW4) cq = f3{plane_bias(bod[0], inv.x), plane_bias(bod[1], inv.y), plane_bias(bod[2], inv.z)};
                    }
                    set_planes(inv.x, inv.y, inv.z);
                    tmin_rel = r.tmin - t_shift;
                    best_rel = r.tmax - t_shift;
                    best_t = r.tmax; bu = 0.0f; bv = 0.0f; best = -1;
                    nv = 0; nt = 0; sp = 0;
                    bot = 0; plo = 0; pk = 31; owner = -1; pending = 0; best_lo = 0;
                    cur = root;
                    busy = true;
                    // an empty tree, or a ray that misses the frame's box
                    // (which bounds every node box, so only a tree with an
                    // internal root), finishes on the first step
                    if (root == kEmptyRoot || (!reach && root >= 0)) cur = 0, sp = -1;
                }
            }
        }
        // ---- drain (no rays left to fetch): idle lanes take work from busy ones
        if constexpr (STEAL) {
            if (drained) {
                // Idle lanes take the bottom stack entries of busy lanes.  The
                // donors' entries were written by other lanes of this wave in
                // earlier steps: keep every LDS access in order.
                __builtin_amdgcn_wave_barrier();
                // Helpers cull against their owner's closest hit as soon as it
                // improves (the subtrees they hold are the ones a single lane
                // would visit last, often culled by then).  Culling only: the
                // closest hit and its ties lie at or before the owner's t, so
                // no box that holds them is dropped, and acceptance still
                // compares with the helper's own best.
                const float owner_rel = shfl_f(best_rel, owner < 0 ? lane : (uint32_t)owner);
                if (busy && owner >= 0) best_rel = fminf(best_rel, owner_rel);
                const uint64_t idle_m = __ballot(!busy && !waiting);
                const uint64_t donor_m = __ballot(busy && pk > 0 && sp - bot >= 1);
                if ((uint32_t)__popcll(idle_m) >= io.steal && donor_m != 0) {
                    const uint32_t k = min((uint32_t)__popcll(idle_m), (uint32_t)__popcll(donor_m));
                    const uint32_t rank_idle = (uint32_t)__popcll(idle_m & lanes_below);
                    const bool thief = ((idle_m >> lane) & 1ull) && rank_idle < k;
                    const bool donor = ((donor_m >> lane) & 1ull) && (uint32_t)__popcll(donor_m & lanes_below) < k;
                    // the i-th idle lane (in lane order) is paired with the
                    // i-th donor; every lane takes part in the transfers
                    const uint32_t src = thief ? nth_set_bit(donor_m, rank_idle) : lane;
                    const float ox = shfl_f(r.o.x, src), oy = shfl_f(r.o.y, src), oz = shfl_f(r.o.z, src);
                    const float dx = shfl_f(r.d.x, src), dy = shfl_f(r.d.y, src), dz = shfl_f(r.d.z, src);
                    const float tmn = shfl_f(r.tmin, src);
                    const float ix = shfl_f(inv.x, src), iy = shfl_f(inv.y, src), iz = shfl_f(inv.z, src);
                    const float cx = shfl_f(cq.x, src), cy = shfl_f(cq.y, src), cz = shfl_f(cq.z, src);
                    const float tsh = shfl_f(t_shift, src), tminr = shfl_f(tmin_rel, src);
                    const float bestr = shfl_f(best_rel, src), bestt = shfl_f(best_t, src);
                    const int v_bot = __shfl(bot, (int)src);
                    const uint32_t v_plo = (uint32_t)__shfl((int)plo, (int)src);
                    const uint32_t v_best_lo = (uint32_t)__shfl((int)best_lo, (int)src);
                    const uint32_t v_pk = (uint32_t)__shfl((int)pk, (int)src);
                    const int32_t v_owner = __shfl(owner < 0 ? (int32_t)lane : owner, (int)src);
                    // batched lanes each carry their item's scene
                    auto shfl_ptr = [&](const void* q) -> uint64_t {
                        const uint64_t v = (uint64_t)q;
                        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, (int)src);
                        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), (int)src);
                        return ((uint64_t)hi << 32) | lo;
                    };
                    uint64_t v_w4 = 0, v_tris = 0, v_triid = 0;
                    if constexpr (BATCHED) {
                        v_w4 = shfl_ptr(l_w4);
                        v_tris = shfl_ptr(l_tris);
                        v_triid = shfl_ptr(l_triid);
                    }
                    if (thief) {
                        if constexpr (BATCHED) {
                            l_w4 = reinterpret_cast<const PNode4*>(v_w4);
                            l_tris = reinterpret_cast<const PTri*>(v_tris);
                            l_triid = reinterpret_cast<const int32_t*>(v_triid);
                        }
                        // the donor's bottom entry: v_bot < donor's sp <= STACK
                        cur = stk[v_bot * BLOCK + (int)(tid - lane + src)];
                        r.o = f3{ox, oy, oz};
                        r.d = f3{dx, dy, dz};
                        r.tmin = tmn;
                        inv = f3{ix, iy, iz};
                        cq = f3{cx, cy, cz};
                        set_planes(ix, iy, iz);
                        t_shift = tsh;
                        tmin_rel = tminr;
                        best_rel = bestr;
                        best_t = bestt;
                        bu = 0.0f; bv = 0.0f; best = -1;
                        sp = 0; bot = 0;
                        pk = v_pk - 1u;
                        plo = v_plo + (1u << pk);
                        // the donor's best keeps the piece it came from: a
                        // best an earlier helper handed over may come from a
                        // piece after this one, and then an equal t found
                        // here must replace it (the cap rule in step)
                        best_lo = v_best_lo;
                        owner = v_owner;
                        pending = 0;
                        busy = true;
                    }
                    if (donor) {
                        bot++;
                        pk--;
                    }
                    // owners count the helpers they gained
                    uint64_t tm = __ballot(thief);
                    while (tm != 0) {
                        const int j = __builtin_ctzll(tm);
                        tm &= tm - 1ull;
                        if ((int32_t)lane == __builtin_amdgcn_readlane(owner, j)) pending++;
                    }
                }
            }
        }
        // (a waiting owner implies a busy helper, so no busy lane means none)
        if (__ballot(busy) == 0) {
            if (drained) break;
            continue;
        }
        step(std::integral_constant<bool, STEAL>{});
        if constexpr (STEAL) {
            if (drained) {
                // ---- helpers that finished hand their result to the owner
                uint64_t fin = __ballot(handoff);
                while (fin != 0) {
                    const int j = __builtin_ctzll(fin);
                    fin &= fin - 1ull;
                    const int32_t o = __builtin_amdgcn_readlane(owner, j);
                    const int32_t hb = __builtin_amdgcn_readlane(best, j);
                    const float ht = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(best_t), j));
                    const float hu = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bu), j));
                    const float hv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(bv), j));
                    const uint32_t hlo = (uint32_t)__builtin_amdgcn_readlane((int)plo, j);
                    if ((int32_t)lane == o) {
                        pending--;
                        if (hb >= 0 && (ht < best_t || (ht == best_t && hlo < best_lo))) {
                            best = hb; best_t = ht; bu = hu; bv = hv; best_lo = hlo;
                            best_rel = ht - t_shift;
                        }
                    }
                }
                handoff = false;
                // an owner whose helpers are all done writes its hit in the
                // next step (sp = -1 finishes a ray at once)
                if (waiting && pending == 0) {
                    waiting = false;
                    busy = true;
                    sp = -1;
                }
            }
        }
    }
    if constexpr (STATS) {
        if (io.sums) {
            // the wave's totals, then one atomic pair per wave
            for (int off = 32; off > 0; off >>= 1) {
                sum_nv += (uint64_t)__shfl_xor((unsigned long long)sum_nv, off);
                sum_nt += (uint64_t)__shfl_xor((unsigned long long)sum_nt, off);
            }
            if (lane == 0) {
                atomicAdd(io.sums, (unsigned long long)sum_nv);
                atomicAdd(io.sums + 1, (unsigned long long)sum_nt);
            }
        }
    }
}

}  // namespace rtamd
