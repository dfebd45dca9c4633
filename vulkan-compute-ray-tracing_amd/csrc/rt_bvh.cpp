// rt_bvh.cpp -- host BVH builder with the reference's split policy, and the
// light-list rule.
//
// Bvh::createBvh (src/ray-tracing/Bvh.h:141-209) pops nodes from a stack,
// boxes the node's triangles (each box padded by eps = 1e-4, Bvh.h:16,94-98),
// draws the split axis with rand() % 3 for EVERY popped node (leaves too),
// std::sorts the node's triangles by box minimum on that axis (Bvh.h:116-137)
// and splits at span/2, numbering children in creation order (left =
// counter, right = counter + 1) and pushing right last so the right subtree
// is expanded first.  The flat output is indexed by those numbers.
//
// This builder produces the identical array without keeping per-node object
// vectors: the node's triangles are always a contiguous range of one index
// array (children are the two halves of the parent's sorted range), so each
// node sorts its range in place -- std::sort sees exactly the sequence the
// reference's copy holds and makes the same moves.  Because the tree shape
// depends only on the counts, each node's position in the pop order (and so
// its rand() draw) and its index are known in advance:
//   pop rank:   right child = r + 1,  left child = r + 2*|right leaves|
//   index:      children of the internal node with rank r, after L leaves
//               were popped, are 1 + 2*(r - L) and 2 + 2*(r - L)
// which lets independent subtrees be built on several threads.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <system_error>
#include <thread>
#include <vector>

#include "rt_common.hpp"
#include "rt_math.hpp"

#pragma clang fp contract(off)

namespace rtamd {
namespace {

// glibc random() TYPE_3 stream (the generator behind rand() when srand was
// not called or called with `seed`): r[0] = seed, r[i] = 16807*r[i-1] mod
// (2^31-1) for i < 31, r[i] = r[i-31] for i < 34, then r[i] = r[i-31] +
// r[i-3] (mod 2^32); output k is r[k+344] >> 1.
struct GlibcRandom {
    uint32_t r[34];
    int pos = 0;
    explicit GlibcRandom(uint32_t seed) {
        if (seed == 0) seed = 1;
        int32_t word = (int32_t)seed;
        r[0] = seed;
        for (int i = 1; i < 31; ++i) {
            int32_t hi = word / 127773, lo = word % 127773;
            word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            r[i] = (uint32_t)word;
        }
        for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
        pos = 0;
        for (int i = 0; i < 310; ++i) step();
    }
    // ring of the last 34 values; pos indexes the oldest
    uint32_t step() {
        uint32_t v = r[(pos + 3) % 34] + r[(pos + 31) % 34];
        r[pos] = v;
        pos = (pos + 1) % 34;
        return v;
    }
    int next() { return (int)(step() >> 1); }
};

struct Key {
    float k[3];     // objectBoundingBox(t).min per axis
    uint32_t idx;
};

inline float glm_min(float x, float y) { return y < x ? y : x; }
inline float glm_max(float x, float y) { return x < y ? y : x; }
const float kEps = 0.0001f;  // Bvh.h:16

inline void tri_box(const rt_triangle& t, float mn[3], float mx[3]) {
    for (int k = 0; k < 3; ++k) {
        mn[k] = glm_min(glm_min(t.v0[k], t.v1[k]), t.v2[k]) - kEps;
        mx[k] = glm_max(glm_max(t.v0[k], t.v1[k]), t.v2[k]) + kEps;
    }
}

struct Builder {
    const rt_triangle* tris;
    Key* items;
    rt_bvh_node* out;
    const uint8_t* axis;        // rand() % 3 by pop rank
    uint32_t spawn_above;       // subtree size above which a thread is used
    int spawn_depth;            // levels that may still spawn

    void run(uint32_t s, uint32_t cnt, int32_t idx, uint32_t rank, uint32_t leaves_before, int depth) {
        const int ax = axis[rank];
        std::sort(items + s, items + s + cnt, [ax](const Key& a, const Key& b) { return a.k[ax] < b.k[ax]; });
        rt_bvh_node& nd = out[idx];
        std::memset(&nd, 0, sizeof nd);
        if (cnt <= 1) {
            tri_box(tris[items[s].idx], nd.min, nd.max);
            nd.leftNodeIndex = -1;
            nd.rightNodeIndex = -1;
            nd.objectIndex = (int32_t)items[s].idx;
            return;
        }
        const uint32_t mid = cnt / 2, rcnt = cnt - mid;
        const int32_t left = (int32_t)(1u + 2u * (rank - leaves_before));
        const int32_t right = left + 1;
        nd.leftNodeIndex = left;
        nd.rightNodeIndex = right;
        nd.objectIndex = -1;
        const uint32_t left_rank = rank + 2u * rcnt;
        if (depth < spawn_depth && cnt > spawn_above) {
            std::thread t([&, s, mid, rcnt, right, rank, leaves_before, depth] {
                run(s + mid, rcnt, right, rank + 1, leaves_before, depth + 1);
            });
            run(s, mid, left, left_rank, leaves_before + rcnt, depth + 1);
            t.join();
        } else {
            run(s + mid, rcnt, right, rank + 1, leaves_before, depth + 1);
            run(s, mid, left, left_rank, leaves_before + rcnt, depth + 1);
        }
        const rt_bvh_node& L = out[left];
        const rt_bvh_node& R = out[right];
        rt_bvh_node& me = out[idx];
        for (int k = 0; k < 3; ++k) {
            me.min[k] = glm_min(L.min[k], R.min[k]);
            me.max[k] = glm_max(L.max[k], R.max[k]);
        }
    }
};

int build(const rt_triangle* tris, uint32_t n, const std::vector<uint8_t>& axis, uint32_t threads, rt_bvh_node* out) {
    std::vector<Key> items(n);
    for (uint32_t i = 0; i < n; ++i) {
        float mn[3], mx[3];
        tri_box(tris[i], mn, mx);
        items[i] = Key{{mn[0], mn[1], mn[2]}, i};
    }
    Builder b{tris, items.data(), out, axis.data(), 1u << 14, 0};
    if (threads > 1) {
        int d = 0;
        while ((1u << d) < threads && d < 10) ++d;
        b.spawn_depth = d;
    }
    b.run(0, n, 0, 0, 0, 0);
    return RT_OK;
}

bool finite_tris(const rt_triangle* tris, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(tris[i].v0[k]) || !std::isfinite(tris[i].v1[k]) || !std::isfinite(tris[i].v2[k]))
                return false;
    return true;
}

}  // namespace
}  // namespace rtamd

using namespace rtamd;

extern "C" int rt_bvh_build(const rt_triangle* tris, uint32_t n, rt_bvh_node* out) {
    if (!tris || !out || n == 0 || n > 0x40000000u) return RT_ERR_INVALID;
    if (!finite_tris(tris, n)) return RT_ERR_INVALID;
    try {
        std::vector<uint8_t> axis(2 * (size_t)n - 1);
        for (auto& a : axis) a = (uint8_t)(std::rand() % 3);
        return build(tris, n, axis, 1, out);
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    }
}

extern "C" int rt_bvh_build_seeded(const rt_triangle* tris, uint32_t n, uint32_t seed, uint32_t threads,
                                   rt_bvh_node* out) {
    if (!tris || !out || n == 0 || n > 0x40000000u) return RT_ERR_INVALID;
    if (!finite_tris(tris, n)) return RT_ERR_INVALID;
    try {
        GlibcRandom g(seed);
        std::vector<uint8_t> axis(2 * (size_t)n - 1);
        for (auto& a : axis) a = (uint8_t)(g.next() % 3);
        return build(tris, n, axis, threads ? threads : 1, out);
    } catch (const std::bad_alloc&) {
        return RT_ERR_NOMEM;
    } catch (const std::system_error&) {
        return RT_ERR_UNSUPPORTED;
    }
}

extern "C" int rt_scene_lights(const rt_triangle* tris, uint32_t n, const rt_material* mats, uint32_t num_mats,
                               rt_light* out, uint32_t cap, uint32_t* count) {
    if (!count || (n && !tris) || (n && !mats)) return RT_ERR_INVALID;
    uint32_t c = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t m = tris[i].materialIndex;
        if (m >= num_mats) return RT_ERR_INVALID;
        if (mats[m].type != (uint32_t)RT_MAT_LIGHT) continue;
        // RtScene.h:93: glm::length(glm::cross(t.v0, t.v1)) * 0.5f
        f3 cr = cross(ld3(tris[i].v0), ld3(tris[i].v1));
        float area = sqrtf(dot(cr, cr)) * 0.5f;
        if (out && c < cap) out[c] = rt_light{i, area};
        ++c;
    }
    *count = c;
    return (out && c > cap) ? RT_ERR_INVALID : RT_OK;
}
