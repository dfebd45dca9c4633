// rt_scene.cpp -- scene validation, repacking and upload to HBM.
//
// Replaces the reference's upload of the RtScene arrays into per-swapchain
// storage buffers (main.cpp:84-106, Buffer.h:83-102) and the descriptor
// binding of ComputeMaterial (ComputeMaterial.cpp:15-26, Material.cpp:258-322).
// Unlike Buffer.h:85, each array is bound with its whole length.
#include <functional>
#include <algorithm>
#include <cmath>
#include <cstring>

#include "rt_common.hpp"
#include "rt_math.hpp"

#pragma clang fp contract(off)

namespace rtamd {

namespace {

inline bool finite3(const float* v) {
    return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]);
}

// Leaf test of the packed traversal's precondition: the triangle's box lies
// strictly inside the node box, so no ray along a box face reaches it.
bool leaf_encloses(const rt_bvh_node& n, const rt_triangle& t) {
    if (!finite3(n.min) || !finite3(n.max) || !finite3(t.v0) || !finite3(t.v1) || !finite3(t.v2)) return false;
    for (int k = 0; k < 3; ++k) {
        float lo = std::min(std::min(t.v0[k], t.v1[k]), t.v2[k]);
        float hi = std::max(std::max(t.v0[k], t.v1[k]), t.v2[k]);
        if (!(n.min[k] < lo) || !(hi < n.max[k])) return false;
    }
    return true;
}

bool box_contains(const rt_bvh_node& p, const rt_bvh_node& c) {
    if (!finite3(p.min) || !finite3(p.max) || !finite3(c.min) || !finite3(c.max)) return false;
    for (int k = 0; k < 3; ++k)
        if (!(p.min[k] <= c.min[k]) || !(c.max[k] <= p.max[k])) return false;
    return true;
}

}  // namespace

SceneCheck check_scene(const rt_scene_desc& d) {
    SceneCheck c;
    auto fail = [&](const std::string& m) { c.error = m; return c; };
    if (d.num_triangles == 0 || !d.triangles) return fail("scene has no triangles (Bvh.h:77 is undefined for an empty list)");
    if (d.num_nodes == 0 || !d.nodes) return fail("scene has no BVH nodes");
    if (d.num_materials == 0 || !d.materials) return fail("scene has no materials");
    if (d.num_lights && !d.lights) return fail("lights pointer is null");
    if (d.num_spheres && !d.spheres) return fail("spheres pointer is null");
    for (uint32_t i = 0; i < d.num_materials; ++i)
        if (d.materials[i].type > 3u) return fail("material " + std::to_string(i) + " has type > 3");
    for (uint32_t i = 0; i < d.num_triangles; ++i)
        if (d.triangles[i].materialIndex >= d.num_materials)
            return fail("triangle " + std::to_string(i) + " names material out of range");
    for (uint32_t i = 0; i < d.num_lights; ++i)
        if (d.lights[i].triangleIndex >= d.num_triangles)
            return fail("light " + std::to_string(i) + " names triangle out of range");
    for (uint32_t i = 0; i < d.num_spheres; ++i)
        if (d.spheres[i].materialIndex >= d.num_materials)
            return fail("sphere " + std::to_string(i) + " names material out of range");
    const int32_t nn = (int32_t)d.num_nodes, nt = (int32_t)d.num_triangles;
    for (int32_t i = 0; i < nn; ++i) {
        const rt_bvh_node& n = d.nodes[i];
        if (n.leftNodeIndex < -1 || n.leftNodeIndex >= nn || n.rightNodeIndex < -1 || n.rightNodeIndex >= nn)
            return fail("node " + std::to_string(i) + " has a child index out of range");
        if (n.objectIndex < -1 || n.objectIndex >= nt)
            return fail("node " + std::to_string(i) + " has an object index out of range");
    }
    // Walk from the root: every node may be reached once (a tree).
    std::vector<uint8_t> seen(d.num_nodes, 0);
    std::vector<std::pair<int32_t, uint32_t>> st;
    st.emplace_back(0, 0u);
    while (!st.empty()) {
        auto [i, depth] = st.back();
        st.pop_back();
        if (seen[i]) return fail("node " + std::to_string(i) + " is reachable twice: the node array is not a tree");
        seen[i] = 1;
        const rt_bvh_node& n = d.nodes[i];
        const bool has_kids = n.leftNodeIndex != -1 || n.rightNodeIndex != -1;
        if (n.objectIndex != -1) {
            if (has_kids) c.standard = false;
            c.leaves++;
            c.height = std::max(c.height, depth);
            if (!leaf_encloses(n, d.triangles[n.objectIndex])) c.enclosing = false;
        }
        for (int32_t ch : {n.leftNodeIndex, n.rightNodeIndex}) {
            if (ch == -1) continue;
            if (!box_contains(n, d.nodes[ch])) c.enclosing = false;
            st.emplace_back(ch, depth + 1);
        }
    }
    return c;
}

void pack_tree(const rt_scene_desc& d, PackedTree& out) {
    const uint32_t nn = d.num_nodes;
    out.nodes.clear();
    out.tris.clear();
    out.tri_id.clear();
    out.nodes.reserve(nn / 2 + 1);
    out.tris.reserve(d.num_triangles);
    out.tri_id.reserve(d.num_triangles);

    // has_tri[i]: the subtree under node i holds at least one triangle.
    // Children may carry any index, so resolve it with a post-order walk.
    std::vector<uint8_t> has_tri(nn, 0);
    {
        std::vector<std::pair<int32_t, bool>> st;
        st.emplace_back(0, false);
        while (!st.empty()) {
            auto [i, done] = st.back();
            st.pop_back();
            const rt_bvh_node& n = d.nodes[i];
            if (done) {
                bool h = n.objectIndex != -1;
                if (n.leftNodeIndex != -1 && has_tri[n.leftNodeIndex]) h = true;
                if (n.rightNodeIndex != -1 && has_tri[n.rightNodeIndex]) h = true;
                has_tri[i] = h;
                continue;
            }
            st.emplace_back(i, true);
            if (n.leftNodeIndex != -1) st.emplace_back(n.leftNodeIndex, false);
            if (n.rightNodeIndex != -1) st.emplace_back(n.rightNodeIndex, false);
        }
    }

    auto add_leaf = [&](int32_t obj) -> int32_t {
        const rt_triangle& t = d.triangles[obj];
        f3 v0 = ld3(t.v0), v1 = ld3(t.v1), v2 = ld3(t.v2);
        f3 e = v0 - v1;        // triIntersect's a
        f3 f = v2 - v0;        // triIntersect's b
        f3 n = cross(f, e);    // n = cross(b, a)
        PTri p;
        p.a = make_float4(v0.x, v0.y, v0.z, e.x);
        p.b = make_float4(e.y, e.z, f.x, f.y);
        p.c = make_float4(f.z, n.x, n.y, n.z);
        int32_t slot = (int32_t)out.tris.size();
        out.tris.push_back(p);
        out.tri_id.push_back(obj);
        return ~slot;
    };

    // Follow single-child chains down to the node that decides a subtree.
    auto resolve = [&](int32_t i) -> int32_t {
        for (;;) {
            const rt_bvh_node& n = d.nodes[i];
            if (n.objectIndex != -1) return i;
            int32_t l = n.leftNodeIndex, r = n.rightNodeIndex;
            bool hl = l != -1 && has_tri[l], hr = r != -1 && has_tri[r];
            if (hl && hr) return i;
            i = hl ? l : r;
        }
    };

    if (!has_tri[0]) { out.root = kEmptyRoot; out.height = 0; return; }

    // Pre-order emission with an explicit stack; each frame fills one
    // child slot of an already-emitted PNode.
    struct Frame { int32_t src; int32_t parent; int slot; uint32_t depth; };
    std::vector<Frame> st;
    int32_t root_src = resolve(0);
    st.push_back({root_src, -1, 0, 1});
    out.height = 0;
    while (!st.empty()) {
        Frame f = st.back();
        st.pop_back();
        const rt_bvh_node& n = d.nodes[f.src];
        int32_t ref;
        if (n.objectIndex != -1) {
            ref = add_leaf(n.objectIndex);
        } else {
            ref = (int32_t)out.nodes.size();
            out.nodes.push_back(PNode{});
            out.height = std::max(out.height, f.depth);
            int32_t l = resolve(n.leftNodeIndex), r = resolve(n.rightNodeIndex);
            const rt_bvh_node& L = d.nodes[n.leftNodeIndex];   // boxes of the direct children
            const rt_bvh_node& R = d.nodes[n.rightNodeIndex];
            PNode& p = out.nodes[ref];
            p.a = make_float4(L.min[0], L.min[1], L.min[2], L.max[0]);
            p.b = make_float4(L.max[1], L.max[2], R.min[0], R.min[1]);
            p.c = make_float4(R.min[2], R.max[0], R.max[1], R.max[2]);
            p.d = make_int4(0, 0, 0, 0);
            // right pushed first so the left subtree follows its parent
            st.push_back({r, ref, 1, f.depth + 1});
            st.push_back({l, ref, 0, f.depth + 1});
        }
        if (f.parent < 0) out.root = ref;
        else if (f.slot == 0) out.nodes[f.parent].d.x = ref;
        else out.nodes[f.parent].d.y = ref;
    }
    quantize_nodes(out);
    widen_nodes(out);
}

void widen_nodes(PackedTree& out) {
    out.w4nodes.clear();
    out.w4nodes.reserve(out.nodes.size() / 2 + 1);
    out.height4 = 0;
    if (out.root == kEmptyRoot || out.root < 0) { out.root4 = out.root; return; }
    // the binary record i holds its two children's boxes (q16) and refs;
    // the 4-wide record of i takes its grandchildren through internal
    // children and the leaf children themselves
    struct Item { uint32_t w0, w1, w2; int32_t ref; };
    auto child = [&](int32_t bi, int k) -> Item {
        const PNodeQ& q = out.qnodes[bi];
        return k == 0 ? Item{q.w[0], q.w[1], q.w[2], q.ref0} : Item{q.w[3], q.w[4], q.w[5], q.ref1};
    };
    struct Frame { int32_t bi; int32_t parent; int slot; uint32_t depth; };
    std::vector<Frame> st;
    st.push_back({out.root, -1, 0, 1});
    while (!st.empty()) {
        Frame f = st.back();
        st.pop_back();
        Item items[4];
        int n = 0;
        for (int k = 0; k < 2; ++k) {
            Item c = child(f.bi, k);
            if (c.ref >= 0) {
                items[n++] = child(c.ref, 0);
                items[n++] = child(c.ref, 1);
            } else {
                items[n++] = c;
            }
        }
        const int32_t me = (int32_t)out.w4nodes.size();
        out.w4nodes.push_back(PNode4{});
        out.height4 = std::max(out.height4, f.depth);
        PNode4& w = out.w4nodes[me];
        for (int k = 0; k < 4; ++k) {
            if (k < n) {
                // the binary record's words hold (lo x, lo y), (lo z, hi x),
                // (hi y, hi z); the 4-wide one pairs each axis' planes,
                // (lo x, hi x), (lo y, hi y), (lo z, hi z), so a ray picks the
                // plane it enters by and the one it leaves by from one word
                const uint32_t lx = items[k].w0 & 0xffffu, ly = items[k].w0 >> 16;
                const uint32_t lz = items[k].w1 & 0xffffu, hx = items[k].w1 >> 16;
                const uint32_t hy = items[k].w2 & 0xffffu, hz = items[k].w2 >> 16;
                w.w[3 * k] = lx | hx << 16;
                w.w[3 * k + 1] = ly | hy << 16;
                w.w[3 * k + 2] = lz | hz << 16;
                w.ref[k] = items[k].ref;     // patched below for internal children
            } else {
                w.w[3 * k] = w.w[3 * k + 1] = w.w[3 * k + 2] = 0;
                w.ref[k] = kNoChild;
            }
        }
        for (int k = n - 1; k >= 0; --k)
            if (items[k].ref >= 0) st.push_back({items[k].ref, me, k, f.depth + 1});
        if (f.parent < 0) out.root4 = me;
        else out.w4nodes[f.parent].ref[f.slot] = me;
    }
}

void quantize_nodes(PackedTree& out) {
    out.qnodes.assign(out.nodes.size(), PNodeQ{});
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (const PNode& p : out.nodes) {
        const float c[12] = {p.a.x, p.a.y, p.a.z, p.a.w, p.b.x, p.b.y, p.b.z, p.b.w, p.c.x, p.c.y, p.c.z, p.c.w};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], (double)std::min(c[k], c[6 + k]));
            hi[k] = std::max(hi[k], (double)std::max(c[3 + k], c[9 + k]));
        }
    }
    if (out.nodes.empty()) for (int k = 0; k < 3; ++k) lo[k] = hi[k] = 0.0;
    // The frame puts the scene's lowest bound 2 quanta above the origin and
    // its highest at most 65002 quanta above, so every quantized bound
    // (rounded outward by one extra quantum below) lies strictly inside
    // [0, 65535] with at least one full quantum between it and the exact one,
    // also at the scene's edges.
    float org[3], scl[3];
    for (int k = 0; k < 3; ++k) {
        const double ext = std::max(hi[k] - lo[k], 1e-30);
        scl[k] = (float)(ext / 65000.0);
        if ((double)scl[k] * 65000.0 < ext) scl[k] = std::nextafter(scl[k], INFINITY);
        org[k] = (float)(lo[k] - 2.0 * (double)scl[k]);
        if ((double)org[k] > lo[k] - 2.0 * (double)scl[k]) org[k] = std::nextafter(org[k], -INFINITY);
    }
    out.frame = QFrame{org[0], org[1], org[2], scl[0], scl[1], scl[2], 1.0f / scl[0], 1.0f / scl[1], 1.0f / scl[2]};
    // q covers [org + q*scl] exactly (double arithmetic of the float frame)
    auto qlo = [&](float v, int k) -> uint32_t {
        double q = std::floor(((double)v - (double)org[k]) / (double)scl[k]) - 1.0;
        return (uint32_t)std::min(65535.0, std::max(0.0, q));
    };
    auto qhi = [&](float v, int k) -> uint32_t {
        double q = std::ceil(((double)v - (double)org[k]) / (double)scl[k]) + 1.0;
        return (uint32_t)std::min(65535.0, std::max(0.0, q));
    };
    for (size_t i = 0; i < out.nodes.size(); ++i) {
        const PNode& p = out.nodes[i];
        const float l0[3] = {p.a.x, p.a.y, p.a.z}, h0[3] = {p.a.w, p.b.x, p.b.y};
        const float l1[3] = {p.b.z, p.b.w, p.c.x}, h1[3] = {p.c.y, p.c.z, p.c.w};
        PNodeQ& q = out.qnodes[i];
        q.w[0] = qlo(l0[0], 0) | qlo(l0[1], 1) << 16;
        q.w[1] = qlo(l0[2], 2) | qhi(h0[0], 0) << 16;
        q.w[2] = qhi(h0[1], 1) | qhi(h0[2], 2) << 16;
        q.w[3] = qlo(l1[0], 0) | qlo(l1[1], 1) << 16;
        q.w[4] = qlo(l1[2], 2) | qhi(h1[0], 0) << 16;
        q.w[5] = qhi(h1[1], 1) | qhi(h1[2], 2) << 16;
        q.ref0 = p.d.x;
        q.ref1 = p.d.y;
    }
}

}  // namespace rtamd

using namespace rtamd;

namespace {
constexpr uint32_t kMaxPackedHeight = 64;

template <typename T>
int upload_array(rt_ctx* ctx, T** dst, const T* src, uint32_t n, uint64_t& bytes) {
    *dst = nullptr;
    if (!n) return RT_OK;
    size_t b = sizeof(T) * (size_t)n;
    RT_HIP(ctx, hipMalloc((void**)dst, b));
    RT_HIP(ctx, hipMemcpy(*dst, src, b, hipMemcpyHostToDevice));
    bytes += b;
    return RT_OK;
}

// The 4-wide node records and the packed triangles, in one allocation when
// they fit in 4 GiB (nodes first, triangles from the next 256-byte boundary;
// the traversal then addresses both by 32-bit offsets from one base), else
// in two.
int upload_records(rt_ctx* ctx, rt_scene* s, const PackedTree& packed, uint64_t& bytes) {
    const size_t nb = (sizeof(PNode4) * packed.w4nodes.size() + 255) / 256 * 256;
    const size_t tb = sizeof(PTri) * packed.tris.size();
    if (nb + tb >= (1ull << 32) || !nb || !tb) {
        int rc = upload_array(ctx, &s->d_w4nodes, packed.w4nodes.data(), (uint32_t)packed.w4nodes.size(), bytes);
        if (!rc) rc = upload_array(ctx, &s->d_ptris, packed.tris.data(), (uint32_t)packed.tris.size(), bytes);
        s->tri_off = kSeparateArrays;
        return rc;
    }
    char* base = nullptr;
    RT_HIP(ctx, hipMalloc((void**)&base, nb + tb));
    s->d_w4nodes = reinterpret_cast<PNode4*>(base);          // owns the allocation
    s->d_ptris = reinterpret_cast<PTri*>(base + nb);
    s->tri_off = (uint32_t)nb;
    RT_HIP(ctx, hipMemcpy(base, packed.w4nodes.data(), sizeof(PNode4) * packed.w4nodes.size(), hipMemcpyHostToDevice));
    RT_HIP(ctx, hipMemcpy(base + nb, packed.tris.data(), tb, hipMemcpyHostToDevice));
    bytes += nb + tb;
    return RT_OK;
}

// Sphere tree: median split on the longest axis of the centres' bounds
// (ties by sphere index), one sphere per leaf, boxes = the spheres' bounds
// padded by 2^-10 (|centre| + radius) + 1e-6 so rounding in the sphere
// test never puts an accepted root outside its box.  Depth <= ceil(log2 n).
// Returns false (keep the plain loop) for non-finite spheres.
bool build_sphere_tree(const rt_sphere* sp, uint32_t n, std::vector<SphereNode>& out) {
    std::vector<float> lo(3 * (size_t)n), hi(3 * (size_t)n);
    for (uint32_t j = 0; j < n; ++j) {
        const float* s4 = sp[j].s;
        if (!std::isfinite(s4[0]) || !std::isfinite(s4[1]) || !std::isfinite(s4[2]) || !std::isfinite(s4[3]))
            return false;
        const float r = std::fabs(s4[3]);
        const float m = std::max(std::max(std::fabs(s4[0]), std::fabs(s4[1])), std::fabs(s4[2]));
        const float pad = (m + r) * 0x1p-10f + 1e-6f;
        for (int a = 0; a < 3; ++a) {
            lo[3 * (size_t)j + a] = s4[a] - r - pad;
            hi[3 * (size_t)j + a] = s4[a] + r + pad;
        }
    }
    std::vector<uint32_t> idx(n);
    for (uint32_t j = 0; j < n; ++j) idx[j] = j;
    out.clear();
    out.reserve(2 * (size_t)n);
    // explicit recursion on [b, e) returning the node index
    std::function<int32_t(uint32_t, uint32_t)> rec = [&](uint32_t b, uint32_t e) -> int32_t {
        const int32_t me = (int32_t)out.size();
        out.push_back(SphereNode{});
        if (e - b == 1) {
            const uint32_t j = idx[b];
            SphereNode& nd = out[me];
            for (int a = 0; a < 3; ++a) { nd.lo[a] = lo[3 * (size_t)j + a]; nd.hi[a] = hi[3 * (size_t)j + a]; }
            nd.left = -1 - (int32_t)j;
            nd.right = -1;
            return me;
        }
        float cmin[3] = {INFINITY, INFINITY, INFINITY}, cmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t k = b; k < e; ++k)
            for (int a = 0; a < 3; ++a) {
                const float c = sp[idx[k]].s[a];
                cmin[a] = std::min(cmin[a], c);
                cmax[a] = std::max(cmax[a], c);
            }
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (cmax[a] - cmin[a] > cmax[axis] - cmin[axis]) axis = a;
        const uint32_t mid = b + (e - b) / 2;
        std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e, [&](uint32_t x, uint32_t y) {
            const float cx = sp[x].s[axis], cy = sp[y].s[axis];
            return cx < cy || (cx == cy && x < y);
        });
        const int32_t l = rec(b, mid);
        const int32_t r = rec(mid, e);
        SphereNode& nd = out[me];
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] = std::min(out[l].lo[a], out[r].lo[a]);
            nd.hi[a] = std::max(out[l].hi[a], out[r].hi[a]);
        }
        nd.left = l;
        nd.right = r;
        return me;
    };
    rec(0, n);
    return true;
}

void free_scene(rt_scene* s) {
    if (!s) return;
    // d_ptris lies inside d_w4nodes' allocation unless the two are separate
    void* ptris = s->tri_off == kSeparateArrays ? (void*)s->d_ptris : nullptr;
    for (void* p : {(void*)s->d_tris, (void*)s->d_mats, (void*)s->d_nodes, (void*)s->d_lights,
                    (void*)s->d_spheres, (void*)s->d_pnodes, (void*)s->d_qnodes, (void*)s->d_w4nodes,
                    ptris, (void*)s->d_snodes,
                    (void*)s->d_ptri_id})
        if (p) (void)hipFree(p);
    delete s;
}
}  // namespace

// The linear BVH of the scene's triangles, built on the device and copied
// back (RT_TREE_DEVICE uploads).
int device_tree(rt_ctx* ctx, const rt_triangle* tris, uint32_t n, std::vector<rt_bvh_node>& out) {
    if (!tris || n == 0) return set_error(ctx, RT_ERR_INVALID, "rt_scene_upload: scene has no triangles");
    if (n > (1u << 30)) return set_error(ctx, RT_ERR_INVALID, "rt_scene_upload: more than 2^30 triangles");
    const size_t tb = sizeof(rt_triangle) * (size_t)n, nb = sizeof(rt_bvh_node) * (2 * (size_t)n - 1);
    rt_triangle* dt = nullptr;
    rt_bvh_node* dn = nullptr;
    int rc = RT_OK;
    hipError_t e = hipMalloc((void**)&dt, tb);
    if (e == hipSuccess) e = hipMalloc((void**)&dn, nb);
    if (e == hipSuccess) e = hipMemcpy(dt, tris, tb, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_error(ctx, e, "rt_scene_upload: device tree");
    if (!rc) rc = rt_bvh_build_device(ctx, dt, n, dn, nullptr);
    if (!rc) {
        try {
            out.resize(2 * (size_t)n - 1);
        } catch (const std::bad_alloc&) {
            rc = set_error(ctx, RT_ERR_NOMEM, "rt_scene_upload: device tree");
        }
    }
    if (!rc) {
        e = hipMemcpy(out.data(), dn, nb, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_error(ctx, e, "rt_scene_upload: device tree");
    }
    if (dt) (void)hipFree(dt);
    if (dn) (void)hipFree(dn);
    return rc;
}

namespace rtamd {

uint64_t packed_record_bytes(const PackedTree& p) {
    const uint64_t nb = (sizeof(PNode4) * (uint64_t)p.w4nodes.size() + 255) / 256 * 256;
    const uint64_t tb = sizeof(PTri) * (uint64_t)p.tris.size();
    if (nb + tb >= (1ull << 32) || !nb || !tb) return sizeof(PNode4) * (uint64_t)p.w4nodes.size() + tb;
    return nb + tb;
}

int scene_prepare(rt_ctx* ctx, const rt_scene_desc* desc, PreparedScene& P) {
    if (!ctx) return RT_ERR_INVALID;
    if (!desc) return set_error(ctx, RT_ERR_INVALID, "rt_scene_upload: null argument");
    // `user` keeps the caller's tree (the shader-order traversal walks it),
    // `packsrc` is the tree the packed traversal is built from
    rt_scene_desc user = *desc, packsrc = *desc;
    P.tree_built = ctx->tuning.tree == RT_TREE_DEVICE;
    if (P.tree_built) {
        RT_HIP(ctx, hipSetDevice(ctx->device));
        int rc = device_tree(ctx, desc->triangles, desc->num_triangles, P.built);
        if (rc) return rc;
        packsrc.nodes = P.built.data();
        packsrc.num_nodes = (uint32_t)P.built.size();
        if (!desc->nodes || desc->num_nodes == 0) user = packsrc;
    }
    SceneCheck pchk;
    try {
        P.chk = check_scene(user);
        pchk = P.tree_built && !P.chk.error.empty() ? P.chk : (P.tree_built ? check_scene(packsrc) : P.chk);
    } catch (const std::exception& e) {
        return set_error(ctx, RT_ERR_NOMEM, std::string("rt_scene_upload: ") + e.what());
    }
    if (!P.chk.error.empty()) return set_error(ctx, RT_ERR_INVALID, "rt_scene_upload: " + P.chk.error);
    if (!pchk.error.empty()) return set_error(ctx, RT_ERR_INVALID, "rt_scene_upload: device tree: " + pchk.error);
    P.fast = pchk.standard && pchk.enclosing;
    if (P.fast) {
        try {
            pack_tree(packsrc, P.packed);
        } catch (const std::exception& e) {
            return set_error(ctx, RT_ERR_NOMEM, std::string("rt_scene_upload: ") + e.what());
        }
        if (P.packed.height > kMaxPackedHeight) P.fast = false;
    }
    P.desc = user;
    if (user.num_spheres > kSphereLinear && user.num_spheres <= kMaxTreeSpheres) {
        try {
            P.sphere_tree = build_sphere_tree(user.spheres, user.num_spheres, P.snodes);
        } catch (const std::bad_alloc&) {
            return set_error(ctx, RT_ERR_NOMEM, "rt_scene_upload: sphere tree");
        }
    }
    // binary node records only when this context's tuning reads them; a
    // later query in such a layout builds them then (scene_binary_records)
    P.with_pnodes = P.fast && ctx->tuning.node_layout == RT_LAYOUT_F32;
    P.with_qnodes = P.fast && ctx->tuning.node_layout == RT_LAYOUT_Q16;
    uint64_t b = sizeof(rt_triangle) * (uint64_t)user.num_triangles + sizeof(rt_material) * (uint64_t)user.num_materials +
                 sizeof(rt_bvh_node) * (uint64_t)user.num_nodes + sizeof(rt_light) * (uint64_t)user.num_lights +
                 sizeof(rt_sphere) * (uint64_t)user.num_spheres;
    if (P.sphere_tree) b += sizeof(SphereNode) * (uint64_t)P.snodes.size();
    if (P.fast) {
        b += packed_record_bytes(P.packed) + sizeof(int32_t) * (uint64_t)P.packed.tri_id.size();
        if (P.with_pnodes) b += sizeof(PNode) * (uint64_t)P.packed.nodes.size();
        if (P.with_qnodes) b += sizeof(PNodeQ) * (uint64_t)P.packed.qnodes.size();
    }
    P.bytes = b;
    return RT_OK;
}

int scene_commit(rt_ctx* ctx, PreparedScene& P, rt_scene** out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    *out = nullptr;
    RT_HIP(ctx, hipSetDevice(ctx->device));
    const rt_scene_desc* desc = &P.desc;
    const PackedTree& packed = P.packed;
    const bool fast = P.fast;
    rt_scene* s = new rt_scene();
    s->ctx = ctx;
    uint64_t bytes = 0;
    int rc = RT_OK;
    do {
        if ((rc = upload_array(ctx, &s->d_tris, desc->triangles, desc->num_triangles, bytes))) break;
        if ((rc = upload_array(ctx, &s->d_mats, desc->materials, desc->num_materials, bytes))) break;
        if ((rc = upload_array(ctx, &s->d_nodes, desc->nodes, desc->num_nodes, bytes))) break;
        if ((rc = upload_array(ctx, &s->d_lights, desc->lights, desc->num_lights, bytes))) break;
        if ((rc = upload_array(ctx, &s->d_spheres, desc->spheres, desc->num_spheres, bytes))) break;
        if (P.sphere_tree &&
            (rc = upload_array(ctx, &s->d_snodes, P.snodes.data(), (uint32_t)P.snodes.size(), bytes))) break;
        if (fast) {
            if (P.with_pnodes &&
                (rc = upload_array(ctx, &s->d_pnodes, packed.nodes.data(), (uint32_t)packed.nodes.size(), bytes))) break;
            if (P.with_qnodes &&
                (rc = upload_array(ctx, &s->d_qnodes, packed.qnodes.data(), (uint32_t)packed.qnodes.size(), bytes))) break;
            if ((rc = upload_records(ctx, s, packed, bytes))) break;
            if ((rc = upload_array(ctx, &s->d_ptri_id, packed.tri_id.data(), (uint32_t)packed.tri_id.size(), bytes))) break;
        }
    } while (0);
    if (rc) {
        free_scene(s);
        return rc;
    }
    s->root_ref = fast ? packed.root : kEmptyRoot;
    s->frame = packed.frame;
    s->root4 = fast ? packed.root4 : kEmptyRoot;
    s->height4 = fast ? packed.height4 : 0;
    s->packed_height = fast ? packed.height : 0;
    s->ref_height = P.chk.height;
    s->fast = fast;
    s->tree_built = P.tree_built;
    rt_scene_info& in = s->info;
    in.num_triangles = desc->num_triangles;
    in.num_nodes = desc->num_nodes;
    in.num_lights = desc->num_lights;
    in.num_spheres = desc->num_spheres;
    in.num_materials = desc->num_materials;
    in.tree_height = P.chk.height;
    in.reachable_leaves = P.chk.leaves;
    in.enclosing = P.chk.enclosing ? 1u : 0u;
    in.fast_path = fast ? 1u : 0u;
    in.device_bytes = bytes;
    *out = s;
    return RT_OK;
}

int scene_binary_records(rt_ctx* ctx, rt_scene* s, uint32_t layout) {
    if (!s->fast) return RT_OK;
    std::lock_guard<std::mutex> lk(s->lazy_mu);
    const bool want_f = layout == RT_LAYOUT_F32 && !s->d_pnodes;
    const bool want_q = layout == RT_LAYOUT_Q16 && !s->d_qnodes;
    if (!want_f && !want_q) return RT_OK;
    RT_HIP(ctx, hipSetDevice(ctx->device));
    // the tree the packed arrays came from, read back: the scene's own node
    // array, or the device builder's tree over its triangles
    const uint32_t nt = s->info.num_triangles;
    std::vector<rt_triangle> tris;
    std::vector<rt_bvh_node> nodes;
    try {
        tris.resize(nt);
        if (!s->tree_built) nodes.resize(s->info.num_nodes);
    } catch (const std::bad_alloc&) {
        return set_error(ctx, RT_ERR_NOMEM, "binary node records: host copy");
    }
    RT_HIP(ctx, hipMemcpy(tris.data(), s->d_tris, sizeof(rt_triangle) * (size_t)nt, hipMemcpyDeviceToHost));
    if (s->tree_built) {
        int rc = device_tree(ctx, tris.data(), nt, nodes);
        if (rc) return rc;
    } else {
        RT_HIP(ctx, hipMemcpy(nodes.data(), s->d_nodes, sizeof(rt_bvh_node) * nodes.size(), hipMemcpyDeviceToHost));
    }
    rt_scene_desc d{};
    d.triangles = tris.data();
    d.num_triangles = nt;
    d.nodes = nodes.data();
    d.num_nodes = (uint32_t)nodes.size();
    PackedTree packed;
    try {
        pack_tree(d, packed);
    } catch (const std::exception& e) {
        return set_error(ctx, RT_ERR_NOMEM, std::string("binary node records: ") + e.what());
    }
    uint64_t bytes = 0;
    int rc = want_f ? upload_array(ctx, &s->d_pnodes, packed.nodes.data(), (uint32_t)packed.nodes.size(), bytes)
                    : upload_array(ctx, &s->d_qnodes, packed.qnodes.data(), (uint32_t)packed.qnodes.size(), bytes);
    if (rc) return rc;
    s->info.device_bytes += bytes;
    return RT_OK;
}

}  // namespace rtamd

extern "C" int rt_scene_upload(rt_ctx* ctx, const rt_scene_desc* desc, rt_scene** out) {
    if (!ctx) return RT_ERR_INVALID;
    if (!desc || !out) return set_error(ctx, RT_ERR_INVALID, "rt_scene_upload: null argument");
    *out = nullptr;
    PreparedScene prep;
    int rc = scene_prepare(ctx, desc, prep);
    if (rc) return rc;
    return scene_commit(ctx, prep, out);
}

extern "C" int rt_scene_check(const rt_scene_desc* desc, rt_scene_info* info, char* err, size_t errlen) {
    if (err && errlen) err[0] = 0;
    if (!desc || !info) return RT_ERR_INVALID;
    SceneCheck chk;
    try {
        chk = check_scene(*desc);
    } catch (const std::exception&) {
        return RT_ERR_NOMEM;
    }
    if (!chk.error.empty()) {
        if (err && errlen) {
            std::strncpy(err, chk.error.c_str(), errlen - 1);
            err[errlen - 1] = 0;
        }
        return RT_ERR_INVALID;
    }
    bool fast = chk.standard && chk.enclosing;
    if (fast) {
        PackedTree packed;
        pack_tree(*desc, packed);
        if (packed.height > kMaxPackedHeight) fast = false;
    }
    std::memset(info, 0, sizeof *info);
    info->num_triangles = desc->num_triangles;
    info->num_nodes = desc->num_nodes;
    info->num_lights = desc->num_lights;
    info->num_spheres = desc->num_spheres;
    info->num_materials = desc->num_materials;
    info->tree_height = chk.height;
    info->reachable_leaves = chk.leaves;
    info->enclosing = chk.enclosing ? 1u : 0u;
    info->fast_path = fast ? 1u : 0u;
    return RT_OK;
}

extern "C" int rt_scene_info_get(const rt_scene* scene, rt_scene_info* info) {
    if (!scene || !info) return RT_ERR_INVALID;
    *info = scene->info;
    return RT_OK;
}

extern "C" int rt_scene_release(rt_ctx* ctx, rt_scene* scene) {
    if (!ctx) return RT_ERR_INVALID;
    if (!scene) return RT_OK;
    if (scene->ctx != ctx) return set_error(ctx, RT_ERR_INVALID, "rt_scene_release: scene belongs to another context");
    (void)hipSetDevice(ctx->device);
    free_scene(scene);
    return RT_OK;
}
