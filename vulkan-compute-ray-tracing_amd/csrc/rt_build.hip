// rt_build.hip -- GPU BVH builder.  A linear BVH after T. Karras,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d
// Trees" (HPG 2012), written straight into the reference's flat node layout
// (BvhNode of src/ray-tracing/GpuModels.h: min, max, leftNodeIndex,
// rightNodeIndex, objectIndex; the shader starts at node 0).  The reference
// builds on the host (Bvh.h:141-209, random axis, median split); this one
// is the device-side alternative for scenes that arrive at run time.
//
// Steps, all on one stream:
//  1. object box of each triangle as the reference builder makes it
//     (Bvh.h:96-97: bounds widened by 0.0001) and its centre
//     c = (lo + hi) * 0.5
//  2. bounds [lo, hi] of the centres (per-block partials, then one block)
//  3. 63-bit Morton code: per axis q = (uint32)min((c - lo) * s, 2^21 - 1)
//     with s = 2^21 / (hi - lo) (s = 0 when hi == lo); bit 3b+2 is x's bit
//     b, 3b+1 y's, 3b z's
//  4. stable radix sort of (code, triangle index) pairs (rocPRIM)
//  5. radix tree: delta(i, j) = common prefix length of the codes at sorted
//     positions i and j, 64 + that of the positions when the codes are
//     equal; internal node i covers the sorted range with one end at i and
//     splits it where the prefix shared with i ends
//  6. boxes bottom-up: every leaf's thread climbs, stopping at a parent it
//     reaches first; the second child to arrive writes the parent's box
// Numbering: internal node k -> node k, the leaf of sorted position k ->
// node n-1+k.  One triangle gives a single leaf at node 0.
#include <rocprim/device/device_radix_sort.hpp>

#include "rt_common.hpp"

namespace rtamd {
namespace {

constexpr int kBoundsBlocks = 1024;

__device__ inline float fmin_(float a, float b) { return b < a ? b : a; }
__device__ inline float fmax_(float a, float b) { return a < b ? b : a; }

struct Box {
    float lo[3], hi[3];
};

// The reference's object box of a triangle (Bvh.h:96-97): its bounds
// widened by eps = 0.0001 on every side, so leaf boxes strictly enclose
// their triangles, flat ones included.
__device__ inline Box tri_box(const rt_triangle* __restrict__ tris, uint32_t i) {
    const float4* p = reinterpret_cast<const float4*>(tris + i);
    const float4 a = p[0], b = p[1], c = p[2];
    const float eps = 0.0001f;
    Box r;
    r.lo[0] = fmin_(fmin_(a.x, b.x), c.x) - eps;
    r.lo[1] = fmin_(fmin_(a.y, b.y), c.y) - eps;
    r.lo[2] = fmin_(fmin_(a.z, b.z), c.z) - eps;
    r.hi[0] = fmax_(fmax_(a.x, b.x), c.x) + eps;
    r.hi[1] = fmax_(fmax_(a.y, b.y), c.y) + eps;
    r.hi[2] = fmax_(fmax_(a.z, b.z), c.z) + eps;
    return r;
}

__device__ inline float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin_(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax_(v, __shfl_xor(v, o, 64));
    return v;
}

// Per-block bounds of the box centres: partial[block * 6 + {0..2 lo, 3..5 hi}].
__global__ __launch_bounds__(256) void k_centre_bounds(const rt_triangle* __restrict__ tris, uint32_t n,
                                                       float* __restrict__ partial) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Box b = tri_box(tris, i);
        for (int a = 0; a < 3; ++a) {
            const float c = (b.lo[a] + b.hi[a]) * 0.5f;
            lo[a] = fmin_(lo[a], c);
            hi[a] = fmax_(hi[a], c);
        }
    }
    __shared__ float red[4][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            red[wave][a] = lo[a];
            red[wave][3 + a] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = red[0][threadIdx.x];
        for (int w = 1; w < 4; ++w) v = threadIdx.x < 3 ? fmin_(v, red[w][threadIdx.x]) : fmax_(v, red[w][threadIdx.x]);
        partial[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

// Folds `blocks` partials into bounds[0..5] (one block of 256 threads).
__global__ __launch_bounds__(256) void k_fold_bounds(const float* __restrict__ partial, uint32_t blocks,
                                                     float* __restrict__ bounds) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x)
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin_(lo[a], partial[b * 6 + a]);
            hi[a] = fmax_(hi[a], partial[b * 6 + 3 + a]);
        }
    __shared__ float red[4][6];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int a = 0; a < 3; ++a) {
        lo[a] = wave_min(lo[a]);
        hi[a] = wave_max(hi[a]);
    }
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            red[wave][a] = lo[a];
            red[wave][3 + a] = hi[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = red[0][threadIdx.x];
        for (int w = 1; w < 4; ++w) v = threadIdx.x < 3 ? fmin_(v, red[w][threadIdx.x]) : fmax_(v, red[w][threadIdx.x]);
        bounds[threadIdx.x] = v;
    }
}

// Spreads the low 21 bits of v to every third bit (bit b -> bit 3b).
__device__ inline uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x001f00000000ffffull;
    x = (x | (x << 16)) & 0x001f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__device__ inline uint32_t quant(float c, float lo, float s) {
    const float q = (c - lo) * s;
    // q >= 0 (lo is the minimum) unless c is NaN; the upper clamp catches
    // q == 2^21 at hi, the lower one maps NaN to 0
    return q >= 2097151.0f ? 2097151u : (q > 0.0f ? (uint32_t)q : 0u);
}

__global__ __launch_bounds__(256) void k_morton(const rt_triangle* __restrict__ tris, uint32_t n,
                                                const float* __restrict__ bounds, uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals) {
    const float lo[3] = {bounds[0], bounds[1], bounds[2]};
    float s[3];
    for (int a = 0; a < 3; ++a) {
        const float ext = bounds[3 + a] - bounds[a];
        const float sv = ext > 0.0f ? 2097152.0f / ext : 0.0f;
        s[a] = sv <= 3.0e38f ? sv : 0.0f;    // a subnormal extent would give inf
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const Box b = tri_box(tris, i);
        uint32_t q[3];
        for (int a = 0; a < 3; ++a) q[a] = quant((b.lo[a] + b.hi[a]) * 0.5f, lo[a], s[a]);
        keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
        vals[i] = i;
    }
}

// Common prefix of sorted positions i and j (j may be out of range: -1).
__device__ inline int delta(const uint64_t* __restrict__ keys, int64_t n, int64_t i, int64_t j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = keys[i], b = keys[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

__global__ __launch_bounds__(256) void k_radix_tree(const uint64_t* __restrict__ keys, uint32_t n,
                                                    rt_bvh_node* __restrict__ nodes, int32_t* __restrict__ parent) {
    const int64_t N = n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N - 1; i += (int64_t)gridDim.x * blockDim.x) {
        const int d = delta(keys, N, i, i + 1) - delta(keys, N, i, i - 1) >= 0 ? 1 : -1;
        const int dmin = delta(keys, N, i, i - d);
        int64_t lmax = 2;
        while (delta(keys, N, i, i + lmax * d) > dmin) lmax *= 2;
        int64_t l = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(keys, N, i, i + (l + t) * d) > dmin) l += t;
        const int64_t j = i + l * d;
        const int dnode = delta(keys, N, i, j);
        int64_t s = 0;
        for (int64_t div = 2;; div *= 2) {
            const int64_t t = (l + div - 1) / div;
            if (delta(keys, N, i, i + (s + t) * d) > dnode) s += t;
            if (t == 1) break;
        }
        const int64_t g = i + s * d + (d < 0 ? -1 : 0);
        const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        const int32_t left = (int32_t)(lo == g ? (N - 1 + g) : g);
        const int32_t right = (int32_t)(hi == g + 1 ? (N - 1 + g + 1) : g + 1);
        rt_bvh_node& nd = nodes[i];
        nd._pad0 = 0.0f;
        nd.leftNodeIndex = left;
        nd.rightNodeIndex = right;
        nd.objectIndex = -1;
        nd._pad1[0] = nd._pad1[1] = 0;
        parent[left] = (int32_t)i;
        parent[right] = (int32_t)i;
        if (i == 0) parent[0] = -1;
    }
}

__device__ inline void store_box(rt_bvh_node* nd, const Box& b) {
    nd->min[0] = b.lo[0];
    nd->min[1] = b.lo[1];
    nd->min[2] = b.lo[2];
    nd->max[0] = b.hi[0];
    nd->max[1] = b.hi[1];
    nd->max[2] = b.hi[2];
}

__device__ inline Box load_box(const rt_bvh_node* nd) {
    Box b;
    for (int a = 0; a < 3; ++a) {
        b.lo[a] = nd->min[a];
        b.hi[a] = nd->max[a];
    }
    return b;
}

// Leaves, then the climb.  `arrivals` holds one zeroed counter per internal node.
__global__ __launch_bounds__(256) void k_boxes(const rt_triangle* __restrict__ tris, const uint32_t* __restrict__ order,
                                               uint32_t n, rt_bvh_node* nodes, const int32_t* __restrict__ parent,
                                               uint32_t* arrivals) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t tri = order[k];
        const int32_t leaf = (int32_t)(n - 1 + k);
        Box b = tri_box(tris, tri);
        rt_bvh_node* nd = nodes + leaf;
        store_box(nd, b);
        nd->_pad0 = 0.0f;
        nd->leftNodeIndex = -1;
        nd->rightNodeIndex = -1;
        nd->objectIndex = (int32_t)tri;
        nd->_pad1[0] = nd->_pad1[1] = 0;
        // a radix tree over 95-bit keys is at most 96 levels deep; the step
        // bound and the index checks only guard against a malformed tree
        const uint32_t n_nodes = 2 * n - 1;
        int32_t p = n > 1 ? parent[leaf] : -1;
        for (int steps = 0; p >= 0 && (uint32_t)p < n - 1 && steps < 128; ++steps) {
            // release this subtree's boxes, acquire the sibling's
            const uint32_t before =
                __hip_atomic_fetch_add(&arrivals[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (before == 0) break;           // the sibling's thread finishes this node
            rt_bvh_node* pn = nodes + p;
            const int32_t li = pn->leftNodeIndex, ri = pn->rightNodeIndex;
            if ((uint32_t)li >= n_nodes || (uint32_t)ri >= n_nodes) break;
            const Box l = load_box(nodes + li), r = load_box(nodes + ri);
            Box u;
            for (int a = 0; a < 3; ++a) {
                u.lo[a] = fmin_(l.lo[a], r.lo[a]);
                u.hi[a] = fmax_(l.hi[a], r.hi[a]);
            }
            store_box(pn, u);
            p = parent[p];
        }
    }
}

__global__ void k_fill_words(uint32_t* p, uint32_t value, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = value;
}

uint32_t grid(uint64_t work, uint32_t cap) {
    uint64_t g = (work + 255) / 256;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace rtamd

using namespace rtamd;

extern "C" int rt_bvh_build_device(rt_ctx* ctx, const rt_triangle* tris_dev, uint32_t n, rt_bvh_node* nodes_dev,
                                   void* stream) {
    if (!ctx) return RT_ERR_INVALID;
    if (n == 0 || !tris_dev || !nodes_dev) return set_error(ctx, RT_ERR_INVALID, "rt_bvh_build_device: empty input");
    if (n > (1u << 30)) return set_error(ctx, RT_ERR_INVALID, "rt_bvh_build_device: more than 2^30 triangles");
    if (((uintptr_t)tris_dev & 15u) || ((uintptr_t)nodes_dev & 15u))
        return set_error(ctx, RT_ERR_INVALID, "rt_bvh_build_device: buffers must be 16-byte aligned");
    RT_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;

    size_t sort_bytes = 0;
    RT_HIP(ctx, rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                          (uint32_t*)nullptr, (uint32_t*)nullptr, n, 0u, 63u, s));
    const size_t n_int = n > 1 ? n - 1 : 1;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    const size_t o_partial = take(sizeof(float) * 6 * kBoundsBlocks);
    const size_t o_bounds = take(sizeof(float) * 6);
    const size_t o_keys = take(sizeof(uint64_t) * n);
    const size_t o_keys2 = take(sizeof(uint64_t) * n);
    const size_t o_vals = take(sizeof(uint32_t) * n);
    const size_t o_vals2 = take(sizeof(uint32_t) * n);
    const size_t o_parent = take(sizeof(int32_t) * (2 * (size_t)n - 1));
    const size_t o_arrive = take(sizeof(uint32_t) * n_int);
    const size_t o_sort = take(sort_bytes);
    int rc = ensure_workspace(ctx, kWsBuild, off);
    if (rc) return rc;
    char* ws = (char*)ctx->ws[kWsBuild];
    float* partial = (float*)(ws + o_partial);
    float* bounds = (float*)(ws + o_bounds);
    uint64_t* keys = (uint64_t*)(ws + o_keys);
    uint64_t* keys2 = (uint64_t*)(ws + o_keys2);
    uint32_t* vals = (uint32_t*)(ws + o_vals);
    uint32_t* vals2 = (uint32_t*)(ws + o_vals2);
    int32_t* parent = (int32_t*)(ws + o_parent);
    uint32_t* arrive = (uint32_t*)(ws + o_arrive);
    void* sort_tmp = ws + o_sort;

    const uint32_t cap = (uint32_t)ctx->num_cus * 16u;
    const uint32_t nb = grid(n, kBoundsBlocks);
    hipLaunchKernelGGL(k_centre_bounds, dim3(nb), dim3(256), 0, s, tris_dev, n, partial);
    RT_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_fold_bounds, dim3(1), dim3(256), 0, s, partial, nb, bounds);
    RT_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_morton, dim3(grid(n, cap)), dim3(256), 0, s, tris_dev, n, bounds, keys, vals);
    RT_HIP(ctx, hipGetLastError());
    RT_HIP(ctx, rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, keys2, vals, vals2, n, 0u, 63u, s));
    hipLaunchKernelGGL(k_fill_words, dim3(grid(n_int, cap)), dim3(256), 0, s, arrive, 0u, (uint32_t)n_int);
    RT_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_fill_words, dim3(grid(2 * (uint64_t)n - 1, cap)), dim3(256), 0, s, (uint32_t*)parent,
                       0xffffffffu, 2 * n - 1);
    RT_HIP(ctx, hipGetLastError());
    if (n > 1) {
        hipLaunchKernelGGL(k_radix_tree, dim3(grid(n - 1, cap)), dim3(256), 0, s, keys2, n, nodes_dev, parent);
        RT_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_boxes, dim3(grid(n, cap)), dim3(256), 0, s, tris_dev, vals2, n, nodes_dev, parent, arrive);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}
