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
//  6. boxes: leaf records, a 64-ary pyramid of box unions over the sorted
//     leaves, then every internal node unions the pyramid entries that tile
//     its leaf range (no thread waits on another, so no cross-XCD cache
//     maintenance is needed)
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
                                                    rt_bvh_node* __restrict__ nodes, int32_t* __restrict__ far) {
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
        far[i] = (int32_t)j;                  // node i covers sorted leaves [min(i,j), max(i,j)]
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

// Leaf records: the leaf of sorted position k is node n-1+k.
__global__ __launch_bounds__(256) void k_leaves(const rt_triangle* __restrict__ tris, const uint32_t* __restrict__ order,
                                                uint32_t n, rt_bvh_node* __restrict__ nodes) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint32_t tri = order[k];
        rt_bvh_node* nd = nodes + (n - 1 + k);
        store_box(nd, tri_box(tris, tri));
        nd->_pad0 = 0.0f;
        nd->leftNodeIndex = -1;
        nd->rightNodeIndex = -1;
        nd->objectIndex = (int32_t)tri;
        nd->_pad1[0] = nd->_pad1[1] = 0;
    }
}

// Box pyramid over the sorted leaves: level 0 is the leaf records, entry m
// of level L+1 is the union of entries [64m, 64m+64) of level L.  Every
// internal node's box is then the union of at most ~126 entries per level
// over its leaf range -- no thread waits on another.
constexpr int kPyramidLevels = 6;             // 64^5 > 2^30 leaves
struct PBox {
    float4 lo, hi;
};
struct Pyramid {
    const rt_bvh_node* leaves;                // level 0: nodes + n - 1
    const PBox* lvl[kPyramidLevels];          // levels 1.. (lvl[0] unused)
    uint32_t count[kPyramidLevels];
    uint32_t levels;                          // levels present, level 0 included
};

__device__ inline Box pyramid_get(const Pyramid& p, uint32_t level, uint32_t k) {
    Box b;
    if (level == 0) {
        const float4* q = reinterpret_cast<const float4*>(p.leaves + k);
        const float4 lo = q[0], hi = q[1];
        b.lo[0] = lo.x; b.lo[1] = lo.y; b.lo[2] = lo.z;
        b.hi[0] = hi.x; b.hi[1] = hi.y; b.hi[2] = hi.z;
    } else {
        const PBox e = p.lvl[level][k];
        b.lo[0] = e.lo.x; b.lo[1] = e.lo.y; b.lo[2] = e.lo.z;
        b.hi[0] = e.hi.x; b.hi[1] = e.hi.y; b.hi[2] = e.hi.z;
    }
    return b;
}

__device__ inline void grow(Box& u, const Box& b) {
    for (int a = 0; a < 3; ++a) {
        u.lo[a] = fmin_(u.lo[a], b.lo[a]);
        u.hi[a] = fmax_(u.hi[a], b.hi[a]);
    }
}

// One wave per 64-entry block of level `level` -> entry of level+1.
__global__ __launch_bounds__(256) void k_pyramid(Pyramid p, uint32_t level, PBox* __restrict__ out) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t blocks = p.count[level + 1];
    for (uint32_t m = (blockIdx.x * blockDim.x + threadIdx.x) / 64u; m < blocks; m += gridDim.x * blockDim.x / 64u) {
        const uint32_t k = m * 64u + lane;
        Box b;
        if (k < p.count[level]) {
            b = pyramid_get(p, level, k);
        } else {
            for (int a = 0; a < 3; ++a) { b.lo[a] = INFINITY; b.hi[a] = -INFINITY; }
        }
        for (int a = 0; a < 3; ++a) {
            b.lo[a] = wave_min(b.lo[a]);
            b.hi[a] = wave_max(b.hi[a]);
        }
        if (lane == 0) out[m] = PBox{make_float4(b.lo[0], b.lo[1], b.lo[2], 0.0f), make_float4(b.hi[0], b.hi[1], b.hi[2], 0.0f)};
    }
}

// Internal node i's box: the union over its sorted leaf range.
__global__ __launch_bounds__(256) void k_internal_boxes(Pyramid p, uint32_t n, const int32_t* __restrict__ far,
                                                        rt_bvh_node* __restrict__ nodes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i + 1 < n; i += gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)far[i];
        uint32_t lo = i < j ? i : j, hi = (i < j ? j : i) + 1;    // [lo, hi) at level 0
        Box u;
        for (int a = 0; a < 3; ++a) { u.lo[a] = INFINITY; u.hi[a] = -INFINITY; }
        for (uint32_t level = 0; lo < hi; ++level) {
            if (level + 1 == p.levels || hi - lo <= 64u) {
                for (uint32_t k = lo; k < hi; ++k) grow(u, pyramid_get(p, level, k));
                break;
            }
            const uint32_t a = min(hi, (lo + 63u) & ~63u);          // head up to a block boundary
            const uint32_t b = max(a, hi & ~63u);                    // tail from the last boundary
            for (uint32_t k = lo; k < a; ++k) grow(u, pyramid_get(p, level, k));
            for (uint32_t k = b; k < hi; ++k) grow(u, pyramid_get(p, level, k));
            lo = a / 64u;                                           // whole blocks [a, b) one level up
            hi = b / 64u;
        }
        store_box(nodes + i, u);
    }
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
    // pyramid level sizes
    Pyramid pyr{};
    pyr.count[0] = n;
    pyr.levels = 1;
    while (pyr.levels < (uint32_t)kPyramidLevels && pyr.count[pyr.levels - 1] > 1) {
        pyr.count[pyr.levels] = (pyr.count[pyr.levels - 1] + 63u) / 64u;
        ++pyr.levels;
    }
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += align256(bytes); return o; };
    const size_t o_partial = take(sizeof(float) * 6 * kBoundsBlocks);
    const size_t o_bounds = take(sizeof(float) * 6);
    const size_t o_keys = take(sizeof(uint64_t) * n);
    const size_t o_keys2 = take(sizeof(uint64_t) * n);
    const size_t o_vals = take(sizeof(uint32_t) * n);
    const size_t o_vals2 = take(sizeof(uint32_t) * n);
    const size_t o_far = take(sizeof(int32_t) * n);
    size_t o_lvl[kPyramidLevels] = {};
    for (uint32_t L = 1; L < pyr.levels; ++L) o_lvl[L] = take(sizeof(PBox) * pyr.count[L]);
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
    int32_t* far = (int32_t*)(ws + o_far);
    void* sort_tmp = ws + o_sort;
    pyr.leaves = nodes_dev + (n - 1);
    for (uint32_t L = 1; L < pyr.levels; ++L) pyr.lvl[L] = (const PBox*)(ws + o_lvl[L]);

    const uint32_t cap = (uint32_t)ctx->num_cus * 16u;
    const uint32_t nb = grid(n, kBoundsBlocks);
    hipLaunchKernelGGL(k_centre_bounds, dim3(nb), dim3(256), 0, s, tris_dev, n, partial);
    RT_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_fold_bounds, dim3(1), dim3(256), 0, s, partial, nb, bounds);
    RT_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_morton, dim3(grid(n, cap)), dim3(256), 0, s, tris_dev, n, bounds, keys, vals);
    RT_HIP(ctx, hipGetLastError());
    RT_HIP(ctx, rocprim::radix_sort_pairs(sort_tmp, sort_bytes, keys, keys2, vals, vals2, n, 0u, 63u, s));
    if (n > 1) {
        hipLaunchKernelGGL(k_radix_tree, dim3(grid(n - 1, cap)), dim3(256), 0, s, keys2, n, nodes_dev, far);
        RT_HIP(ctx, hipGetLastError());
    }
    hipLaunchKernelGGL(k_leaves, dim3(grid(n, cap)), dim3(256), 0, s, tris_dev, vals2, n, nodes_dev);
    RT_HIP(ctx, hipGetLastError());
    for (uint32_t L = 0; L + 1 < pyr.levels; ++L) {
        // one wave (64 threads) per output entry
        hipLaunchKernelGGL(k_pyramid, dim3(grid((uint64_t)pyr.count[L + 1] * 64u, cap)), dim3(256), 0, s, pyr, L,
                           (PBox*)(ws + o_lvl[L + 1]));
        RT_HIP(ctx, hipGetLastError());
    }
    if (n > 1) {
        hipLaunchKernelGGL(k_internal_boxes, dim3(grid(n - 1, cap)), dim3(256), 0, s, pyr, n, far, nodes_dev);
        RT_HIP(ctx, hipGetLastError());
    }
    return RT_OK;
}
