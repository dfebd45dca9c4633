// rt_sphere.hip -- spheres, as a pass after the triangle traversal.
//
// The shader tests spheres only in hit_scene (ray-trace-compute.comp:222-247),
// which its main path does not call, and the reference lists "include
// spheres in bvh" as not done (README.md:27).  Here spheres are an opt-in
// pass over the hits a traversal kernel wrote, in one of two rules:
//   RT_SPHERES_NEAREST   the nearest of the triangle hit and every sphere hit
//                        (hit_sphere's bound is the closest t so far)
//   RT_SPHERES_SCENE     hit_scene's own rule: closest_so_far is reset to
//                        t_max before the sphere loop (comp:236), so any
//                        sphere hit replaces the triangle hit
// In both, sphere j is tested as hit_sphere does (comp:175-203, tMax
// inclusive), so over the loop the winner is the smallest accepted root and,
// between equal roots, the highest index.  Scenes with more than 16 spheres
// carry a sphere tree (built at upload); its walk applies that same rule
// (smaller root, or equal root and higher index), so it returns what the
// loop returns.  Its box test keeps an interval slack of 2^-10 of the
// current bound on top of the boxes' padding, so a root that the sphere
// test accepts is never culled.
//
// The pass reads a ray (32 B) and a hit (16 B) and may write the hit:
// about 64 B per ray beside the kilobytes of the traversal it follows.
#include "rt_common.hpp"
#include "rt_trace_dev.hpp"

namespace rtamd {
namespace {

constexpr int kSphereBlock = 256;

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sphere_pass(const rt_ray* __restrict__ rays, uint32_t n_host,
                                                       const uint32_t* __restrict__ n_dev, rt_hit* __restrict__ hits,
                                                       const rt_sphere* __restrict__ spheres, uint32_t nspheres,
                                                       const SphereNode* __restrict__ nodes, uint32_t ntri, int mode) {
    __shared__ int32_t stk[kSphereStack * BLOCK];
    const uint32_t n = n_dev ? *n_dev : n_host;
    for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        const RayIn r = load_ray(rays, i);
        const float4 h = reinterpret_cast<const float4*>(hits)[i];
        const int32_t id = __float_as_int(h.x);
        float best = (mode == RT_SPHERES_NEAREST && id >= 0) ? h.y : r.tmax;
        int32_t bj = -1;
        int bback = 0;
        if (!nodes) {
            for (uint32_t j = 0; j < nspheres; ++j) {
                float root;
                int back;
                if (sphere_test(spheres[j], r.o, r.d, r.tmin, best, root, back)) {
                    best = root;
                    bj = (int32_t)j;
                    bback = back;
                }
            }
        } else {
            const f3 inv{safe_inv(r.d.x), safe_inv(r.d.y), safe_inv(r.d.z)};
            uint32_t sp = 0;
            stk[threadIdx.x] = 0;
            sp = 1;
            while (sp > 0) {
                --sp;
                const SphereNode nd = nodes[stk[sp * BLOCK + threadIdx.x]];
                const float slack = 0x1p-10f * fabsf(best);
                float tn;
                if (!slab(nd.lo[0], nd.lo[1], nd.lo[2], nd.hi[0], nd.hi[1], nd.hi[2], r.o, inv, r.tmin - slack,
                          best + slack, tn))
                    continue;
                if (nd.left < 0) {
                    const int32_t j = -1 - nd.left;
                    float root;
                    int back;
                    if (sphere_test(spheres[j], r.o, r.d, r.tmin, best, root, back) && (root < best || j > bj)) {
                        best = root;
                        bj = j;
                        bback = back;
                    }
                } else if (sp + 2 <= kSphereStack) {   // depth <= 24 keeps this true
                    stk[sp * BLOCK + threadIdx.x] = nd.right;
                    ++sp;
                    stk[sp * BLOCK + threadIdx.x] = nd.left;
                    ++sp;
                }
            }
        }
        if (bj >= 0) store_hit(hits, i, (int32_t)(ntri + (uint32_t)bj), best, (float)bback, 0.0f);
    }
}

}  // namespace

int launch_sphere_pass(rt_ctx* ctx, const TraceArgs& a, hipStream_t s) {
    const rt_scene* sc = a.scene;
    const uint32_t ns = sc->info.num_spheres;
    if (!a.spheres || ns == 0) return RT_OK;
    uint32_t blocks = (uint32_t)ctx->num_cus * 8u;
    if (!a.n_dev) {
        const uint32_t need = (a.n + kSphereBlock - 1) / kSphereBlock;
        if (need < blocks) blocks = need;
        if (blocks == 0) return RT_OK;
    }
    hipLaunchKernelGGL((k_sphere_pass<kSphereBlock>), dim3(blocks), dim3(kSphereBlock), 0, s, a.rays, a.n, a.n_dev,
                       a.hits, sc->d_spheres, ns, sc->d_snodes, sc->info.num_triangles, (int)a.spheres);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

}  // namespace rtamd
