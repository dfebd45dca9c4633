// rt_noalloc.hip -- rt_closest_hit_noalloc: the closest-hit query for callers
// that cannot allocate device memory (code under stream capture, serving
// loops and harnesses that forbid hipMalloc).
//
// The query's scratch -- the persistent kernel's ray cursors and the trace
// order's sort (origin keys and ray indices, double buffered, plus the radix
// sort's temporary storage) -- lives in one file-scope device array.  Its
// device address is looked up once per device when a context is created
// (noalloc_setup, called by rt_create) and only that address is handed to
// kernels and launchers: in host code the array's name is a host-side
// stand-in, not a device pointer.
//
// Every call on one device shares that device's array and leaves its
// contents behind.  The cursors are cleared by a fill kernel on the call's
// stream before the traversal reads them, and the sort writes every key and
// index it reads, so nothing else needs to start at zero.  Two calls that
// overlap on one device race on the array: order them on one stream.
#include <mutex>

#include "rt_common.hpp"

namespace rtamd {
namespace {

// Largest query: 2^22 rays (4,194,304), the 4M-ray workload in one call.
constexpr uint32_t kMaxRays = 1u << 22;
// Layout: [0, 4 KiB) cursor block; then the sort's four 2^22-word vectors
// (64 MiB) and 16 MiB for the radix sort's temporary storage.
constexpr size_t kCursorBytes = 4096;
constexpr size_t kOrderBytes = 4 * (size_t)kMaxRays * sizeof(uint32_t) + (16u << 20);
constexpr size_t kScratchBytes = kCursorBytes + kOrderBytes;
static_assert(kCursorBlock * sizeof(uint32_t) <= kCursorBytes, "cursor block fits");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};         // scratch's device address, per device
std::string g_err[kMaxDevices];        // why the lookup failed, if it did

char* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

}  // namespace

int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return RT_OK;   // the variant then refuses calls
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return RT_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) {
        g_err[device] = "device not current at set-up";
        return RT_OK;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        (void)hipGetLastError();
        return RT_OK;
    }
    g_ptr[device] = (char*)p;
    return RT_OK;
}

}  // namespace rtamd

using namespace rtamd;

extern "C" int rt_closest_hit_noalloc(rt_ctx* ctx, const rt_scene* scene, const rt_ray* rays_dev, uint32_t n,
                                      rt_hit* hits_dev, const rt_query_params* params, void* stream) {
    if (!ctx) return RT_ERR_INVALID;
    if (!scene_usable(ctx, scene))
        return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit_noalloc: scene not on this context's device");
    if (n == 0) return RT_OK;
    if (n > kMaxRays)
        return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit_noalloc: " + std::to_string(n) +
                                                  " rays, more than the 4,194,304 its scratch holds; split the batch");
    if (!rays_dev || !hits_dev) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit_noalloc: null ray or hit buffer");
    if (((uintptr_t)rays_dev & 15u) || ((uintptr_t)hits_dev & 15u))
        return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit_noalloc: buffers must be 16-byte aligned");
    rt_query_params p{};
    if (params) p = *params;
    if (p.traversal > RT_TRAVERSE_REFERENCE) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit_noalloc: unknown traversal");
    if (p.spheres > RT_SPHERES_SCENE) return set_error(ctx, RT_ERR_INVALID, "rt_closest_hit_noalloc: unknown sphere rule");
    // the shader-order traversal keeps a stack per lane of the whole grid in
    // a workspace sized by the tree's height: not offered without allocation
    if (p.traversal == RT_TRAVERSE_REFERENCE || p.max_stack)
        return set_error(ctx, RT_ERR_UNSUPPORTED, "rt_closest_hit_noalloc: only the packed traversal runs without allocation");
    if (!scene->fast)
        return set_error(ctx, RT_ERR_UNSUPPORTED,
                         "rt_closest_hit_noalloc: the scene's tree does not enclose its triangles (shader-order traversal only)");
    char* base = scratch_of(ctx->device);
    if (!base) {
        std::lock_guard<std::mutex> lk(g_mu);
        const int d = ctx->device;
        return set_error(ctx, RT_ERR_UNSUPPORTED, "rt_closest_hit_noalloc: no scratch on this device" +
                                                      (d >= 0 && d < kMaxDevices && !g_err[d].empty() ? ": " + g_err[d]
                                                                                                       : std::string()));
    }
    RT_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    TraceArgs a{scene, rays_dev, n, nullptr, hits_dev, p.spheres, nullptr, nullptr};
    a.cursor = reinterpret_cast<uint32_t*>(base);
    a.no_alloc = true;
    ctx->last_query_timed = false;
    int rc = ray_order_in(ctx, a, s, base + kCursorBytes, kOrderBytes, &a.order);
    if (rc) return rc;
    ctx->last_query_sorted = a.order != nullptr;
    rc = launch_trace_packed(ctx, a, s);
    if (rc) return rc;
    return launch_sphere_pass(ctx, a, s);
}
