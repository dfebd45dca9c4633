// rt_common.hpp -- internal declarations of librtamd: packed device
// layouts, the context and scene objects, error plumbing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "rtamd.h"

// The C records must keep the reference's std430 layout (GpuModels.h:26-63).
static_assert(sizeof(rt_triangle) == 48, "Triangle is 48 bytes");
static_assert(offsetof(rt_triangle, v1) == 16 && offsetof(rt_triangle, v2) == 32, "Triangle vec3 slots");
static_assert(offsetof(rt_triangle, materialIndex) == 44, "Triangle.materialIndex @44");
static_assert(sizeof(rt_material) == 32 && offsetof(rt_material, albedo) == 16, "Material");
static_assert(sizeof(rt_bvh_node) == 48, "BvhNode is 48 bytes");
static_assert(offsetof(rt_bvh_node, max) == 16 && offsetof(rt_bvh_node, leftNodeIndex) == 28, "BvhNode min/max/left");
static_assert(offsetof(rt_bvh_node, rightNodeIndex) == 32 && offsetof(rt_bvh_node, objectIndex) == 36, "BvhNode right/object");
static_assert(sizeof(rt_light) == 8, "Light");
static_assert(sizeof(rt_sphere) == 32 && offsetof(rt_sphere, materialIndex) == 16, "Sphere");
static_assert(sizeof(rt_ubo) == 32 && offsetof(rt_ubo, currentSample) == 16, "UBO");
static_assert(sizeof(rt_ray) == 32 && sizeof(rt_hit) == 16, "ray/hit");

namespace rtamd {

// ---------------------------------------------------------------------------
// Packed traversal records.
//
// PNode (64 bytes, four 16-byte loads): an internal node of the binary tree
// holding BOTH children's boxes, so one fetch decides which children to
// enter.  ref >= 0 names another PNode; ref < 0 names a leaf whose triangle
// sits in PTri slot ~ref.
//   a = (c0.lo.x, c0.lo.y, c0.lo.z, c0.hi.x)
//   b = (c0.hi.y, c0.hi.z, c1.lo.x, c1.lo.y)
//   c = (c1.lo.z, c1.hi.x, c1.hi.y, c1.hi.z)
//   d = (ref0, ref1, 0, 0)
// PTri (48 bytes, three 16-byte loads): the values triIntersect derives from
// the triangle alone, computed once with the same float operations
// (a = v0-v1, b = v2-v0, n = cross(b, a), ray-trace-compute.comp:159-162):
//   a = (v0.x, v0.y, v0.z, e.x)   e = v0 - v1
//   b = (e.y, e.z, f.x, f.y)      f = v2 - v0
//   c = (f.z, n.x, n.y, n.z)      n = cross(f, e)
// ---------------------------------------------------------------------------
struct alignas(16) PNode { float4 a, b, c; int4 d; };
struct alignas(16) PTri { float4 a, b, c; };
static_assert(sizeof(PNode) == 64 && sizeof(PTri) == 48, "packed records");

// PNodeQ (32 bytes, two 16-byte loads): the same two-children record with the
// box coordinates as 16-bit fixed point over the scene's node bounds,
// world = origin + q * scale per axis.  Lower bounds are rounded down and
// upper bounds up, each by one extra quantum, so a quantized box always
// contains the exact one and the test stays conservative.
//   w[0] = lo0.x | lo0.y << 16    w[3] = lo1.x | lo1.y << 16
//   w[1] = lo0.z | hi0.x << 16    w[4] = lo1.z | hi1.x << 16
//   w[2] = hi0.y | hi0.z << 16    w[5] = hi1.y | hi1.z << 16
//   ref0, ref1 as in PNode
struct alignas(16) PNodeQ { uint32_t w[6]; int32_t ref0, ref1; };
static_assert(sizeof(PNodeQ) == 32, "quantized node record");

// PNode4 (64 bytes, four 16-byte loads): the binary tree collapsed two
// levels at a time -- the children of a node's children, up to four, with
// 16-bit boxes in the same fixed-point frame as PNodeQ.  Child k's box takes
// w[3k..3k+2], one word per axis with that axis' two planes:
//   w[3k] = lo.x | hi.x << 16   w[3k+1] = lo.y | hi.y << 16   w[3k+2] = lo.z | hi.z << 16
// ref[k] as in PNode, kNoChild for an unused slot.  One record (one 64-byte
// line) decides two tree levels.
struct alignas(16) PNode4 { uint32_t w[12]; int32_t ref[4]; };
static_assert(sizeof(PNode4) == 64, "4-wide node record");
constexpr int32_t kNoChild = INT32_MIN + 1;

// Fixed-point frame of the quantized nodes, with 1/scale for the rays.
struct QFrame { float ox, oy, oz, sx, sy, sz, rx, ry, rz; };

// kLayoutW4Joined is kLayoutW4 with its records and packed triangles in one
// allocation (rt_scene::tri_off), addressed by 32-bit offsets.
enum { kLayoutF32 = 1, kLayoutQ16 = 2, kLayoutW4 = 3, kLayoutW4Joined = 4 };

// One item of a batched launch: a resident scene's packed tree (4-wide
// records, and the binary 16-bit ones when the launch reads those) and a ray
// batch with its hit buffer.  Global ray g of the launch belongs to the item
// with base <= g < base + count.
struct BatchEntry {
    const PNode4* nodes;
    const PTri* tris;
    const int32_t* tri_id;
    const rt_ray* rays;
    rt_hit* hits;
    const PNodeQ* qnodes;         // binary 16-bit records, or null
    QFrame frame;
    int32_t root;                 // root of the 4-wide records
    int32_t root_q;               // root of the binary records
    uint32_t base;
    uint32_t count;
    uint32_t win0;                // regrouping: the item's first window
};
struct BatchIo {
    const BatchEntry* table;
    const uint32_t* offsets;      // n_items + 1 prefix sums of counts
    const uint32_t* chunk_item;   // item of global ray 64*c
    const uint32_t* perm;         // optional: global ray -> ray index inside its item
};


constexpr int32_t kEmptyRoot = INT32_MIN;  // tree with no reachable triangle
constexpr uint32_t kSeparateArrays = 0xffffffffu;   // rt_scene::tri_off: not one allocation

// Bytes each traversal step reads, used for the roofline accounting.
constexpr uint32_t kNodeBytes = 64;
constexpr uint32_t kTriBytes = 48;

// Result of repacking a reference node array (host side).
struct PackedTree {
    std::vector<PNode> nodes;
    std::vector<PTri> tris;
    std::vector<int32_t> tri_id;   // slot -> original triangle index
    std::vector<PNodeQ> qnodes;    // the same nodes, 16-bit fixed point
    std::vector<PNode4> w4nodes;   // the same tree, 4-wide
    int32_t root4 = kEmptyRoot;
    uint32_t height4 = 0;          // 4-wide node depth
    QFrame frame{};
    int32_t root = kEmptyRoot;
    uint32_t height = 0;           // packed internal-node depth (stack bound)
};

struct SceneCheck {
    bool standard = true;     // every reachable node is a leaf or an internal node
    bool enclosing = true;    // boxes hold their subtrees (strict at leaves)
    uint32_t height = 0;      // max leaf depth of the reference tree
    uint32_t leaves = 0;
    std::string error;        // non-empty: the scene is invalid
};

// Host helpers (rt_scene.cpp).
SceneCheck check_scene(const rt_scene_desc& d);
void pack_tree(const rt_scene_desc& d, PackedTree& out);
void quantize_nodes(PackedTree& out);
void widen_nodes(PackedTree& out);

}  // namespace rtamd

namespace rtamd {
// Node of a scene's sphere tree (built at upload when a scene holds more
// than kSphereLinear spheres): its box, and two child node indices, or for
// a leaf left = -1 - sphere index and right = -1.
struct SphereNode {
    float lo[3];
    int32_t left;
    float hi[3];
    int32_t right;
};
static_assert(sizeof(SphereNode) == 32, "SphereNode is 32 bytes");
constexpr uint32_t kSphereLinear = 16;     // up to this many: a plain loop
constexpr uint32_t kSphereStack = 32;      // LDS stack entries per lane
constexpr uint32_t kMaxTreeSpheres = 1u << 24;   // median split: depth <= 24

// Host half of an upload: the scene validated and packed, its device bytes
// known before anything is allocated on the device (so a caller with an
// HBM budget can make room first).  scene_commit allocates and copies.
struct PreparedScene {
    rt_scene_desc desc{};               // the caller's arrays (the tree the shader order walks)
    std::vector<rt_bvh_node> built;     // the device-built tree (RT_TREE_DEVICE), or empty
    SceneCheck chk;
    PackedTree packed;
    bool fast = false;
    bool tree_built = false;
    bool sphere_tree = false;
    std::vector<SphereNode> snodes;
    bool with_pnodes = false, with_qnodes = false;   // binary records uploaded now
    uint64_t bytes = 0;                 // device bytes scene_commit allocates
};
int scene_prepare(rt_ctx* ctx, const rt_scene_desc* desc, PreparedScene& out);
int scene_commit(rt_ctx* ctx, PreparedScene& prep, rt_scene** out);
// Device bytes of the 4-wide records and packed triangles as uploaded.
uint64_t packed_record_bytes(const PackedTree& p);
// The binary node records of `layout` (RT_LAYOUT_F32 / RT_LAYOUT_Q16) for a
// scene uploaded without them: rebuilt from the scene's device arrays.
int scene_binary_records(rt_ctx* ctx, rt_scene* s, uint32_t layout);
}  // namespace rtamd

struct rt_scene {
    rt_ctx* ctx = nullptr;
    rt_scene_info info{};
    rt_triangle* d_tris = nullptr;
    rt_material* d_mats = nullptr;
    rt_bvh_node* d_nodes = nullptr;
    rt_light* d_lights = nullptr;
    rt_sphere* d_spheres = nullptr;
    rtamd::PNode* d_pnodes = nullptr;
    rtamd::PNodeQ* d_qnodes = nullptr;
    rtamd::PNode4* d_w4nodes = nullptr;
    rtamd::QFrame frame{};
    int32_t root4 = rtamd::kEmptyRoot;
    uint32_t height4 = 0;
    rtamd::PTri* d_ptris = nullptr;
    // d_w4nodes and d_ptris share one allocation when both fit in 4 GiB:
    // d_ptris = d_w4nodes + tri_off bytes, so a kernel addresses either
    // record by a 32-bit offset from one base.  kSeparateArrays otherwise.
    uint32_t tri_off = rtamd::kSeparateArrays;
    int32_t* d_ptri_id = nullptr;
    int32_t root_ref = rtamd::kEmptyRoot;
    uint32_t packed_height = 0;
    uint32_t ref_height = 0;
    bool fast = false;
    rtamd::SphereNode* d_snodes = nullptr;   // sphere tree, or null: plain loop
    // the packed arrays come from a device-built tree (RT_TREE_DEVICE), not
    // from d_nodes: a later build of the binary records rebuilds that tree
    bool tree_built = false;
    // d_pnodes / d_qnodes are built on the first packed query that needs
    // them (binary layouts); this lock orders those builds
    std::mutex lazy_mu;
};

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int num_cus = 0;
    // grow-only device workspaces: [0] render path state, [1] reference
    // traversal stacks, [2] statistics scratch, [3] traversal stack spill,
    // [4] BVH builder scratch, [5] ray-order keys and permutation
    static constexpr int kWorkspaces = 6;
    void* ws[kWorkspaces] = {};
    size_t ws_bytes[kWorkspaces] = {};
    uint32_t* d_counters = nullptr;   // kRenderCounters words: a render's live rays per bounce
    uint32_t* d_cursors = nullptr;    // ring of persistent-kernel ray cursors
    uint32_t cursor_next = 0;
    rt_tuning tuning{};               // 0 fields = defaults
    uint64_t tuning_gen = 0;          // bumped by every rt_tuning_set
    uint64_t last_primary = 0, last_secondary = 0;
    // opt-in stage timing of renders: events around generate, each bounce's
    // trace and shade, and store (created on the first timed render)
    static constexpr int kRenderEvents = 2 * 63 + 2;
    bool time_render = false;           // stage timing of renders and queries
    hipEvent_t render_ev[kRenderEvents] = {};
    uint32_t last_bounces = 0;
    bool last_timed = false;
    // query stage events: start, ordered, traced, spheres done
    hipEvent_t query_ev[4] = {};
    bool last_query_timed = false;
    bool last_query_sorted = false;
    // opt-in record counting of renders: their traversals run the counting
    // variant and add their node and triangle record totals here
    bool count_render = false;
    bool last_counted = false;
    unsigned long long* d_render_sums = nullptr;
    hipStream_t last_render_stream = nullptr;   // the stream the counted render ran on
};

namespace rtamd {

int set_error(rt_ctx* ctx, int code, const std::string& msg);
// Triangles of one OBJ file (rt_obj.cpp); err receives the reason on failure.
int obj_triangles(const char* path, uint32_t material, std::vector<rt_triangle>& tris, std::string& err);
int hip_error(rt_ctx* ctx, hipError_t e, const char* what);
enum { kWsRender = 0, kWsStack = 1, kWsStats = 2, kWsSpill = 3, kWsBuild = 4, kWsOrder = 5 };
// Persistent launches take their ray cursors from a ring of blocks; a block
// holds kSegments counters kCursorStride words apart (one 128-byte line
// each), one per segment of the ray range.
constexpr uint32_t kXcds = 8;                   // XCDs of the MI355X
constexpr uint32_t kSegments = 8;               // one segment per XCD (32, four per XCD, measured no better)
constexpr uint32_t kSegmentShift = 3;           // log2(kSegments)
static_assert(kSegments == 1u << kSegmentShift, "segments are a power of two");
constexpr uint32_t kCursorStride = 32;
constexpr uint32_t kCursorBlock = kSegments * kCursorStride;
constexpr uint32_t kCursorBlocks = 256;
// A render's per-bounce counters (rt_ctx::d_counters): [b] = live rays entering bounce b.
constexpr uint32_t kRenderCounters = 64;
constexpr uint32_t kCursorSlots = kCursorBlocks * kCursorBlock;
int ensure_workspace(rt_ctx* ctx, int slot, size_t bytes);

#define RT_HIP(ctx, call)                                              \
    do {                                                               \
        hipError_t _e = (call);                                        \
        if (_e != hipSuccess) return ::rtamd::hip_error((ctx), _e, #call); \
    } while (0)

// Launchers (rt_trace.hip).
struct TraceArgs {
    const rt_scene* scene;
    const rt_ray* rays;
    uint32_t n;                 // used when n_dev is null
    const uint32_t* n_dev;      // device-side ray count (render queues)
    rt_hit* hits;
    uint32_t spheres;
    uint32_t* node_visits;      // optional per-ray counters
    uint32_t* tri_tests;
    bool render = false;        // a bounce of the render's wavefront loop
    const uint32_t* order = nullptr;   // packed traversal: trace order, or index order
    uint32_t* cursor = nullptr;        // kCursorBlock words for the ray cursors, or null: ctx's ring
    bool cursor_zeroed = false;        // an earlier kernel on the stream zeroed cursor's segment words
    unsigned long long* sums = nullptr;   // counting variant: record totals (2 words), or null
    bool no_alloc = false;             // a variant that must not allocate: refuse what would
};
int launch_trace_packed(rt_ctx* ctx, const TraceArgs& a, hipStream_t s);
// The next block of the context's ray-cursor ring (kCursorBlock words; the
// traversal reads word k * kCursorStride of segment k).
uint32_t* next_cursor(rt_ctx* ctx);
// The trace order of a packed query (rt_trace.hip): sorts the rays by
// origin cell into ctx's order workspace when ctx->tuning.ray_order asks
// for it (or, by default, for large batches on large scenes); *order stays
// null for index order.
int ray_order(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, const uint32_t** order);
// The same into a caller's workspace of ws_bytes (256-byte aligned); an
// error when the sort does not fit in it.
int ray_order_in(rt_ctx* ctx, const TraceArgs& a, hipStream_t s, char* ws, size_t ws_bytes, const uint32_t** order);
// Workspace bytes ray_order_in needs for n rays (query keys), through
// *bytes; needs the context's device current.
int ray_order_bytes(rt_ctx* ctx, uint32_t n, size_t* bytes);
// Scratch of the allocation-free query (rt_noalloc.hip), set up once per
// device when a context is created.
int noalloc_setup(int device);
// Idle lanes that make a drained wave share its busy lanes' work
// (rt_tuning.steal, default 32).
uint32_t drain_steal(const rt_ctx* ctx);
int resolved_layout(const rt_ctx* ctx);   // RT_LAYOUT_* actually used
int launch_trace_reference(rt_ctx* ctx, const TraceArgs& a, uint32_t max_stack, hipStream_t s);
// Spheres after a traversal (rt_sphere.hip): no-op unless a.spheres is set
// and the scene has spheres.
int launch_sphere_pass(rt_ctx* ctx, const TraceArgs& a, hipStream_t s);
int launch_fill_u32(rt_ctx* ctx, uint32_t* p, uint32_t value, uint32_t count, hipStream_t s);

// Batched traversal from caller-filled device tables (rt_batch.hip): the
// server's per-slot launches.  stack_need = 3 * max 4-wide height + 1;
// `cursor` holds kCursorBlock words (reset by the launch).
int launch_batch_dynamic(rt_ctx* ctx, const BatchIo& bio, uint32_t* cursor, uint32_t total, uint32_t stack_need,
                         uint32_t refill, hipStream_t s);

// A scene is a device resource: any context on its device may trace it.
inline bool scene_usable(const rt_ctx* c, const rt_scene* s);

}  // namespace rtamd

inline bool rtamd::scene_usable(const rt_ctx* c, const rt_scene* s) {
    return c && s && s->ctx && s->ctx->device == c->device;
}
