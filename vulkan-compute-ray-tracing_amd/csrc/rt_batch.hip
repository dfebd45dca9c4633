// rt_batch.hip -- the batched form: many (scene, ray batch) items traced by
// one persistent launch over a global work queue, replayed from a hipGraph,
// with launch parameters chosen by an autotuner.
//
// Work queue: the items' rays are numbered 0..total-1 item after item; a
// lane that needs work takes the next numbers from one cursor (the same
// per-lane refetch as the plain kernel) and finds its item through a
// 64-ray chunk table.  Regrouping (optional) first bins each item's rays --
// by the 16x16x16 origin cell (Morton order) in the scene's box frame, or by
// direction octant and then 8x8x8 origin cell -- within windows of a given
// number of consecutive rays (the regroup interval; the whole item by
// default), so neighbouring lanes start alike and share node fetches; the
// binning permutation only changes which lane traces a ray, never a result.
// The node records read are the 4-wide ones or the binary 16-bit ones.
// Every step is a kernel (counters are cleared by a fill kernel, not a
// memset), so the sequence can be captured into a graph and replayed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "rt_trace_dev.hpp"

#pragma clang fp contract(off)

namespace rtamd {
namespace {

constexpr uint32_t kBins = 4096;         // bins per regroup window: 16^3 origin cells, or 8 octants x 8^3 cells
constexpr uint32_t kChunk = 64;          // rays per chunk-table entry
constexpr uint32_t kMinInterval = 1u << 16;   // smallest regroup window (bounds the histogram)

__device__ __forceinline__ uint32_t item_of(const BatchIo& bio, uint32_t g) {
    uint32_t b = bio.chunk_item[g / kChunk];
    while (g >= bio.offsets[b + 1]) ++b;
    return b;
}

// Bin of every ray in its window, and its rank inside the bin (order of the
// atomics).  mode 1: 16^3 origin cells; mode 2: direction octant (3 high
// bits) over 8^3 origin cells.  Window w of an item holds its local rays
// [w << shift, (w + 1) << shift); shift 31 makes the whole item one window.
__global__ __launch_bounds__(256) void k_regroup_bin(BatchIo bio, uint32_t total, uint32_t mode, uint32_t shift,
                                                     uint32_t* __restrict__ hist, uint32_t* __restrict__ bin_rank) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        const uint32_t b = item_of(bio, g);
        const BatchEntry& e = bio.table[b];
        const uint32_t local = g - e.base;
        const float4* rp = reinterpret_cast<const float4*>(e.rays) + 2 * (size_t)local;
        const float4 o = rp[0];
        // (v - org) * rs spans [2, 65002] over the item's scene
        const float cells = mode == 2 ? 8.0f : 16.0f;
        auto cell = [cells](float v, float org, float rs) -> uint32_t {
            const float q = (v - org) * rs * (cells / 65000.0f);
            return q >= cells - 1.0f ? (uint32_t)cells - 1u : (q > 0.0f ? (uint32_t)q : 0u);
        };
        auto spread4 = [](uint32_t x) -> uint32_t {   // bit k -> bit 3k
            return (x & 1u) | ((x & 2u) << 2) | ((x & 4u) << 4) | ((x & 8u) << 6);
        };
        uint32_t bin = (spread4(cell(o.x, e.frame.ox, e.frame.rx)) << 2) |
                       (spread4(cell(o.y, e.frame.oy, e.frame.ry)) << 1) | spread4(cell(o.z, e.frame.oz, e.frame.rz));
        if (mode == 2) {
            const float4 d = rp[1];
            bin |= ((d.x < 0.0f ? 4u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 1u : 0u)) << 9;
        }
        const uint32_t win = e.win0 + (local >> shift);
        uint32_t rank = atomicAdd(&hist[(size_t)win * kBins + bin], 1u);
        bin_rank[2 * (size_t)g] = bin;
        bin_rank[2 * (size_t)g + 1] = rank;
    }
}

// Exclusive scan of each window's bin counts (one block per window).
__global__ __launch_bounds__(256) void k_regroup_scan(uint32_t* __restrict__ hist, uint32_t n_windows) {
    __shared__ uint32_t part[256];
    const uint32_t b = blockIdx.x;
    if (b >= n_windows) return;
    uint32_t* h = hist + (size_t)b * kBins;
    const uint32_t t = threadIdx.x;
    constexpr uint32_t kPer = kBins / 256;      // consecutive bins per thread
    uint32_t v[kPer];
    uint32_t sum = 0;
    for (uint32_t k = 0; k < kPer; ++k) {
        v[k] = h[t * kPer + k];
        sum += v[k];
    }
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {
        const uint32_t x = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    uint32_t excl = part[t] - sum;
    for (uint32_t k = 0; k < kPer; ++k) {
        h[t * kPer + k] = excl;
        excl += v[k];
    }
}

// Each ray to its place: the item's base, its window's first ray, the
// bin's offset inside the window and the ray's rank inside the bin.
__global__ __launch_bounds__(256) void k_regroup_scatter(BatchIo bio, uint32_t total, uint32_t shift,
                                                         const uint32_t* __restrict__ hist,
                                                         const uint32_t* __restrict__ bin_rank,
                                                         uint32_t* __restrict__ perm) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
        const uint32_t b = item_of(bio, g);
        const BatchEntry& e = bio.table[b];
        const uint32_t local = g - e.base;
        const uint32_t w = local >> shift;
        const uint32_t first = shift >= 31 ? 0u : w << shift;
        const uint32_t bin = bin_rank[2 * (size_t)g], rank = bin_rank[2 * (size_t)g + 1];
        perm[e.base + first + hist[(size_t)(e.win0 + w) * kBins + bin] + rank] = local;
    }
}

template <int BLOCK, int STACK, int RING, int LAYOUT>
__global__ __launch_bounds__(BLOCK) void k_trace_batch(TraceIo io, BatchIo bio, int32_t* spill) {
    __shared__ int32_t stk[(RING ? RING : STACK) * BLOCK];
    TraceScene sc{};
    trace_persistent<BLOCK, STACK, LAYOUT, false, true, RING>(sc, io, bio, stk, spill);
}

__global__ void k_fill(uint32_t* p, uint32_t value, uint32_t count) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) p[i] = value;
}

rt_batch_config default_config() {
    rt_batch_config c{};
    c.block = 256;
    c.stack_lds = 0;
    c.regroup = 0;
    c.refill = 32;
    c.blocks_per_cu = 0;
    c.use_graph = 1;
    c.regroup_interval = 0;
    c.node_layout = RT_LAYOUT_W4;
    return c;
}

bool binary_layout(const rt_batch_config& c) { return c.node_layout == RT_LAYOUT_Q16; }

bool valid_config(const rt_batch_config& c) {
    const bool interval_ok = c.regroup_interval == 0 ||
                             (c.regroup_interval >= kMinInterval && (c.regroup_interval & (c.regroup_interval - 1)) == 0);
    const bool layout_ok = c.node_layout == 0 || c.node_layout == RT_LAYOUT_W4 || c.node_layout == RT_LAYOUT_Q16;
    return (c.block == 128 || c.block == 256) && (c.stack_lds == 0 || c.stack_lds == 8 || c.stack_lds == 16 || c.stack_lds == 32) &&
           c.regroup <= 2 && c.refill >= 1 && c.refill <= 64 && c.blocks_per_cu <= 16 && c.use_graph <= 1 && interval_ok &&
           layout_ok && !(c.node_layout == RT_LAYOUT_Q16 && c.stack_lds != 0);
}

// log2 of the regroup window (31: the whole item)
uint32_t interval_shift(const rt_batch_config& c) {
    if (c.regroup_interval == 0) return 31u;
    uint32_t k = 0;
    while ((1u << k) < c.regroup_interval) ++k;
    return k;
}

}  // namespace
}  // namespace rtamd

using namespace rtamd;

struct rt_batch {
    rt_ctx* ctx = nullptr;
    std::vector<rt_batch_item> items;
    uint32_t total = 0;
    uint32_t stack_need = 0;           // 3 * max 4-wide height + 1
    uint64_t total_tris = 0;
    BatchEntry* d_table = nullptr;
    uint32_t* d_offsets = nullptr;
    uint32_t* d_chunk_item = nullptr;
    uint32_t* d_hist = nullptr;
    uint32_t* d_bin_rank = nullptr;
    uint32_t* d_perm = nullptr;
    uint32_t* d_cursor = nullptr;
    int32_t* d_spill = nullptr;
    size_t spill_bytes = 0;
    rt_batch_config cfg{};
    std::vector<BatchEntry> h_table;   // the item table as last uploaded
    uint32_t hist_windows = 0;         // regroup windows the histogram holds (at the smallest interval)
    uint32_t n_windows = 0;            // regroup windows under cfg's interval
    uint32_t packed_height = 0;        // deepest binary tree among the items (binary records)
    uint32_t grid_blocks = 0;          // persistent blocks for cfg (set by a dry pass)
    bool dry = false;                  // size buffers only, launch nothing
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t tuning_gen = 0;           // the context's tuning the launch was set up under
};

namespace {

template <int BLOCK, int STACK, int RING, int LAYOUT>
int launch_batch_kernel(rt_ctx* ctx, rt_batch* b, hipStream_t s) {
    auto kern = k_trace_batch<BLOCK, STACK, RING, LAYOUT>;
    if (b->dry) {
        // occupancy and the spill buffer are settled outside any capture
        int per_cu = 0;
        RT_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, BLOCK, 0));
        if (per_cu < 1) per_cu = 1;
        if (b->cfg.blocks_per_cu && (int)b->cfg.blocks_per_cu < per_cu) per_cu = (int)b->cfg.blocks_per_cu;
        uint32_t blocks = (uint32_t)ctx->num_cus * (uint32_t)per_cu;
        uint32_t need = (b->total + BLOCK - 1) / BLOCK;
        b->grid_blocks = need < blocks ? need : blocks;
        size_t bytes = RING ? (size_t)(STACK - RING) * b->grid_blocks * BLOCK * sizeof(int32_t) : 0;
        if (bytes > b->spill_bytes) {
            RT_HIP(ctx, hipDeviceSynchronize());
            if (b->d_spill) RT_HIP(ctx, hipFree(b->d_spill));
            b->d_spill = nullptr;
            RT_HIP(ctx, hipMalloc((void**)&b->d_spill, bytes));
            b->spill_bytes = bytes;
        }
        return RT_OK;
    }
    const uint32_t blocks = b->grid_blocks;
    if (!blocks) return RT_OK;
    int32_t* spill = RING ? b->d_spill : nullptr;
    TraceIo io{nullptr, nullptr, nullptr, nullptr, nullptr, b->d_cursor, b->total, b->cfg.refill, nullptr,
               b->total >= (1u << 16) ? kSegments : 1u, drain_steal(ctx)};
    BatchIo bio{b->d_table, b->d_offsets, b->d_chunk_item, b->cfg.regroup ? b->d_perm : nullptr};
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(BLOCK), 0, s, io, bio, spill);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

template <int BLOCK, int STACK>
int launch_batch_ring(rt_ctx* ctx, rt_batch* b, hipStream_t s) {
    if (b->cfg.stack_lds == 8 && STACK > 8) return launch_batch_kernel<BLOCK, STACK, 8, kLayoutW4>(ctx, b, s);
    if (b->cfg.stack_lds == 16 && STACK > 16) return launch_batch_kernel<BLOCK, STACK, 16, kLayoutW4>(ctx, b, s);
    if (b->cfg.stack_lds == 32 && STACK > 32) return launch_batch_kernel<BLOCK, STACK, 32, kLayoutW4>(ctx, b, s);
    return launch_batch_kernel<BLOCK, STACK, 0, kLayoutW4>(ctx, b, s);
}

// Binary 16-bit records: one deferred child per level, whole stack in LDS.
template <int BLOCK>
int launch_batch_binary(rt_ctx* ctx, rt_batch* b, hipStream_t s) {
    const uint32_t h = b->packed_height;
    if (h <= 24) return launch_batch_kernel<BLOCK, 24, 0, kLayoutQ16>(ctx, b, s);
    if (h <= 32) return launch_batch_kernel<BLOCK, 32, 0, kLayoutQ16>(ctx, b, s);
    if (h <= 48) return launch_batch_kernel<BLOCK, 48, 0, kLayoutQ16>(ctx, b, s);
    if (h <= 64) return launch_batch_kernel<BLOCK, 64, 0, kLayoutQ16>(ctx, b, s);
    return set_error(ctx, RT_ERR_UNSUPPORTED, "batch: a binary tree deeper than 64 levels");
}

template <int BLOCK>
int launch_batch_block(rt_ctx* ctx, rt_batch* b, hipStream_t s) {
    if (binary_layout(b->cfg)) return launch_batch_binary<BLOCK>(ctx, b, s);
    const uint32_t need = b->stack_need;
    if (need <= 24) return launch_batch_ring<BLOCK, 24>(ctx, b, s);
    if (need <= 36) return launch_batch_ring<BLOCK, 36>(ctx, b, s);
    if (need <= 48) return launch_batch_ring<BLOCK, 48>(ctx, b, s);
    if (need <= 64) return launch_batch_ring<BLOCK, 64>(ctx, b, s);
    if (need <= 96) return launch_batch_ring<BLOCK, 96>(ctx, b, s);
    return set_error(ctx, RT_ERR_UNSUPPORTED, "batch: a 4-wide tree deeper than 31 levels");
}

// The whole launch sequence on `s` (also what the graph captures).
int launch_sequence(rt_ctx* ctx, rt_batch* b, hipStream_t s) {
    if (b->dry) return b->cfg.block == 128 ? launch_batch_block<128>(ctx, b, s) : launch_batch_block<256>(ctx, b, s);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, b->d_cursor, 0u, kCursorBlock);
    RT_HIP(ctx, hipGetLastError());
    if (b->cfg.regroup) {
        const uint32_t nh = b->n_windows * kBins;
        const uint32_t shift = interval_shift(b->cfg);
        const uint32_t grid = std::max(1u, std::min((b->total + 255u) / 256u, (uint32_t)ctx->num_cus * 16u));
        BatchIo bio{b->d_table, b->d_offsets, b->d_chunk_item, nullptr};
        hipLaunchKernelGGL(k_fill, dim3(std::min((nh + 255u) / 256u, 1024u)), dim3(256), 0, s, b->d_hist, 0u, nh);
        hipLaunchKernelGGL(k_regroup_bin, dim3(grid), dim3(256), 0, s, bio, b->total, b->cfg.regroup, shift, b->d_hist,
                           b->d_bin_rank);
        hipLaunchKernelGGL(k_regroup_scan, dim3(b->n_windows), dim3(256), 0, s, b->d_hist, b->n_windows);
        hipLaunchKernelGGL(k_regroup_scatter, dim3(grid), dim3(256), 0, s, bio, b->total, shift, b->d_hist,
                           b->d_bin_rank, b->d_perm);
        RT_HIP(ctx, hipGetLastError());
    }
    return b->cfg.block == 128 ? launch_batch_block<128>(ctx, b, s) : launch_batch_block<256>(ctx, b, s);
}

void drop_graph(rt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// The item table b->cfg reads: each item's first regroup window under the
// configured interval, and the binary records when the configuration reads
// them (built now for scenes uploaded without).  Uploaded when it changed,
// after the device has finished the launches that read the old one.
int prepare_table(rt_ctx* ctx, rt_batch* b) {
    std::vector<BatchEntry> t = b->h_table;
    const uint32_t shift = interval_shift(b->cfg);
    uint32_t win = 0, hq = 0;
    for (size_t i = 0; i < b->items.size(); ++i) {
        BatchEntry& e = t[i];
        e.win0 = win;
        win += shift >= 31 ? 1u : std::max(1u, (uint32_t)(((uint64_t)e.count + (1u << shift) - 1u) >> shift));
        if (binary_layout(b->cfg)) {
            rt_scene* sc = const_cast<rt_scene*>(b->items[i].scene);
            if (!sc->d_qnodes) {
                const int rc = scene_binary_records(ctx, sc, RT_LAYOUT_Q16);
                if (rc) return rc;
            }
            e.qnodes = sc->d_qnodes;
            e.root_q = sc->root_ref;
            hq = std::max(hq, sc->packed_height);
        }
    }
    if (win > b->hist_windows) return set_error(ctx, RT_ERR_INVALID, "batch: regroup interval below the smallest window");
    b->n_windows = std::max(win, 1u);
    b->packed_height = hq;
    if (t.empty() || std::memcmp(t.data(), b->h_table.data(), sizeof(BatchEntry) * t.size()) == 0) return RT_OK;
    RT_HIP(ctx, hipDeviceSynchronize());
    RT_HIP(ctx, hipMemcpy(b->d_table, t.data(), sizeof(BatchEntry) * t.size(), hipMemcpyHostToDevice));
    b->h_table = t;
    return RT_OK;
}

// Settles the item table, grid and spill buffer for b->cfg, then captures
// the sequence into a graph when the configuration asks for one.
int capture(rt_ctx* ctx, rt_batch* b) {
    drop_graph(b);
    int trc = prepare_table(ctx, b);
    if (trc) return trc;
    b->tuning_gen = ctx->tuning_gen;
    b->dry = true;
    int rc = launch_sequence(ctx, b, nullptr);
    b->dry = false;
    if (rc) return rc;
    if (!b->cfg.use_graph || b->total == 0) return RT_OK;
    RT_HIP(ctx, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    rc = launch_sequence(ctx, b, b->cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_error(ctx, e, "hipStreamEndCapture");
    b->graph = g;
    RT_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return RT_OK;
}

void free_batch(rt_batch* b) {
    drop_graph(b);
    for (void* p : {(void*)b->d_table, (void*)b->d_offsets, (void*)b->d_chunk_item, (void*)b->d_hist,
                    (void*)b->d_bin_rank, (void*)b->d_perm, (void*)b->d_cursor, (void*)b->d_spill})
        if (p) (void)hipFree(p);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
}

uint32_t ilog2(uint64_t v) {
    uint32_t r = 0;
    while (v > 1) { v >>= 1; ++r; }
    return r;
}

std::string cache_key(const rt_batch* b) {
    std::ostringstream k;
    k << "items=" << b->items.size() << " log2rays=" << ilog2(b->total) << " log2tris=" << ilog2(b->total_tris);
    return k.str();
}

template <int STACK>
int launch_dyn_t(rt_ctx* ctx, const BatchIo& bio, uint32_t* cursor, uint32_t total, uint32_t refill, hipStream_t s) {
    auto kern = k_trace_batch<256, STACK, 0, kLayoutW4>;
    int per_cu = 0;
    RT_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0));
    if (per_cu < 1) per_cu = 1;
    uint32_t blocks = (uint32_t)ctx->num_cus * (uint32_t)per_cu;
    uint32_t need = (total + 255u) / 256u;
    if (need < blocks) blocks = need;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, cursor, 0u, kCursorBlock);
    TraceIo io{nullptr, nullptr, nullptr, nullptr, nullptr, cursor, total, refill, nullptr,
               total >= (1u << 16) ? kSegments : 1u, drain_steal(ctx)};
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, io, bio, (int32_t*)nullptr);
    RT_HIP(ctx, hipGetLastError());
    return RT_OK;
}

}  // namespace

namespace rtamd {
int launch_batch_dynamic(rt_ctx* ctx, const BatchIo& bio, uint32_t* cursor, uint32_t total, uint32_t stack_need,
                         uint32_t refill, hipStream_t s) {
    if (!total) return RT_OK;
    if (refill < 1u || refill > 64u) refill = 32u;
    if (stack_need <= 24) return launch_dyn_t<24>(ctx, bio, cursor, total, refill, s);
    if (stack_need <= 36) return launch_dyn_t<36>(ctx, bio, cursor, total, refill, s);
    if (stack_need <= 48) return launch_dyn_t<48>(ctx, bio, cursor, total, refill, s);
    if (stack_need <= 64) return launch_dyn_t<64>(ctx, bio, cursor, total, refill, s);
    if (stack_need <= 96) return launch_dyn_t<96>(ctx, bio, cursor, total, refill, s);
    return set_error(ctx, RT_ERR_UNSUPPORTED, "batch: a 4-wide tree deeper than 31 levels");
}
}  // namespace rtamd

extern "C" int rt_batch_create(rt_ctx* ctx, const rt_batch_item* items, uint32_t n, const rt_batch_config* config,
                               rt_batch** out) {
    if (!ctx) return RT_ERR_INVALID;
    if (!out || (n && !items)) return set_error(ctx, RT_ERR_INVALID, "rt_batch_create: null argument");
    *out = nullptr;
    rt_batch_config cfg = config ? *config : default_config();
    if (!valid_config(cfg)) return set_error(ctx, RT_ERR_INVALID, "rt_batch_create: invalid configuration");
    uint64_t total = 0;
    uint32_t h4 = 0;
    uint64_t tris = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const rt_batch_item& it = items[i];
        if (!scene_usable(ctx, it.scene))
            return set_error(ctx, RT_ERR_INVALID, "rt_batch_create: item " + std::to_string(i) + " has no scene on this device");
        if (!it.scene->fast)
            return set_error(ctx, RT_ERR_UNSUPPORTED, "rt_batch_create: item " + std::to_string(i) +
                                                          " has a scene without the packed traversal");
        if (it.count && (!it.rays_dev || !it.hits_dev || ((uintptr_t)it.rays_dev & 15u) || ((uintptr_t)it.hits_dev & 15u)))
            return set_error(ctx, RT_ERR_INVALID, "rt_batch_create: item " + std::to_string(i) + " buffers null or unaligned");
        total += it.count;
        h4 = std::max(h4, it.scene->height4);
        tris += it.scene->info.num_triangles;
    }
    if (total > 0x7fffffffull) return set_error(ctx, RT_ERR_INVALID, "rt_batch_create: more than 2^31 rays");
    RT_HIP(ctx, hipSetDevice(ctx->device));
    rt_batch* b = new rt_batch();
    b->ctx = ctx;
    b->items.assign(items, items + n);
    b->total = (uint32_t)total;
    b->stack_need = 3 * h4 + 1;
    b->total_tris = tris;
    b->cfg = cfg;
    int rc = RT_OK;
    do {
        std::vector<BatchEntry> table(std::max(n, 1u));
        std::vector<uint32_t> offsets(n + 1, 0);
        for (uint32_t i = 0; i < n; ++i) {
            const rt_scene* sc = items[i].scene;
            BatchEntry& e = table[i];
            e.nodes = sc->d_w4nodes;
            e.tris = sc->d_ptris;
            e.tri_id = sc->d_ptri_id;
            e.rays = items[i].rays_dev;
            e.hits = items[i].hits_dev;
            e.qnodes = sc->d_qnodes;
            e.frame = sc->frame;
            e.root = sc->root4;
            e.root_q = sc->root_ref;
            e.base = offsets[i];
            e.count = items[i].count;
            e.win0 = i;
            offsets[i + 1] = offsets[i] + items[i].count;
            b->hist_windows += std::max(1u, (uint32_t)(((uint64_t)items[i].count + kMinInterval - 1u) / kMinInterval));
        }
        b->hist_windows = std::max(b->hist_windows, 1u);
        b->h_table = table;
        const uint32_t chunks = (b->total + kChunk - 1) / kChunk;
        std::vector<uint32_t> chunk_item(std::max(chunks, 1u), 0);
        for (uint32_t i = 0, c = 0; i < n && c < chunks; ++i)
            while (c < chunks && (uint64_t)c * kChunk < offsets[i + 1]) chunk_item[c++] = i;
        auto up = [&](void** dst, const void* src, size_t bytes) -> int {
            RT_HIP(ctx, hipMalloc(dst, bytes));
            if (src) RT_HIP(ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
            return RT_OK;
        };
        if ((rc = up((void**)&b->d_table, table.data(), sizeof(BatchEntry) * table.size()))) break;
        if ((rc = up((void**)&b->d_offsets, offsets.data(), sizeof(uint32_t) * offsets.size()))) break;
        if ((rc = up((void**)&b->d_chunk_item, chunk_item.data(), sizeof(uint32_t) * chunk_item.size()))) break;
        if ((rc = up((void**)&b->d_cursor, nullptr, sizeof(uint32_t) * kCursorBlock))) break;
        if ((rc = up((void**)&b->d_hist, nullptr, sizeof(uint32_t) * kBins * (size_t)b->hist_windows))) break;
        if ((rc = up((void**)&b->d_bin_rank, nullptr, sizeof(uint32_t) * 2 * std::max(b->total, 1u)))) break;
        if ((rc = up((void**)&b->d_perm, nullptr, sizeof(uint32_t) * std::max(b->total, 1u)))) break;
        if (hipStreamCreate(&b->cap_stream) != hipSuccess) { rc = set_error(ctx, RT_ERR_HIP, "hipStreamCreate"); break; }
        rc = capture(ctx, b);
    } while (0);
    if (rc) {
        free_batch(b);
        return rc;
    }
    *out = b;
    return RT_OK;
}

extern "C" int rt_batch_set_config(rt_ctx* ctx, rt_batch* b, const rt_batch_config* config) {
    if (!ctx || !b || !config || b->ctx != ctx) return ctx ? set_error(ctx, RT_ERR_INVALID, "rt_batch_set_config") : RT_ERR_INVALID;
    if (!valid_config(*config)) return set_error(ctx, RT_ERR_INVALID, "rt_batch_set_config: invalid configuration");
    RT_HIP(ctx, hipSetDevice(ctx->device));
    b->cfg = *config;
    return capture(ctx, b);
}

extern "C" int rt_batch_get_config(const rt_batch* b, rt_batch_config* config) {
    if (!b || !config) return RT_ERR_INVALID;
    *config = b->cfg;
    return RT_OK;
}

extern "C" int rt_batch_launch(rt_ctx* ctx, rt_batch* b, void* stream) {
    if (!ctx || !b || b->ctx != ctx) return ctx ? set_error(ctx, RT_ERR_INVALID, "rt_batch_launch") : RT_ERR_INVALID;
    if (b->total == 0) return RT_OK;
    RT_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    // the context's tuning (the drain's steal threshold) changed since the
    // launch was set up: set it up again, so a replay never runs stale values
    if (b->tuning_gen != ctx->tuning_gen) {
        const int rc = capture(ctx, b);
        if (rc) return rc;
    }
    if (b->exec) {
        RT_HIP(ctx, hipGraphLaunch(b->exec, s));
        return RT_OK;
    }
    return launch_sequence(ctx, b, s);
}

extern "C" int rt_batch_destroy(rt_ctx* ctx, rt_batch* b) {
    if (!ctx) return RT_ERR_INVALID;
    if (!b) return RT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    free_batch(b);
    return RT_OK;
}

extern "C" int rt_batch_autotune(rt_ctx* ctx, rt_batch* b, uint32_t trials, const char* cache_path,
                                 rt_batch_config* best) {
    if (!ctx || !b || b->ctx != ctx) return ctx ? set_error(ctx, RT_ERR_INVALID, "rt_batch_autotune") : RT_ERR_INVALID;
    if (trials == 0) trials = 3;
    RT_HIP(ctx, hipSetDevice(ctx->device));
    const std::string key = cache_key(b);
    if (cache_path && *cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        rt_batch_config found{};
        bool hit = false;
        while (std::getline(in, line)) {
            if (line.compare(0, key.size() + 1, key + " ") != 0) continue;
            std::istringstream ss(line.substr(key.size() + 1));
            std::string tok;
            rt_batch_config c = default_config();
            while (ss >> tok) {
                auto eq = tok.find('=');
                if (eq == std::string::npos) continue;
                std::string k = tok.substr(0, eq);
                uint32_t v = (uint32_t)std::stoul(tok.substr(eq + 1));
                if (k == "block") c.block = v;
                else if (k == "stack_lds") c.stack_lds = v;
                else if (k == "regroup") c.regroup = v;
                else if (k == "refill") c.refill = v;
                else if (k == "blocks_per_cu") c.blocks_per_cu = v;
                else if (k == "regroup_interval") c.regroup_interval = v;
                else if (k == "node_layout") c.node_layout = v;
            }
            if (valid_config(c)) { found = c; hit = true; }
        }
        if (hit) {
            found.use_graph = b->cfg.use_graph;
            int rc = rt_batch_set_config(ctx, b, &found);
            if (rc) return rc;
            if (best) *best = b->cfg;
            return RT_OK;
        }
    }
    hipEvent_t e0, e1;
    RT_HIP(ctx, hipEventCreate(&e0));
    RT_HIP(ctx, hipEventCreate(&e1));
    auto time_cfg = [&](const rt_batch_config& c, float& ms) -> int {
        b->cfg = c;
        int rc = capture(ctx, b);
        if (rc) return rc;
        rc = rt_batch_launch(ctx, b, b->cap_stream);   // warm
        if (rc) return rc;
        std::vector<float> v;
        for (uint32_t t = 0; t < trials; ++t) {
            RT_HIP(ctx, hipEventRecord(e0, b->cap_stream));
            rc = rt_batch_launch(ctx, b, b->cap_stream);
            if (rc) return rc;
            RT_HIP(ctx, hipEventRecord(e1, b->cap_stream));
            RT_HIP(ctx, hipEventSynchronize(e1));
            float x = 0.0f;
            RT_HIP(ctx, hipEventElapsedTime(&x, e0, e1));
            v.push_back(x);
        }
        std::sort(v.begin(), v.end());
        ms = v[v.size() / 2];
        return RT_OK;
    };
    const uint32_t keep_graph = b->cfg.use_graph;
    rt_batch_config cur = b->cfg;
    float cur_ms = 0.0f;
    int rc = time_cfg(cur, cur_ms);
    // coordinate descent over: work-group size, node-fetch layout (4-wide or
    // binary 16-bit records), LDS stack ring, regrouping (off, origin cell,
    // direction octant + origin cell), regroup interval, refill threshold
    // and blocks per CU
    const std::vector<uint32_t> blocks = {256, 128}, layouts = {RT_LAYOUT_W4, RT_LAYOUT_Q16},
                                rings = {0, 32, 16, 8}, regroups = {0, 1, 2}, intervals = {0, 1u << 16, 1u << 18},
                                refills = {8, 16, 24, 32, 48, 64}, bpcs = {0, 2, 3, 4};
    constexpr int kDims = 7;
    for (int pass = 0; pass < 2 && !rc; ++pass) {
        for (int p = 0; p < kDims && !rc; ++p) {
            const std::vector<uint32_t>& vals = p == 0 ? blocks : p == 1 ? layouts : p == 2 ? rings : p == 3 ? regroups
                                              : p == 4 ? intervals : p == 5 ? refills : bpcs;
            for (uint32_t v : vals) {
                rt_batch_config c = cur;
                uint32_t* field = p == 0 ? &c.block : p == 1 ? &c.node_layout : p == 2 ? &c.stack_lds
                                : p == 3 ? &c.regroup : p == 4 ? &c.regroup_interval : p == 5 ? &c.refill
                                : &c.blocks_per_cu;
                if (*field == v) continue;
                *field = v;
                // combinations that cannot run, or equal the current one (an interval without regrouping)
                if (!valid_config(c) || (p == 4 && c.regroup == 0)) continue;
                float ms = 0.0f;
                if ((rc = time_cfg(c, ms))) break;
                if (ms < cur_ms) { cur = c; cur_ms = ms; }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    cur.use_graph = keep_graph;
    rc = rt_batch_set_config(ctx, b, &cur);
    if (rc) return rc;
    if (best) *best = cur;
    if (cache_path && *cache_path) {
        std::ofstream o(cache_path, std::ios::app);
        o << key << " block=" << cur.block << " stack_lds=" << cur.stack_lds << " regroup=" << cur.regroup
          << " refill=" << cur.refill << " blocks_per_cu=" << cur.blocks_per_cu
          << " regroup_interval=" << cur.regroup_interval << " node_layout=" << cur.node_layout << " ms=" << cur_ms
          << "\n";
    }
    return RT_OK;
}
