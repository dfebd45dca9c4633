// rt_server.cpp -- the served form: resident scenes under an HBM budget and
// a request loop that batches closest-hit queries as device slots free up.
//
// One worker thread owns `slots` device slots; each slot is its own context
// (workspaces), a non-blocking stream, pinned staging and device buffers.
// The loop: (1) retire slots whose completion event fired -- copy results to
// the callers' buffers and stamp the completion time; (2) give every free
// slot the oldest queued work -- a render alone, or the run of queued
// queries at the head of the queue (up to max_batch_rays) as one batched
// launch; (3) sleep briefly while slots are busy, or until a submission.
// Launches in different slots overlap on the device, so requests arriving
// while others run are admitted as soon as a slot frees up.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rt_common.hpp"

namespace {

using Clock = std::chrono::steady_clock;
using namespace rtamd;

struct SrvScene {
    rt_scene* scene = nullptr;
    uint64_t bytes = 0;
    uint64_t last_use = 0;
    uint32_t refs = 0;           // queued or running requests
};

struct Request {
    uint64_t id = 0;
    int kind = 0;                // 0 query, 1 render
    uint64_t scene_id = 0;
    rt_scene* scene = nullptr;
    const rt_ray* rays = nullptr;
    uint32_t n = 0;
    rt_hit* hits = nullptr;
    rt_ubo ubo{};
    rt_render_params rp{};
    uint32_t w = 0, h = 0;
    uint8_t* target = nullptr;
    Clock::time_point t0, t1;
    int status = RT_OK;
    bool done = false;
    bool detached = false;       // nobody waits: freed on completion
    bool running = false;        // taken by a slot
    bool direct = false;         // caller buffers are page-locked: DMA in place
    uint32_t offset = 0;         // first ray in the slot's staging
};

// True when `p` is page-locked host memory the device can DMA from.
bool pinned(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

struct Slot {
    rt_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    rt_ray* d_rays = nullptr;
    rt_hit* d_hits = nullptr;
    rt_ray* h_rays = nullptr;
    rt_hit* h_hits = nullptr;
    BatchEntry* d_table = nullptr;
    BatchEntry* h_table = nullptr;
    uint32_t* d_offsets = nullptr;
    uint32_t* h_offsets = nullptr;
    uint32_t* d_chunk = nullptr;
    uint32_t* h_chunk = nullptr;
    uint32_t* d_cursor = nullptr;
    uint8_t* d_target = nullptr;
    uint8_t* d_accum = nullptr;
    uint8_t* h_target = nullptr;
    uint32_t* h_counters = nullptr;
    std::vector<Request*> reqs;
    bool busy = false;
    uint64_t rays = 0;           // query rays of the running launch
    bool render = false;
};

constexpr uint32_t kMaxItems = 4096;

}  // namespace

struct rt_server {
    rt_ctx* ctx = nullptr;               // uploads
    rt_server_config cfg{};
    std::vector<Slot> slots;
    std::thread worker;
    std::mutex mu;
    std::mutex upload_mu;
    std::condition_variable work_cv, done_cv;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> waited;
    std::unordered_map<uint64_t, SrvScene> scenes;
    uint64_t next_request = 1, next_scene = 1, tick = 1;
    uint64_t resident = 0;               // admitted scene bytes (includes one being uploaded)
    uint64_t peak = 0;                   // most of them at once
    uint64_t in_flight = 0;              // submitted and not completed
    bool stop = false;
    rt_server_stats st{};
    std::vector<float> lat_ms;
    std::string err;
};

namespace {

int fail(rt_server* S, int code, const std::string& m) {
    S->err = m;
    return code;
}

void free_slot(Slot& s) {
    for (void* p : {(void*)s.d_rays, (void*)s.d_hits, (void*)s.d_table, (void*)s.d_offsets, (void*)s.d_chunk,
                    (void*)s.d_cursor, (void*)s.d_target, (void*)s.d_accum})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)s.h_rays, (void*)s.h_hits, (void*)s.h_table, (void*)s.h_offsets, (void*)s.h_chunk,
                    (void*)s.h_target, (void*)s.h_counters})
        if (p) (void)hipHostFree(p);
    if (s.ev) (void)hipEventDestroy(s.ev);
    if (s.stream) (void)hipStreamDestroy(s.stream);
    if (s.ctx) rt_destroy(s.ctx);
    s = Slot{};
}

int make_slot(rt_server* S, Slot& s) {
    const uint32_t R = S->cfg.max_batch_rays, P = S->cfg.max_render_pixels;
    if (rt_create(&s.ctx, S->ctx->device) != RT_OK) return fail(S, RT_ERR_HIP, "slot context");
    if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess) return fail(S, RT_ERR_HIP, "slot stream");
    if (hipEventCreateWithFlags(&s.ev, hipEventDisableTiming) != hipSuccess) return fail(S, RT_ERR_HIP, "slot event");
    auto dm = [&](void** p, size_t b) { return hipMalloc(p, b) == hipSuccess; };
    auto hm = [&](void** p, size_t b) { return hipHostMalloc(p, b, hipHostMallocDefault) == hipSuccess; };
    const size_t chunks = R / 64 + 2;
    bool ok = dm((void**)&s.d_rays, (size_t)R * sizeof(rt_ray)) && dm((void**)&s.d_hits, (size_t)R * sizeof(rt_hit)) &&
              hm((void**)&s.h_rays, (size_t)R * sizeof(rt_ray)) && hm((void**)&s.h_hits, (size_t)R * sizeof(rt_hit)) &&
              dm((void**)&s.d_table, kMaxItems * sizeof(BatchEntry)) && hm((void**)&s.h_table, kMaxItems * sizeof(BatchEntry)) &&
              dm((void**)&s.d_offsets, (kMaxItems + 1) * 4) && hm((void**)&s.h_offsets, (kMaxItems + 1) * 4) &&
              dm((void**)&s.d_chunk, chunks * 4) && hm((void**)&s.h_chunk, chunks * 4) && dm((void**)&s.d_cursor, 4 * kCursorBlock) &&
              dm((void**)&s.d_target, (size_t)P * 4) && dm((void**)&s.d_accum, (size_t)P * 4) &&
              hm((void**)&s.h_target, (size_t)P * 4) && hm((void**)&s.h_counters, 64 * 4);
    if (!ok) return fail(S, RT_ERR_NOMEM, "slot buffers");
    if (hipMemset(s.d_accum, 0, (size_t)P * 4) != hipSuccess) return fail(S, RT_ERR_HIP, "slot accumulation image");
    return RT_OK;
}

// Launches the slot's requests (called without the server lock: the
// requests are already out of the queue and belong to this slot).
int launch_slot(rt_server* S, Slot& sl) {
    hipStream_t st = sl.stream;
    if (sl.render) {
        Request* r = sl.reqs[0];
        int rc = rt_render(sl.ctx, r->scene, &r->ubo, &r->rp, sl.d_target, sl.d_accum, r->w, r->h, st);
        if (rc) return rc;
        if (hipMemcpyAsync(r->direct ? r->target : sl.h_target, sl.d_target, (size_t)r->w * r->h * 4,
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(sl.h_counters, sl.ctx->d_counters, 64 * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
            return RT_ERR_HIP;
    } else if (S->cfg.mode == RT_SERVE_FIFO) {
        Request* r = sl.reqs[0];
        const rt_ray* src = r->rays;
        if (!r->direct) {
            std::memcpy(sl.h_rays, r->rays, (size_t)r->n * sizeof(rt_ray));
            src = sl.h_rays;
        }
        if (hipMemcpyAsync(sl.d_rays, src, (size_t)r->n * sizeof(rt_ray), hipMemcpyHostToDevice, st) != hipSuccess)
            return RT_ERR_HIP;
        int rc = rt_closest_hit(sl.ctx, r->scene, sl.d_rays, r->n, sl.d_hits, nullptr, st);
        if (rc) return rc;
        if (hipMemcpyAsync(r->direct ? r->hits : sl.h_hits, sl.d_hits, (size_t)r->n * sizeof(rt_hit),
                           hipMemcpyDeviceToHost, st) != hipSuccess)
            return RT_ERR_HIP;
    } else {
        uint32_t off = 0, need = 0;
        uint64_t long_rays = 0;     // rays on scenes of 2^17 triangles or more
        const uint32_t ni = (uint32_t)sl.reqs.size();
        for (uint32_t i = 0; i < ni; ++i) {
            Request* r = sl.reqs[i];
            r->offset = off;
            if (r->direct) {
                if (hipMemcpyAsync(sl.d_rays + off, r->rays, (size_t)r->n * sizeof(rt_ray), hipMemcpyHostToDevice, st) !=
                    hipSuccess)
                    return RT_ERR_HIP;
            } else {
                std::memcpy(sl.h_rays + off, r->rays, (size_t)r->n * sizeof(rt_ray));
                if (hipMemcpyAsync(sl.d_rays + off, sl.h_rays + off, (size_t)r->n * sizeof(rt_ray),
                                   hipMemcpyHostToDevice, st) != hipSuccess)
                    return RT_ERR_HIP;
            }
            BatchEntry& e = sl.h_table[i];
            const rt_scene* sc = r->scene;
            e.nodes = sc->d_w4nodes;
            e.tris = sc->d_ptris;
            e.tri_id = sc->d_ptri_id;
            e.rays = sl.d_rays + off;
            e.hits = sl.d_hits + off;
            e.frame = sc->frame;
            e.qnodes = nullptr;
            e.root = sc->root4;
            e.root_q = sc->root_ref;
            e.base = off;
            e.count = r->n;
            e.win0 = 0;
            sl.h_offsets[i] = off;
            off += r->n;
            need = std::max(need, 3u * sc->height4 + 1u);
            if (sc->info.num_triangles >= (1u << 17)) long_rays += r->n;
        }
        sl.h_offsets[ni] = off;
        const uint32_t chunks = (off + 63) / 64;
        for (uint32_t i = 0, c = 0; i < ni && c < chunks; ++i)
            while (c < chunks && (uint64_t)c * 64 < sl.h_offsets[i + 1]) sl.h_chunk[c++] = i;
        if (hipMemcpyAsync(sl.d_table, sl.h_table, ni * sizeof(BatchEntry), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(sl.d_offsets, sl.h_offsets, (ni + 1) * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(sl.d_chunk, sl.h_chunk, std::max(chunks, 1u) * 4, hipMemcpyHostToDevice, st) != hipSuccess)
            return RT_ERR_HIP;
        BatchIo bio{sl.d_table, sl.d_offsets, sl.d_chunk, nullptr};
        // refill threshold as for a plain query: early refills when most rays
        // trace a large scene (long traversals), late ones otherwise
        const uint32_t refill = 2ull * long_rays >= off ? 24u : 48u;
        int rc = launch_batch_dynamic(sl.ctx, bio, sl.d_cursor, off, need, refill, st);
        if (rc) return rc;
        for (Request* r : sl.reqs) {
            rt_hit* dst = r->direct ? r->hits : sl.h_hits + r->offset;
            if (hipMemcpyAsync(dst, sl.d_hits + r->offset, (size_t)r->n * sizeof(rt_hit), hipMemcpyDeviceToHost, st) !=
                hipSuccess)
                return RT_ERR_HIP;
        }
    }
    if (hipEventRecord(sl.ev, st) != hipSuccess) return RT_ERR_HIP;
    return RT_OK;
}

// Called with the lock held.
void finish_request(rt_server* S, Request* r, int status) {
    r->status = status;
    r->t1 = Clock::now();
    r->done = true;
    auto it = S->scenes.find(r->scene_id);
    if (it != S->scenes.end() && it->second.refs) it->second.refs--;
    if (status == RT_OK) {
        S->st.completed++;
        S->lat_ms.push_back((float)std::chrono::duration<double, std::milli>(r->t1 - r->t0).count());
    } else if (status == RT_ERR_CANCELLED) {
        S->st.cancelled++;
    } else {
        S->st.failed++;
    }
    S->in_flight--;
    if (r->detached) delete r;
}

// Called with the lock held; the slot's event has fired.
void retire_slot(rt_server* S, Slot& sl) {
    if (sl.render) {
        Request* r = sl.reqs[0];
        if (!r->direct) std::memcpy(r->target, sl.h_target, (size_t)r->w * r->h * 4);
        uint64_t rays = (uint64_t)r->w * r->h;
        for (int i = 1; i < 64; ++i) rays += sl.h_counters[i];
        S->st.rays += rays;
    } else {
        for (Request* r : sl.reqs) {
            if (r->direct) continue;
            const uint32_t off = S->cfg.mode == RT_SERVE_FIFO ? 0u : r->offset;
            std::memcpy(r->hits, sl.h_hits + off, (size_t)r->n * sizeof(rt_hit));
        }
        S->st.rays += sl.rays;
    }
    for (Request* r : sl.reqs) finish_request(S, r, RT_OK);
    sl.reqs.clear();
    sl.busy = false;
    S->done_cv.notify_all();
}

void worker_main(rt_server* S) {
    (void)hipSetDevice(S->ctx->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        bool busy = false;
        for (Slot& sl : S->slots) {
            if (!sl.busy) continue;
            hipError_t q = hipEventQuery(sl.ev);
            if (q == hipSuccess) retire_slot(S, sl);
            else if (q == hipErrorNotReady) busy = true;
            else {
                for (Request* r : sl.reqs) finish_request(S, r, RT_ERR_HIP);
                sl.reqs.clear();
                sl.busy = false;
                S->done_cv.notify_all();
            }
        }
        if (S->stop && S->pending.empty() && !busy) break;
        // requests queued past their deadline are dropped unrun; every
        // request has the same deadline, so they are the oldest, in front
        if (S->cfg.deadline_ms && !S->pending.empty()) {
            const auto now = Clock::now();
            const auto limit = std::chrono::milliseconds(S->cfg.deadline_ms);
            bool dropped = false;
            while (!S->pending.empty() && now - S->pending.front()->t0 >= limit) {
                Request* r = S->pending.front();
                S->pending.pop_front();
                finish_request(S, r, RT_ERR_CANCELLED);
                dropped = true;
            }
            if (dropped) S->done_cv.notify_all();
        }
        for (Slot& sl : S->slots) {
            if (sl.busy || S->pending.empty()) continue;
            sl.render = S->pending.front()->kind == 1;
            sl.rays = 0;
            if (sl.render || S->cfg.mode == RT_SERVE_FIFO) {
                sl.reqs.push_back(S->pending.front());
                sl.rays = sl.reqs[0]->n;
                S->pending.pop_front();
            } else {
                while (!S->pending.empty() && S->pending.front()->kind == 0 && sl.reqs.size() < kMaxItems &&
                       sl.rays + S->pending.front()->n <= S->cfg.max_batch_rays) {
                    sl.rays += S->pending.front()->n;
                    sl.reqs.push_back(S->pending.front());
                    S->pending.pop_front();
                }
            }
            for (Request* r : sl.reqs) r->running = true;
            sl.busy = true;
            S->st.launches++;
            lk.unlock();
            int rc = launch_slot(S, sl);
            lk.lock();
            if (rc) {
                for (Request* r : sl.reqs) finish_request(S, r, rc);
                sl.reqs.clear();
                sl.busy = false;
                S->done_cv.notify_all();
            } else {
                busy = true;
            }
        }
        if (busy) S->work_cv.wait_for(lk, std::chrono::microseconds(20));
        else if (S->pending.empty() && !S->stop) S->work_cv.wait(lk);
    }
}

int submit(rt_server* S, Request* r, uint64_t* request_id) {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) { delete r; return fail(S, RT_ERR_INVALID, "server is stopping"); }
    auto it = S->scenes.find(r->scene_id);
    if (it == S->scenes.end()) { delete r; return fail(S, RT_ERR_NOT_FOUND, "scene is not resident"); }
    r->scene = it->second.scene;
    it->second.refs++;
    it->second.last_use = S->tick++;
    r->id = S->next_request++;
    r->detached = request_id == nullptr;
    r->t0 = Clock::now();
    if (!r->detached) {
        S->waited[r->id] = r;
        *request_id = r->id;
    }
    S->pending.push_back(r);
    S->st.submitted++;
    S->in_flight++;
    S->work_cv.notify_one();
    return RT_OK;
}

}  // namespace

extern "C" int rt_host_alloc(size_t bytes, void** out) {
    if (!out) return RT_ERR_INVALID;
    *out = nullptr;
    if (bytes == 0) return RT_OK;
    return hipHostMalloc(out, bytes, hipHostMallocDefault) == hipSuccess ? RT_OK : RT_ERR_NOMEM;
}

extern "C" int rt_host_free(void* p) {
    if (p) (void)hipHostFree(p);
    return RT_OK;
}

extern "C" int rt_server_create(rt_ctx* ctx, const rt_server_config* config, rt_server** out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    *out = nullptr;
    rt_server_config c = config ? *config : rt_server_config{};
    if (c.mode > RT_SERVE_FIFO) return set_error(ctx, RT_ERR_INVALID, "rt_server_create: unknown mode");
    if (c.slots == 0) c.slots = 4;
    if (c.slots > 8) return set_error(ctx, RT_ERR_INVALID, "rt_server_create: at most 8 slots");
    if (c.mode == RT_SERVE_FIFO) c.slots = 1;
    if (c.max_batch_rays == 0) c.max_batch_rays = 1u << 20;
    if (c.max_render_pixels == 0) c.max_render_pixels = 1u << 20;
    // a slot's batch is one launch: the kernels' 32-bit indices need < 2^31 rays
    if (c.max_batch_rays > (1u << 30) || c.max_render_pixels > (1u << 30))
        return set_error(ctx, RT_ERR_INVALID, "rt_server_create: max_batch_rays and max_render_pixels are at most 2^30");
    RT_HIP(ctx, hipSetDevice(ctx->device));
    rt_server* S = new rt_server();
    S->ctx = ctx;
    S->cfg = c;
    S->slots.resize(c.slots);
    for (Slot& sl : S->slots) {
        int rc = make_slot(S, sl);
        if (rc) {
            std::string m = S->err;
            for (Slot& x : S->slots) free_slot(x);
            delete S;
            return set_error(ctx, rc, "rt_server_create: " + m);
        }
    }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return RT_OK;
}

extern "C" int rt_server_destroy(rt_server* S) {
    if (!S) return RT_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->work_cv.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx->device);
    (void)hipDeviceSynchronize();
    for (Slot& sl : S->slots) free_slot(sl);
    for (auto& kv : S->scenes) rt_scene_release(S->ctx, kv.second.scene);
    for (auto& kv : S->waited) delete kv.second;
    delete S;
    return RT_OK;
}

extern "C" int rt_server_scene_add(rt_server* S, const rt_scene_desc* desc, uint64_t* scene_id) {
    if (!S || !desc || !scene_id) return RT_ERR_INVALID;
    // the host half first: validated and packed, so the scene's device bytes
    // are known before anything is allocated on the device
    PreparedScene prep;
    {
        std::lock_guard<std::mutex> up(S->upload_mu);
        int rc = scene_prepare(S->ctx, desc, prep);
        if (rc) return fail(S, rc, std::string("rt_server_scene_add: ") + rt_last_error(S->ctx));
        if (!prep.fast) return fail(S, RT_ERR_UNSUPPORTED, "rt_server_scene_add: scene without the packed traversal");
    }
    const uint64_t bytes = prep.bytes;
    const uint64_t budget = S->cfg.hbm_budget;
    // the least recently used scene without queued or running work
    auto pick_victim = [&]() -> uint64_t {
        uint64_t victim = 0, oldest = UINT64_MAX;
        for (auto& kv : S->scenes)
            if (kv.second.refs == 0 && kv.second.last_use < oldest) { oldest = kv.second.last_use; victim = kv.first; }
        return victim;
    };
    auto evict = [&](uint64_t victim, std::vector<rt_scene*>& out) {
        out.push_back(S->scenes[victim].scene);
        S->resident -= S->scenes[victim].bytes;
        S->scenes.erase(victim);
        S->st.scenes_evicted++;
    };
    // admission under the budget: evict, then reserve the new scene's bytes
    std::vector<rt_scene*> evicted;
    bool admitted = false;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        while (budget && S->resident + bytes > budget) {
            const uint64_t v = pick_victim();
            if (!v) break;
            evict(v, evicted);
        }
        admitted = !budget || S->resident + bytes <= budget;
        if (admitted) {
            S->resident += bytes;
            S->peak = std::max(S->peak, S->resident);
        } else {
            S->st.scenes_refused++;
        }
    }
    std::lock_guard<std::mutex> up(S->upload_mu);
    for (rt_scene* e : evicted) rt_scene_release(S->ctx, e);      // freed before the new scene is allocated
    evicted.clear();
    if (!admitted) return fail(S, RT_ERR_NOMEM, "rt_server_scene_add: HBM budget exceeded by scenes with pending work");
    // the device itself must hold the scene too (other users share it):
    // make room by evicting further idle scenes while its free memory is short
    const uint64_t margin = 64ull << 20;
    for (;;) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || (uint64_t)free_b >= bytes + margin) break;
        uint64_t v = 0;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            v = pick_victim();
            if (v) evict(v, evicted);
        }
        if (!v) break;
        for (rt_scene* e : evicted) rt_scene_release(S->ctx, e);
        evicted.clear();
    }
    rt_scene* sc = nullptr;
    int rc = scene_commit(S->ctx, prep, &sc);
    std::lock_guard<std::mutex> lk(S->mu);
    if (rc) {
        S->resident -= bytes;
        S->st.scenes_refused++;
        return fail(S, rc == RT_ERR_HIP ? RT_ERR_NOMEM : rc,
                    std::string("rt_server_scene_add: upload: ") + rt_last_error(S->ctx));
    }
    uint64_t id = S->next_scene++;
    S->scenes[id] = SrvScene{sc, bytes, S->tick++, 0};
    S->st.scenes_added++;
    *scene_id = id;
    return RT_OK;
}

extern "C" int rt_server_cancel(rt_server* S, uint64_t id) {
    if (!S) return RT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    for (auto it = S->pending.begin(); it != S->pending.end(); ++it) {
        if ((*it)->id != id) continue;
        Request* r = *it;
        S->pending.erase(it);
        finish_request(S, r, RT_ERR_CANCELLED);
        S->done_cv.notify_all();
        return RT_OK;
    }
    auto w = S->waited.find(id);
    if (w != S->waited.end() && w->second->running && !w->second->done)
        return fail(S, RT_ERR_BUSY, "rt_server_cancel: request is already on the device");
    return fail(S, RT_ERR_NOT_FOUND, "rt_server_cancel: no queued request with this id");
}

extern "C" int rt_server_scene_remove(rt_server* S, uint64_t id) {
    if (!S) return RT_ERR_INVALID;
    rt_scene* sc = nullptr;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->scenes.find(id);
        if (it == S->scenes.end()) return fail(S, RT_ERR_NOT_FOUND, "rt_server_scene_remove: unknown scene");
        if (it->second.refs) return fail(S, RT_ERR_BUSY, "rt_server_scene_remove: scene has pending requests");
        sc = it->second.scene;
        S->resident -= it->second.bytes;
        S->scenes.erase(it);
    }
    std::lock_guard<std::mutex> up(S->upload_mu);
    rt_scene_release(S->ctx, sc);
    return RT_OK;
}

extern "C" int rt_server_scene_list(rt_server* S, uint64_t* ids, uint32_t cap, uint32_t* count) {
    if (!S || !count) return RT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    std::vector<uint64_t> v;
    for (auto& kv : S->scenes) v.push_back(kv.first);
    std::sort(v.begin(), v.end());
    for (uint32_t i = 0; i < v.size() && i < cap && ids; ++i) ids[i] = v[i];
    *count = (uint32_t)v.size();
    return RT_OK;
}

extern "C" int rt_server_submit_query(rt_server* S, uint64_t scene_id, const rt_ray* rays, uint32_t n, rt_hit* hits,
                                      uint64_t* request_id) {
    if (!S) return RT_ERR_INVALID;
    if (n == 0 || !rays || !hits) return fail(S, RT_ERR_INVALID, "rt_server_submit_query: empty request");
    if (n > S->cfg.max_batch_rays) return fail(S, RT_ERR_INVALID, "rt_server_submit_query: more rays than max_batch_rays");
    Request* r = new Request();
    r->kind = 0;
    r->scene_id = scene_id;
    r->rays = rays;
    r->n = n;
    r->hits = hits;
    r->direct = pinned(rays) && pinned(hits);
    return submit(S, r, request_id);
}

extern "C" int rt_server_submit_render(rt_server* S, uint64_t scene_id, const rt_ubo* ubo,
                                       const rt_render_params* params, uint32_t w, uint32_t h, uint8_t* target,
                                       uint64_t* request_id) {
    if (!S) return RT_ERR_INVALID;
    if (!ubo || !params || !target || w == 0 || h == 0) return fail(S, RT_ERR_INVALID, "rt_server_submit_render: bad arguments");
    if ((uint64_t)w * h > S->cfg.max_render_pixels) return fail(S, RT_ERR_INVALID, "rt_server_submit_render: image too large");
    Request* r = new Request();
    r->kind = 1;
    r->scene_id = scene_id;
    r->ubo = *ubo;
    r->ubo.currentSample = ubo->currentSample;
    r->rp = *params;
    r->w = w;
    r->h = h;
    r->target = target;
    r->direct = pinned(target);
    return submit(S, r, request_id);
}

extern "C" int rt_server_wait(rt_server* S, uint64_t id, int timeout_ms, int* status) {
    if (!S) return RT_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->waited.find(id);
    if (it == S->waited.end()) return fail(S, RT_ERR_NOT_FOUND, "rt_server_wait: unknown request");
    Request* r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0) S->done_cv.wait(lk, pred);
    else if (!S->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) return RT_ERR_BUSY;
    if (status) *status = r->status;
    S->waited.erase(it);
    delete r;
    return RT_OK;
}

extern "C" int rt_server_drain(rt_server* S) {
    if (!S) return RT_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    S->done_cv.wait(lk, [&] { return S->in_flight == 0; });
    return RT_OK;
}

extern "C" int rt_server_stats_get(rt_server* S, rt_server_stats* out) {
    if (!S || !out) return RT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    out->resident_bytes = S->resident;
    out->peak_resident_bytes = S->peak;
    if (!S->lat_ms.empty()) {
        std::vector<float> v = S->lat_ms;
        std::sort(v.begin(), v.end());
        auto q = [&](double p) { return (double)v[std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5))]; };
        double sum = 0.0;
        for (float x : v) sum += x;
        out->p50_ms = q(0.50);
        out->p99_ms = q(0.99);
        out->mean_ms = sum / (double)v.size();
        out->max_ms = v.back();
    }
    return RT_OK;
}

extern "C" int rt_server_stats_reset(rt_server* S) {
    if (!S) return RT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->st = rt_server_stats{};
    S->lat_ms.clear();
    return RT_OK;
}

extern "C" const char* rt_server_last_error(const rt_server* S) {
    return S ? S->err.c_str() : "null server";
}
