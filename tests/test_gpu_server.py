"""GPU tests of the served form (rt_server_*): every request's result is the
same alone and in the mix, scenes are evicted least-recently-used under the
HBM budget, and a short Poisson run completes every request."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CAM = [0.0, 10.0, 0.0]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def scenes(rt):
    return [rt.scene_clusters(n, 300 + i) for i, n in enumerate([20000, 70000, 150000])]


@pytest.mark.parametrize("mode,pinned", [(0, False), (1, False), (0, True)])
def test_results_same_alone_and_in_mix(rt, gpu_ctx, scenes, mode, pinned):
    srv = rt.Server(gpu_ctx, mode=mode, slots=3, max_batch_rays=1 << 18, max_render_pixels=128 * 128)
    ids = [srv.add_scene(hs) for hs in scenes]
    alone = [gpu_ctx.upload(hs) for hs in scenes]
    rng = np.random.default_rng(4)
    reqs = []
    for k in range(48):
        s = int(rng.integers(0, 3))
        if k % 6 == 5:
            tgt = rt.pinned_array((96, 128, 4), np.uint8) if pinned else np.zeros((96, 128, 4), np.uint8)
            reqs.append(["render", s, tgt, None])
        else:
            n = int(rng.integers(1, 40000))
            rays = rt.rays_box(n, 500 + k, 4.0)
            hits = np.zeros(n, rt.HIT_DT)
            if pinned:
                pr = rt.pinned_array((n,), rt.RAY_DT)
                pr[...] = rays
                rays = pr
                hits = rt.pinned_array((n,), rt.HIT_DT)
            reqs.append(["query", s, (rays, hits), None])
    print(f"case: server mode {mode} pinned {pinned}, 48 requests", flush=True)
    # inputs are built first so the submits arrive back to back
    for r in reqs:
        kind, s, data, _ = r
        if kind == "render":
            r[3] = srv.submit_render(ids[s], CAM, 128, 96, data, bounces=4)
        else:
            r[3] = srv.submit_query(ids[s], data[0], data[1])
    for kind, s, data, rid in reqs:
        assert srv.wait(rid, 60000) == rt.RT_OK
    for kind, s, data, rid in reqs:
        if kind == "query":
            rays, hits = data
            ref = rt.hits_from_device(gpu_ctx.closest_hit(alone[s], rt.rays_to_device(rays)))
            assert np.array_equal(hits["id"], ref["id"])
            for f in ("t", "u", "v"):
                assert np.array_equal(bits(hits[f]), bits(ref[f]))
        else:
            ref = gpu_ctx.render(alone[s], CAM, 128, 96, vfov=60.0, bounces=4).cpu().numpy()
            assert np.array_equal(data, ref)
    st = srv.stats()
    assert st.completed == 48 and st.failed == 0 and st.p99_ms >= st.p50_ms > 0
    if mode == 0:
        assert st.launches < 48            # queries were batched
    srv.close()
    for a in alone:
        a.release()


def test_lru_eviction_under_budget(rt, gpu_ctx, scenes):
    probe = gpu_ctx.upload(scenes[0])
    b0 = probe.info.device_bytes
    probe.release()
    srv = rt.Server(gpu_ctx, slots=2, hbm_budget=int(b0 * 2.5))
    a = srv.add_scene(scenes[0])
    b = srv.add_scene(scenes[0])
    rays = rt.rays_box(1000, 1, 4.0)
    hits = np.zeros(1000, rt.HIT_DT)
    srv.wait(srv.submit_query(a, rays, hits))          # a used after b: b is least recent
    c = srv.add_scene(scenes[0])
    assert sorted(srv.scenes()) == sorted([a, c])
    st = srv.stats()
    assert st.scenes_evicted == 1 and st.resident_bytes <= int(b0 * 2.5)
    with pytest.raises(rt.RtError) as e:
        srv.submit_query(b, rays, hits)
    assert e.value.status == rt.RT_ERR_NOT_FOUND
    srv.close()
    tiny = rt.Server(gpu_ctx, slots=1, hbm_budget=16)    # no scene fits
    with pytest.raises(rt.RtError) as e:
        tiny.add_scene(scenes[0])
    assert e.value.status == rt.RT_ERR_NOMEM
    tiny.close()


def test_short_poisson_run(rt, gpu_ctx, scenes):
    srv = rt.Server(gpu_ctx, slots=3)
    for hs in scenes:
        srv.add_scene(hs)
    print("case: poisson 500/s for 2 s", flush=True)
    r = srv.poisson(rate=500.0, duration=2.0, render_fraction=0.1, swap_interval=1.0, query_max=65536,
                    render_w=128, render_h=128, swap_triangles=50000, threads=8)
    st = r.stats
    print(f"requests {r.requests} completed {st.completed} p50 {st.p50_ms:.2f} p99 {st.p99_ms:.2f} ms "
          f"launches {st.launches} swaps {r.swaps}", flush=True)
    assert r.requests > 500 and st.completed == r.requests and st.failed == 0
    assert r.swaps >= 1 and st.rays > 0
    srv.close()


def test_admission_before_allocation(rt, gpu_ctx, scenes):
    """A swap never holds more than the budget: the new scene's bytes are
    known before it is allocated, and the evicted scenes are freed first.
    Checked by the server's own peak and by the device's free memory."""
    probe = gpu_ctx.upload(scenes[1])
    b1 = probe.info.device_bytes
    probe.release()
    budget = int(b1 * 2.5)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    srv = rt.Server(gpu_ctx, slots=2, hbm_budget=budget)
    free0, _ = torch.cuda.mem_get_info()
    slack = 96 << 20                  # context workspaces touched by the first queries
    rays = rt.rays_box(2000, 1, 4.0)
    hits = np.zeros(2000, rt.HIT_DT)
    print("case: server admission before allocation, five adds under 2.5 scenes of budget", flush=True)
    for k in range(5):
        sid = srv.add_scene(scenes[1])
        assert srv.wait(srv.submit_query(sid, rays, hits)) == rt.RT_OK
        free, _ = torch.cuda.mem_get_info()
        st = srv.stats()
        assert st.resident_bytes <= budget and st.peak_resident_bytes <= budget
        assert free0 - free <= budget + slack, (k, free0 - free, budget)
    st = srv.stats()
    assert st.scenes_added == 5 and st.scenes_evicted == 3
    srv.close()


def test_binary_records_built_on_demand(rt, gpu_ctx, scenes):
    """Uploads leave out the binary node records unless the context's layout
    reads them; a query in such a layout builds them on first use, with the
    same hits as an upload that made them at once."""
    hs = scenes[0]
    by_layout = {}
    for lay in (rt.LAYOUT_W4, rt.LAYOUT_Q16, rt.LAYOUT_F32):
        gpu_ctx.set_tuning(node_layout=lay)
        try:
            sc = gpu_ctx.upload(hs)
        finally:
            gpu_ctx.set_tuning()
        by_layout[lay] = sc.info.device_bytes
        sc.release()
    dq = by_layout[rt.LAYOUT_Q16] - by_layout[rt.LAYOUT_W4]
    df = by_layout[rt.LAYOUT_F32] - by_layout[rt.LAYOUT_W4]
    assert dq > 0 and df == 2 * dq                       # 32-byte and 64-byte records of one tree
    rays = rt.rays_to_device(rt.rays_box(20000, 5, 4.0))
    lazy = gpu_ctx.upload(hs)
    assert lazy.info.device_bytes == by_layout[rt.LAYOUT_W4]
    for lay in (rt.LAYOUT_Q16, rt.LAYOUT_F32):
        gpu_ctx.set_tuning(node_layout=lay)
        try:
            got = rt.hits_from_device(gpu_ctx.closest_hit(lazy, rays))
            eager = gpu_ctx.upload(hs)
            want = rt.hits_from_device(gpu_ctx.closest_hit(eager, rays))
            eager.release()
        finally:
            gpu_ctx.set_tuning()
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
    import ctypes
    info = rt.SceneInfo()
    assert rt.lib().rt_scene_info_get(lazy.handle, ctypes.byref(info)) == rt.RT_OK
    assert info.device_bytes == by_layout[rt.LAYOUT_W4] + dq + df
    lazy.release()


def test_cancel_and_deadline(rt, gpu_ctx, scenes):
    """Cancellation and deadlines drop queued requests without running them:
    their status is RT_ERR_CANCELLED and the others complete normally."""
    srv = rt.Server(gpu_ctx, mode=rt.SERVE_FIFO)
    sid = srv.add_scene(scenes[2])
    big = np.zeros((1024, 1024, 4), np.uint8)
    first = srv.submit_render(sid, CAM, 1024, 1024, big, bounces=8)
    rays = rt.rays_box(50000, 3, 4.0)
    outs = [np.zeros(50000, rt.HIT_DT) for _ in range(6)]
    rids = [srv.submit_query(sid, rays, h) for h in outs]
    print("case: cancel a queued request behind a long render", flush=True)
    assert srv.cancel(rids[-1]) == rt.RT_OK
    assert srv.cancel(rids[-1]) == rt.RT_ERR_NOT_FOUND
    assert srv.wait(rids[-1]) == rt.RT_ERR_CANCELLED
    assert srv.wait(first) == rt.RT_OK
    alone = gpu_ctx.upload(scenes[2])
    ref = rt.hits_from_device(gpu_ctx.closest_hit(alone, rt.rays_to_device(rays)))
    for rid, h in zip(rids[:-1], outs[:-1]):
        assert srv.wait(rid) == rt.RT_OK
        assert np.array_equal(h.view(np.uint8), ref.view(np.uint8))
    st = srv.stats()
    assert st.cancelled == 1 and st.completed == 6 and st.failed == 0
    srv.close()
    # a 1 ms deadline behind a long render: the queued queries expire
    srv = rt.Server(gpu_ctx, mode=rt.SERVE_FIFO, deadline_ms=1)
    sid = srv.add_scene(scenes[2])
    first = srv.submit_render(sid, CAM, 1024, 1024, big, bounces=8)
    rids = [srv.submit_query(sid, rays, h) for h in outs]
    print("case: 1 ms deadline behind a long render", flush=True)
    status = [srv.wait(r) for r in [first] + rids]
    assert set(status) <= {rt.RT_OK, rt.RT_ERR_CANCELLED} and rt.RT_ERR_CANCELLED in status
    for s, h in zip(status[1:], outs):
        if s == rt.RT_OK:
            assert np.array_equal(h.view(np.uint8), ref.view(np.uint8))
    st = srv.stats()
    assert st.cancelled == status.count(rt.RT_ERR_CANCELLED) and st.completed + st.cancelled == 7
    srv.close()
    alone.release()
