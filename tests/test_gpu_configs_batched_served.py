"""The batched and served workloads at their full sizes, anchored to the oracle.

Batched: the 64 clustered-soup scenes (triangles log-uniform in [1k, 1M],
seeds 100+i) with ray batches log-uniform in [4k, 2M] (seeds 200+i) in one
autotuned persistent launch replayed from a hipGraph.  Every item's hits must
equal its own plain launch byte for byte, and every 256th ray of every item
must match the oracle's ordered traversal up to exact ties.

Served: the 16 resident clustered-soup scenes (log-uniform [10k, 2M], seeds
300-315) under an HBM budget of 1.05x their footprint, Poisson(4000/s)
arrivals for 3 s (10% 512x512 1-spp 4-bounce renders, the rest queries of
log-uniform [1k, 256k] rays) with a 500k-triangle scene swapped in every
second (least recently used eviction).  Every kept request must equal its
result alone, byte for byte, its query rays must match the oracle up to
exact ties, a kept render must match the oracle's render (RMSE <= 1e-3 on
sampled rows), and p99 latency must stay within 20 ms.

The arithmetic all of these reproduce is hit_bvh (ray-trace-compute.comp:263-311)
and, for renders, main/ray_color (ray-trace-compute.comp:314-380).
"""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def oscene(hs):
    return O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_equal_up_to_ties(sc, rays, got, want):
    diff = np.nonzero(got["id"] != want["id"])[0]
    for i in diff:
        ts = []
        for j in (int(got["id"][i]), int(want["id"][i])):
            assert j >= 0, (i, got[i], want[i])
            ts.append(O.hit_triangle(sc.triangles[j:j + 1], rays["o"][i], rays["d"][i])[3])
        assert abs(ts[0] - ts[1]) <= 1e-6 * max(1.0, abs(ts[0])), (i, ts)
    same = got["id"] == want["id"]
    for f in ("t", "u", "v"):
        assert np.array_equal(bits(got[f][same]), bits(want[f][same])), f
    return len(diff)


def test_batched_workload_64_scenes(rt, gpu_ctx):
    items, plain, hosts, rays_np = [], [], [], []
    for i in range(64):
        hs = rt.scene_clusters(rt.log_uniform(100 + i, 1000, 1000000), 100 + i, threads=16)
        sc = gpu_ctx.upload(hs)
        c = rt.log_uniform(200 + i, 4096, 2097152)
        r = rt.rays_box(c, 200 + i, 4.0)
        rd = rt.rays_to_device(r)
        hosts.append(hs)
        rays_np.append(r[::256])
        items.append((sc, rd, torch.full((c, 4), -7.0, device="cuda")))
        plain.append(gpu_ctx.closest_hit(sc, rd))
    total = sum(it[1].shape[0] for it in items)
    tris = sum(len(h.triangles) for h in hosts)
    print(f"case: batched workload, 64 scenes, {tris} triangles, {total} rays", flush=True)
    assert tris > 5_000_000 and total > 15_000_000
    b = rt.Batch(gpu_ctx, items)
    best = b.autotune(trials=2)
    print(f"autotuned: {best.as_dict()}", flush=True)
    for _, _, h in items:
        h.fill_(-7.0)
    b.launch()
    torch.cuda.synchronize()
    ties = 0
    for i, ((sc, rd, hb), hp) in enumerate(zip(items, plain)):
        assert torch.equal(hb.view(torch.int32), hp.view(torch.int32)), f"item {i}: batched != plain"
        got = rt.hits_from_device(hb[::256])
        want = O.closest_hit(oscene(hosts[i]), rays_np[i], O.TRAV_ORDERED, threads=16)
        ties += assert_equal_up_to_ties(oscene(hosts[i]), rays_np[i], got, want)
    print(f"oracle: {sum(len(r) for r in rays_np)} sampled rays, {ties} exact ties", flush=True)
    assert ties <= 16
    b.destroy()
    for sc, _, _ in items:
        sc.release()


def test_served_workload_16_scenes(rt, gpu_ctx):
    hss = [rt.scene_clusters(rt.log_uniform(300 + i, 10000, 2000000), 300 + i, threads=16) for i in range(16)]
    probe = rt.Server(gpu_ctx)
    for h in hss:
        probe.add_scene(h)
    budget = int(probe.stats().resident_bytes * 1.05)
    probe.close()
    srv = rt.Server(gpu_ctx, hbm_budget=budget)
    by_id = {srv.add_scene(h): h for h in hss}
    print(f"case: served workload, 16 scenes, budget {budget} bytes, Poisson 4000/s for 3 s", flush=True)
    srv.poisson(rate=4000.0, duration=1.0, swap_interval=0.0, seed=11, threads=16)      # warm-up
    r, samples, swap_ids = srv.poisson_sampled(rate=4000.0, duration=3.0, every=97, capacity=40, swap_interval=1.0,
                                               swap_triangles=500000, seed=7, threads=16)
    st = r.stats
    print(f"requests {r.requests} completed {st.completed} failed {st.failed} p50 {st.p50_ms:.2f} ms "
          f"p99 {st.p99_ms:.2f} ms, swaps {r.swaps}, evicted {st.scenes_evicted}, kept {len(samples)}", flush=True)
    assert r.requests > 9000 and st.completed == r.requests and st.failed == 0
    assert r.swaps >= 1 and st.scenes_evicted >= 1 and st.resident_bytes <= budget
    assert st.p99_ms <= 20.0
    srv.close()
    for k, sid in enumerate(swap_ids):
        if sid:
            by_id[sid] = rt.scene_clusters(500000, 7 + 1000 + k, threads=16)
    pool = rt.rays_box(262144, 7 ^ 0x5EED, 4.0)
    kinds = {s["kind"] for s in samples}
    assert len(samples) >= 30 and kinds == {"query", "render"}
    uploaded = {}
    rendered_vs_oracle = False
    ties = 0
    for s in samples:
        hs = by_id[s["scene_id"]]
        if s["scene_id"] not in uploaded:
            uploaded[s["scene_id"]] = gpu_ctx.upload(hs)
        sc = uploaded[s["scene_id"]]
        if s["kind"] == "query":
            rays = pool[:s["count"]]
            alone = rt.hits_from_device(gpu_ctx.closest_hit(sc, rt.rays_to_device(rays)))
            assert np.array_equal(s["hits"].view(np.uint8), alone.view(np.uint8)), f"arrival {s['arrival']}"
            sample = rays[::max(1, len(rays) // 128)]
            want = O.closest_hit(oscene(hs), sample, O.TRAV_ORDERED, threads=16)
            ties += assert_equal_up_to_ties(oscene(hs), sample, s["hits"][::max(1, len(rays) // 128)], want)
        else:
            alone = gpu_ctx.render(sc, [0.0, 10.0, 0.0], 512, 512, vfov=60.0, bounces=4).cpu().numpy()
            assert np.array_equal(s["image"], alone), f"arrival {s['arrival']}"
            if not rendered_vs_oracle:
                ref, _ = O.render(oscene(hs), [0.0, 10.0, 0.0], 512, 512, vfov=60.0, bounces=4,
                                  traversal=O.TRAV_ORDERED, row_step=16, threads=16)
                d = (s["image"][::16].astype(np.float64) - ref[::16].astype(np.float64)) / 255.0
                rmse = float(np.sqrt((d * d).mean(axis=(0, 1))).max())
                print(f"render arrival {s['arrival']}: rmse {rmse:.2e} against the oracle", flush=True)
                assert rmse <= 1e-3
                rendered_vs_oracle = True
    assert ties <= 8
    for sc in uploaded.values():
        sc.release()
