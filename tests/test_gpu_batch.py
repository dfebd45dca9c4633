"""GPU tests of the batched form (rt_batch_*): one persistent launch over
many (scene, ray batch) items must give every item exactly the hits the
plain kernel gives it alone, for every launch configuration."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.fixture(scope="module")
def mixed(rt, gpu_ctx):
    """Six scenes of 1k..120k triangles, seven items (one scene twice, one
    empty batch, counts not multiples of 64)."""
    sizes = [1000, 3000, 17000, 60000, 120000, 5000]
    scenes = [gpu_ctx.upload(rt.scene_clusters(n, 100 + i)) for i, n in enumerate(sizes)]
    counts = [4096, 777, 0, 20000, 65536, 3001, 12345]
    which = [0, 1, 2, 3, 4, 5, 3]
    items, refs = [], []
    for k, (c, w) in enumerate(zip(counts, which)):
        rays = rt.rays_to_device(rt.rays_box(c, 200 + k, 4.0)) if c else torch.zeros((0, 8), device="cuda")
        hits = torch.full((c, 4), -7.0, device="cuda")
        items.append((scenes[w], rays, hits))
        refs.append(rt.hits_from_device(gpu_ctx.closest_hit(scenes[w], rays)) if c else None)
    torch.cuda.synchronize()
    return scenes, items, refs


def check(rt, items, refs):
    torch.cuda.synchronize()
    for (scene, rays, hits), ref in zip(items, refs):
        if ref is None:
            continue
        got = rt.hits_from_device(hits)
        assert np.array_equal(got["id"], ref["id"])
        for f in ("t", "u", "v"):
            assert np.array_equal(bits(got[f]), bits(ref[f]))


@pytest.mark.parametrize("cfg", [
    dict(),
    dict(use_graph=0),
    dict(regroup=1),
    dict(stack_lds=8),
    dict(stack_lds=16, regroup=1),
    dict(block=128, refill=16),
    dict(blocks_per_cu=2, refill=64),
    dict(regroup=2),
    dict(node_layout=2),
    dict(node_layout=2, block=128, regroup=2, regroup_interval=1 << 16),
])
def test_batch_matches_plain(rt, gpu_ctx, mixed, cfg):
    _, items, refs = mixed
    print(f"case: batch config {cfg}", flush=True)
    for _, _, hits in items:
        hits.fill_(-7.0)
    b = rt.Batch(gpu_ctx, items, rt.batch_config(**cfg))
    b.launch()
    check(rt, items, refs)
    b.launch()                      # replay
    check(rt, items, refs)
    b.destroy()


def test_batch_autotune_and_cache(rt, gpu_ctx, mixed, tmp_path):
    _, items, refs = mixed
    print("case: batch autotune", flush=True)
    b = rt.Batch(gpu_ctx, items)
    cache = str(tmp_path / "tune.txt")
    best = b.autotune(trials=2, cache_path=cache)
    assert best.block in (128, 256) and best.stack_lds in (0, 8, 16, 32) and 1 <= best.refill <= 64
    assert os.path.getsize(cache) > 0
    b.destroy()
    b2 = rt.Batch(gpu_ctx, items)
    again = b2.autotune(trials=2, cache_path=cache)           # read back from the cache
    assert again.as_dict() == {**best.as_dict(), "use_graph": again.use_graph}
    for _, _, hits in items:
        hits.fill_(-7.0)
    b2.launch()
    check(rt, items, refs)
    b2.destroy()


def test_plain_ring_stack_matches_full_stack(rt, gpu_ctx):
    hs = rt.scene_clusters(200000, 9)
    scene = gpu_ctx.upload(hs)
    rays = rt.rays_to_device(rt.rays_box(50000, 3, 4.0))
    print("case: plain ring stacks", flush=True)
    ref = rt.hits_from_device(gpu_ctx.closest_hit(scene, rays))
    for ring in (8, 16, 32):
        gpu_ctx.set_tuning(stack_lds=ring)
        try:
            got = rt.hits_from_device(gpu_ctx.closest_hit(scene, rays))
        finally:
            gpu_ctx.set_tuning()
        assert np.array_equal(got["id"], ref["id"]) and np.array_equal(bits(got["t"]), bits(ref["t"]))
    scene.release()


def test_batch_rejects_scene_without_packed_path(rt, gpu_ctx):
    hs = rt.scene_soup(64, 2)
    leaf = np.nonzero(hs.nodes["object"] >= 0)[0][:4]
    hs.nodes["max"][leaf] = hs.nodes["min"][leaf]
    scene = gpu_ctx.upload(hs)
    rays = rt.rays_to_device(rt.rays_box(128, 1, 1.0))
    hits = torch.empty((128, 4), device="cuda")
    with pytest.raises(rt.RtError) as e:
        rt.Batch(gpu_ctx, [(scene, rays, hits)])
    assert e.value.status == rt.RT_ERR_UNSUPPORTED
    scene.release()


@pytest.mark.parametrize("steal", [0, 1])
def test_batch_drain_sharing_on_ties(rt, gpu_ctx, steal):
    """Drained waves of a batched launch hand subtrees (with their item's
    scene) to idle lanes.  On the tie grid, where every ray ends on an edge
    or a vertex shared at equal t, and on a clustered soup in the same
    launch, every item's hits equal the counting kernel's (which never
    shares) byte for byte."""
    from test_gpu_query import tie_grid, tie_rays
    tris = tie_grid(rt)
    grid = gpu_ctx.upload(rt.HostScene(tris, np.zeros(1, rt.MAT_DT), rt.bvh_build(tris, seed=1)))
    soup = gpu_ctx.upload(rt.scene_clusters(60000, 31))
    items, refs = [], []
    for k, (scene, rays_np) in enumerate([(grid, tie_rays(rt, 50000, seed=41)),
                                          (soup, rt.rays_box(30000, 42, 4.0)),
                                          (grid, tie_rays(rt, 7777, seed=43))]):
        rays = rt.rays_to_device(rays_np)
        ref, _, _ = gpu_ctx.closest_hit_stats(scene, rays)
        items.append((scene, rays, torch.full((len(rays_np), 4), -7.0, device="cuda")))
        refs.append(ref)
    torch.cuda.synchronize()
    refs = [rt.hits_from_device(r) for r in refs]
    print(f"case: batched drain sharing on the tie grid, steal {steal}", flush=True)
    gpu_ctx.set_tuning(steal=steal)
    try:
        b = rt.Batch(gpu_ctx, items, rt.batch_config(stack_lds=0))
        b.launch()
        check(rt, items, refs)
        b.destroy()
    finally:
        gpu_ctx.set_tuning()
    grid.release()
    soup.release()


def test_batch_regroup_windows_and_layouts(rt, gpu_ctx):
    """Regrouping within windows of consecutive rays (an item spans several
    windows) and the binary 16-bit records: hits equal the plain kernel's."""
    scenes = [gpu_ctx.upload(rt.scene_clusters(n, 500 + i)) for i, n in enumerate((60000, 2000))]
    counts = [300001, 70000, 131072]
    items, refs = [], []
    for k, c in enumerate(counts):
        rays = rt.rays_to_device(rt.rays_box(c, 600 + k, 4.0))
        items.append((scenes[k % 2], rays, torch.full((c, 4), -7.0, device="cuda")))
        refs.append(rt.hits_from_device(gpu_ctx.closest_hit(scenes[k % 2], rays)))
    for cfg in (dict(regroup=1, regroup_interval=1 << 16), dict(regroup=2, regroup_interval=1 << 17),
                dict(regroup=2, regroup_interval=1 << 20), dict(regroup=1, node_layout=2, regroup_interval=1 << 16)):
        print(f"case: windows {cfg}", flush=True)
        for _, _, hits in items:
            hits.fill_(-7.0)
        b = rt.Batch(gpu_ctx, items, rt.batch_config(**cfg))
        b.launch()
        check(rt, items, refs)
        # a change of configuration re-uploads the item table before the next capture
        b.set_config(rt.batch_config(regroup=2, regroup_interval=1 << 16))
        for _, _, hits in items:
            hits.fill_(-7.0)
        b.launch()
        check(rt, items, refs)
        b.destroy()
    for sc in scenes:
        sc.release()


def test_batch_rejects_invalid_configs(rt, gpu_ctx, mixed):
    _, items, _ = mixed
    for cfg in (dict(regroup=3), dict(regroup=1, regroup_interval=1000), dict(regroup=1, regroup_interval=1 << 15),
                dict(node_layout=2, stack_lds=16), dict(node_layout=1)):
        print(f"case: invalid batch config {cfg}", flush=True)
        with pytest.raises(rt.RtError):
            rt.Batch(gpu_ctx, items, rt.batch_config(**cfg))
