"""rt_closest_hit_noalloc: the query that allocates nothing, against rt_closest_hit.

Called through the C ABI on a 2^20-triangle clustered soup (large enough
for the sorted trace order), twice in a row -- first the maximum 4,194,304
rays, then a smaller batch -- plus a batch below the sort threshold.  Each
call's hits must equal rt_closest_hit's byte for byte, the hit slots past n
must stay untouched, and the device's free memory must not move across the
calls (the scratch is set up when the context is created).  Sizes beyond
the maximum and configurations that would need a workspace are refused.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def noalloc(rt, ctx, scene, rays, n, hits):
    p = rt.QueryParams(0, 0, 0, 0)
    return rt.lib().rt_closest_hit_noalloc(ctx._h, scene.handle, rt._dev_ptr(rays), n, rt._dev_ptr(hits),
                                           ctypes.byref(p), None)


def test_noalloc_query_equals_closest_hit(rt, gpu_ctx):
    hs = rt.scene_clusters(1 << 20, 41, threads=16)
    sc = gpu_ctx.upload(hs)
    big = 1 << 22
    rays = rt.rays_to_device(rt.rays_box(big, 42, 4.0))
    # the existing entry point runs first in each round: it loads the kernels
    # and grows its own workspaces, so the free memory measured around the
    # variant's call is what the variant alone changes
    for n in (big, 300001, 1000):
        print(f"case: noalloc query, {n} rays on 2^20 triangles", flush=True)
        want = gpu_ctx.closest_hit(sc, rays[:n])
        hits = torch.full((big, 4), -3.0, device="cuda")
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        rc = noalloc(rt, gpu_ctx, sc, rays, n, hits)
        assert rc == rt.RT_OK, rt.lib().rt_last_error(gpu_ctx._h)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert free0 == free1, (n, free0 - free1)
        assert torch.equal(hits[:n].view(torch.int32), want.view(torch.int32)), f"n={n}: hits differ"
        assert bool((hits[n:] == -3.0).all()), f"n={n}: slots past n written"
    # beyond the maximum: refused before any device work
    hits = torch.full((big, 4), -3.0, device="cuda")
    assert noalloc(rt, gpu_ctx, sc, rays, big + 1, hits) == rt.RT_ERR_INVALID
    assert b"4,194,304" in rt.lib().rt_last_error(gpu_ctx._h)
    # an LDS ring spills to a workspace: refused
    gpu_ctx.set_tuning(stack_lds=16)
    try:
        assert noalloc(rt, gpu_ctx, sc, rays, 1000, hits) == rt.RT_ERR_UNSUPPORTED
    finally:
        gpu_ctx.set_tuning()
    torch.cuda.synchronize()
    assert bool((hits == -3.0).all())
    sc.release()


def test_noalloc_second_context_and_spheres(rt, gpu_ctx):
    """A second context on the same device shares the device's scratch; a
    scene with spheres runs its sphere pass after the traversal."""
    hs = rt.scene_soup(4096, 7)
    sc = gpu_ctx.upload(hs)
    other = rt.Context(0)
    rays = rt.rays_to_device(np.concatenate([rt.rays_shell(20000, 8), rt.rays_box(20000, 9, 1.0)]))
    n = rays.shape[0]
    for spheres in (0, 1, 2):
        want = gpu_ctx.closest_hit(sc, rays, spheres=spheres)
        for c in (gpu_ctx, other):
            hits = torch.empty((n, 4), device="cuda")
            p = rt.QueryParams(0, 0, spheres, 0)
            rc = rt.lib().rt_closest_hit_noalloc(c._h, sc.handle, rt._dev_ptr(rays), n, rt._dev_ptr(hits),
                                                 ctypes.byref(p), None)
            assert rc == rt.RT_OK, rt.lib().rt_last_error(c._h)
            torch.cuda.synchronize()
            assert torch.equal(hits.view(torch.int32), want.view(torch.int32)), spheres
    other.close()
    sc.release()
