"""GPU parity of the closest-hit query (hit_bvh, ray-trace-compute.comp:263-311)
through the C-ABI, against the oracle on the same seeded inputs.

Bar: the packed ordered kernel reproduces the oracle's ordered traversal bit
for bit (id, t, u, v and per-ray node/triangle counts); against the shader's
own visiting order the hit id may differ only where the two nearest hits lie
within 1e-6 in t, and t/u/v are bit-identical whenever the ids agree.  The
reference-order kernel reproduces the shader's order exactly, ties included.
"""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def oscene(hs):
    return O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres)


def run(ctx, scene, rays, **kw):
    import rtamd
    rd = rtamd.rays_to_device(rays)
    h = ctx.closest_hit(scene, rd, **kw)
    torch.cuda.synchronize()
    return rtamd.hits_from_device(h)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_bit_equal(got, want):
    assert np.array_equal(got["id"], want["id"]), np.nonzero(got["id"] != want["id"])[0][:10]
    for f in ("t", "u", "v"):
        assert np.array_equal(bits(got[f]), bits(want[f])), f


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


def mixed_rays(rt, n, seed, box=1.2):
    return np.concatenate([rt.rays_shell(n // 2, seed, 3.0, 0.5), rt.rays_box(n - n // 2, seed + 1, box)])


@pytest.fixture(scope="module")
def soup(rt, gpu_ctx):
    hs = rt.scene_soup(1024, 1)
    print("case: soup 1024 upload", flush=True)
    return hs, gpu_ctx.upload(hs)


def test_soup_packed_matches_oracle_ordered(rt, gpu_ctx, soup):
    """fp32 node layout: the kernel's traversal is the oracle's ordered
    traversal step for step (same box arithmetic), so hits and per-ray
    counts agree exactly."""
    hs, scene = soup
    assert scene.info.fast_path == 1
    rays = mixed_rays(rt, 65536, 11)
    print("case: soup packed fp32 vs ordered, 65536 rays", flush=True)
    gpu_ctx.set_tuning(node_layout=rt.LAYOUT_F32)
    try:
        got = run(gpu_ctx, scene, rays)
        want, nv, tt, _ = O.closest_hit(oscene(hs), rays, O.TRAV_ORDERED, stats=True, threads=8)
        assert_bit_equal(got, want)
        assert (got["id"] >= 0).mean() > 0.1
        hits, gnv, gtt = gpu_ctx.closest_hit_stats(scene, rt.rays_to_device(rays))
        torch.cuda.synchronize()
        assert np.array_equal(gnv.cpu().numpy().astype(np.uint32), nv)
        assert np.array_equal(gtt.cpu().numpy().astype(np.uint32), tt)
    finally:
        gpu_ctx.set_tuning()


@pytest.mark.parametrize("layout,lo,hi", [(2, 1.0, 1.3), (3, 0.3, 0.8)])
def test_soup_quantized_nodes(rt, gpu_ctx, soup, layout, lo, hi):
    """16-bit node boxes (binary 32-byte or 4-wide 64-byte records) contain
    the exact ones: same closest hits (up to exact ties); the binary layout
    reads a few more records than the exact-box traversal, the 4-wide one
    about half as many."""
    hs, scene = soup
    rays = mixed_rays(rt, 65536, 11)
    print(f"case: soup packed layout {layout} vs ordered", flush=True)
    gpu_ctx.set_tuning(node_layout=layout)
    try:
        got = run(gpu_ctx, scene, rays)
        want, nv, tt, _ = O.closest_hit(oscene(hs), rays, O.TRAV_ORDERED, stats=True, threads=8)
        assert assert_equal_up_to_ties(oscene(hs), rays, got, want) <= 2
        _, gnv, gtt = gpu_ctx.closest_hit_stats(scene, rt.rays_to_device(rays))
        torch.cuda.synchronize()
        ratio = gnv.cpu().numpy().astype(np.float64).sum() / nv.astype(np.float64).sum()
        print(f"layout {layout}: node records / exact binary records {ratio:.4f}", flush=True)
        assert lo <= ratio < hi
    finally:
        gpu_ctx.set_tuning()


def test_soup_packed_vs_shader_order(rt, gpu_ctx, soup):
    hs, scene = soup
    rays = mixed_rays(rt, 32768, 12)
    print("case: soup packed vs reference order", flush=True)
    got = run(gpu_ctx, scene, rays)
    want = O.closest_hit(oscene(hs), rays, O.TRAV_REFERENCE, threads=8)
    ties = assert_equal_up_to_ties(oscene(hs), rays, got, want)
    assert ties <= 2


def test_soup_reference_kernel_exact(rt, gpu_ctx, soup):
    import rtamd
    hs, scene = soup
    rays = mixed_rays(rt, 32768, 13)
    print("case: soup reference kernel", flush=True)
    got = run(gpu_ctx, scene, rays, traversal=rtamd.TRAVERSE_REFERENCE)
    want = O.closest_hit(oscene(hs), rays, O.TRAV_REFERENCE, threads=8)
    assert_bit_equal(got, want)


def test_spheres_closest_of_both(rt, gpu_ctx, soup):
    import rtamd
    hs, scene = soup
    rays = np.concatenate([rt.rays_camera_tiles([0, 3, 0], 60.0, 128, 128, 8, 0, 16384), mixed_rays(rt, 8192, 14)])
    print("case: soup spheres", flush=True)
    want = O.closest_hit(oscene(hs), rays, O.TRAV_ORDERED, spheres=1, threads=8)
    assert (want["id"] >= len(hs.triangles)).any()
    assert assert_equal_up_to_ties(oscene(hs), rays, run(gpu_ctx, scene, rays, spheres=1), want) <= 2
    want_ref = O.closest_hit(oscene(hs), rays, O.TRAV_REFERENCE, spheres=1, threads=8)
    assert_bit_equal(run(gpu_ctx, scene, rays, spheres=1, traversal=rtamd.TRAVERSE_REFERENCE), want_ref)


def test_edge_rays(rt, gpu_ctx, soup):
    """Axis-aligned directions (zero components), origins inside boxes,
    tiny tmax (all miss), and rays starting on a triangle."""
    from rtamd import RAY_DT
    hs, scene = soup
    r = np.zeros(4096, RAY_DT)
    rng = np.random.default_rng(5)
    axes = np.eye(3)[rng.integers(0, 3, 4096)] * rng.choice([-1.0, 1.0], 4096)[:, None]
    r["o"] = rng.uniform(-1.2, 1.2, (4096, 3))
    r["d"] = axes
    r["tmin"] = 0.001
    r["tmax"] = np.where(np.arange(4096) % 7 == 0, 0.002, 10000.0)
    t0 = hs.triangles[:512]
    r["o"][:512] = (t0["v0"] + t0["v1"] + t0["v2"]) / 3          # start on a triangle
    print("case: soup edge rays", flush=True)
    want = O.closest_hit(oscene(hs), r, O.TRAV_ORDERED, threads=8)
    assert assert_equal_up_to_ties(oscene(hs), r, run(gpu_ctx, scene, r), want) <= 2
    ref = O.closest_hit(oscene(hs), r, O.TRAV_REFERENCE, threads=8)
    assert_equal_up_to_ties(oscene(hs), r, run(gpu_ctx, scene, r), ref)


@pytest.mark.parametrize("layout", ["w4", "q16"])
def test_far_origins(rt, gpu_ctx, soup, layout):
    """Origins far outside the scene in units of the 16-bit frame's quantum
    (the soup spans ~2.2 units, a quantum ~3.4e-5): near the distance where
    the kernel starts moving the origin along the ray (2^21 quanta, ~70
    units), far past it (1e2 .. 1e5 units, up to ~3e9 quanta), rays aimed at
    the scene from there, rays aimed away, and rays whose tmax ends short of
    it.  The packed kernels must give the oracle's hits (exact float boxes)."""
    from rtamd import RAY_DT
    hs, scene = soup
    rng = np.random.default_rng(17)
    parts = []
    for dist, n in ((60.0, 4096), (75.0, 4096), (1e2, 4096), (1e3, 4096), (1e5, 4096)):
        r = np.zeros(n, RAY_DT)
        u = rng.normal(size=(n, 3))
        o = u / np.linalg.norm(u, axis=1, keepdims=True) * dist
        target = rng.uniform(-1.1, 1.1, (n, 3))
        d = target - o
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        d[: n // 8] *= -1.0                                    # aimed away: all miss
        r["o"], r["d"] = o, d
        r["tmin"] = 0.001
        r["tmax"] = np.where(np.arange(n) % 9 == 0, 0.5 * dist, 10.0 * dist)   # every 9th ends short
        parts.append(r)
    r = np.concatenate(parts)
    print(f"case: far origins, {layout}, {len(r)} rays", flush=True)
    gpu_ctx.set_tuning(node_layout=rt.LAYOUT_W4 if layout == "w4" else rt.LAYOUT_Q16)
    try:
        got = run(gpu_ctx, scene, r)
    finally:
        gpu_ctx.set_tuning()
    want = O.closest_hit(oscene(hs), r, O.TRAV_REFERENCE, threads=8)
    assert (want["id"] >= 0).mean() > 0.3
    assert assert_equal_up_to_ties(oscene(hs), r, got, want) <= 4
    local = np.arange(len(r)) % 4096
    assert (got["id"][local < 512] == -1).all()
    assert (got["id"][local % 9 == 0] == -1).all()


def test_single_triangle_and_empty_batches(rt, gpu_ctx):
    hs = rt.scene_soup(1, 3)
    scene = gpu_ctx.upload(hs)
    t = hs.triangles[0]
    c = (t["v0"] + t["v1"] + t["v2"]) / 3
    from rtamd import RAY_DT
    rays = np.zeros(3, RAY_DT)
    rays["o"] = c + np.array([0, 0, 2.0])
    rays["d"] = [[0, 0, -1], [0, 0, 1], [0.001, 0, -1]]
    rays["tmin"], rays["tmax"] = 0.001, 10000.0
    print("case: single triangle", flush=True)
    got = run(gpu_ctx, scene, rays)
    assert_bit_equal(got, O.closest_hit(oscene(hs), rays, O.TRAV_REFERENCE))
    # n = 0 is a no-op
    empty = torch.zeros((0, 8), dtype=torch.float32, device="cuda")
    gpu_ctx.closest_hit(scene, empty)
    scene.release()


def test_non_enclosing_tree_is_traced_in_shader_order(rt, gpu_ctx):
    hs = rt.scene_soup(256, 8)
    leaf = np.nonzero(hs.nodes["object"] >= 0)[0][:40]
    hs.nodes["max"][leaf] = hs.nodes["min"][leaf] + 0.01     # boxes that miss parts of their triangles
    scene = gpu_ctx.upload(hs)
    assert scene.info.fast_path == 0
    rays = mixed_rays(rt, 16384, 15, box=0.8)
    print("case: non-enclosing tree", flush=True)
    assert_bit_equal(run(gpu_ctx, scene, rays), O.closest_hit(oscene(hs), rays, O.TRAV_REFERENCE, threads=8))
    scene.release()


def test_max_stack_16_truncation_mode(rt, gpu_ctx):
    import rtamd
    hs = rt.scene_soup(8193, 21)
    scene = gpu_ctx.upload(hs)
    rays = mixed_rays(rt, 8192, 16)
    print("case: max_stack 16 truncation", flush=True)
    want, _, _, tr = O.closest_hit(oscene(hs), rays, O.TRAV_REFERENCE, max_stack=16, stats=True, threads=8)
    got = run(gpu_ctx, scene, rays, traversal=rtamd.TRAVERSE_REFERENCE, max_stack=16)
    assert_bit_equal(got, want)
    scene.release()


def test_icosphere_one_million_rays(rt, gpu_ctx):
    """The closest-hit query workload: level-6 displaced icosphere, 1M rays
    (half coherent camera tiles, half incoherent).  Full size: packed kernel
    against the reference-order kernel; a 32k sample against the oracle."""
    import rtamd
    hs = rt.scene_icosphere(6, 2, 0.1, 4.0)
    assert len(hs.triangles) == 81920
    scene = gpu_ctx.upload(hs)
    coh = rt.rays_camera_tiles([0, 3, 0], 60.0, 1024, 1024, 8, 0, 524288)
    inc = rt.rays_shell(524288, 2024, 3.0, 0.5)
    rays = np.concatenate([coh, inc])
    print("case: icosphere 1M packed vs reference kernel", flush=True)
    got = run(gpu_ctx, scene, rays)
    ref = run(gpu_ctx, scene, rays, traversal=rtamd.TRAVERSE_REFERENCE)
    assert (got["id"][524288:] >= 0).all()                  # every shell ray crosses the closed surface
    diff = np.nonzero(got["id"] != ref["id"])[0]
    assert len(diff) <= 50
    same = got["id"] == ref["id"]
    for f in ("t", "u", "v"):
        assert np.array_equal(bits(got[f][same]), bits(ref[f][same]))
    sample = np.concatenate([np.arange(0, 524288, 32), np.arange(524288, 1048576, 32)])
    print("case: icosphere sample vs oracle", flush=True)
    want = O.closest_hit(oscene(hs), rays[sample], O.TRAV_ORDERED, threads=8)
    assert assert_equal_up_to_ties(oscene(hs), rays[sample], got[sample], want) <= 5
    assert_equal_up_to_ties(oscene(hs), rays[sample], got[sample],
                            O.closest_hit(oscene(hs), rays[sample], O.TRAV_REFERENCE, threads=8))
    scene.release()


def test_invalid_scene_rejected_by_upload(rt, gpu_ctx):
    import rtamd
    hs = rt.scene_soup(16, 1)
    hs.nodes["left"][0] = 5000
    with pytest.raises(rtamd.RtError) as e:
        gpu_ctx.upload(hs)
    assert e.value.status == rtamd.RT_ERR_INVALID and "child index" in str(e.value)


def test_ray_count_limits(rt, gpu_ctx):
    """Ray and slot counts at 2^31 and above are refused before any device
    work (the kernels' 32-bit grid-stride indices need fewer).  The buffers
    are null, which would be refused too, so nothing is ever launched."""
    import ctypes
    import rtamd
    L = rtamd.lib()
    hs = rt.scene_soup(16, 1)
    scene = gpu_ctx.upload(hs)
    print("case: ray count limits", flush=True)
    for n in (1 << 31, 0xffffffff):
        rc = L.rt_closest_hit(gpu_ctx.handle, scene.handle, None, n, None, None, None)
        assert rc == rtamd.RT_ERR_INVALID
        assert b"2^31" in L.rt_last_error(gpu_ctx.handle)
    with pytest.raises(rtamd.RtError) as e:
        rtamd.Server(gpu_ctx, max_batch_rays=(1 << 30) + 1)
    assert e.value.status == rtamd.RT_ERR_INVALID and "2^30" in str(e.value)
    scene.release()


def test_trace_order_is_invisible(rt, gpu_ctx):
    """RT_ORDER_SORTED traces the rays in Morton order of their origins:
    hits and per-ray counters are byte-identical to index order."""
    hs = rt.scene_clusters(1 << 18, 21)
    scene = gpu_ctx.upload(hs)
    rays = rt.rays_to_device(rt.rays_box(300000, 22, 4.0))
    print("case: 300k rays, sorted vs given trace order", flush=True)
    out = {}
    for mode in (rt.ORDER_GIVEN, rt.ORDER_SORTED, rt.ORDER_AUTO):
        gpu_ctx.set_tuning(ray_order=mode)
        try:
            h = rt.hits_from_device(gpu_ctx.closest_hit(scene, rays))
            _, nv, tt = gpu_ctx.closest_hit_stats(scene, rays)
            torch.cuda.synchronize()
            out[mode] = (h, nv.cpu().numpy(), tt.cpu().numpy())
        finally:
            gpu_ctx.set_tuning()
    for mode in (rt.ORDER_SORTED, rt.ORDER_AUTO):
        assert np.array_equal(out[mode][0].view(np.uint8), out[rt.ORDER_GIVEN][0].view(np.uint8))
        assert np.array_equal(out[mode][1], out[rt.ORDER_GIVEN][1])
        assert np.array_equal(out[mode][2], out[rt.ORDER_GIVEN][2])


@pytest.mark.parametrize("origins", ["one", "few", "corners"])
def test_sorted_order_traces_every_ray_once(rt, gpu_ctx, origins):
    """The sorted trace order on degenerate origin sets: every ray from one
    origin (one key for all), from 5 origins (long runs of equal keys), and
    from the frame's corners and beyond (clamped cells).  Every ray is traced
    exactly once: hits equal index order's byte for byte over a hit buffer
    pre-filled with a sentinel."""
    hs = rt.scene_clusters(1 << 18, 24)
    scene = gpu_ctx.upload(hs)
    n = 200003                                     # not a multiple of the wave
    r = rt.rays_box(n, 25, 4.0)
    rng = np.random.default_rng(26)
    if origins == "one":
        r["o"] = np.array([0.3, -0.2, 0.1], np.float32)
    elif origins == "few":
        pts = rng.uniform(-3.0, 3.0, (5, 3)).astype(np.float32)
        r["o"] = pts[rng.integers(0, 5, n)]
    else:
        r["o"] = rng.choice([-9.0, 9.0], (n, 3)).astype(np.float32)
        d = -r["o"] / np.linalg.norm(r["o"], axis=1, keepdims=True)
        r["d"] = (d + rng.normal(scale=0.05, size=(n, 3))).astype(np.float32)
    rays = rt.rays_to_device(r)
    print(f"case: sorted trace order, origins {origins}, {n} rays", flush=True)
    out = {}
    for mode in (rt.ORDER_GIVEN, rt.ORDER_SORTED):
        gpu_ctx.set_tuning(ray_order=mode)
        try:
            hits = torch.full((n, 4), -7.0, dtype=torch.float32, device="cuda")
            gpu_ctx.closest_hit(scene, rays, hits)
            torch.cuda.synchronize()
            out[mode] = hits.cpu().numpy()
        finally:
            gpu_ctx.set_tuning()
    assert not (out[rt.ORDER_SORTED][:, 1] == -7.0).any()
    assert np.array_equal(out[rt.ORDER_SORTED].view(np.uint8), out[rt.ORDER_GIVEN].view(np.uint8))
    assert (out[rt.ORDER_GIVEN][:, 0].view(np.int32) >= 0).mean() > 0.05


def test_query_stage_timing(rt, gpu_ctx):
    hs = rt.scene_clusters(1 << 18, 23)
    scene = gpu_ctx.upload(hs)
    rays = rt.rays_to_device(rt.rays_box(200000, 24, 4.0))
    print("case: query stage timing", flush=True)
    gpu_ctx.set_render_timing(True)
    got = {}
    try:
        for mode, sorted_ in ((rt.ORDER_SORTED, 1), (rt.ORDER_GIVEN, 0)):
            gpu_ctx.set_tuning(ray_order=mode)
            gpu_ctx.closest_hit(scene, rays)
            st = gpu_ctx.last_query_stats()
            assert st.timed == 1 and st.sorted == sorted_ and st.trace_ms > 0
            assert st.order_ms + st.trace_ms + st.spheres_ms <= st.total_ms * (1 + 1e-5) + 1e-3
            got[sorted_] = st.order_ms
        assert got[1] > got[0]                      # the key kernel and the sort take time
    finally:
        gpu_ctx.set_render_timing(False)
        gpu_ctx.set_tuning()
    gpu_ctx.closest_hit(scene, rays)
    assert gpu_ctx.last_query_stats().timed == 0


@pytest.mark.parametrize("n", [65535, 65536, 65537, 100003, 262151])
def test_every_ray_traced_once_at_odd_sizes(rt, gpu_ctx, n):
    """Segmented cursors and sorted order: at sizes around the thresholds and
    not divisible by the segment count, every ray gets its own hit (the
    buffer starts as a sentinel) and the hits equal the shader-order ones."""
    hs = rt.scene_clusters(1 << 16, 25)
    scene = gpu_ctx.upload(hs)
    rays_np = rt.rays_box(n, 26 + n % 7, 4.0)
    rays = rt.rays_to_device(rays_np)
    print(f"case: {n} rays, sentinel-filled hits", flush=True)
    for mode in (rt.ORDER_SORTED, rt.ORDER_GIVEN):
        gpu_ctx.set_tuning(ray_order=mode)
        try:
            hits = torch.full((n, 4), float("nan"), device="cuda")
            gpu_ctx.closest_hit(scene, rays, hits)
            got = rt.hits_from_device(hits)
        finally:
            gpu_ctx.set_tuning()
        assert not np.isnan(got["t"]).any()
        ref = rt.hits_from_device(gpu_ctx.closest_hit(scene, rays, traversal=rt.TRAVERSE_REFERENCE))
        assert_equal_up_to_ties(oscene(hs), rays_np, got, ref)


def tie_grid(rt, n=64, layers=3, seed=30):
    """Unit squares on the planes z = 0, -1, ..., each cut into two right
    triangles along alternating diagonals, in shuffled order.  Every normal
    is (0, 0, +-1), so with dyadic rays the triangle test is exact: all
    triangles of one plane give a ray the same t, and a ray through an edge
    or a vertex has two or more hits at exactly equal t."""
    i, j = np.meshgrid(np.arange(n, dtype=np.float32), np.arange(n, dtype=np.float32), indexing="ij")
    i, j = i.ravel(), j.ravel()
    flip = ((i + j) % 2 == 1)[:, None, None]
    parts = []
    for layer in range(layers):
        z = np.full_like(i, -float(layer))
        a, b = np.stack([i, j, z], 1), np.stack([i + 1, j, z], 1)
        c, d = np.stack([i + 1, j + 1, z], 1), np.stack([i, j + 1, z], 1)
        parts.append(np.where(flip, np.stack([a, b, c], 1), np.stack([a, b, d], 1)))
        parts.append(np.where(flip, np.stack([a, c, d], 1), np.stack([b, c, d], 1)))
    v = np.concatenate(parts)[np.random.default_rng(seed).permutation(2 * n * n * layers)]
    t = np.zeros(len(v), rt.TRI_DT)
    t["v0"], t["v1"], t["v2"] = v[:, 0], v[:, 1], v[:, 2]
    return t


def tie_rays(rt, m, n=64, seed=31):
    """Dyadic rays from z = 4 to points on the grid's edges and vertices."""
    rng = np.random.default_rng(seed)
    r = np.zeros(m, rt.RAY_DT)
    o = np.stack([rng.integers(0, 8 * n, m) / 8.0, rng.integers(0, 8 * n, m) / 8.0, np.full(m, 4.0)], 1)
    kind = rng.integers(0, 3, m)
    tx = rng.integers(1, n - 1, m).astype(np.float64)
    ty = rng.integers(1, n - 1, m).astype(np.float64)
    frac = rng.integers(0, 4, m) / 4.0
    tx = np.where(kind == 0, tx + frac, tx)         # on a line y = integer
    ty = np.where(kind == 1, ty + frac, ty)         # on a line x = integer; kind 2: a vertex
    r["o"] = o
    r["d"] = np.stack([tx - o[:, 0], ty - o[:, 1], np.full(m, -4.0)], 1)
    r["tmin"] = 0.001
    r["tmax"] = 10000.0
    return r


@pytest.mark.parametrize("case", ["tie_grid", "clusters"])
def test_drain_work_sharing_is_invisible(rt, gpu_ctx, case):
    """When a wave's rays run out, idle lanes traverse subtrees taken from
    busy lanes' stacks and hand their hits back.  The hits must be the ones
    of the whole traversal in one lane -- the counting variant, which never
    shares -- byte for byte, exact ties included: on the tie grid most rays
    end on an edge or a vertex shared by two or more triangles at equal t."""
    if case == "tie_grid":
        tris = tie_grid(rt)
        mats = np.zeros(1, rt.MAT_DT)
        hs = rt.HostScene(tris, mats, rt.bvh_build(tris, seed=1))
        rays_np = tie_rays(rt, 200000)
    else:
        hs = rt.scene_clusters(1 << 18, 21)
        rays_np = rt.rays_box(300000, 22, 4.0)
    scene = gpu_ctx.upload(hs)
    assert scene.info.fast_path == 1
    rays = rt.rays_to_device(rays_np)
    with pytest.raises(rt.RtError):
        gpu_ctx.set_tuning(steal=rt.STEAL_NEVER + 1)
    alone, _, _ = gpu_ctx.closest_hit_stats(scene, rays)
    torch.cuda.synchronize()
    alone = rt.hits_from_device(alone)
    # steal 1: every idle lane takes work at once; 0: the default threshold
    for steal in (0, 1, rt.STEAL_NEVER):
        for mode in (rt.ORDER_GIVEN, rt.ORDER_SORTED):
            print(f"case: drain work sharing, {case}, {len(rays_np)} rays, steal {steal}, order {mode}", flush=True)
            gpu_ctx.set_tuning(ray_order=mode, steal=steal)
            try:
                shared = rt.hits_from_device(gpu_ctx.closest_hit(scene, rays))
            finally:
                gpu_ctx.set_tuning()
            assert np.array_equal(shared.view(np.uint8), alone.view(np.uint8)), \
                np.nonzero(shared["id"] != alone["id"])[0][:10]
    if case == "tie_grid":
        # every ray hits the top plane on an edge or a vertex; against the
        # oracle's ordered traversal (2-wide visiting order) the ids may
        # differ only by such a tie
        assert (shared["id"] >= 0).all()
        assert (shared["t"] == 1.0).all()
        sample = rays_np[::50]
        want = O.closest_hit(oscene(hs), sample, O.TRAV_ORDERED, threads=8)
        ties = assert_equal_up_to_ties(oscene(hs), sample, shared[::50], want)
        print(f"tie grid: {ties} of {len(sample)} sampled hits differ from the binary order by a tie", flush=True)


@pytest.mark.parametrize("scene_seed,ray_seed,first", [(3, 4, 0), (5, 7, 31)])
def test_bench_configuration_against_the_oracle(rt, gpu_ctx, scene_seed, ray_seed, first):
    """The bench workload at full size -- 4,194,304-triangle clustered soup,
    4,194,304 incoherent rays: the bench's own seeds (3, 4) and a second pair
    -- through the default path (sorted trace order, XCD segments, 4-wide
    nodes), checked against the oracle's ordered traversal on every 64th ray
    (from `first`): ids equal up to near-ties, t/u/v bit-equal where they
    agree."""
    hs = rt.scene_clusters(1 << 22, scene_seed)
    scene = gpu_ctx.upload(hs)
    rays_np = rt.rays_box(1 << 22, ray_seed, 4.0)
    print(f"case: bench-size scene (seed {scene_seed}), 4M rays (seed {ray_seed}), oracle on every 64th from {first}",
          flush=True)
    got = rt.hits_from_device(gpu_ctx.closest_hit(scene, rt.rays_to_device(rays_np)))
    scene.release()
    sample = rays_np[first::64]
    want = O.closest_hit(oscene(hs), sample, O.TRAV_ORDERED, threads=16)
    assert (got["id"][first::64] >= 0).mean() > 0.2
    assert assert_equal_up_to_ties(oscene(hs), sample, got[first::64], want) <= 4
