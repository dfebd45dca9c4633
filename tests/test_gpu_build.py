"""GPU BVH builder (rt_bvh_build_device) against the oracle's restatement
of the same linear BVH, byte for byte, and the closest hits it gives
against those of the reference builder's tree."""
import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def node_bytes(a):
    return np.ascontiguousarray(a).view(np.uint8)


def soup_like(n, seed, degenerate=False):
    rng = np.random.default_rng(seed)
    t = np.zeros(n, O.TRI_DT)
    c = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    for k in ("v0", "v1", "v2"):
        t[k] = c + rng.uniform(-0.05, 0.05, (n, 3)).astype(np.float32)
    if degenerate:
        t["v0"][::2] = t["v0"][0]; t["v1"][::2] = t["v1"][0]; t["v2"][::2] = t["v2"][0]
    return t


@pytest.mark.parametrize("case", ["n1", "n2", "n3", "kat3", "soup1024", "degenerate777", "icosphere5",
                                  "clusters50k", "flat_axis"])
def test_device_tree_matches_oracle_bytes(rt, gpu_ctx, case):
    if case == "n1":
        t = soup_like(1, 1)
    elif case == "n2":
        t = soup_like(2, 2)
    elif case == "n3":
        t = soup_like(3, 3)
    elif case == "kat3":
        t = np.zeros(3, O.TRI_DT)
        t["v0"][1] = t["v1"][1] = t["v2"][1] = [1.0, 1.0, 1.0]
        t["v1"][2] = [0.5, 0.5, 0.5]
    elif case == "soup1024":
        t = rt.scene_soup(1024, 1).triangles
    elif case == "degenerate777":
        t = soup_like(777, 4, degenerate=True)
    elif case == "icosphere5":
        t = rt.scene_icosphere(5, 2, 0.1, 4.0).triangles
    elif case == "clusters50k":
        t = rt.scene_clusters(50000, 5).triangles
    else:                                           # all centres on one plane: zero extent in z
        t = soup_like(4096, 6)
        for k in ("v0", "v1", "v2"):
            t[k][:, 2] = 0.25
    print(f"case: device LBVH {case}, {len(t)} triangles", flush=True)
    got = gpu_ctx.build_bvh(t)
    want, _, _ = O.lbvh(t)
    assert len(got) == 2 * len(t) - 1
    assert np.array_equal(node_bytes(got), node_bytes(want))


def test_device_tree_is_deterministic(rt, gpu_ctx):
    t = rt.scene_clusters(200000, 7).triangles
    print("case: device LBVH twice, 200k triangles", flush=True)
    a = gpu_ctx.build_bvh(t)
    b = gpu_ctx.build_bvh(t)
    assert np.array_equal(node_bytes(a), node_bytes(b))


def test_device_tree_gives_the_same_hits(rt, gpu_ctx):
    """Same scene, reference tree vs device tree: hit ids equal except where
    the two nearest hits lie within 1e-6 in t; t/u/v bit-equal otherwise."""
    hs = rt.scene_clusters(1 << 20, 8)
    print("case: 1M-triangle clusters, reference tree vs device tree, 1M rays", flush=True)
    nodes = gpu_ctx.build_bvh(hs.triangles)
    ref = gpu_ctx.upload(hs)
    dev = gpu_ctx.upload(rt.HostScene(hs.triangles, hs.materials, nodes, hs.lights, hs.spheres))
    assert dev.info.fast_path == 1
    rays = rt.rays_to_device(rt.rays_box(1 << 20, 9, 4.0))
    a = rt.hits_from_device(gpu_ctx.closest_hit(ref, rays))
    b = rt.hits_from_device(gpu_ctx.closest_hit(dev, rays))
    torch.cuda.synchronize()
    diff = np.nonzero(a["id"] != b["id"])[0]
    tri = hs.triangles
    rays_np = rt.rays_box(1 << 20, 9, 4.0)
    for i in diff:
        ts = [O.hit_triangle(tri[j:j + 1], rays_np["o"][i], rays_np["d"][i])[3]
              for j in (int(a["id"][i]), int(b["id"][i]))]
        assert abs(ts[0] - ts[1]) <= 1e-6 * max(1.0, abs(ts[0])), (i, ts)
    same = a["id"] == b["id"]
    for f in ("t", "u", "v"):
        assert np.array_equal(a[f][same].view(np.uint32), b[f][same].view(np.uint32)), f
    assert len(diff) <= 8


def test_device_tree_large_scene_matches_oracle(rt, gpu_ctx):
    hs = rt.scene_clusters(1 << 22, 3)
    print("case: device LBVH 4M-triangle clusters vs oracle", flush=True)
    got = gpu_ctx.build_bvh(hs.triangles)
    want, _, _ = O.lbvh(hs.triangles)
    assert np.array_equal(node_bytes(got), node_bytes(want))


def test_device_build_rejects_bad_input(rt, gpu_ctx):
    t = torch.zeros(48, dtype=torch.uint8, device="cuda")
    nd = torch.zeros(48, dtype=torch.uint8, device="cuda")
    with pytest.raises(rt.RtError):
        gpu_ctx.build_bvh_device(t, 0, nd)


def test_upload_with_device_tree(rt, gpu_ctx):
    """RT_TREE_DEVICE: the packed traversal runs on the device-built tree,
    exactly as if the caller had passed that tree; the shader-order
    traversal keeps the caller's nodes."""
    hs = rt.scene_clusters(200000, 13)
    rays = rt.rays_to_device(rt.rays_box(1 << 18, 14, 4.0))
    print("case: upload with RT_TREE_DEVICE, 200k triangles", flush=True)
    explicit = gpu_ctx.upload(rt.HostScene(hs.triangles, hs.materials, gpu_ctx.build_bvh(hs.triangles), hs.lights))
    given = gpu_ctx.upload(hs)
    gpu_ctx.set_tuning(tree=rt.TREE_DEVICE)
    try:
        dev = gpu_ctx.upload(hs)
    finally:
        gpu_ctx.set_tuning()
    assert dev.info.fast_path == 1 and dev.info.num_nodes == len(hs.nodes)
    a = rt.hits_from_device(gpu_ctx.closest_hit(dev, rays))
    b = rt.hits_from_device(gpu_ctx.closest_hit(explicit, rays))
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    c = rt.hits_from_device(gpu_ctx.closest_hit(dev, rays, traversal=rt.TRAVERSE_REFERENCE))
    d = rt.hits_from_device(gpu_ctx.closest_hit(given, rays, traversal=rt.TRAVERSE_REFERENCE))
    assert np.array_equal(c.view(np.uint8), d.view(np.uint8))


def test_upload_without_nodes(rt, gpu_ctx):
    """A scene given without a BVH: RT_TREE_DEVICE builds one for both
    traversals; RT_TREE_GIVEN rejects it."""
    base = rt.scene_soup(1024, 1)
    bare = rt.HostScene(base.triangles, base.materials, None, base.lights, base.spheres)
    with pytest.raises(rt.RtError):
        gpu_ctx.upload(bare)
    print("case: upload without nodes, RT_TREE_DEVICE", flush=True)
    gpu_ctx.set_tuning(tree=rt.TREE_DEVICE)
    try:
        sc = gpu_ctx.upload(bare)
    finally:
        gpu_ctx.set_tuning()
    nodes = gpu_ctx.build_bvh(base.triangles)
    assert sc.info.num_nodes == len(nodes) and sc.info.fast_path == 1
    rays = np.concatenate([rt.rays_shell(4096, 15, 3.0, 0.5), rt.rays_box(4096, 16, 1.2)])
    want = O.closest_hit(O.Scene(base.triangles, base.materials, nodes, base.lights, base.spheres), rays,
                         O.TRAV_REFERENCE, threads=8)
    got = rt.hits_from_device(gpu_ctx.closest_hit(sc, rt.rays_to_device(rays), traversal=rt.TRAVERSE_REFERENCE))
    assert np.array_equal(got["id"], want["id"])
    assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))


def test_render_on_device_tree_matches_oracle(rt, gpu_ctx):
    """The render on a device-built tree (RT_TREE_DEVICE): the shader's
    paths on another tree, so the image equals the oracle's render on the
    reference builder's tree up to exact-tie choices.  Bar: RMSE <= 1e-3 per
    channel over two accumulated frames."""
    hs = rt.scene_soup(1024, 1)
    print("case: render 256x256 on the device tree", flush=True)
    gpu_ctx.set_tuning(tree=rt.TREE_DEVICE)
    try:
        sc = gpu_ctx.upload(hs)
    finally:
        gpu_ctx.set_tuning()
    tgt, _ = gpu_ctx.render_frames(sc, [0.0, 3.0, 0.0], 256, 256, 2, vfov=60.0, bounces=4)
    torch.cuda.synchronize()
    want, _ = O.render_frames(O.Scene(hs.triangles, hs.materials, hs.nodes, hs.lights, hs.spheres), [0.0, 3.0, 0.0],
                              256, 256, 2, vfov=60.0, bounces=4, threads=8)
    d = (tgt.cpu().numpy().astype(np.float64) - want.astype(np.float64)) / 255.0
    assert float(np.sqrt((d * d).mean(axis=(0, 1))).max()) <= 1e-3
    sc.release()
